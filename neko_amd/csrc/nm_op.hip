// Operator construction (box meshes or Neko-supplied arrays) and the
// C-ABI handle functions around it.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "nm_common.h"
#include "nm_kernels.h"
#include "nm_op.h"
#include "nm_rng.h"

extern "C" {
int nm_gll(int lx, double *z, double *w, double *d, double *dt);
int nm_box_mesh(int nelx, int nely, int nelz, double amp, uint64_t seed,
                double *vxyz);
int nm_dfrmd(int64_t nel, const double *vxyz, int *dfrmd);
int nm_box_gs_sizes(int nelx, int nely, int nelz, int lx, int perx, int pery,
                    int perz, int64_t *nlocal, int64_t *facet_offset);
int nm_box_gs_map(int nelx, int nely, int nelz, int lx, int perx, int pery,
                  int perz, int32_t *dg, int32_t *gd, int32_t *blk_len,
                  int32_t *blk_off, int64_t *nlocal_out,
                  int64_t *facet_offset_out, int64_t *nblks_out);
int64_t nm_box_mask_size(int nelx, int nely, int nelz, int lx, int perx,
                         int pery, int perz);
int nm_box_mask(int nelx, int nely, int nelz, int lx, int perx, int pery,
                int perz, int32_t *msk, int64_t *nmsk);
}

namespace nm {

// Seeded N(0,1) per index: the counter-based generator of the host set-up
// (64-bit mix of seed and index, Box-Muller on two uniforms), evaluated on
// the device; agrees with the host generator to the last few ulps of the
// transcendental functions.
__global__ void rng_normal_kernel(double *__restrict__ f, int64_t n, uint64_t seed) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double u1 = 1.0 - drng_uniform(seed, 2 * (uint64_t)i);
    const double u2 = drng_uniform(seed, 2 * (uint64_t)i + 1);
    f[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
}

Op::~Op() {
  for (void *p : owned) (void)hipFree(p);
}

template <typename T>
T *Op::alloc(int64_t count) {
  void *p = nullptr;
  if (count <= 0) count = 1;
  if (hipMalloc(&p, sizeof(T) * (size_t)count) != hipSuccess) return nullptr;
  owned.push_back(p);
  bytes += (int64_t)(sizeof(T) * (size_t)count);
  return (T *)p;
}
template double *Op::alloc<double>(int64_t);
template float *Op::alloc<float>(int64_t);
template int32_t *Op::alloc<int32_t>(int64_t);
template uint8_t *Op::alloc<uint8_t>(int64_t);

#define NM_ALLOC(var, T, cnt)                                        \
  do {                                                               \
    (var) = op->alloc<T>(cnt);                                       \
    if (!(var)) {                                                    \
      set_error("out of device memory (%lld x %zu B)",               \
                (long long)(cnt), sizeof(T));                        \
      return NM_ERR_NOMEM;                                           \
    }                                                                \
  } while (0)

template <typename T>
int Op::gs(T *u, const int *skip, hipStream_t s) const {
  GsArgs<T> a{u, blk_ptr, gidx, nblks, pair0, npairs, skip};
  a.variant = (gs_variant >= 1 && max_blk <= 8) ? gs_variant : 0;
  if (a.variant == 3) {
    a.seg_idx = seg_idx;
    a.seg_pl = seg_pl;
    a.nseg = nseg;
  }
  if (a.variant == 2 || a.variant == 3) {
    a.faces = gs_faces;
    a.nfaces = nfaces;
    a.lx = lx;
    a.rest = gs_rest;
    a.nrest = nrest;
  }
  return launch_gs<T>(a, s);
}
template int Op::gs<double>(double *, const int *, hipStream_t) const;
template int Op::gs<float>(float *, const int *, hipStream_t) const;

template <typename T>
int Op::ax_code() const {
  return std::is_same<T, double>::value ? epb64 : epb32;
}
template int Op::ax_code<double>() const;
template int Op::ax_code<float>() const;

template <typename T>
int Op::apply(T *w, const T *u, const int *skip, hipStream_t s, bool do_gs,
              bool do_mask, double *dot_part, const MaskRef *mk) const {
  const OpFields<T> &F = fields<T>();
  AxArgs<T> a;
  a.w = w;
  a.u = u;
  a.dx = F.dx;
  a.h1 = F.h1;
  a.h1c = (T)h1c;
  for (int c = 0; c < 6; c++) a.g[c] = F.g[c];
  a.h2b = ifh2 ? F.h2b : nullptr;
  a.nel = nel;
  a.lx = lx;
  a.skip = skip;
  const int code = ax_code<T>();
  if (code >= 200) {
    a.edges = F.edges;
    a.zg = F.zg;
    a.wg = F.wg;
    a.h2c = (T)(ifh2 ? h2c : 0.0);
  }
  a.dot_part = dot_part;
  int st = launch_ax<T>(a, code, s);
  if (st) return st;
  if (do_gs) {
    st = gs<T>(w, skip, s);
    if (st) return st;
  }
  if (do_mask) {
    const MaskRef m = mk ? *mk : mask_ref();
    st = launch_mask<T>(w, m.msk, m.nmsk, skip, s);
  }
  return st;
}
template <typename T>
int Op::apply_cgp(T *w, T *p, const T *r, T *x, const double *beta, const double *alpha,
                  const int *skip, hipStream_t s, double *dot_part, bool defer_faces,
                  const MaskRef *mk) const {
  const MaskRef m = mk ? *mk : mask_ref();
  const OpFields<T> &F = fields<T>();
  const int code = ax_code<T>();
  if (!dot_part) {
    set_error("apply_cgp: needs dot partials");
    return NM_ERR_STATE;
  }
  AxArgs<T> a;
  a.w = w;
  a.u = p;
  a.dx = F.dx;
  a.h1 = F.h1;
  a.h1c = (T)h1c;
  for (int c = 0; c < 6; c++) a.g[c] = F.g[c];
  a.h2b = nullptr;
  if (code >= 200) {
    a.edges = F.edges;
    a.zg = F.zg;
    a.wg = F.wg;
    a.h2c = (T)(ifh2 ? h2c : 0.0);
  } else {
    a.h2b = ifh2 ? F.h2b : nullptr;  // stored geometry: premultiplied h2 B
  }
  a.nel = nel;
  a.lx = lx;
  a.skip = skip;
  a.dot_part = dot_part;
  a.cg_r = r;
  a.cg_dinv = F.dinv;
  a.cg_x = x;
  a.cg_beta = beta;
  a.cg_alpha = alpha;
  int st = launch_ax<T>(a, code, s);
  if (st) return st;
  if (defer_faces) {
    if (!m.face_fuse || !face_fuse_on) {
      set_error("apply_cgp: face sums cannot be deferred on this operator");
      return NM_ERR_STATE;
    }
    // the blocks, then the facet pairs no descriptor covers
    GsArgs<T> b{w, blk_ptr, gidx, nblks, pair0, 0, skip};
    b.variant = max_blk <= 8 ? 1 : 0;
    if (gs_variant == 3 && max_blk <= 8) {
      // the packed-segment blocks (no faces: nfaces stays 0)
      b.variant = 3;
      b.seg_idx = seg_idx;
      b.seg_pl = seg_pl;
      b.nseg = nseg;
      b.lx = lx;
    }
    if (nblks > 0 && (st = launch_gs<T>(b, s))) return st;
    GsArgs<T> c{w, nullptr, gs_rest, 0, 0, nrest, skip};
    c.variant = 1;
    if (nrest > 0 && (st = launch_gs<T>(c, s))) return st;
  } else if ((st = gs<T>(w, skip, s))) {
    return st;
  }
  return launch_mask<T>(w, m.msk, m.nmsk, skip, s);
}
template int Op::apply_cgp<double>(double *, double *, const double *, double *,
                                   const double *, const double *, const int *, hipStream_t,
                                   double *, bool, const MaskRef *) const;
template int Op::apply_cgp<float>(float *, float *, const float *, float *, const double *,
                                  const double *, const int *, hipStream_t, double *,
                                  bool, const MaskRef *) const;

template int Op::apply<double>(double *, const double *, const int *,
                               hipStream_t, bool, bool, double *, const MaskRef *) const;
template int Op::apply<float>(float *, const float *, const int *, hipStream_t,
                              bool, bool, double *, const MaskRef *) const;

int Op::ensure32(hipStream_t s) {
  Op *op = this;
  if (has32) return NM_OK;
  NM_ALLOC(f32.dx, float, lx * lx);
  NM_HIP_TRY(launch_d2f(f64.dx, f32.dx, lx * lx, s) ? hipErrorUnknown : hipSuccess);
  for (int c = 0; c < 6; c++) {
    NM_ALLOC(f32.g[c], float, n);
    if (launch_d2f(f64.g[c], f32.g[c], n, s)) return NM_ERR_HIP;
  }
  if (f64.h1) {
    NM_ALLOC(f32.h1, float, n);
    if (launch_d2f(f64.h1, f32.h1, n, s)) return NM_ERR_HIP;
  }
  if (f64.h2b) {
    NM_ALLOC(f32.h2b, float, n);
    if (launch_d2f(f64.h2b, f32.h2b, n, s)) return NM_ERR_HIP;
  }
  if (f64.edges) {
    NM_ALLOC(f32.edges, float, nel * 36);
    if (launch_d2f(f64.edges, f32.edges, nel * 36, s)) return NM_ERR_HIP;
    NM_ALLOC(f32.zg, float, lx);
    if (launch_d2f(f64.zg, f32.zg, lx, s)) return NM_ERR_HIP;
    NM_ALLOC(f32.wg, float, lx);
    if (launch_d2f(f64.wg, f32.wg, lx, s)) return NM_ERR_HIP;
  }
  NM_ALLOC(f32.mult, float, n);
  if (launch_d2f(f64.mult, f32.mult, n, s)) return NM_ERR_HIP;
  NM_ALLOC(f32.dinv, float, n);
  if (launch_d2f(f64.dinv, f32.dinv, n, s)) return NM_ERR_HIP;
  has32 = true;
  return NM_OK;
}

// Processing order for locality: blocks and facet pairs are independent of
// each other, so the device layout lists them grouped by the element of
// their first (lowest) local dof -- a stable counting sort that keeps
// Neko's order inside each group and never touches the order of the
// entries inside a block (the summation order of a block is unchanged).
int gs_regroup(int64_t nel, int64_t n3, int64_t nblks, int64_t pair0, int64_t npairs,
               std::vector<int32_t> &ptr, std::vector<int32_t> &gi) {
  const int64_t ne = nel > 0 ? nel : 1;
  auto regroup = [&](int64_t count, auto first, auto len, auto emit) {
    std::vector<int64_t> start(ne + 1, 0);
    std::vector<int32_t> key(count);
    for (int64_t b = 0; b < count; b++) {
      int32_t m = INT32_MAX;
      const int64_t f = first(b);
      for (int q = 0; q < len(b); q++) m = std::min(m, gi[f + q]);
      key[b] = (int32_t)(m / n3);
      start[key[b] + 1]++;
    }
    for (int64_t e = 0; e < ne; e++) start[e + 1] += start[e];
    std::vector<int64_t> order(count);
    for (int64_t b = 0; b < count; b++) order[start[key[b]]++] = b;
    for (int64_t k = 0; k < count; k++) emit(order[k]);
  };
  std::vector<int32_t> g2(gi.size()), p2(nblks + 1);
  int64_t o = 0, nb = 0;
  regroup(
      nblks, [&](int64_t b) { return (int64_t)ptr[b]; },
      [&](int64_t b) { return ptr[b + 1] - ptr[b]; },
      [&](int64_t b) {
        p2[nb++] = (int32_t)o;
        for (int32_t q = ptr[b]; q < ptr[b + 1]; q++) g2[o++] = gi[q];
      });
  p2[nblks] = (int32_t)o;
  regroup(
      npairs, [&](int64_t pp) { return pair0 + 2 * pp; }, [&](int64_t) { return 2; },
      [&](int64_t pp) {
        g2[o++] = gi[pair0 + 2 * pp];
        g2[o++] = gi[pair0 + 2 * pp + 1];
      });
  if (o != (int64_t)gi.size()) {
    set_error("gs: regrouping lost entries (%lld of %lld)", (long long)o, (long long)gi.size());
    return NM_ERR_STATE;
  }
  gi.swap(g2);
  ptr.swap(p2);
  return NM_OK;
}

// Whole element faces among the facet pairs (0-based local index pairs):
// the (lx-2)^2 pairs between face fa of element ea and face fb of element
// eb form one descriptor when one of the 8 symmetries of the square maps
// every point (u, v) of face a (1..lx-2 in the face's in-plane element
// coordinates) to its partner on face b.  Pairs that form no such face
// (none on conforming meshes) stay explicit in `rest`.  Descriptors keep
// the order of their first pair (the locality-regrouped order).
void gs_face_pairs(int64_t nel, int lx, const int32_t *pairs, int64_t npairs,
                   std::vector<int32_t> &faces, std::vector<int32_t> &rest) {
  (void)nel;
  const int L = lx - 2, l1 = lx - 1;
  const int64_t n3 = (int64_t)lx * lx * lx;
  faces.clear();
  rest.clear();
  if (L < 1) {
    rest.assign(pairs, pairs + 2 * npairs);
    return;
  }
  // face and in-face coordinates of a local point (interior face points)
  auto locate = [&](int64_t p, int &f, int &u, int &v) {
    const int i = (int)(p % lx), j = (int)((p / lx) % lx), k = (int)(p / ((int64_t)lx * lx));
    const int bi = (i == 0 || i == l1), bj = (j == 0 || j == l1), bk = (k == 0 || k == l1);
    if (bi + bj + bk != 1) return false;
    if (bi) { f = (i == l1); u = j; v = k; }
    else if (bj) { f = 2 + (j == l1); u = i; v = k; }
    else { f = 4 + (k == l1); u = i; v = j; }
    return true;
  };
  struct Group { int64_t ea, eb; int fa, fb; std::vector<int64_t> idx; };
  std::vector<Group> groups;
  std::unordered_map<uint64_t, size_t> key2g;
  for (int64_t q = 0; q < npairs; q++) {
    int64_t a = pairs[2 * q], b = pairs[2 * q + 1];
    if (a > b) std::swap(a, b);
    int fa, ua, va, fb, ub, vb;
    const int64_t ea = a / n3, eb = b / n3;
    if (!locate(a % n3, fa, ua, va) || !locate(b % n3, fb, ub, vb)) {
      rest.push_back(pairs[2 * q]);
      rest.push_back(pairs[2 * q + 1]);
      continue;
    }
    const uint64_t key = ((uint64_t)ea * 6 + fa) * 6 * (uint64_t)nel + (uint64_t)eb * 6 + fb;
    auto it = key2g.find(key);
    size_t g;
    if (it == key2g.end()) {
      g = groups.size();
      key2g.emplace(key, g);
      groups.push_back(Group{ea, eb, fa, fb, {}});
    } else {
      g = it->second;
    }
    groups[g].idx.push_back(q);
  }
  for (const Group &G : groups) {
    int code = -1;
    if ((int64_t)G.idx.size() == (int64_t)L * L) {
      for (int sym = 0; sym < 8 && code < 0; sym++) {
        bool ok = true;
        std::vector<char> seen((size_t)L * L, 0);
        for (int64_t q : G.idx) {
          int64_t a = pairs[2 * q], b = pairs[2 * q + 1];
          if (a > b) std::swap(a, b);
          int fa, ua, va, fb, ub, vb;
          locate(a % n3, fa, ua, va);
          locate(b % n3, fb, ub, vb);
          int tu = (sym & 1) ? va : ua, tv = (sym & 1) ? ua : va;
          if (sym & 2) tu = L + 1 - tu;
          if (sym & 4) tv = L + 1 - tv;
          const size_t slot = (size_t)(ua - 1) + (size_t)L * (va - 1);
          if (tu != ub || tv != vb || seen[slot]) { ok = false; break; }
          seen[slot] = 1;
        }
        if (ok) code = G.fa | (G.fb << 3) | ((sym & 1) << 6) | ((sym & 2) << 6) | ((sym & 4) << 6);
      }
    }
    if (code >= 0) {
      faces.push_back((int32_t)G.ea);
      faces.push_back((int32_t)G.eb);
      faces.push_back(code);
    } else {
      for (int64_t q : G.idx) {
        rest.push_back(pairs[2 * q]);
        rest.push_back(pairs[2 * q + 1]);
      }
    }
  }
}

int op_set_gs(Op *op, int64_t nlocal, int64_t facet_offset, int64_t nblks,
              const int32_t *dg, const int32_t *gd, const int32_t *blk_len,
              const int32_t *blk_off, hipStream_t s) {
  // Device layout: blk_ptr (nblks+1), gidx = gd - 1 (0-based).  Entries
  // from facet_offset on are (nlocal - facet_offset + 1) / 2 pairs.
  op->nlocal = nlocal;
  op->nblks = nblks;
  op->facet_offset = facet_offset;
  op->pair0 = facet_offset - 1;
  op->npairs = (nlocal - (facet_offset - 1)) / 2;
  if ((nlocal - (facet_offset - 1)) % 2 != 0) {
    set_error("gs: odd number of facet entries");
    return NM_ERR_ARG;
  }
  std::vector<int32_t> ptr(nblks + 1), gi(nlocal > 0 ? nlocal : 1);
  op->max_blk = 0;
  for (int64_t b = 0; b < nblks; b++) {
    ptr[b] = blk_off[b];
    if (blk_len[b] > op->max_blk) op->max_blk = blk_len[b];
  }
  ptr[nblks] = nblks > 0 ? blk_off[nblks - 1] + blk_len[nblks - 1] : 0;
  if (ptr[nblks] != facet_offset - 1) {
    set_error("gs: blocks end at %d, facet offset %lld", ptr[nblks],
              (long long)facet_offset);
    return NM_ERR_ARG;
  }
  for (int64_t i = 0; i < nlocal; i++) gi[i] = gd[i] - 1;
  {
    const int st = gs_regroup(op->nel, (int64_t)op->lx * op->lx * op->lx, nblks, op->pair0,
                              op->npairs, ptr, gi);
    if (st) return st;
  }
  {
    std::vector<int32_t> faces, rest;
    gs_face_pairs(op->nel, op->lx, gi.data() + op->pair0, op->npairs, faces, rest);
    op->nfaces = (int64_t)faces.size() / 3;
    op->nrest = (int64_t)rest.size() / 2;
    // the same descriptors per element face; the partner's view of a
    // descriptor swaps the faces and inverts the square symmetry (with the
    // swap the two flips trade places)
    op->h_face_nbr.assign((size_t)op->nel * 12, -1);
    for (int64_t d = 0; d < op->nfaces; d++) {
      const int32_t ea = faces[3 * d], eb = faces[3 * d + 1], code = faces[3 * d + 2];
      const int fa = code & 7, fb = (code >> 3) & 7;
      const int sw = (code >> 6) & 1, fu = (code >> 7) & 1, fv = (code >> 8) & 1;
      const int iu = sw ? fv : fu, iv = sw ? fu : fv;
      const size_t sa = ((size_t)ea * 6 + fa) * 2, sb = ((size_t)eb * 6 + fb) * 2;
      op->h_face_nbr[sa] = eb;
      op->h_face_nbr[sa + 1] = code;
      op->h_face_nbr[sb] = ea;
      op->h_face_nbr[sb + 1] = fb | (fa << 3) | (sw << 6) | (iu << 7) | (iv << 8);
    }
    NM_ALLOC(op->face_nbr, int32_t, op->h_face_nbr.size());
    NM_HIP_TRY(hipMemcpyAsync(op->face_nbr, op->h_face_nbr.data(),
                              sizeof(int32_t) * op->h_face_nbr.size(), hipMemcpyHostToDevice, s));
    NM_ALLOC(op->gs_faces, int32_t, faces.size());
    NM_ALLOC(op->gs_rest, int32_t, rest.size());
    if (!faces.empty())
      NM_HIP_TRY(hipMemcpyAsync(op->gs_faces, faces.data(), sizeof(int32_t) * faces.size(),
                                hipMemcpyHostToDevice, s));
    if (!rest.empty())
      NM_HIP_TRY(hipMemcpyAsync(op->gs_rest, rest.data(), sizeof(int32_t) * rest.size(),
                                hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipStreamSynchronize(s));
  }
  if (op->max_blk <= 8 && nblks > 0) {
    // packed segments: a block's entries on consecutive lanes, never across
    // a 64-lane wavefront (padding entries -1)
    std::vector<int32_t> si;
    std::vector<uint8_t> sp;
    si.reserve((size_t)ptr[nblks] + (size_t)ptr[nblks] / 16 + 64);
    sp.reserve(si.capacity());
    for (int64_t b = 0; b < nblks; b++) {
      const int len = ptr[b + 1] - ptr[b];
      if ((int64_t)(si.size() % 64) + len > 64)
        while (si.size() % 64 != 0) {
          si.push_back(-1);
          sp.push_back(0);
        }
      for (int q = 0; q < len; q++) {
        si.push_back(gi[ptr[b] + q]);
        sp.push_back((uint8_t)(q | (len << 4)));
      }
    }
    op->nseg = (int64_t)si.size();
    NM_ALLOC(op->seg_idx, int32_t, op->nseg);
    NM_ALLOC(op->seg_pl, uint8_t, op->nseg);
    NM_HIP_TRY(hipMemcpyAsync(op->seg_idx, si.data(), sizeof(int32_t) * si.size(),
                              hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipMemcpyAsync(op->seg_pl, sp.data(), si.size(), hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipStreamSynchronize(s));
  }
  NM_ALLOC(op->blk_ptr, int32_t, nblks + 1);
  NM_ALLOC(op->gidx, int32_t, nlocal);
  NM_HIP_TRY(hipMemcpyAsync(op->blk_ptr, ptr.data(), sizeof(int32_t) * (nblks + 1),
                            hipMemcpyHostToDevice, s));
  NM_HIP_TRY(hipMemcpyAsync(op->gidx, gi.data(), sizeof(int32_t) * nlocal,
                            hipMemcpyHostToDevice, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  return NM_OK;
}

bool Op::face_fuse_ok(const int32_t *h_msk, int64_t m) const {
  if (!(nfaces > 0 && lx >= 3 && lx <= NM_FACE_FUSE_MAX_LX &&
        h_face_nbr.size() == (size_t)nel * 12))
    return false;
  const int64_t n3 = (int64_t)lx * lx * lx;
  for (int64_t q = 0; q < m; q++) {
    const int64_t e = h_msk[q] / n3, l = h_msk[q] % n3;
    const int i = (int)(l % lx), j = (int)((l / lx) % lx), k = (int)(l / ((int64_t)lx * lx));
    const bool bi = i == 0 || i == lx - 1, bj = j == 0 || j == lx - 1,
               bk = k == 0 || k == lx - 1;
    if ((int)bi + (int)bj + (int)bk != 1) continue;
    const int f = bi ? (i != 0) : (bj ? 2 + (j != 0) : 4 + (k != 0));
    if (h_face_nbr[((size_t)e * 6 + f) * 2] >= 0) return false;
  }
  return true;
}

// mult = 1/gs(1), volume = sum B, dinv = 1/gs(h1 diag(K) + h2 B)
// (coef.f90:403-419, :1142-1155; pc_jacobi.f90:93-165).
int op_finish(Op *op, hipStream_t s) {
  const int64_t n = op->n;
  // The CG update may take the descriptor face sums unless a masked node
  // lies on a descriptor face interior (the mask acts after the sum).
  op->face_fuse = false;
  {
    std::vector<int32_t> mk(op->nmsk > 0 ? op->nmsk : 1);
    if (op->nmsk > 0)
      NM_HIP_TRY(hipMemcpy(mk.data(), op->msk, sizeof(int32_t) * op->nmsk,
                           hipMemcpyDeviceToHost));
    op->face_fuse = op->face_fuse_ok(mk.data(), op->nmsk);
  }
  if (!op->f64.mult) NM_ALLOC(op->f64.mult, double, n);
  int st = launch_vec<double>(V_CFILL, op->f64.mult, nullptr, nullptr, nullptr, 1.0, n, s);
  if (st) return st;
  if ((st = op->gs<double>(op->f64.mult, nullptr, s))) return st;
  if ((st = launch_vec<double>(V_INVCOL1, op->f64.mult, nullptr, nullptr, nullptr, 0.0, n, s))) return st;
  // the multiplicities as one byte each (mult == 1.0 / count exactly): the
  // fused CG update reads these and looks 1/count up instead of mult
  if (!op->mcnt) {
    NM_ALLOC(op->mcnt, uint8_t, n);
    NM_ALLOC(op->mcnt_bad, int32_t, 1);
  }
  NM_HIP_TRY(hipMemsetAsync(op->mcnt_bad, 0, sizeof(int32_t), s));
  if ((st = launch_mult_count(op->f64.mult, op->mcnt, op->mcnt_bad, n, s))) return st;
  {
    int32_t bad = 1;
    NM_HIP_TRY(hipMemcpyAsync(&bad, op->mcnt_bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NM_HIP_TRY(hipStreamSynchronize(s));
    op->mcnt_ok = bad == 0;
  }
  if (!op->red_part) NM_ALLOC(op->red_part, double, RED_BLOCKS * 4);
  if (!op->red_out) NM_ALLOC(op->red_out, double, 4);
  // volume = sum(B) via glsc3(B, 1, 1); dinv serves as the ones vector
  // before it receives the Jacobi diagonal.
  if (!op->f64.dinv) NM_ALLOC(op->f64.dinv, double, n);
  double *dd = op->f64.dinv;
  if ((st = launch_vec<double>(V_CFILL, dd, nullptr, nullptr, nullptr, 1.0, n, s))) return st;
  if ((st = launch_glsc3_partial(op->B, dd, dd, n, op->red_part, s))) return st;
  if ((st = launch_combine(op->red_part, 1, op->red_out, s))) return st;
  NM_HIP_TRY(hipMemcpyAsync(&op->volume, op->red_out, sizeof(double), hipMemcpyDeviceToHost, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  double *jd = op->pc == 0 ? dd : op->alt64;
  if (op->pc != 0 && (st = launch_vec<double>(V_CFILL, dd, nullptr, nullptr, nullptr, 1.0, n, s)))
    return st;
  if ((st = op_jacobi(op, jd, s))) return st;
  float *jd32 = op->pc == 0 ? op->f32.dinv : op->alt32;
  if (op->has32 && jd32 && launch_d2f(jd, jd32, n, s)) return NM_ERR_HIP;
  NM_HIP_TRY(hipStreamSynchronize(s));
  return NM_OK;
}

// dd = 1/gs(h1 diag(K) + h2 B) with the op's current h1/h2
int op_jacobi(Op *op, double *dd, hipStream_t s) {
  const int64_t n = op->n;
  int st;
  const double *g[6];
  for (int c = 0; c < 6; c++) g[c] = op->f64.g[c];
  if ((st = launch_jacobi_diag(op->nel, op->lx, op->dxt, g, op->dfrmd, dd, s))) return st;
  if (op->f64.h1) {
    if ((st = launch_vec<double>(V_COL2, dd, op->f64.h1, nullptr, nullptr, 0.0, n, s))) return st;
  } else if (op->h1c != 1.0) {
    if ((st = launch_vec<double>(V_CMULT, dd, nullptr, nullptr, nullptr, op->h1c, n, s))) return st;
  }
  if (op->ifh2 && op->f64.h2b) {
    if ((st = launch_vec<double>(V_ADD2, dd, op->f64.h2b, nullptr, nullptr, 0.0, n, s))) return st;
  }
  if ((st = op->gs<double>(dd, nullptr, s))) return st;
  if ((st = launch_vec<double>(V_INVCOL1, dd, nullptr, nullptr, nullptr, 0.0, n, s))) return st;
  NM_HIP_TRY(hipStreamSynchronize(s));
  return NM_OK;
}

// Trilinear GLL coordinates from the 8 element vertices (dofmap_xyzlin).
__global__ void xyz_kernel(int64_t nel, int lx, const double *__restrict__ zg,
                           const double *__restrict__ vxyz,
                           double *__restrict__ X, double *__restrict__ Y,
                           double *__restrict__ Z) {
  const int64_t n3 = (int64_t)lx * lx * lx;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nel * n3) return;
  const int64_t e = idx / n3;
  const int p = (int)(idx % n3);
  const int i = p % lx, j = (p / lx) % lx, k = p / (lx * lx);
  const double ri[2] = {(1.0 - zg[i]) / 2.0, (1.0 + zg[i]) / 2.0};
  const double sj[2] = {(1.0 - zg[j]) / 2.0, (1.0 + zg[j]) / 2.0};
  const double tk[2] = {(1.0 - zg[k]) / 2.0, (1.0 + zg[k]) / 2.0};
  for (int c = 0; c < 3; c++) {
    double acc = 0.0;
    for (int cz = 0; cz < 2; cz++) {
      double ay = 0.0;
      for (int cy = 0; cy < 2; cy++) {
        const double *v = vxyz + (e * 8 + 4 * cz + 2 * cy) * 3 + c;
        const double ax = ri[0] * v[0] + ri[1] * v[3];
        ay += sj[cy] * ax;
      }
      acc += tk[cz] * ay;
    }
    (c == 0 ? X : (c == 1 ? Y : Z))[idx] = acc;
  }
}

// Curved elements: their coordinates, computed on the host, replace the
// trilinear ones (element cel[c] <- buf[(3c + comp) n3 ..]).
__global__ void curve_scatter_kernel(int64_t ncur, int64_t n3, const int32_t *__restrict__ cel,
                                     const double *__restrict__ buf, double *__restrict__ X,
                                     double *__restrict__ Y, double *__restrict__ Z) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncur * n3) return;
  const int64_t c = idx / n3, p = idx % n3, e = cel[c];
  X[e * n3 + p] = buf[(3 * c) * n3 + p];
  Y[e * n3 + p] = buf[(3 * c + 1) * n3 + p];
  Z[e * n3 + p] = buf[(3 * c + 2) * n3 + p];
}

}  // namespace nm

using namespace nm;


extern "C" {


// Build a whole problem from a straight-edged mesh (element vertices in
// Neko's symmetric order) and its host gs maps (Neko's 1-based layout) and
// mask (1-based): device coordinates, geometric factors, mass,
// multiplicity, Jacobi diagonal, and the element edges of the in-kernel
// (trilinear) geometry.  h1 constant; h2 constant (h2 != 0 turns the mass
// term on, ifh2).  On failure the handle is deleted.
static int op_build_linear(nm_op_t *h, int64_t nel, int lx, const std::vector<double> &vxyz,
                           double h1, double h2, int64_t nlocal, int64_t foff, int64_t nb,
                           std::vector<int32_t> &msk, int64_t m, hipStream_t s) {
  Op *op = &h->op;
  op->nel = nel;
  op->lx = lx;
  op->n = nel * lx * lx * lx;
  op->h1c = h1;
  op->ifh2 = h2 != 0.0;
  const int64_t n = op->n;
  int st = NM_OK;
  auto fail = [&](int code) {
    delete h;
    return code;
  };
  // space
  std::vector<double> z(lx), w(lx), d(lx * lx), dt(lx * lx), w3(lx * lx * lx);
  nm_gll(lx, z.data(), w.data(), d.data(), dt.data());
  for (int k = 0; k < lx; k++)
    for (int j = 0; j < lx; j++)
      for (int i = 0; i < lx; i++) w3[i + lx * (j + lx * k)] = w[i] * w[j] * w[k];
  std::vector<int> dfrmd(nel);
  nm_dfrmd(nel, vxyz.data(), dfrmd.data());
  // device arrays
  double *d_z, *d_v, *d_w3;
  if (!(op->f64.dx = op->alloc<double>(lx * lx)) || !(op->dxt = op->alloc<double>(lx * lx)) ||
      !(d_z = op->alloc<double>(lx)) || !(d_w3 = op->alloc<double>(lx * lx * lx)) ||
      !(d_v = op->alloc<double>(nel * 24)) || !(op->dfrmd = op->alloc<int32_t>(nel)))
    return fail(NM_ERR_NOMEM);
  for (int c = 0; c < 3; c++)
    if (!(op->xyz[c] = op->alloc<double>(n))) return fail(NM_ERR_NOMEM);
  for (int c = 0; c < 6; c++)
    if (!(op->f64.g[c] = op->alloc<double>(n))) return fail(NM_ERR_NOMEM);
  if (!(op->B = op->alloc<double>(n)) || !(op->jac = op->alloc<double>(n)))
    return fail(NM_ERR_NOMEM);
  hipMemcpyAsync(op->f64.dx, d.data(), sizeof(double) * lx * lx, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(op->dxt, dt.data(), sizeof(double) * lx * lx, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(d_z, z.data(), sizeof(double) * lx, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(d_w3, w3.data(), sizeof(double) * lx * lx * lx, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(d_v, vxyz.data(), sizeof(double) * nel * 24, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(op->dfrmd, dfrmd.data(), sizeof(int) * nel, hipMemcpyHostToDevice, s);
  hipLaunchKernelGGL(xyz_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     nel, lx, d_z, d_v, op->xyz[0], op->xyz[1], op->xyz[2]);
  if (hipGetLastError() != hipSuccess) return fail(NM_ERR_HIP);
  // curved elements (dofmap_generate_xyz): host coordinates scattered over
  // the trilinear ones; the in-kernel (trilinear) geometry is then off
  bool curved = false;
  if (!op->h_cel.empty()) {
    const int64_t nc = (int64_t)op->h_cel.size(), n3 = (int64_t)lx * lx * lx;
    std::vector<double> buf(nc * 3 * n3);
    for (int64_t c = 0; c < nc; c++) {
      const int32_t *ty = op->h_ctype.data() + c * 12;
      for (int q = 0; q < 12; q++)
        curved |= (q < 8 && ty[q] == 3) || (ty[q] == 4 && lx > 2);
      double *b = buf.data() + c * 3 * n3;
      if ((st = curve_element_xyz(lx, z.data(), vxyz.data() + (int64_t)op->h_cel[c] * 24,
                                  op->h_cdata.data() + c * 60, ty, b, b + n3, b + 2 * n3)))
        return fail(st);
    }
    int32_t *d_cel = op->alloc<int32_t>(nc);
    double *d_buf = nullptr;
    if (!d_cel || hipMalloc(&d_buf, sizeof(double) * buf.size()) != hipSuccess)
      return fail(NM_ERR_NOMEM);
    const bool ok =
        hipMemcpyAsync(d_cel, op->h_cel.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice,
                       s) == hipSuccess &&
        hipMemcpyAsync(d_buf, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice,
                       s) == hipSuccess;
    if (ok)
      hipLaunchKernelGGL(curve_scatter_kernel, dim3((unsigned)((nc * n3 + 255) / 256)),
                         dim3(256), 0, s, nc, n3, d_cel, d_buf, op->xyz[0], op->xyz[1],
                         op->xyz[2]);
    const bool done = ok && hipGetLastError() == hipSuccess &&
                      hipStreamSynchronize(s) == hipSuccess;
    (void)hipFree(d_buf);
    if (!done) return fail(NM_ERR_HIP);
  }
  if ((st = launch_coef(nel, lx, op->xyz[0], op->xyz[1], op->xyz[2], op->f64.dx, d_w3,
                        op->f64.g, op->jac, op->B, s)))
    return fail(st);
  if (op->ifh2) {
    if (!(op->f64.h2b = op->alloc<double>(n))) return fail(NM_ERR_NOMEM);
    launch_vec<double>(V_CMULT2, op->f64.h2b, op->B, nullptr, nullptr, h2, n, s);
  }
  // trilinear geometry for the in-kernel G / B variants: half edge vectors
  // of every element (corner (cx,cy,cz) at vxyz[(8e + 4cz + 2cy + cx)*3])
  {
    std::vector<double> ed(nel * 36);
    for (int64_t e = 0; e < nel; e++) {
      auto V = [&](int cx, int cy, int cz, int c) {
        return vxyz[(e * 8 + 4 * cz + 2 * cy + cx) * 3 + c];
      };
      double *E = ed.data() + e * 36;
      for (int c = 0; c < 3; c++)
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++) {
            E[c * 4 + a + 2 * b] = 0.5 * (V(1, a, b, c) - V(0, a, b, c));
            E[12 + c * 4 + a + 2 * b] = 0.5 * (V(a, 1, b, c) - V(a, 0, b, c));
            E[24 + c * 4 + a + 2 * b] = 0.5 * (V(a, b, 1, c) - V(a, b, 0, c));
          }
    }
    if (!(op->f64.edges = op->alloc<double>(nel * 36)) || !(op->f64.zg = op->alloc<double>(lx)) ||
        !(op->f64.wg = op->alloc<double>(lx)))
      return fail(NM_ERR_NOMEM);
    if (hipMemcpyAsync(op->f64.edges, ed.data(), sizeof(double) * nel * 36,
                       hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(op->f64.zg, z.data(), sizeof(double) * lx, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(op->f64.wg, w.data(), sizeof(double) * lx, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(NM_ERR_HIP);
    op->trilinear = !curved;
    op->h2c = h2;
  }
  // gs maps (already in op->h_*), then the device layout
  op->h_blk_len.resize(nb);
  op->h_blk_off.resize(nb);
  if ((st = op_set_gs(op, nlocal, foff, nb, op->h_dg.data(), op->h_gd.data(),
                      op->h_blk_len.data(), op->h_blk_off.data(), s)))
    return fail(st);
  for (int64_t i = 0; i < m; i++) msk[i] -= 1;
  op->nmsk = m;
  if (!(op->msk = op->alloc<int32_t>(m))) return fail(NM_ERR_NOMEM);
  hipMemcpyAsync(op->msk, msk.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s);
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error("operator set-up: device copies failed");
    return fail(NM_ERR_HIP);
  }
  if ((st = op_finish(op, s))) return fail(st);
  return NM_OK;
}

// Build a whole box problem: host mesh/gs maps/mask, then op_build_linear.
int nm_op_create_box(nm_op_t **out, int nelx, int nely, int nelz, int lx,
                     int perx, int pery, int perz, double amp, uint64_t seed,
                     double h1, double h2, hipStream_t s) {
  *out = nullptr;
  if (lx < 2 || lx > 12) {
    set_error("nm_op_create_box: lx=%d outside 2..12", lx);
    return NM_ERR_ARG;
  }
  if ((perx && nelx < 3) || (pery && nely < 3) || (perz && nelz < 3)) {
    set_error("nm_op_create_box: periodic direction needs >= 3 elements");
    return NM_ERR_ARG;
  }
  nm_op_t *h = new nm_op_t;
  Op *op = &h->op;
  op->box[0] = nelx; op->box[1] = nely; op->box[2] = nelz;
  op->per[0] = perx; op->per[1] = pery; op->per[2] = perz;
  op->amp = amp;
  op->seed = seed;
  const int64_t nel = (int64_t)nelx * nely * nelz;
  std::vector<double> vxyz(nel * 24);
  nm_box_mesh(nelx, nely, nelz, amp, seed, vxyz.data());
  int64_t nlocal, foff, nb;
  int st;
  if ((st = nm_box_gs_sizes(nelx, nely, nelz, lx, perx, pery, perz, &nlocal, &foff))) {
    delete h;
    return st;
  }
  op->h_dg.resize(nlocal);
  op->h_gd.resize(nlocal);
  op->h_blk_len.resize(foff);
  op->h_blk_off.resize(foff);
  if ((st = nm_box_gs_map(nelx, nely, nelz, lx, perx, pery, perz, op->h_dg.data(),
                          op->h_gd.data(), op->h_blk_len.data(), op->h_blk_off.data(),
                          &nlocal, &foff, &nb))) {
    delete h;
    return st;
  }
  const int64_t mcap = nm_box_mask_size(nelx, nely, nelz, lx, perx, pery, perz);
  std::vector<int32_t> msk(mcap > 0 ? mcap : 1);
  int64_t m = 0;
  nm_box_mask(nelx, nely, nelz, lx, perx, pery, perz, msk.data(), &m);
  if ((st = op_build_linear(h, nel, lx, vxyz, h1, h2, nlocal, foff, nb, msk, m, s))) return st;
  *out = h;
  return NM_OK;
}

// Build a problem on a general conforming straight-edged hex mesh (e.g.
// one read by nm_nmsh_read): facet connectivity, gs maps and mask from the
// point ids (nm_mesh_*), Dirichlet on every facet without a neighbour,
// then op_build_linear.  vxyz / pid: host arrays, 24 doubles / 8 ids per
// element.
int nm_op_create_mesh(nm_op_t **out, int64_t nel, int lx, const double *vxyz,
                      const int64_t *pid, double h1, double h2, hipStream_t s) {
  return nm_op_create_mesh_curved(out, nel, lx, vxyz, pid, 0, nullptr, nullptr, nullptr, h1,
                                  h2, s);
}

// The same with curved elements: ncurves records of Neko's curve data
// (element cel[c], 0-based; cdata: 5 x 12 doubles; ctype: 12 edge types),
// as nm_nmsh_curves reads them; at most one record per element.
int nm_op_create_mesh_curved(nm_op_t **out, int64_t nel, int lx, const double *vxyz,
                             const int64_t *pid, int64_t ncurves, const int32_t *cel,
                             const double *cdata, const int32_t *ctype, double h1, double h2,
                             hipStream_t s) {
  *out = nullptr;
  if (nel <= 0 || lx < 2 || lx > 12) {
    set_error("nm_op_create_mesh: nel=%lld lx=%d (lx must be 2..12)", (long long)nel, lx);
    return NM_ERR_ARG;
  }
  if (ncurves < 0 || (ncurves > 0 && (!cel || !cdata || !ctype))) {
    set_error("nm_op_create_mesh_curved: bad curve arguments");
    return NM_ERR_ARG;
  }
  {
    std::vector<char> seen(nel, 0);
    for (int64_t c = 0; c < ncurves; c++) {
      if (cel[c] < 0 || cel[c] >= nel || seen[cel[c]]) {
        set_error("nm_op_create_mesh_curved: curve record %lld names element %d "
                  "(out of range or repeated)", (long long)c, cel[c]);
        return NM_ERR_ARG;
      }
      seen[cel[c]] = 1;
    }
  }
  std::vector<int32_t> fneigh(nel * 6);
  int st;
  if ((st = nm_mesh_facet_neigh(nel, pid, fneigh.data()))) return st;
  int64_t nlocal, foff, nb;
  if ((st = nm_mesh_gs_sizes(nel, lx, fneigh.data(), &nlocal, &foff))) return st;
  nm_op_t *h = new nm_op_t;
  Op *op = &h->op;
  op->h_dg.resize(nlocal);
  op->h_gd.resize(nlocal);
  op->h_blk_len.resize(foff);
  op->h_blk_off.resize(foff);
  if ((st = nm_mesh_gs_map(nel, lx, pid, fneigh.data(), op->h_dg.data(), op->h_gd.data(),
                           op->h_blk_len.data(), op->h_blk_off.data(), &nlocal, &foff, &nb))) {
    delete h;
    return st;
  }
  const int64_t mcap = nm_mesh_mask_size(nel, lx, fneigh.data());
  std::vector<int32_t> msk(mcap > 0 ? mcap : 1);
  int64_t m = 0;
  nm_mesh_mask(nel, lx, fneigh.data(), msk.data(), &m);
  std::vector<double> v(vxyz, vxyz + nel * 24);
  op->h_vxyz = v;
  op->h_pid.assign(pid, pid + nel * 8);
  op->h_fneigh = fneigh;
  op->h_cel.assign(cel, cel + ncurves);
  op->h_cdata.assign(cdata, cdata + ncurves * 60);
  op->h_ctype.assign(ctype, ctype + ncurves * 12);
  if ((st = op_build_linear(h, nel, lx, v, h1, h2, nlocal, foff, nb, msk, m, s))) return st;
  *out = h;
  return NM_OK;
}

// Build a problem from Neko's own host arrays (what a Fortran caller holds
// after space_t / coef_t / gs_t / bc setup).  All pointers are host memory;
// gs maps and mask are 1-based as in Neko.  h1 may be NULL (constant h1c);
// h2 may be NULL when ifh2 == 0.  dfrmd may be NULL (all deformed).
int nm_op_create(nm_op_t **out, int nel, int lx, const double *dx,
                 const double *g11, const double *g22, const double *g33,
                 const double *g12, const double *g13, const double *g23,
                 const double *B, const double *h1, double h1c,
                 const double *h2, int ifh2, const int *dfrmd, int nlocal,
                 int facet_offset, int nblks, const int *dg, const int *gd,
                 const int *blk_len, const int *blk_off, int nmsk,
                 const int *msk, hipStream_t s) {
  *out = nullptr;
  if (nel <= 0 || lx < 2 || lx > 14) {
    set_error("nm_op_create: bad nel/lx");
    return NM_ERR_ARG;
  }
  nm_op_t *hnd = new nm_op_t;
  Op *op = &hnd->op;
  op->nel = nel;
  op->lx = lx;
  op->n = (int64_t)nel * lx * lx * lx;
  op->h1c = h1c;
  op->ifh2 = ifh2 != 0;
  const int64_t n = op->n;
  auto fail = [&](int code) {
    delete hnd;
    return code;
  };
  std::vector<double> dt(lx * lx);
  for (int i = 0; i < lx; i++)
    for (int j = 0; j < lx; j++) dt[j + lx * i] = dx[i + lx * j];
  if (!(op->f64.dx = op->alloc<double>(lx * lx)) || !(op->dxt = op->alloc<double>(lx * lx)) ||
      !(op->B = op->alloc<double>(n)) || !(op->dfrmd = op->alloc<int32_t>(nel)))
    return fail(NM_ERR_NOMEM);
  hipMemcpyAsync(op->f64.dx, dx, sizeof(double) * lx * lx, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(op->dxt, dt.data(), sizeof(double) * lx * lx, hipMemcpyHostToDevice, s);
  const double *gs_[6] = {g11, g22, g33, g12, g13, g23};
  for (int c = 0; c < 6; c++) {
    if (!(op->f64.g[c] = op->alloc<double>(n))) return fail(NM_ERR_NOMEM);
    hipMemcpyAsync(op->f64.g[c], gs_[c], sizeof(double) * n, hipMemcpyHostToDevice, s);
  }
  hipMemcpyAsync(op->B, B, sizeof(double) * n, hipMemcpyHostToDevice, s);
  if (h1) {
    if (!(op->f64.h1 = op->alloc<double>(n))) return fail(NM_ERR_NOMEM);
    hipMemcpyAsync(op->f64.h1, h1, sizeof(double) * n, hipMemcpyHostToDevice, s);
  }
  if (op->ifh2) {
    if (!h2) {
      set_error("nm_op_create: ifh2 set without h2");
      return fail(NM_ERR_ARG);
    }
    std::vector<double> hb(n);
    for (int64_t i = 0; i < n; i++) hb[i] = h2[i] * B[i];
    if (!(op->f64.h2b = op->alloc<double>(n))) return fail(NM_ERR_NOMEM);
    hipMemcpyAsync(op->f64.h2b, hb.data(), sizeof(double) * n, hipMemcpyHostToDevice, s);
    hipStreamSynchronize(s);
  }
  std::vector<int> df(nel, 1);
  if (dfrmd)
    for (int e = 0; e < nel; e++) df[e] = dfrmd[e];
  hipMemcpyAsync(op->dfrmd, df.data(), sizeof(int) * nel, hipMemcpyHostToDevice, s);
  op->h_dg.assign(dg, dg + nlocal);
  op->h_gd.assign(gd, gd + nlocal);
  op->h_blk_len.assign(blk_len, blk_len + nblks);
  op->h_blk_off.assign(blk_off, blk_off + nblks);
  int st = op_set_gs(op, nlocal, facet_offset, nblks, dg, gd, blk_len, blk_off, s);
  if (st) return fail(st);
  std::vector<int32_t> m0(nmsk > 0 ? nmsk : 1);
  for (int i = 0; i < nmsk; i++) m0[i] = msk[i] - 1;
  op->nmsk = nmsk;
  if (!(op->msk = op->alloc<int32_t>(nmsk))) return fail(NM_ERR_NOMEM);
  hipMemcpyAsync(op->msk, m0.data(), sizeof(int32_t) * (nmsk > 0 ? nmsk : 1), hipMemcpyHostToDevice, s);
  if (hipStreamSynchronize(s) != hipSuccess) return fail(NM_ERR_HIP);
  if ((st = op_finish(op, s))) return fail(st);
  *out = hnd;
  return NM_OK;
}

void nm_op_free(nm_op_t *h) { delete h; }

int nm_op_info(const nm_op_t *h, nm_op_info_t *info) {
  const Op &op = h->op;
  info->nel = op.nel;
  info->n = op.n;
  info->nlocal = op.nlocal;
  info->nblks = op.nblks;
  info->npairs = op.npairs;
  info->nmsk = op.nmsk;
  info->facet_offset = op.facet_offset;
  info->lx = op.lx;
  info->volume = op.volume;
  info->h1c = op.h1c;
  info->ifh2 = op.ifh2;
  info->nfaces = op.nfaces;
  info->nrest = op.nrest;
  info->cg_face_sums = op.face_fuse && op.face_fuse_on ? 1 : 0;
  info->mult_counts = op.mcnt_ok ? 1 : 0;
  return NM_OK;
}

// Device pointer of a named field (for zero-copy access by a caller that
// owns no copy): 0..5 G11..G23, 6 B, 7 mult, 8 Jacobi dinv, 9..11 x/y/z,
// 12 h2*B, 13 jac.
const void *nm_op_field(const nm_op_t *h, int which) {
  const Op &op = h->op;
  switch (which) {
    case 0: case 1: case 2: case 3: case 4: case 5: return op.f64.g[which];
    case 6: return op.B;
    case 7: return op.f64.mult;
    case 8: return op.f64.dinv;
    case 9: case 10: case 11: return op.xyz[which - 9];
    case 12: return op.f64.h2b;
    case 13: return op.jac;
    default: return nullptr;
  }
}

// Neko-layout gs maps / mask held by the operator (host copies).
int nm_op_gs_maps(const nm_op_t *h, int32_t *dg, int32_t *gd, int32_t *blk_len,
                  int32_t *blk_off) {
  const Op &op = h->op;
  std::memcpy(dg, op.h_dg.data(), sizeof(int32_t) * op.h_dg.size());
  std::memcpy(gd, op.h_gd.data(), sizeof(int32_t) * op.h_gd.size());
  std::memcpy(blk_len, op.h_blk_len.data(), sizeof(int32_t) * op.h_blk_len.size());
  std::memcpy(blk_off, op.h_blk_off.data(), sizeof(int32_t) * op.h_blk_off.size());
  return NM_OK;
}

int nm_op_set_epb(nm_op_t *h, int epb64, int epb32) {
  const Op &op = h->op;
  const bool tri_ok = op.trilinear && !op.f64.h1;
  if ((epb64 >= 200 || epb32 >= 200) && !tri_ok) {
    set_error("nm_op_set_epb: trilinear variants need a straight-edged mesh "
              "(box operator) and constant h1");
    return NM_ERR_UNSUPPORTED;
  }
  h->op.epb64 = epb64;
  h->op.epb32 = epb32;
  h->op.epoch++;
  return NM_OK;
}

// msh%all_deformed() (mesh.f90:470-473): every element takes the Jacobi
// cross terms; the derived fields (mult, volume, Jacobi) are rebuilt.
int nm_op_set_all_deformed(nm_op_t *h, hipStream_t s) {
  Op *op = &h->op;
  std::vector<int> ones(op->nel, 1);
  NM_HIP_TRY(hipMemcpyAsync(op->dfrmd, ones.data(), sizeof(int) * op->nel,
                            hipMemcpyHostToDevice, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  return op_finish(op, s);
}

int nm_op_set_cg_face_sums(nm_op_t *h, int on) {
  h->op.face_fuse_on = on != 0;
  h->op.epoch++;
  return NM_OK;
}

int nm_op_set_gs_variant(nm_op_t *h, int variant) {
  h->op.gs_variant = variant;
  h->op.epoch++;
  return NM_OK;
}

// New constant Helmholtz coefficients h1, h2 of an operator with constant
// h1 (what the pnpn schemes set before each solve: pressure h1 = 1/rho,
// h2 = 0, pnpn_res_cpu.f90:74-78; velocity h1 = mu, h2 = rho bd / dt,
// :212-216): h2 B and the Jacobi diagonal are rebuilt, FP32 copies too.
int nm_op_set_coeffs(nm_op_t *h, double h1, double h2, hipStream_t s) {
  Op *op = &h->op;
  if (op->f64.h1) {
    set_error("nm_op_set_coeffs: the operator has a pointwise h1 field");
    return NM_ERR_UNSUPPORTED;
  }
  const int64_t n = op->n;
  int st;
  op->h1c = h1;
  op->h2c = h2;
  op->ifh2 = h2 != 0.0;
  if (op->ifh2) {
    if (!op->f64.h2b) NM_ALLOC(op->f64.h2b, double, n);
    if ((st = launch_vec<double>(V_CMULT2, op->f64.h2b, op->B, nullptr, nullptr, h2, n, s)))
      return st;
    if (op->has32) {
      if (!op->f32.h2b) NM_ALLOC(op->f32.h2b, float, n);
      if (launch_d2f(op->f64.h2b, op->f32.h2b, n, s)) return NM_ERR_HIP;
    }
  }
  // the Jacobi array is f64.dinv, or alt64 while the identity is selected
  double *jac64 = op->pc == 0 ? op->f64.dinv : op->alt64;
  float *jac32 = op->pc == 0 ? op->f32.dinv : op->alt32;
  if ((st = op_jacobi(op, jac64, s))) return st;
  if (op->has32 && jac32 && launch_d2f(jac64, jac32, n, s)) return NM_ERR_HIP;
  op->epoch++;
  NM_HIP_TRY(hipStreamSynchronize(s));
  return NM_OK;
}

// Preconditioner of the op's solvers: 0 Jacobi (pc_jacobi, the default),
// 1 identity (the ksp default when no preconditioner is given, krylov.f90
// M_ident: z = r).  The arrays are swapped, not recomputed, so switching
// back and forth is cheap and exact.
int nm_op_set_pc(nm_op_t *h, int kind, hipStream_t s) {
  Op *op = &h->op;
  if (kind != 0 && kind != 1) {
    set_error("nm_op_set_pc: kind %d (0 Jacobi, 1 identity)", kind);
    return NM_ERR_ARG;
  }
  if (kind == op->pc) return NM_OK;
  const int64_t n = op->n;
  if (!op->alt64) {  // first switch: the other array holds ones
    NM_ALLOC(op->alt64, double, n);
    int st = launch_vec<double>(V_CFILL, op->alt64, nullptr, nullptr, nullptr, 1.0, n, s);
    if (st) return st;
  }
  if (op->has32 && !op->alt32) {  // FP32 copy of what becomes the FP64 array
    NM_ALLOC(op->alt32, float, n);
    if (launch_d2f(op->alt64, op->alt32, n, s)) return NM_ERR_HIP;
  }
  std::swap(op->f64.dinv, op->alt64);
  if (op->has32) std::swap(op->f32.dinv, op->alt32);
  op->pc = kind;
  op->epoch++;
  NM_HIP_TRY(hipStreamSynchronize(s));
  return NM_OK;
}

// Autotuner: time every Ax variant (elements per workgroup x G prefetch)
// in FP64 and FP32 and every gs variant on this problem's own data, keep
// the fastest.  Results (codes and ms) are reported through the arrays
// (up to 16 entries each, may be NULL).  All variants compute identical
// results (gs) or results within rounding of each other (Ax).
int nm_op_autotune(nm_op_t *h, int reps, int *best64, int *best32,
                   int *best_gs, hipStream_t s) {
  Op &op = h->op;
  int st = op.ensure32(s);
  if (st) return st;
  if (reps <= 0) reps = 5;
  const int64_t n = op.n;
  double *u = nullptr, *w = nullptr;
  NM_HIP_TRY(hipMalloc(&u, sizeof(double) * n));
  NM_HIP_TRY(hipMalloc(&w, sizeof(double) * n));
  hipEvent_t e0, e1;
  NM_HIP_TRY(hipEventCreate(&e0));
  NM_HIP_TRY(hipEventCreate(&e1));
  launch_vec<double>(V_CFILL, u, nullptr, nullptr, nullptr, 1.0, n, s);
  std::vector<int> codes = {0, 2, 4, 100, 102, 104};
  if (op.trilinear && !op.f64.h1) {
    codes.insert(codes.end(), {201, 202, 204});
    if (op.lx == 4 || op.lx == 8) codes.push_back(300);  // FP64 on the matrix cores
  }
  auto timeit = [&](auto fn) -> float {
    for (int r = 0; r < 2; r++) fn();
    hipEventRecord(e0, s);
    for (int r = 0; r < reps; r++) fn();
    hipEventRecord(e1, s);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    return ms / reps;
  };
  float b64 = 1e30f, b32 = 1e30f;
  int c64 = 0, c32 = 0;
  for (int c : codes) {
    op.epb64 = c;
    op.epb32 = c;
    if (op.apply<double>(w, u, nullptr, s, false, false) != NM_OK) continue;
    const float t64 = timeit([&] { op.apply<double>(w, u, nullptr, s, false, false); });
    const float t32 = timeit([&] {
      op.apply<float>((float *)w, (const float *)u, nullptr, s, false, false);
    });
    if (t64 < b64) { b64 = t64; c64 = c; }
    if (t32 < b32) { b32 = t32; c32 = c; }
  }
  op.epb64 = c64;
  op.epb32 = c32;
  float bg = 1e30f;
  int cg = 0;
  for (int v = 0; v < 4; v++) {
    if (v >= 1 && op.max_blk > 8) continue;
    op.gs_variant = v;
    const float tg = timeit([&] { op.gs<double>(w, nullptr, s); });
    if (tg < bg) { bg = tg; cg = v; }
  }
  op.epoch++;
  op.gs_variant = cg;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(u);
  hipFree(w);
  if (best64) *best64 = c64;
  if (best32) *best32 = c32;
  if (best_gs) *best_gs = cg;
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// w = A_local u (no gs, no mask), FP64, device pointers.
int nm_op_ax_local(const nm_op_t *h, double *w, const double *u, hipStream_t s) {
  return h->op.apply<double>(w, u, nullptr, s, false, false);
}
// ax_t%compute_vector (src/math/ax.f90:98-115): the local operator on the
// three components of a vector field
int nm_op_ax_local_vector(const nm_op_t *h, double *au, double *av, double *aw,
                          const double *u, const double *v, const double *w, hipStream_t s) {
  int st = h->op.apply<double>(au, u, nullptr, s, false, false);
  if (!st) st = h->op.apply<double>(av, v, nullptr, s, false, false);
  if (!st) st = h->op.apply<double>(aw, w, nullptr, s, false, false);
  return st;
}
// w = mask(gs(A_local u)) (the operator every Krylov iteration applies).
int nm_op_ax(const nm_op_t *h, double *w, const double *u, hipStream_t s) {
  return h->op.apply<double>(w, u, nullptr, s, true, true);
}
// w = gs(A_local u) (Ax+gs, no mask) FP64
int nm_op_ax_gs(const nm_op_t *h, double *w, const double *u, hipStream_t s) {
  return h->op.apply<double>(w, u, nullptr, s, true, false);
}
int nm_op_ax_gs32(nm_op_t *h, float *w, const float *u, hipStream_t s) {
  int st = h->op.ensure32(s);
  if (st) return st;
  return h->op.apply<float>(w, u, nullptr, s, true, false);
}
int nm_op_ax_local32(nm_op_t *h, float *w, const float *u, hipStream_t s) {
  int st = h->op.ensure32(s);
  if (st) return st;
  return h->op.apply<float>(w, u, nullptr, s, false, false);
}
int nm_op_gs(const nm_op_t *h, double *u, hipStream_t s) {
  return h->op.gs<double>(u, nullptr, s);
}
int nm_op_mask(const nm_op_t *h, double *u, hipStream_t s) {
  return launch_mask<double>(u, h->op.msk, h->op.nmsk, nullptr, s);
}

int nm_op_rhs_normal(const nm_op_t *h, uint64_t seed, double *f, hipStream_t s) {
  const Op &op = h->op;
  const int64_t n = op.n;
  const int64_t nb = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(rng_normal_kernel, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(256), 0,
                     s, f, n, seed);
  NM_HIP_TRY(hipGetLastError());
  int st = op.gs<double>(f, nullptr, s);
  if (st) return st;
  return launch_mask<double>(f, op.msk, op.nmsk, nullptr, s);
}

}  // extern "C"
