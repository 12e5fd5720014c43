// Batched solves: many independent box problems of uneven size and order
// advance together, one FP32 CG iteration of all of them per captured
// graph step, with FP64 residual replacement per problem (the mixed
// scheme of nm_cg.hip).
//
// Layout: every problem's vectors are one slice of shared batch vectors,
// padded to whole 256-dof chunks (padding stays zero: dinv = mult = 0 and
// multiplicity count 0), so
// a chunk belongs to exactly one problem.  Per iteration:
//   p = dinv r + beta_i p            one launch over all chunks
//   w = A_i p (FP32, trilinear)      one launch per lx, workgroups tagged
//                                    with their problem; fused p.Ap partials
//   w = mask(gs(w))                  merged gs maps (offsets added), one
//                                    gs + one mask launch for all problems
//   alpha_i, x32 += a p, r -= a w, rtr_i, rtz_i, beta_i, stop test
// Problems that have stopped skip their work chunk by chunk.  Reductions
// are per problem in a fixed order (partials of runs of up to 16 chunks
// starting at the problem's first chunk, then one workgroup per problem),
// so results do not depend on the other problems of the batch.
// Neko solves one system at a time (krylov.f90 ksp_t%solve); this has no
// counterpart there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "nm_cg.h"
#include "nm_common.h"
#include "nm_kernels.h"
#include "nm_op.h"

namespace nm {

constexpr int CHUNK = 256;

__device__ inline double bwsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// block of CHUNK threads: sums of a and b to thread 0
__device__ inline void bsum_pair(double &a, double &b) {
  __shared__ double sh[2][CHUNK / 64];
  a = bwsum(a);
  b = bwsum(b);
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ra = 0.0, rb = 0.0;
    for (int q = 0; q < CHUNK / 64; q++) {
      ra += sh[0][q];
      rb += sh[1][q];
    }
    a = ra;
    b = rb;
  }
  __syncthreads();
}

// Start of a step: a problem whose residual replacement has finished
// beside the graph (rearm, stored with release order after its scalars)
// runs again from this step on; within a step done never goes 1 -> 0.
__global__ void brearm_kernel(CgScal *__restrict__ sc, int ni) {
  // a relaxed look first: the acquire (a cache invalidation at device
  // scope) only where a flag is up
  for (int i = threadIdx.x; i < ni; i += blockDim.x)
    if (__hip_atomic_load(&sc[i].rearm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      sc[i].rearm = 0;
      sc[i].done = 0;
    }
}

// alpha_i from the fused p.Ap partials of the problem's workgroups
__global__ void __launch_bounds__(CHUNK)
    balpha_kernel(const double *__restrict__ dot_part, const BatchInst *__restrict__ inst,
                  CgScal *__restrict__ sc) {
  CgScal &S = sc[blockIdx.x];
  if (S.done) return;
  const BatchInst I = inst[blockIdx.x];
  double a = 0.0, b = 0.0;
  for (int64_t q = I.blk0 + threadIdx.x; q < I.blk0 + I.nblk; q += CHUNK) a += dot_part[q];
  bsum_pair(a, b);
  if (threadIdx.x == 0) {
    S.pap = a;
    S.alpha = S.rtz / a;
  }
}

// fused schedule (p and x updates in the Ax prologue): r -= alpha w and
// the rtr/rtz partials.  Workgroup u takes the run of up to RUN chunks of
// one problem that starts at chunk ub_c0[u] (a problem's runs start at its
// first chunk, so its partials do not depend on the rest of the batch);
// four points per thread (16-byte accesses), one reduction per run.  With
// CNT the multiplicity weight is looked up from the one-byte count (mult
// is exactly 1/count: the same bits as mult itself).
constexpr int RUN = 16;
typedef float bnt_f4 __attribute__((ext_vector_type(4)));
template <bool CNT>
__global__ void __launch_bounds__(CHUNK)
    bupd3_kernel(float *__restrict__ r, const float *__restrict__ w,
                 const float *__restrict__ mult, const uint8_t *__restrict__ mcnt,
                 const float *__restrict__ dinv, const int32_t *__restrict__ ub_c0,
                 const int32_t *__restrict__ ub_inst, int64_t nub,
                 const BatchInst *__restrict__ inst, const CgScal *__restrict__ sc,
                 double *__restrict__ part) {
  const int64_t u = blockIdx.x;
  const int ii = ub_inst[u];
  const CgScal &S = sc[ii];
  if (S.done) return;  // uniform over the workgroup
  __shared__ float tab[CNT ? 256 : 1];
  if (CNT) {
    tab[threadIdx.x] = threadIdx.x ? (float)(1.0 / (double)threadIdx.x) : 0.0f;
    __syncthreads();
  }
  const float alpha = (float)S.alpha;
  const int64_t cb = ub_c0[u];
  const int64_t ce = cb + RUN < inst[ii].c1 ? cb + RUN : inst[ii].c1;
  double a = 0.0, b = 0.0;
  // every array is touched once per step: streamed past the caches
  for (int64_t i = cb * CHUNK + 4 * threadIdx.x; i < ce * CHUNK; i += 4 * CHUNK) {
    const float4 wv = *reinterpret_cast<const float4 *>(w + i);
    const bnt_f4 rv = __builtin_nontemporal_load(reinterpret_cast<const bnt_f4 *>(r + i));
    const bnt_f4 dv = __builtin_nontemporal_load(reinterpret_cast<const bnt_f4 *>(dinv + i));
    float m[4];
    if (CNT) {
      const uchar4 cv = *reinterpret_cast<const uchar4 *>(mcnt + i);
      m[0] = tab[cv.x];
      m[1] = tab[cv.y];
      m[2] = tab[cv.z];
      m[3] = tab[cv.w];
    } else {
      const float4 mv = *reinterpret_cast<const float4 *>(mult + i);
      m[0] = mv.x;
      m[1] = mv.y;
      m[2] = mv.z;
      m[3] = mv.w;
    }
    const float wq[4] = {wv.x, wv.y, wv.z, wv.w};
    bnt_f4 ro;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const float ri = rv[c] - alpha * wq[c];
      ro[c] = ri;
      const double rm = (double)ri * (double)ri * (double)m[c];
      a += rm;
      b += rm * (double)dv[c];
    }
    __builtin_nontemporal_store(ro, reinterpret_cast<bnt_f4 *>(r + i));
  }
  bsum_pair(a, b);
  if (threadIdx.x == 0) {
    part[u] = a;
    part[nub + u] = b;
  }
}

// The same update with the face sums of gs (the batch's gs then runs the
// blocks only): a face-interior point of an element face that a descriptor
// covers adds its partner's value, w_a + w_b, the bits the facet-pair pass
// would have stored (nm_cg.hip, cg_update2f_kernel).  Launched per lx
// group; workgroup u takes a run of up to RK elements of one problem
// (runs start at the problem's first element); fnb holds every problem's
// {neighbour element, code} per element face from the problem's e0 on.
template <int LX>
__global__ void __launch_bounds__(CHUNK)
    bupd_face_kernel(float *__restrict__ r, const float *__restrict__ w,
                     const uint8_t *__restrict__ mcnt, const float *__restrict__ dinv,
                     const int2 *__restrict__ fnb, const int32_t *__restrict__ run_inst,
                     const int32_t *__restrict__ run_k0, const int32_t *__restrict__ run_map,
                     int64_t r_base, int64_t nrun, const BatchInst *__restrict__ inst,
                     const CgScal *__restrict__ sc, double *__restrict__ part) {
  constexpr int N3 = LX * LX * LX;
  // four neighbouring points per thread (16-byte accesses) where lx^3 is a
  // multiple of 4 (element starts stay 4-aligned), else one
  constexpr int V = N3 % 4 == 0 ? 4 : 1;
  constexpr int NH = N3 / V;                                     // vectors per element
  constexpr int EPI = NH >= CHUNK ? 1 : CHUNK / NH;              // elements per step
  constexpr int NQ = NH >= CHUNK ? (NH + CHUNK - 1) / CHUNK : 1;  // vectors per thread
  constexpr int RK = N3 >= 4096 ? 1 : 4096 / N3;
  const int64_t u = run_map[r_base + blockIdx.x];
  const int ii = run_inst[u];
  const CgScal &S = sc[ii];
  if (S.done) return;  // uniform over the workgroup
  __shared__ float tab[256];
  tab[threadIdx.x] = threadIdx.x ? (float)(1.0 / (double)threadIdx.x) : 0.0f;
  __syncthreads();
  const BatchInst I = inst[ii];
  const float alpha = (float)S.alpha;
  const int eo = EPI > 1 ? (int)threadIdx.x / NH : 0;
  const int h0 = EPI > 1 ? (int)threadIdx.x % NH : (int)threadIdx.x;
  int fq[NQ][V], uq[NQ][V], vq[NQ][V];
#pragma unroll
  for (int t = 0; t < NQ; t++)
#pragma unroll
    for (int c = 0; c < V; c++) {
      const int q = V * (h0 + t * CHUNK) + c;
      const int i2 = q % LX, jj = (q / LX) % LX, kk = q / (LX * LX);
      const bool bi = i2 == 0 || i2 == LX - 1, bj = jj == 0 || jj == LX - 1,
                 bk = kk == 0 || kk == LX - 1;
      fq[t][c] = (q < N3 && (int)bi + (int)bj + (int)bk == 1)
                     ? (bi ? (i2 != 0) : (bj ? 2 + (jj != 0) : 4 + (kk != 0)))
                     : -1;
      uq[t][c] = bi ? jj : i2;
      vq[t][c] = bk ? jj : kk;
    }
  const int64_t k0 = run_k0[u];
  const int64_t k1 = k0 + RK < I.nel ? k0 + RK : I.nel;
  double a = 0.0, b = 0.0;
  if (eo < EPI) {
    for (int64_t k = k0 + eo; k < k1; k += EPI) {
#pragma unroll
      for (int t = 0; t < NQ; t++) {
        const int h = h0 + t * CHUNK;
        if (h >= NH) break;
        const int64_t i = I.voff + k * N3 + (int64_t)V * h;
        float wv[V], rv[V], dv[V], mv[V];
        if constexpr (V == 4) {
          const float4 w4 = *reinterpret_cast<const float4 *>(w + i);
          const bnt_f4 r4 = __builtin_nontemporal_load(reinterpret_cast<const bnt_f4 *>(r + i));
          const bnt_f4 d4 = __builtin_nontemporal_load(reinterpret_cast<const bnt_f4 *>(dinv + i));
          const uchar4 c4 = *reinterpret_cast<const uchar4 *>(mcnt + i);
          wv[0] = w4.x; wv[1] = w4.y; wv[2] = w4.z; wv[3] = w4.w;
          mv[0] = tab[c4.x]; mv[1] = tab[c4.y]; mv[2] = tab[c4.z]; mv[3] = tab[c4.w];
#pragma unroll
          for (int c = 0; c < 4; c++) {
            rv[c] = r4[c];
            dv[c] = d4[c];
          }
        } else {
          wv[0] = w[i];
          rv[0] = __builtin_nontemporal_load(r + i);
          dv[0] = __builtin_nontemporal_load(dinv + i);
          mv[0] = tab[__builtin_nontemporal_load(mcnt + i)];
        }
        float ro[V];
#pragma unroll
        for (int c = 0; c < V; c++) {
          float wi = wv[c];
          if (fq[t][c] >= 0) {
            const int2 d = fnb[(I.e0 + k) * 6 + fq[t][c]];
            if (d.x >= 0) {
              int ub = (d.y & 64) ? vq[t][c] : uq[t][c], vb = (d.y & 64) ? uq[t][c] : vq[t][c];
              if (d.y & 128) ub = LX - 1 - ub;
              if (d.y & 256) vb = LX - 1 - vb;
              wi = wi + w[I.voff + (int64_t)d.x * N3 + face_point((d.y >> 3) & 7, ub, vb, LX)];
            }
          }
          ro[c] = rv[c] - alpha * wi;
          const double rm = (double)ro[c] * (double)ro[c] * (double)mv[c];
          a += rm;
          b += rm * (double)dv[c];
        }
        if constexpr (V == 4) {
          bnt_f4 o;
#pragma unroll
          for (int c = 0; c < 4; c++) o[c] = ro[c];
          __builtin_nontemporal_store(o, reinterpret_cast<bnt_f4 *>(r + i));
        } else {
          __builtin_nontemporal_store(ro[0], r + i);
        }
      }
    }
  }
  bsum_pair(a, b);
  if (threadIdx.x == 0) {
    part[u] = a;
    part[nrun + u] = b;
  }
}

// elements per update run of an lx (bupd_face_kernel's RK)
static inline int64_t face_run(int lx) {
  const int64_t n3 = (int64_t)lx * lx * lx;
  return n3 >= 4096 ? 1 : 4096 / n3;
}

// per problem: rtr, rtz, beta, norm, stop test (cg_fin semantics) over the
// problem's run partials
__global__ void __launch_bounds__(CHUNK)
    bfin_kernel(const double *__restrict__ part, int64_t nub,
                const BatchInst *__restrict__ inst, CgScal *__restrict__ sc) {
  CgScal &S = sc[blockIdx.x];
  if (S.done) return;
  const BatchInst I = inst[blockIdx.x];
  double a = 0.0, b = 0.0;
  for (int64_t c = I.u0 + threadIdx.x; c < I.u1; c += CHUNK) {
    a += part[c];
    b += part[nub + c];
  }
  bsum_pair(a, b);
  if (threadIdx.x == 0) {
    S.rtr = a;
    S.rtz_old = S.rtz;
    S.rtz = b;
    S.beta = b / S.rtz_old;
    S.rnorm = sqrt(a) * S.norm_fac;
    S.iter = S.iter + 1;
    if (S.rnorm < S.tol || S.iter >= S.max_iter) S.done = 1;
  }
}

// float copy of a double field into a batch slice
__global__ void d2f_slice_kernel(float *__restrict__ dst, const double *__restrict__ src,
                                 int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = (float)src[i];
}

struct LxGroup {
  int lx = 0, epb = 1;
  int64_t blk_base = 0, nblocks = 0;
  int64_t r0 = 0, r1 = 0;  // its update runs (face-sum schedule)
  int64_t m0 = 0, m1 = 0;  // its running problems' runs: run_map[m0, m1)
  std::vector<int> members;
  const float *dx = nullptr, *zg = nullptr, *wg = nullptr;
};

struct Batch {
  std::vector<Op *> ops;
  std::vector<int64_t> voff, vlen;   // slice offsets / padded lengths
  int64_t ntot = 0, nchunk = 0, nblk_total = 0;
  std::vector<LxGroup> groups;
  std::vector<BatchInst> h_inst;
  hipStream_t s = nullptr;
  // device
  BatchInst *inst = nullptr;
  int32_t *blk_inst = nullptr;
  // update runs: first chunk and problem of every update workgroup
  int32_t *ub_c0 = nullptr, *ub_inst = nullptr;
  int64_t nub = 0;
  uint8_t *mcnt = nullptr;  // multiplicity counts (every problem's mcnt_ok)
  bool cnt = false;
  // face-sum schedule (every problem can defer its face sums past its mask
  // and has no facet pair outside a whole-face descriptor): per-element face
  // table and update runs of elements
  bool face = false;
  int2 *fnb = nullptr;
  int32_t *run_inst = nullptr, *run_k0 = nullptr;
  int64_t nrun = 0;
  int32_t *gs_ptr = nullptr, *gs_idx = nullptr, *msk = nullptr;
  int64_t gs_nblks = 0, gs_pair0 = 0, gs_npairs = 0, nmsk = 0;
  int gs_maxblk = 0;
  // lane-parallel gs (blocks <= 8 entries): each problem's blocks padded to
  // whole workgroups of 32 blocks and its pairs to 128 pairs, with the
  // problem of every logical workgroup, so a converged or replacing
  // problem's gs workgroups exit at once
  int32_t *wg_blk = nullptr, *wg_pair = nullptr, *wg_msk = nullptr;
  // the workgroups of the problems still running (rebuilt with the Ax
  // tables): gs block and mask workgroups, update runs; host copies of the
  // full problem-per-workgroup tables they are drawn from
  int32_t *map_blk = nullptr, *map_msk = nullptr, *map_run = nullptr;
  int64_t nmap_blk = 0, nmap_msk = 0;
  std::vector<int32_t> h_wg_blk, h_wg_msk, h_run_inst;
  // the gs blocks as packed segments (one lane per entry, a block on
  // consecutive lanes of one wavefront), each problem's padded to whole
  // workgroups of 256 lanes: problem of each workgroup, running ones' map
  int32_t *seg_idx = nullptr, *wg_seg = nullptr, *map_seg = nullptr;
  uint8_t *seg_pl = nullptr;
  int64_t nseg = 0, nmap_seg = 0;
  std::vector<int32_t> h_wg_seg;
  double *x = nullptr, *f = nullptr, *w64 = nullptr, *part = nullptr, *dot_part = nullptr;
  double *rpart = nullptr;  // scratch of the per-problem replacement
  float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *w32 = nullptr;
  float *dinv = nullptr, *mult = nullptr;
  CgScal *sc = nullptr, *h_sc = nullptr;  // h_sc: two read-back slots of ninst
  // residual replacements run on s2 beside the iteration graph on s: each
  // problem's decided scalars land in rep_out (device) and h_rep (pinned),
  // its completion in rep_ev; gev marks each read-back slot
  hipStream_t s2 = nullptr;
  CgScal *rep_out = nullptr, *h_rep = nullptr;
  std::vector<hipEvent_t> rep_ev;
  hipEvent_t gev[2] = {nullptr, nullptr}, s2_ev = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_iters = 4;
  bool compacted = false;   // tables currently cover a subset
  int64_t covered_dof = 0;  // local dofs of the problems the tables cover
  // statistics of the last solve: graph steps, and dof-iterations the Ax
  // launches covered (finished problems inside still-covered tables count)
  int64_t stat_steps = 0;
  double stat_launched = 0.0;
  int64_t bytes = 0;
  std::vector<void *> owned;
  ~Batch() {
    if (s) (void)hipStreamSynchronize(s);
    if (s2) (void)hipStreamSynchronize(s2);
    if (graph) (void)hipGraphExecDestroy(graph);
    for (void *q : owned) (void)hipFree(q);
    if (h_sc) (void)hipHostFree(h_sc);
    if (h_rep) (void)hipHostFree(h_rep);
    for (hipEvent_t e : rep_ev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {gev[0], gev[1], s2_ev})
      if (e) (void)hipEventDestroy(e);
    if (s2) (void)hipStreamDestroy(s2);
    if (s) (void)hipStreamDestroy(s);
  }
  template <typename T>
  T *alloc(int64_t cnt) {
    void *q = nullptr;
    const size_t b = sizeof(T) * (size_t)std::max<int64_t>(cnt, 1);
    if (hipMalloc(&q, b) != hipSuccess) return nullptr;
    owned.push_back(q);
    bytes += (int64_t)b;
    return (T *)q;
  }
};

// One FP32 CG step of every unfinished problem.  Fused schedule: the p
// update (p = dinv r + beta p) and the previous step's x update ride in the
// batched Ax kernel (x is one step behind; a replacement flushes it), the
// update kernel does r -= alpha w and the rtr/rtz partials.
static int batch_iteration(Batch *B, hipStream_t s) {
  hipLaunchKernelGGL(brearm_kernel, dim3(1), dim3(256), 0, s, B->sc, (int)B->ops.size());
  for (const LxGroup &g : B->groups) {
    AxBatchArgs a;
    a.w = B->w32;
    a.u = B->p32;
    a.dx = g.dx;
    a.zg = g.zg;
    a.wg = g.wg;
    a.inst = B->inst;
    a.blk_inst = B->blk_inst;
    a.blk_base = g.blk_base;
    a.nblocks = g.nblocks;
    a.lx = g.lx;
    a.epb = g.epb;
    a.scal = B->sc;
    a.scal_size = (int)sizeof(CgScal);
    a.done_off = (int)offsetof(CgScal, done);
    a.dot_part = B->dot_part;
    a.cg_r = B->r32;
    a.cg_dinv = B->dinv;
    a.cg_x = B->x32;
    a.beta_off = (int)offsetof(CgScal, beta);
    a.alpha_off = (int)offsetof(CgScal, alpha);
    int st = launch_ax_batch(a, s);
    if (st) return st;
  }
  GsArgs<float> ga{B->w32, B->gs_ptr, B->gs_idx, B->gs_nblks, B->gs_pair0,
                   B->face ? 0 : B->gs_npairs, nullptr};
  ga.variant = B->gs_maxblk <= 8 ? 1 : 0;
  if (ga.variant == 1) {
    ga.wg_blk = B->wg_blk;
    ga.wg_pair = B->wg_pair;
    ga.scal = (const char *)B->sc;
    ga.scal_size = (int)sizeof(CgScal);
    ga.done_off = (int)offsetof(CgScal, done);
    ga.wg_map = B->map_blk;
    ga.nwg_map = B->nmap_blk;
    if (B->seg_idx) {
      ga.seg_idx = B->seg_idx;
      ga.seg_pl = B->seg_pl;
      ga.nseg = B->nseg;
      ga.wg_blk = B->wg_seg;
      ga.wg_map = B->map_seg;
      ga.nwg_map = B->nmap_seg;
    }
  }
  int st = launch_gs<float>(ga, s);
  if (st) return st;
  if (B->wg_msk)
    st = launch_mask_wg(B->w32, B->msk, B->nmsk, B->wg_msk, B->sc, (int)sizeof(CgScal),
                        (int)offsetof(CgScal, done), s, B->map_msk, B->nmap_msk);
  else
    st = launch_mask<float>(B->w32, B->msk, B->nmsk, nullptr, s);
  if (st) return st;
  const unsigned ni = (unsigned)B->ops.size();
  hipLaunchKernelGGL(balpha_kernel, dim3(ni), dim3(CHUNK), 0, s, B->dot_part, B->inst, B->sc);
  if (B->face) {
    for (const LxGroup &g : B->groups) {
      if (g.m1 <= g.m0) continue;
      const dim3 grid((unsigned)(g.m1 - g.m0)), block(CHUNK);
#define NM_BF_CASE(L)                                                                      \
  case L:                                                                                  \
    hipLaunchKernelGGL(bupd_face_kernel<L>, grid, block, 0, s, B->r32, B->w32, B->mcnt,      \
                       B->dinv, B->fnb, B->run_inst, B->run_k0, B->map_run, g.m0, B->nrun,   \
                       B->inst, B->sc, B->part);                                           \
    break;
      switch (g.lx) {
        NM_BF_CASE(3) NM_BF_CASE(4) NM_BF_CASE(5) NM_BF_CASE(6) NM_BF_CASE(7) NM_BF_CASE(8)
        NM_BF_CASE(9) NM_BF_CASE(10) NM_BF_CASE(11) NM_BF_CASE(12)
        default:
          set_error("batch: face-sum update unsupported for lx=%d", g.lx);
          return NM_ERR_UNSUPPORTED;
      }
#undef NM_BF_CASE
    }
    hipLaunchKernelGGL(bfin_kernel, dim3(ni), dim3(CHUNK), 0, s, B->part, B->nrun, B->inst,
                       B->sc);
    NM_HIP_TRY(hipGetLastError());
    return NM_OK;
  }
  if (B->cnt)
    hipLaunchKernelGGL(bupd3_kernel<true>, dim3((unsigned)B->nub), dim3(CHUNK), 0, s, B->r32,
                       B->w32, B->mult, B->mcnt, B->dinv, B->ub_c0, B->ub_inst, B->nub, B->inst,
                       B->sc, B->part);
  else
    hipLaunchKernelGGL(bupd3_kernel<false>, dim3((unsigned)B->nub), dim3(CHUNK), 0, s, B->r32,
                       B->w32, B->mult, B->mcnt, B->dinv, B->ub_c0, B->ub_inst, B->nub, B->inst,
                       B->sc, B->part);
  hipLaunchKernelGGL(bfin_kernel, dim3(ni), dim3(CHUNK), 0, s, B->part, B->nub, B->inst, B->sc);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

static int batch_capture(Batch *B) {
  if (B->graph) {
    NM_HIP_TRY(hipGraphExecDestroy(B->graph));
    B->graph = nullptr;
  }
  hipGraph_t g;
  NM_HIP_TRY(hipStreamBeginCapture(B->s, hipStreamCaptureModeThreadLocal));
  int st = NM_OK;
  for (int k = 0; k < B->graph_iters && st == NM_OK; k++) st = batch_iteration(B, B->s);
  hipError_t e = hipStreamEndCapture(B->s, &g);
  if (st) return st;
  NM_HIP_TRY(e);
  NM_HIP_TRY(hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0));
  NM_HIP_TRY(hipGraphDestroy(g));
  return NM_OK;
}

// Build the per-group workgroup tables for the current elements-per-
// workgroup choices.
// With `finished`, problems marked there get no workgroups (compaction of
// the Ax launches once part of the batch has converged; a problem's own
// workgroups keep their relative order, so its results do not change).
static int batch_tables(Batch *B, const std::vector<char> *finished = nullptr) {
  const int ni = (int)B->ops.size();
  int64_t gb = 0;
  std::vector<int32_t> bi;
  B->covered_dof = 0;
  for (LxGroup &g : B->groups) {
    g.blk_base = gb;
    for (int m : g.members) {
      if (finished && (*finished)[m]) continue;
      B->covered_dof += B->ops[m]->n;
      BatchInst &I = B->h_inst[m];
      I.blk0 = gb;
      I.nblk = (I.nel + g.epb - 1) / g.epb;
      for (int64_t q = 0; q < I.nblk; q++) bi.push_back(m);
      gb += I.nblk;
    }
    g.nblocks = gb - g.blk_base;
  }
  if (gb > B->nblk_total) {
    set_error("batch: workgroup table grew beyond its allocation");
    return NM_ERR_STATE;
  }
  // the gs block, mask and update-run workgroups of the running problems
  auto running = [&](int32_t m) { return !finished || !(*finished)[m]; };
  std::vector<int32_t> mb, mm, mr;
  for (size_t b = 0; b < B->h_wg_blk.size(); b++)
    if (running(B->h_wg_blk[b])) mb.push_back((int32_t)b);
  for (size_t b = 0; b < B->h_wg_msk.size(); b++)
    if (running(B->h_wg_msk[b])) mm.push_back((int32_t)b);
  std::vector<int32_t> ms2;
  for (size_t b = 0; b < B->h_wg_seg.size(); b++)
    if (running(B->h_wg_seg[b])) ms2.push_back((int32_t)b);
  B->nmap_seg = (int64_t)ms2.size();
  if (B->map_seg && !ms2.empty())
    NM_HIP_TRY(hipMemcpyAsync(B->map_seg, ms2.data(), sizeof(int32_t) * ms2.size(),
                              hipMemcpyHostToDevice, B->s));
  for (LxGroup &g : B->groups) {
    g.m0 = (int64_t)mr.size();
    for (int64_t u = g.r0; u < g.r1; u++)
      if (running(B->h_run_inst[u])) mr.push_back((int32_t)u);
    g.m1 = (int64_t)mr.size();
  }
  B->nmap_blk = (int64_t)mb.size();
  B->nmap_msk = (int64_t)mm.size();
  if (B->map_blk && !mb.empty())
    NM_HIP_TRY(hipMemcpyAsync(B->map_blk, mb.data(), sizeof(int32_t) * mb.size(),
                              hipMemcpyHostToDevice, B->s));
  if (B->map_msk && !mm.empty())
    NM_HIP_TRY(hipMemcpyAsync(B->map_msk, mm.data(), sizeof(int32_t) * mm.size(),
                              hipMemcpyHostToDevice, B->s));
  if (B->map_run && !mr.empty())
    NM_HIP_TRY(hipMemcpyAsync(B->map_run, mr.data(), sizeof(int32_t) * mr.size(),
                              hipMemcpyHostToDevice, B->s));
  NM_HIP_TRY(hipMemcpyAsync(B->blk_inst, bi.data(), sizeof(int32_t) * bi.size(),
                            hipMemcpyHostToDevice, B->s));
  NM_HIP_TRY(hipMemcpyAsync(B->inst, B->h_inst.data(), sizeof(BatchInst) * ni,
                            hipMemcpyHostToDevice, B->s));
  NM_HIP_TRY(hipStreamSynchronize(B->s));
  return NM_OK;
}

}  // namespace nm

using namespace nm;

struct nm_batch_s {
  Batch b;
};

extern "C" {

int nm_batch_create(nm_batch_t **out, int ninst, nm_op_t *const *ops, hipStream_t s) {
  *out = nullptr;
  if (ninst < 1) {
    set_error("nm_batch_create: no problems");
    return NM_ERR_ARG;
  }
  nm_batch_t *h = new nm_batch_t;
  Batch *B = &h->b;
  auto fail = [&](int c) {
    delete h;
    return c;
  };
  if (hipStreamCreateWithFlags(&B->s, hipStreamNonBlocking) != hipSuccess) return fail(NM_ERR_HIP);
  int64_t off = 0, ms = 0, nblk_max = 0;
  int64_t gs_entries = 0, gs_blocks = 0, gs_pairs = 0;
  int maxblk_all = 0;
  for (int i = 0; i < ninst; i++) {
    Op *op = &ops[i]->op;
    if (!op->trilinear || op->f64.h1) {
      set_error("nm_batch_create: problem %d is not a straight-edged operator with constant h1", i);
      return fail(NM_ERR_ARG);
    }
    int st = op->ensure32(s);
    if (st) return fail(st);
    B->ops.push_back(op);
    B->voff.push_back(off);
    const int64_t len = (op->n + CHUNK - 1) / CHUNK * CHUNK;
    B->vlen.push_back(len);
    off += len;
    ms += op->nmsk;
    nblk_max += op->nel;  // epb >= 1
    gs_entries += op->nlocal;
    gs_blocks += op->nblks;
    gs_pairs += op->npairs;
    maxblk_all = std::max(maxblk_all, op->max_blk);
  }
  if (off >= ((int64_t)1 << 31)) {
    set_error("nm_batch_create: %lld dofs exceed the 32-bit gs index range", (long long)off);
    return fail(NM_ERR_ARG);
  }
  B->ntot = off;
  B->nchunk = off / CHUNK;
  B->nblk_total = nblk_max;
  // groups by lx
  for (int i = 0; i < ninst; i++) {
    const int lx = B->ops[i]->lx;
    auto it = std::find_if(B->groups.begin(), B->groups.end(),
                           [&](const LxGroup &g) { return g.lx == lx; });
    if (it == B->groups.end()) {
      LxGroup g;
      g.lx = lx;
      g.dx = B->ops[i]->f32.dx;
      g.zg = B->ops[i]->f32.zg;
      g.wg = B->ops[i]->f32.wg;
      g.epb = 1;
      B->groups.push_back(g);
      it = B->groups.end() - 1;
    }
    it->members.push_back(i);
  }
  // device vectors
  const int64_t n = B->ntot;
  if (!(B->x = B->alloc<double>(n)) || !(B->f = B->alloc<double>(n)) ||
      !(B->w64 = B->alloc<double>(n)) || !(B->x32 = B->alloc<float>(n)) ||
      !(B->r32 = B->alloc<float>(n)) || !(B->p32 = B->alloc<float>(n)) ||
      !(B->w32 = B->alloc<float>(n)) || !(B->dinv = B->alloc<float>(n)) ||
      !(B->mult = B->alloc<float>(n)) || !(B->part = B->alloc<double>(2 * B->nchunk)) ||
      !(B->dot_part = B->alloc<double>(nblk_max)) ||
      !(B->rpart = B->alloc<double>(2 * RED_BLOCKS)) || !(B->sc = B->alloc<CgScal>(ninst)) ||
      !(B->inst = B->alloc<BatchInst>(ninst)) || !(B->blk_inst = B->alloc<int32_t>(nblk_max)) ||
      !(B->ub_c0 = B->alloc<int32_t>(B->nchunk)) || !(B->ub_inst = B->alloc<int32_t>(B->nchunk)) ||
      !(B->mcnt = B->alloc<uint8_t>(n)) ||
      !(B->gs_ptr = B->alloc<int32_t>(gs_blocks + 32 * (int64_t)ninst + 1)) ||
      !(B->gs_idx = B->alloc<int32_t>(gs_entries + 256 * (int64_t)ninst)) ||
      !(B->msk = B->alloc<int32_t>(ms + 256 * (int64_t)ninst))) {
    set_error("nm_batch_create: out of device memory");
    return fail(NM_ERR_NOMEM);
  }
  if (hipHostMalloc((void **)&B->h_sc, 2 * sizeof(CgScal) * ninst, hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void **)&B->h_rep, sizeof(CgScal) * ninst, hipHostMallocDefault) !=
          hipSuccess ||
      !(B->rep_out = B->alloc<CgScal>(ninst)))
    return fail(NM_ERR_NOMEM);
  if (hipStreamCreateWithFlags(&B->s2, hipStreamNonBlocking) != hipSuccess)
    return fail(NM_ERR_HIP);
  B->rep_ev.assign(ninst, nullptr);
  for (hipEvent_t *e : {&B->gev[0], &B->gev[1], &B->s2_ev})
    if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) return fail(NM_ERR_HIP);
  for (int i = 0; i < ninst; i++)
    if (hipEventCreateWithFlags(&B->rep_ev[i], hipEventDisableTiming) != hipSuccess)
      return fail(NM_ERR_HIP);
  // zero every vector (padding must stay zero), then the FP32 fields
  for (void *v : {(void *)B->x, (void *)B->f, (void *)B->w64})
    if (launch_vec<double>(V_RZERO, (double *)v, nullptr, nullptr, nullptr, 0.0, n, B->s))
      return fail(NM_ERR_HIP);
  for (void *v : {(void *)B->x32, (void *)B->r32, (void *)B->p32, (void *)B->w32,
                  (void *)B->dinv, (void *)B->mult})
    if (launch_vec<float>(V_RZERO, (float *)v, nullptr, nullptr, nullptr, 0.0f, n, B->s))
      return fail(NM_ERR_HIP);
  // instance table, chunk table, merged gs and mask (offsets added)
  B->h_inst.resize(ninst);
  std::vector<int32_t> ubc0, ubinst, gptr, gidx, gpairs, msk, wgb, wgp, wgm, sidx, wgs;
  std::vector<uint8_t> spl;
  B->cnt = true;
  for (int i = 0; i < ninst; i++) B->cnt = B->cnt && B->ops[i]->mcnt_ok;
  B->face = B->cnt;
  for (int i = 0; i < ninst; i++) {
    const Op *op = B->ops[i];
    B->face = B->face && op->face_fuse && op->face_fuse_on && op->nrest == 0 &&
              op->lx >= 3 && op->lx <= 12 && op->h_face_nbr.size() == (size_t)op->nel * 12;
  }
  if (B->cnt && hipMemsetAsync(B->mcnt, 0, (size_t)n, B->s) != hipSuccess) return fail(NM_ERR_HIP);
  gptr.reserve(gs_blocks + 1);
  gidx.reserve(gs_entries);
  gpairs.reserve(2 * gs_pairs);
  msk.reserve(ms);
  int maxblk = 0;
  const bool pad = maxblk_all <= 8;  // the lane-parallel kernels run
  for (int i = 0; i < ninst; i++) {
    Op *op = B->ops[i];
    BatchInst &I = B->h_inst[i];
    I.voff = B->voff[i];
    I.nel = op->nel;
    I.edges = op->f32.edges;
    I.h1c = (float)op->h1c;
    I.h2c = (float)(op->ifh2 ? op->h2c : 0.0);
    I.c0 = B->voff[i] / CHUNK;
    I.c1 = (B->voff[i] + B->vlen[i]) / CHUNK;
    I.u0 = (int64_t)ubc0.size();
    for (int64_t c = I.c0; c < I.c1; c += RUN) {
      ubc0.push_back((int32_t)c);
      ubinst.push_back(i);
    }
    I.u1 = (int64_t)ubc0.size();
    if (B->cnt && hipMemcpyAsync(B->mcnt + I.voff, op->mcnt, (size_t)op->n,
                                 hipMemcpyDeviceToDevice, B->s) != hipSuccess)
      return fail(NM_ERR_HIP);
    std::vector<int32_t> ptr(op->nblks + 1), gi(op->nlocal), mk(op->nmsk);
    if (hipMemcpy(ptr.data(), op->blk_ptr, sizeof(int32_t) * (op->nblks + 1),
                  hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(gi.data(), op->gidx, sizeof(int32_t) * op->nlocal, hipMemcpyDeviceToHost) !=
            hipSuccess ||
        (op->nmsk > 0 && hipMemcpy(mk.data(), op->msk, sizeof(int32_t) * op->nmsk,
                                   hipMemcpyDeviceToHost) != hipSuccess))
      return fail(NM_ERR_HIP);
    if (pad) {
      // packed segments of this problem's blocks, then padding to a whole
      // workgroup (a block never crosses a 64-lane wavefront)
      for (int64_t b = 0; b < op->nblks; b++) {
        const int len = ptr[b + 1] - ptr[b];
        if ((int64_t)(sidx.size() % 64) + len > 64)
          while (sidx.size() % 64 != 0) {
            sidx.push_back(-1);
            spl.push_back(0);
          }
        for (int q = 0; q < len; q++) {
          sidx.push_back(gi[ptr[b] + q] + (int32_t)I.voff);
          spl.push_back((uint8_t)(q | (len << 4)));
        }
      }
      while (sidx.size() % 256 != 0) {
        sidx.push_back(-1);
        spl.push_back(0);
      }
      while ((int64_t)wgs.size() * 256 < (int64_t)sidx.size()) wgs.push_back(i);
    }
    const int32_t base = (int32_t)gidx.size();
    for (int64_t b = 0; b < op->nblks; b++) gptr.push_back(base + ptr[b]);
    for (int64_t q = 0; q < ptr[op->nblks]; q++) gidx.push_back(gi[q] + (int32_t)I.voff);
    for (int64_t q = op->pair0; q < op->nlocal; q++) gpairs.push_back(gi[q] + (int32_t)I.voff);
    if (pad) {
      // empty blocks and (-1, -1) pairs up to whole workgroups
      while (gptr.size() % 32 != 0) gptr.push_back((int32_t)gidx.size());
      while (gpairs.size() % 256 != 0) gpairs.push_back(-1);
      while ((int64_t)wgb.size() * 32 < (int64_t)gptr.size()) wgb.push_back(i);
      while ((int64_t)wgp.size() * 256 < (int64_t)gpairs.size()) wgp.push_back(i);
    }
    for (int64_t q = 0; q < op->nmsk; q++) msk.push_back(mk[q] + (int32_t)I.voff);
    // mask entries padded to whole workgroups of 256, one problem each
    while (msk.size() % 256 != 0) msk.push_back(-1);
    while ((int64_t)wgm.size() * 256 < (int64_t)msk.size()) wgm.push_back(i);
    maxblk = std::max(maxblk, op->max_blk);
    // FP32 Jacobi and multiplicity into the slices
    const unsigned g = (unsigned)std::min<int64_t>((op->n + 255) / 256, 65536);
    hipLaunchKernelGGL(d2f_slice_kernel, dim3(g), dim3(256), 0, B->s, B->dinv + I.voff,
                       op->f64.dinv, op->n);
    hipLaunchKernelGGL(d2f_slice_kernel, dim3(g), dim3(256), 0, B->s, B->mult + I.voff,
                       op->f64.mult, op->n);
  }
  B->gs_nblks = (int64_t)gptr.size();
  gptr.push_back((int32_t)gidx.size());
  B->gs_pair0 = (int64_t)gidx.size();
  B->gs_npairs = (int64_t)gpairs.size() / 2;
  B->gs_maxblk = maxblk;
  gidx.insert(gidx.end(), gpairs.begin(), gpairs.end());
  B->nmsk = (int64_t)msk.size();
  if (!wgm.empty()) {
    B->h_wg_msk = wgm;
    if (!(B->map_msk = B->alloc<int32_t>((int64_t)wgm.size()))) return fail(NM_ERR_NOMEM);
    if (!(B->wg_msk = B->alloc<int32_t>((int64_t)wgm.size()))) return fail(NM_ERR_NOMEM);
    if (hipMemcpyAsync(B->wg_msk, wgm.data(), sizeof(int32_t) * wgm.size(),
                       hipMemcpyHostToDevice, B->s) != hipSuccess)
      return fail(NM_ERR_HIP);
  }
  if (pad) {
    // tables over every logical workgroup of the two launches (the XCD
    // order rounds the count up to a multiple of 8; those extra workgroups
    // have no entries)
    const int64_t nb = NM_NXCD * ((B->gs_nblks / 32 + NM_NXCD - 1) / NM_NXCD);
    const int64_t np = NM_NXCD * ((B->gs_npairs / 128 + NM_NXCD - 1) / NM_NXCD);
    wgb.resize(std::max<int64_t>(nb, 1), 0);
    wgp.resize(std::max<int64_t>(np, 1), 0);
    B->h_wg_blk = wgb;
    if (!(B->map_blk = B->alloc<int32_t>((int64_t)wgb.size())) ||
        !(B->wg_blk = B->alloc<int32_t>((int64_t)wgb.size())) ||
        !(B->wg_pair = B->alloc<int32_t>((int64_t)wgp.size())))
      return fail(NM_ERR_NOMEM);
    if (hipMemcpyAsync(B->wg_blk, wgb.data(), sizeof(int32_t) * wgb.size(),
                       hipMemcpyHostToDevice, B->s) != hipSuccess ||
        hipMemcpyAsync(B->wg_pair, wgp.data(), sizeof(int32_t) * wgp.size(),
                       hipMemcpyHostToDevice, B->s) != hipSuccess)
      return fail(NM_ERR_HIP);
    if (!sidx.empty()) {
      // the segment tables, padded up to a multiple of 8 workgroups (the
      // XCD order's rounding; those workgroups hold only padding)
      const int64_t nwg = NM_NXCD * (((int64_t)wgs.size() + NM_NXCD - 1) / NM_NXCD);
      sidx.resize((size_t)nwg * 256, -1);
      spl.resize((size_t)nwg * 256, 0);
      wgs.resize((size_t)nwg, 0);
      B->nseg = (int64_t)sidx.size();
      B->h_wg_seg = wgs;
      if (!(B->seg_idx = B->alloc<int32_t>(B->nseg)) || !(B->seg_pl = B->alloc<uint8_t>(B->nseg)) ||
          !(B->wg_seg = B->alloc<int32_t>(nwg)) || !(B->map_seg = B->alloc<int32_t>(nwg)))
        return fail(NM_ERR_NOMEM);
      if (hipMemcpyAsync(B->seg_idx, sidx.data(), sizeof(int32_t) * sidx.size(),
                         hipMemcpyHostToDevice, B->s) != hipSuccess ||
          hipMemcpyAsync(B->seg_pl, spl.data(), spl.size(), hipMemcpyHostToDevice, B->s) !=
              hipSuccess ||
          hipMemcpyAsync(B->wg_seg, wgs.data(), sizeof(int32_t) * wgs.size(),
                         hipMemcpyHostToDevice, B->s) != hipSuccess)
        return fail(NM_ERR_HIP);
    }
  }
  B->nub = (int64_t)ubc0.size();
  if (B->face) {
    // face tables from element e0 of each problem, update runs per lx group
    // (a problem's runs contiguous: [u0, u1) into the run partials)
    std::vector<int32_t> fn, rinst, rk0;
    int64_t e0 = 0;
    for (int i = 0; i < ninst; i++) {
      B->h_inst[i].e0 = e0;
      fn.insert(fn.end(), B->ops[i]->h_face_nbr.begin(), B->ops[i]->h_face_nbr.end());
      e0 += B->ops[i]->nel;
    }
    for (LxGroup &g : B->groups) {
      g.r0 = (int64_t)rinst.size();
      const int64_t rk = face_run(g.lx);
      for (int m : g.members) {
        BatchInst &I = B->h_inst[m];
        I.u0 = (int64_t)rinst.size();
        for (int64_t k = 0; k < I.nel; k += rk) {
          rinst.push_back(m);
          rk0.push_back((int32_t)k);
        }
        I.u1 = (int64_t)rinst.size();
      }
      g.r1 = (int64_t)rinst.size();
    }
    B->nrun = (int64_t)rinst.size();
    if (B->nrun > B->nchunk) {  // a run holds at least one chunk's worth of points
      set_error("nm_batch_create: %lld update runs exceed the partial buffer",
                (long long)B->nrun);
      return fail(NM_ERR_STATE);
    }
    B->h_run_inst = rinst;
    if (!(B->fnb = B->alloc<int2>((int64_t)fn.size() / 2)) ||
        !(B->run_inst = B->alloc<int32_t>(B->nrun)) || !(B->run_k0 = B->alloc<int32_t>(B->nrun)) ||
        !(B->map_run = B->alloc<int32_t>(B->nrun))) {
      set_error("nm_batch_create: out of device memory");
      return fail(NM_ERR_NOMEM);
    }
    if (hipMemcpyAsync(B->fnb, fn.data(), sizeof(int32_t) * fn.size(), hipMemcpyHostToDevice,
                       B->s) != hipSuccess ||
        hipMemcpyAsync(B->run_inst, rinst.data(), sizeof(int32_t) * rinst.size(),
                       hipMemcpyHostToDevice, B->s) != hipSuccess ||
        hipMemcpyAsync(B->run_k0, rk0.data(), sizeof(int32_t) * rk0.size(),
                       hipMemcpyHostToDevice, B->s) != hipSuccess)
      return fail(NM_ERR_HIP);
  }
  if (hipMemcpyAsync(B->ub_c0, ubc0.data(), sizeof(int32_t) * ubc0.size(),
                     hipMemcpyHostToDevice, B->s) != hipSuccess ||
      hipMemcpyAsync(B->ub_inst, ubinst.data(), sizeof(int32_t) * ubinst.size(),
                     hipMemcpyHostToDevice, B->s) != hipSuccess ||
      hipMemcpyAsync(B->gs_ptr, gptr.data(), sizeof(int32_t) * gptr.size(),
                     hipMemcpyHostToDevice, B->s) != hipSuccess ||
      hipMemcpyAsync(B->gs_idx, gidx.data(), sizeof(int32_t) * gidx.size(),
                     hipMemcpyHostToDevice, B->s) != hipSuccess ||
      (B->nmsk > 0 && hipMemcpyAsync(B->msk, msk.data(), sizeof(int32_t) * msk.size(),
                                     hipMemcpyHostToDevice, B->s) != hipSuccess) ||
      hipStreamSynchronize(B->s) != hipSuccess)
    return fail(NM_ERR_HIP);
  int st = batch_tables(B);
  if (st) return fail(st);
  if ((st = batch_capture(B))) return fail(st);
  *out = h;
  return NM_OK;
}

void nm_batch_free(nm_batch_t *h) { delete h; }

// Autotuner: per lx group, time the batched FP32 Ax for every
// elements-per-workgroup choice on the batch's own data, keep the fastest
// and recapture the iteration graph.  epb_out[g] receives the choice of
// group g (groups in order of first appearance; up to 16).
int nm_batch_autotune(nm_batch_t *h, int reps, int *lx_out, int *epb_out, int *ngroups) {
  Batch *B = &h->b;
  if (reps <= 0) reps = 3;
  std::vector<CgScal> zero(B->ops.size());
  std::memset(zero.data(), 0, sizeof(CgScal) * zero.size());
  NM_HIP_TRY(hipMemcpyAsync(B->sc, zero.data(), sizeof(CgScal) * zero.size(),
                            hipMemcpyHostToDevice, B->s));
  hipEvent_t e0, e1;
  NM_HIP_TRY(hipEventCreate(&e0));
  NM_HIP_TRY(hipEventCreate(&e1));
  for (size_t gi = 0; gi < B->groups.size(); gi++) {
    LxGroup &g = B->groups[gi];
    float best = 1e30f;
    int bepb = 1;
    for (int epb : {1, 2, 4, 8}) {
      if (g.lx * g.lx * epb > 256) continue;
      g.epb = epb;
      int st = batch_tables(B);
      if (st) return st;
      AxBatchArgs a;
      a.w = B->w32; a.u = B->p32; a.dx = g.dx; a.zg = g.zg; a.wg = g.wg;
      a.inst = B->inst; a.blk_inst = B->blk_inst; a.blk_base = g.blk_base;
      a.nblocks = g.nblocks; a.lx = g.lx; a.epb = epb; a.scal = B->sc;
      a.scal_size = (int)sizeof(CgScal); a.done_off = (int)offsetof(CgScal, done);
      a.dot_part = B->dot_part;
      for (int r = 0; r < 2; r++)
        if ((st = launch_ax_batch(a, B->s))) return st;
      NM_HIP_TRY(hipEventRecord(e0, B->s));
      for (int r = 0; r < reps; r++) launch_ax_batch(a, B->s);
      NM_HIP_TRY(hipEventRecord(e1, B->s));
      NM_HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.f;
      NM_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        bepb = epb;
      }
    }
    g.epb = bepb;
    if (gi < 16) {
      if (lx_out) lx_out[gi] = g.lx;
      if (epb_out) epb_out[gi] = g.epb;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ngroups) *ngroups = (int)B->groups.size();
  int st = batch_tables(B);
  if (st) return st;
  return batch_capture(B);
}

// Solve every problem i: A_i x_i = f_i to rel_tol of its initial residual
// (mixed precision, FP64 residual replacement with reduction `delta` per
// segment; <= 0: 0.1).  x, f: arrays of per-problem device pointers (host
// arrays).  mon: per-problem monitors.  Ordered after / before `stream`.
int nm_batch_solve(nm_batch_t *h, double *const *x, const double *const *f, double rel_tol,
                   int max_iter, double delta, nm_monitor_t *mon, hipStream_t stream) {
  Batch *B = &h->b;
  const int ni = (int)B->ops.size();
  if (delta <= 0.0) delta = 0.1;
  // the ordering event is destroyed on every exit path, errors included
  struct EventGuard {
    hipEvent_t ev = nullptr;
    ~EventGuard() {
      if (ev) (void)hipEventDestroy(ev);
    }
  } guard;
  NM_HIP_TRY(hipEventCreateWithFlags(&guard.ev, hipEventDisableTiming));
  hipEvent_t ev = guard.ev;
  NM_HIP_TRY(hipEventRecord(ev, stream));
  NM_HIP_TRY(hipStreamWaitEvent(B->s, ev, 0));
  // replacements a failed earlier solve may have left on s2 finish first
  NM_HIP_TRY(hipEventRecord(B->s2_ev, B->s2));
  NM_HIP_TRY(hipStreamWaitEvent(B->s, B->s2_ev, 0));
  for (int i = 0; i < ni; i++) {
    CgScal init{};
    init.norm_fac = 1.0 / std::sqrt(B->ops[i]->volume);
    init.scale = 1.0;
    B->h_sc[i] = init;
  }
  NM_HIP_TRY(hipMemcpyAsync(B->sc, B->h_sc, sizeof(CgScal) * ni, hipMemcpyHostToDevice, B->s));
  int st;
  for (int i = 0; i < ni; i++) {
    const int64_t o = B->voff[i];
    NM_HIP_TRY(hipMemcpyAsync(B->f + o, f[i], sizeof(double) * B->ops[i]->n,
                              hipMemcpyDeviceToDevice, B->s));
    if ((st = mx_slice_begin(*B->ops[i], B->x + o, B->x32 + o, B->r32 + o, B->p32 + o,
                             B->f + o, B->sc + i, B->rpart, 0.0, rel_tol, delta, max_iter,
                             B->s)))
      return st;
  }
  if (B->compacted) {
    if ((st = batch_tables(B)) || (st = batch_capture(B))) return st;
    B->compacted = false;
  }
  // Pipelined: graph k+1 is launched before the host reads graph k's
  // scalars, and a problem whose FP32 slice has stopped gets its residual
  // replacement on s2 while the graphs go on (its done flag keeps every
  // graph kernel off its slices until the replacement rearms it).  A done,
  // unfinished problem read at the same iteration count as its last
  // replacement is one whose rearm is still pending, not a new stop.
  std::vector<char> state(ni, 0);  // 0 running, 1 finished, 2 replacing on s2
  std::vector<int> rep_iter(ni, -1);
  std::vector<CgScal> last(ni);
  int left = ni, launches = 0, covered = ni;
  const int limit = max_iter / B->graph_iters + 64;
  B->stat_steps = 0;
  B->stat_launched = 0.0;
  auto launch = [&]() -> int {
    NM_HIP_TRY(hipGraphLaunch(B->graph, B->s));
    NM_HIP_TRY(hipMemcpyAsync(B->h_sc + (size_t)(launches % 2) * ni, B->sc,
                              sizeof(CgScal) * ni, hipMemcpyDeviceToHost, B->s));
    NM_HIP_TRY(hipEventRecord(B->gev[launches % 2], B->s));
    launches++;
    B->stat_steps += B->graph_iters;
    B->stat_launched += (double)B->graph_iters * (double)B->covered_dof;
    return NM_OK;
  };
  // a graph is launched only while some problem runs; with every
  // unfinished problem replacing, the host waits for the second stream
  auto any_running = [&]() {
    for (int i = 0; i < ni; i++)
      if (state[i] == 0) return true;
    return false;
  };
  int inflight = -1;  // the launch whose scalars are still to be read
  while (left > 0) {
    if (inflight < 0 && any_running()) {
      if ((st = launch())) return st;
      inflight = launches - 1;
    }
    const int cur = inflight;
    inflight = -1;
    if (cur >= 0 && any_running()) {
      if ((st = launch())) return st;
      inflight = launches - 1;
    }
    const CgScal *H = nullptr;
    if (cur >= 0) {
      NM_HIP_TRY(hipEventSynchronize(B->gev[cur % 2]));
      H = B->h_sc + (size_t)(cur % 2) * ni;
    } else {
      NM_HIP_TRY(hipStreamSynchronize(B->s2));
    }
    for (int i = 0; i < ni; i++) {
      if (state[i] == 1) continue;
      if (state[i] == 2) {
        const hipError_t q = hipEventQuery(B->rep_ev[i]);
        if (q == hipErrorNotReady) continue;
        NM_HIP_TRY(q);
        if (B->h_rep[i].fin) {
          state[i] = 1;
          last[i] = B->h_rep[i];
          left--;
        } else {
          state[i] = 0;
        }
        continue;  // H may predate the replacement's end
      }
      if (!H) continue;
      const CgScal &S = H[i];
      if (S.fin) {
        state[i] = 1;
        last[i] = S;
        left--;
      } else if (S.done && S.iter != rep_iter[i]) {
        const int64_t o = B->voff[i];
        if ((st = mx_slice_flush(B->x32 + o, B->p32 + o, B->ops[i]->n, B->sc + i, B->s2)) ||
            (st = mx_slice_replace(*B->ops[i], B->x + o, B->x32 + o, B->r32 + o, B->f + o,
                                   B->w64 + o, B->sc + i, B->rpart, B->s2, nullptr,
                                   B->rep_out + i)))
          return st;
        NM_HIP_TRY(hipMemcpyAsync(B->h_rep + i, B->rep_out + i, sizeof(CgScal),
                                  hipMemcpyDeviceToHost, B->s2));
        NM_HIP_TRY(hipEventRecord(B->rep_ev[i], B->s2));
        rep_iter[i] = S.iter;
        state[i] = 2;
      }
    }
    if (launches > limit) {
      set_error("nm_batch_solve: problems did not stop after %d graph launches", launches);
      return NM_ERR_STATE;
    }
    // drop finished problems from the Ax launches once a quarter of the
    // covered ones are done (a re-capture costs far less than the Ax work
    // of the finished problems it removes); batch_tables waits for the
    // graph in flight before the graph is replaced
    if (left > 0 && 4 * left <= 3 * covered) {
      std::vector<char> fin(ni);
      for (int i = 0; i < ni; i++) fin[i] = state[i] == 1;
      if ((st = batch_tables(B, &fin)) || (st = batch_capture(B))) return st;
      B->compacted = true;
      covered = left;
      // batch_tables waited for the graph in flight; its scalars are read
      // by the next pass as usual (the slot and event stay valid)
    }
  }
  // the last replacements (all observed finished) precede the copies out
  NM_HIP_TRY(hipEventRecord(B->s2_ev, B->s2));
  NM_HIP_TRY(hipStreamWaitEvent(B->s, B->s2_ev, 0));
  for (int i = 0; i < ni; i++) {
    const CgScal &S = last[i];
    mon[i].iter = S.iter;
    mon[i].outer = S.updates;
    mon[i].res_start = S.res_start;
    mon[i].res_final = S.true_rnorm;
    mon[i].converged = S.converged;
    NM_HIP_TRY(hipMemcpyAsync(x[i], B->x + B->voff[i], sizeof(double) * B->ops[i]->n,
                              hipMemcpyDeviceToDevice, B->s));
  }
  NM_HIP_TRY(hipEventRecord(ev, B->s));
  NM_HIP_TRY(hipStreamWaitEvent(stream, ev, 0));
  return NM_OK;
}

int nm_batch_stats(const nm_batch_t *h, int64_t *steps, double *launched_dof_iter) {
  *steps = h->b.stat_steps;
  *launched_dof_iter = h->b.stat_launched;
  return NM_OK;
}

int nm_batch_info(const nm_batch_t *h, int64_t *total_dof, int64_t *device_bytes) {
  *total_dof = h->b.ntot;
  *device_bytes = h->b.bytes;
  return NM_OK;
}

}  // extern "C"
