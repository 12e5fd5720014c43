// Host-side setup of the solve path: GLL space, synthetic box meshes, the
// structured dof numbering of a box, Neko's gs maps (bit-exact layout) and
// the Dirichlet mask.  This plays the role of Neko's own setup (space_t,
// mesh_t, dofmap_t, gs_init) for the standalone harnesses; when the library
// runs behind Neko, Neko supplies these arrays itself.
#include "neko_amd.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "nm_common.h"

namespace nm {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char *last_error() { return g_err; }

// Run f(i) for i in [0, n) on up to `threads` host threads (static chunks).
template <class F>
static void parallel_for(int64_t n, F f, int threads = 0) {
  if (threads <= 0) {
    threads = (int)std::thread::hardware_concurrency();
    if (threads > 16) threads = 16;
    if (threads < 1) threads = 1;
  }
  if (n < 4096 || threads == 1) {
    for (int64_t i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; t++) {
    const int64_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    pool.emplace_back([=] {
      for (int64_t i = a; i < b; i++) f(i);
    });
  }
  for (auto &th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Seeded synthetic inputs.  The counter-based generator below is this
// project's input specification (it is restated in the oracle so that both
// sides see identical meshes and right-hand sides).
// ---------------------------------------------------------------------------
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static inline uint64_t rng_u64(uint64_t seed, uint64_t idx) {
  const uint64_t k = mix64(seed + 0x9e3779b97f4a7c15ULL);
  return mix64(k ^ (idx * 0xd1b54a32d192ed03ULL + 0x9e3779b97f4a7c15ULL));
}
static inline double rng_uniform(uint64_t seed, uint64_t idx) {
  return (double)(rng_u64(seed, idx) >> 11) * 0x1.0p-53;
}
static inline double rng_normal(uint64_t seed, uint64_t idx) {
  const double u1 = 1.0 - rng_uniform(seed, 2 * idx);
  const double u2 = rng_uniform(seed, 2 * idx + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// ---------------------------------------------------------------------------
// GLL nodes, weights and derivative matrix.  Nodes are the roots of
// (1-x^2) P'_N(x), found by Newton iteration on P'_N from Chebyshev-Lobatto
// starting points; weights 2/(N(N+1) P_N(x)^2); D(i,j) = P_N(x_i) /
// (P_N(x_j) (x_i - x_j)), D(0,0) = -N(N+1)/4, D(N,N) = N(N+1)/4.
// (Same quantities as Neko's speclib.f90 ZWGLL / DGLL, own algorithm.)
// ---------------------------------------------------------------------------
static void legendre(int n, double x, double &p, double &dp, double &d2p) {
  double p0 = 1.0, p1 = x;
  if (n == 0) {
    p = 1.0; dp = 0.0; d2p = 0.0;
    return;
  }
  for (int k = 2; k <= n; k++) {
    const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = pk;
  }
  p = p1;
  // derivatives from the Legendre ODE where |x| < 1
  const double om = 1.0 - x * x;
  dp = n * (p0 - x * p1) / om;
  d2p = (2.0 * x * dp - n * (n + 1.0) * p1) / om;
}

static double legendre_p(int n, double x) {
  double p0 = 1.0, p1 = x;
  if (n == 0) return 1.0;
  for (int k = 2; k <= n; k++) {
    const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = pk;
  }
  return p1;
}

}  // namespace nm

using namespace nm;

extern "C" {

const char *nm_last_error(void) { return nm::last_error(); }

int nm_gll(int lx, double *z, double *w, double *d, double *dt) {
  if (lx < 2 || lx > 32) {
    set_error("nm_gll: lx=%d outside [2,32]", lx);
    return NM_ERR_ARG;
  }
  const int N = lx - 1;
  const double pi = 3.14159265358979323846;
  z[0] = -1.0;
  z[N] = 1.0;
  for (int i = 1; i <= N / 2; i++) {
    double x = -std::cos(pi * i / N);
    for (int it = 0; it < 100; it++) {
      double p, dp, d2p;
      legendre(N, x, p, dp, d2p);
      const double dx = dp / d2p;
      x -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    z[i] = x;
    z[N - i] = -x;
  }
  if (N % 2 == 0) z[N / 2] = 0.0;
  const double nn1 = N * (N + 1.0);
  for (int i = 0; i <= N; i++) {
    const double p = legendre_p(N, z[i]);
    w[i] = 2.0 / (nn1 * p * p);
  }
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= N; j++) {
      double v;
      if (i != j)
        v = legendre_p(N, z[i]) / (legendre_p(N, z[j]) * (z[i] - z[j]));
      else if (i == 0)
        v = -nn1 / 4.0;
      else if (i == N)
        v = nn1 / 4.0;
      else
        v = 0.0;
      d[i + lx * j] = v;   // D(i,j), column-major
      dt[j + lx * i] = v;  // D^T
    }
  return NM_OK;
}

double nm_rng_uniform(uint64_t seed, uint64_t idx) { return rng_uniform(seed, idx); }

int nm_rng_normal_fill(uint64_t seed, int64_t n, double *out) {
  parallel_for(n, [&](int64_t i) { out[i] = rng_normal(seed, (uint64_t)i); });
  return NM_OK;
}

// Box mesh on the unit cube, elements x-fastest, vertices in Neko's
// lexicographic hex order (contrib/genmeshbox/genmeshbox.f90:200-216).
// Vertices strictly inside the box are moved by amp*h*(2U-1) per
// coordinate with U = uniform(seed, 3*lattice_vertex + c).
int nm_box_mesh(int nelx, int nely, int nelz, double amp, uint64_t seed,
                double *vxyz) {
  if (nelx < 1 || nely < 1 || nelz < 1) {
    set_error("nm_box_mesh: bad dims");
    return NM_ERR_ARG;
  }
  const int64_t nel = (int64_t)nelx * nely * nelz;
  const int64_t nx1 = nelx + 1, ny1 = nely + 1;
  parallel_for(nel, [&](int64_t e) {
    const int ex = (int)(e % nelx), ey = (int)((e / nelx) % nely),
              ez = (int)(e / ((int64_t)nelx * nely));
    for (int v = 0; v < 8; v++) {
      const int gx = ex + (v & 1), gy = ey + ((v >> 1) & 1), gz = ez + (v >> 2);
      const int64_t lat = gx + nx1 * (gy + ny1 * (int64_t)gz);
      double c[3] = {(double)gx / nelx, (double)gy / nely, (double)gz / nelz};
      const bool interior = gx > 0 && gx < nelx && gy > 0 && gy < nely &&
                            gz > 0 && gz < nelz;
      if (interior && amp != 0.0) {
        const double h[3] = {1.0 / nelx, 1.0 / nely, 1.0 / nelz};
        for (int cc = 0; cc < 3; cc++)
          c[cc] += amp * h[cc] *
                   (2.0 * rng_uniform(seed, 3 * (uint64_t)lat + cc) - 1.0);
      }
      for (int cc = 0; cc < 3; cc++) vxyz[(e * 8 + v) * 3 + cc] = c[cc];
    }
  });
  return NM_OK;
}

// mesh_generate_flags semantics (mesh.f90:437-467): an element is flagged
// deformed unless the edges meeting at vertices 1 and 8 are orthogonal.
int nm_dfrmd(int64_t nel, const double *vxyz, int *dfrmd) {
  const double eps = 2.220446049250313e-16;
  parallel_for(nel, [&](int64_t e) {
    const double *p = vxyz + e * 24;
    auto P = [&](int n, int c) { return p[(n - 1) * 3 + c]; };
    double u[3], v[3], w[3];
    int d = 0;
    for (int c = 0; c < 3; c++) {
      u[c] = P(2, c) - P(1, c);
      v[c] = P(3, c) - P(1, c);
      w[c] = P(5, c) - P(1, c);
    }
    if (!(std::fabs(u[0] * v[0] + u[1] * v[1] + u[2] * v[2]) < eps)) d = 1;
    if (!(std::fabs(u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) < eps)) d = 1;
    for (int c = 0; c < 3; c++) {
      u[c] = P(7, c) - P(8, c);
      v[c] = P(6, c) - P(8, c);
      w[c] = P(4, c) - P(8, c);
    }
    if (!(std::fabs(u[0] * v[0] + u[1] * v[1] + u[2] * v[2]) < eps)) d = 1;
    if (!(std::fabs(u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) < eps)) d = 1;
    dfrmd[e] = d;
  });
  return NM_OK;
}

// Trilinear GLL point coordinates (dofmap_xyzlin, dofmap.f90:866-939).
int nm_xyz(int64_t nel, int lx, const double *zg, const double *vxyz,
           double *x, double *y, double *z) {
  std::vector<double> jx(2 * lx);
  for (int i = 0; i < lx; i++) {
    jx[2 * i] = (1.0 - zg[i]) / 2.0;
    jx[2 * i + 1] = (1.0 + zg[i]) / 2.0;
  }
  const int64_t n3 = (int64_t)lx * lx * lx;
  parallel_for(nel, [&](int64_t e) {
    for (int c = 0; c < 3; c++) {
      double xb[2][2][2];
      for (int v = 0; v < 8; v++) xb[v >> 2][(v >> 1) & 1][v & 1] = vxyz[(e * 8 + v) * 3 + c];
      double *out = (c == 0 ? x : (c == 1 ? y : z)) + e * n3;
      for (int k = 0; k < lx; k++)
        for (int j = 0; j < lx; j++)
          for (int i = 0; i < lx; i++) {
            double acc = 0.0;
            for (int cz = 0; cz < 2; cz++) {
              double ay = 0.0;
              for (int cy = 0; cy < 2; cy++) {
                const double ax = jx[2 * i] * xb[cz][cy][0] + jx[2 * i + 1] * xb[cz][cy][1];
                ay += jx[2 * j + cy] * ax;
              }
              acc += jx[2 * k + cz] * ay;
            }
            out[i + lx * (j + lx * k)] = acc;
          }
    }
  });
  return NM_OK;
}

// Number of gs entries of a box (single rank): 8 vertices and 12 edges of
// every element, plus the interior points of every face with a neighbour.
int nm_box_gs_sizes(int nelx, int nely, int nelz, int lx, int perx, int pery,
                    int perz, int64_t *nlocal, int64_t *facet_offset) {
  const int64_t nel = (int64_t)nelx * nely * nelz;
  const int64_t L = lx - 2;
  const int64_t nfx = (int64_t)(perx ? nelx : nelx - 1) * nely * nelz;
  const int64_t nfy = (int64_t)nelx * (pery ? nely : nely - 1) * nelz;
  const int64_t nfz = (int64_t)nelx * nely * (perz ? nelz : nelz - 1);
  const int64_t nv = nel * (8 + 12 * L);
  *facet_offset = nv + 1;
  *nlocal = nv + 2 * (nfx + nfy + nfz) * L * L;
  if (*nlocal > INT32_MAX) {
    set_error("nm_box_gs_sizes: %lld gs entries exceed int32 maps",
              (long long)*nlocal);
    return NM_ERR_UNSUPPORTED;
  }
  return NM_OK;
}

}  // extern "C"

namespace nm {

// Hoare quicksort of (dg, gd) by dg with the middle pivot, exactly the
// partition scheme of gs_qsort_dofmap (gather_scatter.f90:1291-1330), so
// that the (unstable) order inside equal-id runs is Neko's.  The two
// recursive calls work on disjoint ranges, so the top levels run on
// separate threads without changing the result.  1-based lo/hi.
static void hoare_sort(int32_t *dg, int32_t *gd, int64_t lo, int64_t hi,
                       int depth) {
  for (;;) {
    int64_t i = lo - 1, j = hi + 1;
    const int32_t pivot = dg[(lo + hi) / 2 - 1];
    for (;;) {
      do { i++; } while (dg[i - 1] < pivot);
      do { j--; } while (dg[j - 1] > pivot);
      if (i < j) {
        std::swap(dg[i - 1], dg[j - 1]);
        std::swap(gd[i - 1], gd[j - 1]);
      } else if (i == j) {
        i++;
        break;
      } else {
        break;
      }
    }
    const bool left = lo < j, right = i < hi;
    if (left && right && depth < 5 && (j - lo) > 200000 && (hi - i) > 200000) {
      std::thread t([=] { hoare_sort(dg, gd, lo, j, depth + 1); });
      hoare_sort(dg, gd, i, hi, depth + 1);
      t.join();
      return;
    }
    if (left) hoare_sort(dg, gd, lo, j, depth + 1);
    if (right) {
      lo = i;
      continue;
    }
    return;
  }
}

void hoare_sort_dofs(int32_t *dg, int32_t *gd, int64_t lo, int64_t hi) {
  hoare_sort(dg, gd, lo, hi, 0);
}

}  // namespace nm

extern "C" {

// Neko's local gs maps for a box mesh, in Neko's own layout (1-based dg =
// gs id, gd = linear local index, blk_len/blk_off; gather_scatter.f90:
// 513-1375, single rank).  Coincidence of local dofs comes from the
// structured GLL lattice of the box (periodic directions wrap), which is
// the same relation Neko's dofmap numbering encodes.  Arrays must hold
// nlocal entries (nm_box_gs_sizes); blk arrays facet_offset-1 entries.
int nm_box_gs_map(int nelx, int nely, int nelz, int lx, int perx, int pery,
                  int perz, int32_t *dg, int32_t *gd, int32_t *blk_len,
                  int32_t *blk_off, int64_t *nlocal_out,
                  int64_t *facet_offset_out, int64_t *nblks_out) {
  int64_t nlocal, foff;
  int st = nm_box_gs_sizes(nelx, nely, nelz, lx, perx, pery, perz, &nlocal, &foff);
  if (st) return st;
  const int64_t nel = (int64_t)nelx * nely * nelz;
  const int64_t Gx = (int64_t)nelx * (lx - 1) + (perx ? 0 : 1);
  const int64_t Gy = (int64_t)nely * (lx - 1) + (pery ? 0 : 1);
  const int64_t Gz = (int64_t)nelz * (lx - 1) + (perz ? 0 : 1);
  const int64_t wx = (int64_t)nelx * (lx - 1), wy = (int64_t)nely * (lx - 1),
                wz = (int64_t)nelz * (lx - 1);
  // gs id of each lattice point, 0 = not yet registered
  std::vector<int32_t> gsid((size_t)(Gx * Gy * Gz), 0);
  int32_t max_id = 0;
  int64_t n = 0;
  const int l1 = lx - 1;
  auto push = [&](int64_t e, int i, int j, int k) {
    const int64_t ex = e % nelx, ey = (e / nelx) % nely, ez = e / ((int64_t)nelx * nely);
    int64_t gx = ex * l1 + i, gy = ey * l1 + j, gz = ez * l1 + k;
    if (perx && gx == wx) gx = 0;
    if (pery && gy == wy) gy = 0;
    if (perz && gz == wz) gz = 0;
    int32_t &s = gsid[(size_t)(gx + Gx * (gy + Gy * gz))];
    if (s == 0) s = ++max_id;
    dg[n] = s;
    gd[n] = (int32_t)(i + lx * (j + (int64_t)lx * (k + (int64_t)lx * e)) + 1);
    n++;
  };
  // vertices (gather_scatter.f90:555-650)
  for (int64_t e = 0; e < nel; e++) {
    push(e, 0, 0, 0); push(e, l1, 0, 0); push(e, 0, l1, 0); push(e, l1, l1, 0);
    push(e, 0, 0, l1); push(e, l1, 0, l1); push(e, 0, l1, l1); push(e, l1, l1, l1);
  }
  // edges (:656-853)
  for (int64_t e = 0; e < nel; e++) {
    for (int a = 1; a < l1; a++) push(e, a, 0, 0);
    for (int a = 1; a < l1; a++) push(e, a, 0, l1);
    for (int a = 1; a < l1; a++) push(e, a, l1, 0);
    for (int a = 1; a < l1; a++) push(e, a, l1, l1);
    for (int a = 1; a < l1; a++) push(e, 0, a, 0);
    for (int a = 1; a < l1; a++) push(e, 0, a, l1);
    for (int a = 1; a < l1; a++) push(e, l1, a, 0);
    for (int a = 1; a < l1; a++) push(e, l1, a, l1);
    for (int a = 1; a < l1; a++) push(e, 0, 0, a);
    for (int a = 1; a < l1; a++) push(e, l1, 0, a);
    for (int a = 1; a < l1; a++) push(e, 0, l1, a);
    for (int a = 1; a < l1; a++) push(e, l1, l1, a);
  }
  // faces with a neighbour (:960-1062)
  for (int64_t e = 0; e < nel; e++) {
    const int64_t ex = e % nelx, ey = (e / nelx) % nely, ez = e / ((int64_t)nelx * nely);
    const bool nb[6] = {perx || ex > 0, perx || ex < nelx - 1,
                        pery || ey > 0, pery || ey < nely - 1,
                        perz || ez > 0, perz || ez < nelz - 1};
    if (nb[0]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, 0, a, b);
    if (nb[1]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, l1, a, b);
    if (nb[2]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, a, 0, b);
    if (nb[3]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, a, l1, b);
    if (nb[4]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, a, b, 0);
    if (nb[5]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, a, b, l1);
  }
  gsid.clear();
  gsid.shrink_to_fit();
  if (n != nlocal) {
    set_error("nm_box_gs_map: internal count %lld != %lld", (long long)n,
              (long long)nlocal);
    return NM_ERR_STATE;
  }
  if (nlocal > 0) nm::hoare_sort(dg, gd, 1, nlocal, 0);
  // gs_find_blks (:1333-1373)
  int64_t nb = 0, i = 1;
  while (i < foff) {
    const int32_t id = dg[i - 1];
    int64_t j = i, cnt = 1;
    while (j + 1 <= nlocal && dg[j] == id) {
      j++;
      cnt++;
    }
    blk_len[nb++] = (int32_t)cnt;
    i = j + 1;
  }
  if (nb > 0) {
    blk_off[0] = 0;
    for (int64_t q = 1; q < nb; q++) blk_off[q] = blk_off[q - 1] + blk_len[q - 1];
  }
  *nlocal_out = nlocal;
  *facet_offset_out = foff;
  *nblks_out = nb;
  return NM_OK;
}

// Number of Dirichlet-masked local dofs of a box: every point of every
// element face on a non-periodic boundary (mark_facet on facet_neigh == 0).
int64_t nm_box_mask_size(int nelx, int nely, int nelz, int lx, int perx,
                         int pery, int perz) {
  int64_t nf = 0;
  if (!perx) nf += 2LL * nely * nelz;
  if (!pery) nf += 2LL * nelx * nelz;
  if (!perz) nf += 2LL * nelx * nely;
  return nf * lx * lx;
}

// Mask list (1-based local indices, as Neko's zero_dirichlet msk(1:m)).
int nm_box_mask(int nelx, int nely, int nelz, int lx, int perx, int pery,
                int perz, int32_t *msk, int64_t *nmsk) {
  const int64_t nel = (int64_t)nelx * nely * nelz;
  int64_t m = 0;
  const int l1 = lx - 1;
  for (int64_t e = 0; e < nel; e++) {
    const int64_t ex = e % nelx, ey = (e / nelx) % nely, ez = e / ((int64_t)nelx * nely);
    const bool bnd[6] = {!perx && ex == 0, !perx && ex == nelx - 1,
                         !pery && ey == 0, !pery && ey == nely - 1,
                         !perz && ez == 0, !perz && ez == nelz - 1};
    for (int f = 0; f < 6; f++) {
      if (!bnd[f]) continue;
      for (int b = 0; b < lx; b++)
        for (int a = 0; a < lx; a++) {
          int i, j, k;
          if (f < 2) { i = (f == 1) ? l1 : 0; j = a; k = b; }
          else if (f < 4) { j = (f == 3) ? l1 : 0; i = a; k = b; }
          else { k = (f == 5) ? l1 : 0; i = a; j = b; }
          msk[m++] = (int32_t)(i + lx * (j + (int64_t)lx * (k + (int64_t)lx * e)) + 1);
        }
    }
  }
  *nmsk = m;
  return NM_OK;
}

}  // extern "C"
