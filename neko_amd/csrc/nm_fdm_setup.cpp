// Host setup of the fast-diagonalisation (FDM) local solves of the
// overlapping Schwarz smoother: per element and direction, the 1-D
// stiffness/mass pencil of the element extended by one GLL point into each
// neighbour, its B-orthonormal eigenvectors and the inverse eigenvalue sums.
// Semantics of Neko's fdm_init / swap_lengths / fdm_setup_fast
// (src/math/fdm.f90:110-427); the eigensolver (LAPACK dsygv in Neko) is this
// project's own cyclic-Jacobi solver, exploiting that the GLL mass matrix of
// the pencil is diagonal.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "nm_common.h"

namespace {

inline int64_t ix3(int i, int j, int k, int n) {
  return i + (int64_t)n * (j + (int64_t)n * k);
}

// Symmetric-definite pencil (A, diag(b)): A s = lam B s.  On return a holds
// the B-orthonormal eigenvectors (columns, ascending eigenvalues).
void eig_pencil_diag(double *a, const double *b, double *lam, int n) {
  std::vector<double> c(n * n), q(n * n, 0.0), sb(n);
  for (int i = 0; i < n; i++) sb[i] = 1.0 / std::sqrt(b[i]);
  for (int j = 0; j < n; j++) {
    q[j + n * j] = 1.0;
    for (int i = 0; i < n; i++) {
      const double aij = i <= j ? a[i + n * j] : a[j + n * i];
      c[i + n * j] = sb[i] * aij * sb[j];
    }
  }
  for (int sweep = 0; sweep < 64; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const double v = c[i + n * j] * c[i + n * j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-32 * tot) break;
    for (int p = 0; p < n - 1; p++)
      for (int r = p + 1; r < n; r++) {
        const double apr = c[p + n * r];
        if (std::fabs(apr) < 1e-300) continue;
        const double th = (c[r + n * r] - c[p + n * p]) / (2.0 * apr);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; k++) {
          const double x = c[k + n * p], y = c[k + n * r];
          c[k + n * p] = cs * x - sn * y;
          c[k + n * r] = sn * x + cs * y;
        }
        for (int k = 0; k < n; k++) {
          const double x = c[p + n * k], y = c[r + n * k];
          c[p + n * k] = cs * x - sn * y;
          c[r + n * k] = sn * x + cs * y;
        }
        for (int k = 0; k < n; k++) {
          const double x = q[k + n * p], y = q[k + n * r];
          q[k + n * p] = cs * x - sn * y;
          q[k + n * r] = sn * x + cs * y;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; i++) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int x, int y) { return c[x + n * x] < c[y + n * y]; });
  for (int jj = 0; jj < n; jj++) {
    const int j = ord[jj];
    lam[jj] = c[j + n * j];
    for (int i = 0; i < n; i++) a[i + n * jj] = sb[i] * q[i + n * j];
  }
}

// 1-D extended-element pencil (fdm.f90:478-560), nl = n + 3 points.
void pencil_1d(double *a, double *b, int lbc, int rbc, double ll, double lm,
               double lr, const double *ah, const double *bh, int n) {
  const int N = n + 3, np = n + 1;
  auto A = [&](int i, int j) -> double & { return a[i + N * j]; };
  auto AH = [&](int i, int j) { return ah[i + np * j]; };
  const int i0 = lbc == 1 ? 1 : 0, i1 = rbc == 1 ? n - 1 : n;
  std::fill(a, a + N * N, 0.0);
  std::fill(b, b + N, 0.0);
  double fac = 2.0 / lm;
  A(1, 1) = 1.0;
  A(n + 1, n + 1) = 1.0;
  for (int j = i0; j <= i1; j++)
    for (int i = i0; i <= i1; i++) A(i + 1, j + 1) = fac * AH(i, j);
  if (lbc == 0) {
    fac = 2.0 / ll;
    A(0, 0) = fac * AH(n - 1, n - 1);
    A(1, 0) = fac * AH(n, n - 1);
    A(0, 1) = fac * AH(n - 1, n);
    A(1, 1) = A(1, 1) + fac * AH(n, n);
  } else {
    A(0, 0) = 1.0;
  }
  if (rbc == 0) {
    fac = 2.0 / lr;
    A(n + 1, n + 1) = A(n + 1, n + 1) + fac * AH(0, 0);
    A(n + 2, n + 1) = fac * AH(1, 0);
    A(n + 1, n + 2) = fac * AH(0, 1);
    A(n + 2, n + 2) = fac * AH(1, 1);
  } else {
    A(n + 2, n + 2) = 1.0;
  }
  fac = 0.5 * lm;
  b[1] = 1.0;
  b[n + 1] = 1.0;
  for (int i = i0; i <= i1; i++) b[i + 1] = fac * bh[i];
  if (lbc == 0) {
    fac = 0.5 * ll;
    b[0] = fac * bh[n - 1];
    b[1] = b[1] + fac * bh[n];
  } else {
    b[0] = 1.0;
  }
  if (rbc == 0) {
    fac = 0.5 * lr;
    b[n + 1] = b[n + 1] + fac * bh[0];
    b[n + 2] = fac * bh[1];
  } else {
    b[n + 2] = 1.0;
  }
}

void setup_1d(double *s, double *lam, int nl, int lbc, int rbc, double ll,
              double lm, double lr, const double *ah, const double *bh, int n) {
  std::vector<double> b(nl);
  pencil_1d(s, b.data(), lbc, rbc, ll, lm, lr, ah, bh, n);
  eig_pencil_diag(s, b.data(), lam, nl);
  auto zero_row = [&](int r) {
    for (int j = 0; j < nl; j++) s[r + nl * j] = 0.0;
  };
  if (lbc > 0) zero_row(0);
  if (lbc == 1) zero_row(1);
  if (rbc > 0) zero_row(nl - 1);
  if (rbc == 1) zero_row(nl - 2);
  double *st = s + nl * nl;
  for (int j = 0; j < nl; j++)
    for (int i = 0; i < nl; i++) st[i + nl * j] = s[j + nl * i];
}

}  // namespace

extern "C" {

// One-level FDM setup.  x,y,z: GLL coordinates of the level (host, nel*lx^3);
// w: GLL weights; d1: GLL derivative matrix D(i,j) (lx*lx); facet_type:
// 6 per element (0 = neighbour across the facet); gs_l: host callback that
// applies the level's gs ADD in place (neighbour lengths come from a gs
// sum).  Outputs s (nel * 6 * nl^2: S_r, S_r^T, S_s, S_s^T, S_t, S_t^T) and
// d (nel * nl^3), nl = lx + 2.
int nm_fdm_setup_host(int64_t nel, int lx, const double *x, const double *y,
                      const double *z, const double *w, const double *d1,
                      const int *facet_type,
                      void (*gs_l)(void *ctx, double *u), void *gs_ctx,
                      double *s, double *d) {
  const int n = lx - 1, nl = lx + 2, n2 = lx - 1;
  const int64_t n3 = (int64_t)lx * lx * lx;
  // semhat a = D^T W D, b = W
  std::vector<double> ah(lx * lx, 0.0), bh(w, w + lx);
  for (int j = 0; j < lx; j++)
    for (int i = 0; i < lx; i++)
      for (int k = 0; k < lx; k++) ah[i + lx * j] += d1[k + lx * i] * bh[k] * d1[k + lx * j];
  // plane_space lengths
  std::vector<double> lmr(nel), lms(nel), lmt(nel);
  const int nx = lx - 2;
  auto plane = [&](int64_t e) {
    const int64_t o = e * n3;
    auto d2 = [&](int a, int b, int c, int p, int q, int r) {
      const double dx = x[o + ix3(a, b, c, lx)] - x[o + ix3(p, q, r, lx)];
      const double dy = y[o + ix3(a, b, c, lx)] - y[o + ix3(p, q, r, lx)];
      const double dz = z[o + ix3(a, b, c, lx)] - z[o + ix3(p, q, r, lx)];
      return dx * dx + dy * dy + dz * dz;
    };
    double s2 = 0.0, ws = 0.0;
    for (int k = 1; k <= nx; k++)
      for (int j = 1; j <= nx; j++) {
        const double wt = w[j - 1] * w[k - 1];
        s2 += wt / d2(n2, j, k, 0, j, k);
        ws += wt;
      }
    lmr[e] = 1.0 / std::sqrt(s2 / ws);
    s2 = ws = 0.0;
    for (int k = 1; k <= nx; k++)
      for (int i = 1; i <= nx; i++) {
        const double wt = w[i - 1] * w[k - 1];
        s2 += wt / d2(i, n2, k, i, 0, k);
        ws += wt;
      }
    lms[e] = 1.0 / std::sqrt(s2 / ws);
    s2 = ws = 0.0;
    for (int j = 1; j <= nx; j++)
      for (int i = 1; i <= nx; i++) {
        const double wt = w[i - 1] * w[j - 1];
        s2 += wt / d2(i, j, n2, i, j, 0);
        ws += wt;
      }
    lmt[e] = 1.0 / std::sqrt(s2 / ws);
  };
  for (int64_t e = 0; e < nel; e++) plane(e);
  // neighbour lengths through one gs sum of face-interior markers
  std::vector<double> l((size_t)(nel * n3), 0.0);
  for (int64_t e = 0; e < nel; e++) {
    double *le = l.data() + e * n3;
    for (int j = 1; j < n2; j++)
      for (int k = 1; k < n2; k++) {
        le[ix3(0, k, j, lx)] = lmr[e];
        le[ix3(n2, k, j, lx)] = lmr[e];
        le[ix3(k, 0, j, lx)] = lms[e];
        le[ix3(k, n2, j, lx)] = lms[e];
        le[ix3(k, j, 0, lx)] = lmt[e];
        le[ix3(k, j, n2, lx)] = lmt[e];
      }
  }
  gs_l(gs_ctx, l.data());
  // per-element pencils, threaded
  auto work = [&](int64_t e0, int64_t e1) {
    std::vector<double> lam(3 * nl);
    for (int64_t e = e0; e < e1; e++) {
      const double *le = l.data() + e * n3;
      const double llr = le[ix3(0, 1, 1, lx)] - lmr[e], lrr = le[ix3(n2, 1, 1, lx)] - lmr[e];
      const double lls = le[ix3(1, 0, 1, lx)] - lms[e], lrs = le[ix3(1, n2, 1, lx)] - lms[e];
      const double llt = le[ix3(1, 1, 0, lx)] - lmt[e], lrt = le[ix3(1, 1, n2, lx)] - lmt[e];
      const int *ft = facet_type + 6 * e;
      double *se = s + e * 6 * nl * nl;
      setup_1d(se, lam.data(), nl, ft[0], ft[1], llr, lmr[e], lrr, ah.data(), bh.data(), n);
      setup_1d(se + 2 * nl * nl, lam.data() + nl, nl, ft[2], ft[3], lls, lms[e], lrs, ah.data(), bh.data(), n);
      setup_1d(se + 4 * nl * nl, lam.data() + 2 * nl, nl, ft[4], ft[5], llt, lmt[e], lrt, ah.data(), bh.data(), n);
      double m0 = -99e20, m1 = -99e20, m2 = -99e20;
      for (int i = 1; i < nl - 1; i++) {
        m0 = std::max(m0, lam[i]);
        m1 = std::max(m1, lam[nl + i]);
        m2 = std::max(m2, lam[2 * nl + i]);
      }
      const double eps = 1e-5 * (m0 + m1 + m2);
      double *de = d + e * (int64_t)nl * nl * nl;
      int64_t il = 0;
      for (int k = 0; k < nl; k++)
        for (int j = 0; j < nl; j++)
          for (int i = 0; i < nl; i++) {
            const double dg = lam[i] + lam[nl + j] + lam[2 * nl + k];
            de[il++] = dg > eps ? 1.0 / dg : 0.0;
          }
    }
  };
  int nt = (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  if (nel < 256) nt = 1;
  std::vector<std::thread> pool;
  const int64_t chunk = (nel + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    const int64_t a = t * chunk, b = std::min(nel, a + chunk);
    if (a < b) pool.emplace_back(work, a, b);
  }
  for (auto &th : pool) th.join();
  return nm::NM_OK;
}

// Lagrange interpolation matrix J(i,j) = l_j(z_to[i]) through the points
// z_from (column-major n_to x n_from) and its transpose.
int nm_interp_matrix(const double *z_to, int n_to, const double *z_from,
                     int n_from, double *j, double *jt) {
  for (int i = 0; i < n_to; i++)
    for (int q = 0; q < n_from; q++) {
      double v = 1.0;
      for (int m = 0; m < n_from; m++)
        if (m != q) v *= (z_to[i] - z_from[m]) / (z_from[q] - z_from[m]);
      j[i + n_to * q] = v;
      jt[q + n_from * i] = v;
    }
  return nm::NM_OK;
}

}  // extern "C"
