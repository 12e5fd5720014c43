/*
 * neko_oracle_hsmg.c -- CPU FP64 restatement of Neko's hsmg preconditioner
 * (overlapping Schwarz with FDM local solves, p-multigrid transfers, coarse
 * Jacobi-CG) and of GMRES with an arbitrary preconditioner.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Reference files (Neko source tree):
 *   src/math/fast3d.f90        fd_weights_full :105-146, semhat :169-220,
 *                              setup_intp :245-263
 *   src/math/bcknd/cpu/tensor_cpu.f90  tnsr3d_el_nvnu_cpu :80-130
 *   src/sem/interpolation.f90  interpolator_init/map :88-187
 *   src/math/fdm.f90           fdm_init :110-159, swap_lengths, plane_space,
 *                              fdm_setup_fast :337-427, 1d_a/_b :478-560,
 *                              generalev :432-458 (LAPACK dsygv)
 *   src/math/bcknd/cpu/fdm_cpu.f90     fdm_do_fast_cpu :12-46
 *   src/math/schwarz.f90       init :112-171, setup_wt :217-283,
 *                              toreg/toext/extrude :321-500, compute :506-596,
 *                              wt3d :599-633
 *   src/krylov/pc_hsmg.f90     init :147-378, solve (CPU branch) :574-626
 *
 * Third-party boundary: fdm.f90 calls LAPACK dsygv (no version pinned).  The
 * 1-D mass matrix of the extended element is diagonal, so the generalized
 * problem A s = lam B s is solved here as the symmetric problem
 * B^-1/2 A B^-1/2 (cyclic Jacobi rotations), s = B^-1/2 q, ascending
 * eigenvalues, B-orthonormal eigenvectors -- dsygv's conventions.  The FDM
 * operator S D S^T does not depend on the eigenvector signs, so any correct
 * eigensolver reproduces it to rounding.
 *
 * Pinned (tests/test_oracle_hsmg.py): semhat by exact monomial integrals,
 * setup_intp by polynomial reproduction, tnsr3d by exact prolongation of a
 * separable polynomial and the restriction/prolongation adjoint identity,
 * the 1-D FDM pencil by generic finite-element assembly of the three
 * elements, the eigen-decomposition by scipy.linalg.eigh, plane_space by
 * undeformed element lengths, the whole Schwarz/FDM path by the exact
 * inverse on one undeformed Dirichlet element, the overlap weights by a
 * brute-force count of the extended subdomains holding each point, and the
 * V-cycle by the additive two-level operator S_f + P (S_m + P_c A_c^-1
 * P_c^T) P^T with an exact coarse solve (dense matrices, numpy
 * interpolation), and by its effect on GMRES.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "neko_oracle.h"
#include "neko_oracle_hsmg.h"

#define IX3(i, j, k, n) ((int64_t)(i) + (int64_t)(n) * ((j) + (int64_t)(n) * (k)))

/* fast3d.f90:105-146 (Fornberg), c[j + (n+1)*k], k = derivative order */
void or_fd_weights(double xi, const double *x, int n, int m, double *c) {
  double c1 = 1.0, c2, c3, c4 = x[0] - xi, c5;
  const int N = n + 1;
  for (int k = 0; k <= m; k++)
    for (int j = 0; j <= n; j++) c[j + N * k] = 0.0;
  c[0] = 1.0;
  for (int i = 1; i <= n; i++) {
    int mn = i < m ? i : m;
    c2 = 1.0;
    c5 = c4;
    c4 = x[i] - xi;
    for (int j = 0; j <= i - 1; j++) {
      c3 = x[i] - x[j];
      c2 = c2 * c3;
      for (int k = mn; k >= 1; k--)
        c[i + N * k] = c1 * (k * c[(i - 1) + N * (k - 1)] - c5 * c[(i - 1) + N * k]) / c2;
      c[i] = -c1 * c5 * c[i - 1] / c2;
      for (int k = mn; k >= 1; k--)
        c[j + N * k] = (c4 * c[j + N * k] - k * c[j + N * (k - 1)]) / c3;
      c[j] = c4 * c[j] / c3;
    }
    c1 = c2;
  }
}

/* fast3d.f90:245-263: jh(i,j) (n_to x n_from, column-major) and jht */
void or_setup_intp(double *jh, double *jht, const double *z_to,
                   const double *z_from, int n_to, int n_from, int deriv) {
  double *w = (double *)malloc(sizeof(double) * n_from * (deriv + 1));
  for (int i = 0; i < n_to; i++) {
    or_fd_weights(z_to[i], z_from, n_from - 1, deriv, w);
    for (int j = 0; j < n_from; j++) {
      jh[i + n_to * j] = w[j + n_from * deriv];
      jht[j + n_from * i] = w[j + n_from * deriv];
    }
  }
  free(w);
}

/* tensor_cpu.f90:80-130: v(i,j,l) = sum A(i,a) Bt(b,j) Ct(c,l) u(a,b,c)
 * with A (nv x nu), Bt, Ct (nu x nv), column-major; per element. */
static void tnsr3d_el(double *v, int nv, const double *u, int nu,
                      const double *A, const double *Bt, const double *Ct,
                      double *work, double *work2) {
  const int nunu = nu * nu, nvnu = nv * nu, nvnv = nv * nv;
  for (int j = 0; j < nunu; j++)
    for (int i = 0; i < nv; i++) {
      double t = 0.0;
      for (int k = 0; k < nu; k++) t = t + A[i + nv * k] * u[k + nu * j];
      work[i + nv * j] = t;
    }
  for (int i = 0; i < nu; i++)
    for (int j = 0; j < nv; j++)
      for (int l = 0; l < nv; l++) {
        double t = 0.0;
        for (int k = 0; k < nu; k++)
          t = t + work[l + nv * k + nvnu * i] * Bt[k + nu * j];
        work2[l + nv * j + nvnv * i] = t;
      }
  for (int j = 0; j < nv; j++)
    for (int i = 0; i < nvnv; i++) {
      double t = 0.0;
      for (int k = 0; k < nu; k++) t = t + work2[i + nvnv * k] * Ct[k + nu * j];
      v[i + nvnv * j] = t;
    }
}

void or_tnsr3d(double *v, int nv, const double *u, int nu, const double *A,
               const double *Bt, const double *Ct, int nel) {
  const int mx = (nv > nu ? nv : nu);
  double *work = (double *)malloc(sizeof(double) * mx * mx * mx);
  double *work2 = (double *)malloc(sizeof(double) * mx * mx * mx);
  for (int e = 0; e < nel; e++)
    tnsr3d_el(v + (int64_t)e * nv * nv * nv, nv, u + (int64_t)e * nu * nu * nu,
              nu, A, Bt, Ct, work, work2);
  free(work);
  free(work2);
}

/* fast3d.f90:169-220 (only a = D^T B D and b needed by fdm) */
void or_semhat(int n, double *ah, double *bh) {
  const int np = n + 1;
  double *z = (double *)malloc(sizeof(double) * np);
  double *w = (double *)malloc(sizeof(double) * 2 * np);
  double *d = (double *)malloc(sizeof(double) * np * np);
  or_zwgll(z, bh, np);
  for (int i = 0; i <= n; i++) {
    or_fd_weights(z[i], z, n, 1, w);
    for (int j = 0; j <= n; j++) d[i + np * j] = w[j + np];
  }
  for (int q = 0; q < np * np; q++) ah[q] = 0.0;
  for (int j = 0; j <= n; j++)
    for (int i = 0; i <= n; i++)
      for (int k = 0; k <= n; k++)
        ah[i + np * j] = ah[i + np * j] + d[k + np * i] * bh[k] * d[k + np * j];
  free(z);
  free(w);
  free(d);
}

/* ---- generalized symmetric eigenproblem with diagonal B (dsygv role) ---- */
static void geneig_diagb(double *a /* n x n, overwritten by eigenvectors */,
                         const double *bdiag, double *lam, int n) {
  double *c = (double *)malloc(sizeof(double) * n * n);
  double *q = (double *)malloc(sizeof(double) * n * n);
  double *isb = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) isb[i] = 1.0 / sqrt(bdiag[i]);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      /* symmetric from the upper triangle, as dsygv('U') reads it */
      double aij = (i <= j) ? a[i + n * j] : a[j + n * i];
      c[i + n * j] = isb[i] * aij * isb[j];
      q[i + n * j] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        tot += c[i + n * j] * c[i + n * j];
        if (i != j) off += c[i + n * j] * c[i + n * j];
      }
    if (off <= 1e-32 * tot) break;
    for (int p = 0; p < n - 1; p++)
      for (int r = p + 1; r < n; r++) {
        const double apr = c[p + n * r];
        if (fabs(apr) < 1e-300) continue;
        const double app = c[p + n * p], arr = c[r + n * r];
        const double theta = (arr - app) / (2.0 * apr);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; k++) { /* columns p, r */
          const double ckp = c[k + n * p], ckr = c[k + n * r];
          c[k + n * p] = cs * ckp - sn * ckr;
          c[k + n * r] = sn * ckp + cs * ckr;
        }
        for (int k = 0; k < n; k++) { /* rows p, r */
          const double cpk = c[p + n * k], crk = c[r + n * k];
          c[p + n * k] = cs * cpk - sn * crk;
          c[r + n * k] = sn * cpk + cs * crk;
        }
        for (int k = 0; k < n; k++) {
          const double qkp = q[k + n * p], qkr = q[k + n * r];
          q[k + n * p] = cs * qkp - sn * qkr;
          q[k + n * r] = sn * qkp + cs * qkr;
        }
      }
  }
  /* ascending order */
  int *ord = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) ord[i] = i;
  for (int i = 1; i < n; i++) {
    int v = ord[i], j = i - 1;
    while (j >= 0 && c[ord[j] * (n + 1)] > c[v * (n + 1)]) {
      ord[j + 1] = ord[j];
      j--;
    }
    ord[j + 1] = v;
  }
  for (int jj = 0; jj < n; jj++) {
    const int j = ord[jj];
    lam[jj] = c[j + n * j];
    for (int i = 0; i < n; i++) a[i + n * jj] = isb[i] * q[i + n * j];
  }
  free(ord);
  free(c);
  free(q);
  free(isb);
}

/* fdm.f90:478-520 / :522-560 (0-based (n+3)x(n+3) matrices) */
static void fdm_1d_ab(double *a, double *b, int lbc, int rbc, double ll,
                      double lm, double lr, const double *ah, const double *bh,
                      int n) {
  const int N = n + 3, np = n + 1;
#define A(i, j) a[(i) + N * (j)]
#define B(i, j) b[(i) + N * (j)]
#define AH(i, j) ah[(i) + np * (j)]
  int i0 = 0, i1 = n;
  if (lbc == 1) i0 = 1;
  if (rbc == 1) i1 = n - 1;
  for (int q = 0; q < N * N; q++) { a[q] = 0.0; b[q] = 0.0; }
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
  B(1, 1) = 1.0;
  B(n + 1, n + 1) = 1.0;
  for (int i = i0; i <= i1; i++) B(i + 1, i + 1) = fac * bh[i];
  if (lbc == 0) {
    fac = 0.5 * ll;
    B(0, 0) = fac * bh[n - 1];
    B(1, 1) = B(1, 1) + fac * bh[n];
  } else {
    B(0, 0) = 1.0;
  }
  if (rbc == 0) {
    fac = 0.5 * lr;
    B(n + 1, n + 1) = B(n + 1, n + 1) + fac * bh[0];
    B(n + 2, n + 2) = fac * bh[1];
  } else {
    B(n + 2, n + 2) = 1.0;
  }
#undef A
#undef B
#undef AH
}

/* fdm.f90:410-426 fdm_setup_fast1d: s(:,:,1) eigenvectors with boundary
 * rows zeroed, s(:,:,2) its transpose, lam eigenvalues */
static void fdm_setup_fast1d(double *s, double *lam, int nl, int lbc, int rbc,
                             double ll, double lm, double lr, const double *ah,
                             const double *bh, int n) {
  double *b = (double *)malloc(sizeof(double) * nl * nl);
  double *bd = (double *)malloc(sizeof(double) * nl);
  fdm_1d_ab(s, b, lbc, rbc, ll, lm, lr, ah, bh, n);
  for (int i = 0; i < nl; i++) bd[i] = b[i + nl * i];
  geneig_diagb(s, bd, lam, nl);
  /* row_zero(s, nl, nl, e): zero row e (1-based) */
  if (lbc > 0) for (int j = 0; j < nl; j++) s[0 + nl * j] = 0.0;
  if (lbc == 1) for (int j = 0; j < nl; j++) s[1 + nl * j] = 0.0;
  if (rbc > 0) for (int j = 0; j < nl; j++) s[(nl - 1) + nl * j] = 0.0;
  if (rbc == 1) for (int j = 0; j < nl; j++) s[(nl - 2) + nl * j] = 0.0;
  double *st = s + nl * nl;
  for (int j = 0; j < nl; j++)
    for (int i = 0; i < nl; i++) st[i + nl * j] = s[j + nl * i];
  free(b);
  free(bd);
}

/* Test hook: the 1-D pencil (a, b, (n+3)^2 each) of fdm_1d_ab and the
 * eigen-decomposition of fdm_setup_fast1d (s: eigenvectors with boundary
 * rows zeroed, then its transpose; lam: eigenvalues). */
void or_fdm_1d(double *a, double *b, double *s, double *lam, int lbc, int rbc,
               double ll, double lm, double lr, int n) {
  const int np = n + 1, nl = n + 3;
  double *ah = (double *)malloc(sizeof(double) * np * np);
  double *bh = (double *)malloc(sizeof(double) * np);
  or_semhat(n, ah, bh);
  fdm_1d_ab(a, b, lbc, rbc, ll, lm, lr, ah, bh, n);
  fdm_setup_fast1d(s, lam, nl, lbc, rbc, ll, lm, lr, ah, bh, n);
  free(ah);
  free(bh);
}

/* fdm.f90 plane_space (3-D branch): weighted mean of inverse squared
 * distances between the two planes i1,i2 (0-based GLL indices) */
static void plane_space(double *lr, double *ls, double *lt, int i1, int i2,
                        const double *w, const double *x, const double *y,
                        const double *z, int lx, int nel) {
  const int nx = lx - 2;
  for (int e = 0; e < nel; e++) {
    const int64_t o = (int64_t)e * lx * lx * lx;
#define P(arr, i, j, k) arr[o + IX3(i, j, k, lx)]
#define D2(a, b, c, d, e_, f)                                                  \
  ((P(x, a, b, c) - P(x, d, e_, f)) * (P(x, a, b, c) - P(x, d, e_, f)) +        \
   (P(y, a, b, c) - P(y, d, e_, f)) * (P(y, a, b, c) - P(y, d, e_, f)) +        \
   (P(z, a, b, c) - P(z, d, e_, f)) * (P(z, a, b, c) - P(z, d, e_, f)))
    double s2 = 0.0, ws = 0.0;
    for (int k = 1; k <= nx; k++)
      for (int j = 1; j <= nx; j++) {
        const double wt = w[j - 1] * w[k - 1];
        s2 = s2 + wt / D2(i2, j, k, i1, j, k);
        ws = ws + wt;
      }
    lr[e] = 1.0 / sqrt(s2 / ws);
    s2 = 0.0; ws = 0.0;
    for (int k = 1; k <= nx; k++)
      for (int i = 1; i <= nx; i++) {
        const double wt = w[i - 1] * w[k - 1];
        s2 = s2 + wt / D2(i, i2, k, i, i1, k);
        ws = ws + wt;
      }
    ls[e] = 1.0 / sqrt(s2 / ws);
    s2 = 0.0; ws = 0.0;
    for (int j = 1; j <= nx; j++)
      for (int i = 1; i <= nx; i++) {
        const double wt = w[i - 1] * w[j - 1];
        s2 = s2 + wt / D2(i, j, i2, i, j, i1);
        ws = ws + wt;
      }
    lt[e] = 1.0 / sqrt(s2 / ws);
#undef P
#undef D2
  }
}

/* Test hook: plane_space lengths (first/last GLL planes) of every element */
void or_plane_space(double *lr, double *ls, double *lt, const double *w,
                    const double *x, const double *y, const double *z, int lx,
                    int nel) {
  plane_space(lr, ls, lt, 0, lx - 1, w, x, y, z, lx, nel);
}

void or_gs_apply(const or_gs_t *g, double *u) {
  or_gs_op_add(u, g->n, g->m, g->o, g->dg, g->gd, g->nb, g->b, g->bo, g->buf);
}

/* fdm_init + swap_lengths + fdm_setup_fast (3-D).  wx: GLL weights of the
 * level (the first lx-2 entries are used, as plane_space receives
 * Xh%wx with length nx).  facet_type: 6 per element (0 = neighbour). */
void or_fdm_setup(int nel, int lx, const double *x, const double *y,
                  const double *z, const double *wx, const int *facet_type,
                  const or_gs_t *gs, double *s, double *d) {
  const int n = lx - 1, nl = lx + 2, n2 = lx - 1;
  const int64_t n3 = (int64_t)lx * lx * lx;
  double *ah = (double *)malloc(sizeof(double) * lx * lx);
  double *bh = (double *)malloc(sizeof(double) * lx);
  or_semhat(n, ah, bh);
  double *lmr = malloc(sizeof(double) * nel), *lms = malloc(sizeof(double) * nel),
         *lmt = malloc(sizeof(double) * nel);
  plane_space(lmr, lms, lmt, 0, n2, wx, x, y, z, lx, nel);
  double *l = (double *)calloc((size_t)nel * n3, sizeof(double));
  for (int e = 0; e < nel; e++) {
    double *le = l + e * n3;
    for (int j = 1; j < n2; j++)
      for (int k = 1; k < n2; k++) {
        le[IX3(0, k, j, lx)] = lmr[e];
        le[IX3(n2, k, j, lx)] = lmr[e];
        le[IX3(k, 0, j, lx)] = lms[e];
        le[IX3(k, n2, j, lx)] = lms[e];
        le[IX3(k, j, 0, lx)] = lmt[e];
        le[IX3(k, j, n2, lx)] = lmt[e];
      }
  }
  or_gs_apply(gs, l);
  double *lam = (double *)malloc(sizeof(double) * 3 * nl);
  for (int e = 0; e < nel; e++) {
    const double *le = l + e * n3;
    const double llr = le[IX3(0, 1, 1, lx)] - lmr[e];
    const double lrr = le[IX3(n2, 1, 1, lx)] - lmr[e];
    const double lls = le[IX3(1, 0, 1, lx)] - lms[e];
    const double lrs = le[IX3(1, n2, 1, lx)] - lms[e];
    const double llt = le[IX3(1, 1, 0, lx)] - lmt[e];
    const double lrt = le[IX3(1, 1, n2, lx)] - lmt[e];
    const int *ft = facet_type + 6 * e;
    double *se = s + (int64_t)e * nl * nl * 2 * 3;
    fdm_setup_fast1d(se, lam, nl, ft[0], ft[1], llr, lmr[e], lrr, ah, bh, n);
    fdm_setup_fast1d(se + 2 * nl * nl, lam + nl, nl, ft[2], ft[3], lls, lms[e], lrs, ah, bh, n);
    fdm_setup_fast1d(se + 4 * nl * nl, lam + 2 * nl, nl, ft[4], ft[5], llt, lmt[e], lrt, ah, bh, n);
    double vmax = -99e20, smax = -99e20, tmax = -99e20;
    for (int i = 1; i < nl - 1; i++) {
      if (lam[i] > vmax) vmax = lam[i];
      if (lam[nl + i] > smax) smax = lam[nl + i];
      if (lam[2 * nl + i] > tmax) tmax = lam[2 * nl + i];
    }
    const double eps = 1e-5 * (vmax + smax + tmax);
    double *de = d + (int64_t)e * nl * nl * nl;
    int il = 0;
    for (int k = 0; k < nl; k++)
      for (int j = 0; j < nl; j++)
        for (int i = 0; i < nl; i++) {
          const double diag = lam[i] + lam[nl + j] + lam[2 * nl + k];
          de[il++] = diag > eps ? 1.0 / diag : 0.0;
        }
  }
  free(lam);
  free(l);
  free(lmr); free(lms); free(lmt);
  free(ah);
  free(bh);
}

/* fdm_cpu.f90:33-44 (r is overwritten, as in the reference) */
void or_fdm_compute(int nel, int nl, double *e, double *r, const double *s,
                    const double *d) {
  const int64_t nn = (int64_t)nl * nl * nl;
  double *work = (double *)malloc(sizeof(double) * nn);
  double *work2 = (double *)malloc(sizeof(double) * nn);
  for (int ie = 0; ie < nel; ie++) {
    const double *se = s + (int64_t)ie * nl * nl * 6;
    const double *S1 = se, *S1t = se + nl * nl, *S2 = se + 2 * nl * nl,
                 *S2t = se + 3 * nl * nl, *S3 = se + 4 * nl * nl, *S3t = se + 5 * nl * nl;
    double *ee = e + ie * nn, *rr = r + ie * nn;
    tnsr3d_el(ee, nl, rr, nl, S1t, S2, S3, work, work2);
    for (int64_t i = 0; i < nn; i++) rr[i] = d[ie * nn + i] * ee[i];
    tnsr3d_el(ee, nl, rr, nl, S1, S2t, S3t, work, work2);
  }
  free(work);
  free(work2);
}

/* schwarz.f90:378-440 extrude: arr1 face layers (offset l1) =
 * f1*arr1 + f2*arr2 (offset l2), face interiors only, 3-D */
void or_schwarz_extrude(double *arr1, int l1, double f1, const double *arr2,
                        int l2, double f2, int nx, int nel) {
  const int64_t n3 = (int64_t)nx * nx * nx;
  const int i0 = 1, i1 = nx - 2; /* 0-based face-interior range */
  for (int e = 0; e < nel; e++) {
    double *a = arr1 + e * n3;
    const double *b = arr2 + e * n3;
#define A1(i, j, k) a[IX3(i, j, k, nx)]
#define A2(i, j, k) b[IX3(i, j, k, nx)]
    for (int k = i0; k <= i1; k++)
      for (int j = i0; j <= i1; j++) {
        A1(l1, j, k) = f1 * A1(l1, j, k) + f2 * A2(l2, j, k);
        A1(nx - 1 - l1, j, k) = f1 * A1(nx - 1 - l1, j, k) + f2 * A2(nx - 1 - l2, j, k);
      }
    for (int k = i0; k <= i1; k++)
      for (int i = i0; i <= i1; i++) {
        A1(i, l1, k) = f1 * A1(i, l1, k) + f2 * A2(i, l2, k);
        A1(i, nx - 1 - l1, k) = f1 * A1(i, nx - 1 - l1, k) + f2 * A2(i, nx - 1 - l2, k);
      }
    for (int j = i0; j <= i1; j++)
      for (int i = i0; i <= i1; i++) {
        A1(i, j, l1) = f1 * A1(i, j, l1) + f2 * A2(i, j, l2);
        A1(i, j, nx - 1 - l1) = f1 * A1(i, j, nx - 1 - l1) + f2 * A2(i, j, nx - 1 - l2);
      }
#undef A1
#undef A2
  }
}

/* extrude_single: same with arr2 == arr1 (schwarz.f90:442-503) */
void or_schwarz_extrude_single(double *arr, int l1, double f1, int l2,
                               double f2, int nx, int nel) {
  or_schwarz_extrude(arr, l1, f1, arr, l2, f2, nx, nel);
}

void or_schwarz_toext3d(double *a, const double *b, int n, int nel) {
  const int m = n + 2;
  memset(a, 0, sizeof(double) * (size_t)nel * m * m * m);
  for (int e = 0; e < nel; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
          a[(int64_t)e * m * m * m + IX3(i + 1, j + 1, k + 1, m)] =
              b[(int64_t)e * n * n * n + IX3(i, j, k, n)];
}

void or_schwarz_toreg3d(double *b, const double *a, int n, int nel) {
  const int m = n + 2;
  for (int e = 0; e < nel; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
          b[(int64_t)e * n * n * n + IX3(i, j, k, n)] =
              a[(int64_t)e * m * m * m + IX3(i + 1, j + 1, k + 1, m)];
}

/* schwarz_setup_wt (schwarz.f90:217-283) + wt3d_2 layout
 * wt(n, n, 4, 3, nel) */
void or_schwarz_setup_wt(int nel, int lx, const or_gs_t *gs_ext,
                         const or_gs_t *gs, double *wt) {
  const int nx = lx + 2;
  const int64_t ns = (int64_t)nx * nx * nx * nel, n = (int64_t)lx * lx * lx * nel;
  double *w1 = (double *)calloc((size_t)ns, sizeof(double));
  double *w2 = (double *)malloc(sizeof(double) * ns);
  double *e1 = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < ns; i++) w2[i] = 1.0;
  or_schwarz_extrude(w1, 0, 0.0, w2, 0, 1.0, nx, nel);
  or_gs_apply(gs_ext, w2);
  or_schwarz_extrude(w2, 0, 1.0, w1, 0, -1.0, nx, nel);
  or_schwarz_extrude_single(w2, 2, 1.0, 0, 1.0, nx, nel);
  or_schwarz_toreg3d(e1, w2, lx, nel);
  or_gs_apply(gs, e1);
  const int N = lx;
  for (int e = 0; e < nel; e++) {
    const double *wk = e1 + (int64_t)e * N * N * N;
    double *we = wt + (int64_t)e * N * N * 4 * 3;
#define WK(i, j, k) wk[IX3(i, j, k, N)]
#define WT(a, b, c, d) we[(a) + N * ((b) + N * ((c) + 4 * (d)))]
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++) {
        WT(j, k, 0, 0) = 1.0 / WK(0, j, k);
        WT(j, k, 1, 0) = 1.0 / WK(1, j, k);
        WT(j, k, 2, 0) = 1.0 / WK(N - 2, j, k);
        WT(j, k, 3, 0) = 1.0 / WK(N - 1, j, k);
      }
    for (int k = 0; k < N; k++)
      for (int i = 0; i < N; i++) {
        WT(i, k, 0, 1) = 1.0 / WK(i, 0, k);
        WT(i, k, 1, 1) = 1.0 / WK(i, 1, k);
        WT(i, k, 2, 1) = 1.0 / WK(i, N - 2, k);
        WT(i, k, 3, 1) = 1.0 / WK(i, N - 1, k);
      }
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        WT(i, j, 0, 2) = 1.0 / WK(i, j, 0);
        WT(i, j, 1, 2) = 1.0 / WK(i, j, 1);
        WT(i, j, 2, 2) = 1.0 / WK(i, j, N - 2);
        WT(i, j, 3, 2) = 1.0 / WK(i, j, N - 1);
      }
#undef WK
#undef WT
  }
  free(w1);
  free(w2);
  free(e1);
}

/* schwarz_wt3d (schwarz.f90:599-633) */
void or_schwarz_wt3d(double *e, const double *wt, int n, int nel) {
  for (int ie = 0; ie < nel; ie++) {
    double *ee = e + (int64_t)ie * n * n * n;
    const double *we = wt + (int64_t)ie * n * n * 4 * 3;
#define E(i, j, k) ee[IX3(i, j, k, n)]
#define WT(a, b, c, d) we[(a) + n * ((b) + n * ((c) + 4 * (d)))]
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++) {
        E(0, j, k) = E(0, j, k) * WT(j, k, 0, 0);
        E(1, j, k) = E(1, j, k) * WT(j, k, 1, 0);
        E(n - 2, j, k) = E(n - 2, j, k) * WT(j, k, 2, 0);
        E(n - 1, j, k) = E(n - 1, j, k) * WT(j, k, 3, 0);
      }
    for (int k = 0; k < n; k++)
      for (int i = 2; i < n - 2; i++) {
        E(i, 0, k) = E(i, 0, k) * WT(i, k, 0, 1);
        E(i, 1, k) = E(i, 1, k) * WT(i, k, 1, 1);
        E(i, n - 2, k) = E(i, n - 2, k) * WT(i, k, 2, 1);
        E(i, n - 1, k) = E(i, n - 1, k) * WT(i, k, 3, 1);
      }
    for (int j = 2; j < n - 2; j++)
      for (int i = 2; i < n - 2; i++) {
        E(i, j, 0) = E(i, j, 0) * WT(i, j, 0, 2);
        E(i, j, 1) = E(i, j, 1) * WT(i, j, 1, 2);
        E(i, j, n - 2) = E(i, j, n - 2) * WT(i, j, 2, 2);
        E(i, j, n - 1) = E(i, j, n - 1) * WT(i, j, 3, 2);
      }
#undef E
#undef WT
  }
}

/* schwarz_compute, CPU branch (schwarz.f90:565-595).  r is masked in place
 * (bclst%apply_scalar) as in the reference. */
void or_schwarz_compute(const or_schwarz_t *sw, double *e, double *r) {
  const int lx = sw->lx, nx = lx + 2, nel = sw->nel;
  or_mask_apply(r, sw->msk, sw->nmsk);
  or_schwarz_toext3d(sw->work1, r, lx, nel);
  or_schwarz_extrude_single(sw->work1, 0, 0.0, 2, 1.0, nx, nel);
  or_gs_apply(&sw->gs_ext, sw->work1);
  or_schwarz_extrude_single(sw->work1, 0, 1.0, 2, -1.0, nx, nel);
  or_fdm_compute(nel, nx, sw->work2, sw->work1, sw->s, sw->d);
  or_schwarz_extrude(sw->work1, 0, 0.0, sw->work2, 0, 1.0, nx, nel);
  or_gs_apply(&sw->gs_ext, sw->work2);
  or_schwarz_extrude(sw->work2, 0, 1.0, sw->work1, 0, -1.0, nx, nel);
  or_schwarz_extrude_single(sw->work2, 2, 1.0, 0, 1.0, nx, nel);
  or_schwarz_toreg3d(e, sw->work2, lx, nel);
  or_gs_apply(&sw->gs, e);
  or_mask_apply(e, sw->msk, sw->nmsk);
  or_schwarz_wt3d(e, sw->wt, lx, nel);
}

/* hsmg_solve, CPU branch (pc_hsmg.f90:601-626) */
void or_hsmg_solve(or_hsmg_t *h, double *z, const double *rin) {
  const int nel = h->nel;
  const int64_t nf = h->nf, nm = h->nm, nc = h->nc;
  double *r = h->r, *w = h->w;
  memcpy(r, rin, sizeof(double) * nf);
  or_schwarz_compute(&h->sw_fine, z, r);
  for (int64_t i = 0; i < nf; i++) r[i] = r[i] * h->mult_f[i];
  /* fine -> mid restriction: tnsr3d(w, lm, r, lf, J_fm^T, J_fm, J_fm) */
  or_tnsr3d(w, h->lm, r, h->lf, h->jfm_t, h->jfm, h->jfm, nel);
  or_gs_apply(&h->sw_mid.gs, w);
  or_schwarz_compute(&h->sw_mid, h->e_mid, w);
  for (int64_t i = 0; i < nm; i++) w[i] = w[i] * h->mult_m[i];
  or_tnsr3d(r, h->lc, w, h->lm, h->jmc_t, h->jmc, h->jmc, nel);
  or_gs_apply(&h->gs_crs, r);
  or_mask_apply(r, h->msk_c, h->nmsk_c);
  or_monitor_t mon = or_cg_solve(&h->crs, h->pc_crs, h->e_crs, r, h->niter, h->crs_tol);
  h->last_crs_iter = mon.iter;
  or_mask_apply(h->e_crs, h->msk_c, h->nmsk_c);
  /* prolongation: mid <- crs, fine <- mid */
  or_tnsr3d(w, h->lm, h->e_crs, h->lc, h->jmc, h->jmc_t, h->jmc_t, nel);
  for (int64_t i = 0; i < nm; i++) h->e_mid[i] = h->e_mid[i] + w[i];
  or_tnsr3d(w, h->lf, h->e_mid, h->lm, h->jfm, h->jfm_t, h->jfm_t, nel);
  for (int64_t i = 0; i < nf; i++) z[i] = z[i] + w[i];
  or_gs_apply(&h->sw_fine.gs, z);
  for (int64_t i = 0; i < nf; i++) z[i] = z[i] * h->mult_f[i];
  (void)nc;
}

/* GMRES (gmres.f90:175-411) with a preconditioner callback */
static void pc_hsmg_cb(void *ctx, double *z, const double *v, int64_t n) {
  (void)n;
  or_hsmg_solve((or_hsmg_t *)ctx, z, v);
}

or_monitor_t or_gmres_hsmg_solve(const or_problem_t *p, or_hsmg_t *h,
                                 double *x, const double *f, int max_iter,
                                 double abs_tol, int lgmres) {
  return or_gmres_solve_pc(p, pc_hsmg_cb, h, x, f, max_iter, abs_tol, lgmres);
}
