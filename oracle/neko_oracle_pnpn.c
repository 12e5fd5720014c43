/* TEST INFRASTRUCTURE ONLY (oracle): CPU FP64 restatement of the operator
 * pieces of Neko's pnpn residuals.  The HIP product is compared against it;
 * the product never links it.
 *
 *  or_cofactors  coef_generate_geo's drdx..dtdz (src/sem/coef.f90:903-1030):
 *                the cofactor entries of the map's Jacobian from the GLL
 *                coordinates, derivatives by D in r, s and t.
 *  or_opgrad     cpu_opgrad_lx (src/math/bcknd/cpu/cpu_opgrad.f90:263-329):
 *                ux = w3 (drdx ur + dsdx us + dtdx ut), uy, uz likewise.
 *  or_dudxyz     cpu_dudxyz_lx (cpu_dudxyz.f90:97-159), used by curl.
 *  or_cdtp       cpu_cdtp_lx (cpu_cdtp.f90:191-256), the weak divergence.
 *
 * Pinned in tests/test_oracle_pnpn.py: a constant field has zero gradient;
 * a linear field a.x gives ux = B a_x exactly on trilinear elements (the
 * cofactors are jac times the inverse Jacobian); the cofactors agree with
 * the closed-form cross products of the trilinear map's edge derivatives;
 * dudxyz of a linear field is its slope; cdtp is the adjoint of opgrad
 * (<cdtp_x(x), y> = <x, opgrad_x(y)>); curl of a gradient vanishes and
 * curl(-y, x, 0) = (0, 0, 2).
 */
#include <stdint.h>
#include <stdlib.h>

#include "neko_oracle.h"

void or_cofactors(int nel, int lx, const double *x, const double *y, const double *z,
                  const double *d, double *cof) {
  /* cof: 9 arrays of nel*lx^3, order drdx drdy drdz dsdx dsdy dsdz dtdx
   * dtdy dtdz */
  const int n3 = lx * lx * lx;
  const int64_t n = (int64_t)nel * n3;
  for (int e = 0; e < nel; e++) {
    const double *f3[3] = {x + (int64_t)e * n3, y + (int64_t)e * n3, z + (int64_t)e * n3};
    for (int k = 0; k < lx; k++)
      for (int j = 0; j < lx; j++)
        for (int i = 0; i < lx; i++) {
          double dr[3], ds[3], dt[3];
          for (int c = 0; c < 3; c++) {
            const double *f = f3[c];
            double a = 0.0, b = 0.0, cc = 0.0;
            for (int l = 0; l < lx; l++) {
              a += d[i + lx * l] * f[l + lx * (j + lx * k)];
              b += d[j + lx * l] * f[i + lx * (l + lx * k)];
              cc += d[k + lx * l] * f[i + lx * (j + lx * l)];
            }
            dr[c] = a;
            ds[c] = b;
            dt[c] = cc;
          }
          const double xr = dr[0], yr = dr[1], zr = dr[2];
          const double xs = ds[0], ys = ds[1], zs = ds[2];
          const double xt = dt[0], yt = dt[1], zt = dt[2];
          const int64_t q = (int64_t)e * n3 + i + lx * (j + lx * k);
          cof[0 * n + q] = ys * zt - zs * yt; /* drdx */
          cof[1 * n + q] = xt * zs - xs * zt; /* drdy */
          cof[2 * n + q] = xs * yt - xt * ys; /* drdz */
          cof[3 * n + q] = yt * zr - yr * zt; /* dsdx */
          cof[4 * n + q] = xr * zt - xt * zr; /* dsdy */
          cof[5 * n + q] = xt * yr - xr * yt; /* dsdz */
          cof[6 * n + q] = yr * zs - ys * zr; /* dtdx */
          cof[7 * n + q] = xs * zr - xr * zs; /* dtdy */
          cof[8 * n + q] = xr * ys - xs * yr; /* dtdz */
        }
  }
}

void or_opgrad(int nel, int lx, const double *u, const double *d, const double *w3,
               const double *cof, double *ux, double *uy, double *uz) {
  const int n3 = lx * lx * lx;
  const int64_t n = (int64_t)nel * n3;
  const double *drdx = cof, *drdy = cof + n, *drdz = cof + 2 * n;
  const double *dsdx = cof + 3 * n, *dsdy = cof + 4 * n, *dsdz = cof + 5 * n;
  const double *dtdx = cof + 6 * n, *dtdy = cof + 7 * n, *dtdz = cof + 8 * n;
  for (int e = 0; e < nel; e++) {
    const double *ue = u + (int64_t)e * n3;
    for (int k = 0; k < lx; k++)
      for (int j = 0; j < lx; j++)
        for (int i = 0; i < lx; i++) {
          double ur = 0.0, us = 0.0, ut = 0.0;
          for (int l = 0; l < lx; l++) {
            ur += d[i + lx * l] * ue[l + lx * (j + lx * k)];
            us += d[j + lx * l] * ue[i + lx * (l + lx * k)];
            ut += d[k + lx * l] * ue[i + lx * (j + lx * l)];
          }
          const int p = i + lx * (j + lx * k);
          const int64_t q = (int64_t)e * n3 + p;
          ux[q] = w3[p] * (drdx[q] * ur + dsdx[q] * us + dtdx[q] * ut);
          uy[q] = w3[p] * (dsdy[q] * us + drdy[q] * ur + dtdy[q] * ut);
          uz[q] = w3[p] * (dtdz[q] * ut + drdz[q] * ur + dsdz[q] * us);
        }
  }
}

/* cpu_dudxyz_lx (src/math/bcknd/cpu/cpu_dudxyz.f90:97-159):
 * du = ((D_r u) dr + (D_s u) ds + (D_t u) dt) * jacinv, jacinv = 1/jac */
void or_dudxyz(int nel, int lx, const double *u, const double *dr, const double *ds,
               const double *dt, const double *jac, const double *d, double *du) {
  const int n3 = lx * lx * lx;
  for (int e = 0; e < nel; e++) {
    const double *ue = u + (int64_t)e * n3;
    for (int k = 0; k < lx; k++)
      for (int j = 0; j < lx; j++)
        for (int i = 0; i < lx; i++) {
          double a = 0.0, b = 0.0, c = 0.0;
          for (int l = 0; l < lx; l++) {
            a += d[i + lx * l] * ue[l + lx * (j + lx * k)];
            b += d[j + lx * l] * ue[i + lx * (l + lx * k)];
            c += d[k + lx * l] * ue[i + lx * (j + lx * l)];
          }
          const int64_t q = (int64_t)e * n3 + i + lx * (j + lx * k);
          double v = a * dr[q];
          v = v + b * ds[q];
          v = v + c * dt[q];
          du[q] = v * (1.0 / jac[q]);
        }
  }
}

/* cpu_cdtp_lx (src/math/bcknd/cpu/cpu_cdtp.f90:191-256): with wx = x w3,
 * dtx = D^T_r (wx dr) + D^T_s (wx ds) + D^T_t (wx dt) */
void or_cdtp(int nel, int lx, const double *x, const double *dr, const double *ds,
             const double *dt, const double *d, const double *w3, double *dtx) {
  const int n3 = lx * lx * lx;
  double *ta = (double *)malloc(sizeof(double) * n3 * 3);
  for (int e = 0; e < nel; e++) {
    const int64_t o = (int64_t)e * n3;
    for (int p = 0; p < n3; p++) {
      const double wx = x[o + p] * w3[p];
      ta[p] = wx * dr[o + p];
      ta[n3 + p] = wx * ds[o + p];
      ta[2 * n3 + p] = wx * dt[o + p];
    }
    for (int k = 0; k < lx; k++)
      for (int j = 0; j < lx; j++)
        for (int i = 0; i < lx; i++) {
          double a = 0.0, b = 0.0, c = 0.0;
          for (int l = 0; l < lx; l++) {
            a += d[l + lx * i] * ta[l + lx * (j + lx * k)];          /* dxt(i,l) */
            b += d[l + lx * j] * ta[n3 + i + lx * (l + lx * k)];     /* dyt(j,l) */
            c += d[l + lx * k] * ta[2 * n3 + i + lx * (j + lx * l)]; /* dzt(k,l) */
          }
          dtx[o + i + lx * (j + lx * k)] = (a + b) + c;
        }
  }
  free(ta);
}
