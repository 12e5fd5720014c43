// Operator glue of Neko's pnpn scheme around the velocity solve: the weak
// gradient opgrad (src/math/bcknd/cpu/cpu_opgrad.f90:263-329) and the
// velocity residual (pnpn_vel_res_cpu_compute, src/fluid/bcknd/cpu/
// pnpn_res_cpu.f90:182-241), on straight-edged (trilinear) operators whose
// geometry is evaluated in the kernel from the element edges, as in the
// trilinear Ax.
//
// opgrad at point (i,j,k) of an element: with the GLL derivatives
// (ur, us, ut) of u and the cofactor rows r = X_s x X_t, s = X_t x X_r,
// t = X_r x X_s (Neko's drdx..dtdz), ux = w3 (r_x ur + s_x us + t_x ut) and
// likewise uy, uz.  The residual kernel fuses the combination
// res_c = -res_c - opgrad_c(p) + f_c, where res_c holds A u_c on entry.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "neko_amd.h"
#include "nm_common.h"
#include "nm_kernels.h"
#include "nm_op.h"

namespace nm {

namespace {

template <int LX, bool RES>
__global__ void __launch_bounds__(LX *LX)
    opgrad_tri_kernel(const double *__restrict__ u, const double *__restrict__ dx,
                      const double *__restrict__ zg, const double *__restrict__ wg,
                      const double *__restrict__ edges, int64_t nel,
                      double *__restrict__ ox, double *__restrict__ oy,
                      double *__restrict__ oz, const double *__restrict__ fx,
                      const double *__restrict__ fy, const double *__restrict__ fz) {
  constexpr int P = LX * LX;
  __shared__ double su[LX * P];
  __shared__ double sa[LX][LX][3];
  __shared__ double sb[LX][LX][3];
  __shared__ double se[36];
  const int64_t e = blockIdx.x;
  if (e >= nel) return;
  const int t = threadIdx.x, i = t % LX, j = t / LX;
  const int64_t base = e * (int64_t)(LX * P) + t;
  double uc[LX];
#pragma unroll
  for (int k = 0; k < LX; k++) {
    uc[k] = u[base + k * P];
    su[k * P + t] = uc[k];
  }
  for (int x = t; x < 36; x += P) se[x] = edges[e * 36 + x];
  __syncthreads();
  // dX/dr(s_j', t_i') and dX/ds(r_j', t_i') tables, one pair per thread
  {
    const double p1 = (1.0 + zg[j]) * 0.5, p0 = 1.0 - p1;
    const double t1 = (1.0 + zg[i]) * 0.5, t0 = 1.0 - t1;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      sa[i][j][c] = (se[c * 4 + 0] * p0 + se[c * 4 + 1] * p1) * t0 +
                    (se[c * 4 + 2] * p0 + se[c * 4 + 3] * p1) * t1;
      sb[i][j][c] = (se[12 + c * 4 + 0] * p0 + se[12 + c * 4 + 1] * p1) * t0 +
                    (se[12 + c * 4 + 2] * p0 + se[12 + c * 4 + 3] * p1) * t1;
    }
  }
  const double ri1 = (1.0 + zg[i]) * 0.5, ri0 = 1.0 - ri1;
  const double sj1 = (1.0 + zg[j]) * 0.5, sj0 = 1.0 - sj1;
  double ct[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    ct[c] = (se[24 + c * 4 + 0] * ri0 + se[24 + c * 4 + 1] * ri1) * sj0 +
            (se[24 + c * 4 + 2] * ri0 + se[24 + c * 4 + 3] * ri1) * sj1;
  double Di[LX], Dj[LX];
#pragma unroll
  for (int l = 0; l < LX; l++) {
    Di[l] = dx[i + LX * l];
    Dj[l] = dx[j + LX * l];
  }
  const double wij = wg[i] * wg[j];
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < LX; k++) {
    const double *pl = &su[k * P];
    double ur = 0.0, us = 0.0, ut = 0.0;
#pragma unroll
    for (int l = 0; l < LX; l++) {
      ur += Di[l] * pl[l + LX * j];
      us += Dj[l] * pl[i + LX * l];
      ut += dx[k + LX * l] * uc[l];
    }
    const double a0 = sa[k][j][0], a1 = sa[k][j][1], a2 = sa[k][j][2];
    const double b0 = sb[k][i][0], b1 = sb[k][i][1], b2 = sb[k][i][2];
    const double rx = b1 * ct[2] - ct[1] * b2, ry = ct[0] * b2 - b0 * ct[2],
                 rz = b0 * ct[1] - ct[0] * b1;
    const double sx = ct[1] * a2 - a1 * ct[2], sy = a0 * ct[2] - ct[0] * a2,
                 sz = ct[0] * a1 - a0 * ct[1];
    const double tx = a1 * b2 - b1 * a2, ty = b0 * a2 - a0 * b2, tz = a0 * b1 - b0 * a1;
    const double w3 = wij * wg[k];
    const double gx = w3 * (rx * ur + sx * us + tx * ut);
    const double gy = w3 * (sy * us + ry * ur + ty * ut);
    const double gz = w3 * (tz * ut + rz * ur + sz * us);
    const int64_t q = base + k * P;
    if (RES) {
      ox[q] = (-ox[q]) - gx + fx[q];
      oy[q] = (-oy[q]) - gy + fy[q];
      oz[q] = (-oz[q]) - gz + fz[q];
    } else {
      ox[q] = gx;
      oy[q] = gy;
      oz[q] = gz;
    }
  }
}

template <bool RES>
int launch_opgrad(const Op &op, const double *u, double *ox, double *oy, double *oz,
                  const double *fx, const double *fy, const double *fz, hipStream_t s) {
  const dim3 g((unsigned)op.nel);
  const OpFields<double> &F = op.f64;
#define NM_OPG(L)                                                                         \
  case L:                                                                                 \
    hipLaunchKernelGGL((opgrad_tri_kernel<L, RES>), g, dim3(L * L), 0, s, u, F.dx, F.zg,  \
                       F.wg, F.edges, op.nel, ox, oy, oz, fx, fy, fz);                    \
    break;
  switch (op.lx) {
    NM_OPG(2) NM_OPG(3) NM_OPG(4) NM_OPG(5) NM_OPG(6) NM_OPG(7) NM_OPG(8) NM_OPG(9)
    NM_OPG(10) NM_OPG(11) NM_OPG(12)
    default:
      set_error("opgrad: lx=%d outside 2..12", op.lx);
      return NM_ERR_UNSUPPORTED;
  }
#undef NM_OPG
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

int check_tri(const Op &op, const char *who) {
  if (!op.trilinear || !op.f64.edges) {
    set_error("%s: needs an operator on a straight-edged mesh (box or mesh constructor)",
              who);
    return NM_ERR_UNSUPPORTED;
  }
  return NM_OK;
}

}  // namespace

}  // namespace nm

using namespace nm;

extern "C" {

// opgrad(ux, uy, uz, u, coef) (src/math/operators.f90:307-343), device
// pointers, local (no gs)
int nm_opgrad(nm_op_t *h, double *ux, double *uy, double *uz, const double *u,
              hipStream_t s) {
  const Op &op = h->op;
  int st = check_tri(op, "nm_opgrad");
  if (st) return st;
  return launch_opgrad<false>(op, u, ux, uy, uz, nullptr, nullptr, nullptr, s);
}

// pnpn_vel_res_cpu_compute: sets the op's coefficients to h1 = mu,
// h2 = rho bd / dt (if they differ), then res_c = -A u_c - opgrad_c(p) + f_c
// for c = x, y, z with the local operator (gs and boundary conditions are
// the caller's, as in the reference).
int nm_vel_res(nm_op_t *h, double mu, double rho, double bd, double dt, const double *u,
               const double *v, const double *w, const double *p, const double *fx,
               const double *fy, const double *fz, double *ures, double *vres, double *wres,
               hipStream_t s) {
  Op &op = h->op;
  int st = check_tri(op, "nm_vel_res");
  if (st) return st;
  const double h2 = rho * bd / dt;
  if (!(op.h1c == mu && op.ifh2 && op.h2c == h2) &&
      (st = nm_op_set_coeffs(h, mu, h2, s)))
    return st;
  if ((st = op.apply<double>(ures, u, nullptr, s, false, false))) return st;
  if ((st = op.apply<double>(vres, v, nullptr, s, false, false))) return st;
  if ((st = op.apply<double>(wres, w, nullptr, s, false, false))) return st;
  return launch_opgrad<true>(op, p, ures, vres, wres, fx, fy, fz, s);
}

}  // extern "C"
