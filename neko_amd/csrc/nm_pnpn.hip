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

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <unordered_map>
#include <vector>

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

// ---------------------------------------------------------------------------
// Pressure residual (pnpn_prs_res_cpu_compute, pnpn_res_cpu.f90:31-180),
// without the facet-normal surface terms (bc_prs_surface / bc_sym_surface),
// which vanish when the velocity is zero on the marked facets (walls) and
// there are no symmetry facets; curl (opr_cpu_curl, opr_cpu.f90:126-180),
// dudxyz (cpu_dudxyz.f90:97-159) and cdtp (cpu_cdtp.f90:191-256) on
// trilinear operators with the geometry evaluated in the kernels.
// ---------------------------------------------------------------------------
namespace nm {
namespace {

// Geometry of one element for a thread (i,j): the dX/dr and dX/ds tables in
// LDS and dX/dt in registers, as in the trilinear Ax.
template <int LX>
struct TriGeo {
  double (*sa)[LX][3];
  double (*sb)[LX][3];
  double ct[3];
  __device__ void cof(int k, int i, int j, double r[3], double s[3], double t[3],
                      double &jac) const {
    const double a0 = sa[k][j][0], a1 = sa[k][j][1], a2 = sa[k][j][2];
    const double b0 = sb[k][i][0], b1 = sb[k][i][1], b2 = sb[k][i][2];
    r[0] = b1 * ct[2] - ct[1] * b2; r[1] = ct[0] * b2 - b0 * ct[2]; r[2] = b0 * ct[1] - ct[0] * b1;
    s[0] = ct[1] * a2 - a1 * ct[2]; s[1] = a0 * ct[2] - ct[0] * a2; s[2] = ct[0] * a1 - a0 * ct[1];
    t[0] = a1 * b2 - b1 * a2; t[1] = b0 * a2 - a0 * b2; t[2] = a0 * b1 - b0 * a1;
    jac = ct[0] * t[0] + ct[1] * t[1] + ct[2] * t[2];
  }
};

template <int LX>
__device__ void tri_tables(const double *__restrict__ zg, const double *se, double (*sa)[LX][3], double (*sb)[LX][3], int i, int j,
                           double ct[3]) {
  const double p1 = (1.0 + zg[j]) * 0.5, p0 = 1.0 - p1;
  const double t1 = (1.0 + zg[i]) * 0.5, t0 = 1.0 - t1;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    sa[i][j][c] = (se[c * 4 + 0] * p0 + se[c * 4 + 1] * p1) * t0 +
                  (se[c * 4 + 2] * p0 + se[c * 4 + 3] * p1) * t1;
    sb[i][j][c] = (se[12 + c * 4 + 0] * p0 + se[12 + c * 4 + 1] * p1) * t0 +
                  (se[12 + c * 4 + 2] * p0 + se[12 + c * 4 + 3] * p1) * t1;
  }
  const double ri1 = (1.0 + zg[i]) * 0.5, ri0 = 1.0 - ri1;
  const double sj1 = (1.0 + zg[j]) * 0.5, sj0 = 1.0 - sj1;
#pragma unroll
  for (int c = 0; c < 3; c++)
    ct[c] = (se[24 + c * 4 + 0] * ri0 + se[24 + c * 4 + 1] * ri1) * sj0 +
            (se[24 + c * 4 + 2] * ri0 + se[24 + c * 4 + 3] * ri1) * sj1;
}

// w = B curl(u) per point (local): the curl of opr_cpu_curl before its gs
// (w1 = dw/dy - dv/dz, w2 = du/dz - dw/dx, w3 = dv/dx - du/dy, each
// derivative ((f_r dr + f_s ds) + f_t dt) / jac)
template <int LX>
__global__ void __launch_bounds__(LX *LX)
    curl_tri_kernel(const double *__restrict__ u1, const double *__restrict__ u2,
                    const double *__restrict__ u3, const double *__restrict__ dx,
                    const double *__restrict__ zg, const double *__restrict__ wg,
                    const double *__restrict__ edges, int64_t nel, double *__restrict__ w1,
                    double *__restrict__ w2, double *__restrict__ w3o) {
  constexpr int P = LX * LX;
  __shared__ double su[3][LX * P];
  __shared__ double sa[LX][LX][3], sb[LX][LX][3], se[36];
  const int64_t e = blockIdx.x;
  if (e >= nel) return;
  const int t = threadIdx.x, i = t % LX, j = t / LX;
  const int64_t base = e * (int64_t)(LX * P) + t;
  const double *uf[3] = {u1, u2, u3};
  double uc[3][LX];
#pragma unroll
  for (int f = 0; f < 3; f++)
#pragma unroll
    for (int k = 0; k < LX; k++) {
      uc[f][k] = uf[f][base + k * P];
      su[f][k * P + t] = uc[f][k];
    }
  for (int x = t; x < 36; x += P) se[x] = edges[e * 36 + x];
  __syncthreads();
  TriGeo<LX> G{sa, sb, {0, 0, 0}};
  tri_tables<LX>(zg, se, sa, sb, i, j, G.ct);
  const double wij = wg[i] * wg[j];
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < LX; k++) {
    double d[3][3];  // d[field][r,s,t]
#pragma unroll
    for (int f = 0; f < 3; f++) {
      const double *pl = &su[f][k * P];
      double fr = 0.0, fs = 0.0, ft = 0.0;
#pragma unroll
      for (int l = 0; l < LX; l++) {
        fr += dx[i + LX * l] * pl[l + LX * j];
        fs += dx[j + LX * l] * pl[i + LX * l];
        ft += dx[k + LX * l] * uc[f][l];
      }
      d[f][0] = fr; d[f][1] = fs; d[f][2] = ft;
    }
    double r[3], s[3], tv[3], jac;
    G.cof(k, i, j, r, s, tv, jac);
    const double jinv = 1.0 / jac;
    const double B = jac * (wij * wg[k]);
    // dudxyz(f, c): ((f_r r_c + f_s s_c) + f_t t_c) / jac
    auto dd = [&](int f, int c) { return ((d[f][0] * r[c] + d[f][1] * s[c]) + d[f][2] * tv[c]) * jinv; };
    const int64_t q = base + k * P;
    w1[q] = (dd(2, 1) - dd(1, 2)) * B;
    w2[q] = (dd(0, 2) - dd(2, 0)) * B;
    w3o[q] = (dd(1, 0) - dd(0, 1)) * B;
  }
}

// out = -out + sum_c cdtp(t_c, r_c, s_c, t_c): with X_r = w3 (r . tv),
// X_s = w3 (s . tv), X_t = w3 (t . tv) per point, out(i,j,k) gains
// sum_l D(l,i) X_r(l,j,k) + D(l,j) X_s(i,l,k) + D(l,k) X_t(i,j,l)
template <int LX>
__global__ void __launch_bounds__(LX *LX)
    cdtp_sum_tri_kernel(const double *__restrict__ t1, const double *__restrict__ t2,
                        const double *__restrict__ t3, const double *__restrict__ dx,
                        const double *__restrict__ zg, const double *__restrict__ wg,
                        const double *__restrict__ edges, int64_t nel, double *__restrict__ out) {
  constexpr int P = LX * LX;
  __shared__ double sr[LX * P], ss[LX * P];
  __shared__ double sa[LX][LX][3], sb[LX][LX][3], se[36];
  const int64_t e = blockIdx.x;
  if (e >= nel) return;
  const int t = threadIdx.x, i = t % LX, j = t / LX;
  const int64_t base = e * (int64_t)(LX * P) + t;
  for (int x = t; x < 36; x += P) se[x] = edges[e * 36 + x];
  __syncthreads();
  TriGeo<LX> G{sa, sb, {0, 0, 0}};
  tri_tables<LX>(zg, se, sa, sb, i, j, G.ct);
  const double wij = wg[i] * wg[j];
  __syncthreads();
  double xt[LX];
#pragma unroll
  for (int k = 0; k < LX; k++) {
    double r[3], s[3], tv[3], jac;
    G.cof(k, i, j, r, s, tv, jac);
    const int64_t q = base + k * P;
    const double a0 = t1[q], a1 = t2[q], a2 = t3[q];
    const double w3 = wij * wg[k];
    sr[k * P + t] = w3 * (a0 * r[0] + a1 * r[1] + a2 * r[2]);
    ss[k * P + t] = w3 * (a0 * s[0] + a1 * s[1] + a2 * s[2]);
    xt[k] = w3 * (a0 * tv[0] + a1 * tv[1] + a2 * tv[2]);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < LX; k++) {
    double acc = 0.0;
#pragma unroll
    for (int l = 0; l < LX; l++) {
      acc += dx[l + LX * i] * sr[k * P + l + LX * j];
      acc += dx[l + LX * j] * ss[k * P + i + LX * l];
      acc += dx[l + LX * k] * xt[l];
    }
    const int64_t q = base + k * P;
    out[q] = (-out[q]) + acc;
  }
}

// y_c *= binv (c = 1..3)
__global__ void scale3_kernel(double *__restrict__ a, double *__restrict__ b,
                              double *__restrict__ c, const double *__restrict__ binv,
                              int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double v = binv[i];
    a[i] = a[i] * v;
    b[i] = b[i] * v;
    c[i] = c[i] * v;
  }
}

// wa <- wa binv (the curl's Binv), then ta_c = f_c / rho - ((wa_c mu/rho) B)
__global__ void prs_src_kernel(double *__restrict__ ta1, double *__restrict__ ta2,
                               double *__restrict__ ta3, const double *__restrict__ wa1,
                               const double *__restrict__ wa2, const double *__restrict__ wa3,
                               const double *__restrict__ fx, const double *__restrict__ fy,
                               const double *__restrict__ fz, const double *__restrict__ binv,
                               const double *__restrict__ B, double rho, double mu, int64_t n) {
  const double mr = mu / rho;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double bi = binv[i], b = B[i];
    ta1[i] = fx[i] / rho - ((wa1[i] * bi) * mr) * b;
    ta2[i] = fy[i] / rho - ((wa2[i] * bi) * mr) * b;
    ta3[i] = fz[i] / rho - ((wa3[i] * bi) * mr) * b;
  }
}

int ensure_binv(Op *op, hipStream_t s) {
  if (op->binv) return NM_OK;
  const int64_t n = op->n;
  if (!(op->binv = op->alloc<double>(n))) {
    set_error("pressure residual: out of device memory");
    return NM_ERR_NOMEM;
  }
  int st;
  if ((st = launch_vec<double>(V_COPY, op->binv, op->B, nullptr, nullptr, 0.0, n, s))) return st;
  if ((st = op->gs<double>(op->binv, nullptr, s))) return st;
  return launch_vec<double>(V_INVCOL1, op->binv, nullptr, nullptr, nullptr, 0.0, n, s);
}

int launch_curl(const Op &op, const double *u1, const double *u2, const double *u3,
                double *w1, double *w2, double *w3, hipStream_t s) {
  const OpFields<double> &F = op.f64;
  const dim3 g((unsigned)op.nel);
#define NM_CURL(L)                                                                        \
  case L:                                                                                 \
    hipLaunchKernelGGL(curl_tri_kernel<L>, g, dim3(L * L), 0, s, u1, u2, u3, F.dx, F.zg,  \
                       F.wg, F.edges, op.nel, w1, w2, w3);                                \
    break;
  switch (op.lx) {
    NM_CURL(2) NM_CURL(3) NM_CURL(4) NM_CURL(5) NM_CURL(6) NM_CURL(7) NM_CURL(8) NM_CURL(9)
    NM_CURL(10)
    default:
      set_error("curl: lx=%d outside 2..10", op.lx);
      return NM_ERR_UNSUPPORTED;
  }
#undef NM_CURL
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

int launch_cdtp_sum(const Op &op, const double *t1, const double *t2, const double *t3,
                    double *out, hipStream_t s) {
  const OpFields<double> &F = op.f64;
  const dim3 g((unsigned)op.nel);
#define NM_CDTP(L)                                                                        \
  case L:                                                                                 \
    hipLaunchKernelGGL(cdtp_sum_tri_kernel<L>, g, dim3(L * L), 0, s, t1, t2, t3, F.dx,    \
                       F.zg, F.wg, F.edges, op.nel, out);                                 \
    break;
  switch (op.lx) {
    NM_CDTP(2) NM_CDTP(3) NM_CDTP(4) NM_CDTP(5) NM_CDTP(6) NM_CDTP(7) NM_CDTP(8) NM_CDTP(9)
    NM_CDTP(10)
    default:
      set_error("cdtp: lx=%d outside 2..10", op.lx);
      return NM_ERR_UNSUPPORTED;
  }
#undef NM_CDTP
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// curl including its gs and Binv (opr_cpu_curl): w = Binv gs(B curl u)
int curl_full(Op *op, const double *u1, const double *u2, const double *u3, double *w1,
              double *w2, double *w3, hipStream_t s) {
  int st;
  if ((st = launch_curl(*op, u1, u2, u3, w1, w2, w3, s))) return st;
  for (double *w : {w1, w2, w3})
    if ((st = op->gs<double>(w, nullptr, s))) return st;
  hipLaunchKernelGGL(scale3_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, w1, w2, w3,
                     (const double *)op->binv, op->n);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

}  // namespace
}  // namespace nm

// p_res = -A p + sum_c cdtp(ta_c) with h1 = 1/rho, h2 = 0, where ta =
// Binv gs(f / rho - (mu/rho) B curl(curl(u_e))) (pnpn_prs_res_cpu_compute);
// with surfaces also the facet-normal terms.  Local like the reference (gs
// of p_res is the solver's).
static int prs_res_impl(nm_op_t *h, double mu, double rho, const double *p, const double *ue,
                        const double *ve, const double *we, const double *fx,
                        const double *fy, const double *fz, double *p_res, hipStream_t s,
                        const char *who, double dtbd = 0.0, const double *u = nullptr,
                        const double *v = nullptr, const double *w = nullptr,
                        const nm_surface_t *prs = nullptr, const nm_surface_t *sym = nullptr);

extern "C" {

// curl (src/math/operators.f90 curl; opr_cpu_curl): w = Binv gs(B curl u)
int nm_curl(nm_op_t *h, double *w1, double *w2, double *w3, const double *u1,
            const double *u2, const double *u3, hipStream_t s) {
  Op *op = &h->op;
  int st = check_tri(*op, "nm_curl");
  if (st || (st = ensure_binv(op, s))) return st;
  return curl_full(op, u1, u2, u3, w1, w2, w3, s);
}

int nm_prs_res(nm_op_t *h, double mu, double rho, const double *p, const double *ue,
               const double *ve, const double *we, const double *fx, const double *fy,
               const double *fz, double *p_res, hipStream_t s) {
  return prs_res_impl(h, mu, rho, p, ue, ve, we, fx, fy, fz, p_res, s, "nm_prs_res");
}

int nm_prs_res_surf(nm_op_t *h, double mu, double rho, double bd, double dt, const double *p,
                    const double *ue, const double *ve, const double *we, const double *u,
                    const double *v, const double *w, const double *fx, const double *fy,
                    const double *fz, const nm_surface_t *prs, const nm_surface_t *sym,
                    double *p_res, hipStream_t s) {
  if (!(dt != 0.0)) {
    set_error("nm_prs_res_surf: dt must be nonzero");
    return NM_ERR_ARG;
  }
  if (prs && (!u || !v || !w)) {
    set_error("nm_prs_res_surf: the velocity-prescribed surface needs u, v, w");
    return NM_ERR_ARG;
  }
  return prs_res_impl(h, mu, rho, p, ue, ve, we, fx, fy, fz, p_res, s, "nm_prs_res_surf",
                      bd / dt, u, v, w, prs, sym);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Facet-normal surfaces (facet_normal_t, src/bc/facet_normal.f90:233-326)
// ---------------------------------------------------------------------------
struct nm_surface_s {
  nm::Op *op = nullptr;
  int64_t np = 0;
  int32_t *pts = nullptr;      // unique local points (0-based), first-seen order
  double *nrm[3] = {nullptr, nullptr, nullptr};  // area-weighted normals
  std::vector<int32_t> h_pts;
  std::vector<double> h_nrm[3];
  ~nm_surface_s() {
    if (pts) (void)hipFree(pts);
    for (double *q : nrm)
      if (q) (void)hipFree(q);
  }
};

namespace nm {
namespace {

// p_res[k] -= c (a0 n0 + a1 n1 + a2 n2)[k] over a surface's points, the sum
// in Neko's order ((a0 n0 + a1 n1) + a2 n2); c == 1 skips the product
__global__ void surf_sub_kernel(double *__restrict__ p_res, const int32_t *__restrict__ pts,
                                const double *__restrict__ n0, const double *__restrict__ n1,
                                const double *__restrict__ n2, const double *__restrict__ a0,
                                const double *__restrict__ a1, const double *__restrict__ a2,
                                double c, int scale, int64_t np) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int32_t k = pts[q];
  const double t = a0[k] * n0[q] + a1[k] * n1[q] + a2[k] * n2[q];
  p_res[k] = p_res[k] - (scale ? c * t : t);
}

// coordinates of the listed elements, packed: out[(q * 3 + c) * n3 + l]
__global__ void gather_elems_kernel(double *__restrict__ out, const double *__restrict__ x,
                                    const double *__restrict__ y, const double *__restrict__ z,
                                    const int32_t *__restrict__ el, int64_t ne, int64_t n3) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * n3) return;
  const int64_t q = t / n3, l = t - q * n3;
  const int64_t src = (int64_t)el[q] * n3 + l;
  out[(q * 3 + 0) * n3 + l] = x[src];
  out[(q * 3 + 1) * n3 + l] = y[src];
  out[(q * 3 + 2) * n3 + l] = z[src];
}

int surf_sub(double *p_res, const nm_surface_t *sf, const double *a0, const double *a1,
             const double *a2, double c, bool scale, hipStream_t s) {
  if (sf->np <= 0) return NM_OK;
  hipLaunchKernelGGL(surf_sub_kernel, dim3((unsigned)((sf->np + 255) / 256)), dim3(256), 0, s,
                     p_res, sf->pts, sf->nrm[0], sf->nrm[1], sf->nrm[2], a0, a1, a2, c,
                     scale ? 1 : 0, sf->np);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

}  // namespace
}  // namespace nm

namespace nm {
namespace {

// One point of a facet: local index, unit normal, area weight and x, y.
struct FacetPoint {
  int64_t loc;
  double nx, ny, nz, area, x, y;
};

// The points of the given facets with the normal and area of each as
// coef_generate_area_and_normal (src/sem/coef.f90:1195-1370) forms them:
// tangent derivatives (D along the face's two directions), their cross
// product with sign -,+ / +,- / -,+ on facets 1,2 / 3,4 / 5,6, area =
// |cross| w_a w_b, unit normal above machine epsilon.  Facets in list
// order, a facet's points with the first index fastest (the order of
// Neko's i, j, k loops over the facet).
int facet_points(Op *op, const int32_t *facets, int64_t nfacets, const char *who,
                 std::vector<FacetPoint> &out, hipStream_t s) {
  if (!op->xyz[0]) {
    set_error("%s: the operator carries no coordinates (box or mesh operators do)", who);
    return NM_ERR_STATE;
  }
  const int lx = op->lx;
  const int64_t n3 = (int64_t)lx * lx * lx;
  std::vector<double> D(lx * lx), z(lx), wq(lx), dd(lx * lx), ddt(lx * lx);
  NM_HIP_TRY(hipMemcpyAsync(D.data(), op->f64.dx, sizeof(double) * lx * lx,
                            hipMemcpyDeviceToHost, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  if (nm_gll(lx, z.data(), wq.data(), dd.data(), ddt.data())) {
    set_error("%s: no GLL rule for lx = %d", who, lx);
    return NM_ERR_ARG;
  }
  auto Dm = [&](int i, int m) { return D[i + lx * m]; };
  // the facets' elements, their coordinates gathered on the device and
  // copied back in one transfer
  std::unordered_map<int64_t, int64_t> epos;
  std::vector<int32_t> elist;
  for (int64_t q = 0; q < nfacets; q++) {
    const int64_t e = facets[2 * q];
    const int f = facets[2 * q + 1];
    if (e < 0 || e >= op->nel || f < 1 || f > 6) {
      set_error("%s: facet %lld is (%lld, %d); elements 0..%lld, facets 1..6", who,
                (long long)q, (long long)e, f, (long long)op->nel - 1);
      return NM_ERR_ARG;
    }
    if (epos.emplace(e, (int64_t)elist.size()).second) elist.push_back((int32_t)e);
  }
  const int64_t ne = (int64_t)elist.size();
  std::vector<double> XE((size_t)std::max<int64_t>(ne, 1) * 3 * n3);
  if (ne > 0) {
    int32_t *d_el = nullptr;
    double *d_x = nullptr;
    if (hipMalloc(&d_el, sizeof(int32_t) * ne) != hipSuccess ||
        hipMalloc(&d_x, sizeof(double) * ne * 3 * n3) != hipSuccess) {
      if (d_el) (void)hipFree(d_el);
      set_error("%s: out of device memory", who);
      return NM_ERR_NOMEM;
    }
    hipError_t e1 = hipMemcpyAsync(d_el, elist.data(), sizeof(int32_t) * ne,
                                   hipMemcpyHostToDevice, s);
    if (e1 == hipSuccess) {
      hipLaunchKernelGGL(gather_elems_kernel, dim3((unsigned)((ne * n3 + 255) / 256)), dim3(256),
                         0, s, d_x, (const double *)op->xyz[0], (const double *)op->xyz[1],
                         (const double *)op->xyz[2], (const int32_t *)d_el, ne, n3);
      e1 = hipGetLastError();
    }
    if (e1 == hipSuccess)
      e1 = hipMemcpyAsync(XE.data(), d_x, sizeof(double) * ne * 3 * n3, hipMemcpyDeviceToHost, s);
    if (e1 == hipSuccess) e1 = hipStreamSynchronize(s);
    (void)hipFree(d_el);
    (void)hipFree(d_x);
    NM_HIP_TRY(e1);
  }
  const double eps = 2.220446049250313e-16;
  out.clear();
  out.reserve((size_t)nfacets * lx * lx);
  for (int64_t q = 0; q < nfacets; q++) {
    const int64_t e = facets[2 * q];
    const int f = facets[2 * q + 1];
    const double *xe = XE.data() + epos[e] * 3 * n3;
    auto at = [&](int c, int i, int j, int k) { return xe[c * n3 + i + lx * (j + lx * k)]; };
    const int fixed = (f % 2 == 1) ? 0 : lx - 1;
    const int dir = (f - 1) / 2;  // 0: i fixed, 1: j fixed, 2: k fixed
    const double sgn = (f == 1 || f == 4 || f == 5) ? -1.0 : 1.0;
    for (int b = 0; b < lx; b++)
      for (int a = 0; a < lx; a++) {
        // face point: (a, b) = (j, k) / (i, k) / (i, j)
        const int i = dir == 0 ? fixed : a;
        const int j = dir == 0 ? a : (dir == 1 ? fixed : b);
        const int k = dir == 2 ? fixed : b;
        double t1[3], t2[3];
        for (int c = 0; c < 3; c++) {
          double dr = 0.0, ds = 0.0, dt = 0.0;
          for (int m = 0; m < lx; m++) {
            dr += Dm(i, m) * at(c, m, j, k);
            ds += Dm(j, m) * at(c, i, m, k);
            dt += Dm(k, m) * at(c, i, j, m);
          }
          t1[c] = dir == 0 ? ds : dr;
          t2[c] = dir == 2 ? ds : dt;
        }
        const double A = t1[1] * t2[2] - t1[2] * t2[1];
        const double B = t1[2] * t2[0] - t1[0] * t2[2];
        const double C = t1[0] * t2[1] - t1[1] * t2[0];
        FacetPoint fp;
        fp.area = std::sqrt(A * A + B * B + C * C) * (wq[a] * wq[b]);
        fp.nx = sgn * A;
        fp.ny = sgn * B;
        fp.nz = sgn * C;
        const double len = std::sqrt(fp.nx * fp.nx + fp.ny * fp.ny + fp.nz * fp.nz);
        if (len > eps) {
          fp.nx = fp.nx / len;
          fp.ny = fp.ny / len;
          fp.nz = fp.nz / len;
        }
        fp.loc = e * n3 + i + lx * (j + (int64_t)lx * k);
        fp.x = at(0, i, j, k);
        fp.y = at(1, i, j, k);
        out.push_back(fp);
      }
  }
  return NM_OK;
}

}  // namespace
}  // namespace nm

extern "C" {

// Area-weighted normals of the given facets' points (facet_points), summed
// per unique local point over the facets in list order
// (facet_normal_finalize).
int nm_surface_create(nm_surface_t **out, nm_op_t *h, const int32_t *facets, int64_t nfacets,
                      hipStream_t s) {
  using namespace nm;
  *out = nullptr;
  Op *op = &h->op;
  std::vector<FacetPoint> fps;
  int st = facet_points(op, facets, nfacets, "nm_surface_create", fps, s);
  if (st) return st;
  // owned until the end: every early return below frees it and its buffers
  std::unique_ptr<nm_surface_t> sf(new nm_surface_t);
  sf->op = op;
  std::unordered_map<int64_t, int64_t> slot;
  for (const FacetPoint &fp : fps) {
    auto it = slot.find(fp.loc);
    int64_t sl;
    if (it == slot.end()) {
      sl = (int64_t)sf->h_pts.size();
      slot.emplace(fp.loc, sl);
      sf->h_pts.push_back((int32_t)fp.loc);
      for (int c = 0; c < 3; c++) sf->h_nrm[c].push_back(0.0);
    } else {
      sl = it->second;
    }
    sf->h_nrm[0][sl] = sf->h_nrm[0][sl] + fp.nx * fp.area;
    sf->h_nrm[1][sl] = sf->h_nrm[1][sl] + fp.ny * fp.area;
    sf->h_nrm[2][sl] = sf->h_nrm[2][sl] + fp.nz * fp.area;
  }
  sf->np = (int64_t)sf->h_pts.size();
  const size_t np1 = (size_t)std::max<int64_t>(sf->np, 1);
  if (hipMalloc(&sf->pts, sizeof(int32_t) * np1) != hipSuccess) {
    set_error("nm_surface_create: out of device memory");
    return NM_ERR_NOMEM;
  }
  for (int c = 0; c < 3; c++)
    if (hipMalloc(&sf->nrm[c], sizeof(double) * np1) != hipSuccess) {
      set_error("nm_surface_create: out of device memory");
      return NM_ERR_NOMEM;
    }
  if (sf->np > 0) {
    NM_HIP_TRY(hipMemcpyAsync(sf->pts, sf->h_pts.data(), sizeof(int32_t) * sf->np,
                              hipMemcpyHostToDevice, s));
    for (int c = 0; c < 3; c++)
      NM_HIP_TRY(hipMemcpyAsync(sf->nrm[c], sf->h_nrm[c].data(), sizeof(double) * sf->np,
                                hipMemcpyHostToDevice, s));
  }
  NM_HIP_TRY(hipStreamSynchronize(s));
  *out = sf.release();
  return NM_OK;
}

void nm_surface_free(nm_surface_t *sf) { delete sf; }

int nm_surface_info(const nm_surface_t *sf, int64_t *npoints) {
  *npoints = sf->np;
  return NM_OK;
}

// host copies: points (0-based local), area-weighted normals (npoints each)
int nm_surface_normals(const nm_surface_t *sf, int32_t *pts, double *nx, double *ny,
                       double *nz) {
  for (int64_t q = 0; q < sf->np; q++) {
    pts[q] = sf->h_pts[q];
    nx[q] = sf->h_nrm[0][q];
    ny[q] = sf->h_nrm[1][q];
    nz[q] = sf->h_nrm[2][q];
  }
  return NM_OK;
}

}  // extern "C"

static int prs_res_impl(nm_op_t *h, double mu, double rho, const double *p, const double *ue,
                        const double *ve, const double *we, const double *fx,
                        const double *fy, const double *fz, double *p_res, hipStream_t s,
                        const char *who, double dtbd, const double *u, const double *v,
                        const double *w, const nm_surface_t *prs, const nm_surface_t *sym) {
  using namespace nm;
  Op *op = &h->op;
  int st = check_tri(*op, who);
  if (st || (st = ensure_binv(op, s))) return st;
  if ((prs && prs->op != op) || (sym && sym->op != op)) {
    set_error("%s: a surface belongs to another operator", who);
    return NM_ERR_ARG;
  }
  const int64_t n = op->n;
  if (!op->pnpn_work && !(op->pnpn_work = op->alloc<double>(6 * n))) {
    set_error("%s: out of device memory", who);
    return NM_ERR_NOMEM;
  }
  double *ta[3] = {op->pnpn_work, op->pnpn_work + n, op->pnpn_work + 2 * n};
  double *wa[3] = {op->pnpn_work + 3 * n, op->pnpn_work + 4 * n, op->pnpn_work + 5 * n};
  const double h1 = 1.0 / rho;
  if (!(op->h1c == h1 && !op->ifh2) && (st = nm_op_set_coeffs(h, h1, 0.0, s))) return st;
  if ((st = curl_full(op, ue, ve, we, ta[0], ta[1], ta[2], s))) return st;
  // second curl up to its gs; its Binv is applied inside prs_src_kernel
  if ((st = launch_curl(*op, ta[0], ta[1], ta[2], wa[0], wa[1], wa[2], s))) return st;
  for (double *x : wa)
    if ((st = op->gs<double>(x, nullptr, s))) return st;
  hipLaunchKernelGGL(prs_src_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, ta[0], ta[1], ta[2],
                     (const double *)wa[0], (const double *)wa[1], (const double *)wa[2], fx,
                     fy, fz, (const double *)op->binv, (const double *)op->B, rho, mu, n);
  NM_HIP_TRY(hipGetLastError());
  for (double *t : ta)
    if ((st = op->gs<double>(t, nullptr, s))) return st;
  hipLaunchKernelGGL(scale3_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, ta[0], ta[1], ta[2],
                     (const double *)op->binv, n);
  NM_HIP_TRY(hipGetLastError());
  if ((st = op->apply<double>(p_res, p, nullptr, s, false, false))) return st;
  if ((st = launch_cdtp_sum(*op, ta[0], ta[1], ta[2], p_res, s))) return st;
  // surface terms (pnpn_res_cpu.f90:134-174): p_res - (bd/dt)(u . nA) on the
  // velocity-prescribed surface, then - (ta . nA) on the symmetry surface
  if (prs && (st = surf_sub(p_res, prs, u, v, w, dtbd, true, s))) return st;
  if (sym && (st = surf_sub(p_res, sym, ta[0], ta[1], ta[2], 1.0, false, s))) return st;
  return NM_OK;
}

// ---------------------------------------------------------------------------
// Cyclic (rotationally periodic about z) velocity components: rotate_cyc
// (src/math/operators.f90:1078-1140, opr_cpu_rotate_cyc_r1,
// src/math/bcknd/cpu/opr_cpu.f90:333-357) with the point list and rotations
// of coef_cyclic (src/sem/coef.f90:1380-1430).
// ---------------------------------------------------------------------------
namespace nm {
namespace {

// idir 1: (vx, vy) -> (normal, tangential) components; 0: back.  Each
// thread takes one local point and applies its entries in list order (a
// point on two cyclic facets of one element is listed twice and rotated
// twice, as the reference's sequential loop does).
__global__ void rotate_cyc_chain_kernel(double *__restrict__ vx, double *__restrict__ vy,
                                        const int32_t *__restrict__ upts,
                                        const int32_t *__restrict__ uptr,
                                        const double *__restrict__ r11,
                                        const double *__restrict__ r12, int64_t nu, int idir) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nu) return;
  const int64_t j = upts[u];
  double x = vx[j], y = vy[j];
  for (int32_t q = uptr[u]; q < uptr[u + 1]; q++) {
    const double a = r11[q], b = r12[q];
    double vn, vt;
    if (idir == 1) {
      vn = x * a + y * b;
      vt = -x * b + y * a;
    } else {
      vn = x * a - y * b;
      vt = x * b + y * a;
    }
    x = vn;
    y = vt;
  }
  vx[j] = x;
  vy[j] = y;
}

// Neko's device layout (hip_rotate_cyc): cyc_msk[0] = ncyc + 1, then the
// 1-based points; one thread per entry
__global__ void rotate_cyc_entry_kernel(double *__restrict__ vx, double *__restrict__ vy,
                                        const int *__restrict__ cyc_msk,
                                        const double *__restrict__ r11,
                                        const double *__restrict__ r12, int ncyc, int idir) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncyc; i += stride) {
    const int64_t j = cyc_msk[i + 1] - 1;
    const double x = vx[j], y = vy[j], a = r11[i], b = r12[i];
    if (idir == 1) {
      vx[j] = x * a + y * b;
      vy[j] = -x * b + y * a;
    } else {
      vx[j] = x * a - y * b;
      vy[j] = x * b + y * a;
    }
  }
}

}  // namespace
}  // namespace nm

struct nm_cyclic_s {
  int64_t ncyc = 0, nu = 0;
  std::vector<int32_t> h_msk;           // 0-based points, entry order
  std::vector<double> h_r11, h_r12;
  int32_t *upts = nullptr, *uptr = nullptr;
  double *r11 = nullptr, *r12 = nullptr;  // entries grouped by point
  ~nm_cyclic_s() {
    for (void *q : {(void *)upts, (void *)uptr, (void *)r11, (void *)r12})
      if (q) (void)hipFree(q);
  }
};

extern "C" {

// coef_cyclic: for each cyclic facet (element 0-based, facet 1..6, in list
// order, as msh%periodic lists them) and each of its points, the unit
// normal's x-y direction n / |n| signed by d = y n_x - x n_y gives R11,
// R12; a normal without an x-y part is an error (rotations about z only).
int nm_cyclic_create(nm_cyclic_t **out, nm_op_t *h, const int32_t *facets, int64_t nfacets,
                     hipStream_t s) {
  using namespace nm;
  *out = nullptr;
  std::vector<FacetPoint> fps;
  int st = facet_points(&h->op, facets, nfacets, "nm_cyclic_create", fps, s);
  if (st) return st;
  std::unique_ptr<nm_cyclic_t> cy(new nm_cyclic_t);
  const double eps = 2.220446049250313e-16;
  for (const FacetPoint &fp : fps) {
    const double len = std::sqrt(fp.nx * fp.nx + fp.ny * fp.ny);
    if (!(len > eps)) {
      set_error("nm_cyclic_create: a cyclic facet point's normal has no x-y part "
                "(cyclic rotations are about z)");
      return NM_ERR_ARG;
    }
    const double d = fp.y * fp.nx - fp.x * fp.ny;
    const double sg = std::copysign(1.0, d);
    cy->h_msk.push_back((int32_t)fp.loc);
    cy->h_r11.push_back(fp.nx / len * sg);
    cy->h_r12.push_back(fp.ny / len * sg);
  }
  cy->ncyc = (int64_t)cy->h_msk.size();
  // group the entries by point, list order kept within a point
  std::vector<int32_t> order(cy->ncyc);
  for (int64_t q = 0; q < cy->ncyc; q++) order[q] = (int32_t)q;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return cy->h_msk[a] < cy->h_msk[b]; });
  std::vector<int32_t> upts, uptr;
  std::vector<double> r11(cy->ncyc), r12(cy->ncyc);
  for (int64_t q = 0; q < cy->ncyc; q++) {
    const int32_t e = order[q];
    if (q == 0 || cy->h_msk[e] != cy->h_msk[order[q - 1]]) {
      upts.push_back(cy->h_msk[e]);
      uptr.push_back((int32_t)q);
    }
    r11[q] = cy->h_r11[e];
    r12[q] = cy->h_r12[e];
  }
  uptr.push_back((int32_t)cy->ncyc);
  cy->nu = (int64_t)upts.size();
  const size_t n1 = (size_t)std::max<int64_t>(cy->ncyc, 1), u1 = (size_t)std::max<int64_t>(cy->nu, 1);
  if (hipMalloc(&cy->upts, sizeof(int32_t) * u1) != hipSuccess ||
      hipMalloc(&cy->uptr, sizeof(int32_t) * (u1 + 1)) != hipSuccess ||
      hipMalloc(&cy->r11, sizeof(double) * n1) != hipSuccess ||
      hipMalloc(&cy->r12, sizeof(double) * n1) != hipSuccess) {
    set_error("nm_cyclic_create: out of device memory");
    return NM_ERR_NOMEM;
  }
  if (cy->ncyc > 0) {
    NM_HIP_TRY(hipMemcpyAsync(cy->upts, upts.data(), sizeof(int32_t) * cy->nu,
                              hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipMemcpyAsync(cy->uptr, uptr.data(), sizeof(int32_t) * (cy->nu + 1),
                              hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipMemcpyAsync(cy->r11, r11.data(), sizeof(double) * cy->ncyc,
                              hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipMemcpyAsync(cy->r12, r12.data(), sizeof(double) * cy->ncyc,
                              hipMemcpyHostToDevice, s));
  }
  NM_HIP_TRY(hipStreamSynchronize(s));
  *out = cy.release();
  return NM_OK;
}

void nm_cyclic_free(nm_cyclic_t *cy) { delete cy; }

// host copies: entry count, and (when msk is not NULL) Neko's arrays --
// cyc_msk(1..ncyc) 1-based, R11, R12 -- in entry order
int nm_cyclic_info(const nm_cyclic_t *cy, int64_t *ncyc, int32_t *msk, double *r11,
                   double *r12) {
  *ncyc = cy->ncyc;
  if (msk)
    for (int64_t q = 0; q < cy->ncyc; q++) {
      msk[q] = cy->h_msk[q] + 1;
      r11[q] = cy->h_r11[q];
      r12[q] = cy->h_r12[q];
    }
  return NM_OK;
}

int nm_rotate_cyc(const nm_cyclic_t *cy, double *vx, double *vy, double *vz, int idir,
                  hipStream_t s) {
  using namespace nm;
  (void)vz;
  if (idir != 0 && idir != 1) {
    set_error("nm_rotate_cyc: idir %d (1 to normal/tangential, 0 back)", idir);
    return NM_ERR_ARG;
  }
  if (cy->nu == 0) return NM_OK;
  hipLaunchKernelGGL(rotate_cyc_chain_kernel, dim3((unsigned)((cy->nu + 255) / 256)),
                     dim3(256), 0, s, vx, vy, cy->upts, cy->uptr, cy->r11, cy->r12, cy->nu,
                     idir);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// Drop-in for Neko's HIP backend (opr_device.F90:159-168): device
// pointers by value, integers by reference, on glb_cmd_queue.  One thread
// per entry, as the reference kernel: a point listed twice must not occur.
void hip_rotate_cyc(void *vx, void *vy, void *vz, void *x, void *y, void *z, void *cyc_msk,
                    void *R11, void *R12, int *ncyc, int *idir) {
  using namespace nm;
  (void)vz;
  (void)x;
  (void)y;
  (void)z;
  if (*ncyc <= 0) return;
  const int nb = std::min(65535, (*ncyc + 255) / 256);
  hipLaunchKernelGGL(rotate_cyc_entry_kernel, dim3((unsigned)nb), dim3(256), 0,
                     (hipStream_t)glb_cmd_queue, (double *)vx, (double *)vy,
                     (const int *)cyc_msk, (const double *)R11, (const double *)R12, *ncyc,
                     *idir);
}

}  // extern "C"
