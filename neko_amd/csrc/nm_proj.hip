// Solution-space projection around a solve (Neko's projection_t,
// src/common/projection.f90, CPU path): the last L solutions X and their
// images B = A X are kept A-orthonormal in HBM; before a solve the RHS is
// projected onto span(X)'s A-complement (project_on, two rounds of
// classical Gram-Schmidt, :476-551) and the projected part of the solution
// is kept in xbar; after the solve xbar is added back, the new solution
// joins the basis (restarting at one vector once L are held) and is
// A-orthonormalised against it (project_back :317-358, cpu_proj_ortho
// :717-834: two CGS rounds, symmetric weights, dropped below tol * its
// original A-norm).
//
// Device layout: X and B are one allocation each, L columns of n doubles;
// every pass streams the columns it needs once, with one accumulator per
// column in registers (L <= 32), and reduces in the fixed order of the
// library's other reductions (RED_BLOCKS partials, one combine block).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "neko_amd.h"
#include "nm_common.h"
#include "nm_kernels.h"
#include "nm_op.h"

namespace nm {

namespace {

constexpr int PJ_LMAX = 32;

__device__ inline double pj_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sums of v[0..nv) into part[k * RED_BLOCKS + blockIdx.x].
template <int NV>
__device__ inline void pj_block_sums(const double (&v)[NV], int nv, double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64][NV];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    if (k < nv) {
      const double x = pj_wave_sum(v[k]);
      if (lane == 0) sh[wid][k] = x;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    double r = 0.0;
    for (int q = 0; q < RED_THREADS / 64; q++) r += sh[q][threadIdx.x];
    part[(int64_t)threadIdx.x * RED_BLOCKS + blockIdx.x] = r;
  }
}

// alpha_k partials: sum_i X_k w b, k < m  (first CGS round of project_on)
__global__ void __launch_bounds__(RED_THREADS)
    pj_dots_kernel(const double *__restrict__ X, int64_t n, int m,
                   const double *__restrict__ b, const double *__restrict__ w,
                   double *__restrict__ part) {
  double acc[PJ_LMAX];
#pragma unroll
  for (int k = 0; k < PJ_LMAX; k++) acc[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double wi = w[i], bi = b[i];
#pragma unroll
    for (int k = 0; k < PJ_LMAX; k++)
      if (k < m) acc[k] += X[k * n + i] * wi * bi;
  }
  pj_block_sums(acc, m, part);
}

// project_on, second pass: xbar = sum_k a_k X_k, b -= sum_k a_k B_k (in k
// order), then the second round's partials sum_i X_k w b_new
__global__ void __launch_bounds__(RED_THREADS)
    pj_on2_kernel(const double *__restrict__ X, const double *__restrict__ B, int64_t n,
                  int m, const double *__restrict__ a, double *__restrict__ b,
                  double *__restrict__ xbar, const double *__restrict__ w,
                  double *__restrict__ part) {
  double acc[PJ_LMAX];
#pragma unroll
  for (int k = 0; k < PJ_LMAX; k++) acc[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double xb = a[0] * X[i];
    double bn = b[i] - a[0] * B[i];
    for (int k = 1; k < m; k++) {
      xb = xb + a[k] * X[k * n + i];
      bn = bn - a[k] * B[k * n + i];
    }
    xbar[i] = xb;
    b[i] = bn;
    const double wi = w[i];
#pragma unroll
    for (int k = 0; k < PJ_LMAX; k++)
      if (k < m) acc[k] += X[k * n + i] * wi * bn;
  }
  pj_block_sums(acc, m, part);
}

// y += sign * sum_k c_k Y_k for (xbar, X) with sign +1 and (b, B) with -1
__global__ void __launch_bounds__(RED_THREADS)
    pj_on3_kernel(const double *__restrict__ X, const double *__restrict__ B, int64_t n,
                  int m, const double *__restrict__ c, double *__restrict__ b,
                  double *__restrict__ xbar) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double xb = xbar[i], bn = b[i];
    for (int k = 0; k < m; k++) {
      xb = xb + c[k] * X[k * n + i];
      bn = bn - c[k] * B[k * n + i];
    }
    xbar[i] = xb;
    b[i] = bn;
  }
}

// proj_ortho partials for k < nk: 0.5 (sum X_k w Bm + sum B_k w Xm), after
// first subtracting sum_{k < nsub} c_k (X_k, B_k) from (Xm, Bm) when c is
// given.  Xm, Bm = column m-1.
__global__ void __launch_bounds__(RED_THREADS)
    pj_ortho_kernel(double *__restrict__ X, double *__restrict__ B, int64_t n, int m,
                    const double *__restrict__ c, int nsub, int nk,
                    const double *__restrict__ w, double *__restrict__ part) {
  double s[PJ_LMAX], t[PJ_LMAX];
#pragma unroll
  for (int k = 0; k < PJ_LMAX; k++) s[k] = t[k] = 0.0;
  double *Xm = X + (int64_t)(m - 1) * n, *Bm = B + (int64_t)(m - 1) * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double xm = Xm[i], bm = Bm[i];
    if (c != nullptr) {
      for (int k = 0; k < nsub; k++) {
        xm = xm - c[k] * X[k * n + i];
        bm = bm - c[k] * B[k * n + i];
      }
      Xm[i] = xm;
      Bm[i] = bm;
    }
    const double wi = w[i];
#pragma unroll
    for (int k = 0; k < PJ_LMAX; k++) {
      if (k < nk) {
        const bool last = k == m - 1;
        const double xk = last ? xm : X[k * n + i], bk = last ? bm : B[k * n + i];
        s[k] += xk * wi * bm;
        t[k] += bk * wi * xm;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PJ_LMAX; k++) s[k] = 0.5 * (s[k] + t[k]);
  // (0.5 (s + t) per thread sums to 0.5 (S + T) per block up to rounding)
  pj_block_sums(s, nk, part);
}

// column m-1 subtracted by sum_{k<nsub} c_k (X_k, B_k); partial X_m w B_m
__global__ void __launch_bounds__(RED_THREADS)
    pj_ortho_last_kernel(double *__restrict__ X, double *__restrict__ B, int64_t n, int m,
                         const double *__restrict__ c, int nsub,
                         const double *__restrict__ w, double *__restrict__ part) {
  double acc[1] = {0.0};
  double *Xm = X + (int64_t)(m - 1) * n, *Bm = B + (int64_t)(m - 1) * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double xm = Xm[i], bm = Bm[i];
    for (int k = 0; k < nsub; k++) {
      xm = xm - c[k] * X[k * n + i];
      bm = bm - c[k] * B[k * n + i];
    }
    Xm[i] = xm;
    Bm[i] = bm;
    acc[0] += xm * w[i] * bm;
  }
  pj_block_sums(acc, 1, part);
}

// keep the new column iff sqrt(a_m) > tol * sqrt(alpha_m of the first
// round); sc = {1/sqrt(a_m) or 0, keep}
__global__ void pj_decide_kernel(const double *__restrict__ a1, int m,
                                 const double *__restrict__ am, double tol,
                                 double *__restrict__ sc) {
  const double nrm = sqrt(a1[m - 1]);
  const double a = sqrt(am[0]);
  const bool keep = a > tol * nrm;
  sc[0] = keep ? 1.0 / a : 0.0;
  sc[1] = keep ? 1.0 : 0.0;
}

__global__ void pj_scale_kernel(double *__restrict__ Xm, double *__restrict__ Bm, int64_t n,
                                const double *__restrict__ sc) {
  if (sc[1] == 0.0) return;
  const double f = sc[0];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    Xm[i] = f * Xm[i];
    Bm[i] = f * Bm[i];
  }
}

}  // namespace

struct Proj {
  const Op *op = nullptr;
  int L = 0, m = 0, activ = 5;
  double tol = 1e-7;
  int64_t n = 0;
  double *X = nullptr, *B = nullptr, *xbar = nullptr;
  double *part = nullptr;   // RED_BLOCKS * PJ_LMAX
  double *scal = nullptr;   // a[32] work[32] beta[32] am res sc[2]
  double proj_res_sq = 0.0;
  int proj_m = 0;
  bool res_pending = false;
};

static void proj_release(Proj *p) {
  for (void *q : {(void *)p->X, (void *)p->B, (void *)p->xbar, (void *)p->part, (void *)p->scal})
    if (q) (void)hipFree(q);
  delete p;
}

static int proj_clear(Proj *p, hipStream_t s) {
  p->m = 0;
  p->proj_m = 0;
  if (p->L > 0) {
    NM_HIP_TRY(hipMemsetAsync(p->X, 0, sizeof(double) * p->L * p->n, s));
    NM_HIP_TRY(hipMemsetAsync(p->B, 0, sizeof(double) * p->L * p->n, s));
  }
  return NM_OK;
}

static int proj_project_on(Proj *p, double *b, hipStream_t s) {
  if (p->m <= 0) return NM_OK;
  const Op &op = *p->op;
  const double *w = op.f64.mult;
  double *a = p->scal, *work = p->scal + 32, *res = p->scal + 97;
  int st;
  if ((st = launch_glsc3_partial(b, b, w, p->n, p->part, s))) return st;
  if ((st = launch_combine(p->part, 1, res, s))) return st;
  p->res_pending = true;
  p->proj_m = p->m;
  hipLaunchKernelGGL(pj_dots_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, p->X, p->n,
                     p->m, b, w, p->part);
  NM_HIP_TRY(hipGetLastError());
  if ((st = launch_combine(p->part, p->m, a, s))) return st;
  hipLaunchKernelGGL(pj_on2_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, p->X, p->B,
                     p->n, p->m, a, b, p->xbar, w, p->part);
  NM_HIP_TRY(hipGetLastError());
  if ((st = launch_combine(p->part, p->m, work, s))) return st;
  hipLaunchKernelGGL(pj_on3_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, p->X, p->B,
                     p->n, p->m, work, b, p->xbar);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

static int proj_project_back(Proj *p, double *x, hipStream_t s) {
  const Op &op = *p->op;
  const int64_t n = p->n;
  int st;
  if (p->m > 0 &&
      (st = launch_vec<double>(V_ADD2, x, p->xbar, nullptr, nullptr, 0.0, n, s)))
    return st;
  p->m = (p->m == p->L) ? 1 : std::min(p->m + 1, p->L);
  const int m = p->m;
  double *Xm = p->X + (int64_t)(m - 1) * n, *Bm = p->B + (int64_t)(m - 1) * n;
  NM_HIP_TRY(hipMemcpyAsync(Xm, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  if ((st = op.apply<double>(Bm, x, nullptr, s))) return st;
  // cpu_proj_ortho
  const double *w = op.f64.mult;
  double *a = p->scal, *beta = p->scal + 64, *am = p->scal + 96, *sc = p->scal + 98;
  hipLaunchKernelGGL(pj_ortho_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, p->X,
                     p->B, n, m, (const double *)nullptr, 0, m, w, p->part);
  NM_HIP_TRY(hipGetLastError());
  if ((st = launch_combine(p->part, m, a, s))) return st;
  if (m > 1) {
    hipLaunchKernelGGL(pj_ortho_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, p->X,
                       p->B, n, m, (const double *)a, m - 1, m - 1, w, p->part);
    NM_HIP_TRY(hipGetLastError());
    if ((st = launch_combine(p->part, m - 1, beta, s))) return st;
  }
  hipLaunchKernelGGL(pj_ortho_last_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, p->X,
                     p->B, n, m, (const double *)beta, m - 1, w, p->part);
  NM_HIP_TRY(hipGetLastError());
  if ((st = launch_combine(p->part, 1, am, s))) return st;
  hipLaunchKernelGGL(pj_decide_kernel, dim3(1), dim3(1), 0, s, (const double *)a, m,
                     (const double *)am, p->tol, sc);
  hipLaunchKernelGGL(pj_scale_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, Xm, Bm, n,
                     (const double *)sc);
  NM_HIP_TRY(hipGetLastError());
  double keep = 0.0;
  NM_HIP_TRY(hipMemcpyAsync(&keep, sc + 1, sizeof(double), hipMemcpyDeviceToHost, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  if (keep == 0.0) p->m -= 1;  // not linearly independent: forget it
  return NM_OK;
}

}  // namespace nm

using namespace nm;

struct nm_proj_s {
  Proj *p;
};

extern "C" {

int nm_proj_create(nm_proj_t **out, nm_op_t *op, int L, int activ_step) {
  *out = nullptr;
  if (L < 0 || L > PJ_LMAX) {
    set_error("nm_proj_create: L=%d outside 0..%d", L, PJ_LMAX);
    return NM_ERR_ARG;
  }
  Proj *p = new Proj;
  p->op = &op->op;
  p->L = L;
  p->activ = activ_step;
  p->n = op->op.n;
  const int64_t n = p->n;
  bool ok = true;
  if (L > 0) {
    ok = hipMalloc((void **)&p->X, sizeof(double) * L * n) == hipSuccess &&
         hipMalloc((void **)&p->B, sizeof(double) * L * n) == hipSuccess &&
         hipMalloc((void **)&p->xbar, sizeof(double) * n) == hipSuccess &&
         hipMalloc((void **)&p->part, sizeof(double) * RED_BLOCKS * PJ_LMAX) == hipSuccess &&
         hipMalloc((void **)&p->scal, sizeof(double) * 128) == hipSuccess;
  }
  if (!ok) {
    proj_release(p);
    set_error("nm_proj_create: out of device memory (L=%d, n=%lld)", L, (long long)n);
    return NM_ERR_NOMEM;
  }
  if (L > 0) {
    hipStream_t s = nullptr;
    int st = proj_clear(p, s);
    if (!st && (hipMemset(p->xbar, 0, sizeof(double) * n) != hipSuccess ||
                hipMemset(p->scal, 0, sizeof(double) * 128) != hipSuccess ||
                hipDeviceSynchronize() != hipSuccess))
      st = NM_ERR_HIP;
    if (st) {
      proj_release(p);
      return st;
    }
  }
  *out = new nm_proj_s{p};
  return NM_OK;
}

void nm_proj_free(nm_proj_t *h) {
  if (!h) return;
  (void)hipDeviceSynchronize();
  proj_release(h->p);
  delete h;
}

int nm_proj_clear(nm_proj_t *h, hipStream_t s) { return proj_clear(h->p, s); }

// pre_solving (projection.f90:236-280): dt_last_change < 0 for a constant
// time step; otherwise the steps since the last change of dt (0 clears the
// basis, projection resumes after activ_step steps).
int nm_proj_pre(nm_proj_t *h, double *b, int tstep, int dt_last_change, hipStream_t s) {
  Proj *p = h->p;
  if (!(tstep > p->activ && p->L > 0)) return NM_OK;
  if (dt_last_change >= 0) {
    if (dt_last_change == 0) return proj_clear(p, s);
    if (!(dt_last_change > p->activ - 1)) return NM_OK;
  }
  return proj_project_on(p, b, s);
}

// post_solving (projection.f90:282-301)
int nm_proj_post(nm_proj_t *h, double *x, int tstep, int dt_last_change, hipStream_t s) {
  Proj *p = h->p;
  if (!(tstep > p->activ && p->L > 0)) return NM_OK;
  if (dt_last_change >= 0 && !(dt_last_change > p->activ - 1)) return NM_OK;
  return proj_project_back(p, x, s);
}

// basis size and, for the last project_on, the vectors used and the
// original residual sqrt(b.mult.b / volume) (print_proj_info)
int nm_proj_info(nm_proj_t *h, int *m, int *proj_m, double *proj_res, hipStream_t s) {
  Proj *p = h->p;
  if (p->res_pending) {
    NM_HIP_TRY(hipMemcpyAsync(&p->proj_res_sq, p->scal + 97, sizeof(double),
                              hipMemcpyDeviceToHost, s));
    NM_HIP_TRY(hipStreamSynchronize(s));
    p->res_pending = false;
  }
  if (m) *m = p->m;
  if (proj_m) *proj_m = p->proj_m;
  if (proj_res) *proj_res = std::sqrt(p->proj_res_sq / p->op->volume);
  return NM_OK;
}

}  // extern "C"
