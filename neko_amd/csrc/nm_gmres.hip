// Restarted GMRES(m) on the device (gmres.f90:175-411 semantics): right
// preconditioning (z_j = M v_j, x += sum c_j z_j), classical Gram-Schmidt
// with the j dot products of one iteration fused into one pass over w,
// the orthogonalisation update fused with the new norm (w -= sum h_l v_l,
// alpha^2 = w.mult.w), Givens rotations and the back substitution on the
// host.  One host synchronisation per iteration (the j+1 scalars).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nm_common.h"
#include "nm_kernels.h"
#include "nm_op.h"


namespace nm {

// w -= sum_{l<j} h_l v_l ; partial sums of w*w*mult
__global__ void __launch_bounds__(RED_THREADS)
    gmres_part2_kernel(double *__restrict__ w, const double *const *__restrict__ v,
                       const double *__restrict__ h, const double *__restrict__ mult,
                       int j, int64_t n, double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  __shared__ double hs[64];
  if (threadIdx.x < j) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // the basis vectors eight at a time (their loads in flight together) and
  // two points per trip; every sum runs in the order of the plain loop
  // over l and over the thread's points
  constexpr int G = 8;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + stride < n; i += 2 * stride) {
    const int64_t i2 = i + stride;
    double t = 0.0, t2 = 0.0;
    for (int l0 = 0; l0 < j; l0 += G) {
      double a[G], b[G];
#pragma unroll
      for (int g = 0; g < G; g++) {
        a[g] = l0 + g < j ? v[l0 + g][i] : 0.0;
        b[g] = l0 + g < j ? v[l0 + g][i2] : 0.0;
      }
#pragma unroll
      for (int g = 0; g < G; g++)
        if (l0 + g < j) {
          t -= hs[l0 + g] * a[g];
          t2 -= hs[l0 + g] * b[g];
        }
    }
    const double wi = w[i] + t, wi2 = w[i2] + t2;
    w[i] = wi;
    w[i2] = wi2;
    acc += wi * wi * mult[i];
    acc += wi2 * wi2 * mult[i2];
  }
  if (i < n) {
    double t = 0.0;
    for (int l = 0; l < j; l++) t -= hs[l] * v[l][i];
    const double wi = w[i] + t;
    w[i] = wi;
    acc += wi * wi * mult[i];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) r += sh[q];
    part[blockIdx.x] = r;
  }
}

// x += sum_{l<j} c_l z_l
__global__ void add2s2_many_kernel(double *__restrict__ x, const double *const *__restrict__ z,
                                   const double *__restrict__ c, int j, int64_t n) {
  __shared__ double cs[64];
  if (threadIdx.x < j) cs[threadIdx.x] = c[threadIdx.x];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double t = 0.0;
    for (int l = 0; l < j; l++) t += cs[l] * z[l][i];
    x[i] += t;
  }
}

// r = f - w, partial r.mult.r
__global__ void __launch_bounds__(RED_THREADS)
    sub_norm_kernel(double *__restrict__ r, const double *__restrict__ f,
                    const double *__restrict__ w, const double *__restrict__ mult,
                    int64_t n, int use_w, double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ri = use_w ? f[i] - w[i] : f[i];
    r[i] = ri;
    acc += ri * ri * mult[i];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += sh[q];
    part[blockIdx.x] = s;
  }
}

static inline unsigned vb(int64_t n) {
  int64_t nb = (n + 255) / 256;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(nb, 65536));
}

struct Gmres {
  Op *op = nullptr;
  int lg = 30;
  int pc = 0;  // 0 Jacobi, 1 hsmg, 2 none, 3 phmg
  void *pch = nullptr;
  double *V = nullptr, *Z = nullptr, *w = nullptr, *r = nullptr, *f = nullptr;
  double *x = nullptr;
  const double **vptr = nullptr, **zptr = nullptr;
  double *part = nullptr, *scal = nullptr, *hc = nullptr;  // scal: lg+1
  double *h_scal = nullptr;                                 // pinned
  std::vector<void *> owned;
  ~Gmres() {
    for (void *q : owned) (void)hipFree(q);
    if (h_scal) (void)hipHostFree(h_scal);
  }
  template <typename T>
  T *alloc(int64_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess) return nullptr;
    owned.push_back(q);
    return (T *)q;
  }
};

static int gmres_pc(Gmres *g, double *z, const double *v, hipStream_t s) {
  const Op &op = *g->op;
  if (g->pc == 1) return nm_hsmg_apply((nm_hsmg_t *)g->pch, z, v, s);
  if (g->pc == 3) return nm_phmg_apply((nm_phmg_t *)g->pch, z, v, s);
  if (g->pc == 0)
    return launch_vec<double>(V_COL3, z, v, op.f64.dinv, nullptr, 0.0, op.n, s);
  NM_HIP_TRY(hipMemcpyAsync(z, v, sizeof(double) * op.n, hipMemcpyDeviceToDevice, s));
  return NM_OK;
}

}  // namespace nm

using namespace nm;

struct nm_gmres_s {
  Gmres g;
};

extern "C" {

// pc: 0 Jacobi (the operator's diagonal), 1 hsmg, 3 phmg (pass the
// handle), 2 none.
int nm_gmres_create(nm_gmres_t **out, nm_op_t *oph, int pc, void *pch, int lgmres) {
  *out = nullptr;
  if (lgmres < 1 || lgmres > 64) {
    set_error("nm_gmres_create: restart length %d outside 1..64", lgmres);
    return NM_ERR_ARG;
  }
  if ((pc == 1 || pc == 3) && !pch) {
    set_error("nm_gmres_create: multigrid preconditioner without handle");
    return NM_ERR_ARG;
  }
  if (pc < 0 || pc > 3) {
    set_error("nm_gmres_create: unknown preconditioner %d", pc);
    return NM_ERR_ARG;
  }
  nm_gmres_t *h = new nm_gmres_t;
  Gmres *g = &h->g;
  g->op = &oph->op;
  g->lg = lgmres;
  g->pc = pc;
  g->pch = pch;
  const int64_t n = g->op->n;
  if (!(g->V = g->alloc<double>(n * lgmres)) || !(g->Z = g->alloc<double>(n * lgmres)) ||
      !(g->w = g->alloc<double>(n)) || !(g->r = g->alloc<double>(n)) ||
      !(g->f = g->alloc<double>(n)) || !(g->x = g->alloc<double>(n)) ||
      !(g->vptr = g->alloc<const double *>(lgmres)) ||
      !(g->zptr = g->alloc<const double *>(lgmres)) ||
      !(g->part = g->alloc<double>(RED_BLOCKS * (lgmres + 1))) ||
      !(g->scal = g->alloc<double>(lgmres + 2)) || !(g->hc = g->alloc<double>(lgmres + 2))) {
    delete h;
    set_error("nm_gmres_create: out of device memory");
    return NM_ERR_NOMEM;
  }
  std::vector<const double *> vp(lgmres), zp(lgmres);
  for (int l = 0; l < lgmres; l++) {
    vp[l] = g->V + (int64_t)l * n;
    zp[l] = g->Z + (int64_t)l * n;
  }
  if (hipMemcpy(g->vptr, vp.data(), sizeof(double *) * lgmres, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->zptr, zp.data(), sizeof(double *) * lgmres, hipMemcpyHostToDevice) != hipSuccess ||
      hipHostMalloc((void **)&g->h_scal, sizeof(double) * (lgmres + 2), hipHostMallocDefault) != hipSuccess) {
    delete h;
    set_error("nm_gmres_create: setup copies failed");
    return NM_ERR_HIP;
  }
  *out = h;
  return NM_OK;
}

void nm_gmres_free(nm_gmres_t *h) { delete h; }

int nm_gmres_solve(nm_gmres_t *hh, double *x_out, const double *f_in,
                   double abs_tol, int max_iter, nm_gmres_monitor_t *mon,
                   hipStream_t s) {
  Gmres *g = &hh->g;
  const Op &op = *g->op;
  const int64_t n = op.n;
  const int lg = g->lg;
  std::memset(mon, 0, sizeof(*mon));
  const double norm_fac = 1.0 / std::sqrt(op.volume);
  std::vector<double> H((size_t)lg * lg, 0.0), c(lg, 1.0), sn(lg, 1.0), gam(lg + 1, 0.0);
  auto Hm = [&](int a, int b) -> double & { return H[(a - 1) + (size_t)lg * (b - 1)]; };
  int st;
  NM_HIP_TRY(hipMemcpyAsync(g->f, f_in, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  if ((st = launch_vec<double>(V_RZERO, g->x, nullptr, nullptr, nullptr, 0.0, n, s))) return st;
  auto fetch = [&](int cnt) -> int {
    NM_HIP_TRY(hipMemcpyAsync(g->h_scal, g->scal, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
    NM_HIP_TRY(hipStreamSynchronize(s));
    return NM_OK;
  };
  int iter = 0, j = 0;
  bool conv = false;
  double rnorm = 0.0;
  while (!conv && iter < max_iter) {
    if (iter != 0 && (st = op.apply<double>(g->w, g->x, nullptr, s, true, true))) return st;
    hipLaunchKernelGGL(sub_norm_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, g->r,
                       g->f, g->w, op.f64.mult, n, iter != 0 ? 1 : 0, g->part);
    if ((st = launch_combine(g->part, 1, g->scal, s))) return st;
    if ((st = fetch(1))) return st;
    gam[0] = std::sqrt(g->h_scal[0]);
    if (iter == 0) mon->res_start = gam[0] * norm_fac;
    if (std::fabs(gam[0]) < 2.220446049250313e-16) break;
    rnorm = 0.0;
    if ((st = launch_vec<double>(V_CMULT2, g->V, g->r, nullptr, nullptr, 1.0 / gam[0], n, s)))
      return st;
    for (j = 1; j <= lg; j++) {
      iter++;
      double *zj = g->Z + (int64_t)(j - 1) * n;
      const double *vj = g->V + (int64_t)(j - 1) * n;
      if ((st = gmres_pc(g, zj, vj, s))) return st;
      if ((st = op.apply<double>(g->w, zj, nullptr, s, true, true))) return st;
      if ((st = launch_glsc3_many_partial(g->w, g->vptr, op.f64.mult, j, n, g->part, s)))
        return st;
      if ((st = launch_combine(g->part, j, g->scal, s))) return st;
      hipLaunchKernelGGL(gmres_part2_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, g->w,
                         g->vptr, g->scal, op.f64.mult, j, n, g->part);
      if ((st = launch_combine(g->part, 1, g->scal + j, s))) return st;
      if ((st = fetch(j + 1))) return st;
      for (int l = 1; l <= j; l++) Hm(l, j) = g->h_scal[l - 1];
      const double alpha2 = g->h_scal[j];
      const double alpha = std::sqrt(alpha2);
      for (int i = 1; i <= j - 1; i++) {
        const double t = Hm(i, j);
        Hm(i, j) = c[i - 1] * t + sn[i - 1] * Hm(i + 1, j);
        Hm(i + 1, j) = -sn[i - 1] * t + c[i - 1] * Hm(i + 1, j);
      }
      rnorm = 0.0;
      if (std::fabs(alpha) < 2.220446049250313e-16) {
        conv = true;
        break;
      }
      const double lr = std::sqrt(Hm(j, j) * Hm(j, j) + alpha2);
      const double t = 1.0 / lr;
      c[j - 1] = Hm(j, j) * t;
      sn[j - 1] = alpha * t;
      Hm(j, j) = lr;
      gam[j] = -sn[j - 1] * gam[j - 1];
      gam[j - 1] = c[j - 1] * gam[j - 1];
      rnorm = std::fabs(gam[j]) * norm_fac;
      if (rnorm < abs_tol) {
        conv = true;
        break;
      }
      if (iter + 1 > max_iter) break;
      if (j < lg) {
        if ((st = launch_vec<double>(V_CMULT2, g->V + (int64_t)j * n, g->w, nullptr, nullptr,
                                     1.0 / alpha, n, s)))
          return st;
      }
    }
    if (j > lg) j = lg;
    for (int k = j; k >= 1; k--) {
      double t = gam[k - 1];
      for (int i = j; i >= k + 1; i--) t -= Hm(k, i) * c[i - 1];
      c[k - 1] = t / Hm(k, k);
    }
    NM_HIP_TRY(hipMemcpyAsync(g->hc, c.data(), sizeof(double) * j, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(add2s2_many_kernel, dim3(vb(n)), dim3(256), 0, s, g->x, g->zptr, g->hc,
                       j, n);
    NM_HIP_TRY(hipStreamSynchronize(s));  // c is a host array reused next cycle
  }
  mon->iter = iter;
  mon->res_final = rnorm;
  mon->converged = !(rnorm > abs_tol) && iter < max_iter;
  NM_HIP_TRY(hipMemcpyAsync(x_out, g->x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  return NM_OK;
}

}  // extern "C"
