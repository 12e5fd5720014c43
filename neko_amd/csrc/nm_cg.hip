// Jacobi-preconditioned CG on the device (FP64, and FP32-inner / FP64
// iterative refinement).  Semantics of Neko's ksp_t CG (krylov.f90:56-66,
// 411-421; cg.f90:140-279): residual norm sqrt(r.mult.r)/sqrt(volume), stop
// on an absolute tolerance, iteration count reported as Neko does.
//
// Unlike Neko's device CG (one host synchronisation per dot product), every
// scalar stays on the device: alpha, beta and the stopping test are
// computed by one-block "finish" kernels, every kernel of an iteration
// exits early once the device flag `done` is set, and the host only polls
// the flag between batches of graph-captured iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "nm_common.h"
#include "nm_kernels.h"
#include "nm_cg.h"
#include "nm_op.h"

namespace nm {

// ---- kernels ---------------------------------------------------------------
__device__ inline double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline void bsum2(double &a, double &b, double *sh) {
  a = wsum(a);
  b = wsum(b);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    sh[wid] = a;
    sh[8 + wid] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ra = 0.0, rb = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
      ra += sh[q];
      rb += sh[8 + q];
    }
    a = ra;
    b = rb;
  }
  __syncthreads();
}

// alpha = rtz / p.Ap from the p.Ap partials exactly as cg_alpha_kernel forms
// it (same partials, same per-thread order, same tree), taken by every
// workgroup of an update kernel itself instead of a separate launch;
// workgroup 0 also stores pap and alpha for the kernels that follow (the
// next Ax's x update, a flush).  Blocks of RED_THREADS threads.
__device__ inline double upd_alpha(const double *__restrict__ apart, int64_t napart,
                                   CgScal *__restrict__ sc, double *sh) {
  __shared__ double s_alpha;
  double a = 0.0, b = 0.0;
  for (int64_t q = threadIdx.x; q < napart; q += blockDim.x) a += apart[q];
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    const double al = sc->rtz / a;
    s_alpha = al;
    if (blockIdx.x == 0) {
      sc->pap = a;
      sc->alpha = al;
    }
  }
  __syncthreads();
  return s_alpha;
}

// init: x = 0, p = 0, r = scale_in * f; partial rtr = sum r r mult,
// rtz = sum r r mult dinv.  scale_in comes from the device (refinement) or 1.
template <typename T, typename TX, typename TF>
__global__ void __launch_bounds__(RED_THREADS)
    cg_init_kernel(TX *__restrict__ x, T *__restrict__ r, T *__restrict__ p,
                   const TF *__restrict__ f, const T *__restrict__ mult,
                   const T *__restrict__ dinv, int64_t n,
                   const CgScal *__restrict__ sc_in, double *__restrict__ part) {
  __shared__ double sh[16];
  const double sf = sc_in ? sc_in->scale : 1.0;
  double a = 0.0, b = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const T ri = (T)((double)f[i] * sf);
    x[i] = TX(0);
    p[i] = T(0);
    r[i] = ri;
    const double rm = (double)ri * (double)ri * (double)mult[i];
    a += rm;
    b += rm * (double)dinv[i];
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[RED_BLOCKS + blockIdx.x] = b;
  }
}

// finish of init: rnorm, res_start, tolerance (absolute, or relative to the
// start residual when rel > 0), done flag.
__global__ void cg_init_fin_kernel(const double *__restrict__ part,
                                   CgScal *__restrict__ sc, double abs_tol,
                                   double rel, int max_iter) {
  __shared__ double sh[16];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < RED_BLOCKS; q += blockDim.x) {
    a += part[q];
    b += part[RED_BLOCKS + q];
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    sc->rtr = a;
    sc->rtz = b;
    sc->rtz_old = 1.0;
    sc->beta = 0.0;
    sc->rnorm = sqrt(a) * sc->norm_fac;
    sc->res_start = sc->rnorm;
    sc->tol = rel > 0.0 ? rel * sc->rnorm : abs_tol;
    sc->iter = 0;
    sc->max_iter = max_iter;
    sc->converged = 0;
    sc->done = (fabs(sc->rnorm) < 2.220446049250313e-16) ? 1 : 0;
    if (sc->done) sc->converged = 1;
  }
}

// p = dinv * r + beta * p
template <typename T>
__global__ void cg_p_kernel(T *__restrict__ p, const T *__restrict__ r,
                            const T *__restrict__ dinv, int64_t n,
                            const CgScal *__restrict__ sc) {
  if (sc->done) return;
  const T beta = (T)sc->beta;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = dinv[i] * r[i] + beta * p[i];
}

// partial pap = sum w mult p
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
    cg_pap_kernel(const T *__restrict__ w, const T *__restrict__ mult,
                  const T *__restrict__ p, int64_t n, const CgScal *__restrict__ sc,
                  double *__restrict__ part) {
  if (sc->done) return;
  __shared__ double sh[16];
  double a = 0.0, b = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    a += (double)w[i] * (double)mult[i] * (double)p[i];
  bsum2(a, b, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ void cg_alpha_kernel(const double *__restrict__ part, CgScal *__restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[16];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < RED_BLOCKS; q += blockDim.x) a += part[q];
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    sc->pap = a;
    sc->alpha = sc->rtz / a;
  }
}

// x += alpha p; r -= alpha w; partial rtr and rtz of the new r
template <typename T, typename TX>
__global__ void __launch_bounds__(RED_THREADS)
    cg_update_kernel(TX *__restrict__ x, T *__restrict__ r, const T *__restrict__ p,
                     const T *__restrict__ w, const T *__restrict__ mult,
                     const T *__restrict__ dinv, int64_t n,
                     const CgScal *__restrict__ sc, double *__restrict__ part) {
  if (sc->done) return;
  __shared__ double sh[16];
  const double alpha = sc->alpha;
  const T alpha_t = (T)alpha;
  double a = 0.0, b = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    x[i] = x[i] + (TX)(alpha_t * p[i]);
    const T ri = r[i] - alpha_t * w[i];
    r[i] = ri;
    const double rm = (double)ri * (double)ri * (double)mult[i];
    a += rm;
    b += rm * (double)dinv[i];
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[RED_BLOCKS + blockIdx.x] = b;
  }
}

// part holds npart partial rtr sums, then npart partial rtz sums
__global__ void cg_fin_kernel(const double *__restrict__ part, int npart,
                              CgScal *__restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[16];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < npart; q += blockDim.x) {
    a += part[q];
    b += part[npart + q];
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    sc->rtr = a;
    sc->rtz_old = sc->rtz;
    sc->rtz = b;
    sc->beta = b / sc->rtz_old;
    sc->rnorm = sqrt(a) * sc->norm_fac;
    sc->iter = sc->iter + 1;
    if (sc->rnorm < sc->tol) {
      sc->done = 1;
      sc->converged = 1;
    } else if (sc->iter >= sc->max_iter) {
      sc->done = 1;
    }
  }
}

// fused schedule (CG prologue inside the Ax kernel): r -= alpha w; partial
// rtr and rtz of the new r
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
    cg_update2_kernel(T *__restrict__ r, const T *__restrict__ w, const T *__restrict__ mult,
                      const T *__restrict__ dinv, int64_t n, CgScal *__restrict__ sc,
                      double *__restrict__ part, const double *__restrict__ apart,
                      int64_t napart) {
  if (sc->done) return;
  __shared__ double sh[16];
  const T alpha_t = (T)upd_alpha(apart, napart, sc, sh);
  double a = 0.0, b = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const T ri = r[i] - alpha_t * w[i];
    r[i] = ri;
    const double rm = (double)ri * (double)ri * (double)mult[i];
    a += rm;
    b += rm * (double)dinv[i];
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[RED_BLOCKS + blockIdx.x] = b;
  }
}

// As cg_update2_kernel, with the facet-pair sums of gs taken here: w holds
// the local Ax result after the block sums, the facet pairs no descriptor
// covers and the mask, while the interior points of descriptor faces still
// hold their own element's value.  Such a point adds its partner's value
// (w_a + w_b, the same bits the face kernel would have stored, since the
// sum commutes exactly) where w is read anyway, which saves the face
// kernel's read and write of w.  One element per loop step, so the point's
// (i, j, k) come from the thread index.  UPF_BLOCKS workgroups (more than
// the plain update's RED_BLOCKS: the partner load waits on the face-table
// load, and the extra waves hide that chain; 2048 timed best against 1024,
// 4096 and 8192, and against keeping 2 or 4 elements in flight per thread;
// fewer on problems with fewer element steps), workgroup b of G writing its
// partial rtr to part[b] and rtz to part[G + b].
// (Walking chunks of consecutive elements, dealt out so that a point's
// face partners run on the same XCD at the same time, timed slower: the
// partner lines cost less than the lost spread of the element walk.)
constexpr int UPF_BLOCKS = 2048;
// The grid is one round of the chip at 8 waves per SIMD (2048 workgroups of
// four waves on 256 CUs): the kernels below are held to 64 VGPRs up to lx =
// 10, since a second, partial round of workgroups (each with an equal share
// of the elements) would leave most of the chip idle at the end; beyond,
// the points per thread need more registers than that without spilling.
// With CNT the multiplicity weight comes from the one-byte count (mult is
// exactly 1/count): tab[count] = (T)(1.0 / count), the bits of mult itself.
template <typename T, int LX, bool CNT>
__global__ void __launch_bounds__(RED_THREADS) __attribute__((amdgpu_waves_per_eu(LX <= 10 ? 8 : 1)))
    cg_update2f_kernel(T *__restrict__ r, const T *__restrict__ w, const T *__restrict__ mult,
                       const uint8_t *__restrict__ mcnt, const T *__restrict__ dinv,
                       const int2 *__restrict__ fnb, int64_t nel, CgScal *__restrict__ sc,
                       double *__restrict__ part, const double *__restrict__ apart,
                       int64_t napart) {
  if (sc->done) return;
  constexpr int N3 = LX * LX * LX;
  constexpr int NQ = (N3 + RED_THREADS - 1) / RED_THREADS;  // points per thread and element
  __shared__ double sh[16];
  __shared__ T tab[CNT ? 256 : 1];
  if (CNT) {
    for (int c = threadIdx.x; c < 256; c += RED_THREADS) tab[c] = c ? (T)(1.0 / (double)c) : T(0);
    __syncthreads();
  }
  const T alpha_t = (T)upd_alpha(apart, napart, sc, sh);
  // where this thread's points sit in an element (the same for every
  // element): face f of a face-interior point and its (u, v), else f = -1
  int fq[NQ], uq[NQ], vq[NQ];
#pragma unroll
  for (int t = 0; t < NQ; t++) {
    const int q = threadIdx.x + t * RED_THREADS;
    const int ii = q % LX, jj = (q / LX) % LX, kk = q / (LX * LX);
    const bool bi = ii == 0 || ii == LX - 1, bj = jj == 0 || jj == LX - 1,
               bk = kk == 0 || kk == LX - 1;
    fq[t] = (q < N3 && (int)bi + (int)bj + (int)bk == 1)
                ? (bi ? (ii != 0) : (bj ? 2 + (jj != 0) : 4 + (kk != 0)))
                : -1;
    uq[t] = bi ? jj : ii;
    vq[t] = bk ? jj : kk;
  }
  double a = 0.0, b = 0.0;
  for (int64_t e = blockIdx.x; e < nel; e += gridDim.x) {
#pragma unroll
    for (int t = 0; t < NQ; t++) {
      const int q = threadIdx.x + t * RED_THREADS;
      if (q >= N3) break;
      const int64_t i = e * N3 + q;
      T wi = w[i];
      if (fq[t] >= 0) {
        const int2 d = fnb[e * 6 + fq[t]];
        if (d.x >= 0) {
          int ub = (d.y & 64) ? vq[t] : uq[t], vb = (d.y & 64) ? uq[t] : vq[t];
          if (d.y & 128) ub = LX - 1 - ub;
          if (d.y & 256) vb = LX - 1 - vb;
          wi = wi + w[(int64_t)d.x * N3 + face_point((d.y >> 3) & 7, ub, vb, LX)];
        }
      }
      const T ri = __builtin_nontemporal_load(r + i) - alpha_t * wi;
      __builtin_nontemporal_store(ri, r + i);
      const T mi = CNT ? tab[__builtin_nontemporal_load(mcnt + i)] : mult[i];
      const double rm = (double)ri * (double)ri * (double)mi;
      a += rm;
      b += rm * (double)__builtin_nontemporal_load(dinv + i);
    }
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = b;
  }
}

// Even lx: the same update with a point pair per thread (2h, 2h + 1 share
// j and k, and sit at an even offset), so w, r and dinv move as 8-byte
// (FP32) or 16-byte (FP64) vectors and the counts as 2-byte pairs: half
// the memory instructions of the one-point form.  Used for FP32 (64^3,
// lx 8: mixed iteration 1.81 -> 1.74 ms); the FP64 pair form timed 1 %
// slower than the one-point form, which FP64 keeps.
template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

template <typename T, int LX, bool CNT>
__global__ void __launch_bounds__(RED_THREADS) __attribute__((amdgpu_waves_per_eu(LX <= 10 ? 8 : 1)))
    cg_update2f_pair_kernel(T *__restrict__ r, const T *__restrict__ w,
                            const T *__restrict__ mult, const uint8_t *__restrict__ mcnt,
                            const T *__restrict__ dinv, const int2 *__restrict__ fnb,
                            int64_t nel, CgScal *__restrict__ sc,
                            double *__restrict__ part, const double *__restrict__ apart,
                            int64_t napart) {
  static_assert(LX % 2 == 0, "point pairs need an even lx");
  using T2 = typename Vec2<T>::type;
  if (sc->done) return;
  constexpr int N3 = LX * LX * LX, NH = N3 / 2;
  constexpr int NQ = (NH + RED_THREADS - 1) / RED_THREADS;  // pairs per thread and element
  __shared__ double sh[16];
  __shared__ T tab[CNT ? 256 : 1];
  if (CNT) {
    for (int c = threadIdx.x; c < 256; c += RED_THREADS) tab[c] = c ? (T)(1.0 / (double)c) : T(0);
    __syncthreads();
  }
  const T alpha_t = (T)upd_alpha(apart, napart, sc, sh);
  // per pair slot and point: face of a face-interior point (else -1), (u, v)
  int fq[NQ][2], uq[NQ][2], vq[NQ][2];
#pragma unroll
  for (int t = 0; t < NQ; t++)
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int q = 2 * (threadIdx.x + t * RED_THREADS) + c;
      const int ii = q % LX, jj = (q / LX) % LX, kk = q / (LX * LX);
      const bool bi = ii == 0 || ii == LX - 1, bj = jj == 0 || jj == LX - 1,
                 bk = kk == 0 || kk == LX - 1;
      fq[t][c] = (q < N3 && (int)bi + (int)bj + (int)bk == 1)
                     ? (bi ? (ii != 0) : (bj ? 2 + (jj != 0) : 4 + (kk != 0)))
                     : -1;
      uq[t][c] = bi ? jj : ii;
      vq[t][c] = bk ? jj : kk;
    }
  double a = 0.0, b = 0.0;
  for (int64_t e = blockIdx.x; e < nel; e += gridDim.x) {
#pragma unroll
    for (int t = 0; t < NQ; t++) {
      const int h = threadIdx.x + t * RED_THREADS;
      if (h >= NH) break;
      const int64_t i = e * N3 + 2 * h;
      const T2 wv = *reinterpret_cast<const T2 *>(w + i);
      const T2 rv = *reinterpret_cast<const T2 *>(r + i);
      const T2 dv = *reinterpret_cast<const T2 *>(dinv + i);
      T m0, m1;
      if (CNT) {
        const uchar2 cv = *reinterpret_cast<const uchar2 *>(mcnt + i);
        m0 = tab[cv.x];
        m1 = tab[cv.y];
      } else {
        const T2 mv = *reinterpret_cast<const T2 *>(mult + i);
        m0 = mv.x;
        m1 = mv.y;
      }
      T wi[2] = {wv.x, wv.y};
#pragma unroll
      for (int c = 0; c < 2; c++) {
        if (fq[t][c] >= 0) {
          const int2 d = fnb[e * 6 + fq[t][c]];
          if (d.x >= 0) {
            int ub = (d.y & 64) ? vq[t][c] : uq[t][c], vb = (d.y & 64) ? uq[t][c] : vq[t][c];
            if (d.y & 128) ub = LX - 1 - ub;
            if (d.y & 256) vb = LX - 1 - vb;
            wi[c] = wi[c] + w[(int64_t)d.x * N3 + face_point((d.y >> 3) & 7, ub, vb, LX)];
          }
        }
      }
      T2 ro;
      ro.x = rv.x - alpha_t * wi[0];
      ro.y = rv.y - alpha_t * wi[1];
      *reinterpret_cast<T2 *>(r + i) = ro;
      const double r0 = (double)ro.x * (double)ro.x * (double)m0;
      const double r1 = (double)ro.y * (double)ro.y * (double)m1;
      a += r0;
      b += r0 * (double)dv.x;
      a += r1;
      b += r1 * (double)dv.y;
    }
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = b;
  }
}

// four floats as a native vector, for the streaming (nontemporal) loads
// and stores: r and dinv are touched once per iteration, so they bypass
// the caches that the partner values of w are reused from
typedef float nt_f4 __attribute__((ext_vector_type(4)));

// lx divisible by 4, FP32: four neighbouring points per thread (16-byte
// loads and stores); with fewer quads per element than threads, a
// workgroup takes 256 / (N3 / 4) elements per step.
template <int LX, bool CNT>
__global__ void __launch_bounds__(RED_THREADS) __attribute__((amdgpu_waves_per_eu(LX <= 10 ? 8 : 1)))
    cg_update2f_quad_kernel(float *__restrict__ r, const float *__restrict__ w,
                            const float *__restrict__ mult, const uint8_t *__restrict__ mcnt,
                            const float *__restrict__ dinv, const int2 *__restrict__ fnb,
                            int64_t nel, CgScal *__restrict__ sc,
                            double *__restrict__ part, const double *__restrict__ apart,
                            int64_t napart) {
  static_assert(LX % 4 == 0, "point quads need lx divisible by 4");
  constexpr int N3 = LX * LX * LX, NH = N3 / 4;
  constexpr int EPI = NH >= RED_THREADS ? 1 : RED_THREADS / NH;  // elements per step
  constexpr int NQ = NH >= RED_THREADS ? (NH + RED_THREADS - 1) / RED_THREADS : 1;
  if (sc->done) return;
  __shared__ double sh[16];
  __shared__ float tab[CNT ? 256 : 1];
  if (CNT) {
    for (int c = threadIdx.x; c < 256; c += RED_THREADS) tab[c] = c ? (float)(1.0 / (double)c) : 0.0f;
    __syncthreads();
  }
  const float alpha_t = (float)upd_alpha(apart, napart, sc, sh);
  const int eo = EPI > 1 ? (int)threadIdx.x / NH : 0;  // element offset in the step
  const int h0 = EPI > 1 ? (int)threadIdx.x % NH : (int)threadIdx.x;
  int fq[NQ][4], uq[NQ][4], vq[NQ][4];
#pragma unroll
  for (int t = 0; t < NQ; t++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int q = 4 * (h0 + t * RED_THREADS) + c;
      const int ii = q % LX, jj = (q / LX) % LX, kk = q / (LX * LX);
      const bool bi = ii == 0 || ii == LX - 1, bj = jj == 0 || jj == LX - 1,
                 bk = kk == 0 || kk == LX - 1;
      fq[t][c] = (q < N3 && (int)bi + (int)bj + (int)bk == 1)
                     ? (bi ? (ii != 0) : (bj ? 2 + (jj != 0) : 4 + (kk != 0)))
                     : -1;
      uq[t][c] = bi ? jj : ii;
      vq[t][c] = bk ? jj : kk;
    }
  double a = 0.0, b = 0.0;
  if (eo < EPI) {
    for (int64_t e = (int64_t)blockIdx.x * EPI + eo; e < nel; e += (int64_t)gridDim.x * EPI) {
#pragma unroll
      for (int t = 0; t < NQ; t++) {
        const int h = h0 + t * RED_THREADS;
        if (h >= NH) break;
        const int64_t i = e * N3 + 4 * h;
        const float4 wv = *reinterpret_cast<const float4 *>(w + i);
        const nt_f4 rn = __builtin_nontemporal_load(reinterpret_cast<const nt_f4 *>(r + i));
        const nt_f4 dn = __builtin_nontemporal_load(reinterpret_cast<const nt_f4 *>(dinv + i));
        const float4 rv = make_float4(rn.x, rn.y, rn.z, rn.w);
        const float4 dv = make_float4(dn.x, dn.y, dn.z, dn.w);
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
        float wi[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
        for (int c = 0; c < 4; c++) {
          if (fq[t][c] >= 0) {
            const int2 d = fnb[e * 6 + fq[t][c]];
            if (d.x >= 0) {
              int ub = (d.y & 64) ? vq[t][c] : uq[t][c], vb = (d.y & 64) ? uq[t][c] : vq[t][c];
              if (d.y & 128) ub = LX - 1 - ub;
              if (d.y & 256) vb = LX - 1 - vb;
              wi[c] = wi[c] + w[(int64_t)d.x * N3 + face_point((d.y >> 3) & 7, ub, vb, LX)];
            }
          }
        }
        float4 ro;
        ro.x = rv.x - alpha_t * wi[0];
        ro.y = rv.y - alpha_t * wi[1];
        ro.z = rv.z - alpha_t * wi[2];
        ro.w = rv.w - alpha_t * wi[3];
        nt_f4 rs;
        rs.x = ro.x;
        rs.y = ro.y;
        rs.z = ro.z;
        rs.w = ro.w;
        __builtin_nontemporal_store(rs, reinterpret_cast<nt_f4 *>(r + i));
        const float ro_[4] = {ro.x, ro.y, ro.z, ro.w};
        const float dv_[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const double rr = (double)ro_[c] * (double)ro_[c] * (double)m[c];
          a += rr;
          b += rr * (double)dv_[c];
        }
      }
    }
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = b;
  }
}

// apart/napart: the p.Ap partials alpha is formed from (never `part`, the
// update's own partials, which other workgroups write meanwhile)
// Workgroups of the face-summing update: UPF_BLOCKS, or fewer when the
// operator has fewer element steps than that (small problems: no idle
// workgroups to launch and no zero partials to sum)
template <typename T, int LX>
static int upf_grid(int64_t nel) {
  int64_t epi = 1;
  if constexpr (LX % 4 == 0 && std::is_same<T, float>::value) {
    constexpr int NH = LX * LX * LX / 4;
    epi = NH >= RED_THREADS ? 1 : RED_THREADS / NH;
  }
  return (int)std::max<int64_t>(1, std::min<int64_t>(UPF_BLOCKS, (nel + epi - 1) / epi));
}

template <typename T, int LX, bool CNT>
static int launch_update2f_lx(const Op &op, T *r, const T *w, const int2 *fnb, CgScal *sc,
                              double *part, const double *apart, int64_t napart,
                              hipStream_t s) {
  const OpFields<T> &F = op.fields<T>();
  const int grid = upf_grid<T, LX>(op.nel);
  if constexpr (LX % 4 == 0 && std::is_same<T, float>::value)
    hipLaunchKernelGGL((cg_update2f_quad_kernel<LX, CNT>), dim3(grid), dim3(RED_THREADS), 0, s,
                       r, w, F.mult, op.mcnt, F.dinv, fnb, op.nel, sc, part, apart, napart);
  else if constexpr (LX % 2 == 0 && std::is_same<T, float>::value)
    hipLaunchKernelGGL((cg_update2f_pair_kernel<T, LX, CNT>), dim3(grid), dim3(RED_THREADS), 0,
                       s, r, w, F.mult, op.mcnt, F.dinv, fnb, op.nel, sc, part, apart, napart);
  else
    hipLaunchKernelGGL((cg_update2f_kernel<T, LX, CNT>), dim3(grid), dim3(RED_THREADS), 0, s,
                       r, w, F.mult, op.mcnt, F.dinv, fnb, op.nel, sc, part, apart, napart);
  return grid;
}

template <typename T>
// *grid: the workgroups launched (= partial pairs written to part)
static int launch_update2f(const Op &op, T *r, const T *w, CgScal *sc, double *part,
                           const double *apart, int64_t napart, hipStream_t s, int *grid) {
  const int2 *fnb = (const int2 *)op.face_nbr;
#define NM_U2F(L)                                                                           \
  case L:                                                                                   \
    if (op.mcnt_ok)                                                                         \
      *grid = launch_update2f_lx<T, L, true>(op, r, w, fnb, sc, part, apart, napart, s);    \
    else                                                                                    \
      *grid = launch_update2f_lx<T, L, false>(op, r, w, fnb, sc, part, apart, napart, s);   \
    break;
  switch (op.lx) {
    NM_U2F(3) NM_U2F(4) NM_U2F(5) NM_U2F(6) NM_U2F(7) NM_U2F(8) NM_U2F(9) NM_U2F(10)
    NM_U2F(11) NM_U2F(12) NM_U2F(13) NM_U2F(14)
    default:
      set_error("cg: face sums in the update need 3 <= lx <= %d", NM_FACE_FUSE_MAX_LX);
      return NM_ERR_UNSUPPORTED;
  }
#undef NM_U2F
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// the x update the fused schedule still owes: x += alpha p
template <typename T>
__global__ void cg_flush_kernel(T *__restrict__ x, const T *__restrict__ p, int64_t n,
                                const CgScal *__restrict__ sc) {
  const T alpha_t = (T)sc->alpha;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    x[i] = x[i] + alpha_t * p[i];
}
__global__ void cg_clear_alpha_kernel(CgScal *sc) { sc->alpha = 0.0; }

// mixed precision: FP32 CG with FP64 residual replacement ----------------------
// The FP32 recurrence runs on (x32, r32, p32, w32) with FP32 operator and
// vectors and FP64 dot-product accumulation.  Whenever its residual has
// dropped by `delta` below the last true residual (or below the final
// tolerance), the FP64 solution absorbs x32 and the true residual
// r = f - A x is recomputed with the FP64 operator and replaces r32; the
// search direction p32 is kept, so the Krylov process continues instead of
// restarting.

// start: x = 0, x32 = 0, p32 = 0, r32 = (float) f; partial rtr (FP64 f) and
// rtz (r32 with the FP32 Jacobi diagonal)
__global__ void __launch_bounds__(RED_THREADS)
    mx_init_kernel(double *__restrict__ x, float *__restrict__ x32, float *__restrict__ r32,
                   float *__restrict__ p32, const double *__restrict__ f,
                   const double *__restrict__ mult, const float *__restrict__ dinv32,
                   int64_t n, double *__restrict__ part) {
  __shared__ double sh[16];
  double a = 0.0, b = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double fi = f[i];
    const float ri = (float)fi;
    x[i] = 0.0;
    x32[i] = 0.0f;
    p32[i] = 0.0f;
    r32[i] = ri;
    a += fi * fi * mult[i];
    b += (double)ri * (double)ri * mult[i] * (double)dinv32[i];
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[RED_BLOCKS + blockIdx.x] = b;
  }
}

__global__ void mx_init_fin_kernel(const double *__restrict__ part, CgScal *__restrict__ sc,
                                   double abs_tol, double rel, double delta, int max_iter) {
  __shared__ double sh[16];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < RED_BLOCKS; q += blockDim.x) {
    a += part[q];
    b += part[RED_BLOCKS + q];
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    sc->rtr = a;
    sc->rtz = b;
    sc->rtz_old = 1.0;
    sc->beta = 0.0;
    sc->rnorm = sqrt(a) * sc->norm_fac;
    sc->res_start = sc->rnorm;
    sc->tol_final = rel > 0.0 ? rel * sc->rnorm : abs_tol;
    sc->delta = delta;
    sc->true_rnorm = sc->rnorm;
    sc->tol = fmax(delta * sc->rnorm, sc->tol_final);
    sc->iter = 0;
    sc->updates = 0;
    sc->max_iter = max_iter;
    sc->converged = 0;
    sc->done = (fabs(sc->rnorm) < 2.220446049250313e-16 || sc->rnorm < sc->tol_final) ? 1 : 0;
    if (sc->done) sc->converged = 1;
    sc->fin = sc->done;
  }
}

// x += x32, x32 = 0
__global__ void mx_absorb_kernel(double *__restrict__ x, float *__restrict__ x32, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    x[i] += (double)x32[i];
    x32[i] = 0.0f;
  }
}

// true residual r = f - w (w = A x, FP64), r32 = (float) r; partial rtr
// (FP64) and rtz (r32, FP32 Jacobi)
__global__ void __launch_bounds__(RED_THREADS)
    mx_resid_kernel(float *__restrict__ r32, const double *__restrict__ f,
                    const double *__restrict__ w, const double *__restrict__ mult,
                    const float *__restrict__ dinv32, int64_t n, double *__restrict__ part) {
  __shared__ double sh[16];
  double a = 0.0, b = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ri = f[i] - w[i];
    const float r3 = (float)ri;
    r32[i] = r3;
    a += ri * ri * mult[i];
    b += (double)r3 * (double)r3 * mult[i] * (double)dinv32[i];
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[RED_BLOCKS + blockIdx.x] = b;
  }
}

// replacement scalars: true norm, new rtz and beta (the next direction is
// p = z + beta p with z from the replaced residual), next threshold.
// With out (a replacement beside the iteration graph): a continuing
// problem keeps done = 1, its scalars go to *out, then rearm = 1 is stored
// with release order at device scope (the graph's step that clears it sees
// every write before it); a finished one copies its scalars to *out.
__global__ void mx_resid_fin_kernel(const double *__restrict__ part, CgScal *__restrict__ sc,
                                    CgScal *__restrict__ out) {
  __shared__ double sh[16];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < RED_BLOCKS; q += blockDim.x) {
    a += part[q];
    b += part[RED_BLOCKS + q];
  }
  bsum2(a, b, sh);
  if (threadIdx.x == 0) {
    sc->rtr = a;
    sc->rnorm = sqrt(a) * sc->norm_fac;
    sc->true_rnorm = sc->rnorm;
    sc->updates = sc->updates + 1;
    if (sc->iter > 0) {
      sc->rtz = b;
      sc->beta = b / sc->rtz_old;
    }
    sc->tol = fmax(sc->delta * sc->rnorm, sc->tol_final);
    if (sc->rnorm < sc->tol_final || fabs(sc->rnorm) < 2.220446049250313e-16) {
      sc->done = 1;
      sc->converged = 1;
      sc->fin = 1;
    } else if (sc->iter >= sc->max_iter) {
      sc->done = 1;
      sc->converged = 0;
      sc->fin = 1;
    } else {
      sc->done = out ? 1 : 0;
      sc->converged = 0;
    }
    if (out) {
      *out = *sc;
      if (!sc->fin) __hip_atomic_store(&sc->rearm, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

static inline unsigned vec_blocks(int64_t n) {
  int64_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  return (unsigned)nb;
}

// ---- solver object -----------------------------------------------------------
// fold `npart` per-workgroup partials of the fused p.Ap into RED_BLOCKS
// partials (block b sums one contiguous chunk; fixed order)
__global__ void __launch_bounds__(RED_THREADS)
    fold_partials_kernel(const double *__restrict__ in, int64_t npart,
                         double *__restrict__ out, const CgScal *__restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[16];
  const int64_t chunk = (npart + RED_BLOCKS - 1) / RED_BLOCKS;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < npart ? lo + chunk : npart;
  double a = 0.0, b = 0.0;
  for (int64_t q = lo + threadIdx.x; q < hi; q += blockDim.x) a += in[q];
  bsum2(a, b, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = a;
}

// One CG iteration (T = vector type, TX = solution type) on stream s.
// With dpart (npart > 0) p.Ap comes fused out of the Ax kernel as
// sum_i p_i (A_local p)_i, which equals glsc3(w, mult, p) for the
// continuous, masked p of CG (the multiplicity weights undo the gs sum);
// otherwise it is a separate pass over w, mult and p.
template <typename T, typename TX>
static int cg_iteration(const Op &op, TX *x, T *r, T *p, T *w, CgScal *sc,
                        double *part, hipStream_t s, double *dpart = nullptr,
                        int64_t npart = 0, bool cgp = false, const MaskRef *mk = nullptr) {
  const MaskRef m = mk ? *mk : op.mask_ref();
  const OpFields<T> &F = op.fields<T>();
  const int64_t n = op.n;
  if constexpr (std::is_same<T, TX>::value) {
    if (cgp) {
      // p update and the previous x update ride in the Ax kernel, the
      // descriptor face sums of gs in the update kernel
      const bool ff = m.face_fuse && op.face_fuse_on;
      int st = op.apply_cgp<T>(w, p, r, x, &sc->beta, &sc->alpha, &sc->done, s, dpart, ff, &m);
      if (st) return st;
      // alpha is formed inside the update kernel from apart.  Up to
      // RED_BLOCKS Ax partials it reads them directly (folding them would
      // copy them unchanged) and the update's partials go to `part`;
      // beyond, they are folded into `part` first and the update's
      // partials go to dpart.
      const bool small = npart <= RED_BLOCKS;
      const double *apart = small ? dpart : part;
      const int64_t napart = small ? npart : RED_BLOCKS;
      double *upart = small ? part : dpart;
      if (!small)
        hipLaunchKernelGGL(fold_partials_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s,
                           dpart, npart, part, sc);
      if (ff) {
        int ug = 0;
        if ((st = launch_update2f<T>(op, r, w, sc, upart, apart, napart, s, &ug))) return st;
        hipLaunchKernelGGL(cg_fin_kernel, dim3(1), dim3(RED_THREADS), 0, s, upart, ug, sc);
        NM_HIP_TRY(hipGetLastError());
        return NM_OK;
      }
      hipLaunchKernelGGL(cg_update2_kernel<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, r, w,
                         F.mult, F.dinv, n, sc, upart, apart, napart);
      hipLaunchKernelGGL(cg_fin_kernel, dim3(1), dim3(RED_THREADS), 0, s, upart, RED_BLOCKS, sc);
      NM_HIP_TRY(hipGetLastError());
      return NM_OK;
    }
  }
  hipLaunchKernelGGL(cg_p_kernel<T>, dim3(vec_blocks(n)), dim3(256), 0, s, p, r, F.dinv, n, sc);
  const bool fused = npart > 0 && dpart != nullptr;
  int st = op.apply<T>(w, p, &sc->done, s, true, true, fused ? dpart : nullptr, &m);
  if (st) return st;
  if (fused)
    hipLaunchKernelGGL(fold_partials_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, dpart,
                       npart, part, sc);
  else
    hipLaunchKernelGGL(cg_pap_kernel<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, w, F.mult,
                       p, n, sc, part);
  hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(RED_THREADS), 0, s, part, sc);
  hipLaunchKernelGGL((cg_update_kernel<T, TX>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, x,
                     r, p, w, F.mult, F.dinv, n, sc, part);
  hipLaunchKernelGGL(cg_fin_kernel, dim3(1), dim3(RED_THREADS), 0, s, part, RED_BLOCKS, sc);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// Fixed-budget Jacobi-PCG without any host synchronisation: init plus
// `niter` iterations are enqueued; every kernel turns into a no-op once the
// device flag is set (converged below abs_tol or niter reached).  This is
// the coarse-grid solve of hsmg (pc_hsmg.f90:598-608: CG, Jacobi, 10
// iterations, Neko's default absolute tolerance).
struct CgFixed {
  const Op *op = nullptr;
  double *r = nullptr, *p = nullptr, *w = nullptr, *part = nullptr;
  CgScal *sc = nullptr;
  std::vector<void *> owned;
  ~CgFixed() {
    for (void *q : owned) (void)hipFree(q);
  }
};

int cg_fixed_create(CgFixed **out, const Op *op) {
  CgFixed *c = new CgFixed;
  c->op = op;
  auto al = [&](void **q, size_t bytes) {
    if (hipMalloc(q, bytes) != hipSuccess) return false;
    c->owned.push_back(*q);
    return true;
  };
  const size_t nb = sizeof(double) * (size_t)std::max<int64_t>(op->n, 1);
  if (!al((void **)&c->r, nb) || !al((void **)&c->p, nb) || !al((void **)&c->w, nb) ||
      !al((void **)&c->part, sizeof(double) * 2 * RED_BLOCKS) ||
      !al((void **)&c->sc, sizeof(CgScal))) {
    delete c;
    set_error("cg_fixed_create: out of device memory");
    return NM_ERR_NOMEM;
  }
  *out = c;
  return NM_OK;
}

void cg_fixed_free(CgFixed *c) { delete c; }

// x = CG(A, f) with at most niter iterations, all on stream s.
int cg_fixed_solve(CgFixed *c, double *x, const double *f, int niter,
                   double abs_tol, hipStream_t s) {
  const Op &op = *c->op;
  const int64_t n = op.n;
  CgScal init{};
  init.norm_fac = 1.0 / std::sqrt(op.volume);
  init.scale = 1.0;
  NM_HIP_TRY(hipMemcpyAsync(c->sc, &init, sizeof(CgScal), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL((cg_init_kernel<double, double, double>), dim3(RED_BLOCKS),
                     dim3(RED_THREADS), 0, s, x, c->r, c->p, f, op.f64.mult,
                     op.f64.dinv, n, (const CgScal *)nullptr, c->part);
  hipLaunchKernelGGL(cg_init_fin_kernel, dim3(1), dim3(RED_THREADS), 0, s, c->part,
                     c->sc, abs_tol, 0.0, niter);
  for (int k = 0; k < niter; k++) {
    int st = cg_iteration<double, double>(op, x, c->r, c->p, c->w, c->sc, c->part, s);
    if (st) return st;
  }
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

CgSolver::~CgSolver() {
  if (g64) (void)hipGraphExecDestroy(g64);
  if (g32) (void)hipGraphExecDestroy(g32);
  if (ev) (void)hipEventDestroy(ev);
  if (s) (void)hipStreamSynchronize(s);
  for (void *q : owned) (void)hipFree(q);
  if (h_sc) (void)hipHostFree(h_sc);
  if (s) (void)hipStreamDestroy(s);
}

template <typename T>
T *CgSolver::alloc(int64_t cnt) {
  void *q = nullptr;
  const size_t b = sizeof(T) * (size_t)std::max<int64_t>(cnt, 1);
  if (hipMalloc(&q, b) != hipSuccess) return nullptr;
  owned.push_back(q);
  bytes += (int64_t)b;
  return (T *)q;
}

template <typename T, typename TX>
static int capture_iters(CgSolver *cg, TX *x, T *r, T *p, T *w, CgScal *sc,
                         hipGraphExec_t *out) {
  hipGraph_t graph;
  NM_HIP_TRY(hipStreamBeginCapture(cg->s, hipStreamCaptureModeThreadLocal));
  int st = NM_OK;
  const bool fused = cg->op->fused_dot<T>();
  cg->cgp = fused && std::is_same<T, TX>::value;
  for (int k = 0; k < cg->graph_iters && st == NM_OK; k++)
    st = cg_iteration<T, TX>(*cg->op, x, r, p, w, sc, cg->part, cg->s,
                             fused ? cg->dpart : nullptr, fused ? cg->npart : 0, cg->cgp,
                             cg->own_mask ? &cg->mk : nullptr);
  hipError_t e = hipStreamEndCapture(cg->s, &graph);
  if (st) return st;
  NM_HIP_TRY(e);
  NM_HIP_TRY(hipGraphInstantiate(out, graph, nullptr, nullptr, 0));
  NM_HIP_TRY(hipGraphDestroy(graph));
  return NM_OK;
}

// (Re)capture the iteration graph for the op's current variant codes,
// coefficients and preconditioner arrays (kernel arguments are baked into
// a captured graph); cg_begin calls it again whenever op->epoch moved.
static int cg_capture(CgSolver *cg) {
  Op *op = cg->op;
  // per-workgroup partials of the fused p.Ap
  const int code = cg->mixed ? op->ax_code<float>() : op->ax_code<double>();
  cg->npart = 0;
  {
    cg->npart = ax_grid(op->nel, op->lx, code);
    // the buffer also takes the face-sum update's 2 UPF_BLOCKS partials
    const int64_t need = std::max<int64_t>(cg->npart, 2 * UPF_BLOCKS);
    if (need > cg->npart_cap) {
      if (!(cg->dpart = cg->alloc<double>(need))) {
        set_error("cg: out of device memory");
        return NM_ERR_NOMEM;
      }
      cg->npart_cap = need;
    }
  }
  hipGraphExec_t *g = cg->mixed ? &cg->g32 : &cg->g64;
  if (*g) {
    NM_HIP_TRY(hipStreamSynchronize(cg->s));
    NM_HIP_TRY(hipGraphExecDestroy(*g));
    *g = nullptr;
  }
  int st;
  if (!cg->mixed)
    st = capture_iters<double, double>(cg, cg->x, cg->r, cg->p, cg->w, cg->sc, g);
  else
    st = capture_iters<float, float>(cg, cg->e32, cg->r32, cg->p32, cg->w32, cg->sc, g);
  if (st) return st;
  cg->epoch = op->epoch;
  return NM_OK;
}

int cg_solver_create(CgSolver **out, Op *op, int precision) {
  *out = nullptr;
  CgSolver *cg = new CgSolver;
  cg->op = op;
  cg->mixed = precision;
  const int64_t n = op->n;
  auto fail = [&](int code) {
    delete cg;
    return code;
  };
  if (hipStreamCreateWithFlags(&cg->s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&cg->ev, hipEventDisableTiming) != hipSuccess) {
    set_error("cg: stream/event creation failed");
    return fail(NM_ERR_HIP);
  }
  if (!(cg->x = cg->alloc<double>(n)) || !(cg->r = cg->alloc<double>(n)) ||
      !(cg->w = cg->alloc<double>(n)) || !(cg->f = cg->alloc<double>(n)) ||
      !(cg->part = cg->alloc<double>(2 * std::max(RED_BLOCKS, UPF_BLOCKS))) ||
      !(cg->sc = cg->alloc<CgScal>(2))) {
    set_error("cg: out of device memory");
    return fail(NM_ERR_NOMEM);
  }
  if (!precision) {
    if (!(cg->p = cg->alloc<double>(n))) {
      set_error("cg: out of device memory");
      return fail(NM_ERR_NOMEM);
    }
  } else {
    if (!(cg->e32 = cg->alloc<float>(n)) || !(cg->r32 = cg->alloc<float>(n)) ||
        !(cg->p32 = cg->alloc<float>(n)) || !(cg->w32 = cg->alloc<float>(n))) {
      set_error("cg: out of device memory");
      return fail(NM_ERR_NOMEM);
    }
    int st = op->ensure32(cg->s);
    if (st) return fail(st);
  }
  if (hipHostMalloc((void **)&cg->h_sc, sizeof(CgScal) * 2, hipHostMallocDefault) != hipSuccess) {
    set_error("cg: pinned host allocation failed");
    return fail(NM_ERR_NOMEM);
  }
  int st = cg_capture(cg);
  if (st) return fail(st);
  if (hipStreamSynchronize(cg->s) != hipSuccess) return fail(NM_ERR_HIP);
  *out = cg;
  return NM_OK;
}

void cg_solver_destroy(CgSolver *cg) { delete cg; }

// enqueue `chunk` graph launches, the scalar read-back and the event
static int enqueue_chunk(CgSolver *cg, hipGraphExec_t g) {
  for (int b = 0; b < cg->chunk; b++) {
    NM_HIP_TRY(hipGraphLaunch(g, cg->s));
    cg->launched += cg->graph_iters;
  }
  if (cg->chunk < cg->chunk_max) cg->chunk *= 2;
  NM_HIP_TRY(hipMemcpyAsync(cg->h_sc, cg->sc, 2 * sizeof(CgScal), hipMemcpyDeviceToHost, cg->s));
  NM_HIP_TRY(hipEventRecord(cg->ev, cg->s));
  cg->pending = true;
  return NM_OK;
}

// residual replacement: x += x32, r = f - A x (FP64 operator), r32 = r,
// new scalars, read-back
int mx_slice_begin(const Op &op, double *x, float *x32, float *r32, float *p32,
                   const double *f, CgScal *sc, double *part, double abs_tol, double rel,
                   double delta, int max_iter, hipStream_t s) {
  hipLaunchKernelGGL(mx_init_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, x, x32, r32,
                     p32, f, op.f64.mult, op.f32.dinv, op.n, part);
  hipLaunchKernelGGL(mx_init_fin_kernel, dim3(1), dim3(RED_THREADS), 0, s, part, sc, abs_tol,
                     rel, delta, max_iter);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// the x update a fused (CG prologue) schedule still owes before a
// replacement: x32 += alpha p32, then alpha = 0 (the next prologue adds
// nothing for the kept p)
int mx_slice_flush(float *x32, const float *p32, int64_t n, CgScal *sc, hipStream_t s) {
  hipLaunchKernelGGL(cg_flush_kernel<float>, dim3(vec_blocks(n)), dim3(256), 0, s, x32, p32, n,
                     sc);
  hipLaunchKernelGGL(cg_clear_alpha_kernel, dim3(1), dim3(1), 0, s, sc);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

int mx_slice_replace(const Op &op, double *x, float *x32, float *r32, const double *f,
                     double *w, CgScal *sc, double *part, hipStream_t s, const MaskRef *mk,
                     CgScal *rep_out) {
  const int64_t n = op.n;
  hipLaunchKernelGGL(mx_absorb_kernel, dim3(vec_blocks(n)), dim3(256), 0, s, x, x32, n);
  int st = op.apply<double>(w, x, nullptr, s, true, true, nullptr, mk);
  if (st) return st;
  hipLaunchKernelGGL(mx_resid_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, r32, f, w,
                     op.f64.mult, op.f32.dinv, n, part);
  hipLaunchKernelGGL(mx_resid_fin_kernel, dim3(1), dim3(RED_THREADS), 0, s, part, sc, rep_out);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

static int enqueue_replace(CgSolver *cg) {
  if (cg->cgp) {
    const int64_t n = cg->op->n;
    hipLaunchKernelGGL(cg_flush_kernel<float>, dim3(vec_blocks(n)), dim3(256), 0, cg->s, cg->e32,
                       cg->p32, n, cg->sc);
    hipLaunchKernelGGL(cg_clear_alpha_kernel, dim3(1), dim3(1), 0, cg->s, cg->sc);
    NM_HIP_TRY(hipGetLastError());
  }
  int st = mx_slice_replace(*cg->op, cg->x, cg->e32, cg->r32, cg->f, cg->w, cg->sc, cg->part,
                            cg->s, cg->own_mask ? &cg->mk : nullptr);
  if (st) return st;
  NM_HIP_TRY(hipMemcpyAsync(cg->h_sc, cg->sc, 2 * sizeof(CgScal), hipMemcpyDeviceToHost, cg->s));
  NM_HIP_TRY(hipEventRecord(cg->ev, cg->s));
  cg->pending = true;
  return NM_OK;
}

int cg_begin(CgSolver *cg, const double *f, double abs_tol, double rel_tol,
             int max_iter, double inner_rel) {
  const Op &op = *cg->op;
  const int64_t n = op.n;
  if (cg->pending) {
    set_error("cg: a solve is still in flight on this solver");
    return NM_ERR_STATE;
  }
  if (cg->epoch != op.epoch) {
    if (cg->mixed) {
      int st0 = cg->op->ensure32(cg->s);
      if (st0) return st0;
    }
    int st1 = cg_capture(cg);
    if (st1) return st1;
  }
  cg->abs_tol = abs_tol;
  cg->rel_tol = rel_tol;
  cg->inner_rel = inner_rel > 0.0 ? inner_rel : 0.1;
  cg->max_iter = max_iter;
  cg->total = cg->sweeps = cg->launched = 0;
  cg->chunk = 1;
  cg->res_start = -1.0;
  cg->rnorm = 0.0;
  cg->converged = 0;
  CgScal init{};
  init.norm_fac = 1.0 / std::sqrt(op.volume);
  init.scale = 1.0;
  std::memcpy(&cg->h_sc[0], &init, sizeof(CgScal));
  std::memcpy(&cg->h_sc[1], &init, sizeof(CgScal));
  NM_HIP_TRY(hipMemcpyAsync(cg->sc, cg->h_sc, 2 * sizeof(CgScal), hipMemcpyHostToDevice, cg->s));
  if (f) NM_HIP_TRY(hipMemcpyAsync(cg->f, f, sizeof(double) * n, hipMemcpyDeviceToDevice, cg->s));
  if (!cg->mixed) {
    hipLaunchKernelGGL((cg_init_kernel<double, double, double>), dim3(RED_BLOCKS),
                       dim3(RED_THREADS), 0, cg->s, cg->x, cg->r, cg->p, cg->f,
                       op.f64.mult, op.f64.dinv, n, (const CgScal *)nullptr, cg->part);
    hipLaunchKernelGGL(cg_init_fin_kernel, dim3(1), dim3(RED_THREADS), 0, cg->s, cg->part,
                       cg->sc, abs_tol, rel_tol, max_iter);
    NM_HIP_TRY(hipGetLastError());
    cg->phase = CgSolver::RUN64;
    return enqueue_chunk(cg, cg->g64);
  }
  int st = mx_slice_begin(op, cg->x, cg->e32, cg->r32, cg->p32, cg->f, cg->sc, cg->part,
                          abs_tol, rel_tol, cg->inner_rel, max_iter, cg->s);
  if (st) return st;
  cg->phase = CgSolver::MIX;
  return enqueue_chunk(cg, cg->g32);
}

int cg_poll(CgSolver *cg) {
  if (cg->phase == CgSolver::DONE) return 1;
  if (cg->phase == CgSolver::IDLE) {
    set_error("cg: poll without a solve");
    return NM_ERR_STATE;
  }
  if (cg->pending) {
    const hipError_t q = hipEventQuery(cg->ev);
    if (q == hipErrorNotReady) return 0;
    if (q != hipSuccess) {
      set_error("cg: %s", hipGetErrorString(q));
      return NM_ERR_HIP;
    }
    cg->pending = false;
  }
  const Op &op = *cg->op;
  const int64_t n = op.n;
  const CgScal &in = cg->h_sc[0];
  switch (cg->phase) {
    case CgSolver::RUN64:
      if (in.done) {
        if (cg->cgp) {
          hipLaunchKernelGGL(cg_flush_kernel<double>, dim3(vec_blocks(n)), dim3(256), 0, cg->s,
                             cg->x, cg->p, n, cg->sc);
          hipLaunchKernelGGL(cg_clear_alpha_kernel, dim3(1), dim3(1), 0, cg->s, cg->sc);
          NM_HIP_TRY(hipGetLastError());
        }
        cg->total = in.iter;
        cg->res_start = in.res_start;
        cg->rnorm = in.rnorm;
        cg->converged = in.converged;
        cg->phase = CgSolver::DONE;
        return 1;
      }
      if (cg->launched > cg->max_iter + 8 * cg->graph_iters) {
        set_error("cg: device did not stop after %d iterations", cg->launched);
        return NM_ERR_STATE;
      }
      {
        const int st = enqueue_chunk(cg, cg->g64);
        return st ? st : 0;
      }
    case CgSolver::MIX: {
      // FP32 recurrence stopped (or still running): replace the residual
      if (!in.done) {
        if (cg->launched > cg->max_iter + 8 * cg->graph_iters) {
          set_error("cg: device did not stop after %d iterations", cg->launched);
          return NM_ERR_STATE;
        }
        const int st = enqueue_chunk(cg, cg->g32);
        return st ? st : 0;
      }
      cg->phase = CgSolver::REPLACE;
      const int st = enqueue_replace(cg);
      return st ? st : 0;
    }
    case CgSolver::REPLACE: {
      cg->res_start = in.res_start;
      cg->rnorm = in.true_rnorm;
      cg->total = in.iter;
      cg->sweeps = in.updates;
      if (in.done) {
        cg->converged = in.converged;
        cg->phase = CgSolver::DONE;
        return 1;
      }
      cg->phase = CgSolver::MIX;
      const int st = enqueue_chunk(cg, cg->g32);
      return st ? st : 0;
    }
    default:
      set_error("cg: bad phase");
      return NM_ERR_STATE;
  }
}

int cg_wait(CgSolver *cg) {
  for (;;) {
    if (cg->pending) NM_HIP_TRY(hipEventSynchronize(cg->ev));
    const int st = cg_poll(cg);
    if (st < 0 || st > 0) return st < 0 ? st : NM_OK;
    if (st == 0 && !cg->pending) {
      set_error("cg: solver stalled");
      return NM_ERR_STATE;
    }
  }
}

}  // namespace nm

using namespace nm;

struct nm_cg_s {
  CgSolver *cg = nullptr;
  ~nm_cg_s() {
    if (cg) cg_solver_destroy(cg);
  }
};

extern "C" {

// precision: 0 = FP64 throughout; 1 = FP32 inner CG inside FP64 refinement.
int nm_cg_create(nm_cg_t **out, nm_op_t *oph, int precision) {
  *out = nullptr;
  nm_cg_t *h = new nm_cg_t;
  const int st = cg_solver_create(&h->cg, &oph->op, precision);
  if (st) {
    delete h;
    return st;
  }
  *out = h;
  return NM_OK;
}

void nm_cg_free(nm_cg_t *h) { delete h; }

// The solver's own zero-Dirichlet mask (1-based msk, nmsk entries), used in
// place of the operator's by every operator application of its solves;
// NULL restores the operator's.  Captured graphs are re-captured.
int nm_cg_set_mask(nm_cg_t *h, const int32_t *msk, int64_t nmsk, hipStream_t s) {
  CgSolver *cg = h->cg;
  const Op &op = *cg->op;
  if (!msk) {
    cg->own_mask = false;
    cg->epoch = ~(uint64_t)0;
    return NM_OK;
  }
  if (nmsk < 0) {
    set_error("nm_cg_set_mask: nmsk = %lld", (long long)nmsk);
    return NM_ERR_ARG;
  }
  std::vector<int32_t> m0((size_t)std::max<int64_t>(nmsk, 1));
  for (int64_t q = 0; q < nmsk; q++) {
    if (msk[q] < 1 || msk[q] > op.n) {
      set_error("nm_cg_set_mask: entry %lld is %d, outside 1..%lld", (long long)q, msk[q],
                (long long)op.n);
      return NM_ERR_ARG;
    }
    m0[q] = msk[q] - 1;
  }
  if (nmsk > cg->mask_cap) {
    if (!(cg->mask_dev = cg->alloc<int32_t>(nmsk))) {
      set_error("nm_cg_set_mask: out of device memory");
      return NM_ERR_NOMEM;
    }
    cg->mask_cap = nmsk;
  }
  NM_HIP_TRY(hipStreamSynchronize(cg->s));
  if (nmsk > 0)
    NM_HIP_TRY(hipMemcpyAsync(cg->mask_dev, m0.data(), sizeof(int32_t) * nmsk,
                              hipMemcpyHostToDevice, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  cg->mk = MaskRef{cg->mask_dev, nmsk, op.face_fuse_ok(m0.data(), nmsk)};
  cg->own_mask = true;
  cg->epoch = ~(uint64_t)0;  // re-capture on the next solve
  return NM_OK;
}

// Solve A x = f.  x, f: device FP64 arrays of op->n entries (caller owned).
// abs_tol: Neko absolute tolerance on sqrt(r.mult.r)/sqrt(volume).
// Mixed mode: FP32 CG (FP32 operator and vectors, FP64 dot products) whose
// residual is replaced by the true FP64 residual f - A x each time the FP32
// recurrence has dropped by `inner_rel` (<= 0: default 0.1) below the last
// true residual; stops on the true residual.  mon->outer counts the
// replacements.  `stream` is ordered before and after the solve.
int nm_cg_solve(nm_cg_t *h, double *x, const double *f, double abs_tol,
                int max_iter, double inner_rel, nm_monitor_t *mon,
                hipStream_t stream) {
  CgSolver *cg = h->cg;
  const int64_t n = cg->op->n;
  std::memset(mon, 0, sizeof(*mon));
  hipEvent_t ev;
  NM_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  NM_HIP_TRY(hipEventRecord(ev, stream));
  NM_HIP_TRY(hipStreamWaitEvent(cg->s, ev, 0));
  int st = cg_begin(cg, f, abs_tol, 0.0, max_iter, inner_rel);
  if (!st) st = cg_wait(cg);
  if (st) {
    (void)hipEventDestroy(ev);
    return st;
  }
  // an exhausted run reports max_iter + 1, the value cg.f90's
  // `do iter = 1, max_iter` loop variable has after the loop (ksp_results%iter)
  mon->iter = (!cg->converged && cg->total >= max_iter) ? max_iter + 1 : cg->total;
  mon->outer = cg->sweeps;
  mon->res_start = cg->res_start;
  mon->res_final = cg->rnorm;
  mon->converged = cg->converged;
  NM_HIP_TRY(hipMemcpyAsync(x, cg->x, sizeof(double) * n, hipMemcpyDeviceToDevice, cg->s));
  NM_HIP_TRY(hipEventRecord(ev, cg->s));
  NM_HIP_TRY(hipStreamWaitEvent(stream, ev, 0));
  NM_HIP_TRY(hipEventDestroy(ev));
  return NM_OK;
}

// cg_solve_coupled (src/krylov/bcknd/cpu/cg.f90:282-305): the three
// velocity components solved one after the other with the same operator
// and preconditioner, component c by cg[c] (whose mask, nm_cg_set_mask, is
// that component's bc list blstx / blsty / blstz), each to abs_tol.
int nm_cg_solve_coupled(nm_cg_t *const *cg, double *const *x, const double *const *f,
                        double abs_tol, int max_iter, double inner_rel, nm_monitor_t *mon,
                        hipStream_t stream) {
  for (int c = 0; c < 3; c++) {
    if (!cg[c] || !x[c] || !f[c]) {
      set_error("nm_cg_solve_coupled: component %d has a null solver or vector", c);
      return NM_ERR_ARG;
    }
    if (cg[c]->cg->op->n != cg[0]->cg->op->n) {
      set_error("nm_cg_solve_coupled: the three solvers work on different sizes");
      return NM_ERR_ARG;
    }
  }
  for (int c = 0; c < 3; c++) {
    const int st = nm_cg_solve(cg[c], x[c], f[c], abs_tol, max_iter, inner_rel, &mon[c], stream);
    if (st) return st;
  }
  return NM_OK;
}

}  // extern "C"
