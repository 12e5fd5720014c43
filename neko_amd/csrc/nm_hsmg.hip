// hsmg preconditioner on the device and GMRES(m) around it.
//
// hsmg: three-level p-multigrid V-cycle (lx -> lx_mid -> 2) with overlapping
// Schwarz smoothing (overlap one GLL point, fast-diagonalisation local
// solves on the (lx+2)^3 extended element) on the fine and middle levels,
// and a fixed-budget Jacobi-CG on the coarse level (pc_hsmg.f90:147-627,
// schwarz.f90:112-634, fdm.f90:110-665).  GMRES(m) with right
// preconditioning and classical Gram-Schmidt (gmres.f90:175-411).
//
// Kernels: one workgroup per element for the element-local tensor work
// (Schwarz extrusion, FDM, interpolation), LDS-staged; gs and BLAS-1 reuse
// the operator kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "nm_common.h"
#include "nm_kernels.h"
#include "nm_op.h"
#include "nm_rng.h"
#include "nm_tamg.h"

// host set-up helpers of nm_fdm_setup.cpp (internal)
extern "C" {
int nm_fdm_setup_host(int64_t nel, int lx, const double *x, const double *y,
                      const double *z, const double *w, const double *d1,
                      const int *facet_type, void (*gs_l)(void *, double *),
                      void *gs_ctx, double *s, double *d);
int nm_interp_matrix(const double *z_to, int n_to, const double *z_from,
                     int n_from, double *j, double *jt);
}

namespace nm {

// ------------------------------------------------------------------ kernels
__device__ inline int64_t i3(int i, int j, int k, int n) {
  return i + (int64_t)n * (j + (int64_t)n * k);
}

// a (ext, (n+2)^3) = 0 with the interior set to b (n^3)  (schwarz_toext3d)
__global__ void toext_kernel(double *__restrict__ a, const double *__restrict__ b,
                             int n, int64_t nel) {
  const int m = n + 2;
  const int64_t m3 = (int64_t)m * m * m;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nel * m3) return;
  const int64_t e = idx / m3;
  const int p = (int)(idx % m3);
  const int i = p % m, j = (p / m) % m, k = p / (m * m);
  double v = 0.0;
  if (i >= 1 && i <= n && j >= 1 && j <= n && k >= 1 && k <= n)
    v = b[e * (int64_t)n * n * n + i3(i - 1, j - 1, k - 1, n)];
  a[idx] = v;
}

// b (n^3) = interior of a  (schwarz_toreg3d)
__global__ void toreg_kernel(double *__restrict__ b, const double *__restrict__ a,
                             int n, int64_t nel) {
  const int m = n + 2;
  const int64_t n3 = (int64_t)n * n * n;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nel * n3) return;
  const int64_t e = idx / n3;
  const int p = (int)(idx % n3);
  const int i = p % n, j = (p / n) % n, k = p / (n * n);
  b[idx] = a[e * (int64_t)m * m * m + i3(i + 1, j + 1, k + 1, m)];
}

// schwarz_extrude: face layers of arr1 at depth l1 = f1*arr1 + f2*arr2 at
// depth l2, face interiors only; r-faces, then s-faces, then t-faces (the
// reference's loop order, which matters where the layers of two
// directions meet).  One workgroup per element.
__global__ void __launch_bounds__(256)
    extrude_kernel(double *arr1, int l1, double f1, const double *arr2, int l2,
                   double f2, int nx, int64_t nel) {
  const int64_t e = blockIdx.x;
  if (e >= nel) return;
  double *a = arr1 + e * (int64_t)nx * nx * nx;
  const double *b = arr2 + e * (int64_t)nx * nx * nx;
  const int w = nx - 2;  // face-interior width
  const int nf = w * w;
  const int h1 = nx - 1 - l1, h2 = nx - 1 - l2;
  for (int q = threadIdx.x; q < nf; q += blockDim.x) {
    const int j = 1 + q % w, k = 1 + q / w;
    a[i3(l1, j, k, nx)] = f1 * a[i3(l1, j, k, nx)] + f2 * b[i3(l2, j, k, nx)];
    a[i3(h1, j, k, nx)] = f1 * a[i3(h1, j, k, nx)] + f2 * b[i3(h2, j, k, nx)];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nf; q += blockDim.x) {
    const int i = 1 + q % w, k = 1 + q / w;
    a[i3(i, l1, k, nx)] = f1 * a[i3(i, l1, k, nx)] + f2 * b[i3(i, l2, k, nx)];
    a[i3(i, h1, k, nx)] = f1 * a[i3(i, h1, k, nx)] + f2 * b[i3(i, h2, k, nx)];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nf; q += blockDim.x) {
    const int i = 1 + q % w, j = 1 + q / w;
    a[i3(i, j, l1, nx)] = f1 * a[i3(i, j, l1, nx)] + f2 * b[i3(i, j, l2, nx)];
    a[i3(i, j, h1, nx)] = f1 * a[i3(i, j, h1, nx)] + f2 * b[i3(i, j, h2, nx)];
  }
}

// Element-local tensor product v = (C (x) B (x) A) u with per-call small
// matrices: v(i,j,l) = sum A(i,a) Bt(b,j) Ct(c,l) u(a,b,c) (tnsr3d);
// matrices may be per element (stride > 0).  One workgroup per element,
// everything staged in LDS.
template <int MAXN>
__global__ void __launch_bounds__(256)
    tnsr3d_kernel(double *__restrict__ v, int nv, const double *__restrict__ u,
                  int nu, const double *__restrict__ A, const double *__restrict__ Bt,
                  const double *__restrict__ Ct, int64_t mstride, int64_t nel) {
  __shared__ double su[MAXN * MAXN * MAXN];
  __shared__ double s1[MAXN * MAXN * MAXN];
  __shared__ double sA[MAXN * MAXN], sB[MAXN * MAXN], sC[MAXN * MAXN];
  const int64_t e = blockIdx.x;
  if (e >= nel) return;
  const double *ue = u + e * (int64_t)nu * nu * nu;
  double *ve = v + e * (int64_t)nv * nv * nv;
  const int nn = nv * nu, nu3 = nu * nu * nu;
  for (int q = threadIdx.x; q < nn; q += blockDim.x) {
    sA[q] = A[e * mstride + q];
    sB[q] = Bt[e * mstride + q];
    sC[q] = Ct[e * mstride + q];
  }
  {
    // every load of u issued before the first LDS store (blockDim.x 256)
    constexpr int NR = (MAXN * MAXN * MAXN + 255) / 256;
    double lu[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const int q = threadIdx.x + 256 * k;
      lu[k] = q < nu3 ? ue[q] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const int q = threadIdx.x + 256 * k;
      if (q < nu3) su[q] = lu[k];
    }
  }
  __syncthreads();
  // pass 1: w1(i,b,c) = sum_a A(i,a) u(a,b,c)        (nv x nu x nu)
  for (int q = threadIdx.x; q < nv * nu * nu; q += blockDim.x) {
    const int i = q % nv, bc = q / nv;
    double t = 0.0;
    for (int a = 0; a < nu; a++) t += sA[i + nv * a] * su[a + nu * bc];
    s1[q] = t;
  }
  __syncthreads();
  // pass 2: w2(i,j,c) = sum_b w1(i,b,c) Bt(b,j)      (nv x nv x nu) -> su
  for (int q = threadIdx.x; q < nv * nv * nu; q += blockDim.x) {
    const int i = q % nv, j = (q / nv) % nv, c = q / (nv * nv);
    double t = 0.0;
    for (int b = 0; b < nu; b++) t += s1[i + nv * (b + nu * c)] * sB[b + nu * j];
    su[q] = t;
  }
  __syncthreads();
  // pass 3: v(i,j,l) = sum_c w2(i,j,c) Ct(c,l)
  for (int q = threadIdx.x; q < nv * nv * nv; q += blockDim.x) {
    const int ij = q % (nv * nv), l = q / (nv * nv);
    double t = 0.0;
    for (int c = 0; c < nu; c++) t += su[ij + nv * nv * c] * sC[c + nu * l];
    ve[q] = t;
  }
}

// face-interior point q of the 6 faces of an NL^3 block: face f (0,1: r low
// / high, 2,3: s, 4,5: t), (u,v) in [1, NL-2]^2; returns the point index at
// layer `depth` from that face
template <int NL>
__device__ inline int face_point(int f, int u, int v, int depth) {
  const int l = (f & 1) ? NL - 1 - depth : depth;
  switch (f >> 1) {
    case 0: return l + NL * (u + NL * v);
    case 1: return u + NL * (l + NL * v);
    default: return u + NL * (v + NL * l);
  }
}

// fdm_do_fast (3-D): e = (S_t (x) S_s (x) S_r) D (S_t^T (x) S_s^T (x) S_r^T) r
// with per-element S (nl x nl) and D (nl^3); T = double, or float for the
// FP32 smoothers.
template <int NL, bool FUSED, typename T>
__global__ void __launch_bounds__(256)
    fdm_kernel(T *__restrict__ e, T *r, const T *__restrict__ S, const T *__restrict__ D,
               int64_t nel) {
  constexpr int N3 = NL * NL * NL, N2 = NL * NL, W = NL - 2, NF = 6 * W * W;
  __shared__ T a[N3], b[N3];
  __shared__ T s[6 * N2];
  const int64_t el = blockIdx.x;
  if (el >= nel) return;
  T *re = r + el * N3;
  const T *se = S + el * 6 * N2;
  for (int q = threadIdx.x; q < 6 * N2; q += blockDim.x) s[q] = se[q];
  for (int q = threadIdx.x; q < N3; q += blockDim.x) a[q] = re[q];
  __syncthreads();
  if (FUSED) {
    // the neighbour's overlap only: face layer 0 -= layer 2 (face interiors;
    // the six faces are disjoint and read interior layers only)
    for (int q = threadIdx.x; q < NF; q += blockDim.x) {
      const int f = q / (W * W), u = 1 + (q % (W * W)) % W, v = 1 + (q % (W * W)) / W;
      const int p0 = face_point<NL>(f, u, v, 0), p2 = face_point<NL>(f, u, v, 2);
      a[p0] = a[p0] - a[p2];
    }
    __syncthreads();
  }
  const T *S1 = s, *S1t = s + N2, *S2 = s + 2 * N2, *S2t = s + 3 * N2,
               *S3 = s + 4 * N2, *S3t = s + 5 * N2;
  // forward: A = S1^T (as stored in S1t: A(i,a) = S1t[i + NL a]),
  //          Bt = S2, Ct = S3
  for (int q = threadIdx.x; q < N3; q += blockDim.x) {
    const int i = q % NL, bc = q / NL;
    T t = T(0);
    for (int x = 0; x < NL; x++) t += S1t[i + NL * x] * a[x + NL * bc];
    b[q] = t;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < N3; q += blockDim.x) {
    const int i = q % NL, j = (q / NL) % NL, c = q / N2;
    T t = T(0);
    for (int y = 0; y < NL; y++) t += b[i + NL * (y + NL * c)] * S2[y + NL * j];
    a[q] = t;
  }
  __syncthreads();
  const T *de = D + el * N3;
  for (int q = threadIdx.x; q < N3; q += blockDim.x) {
    const int ij = q % N2, l = q / N2;
    T t = T(0);
    for (int z = 0; z < NL; z++) t += a[ij + N2 * z] * S3[z + NL * l];
    b[q] = de[q] * t;
  }
  __syncthreads();
  // backward: A = S1, Bt = S2^T, Ct = S3^T
  for (int q = threadIdx.x; q < N3; q += blockDim.x) {
    const int i = q % NL, bc = q / NL;
    T t = T(0);
    for (int x = 0; x < NL; x++) t += S1[i + NL * x] * b[x + NL * bc];
    a[q] = t;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < N3; q += blockDim.x) {
    const int i = q % NL, j = (q / NL) % NL, c = q / N2;
    T t = T(0);
    for (int y = 0; y < NL; y++) t += a[i + NL * (y + NL * c)] * S2t[y + NL * j];
    b[q] = t;
  }
  __syncthreads();
  T *ee = e + el * N3;
  for (int q = threadIdx.x; q < N3; q += blockDim.x) {
    const int ij = q % N2, l = q / N2;
    T t = T(0);
    for (int z = 0; z < NL; z++) t += b[ij + N2 * z] * S3t[z + NL * l];
    ee[q] = t;
    if (FUSED) a[q] = t;
  }
  if (FUSED) {
    // keep this element's own face layer 0 (to subtract after the gs)
    __syncthreads();
    for (int q = threadIdx.x; q < NF; q += blockDim.x) {
      const int f = q / (W * W), u = 1 + (q % (W * W)) % W, v = 1 + (q % (W * W)) / W;
      const int p0 = face_point<NL>(f, u, v, 0);
      re[p0] = a[p0];
    }
  }
}

// The same FDM solve with one thread per grid line (NL^2 threads per
// element): a thread holds its line in registers and contracts it with the
// element's 1-D matrices, whose entries are wave-uniform (broadcast LDS
// reads; as scalar loads the compiler serialised them on scalar-memory
// latency), so the data moves through LDS only for the line transposes
// between directions.  The forward
// x, y, z sweeps are followed by the backward z, y, x sweeps (the three
// directions commute), so the z-lines of the forward sweep stay in
// registers through the D scaling and the first backward contraction.
// Used for the FP32 smoothers (48^3, lx 8: 0.93 -> 0.83 ms per solve); in
// FP64 the lines need 112 VGPRs and the strided 8-byte transposes meet
// more bank conflicts, and it timed 5 % slower than fdm_kernel, which
// FP64 keeps.
template <int NL, bool FUSED, typename T>
__global__ void __launch_bounds__(NL *NL)
    fdm_line_kernel(T *__restrict__ e, T *r, const T *__restrict__ S,
                    const T *__restrict__ D, int64_t nel) {
  constexpr int N3 = NL * NL * NL, N2 = NL * NL, W = NL - 2, NF = 6 * W * W;
  __shared__ T a[N3], b[N3];
  __shared__ T sm[6 * N2];
  const int64_t el = blockIdx.x;
  if (el >= nel) return;
  const int L = threadIdx.x;
  T *re = r + el * N3;
  const T *se = S + el * 6 * N2;
  for (int q = L; q < 6 * N2; q += N2) sm[q] = se[q];
  const T *S1 = sm, *S1t = sm + N2, *S2 = sm + 2 * N2, *S2t = sm + 3 * N2, *S3 = sm + 4 * N2,
          *S3t = sm + 5 * N2;
  for (int q = L; q < N3; q += N2) a[q] = re[q];
  __syncthreads();
  if (FUSED) {
    // the neighbour's overlap only: face layer 0 -= layer 2 (face interiors)
    for (int q = L; q < NF; q += N2) {
      const int f = q / (W * W), u = 1 + (q % (W * W)) % W, v = 1 + (q % (W * W)) / W;
      const int p0 = face_point<NL>(f, u, v, 0), p2 = face_point<NL>(f, u, v, 2);
      a[p0] = a[p0] - a[p2];
    }
    __syncthreads();
  }
  T v[NL], w[NL];
  // forward x: line (y, z) = L; b(i, y, z) = sum_x S1t(i + NL x) a(x, y, z)
#pragma unroll
  for (int x = 0; x < NL; x++) v[x] = a[x + NL * L];
#pragma unroll
  for (int i = 0; i < NL; i++) {
    T t = T(0);
#pragma unroll
    for (int x = 0; x < NL; x++) t += S1t[i + NL * x] * v[x];
    w[i] = t;
  }
#pragma unroll
  for (int i = 0; i < NL; i++) b[i + NL * L] = w[i];
  __syncthreads();
  const int li = L % NL, lz = L / NL;
  // forward y: line (i, z); a(i, j, z) = sum_y b(i, y, z) S2(y + NL j)
#pragma unroll
  for (int y = 0; y < NL; y++) v[y] = b[li + NL * (y + NL * lz)];
#pragma unroll
  for (int j = 0; j < NL; j++) {
    T t = T(0);
#pragma unroll
    for (int y = 0; y < NL; y++) t += v[y] * S2[y + NL * j];
    w[j] = t;
  }
#pragma unroll
  for (int j = 0; j < NL; j++) a[li + NL * (j + NL * lz)] = w[j];
  __syncthreads();
  // forward z: line (i, j) = L, scaled by D; then backward z with S3t
  const T *de = D + el * N3;
#pragma unroll
  for (int z = 0; z < NL; z++) v[z] = a[L + N2 * z];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    T t = T(0);
#pragma unroll
    for (int z = 0; z < NL; z++) t += v[z] * S3[z + NL * l];
    w[l] = de[L + N2 * l] * t;
  }
#pragma unroll
  for (int l = 0; l < NL; l++) {
    T t = T(0);
#pragma unroll
    for (int z = 0; z < NL; z++) t += w[z] * S3t[z + NL * l];
    v[l] = t;
  }
#pragma unroll
  for (int l = 0; l < NL; l++) b[L + N2 * l] = v[l];
  __syncthreads();
  // backward y: line (i, z); a(i, j, z) = sum_y b(i, y, z) S2t(y + NL j)
#pragma unroll
  for (int y = 0; y < NL; y++) v[y] = b[li + NL * (y + NL * lz)];
#pragma unroll
  for (int j = 0; j < NL; j++) {
    T t = T(0);
#pragma unroll
    for (int y = 0; y < NL; y++) t += v[y] * S2t[y + NL * j];
    w[j] = t;
  }
#pragma unroll
  for (int j = 0; j < NL; j++) a[li + NL * (j + NL * lz)] = w[j];
  __syncthreads();
  // backward x: line (y, z) = L; e(i, y, z) = sum_x S1(i + NL x) a(x, y, z)
#pragma unroll
  for (int x = 0; x < NL; x++) v[x] = a[x + NL * L];
  T *ee = e + el * N3;
  const int ly = L % NL;
  const bool yb = ly == 0 || ly == NL - 1, zb = lz == 0 || lz == NL - 1;
#pragma unroll
  for (int i = 0; i < NL; i++) {
    T t = T(0);
#pragma unroll
    for (int x = 0; x < NL; x++) t += S1[i + NL * x] * v[x];
    ee[i + NL * L] = t;
    if (FUSED) {
      // this element's own face layer 0 goes back into r (subtracted after
      // the gs): the face-interior points at depth 0 on this line
      const bool xb = i == 0 || i == NL - 1;
      if ((int)xb + (int)yb + (int)zb == 1) re[i + NL * L] = t;
    }
  }
}

// ext (NL^3, zero) <- interior r (lx^3) plus face layer 0 = interior layer
// 2 (schwarz_toext3d + extrude(.., 0, 0, .., 2, 1)).  One workgroup per
// element: r staged in LDS by coalesced loads, the extended element
// written coalesced, one thread per extended point.
// With SUB the interior is f - r instead of r (the Chebyshev residual
// formed on load).
template <typename T, int n, bool SUB>
__global__ void __launch_bounds__(256)
    toext_el_kernel(T *__restrict__ a, const T *__restrict__ r, const T *__restrict__ f,
                    int64_t nel) {
  constexpr int m = n + 2, n3 = n * n * n, m3 = m * m * m;
  constexpr int NR = (n3 + 255) / 256;
  __shared__ T sr[n3];
  const int64_t e = blockIdx.x;
  if (e >= nel) return;
  // every load of the element issued before the first LDS store
  T ld[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const int q = threadIdx.x + 256 * k;
    ld[k] = q < n3 ? (SUB ? f[e * n3 + q] - r[e * n3 + q] : r[e * n3 + q]) : T(0);
  }
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const int q = threadIdx.x + 256 * k;
    if (q < n3) sr[q] = ld[k];
  }
  __syncthreads();
  for (int p = threadIdx.x; p < m3; p += 256) {
    int c[3] = {p % m, (p / m) % m, p / (m * m)};
    int nb = 0, dir = -1;
#pragma unroll
    for (int d = 0; d < 3; d++)
      if (c[d] == 0 || c[d] == m - 1) {
        nb++;
        dir = d;
      }
    T v = T(0);
    if (nb <= 1) {
      if (nb == 1) c[dir] = c[dir] == 0 ? 2 : m - 3;
      v = sr[(c[0] - 1) + n * ((c[1] - 1) + n * (c[2] - 1))];
    }
    a[e * m3 + p] = v;
  }
}

// e (lx^3) <- interior of work2 after the overlap exchange: layer 2 of
// every face += (work2 - work1) at layer 0 (extrude(w2, 0, 1, w1, 0, -1)
// then extrude(w2, 2, 1, w2, 0, 1), r faces then s then t, low before
// high), then schwarz_toreg3d.  One workgroup per element: the
// extended element of work2 and the face interiors of work1 staged in LDS
// by coalesced loads, the element written coalesced.
template <typename T, int n>
__global__ void __launch_bounds__(256)
    toreg_el_kernel(T *__restrict__ e, const T *__restrict__ w2, const T *__restrict__ w1,
                    int64_t nel) {
  constexpr int m = n + 2, n3 = n * n * n, m3 = m * m * m;
  __shared__ T sa[m3], sb[m3];
  const int64_t el = blockIdx.x;
  if (el >= nel) return;
  const T *A = w2 + el * m3, *B = w1 + el * m3;
  // every load of the element issued before the first LDS store
  constexpr int NR = (m3 + 255) / 256;
  T la[NR], lb[NR];
#pragma unroll
  for (int q = 0; q < NR; q++) {
    const int p = threadIdx.x + 256 * q;
    const int i = p % m, j = (p / m) % m, k = p / (m * m);
    const int nb = (i == 0 || i == m - 1) + (j == 0 || j == m - 1) + (k == 0 || k == m - 1);
    la[q] = p < m3 ? A[p] : T(0);
    lb[q] = (p < m3 && nb == 1) ? B[p] : T(0);
  }
#pragma unroll
  for (int q = 0; q < NR; q++) {
    const int p = threadIdx.x + 256 * q;
    if (p < m3) {
      sa[p] = la[q];
      sb[p] = lb[q];
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n3; q += 256) {
    const int I = 1 + q % n, J = 1 + (q / n) % n, K = 1 + q / (n * n);
    T v = sa[I + m * (J + m * K)];
    if (I == 2) v += sa[0 + m * (J + m * K)] - sb[0 + m * (J + m * K)];
    if (I == m - 3) v += sa[(m - 1) + m * (J + m * K)] - sb[(m - 1) + m * (J + m * K)];
    if (J == 2) v += sa[I + m * (0 + m * K)] - sb[I + m * (0 + m * K)];
    if (J == m - 3) v += sa[I + m * ((m - 1) + m * K)] - sb[I + m * ((m - 1) + m * K)];
    if (K == 2) v += sa[I + m * (J + m * 0)] - sb[I + m * (J + m * 0)];
    if (K == m - 3) v += sa[I + m * (J + m * (m - 1))] - sb[I + m * (J + m * (m - 1))];
    e[el * n3 + q] = v;
  }
}

// schwarz_wt3d (schwarz.f90:597-634): weights on the two outer layers of
// every face.  The weights a point takes (r-faces for all (j,k); s-faces
// for i in 3..n-2 (1-based); t-faces for i,j in 3..n-2; a point on several
// layers takes every weight, in the reference's order) are multiplied
// into one per-point factor at set-up, so the application is a coalesced
// e *= wtp.
__global__ void wtp_setup_kernel(double *__restrict__ wtp, const double *__restrict__ wt,
                                 int n, int64_t nel) {
  const int64_t n3 = (int64_t)n * n * n;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nel * n3) return;
  const int64_t el = idx / n3;
  const int p = (int)(idx % n3);
  const int i = p % n, j = (p / n) % n, k = p / (n * n);
  const double *we = wt + el * (int64_t)n * n * 12;
  auto W = [&](int a, int b, int c, int d) { return we[a + n * (b + n * (c + 4 * d))]; };
  double v = 1.0;
  auto apply = [&](int x, int a, int b, int dir) {
    if (x == 0) v *= W(a, b, 0, dir);
    if (x == 1) v *= W(a, b, 1, dir);
    if (x == n - 2) v *= W(a, b, 2, dir);
    if (x == n - 1) v *= W(a, b, 3, dir);
  };
  apply(i, j, k, 0);
  const bool mid_i = (i >= 2 && i <= n - 3), mid_j = (j >= 2 && j <= n - 3);
  if (mid_i) apply(j, i, k, 1);
  if (mid_i && mid_j) apply(k, i, j, 2);
  wtp[idx] = v;
}

template <typename T>
__global__ void wtp_mul_kernel(T *__restrict__ e, const T *__restrict__ wtp, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    e[i] *= __builtin_nontemporal_load(wtp + i);
}

// wt from the summed overlap counts (schwarz_setup_schwarz_wt3d_2)
__global__ void wt_setup_kernel(double *__restrict__ wt, const double *__restrict__ c,
                                int n, int64_t nel) {
  const int64_t per = (int64_t)n * n * 12;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nel * per) return;
  const int64_t el = idx / per;
  const int q = (int)(idx % per);
  const int a = q % n, b = (q / n) % n, lay = (q / (n * n)) % 4, dir = q / (n * n * 4);
  const int pos = lay == 0 ? 0 : (lay == 1 ? 1 : (lay == 2 ? n - 2 : n - 1));
  const double *ce = c + el * (int64_t)n * n * n;
  double v;
  if (dir == 0) v = ce[i3(pos, a, b, n)];
  else if (dir == 1) v = ce[i3(a, pos, b, n)];
  else v = ce[i3(a, b, pos, n)];
  wt[idx] = 1.0 / v;
}

static inline unsigned blocks_for(int64_t n, int bs = 256) {
  return (unsigned)((n + bs - 1) / bs);
}

// --------------------------------------------------------------- structures
struct GsDev {
  int32_t *blk_ptr = nullptr, *gidx = nullptr;
  int64_t nblks = 0, pair0 = 0, npairs = 0;
  int max_blk = 0;
  // the facet pairs as whole-face descriptors (gs variant 2) when every
  // pair lies on one, with the element size lx of this layout
  int32_t *faces = nullptr;
  int64_t nfaces = 0;
  int lx = 0;
  // every block entry lies on two or more element boundaries (an edge or
  // vertex of the element): where a field is zero on those points, as the
  // Schwarz extension leaves it, the block sums are zeros and can be skipped
  bool blocks_on_edges = false;
};

struct Schwarz {
  int lx = 0, nl = 0;
  int64_t nel = 0;
  const Op *op = nullptr;
  GsDev ext;
  double *S = nullptr, *D = nullptr, *wt = nullptr;  // wt: per-point weight product
  double *work1 = nullptr, *work2 = nullptr;
  // FP32 copies for the FP32 smoothers (allocated on request)
  float *S32 = nullptr, *D32 = nullptr, *wt32 = nullptr;
  float *work1_32 = nullptr, *work2_32 = nullptr;
  template <typename T>
  void arrays(const T *&s, const T *&d, const T *&w, T *&w1, T *&w2) const {
    if constexpr (std::is_same<T, double>::value) {
      s = S; d = D; w = wt; w1 = work1; w2 = work2;
    } else {
      s = S32; d = D32; w = wt32; w1 = work1_32; w2 = work2_32;
    }
  }
};

struct Hsmg {
  Op *fine = nullptr;
  nm_op_t *mid_h = nullptr, *crs_h = nullptr;
  Op *mid = nullptr, *crs = nullptr;
  Schwarz sw_f, sw_m;
  double *jfm = nullptr, *jfm_t = nullptr, *jmc = nullptr, *jmc_t = nullptr;
  double *r = nullptr, *w = nullptr, *e_mid = nullptr, *e_crs = nullptr, *r_crs = nullptr;
  CgFixed *cg = nullptr;
  int niter = 10;
  double crs_tol = 1e-9;
  std::vector<void *> owned;
  ~Hsmg() {
    if (cg) cg_fixed_free(cg);
    for (void *q : owned) (void)hipFree(q);
    if (mid_h) nm_op_free(mid_h);
    if (crs_h) nm_op_free(crs_h);
  }
  template <typename T>
  T *alloc(int64_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess) return nullptr;
    owned.push_back(q);
    return (T *)q;
  }
};

template <typename T>
static int gs_dev(const GsDev &g, T *u, hipStream_t s, bool blocks = true) {
  GsArgs<T> a{u, g.blk_ptr, g.gidx, blocks ? g.nblks : 0, g.pair0, g.npairs, nullptr};
  // the block length limit of the lane-parallel kernels matters only when
  // the blocks run
  a.variant = (!blocks || g.max_blk <= 8) ? (g.faces ? 2 : 1) : 0;
  a.faces = g.faces;
  a.nfaces = g.nfaces;
  a.lx = g.lx;
  return launch_gs<T>(a, s);
}

static int extrude(double *a1, int l1, double f1, const double *a2, int l2,
                   double f2, int nx, int64_t nel, hipStream_t s) {
  hipLaunchKernelGGL(extrude_kernel, dim3((unsigned)nel), dim3(128), 0, s, a1, l1,
                     f1, a2, l2, f2, nx, nel);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

static int tnsr3d(double *v, int nv, const double *u, int nu, const double *A,
                  const double *Bt, const double *Ct, int64_t nel, hipStream_t s) {
  // LDS sized for the larger of the two extents: at 14^3 a workgroup holds
  // 49 KB (three per CU), at 8^3 10 KB
  switch (std::max(nv, nu)) {
#define NM_TNSR(N)                                                                     \
  case N:                                                                              \
    hipLaunchKernelGGL((tnsr3d_kernel<N>), dim3((unsigned)nel), dim3(256), 0, s, v, nv, \
                       u, nu, A, Bt, Ct, (int64_t)0, nel);                             \
    break;
    NM_TNSR(2) NM_TNSR(3) NM_TNSR(4) NM_TNSR(5) NM_TNSR(6) NM_TNSR(7) NM_TNSR(8)
    NM_TNSR(9) NM_TNSR(10)
#undef NM_TNSR
    case 11: case 12: case 13: case 14:
      hipLaunchKernelGGL((tnsr3d_kernel<14>), dim3((unsigned)nel), dim3(256), 0, s, v, nv,
                         u, nu, A, Bt, Ct, (int64_t)0, nel);
      break;
    default:
      set_error("tnsr3d: size %d unsupported", std::max(nv, nu));
      return NM_ERR_UNSUPPORTED;
  }
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// The same FDM solve on the matrix cores: one wave per element, each 1-D
// transform an (NL x NL) by (NL x NL^2) product in 16x16x4 tiles
// (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64): the tile's rows are the
// transform's output index (NL of 16 used), its four k the summed index
// (zero-padded to a multiple of 4), its 16 columns 16 of the NL^2 lines.
// A lane holds A(row = lane % 16, k = 4 kk + lane / 16) and B(k, col =
// lane % 16); the 4 accumulator entries of lane l are rows 4 (l / 16) + v
// (FP32) or l / 16 + 4 v (FP64) of column l % 16 (checked on the device).
// The block moves between the LDS arrays a and b once per direction; the
// forward x, y, z, the D scaling and the backward z, y, x run in that
// order, as in fdm_line_kernel.
template <typename T> struct MfmaTile;
template <> struct MfmaTile<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t fma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct MfmaTile<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t fma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

// index of (transform index k, line col) in the NL^3 block for direction
// DIR (0: x, lines (y,z); 1: y, lines (x,z); 2: z, lines (x,y))
template <int NL, int DIR>
__device__ __forceinline__ int fdm_at(int k, int col) {
  if (DIR == 0) return k + NL * col;
  if (DIR == 1) return col % NL + NL * k + NL * NL * (col / NL);
  return col + NL * NL * k;
}

// out(row, col) = sum_k A(row, k) in(k, col), A(row, k) = M[row + NL k]
// (TRANS false) or M[k + NL row], over this wave's tiles (tile tt = wave,
// wave + WPE, ...); SCALE multiplies by the D values prefetched in dsc
// (per tile of the wave, per accumulator entry).  All of a tile's B
// operands are read before its matrix instructions, so the LDS reads of
// one tile overlap the products of the previous one.
template <int NL, int WPE>
struct FdmTiles {
  static constexpr int N2 = NL * NL, KS = (NL + 3) / 4, NT = (N2 + 15) / 16,
                       TPW = (NT + WPE - 1) / WPE;
};

template <int NL, int WPE, int DIR, bool TRANS, bool SCALE, typename T>
__device__ __forceinline__ void fdm_mfma_pass(T *__restrict__ out, const T *__restrict__ in,
                                              const T *__restrict__ M,
                                              const T (*dsc)[4], int lane, int wave) {
  using TL = FdmTiles<NL, WPE>;
  constexpr int N2 = TL::N2, KS = TL::KS, NT = TL::NT, TPW = TL::TPW;
  using acc_t = typename MfmaTile<T>::acc_t;
  const int lr = lane & 15, lk = lane >> 4;
  T am[KS];
#pragma unroll
  for (int kk = 0; kk < KS; kk++) {
    const int k = 4 * kk + lk;
    am[kk] = (lr < NL && k < NL) ? (TRANS ? M[k + NL * lr] : M[lr + NL * k]) : T(0);
  }
#pragma unroll
  for (int m = 0; m < TPW; m++) {
    const int tt = wave + WPE * m;
    if (tt >= NT) break;
    const int col = 16 * tt + lr;
    T bv[KS];
#pragma unroll
    for (int kk = 0; kk < KS; kk++) {
      const int k = 4 * kk + lk;
      bv[kk] = (k < NL && col < N2) ? in[fdm_at<NL, DIR>(k, col)] : T(0);
    }
    acc_t acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int kk = 0; kk < KS; kk++) acc = MfmaTile<T>::fma(am[kk], bv[kk], acc);
    if (col < N2) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = MfmaTile<T>::row(lane, v);
        if (row < NL) out[fdm_at<NL, DIR>(row, col)] = SCALE ? dsc[m][v] * acc[v] : acc[v];
      }
    }
  }
}

// WPE waves per element (the tiles of every transform dealt out over them)
template <int NL, int WPE, bool FUSED, typename T>
__global__ void __launch_bounds__(64 * WPE)
    fdm_mfma_kernel(T *__restrict__ e, T *r, const T *__restrict__ S, const T *__restrict__ D,
                    int64_t nel) {
  static_assert(NL <= 16, "one 16-row tile per transform");
  using TL = FdmTiles<NL, WPE>;
  constexpr int N3 = NL * NL * NL, N2 = NL * NL, W = NL - 2, NF = 6 * W * W;
  constexpr int NTH = 64 * WPE, TPW = TL::TPW, NT = TL::NT;
  __shared__ T a[N3], b[N3];
  __shared__ T sm[6 * N2];
  const int64_t el = blockIdx.x;
  if (el >= nel) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  T *re = r + el * N3;
  const T *se = S + el * 6 * N2;
  // this lane's D values of the z-forward pass (its output points there)
  T dsc[TPW][4];
  {
    const T *de = D + el * N3;
    const int lr = lane & 15;
#pragma unroll
    for (int m = 0; m < TPW; m++) {
      const int tt = wave + WPE * m, col = 16 * tt + lr;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = MfmaTile<T>::row(lane, v);
        dsc[m][v] = (tt < NT && col < N2 && row < NL) ? de[fdm_at<NL, 2>(row, col)] : T(0);
      }
    }
  }
  {
    // every load of the element issued before the first LDS store
    constexpr int NS = (6 * N2 + NTH - 1) / NTH, NA = (N3 + NTH - 1) / NTH;
    T ls[NS], la[NA];
#pragma unroll
    for (int k = 0; k < NS; k++) ls[k] = tid + NTH * k < 6 * N2 ? se[tid + NTH * k] : T(0);
#pragma unroll
    for (int k = 0; k < NA; k++) la[k] = tid + NTH * k < N3 ? re[tid + NTH * k] : T(0);
#pragma unroll
    for (int k = 0; k < NS; k++)
      if (tid + NTH * k < 6 * N2) sm[tid + NTH * k] = ls[k];
#pragma unroll
    for (int k = 0; k < NA; k++)
      if (tid + NTH * k < N3) a[tid + NTH * k] = la[k];
  }
  __syncthreads();
  if (FUSED) {
    // the neighbour's overlap only: face layer 0 -= layer 2 (face interiors)
    for (int q = tid; q < NF; q += NTH) {
      const int f = q / (W * W), u = 1 + (q % (W * W)) % W, v = 1 + (q % (W * W)) / W;
      const int p0 = face_point<NL>(f, u, v, 0), p2 = face_point<NL>(f, u, v, 2);
      a[p0] = a[p0] - a[p2];
    }
    __syncthreads();
  }
  const T *S1 = sm, *S1t = sm + N2, *S2 = sm + 2 * N2, *S2t = sm + 3 * N2, *S3 = sm + 4 * N2,
          *S3t = sm + 5 * N2;
  fdm_mfma_pass<NL, WPE, 0, false, false>(b, a, S1t, dsc, lane, wave);
  __syncthreads();
  fdm_mfma_pass<NL, WPE, 1, true, false>(a, b, S2, dsc, lane, wave);
  __syncthreads();
  fdm_mfma_pass<NL, WPE, 2, true, true>(b, a, S3, dsc, lane, wave);
  __syncthreads();
  fdm_mfma_pass<NL, WPE, 2, true, false>(a, b, S3t, dsc, lane, wave);
  __syncthreads();
  fdm_mfma_pass<NL, WPE, 1, true, false>(b, a, S2t, dsc, lane, wave);
  __syncthreads();
  fdm_mfma_pass<NL, WPE, 0, false, false>(a, b, S1, dsc, lane, wave);
  __syncthreads();
  T *ee = e + el * N3;
  for (int q = tid; q < N3; q += NTH) ee[q] = a[q];
  if (FUSED) {
    // this element's own face layer 0 goes back into r (subtracted after
    // the gs)
    for (int q = tid; q < NF; q += NTH) {
      const int f = q / (W * W), u = 1 + (q % (W * W)) % W, v = 1 + (q % (W * W)) / W;
      const int p0 = face_point<NL>(f, u, v, 0);
      re[p0] = a[p0];
    }
  }
}

// FDM kernel family: 0 matrix cores where they are faster (fdm_mfma_kernel),
// 1 vector ALUs everywhere (the line kernel in FP32, the plane kernel in
// FP64); process-wide
static int g_fdm_variant = 0;
constexpr int FDM_WPE = 4;  // waves per element of the matrix-core kernel

// fused: r is the extended work array after its gs (face layer 0 is
// corrected on load and receives e's layer 0 afterwards)
template <typename T>
static int fdm(T *e, T *r, const T *S, const T *D, int nl, int64_t nel, hipStream_t s,
               bool fused = false) {
  // matrix cores for FP64 at every size and for FP32 from NL = 10 (timed on
  // 48^3: NL 10 FP64 1.49 -> 0.97 ms, FP32 0.82 -> 0.65 ms; NL 6 FP64 0.29
  // -> 0.27 ms, but FP32 0.18 -> 0.23 ms, where the line kernel stays)
  if (g_fdm_variant == 0 && (std::is_same<T, double>::value || nl >= 10)) {
    // FP32: two waves per element (a workgroup's LDS is half the FP64 one,
    // so twice as many elements fit on a CU)
    constexpr int WP = std::is_same<T, float>::value ? 2 : FDM_WPE;
    switch (nl) {
#define NM_FDMM(N)                                                                       \
  case N:                                                                                \
    if (fused)                                                                           \
      hipLaunchKernelGGL((fdm_mfma_kernel<N, WP, true, T>), dim3((unsigned)nel),         \
                         dim3(64 * WP), 0, s, e, r, S, D, nel);                          \
    else                                                                                 \
      hipLaunchKernelGGL((fdm_mfma_kernel<N, WP, false, T>), dim3((unsigned)nel),        \
                         dim3(64 * WP), 0, s, e, r, S, D, nel);                          \
    NM_HIP_TRY(hipGetLastError());                                                       \
    return NM_OK;
      NM_FDMM(5) NM_FDMM(6) NM_FDMM(7) NM_FDMM(8) NM_FDMM(9) NM_FDMM(10) NM_FDMM(11)
      NM_FDMM(12)
#undef NM_FDMM
      default:
        set_error("fdm: extended size %d unsupported (lx 3..10)", nl);
        return NM_ERR_UNSUPPORTED;
    }
  }
  switch (nl) {
#define NM_FDM(N)                                                                       \
  case N:                                                                               \
    if (std::is_same<T, float>::value) {                                                \
      if (fused)                                                                        \
        hipLaunchKernelGGL((fdm_line_kernel<N, true, T>), dim3((unsigned)nel),          \
                           dim3(N * N), 0, s, e, r, S, D, nel);                         \
      else                                                                              \
        hipLaunchKernelGGL((fdm_line_kernel<N, false, T>), dim3((unsigned)nel),         \
                           dim3(N * N), 0, s, e, r, S, D, nel);                         \
    } else if (fused) {                                                                 \
      hipLaunchKernelGGL((fdm_kernel<N, true, T>), dim3((unsigned)nel), dim3(256), 0, s,  \
                         e, r, S, D, nel);                                              \
    } else {                                                                            \
      hipLaunchKernelGGL((fdm_kernel<N, false, T>), dim3((unsigned)nel), dim3(256), 0, s, \
                         e, r, S, D, nel);                                              \
    }                                                                                   \
    break;
    NM_FDM(5) NM_FDM(6) NM_FDM(7) NM_FDM(8) NM_FDM(9) NM_FDM(10) NM_FDM(11) NM_FDM(12)
#undef NM_FDM
    default:
      set_error("fdm: extended size %d unsupported (lx 3..10)", nl);
      return NM_ERR_UNSUPPORTED;
  }
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// w[msk] = f[msk]: f - w is then +0 at every masked point, as masking
// f - w after the subtraction leaves it
template <typename T>
__global__ void mask_copy_kernel(T *__restrict__ w, const T *__restrict__ f,
                                 const int32_t *__restrict__ msk, int64_t m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) w[msk[i]] = f[msk[i]];
}

// schwarz_compute (CPU branch semantics, schwarz.f90:565-595); r masked in
// place as in the reference.  weight = false leaves out the final multiply
// by the Schwarz weights, for a caller that applies them in its own next
// pass (the Chebyshev step).  With f the input residual is mask(f - r),
// formed inside the extension (r is overwritten at the masked points).
template <typename T>
static int schwarz_apply(const Schwarz &sw, T *e, T *r, hipStream_t s, bool weight = true,
                         const T *f = nullptr) {
  const Op &op = *sw.op;
  const int nx = sw.nl, lx = sw.lx;
  const int64_t nel = sw.nel, ns = nel * nx * nx * nx;
  const T *S, *D, *wt;
  T *w1, *w2;
  sw.arrays<T>(S, D, wt, w1, w2);
  if (!S) {
    set_error("schwarz: FP32 arrays not set up");
    return NM_ERR_STATE;
  }
  int st;
  // the reference's sequence mask, toext3d, extrude x4, fdm, toreg3d runs
  // here as three fused passes around the two extended-layout gs
  if (f) {
    if (op.nmsk > 0)
      hipLaunchKernelGGL(mask_copy_kernel<T>, dim3((unsigned)((op.nmsk + 255) / 256)),
                         dim3(256), 0, s, r, f, op.msk, op.nmsk);
  } else if ((st = launch_mask<T>(r, op.msk, op.nmsk, nullptr, s))) {
    return st;
  }
  // the extents are template constants, so the point index splits by
  // multiplications instead of 64-bit divisions
  switch (lx) {
#define NM_TOEXT(N)                                                                      \
  case N:                                                                                \
    if (f)                                                                               \
      hipLaunchKernelGGL((toext_el_kernel<T, N, true>), dim3((unsigned)nel), dim3(256),  \
                         0, s, w1, r, f, nel);                                           \
    else                                                                                 \
      hipLaunchKernelGGL((toext_el_kernel<T, N, false>), dim3((unsigned)nel), dim3(256), \
                         0, s, w1, r, nullptr, nel);                                     \
    break;
    NM_TOEXT(3) NM_TOEXT(4) NM_TOEXT(5) NM_TOEXT(6) NM_TOEXT(7) NM_TOEXT(8) NM_TOEXT(9)
    NM_TOEXT(10)
#undef NM_TOEXT
    default:
      set_error("schwarz: lx %d unsupported (3..10)", lx);
      return NM_ERR_UNSUPPORTED;
  }
  // the extension leaves the edges and vertices of every extended element
  // zero, so the first exchange's block sums are zeros: faces only
  if ((st = gs_dev<T>(sw.ext, w1, s, !sw.ext.blocks_on_edges))) return st;
  if ((st = fdm<T>(w2, w1, S, D, nx, nel, s, true))) return st;
  // the restriction below reads the extended face interiors and the element
  // interior only, never an extended edge or vertex: faces only again
  if ((st = gs_dev<T>(sw.ext, w2, s, !sw.ext.blocks_on_edges))) return st;
  switch (lx) {
#define NM_TOREG(N)                                                                      \
  case N:                                                                                \
    hipLaunchKernelGGL((toreg_el_kernel<T, N>), dim3((unsigned)nel), dim3(256), 0, s,   \
                       e, w2, w1, nel);                                                  \
    break;
    NM_TOREG(3) NM_TOREG(4) NM_TOREG(5) NM_TOREG(6) NM_TOREG(7) NM_TOREG(8) NM_TOREG(9)
    NM_TOREG(10)
#undef NM_TOREG
    default:
      set_error("schwarz: lx %d unsupported (3..10)", lx);
      return NM_ERR_UNSUPPORTED;
  }
  if ((st = op.gs<T>(e, nullptr, s))) return st;
  if ((st = launch_mask<T>(e, op.msk, op.nmsk, nullptr, s))) return st;
  if (weight)
    hipLaunchKernelGGL(wtp_mul_kernel<T>,
                       dim3((unsigned)std::min<int64_t>(blocks_for(op.n), 16384)), dim3(256), 0,
                       s, e, wt, op.n);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// host gs ADD with Neko's 1-based maps (setup only)
struct HostGs {
  const Op *op;
};
static void host_gs(void *ctx, double *u) {
  const Op &op = *((HostGs *)ctx)->op;
  const int32_t *dg = op.h_dg.data(), *gd = op.h_gd.data();
  const int64_t nb = (int64_t)op.h_blk_len.size(), m = (int64_t)op.h_dg.size();
  for (int64_t b = 0; b < nb; b++) {
    const int32_t k = op.h_blk_off[b], len = op.h_blk_len[b];
    double t = u[gd[k] - 1];
    for (int q = 1; q < len; q++) t += u[gd[k + q] - 1];
    for (int q = 0; q < len; q++) u[gd[k + q] - 1] = t;
  }
  for (int64_t i = op.facet_offset; i + 1 <= m; i += 2) {
    const double t = u[gd[i - 1] - 1] + u[gd[i] - 1];
    u[gd[i - 1] - 1] = t;
    u[gd[i] - 1] = t;
  }
  (void)dg;
}

// Level operators and the Schwarz extended gs on the fine operator's mesh:
// a box (nm_op_create_box) or a general mesh (nm_op_create_mesh, whose
// topology the operator keeps).
static bool has_topology(const Op *op) { return op->box[0] != 0 || !op->h_pid.empty(); }

static int level_op(const Op *fine, int lx, double h1, double h2, nm_op_t **out,
                    hipStream_t s) {
  if (fine->box[0] != 0) {
    const int *b = fine->box, *p = fine->per;
    return nm_op_create_box(out, b[0], b[1], b[2], lx, p[0], p[1], p[2], fine->amp,
                            fine->seed, h1, h2, s);
  }
  return nm_op_create_mesh_curved(out, fine->nel, lx, fine->h_vxyz.data(), fine->h_pid.data(),
                                  (int64_t)fine->h_cel.size(), fine->h_cel.data(),
                                  fine->h_cdata.data(), fine->h_ctype.data(), h1, h2, s);
}

// 1 where element facet e*6+f has a neighbour (box lattice or mesh)
static std::vector<char> facet_has_neighbour(const Op *op) {
  const int64_t nel = op->nel;
  std::vector<char> nb(nel * 6, 0);
  if (op->box[0] == 0) {
    for (int64_t q = 0; q < nel * 6; q++) nb[q] = op->h_fneigh[q] != 0;
    return nb;
  }
  const int64_t nx_ = op->box[0], ny_ = op->box[1], nz_ = op->box[2];
  for (int64_t e = 0; e < nel; e++) {
    const int64_t ex = e % nx_, ey = (e / nx_) % ny_, ez = e / (nx_ * ny_);
    const bool f[6] = {op->per[0] || ex > 0, op->per[0] || ex < nx_ - 1,
                       op->per[1] || ey > 0, op->per[1] || ey < ny_ - 1,
                       op->per[2] || ez > 0, op->per[2] || ez < nz_ - 1};
    for (int q = 0; q < 6; q++) nb[e * 6 + q] = f[q];
  }
  return nb;
}

// element facet neighbours (msh%facet_neigh: 6 per element, 1-based, 0 on
// the boundary) of a box lattice or a mesh operator
static std::vector<int32_t> facet_neighbours(const Op *op) {
  const int64_t nel = op->nel;
  if (op->box[0] == 0) return op->h_fneigh;
  std::vector<int32_t> nb(nel * 6, 0);
  const int64_t nx_ = op->box[0], ny_ = op->box[1], nz_ = op->box[2];
  for (int64_t e = 0; e < nel; e++) {
    const int64_t c[3] = {e % nx_, (e / nx_) % ny_, e / (nx_ * ny_)};
    const int64_t dim[3] = {nx_, ny_, nz_}, stride[3] = {1, nx_, nx_ * ny_};
    for (int d = 0; d < 3; d++)
      for (int side = 0; side < 2; side++) {
        int64_t cc = c[d] + (side ? 1 : -1);
        if (cc < 0 || cc >= dim[d]) {
          if (!op->per[d]) continue;
          cc = (cc + dim[d]) % dim[d];
        }
        nb[e * 6 + 2 * d + side] = (int32_t)(e + (cc - c[d]) * stride[d] + 1);
      }
  }
  return nb;
}

template <typename H>
static int gsdev_level(H *own, GsDev *g, const Op *op, int lx, hipStream_t s) {
  int64_t nl, fo, nb;
  const int *box = op->box, *per = op->per;
  int st;
  std::vector<int32_t> dg, gd, bl, bo;
  if (box[0] != 0) {
    if ((st = nm_box_gs_sizes(box[0], box[1], box[2], lx, per[0], per[1], per[2], &nl, &fo)))
      return st;
    dg.resize(nl); gd.resize(nl); bl.resize(fo); bo.resize(fo);
    st = nm_box_gs_map(box[0], box[1], box[2], lx, per[0], per[1], per[2], dg.data(),
                       gd.data(), bl.data(), bo.data(), &nl, &fo, &nb);
  } else {
    if ((st = nm_mesh_gs_sizes(op->nel, lx, op->h_fneigh.data(), &nl, &fo))) return st;
    dg.resize(nl); gd.resize(nl); bl.resize(fo); bo.resize(fo);
    st = nm_mesh_gs_map(op->nel, lx, op->h_pid.data(), op->h_fneigh.data(), dg.data(),
                        gd.data(), bl.data(), bo.data(), &nl, &fo, &nb);
  }
  if (st) return st;
  std::vector<int32_t> ptr(nb + 1), gi(nl);
  g->max_blk = 0;
  for (int64_t b = 0; b < nb; b++) {
    ptr[b] = bo[b];
    g->max_blk = std::max<int>(g->max_blk, bl[b]);
  }
  ptr[nb] = nb > 0 ? bo[nb - 1] + bl[nb - 1] : 0;
  for (int64_t i = 0; i < nl; i++) gi[i] = gd[i] - 1;
  g->nblks = nb;
  g->pair0 = fo - 1;
  g->npairs = (nl - (fo - 1)) / 2;
  ptr.resize(nb + 1);
  {
    const int st2 = gs_regroup(op->nel, (int64_t)lx * lx * lx, nb, g->pair0, g->npairs, ptr, gi);
    if (st2) return st2;
  }
  g->blocks_on_edges = true;
  {
    const int64_t n3 = (int64_t)lx * lx * lx;
    for (int64_t q = 0; q < ptr[nb] && g->blocks_on_edges; q++) {
      const int64_t l = gi[q] % n3;
      const int c[3] = {(int)(l % lx), (int)((l / lx) % lx), (int)(l / ((int64_t)lx * lx))};
      int on = 0;
      for (int d = 0; d < 3; d++) on += (c[d] == 0 || c[d] == lx - 1);
      g->blocks_on_edges = on >= 2;
    }
  }
  if (!(g->blk_ptr = own->template alloc<int32_t>(nb + 1)) ||
      !(g->gidx = own->template alloc<int32_t>(nl))) {
    set_error("hsmg: out of device memory");
    return NM_ERR_NOMEM;
  }
  if (lx > 2) {
    // (used also where blocks exceed 8 entries, by exchanges that skip the
    // block sums)
    std::vector<int32_t> faces, rest;
    gs_face_pairs(op->nel, lx, gi.data() + g->pair0, g->npairs, faces, rest);
    if (rest.empty() && !faces.empty()) {
      if (!(g->faces = own->template alloc<int32_t>((int64_t)faces.size()))) {
        set_error("hsmg: out of device memory");
        return NM_ERR_NOMEM;
      }
      g->nfaces = (int64_t)faces.size() / 3;
      g->lx = lx;
      NM_HIP_TRY(hipMemcpyAsync(g->faces, faces.data(), sizeof(int32_t) * faces.size(),
                                hipMemcpyHostToDevice, s));
      NM_HIP_TRY(hipStreamSynchronize(s));
    }
  }
  NM_HIP_TRY(hipMemcpyAsync(g->blk_ptr, ptr.data(), sizeof(int32_t) * (nb + 1), hipMemcpyHostToDevice, s));
  NM_HIP_TRY(hipMemcpyAsync(g->gidx, gi.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  return NM_OK;
}

template <typename H>
static int schwarz_setup(H *h, Schwarz *sw, const Op *op, int facet_bc, hipStream_t s) {
  const int lx = op->lx, nl = lx + 2;
  const int64_t nel = op->nel, n = op->n;
  sw->lx = lx;
  sw->nl = nl;
  sw->nel = nel;
  sw->op = op;
  int st = gsdev_level(h, &sw->ext, op, nl, s);
  if (st) return st;
  const int64_t ns = nel * nl * nl * nl;
  if (!(sw->S = h->template alloc<double>(nel * 6 * nl * nl)) || !(sw->D = h->template alloc<double>(ns)) ||
      !(sw->wt = h->template alloc<double>(n)) || !(sw->work1 = h->template alloc<double>(ns)) ||
      !(sw->work2 = h->template alloc<double>(ns))) {
    set_error("hsmg: out of device memory");
    return NM_ERR_NOMEM;
  }
  // host geometry of the level + FDM pencils
  std::vector<double> x(n), y(n), z(n), zg(lx), wg(lx), d1(lx * lx), dt(lx * lx);
  NM_HIP_TRY(hipMemcpyAsync(x.data(), op->xyz[0], sizeof(double) * n, hipMemcpyDeviceToHost, s));
  NM_HIP_TRY(hipMemcpyAsync(y.data(), op->xyz[1], sizeof(double) * n, hipMemcpyDeviceToHost, s));
  NM_HIP_TRY(hipMemcpyAsync(z.data(), op->xyz[2], sizeof(double) * n, hipMemcpyDeviceToHost, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  nm_gll(lx, zg.data(), wg.data(), d1.data(), dt.data());
  // facet types: 0 where a neighbour exists, facet_bc on the boundary
  std::vector<int> ft(nel * 6);
  {
    const std::vector<char> nb = facet_has_neighbour(op);
    for (int64_t q = 0; q < nel * 6; q++) ft[q] = nb[q] ? 0 : facet_bc;
  }
  std::vector<double> hS(nel * 6 * nl * nl), hD(ns);
  HostGs hg{op};
  st = nm_fdm_setup_host(nel, lx, x.data(), y.data(), z.data(), wg.data(), d1.data(),
                         ft.data(), host_gs, &hg, hS.data(), hD.data());
  if (st) return st;
  NM_HIP_TRY(hipMemcpyAsync(sw->S, hS.data(), sizeof(double) * hS.size(), hipMemcpyHostToDevice, s));
  NM_HIP_TRY(hipMemcpyAsync(sw->D, hD.data(), sizeof(double) * hD.size(), hipMemcpyHostToDevice, s));
  // weights (schwarz_setup_wt): counts of overlapping copies
  double *e1;
  if (!(e1 = h->template alloc<double>(n))) return NM_ERR_NOMEM;
  if ((st = launch_vec<double>(V_CFILL, sw->work2, nullptr, nullptr, nullptr, 1.0, ns, s))) return st;
  if ((st = launch_vec<double>(V_RZERO, sw->work1, nullptr, nullptr, nullptr, 0.0, ns, s))) return st;
  if ((st = extrude(sw->work1, 0, 0.0, sw->work2, 0, 1.0, nl, nel, s))) return st;
  if ((st = gs_dev(sw->ext, sw->work2, s))) return st;
  if ((st = extrude(sw->work2, 0, 1.0, sw->work1, 0, -1.0, nl, nel, s))) return st;
  if ((st = extrude(sw->work2, 2, 1.0, sw->work2, 0, 1.0, nl, nel, s))) return st;
  hipLaunchKernelGGL(toreg_kernel, dim3(blocks_for(n)), dim3(256), 0, s, e1, sw->work2, lx, nel);
  if ((st = op->gs<double>(e1, nullptr, s))) return st;
  // the reference's face-layer weights (12 lx^2 per element, set-up only),
  // then their product per point
  double *wface = nullptr;
  NM_HIP_TRY(hipMalloc(&wface, sizeof(double) * nel * lx * lx * 12));
  hipLaunchKernelGGL(wt_setup_kernel, dim3(blocks_for(nel * lx * lx * 12)), dim3(256), 0, s,
                     wface, e1, lx, nel);
  hipLaunchKernelGGL(wtp_setup_kernel, dim3(blocks_for(n)), dim3(256), 0, s, sw->wt, wface, lx,
                     nel);
  const bool ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
  (void)hipFree(wface);
  if (!ok) {
    set_error("hsmg: Schwarz weight set-up failed");
    return NM_ERR_HIP;
  }
  return NM_OK;
}

// ------------------------------------------------------------- Chebyshev
// cheby_t (cheby.f90): power estimate of lambda_max of M^-1 A (M = the
// level's Schwarz smoother), then the bound solve cheby_impl.
__global__ void cheby_start_kernel(double *__restrict__ d, int64_t n, uint64_t seed) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    d[i] = drng_uniform(seed, (uint64_t)i) + 10.0;
}
// d = a d + b w; x += d.  The first step (first != 0) sets d = b w without
// reading d (which holds no value yet).  With wt, w is the Schwarz result
// before its weights: w wt, rounded as the weight pass would store it.
template <typename T>
__global__ void cheby_step_kernel(T *__restrict__ d, T *__restrict__ x, const T *__restrict__ w,
                                  T a, T b, int64_t n, int first, const T *__restrict__ wt) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const T wi = wt ? w[i] * wt[i] : w[i];
    const T di = first ? b * wi : a * d[i] + b * wi;
    d[i] = di;
    x[i] = x[i] + di;
  }
}
// FP64 <-> FP32 copies at the smoother boundary
__global__ void f2d_kernel(double *__restrict__ y, const float *__restrict__ x, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = (double)x[i];
}
// w = (r - w) * mult
__global__ void resid_mult_kernel(double *__restrict__ w, const double *__restrict__ r,
                                  const double *__restrict__ mult, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    w[i] = (r[i] - w[i]) * mult[i];
}

struct Cheby {
  const Op *op = nullptr;
  const Schwarz *sw = nullptr;
  double tha = 0.0, dlt = 0.0;
  double *d = nullptr, *w = nullptr, *r = nullptr;
  float *d32 = nullptr, *w32 = nullptr, *r32 = nullptr;  // FP32 smoothing
  template <typename T>
  void work(T *&dd, T *&ww, T *&rr) {
    if constexpr (std::is_same<T, double>::value) {
      dd = d; ww = w; rr = r;
    } else {
      dd = d32; ww = w32; rr = r32;
    }
  }
};

static unsigned vblocks(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

static int dot_host(const Op &op, const double *a, const double *b, double *part, double *out,
                    double *h, hipStream_t s) {
  int st = launch_glsc3_partial(a, op.f64.mult, b, op.n, part, s);
  if (st) return st;
  if ((st = launch_combine(part, 1, out, s))) return st;
  NM_HIP_TRY(hipMemcpyAsync(h, out, sizeof(double), hipMemcpyDeviceToHost, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  return NM_OK;
}

// cheby_power (cheby.f90:136-231); part/out: reduction scratch
static int cheby_power(Cheby &c, uint64_t seed, int its, double *part, double *out,
                       hipStream_t s) {
  const Op &op = *c.op;
  const int64_t n = op.n;
  int st;
  hipLaunchKernelGGL(cheby_start_kernel, dim3(vblocks(n)), dim3(256), 0, s, c.d, n, seed);
  if ((st = op.gs<double>(c.d, nullptr, s))) return st;
  if ((st = launch_mask<double>(c.d, op.msk, op.nmsk, nullptr, s))) return st;
  double wtw = 0.0, dtw = 0.0, dtd = 0.0;
  for (int it = 0; it <= its; it++) {
    if ((st = op.apply<double>(c.w, c.d, nullptr, s, true, true))) return st;
    if ((st = schwarz_apply(*c.sw, c.r, c.w, s))) return st;
    NM_HIP_TRY(hipMemcpyAsync(c.w, c.r, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (it == its) break;
    if ((st = dot_host(op, c.w, c.w, part, out, &wtw, s))) return st;
    if ((st = launch_vec<double>(V_CMULT2, c.d, c.w, nullptr, nullptr, 1.0 / std::sqrt(wtw), n,
                                 s)))
      return st;
    if ((st = launch_mask<double>(c.d, op.msk, op.nmsk, nullptr, s))) return st;
  }
  if ((st = dot_host(op, c.d, c.w, part, out, &dtw, s))) return st;
  if ((st = dot_host(op, c.d, c.d, part, out, &dtd, s))) return st;
  const double lam = dtw / dtd;
  const double b = lam * 1.1, a = lam / 30.0;
  c.tha = (b + a) / 2.0;
  c.dlt = (b - a) / 2.0;
  return NM_OK;
}

// cheby_impl (cheby.f90:317-397): niter steps on A x = f, all on stream s;
// T = float runs the whole smoother (operator, gs, Schwarz) in FP32
template <typename T>
static int cheby_solve(Cheby &c, T *x, const T *f, int niter, bool zero_init, hipStream_t s) {
  const Op &op = *c.op;
  const int64_t n = op.n;
  const unsigned g = vblocks(n);
  T *d, *w, *r;
  c.work<T>(d, w, r);
  int st;
  // the Schwarz weights ride in the Chebyshev step (one pass over w less)
  const T *S_, *D_, *wt;
  T *w1_, *w2_;
  c.sw->arrays<T>(S_, D_, wt, w1_, w2_);
  // the residual f - A x is formed inside the Schwarz extension (w holds
  // A x until the Schwarz result replaces it)
  if (!zero_init) {
    if ((st = op.apply<T>(w, x, nullptr, s, true, true))) return st;
    if ((st = schwarz_apply<T>(*c.sw, w, w, s, false, f))) return st;
  } else {
    NM_HIP_TRY(hipMemcpyAsync(r, f, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
    if ((st = schwarz_apply<T>(*c.sw, w, r, s, false))) return st;
  }
  hipLaunchKernelGGL(cheby_step_kernel<T>, dim3(g), dim3(256), 0, s, d, x, w, T(0),
                     T(1.0 / c.tha), n, 1, wt);
  const double sig1 = c.tha / c.dlt;
  double rhok = 1.0 / sig1;
  for (int iter = 2; iter <= niter; iter++) {
    const double rhokp1 = 1.0 / (2.0 * sig1 - rhok);
    const double tmp1 = rhokp1 * rhok;
    const double tmp2 = 2.0 * rhokp1 / c.dlt;
    rhok = rhokp1;
    if ((st = op.apply<T>(w, x, nullptr, s, true, true))) return st;
    if ((st = schwarz_apply<T>(*c.sw, w, w, s, false, f))) return st;
    hipLaunchKernelGGL(cheby_step_kernel<T>, dim3(g), dim3(256), 0, s, d, x, w, (T)tmp1,
                       (T)tmp2, n, 0, wt);
  }
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// ------------------------------------------------------------------ phmg
// phmg_mg_cycle (phmg.f90:605-757) on hsmg's levels lx -> lx_mid -> 2 with
// Chebyshev smoothing accelerated by Schwarz on the two finer levels and
// Jacobi-CG on the coarse level.
struct Phmg {
  Op *lev[3] = {nullptr, nullptr, nullptr};
  nm_op_t *mid_h = nullptr, *crs_h = nullptr;
  Schwarz sw[2];
  Cheby ch[2];
  double *jfm = nullptr, *jfm_t = nullptr, *jmc = nullptr, *jmc_t = nullptr;
  double *r[3] = {}, *z[3] = {}, *w[3] = {};
  double *part = nullptr, *out = nullptr;
  CgFixed *cg = nullptr;
  Tamg *amg = nullptr;  // the tree-AMG coarse solve, when set (else cg)
  int smooth_its = 4, crs_niter = 10;
  double crs_tol = 1e-9;
  bool fp32 = false;                      // FP32 smoothers
  float *z32[2] = {}, *f32[2] = {};       // level solution / rhs in FP32
  std::vector<void *> owned;
  ~Phmg() {
    delete amg;
    if (cg) cg_fixed_free(cg);
    for (void *q : owned) (void)hipFree(q);
    if (mid_h) nm_op_free(mid_h);
    if (crs_h) nm_op_free(crs_h);
  }
  template <typename T>
  T *alloc(int64_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess) return nullptr;
    owned.push_back(q);
    return (T *)q;
  }
};

// Chebyshev smoothing of level l (pre: from z = 0).  FP32 smoothers take
// FP32 copies of the level's rhs and iterate and return z in FP64.
static int phmg_smooth(Phmg *h, int l, bool pre, hipStream_t s) {
  if (!h->fp32) return cheby_solve<double>(h->ch[l], h->z[l], h->r[l], h->smooth_its, pre, s);
  const int64_t n = h->lev[l]->n;
  const unsigned g = vblocks(n);
  int st;
  if (pre) {
    if ((st = launch_d2f(h->r[l], h->f32[l], n, s))) return st;
    if ((st = launch_vec<float>(V_RZERO, h->z32[l], nullptr, nullptr, nullptr, 0.0f, n, s)))
      return st;
  } else if ((st = launch_d2f(h->z[l], h->z32[l], n, s))) {
    return st;
  }
  if ((st = cheby_solve<float>(h->ch[l], h->z32[l], h->f32[l], h->smooth_its, pre, s))) return st;
  hipLaunchKernelGGL(f2d_kernel, dim3(g), dim3(256), 0, s, h->z[l], h->z32[l], n);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

static int phmg_cycle(Phmg *h, double *zout, const double *rin, hipStream_t s) {
  Op **L = h->lev;
  const int64_t nel = L[0]->nel;
  const double *jt[2] = {h->jfm_t, h->jmc_t}, *jj[2] = {h->jfm, h->jmc};
  int st;
  NM_HIP_TRY(hipMemcpyAsync(h->r[0], rin, sizeof(double) * L[0]->n, hipMemcpyDeviceToDevice, s));
  if ((st = launch_vec<double>(V_RZERO, h->z[0], nullptr, nullptr, nullptr, 0.0, L[0]->n, s)))
    return st;
  for (int l = 0; l < 2; l++) {
    const Op &op = *L[l];
    if ((st = phmg_smooth(h, l, true, s))) return st;
    if ((st = op.apply<double>(h->w[l], h->z[l], nullptr, s, true, true))) return st;
    hipLaunchKernelGGL(resid_mult_kernel, dim3(vblocks(op.n)), dim3(256), 0, s, h->w[l],
                       h->r[l], op.f64.mult, op.n);
    if ((st = tnsr3d(h->r[l + 1], L[l + 1]->lx, h->w[l], op.lx, jt[l], jj[l], jj[l], nel, s)))
      return st;
    if ((st = L[l + 1]->gs<double>(h->r[l + 1], nullptr, s))) return st;
    if ((st = launch_mask<double>(h->r[l + 1], L[l + 1]->msk, L[l + 1]->nmsk, nullptr, s)))
      return st;
    if ((st = launch_vec<double>(V_RZERO, h->z[l + 1], nullptr, nullptr, nullptr, 0.0,
                                 L[l + 1]->n, s)))
      return st;
  }
  if (h->amg)  // phmg's coarse solve: the tree-AMG (phmg.f90:693-700)
    st = tamg_solve(h->amg, h->z[2], h->r[2], s);
  else
    st = cg_fixed_solve(h->cg, h->z[2], h->r[2], h->crs_niter, h->crs_tol, s);
  if (st) return st;
  for (int l = 1; l >= 0; l--) {
    const Op &op = *L[l];
    if ((st = tnsr3d(h->w[l], op.lx, h->z[l + 1], L[l + 1]->lx, jj[l], jt[l], jt[l], nel, s)))
      return st;
    if ((st = op.gs<double>(h->w[l], nullptr, s))) return st;
    if ((st = launch_vec<double>(V_ADDCOL3, h->z[l], h->w[l], op.f64.mult, nullptr, 0.0, op.n,
                                 s)))
      return st;
    if ((st = phmg_smooth(h, l, false, s))) return st;
  }
  NM_HIP_TRY(hipMemcpyAsync(zout, h->z[0], sizeof(double) * L[0]->n, hipMemcpyDeviceToDevice, s));
  return NM_OK;
}

}  // namespace nm

using namespace nm;

struct nm_hsmg_s {
  Hsmg h;
};
struct nm_gmres_s;

extern "C" {

// Kernel family of every Schwarz FDM solve that follows (hsmg, phmg):
// 0 matrix cores (default), 1 vector ALUs.  Both give the same solve to
// rounding; the choice is for timing.
int nm_fdm_set_variant(int variant) {
  if (variant != 0 && variant != 1) {
    set_error("nm_fdm_set_variant: %d (0 matrix cores, 1 vector ALUs)", variant);
    return NM_ERR_ARG;
  }
  g_fdm_variant = variant;
  return NM_OK;
}

// hsmg on an operator built on a box (nm_op_create_box) or a general mesh
// (nm_op_create_mesh).  facet_bc: Neko's facet_type of the boundary facets
// (1 strong Dirichlet as set for pressure, 2 as set for velocity;
// fluid_pnpn_bc_fctry.f90:162,259).
int nm_hsmg_create(nm_hsmg_t **out, nm_op_t *fine_h, int facet_bc, int niter,
                   hipStream_t s) {
  *out = nullptr;
  Op *fine = &fine_h->op;
  if (!has_topology(fine)) {
    set_error("nm_hsmg_create: operator has no mesh topology (build it with "
              "nm_op_create_box or nm_op_create_mesh)");
    return NM_ERR_ARG;
  }
  const int lf = fine->lx;
  if (lf < 3 || lf > 10) {
    set_error("nm_hsmg_create: lx=%d outside 3..10", lf);
    return NM_ERR_ARG;
  }
  const int lm = lf < 5 ? std::max(lf - 1, 3) : 4, lc = 2;
  nm_hsmg_t *hh = new nm_hsmg_t;
  Hsmg *h = &hh->h;
  h->fine = fine;
  h->niter = niter > 0 ? niter : 10;
  auto fail = [&](int c) {
    delete hh;
    return c;
  };
  // level operators on the same mesh
  int st = level_op(fine, lm, 1.0, 0.0, &h->mid_h, s);
  if (st) return fail(st);
  if ((st = level_op(fine, lc, fine->h1c, 0.0, &h->crs_h, s))) return fail(st);
  h->mid = &h->mid_h->op;
  h->crs = &h->crs_h->op;
  if ((st = nm_op_set_all_deformed(h->crs_h, s))) return fail(st);
  if ((st = schwarz_setup(h, &h->sw_f, fine, facet_bc, s))) return fail(st);
  if ((st = schwarz_setup(h, &h->sw_m, h->mid, facet_bc, s))) return fail(st);
  // interpolation matrices (mid -> fine: lf x lm, crs -> mid: lm x lc)
  std::vector<double> zf(lf), zm(lm), zc(lc), tmp(lf * lf * 3);
  std::vector<double> wdum(lf), ddum(lf * lf), dtdum(lf * lf);
  nm_gll(lf, zf.data(), wdum.data(), ddum.data(), dtdum.data());
  nm_gll(lm, zm.data(), wdum.data(), ddum.data(), dtdum.data());
  nm_gll(lc, zc.data(), wdum.data(), ddum.data(), dtdum.data());
  std::vector<double> jfm(lf * lm), jfmt(lf * lm), jmc(lm * lc), jmct(lm * lc);
  nm_interp_matrix(zf.data(), lf, zm.data(), lm, jfm.data(), jfmt.data());
  nm_interp_matrix(zm.data(), lm, zc.data(), lc, jmc.data(), jmct.data());
  if (!(h->jfm = h->alloc<double>(lf * lm)) || !(h->jfm_t = h->alloc<double>(lf * lm)) ||
      !(h->jmc = h->alloc<double>(lm * lc)) || !(h->jmc_t = h->alloc<double>(lm * lc)) ||
      !(h->r = h->alloc<double>(fine->n)) || !(h->w = h->alloc<double>(fine->n)) ||
      !(h->e_mid = h->alloc<double>(h->mid->n)) || !(h->e_crs = h->alloc<double>(h->crs->n)) ||
      !(h->r_crs = h->alloc<double>(h->crs->n)))
    return fail(NM_ERR_NOMEM);
  if (hipMemcpyAsync(h->jfm, jfm.data(), sizeof(double) * jfm.size(), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(h->jfm_t, jfmt.data(), sizeof(double) * jfmt.size(), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(h->jmc, jmc.data(), sizeof(double) * jmc.size(), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(h->jmc_t, jmct.data(), sizeof(double) * jmct.size(), hipMemcpyHostToDevice, s))
    return fail(NM_ERR_HIP);
  if ((st = cg_fixed_create(&h->cg, h->crs))) return fail(st);
  if (hipStreamSynchronize(s) != hipSuccess) return fail(NM_ERR_HIP);
  *out = hh;
  return NM_OK;
}

void nm_hsmg_free(nm_hsmg_t *h) { delete h; }

// z = hsmg(r) (pc_hsmg.f90:601-626, CPU branch semantics).  Device FP64.
int nm_hsmg_apply(nm_hsmg_t *hh, double *z, const double *rin, hipStream_t s) {
  Hsmg *h = &hh->h;
  const Op &F = *h->fine, &M = *h->mid, &Cc = *h->crs;
  const int lf = F.lx, lm = M.lx, lc = Cc.lx;
  const int64_t nel = F.nel;
  int st;
  NM_HIP_TRY(hipMemcpyAsync(h->r, rin, sizeof(double) * F.n, hipMemcpyDeviceToDevice, s));
  if ((st = schwarz_apply(h->sw_f, z, h->r, s))) return st;
  if ((st = launch_vec<double>(V_COL2, h->r, F.f64.mult, nullptr, nullptr, 0.0, F.n, s))) return st;
  double *wm = h->w;  // mid-size work in the fine-size buffer
  if ((st = tnsr3d(wm, lm, h->r, lf, h->jfm_t, h->jfm, h->jfm, nel, s))) return st;
  if ((st = M.gs<double>(wm, nullptr, s))) return st;
  if ((st = schwarz_apply(h->sw_m, h->e_mid, wm, s))) return st;
  if ((st = launch_vec<double>(V_COL2, wm, M.f64.mult, nullptr, nullptr, 0.0, M.n, s))) return st;
  if ((st = tnsr3d(h->r_crs, lc, wm, lm, h->jmc_t, h->jmc, h->jmc, nel, s))) return st;
  if ((st = Cc.gs<double>(h->r_crs, nullptr, s))) return st;
  if ((st = launch_mask<double>(h->r_crs, Cc.msk, Cc.nmsk, nullptr, s))) return st;
  if ((st = cg_fixed_solve(h->cg, h->e_crs, h->r_crs, h->niter, h->crs_tol, s))) return st;
  if ((st = launch_mask<double>(h->e_crs, Cc.msk, Cc.nmsk, nullptr, s))) return st;
  if ((st = tnsr3d(wm, lm, h->e_crs, lc, h->jmc, h->jmc_t, h->jmc_t, nel, s))) return st;
  if ((st = launch_vec<double>(V_ADD2, h->e_mid, wm, nullptr, nullptr, 0.0, M.n, s))) return st;
  if ((st = tnsr3d(h->r, lf, h->e_mid, lm, h->jfm, h->jfm_t, h->jfm_t, nel, s))) return st;
  if ((st = launch_vec<double>(V_ADD2, z, h->r, nullptr, nullptr, 0.0, F.n, s))) return st;
  if ((st = F.gs<double>(z, nullptr, s))) return st;
  return launch_vec<double>(V_COL2, z, F.f64.mult, nullptr, nullptr, 0.0, F.n, s);
}

struct nm_phmg_s {
  Phmg h;
};

// phmg-style preconditioner on a box operator (see include/neko_amd.h).
int nm_phmg_create(nm_phmg_t **out, nm_op_t *fine_h, int facet_bc, int smooth_its,
                   int level_h2_fine, int smoother_fp32, hipStream_t s) {
  *out = nullptr;
  Op *fine = &fine_h->op;
  if (!has_topology(fine)) {
    set_error("nm_phmg_create: operator has no mesh topology (build it with "
              "nm_op_create_box or nm_op_create_mesh)");
    return NM_ERR_ARG;
  }
  const int lf = fine->lx;
  if (lf < 3 || lf > 10) {
    set_error("nm_phmg_create: lx=%d outside 3..10", lf);
    return NM_ERR_ARG;
  }
  if (fine->f64.h1) {
    set_error("nm_phmg_create: pointwise h1 unsupported (the levels take a constant h1)");
    return NM_ERR_UNSUPPORTED;
  }
  const int lm = lf < 5 ? std::max(lf - 1, 3) : 4, lc = 2;
  nm_phmg_t *hh = new nm_phmg_t;
  Phmg *h = &hh->h;
  h->smooth_its = smooth_its > 0 ? smooth_its : 4;
  auto fail = [&](int c) {
    delete hh;
    return c;
  };
  // level operators: h1 and ifh2 of the fine level; h2 = 1 (Neko's coef
  // default, phmg.f90:252-254 copies h1 and ifh2 only) or the fine h2
  const double h2l = fine->ifh2 ? (level_h2_fine ? fine->h2c : 1.0) : 0.0;
  int st = level_op(fine, lm, fine->h1c, h2l, &h->mid_h, s);
  if (st) return fail(st);
  if ((st = level_op(fine, lc, fine->h1c, h2l, &h->crs_h, s))) return fail(st);
  h->lev[0] = fine;
  h->lev[1] = &h->mid_h->op;
  h->lev[2] = &h->crs_h->op;
  for (int l = 0; l < 2; l++)
    if ((st = schwarz_setup(h, &h->sw[l], h->lev[l], facet_bc, s))) return fail(st);
  std::vector<double> zf(lf), zm(lm), zc(lc), wd(lf), dd(lf * lf), dtd(lf * lf);
  nm_gll(lf, zf.data(), wd.data(), dd.data(), dtd.data());
  nm_gll(lm, zm.data(), wd.data(), dd.data(), dtd.data());
  nm_gll(lc, zc.data(), wd.data(), dd.data(), dtd.data());
  std::vector<double> jfm(lf * lm), jfmt(lf * lm), jmc(lm * lc), jmct(lm * lc);
  nm_interp_matrix(zf.data(), lf, zm.data(), lm, jfm.data(), jfmt.data());
  nm_interp_matrix(zm.data(), lm, zc.data(), lc, jmc.data(), jmct.data());
  if (!(h->jfm = h->alloc<double>(lf * lm)) || !(h->jfm_t = h->alloc<double>(lf * lm)) ||
      !(h->jmc = h->alloc<double>(lm * lc)) || !(h->jmc_t = h->alloc<double>(lm * lc)) ||
      !(h->part = h->alloc<double>(RED_BLOCKS)) || !(h->out = h->alloc<double>(1)))
    return fail(NM_ERR_NOMEM);
  for (int l = 0; l < 3; l++)
    if (!(h->r[l] = h->alloc<double>(h->lev[l]->n)) || !(h->z[l] = h->alloc<double>(h->lev[l]->n)) ||
        !(h->w[l] = h->alloc<double>(h->lev[l]->n)))
      return fail(NM_ERR_NOMEM);
  if (hipMemcpyAsync(h->jfm, jfm.data(), sizeof(double) * jfm.size(), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(h->jfm_t, jfmt.data(), sizeof(double) * jfmt.size(), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(h->jmc, jmc.data(), sizeof(double) * jmc.size(), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(h->jmc_t, jmct.data(), sizeof(double) * jmct.size(), hipMemcpyHostToDevice, s))
    return fail(NM_ERR_HIP);
  for (int l = 0; l < 2; l++) {
    Cheby &c = h->ch[l];
    c.op = h->lev[l];
    c.sw = &h->sw[l];
    const int64_t n = c.op->n;
    if (!(c.d = h->alloc<double>(n)) || !(c.w = h->alloc<double>(n)) ||
        !(c.r = h->alloc<double>(n)))
      return fail(NM_ERR_NOMEM);
    if ((st = cheby_power(c, 3901, 150, h->part, h->out, s))) return fail(st);
  }
  if ((st = cg_fixed_create(&h->cg, h->lev[2]))) return fail(st);
  if (smoother_fp32) {
    h->fp32 = true;
    for (int l = 0; l < 2; l++) {
      Op *op = h->lev[l];
      if ((st = op->ensure32(s))) return fail(st);
      Schwarz &w = h->sw[l];
      const int64_t n = op->n, nl = w.nl, ns = w.nel * nl * nl * nl;
      const int64_t nS = w.nel * 6 * nl * nl, nW = n;
      Cheby &c = h->ch[l];
      if (!(w.S32 = h->alloc<float>(nS)) || !(w.D32 = h->alloc<float>(ns)) ||
          !(w.wt32 = h->alloc<float>(nW)) || !(w.work1_32 = h->alloc<float>(ns)) ||
          !(w.work2_32 = h->alloc<float>(ns)) || !(c.d32 = h->alloc<float>(n)) ||
          !(c.w32 = h->alloc<float>(n)) || !(c.r32 = h->alloc<float>(n)) ||
          !(h->z32[l] = h->alloc<float>(n)) || !(h->f32[l] = h->alloc<float>(n)))
        return fail(NM_ERR_NOMEM);
      if ((st = launch_d2f(w.S, w.S32, nS, s)) || (st = launch_d2f(w.D, w.D32, ns, s)) ||
          (st = launch_d2f(w.wt, w.wt32, nW, s)))
        return fail(st);
    }
  }
  if (hipStreamSynchronize(s) != hipSuccess) return fail(NM_ERR_HIP);
  *out = hh;
  return NM_OK;
}

void nm_phmg_free(nm_phmg_t *h) { delete h; }

// phmg's own coarse solve, the tree-AMG (phmg.f90:150-157, coarse_grid
// levels / iterations / cheby_degree), on the coarse level; levels 0
// restores hsmg's coarse Jacobi-CG.
int nm_phmg_set_coarse_amg(nm_phmg_t *hh, int levels, int iterations, int cheby_degree,
                           hipStream_t s) {
  Phmg *h = &hh->h;
  delete h->amg;
  h->amg = nullptr;
  if (levels == 0) return NM_OK;
  const Op *crs = h->lev[2];
  return tamg_create(&h->amg, crs, facet_neighbours(crs), levels, iterations, cheby_degree,
                     250, s);
}

int nm_phmg_coarse_amg_levels(const nm_phmg_t *hh, int64_t *nn, int cap) {
  const Tamg *a = hh->h.amg;
  if (!a) return 0;
  for (int l = 0; l <= a->nlvls && l < cap; l++) nn[l] = a->nn[l];
  return a->nlvls;
}

int nm_phmg_coarse_amg_map(const nm_phmg_t *hh, int lvl, int32_t *out) {
  const Tamg *a = hh->h.amg;
  if (!a || lvl < 1 || lvl > a->nlvls) {
    set_error("nm_phmg_coarse_amg_map: no AMG level %d", lvl);
    return NM_ERR_ARG;
  }
  std::copy(a->h_map[lvl].begin(), a->h_map[lvl].end(), out);
  return NM_OK;
}

int nm_phmg_apply(nm_phmg_t *h, double *z, const double *r, hipStream_t s) {
  return phmg_cycle(&h->h, z, r, s);
}

int nm_phmg_eigs(const nm_phmg_t *h, double *tha_dlt) {
  for (int l = 0; l < 2; l++) {
    tha_dlt[2 * l] = h->h.ch[l].tha;
    tha_dlt[2 * l + 1] = h->h.ch[l].dlt;
  }
  return NM_OK;
}

// Schwarz smoother of the fine level alone (e = schwarz(r); r masked).
int nm_hsmg_schwarz_fine(nm_hsmg_t *hh, double *e, double *r, hipStream_t s) {
  return schwarz_apply(hh->h.sw_f, e, r, s);
}

}  // extern "C"
