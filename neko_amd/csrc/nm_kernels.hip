// gfx950 kernels of the solve path: tensor-product Helmholtz operator,
// gather-scatter, Dirichlet mask, deterministic reductions, BLAS-1 updates,
// geometric factors and the Jacobi diagonal.
//
// Field layout everywhere: Neko's column-major (lx,lx,lx,nel), i.e. point
// (i,j,k) of element e at i + lx*(j + lx*(k + lx*e)).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "nm_common.h"
#include "nm_kernels.h"

namespace nm {

// ---------------------------------------------------------------------------
// Ax: "k-step" tensor-product kernel.
//
// One lx*lx thread plane per element (EPB elements per workgroup).  Thread
// (i,j) keeps the u column u(i,j,:) in registers, so the t-derivative and
// the final D^T contraction in t never touch LDS; the r- and s-derivatives
// and their transposes go through one lx*lx LDS plane per element.  The six
// geometric factors are streamed once, coalesced (consecutive threads read
// consecutive addresses of each plane).  Reference behaviour:
// ax_helm_cpu.f90:180-290 (generic) / :1211-1341 (lx=8); device twin
// ax_helm_kernel.h:149-355.
// ---------------------------------------------------------------------------
// Fixed-order workgroup sum of one double per thread (the tree does not
// need whole waves); thread 0 stores it.  Every thread must call it.
template <int NT>
__device__ __forceinline__ void wg_sum_store(double v, double *sdot, int flat,
                                             double *__restrict__ out) {
  constexpr int H = NT <= 2 ? 1 : (NT <= 4 ? 2 : (NT <= 8 ? 4 : (NT <= 16 ? 8 :
                    (NT <= 32 ? 16 : (NT <= 64 ? 32 : (NT <= 128 ? 64 : 128))))));
  sdot[flat] = v;
  __syncthreads();
  for (int h = H; h > 0; h >>= 1) {
    if (flat < h && flat + h < NT) sdot[flat] += sdot[flat + h];
    __syncthreads();
  }
  if (flat == 0) *out = sdot[0];
}

// CG: the CG prologue and the p.Ap partial ride in the operator, as in the
// in-kernel-geometry kernel below: the operand is the new search direction
// p = dinv r + beta p_old (formed while loading, written back over p_old),
// the previous step's x += alpha p_old uses the same loads, and each
// workgroup stores sum_i p_i (A_local p)_i (mass term included) to
// dot_part[blockIdx.x].
template <typename T, int LX, int EPB, bool PF, bool CG>
__global__ void __launch_bounds__(LX *LX *EPB)
    ax_kstep_kernel(T *__restrict__ w, const T *__restrict__ u,
                    const T *__restrict__ dx, const T *__restrict__ h1, T h1c,
                    const T *__restrict__ g11, const T *__restrict__ g22,
                    const T *__restrict__ g33, const T *__restrict__ g12,
                    const T *__restrict__ g13, const T *__restrict__ g23,
                    const T *__restrict__ h2b, int64_t nel,
                    const int *__restrict__ skip, double *__restrict__ dot_part,
                    const T *__restrict__ cg_r, const T *__restrict__ cg_dinv,
                    T *__restrict__ cg_x, const double *__restrict__ cg_beta,
                    const double *__restrict__ cg_alpha) {
  constexpr int P = LX * LX;
  constexpr int N3 = LX * P;
  __shared__ T su[EPB][N3];     // element u; plane k then holds the r/s part of w
  __shared__ T sr[EPB][2][P];   // r-flux plane, double buffered
  __shared__ T ss[EPB][2][P];   // s-flux plane, double buffered
  __shared__ T sD[P];           // D, for the columns of the transposed contraction
  __shared__ double sdot[CG ? P * EPB : 1];

  if (skip != nullptr && *skip) return;

  const int t = threadIdx.x;
  const int q = threadIdx.y;
  const int i = t % LX;
  const int j = t / LX;
  const int64_t e = (int64_t)blockIdx.x * EPB + q;
  const bool active = e < nel;
  const int64_t base = e * (int64_t)N3 + t;

  // u column in registers (static indices only) and the whole element in
  // LDS for the r/s derivatives; plane k is P consecutive values, so both
  // the loads and the LDS stores are coalesced.
  T uc[LX];
  if (CG) {
    // all loads of the column before any store (the arrays are distinct,
    // which the compiler cannot know)
    T *p = const_cast<T *>(u);
    const T beta = (T)*cg_beta, alpha = (T)*cg_alpha;
    T po[LX], pr[LX], pd[LX], px[LX];
#pragma unroll
    for (int k = 0; k < LX; k++) {
      const int64_t ix = base + k * P;
      po[k] = active ? __builtin_nontemporal_load(p + ix) : T(0);
      pr[k] = active ? __builtin_nontemporal_load(cg_r + ix) : T(0);
      pd[k] = active ? __builtin_nontemporal_load(cg_dinv + ix) : T(0);
      px[k] = active ? __builtin_nontemporal_load(cg_x + ix) : T(0);
    }
#pragma unroll
    for (int k = 0; k < LX; k++) {
      uc[k] = pd[k] * pr[k] + beta * po[k];
      su[q][k * P + t] = uc[k];
    }
    if (active) {
#pragma unroll
      for (int k = 0; k < LX; k++) {
        const int64_t ix = base + k * P;
        __builtin_nontemporal_store(uc[k], p + ix);
        __builtin_nontemporal_store(px[k] + alpha * po[k], cg_x + ix);
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < LX; k++) {
      uc[k] = active ? __builtin_nontemporal_load(u + base + k * P) : T(0);
      su[q][k * P + t] = uc[k];
    }
  }

  // rows D(i,:), D(j,:) in registers; the columns D(:,i), D(:,j) of the
  // transposed contraction from LDS (broadcast-free, conflict-light reads),
  // which leaves the registers to the streamed geometry
  if (q == 0) sD[t] = dx[t];
  T Di[LX], Dj[LX];
#pragma unroll
  for (int l = 0; l < LX; l++) {
    Di[l] = dx[i + LX * l];
    Dj[l] = dx[j + LX * l];
  }
  // t-direction transpose accumulates in registers (scatter form)
  T wc[LX];
#pragma unroll
  for (int m = 0; m < LX; m++) wc[m] = T(0);
  __syncthreads();

  // Plane loop kept rolled: the working set per iteration is one plane, so
  // the register footprint stays small enough for several waves per SIMD.
  T nG[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  T nh = h1c;
  if (PF && active) {
    nG[0] = __builtin_nontemporal_load(g11 + base); nG[1] = __builtin_nontemporal_load(g22 + base);
    nG[2] = __builtin_nontemporal_load(g33 + base); nG[3] = __builtin_nontemporal_load(g12 + base);
    nG[4] = __builtin_nontemporal_load(g13 + base); nG[5] = __builtin_nontemporal_load(g23 + base);
    if (h1 != nullptr) nh = h1[base];
  }
#pragma unroll 1
  for (int k = 0; k < LX; k++) {
    const int b = k & 1;
    const int64_t p = base + (int64_t)k * P;
    T G11 = T(0), G22 = T(0), G33 = T(0), G12 = T(0), G13 = T(0), G23 = T(0);
    T hh = h1c;
    if (PF) {
      // software pipelining: this plane's factors were loaded one plane
      // ahead; issue the next plane's loads now
      G11 = nG[0]; G22 = nG[1]; G33 = nG[2]; G12 = nG[3]; G13 = nG[4]; G23 = nG[5];
      hh = nh;
      if (active && k + 1 < LX) {
        const int64_t pn = p + P;
        nG[0] = __builtin_nontemporal_load(g11 + pn); nG[1] = __builtin_nontemporal_load(g22 + pn);
        nG[2] = __builtin_nontemporal_load(g33 + pn); nG[3] = __builtin_nontemporal_load(g12 + pn);
        nG[4] = __builtin_nontemporal_load(g13 + pn); nG[5] = __builtin_nontemporal_load(g23 + pn);
        if (h1 != nullptr) nh = h1[pn];
      }
    } else if (active) {
      G11 = __builtin_nontemporal_load(g11 + p); G22 = __builtin_nontemporal_load(g22 + p);
      G33 = __builtin_nontemporal_load(g33 + p); G12 = __builtin_nontemporal_load(g12 + p);
      G13 = __builtin_nontemporal_load(g13 + p); G23 = __builtin_nontemporal_load(g23 + p);
      if (h1 != nullptr) hh = h1[p];
    }
    // D(k,:) is wave-uniform (scalar loads)
    T dk[LX];
#pragma unroll
    for (int l = 0; l < LX; l++) dk[l] = dx[k + LX * l];
    T vt = T(0), ur = T(0), us = T(0);
    const T *plane = &su[q][k * P];
#pragma unroll
    for (int l = 0; l < LX; l++) {
      vt += dk[l] * uc[l];
      ur += Di[l] * plane[l + LX * j];
      us += Dj[l] * plane[i + LX * l];
    }
    const T rr = hh * (G11 * ur + G12 * us + G13 * vt);
    const T rs = hh * (G12 * ur + G22 * us + G23 * vt);
    const T rt = hh * (G13 * ur + G23 * us + G33 * vt);
    sr[q][b][t] = rr;
    ss[q][b][t] = rs;
#pragma unroll
    for (int m = 0; m < LX; m++) wc[m] += dk[m] * rt;
    __syncthreads();
    T acc = T(0);
#pragma unroll
    for (int l = 0; l < LX; l++) acc += sD[l + LX * i] * sr[q][b][l + LX * j];
#pragma unroll
    for (int l = 0; l < LX; l++) acc += sD[l + LX * j] * ss[q][b][i + LX * l];
    su[q][k * P + t] = acc;  // plane k is dead: its r/s reads precede the barrier
  }

  double dsum = 0.0;
  if (active) {
#pragma unroll
    for (int m = 0; m < LX; m++) {
      T r = wc[m] + su[q][m * P + t];
      if (h2b != nullptr) r += h2b[base + m * P] * uc[m];
      __builtin_nontemporal_store(r, w + base + m * P);
      if (CG) dsum += (double)r * (double)uc[m];
    }
  }
  if (CG) wg_sum_store<P * EPB>(dsum, sdot, q * P + t, dot_part + blockIdx.x);
}

// ---------------------------------------------------------------------------
// Ax with the geometry evaluated in the kernel ("trilinear" variant).
//
// For an element with straight edges Neko's GLL coordinates are the
// trilinear interpolant of its 8 vertices (dofmap_xyzlin, dofmap.f90:
// 799-856), so the spectral derivatives coef.f90 takes of them are the
// exact derivatives of that map: dx/dr is bilinear in (s,t), dx/ds in
// (r,t), dx/dt in (r,s).  From 36 half edge vectors per element the kernel
// forms the Jacobian, its cofactors (Neko's drdx.. = cofactor entries),
// jac, G = h1 w3 C C^T / jac and B = jac w3 at every point (the formulas of
// coef_generate_geo / coef_generate_mass, coef.f90:903-1030,1112-1156), so
// per point it streams only u in and w out (16 B FP64) instead of 64 B.
// Same thread layout as ax_kstep_kernel.
//
// edges per element: [0,12) r-edges Er[c][cy][cz] at c*4 + cy + 2*cz,
// [12,24) s-edges Es[c][cx][cz] at 12 + c*4 + cx + 2*cz, [24,36) t-edges
// Et[c][cx][cy] at 24 + c*4 + cx + 2*cy, each (X(hi) - X(lo)) / 2.
// DOT: also writes the workgroup's partial sum_i u_i w_i (local w, before
// gs), which equals the assembled p.A p of CG for a continuous, masked u.
// ---------------------------------------------------------------------------
// Body shared by the single-problem and the batched kernels.  Element q of
// the workgroup: `active`, its first dof at w/u + `ebase`, its 36 edge
// values at `ed`.  With DOT the workgroup sum lands in *dot_out.
// CG prologue fused into the operator (CGP): the operand is the new search
// direction p = dinv r + beta p_old, formed while loading and written back
// over p_old, and the previous step's x += alpha p_old is applied from the
// same loads.
template <typename T>
struct CgPro {
  T *p;
  const T *r, *dinv;
  T *x;
  T beta, alpha;
};

// 1/a for the Jacobian: the hardware reciprocal estimate refined by
// Newton-Raphson steps y <- y (2 - a y), each of which squares the relative
// error -- two in FP64, one in FP32 (whose estimate is already good to
// about an ulp) -- so the result is within an ulp or two of 1/a (a > 0 and
// normal here); it skips the scaling and fix-up steps of a full IEEE
// division.
__device__ __forceinline__ double recip(double a) {
  double y = __builtin_amdgcn_rcp(a);
  y = y * (2.0 - a * y);
  return y * (2.0 - a * y);
}
__device__ __forceinline__ float recip(float a) {
  const float y = __builtin_amdgcn_rcpf(a);
  return y * (2.0f - a * y);
}

template <typename T, int LX, int EPB, bool DOT, bool CGP = false>
__device__ __forceinline__ void ax_tri_body(T *__restrict__ w, const T *__restrict__ u,
                                            const T *__restrict__ dx,
                                            const T *__restrict__ zg,
                                            const T *__restrict__ wg,
                                            const T *__restrict__ ed, T h1c, T h2c,
                                            bool active, int64_t ebase,
                                            double *__restrict__ dot_out,
                                            CgPro<T> cp = CgPro<T>{}) {
  constexpr int P = LX * LX;
  constexpr int N3 = LX * P;
  __shared__ T su[EPB][N3];
  __shared__ T sr[EPB][2][P];
  __shared__ T ss[EPB][2][P];
  __shared__ T sedge[EPB][36];
  __shared__ double sdot[DOT ? LX * LX * EPB : 1];

  const int t = threadIdx.x;
  const int q = threadIdx.y;
  const int i = t % LX;
  const int j = t / LX;
  const int64_t base = ebase + t;

  T uc[LX];
  if (CGP) {
    // all loads of the column first: p, r, dinv and x are distinct arrays,
    // but the compiler cannot know it, and interleaved stores would make
    // every plane wait for the previous plane's stores (one HBM round trip
    // per plane instead of one per column)
    T po[LX], pr[LX], pd[LX], px[LX];
#pragma unroll
    for (int k = 0; k < LX; k++) {
      const int64_t ix = base + k * P;
      po[k] = active ? __builtin_nontemporal_load(cp.p + ix) : T(0);
      pr[k] = active ? __builtin_nontemporal_load(cp.r + ix) : T(0);
      pd[k] = active ? __builtin_nontemporal_load(cp.dinv + ix) : T(0);
      px[k] = active ? __builtin_nontemporal_load(cp.x + ix) : T(0);
    }
#pragma unroll
    for (int k = 0; k < LX; k++) {
      uc[k] = pd[k] * pr[k] + cp.beta * po[k];
      su[q][k * P + t] = uc[k];
    }
    if (active) {
#pragma unroll
      for (int k = 0; k < LX; k++) {
        const int64_t ix = base + k * P;
        __builtin_nontemporal_store(uc[k], cp.p + ix);
        __builtin_nontemporal_store(px[k] + cp.alpha * po[k], cp.x + ix);
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < LX; k++) {
      uc[k] = active ? __builtin_nontemporal_load(u + base + k * P) : T(0);
      su[q][k * P + t] = uc[k];
    }
  }
  for (int x = t; x < 36; x += P) sedge[q][x] = active ? ed[x] : T(0);

  // rows D(i,:), D(j,:) in registers; the columns D(:,i), D(:,j) of the
  // transposed contraction are read from LDS (in registers they would cost
  // a wave per SIMD, which is worth more than the LDS reads)
  __shared__ T sD[LX * LX];
  if (q == 0) sD[t] = dx[t];
  T Di[LX], Dj[LX];
#pragma unroll
  for (int l = 0; l < LX; l++) {
    Di[l] = dx[i + LX * l];
    Dj[l] = dx[j + LX * l];
  }
  T wc[LX];
#pragma unroll
  for (int m = 0; m < LX; m++) wc[m] = T(0);
  __syncthreads();

  // Geometry.  dX/dt is bilinear in (r,s), so constant along a thread's
  // column: kept in registers.  dX/dr is bilinear in (s,t) and dX/ds in
  // (r,t): tabulated per element in LDS, one (plane, row) pair per thread,
  // sa[plane k][row j] = dX/dr(s_j, t_k), sb[plane k][row i] = dX/ds(r_i,
  // t_k), and read per plane (broadcast reads).
  __shared__ T sa[EPB][LX][LX][3];
  __shared__ T sb[EPB][LX][LX][3];
  const T *E = sedge[q];
  {
    const T p1 = (T(1) + zg[j]) * T(0.5), p0 = T(1) - p1;  // row coordinate
    const T t1 = (T(1) + zg[i]) * T(0.5), t0 = T(1) - t1;  // plane coordinate
#pragma unroll
    for (int c = 0; c < 3; c++) {
      sa[q][i][j][c] = (E[c * 4 + 0] * p0 + E[c * 4 + 1] * p1) * t0 +
                       (E[c * 4 + 2] * p0 + E[c * 4 + 3] * p1) * t1;
      sb[q][i][j][c] = (E[12 + c * 4 + 0] * p0 + E[12 + c * 4 + 1] * p1) * t0 +
                       (E[12 + c * 4 + 2] * p0 + E[12 + c * 4 + 3] * p1) * t1;
    }
  }
  const T ri1 = (T(1) + zg[i]) * T(0.5), ri0 = T(1) - ri1;
  const T sj1 = (T(1) + zg[j]) * T(0.5), sj0 = T(1) - sj1;
  const T wij = wg[i] * wg[j];
  const T h1w = h1c * wij, h2w = h2c * wij;
  const bool mass = h2c != T(0);
  T ct[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    ct[c] = (E[24 + c * 4 + 0] * ri0 + E[24 + c * 4 + 1] * ri1) * sj0 +
            (E[24 + c * 4 + 2] * ri0 + E[24 + c * 4 + 3] * ri1) * sj1;
  __syncthreads();

  // FP64: two planes per trip lets the scheduler overlap one plane's
  // geometry with the other's LDS traffic (168 VGPRs, 3 waves/SIMD); FP32
  // keeps one plane per trip (83 VGPRs, 5 waves/SIMD)
  constexpr int kPlaneUnroll = std::is_same<T, float>::value ? 1 : 2;
#pragma unroll kPlaneUnroll
  for (int k = 0; k < LX; k++) {
    const int b = k & 1;
    const T a0 = sa[q][k][j][0], a1 = sa[q][k][j][1], a2 = sa[q][k][j][2];
    const T b0 = sb[q][k][i][0], b1 = sb[q][k][i][1], b2 = sb[q][k][i][2];
    // rows of the cofactor matrix (Neko's drdx,drdy,drdz / dsdx.. / dtdx..):
    // r = X_s x X_t, s = X_t x X_r, t = X_r x X_s; jac = X_t . (X_r x X_s)
    const T rx = b1 * ct[2] - ct[1] * b2;
    const T ry = ct[0] * b2 - b0 * ct[2];
    const T rz = b0 * ct[1] - ct[0] * b1;
    const T sx = ct[1] * a2 - a1 * ct[2];
    const T sy = a0 * ct[2] - ct[0] * a2;
    const T sz = ct[0] * a1 - a0 * ct[1];
    const T tx = a1 * b2 - b1 * a2;
    const T ty = b0 * a2 - a0 * b2;
    const T tz = a0 * b1 - b0 * a1;
    const T jac = ct[0] * tx + ct[1] * ty + ct[2] * tz;
    const T sc = h1w * wg[k] * recip(jac);

    T dk[LX];
#pragma unroll
    for (int l = 0; l < LX; l++) dk[l] = dx[k + LX * l];
    T vt = T(0), ur = T(0), us = T(0);
    const T *plane = &su[q][k * P];
#pragma unroll
    for (int l = 0; l < LX; l++) {
      vt += dk[l] * uc[l];
      ur += Di[l] * plane[l + LX * j];
      us += Dj[l] * plane[i + LX * l];
    }
    // G (ur,us,ut) with G = sc C C^T: g = sc C^T (ur,us,ut), then C g
    const T gx = sc * (ur * rx + us * sx + vt * tx);
    const T gy = sc * (ur * ry + us * sy + vt * ty);
    const T gz = sc * (ur * rz + us * sz + vt * tz);
    const T rr = rx * gx + ry * gy + rz * gz;
    const T rs = sx * gx + sy * gy + sz * gz;
    const T rt = tx * gx + ty * gy + tz * gz;
    sr[q][b][t] = rr;
    ss[q][b][t] = rs;
#pragma unroll
    for (int m = 0; m < LX; m++) wc[m] += dk[m] * rt;
    __syncthreads();
    T acc = mass ? h2w * wg[k] * jac * plane[t] : T(0);
#pragma unroll
    for (int l = 0; l < LX; l++)
      acc += sD[l + LX * i] * sr[q][b][l + LX * j];
#pragma unroll
    for (int l = 0; l < LX; l++)
      acc += sD[l + LX * j] * ss[q][b][i + LX * l];
    // plane k of su is dead once every thread passed this plane's barrier
    // (its r/s reads come before it): the point's r/s part of w goes there
    su[q][k * P + t] = acc;
  }

  double dsum = 0.0;
  if (active) {
#pragma unroll
    for (int m = 0; m < LX; m++) {
      const T r = wc[m] + su[q][m * P + t];
      __builtin_nontemporal_store(r, w + base + m * P);
      if (DOT) dsum += (double)r * (double)uc[m];
    }
  }
  if (DOT) wg_sum_store<LX * LX * EPB>(dsum, sdot, q * P + t, dot_out);
}

template <typename T, int LX, int EPB, bool DOT, bool CGP>
__global__ void __launch_bounds__(LX *LX *EPB)
    ax_tri_kernel(T *__restrict__ w, const T *__restrict__ u,
                  const T *__restrict__ dx, const T *__restrict__ zg,
                  const T *__restrict__ wg, const T *__restrict__ edges, T h1c,
                  T h2c, int64_t nel, const int *__restrict__ skip,
                  double *__restrict__ dot_part, const T *__restrict__ cg_r,
                  const T *__restrict__ cg_dinv, T *__restrict__ cg_x,
                  const double *__restrict__ cg_beta, const double *__restrict__ cg_alpha) {
  if (skip != nullptr && *skip) return;
  const int64_t e = (int64_t)blockIdx.x * EPB + threadIdx.y;
  const bool active = e < nel;
  CgPro<T> cp{};
  if (CGP) cp = CgPro<T>{const_cast<T *>(u), cg_r, cg_dinv, cg_x, (T)*cg_beta, (T)*cg_alpha};
  ax_tri_body<T, LX, EPB, DOT, CGP>(w, CGP ? nullptr : u, dx, zg, wg,
                                    edges + (active ? e : 0) * 36, h1c,
                                    h2c, active, e * (int64_t)(LX * LX * LX),
                                    DOT ? dot_part + blockIdx.x : nullptr, cp);
}

// ---------------------------------------------------------------------------
// Ax with in-kernel trilinear geometry, contractions on the matrix cores
// (FP64, lx a multiple of 4).
//
// v_mfma_f64_4x4x4f64 computes, in each of four 4x4 blocks, C += A B with
// K = 4, where lane l holds C(row l/16, column l%4) of block (l%16)/4,
// B(k = l/16, column l%4) and A(row l%4, k = l/16).  A line of lx points
// (the values a derivative contracts over) therefore fits in the four lanes
// n, n+16, n+32, n+48 (n = l % 16), lane n + 16h holding the line's entries
// h, h+4, h+8, ..., and the product D x line lands in the same lanes with
// the same entries: D (or D^T) is the A operand, shared by the blocks, and
// sixteen lines go through one instruction.  Lines are indexed c = 16g + n,
// g < lx^2/16.  The t-lines (c = i + lx j) are how the element is loaded,
// kept and stored; the r-lines (c = j + lx k) and s-lines (c = i + lx k) go
// through LDS both ways.  Per element a derivative takes lx^4/256
// instructions (16 at lx = 8) and no vector FMA; the geometric factors of
// ax_tri_body (cofactors of the trilinear map, jac, h1 w3 / jac) are formed
// per point on the vector ALUs in the t-line layout.  One wave per element.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mfma_f64_4x4x4(double a, double b, double c) {
  return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// Exchange point between the lanes of a one-wave workgroup: a wave's LDS
// instructions execute in issue order, so ordering the accesses in the
// compiler (wavefront-scope fences) is enough; no wait for the LDS queue
// to drain, no s_barrier.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LX, bool MASS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
    ax_tri_mfma_kernel(double *__restrict__ w, const double *__restrict__ u,
                       const double *__restrict__ dx, const double *__restrict__ zg,
                       const double *__restrict__ wg, const double *__restrict__ edges,
                       double h1c, double h2c, int64_t nel, const int *__restrict__ skip) {
  static_assert(LX % 4 == 0 && LX <= 8, "lines of 4-entry chunks, registers for lx <= 8");
  constexpr int P = LX * LX, N3 = LX * P;
  constexpr int NL = LX / 4;   // entries of a line per lane
  constexpr int NG = P / 16;   // line groups (16 lines each)
  // element arrays in LDS with point (i,j,k) at (P + LX) k + LX j + (i ^ j):
  // the padded plane stride moves odd planes to the other half of the
  // banks and the xor spreads a row, so the 16 lanes that one LDS cycle
  // serves touch 16 distinct bank pairs along t-lines, r-lines and s-lines
  constexpr int NS = (P + LX) * LX;
  __shared__ double sa[NS];    // the element's u, then s-parts
  __shared__ double sb[NS];    // r-parts
  __shared__ double tr[LX][LX][3], ts[LX][LX][3];
  __shared__ double sedge[36], swg[LX];
  if (skip != nullptr && *skip) return;
  const int64_t e = blockIdx.x;
  if (e >= nel) return;  // whole wave: no exchange is skipped by part of it
  const int64_t base = e * (int64_t)N3;
  const int l = threadIdx.x;
  const int n = l & 15, h = l >> 4;
  // LDS positions of the lane's entries: t-lines (c = i + LX j, entry k),
  // r-lines (c = j + LX k, entry i), s-lines (c = i + LX k, entry j); each
  // position is owned by one lane in each layout
  auto pos = [&](int i, int j, int k) { return (P + LX) * k + LX * j + (i ^ j); };
  auto pt = [&](int g, int s) {
    const int c = 16 * g + n;
    return pos(c % LX, c / LX, h + 4 * s);
  };
  auto pr = [&](int g, int s) {
    const int c = 16 * g + n;
    return pos(h + 4 * s, c % LX, c / LX);
  };
  auto ps = [&](int g, int s) {
    const int c = 16 * g + n;
    return pos(c % LX, h + 4 * s, c / LX);
  };
  // offset in the element (global memory) of the lane's t-line entries
  auto gt = [&](int g, int s) { return 16 * g + n + P * (h + 4 * s); };

  if (l < LX) swg[l] = wg[l];
  // A operands: D(4 mb + m, 4 ks + kk) and D^T, m = l % 4, kk = l / 16
  double Ad[NL][NL], At[NL][NL];
#pragma unroll
  for (int mb = 0; mb < NL; mb++)
#pragma unroll
    for (int ks = 0; ks < NL; ks++) {
      Ad[mb][ks] = dx[(4 * mb + (l & 3)) + LX * (4 * ks + h)];
      At[mb][ks] = dx[(4 * ks + h) + LX * (4 * mb + (l & 3))];
    }
  double uc[NG][NL];
#pragma unroll
  for (int g = 0; g < NG; g++)
#pragma unroll
    for (int s = 0; s < NL; s++) {
      uc[g][s] = __builtin_nontemporal_load(u + base + gt(g, s));
      sa[pt(g, s)] = uc[g][s];
    }
  if (l < 36) sedge[l] = edges[e * 36 + l];
  wave_lds_sync();

  // geometry tables: tr[k][j] = dX/dr(s_j, t_k), ts[k][i] = dX/ds(r_i, t_k)
  // (bilinear interpolants of the half edge vectors: r-edges [0,12),
  // s-edges [12,24), t-edges [24,36), component c at c*4 + (low/high of the
  // first other coordinate) + 2*(of the second)); dX/dt(r_i, s_j) per lane
  for (int q = l; q < P; q += 64) {
    const int a = q % LX, b = q / LX;
    const double pa1 = (1.0 + zg[a]) * 0.5, pa0 = 1.0 - pa1;
    const double pb1 = (1.0 + zg[b]) * 0.5, pb0 = 1.0 - pb1;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      tr[b][a][c] = (sedge[c * 4 + 0] * pa0 + sedge[c * 4 + 1] * pa1) * pb0 +
                    (sedge[c * 4 + 2] * pa0 + sedge[c * 4 + 3] * pa1) * pb1;
      ts[b][a][c] = (sedge[12 + c * 4 + 0] * pa0 + sedge[12 + c * 4 + 1] * pa1) * pb0 +
                    (sedge[12 + c * 4 + 2] * pa0 + sedge[12 + c * 4 + 3] * pa1) * pb1;
    }
  }
  // the lane's t-line columns share i = n % LX (16 g is a multiple of LX):
  // the t-edges interpolated in r once, then in s per column
  double tlo[3], thi[3];
  {
    const int i = n % LX;
    const double r1 = (1.0 + zg[i]) * 0.5, r0 = 1.0 - r1;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      tlo[c] = sedge[24 + c * 4 + 0] * r0 + sedge[24 + c * 4 + 1] * r1;
      thi[c] = sedge[24 + c * 4 + 2] * r0 + sedge[24 + c * 4 + 3] * r1;
    }
  }

  // ur on the r-lines straight to sb; us on the s-lines, stored over u
  // once every lane has read its lines
  double usl[NG][NL];
#pragma unroll
  for (int g = 0; g < NG; g++) {
    double lr[NL], ls[NL];
#pragma unroll
    for (int s = 0; s < NL; s++) {
      lr[s] = sa[pr(g, s)];
      ls[s] = sa[ps(g, s)];
    }
#pragma unroll
    for (int mb = 0; mb < NL; mb++) {
      double cr = 0.0, cs = 0.0;
#pragma unroll
      for (int ks = 0; ks < NL; ks++) {
        cr = mfma_f64_4x4x4(Ad[mb][ks], lr[ks], cr);
        cs = mfma_f64_4x4x4(Ad[mb][ks], ls[ks], cs);
      }
      sb[pr(g, mb)] = cr;
      usl[g][mb] = cs;
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int g = 0; g < NG; g++)
#pragma unroll
    for (int s = 0; s < NL; s++) sa[ps(g, s)] = usl[g][s];
  wave_lds_sync();

  // per t-line group: ut on the matrix cores, the geometry per point, the
  // r- and s-parts over ur/us (the lane's own positions), the t-part
  // contracted at once (its accumulator starts at the mass term)
  double wt[NG][NL];
#pragma unroll
  for (int g = 0; g < NG; g++) {
    double vt[NL], rt[NL], m0[NL], cv[3], h1ij, h2ij;
    {
      const int c = 16 * g + n, i = c % LX, j = c / LX;
      const double s1 = (1.0 + zg[j]) * 0.5, s0 = 1.0 - s1;
#pragma unroll
      for (int q = 0; q < 3; q++) cv[q] = tlo[q] * s0 + thi[q] * s1;
      const double wij = swg[i] * swg[j];
      h1ij = h1c * wij;
      h2ij = h2c * wij;
    }
#pragma unroll
    for (int mb = 0; mb < NL; mb++) {
      double ct = 0.0;
#pragma unroll
      for (int ks = 0; ks < NL; ks++) ct = mfma_f64_4x4x4(Ad[mb][ks], uc[g][ks], ct);
      vt[mb] = ct;
    }
#pragma unroll
    for (int s = 0; s < NL; s++) {
      const int c = 16 * g + n, i = c % LX, j = c / LX, k = h + 4 * s;
      const int p = pt(g, s);
      const double vr = sb[p], vs = sa[p];
      const double a0 = tr[k][j][0], a1 = tr[k][j][1], a2 = tr[k][j][2];
      const double b0 = ts[k][i][0], b1 = ts[k][i][1], b2 = ts[k][i][2];
      const double c0 = cv[0], c1 = cv[1], c2 = cv[2];
      const double rx = b1 * c2 - c1 * b2;
      const double ry = c0 * b2 - b0 * c2;
      const double rz = b0 * c1 - c0 * b1;
      const double sx = c1 * a2 - a1 * c2;
      const double sy = a0 * c2 - c0 * a2;
      const double sz = c0 * a1 - a0 * c1;
      const double tx = a1 * b2 - b1 * a2;
      const double ty = b0 * a2 - a0 * b2;
      const double tz = a0 * b1 - b0 * a1;
      const double jac = c0 * tx + c1 * ty + c2 * tz;
      const double wk = swg[k];
      const double sc = h1ij * wk * recip(jac);
      const double gx = sc * (vr * rx + vs * sx + vt[s] * tx);
      const double gy = sc * (vr * ry + vs * sy + vt[s] * ty);
      const double gz = sc * (vr * rz + vs * sz + vt[s] * tz);
      sb[p] = rx * gx + ry * gy + rz * gz;
      sa[p] = sx * gx + sy * gy + sz * gz;
      rt[s] = tx * gx + ty * gy + tz * gz;
      m0[s] = MASS ? h2ij * wk * jac * uc[g][s] : 0.0;
    }
#pragma unroll
    for (int mb = 0; mb < NL; mb++) {
      double acc = m0[mb];
#pragma unroll
      for (int ks = 0; ks < NL; ks++) acc = mfma_f64_4x4x4(At[mb][ks], rt[ks], acc);
      wt[g][mb] = acc;
    }
  }
  wave_lds_sync();
  // D^T contractions of the r-parts (r-lines) and s-parts (s-lines), the
  // results over their inputs (the lane's own positions)
#pragma unroll
  for (int g = 0; g < NG; g++) {
    double lr[NL], ls[NL];
#pragma unroll
    for (int s = 0; s < NL; s++) {
      lr[s] = sb[pr(g, s)];
      ls[s] = sa[ps(g, s)];
    }
#pragma unroll
    for (int mb = 0; mb < NL; mb++) {
      double cr = 0.0, cs = 0.0;
#pragma unroll
      for (int ks = 0; ks < NL; ks++) {
        cr = mfma_f64_4x4x4(At[mb][ks], lr[ks], cr);
        cs = mfma_f64_4x4x4(At[mb][ks], ls[ks], cs);
      }
      sb[pr(g, mb)] = cr;
      sa[ps(g, mb)] = cs;
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int g = 0; g < NG; g++)
#pragma unroll
    for (int s = 0; s < NL; s++) {
      const int p = pt(g, s);
      __builtin_nontemporal_store(wt[g][s] + (sb[p] + sa[p]), w + base + gt(g, s));
    }
}

// Batched: workgroup b of a launch is global workgroup blk_base + b, which
// belongs to instance blk_inst[.] (instances never share a workgroup); the
// instance's flag `done` (CgScal-compatible layout: int at done_off bytes)
// turns its workgroups into no-ops.
template <int LX, int EPB, bool CGP>
__global__ void __launch_bounds__(LX *LX *EPB)
    ax_tri_batch_kernel(float *__restrict__ w, const float *__restrict__ u,
                        const float *__restrict__ dx, const float *__restrict__ zg,
                        const float *__restrict__ wg, const BatchInst *__restrict__ inst,
                        const int32_t *__restrict__ blk_inst, int64_t blk_base,
                        const char *__restrict__ scal, int scal_size, int done_off,
                        double *__restrict__ dot_part, const float *__restrict__ cg_r,
                        const float *__restrict__ cg_dinv, float *__restrict__ cg_x,
                        int beta_off, int alpha_off) {
  const int64_t gb = blk_base + blockIdx.x;
  const int ii = blk_inst[gb];
  const char *rec = scal + (int64_t)ii * scal_size;
  if (*(const int *)(rec + done_off)) return;
  const BatchInst I = inst[ii];
  const int64_t e = (gb - I.blk0) * EPB + threadIdx.y;
  const bool active = e < I.nel;
  CgPro<float> cp{};
  if (CGP)
    cp = CgPro<float>{const_cast<float *>(u), cg_r, cg_dinv, cg_x,
                      (float)*(const double *)(rec + beta_off),
                      (float)*(const double *)(rec + alpha_off)};
  ax_tri_body<float, LX, EPB, true, CGP>(w, CGP ? nullptr : u, dx, zg, wg,
                                         I.edges + (active ? e : 0) * 36, I.h1c, I.h2c, active,
                                         I.voff + e * (int64_t)(LX * LX * LX), dot_part + gb,
                                         cp);
}

template <int LX>
static int ax_batch_lx(const AxBatchArgs &a, hipStream_t s) {
  constexpr int P = LX * LX;
#define NM_AXB_CASE(E)                                                                   \
  case E:                                                                                \
    if constexpr (P * E <= 256) {                                                        \
      if (a.cg_r)                                                                        \
        hipLaunchKernelGGL((ax_tri_batch_kernel<LX, E, true>), dim3((unsigned)a.nblocks),  \
                           dim3(P, E), 0, s, a.w, a.u, a.dx, a.zg, a.wg, a.inst,         \
                           a.blk_inst, a.blk_base, (const char *)a.scal, a.scal_size,    \
                           a.done_off, a.dot_part, a.cg_r, a.cg_dinv, a.cg_x,            \
                           a.beta_off, a.alpha_off);                                     \
      else                                                                               \
        hipLaunchKernelGGL((ax_tri_batch_kernel<LX, E, false>), dim3((unsigned)a.nblocks), \
                           dim3(P, E), 0, s, a.w, a.u, a.dx, a.zg, a.wg, a.inst,         \
                           a.blk_inst, a.blk_base, (const char *)a.scal, a.scal_size,    \
                           a.done_off, a.dot_part, nullptr, nullptr, nullptr, 0, 0);     \
      NM_HIP_TRY(hipGetLastError());                                                     \
      return NM_OK;                                                                      \
    }                                                                                    \
    break;
  switch (a.epb) {
    NM_AXB_CASE(1) NM_AXB_CASE(2) NM_AXB_CASE(4) NM_AXB_CASE(8) NM_AXB_CASE(16)
    default: break;
  }
#undef NM_AXB_CASE
  set_error("ax batch: elements-per-workgroup %d unsupported for lx=%d", a.epb, LX);
  return NM_ERR_UNSUPPORTED;
}

int launch_ax_batch(const AxBatchArgs &a, hipStream_t s) {
  if (a.nblocks <= 0) return NM_OK;
  switch (a.lx) {
    case 2: return ax_batch_lx<2>(a, s);
    case 3: return ax_batch_lx<3>(a, s);
    case 4: return ax_batch_lx<4>(a, s);
    case 5: return ax_batch_lx<5>(a, s);
    case 6: return ax_batch_lx<6>(a, s);
    case 7: return ax_batch_lx<7>(a, s);
    case 8: return ax_batch_lx<8>(a, s);
    case 9: return ax_batch_lx<9>(a, s);
    case 10: return ax_batch_lx<10>(a, s);
    case 11: return ax_batch_lx<11>(a, s);
    case 12: return ax_batch_lx<12>(a, s);
    default:
      set_error("ax batch: lx=%d unsupported (2..12)", a.lx);
      return NM_ERR_UNSUPPORTED;
  }
}

template <typename T, int LX, int EPB>
static int ax_tri_go(const AxArgs<T> &a, hipStream_t s) {
  dim3 block(LX * LX, EPB);
  dim3 grid((unsigned)((a.nel + EPB - 1) / EPB));
  if (a.nel <= 0) return NM_OK;
  if (a.cg_r) {
    if (!a.dot_part || !a.cg_dinv || !a.cg_x || !a.cg_beta || !a.cg_alpha) {
      set_error("ax: fused CG prologue needs dot partials, dinv, x and scalars");
      return NM_ERR_ARG;
    }
    hipLaunchKernelGGL((ax_tri_kernel<T, LX, EPB, true, true>), grid, block, 0, s, a.w, a.u,
                       a.dx, a.zg, a.wg, a.edges, a.h1c, a.h2c, a.nel, a.skip, a.dot_part,
                       a.cg_r, a.cg_dinv, a.cg_x, a.cg_beta, a.cg_alpha);
  } else if (a.dot_part) {
    hipLaunchKernelGGL((ax_tri_kernel<T, LX, EPB, true, false>), grid, block, 0, s, a.w, a.u,
                       a.dx, a.zg, a.wg, a.edges, a.h1c, a.h2c, a.nel, a.skip, a.dot_part,
                       nullptr, nullptr, nullptr, nullptr, nullptr);
  } else {
    hipLaunchKernelGGL((ax_tri_kernel<T, LX, EPB, false, false>), grid, block, 0, s, a.w, a.u,
                       a.dx, a.zg, a.wg, a.edges, a.h1c, a.h2c, a.nel, a.skip, a.dot_part,
                       nullptr, nullptr, nullptr, nullptr, nullptr);
  }
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

template <typename T, int LX, int EPB, bool PF>
static int ax_go(const AxArgs<T> &a, hipStream_t s) {
  dim3 block(LX * LX, EPB);
  dim3 grid((unsigned)((a.nel + EPB - 1) / EPB));
  if (a.nel <= 0) return NM_OK;
  if (a.cg_r) {
    if (!a.dot_part || !a.cg_dinv || !a.cg_x || !a.cg_beta || !a.cg_alpha) {
      set_error("ax: fused CG prologue needs dot partials, dinv, x and scalars");
      return NM_ERR_ARG;
    }
    hipLaunchKernelGGL((ax_kstep_kernel<T, LX, EPB, PF, true>), grid, block, 0, s, a.w, a.u,
                       a.dx, a.h1, a.h1c, a.g[0], a.g[1], a.g[2], a.g[3], a.g[4], a.g[5],
                       a.h2b, a.nel, a.skip, a.dot_part, a.cg_r, a.cg_dinv, a.cg_x,
                       a.cg_beta, a.cg_alpha);
  } else if (a.dot_part) {
    set_error("ax: the stored-geometry variants fuse p.Ap only with the CG prologue");
    return NM_ERR_UNSUPPORTED;
  } else {
    hipLaunchKernelGGL((ax_kstep_kernel<T, LX, EPB, PF, false>), grid, block, 0, s, a.w, a.u,
                       a.dx, a.h1, a.h1c, a.g[0], a.g[1], a.g[2], a.g[3], a.g[4], a.g[5],
                       a.h2b, a.nel, a.skip, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr);
  }
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// Variant code: epb = elements per workgroup (1,2,4,8,16; 0 = default),
// plus 100 for the G-prefetching (software-pipelined) variant.
static inline int ax_epb_of(int lx, int code) {
  int epb = code % 100;
  if (epb <= 0) {
    epb = 1;
    while (lx * lx * epb * 2 <= 64 && epb < 16) epb *= 2;
  }
  return epb;
}

int64_t ax_grid(int64_t nel, int lx, int code) {
  const int epb = ax_epb_of(lx, code);
  return (nel + epb - 1) / epb;
}

template <typename T, int LX>
static int ax_epb(const AxArgs<T> &a, int code, hipStream_t s) {
  constexpr int P = LX * LX;
  const bool pf = code >= 100 && code < 200;
  const bool tri = code >= 200;
  const int epb = ax_epb_of(LX, code);
  if (tri != (a.edges != nullptr)) {
    set_error("ax: variant %d needs %s geometry", code, tri ? "trilinear" : "stored");
    return NM_ERR_UNSUPPORTED;
  }
  if (tri) {
    switch (epb) {
#define NM_AXT_CASE(E)                                 \
  case E:                                              \
    if constexpr (P * E <= 256) return ax_tri_go<T, LX, E>(a, s); \
    break;
      NM_AXT_CASE(1) NM_AXT_CASE(2) NM_AXT_CASE(4) NM_AXT_CASE(8) NM_AXT_CASE(16)
#undef NM_AXT_CASE
      default: break;
    }
    set_error("ax: elements-per-workgroup %d unsupported for lx=%d", epb, LX);
    return NM_ERR_UNSUPPORTED;
  }
#define NM_AX_CASE(E)                                         \
  case E:                                                     \
    if constexpr (P * E <= 256) {                             \
      return pf ? ax_go<T, LX, E, true>(a, s)                 \
                : ax_go<T, LX, E, false>(a, s);               \
    }                                                         \
    break;
  switch (epb) {
    NM_AX_CASE(1) NM_AX_CASE(2) NM_AX_CASE(4) NM_AX_CASE(8) NM_AX_CASE(16)
    default: break;
  }
#undef NM_AX_CASE
  set_error("ax: elements-per-workgroup %d unsupported for lx=%d", epb, LX);
  return NM_ERR_UNSUPPORTED;
}

// Variant 300 + epb: the matrix-core kernel where it applies (FP64, lx 4 or
// 8, no fused dot or CG prologue), else the vector kernel 200 + epb.
static bool ax_mfma_applies(int lx, bool fp64, bool fused) {
  return fp64 && (lx == 4 || lx == 8) && !fused;
}

template <typename T>
int launch_ax(const AxArgs<T> &a, int epb, hipStream_t s) {
  if (epb >= 300) {
    const bool fused = a.dot_part != nullptr || a.cg_r != nullptr;
    if (!ax_mfma_applies(a.lx, std::is_same<T, double>::value, fused)) {
      epb -= 100;
    } else if constexpr (std::is_same<T, double>::value) {
      if (a.nel <= 0) return NM_OK;
      if (!a.edges || !a.zg || !a.wg) {
        set_error("ax: the matrix-core variant needs trilinear geometry");
        return NM_ERR_ARG;
      }
      const dim3 grid((unsigned)a.nel), block(64);
      const bool mass = a.h2c != 0.0;
      if (a.lx == 8 && mass)
        hipLaunchKernelGGL((ax_tri_mfma_kernel<8, true>), grid, block, 0, s, a.w, a.u, a.dx,
                           a.zg, a.wg, a.edges, a.h1c, a.h2c, a.nel, a.skip);
      else if (a.lx == 8)
        hipLaunchKernelGGL((ax_tri_mfma_kernel<8, false>), grid, block, 0, s, a.w, a.u, a.dx,
                           a.zg, a.wg, a.edges, a.h1c, a.h2c, a.nel, a.skip);
      else if (mass)
        hipLaunchKernelGGL((ax_tri_mfma_kernel<4, true>), grid, block, 0, s, a.w, a.u, a.dx,
                           a.zg, a.wg, a.edges, a.h1c, a.h2c, a.nel, a.skip);
      else
        hipLaunchKernelGGL((ax_tri_mfma_kernel<4, false>), grid, block, 0, s, a.w, a.u, a.dx,
                           a.zg, a.wg, a.edges, a.h1c, a.h2c, a.nel, a.skip);
      NM_HIP_TRY(hipGetLastError());
      return NM_OK;
    }
  }
  switch (a.lx) {
    case 2: return ax_epb<T, 2>(a, epb, s);
    case 3: return ax_epb<T, 3>(a, epb, s);
    case 4: return ax_epb<T, 4>(a, epb, s);
    case 5: return ax_epb<T, 5>(a, epb, s);
    case 6: return ax_epb<T, 6>(a, epb, s);
    case 7: return ax_epb<T, 7>(a, epb, s);
    case 8: return ax_epb<T, 8>(a, epb, s);
    case 9: return ax_epb<T, 9>(a, epb, s);
    case 10: return ax_epb<T, 10>(a, epb, s);
    case 11: return ax_epb<T, 11>(a, epb, s);
    case 12: return ax_epb<T, 12>(a, epb, s);
    case 13: return ax_epb<T, 13>(a, epb, s);
    case 14: return ax_epb<T, 14>(a, epb, s);
    default:
      set_error("ax: lx=%d unsupported (2..14)", a.lx);
      return NM_ERR_UNSUPPORTED;
  }
}
template int launch_ax<double>(const AxArgs<double> &, int, hipStream_t);
template int launch_ax<float>(const AxArgs<float> &, int, hipStream_t);

// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs (each with its own L2), so blockIdx b runs on XCD b % 8.  Remapping
// b -> (b % 8) * per + b / 8 hands every XCD one contiguous range of the
// logical work, so the element cache lines a range of gs entries shares stay
// in one L2 instead of being fetched by all eight.  The launch must use a
// grid of 8 * per blocks; logical blocks >= the real count exit.
__device__ inline int64_t xcd_block(int64_t per) {
  const int64_t b = blockIdx.x;
  return (b % NM_NXCD) * per + b / NM_NXCD;
}
static inline int64_t xcd_per(int64_t nblocks) {
  return (nblocks + NM_NXCD - 1) / NM_NXCD;
}

// ---------------------------------------------------------------------------
// Gather-scatter ADD, fused: one thread per gs block (a run of local dofs
// with the same gs id) or per facet pair.  Block sums are taken in the
// order of Neko's sorted map (tmp = u1; tmp += u2; ...), which makes the
// result bit-identical to gs_gather_kernel_add + gs_scatter_kernel
// (gs_cpu.f90:98-138, :292-322).  No atomics, no intermediate v buffer.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
    gs_fused_kernel(T *__restrict__ u, const int32_t *__restrict__ blk_ptr,
                    const int32_t *__restrict__ gidx, int64_t nblks,
                    int64_t pair0, int64_t npairs,
                    const int *__restrict__ skip, int64_t per) {
  if (skip != nullptr && *skip) return;
  const int64_t t = xcd_block(per) * blockDim.x + threadIdx.x;
  if (t < nblks) {
    const int32_t s0 = blk_ptr[t], s1 = blk_ptr[t + 1];
    T acc = u[gidx[s0]];
    for (int32_t q = s0 + 1; q < s1; q++) acc += u[gidx[q]];
    for (int32_t q = s0; q < s1; q++) u[gidx[q]] = acc;
  } else if (t < nblks + npairs) {
    const int64_t p = pair0 + 2 * (t - nblks);
    const int32_t a = gidx[p], b = gidx[p + 1];
    const T acc = u[a] + u[b];
    u[a] = acc;
    u[b] = acc;
  }
}

// Lane-parallel variant: one 8-lane group per block (blocks of at most 8
// entries) and one lane per facet-pair entry, so the index loads are
// coalesced and every entry's value load is issued by its own lane.  The
// group sums its values in map order through lane shuffles (same order as
// the thread-per-block kernel, hence the same bits).
// done flag of the problem that owns logical workgroup lb (batched gs)
__device__ inline bool wg_done(const int32_t *__restrict__ wg, const char *__restrict__ scal,
                               int scal_size, int done_off, int64_t lb) {
  return wg != nullptr && *(const int *)(scal + (int64_t)wg[lb] * scal_size + done_off) != 0;
}

template <typename T>
__global__ void __launch_bounds__(256)
    gs_blk8_kernel(T *__restrict__ u, const int32_t *__restrict__ blk_ptr,
                   const int32_t *__restrict__ gidx, int64_t nblks,
                   const int *__restrict__ skip, int64_t per, const int32_t *__restrict__ wg,
                   const char *__restrict__ scal, int scal_size, int done_off,
                   const int32_t *__restrict__ wg_map, int64_t nmap) {
  if (skip != nullptr && *skip) return;
  int64_t lb = xcd_block(per);
  if (wg_map != nullptr) {
    if (lb >= nmap) return;
    lb = wg_map[lb];
  }
  if (wg_done(wg, scal, scal_size, done_off, lb)) return;
  const int64_t t = lb * blockDim.x + threadIdx.x;
  const int64_t g = t >> 3;
  const int m = (int)(t & 7);
  const int lane0 = (int)(threadIdx.x & 63) & ~7;
  int32_t s0 = 0, len = 0, idx = 0;
  T v = T(0);
  if (g < nblks) {
    s0 = blk_ptr[g];
    len = blk_ptr[g + 1] - s0;
    if (m < len) {
      idx = gidx[s0 + m];
      v = u[idx];
    }
  }
  T acc = __shfl(v, lane0, 64);
#pragma unroll
  for (int q = 1; q < 8; q++) {
    const T o = __shfl(v, lane0 + q, 64);
    if (q < len) acc += o;
  }
  if (g < nblks && m < len) __builtin_nontemporal_store(acc, u + idx);
}

// Blocks of <= 8 entries as packed segments: one lane per entry (the
// lane-per-slot form above leaves about half its lanes idle on edge
// blocks of 4), a block's entries on consecutive lanes of one wavefront;
// each lane gathers its block's values in block order, so the sum is the
// same bits as the other block kernels'.
template <typename T>
__global__ void __launch_bounds__(256)
    gs_seg_kernel(T *__restrict__ u, const int32_t *__restrict__ seg_idx,
                  const uint8_t *__restrict__ seg_pl, int64_t nseg,
                  const int *__restrict__ skip, int64_t per, const int32_t *__restrict__ wg,
                  const char *__restrict__ scal, int scal_size, int done_off,
                  const int32_t *__restrict__ wg_map, int64_t nmap) {
  if (skip != nullptr && *skip) return;
  int64_t lb = xcd_block(per);
  if (wg_map != nullptr) {
    if (lb >= nmap) return;
    lb = wg_map[lb];
  }
  if (wg_done(wg, scal, scal_size, done_off, lb)) return;
  const int64_t t = lb * blockDim.x + threadIdx.x;
  int32_t idx = -1;
  int off = 0, len = 0;
  if (t < nseg) {
    idx = seg_idx[t];
    const int pl = seg_pl[t];
    off = pl & 15;
    len = pl >> 4;
  }
  const T v = idx >= 0 ? u[idx] : T(0);
  const int lane0 = (int)(threadIdx.x & 63) - off;
  T acc = __shfl(v, lane0, 64);
#pragma unroll
  for (int q = 1; q < 8; q++) {
    const T o = __shfl(v, lane0 + q, 64);
    if (q < len) acc += o;
  }
  if (idx >= 0) __builtin_nontemporal_store(acc, u + idx);
}

template <typename T>
__global__ void __launch_bounds__(256)
    gs_pair_kernel(T *__restrict__ u, const int32_t *__restrict__ gidx,
                   int64_t pair0, int64_t npairs, const int *__restrict__ skip,
                   int64_t per, const int32_t *__restrict__ wg, const char *__restrict__ scal,
                   int scal_size, int done_off) {
  if (skip != nullptr && *skip) return;
  const int64_t lb = xcd_block(per);
  if (wg_done(wg, scal, scal_size, done_off, lb)) return;
  const int64_t t = lb * blockDim.x + threadIdx.x;
  bool act = t < 2 * npairs;
  int32_t idx = 0;
  T v = T(0);
  if (act) {
    idx = gidx[pair0 + t];
    act = idx >= 0;  // padding pairs (batched gs) are (-1, -1)
    if (act) v = u[idx];
  }
  const T o = __shfl_xor(v, 1, 64);
  if (act) __builtin_nontemporal_store(v + o, u + idx);  // a + b == b + a exactly
}

// Facet pairs of one element face shared by two elements, from a
// descriptor: thread (d, q) handles interior face point q = (u-1) + L (v-1)
// of descriptor d; its partner in element b is the point the descriptor's
// square symmetry maps (u, v) to.  Both get u_a + u_b (== u_b + u_a
// exactly, so the sum is the same bits as the pair kernel's).  The indices
// are computed, not loaded: no index traffic and no dependent load.
template <typename T>
__global__ void __launch_bounds__(256)
    gs_face_kernel(T *__restrict__ u, const int32_t *__restrict__ faces, int64_t nfaces,
                   int lx, const int *__restrict__ skip, int64_t per) {
  if (skip != nullptr && *skip) return;
  const int L = lx - 2, LL = L * L;
  const int64_t t = xcd_block(per) * blockDim.x + threadIdx.x;
  const int64_t d = t / LL;
  if (d >= nfaces) return;
  const int q = (int)(t - d * LL);
  const int uu = q % L + 1, vv = q / L + 1;
  const int32_t ea = faces[3 * d], eb = faces[3 * d + 1], code = faces[3 * d + 2];
  const int fa = code & 7, fb = (code >> 3) & 7;
  int ub = (code & 64) ? vv : uu, vb = (code & 64) ? uu : vv;
  if (code & 128) ub = L + 1 - ub;
  if (code & 256) vb = L + 1 - vb;
  const int64_t n3 = (int64_t)lx * lx * lx;
  const int64_t ia = ea * n3 + face_point(fa, uu, vv, lx);
  const int64_t ib = eb * n3 + face_point(fb, ub, vb, lx);
  const T acc = u[ia] + u[ib];
  __builtin_nontemporal_store(acc, u + ia);
  __builtin_nontemporal_store(acc, u + ib);
}

__global__ void mult_count_kernel(const double *__restrict__ mult, uint8_t *__restrict__ cnt,
                                  int *__restrict__ bad, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double m = mult[i];
    const double c = m > 0.0 ? rint(1.0 / m) : 0.0;
    const bool ok = c >= 1.0 && c <= 255.0 && 1.0 / c == m;
    cnt[i] = ok ? (uint8_t)c : (uint8_t)0;
    if (!ok) *bad = 1;
  }
}

int launch_mult_count(const double *mult, uint8_t *cnt, int *bad, int64_t n, hipStream_t s) {
  if (n <= 0) return NM_OK;
  const int64_t nb = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(mult_count_kernel, dim3((unsigned)nb), dim3(256), 0, s, mult, cnt, bad, n);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

template <typename T>
int launch_gs(const GsArgs<T> &a, hipStream_t s) {
  const int64_t items = a.nblks + a.npairs;
  if (items <= 0) return NM_OK;
  const int bs = 256;
  if (a.variant == 2 || a.variant == 3) {
    if (a.variant == 3 && a.nseg > 0) {
      const int64_t per = xcd_per((a.nseg + bs - 1) / bs);
      hipLaunchKernelGGL(gs_seg_kernel<T>, dim3((unsigned)(NM_NXCD * per)), dim3(bs), 0, s,
                         a.u, a.seg_idx, a.seg_pl, a.nseg, a.skip, per, nullptr, nullptr, 0, 0,
                         nullptr, (int64_t)0);
    } else if (a.nblks > 0) {
      const int64_t per = xcd_per((8 * a.nblks + bs - 1) / bs);
      hipLaunchKernelGGL(gs_blk8_kernel<T>, dim3((unsigned)(NM_NXCD * per)), dim3(bs), 0, s,
                         a.u, a.blk_ptr, a.gidx, a.nblks, a.skip, per, nullptr, nullptr, 0, 0,
                         nullptr, (int64_t)0);
    }
    if (a.nfaces > 0) {
      const int64_t LL = (int64_t)(a.lx - 2) * (a.lx - 2);
      const int64_t per = xcd_per((a.nfaces * LL + bs - 1) / bs);
      hipLaunchKernelGGL(gs_face_kernel<T>, dim3((unsigned)(NM_NXCD * per)), dim3(bs), 0, s,
                         a.u, a.faces, a.nfaces, a.lx, a.skip, per);
    }
    if (a.nrest > 0) {
      const int64_t per = xcd_per((2 * a.nrest + bs - 1) / bs);
      hipLaunchKernelGGL(gs_pair_kernel<T>, dim3((unsigned)(NM_NXCD * per)), dim3(bs), 0, s,
                         a.u, a.rest, (int64_t)0, a.nrest, a.skip, per, nullptr, nullptr, 0, 0);
    }
  } else if (a.variant == 1) {
    if (a.seg_idx != nullptr) {
      // batched: the blocks as packed segments, whole workgroups of 256
      // lanes per problem (wg_blk: problem of each), optionally only the
      // workgroups in wg_map
      if (a.nseg > 0 && (a.wg_map == nullptr || a.nwg_map > 0)) {
        const int64_t per = a.wg_map ? xcd_per(a.nwg_map) : xcd_per((a.nseg + bs - 1) / bs);
        hipLaunchKernelGGL(gs_seg_kernel<T>, dim3((unsigned)(NM_NXCD * per)), dim3(bs), 0, s,
                           a.u, a.seg_idx, a.seg_pl, a.nseg, a.skip, per, a.wg_blk, a.scal,
                           a.scal_size, a.done_off, a.wg_map, a.nwg_map);
      }
    } else if (a.nblks > 0 && (a.wg_map == nullptr || a.nwg_map > 0)) {
      const int64_t per = a.wg_map ? xcd_per(a.nwg_map) : xcd_per((8 * a.nblks + bs - 1) / bs);
      hipLaunchKernelGGL(gs_blk8_kernel<T>, dim3((unsigned)(NM_NXCD * per)), dim3(bs), 0, s,
                         a.u, a.blk_ptr, a.gidx, a.nblks, a.skip, per, a.wg_blk, a.scal,
                         a.scal_size, a.done_off, a.wg_map, a.nwg_map);
    }
    if (a.npairs > 0) {
      const int64_t per = xcd_per((2 * a.npairs + bs - 1) / bs);
      hipLaunchKernelGGL(gs_pair_kernel<T>, dim3((unsigned)(NM_NXCD * per)), dim3(bs), 0, s,
                         a.u, a.gidx, a.pair0, a.npairs, a.skip, per, a.wg_pair, a.scal,
                         a.scal_size, a.done_off);
    }
  } else {
    const int64_t per = xcd_per((items + bs - 1) / bs);
    hipLaunchKernelGGL(gs_fused_kernel<T>, dim3((unsigned)(NM_NXCD * per)), dim3(bs), 0, s,
                       a.u, a.blk_ptr, a.gidx, a.nblks, a.pair0, a.npairs, a.skip, per);
  }
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}
template int launch_gs<double>(const GsArgs<double> &, hipStream_t);
template int launch_gs<float>(const GsArgs<float> &, hipStream_t);

// Neko-layout gather: v(dg(k+1)) = op over u(gd(k+j)) for each block, then
// the facet pairs (o > 0) or the plain shared copy (o < 0); gs_cpu.f90
// semantics for ADD/MUL/MIN/MAX (gs_ops.f90:35-36).  One thread per block
// and one per facet pair / copied entry.
template <typename T>
__device__ inline T gs_apply(int op, T a, T b) {
  switch (op) {
    case 2: return a * b;
    case 3: return a < b ? a : b;
    case 4: return a > b ? a : b;
    default: return a + b;
  }
}

template <typename T>
__global__ void neko_gather_kernel(T *__restrict__ v, int m, int o,
                                   const int *__restrict__ dg,
                                   const T *__restrict__ u,
                                   const int *__restrict__ gd, int nb,
                                   const int *__restrict__ b,
                                   const int *__restrict__ bo, int op) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nb) {
    const int k = bo[t], len = b[t];
    T acc = u[gd[k] - 1];
    for (int q = 1; q < len; q++) acc = gs_apply(op, acc, u[gd[k + q] - 1]);
    v[dg[k] - 1] = acc;
    return;
  }
  const int64_t r = t - nb;
  if (o < 0) {
    const int64_t i = (int64_t)(-o) + r;  // 1-based
    if (i <= m) v[dg[i - 1] - 1] = u[gd[i - 1] - 1];
  } else {
    const int64_t i = (int64_t)o + 2 * r;  // 1-based, pairs
    if (i + 1 <= m) v[dg[i - 1] - 1] = gs_apply(op, u[gd[i - 1] - 1], u[gd[i] - 1]);
  }
}

template <typename T>
int launch_neko_gather(T *v, int m, int o, const int *dg, const T *u,
                       const int *gd, int nb, const int *b, const int *bo,
                       int op, hipStream_t s) {
  if (m <= 0) return NM_OK;
  int64_t tail = (o < 0) ? (int64_t)m - (-o) + 1 : ((int64_t)m - o + 1 + 1) / 2;
  if (tail < 0) tail = 0;
  const int64_t items = nb + tail;
  if (items <= 0) return NM_OK;
  hipLaunchKernelGGL(neko_gather_kernel<T>, dim3((unsigned)((items + 255) / 256)),
                     dim3(256), 0, s, v, m, o, dg, u, gd, nb, b, bo, op);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}
template int launch_neko_gather<double>(double *, int, int, const int *,
                                        const double *, const int *, int,
                                        const int *, const int *, int,
                                        hipStream_t);
template int launch_neko_gather<float>(float *, int, int, const int *,
                                       const float *, const int *, int,
                                       const int *, const int *, int,
                                       hipStream_t);

// Neko-layout scatter: u(gd(k+j)) = v(dg(k+1)) for blocks, then
// u(gd(i)) = v(dg(i)) for the remaining entries (gs_cpu.f90:292-322).
template <typename T>
__global__ void neko_scatter_kernel(const T *__restrict__ v, int m,
                                    const int *__restrict__ dg,
                                    T *__restrict__ u,
                                    const int *__restrict__ gd, int nb,
                                    const int *__restrict__ b,
                                    const int *__restrict__ bo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int start = nb > 0 ? bo[nb - 1] + b[nb - 1] + 1 : 1;
  if (t < nb) {
    const int k = bo[t], len = b[t];
    const T val = v[dg[k] - 1];
    for (int q = 0; q < len; q++) u[gd[k + q] - 1] = val;
    return;
  }
  const int64_t i = (int64_t)start + (t - nb);  // 1-based
  if (i <= m) u[gd[i - 1] - 1] = v[dg[i - 1] - 1];
}

template <typename T>
int launch_neko_scatter(const T *v, int m, const int *dg, T *u, const int *gd,
                        int nb, const int *b, const int *bo, hipStream_t s) {
  if (m <= 0) return NM_OK;
  // The first entry after the blocks, bo(nb) + b(nb) + 1, is read by the
  // kernel itself; launch an upper bound of nb + m items.
  const int64_t items = (int64_t)nb + m;
  hipLaunchKernelGGL(neko_scatter_kernel<T>, dim3((unsigned)((items + 255) / 256)),
                     dim3(256), 0, s, v, m, dg, u, gd, nb, b, bo);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}
template int launch_neko_scatter<double>(const double *, int, const int *,
                                         double *, const int *, int,
                                         const int *, const int *, hipStream_t);
template int launch_neko_scatter<float>(const float *, int, const int *,
                                        float *, const int *, int,
                                        const int *, const int *, hipStream_t);

// ---------------------------------------------------------------------------
// Zero-Dirichlet mask: x(msk(i)) = 0 (zero_dirichlet.f90:92-118).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void mask_kernel(T *__restrict__ x, const int32_t *__restrict__ msk,
                            int64_t m, const int *__restrict__ skip) {
  if (skip != nullptr && *skip) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) x[msk[t]] = T(0);
}

// Batched mask: workgroup b's 256 entries belong to problem wg[b] (entries
// padded with -1), a no-op while that problem's int flag at
// scal + problem * scal_size + done_off is set.
__global__ void mask_wg_kernel(float *__restrict__ x, const int32_t *__restrict__ msk,
                               int64_t m, const int32_t *__restrict__ wg,
                               const char *__restrict__ scal, int scal_size, int done_off,
                               const int32_t *__restrict__ map) {
  const int64_t b = map ? (int64_t)map[blockIdx.x] : (int64_t)blockIdx.x;
  if (*(const int *)(scal + (int64_t)wg[b] * scal_size + done_off)) return;
  const int64_t t = b * blockDim.x + threadIdx.x;
  if (t < m) {
    const int32_t k = msk[t];
    if (k >= 0) x[k] = 0.0f;
  }
}

int launch_mask_wg(float *x, const int32_t *msk, int64_t m, const int32_t *wg,
                   const void *scal, int scal_size, int done_off, hipStream_t s,
                   const int32_t *map, int64_t nmap) {
  const int64_t nwg = map ? nmap : (m + 255) / 256;
  if (m <= 0 || nwg <= 0) return NM_OK;
  hipLaunchKernelGGL(mask_wg_kernel, dim3((unsigned)nwg), dim3(256), 0, s, x, msk, m, wg,
                     (const char *)scal, scal_size, done_off, map);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

template <typename T>
int launch_mask(T *x, const int32_t *msk0, int64_t m, const int *skip,
                hipStream_t s) {
  if (m <= 0) return NM_OK;
  hipLaunchKernelGGL(mask_kernel<T>, dim3((unsigned)((m + 255) / 256)), dim3(256),
                     0, s, x, msk0, m, skip);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}
template int launch_mask<double>(double *, const int32_t *, int64_t, const int *, hipStream_t);
template int launch_mask<float>(float *, const int32_t *, int64_t, const int *, hipStream_t);

// ---------------------------------------------------------------------------
// Deterministic reductions.  Wave64 butterfly (DPP/shuffle) then LDS across
// the 4 waves of a 256-thread block; fixed grid, fixed combine order.
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int q = 0; q < nw; q++) r += sh[q];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

__global__ void __launch_bounds__(RED_THREADS)
    glsc3_partial_kernel(const double *__restrict__ a, const double *__restrict__ b,
                         const double *__restrict__ c, int64_t n,
                         double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc += a[i] * b[i] * c[i];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

int launch_glsc3_partial(const double *a, const double *b, const double *c,
                         int64_t n, double *part, hipStream_t s) {
  hipLaunchKernelGGL(glsc3_partial_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS),
                     0, s, a, b, c, n, part);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// j weighted dot products w.(v_l * mult), l < j  (glsc3_many semantics)
__global__ void __launch_bounds__(RED_THREADS)
    glsc3_many_partial_kernel(const double *__restrict__ w,
                              const double *const *__restrict__ v,
                              const double *__restrict__ mult, int j, int64_t n,
                              double *__restrict__ part) {
  // G vectors per sweep, so w and mult are read once per G vectors instead
  // of once per vector; every vector's per-thread sum runs over the same
  // points in the same order as a one-vector sweep, so the partials are
  // the same bits whatever G is
  constexpr int G = 8;
  __shared__ double sh[RED_THREADS / 64];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int l0 = 0; l0 < j; l0 += G) {
    const int ng = j - l0 < G ? j - l0 : G;
    const double *vp[G];
#pragma unroll
    for (int g = 0; g < G; g++) vp[g] = v[l0 + (g < ng ? g : 0)];
    double acc[G];
#pragma unroll
    for (int g = 0; g < G; g++) acc[g] = 0.0;
    // two points per trip (both points' loads in flight together), summed
    // in the same order as one point per trip
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + stride < n; i += 2 * stride) {
      const int64_t i2 = i + stride;
      const double wi = w[i], mi = mult[i], wi2 = w[i2], mi2 = mult[i2];
      double a[G], b[G];
#pragma unroll
      for (int g = 0; g < G; g++) {
        a[g] = g < ng ? __builtin_nontemporal_load(vp[g] + i) : 0.0;
        b[g] = g < ng ? __builtin_nontemporal_load(vp[g] + i2) : 0.0;
      }
#pragma unroll
      for (int g = 0; g < G; g++)
        if (g < ng) {
          acc[g] += wi * a[g] * mi;
          acc[g] += wi2 * b[g] * mi2;
        }
    }
    if (i < n) {
      const double wi = w[i], mi = mult[i];
#pragma unroll
      for (int g = 0; g < G; g++)
        if (g < ng) acc[g] += wi * __builtin_nontemporal_load(vp[g] + i) * mi;
    }
    for (int g = 0; g < ng; g++) {
      const double r = block_sum(acc[g], sh);
      if (threadIdx.x == 0) part[(int64_t)(l0 + g) * RED_BLOCKS + blockIdx.x] = r;
    }
  }
}

int launch_glsc3_many_partial(const double *w, const double *const *v,
                              const double *mult, int j, int64_t n,
                              double *part, hipStream_t s) {
  hipLaunchKernelGGL(glsc3_many_partial_kernel, dim3(RED_BLOCKS),
                     dim3(RED_THREADS), 0, s, w, v, mult, j, n, part);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

__global__ void __launch_bounds__(RED_THREADS)
    combine_kernel(const double *__restrict__ part, int nv, double *__restrict__ out) {
  __shared__ double sh[RED_THREADS / 64];
  for (int l = 0; l < nv; l++) {
    double acc = 0.0;
    for (int q = threadIdx.x; q < RED_BLOCKS; q += blockDim.x)
      acc += part[(int64_t)l * RED_BLOCKS + q];
    const double r = block_sum(acc, sh);
    if (threadIdx.x == 0) out[l] = r;
  }
}

int launch_combine(const double *part, int nv, double *out, hipStream_t s) {
  hipLaunchKernelGGL(combine_kernel, dim3(1), dim3(RED_THREADS), 0, s, part,
                     nv, out);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// ---------------------------------------------------------------------------
// Elementwise updates (device_math.F90 semantics).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void vec_kernel(int op, T *__restrict__ a, const T *__restrict__ b,
                           const T *__restrict__ c, const T *__restrict__ d,
                           T s1, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    switch (op) {
      case V_COPY: a[i] = b[i]; break;
      case V_RZERO: a[i] = T(0); break;
      case V_ADD2S1: a[i] = s1 * a[i] + b[i]; break;       // a = c1 a + b
      case V_ADD2S2: a[i] = a[i] + s1 * b[i]; break;       // a = a + c1 b
      case V_COL2: a[i] = a[i] * b[i]; break;
      case V_COL3: a[i] = b[i] * c[i]; break;
      case V_SUB2: a[i] = a[i] - b[i]; break;
      case V_CMULT2: a[i] = b[i] * s1; break;
      case V_INVCOL1: a[i] = T(1) / a[i]; break;
      case V_ADDCOL3: a[i] = a[i] + b[i] * c[i]; break;
      case V_ADDCOL4: a[i] = a[i] + b[i] * c[i] * d[i]; break;
      case V_CMULT: a[i] = a[i] * s1; break;
      case V_ADD2: a[i] = a[i] + b[i]; break;
      case V_CFILL: a[i] = s1; break;
      default: break;
    }
  }
}

template <typename T>
int launch_vec(int op, T *a, const T *b, const T *c, const T *d, T s1,
               int64_t n, hipStream_t s) {
  if (n <= 0) return NM_OK;
  int64_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(vec_kernel<T>, dim3((unsigned)nb), dim3(256), 0, s, op, a,
                     b, c, d, s1, n);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}
template int launch_vec<double>(int, double *, const double *, const double *,
                                const double *, double, int64_t, hipStream_t);
template int launch_vec<float>(int, float *, const float *, const float *,
                               const float *, float, int64_t, hipStream_t);

__global__ void d2f_kernel(const double *__restrict__ a, float *__restrict__ b,
                           int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    b[i] = (float)a[i];
}
int launch_d2f(const double *a, float *b, int64_t n, hipStream_t s) {
  if (n <= 0) return NM_OK;
  int64_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(d2f_kernel, dim3((unsigned)nb), dim3(256), 0, s, a, b, n);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// ---------------------------------------------------------------------------
// Geometric factors (coef.f90:714-1030 and :1112-1156): dx/dr via D in the
// three directions, Jacobian, cofactor metrics, G = (grad r_i . grad r_j) *
// jacinv * w3, B = jac * w3.  One thread per point, element staged in LDS.
// ---------------------------------------------------------------------------
template <int LX>
__global__ void __launch_bounds__(LX *LX *LX <= 1024 ? LX *LX *LX : 1024)
    coef_kernel(const double *__restrict__ X, const double *__restrict__ Y,
                const double *__restrict__ Z, const double *__restrict__ dx,
                const double *__restrict__ w3, double *__restrict__ g11,
                double *__restrict__ g22, double *__restrict__ g33,
                double *__restrict__ g12, double *__restrict__ g13,
                double *__restrict__ g23, double *__restrict__ jac,
                double *__restrict__ B) {
  constexpr int N3 = LX * LX * LX;
  __shared__ double sx[3][N3];
  __shared__ double sD[LX * LX];
  const int64_t e = blockIdx.x;
  const int64_t off = e * N3;
  for (int p = threadIdx.x; p < N3; p += blockDim.x) {
    sx[0][p] = X[off + p];
    sx[1][p] = Y[off + p];
    sx[2][p] = Z[off + p];
  }
  for (int p = threadIdx.x; p < LX * LX; p += blockDim.x) sD[p] = dx[p];
  __syncthreads();
  for (int p = threadIdx.x; p < N3; p += blockDim.x) {
    const int i = p % LX, j = (p / LX) % LX, k = p / (LX * LX);
    double dr[3], ds[3], dt[3];
    for (int c = 0; c < 3; c++) {
      double a = 0.0, b = 0.0, cc = 0.0;
      for (int l = 0; l < LX; l++) {
        a += sD[i + LX * l] * sx[c][l + LX * (j + LX * k)];
        b += sx[c][i + LX * (l + LX * k)] * sD[j + LX * l];
        cc += sx[c][i + LX * (j + LX * l)] * sD[k + LX * l];
      }
      dr[c] = a; ds[c] = b; dt[c] = cc;
    }
    const double xr = dr[0], yr = dr[1], zr = dr[2];
    const double xs = ds[0], ys = ds[1], zs = ds[2];
    const double xt = dt[0], yt = dt[1], zt = dt[2];
    double jc = 0.0;
    jc = jc + (xr * ys * zt);
    jc = jc + (xt * yr * zs);
    jc = jc + (xs * yt * zr);
    jc = jc - (xr * yt * zs);
    jc = jc - (xs * yr * zt);
    jc = jc - (xt * ys * zr);
    const double rx = ys * zt - yt * zs, ry = xt * zs - xs * zt, rz = xs * yt - xt * ys;
    const double sxx = yt * zr - yr * zt, sy = xr * zt - xt * zr, sz = xt * yr - xr * yt;
    const double tx = yr * zs - ys * zr, ty = xs * zr - xr * zs, tz = xr * ys - xs * yr;
    const double jinv = 1.0 / jc;
    const double wq = w3[p];
    const int64_t q = off + p;
    g11[q] = (rx * rx + ry * ry + rz * rz) * jinv * wq;
    g22[q] = (sxx * sxx + sy * sy + sz * sz) * jinv * wq;
    g33[q] = (tx * tx + ty * ty + tz * tz) * jinv * wq;
    g12[q] = (rx * sxx + ry * sy + rz * sz) * jinv * wq;
    g13[q] = (rx * tx + ry * ty + rz * tz) * jinv * wq;
    g23[q] = (sxx * tx + sy * ty + sz * tz) * jinv * wq;
    jac[q] = jc;
    B[q] = jc * wq;
  }
}

int launch_coef(int64_t nel, int lx, const double *x, const double *y,
                const double *z, const double *dx, const double *w3,
                double *g[6], double *jac, double *B, hipStream_t s) {
  if (nel <= 0) return NM_OK;
  const int threads = lx * lx * lx <= 1024 ? lx * lx * lx : 1024;
#define NM_COEF_CASE(L)                                                          \
  case L:                                                                        \
    hipLaunchKernelGGL(coef_kernel<L>, dim3((unsigned)nel), dim3(threads), 0, s, \
                       x, y, z, dx, w3, g[0], g[1], g[2], g[3], g[4], g[5], jac, B); \
    break;
  switch (lx) {
    NM_COEF_CASE(2) NM_COEF_CASE(3) NM_COEF_CASE(4) NM_COEF_CASE(5)
    NM_COEF_CASE(6) NM_COEF_CASE(7) NM_COEF_CASE(8) NM_COEF_CASE(9)
    NM_COEF_CASE(10) NM_COEF_CASE(11) NM_COEF_CASE(12)
    default:
      set_error("coef: lx=%d unsupported (2..12)", lx);
      return NM_ERR_UNSUPPORTED;
  }
#undef NM_COEF_CASE
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// ---------------------------------------------------------------------------
// Jacobi diagonal, CPU formula (pc_jacobi.f90:168-264), or with dfrmd ==
// nullptr the device formula (pc_jacobi.hip:42-104): sum_l of G11(l,j,k)
// dxt(i,l)^2 + G22(i,l,k) dyt(j,l)^2 + G33(i,j,l) dzt(k,l)^2, plus the
// G12/G13/G23 corner corrections on deformed elements only.
// ---------------------------------------------------------------------------
__global__ void jacobi_diag_kernel(int64_t nel, int lx,
                                   const double *__restrict__ dxt,
                                   const double *__restrict__ G11,
                                   const double *__restrict__ G22,
                                   const double *__restrict__ G33,
                                   const double *__restrict__ G12,
                                   const double *__restrict__ G13,
                                   const double *__restrict__ G23,
                                   const int *__restrict__ dfrmd,
                                   double *__restrict__ d) {
  const int64_t n3 = (int64_t)lx * lx * lx;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nel * n3) return;
  const int64_t e = idx / n3;
  const int p = (int)(idx % n3);
  const int i = p % lx, j = (p / lx) % lx, k = p / (lx * lx);
  const int64_t o = e * n3;
#define DT(a, b) dxt[(a) + lx * (b)]
#define A(arr, a, b, c) arr[o + (a) + lx * ((b) + lx * (c))]
  double acc = 0.0;
  if (dfrmd == nullptr) {
    // the device kernel's products (g t) t
    for (int l = 0; l < lx; l++) acc += A(G11, l, j, k) * DT(i, l) * DT(i, l);
    for (int l = 0; l < lx; l++) acc += A(G22, i, l, k) * DT(j, l) * DT(j, l);
    for (int l = 0; l < lx; l++) acc += A(G33, i, j, l) * DT(k, l) * DT(k, l);
  } else {
    // the CPU kernel's G dxt**2
    for (int l = 0; l < lx; l++) acc = acc + A(G11, l, j, k) * (DT(i, l) * DT(i, l));
    for (int l = 0; l < lx; l++) acc = acc + A(G22, i, l, k) * (DT(j, l) * DT(j, l));
    for (int l = 0; l < lx; l++) acc = acc + A(G33, i, j, l) * (DT(k, l) * DT(k, l));
  }
  const int L = lx - 1;
  const bool ci = (i == 0 || i == L), cj = (j == 0 || j == L), ck = (k == 0 || k == L);
  if (dfrmd == nullptr) {
    // device formula (pc_jacobi.hip:86-99): the cross terms of every face
    // point, each branch in the reference's order; D(i,i) vanishes inside
    // the element, so only edge and corner points receive nonzero terms
    if (ci) {
      acc += A(G12, i, j, k) * DT(i, i) * DT(j, j);
      acc += A(G13, i, j, k) * DT(i, i) * DT(k, k);
    }
    if (cj) {
      acc += A(G12, i, j, k) * DT(j, j) * DT(i, i);
      acc += A(G23, i, j, k) * DT(j, j) * DT(k, k);
    }
    if (ck) {
      acc += A(G13, i, j, k) * DT(k, k) * DT(i, i);
      acc += A(G23, i, j, k) * DT(k, k) * DT(j, j);
    }
  } else if (dfrmd[e] && ci && cj && ck) {
    // the three correction loops of the reference visit each corner once
    acc = acc + A(G12, i, j, k) * DT(i, i) * DT(j, j) + A(G13, i, j, k) * DT(i, i) * DT(k, k);
    acc = acc + A(G12, i, j, k) * DT(j, j) * DT(i, i) + A(G23, i, j, k) * DT(j, j) * DT(k, k);
    acc = acc + A(G13, i, j, k) * DT(k, k) * DT(i, i) + A(G23, i, j, k) * DT(k, k) * DT(j, j);
  }
  d[o + p] = acc;
#undef DT
#undef A
}

int launch_jacobi_diag(int64_t nel, int lx, const double *dxt,
                       const double *const g[6], const int *dfrmd, double *d,
                       hipStream_t s) {
  const int64_t n = nel * lx * lx * lx;
  if (n <= 0) return NM_OK;
  hipLaunchKernelGGL(jacobi_diag_kernel, dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, s, nel, lx, dxt, g[0], g[1], g[2], g[3],
                     g[4], g[5], dfrmd, d);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

}  // namespace nm
