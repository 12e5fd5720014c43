// Launchers of the gfx950 kernels (implemented in nm_kernels.hip).
#pragma once
#include <cstdint>

#include <hip/hip_runtime.h>

namespace nm {

// Reduction geometry: every reduction uses exactly RED_BLOCKS x RED_THREADS
// threads and a fixed combine order, so a result depends only on the data
// (deterministic across runs and across concurrent problems).
constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;

// XCDs of the chip; workgroup b of a launch runs on XCD b % NM_NXCD.
constexpr int NM_NXCD = 8;

// Largest lx for which the CG update kernel can take the descriptor face
// sums of gs itself.
constexpr int NM_FACE_FUSE_MAX_LX = 14;

// Multiplicity counts: cnt[i] = c with mult[i] == 1.0 / c exactly and
// 1 <= c <= 255, else *bad = 1 (cnt[i] then 0).
int launch_mult_count(const double *mult, uint8_t *cnt, int *bad, int64_t n, hipStream_t s);

// Local offset of point (u, v) of element face f (f = 2 dir + side, dir 0:
// i fixed, u = j, v = k; dir 1: j fixed, u = i, v = k; dir 2: k fixed, u = i,
// v = j; side 1 is the index lx - 1).
__device__ inline int64_t face_point(int f, int u, int v, int lx) {
  const int fixed = (f & 1) ? lx - 1 : 0;
  const int dir = f >> 1;
  const int i = dir == 0 ? fixed : u;
  const int j = dir == 0 ? u : (dir == 1 ? fixed : v);
  const int k = dir == 2 ? fixed : v;
  return i + (int64_t)lx * (j + (int64_t)lx * k);
}

// Arguments of the tensor-product Helmholtz operator
//   w = sum_{rst} D^T (h1 G) D u  (+ h2B * u)
template <typename T>
struct AxArgs {
  T *w;
  const T *u;
  const T *dx;         // D(i,l) column-major, lx*lx
  const T *h1;         // pointwise h1 field, or nullptr
  T h1c;               // constant h1 when h1 == nullptr
  const T *g[6];       // G11 G22 G33 G12 G13 G23
  const T *h2b;        // premultiplied h2*B field, or nullptr (ifh2 false)
  int64_t nel;
  int lx;
  const int *skip;     // device flag: launch is a no-op when *skip != 0
  // trilinear geometry (elements with straight edges): G and B evaluated
  // in the kernel from 36 half edge vectors per element instead of read
  const T *edges = nullptr;
  const T *zg = nullptr;   // GLL nodes
  const T *wg = nullptr;   // GLL weights
  T h2c = T(0);            // constant h2 (mass term h2*B*u when nonzero)
  // optional fused partial dot sum_i u_i (A_local u)_i, one double per
  // workgroup (grid size returned by ax_grid); the stored-geometry
  // variants take it only together with the CG prologue
  double *dot_part = nullptr;
  // optional fused CG prologue (with dot_part): u is
  // the search direction p, replaced by dinv r + beta p while loading, and
  // x += alpha p_old; beta/alpha read from the device
  const T *cg_r = nullptr, *cg_dinv = nullptr;
  T *cg_x = nullptr;
  const double *cg_beta = nullptr, *cg_alpha = nullptr;
};

// Workgroups of an Ax launch for a variant code (for sizing dot_part)
int64_t ax_grid(int64_t nel, int lx, int code);

// Batched trilinear Ax (FP32) over many box problems of one lx: instance
// table entry and launch arguments.
struct BatchInst {
  int64_t voff;          // offset of the instance in the batch vectors
  int64_t nel;
  const float *edges;    // its 36 half edge vectors per element
  float h1c, h2c;
  int64_t blk0, nblk;    // its Ax workgroups (global numbering)
  int64_t c0, c1;        // its 256-dof chunks [c0, c1)
  int64_t u0, u1;        // its update workgroups [u0, u1) (runs of chunks or elements)
  int64_t e0;            // its first element in the batch's face table
};
struct AxBatchArgs {
  float *w;
  const float *u;
  const float *dx, *zg, *wg;  // of this lx
  const BatchInst *inst;
  const int32_t *blk_inst;    // instance of every global workgroup
  int64_t blk_base, nblocks;  // this launch: global workgroups [base, base+n)
  int lx, epb;
  const void *scal;           // per-instance scalar records
  int scal_size, done_off;    // record size, offset of the int `done`
  double *dot_part;           // per global workgroup: sum u.w
  // CG prologue (cg_r != nullptr): u is p, updated in place to
  // dinv r + beta p while x += alpha p_old, with the instance's double
  // beta and alpha at these record offsets
  const float *cg_r = nullptr, *cg_dinv = nullptr;
  float *cg_x = nullptr;
  int beta_off = 0, alpha_off = 0;
};
int launch_ax_batch(const AxBatchArgs &a, hipStream_t s);

template <typename T>
int launch_ax(const AxArgs<T> &a, int epb, hipStream_t s);

// Fused local gather-scatter ADD on the device layout:
//   blocks  : entries gidx[blk_ptr[b] .. blk_ptr[b+1]) summed in order
//   pairs   : entries gidx[pair0 + 2p], gidx[pair0 + 2p + 1]
template <typename T>
struct GsArgs {
  T *u;
  const int32_t *blk_ptr;  // nblks + 1 offsets (0-based)
  const int32_t *gidx;     // nlocal 0-based local indices
  int64_t nblks;
  int64_t pair0;           // offset of the first facet pair in gidx
  int64_t npairs;
  const int *skip;
  int variant = 0;         // 0 thread per block/pair, 1 lane-parallel (blocks <= 8),
                           // 2 as 1 with facet pairs from face descriptors
  // variant 2: whole element faces shared by two elements, one descriptor
  // each ({element a, element b, code}, code = face a | face b << 3 |
  // swap << 6 | flip u << 7 | flip v << 8), and the facet pairs no
  // descriptor covers (pairs of 0-based indices)
  const int32_t *faces = nullptr;
  int64_t nfaces = 0;
  int lx = 0;
  const int32_t *rest = nullptr;
  int64_t nrest = 0;
  // batched problems (variant 1): every workgroup's blocks / pairs belong
  // to one problem, wg_blk[b] / wg_pair[b] (logical workgroup b); the
  // workgroup is a no-op while that problem's int flag at
  // scal + problem * scal_size + done_off is set.  Pairs of -1 are padding.
  const int32_t *wg_blk = nullptr, *wg_pair = nullptr;
  const char *scal = nullptr;
  int scal_size = 0, done_off = 0;
  // variant 3 (as 2 for the faces): the blocks as packed segments, one lane
  // per entry: seg_idx[t] the local index (-1: padding), seg_pl[t] its
  // place in its block | block length << 4, no block across a wavefront
  const int32_t *seg_idx = nullptr;
  const uint8_t *seg_pl = nullptr;
  int64_t nseg = 0;
  // optional (variant 1 blocks): only the nwg_map logical workgroups listed
  // in wg_map are launched (those of the problems still running)
  const int32_t *wg_map = nullptr;
  int64_t nwg_map = 0;
};
template <typename T>
int launch_gs(const GsArgs<T> &a, hipStream_t s);

// Neko-layout gather / scatter (1-based maps, gs_kernels semantics).
template <typename T>
int launch_neko_gather(T *v, int m, int o, const int *dg, const T *u,
                       const int *gd, int nb, const int *b, const int *bo,
                       int op, hipStream_t s);
template <typename T>
int launch_neko_scatter(const T *v, int m, const int *dg, T *u, const int *gd,
                        int nb, const int *b, const int *bo, hipStream_t s);

template <typename T>
int launch_mask(T *x, const int32_t *msk0, int64_t m, const int *skip,
                hipStream_t s);
// batched FP32 mask with per-workgroup problem skip (entries padded to 256
// per problem with -1; wg[b] = problem of workgroup b); with map only the
// nmap workgroups listed there are launched
int launch_mask_wg(float *x, const int32_t *msk, int64_t m, const int32_t *wg,
                   const void *scal, int scal_size, int done_off, hipStream_t s,
                   const int32_t *map = nullptr, int64_t nmap = 0);

// Deterministic reductions: partial sums into part[RED_BLOCKS * nv] then a
// one-block combine into out[nv] (double).
int launch_glsc3_partial(const double *a, const double *b, const double *c,
                         int64_t n, double *part, hipStream_t s);
int launch_glsc3_many_partial(const double *w, const double *const *v,
                              const double *mult, int j, int64_t n,
                              double *part, hipStream_t s);
int launch_combine(const double *part, int nv, double *out, hipStream_t s);

// Elementwise (Neko device_math semantics, 0-based, any n)
enum VecOp : int {
  V_COPY, V_RZERO, V_ADD2S1, V_ADD2S2, V_COL2, V_COL3, V_SUB2, V_CMULT2,
  V_INVCOL1, V_ADDCOL3, V_ADDCOL4, V_CMULT, V_ADD2, V_CFILL
};
template <typename T>
int launch_vec(int op, T *a, const T *b, const T *c, const T *d, T s1,
               int64_t n, hipStream_t s);

// Geometric factors from GLL point coordinates (coef.f90 formulas).
int launch_coef(int64_t nel, int lx, const double *x, const double *y,
                const double *z, const double *dx, const double *w3,
                double *g[6], double *jac, double *B, hipStream_t s);

// Jacobi diagonal before h1/h2/gs/invert: CPU formula of pc_jacobi.f90
// with per-element flags dfrmd, or (dfrmd == nullptr) the formula of the
// reference's device kernel (pc_jacobi.hip), cross terms at all face points.
int launch_jacobi_diag(int64_t nel, int lx, const double *dxt,
                       const double *const g[6], const int *dfrmd,
                       double *d, hipStream_t s);

// Convert / cast
int launch_d2f(const double *a, float *b, int64_t n, hipStream_t s);

}  // namespace nm
