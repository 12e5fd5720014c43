// Device-resident discretisation ("operator") shared by the solvers:
// geometry, gs maps, mask, mass, multiplicity and the Jacobi diagonal of one
// problem, in FP64 and (on demand) FP32.
#pragma once
#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

#include "neko_amd.h"
#include "nm_common.h"
#include "nm_kernels.h"

namespace nm {

// GLL coordinates of one curved element (nm_curve.cpp)
int curve_element_xyz(int lx, const double *zg, const double *v, const double *cd,
                      const int32_t *type, double *x, double *y, double *z);

template <typename T>
struct OpFields {
  T *dx = nullptr;        // D(i,l), lx*lx
  T *g[6] = {nullptr};    // G11 G22 G33 G12 G13 G23
  T *h1 = nullptr;        // pointwise h1 (nullptr: constant h1c)
  T *h2b = nullptr;       // h2*B (nullptr: no h2 term)
  T *mult = nullptr;
  T *dinv = nullptr;      // Jacobi: 1/gs(h1 diag(K) + h2 B)
  T *edges = nullptr;     // trilinear geometry: 36 half edge vectors/element
  T *zg = nullptr, *wg = nullptr;  // GLL nodes and weights
};

// A zero-Dirichlet mask (0-based device indices) and whether the fused CG
// update may take the descriptor face sums under it (no masked node on a
// descriptor face interior).
struct MaskRef {
  const int32_t *msk = nullptr;
  int64_t nmsk = 0;
  bool face_fuse = false;
};

struct Op {
  int64_t nel = 0, n = 0;
  int lx = 0;
  double h1c = 1.0;
  bool ifh2 = false;
  double volume = 0.0;
  // gs (device layout)
  int64_t nlocal = 0, nblks = 0, npairs = 0, pair0 = 0;
  int32_t *blk_ptr = nullptr, *gidx = nullptr;
  // the blocks as packed segments (gs variant 3; blocks of <= 8 entries)
  int32_t *seg_idx = nullptr;
  uint8_t *seg_pl = nullptr;
  int64_t nseg = 0;
  // facet pairs as whole-face descriptors (3 int32 each) + the rest
  int32_t *gs_faces = nullptr, *gs_rest = nullptr;
  int64_t nfaces = 0, nrest = 0;
  // per element face (6 per element): {neighbour element, code} of the
  // descriptor covering it, seen from this element (code as in GsArgs::faces:
  // own face in bits 0-2, the neighbour's face in bits 3-5), neighbour -1
  // where no descriptor covers the face
  int32_t *face_nbr = nullptr;
  std::vector<int32_t> h_face_nbr;
  // may the CG update take the descriptor face sums (no masked node on a
  // descriptor face interior)?
  bool face_fuse = false;
  bool face_fuse_on = true;  // caller's switch (nm_op_set_cg_face_sums)
  // multiplicity counts (mult == 1/count, count <= 255 everywhere: mcnt_ok)
  uint8_t *mcnt = nullptr;
  int32_t *mcnt_bad = nullptr;
  bool mcnt_ok = false;
  // Neko-layout copies (1-based) kept for the drop-in gather/scatter tests
  std::vector<int32_t> h_dg, h_gd, h_blk_len, h_blk_off;
  int64_t facet_offset = 0;
  // mask
  int64_t nmsk = 0;
  int32_t *msk = nullptr;  // 0-based
  int *dfrmd = nullptr;
  double *dxt = nullptr;
  double *B = nullptr, *jac = nullptr;
  double *binv = nullptr;      // 1/gs(B) (coef%Binv), built on first use
  double *pnpn_work = nullptr; // 6 n-vectors of pressure-residual scratch
  double *xyz[3] = {nullptr, nullptr, nullptr};
  OpFields<double> f64;
  OpFields<float> f32;
  bool has32 = false;
  int epb64 = 0, epb32 = 0;  // Ax variant codes (0 = default)
  // gather-scatter kernel variant (3 by default; variant 0 wherever a
  // block has more than 8 entries)
  int gs_variant = 3;
  int max_blk = 0;           // longest gs block
  double *red_part = nullptr;   // RED_BLOCKS * 4 doubles
  double *red_out = nullptr;    // 4 doubles
  std::vector<void *> owned;    // every device allocation of this op
  int64_t bytes = 0;            // their total size
  int box[3] = {0, 0, 0}, per[3] = {0, 0, 0};
  // general-mesh operators (nm_op_create_mesh): the topology the
  // multigrid levels and the Schwarz extended gs are rebuilt from
  std::vector<double> h_vxyz;
  std::vector<int64_t> h_pid;
  std::vector<int32_t> h_fneigh;
  // curved elements of a mesh operator (Neko's curve records: element,
  // 5 x 12 curve data, 12 edge types), kept for the multigrid levels
  std::vector<int32_t> h_cel, h_ctype;
  std::vector<double> h_cdata;
  bool trilinear = false;  // straight-edged elements: geometry in-kernel
  // diagonal preconditioner of the CG/GMRES solvers: 0 Jacobi (f*.dinv
  // holds 1/diag), 1 identity (f*.dinv holds ones); alt* keep the other
  int pc = 0;
  // bumped whenever kernel arguments a captured solver graph bakes in
  // change (variant codes, coefficients, preconditioner arrays)
  uint64_t epoch = 0;
  double *alt64 = nullptr;
  float *alt32 = nullptr;
  double h2c = 0.0;        // constant h2 (with trilinear geometry)
  double amp = 0.0;     // box perturbation amplitude (box operators)
  uint64_t seed = 0;    // box perturbation seed

  ~Op();
  template <typename T>
  T *alloc(int64_t count);
  int ensure32(hipStream_t s);
  template <typename T>
  const OpFields<T> &fields() const;
  // the operator's own mask
  MaskRef mask_ref() const { return MaskRef{msk, nmsk, face_fuse}; }
  // may the descriptor face sums be deferred past a mask of these host
  // (0-based) indices?
  bool face_fuse_ok(const int32_t *h_msk, int64_t m) const;
  // w = mask(gs(A_local u)); mk replaces the operator's mask
  template <typename T>
  int apply(T *w, const T *u, const int *skip, hipStream_t s, bool gs = true,
            bool mask = true, double *dot_part = nullptr,
            const MaskRef *mk = nullptr) const;
  // w = mask(gs(A p)) with the CG prologue fused into the trilinear Ax:
  // p <- dinv r + (*beta) p and x += (*alpha) p_old while loading p
  // With defer_faces the interior points of descriptor faces are left
  // unsummed (the caller's update kernel adds the partner values).
  template <typename T>
  int apply_cgp(T *w, T *p, const T *r, T *x, const double *beta, const double *alpha,
                const int *skip, hipStream_t s, double *dot_part,
                bool defer_faces = false, const MaskRef *mk = nullptr) const;
  // variant code of the Ax kernel used for precision T
  template <typename T>
  int ax_code() const;
  // can the fused p.Ap dot run?  Every Ax variant fuses it together with
  // the CG prologue.
  template <typename T>
  bool fused_dot() const { return true; }
  template <typename T>
  int gs(T *u, const int *skip, hipStream_t s) const;
};

template <>
inline const OpFields<double> &Op::fields<double>() const { return f64; }
template <>
inline const OpFields<float> &Op::fields<float>() const { return f32; }

// Build the device layout of Neko's (1-based) local gs maps and finish the
// derived fields (mult, volume, Jacobi) on the device.
int op_set_gs(Op *op, int64_t nlocal, int64_t facet_offset, int64_t nblks,
              const int32_t *dg, const int32_t *gd, const int32_t *blk_len,
              const int32_t *blk_off, hipStream_t s);
int op_finish(Op *op, hipStream_t s);
// dd = 1/gs(h1 diag(K) + h2 B) with the op's current coefficients
int op_jacobi(Op *op, double *dd, hipStream_t s);
// Regroup a device gs layout (blk_ptr/gidx, pairs from pair0) by the element
// of each block's / pair's lowest dof; the order inside blocks is kept.
// Facet pairs (0-based index pairs) -> whole-face descriptors (3 int32:
// element a, element b, code) + the pairs no descriptor covers.
void gs_face_pairs(int64_t nel, int lx, const int32_t *pairs, int64_t npairs,
                   std::vector<int32_t> &faces, std::vector<int32_t> &rest);
int gs_regroup(int64_t nel, int64_t n3, int64_t nblks, int64_t pair0, int64_t npairs,
               std::vector<int32_t> &ptr, std::vector<int32_t> &gi);

// Coarse-grid CG with a fixed budget and no host synchronisation
// (nm_cg.hip).
struct CgFixed;
int cg_fixed_create(CgFixed **out, const Op *op);
void cg_fixed_free(CgFixed *c);
int cg_fixed_solve(CgFixed *c, double *x, const double *f, int niter,
                   double abs_tol, hipStream_t s);

}  // namespace nm

// Opaque handle of the C-ABI (include/neko_amd.h: nm_op_t).
struct nm_op_s {
  nm::Op op;
};
