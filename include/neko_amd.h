/*
 * neko_amd.h -- C ABI of libneko_amd.so, the MI355X (gfx950) implementation
 * of Neko's pressure/velocity linear-solve hot path.
 *
 * Two layers:
 *
 *  1. Drop-in entry points with the exact signatures Neko's Fortran binds
 *     for its HIP backend (device pointers by value, integers by reference,
 *     FP64 "real", 1-based gs maps, stream = Neko's glb_cmd_queue where the
 *     reference takes none).  Each names the reference function it replaces.
 *     Like the reference they return nothing and abort on a HIP error.
 *
 *  2. A handle API (nm_*) that owns device copies of one problem's geometry,
 *     gs maps and mask, and runs whole solves with device-resident scalars.
 *     Every nm_* function returns 0 on success or a status code; the message
 *     of the last failure on the calling thread is nm_last_error().
 *
 * All fields use Neko's layout: column-major (lx,lx,lx,nel), i.e. point
 * (i,j,k) of element e at i + lx*(j + lx*(k + lx*e)).  D(i,j) = dx[i+lx*j].
 * Reference paths below are relative to the Neko source tree.
 */
#ifndef NEKO_AMD_H
#define NEKO_AMD_H

#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================
 * 1. Drop-in entry points
 * ====================================================================== */

/* Neko's global command queue (src/device/device.F90:52, bind(c)); weak
 * definition in the library, Neko's own wins when linked into Neko. */
extern void *glb_cmd_queue;

/* replaces hip_ax_helm, src/math/bcknd/device/hip/ax_helm.hip:64-67
 * (Fortran interface src/math/bcknd/device/ax_helm_device.F90:53-64).
 * w = sum_rst D^T (h1 G) D u per element; stream glb_cmd_queue. */
void hip_ax_helm(void *w, void *u, void *dx, void *dy, void *dz, void *dxt,
                 void *dyt, void *dzt, void *h1, void *g11, void *g22,
                 void *g33, void *g12, void *g13, void *g23, int *nelv,
                 int *lx);

/* FP32 twin for Neko's single-precision build (--enable-real=sp); no
 * reference counterpart with this name (the reference templates on real). */
void hip_ax_helm_sp(void *w, void *u, void *dx, void *h1, void *g11,
                    void *g22, void *g33, void *g12, void *g13, void *g23,
                    int *nelv, int *lx, hipStream_t strm);

/* replaces hip_gather_kernel / hip_scatter_kernel,
 * src/gs/bcknd/device/hip/gs.hip:52-93 and :98-118 (bound in
 * src/gs/bcknd/device/gs_device.F90).  op: 1 ADD, 2 MUL, 3 MIN, 4 MAX
 * (src/gs/gs_ops.f90:35-36); o < 0 selects the plain shared-copy tail. */
void hip_gather_kernel(void *v, int *m, int *o, void *dg, void *u, int *n,
                       void *gd, int *nb, void *b, void *bo, int *op,
                       hipStream_t stream);
void hip_scatter_kernel(void *v, int *m, void *dg, void *u, int *n, void *gd,
                        int *nb, void *b, void *bo, hipStream_t stream);

/* replaces hip_glsc3, src/math/bcknd/device/hip/math.hip:891 */
double hip_glsc3(void *a, void *b, void *c, int *n, hipStream_t stream);
/* replaces hip_glsc3_many, math.hip:922 (v: device array of j pointers) */
void hip_glsc3_many(double *h, void *w, void *v, void *mult, int *j, int *n,
                    hipStream_t stream);

/* replace the BLAS-1 wrappers of math.hip:55-720 used by the Krylov
 * solvers (add2s1 :440, add2s2 :456, cmult2 :265, copy :55, rzero :231,
 * cmult :251, cfill :377, add2 :394, col2 :601, col3 :616, sub2 :646,
 * invcol1 :558, addcol3 :676, addcol4 :691) */
void hip_add2s1(void *a, void *b, double *c1, int *n, hipStream_t strm);
void hip_add2s2(void *a, void *b, double *c1, int *n, hipStream_t strm);
void hip_cmult2(void *a, void *b, double *c1, int *n, hipStream_t strm);
void hip_copy(void *a, void *b, int *n, hipStream_t strm);
void hip_rzero(void *a, int *n, hipStream_t strm);
void hip_cmult(void *a, double *c, int *n, hipStream_t strm);
void hip_cfill(void *a, double *c, int *n, hipStream_t strm);
void hip_add2(void *a, void *b, int *n, hipStream_t strm);
void hip_col2(void *a, void *b, int *n, hipStream_t strm);
void hip_col3(void *a, void *b, void *c, int *n, hipStream_t strm);
void hip_sub2(void *a, void *b, int *n, hipStream_t strm);
void hip_invcol1(void *a, int *n, hipStream_t strm);
void hip_addcol3(void *a, void *b, void *c, int *n, hipStream_t strm);
void hip_addcol4(void *a, void *b, void *c, void *d, int *n,
                 hipStream_t strm);

/* replaces hip_jacobi_update, src/krylov/bcknd/device/hip/pc_jacobi.hip:108,
 * with that kernel's formula (:42-104: G12/G13/G23 cross terms at every face
 * point; they vanish except on element edges and corners); glb_cmd_queue. */
void hip_jacobi_update(void *d, void *dxt, void *dyt, void *dzt, void *G11,
                       void *G22, void *G33, void *G12, void *G13, void *G23,
                       int *nel, int *lxp);
/* Same with Neko's per-element flags msh%dfrmd_el (device int32 array):
 * exact CPU semantics (pc_jacobi.f90:922-955 cross terms only on deformed
 * elements).  No reference entry point carries the flags. */
void nm_jacobi_update(void *d, void *dxt, void *G11, void *G22, void *G33,
                      void *G12, void *G13, void *G23, void *dfrmd, int *nel,
                      int *lxp, hipStream_t strm);

/* No counterpart in the reference: hip_glsc3_many (above) for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  h_dev[l] = sum_i w_i
 * v_l[i] mult_i for l < j, written to device memory on stream s (nothing
 * is copied to the host); v: device array of j device pointers.  Calls no
 * allocating routine: the partial sums live in a static device array of
 * the library (512 KiB per GPU, held until the process ends) that bounds
 * j to 1..64 (else NM_ERR_ARG).  Calls that overlap on one GPU share that
 * array and race. */
int nm_glsc3_many_noalloc(double *h_dev, const double *w,
                          const double *const *v, const double *mult, int j,
                          int n, hipStream_t s);

/* ======================================================================
 * 2. Handle API
 * ====================================================================== */
typedef struct nm_op_s nm_op_t;
typedef struct nm_cg_s nm_cg_t;
typedef struct nm_hsmg_s nm_hsmg_t;
typedef struct nm_gmres_s nm_gmres_t;
typedef struct nm_server_s nm_server_t;
typedef struct nm_batch_s nm_batch_t;
typedef struct nm_phmg_s nm_phmg_t;
typedef struct nm_proj_s nm_proj_t;
typedef struct nm_surface_s nm_surface_t;

typedef struct {
  int64_t nel, n, nlocal, nblks, npairs, nmsk, facet_offset;
  int lx;
  double volume;
  double h1c;
  int ifh2;
  /* facet pairs the face-descriptor gs covers (whole shared faces) and the
   * pairs it leaves to the pair kernel */
  int64_t nfaces, nrest;
  /* 1 when the fused CG update takes the descriptor face sums of gs (no
   * masked node on a descriptor face interior, and not switched off) */
  int cg_face_sums;
  /* 1 when every multiplicity weight is 1/c for a count c <= 255 (the fused
   * CG update then reads one byte per point instead of mult) */
  int mult_counts;
} nm_op_info_t;

typedef struct {
  int iter, outer;
  double res_start, res_final;
  int converged;
} nm_gmres_monitor_t;

typedef struct {
  int iter;          /* CG iterations (mixed: total inner FP32 iterations) */
  int outer;         /* refinement sweeps (mixed), 0 for FP64 */
  double res_start;  /* Neko-normalised residuals sqrt(r.mult.r/volume) */
  double res_final;
  int converged;
} nm_monitor_t;

const char *nm_last_error(void);

/* --- setup stand-ins for Neko's space_t / mesh_t / dofmap_t / gs_init --- */
/* GLL nodes, weights, D and D^T (space.f90:149-320; speclib.f90 ZWGLL/DGLL) */
int nm_gll(int lx, double *z, double *w, double *d, double *dt);
/* seeded synthetic inputs (this project's input specification) */
double nm_rng_uniform(uint64_t seed, uint64_t idx);
int nm_rng_normal_fill(uint64_t seed, int64_t n, double *out);
/* unit-cube box mesh, element vertices (nel*8*3) in Neko's hex order
 * (contrib/genmeshbox/genmeshbox.f90:200-216), interior vertices perturbed
 * by amp*h*(2U-1) */
int nm_box_mesh(int nelx, int nely, int nelz, double amp, uint64_t seed,
                double *vxyz);
/* msh%dfrmd_el (src/mesh/mesh.f90:437-467) */
int nm_dfrmd(int64_t nel, const double *vxyz, int *dfrmd);
/* trilinear GLL coordinates (src/sem/dofmap.f90:866-939 dofmap_xyzlin) */
int nm_xyz(int64_t nel, int lx, const double *zg, const double *vxyz,
           double *x, double *y, double *z);
/* Neko's local gs maps of a box, bit-exact layout of gs_init_mapping
 * (src/gs/gather_scatter.f90:513-1375): 1-based dg/gd, blk_len/blk_off */
int nm_box_gs_sizes(int nelx, int nely, int nelz, int lx, int perx, int pery,
                    int perz, int64_t *nlocal, int64_t *facet_offset);
int nm_box_gs_map(int nelx, int nely, int nelz, int lx, int perx, int pery,
                  int perz, int32_t *dg, int32_t *gd, int32_t *blk_len,
                  int32_t *blk_off, int64_t *nlocal, int64_t *facet_offset,
                  int64_t *nblks);
/* zero-Dirichlet mask of a box (1-based, zero_dirichlet.f90 msk(1:m)) */
int64_t nm_box_mask_size(int nelx, int nely, int nelz, int lx, int perx,
                         int pery, int perz);
int nm_box_mask(int nelx, int nely, int nelz, int lx, int perx, int pery,
                int perz, int32_t *msk, int64_t *nmsk);

/* --- general conforming hex meshes ------------------------------------- */
/* Neko's binary mesh file (src/io/nmsh_file.f90:70-207, records of
 * src/io/format/nmsh.f90): 3-D hex meshes with straight edges.  info gives
 * the element, zone, periodic-zone and curve counts; read returns the
 * element vertices in Neko's symmetric order (the p1,p2,p4,p3,p5,p6,p8,p7
 * swap of nmsh_file.f90:159) and their point ids after the periodic zones'
 * id replacement (mesh_apply_periodic_facet, src/mesh/mesh.f90:1837-1866).
 * The vertices only: the curve records of a curved mesh come from
 * nm_nmsh_curves and go to nm_op_create_mesh_curved. */
int nm_nmsh_info(const char *path, int64_t *nel, int64_t *nzones,
                 int64_t *nperiodic, int64_t *ncurves);
int nm_nmsh_read(const char *path, int64_t nel, double *vxyz, int64_t *pid,
                 int32_t *el_idx);
/* facet neighbours, slot e*6+f, 1-based element or 0 on the boundary
 * (msh%facet_neigh, src/mesh/mesh.f90:590-628) */
int nm_mesh_facet_neigh(int64_t nel, const int64_t *pid, int32_t *fneigh);
/* Neko's local gs maps of a general mesh (gather_scatter.f90:513-1375),
 * same layout as nm_box_gs_map; shared points identified from the point
 * ids as Neko's dofmap numbers them (src/sem/dofmap.f90:264-797) */
int nm_mesh_gs_sizes(int64_t nel, int lx, const int32_t *fneigh,
                     int64_t *nlocal, int64_t *facet_offset);
int nm_mesh_gs_map(int64_t nel, int lx, const int64_t *pid,
                   const int32_t *fneigh, int32_t *dg, int32_t *gd,
                   int32_t *blk_len, int32_t *blk_off, int64_t *nlocal,
                   int64_t *facet_offset, int64_t *nblks);
/* zero-Dirichlet mask of every facet without a neighbour (1-based; the
 * Poisson example's set_bc, examples/poisson/setup.f90) */
int64_t nm_mesh_mask_size(int64_t nel, int lx, const int32_t *fneigh);
int nm_mesh_mask(int64_t nel, int lx, const int32_t *fneigh, int32_t *msk,
                 int64_t *nmsk);

/* --- problem (operator) handles ----------------------------------------- */
/* Whole box problem built on the device (coef on the GPU). */
int nm_op_create_box(nm_op_t **op, int nelx, int nely, int nelz, int lx,
                     int perx, int pery, int perz, double amp, uint64_t seed,
                     double h1, double h2, hipStream_t s);
/* zone records of a .nmsh file (nmsh_file.f90:326-380), 4 int32 each:
 * element (0-based, file order), facet 1..6, type (5 periodic, 7 labelled
 * boundary), label (p_f of the record: the boundary label of a type-7
 * zone); nzones as nm_nmsh_info reports */
int nm_nmsh_zones(const char *path, int64_t nzones, int32_t *out);
/* curve records of a .nmsh file (nmsh_file_read_curves,
 * src/io/nmsh_file.f90:575-639; record nmsh_curve_el_t,
 * src/io/format/nmsh.f90:39-43): per record the element (0-based, file
 * order), curve_data(5,12) as 60 doubles (5 per edge, Nek5000's cyclic edge
 * order) and the 12 edge types (3 circular arc, 4 midpoint; others change
 * nothing, as in dofmap_generate_xyz); ncurves as nm_nmsh_info reports */
int nm_nmsh_curves(const char *path, int64_t ncurves, int32_t *el, double *data,
                   int32_t *type);
/* GLL coordinates (lx^3 each) of one element from its 8 vertices (Neko's
 * symmetric order) and one curve record (dofmap_generate_xyz with
 * dofmap_xyzquad, gh_face_extend_3d and arc_surface,
 * src/sem/dofmap.f90:797-1285) */
int nm_curve_xyz(int lx, const double *vxyz, const double *data, const int32_t *type,
                 double *x, double *y, double *z);
/* Whole problem on a general straight-edged hex mesh (e.g. from
 * nm_nmsh_read): the set-up of examples/poisson/driver.f90 (dofmap, gs,
 * coef, Dirichlet on every facet without a neighbour), constant h1/h2. */
int nm_op_create_mesh(nm_op_t **op, int64_t nel, int lx, const double *vxyz,
                      const int64_t *pid, double h1, double h2, hipStream_t s);
/* the same with curved elements (dofmap_generate_xyz,
 * src/sem/dofmap.f90:797-864): ncurves records as nm_nmsh_curves returns
 * them (at most one per element); the geometric factors then come from
 * the curved coordinates (coef.f90:248-1156) and the in-kernel (trilinear)
 * Ax variants are off.  Multigrid levels built on it are curved too. */
int nm_op_create_mesh_curved(nm_op_t **op, int64_t nel, int lx, const double *vxyz,
                             const int64_t *pid, int64_t ncurves, const int32_t *cel,
                             const double *cdata, const int32_t *ctype, double h1,
                             double h2, hipStream_t s);
/* Problem from Neko's own host arrays (coef_t, gs_t, bc mask) -- what the
 * Fortran ksp_t/pc_t plug-ins pass at init (src/krylov/krylov.f90:73-201,
 * src/krylov/precon.f90:40-65). */
int nm_op_create(nm_op_t **op, int nel, int lx, const double *dx,
                 const double *g11, const double *g22, const double *g33,
                 const double *g12, const double *g13, const double *g23,
                 const double *B, const double *h1, double h1c,
                 const double *h2, int ifh2, const int *dfrmd, int nlocal,
                 int facet_offset, int nblks, const int *dg, const int *gd,
                 const int *blk_len, const int *blk_off, int nmsk,
                 const int *msk, hipStream_t s);
void nm_op_free(nm_op_t *op);
int nm_op_info(const nm_op_t *op, nm_op_info_t *info);
/* device pointer of a field: 0..5 G11..G23, 6 B, 7 mult, 8 Jacobi 1/diag,
 * 9..11 x,y,z, 12 h2*B, 13 jac */
const void *nm_op_field(const nm_op_t *op, int which);
int nm_op_gs_maps(const nm_op_t *op, int32_t *dg, int32_t *gd,
                  int32_t *blk_len, int32_t *blk_off);
/* Ax variant code: elements per workgroup (1,2,4,8,16; 0 = default) plus
 * 100 for the G-prefetching variant, 200 for the geometry evaluated in the
 * kernel (straight-edged operators); gs variant 0 (thread per block), 1
 * (lane-parallel, blocks <= 8 entries), 2 (as 1, facet pairs as whole
 * shared faces with computed indices) or 3 (as 2, the blocks as packed
 * segments with one lane per entry; the default, 0 where a block has more
 * than 8 entries). */
int nm_op_set_epb(nm_op_t *op, int epb64, int epb32);
int nm_op_set_gs_variant(nm_op_t *op, int variant);
/* Fused CG schedule: let (1, the default) or forbid (0) its update kernel
 * to add the partner values of descriptor-face points itself instead of a
 * face gs pass (same sums; no counterpart in the reference, whose CG runs
 * gs as one pass). */
int nm_op_set_cg_face_sums(nm_op_t *op, int on);
/* new constant coefficients h1, h2 (h2 != 0: mass term on) of an operator
 * with constant h1, as the pnpn schemes set them before each solve
 * (pressure h1 = 1/rho, h2 = 0: src/fluid/bcknd/cpu/pnpn_res_cpu.f90:74-78;
 * velocity h1 = mu, h2 = rho bd/dt: :212-216); h2 B and the Jacobi diagonal
 * are rebuilt.  Solvers created before re-capture their graphs on their
 * next solve; batches keep the coefficients they were created with. */
int nm_op_set_coeffs(nm_op_t *op, double h1, double h2, hipStream_t s);
/* weak gradient opgrad (src/math/operators.f90:307-343, CPU kernel
 * src/math/bcknd/cpu/cpu_opgrad.f90:263-329), device pointers, local;
 * straight-edged operators (geometry from the element edges) */
int nm_opgrad(nm_op_t *op, double *ux, double *uy, double *uz, const double *u,
              hipStream_t s);
/* velocity residual pnpn_vel_res_cpu_compute (pnpn_res_cpu.f90:182-241):
 * sets h1 = mu, h2 = rho bd/dt, then res_c = -A u_c - opgrad_c(p) + f_c
 * (local operator; gs and boundary conditions stay with the caller) */
int nm_vel_res(nm_op_t *op, double mu, double rho, double bd, double dt,
               const double *u, const double *v, const double *w, const double *p,
               const double *fx, const double *fy, const double *fz,
               double *ures, double *vres, double *wres, hipStream_t s);
/* curl (src/math/operators.f90 curl, CPU kernel opr_cpu_curl,
 * src/math/bcknd/cpu/opr_cpu.f90:126-180): w = Binv gs(B curl u), no cyclic
 * rotation; straight-edged operators */
int nm_curl(nm_op_t *op, double *w1, double *w2, double *w3, const double *u1,
            const double *u2, const double *u3, hipStream_t s);
/* pressure residual pnpn_prs_res_cpu_compute (pnpn_res_cpu.f90:31-180):
 * sets h1 = 1/rho, h2 = 0; p_res = -A p + sum_c cdtp(ta_c) with
 * ta = Binv gs(f/rho - (mu/rho) B curl(curl(u_e))), without the
 * facet-normal surface terms (exact when the velocity vanishes on the
 * velocity-Dirichlet facets and there are no symmetry facets; with them:
 * nm_prs_res_surf). */
int nm_prs_res(nm_op_t *op, double mu, double rho, const double *p,
               const double *ue, const double *ve, const double *we,
               const double *fx, const double *fy, const double *fz,
               double *p_res, hipStream_t s);
/* facet-normal surface of a boundary zone (facet_normal_t,
 * src/bc/facet_normal.f90:233-326; normals and areas as
 * coef_generate_area_and_normal, src/sem/coef.f90:1195-1370): facets =
 * nfacets (element 0-based, facet 1..6 in Neko's numbering) pairs, in the
 * zone's order; box and mesh operators (they carry coordinates) */
int nm_surface_create(nm_surface_t **out, nm_op_t *op, const int32_t *facets,
                      int64_t nfacets, hipStream_t s);
void nm_surface_free(nm_surface_t *sf);
int nm_surface_info(const nm_surface_t *sf, int64_t *npoints);
/* host copies of the unique points (0-based local) and their area-weighted
 * normals, npoints each */
int nm_surface_normals(const nm_surface_t *sf, int32_t *pts, double *nx,
                       double *ny, double *nz);
/* Cyclic (rotationally periodic about z) boundaries: coef_cyclic
 * (src/sem/coef.f90:1380-1430) over the cyclic facets (element 0-based,
 * facet 1..6, 2 x nfacets int32, msh%periodic's order): per facet point
 * R11, R12 = the unit normal's x-y direction signed by y n_x - x n_y. */
typedef struct nm_cyclic_s nm_cyclic_t;
int nm_cyclic_create(nm_cyclic_t **cy, nm_op_t *op, const int32_t *facets,
                     int64_t nfacets, hipStream_t s);
void nm_cyclic_free(nm_cyclic_t *cy);
/* entry count; with msk != NULL also Neko's arrays (cyc_msk(1..ncyc),
 * 1-based points, R11, R12) in entry order */
int nm_cyclic_info(const nm_cyclic_t *cy, int64_t *ncyc, int32_t *msk,
                   double *r11, double *r12);
/* rotate_cyc (src/math/operators.f90:1078-1140, opr_cpu_rotate_cyc_r1,
 * src/math/bcknd/cpu/opr_cpu.f90:333-357): idir 1 (vx, vy) -> (normal,
 * tangential) at the cyclic points, 0 back; vz untouched */
int nm_rotate_cyc(const nm_cyclic_t *cy, double *vx, double *vy, double *vz,
                  int idir, hipStream_t s);
/* drop-in for Neko's HIP rotate_cyc (opr_device.F90:159-168,
 * opr_rotate_cyc.hip): cyc_msk[0] = ncyc + 1, then 1-based points */
void hip_rotate_cyc(void *vx, void *vy, void *vz, void *x, void *y, void *z,
                    void *cyc_msk, void *R11, void *R12, int *ncyc, int *idir);
/* the whole pressure residual pnpn_prs_res_cpu_compute (pnpn_res_cpu.f90:
 * 31-180): nm_prs_res, then p_res -= (bd/dt) (u . n A) on the points of
 * prs (bc_prs_surface: the velocity-prescribed facets, u the current
 * velocity) and -= (ta . n A) on those of sym (bc_sym_surface); either
 * surface may be NULL */
int nm_prs_res_surf(nm_op_t *op, double mu, double rho, double bd, double dt,
                    const double *p, const double *ue, const double *ve,
                    const double *we, const double *u, const double *v,
                    const double *w, const double *fx, const double *fy,
                    const double *fz, const nm_surface_t *prs,
                    const nm_surface_t *sym, double *p_res, hipStream_t s);
/* preconditioner of the op's CG/GMRES/batch/serve solvers: 0 Jacobi
 * (src/krylov/bcknd/cpu/pc_jacobi.f90, default), 1 identity (the ksp
 * default without a preconditioner, src/krylov/pc_identity.f90) */
int nm_op_set_pc(nm_op_t *op, int kind, hipStream_t s);
/* autotuner: times all Ax (FP64, FP32) and gs variants on the problem's own
 * data and keeps the fastest; returns the chosen codes */
int nm_op_autotune(nm_op_t *op, int reps, int *best64, int *best32,
                   int *best_gs, hipStream_t s);
/* w = A_local u ; w = gs(A_local u) ; w = mask(gs(A_local u)) (FP64) */
int nm_op_ax_local(const nm_op_t *op, double *w, const double *u,
                   hipStream_t s);
int nm_op_ax_gs(const nm_op_t *op, double *w, const double *u, hipStream_t s);
int nm_op_ax(const nm_op_t *op, double *w, const double *u, hipStream_t s);
/* ax_t%compute_vector (src/math/ax.f90:98-115): au, av, aw = A_local of the
 * three components u, v, w (no gs, no mask) */
int nm_op_ax_local_vector(const nm_op_t *op, double *au, double *av,
                          double *aw, const double *u, const double *v,
                          const double *w, hipStream_t s);
int nm_op_ax_local32(nm_op_t *op, float *w, const float *u, hipStream_t s);
int nm_op_ax_gs32(nm_op_t *op, float *w, const float *u, hipStream_t s);
int nm_op_gs(const nm_op_t *op, double *u, hipStream_t s);
int nm_op_mask(const nm_op_t *op, double *u, hipStream_t s);
/* f = mask(gs(g)) with g_i ~ N(0,1) from (seed, i): the seeded right-hand
 * side of the synthetic workloads, generated on the device */
int nm_op_rhs_normal(const nm_op_t *op, uint64_t seed, double *f,
                     hipStream_t s);
/* Jacobi diagonal with every element treated as deformed (the coarse grid
 * of pc_hsmg.f90:312-318 sets msh%dfrmd_el = .true.) */
int nm_op_set_all_deformed(nm_op_t *op, hipStream_t s);

/* --- solution-space projection (projection_t, src/common/projection.f90) */
/* The last L (<= 32) solutions and their images A x kept A-orthonormal on
 * the device.  pre: project the RHS b in place before a solve
 * (pre_solving/project_on, :236-280,:476-551); post: add the projected part
 * back to the solution x and take x into the basis (post_solving/
 * project_back/proj_ortho, :282-358,:717-834).  tstep and activ_step as in
 * Neko (projection from step activ_step + 1); dt_last_change < 0 for a
 * constant time step (time_step_controller_t%dt_last_change otherwise). */
int nm_proj_create(nm_proj_t **pj, nm_op_t *op, int L, int activ_step);
void nm_proj_free(nm_proj_t *pj);
int nm_proj_clear(nm_proj_t *pj, hipStream_t s);
int nm_proj_pre(nm_proj_t *pj, double *b, int tstep, int dt_last_change,
                hipStream_t s);
int nm_proj_post(nm_proj_t *pj, double *x, int tstep, int dt_last_change,
                 hipStream_t s);
/* basis size m; proj_m and proj_res (original residual) of the last
 * project_on, as print_proj_info logs them */
int nm_proj_info(nm_proj_t *pj, int *m, int *proj_m, double *proj_res,
                 hipStream_t s);

/* --- Jacobi-PCG (ksp_t 'cg' + pc_t 'jacobi') ----------------------------- */
/* precision 0: FP64; 1: FP32 inner CG inside FP64 iterative refinement */
int nm_cg_create(nm_cg_t **cg, nm_op_t *op, int precision);
void nm_cg_free(nm_cg_t *cg);
int nm_cg_solve(nm_cg_t *cg, double *x, const double *f, double abs_tol,
                int max_iter, double inner_rel, nm_monitor_t *mon,
                hipStream_t stream);
/* the solver's own zero-Dirichlet mask (1-based msk, nmsk entries) in place
 * of its operator's, for every operator application of its solves; NULL
 * restores the operator's.  It carries one component's bc list (blstx,
 * blsty or blstz) of the velocity solve, src/fluid/fluid_pnpn.f90:899. */
int nm_cg_set_mask(nm_cg_t *cg, const int32_t *msk, int64_t nmsk,
                   hipStream_t s);
/* ksp_t%solve_coupled of cg (src/krylov/bcknd/cpu/cg.f90:282-305): the
 * three components solved one after the other, component c (x[c], f[c],
 * mon[c]) by cg[c], all on the same operator size, each to abs_tol */
int nm_cg_solve_coupled(nm_cg_t *const *cg, double *const *x,
                        const double *const *f, double abs_tol, int max_iter,
                        double inner_rel, nm_monitor_t *mon,
                        hipStream_t stream);

/* --- hsmg preconditioner (pc_t 'hsmg', src/krylov/pc_hsmg.f90) ---------- */
/* Kernel family of the Schwarz FDM solves of hsmg and phmg (no
 * counterpart in Neko, whose fdm_do_fast has one implementation per
 * backend, src/math/fdm.f90:635-662): 0 the matrix cores (default), 1 the
 * vector ALUs; the same solve to rounding, process-wide, for timing. */
int nm_fdm_set_variant(int variant);
/* Three-level V-cycle lx -> lx_mid (4, or max(lx-1,3) below lx 5) -> 2 on
 * the mesh of a box or mesh operator: overlapping Schwarz with FDM local solves
 * (src/math/schwarz.f90:112-634, src/math/fdm.f90:110-665) on the fine and
 * middle levels, `niter` Jacobi-CG iterations (abs. tol 1e-9) on the
 * coarse level (pc_hsmg.f90:147-365 init, :481-627 solve, CPU branch).
 * facet_bc: facet_type of the boundary facets (1 pressure, 2 velocity;
 * src/fluid/fluid_pnpn.f90 boundary set-up). */
int nm_hsmg_create(nm_hsmg_t **h, nm_op_t *op, int facet_bc, int niter,
                   hipStream_t s);
void nm_hsmg_free(nm_hsmg_t *h);
/* z = M^{-1} r (precon.f90 pc_solve interface: solve(z, r, n)); r is used
 * as scratch copy only, the input stays untouched */
int nm_hsmg_apply(nm_hsmg_t *h, double *z, const double *r, hipStream_t s);
/* fine-level Schwarz smoother alone: e = schwarz(r), r masked in place
 * (schwarz_compute, schwarz.f90:565-595) */
int nm_hsmg_schwarz_fine(nm_hsmg_t *h, double *e, double *r, hipStream_t s);

/* --- phmg-style V-cycle with Chebyshev smoothing -------------------------
 * Chebyshev smoother cheby_t (src/krylov/bcknd/cpu/cheby.f90: power
 * estimate :136-231 with 150 iterations, boost 1.1, lambda_max/30; solve
 * cheby_impl :317-397) accelerated by the level's Schwarz smoother, in the
 * V-cycle of phmg_mg_cycle (src/multigrid/phmg.f90:605-757) on hsmg's
 * levels lx -> lx_mid -> 2 with hsmg's coarse Jacobi-CG (10 iterations) in
 * place of phmg's tree-AMG.  smooth_its: Chebyshev iterations per
 * smoothing (phmg 'smoother_iterations'; 4 = Chebyshev(4)).  Level
 * operators take the fine h1 and ifh2; their h2 is Neko's coef default 1
 * (level_h2_fine = 0, phmg.f90:252-254) or the fine level's h2 (1).  The
 * power method starts from this project's generator (seed 3901), not
 * gfortran's random_number.  smoother_fp32: the Chebyshev/Schwarz
 * smoothing (operator, gs, FDM) runs in FP32 on FP32 copies of the level
 * vectors; transfers, eigenvalue estimates and the coarse CG stay FP64. */
int nm_phmg_create(nm_phmg_t **h, nm_op_t *op, int facet_bc, int smooth_its,
                   int level_h2_fine, int smoother_fp32, hipStream_t s);
void nm_phmg_free(nm_phmg_t *h);
int nm_phmg_apply(nm_phmg_t *h, double *z, const double *r, hipStream_t s);
/* phmg's own coarse solve, the tree-AMG tamg_solver_t
 * (src/multigrid/tree_amg_multigrid.f90:111-227 set-up, tamg_mg_cycle;
 * tree_amg_aggregate.f90; tree_amg.f90; tree_amg_smoother.f90) on the
 * coarse level in place of the Jacobi-CG: `levels` hierarchy levels
 * (elements as aggregates, then greedy aggregation of the element graph,
 * one aggregate last), `iterations` V-cycles per coarse solve, Chebyshev
 * smoothing of degree `cheby_degree` on every level (phmg.f90:150-157
 * defaults 3, 1, 4).  The aggregation order and the power method's start
 * vector come from this project's generator (seed 3901).  levels 0
 * restores the Jacobi-CG. */
int nm_phmg_set_coarse_amg(nm_phmg_t *h, int levels, int iterations,
                           int cheby_degree, hipStream_t s);
/* level sizes of the coarse AMG (nn[0] coarse-level dofs, nn[l] level-l
 * aggregates, up to cap entries); returns the level count, 0 without AMG */
int nm_phmg_coarse_amg_levels(const nm_phmg_t *h, int64_t *nn, int cap);
/* the AMG's map coarse-level dof -> level-lvl aggregate (0-based), nn[0]
 * entries */
int nm_phmg_coarse_amg_map(const nm_phmg_t *h, int lvl, int32_t *out);
/* (tha, dlt) of the fine and the middle level smoothers */
int nm_phmg_eigs(const nm_phmg_t *h, double *tha_dlt);

/* --- GMRES(m) (ksp_t 'gmres', src/krylov/bcknd/cpu/gmres.f90:175-411) --- */
/* pc: 0 Jacobi (the operator's diagonal), 1 hsmg (nm_hsmg_t handle),
 * 2 identity, 3 phmg (nm_phmg_t handle) */
int nm_gmres_create(nm_gmres_t **g, nm_op_t *op, int pc, void *pc_handle,
                    int lgmres);
void nm_gmres_free(nm_gmres_t *g);
int nm_gmres_solve(nm_gmres_t *g, double *x, const double *f, double abs_tol,
                   int max_iter, nm_gmres_monitor_t *mon, hipStream_t s);

/* --- batched solves (no counterpart in Neko; same CG semantics) -------- */
/* Many straight-edged operators (box or mesh constructors, constant h1) of uneven size and
 * order solved together: one FP32 CG iteration of every problem per
 * captured graph step (one batched Ax launch per lx, one merged gs),
 * FP64 residual replacement per problem. */
int nm_batch_create(nm_batch_t **b, int nproblems, nm_op_t *const *ops,
                    hipStream_t s);
void nm_batch_free(nm_batch_t *b);
/* per-lx elements-per-workgroup autotuner (recaptures the graph) */
int nm_batch_autotune(nm_batch_t *b, int reps, int *lx_out, int *epb_out,
                      int *ngroups);
/* x[i], f[i]: device FP64 arrays of problem i; rel_tol relative to each
 * problem's initial residual; delta: FP32 reduction between residual
 * replacements (<= 0: 0.1); mon[i] per problem */
int nm_batch_solve(nm_batch_t *b, double *const *x, const double *const *f,
                   double rel_tol, int max_iter, double delta,
                   nm_monitor_t *mon, hipStream_t stream);
int nm_batch_info(const nm_batch_t *b, int64_t *total_dof,
                  int64_t *device_bytes);
/* statistics of the last nm_batch_solve: graph steps (one FP32 iteration
 * of every unfinished problem each) and the local dof-iterations its Ax
 * launches covered (problems that finished stay covered until the next
 * table compaction, after a quarter of the covered ones have finished) */
int nm_batch_stats(const nm_batch_t *b, int64_t *steps,
                   double *launched_dof_iter);

/* --- serving loop (no counterpart in Neko; drives the CG above) --------- */
typedef struct {
  int state;            /* 0 queued, 1 running, 2 done, 3 failed */
  int iter, outer, converged;
  double res_start, res_final;
  double t_arrive, t_admit, t_done;  /* seconds from the stream's start */
  int64_t n;            /* local dofs of the request's mesh */
} nm_serve_result_t;
/* precision 0 FP64 / 1 mixed; chunk_iters: CG iterations issued per
 * request between host polls; keep_solutions: copy each x aside */
int nm_server_create(nm_server_t **sv, double hbm_budget_bytes,
                     int max_inflight, int precision, int chunk_iters,
                     int keep_solutions);
void nm_server_free(nm_server_t *sv);
/* register a resident Dirichlet box mesh (Poisson, h1 = 1); tune_reps > 0
 * runs the kernel-variant autotuner on it first; prewarm: CG states to
 * allocate up front */
int nm_server_add_box_mesh(nm_server_t *sv, int nelx, int nely, int nelz,
                           int lx, double amp, uint64_t seed, int tune_reps,
                           int prewarm, int *mesh_id, hipStream_t s);
/* the same for a general straight-edged mesh (vertices in Neko's symmetric
 * order and point ids, e.g. from nm_nmsh_read) */
int nm_server_add_mesh(nm_server_t *sv, int64_t nel, int lx, const double *vxyz,
                       const int64_t *pid, int tune_reps, int prewarm,
                       int *mesh_id, hipStream_t s);
int nm_server_mesh_op(nm_server_t *sv, int mesh_id, nm_op_t **op);
/* a request: mesh, RHS seed (nm_op_rhs_normal), relative tolerance,
 * arrival time in seconds after the start of nm_server_run (or of
 * nm_server_serve; negative: now, on a live stream's clock).  Safe to call
 * from another thread while nm_server_serve runs. */
int nm_server_submit(nm_server_t *sv, int mesh_id, uint64_t rhs_seed,
                     double rel_tol, int max_iter, double t_arrive,
                     int *req_id);
/* serve the requests submitted so far (pre-timed arrivals); returns when
 * all have completed */
int nm_server_run(nm_server_t *sv, double *wall_s);
/* live stream: serve requests as other threads submit them, until
 * nm_server_close has been called and every request submitted before it
 * has completed; meshes are registered before it starts.  nm_server_result
 * may be polled meanwhile: a request's fields are final once its state
 * reads 2. */
int nm_server_serve(nm_server_t *sv, double *wall_s);
int nm_server_close(nm_server_t *sv);
/* seconds on the running live stream's clock (0 when none runs) */
double nm_server_clock(nm_server_t *sv);
int nm_server_result(const nm_server_t *sv, int req_id,
                     nm_serve_result_t *r);
int nm_server_solution(const nm_server_t *sv, int req_id, double *x,
                       hipStream_t s);
int nm_server_memory(const nm_server_t *sv, int64_t *used, int64_t *peak);

/* ordered copy helper for host harnesses (kind: hipMemcpyKind) */
int nm_memcpy(void *dst, const void *src, size_t bytes, int kind,
              hipStream_t strm);

#ifdef __cplusplus
}
#endif
#endif /* NEKO_AMD_H */
