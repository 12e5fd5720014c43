/*
 * neko_oracle.h -- CPU FP64 restatement of Neko's linear-solve hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never links
 * or calls it.
 *
 * Conventions (identical to Neko's Fortran): every field is column-major
 * (lx,lx,lx,nel), i.e. C index i + lx*(j + lx*(k + lx*e)), 0-based here.
 * Derivative matrices are column-major lx*lx: D(i,j) = d[i + lx*j].
 * gs maps (dg, gd, blk_len, blk_off) are 1-based, exactly as Neko stores
 * them.
 */
#ifndef NEKO_ORACLE_H
#define NEKO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- synthetic input generator (shared specification with the product) */
uint64_t or_rng_u64(uint64_t seed, uint64_t idx);
double or_rng_uniform(uint64_t seed, uint64_t idx);
double or_rng_normal(uint64_t seed, uint64_t idx);

/* ---- a1: GLL space (speclib.f90 / space.f90) */
void or_zwgll(double *z, double *w, int np);
void or_dgll(double *d, double *dt, const double *z, int nz);

/* ---- a2: box mesh, connectivity, dofmap */
void or_box_mesh(int nelx, int nely, int nelz, int perx, int pery, int perz,
                 double amp, uint64_t seed, double *vxyz, int64_t *pid);
void or_facet_neigh(int nel, const int64_t *pid, int *fneigh);
void or_dfrmd_el(int nel, const double *vxyz, int *dfrmd);
void or_dof_ids(int nel, int lx, const int64_t *pid, int64_t *dof);
void or_xyz(int nel, int lx, const double *zg, const double *vxyz,
            double *x, double *y, double *z);
/* curved elements (dofmap_generate_xyz, dofmap.f90:797-1285): type 4
 * (midpoint, Gordon-Hall quadratic) and type 3 (circular arc) edges */
int or_curve_element(int lx, const double *zg, const double *v, const double *cd,
                     const int *type, double *x, double *y, double *z);
void or_xyz_curved(int nel, int lx, const double *zg, const double *vxyz, int ncurves,
                   const int *cel, const double *cdata, const int *ctype, double *x,
                   double *y, double *z, int *status);

/* ---- a6: gs mapping (gather_scatter.f90: gs_init_mapping) */
int or_gs_register(int nel, int lx, const int64_t *dof, const int *fneigh,
                   int *dg, int *gd, int *facet_offset_out);
void or_gs_qsort_dofmap(int *dg, int *gd, int n);
int or_gs_map(int nel, int lx, const int64_t *dof, const int *fneigh,
              int *dg, int *gd, int *nlocal, int *facet_offset,
              int *blk_len, int *blk_off, int *nblks);

/* ---- a5: gs ADD on the local path (gs_cpu.f90) */
void or_gs_op_add(double *u, int n, int m, int o, const int *dg,
                  const int *gd, int nb, const int *b, const int *bo,
                  double *v);

/* ---- a3: coef */
void or_coef(int nel, int lx, const double *x, const double *y,
             const double *z, const double *d, const double *w3,
             double *g11, double *g22, double *g33, double *g12,
             double *g13, double *g23, double *jac, double *B);

/* ---- a4: ax_helm (ax_helm_cpu.f90) */
void or_ax_helm(double *w, const double *u, const double *d, const double *dt,
                const double *h1, const double *g11, const double *g22,
                const double *g33, const double *g12, const double *g13,
                const double *g23, int nel, int lx,
                const double *h2, const double *B, int ifh2);

/* ---- a7: Dirichlet mask */
int or_mask(int nel, int lx, const int *fneigh, int *msk);
void or_mask_apply(double *x, const int *msk, int m);

/* ---- a8: glsc3 */
double or_glsc3(const double *a, const double *b, const double *c, int64_t n);

/* ---- a12: Jacobi (pc_jacobi.f90 CPU formula) */
/* ---- pnpn residual pieces (neko_oracle_pnpn.c) */
void or_cofactors(int nel, int lx, const double *x, const double *y, const double *z,
                  const double *d, double *cof);
void or_opgrad(int nel, int lx, const double *u, const double *d, const double *w3,
               const double *cof, double *ux, double *uy, double *uz);
void or_dudxyz(int nel, int lx, const double *u, const double *dr, const double *ds,
               const double *dt, const double *jac, const double *d, double *du);
void or_cdtp(int nel, int lx, const double *x, const double *dr, const double *ds,
             const double *dt, const double *d, const double *w3, double *dtx);

void or_jacobi_diag(double *d, const double *dt, const double *g11,
                    const double *g22, const double *g33, const double *g12,
                    const double *g13, const double *g23, const int *dfrmd,
                    int nel, int lx);
/* diagonal of the reference's device kernel (pc_jacobi.hip:42-104) */
void or_jacobi_diag_dev(double *d, const double *dt, const double *g11,
                        const double *g22, const double *g33, const double *g12,
                        const double *g13, const double *g23, int nel, int lx);

/* ---- a10: Jacobi-PCG (cg.f90) */
typedef struct {
  int nel, lx;
  const double *d, *dt;
  const double *h1, *h2, *B;
  int ifh2;
  const double *g11, *g22, *g33, *g12, *g13, *g23;
  const double *mult;
  int gs_m, gs_o, gs_nb;
  const int *dg, *gd, *blk_len, *blk_off;
  double *gs_buf;
  const int *msk;
  int nmsk;
  double volume;
} or_problem_t;

typedef struct {
  int iter;
  double res_start, res_final;
  int converged;
} or_monitor_t;

/* record rnorm of every iteration of the next or_cg_solve calls (NULL: off) */
void or_cg_history(double *hist, int cap);
or_monitor_t or_cg_solve(const or_problem_t *p, const double *pcd,
                         double *x, const double *f, int max_iter,
                         double abs_tol);
void or_ax_full(const or_problem_t *p, double *w, const double *u);

/* ---- a11: GMRES(m) (gmres.f90) with a Jacobi or no preconditioner */
or_monitor_t or_gmres_solve(const or_problem_t *p, const double *pcd,
                            double *x, const double *f, int max_iter,
                            double abs_tol, int lgmres);

#ifdef __cplusplus
}
#endif
#endif
