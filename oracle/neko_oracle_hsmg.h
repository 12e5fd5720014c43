/* neko_oracle_hsmg.h -- oracle restatement of hsmg (test infrastructure). */
#ifndef NEKO_ORACLE_HSMG_H
#define NEKO_ORACLE_HSMG_H
#include <stdint.h>

#include "neko_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n, m, o, nb;  /* field size, gs entries, facet offset, blocks */
  const int *dg, *gd, *b, *bo;
  double *buf;      /* m entries */
} or_gs_t;

typedef struct {
  int lx, nel;
  or_gs_t gs;       /* the level's gs */
  or_gs_t gs_ext;   /* gs of the (lx+2)^3 extended dofmap */
  const int *msk;
  int nmsk;
  const double *s, *d, *wt;
  double *work1, *work2;  /* (lx+2)^3 * nel each */
} or_schwarz_t;

typedef struct {
  int nel, lf, lm, lc;
  int64_t nf, nm, nc;
  or_schwarz_t sw_fine, sw_mid;
  or_gs_t gs_crs;
  const int *msk_c;
  int nmsk_c;
  const double *mult_f, *mult_m;
  const double *jfm, *jfm_t, *jmc, *jmc_t;
  or_problem_t crs;          /* coarse operator (h1 from the fine level) */
  const double *pc_crs;      /* coarse Jacobi 1/diag */
  int niter;
  double crs_tol;
  double *r, *w, *e_mid, *e_crs;
  int last_crs_iter;
} or_hsmg_t;

typedef void (*or_pc_fn)(void *ctx, double *z, const double *r, int64_t n);

void or_fd_weights(double xi, const double *x, int n, int m, double *c);
void or_setup_intp(double *jh, double *jht, const double *z_to,
                   const double *z_from, int n_to, int n_from, int deriv);
void or_tnsr3d(double *v, int nv, const double *u, int nu, const double *A,
               const double *Bt, const double *Ct, int nel);
void or_semhat(int n, double *ah, double *bh);
void or_gs_apply(const or_gs_t *g, double *u);
void or_fdm_1d(double *a, double *b, double *s, double *lam, int lbc, int rbc,
               double ll, double lm, double lr, int n);
void or_plane_space(double *lr, double *ls, double *lt, const double *w,
                    const double *x, const double *y, const double *z, int lx,
                    int nel);
void or_fdm_setup(int nel, int lx, const double *x, const double *y,
                  const double *z, const double *wx, const int *facet_type,
                  const or_gs_t *gs, double *s, double *d);
void or_fdm_compute(int nel, int nl, double *e, double *r, const double *s,
                    const double *d);
void or_schwarz_extrude(double *arr1, int l1, double f1, const double *arr2,
                        int l2, double f2, int nx, int nel);
void or_schwarz_extrude_single(double *arr, int l1, double f1, int l2,
                               double f2, int nx, int nel);
void or_schwarz_toext3d(double *a, const double *b, int n, int nel);
void or_schwarz_toreg3d(double *b, const double *a, int n, int nel);
void or_schwarz_setup_wt(int nel, int lx, const or_gs_t *gs_ext,
                         const or_gs_t *gs, double *wt);
void or_schwarz_wt3d(double *e, const double *wt, int n, int nel);
void or_schwarz_compute(const or_schwarz_t *sw, double *e, double *r);
void or_hsmg_solve(or_hsmg_t *h, double *z, const double *r);
or_monitor_t or_gmres_solve_pc(const or_problem_t *p, or_pc_fn pc, void *ctx,
                               double *x, const double *f, int max_iter,
                               double abs_tol, int lgmres);
or_monitor_t or_gmres_hsmg_solve(const or_problem_t *p, or_hsmg_t *h,
                                 double *x, const double *f, int max_iter,
                                 double abs_tol, int lgmres);

#ifdef __cplusplus
}
#endif
#endif
