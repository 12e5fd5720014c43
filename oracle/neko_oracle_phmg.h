/*
 * neko_oracle_phmg.h -- CPU FP64 restatement of Neko's Chebyshev smoother
 * (cheby_t) and of a p-multigrid V-cycle with Chebyshev/Schwarz smoothing
 * (phmg_t's cycle) whose coarse solve is hsmg's Jacobi-CG.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 */
#ifndef NEKO_ORACLE_PHMG_H
#define NEKO_ORACLE_PHMG_H
#include "neko_oracle.h"
#include "neko_oracle_hsmg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  const or_problem_t *pb; /* level operator: ax, gs, mask, mult */
  or_schwarz_t *sw;       /* Schwarz accelerator, or NULL: Jacobi */
  const double *pcd;      /* Jacobi 1/diag (sw == NULL) */
  double tha, dlt;
  double *d, *w, *r;      /* n each */
} or_cheby_t;

/* cheby_power (cheby.f90:136-231) from the start vector U(seed, i) + 10 */
void or_cheby_power(or_cheby_t *c, uint64_t seed, int its);
/* cheby_impl (cheby.f90:317-397): niter iterations on A x = f */
void or_cheby_solve(or_cheby_t *c, double *x, const double *f, int niter,
                    int zero_init);

typedef struct {
  int nel, lf, lm, lc;
  const or_problem_t *pb[3];
  or_cheby_t ch[2];
  int smooth_its;
  const double *jfm, *jfm_t, *jmc, *jmc_t;
  const double *pc_crs;
  int crs_niter;
  double crs_tol;
  double *r[3], *z[3], *w[3];
  void *amg;  /* or_tamg_t on the coarse level: the coarse solve (else CG) */
} or_phmg_t;

/* z = M^{-1} r: phmg_mg_cycle (phmg.f90:605-757) with levels lf -> lm ->
 * lc, Chebyshev pre/post smoothing, coarse Jacobi-CG */
void or_phmg_solve(or_phmg_t *h, double *z, const double *r);
or_monitor_t or_gmres_phmg_solve(const or_problem_t *p, or_phmg_t *h,
                                 double *x, const double *f, int max_iter,
                                 double abs_tol, int lgmres);

#ifdef __cplusplus
}
#endif
#endif
