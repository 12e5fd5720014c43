/*
 * neko_oracle_tamg.h -- CPU FP64 restatement of Neko's tree-AMG coarse
 * solver (tamg_solver_t), the coarse solve of phmg.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 */
#ifndef NEKO_ORACLE_TAMG_H
#define NEKO_ORACLE_TAMG_H
#include "neko_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OR_TAMG_MAXL 8

typedef struct {
  const or_problem_t *pb;  /* finest level: the operator (ax, gs, mask, mult) */
  int nlvls;               /* levels of the hierarchy after aggregation */
  int64_t n0;              /* finest local dofs */
  int64_t nn[OR_TAMG_MAXL + 1];     /* level l nodes (l >= 1); nn[0] = n0 */
  /* level l >= 1: node j's dofs (0-based indices into level l-1) in CSR */
  int64_t *node_ptr[OR_TAMG_MAXL + 1];
  int32_t *node_dof[OR_TAMG_MAXL + 1];
  int32_t *parent[OR_TAMG_MAXL + 1];  /* level l-1 index -> its level l node */
  int32_t *map[OR_TAMG_MAXL + 1];     /* finest dof -> level l node */
  int64_t *agg_ptr[OR_TAMG_MAXL + 1]; /* CSR of map: finest dofs ascending */
  int32_t *agg_dof[OR_TAMG_MAXL + 1];
  int cheby_degree, max_iter, power_its;
  double tha[OR_TAMG_MAXL + 1], dlt[OR_TAMG_MAXL + 1];
  double *wrk_in, *wrk_out;                        /* finest */
  double *x[OR_TAMG_MAXL + 1], *b[OR_TAMG_MAXL + 1], *r[OR_TAMG_MAXL + 1];
  double *cd[OR_TAMG_MAXL + 1], *cw[OR_TAMG_MAXL + 1], *cr[OR_TAMG_MAXL + 1];
} or_tamg_t;

/* tamg_mg_init (tree_amg_multigrid.f90:111-227): elements as aggregates,
 * greedy aggregation of the element graph fneigh (6 per element, 1-based,
 * 0 = none) on levels 2..nlvls-1, one aggregate on the last level;
 * Chebyshev smoothers of `cheby_degree` iterations with the power estimate
 * (power_its iterations) from U(seed 3901, i) + 10.  Returns 0 or -1. */
int or_tamg_init(or_tamg_t *h, const or_problem_t *pb, const int *fneigh,
                 int nlvls, int max_iter, int cheby_degree, int power_its);
void or_tamg_free(or_tamg_t *h);
/* heap-allocated handle with accessors (ctypes) */
or_tamg_t *or_tamg_new(const or_problem_t *pb, const int *fneigh, int nlvls, int max_iter,
                       int cheby_degree, int power_its);
void or_tamg_delete(or_tamg_t *h);
int or_tamg_levels(const or_tamg_t *h, int64_t *nn);           /* nn[0..nlvls] */
void or_tamg_map(const or_tamg_t *h, int lvl, int32_t *out);   /* finest -> lvl */
void or_tamg_eigs(const or_tamg_t *h, double *tha, double *dlt);
/* tamg_mg_solve: z = max_iter V-cycles from zero on A z = r (finest) */
void or_tamg_solve(or_tamg_t *h, double *z, const double *r);
/* the level-l operator (tamg_matvec_flat_impl) */
void or_tamg_matvec(or_tamg_t *h, double *out, const double *in, int lvl);
/* one Chebyshev smoothing of level l (amg_cheby_solve) */
void or_tamg_cheby(or_tamg_t *h, double *x, const double *f, int lvl, int zero_init);

#ifdef __cplusplus
}
#endif
#endif
