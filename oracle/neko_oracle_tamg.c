/*
 * neko_oracle_tamg.c -- CPU FP64 restatement of Neko's tree-AMG solver,
 * the coarse solve of phmg (see neko_oracle_tamg.h).
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Reference files (Neko source tree, src/multigrid/):
 *   tree_amg_multigrid.f90   tamg_mg_init :111-227 (level loop, target
 *                            nnodes/8, stop below 4), tamg_mg_solve,
 *                            tamg_mg_cycle (the V-cycle), calc_resid,
 *                            fill_lvl_map, build_agg_csr
 *   tree_amg_aggregate.f90   aggregate_finest_level (elements as
 *                            aggregates), agg_greedy_first_pass /
 *                            _second_pass, agg_fill_nhbr_info,
 *                            aggregate_greedy, aggregate_end
 *   tree_amg.f90             tamg_matvec_flat_impl (a coarse operator
 *                            applied through the finest level: copy, gs,
 *                            mult, mask, Ax, gs, mask, mult, sum over the
 *                            aggregate), tamg_restriction_operator,
 *                            tamg_prolongation_operator
 *   tree_amg_smoother.f90    amg_cheby_power (250 iterations, boost 1.1,
 *                            lambda_max / 30), amg_cheby_solve
 *   phmg.f90:150-157         defaults: 3 levels, 1 iteration, degree 4
 *
 * Deviations, by construction of the workload: the greedy aggregation's
 * random element order and the power method's start vector come from
 * this project's seeded uniform generator (seed 3901) instead of
 * gfortran's random_number with seed 3901, which is not reproducible
 * outside gfortran (parity with the reference unpinned for those two
 * draws; the product uses the same generator).
 */
#include "neko_oracle_tamg.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SEED_TAMG 3901u

static int64_t pb_n(const or_problem_t *p) {
  return (int64_t)p->nel * p->lx * p->lx * p->lx;
}

static void pb_gs(const or_problem_t *p, double *u) {
  or_gs_op_add(u, (int)pb_n(p), p->gs_m, p->gs_o, p->dg, p->gd, p->gs_nb, p->blk_len,
               p->blk_off, p->gs_buf);
}

/* ---- aggregation -------------------------------------------------------- */

/* agg_greedy_first_pass / _second_pass over a graph of n elements with
 * ns neighbours each (nb[i*ns + side], 1-based, <= 0 for none); agg[i]
 * 1-based aggregate (out), returns the number of aggregates */
static int greedy(int n, int ns, const int *nb, int *agg, int **size_out) {
  int cap = 64, naggs = 0;
  int *size = (int *)malloc(sizeof(int) * cap);
  for (int i = 0; i < n; i++) agg[i] = -1;
  /* random order: Fisher-Yates from the end, j = int(r i) + 1 */
  int *order = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) order[i] = i + 1;
  uint64_t draw = 0;
  for (int i = n; i >= 2; i--) {
    const double r = or_rng_uniform(SEED_TAMG, draw++);
    const int j = (int)(r * (double)i) + 1;
    const int t = order[i - 1];
    order[i - 1] = order[j - 1];
    order[j - 1] = t;
  }
#define NB(i, s) nb[(int64_t)((i) - 1) * ns + (s)]
#define OK(v) ((v) > 0 && (v) <= n)
  for (int t = 0; t < n; t++) {
    const int i = order[t];
    if (agg[i - 1] != -1) continue;
    int free_nbhd = 1;
    for (int s = 0; s < ns; s++) {
      const int v = NB(i, s);
      if (OK(v) && agg[v - 1] != -1) free_nbhd = 0;
    }
    if (!free_nbhd) continue;
    naggs++;
    if (naggs > cap) {
      cap *= 2;
      size = (int *)realloc(size, sizeof(int) * cap);
    }
    agg[i - 1] = naggs;
    size[naggs - 1] = 1;
    for (int s = 0; s < ns; s++) {
      const int v = NB(i, s);
      if (OK(v) && agg[v - 1] == -1) {
        agg[v - 1] = naggs;
        size[naggs - 1]++;
      }
    }
  }
  for (int i = 1; i <= n; i++) {
    if (agg[i - 1] != -1) continue;
    int best = -1, best_size = 2147483647;
    for (int s = 0; s < ns; s++) {
      const int v = NB(i, s);
      if (OK(v) && agg[v - 1] != -1) {
        const int a = agg[v - 1];
        if (size[a - 1] < best_size) {
          best_size = size[a - 1];
          best = a;
        }
      }
    }
    if (best != -1) {
      agg[i - 1] = best;
      size[best - 1]++;
    } else {
      naggs++;
      if (naggs > cap) {
        cap *= 2;
        size = (int *)realloc(size, sizeof(int) * cap);
      }
      agg[i - 1] = naggs;
      size[naggs - 1] = 1;
      for (int s = 0; s < ns; s++) {
        const int v = NB(i, s);
        if (OK(v) && agg[v - 1] == -1) {
          size[naggs - 1]++;
          agg[v - 1] = naggs;
        }
      }
    }
  }
#undef NB
#undef OK
  free(order);
  *size_out = size;
  return naggs;
}

/* agg_fill_nhbr_info: the distinct neighbouring aggregates of every
 * aggregate in discovery order (element ascending, side ascending); -1
 * pads; returns the column count */
static int fill_nhbr(int n, int ns, const int *nb, const int *agg, int naggs, int **out) {
  int *cnt = (int *)calloc((size_t)naggs, sizeof(int));
  int m = 0;
  for (int i = 1; i <= n; i++) {
    const int a = agg[i - 1];
    for (int s = 0; s < ns; s++) {
      const int v = nb[(int64_t)(i - 1) * ns + s];
      if (v > 0 && v <= n && agg[v - 1] != a) cnt[a - 1]++;
    }
    if (cnt[a - 1] > m) m = cnt[a - 1];
  }
  if (m < 1) m = 1;
  int *g = (int *)malloc(sizeof(int) * (size_t)naggs * m);
  for (int64_t q = 0; q < (int64_t)naggs * m; q++) g[q] = -1;
  for (int i = 1; i <= n; i++) {
    const int a = agg[i - 1];
    for (int s = 0; s < ns; s++) {
      const int v = nb[(int64_t)(i - 1) * ns + s];
      if (!(v > 0 && v <= n)) continue;
      const int t = agg[v - 1];
      if (t == a) continue;
      int added = 0;
      for (int j = 0; j < m; j++) {
        int *slot = &g[(int64_t)(a - 1) * m + j];
        if (*slot == t) {
          added = 1;
        } else if (*slot == -1 && !added) {
          *slot = t;
          added = 1;
        }
      }
    }
  }
  free(cnt);
  *out = g;
  return m;
}

static void csr_from_parent(const int32_t *parent, int64_t n, int64_t nn, int64_t **ptr_out,
                            int32_t **dof_out) {
  int64_t *ptr = (int64_t *)calloc((size_t)nn + 1, sizeof(int64_t));
  int32_t *dof = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) ptr[parent[i] + 1]++;
  for (int64_t c = 0; c < nn; c++) ptr[c + 1] += ptr[c];
  int64_t *fill = (int64_t *)calloc((size_t)nn + 1, sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) {
    const int32_t c = parent[i];
    dof[ptr[c] + fill[c]++] = (int32_t)i;
  }
  free(fill);
  *ptr_out = ptr;
  *dof_out = dof;
}

int or_tamg_init(or_tamg_t *h, const or_problem_t *pb, const int *fneigh, int nlvls,
                 int max_iter, int cheby_degree, int power_its) {
  memset(h, 0, sizeof(*h));
  if (nlvls < 2 || nlvls > OR_TAMG_MAXL) return -1;
  h->pb = pb;
  h->max_iter = max_iter;
  h->cheby_degree = cheby_degree;
  h->power_its = power_its;
  const int nel = pb->nel;
  const int64_t n3 = (int64_t)pb->lx * pb->lx * pb->lx;
  const int64_t n0 = pb_n(pb);
  h->n0 = n0;
  h->nn[0] = n0;
  /* level 1: elements */
  h->nn[1] = nel;
  h->parent[1] = (int32_t *)malloc(sizeof(int32_t) * n0);
  for (int64_t i = 0; i < n0; i++) h->parent[1][i] = (int32_t)(i / n3);
  /* levels 2 .. nlvls-1: greedy on the level below's graph */
  int ns = 6;
  int *graph = (int *)malloc(sizeof(int) * (size_t)nel * 6);
  memcpy(graph, fneigh, sizeof(int) * (size_t)nel * 6);
  int levels = nlvls;
  for (int l = 2; l <= nlvls - 1; l++) {
    const int nprev = (int)h->nn[l - 1];
    if (nprev / 8 < 4) {  /* too few nodes for another level */
      levels = l;
      break;
    }
    int *agg = (int *)malloc(sizeof(int) * nprev), *size = NULL;
    const int naggs = greedy(nprev, ns, graph, agg, &size);
    int *ng = NULL;
    const int m = fill_nhbr(nprev, ns, graph, agg, naggs, &ng);
    free(graph);
    graph = ng;
    ns = m;
    h->nn[l] = naggs;
    h->parent[l] = (int32_t *)malloc(sizeof(int32_t) * nprev);
    for (int i = 0; i < nprev; i++) h->parent[l][i] = agg[i] - 1;
    free(agg);
    free(size);
  }
  free(graph);
  /* last level: one aggregate of every node of the level below */
  h->nlvls = levels;
  h->nn[levels] = 1;
  h->parent[levels] = (int32_t *)calloc((size_t)h->nn[levels - 1], sizeof(int32_t));
  /* node CSR per level (dofs ascending = the reference's element loop order) */
  for (int l = 1; l <= levels; l++)
    csr_from_parent(h->parent[l], h->nn[l - 1], h->nn[l], &h->node_ptr[l], &h->node_dof[l]);
  /* finest -> level maps and their CSR (fill_lvl_map, build_agg_csr) */
  for (int l = 1; l <= levels; l++) {
    h->map[l] = (int32_t *)malloc(sizeof(int32_t) * n0);
    for (int64_t i = 0; i < n0; i++)
      h->map[l][i] = l == 1 ? h->parent[1][i] : h->parent[l][h->map[l - 1][i]];
    csr_from_parent(h->map[l], n0, h->nn[l], &h->agg_ptr[l], &h->agg_dof[l]);
  }
  h->wrk_in = (double *)malloc(sizeof(double) * n0);
  h->wrk_out = (double *)malloc(sizeof(double) * n0);
  for (int l = 0; l < levels; l++) {
    const int64_t n = h->nn[l];
    h->x[l] = (double *)calloc((size_t)n, sizeof(double));
    h->b[l] = (double *)calloc((size_t)n, sizeof(double));
    h->r[l] = (double *)calloc((size_t)n, sizeof(double));
    h->cd[l] = (double *)calloc((size_t)n, sizeof(double));
    h->cw[l] = (double *)calloc((size_t)n, sizeof(double));
    h->cr[l] = (double *)calloc((size_t)n, sizeof(double));
  }
  /* amg_cheby_power on every smoother level 0 .. levels-1 */
  for (int l = 0; l < levels; l++) {
    const int64_t n = h->nn[l];
    double *d = h->cd[l], *w = h->cw[l];
    for (int64_t i = 0; i < n; i++) d[i] = or_rng_uniform(SEED_TAMG, (uint64_t)i) + 10.0;
    if (l == 0) {
      pb_gs(pb, d);
      or_mask_apply(d, pb->msk, pb->nmsk);
    }
    for (int it = 0; it < power_its; it++) {
      or_tamg_matvec(h, w, d, l);
      double wtw = 0.0;
      if (l == 0) wtw = or_glsc3(w, pb->mult, w, n);
      else
        for (int64_t i = 0; i < n; i++) wtw = wtw + w[i] * w[i];
      const double s = 1.0 / sqrt(wtw);
      for (int64_t i = 0; i < n; i++) d[i] = w[i] * s;
    }
    or_tamg_matvec(h, w, d, l);
    double dtw = 0.0, dtd = 0.0;
    if (l == 0) {
      dtw = or_glsc3(d, pb->mult, w, n);
      dtd = or_glsc3(d, pb->mult, d, n);
    } else {
      for (int64_t i = 0; i < n; i++) dtw = dtw + d[i] * w[i];
      for (int64_t i = 0; i < n; i++) dtd = dtd + d[i] * d[i];
    }
    const double lam = dtw / dtd, bb = lam * 1.1, a = lam / 30.0;
    h->tha[l] = (bb + a) / 2.0;
    h->dlt[l] = (bb - a) / 2.0;
  }
  return 0;
}

void or_tamg_free(or_tamg_t *h) {
  for (int l = 0; l <= OR_TAMG_MAXL; l++) {
    free(h->node_ptr[l]);
    free(h->node_dof[l]);
    free(h->parent[l]);
    free(h->map[l]);
    free(h->agg_ptr[l]);
    free(h->agg_dof[l]);
    free(h->x[l]);
    free(h->b[l]);
    free(h->r[l]);
    free(h->cd[l]);
    free(h->cw[l]);
    free(h->cr[l]);
  }
  free(h->wrk_in);
  free(h->wrk_out);
  memset(h, 0, sizeof(*h));
}

/* tamg_matvec_flat_impl */
void or_tamg_matvec(or_tamg_t *h, double *out, const double *in, int lvl) {
  const or_problem_t *pb = h->pb;
  const int64_t n0 = h->n0;
  if (lvl == 0) {
    or_ax_full(pb, out, in);
    return;
  }
  const int64_t nagg = h->nn[lvl];
  const int64_t *ap = h->agg_ptr[lvl];
  const int32_t *ad = h->agg_dof[lvl];
  for (int64_t c = 0; c < nagg; c++)
    for (int64_t k = ap[c]; k < ap[c + 1]; k++) h->wrk_in[ad[k]] = in[c];
  pb_gs(pb, h->wrk_in);
  for (int64_t i = 0; i < n0; i++) h->wrk_in[i] = h->wrk_in[i] * pb->mult[i];
  or_mask_apply(h->wrk_in, pb->msk, pb->nmsk);
  or_ax_full(pb, h->wrk_out, h->wrk_in);
  for (int64_t i = 0; i < n0; i++) h->wrk_out[i] = h->wrk_out[i] * pb->mult[i];
  for (int64_t c = 0; c < nagg; c++) {
    double acc = 0.0;
    for (int64_t k = ap[c]; k < ap[c + 1]; k++) acc = acc + h->wrk_out[ad[k]];
    out[c] = acc;
  }
}

/* tamg_restriction_operator: out (level l) from in (level l-1) */
static void restrict_to(or_tamg_t *h, double *out, double *in, int l) {
  if (l - 1 == 0)
    for (int64_t i = 0; i < h->n0; i++) in[i] = in[i] * h->pb->mult[i];
  const int64_t *np = h->node_ptr[l];
  const int32_t *nd = h->node_dof[l];
  for (int64_t j = 0; j < h->nn[l]; j++) {
    double acc = 0.0;
    for (int64_t k = np[j]; k < np[j + 1]; k++) acc = acc + in[nd[k]];
    out[j] = acc;
  }
}

/* tamg_prolongation_operator: out (level l-1) from in (level l) */
static void prolong_from(or_tamg_t *h, double *out, const double *in, int l) {
  const int64_t n = h->nn[l - 1];
  for (int64_t i = 0; i < n; i++) out[i] = 0.0;
  const int64_t *np = h->node_ptr[l];
  const int32_t *nd = h->node_dof[l];
  for (int64_t j = 0; j < h->nn[l]; j++)
    for (int64_t k = np[j]; k < np[j + 1]; k++) out[nd[k]] = out[nd[k]] + in[j];
  if (l - 1 == 0) {
    pb_gs(h->pb, out);
    for (int64_t i = 0; i < n; i++) out[i] = out[i] * h->pb->mult[i];
    or_mask_apply(out, h->pb->msk, h->pb->nmsk);
  }
}

/* amg_cheby_solve */
void or_tamg_cheby(or_tamg_t *h, double *x, const double *f, int lvl, int zero_init) {
  const int64_t n = h->nn[lvl];
  double *w = h->cw[lvl], *r = h->cr[lvl], *d = h->cd[lvl];
  for (int64_t i = 0; i < n; i++) r[i] = f[i];
  if (!zero_init) {
    or_tamg_matvec(h, w, x, lvl);
    for (int64_t i = 0; i < n; i++) r[i] = r[i] - w[i];
  }
  const double thet = h->tha[lvl], delt = h->dlt[lvl];
  const double s1 = thet / delt;
  double rhok = 1.0 / s1;
  for (int64_t i = 0; i < n; i++) {
    d[i] = 1.0 / thet * r[i];
    x[i] = x[i] + d[i];
  }
  for (int iter = 2; iter <= h->cheby_degree; iter++) {
    or_tamg_matvec(h, w, d, lvl);
    const double rhokp1 = 1.0 / (2.0 * s1 - rhok);
    const double tmp1 = rhokp1 * rhok;
    const double tmp2 = 2.0 * rhokp1 / delt;
    rhok = rhokp1;
    for (int64_t i = 0; i < n; i++) {
      r[i] = r[i] - w[i];
      d[i] = tmp1 * d[i] + tmp2 * r[i];
      x[i] = x[i] + d[i];
    }
  }
}

/* tamg_mg_cycle */
static void cycle(or_tamg_t *h, int *zero_init) {
  const int max_lvl = h->nlvls - 1;
  for (int l = 0; l < max_lvl; l++) {
    const int64_t n = h->nn[l];
    or_tamg_cheby(h, h->x[l], h->b[l], l, *zero_init);
    or_tamg_matvec(h, h->r[l], h->x[l], l);
    for (int64_t i = 0; i < n; i++) h->r[l][i] = -1.0 * h->r[l][i] + h->b[l][i];
    restrict_to(h, h->b[l + 1], h->r[l], l + 1);
    for (int64_t i = 0; i < h->nn[l + 1]; i++) h->x[l + 1][i] = 0.0;
    *zero_init = 1;
  }
  or_tamg_cheby(h, h->x[max_lvl], h->b[max_lvl], max_lvl, *zero_init);
  *zero_init = 0;
  for (int l = max_lvl - 1; l >= 0; l--) {
    const int64_t n = h->nn[l];
    prolong_from(h, h->r[l], h->x[l + 1], l + 1);
    for (int64_t i = 0; i < n; i++) h->x[l][i] = h->x[l][i] + h->r[l][i];
    or_tamg_cheby(h, h->x[l], h->b[l], l, 0);
  }
}

void or_tamg_solve(or_tamg_t *h, double *z, const double *r) {
  const int64_t n = h->n0;
  for (int64_t i = 0; i < n; i++) {
    h->x[0][i] = 0.0;
    h->b[0][i] = r[i];
  }
  int zero_init = 1;
  for (int it = 0; it < h->max_iter; it++) {
    cycle(h, &zero_init);
    zero_init = 0;
  }
  memcpy(z, h->x[0], sizeof(double) * n);
}

or_tamg_t *or_tamg_new(const or_problem_t *pb, const int *fneigh, int nlvls, int max_iter,
                       int cheby_degree, int power_its) {
  or_tamg_t *h = (or_tamg_t *)calloc(1, sizeof(or_tamg_t));
  if (or_tamg_init(h, pb, fneigh, nlvls, max_iter, cheby_degree, power_its)) {
    free(h);
    return NULL;
  }
  return h;
}

void or_tamg_delete(or_tamg_t *h) {
  if (!h) return;
  or_tamg_free(h);
  free(h);
}

int or_tamg_levels(const or_tamg_t *h, int64_t *nn) {
  for (int l = 0; l <= h->nlvls; l++) nn[l] = h->nn[l];
  return h->nlvls;
}

void or_tamg_map(const or_tamg_t *h, int lvl, int32_t *out) {
  memcpy(out, h->map[lvl], sizeof(int32_t) * h->n0);
}

void or_tamg_eigs(const or_tamg_t *h, double *tha, double *dlt) {
  for (int l = 0; l < h->nlvls; l++) {
    tha[l] = h->tha[l];
    dlt[l] = h->dlt[l];
  }
}
