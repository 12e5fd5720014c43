// Tree-AMG solver (tamg_solver_t, src/multigrid/tree_amg_multigrid.f90):
// the coarse solve of the p-multigrid V-cycle.
#pragma once
#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

#include "nm_op.h"

namespace nm {

constexpr int NM_TAMG_MAXL = 8;

struct Tamg {
  const Op *op = nullptr;  // finest level (the p-multigrid's coarse level)
  int nlvls = 0;
  int64_t n0 = 0;
  int64_t nn[NM_TAMG_MAXL + 1] = {};  // nodes per level (nn[0] = finest dofs)
  int cheby_degree = 4, max_iter = 1;
  double tha[NM_TAMG_MAXL + 1] = {}, dlt[NM_TAMG_MAXL + 1] = {};
  // device, level l >= 1: finest dof -> node (map), its CSR with finest
  // dofs ascending (agg_*), the node's level l-1 members (node_*, members
  // ascending) and each level l-1 member's node (parent)
  int32_t *map[NM_TAMG_MAXL + 1] = {}, *parent[NM_TAMG_MAXL + 1] = {};
  int32_t *agg_ptr[NM_TAMG_MAXL + 1] = {}, *agg_dof[NM_TAMG_MAXL + 1] = {};
  int32_t *node_ptr[NM_TAMG_MAXL + 1] = {}, *node_dof[NM_TAMG_MAXL + 1] = {};
  std::vector<int32_t> h_map[NM_TAMG_MAXL + 1];  // host copies (info)
  double *wrk_in = nullptr, *wrk_out = nullptr, *ones = nullptr;
  double *x[NM_TAMG_MAXL + 1] = {}, *b[NM_TAMG_MAXL + 1] = {}, *r[NM_TAMG_MAXL + 1] = {};
  double *cd[NM_TAMG_MAXL + 1] = {}, *cw[NM_TAMG_MAXL + 1] = {}, *cr[NM_TAMG_MAXL + 1] = {};
  double *part = nullptr, *out = nullptr, *h_out = nullptr;
  std::vector<void *> owned;
  ~Tamg();
  template <typename T>
  T *alloc(int64_t n);
};

// Build the hierarchy on `op` (a box or mesh operator, whose element facet
// neighbours fneigh (6 per element, 1-based, 0 none) are given): levels,
// V-cycles per solve, Chebyshev degree, power iterations.
int tamg_create(Tamg **out, const Op *op, const std::vector<int32_t> &fneigh, int nlvls,
                int max_iter, int cheby_degree, int power_its, hipStream_t s);
// z = AMG(r) on the finest level, enqueued on s (no host synchronisation)
int tamg_solve(Tamg *h, double *z, const double *r, hipStream_t s);

}  // namespace nm
