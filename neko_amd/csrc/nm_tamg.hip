// Tree-AMG (tamg_solver_t, src/multigrid/tree_amg_multigrid.f90:111-227
// set-up, tamg_mg_solve / tamg_mg_cycle; tree_amg_aggregate.f90;
// tree_amg.f90 tamg_matvec_flat_impl, restriction / prolongation;
// tree_amg_smoother.f90 amg_cheby_power / amg_cheby_solve): the coarse
// solve of the p-multigrid V-cycle.
//
// Level 1 takes the elements as aggregates, levels 2 .. L-1 aggregate the
// level below greedily over its neighbour graph (element faces first, then
// the aggregates' neighbour lists), level L is one aggregate.  No coarse
// matrix is formed: a level-l product runs through the finest level --
// copy each aggregate's value to its finest dofs, gs, mult, mask, the
// finest operator (Ax, gs, mask), mult, sum over the aggregate.  Every
// level smooths with Chebyshev of fixed degree (eigenvalue bounds from a
// power method at set-up), the last level included; transfers between
// levels are piecewise constant.  The aggregation's random element order
// and the power method's start vector use this project's seeded generator
// (seed 3901), where Neko draws from gfortran's random_number.
//
// Device work per product: gathers, the operator, and sums over
// aggregates (one wave per aggregate, a fixed lane order; the single
// aggregate of the last level by the fixed-grid reduction), so results
// are deterministic run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "neko_amd.h"
#include "nm_common.h"
#include "nm_kernels.h"
#include "nm_tamg.h"

namespace nm {

Tamg::~Tamg() {
  for (void *p : owned) (void)hipFree(p);
  if (h_out) (void)hipHostFree(h_out);
}

template <typename T>
T *Tamg::alloc(int64_t n) {
  void *p = nullptr;
  if (hipMalloc(&p, sizeof(T) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess) return nullptr;
  owned.push_back(p);
  return (T *)p;
}

namespace {

constexpr uint64_t kSeed = 3901;

// ---- host: aggregation ------------------------------------------------------

// Greedy aggregation of a graph of n nodes with `ns` neighbour slots each
// (1-based neighbour, <= 0 or > n: none): first pass in a random order
// makes a node whose neighbours are all free the root of a new aggregate
// with them; the second pass gives each remaining node to its smallest
// neighbouring aggregate (the first of equals), or roots a new one.
// Returns the 0-based aggregate of every node.
std::vector<int32_t> greedy_aggregate(int n, int ns, const std::vector<int32_t> &nbr,
                                      int *naggs_out) {
  std::vector<int32_t> agg(n, -1);
  std::vector<int32_t> sz;
  std::vector<int32_t> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  uint64_t draw = 0;
  for (int i = n; i >= 2; i--) {  // Fisher-Yates, j = int(u i) + 1 (1-based)
    const int j = (int)(nm_rng_uniform(kSeed, draw++) * (double)i);
    std::swap(order[i - 1], order[j]);
  }
  auto nb = [&](int i, int s) {
    const int32_t v = nbr[(size_t)i * ns + s];
    return (v > 0 && v <= n) ? v - 1 : -1;
  };
  for (int i : order) {
    if (agg[i] >= 0) continue;
    bool free_nbhd = true;
    for (int s = 0; s < ns && free_nbhd; s++) {
      const int v = nb(i, s);
      if (v >= 0 && agg[v] >= 0) free_nbhd = false;
    }
    if (!free_nbhd) continue;
    const int a = (int)sz.size();
    agg[i] = a;
    sz.push_back(1);
    for (int s = 0; s < ns; s++) {
      const int v = nb(i, s);
      if (v >= 0 && agg[v] < 0) {
        agg[v] = a;
        sz[a]++;
      }
    }
  }
  for (int i = 0; i < n; i++) {
    if (agg[i] >= 0) continue;
    int best = -1;
    for (int s = 0; s < ns; s++) {
      const int v = nb(i, s);
      if (v >= 0 && agg[v] >= 0 && (best < 0 || sz[agg[v]] < sz[best])) best = agg[v];
    }
    if (best >= 0) {
      agg[i] = best;
      sz[best]++;
      continue;
    }
    const int a = (int)sz.size();
    agg[i] = a;
    sz.push_back(1);
    for (int s = 0; s < ns; s++) {
      const int v = nb(i, s);
      if (v >= 0 && agg[v] < 0) {
        agg[v] = a;
        sz[a]++;
      }
    }
  }
  *naggs_out = (int)sz.size();
  return agg;
}

// The neighbour graph of the aggregates: each aggregate's distinct
// neighbouring aggregates (1-based) in the order the node loop (node
// ascending, slot ascending) meets them, padded with -1 to the widest.
std::vector<int32_t> aggregate_graph(int n, int ns, const std::vector<int32_t> &nbr,
                                     const std::vector<int32_t> &agg, int naggs, int *ns_out) {
  std::vector<std::vector<int32_t>> lists(naggs);
  std::vector<int32_t> counted(naggs, 0);
  int width = 0;
  for (int i = 0; i < n; i++) {
    const int a = agg[i];
    for (int s = 0; s < ns; s++) {
      const int32_t v = nbr[(size_t)i * ns + s];
      if (!(v > 0 && v <= n)) continue;
      const int t = agg[v - 1];
      if (t == a) continue;
      counted[a]++;
      if (std::find(lists[a].begin(), lists[a].end(), t + 1) == lists[a].end())
        lists[a].push_back(t + 1);
    }
    width = std::max(width, counted[a]);
  }
  width = std::max(width, 1);
  std::vector<int32_t> g((size_t)naggs * width, -1);
  for (int a = 0; a < naggs; a++)
    std::copy(lists[a].begin(), lists[a].end(), g.begin() + (size_t)a * width);
  *ns_out = width;
  return g;
}

// members of each group in ascending order: CSR (ptr, idx) of `of`
void group_csr(const std::vector<int32_t> &of, int64_t ngroups, std::vector<int32_t> &ptr,
               std::vector<int32_t> &idx) {
  ptr.assign(ngroups + 1, 0);
  for (int32_t g : of) ptr[g + 1]++;
  for (int64_t g = 0; g < ngroups; g++) ptr[g + 1] += ptr[g];
  idx.resize(of.size());
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (size_t i = 0; i < of.size(); i++) idx[fill[of[i]]++] = (int32_t)i;
}

// ---- device kernels ------------------------------------------------------------

__global__ void gather_kernel(double *__restrict__ out, const double *__restrict__ in,
                              const int32_t *__restrict__ idx, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = in[idx[i]];
}

// out[g] = sum of src[idx[k]], k in [ptr[g], ptr[g+1]): one wave per group,
// lane l sums k = ptr[g] + l, + 64, ..., then a fixed shuffle tree
__global__ void __launch_bounds__(256)
    group_sum_kernel(double *__restrict__ out, const double *__restrict__ src,
                     const int32_t *__restrict__ ptr, const int32_t *__restrict__ idx,
                     int64_t ngroups) {
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  if (g < ngroups)
    for (int32_t k = ptr[g] + lane; k < ptr[g + 1]; k += 64) acc += src[idx[k]];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (g < ngroups && lane == 0) out[g] = acc;
}

// Chebyshev: d = r / theta, x += d
__global__ void cheby_first_kernel(double *__restrict__ x, double *__restrict__ d,
                                   const double *__restrict__ r, double inv_theta, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double v = inv_theta * r[i];
    d[i] = v;
    x[i] = x[i] + v;
  }
}

// Chebyshev: r -= w, d = c1 d + c2 r, x += d
__global__ void cheby_next_kernel(double *__restrict__ x, double *__restrict__ d,
                                  double *__restrict__ r, const double *__restrict__ w,
                                  double c1, double c2, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ri = r[i] - w[i];
    r[i] = ri;
    const double di = c1 * d[i] + c2 * ri;
    d[i] = di;
    x[i] = x[i] + di;
  }
}

unsigned nblocks(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

// out[g] = sum over group g (the last level's single group by the fixed
// grid reduction)
int group_sum(Tamg *h, double *out, const double *src, const int32_t *ptr, const int32_t *idx,
              int64_t ngroups, int64_t nsrc, hipStream_t s) {
  if (ngroups == 1) {
    int st = launch_glsc3_partial(src, h->ones, h->ones, nsrc, h->part, s);
    if (!st) st = launch_combine(h->part, 1, out, s);
    return st;
  }
  const int64_t threads = ngroups * 64;
  hipLaunchKernelGGL(group_sum_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                     s, out, src, ptr, idx, ngroups);
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// the level-l operator (tamg_matvec_flat_impl)
int matvec(Tamg *h, double *out, const double *in, int l, hipStream_t s) {
  const Op &op = *h->op;
  if (l == 0) return op.apply<double>(out, in, nullptr, s, true, true);
  const int64_t n0 = h->n0;
  hipLaunchKernelGGL(gather_kernel, dim3(nblocks(n0)), dim3(256), 0, s, h->wrk_in, in,
                     h->map[l], n0);
  int st = op.gs<double>(h->wrk_in, nullptr, s);
  if (!st) st = launch_vec<double>(V_COL2, h->wrk_in, op.f64.mult, nullptr, nullptr, 0.0, n0, s);
  if (!st) st = launch_mask<double>(h->wrk_in, op.msk, op.nmsk, nullptr, s);
  if (!st) st = op.apply<double>(h->wrk_out, h->wrk_in, nullptr, s, true, true);
  if (!st)
    st = launch_vec<double>(V_COL2, h->wrk_out, op.f64.mult, nullptr, nullptr, 0.0, n0, s);
  if (!st) st = group_sum(h, out, h->wrk_out, h->agg_ptr[l], h->agg_dof[l], h->nn[l], n0, s);
  return st;
}

// amg_cheby_solve: cheby_degree steps on level l from x (zero_init: x = 0)
int cheby(Tamg *h, double *x, const double *f, int l, bool zero_init, hipStream_t s) {
  const int64_t n = h->nn[l];
  double *w = h->cw[l], *r = h->cr[l], *d = h->cd[l];
  NM_HIP_TRY(hipMemcpyAsync(r, f, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  int st;
  if (!zero_init) {
    if ((st = matvec(h, w, x, l, s))) return st;
    if ((st = launch_vec<double>(V_SUB2, r, w, nullptr, nullptr, 0.0, n, s))) return st;
  }
  const double thet = h->tha[l], delt = h->dlt[l];
  const double s1 = thet / delt;
  double rhok = 1.0 / s1;
  hipLaunchKernelGGL(cheby_first_kernel, dim3(nblocks(n)), dim3(256), 0, s, x, d, r,
                     1.0 / thet, n);
  for (int it = 2; it <= h->cheby_degree; it++) {
    if ((st = matvec(h, w, d, l, s))) return st;
    const double rhokp1 = 1.0 / (2.0 * s1 - rhok);
    const double c1 = rhokp1 * rhok, c2 = 2.0 * rhokp1 / delt;
    rhok = rhokp1;
    hipLaunchKernelGGL(cheby_next_kernel, dim3(nblocks(n)), dim3(256), 0, s, x, d, r, w, c1,
                       c2, n);
  }
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

// host value of sum a_i b_i c_i (c = ones: the plain dot)
int dot(Tamg *h, const double *a, const double *b, const double *c, int64_t n, double *v,
        hipStream_t s) {
  int st = launch_glsc3_partial(a, b, c, n, h->part, s);
  if (!st) st = launch_combine(h->part, 1, h->out, s);
  if (st) return st;
  NM_HIP_TRY(hipMemcpyAsync(h->h_out, h->out, sizeof(double), hipMemcpyDeviceToHost, s));
  NM_HIP_TRY(hipStreamSynchronize(s));
  *v = h->h_out[0];
  return NM_OK;
}

// amg_cheby_power on level l
int power(Tamg *h, int l, int its, hipStream_t s) {
  const Op &op = *h->op;
  const int64_t n = h->nn[l];
  double *d = h->cd[l], *w = h->cw[l];
  std::vector<double> d0(n);
  for (int64_t i = 0; i < n; i++) d0[i] = nm_rng_uniform(kSeed, (uint64_t)i) + 10.0;
  NM_HIP_TRY(hipMemcpyAsync(d, d0.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  int st;
  if (l == 0) {
    if ((st = op.gs<double>(d, nullptr, s))) return st;
    if ((st = launch_mask<double>(d, op.msk, op.nmsk, nullptr, s))) return st;
  }
  const double *wt = l == 0 ? op.f64.mult : h->ones;
  for (int it = 0; it < its; it++) {
    if ((st = matvec(h, w, d, l, s))) return st;
    double wtw;
    if ((st = dot(h, w, wt, w, n, &wtw, s))) return st;
    if ((st = launch_vec<double>(V_CMULT2, d, w, nullptr, nullptr, 1.0 / std::sqrt(wtw), n, s)))
      return st;
  }
  if ((st = matvec(h, w, d, l, s))) return st;
  double dtw, dtd;
  if ((st = dot(h, d, wt, w, n, &dtw, s)) || (st = dot(h, d, wt, d, n, &dtd, s))) return st;
  const double lam = dtw / dtd, bb = lam * 1.1, a = lam / 30.0;
  h->tha[l] = (bb + a) / 2.0;
  h->dlt[l] = (bb - a) / 2.0;
  return NM_OK;
}

}  // namespace

int tamg_create(Tamg **out, const Op *op, const std::vector<int32_t> &fneigh, int nlvls,
                int max_iter, int cheby_degree, int power_its, hipStream_t s) {
  *out = nullptr;
  if (nlvls < 2 || nlvls > NM_TAMG_MAXL || max_iter < 1 || cheby_degree < 1) {
    set_error("tamg: levels %d (2..%d), iterations %d, degree %d", nlvls, NM_TAMG_MAXL,
              max_iter, cheby_degree);
    return NM_ERR_ARG;
  }
  const int64_t nel = op->nel, n0 = op->n, n3 = (int64_t)op->lx * op->lx * op->lx;
  if ((int64_t)fneigh.size() != nel * 6 || n0 >= INT32_MAX) {
    set_error("tamg: %lld facet neighbours for %lld elements", (long long)fneigh.size(),
              (long long)nel);
    return NM_ERR_ARG;
  }
  // owned until the end: every early return (NM_HIP_TRY included) frees it
  std::unique_ptr<Tamg> hold(new Tamg);
  Tamg *h = hold.get();
  auto fail = [](int c) { return c; };
  h->op = op;
  h->n0 = n0;
  h->max_iter = max_iter;
  h->cheby_degree = cheby_degree;
  // parents per level
  std::vector<std::vector<int32_t>> parent(NM_TAMG_MAXL + 1);
  h->nn[0] = n0;
  h->nn[1] = nel;
  parent[1].resize(n0);
  for (int64_t i = 0; i < n0; i++) parent[1][i] = (int32_t)(i / n3);
  std::vector<int32_t> graph(fneigh);
  int ns = 6, levels = nlvls;
  for (int l = 2; l <= nlvls - 1; l++) {
    const int nprev = (int)h->nn[l - 1];
    if (nprev / 8 < 4) {  // too few nodes for another level
      levels = l;
      break;
    }
    int naggs = 0;
    parent[l] = greedy_aggregate(nprev, ns, graph, &naggs);
    int ns2 = 0;
    std::vector<int32_t> g2 = aggregate_graph(nprev, ns, graph, parent[l], naggs, &ns2);
    graph.swap(g2);
    ns = ns2;
    h->nn[l] = naggs;
  }
  h->nlvls = levels;
  h->nn[levels] = 1;
  parent[levels].assign(h->nn[levels - 1], 0);
  // device tables
  std::vector<int32_t> ptr, idx;
  for (int l = 1; l <= levels; l++) {
    std::vector<int32_t> &m = h->h_map[l];
    m.resize(n0);
    for (int64_t i = 0; i < n0; i++) m[i] = l == 1 ? parent[1][i] : parent[l][h->h_map[l - 1][i]];
    if (!(h->map[l] = h->alloc<int32_t>(n0)) || !(h->parent[l] = h->alloc<int32_t>(h->nn[l - 1])))
      return fail(NM_ERR_NOMEM);
    NM_HIP_TRY(hipMemcpyAsync(h->map[l], m.data(), sizeof(int32_t) * n0, hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipMemcpyAsync(h->parent[l], parent[l].data(), sizeof(int32_t) * h->nn[l - 1],
                              hipMemcpyHostToDevice, s));
    group_csr(m, h->nn[l], ptr, idx);
    if (!(h->agg_ptr[l] = h->alloc<int32_t>(ptr.size())) ||
        !(h->agg_dof[l] = h->alloc<int32_t>(idx.size())))
      return fail(NM_ERR_NOMEM);
    NM_HIP_TRY(hipMemcpyAsync(h->agg_ptr[l], ptr.data(), sizeof(int32_t) * ptr.size(),
                              hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipMemcpyAsync(h->agg_dof[l], idx.data(), sizeof(int32_t) * idx.size(),
                              hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipStreamSynchronize(s));
    group_csr(parent[l], h->nn[l], ptr, idx);
    if (!(h->node_ptr[l] = h->alloc<int32_t>(ptr.size())) ||
        !(h->node_dof[l] = h->alloc<int32_t>(idx.size())))
      return fail(NM_ERR_NOMEM);
    NM_HIP_TRY(hipMemcpyAsync(h->node_ptr[l], ptr.data(), sizeof(int32_t) * ptr.size(),
                              hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipMemcpyAsync(h->node_dof[l], idx.data(), sizeof(int32_t) * idx.size(),
                              hipMemcpyHostToDevice, s));
    NM_HIP_TRY(hipStreamSynchronize(s));
  }
  if (!(h->wrk_in = h->alloc<double>(n0)) || !(h->wrk_out = h->alloc<double>(n0)) ||
      !(h->ones = h->alloc<double>(n0)) || !(h->part = h->alloc<double>(RED_BLOCKS)) ||
      !(h->out = h->alloc<double>(1)))
    return fail(NM_ERR_NOMEM);
  if (hipHostMalloc(&h->h_out, sizeof(double), hipHostMallocDefault) != hipSuccess) {
    h->h_out = nullptr;
    return fail(NM_ERR_NOMEM);
  }
  int st = launch_vec<double>(V_CFILL, h->ones, nullptr, nullptr, nullptr, 1.0, n0, s);
  if (st) return fail(st);
  for (int l = 0; l < levels; l++) {
    const int64_t n = h->nn[l];
    if (!(h->x[l] = h->alloc<double>(n)) || !(h->b[l] = h->alloc<double>(n)) ||
        !(h->r[l] = h->alloc<double>(n)) || !(h->cd[l] = h->alloc<double>(n)) ||
        !(h->cw[l] = h->alloc<double>(n)) || !(h->cr[l] = h->alloc<double>(n)))
      return fail(NM_ERR_NOMEM);
  }
  for (int l = 0; l < levels; l++)
    if ((st = power(h, l, power_its, s))) return fail(st);
  *out = hold.release();
  return NM_OK;
}

int tamg_solve(Tamg *h, double *z, const double *rin, hipStream_t s) {
  const Op &op = *h->op;
  const int64_t n0 = h->n0;
  const int top = h->nlvls - 1;
  int st = launch_vec<double>(V_RZERO, h->x[0], nullptr, nullptr, nullptr, 0.0, n0, s);
  if (st) return st;
  NM_HIP_TRY(hipMemcpyAsync(h->b[0], rin, sizeof(double) * n0, hipMemcpyDeviceToDevice, s));
  bool zero = true;
  for (int it = 0; it < h->max_iter; it++) {
    // tamg_mg_cycle: down
    for (int l = 0; l < top; l++) {
      const int64_t n = h->nn[l];
      if ((st = cheby(h, h->x[l], h->b[l], l, zero, s))) return st;
      if ((st = matvec(h, h->r[l], h->x[l], l, s))) return st;
      // r = -r + b, then the restriction (mult first from the finest level)
      if ((st = launch_vec<double>(V_ADD2S1, h->r[l], h->b[l], nullptr, nullptr, -1.0, n, s)))
        return st;
      if (l == 0 &&
          (st = launch_vec<double>(V_COL2, h->r[0], op.f64.mult, nullptr, nullptr, 0.0, n0, s)))
        return st;
      if ((st = group_sum(h, h->b[l + 1], h->r[l], h->node_ptr[l + 1], h->node_dof[l + 1],
                          h->nn[l + 1], n, s)))
        return st;
      if ((st = launch_vec<double>(V_RZERO, h->x[l + 1], nullptr, nullptr, nullptr, 0.0,
                                   h->nn[l + 1], s)))
        return st;
      zero = true;
    }
    if ((st = cheby(h, h->x[top], h->b[top], top, zero, s))) return st;
    zero = false;
    // up: prolongation (piecewise constant; averaged into the continuous
    // space at the finest level), correction, smoothing
    for (int l = top - 1; l >= 0; l--) {
      const int64_t n = h->nn[l];
      hipLaunchKernelGGL(gather_kernel, dim3(nblocks(n)), dim3(256), 0, s, h->r[l], h->x[l + 1],
                         h->parent[l + 1], n);
      if (l == 0) {
        if ((st = op.gs<double>(h->r[0], nullptr, s))) return st;
        if ((st = launch_vec<double>(V_COL2, h->r[0], op.f64.mult, nullptr, nullptr, 0.0, n0,
                                     s)))
          return st;
        if ((st = launch_mask<double>(h->r[0], op.msk, op.nmsk, nullptr, s))) return st;
      }
      if ((st = launch_vec<double>(V_ADD2, h->x[l], h->r[l], nullptr, nullptr, 0.0, n, s)))
        return st;
      if ((st = cheby(h, h->x[l], h->b[l], l, false, s))) return st;
    }
  }
  NM_HIP_TRY(hipMemcpyAsync(z, h->x[0], sizeof(double) * n0, hipMemcpyDeviceToDevice, s));
  NM_HIP_TRY(hipGetLastError());
  return NM_OK;
}

}  // namespace nm
