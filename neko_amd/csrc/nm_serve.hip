// Serving loop for streams of Poisson solves.
//
// Requests name a registered (device-resident) box mesh, a right-hand-side
// seed and a relative tolerance, and carry an arrival time.  The loop
// admits requests in arrival order once they have arrived, while fewer
// than `max_inflight` are running and the HBM budget covers their Krylov
// state; geometry, gs maps and the CG state of finished requests stay
// resident and are reused.  Every in-flight request advances on its own
// stream in chunks of graph-captured CG iterations; the host polls events
// and never blocks on one request, so iterations of different requests
// interleave on the device.  Latency = completion - arrival.
//
// This has no counterpart in Neko (which solves one system at a time from
// its time-stepper, fluid_pnpn.f90:849,899); it drives the same solver as
// the single-solve entry point (nm_cg_solve) and reproduces its results
// bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "nm_cg.h"
#include "nm_common.h"
#include "nm_op.h"

namespace nm {

struct ServeMesh {
  nm_op_t *op = nullptr;
  std::vector<CgSolver *> idle;
  int64_t solver_bytes = 0;  // estimate per solver (refined after creation)
};

struct ServeReq {
  int mesh = 0;
  uint64_t seed = 0;
  double rel_tol = 1e-8;
  int max_iter = 10000;
  double t_arrive = 0.0, t_admit = -1.0, t_done = -1.0;
  // 0 queued, 1 running, 2 done, 3 failed; stored last (release) by the
  // serving thread, so a reader that sees 2 sees every result field final
  std::atomic<int> state{0};
  CgSolver *cg = nullptr;
  double *x_keep = nullptr;
  int iter = 0, outer = 0, converged = 0;
  double res_start = 0.0, res_final = 0.0;
};

struct Server {
  int64_t budget = 0, used = 0, peak = 0;
  int precision = 1, max_inflight = 8, chunk_max = 2, keep = 0;
  std::vector<ServeMesh> meshes;
  // requests: records never move once submitted; the vector of them is
  // grown by nm_server_submit, possibly from another thread while a live
  // stream runs, so it is touched only under `mu`
  std::vector<std::unique_ptr<ServeReq>> reqs;
  std::mutex mu;
  std::vector<int> pending;  // submitted, not yet in `order`
  bool live = false, closed = false;
  double t0 = 0.0;           // stream clock origin (steady clock, seconds)
  std::vector<int> order, inflight;
  size_t next = 0;
  int completed = 0, failed = 0;
  double t_end = 0.0;
  ServeReq *req(int id) {
    std::lock_guard<std::mutex> g(mu);
    return reqs[id].get();
  }
  ~Server() {
    for (auto &qp : reqs) {
      if (qp->cg) cg_solver_destroy(qp->cg);
      if (qp->x_keep) (void)hipFree(qp->x_keep);
    }
    for (ServeMesh &m : meshes) {
      for (CgSolver *c : m.idle) cg_solver_destroy(c);
      if (m.op) nm_op_free(m.op);
    }
  }
};

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static int64_t solver_estimate(const Op &op, int precision) {
  const int64_t n = op.n;
  int64_t b = n * 8 * 4 + (int64_t)2 * RED_BLOCKS * 8 + 4096;
  b += precision ? n * 4 * 4 : n * 8;
  if (precision && !op.has32) b += n * 4 * 10;  // FP32 copies of the fields
  return b;
}

// Release idle solvers of other meshes until `need` bytes fit.
static void evict_for(Server *sv, int64_t need, int keep_mesh) {
  for (size_t m = 0; m < sv->meshes.size() && sv->used + need > sv->budget; m++) {
    if ((int)m == keep_mesh) continue;
    ServeMesh &M = sv->meshes[m];
    while (!M.idle.empty() && sv->used + need > sv->budget) {
      CgSolver *c = M.idle.back();
      M.idle.pop_back();
      sv->used -= c->bytes;
      cg_solver_destroy(c);
    }
  }
}

static int get_solver(Server *sv, int mesh, CgSolver **out, bool *deferred) {
  ServeMesh &M = sv->meshes[mesh];
  *deferred = false;
  if (!M.idle.empty()) {
    *out = M.idle.back();
    M.idle.pop_back();
    return NM_OK;
  }
  Op &op = M.op->op;
  const int64_t need = solver_estimate(op, sv->precision);
  if (sv->used + need > sv->budget) evict_for(sv, need, mesh);
  if (sv->used + need > sv->budget) {
    *deferred = true;
    return NM_OK;
  }
  const int64_t op_before = op.bytes;
  int st = cg_solver_create(out, &op, sv->precision);
  if (st) return st;
  sv->used += (*out)->bytes + (op.bytes - op_before);
  sv->peak = std::max(sv->peak, sv->used);
  return NM_OK;
}

static int admit(Server *sv, ServeReq &q, double t, bool *deferred) {
  CgSolver *cg = nullptr;
  int st = get_solver(sv, q.mesh, &cg, deferred);
  if (st || *deferred) return st;
  cg->chunk_max = sv->chunk_max;
  st = nm_op_rhs_normal(sv->meshes[q.mesh].op, q.seed, cg->f, cg->s);
  if (!st) st = cg_begin(cg, nullptr, 0.0, q.rel_tol, q.max_iter, 0.0);
  if (st) {
    sv->meshes[q.mesh].idle.push_back(cg);
    return st;
  }
  q.cg = cg;
  q.t_admit = t;
  q.state.store(1, std::memory_order_release);
  return NM_OK;
}

static int finish(Server *sv, ServeReq &q, double t) {
  CgSolver *cg = q.cg;
  q.iter = cg->total;
  q.outer = cg->sweeps;
  q.converged = cg->converged;
  q.res_start = cg->res_start;
  q.res_final = cg->rnorm;
  q.t_done = t;
  if (sv->keep) {
    const int64_t nb = sizeof(double) * cg->op->n;
    if (hipMalloc((void **)&q.x_keep, nb) != hipSuccess) {
      set_error("serve: out of device memory for a kept solution");
      return NM_ERR_NOMEM;
    }
    sv->used += nb;
    sv->peak = std::max(sv->peak, sv->used);
    NM_HIP_TRY(hipMemcpyAsync(q.x_keep, cg->x, nb, hipMemcpyDeviceToDevice, cg->s));
    NM_HIP_TRY(hipStreamSynchronize(cg->s));
  }
  cg->phase = CgSolver::IDLE;
  sv->meshes[q.mesh].idle.push_back(cg);
  q.cg = nullptr;
  q.state.store(2, std::memory_order_release);
  return NM_OK;
}

}  // namespace nm

using namespace nm;

struct nm_server_s {
  Server sv;
};

extern "C" {

int nm_server_create(nm_server_t **out, double hbm_budget_bytes, int max_inflight,
                     int precision, int chunk_iters, int keep_solutions) {
  *out = nullptr;
  if (max_inflight < 1 || hbm_budget_bytes <= 0.0 || (precision != 0 && precision != 1)) {
    set_error("nm_server_create: bad arguments");
    return NM_ERR_ARG;
  }
  nm_server_t *h = new nm_server_t;
  Server &sv = h->sv;
  sv.budget = (int64_t)hbm_budget_bytes;
  sv.max_inflight = max_inflight;
  sv.precision = precision;
  sv.chunk_max = std::max(1, chunk_iters / 4);
  sv.keep = keep_solutions;
  *out = h;
  return NM_OK;
}

void nm_server_free(nm_server_t *h) { delete h; }

// Make an operator resident: autotune it (kernel variants are fixed
// before any CG state captures its graphs), charge it to the budget and
// pre-create `prewarm` CG states.  On failure the operator is freed.
static int register_mesh(nm_server_t *h, nm_op_t *op, int tune_reps, int prewarm,
                         int *mesh_id, hipStream_t s) {
  Server &sv = h->sv;
  {
    std::lock_guard<std::mutex> g(sv.mu);
    if (sv.live) {
      nm_op_free(op);
      set_error("server: meshes are registered before a live stream starts");
      return NM_ERR_STATE;
    }
  }
  ServeMesh m;
  m.op = op;
  int st;
  if (tune_reps > 0 && (st = nm_op_autotune(m.op, tune_reps, nullptr, nullptr, nullptr, s))) {
    nm_op_free(m.op);
    return st;
  }
  if (sv.used + m.op->op.bytes > sv.budget) {
    nm_op_free(m.op);
    set_error("server: mesh does not fit the HBM budget");
    return NM_ERR_NOMEM;
  }
  sv.used += m.op->op.bytes;
  sv.peak = std::max(sv.peak, sv.used);
  sv.meshes.push_back(m);
  *mesh_id = (int)sv.meshes.size() - 1;
  for (int k = 0; k < prewarm; k++) {
    CgSolver *cg = nullptr;
    bool deferred = false;
    if ((st = get_solver(&sv, *mesh_id, &cg, &deferred))) return st;
    if (deferred) break;
    sv.meshes[*mesh_id].idle.push_back(cg);
  }
  return NM_OK;
}

int nm_server_add_box_mesh(nm_server_t *h, int nelx, int nely, int nelz, int lx,
                           double amp, uint64_t seed, int tune_reps, int prewarm,
                           int *mesh_id, hipStream_t s) {
  nm_op_t *op = nullptr;
  int st = nm_op_create_box(&op, nelx, nely, nelz, lx, 0, 0, 0, amp, seed, 1.0, 0.0, s);
  if (st) return st;
  return register_mesh(h, op, tune_reps, prewarm, mesh_id, s);
}

// A general straight-edged mesh (e.g. from nm_nmsh_read), Poisson operator
// with Dirichlet on every facet without a neighbour.
int nm_server_add_mesh(nm_server_t *h, int64_t nel, int lx, const double *vxyz,
                       const int64_t *pid, int tune_reps, int prewarm, int *mesh_id,
                       hipStream_t s) {
  nm_op_t *op = nullptr;
  int st = nm_op_create_mesh(&op, nel, lx, vxyz, pid, 1.0, 0.0, s);
  if (st) return st;
  return register_mesh(h, op, tune_reps, prewarm, mesh_id, s);
}

int nm_server_mesh_op(nm_server_t *h, int mesh_id, nm_op_t **op) {
  if (mesh_id < 0 || mesh_id >= (int)h->sv.meshes.size()) {
    set_error("nm_server_mesh_op: no mesh %d", mesh_id);
    return NM_ERR_ARG;
  }
  *op = h->sv.meshes[mesh_id].op;
  return NM_OK;
}

int nm_server_submit(nm_server_t *h, int mesh_id, uint64_t rhs_seed, double rel_tol,
                     int max_iter, double t_arrive, int *req_id) {
  Server &sv = h->sv;
  if (mesh_id < 0 || mesh_id >= (int)sv.meshes.size() || rel_tol <= 0.0 || max_iter < 1) {
    set_error("nm_server_submit: bad arguments");
    return NM_ERR_ARG;
  }
  auto q = std::make_unique<ServeReq>();
  q->mesh = mesh_id;
  q->seed = rhs_seed;
  q->rel_tol = rel_tol;
  q->max_iter = max_iter;
  std::lock_guard<std::mutex> g(sv.mu);
  if (sv.live && sv.closed) {
    set_error("nm_server_submit: the live stream is closed");
    return NM_ERR_STATE;
  }
  // negative: arrives now (live streams), on the stream clock
  q->t_arrive = t_arrive >= 0.0 ? t_arrive : (sv.live ? now_s() - sv.t0 : 0.0);
  sv.reqs.push_back(std::move(q));
  *req_id = (int)sv.reqs.size() - 1;
  sv.pending.push_back(*req_id);
  return NM_OK;
}

// One pass of the request loop: admission in arrival order, then polling
// of every in-flight request.  Returns the number of requests finished in
// the pass (>= 0) or an error status (< 0).
static int serve_pass(Server &sv, double t0, bool *progressed) {
  const double t = now_s() - t0;
  *progressed = false;
  int done = 0;
  while (sv.next < sv.order.size() && (int)sv.inflight.size() < sv.max_inflight) {
    ServeReq &q = *sv.req(sv.order[sv.next]);
    if (q.t_arrive > t) break;
    bool deferred = false;
    int st = admit(&sv, q, t, &deferred);
    if (st) return -st;
    if (deferred) {
      if (sv.inflight.empty()) {
        set_error("nm_server: a request does not fit the HBM budget");
        return -NM_ERR_NOMEM;
      }
      break;
    }
    sv.inflight.push_back(sv.order[sv.next]);
    sv.next++;
    *progressed = true;
  }
  for (size_t k = 0; k < sv.inflight.size();) {
    ServeReq &q = *sv.req(sv.inflight[k]);
    const int st = cg_poll(q.cg);
    if (st < 0) return st;
    if (st == 1) {
      const int fs = finish(&sv, q, now_s() - t0);
      if (fs) return -fs;
      sv.inflight.erase(sv.inflight.begin() + (long)k);
      done++;
      *progressed = true;
    } else {
      k++;
    }
  }
  return done;
}

// move newly submitted requests into the admission order (the not yet
// admitted tail stays sorted by arrival time)
static void take_pending(Server &sv) {
  std::lock_guard<std::mutex> g(sv.mu);
  if (sv.pending.empty()) return;
  for (int id : sv.pending) sv.order.push_back(id);
  sv.pending.clear();
  std::stable_sort(sv.order.begin() + (long)sv.next, sv.order.end(), [&](int a, int b) {
    return sv.reqs[a]->t_arrive < sv.reqs[b]->t_arrive;
  });
}

// Run the stream: arrival times are seconds after this call starts.
// Returns when every submitted request has completed (or failed).
int nm_server_run(nm_server_t *h, double *wall_s) {
  Server &sv = h->sv;
  {
    std::lock_guard<std::mutex> g(sv.mu);
    if (sv.live) {
      set_error("nm_server_run: a live stream is running");
      return NM_ERR_STATE;
    }
  }
  sv.order.clear();
  sv.next = 0;
  take_pending(sv);
  const int todo = (int)sv.order.size();
  int done = 0;
  const double t0 = now_s();
  while (done < todo) {
    bool progressed = false;
    const int st = serve_pass(sv, t0, &progressed);
    if (st < 0) return -st;
    done += st;
    if (!progressed) std::this_thread::yield();
  }
  sv.t_end = now_s() - t0;
  if (wall_s) *wall_s = sv.t_end;
  return NM_OK;
}

// Live stream: nm_server_submit may be called from other threads while
// this runs (t_arrive < 0 stamps the submission time on the stream
// clock, which starts here); returns once nm_server_close has been called
// and every request submitted before it has completed.
int nm_server_serve(nm_server_t *h, double *wall_s) {
  Server &sv = h->sv;
  {
    std::lock_guard<std::mutex> g(sv.mu);
    if (sv.live) {
      set_error("nm_server_serve: a live stream is already running");
      return NM_ERR_STATE;
    }
    sv.live = true;  // a close() made before this call ends this stream
    sv.t0 = now_s();
  }
  sv.order.clear();
  sv.next = 0;
  int rc = NM_OK;
  for (;;) {
    bool closed;
    {
      std::lock_guard<std::mutex> g(sv.mu);
      closed = sv.closed;
    }
    take_pending(sv);
    bool progressed = false;
    const int st = serve_pass(sv, sv.t0, &progressed);
    if (st < 0) {
      rc = -st;
      break;
    }
    if (closed && sv.next == sv.order.size() && sv.inflight.empty()) {
      std::lock_guard<std::mutex> g(sv.mu);
      if (sv.pending.empty()) break;
      continue;
    }
    if (!progressed) std::this_thread::yield();
  }
  sv.t_end = now_s() - sv.t0;
  if (wall_s) *wall_s = sv.t_end;
  std::lock_guard<std::mutex> g(sv.mu);
  sv.live = false;
  sv.closed = false;
  return rc;
}

// End a live stream: no further submissions; nm_server_serve returns once
// the submitted requests have completed.
int nm_server_close(nm_server_t *h) {
  std::lock_guard<std::mutex> g(h->sv.mu);
  h->sv.closed = true;
  return NM_OK;
}

// Seconds on the stream clock of the running live stream (0 when none).
double nm_server_clock(nm_server_t *h) {
  std::lock_guard<std::mutex> g(h->sv.mu);
  return h->sv.live ? now_s() - h->sv.t0 : 0.0;
}

int nm_server_result(const nm_server_t *h, int req_id, nm_serve_result_t *r) {
  Server &sv = const_cast<nm_server_t *>(h)->sv;
  std::lock_guard<std::mutex> g(sv.mu);
  if (req_id < 0 || req_id >= (int)sv.reqs.size()) {
    set_error("nm_server_result: no request %d", req_id);
    return NM_ERR_ARG;
  }
  const ServeReq &q = *sv.reqs[req_id];
  r->state = q.state.load(std::memory_order_acquire);
  r->iter = q.iter;
  r->outer = q.outer;
  r->converged = q.converged;
  r->res_start = q.res_start;
  r->res_final = q.res_final;
  r->t_arrive = q.t_arrive;
  r->t_admit = q.t_admit;
  r->t_done = q.t_done;
  r->n = sv.meshes[q.mesh].op->op.n;
  return NM_OK;
}

int nm_server_solution(const nm_server_t *h, int req_id, double *x, hipStream_t s) {
  Server &sv = const_cast<nm_server_t *>(h)->sv;
  ServeReq *qp = nullptr;
  {
    std::lock_guard<std::mutex> g(sv.mu);
    if (req_id >= 0 && req_id < (int)sv.reqs.size()) qp = sv.reqs[req_id].get();
  }
  if (!qp || !qp->x_keep) {
    set_error("nm_server_solution: no kept solution for request %d", req_id);
    return NM_ERR_ARG;
  }
  const ServeReq &q = *qp;
  NM_HIP_TRY(hipMemcpyAsync(x, q.x_keep, sizeof(double) * sv.meshes[q.mesh].op->op.n,
                            hipMemcpyDeviceToDevice, s));
  return NM_OK;
}

int nm_server_memory(const nm_server_t *h, int64_t *used, int64_t *peak) {
  *used = h->sv.used;
  *peak = h->sv.peak;
  return NM_OK;
}

}  // extern "C"
