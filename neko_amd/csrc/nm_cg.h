// Device Jacobi-PCG as a non-blocking state machine: begin() enqueues the
// start of a solve on the solver's own stream, poll() advances it by one
// chunk of graph-captured iterations whenever the previous chunk has
// finished (never blocks), wait() drives it to the end.  The blocking C-ABI
// solve and the serving loop share it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "nm_op.h"

namespace nm {

// Device-resident scalars of one CG solve.
struct CgScal {
  double rtz, rtz_old, pap, rtr, alpha, beta, rnorm, tol, norm_fac, res_start;
  double scale;       // start: r = scale * f
  double tol_final;   // mixed: final tolerance on the true residual
  double delta;       // mixed: recurrence reduction between replacements
  double true_rnorm;  // mixed: norm of the last true residual
  int iter, done, max_iter, converged, updates;
  int fin;            // mixed: solve finished (set with the true residual)
  // batched mixed solves: a replacement that ran beside the iteration
  // graph leaves done set and raises rearm; the graph's next step clears
  // both, so a step never sees a problem resume half-way
  int rearm;
};

struct CgSolver {
  Op *op = nullptr;
  int mixed = 0;
  hipStream_t s = nullptr;
  // FP64 state
  double *x = nullptr, *r = nullptr, *p = nullptr, *w = nullptr, *f = nullptr;
  // FP32 inner state
  float *e32 = nullptr, *r32 = nullptr, *p32 = nullptr, *w32 = nullptr;
  double *part = nullptr;
  double *dpart = nullptr;   // fused p.Ap partials (trilinear Ax), npart
  int64_t npart = 0;
  bool cgp = false;          // CG prologue fused into Ax (x one step behind)
  CgScal *sc = nullptr;      // device: [0] main/inner, [1] outer
  CgScal *h_sc = nullptr;    // pinned host mirror
  hipGraphExec_t g64 = nullptr, g32 = nullptr;
  uint64_t epoch = 0;        // op->epoch the graphs were captured at
  // the solver's own zero-Dirichlet mask (nm_cg_set_mask), in place of the
  // operator's
  bool own_mask = false;
  MaskRef mk;
  int32_t *mask_dev = nullptr;
  int64_t mask_cap = 0;
  int64_t npart_cap = 0;     // capacity of dpart
  int graph_iters = 4;
  int64_t bytes = 0;         // device memory held by this solver
  std::vector<void *> owned;
  // run state
  enum Phase { IDLE, RUN64, MIX, REPLACE, DONE } phase = IDLE;
  hipEvent_t ev = nullptr;
  bool pending = false;      // an event is recorded and not yet consumed
  double abs_tol = 0.0, rel_tol = 0.0, inner_rel = 0.1, tol = 0.0;
  int max_iter = 0, total = 0, sweeps = 0, launched = 0, chunk = 1;
  int chunk_max = 16;        // graph launches per chunk (doubling up to this)
  double res_start = -1.0, rnorm = 0.0;
  int converged = 0;
  ~CgSolver();
  template <typename T>
  T *alloc(int64_t cnt);
};

// precision 0: FP64; 1: FP32 CG with FP64 residual replacement
int cg_solver_create(CgSolver **out, Op *op, int precision);
void cg_solver_destroy(CgSolver *cg);
// Start a solve of A x = f (f copied; f == nullptr: cg->f already holds
// the right-hand side).  Stops when the Neko residual norm drops below
// rel_tol * (initial norm) if rel_tol > 0, else below abs_tol.
int cg_begin(CgSolver *cg, const double *f, double abs_tol, double rel_tol,
             int max_iter, double inner_rel);
// Non-blocking: 1 when the solve has finished, 0 while running, < 0 error.
int cg_poll(CgSolver *cg);
int cg_wait(CgSolver *cg);

// Mixed-precision building blocks on one problem's slice of (possibly
// shared) vectors: start (x = 0, r32 = f, scalars) and residual
// replacement (x += x32, r32 = f - A x with the FP64 operator).
int mx_slice_begin(const Op &op, double *x, float *x32, float *r32, float *p32,
                   const double *f, CgScal *sc, double *part, double abs_tol, double rel,
                   double delta, int max_iter, hipStream_t s);
// x32 += alpha p32, alpha = 0 (the fused schedule's owed x update)
int mx_slice_flush(float *x32, const float *p32, int64_t n, CgScal *sc, hipStream_t s);
// With rep_out the replacement may run beside the iteration graph: a
// continuing problem keeps done set and raises rearm (with release order
// after its other writes), and the scalars it decided on are copied to
// *rep_out, which nothing else writes.
int mx_slice_replace(const Op &op, double *x, float *x32, float *r32, const double *f,
                     double *w, CgScal *sc, double *part, hipStream_t s,
                     const MaskRef *mk = nullptr, CgScal *rep_out = nullptr);

}  // namespace nm
