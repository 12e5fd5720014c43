"""Benchmark of the Neko linear-solve hot path on one MI355X.

Workload (default): the large Poisson problem -- 64x64x64 = 262,144 hex
elements, lx = 8 (134,217,728 local DOF), interior vertices perturbed by
seeded U(-0.1h, 0.1h), trilinear geometry, homogeneous Dirichlet on all
faces, RHS i.i.d. N(0,1) made consistent by one gs sum.

One "step" is one FP64 Ax+gs apply, w = gs(A_local u), on inputs resident
in HBM.  `value` = GDOF/s of that step (local DOF / s).  The same run also
times Jacobi-PCG solves to 1e-10 relative residual in FP64 and in FP32-inner
/ FP64-refinement mode, and the CPU oracle on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--elems 64]
                    [--no-solve] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def log(msg):
    print(msg, file=sys.stderr, flush=True)


FP64_PEAK_TFS = 78.6   # MI355X FP64 (vector and matrix) peak


def ax_bytes_per_dof(lx, code, h1_field=False, h2=False):
    """Algorithmic HBM bytes per local DOF of one FP64 Ax launch.
    Stored geometry (codes < 200): read u + G11..G23 + write w (+ h1 field,
    + h2*B field).  In-kernel trilinear geometry (codes >= 200): read u +
    write w + 36 edge values per element."""
    if code >= 200:
        return 8 * 2 + 36 * 8 / lx ** 3
    return 8 * (1 + 6 + 1) + (8 if h1_field else 0) + (8 if h2 else 0)


def ax_flops_per_dof(lx, code):
    """FP64 flops per local DOF of one Ax launch (FMA = 2), counted from
    the kernel source: derivatives 6 lx, geometric flux 15, transposes
    6 lx, final sum 2 (113 at lx = 8; Neko counts 15 + 12 lx).  The
    trilinear variant evaluates the geometry per point instead of reading
    it: cofactor rows as three cross products (27), jac (5), 1/jac by the
    reciprocal estimate and two Newton steps (7), the weight w3 h1 / jac
    (2), the flux as sc C^T grad u then C g (33 instead of 15), and the
    per-element edge tables amortised over the element's points (~8):
    67 more."""
    base = 2 * 3 * lx + 15 + 2 * lx + 2 * 2 * lx + 2
    return base + (67 if code >= 200 else 0)


def gs_bytes(info, word=8):
    # per gs entry: read u, write u, read its 0-based local index
    # (+ one 4-byte block offset per block)
    return info.nlocal * (word + word + 4) + info.nblks * 4


def pcg_bytes_per_dof(info, n, word=8):
    """HBM bytes per local DOF of one Jacobi-PCG iteration in the minimal
    fused schedule the roofline is graded against: p update (read r, dinv,
    mult, p; write p) 40, stored-G Ax 64, gs, p.Ap (read w, mult, p) 24,
    x/r update with rtr (read x, p, r, w, mult; write x, r) 56, all in FP64
    words (195.6 B/DOF at 64^3, lx 8); FP32 halves the words, not the gs
    indices."""
    return (40 + 64 + 24 + 56) * word / 8 + gs_bytes(info, word) / n


def replacement_bytes_per_dof(info, n):
    """One FP64 residual replacement of the mixed solve: Ax+gs in FP64 plus
    r = f - A x and x += x32, r32 = (float) r (read f, w, x, x32; write x,
    r32) 40 B/DOF."""
    return 64 + gs_bytes(info) / n + 40


def fused_pcg_bytes_per_dof(info, n, lx, word=8, stored=False, fields=0):
    """Algorithmic HBM bytes per local DOF of one iteration of the schedule
    this library runs: Ax with the CG prologue (read p, r, dinv, x; write p,
    x, w; plus the six geometric factors and `fields` more pointwise fields
    (h1, h2 B) when the geometry is stored, or 36 edge-vector words per
    element when it is evaluated in the kernel); the gs block sums on the
    edge and vertex entries; the update (read r, w, dinv, a one-byte
    multiplicity count; write r; plus the partner value of every face-pair
    entry)."""
    face = 2 * info.npairs
    if stored:
        ax = (7 + 6 + fields) * word
    else:
        ax = 7 * word + 36 * word / lx ** 3
    blocks = ((info.nlocal - face) * (2 * word + 4) + 4 * info.nblks) / n
    update = 4 * word + 1 + word * face / n
    return ax + blocks + update


def pcg_roofline(prec, mon, dt, info, n, lx, stored=False, fields=0):
    """PCG iterations/s and solves/s of a finished solve as fractions of the
    HBM roofline of the bytes of the schedule run here (`schedule_run`;
    the fused face sums need cg_face_sums) and, for the stored-geometry
    operator, of the minimal fused schedule of the reference's kernels
    (`minimal_schedule`, the grading basis: 195.6 B/DOF at 64^3, lx 8).
    The in-kernel-geometry operator streams no geometric factors, so only
    its own bytes are a meaningful basis for it."""
    bw = HBM_PEAK_GBS * 1e9
    word = 8 if prec == "fp64" else 4
    rep = 0.0 if prec == "fp64" else mon.outer * replacement_bytes_per_dof(info, n)
    out = {}
    cases = [("schedule_run", fused_pcg_bytes_per_dof(info, n, lx, word, stored, fields)
              if info.cg_face_sums else None)]
    if stored:
        cases.insert(0, ("minimal_schedule", pcg_bytes_per_dof(info, n, word)))
    for key, b_it in cases:
        if b_it is None:
            continue
        t_min = (mon.iter * b_it + rep) * n / bw
        it_roof = bw / (b_it * n)
        out[key] = {"bytes_per_dof_iteration": round(b_it, 2),
                    "it_per_s_roofline": round(it_roof, 1),
                    "it_frac": round(mon.iter / dt / it_roof, 4),
                    "solves_per_s_roofline": round(1.0 / t_min, 4),
                    "solves_frac": round(t_min / dt, 4)}
    return out


def hbm_stream(n=134217728):
    """STREAM-style copy (16 B/element) and triad (24 B/element) over FP64
    arrays of n elements with torch's own kernels: the practical HBM
    bandwidth of this device beside the 8 TB/s nominal peak."""
    import torch
    a = torch.full((n,), 1.0, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    c = torch.full((n,), 2.0, dtype=torch.float64, device="cuda")

    def t(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    copy = 16 * n / t(lambda: b.copy_(a)) / 1e9
    triad = 24 * n / t(lambda: torch.add(a, c, alpha=3.0, out=b)) / 1e9
    del a, b, c
    torch.cuda.empty_cache()
    return {"copy_GBs": round(copy, 1), "triad_GBs": round(triad, 1),
            "triad_frac_of_peak": round(triad / HBM_PEAK_GBS, 4),
            "array_bytes": 8 * n}


def c1_solve_section(S, reps=20):
    """The small Poisson case on the GPU: 8^3 undeformed box, lx = 8,
    RHS B 3 pi^2 sin(pi x) sin(pi y) sin(pi z), Jacobi-PCG to 1e-10
    relative, FP64 and mixed; mean over `reps` solves from x = 0 after one
    untimed solve, with the max-norm error against the exact solution."""
    import numpy as np
    import torch
    p = S.Problem.box(8, 8, 8, 8)
    f = p.rhs_sine()
    mult = p.field(S.FIELD_MULT)
    res0 = float(torch.sqrt(torch.sum(f * f * mult)) / np.sqrt(p.volume))
    exact = (torch.sin(np.pi * p.field(S.FIELD_X)) * torch.sin(np.pi * p.field(S.FIELD_Y))
             * torch.sin(np.pi * p.field(S.FIELD_Z)))
    out = {"workload": "8^3 undeformed box, lx=8, sine RHS, Jacobi-PCG 1e-10",
           "local_dof": p.n}
    for prec in ("fp64", "mixed"):
        cg = S.CG(p, prec)
        x = torch.zeros_like(f)
        cg.solve(x, f, 1e-10 * res0)
        torch.cuda.synchronize()
        t1 = time.time()
        for _ in range(reps):
            x.zero_()
            mon = cg.solve(x, f, 1e-10 * res0)
        torch.cuda.synchronize()
        dt = (time.time() - t1) / reps
        out[prec] = {"iterations": mon.iter, "outer": mon.outer, "converged": mon.converged,
                     "seconds": round(dt, 5), "solves_per_s": round(1.0 / dt, 2),
                     "max_err_vs_exact": float(torch.max(torch.abs(x - exact)))}
        cg.free()
    p.free()
    log(f"[bench] c1 solve {out}")
    return out


def neko_array_operator(S, p, E, lx):
    """The same 64^3 problem as an operator built from Neko's own arrays
    (nm_op_create: the geometric factors G11..G23, mass B, Jacobi
    deformation flags, the 1-based gs maps and the Dirichlet mask, as a
    ksp_t/pc_t plug-in receives them from coef_t / gs_t / the bc list): the
    path that streams the six geometric factors from HBM."""
    import numpy as np
    sp = S.Space(lx)
    g = [p.field(k).cpu().numpy() for k in range(6)]
    B = p.field(S.FIELD_B).cpu().numpy()
    gs = p.gs_maps()
    msk = S.box_mask(E, E, E, lx)
    dfrmd = np.ones(E ** 3, dtype=np.int32)  # every perturbed element
    q = S.Problem.from_arrays(E ** 3, lx, sp.dx, g, B, None, gs, msk, h1c=1.0,
                              dfrmd=dfrmd)
    del g, B, gs, msk
    return q


def axgs_timing(p, steps, warmup):
    """ms of Ax+gs, Ax alone and gs alone (HIP events on the current stream)."""
    import torch
    stream = torch.cuda.current_stream()
    u = p.rhs_normal(7)
    w = torch.empty_like(u)

    def run(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    out = {"axgs": run(lambda: p.ax_gs(w, u)), "ax": run(lambda: p.ax_local(w, u)),
           "gs": run(lambda: p.gs_op(w))}
    del u, w
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--elems", type=int, default=64, help="elements per side")
    ap.add_argument("--lx", type=int, default=8)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-helmholtz", action="store_true")
    ap.add_argument("--max-iter", type=int, default=20000,
                    help="iteration cap of the solves (profiling runs)")
    ap.add_argument("--pmg-default-cap", type=int, default=300,
                    help="iteration cap of GMRES with the default-h2 level operators")
    ap.add_argument("--batch-size", type=int, default=256)
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rates", default="10,2000",
                    help="arrival rates (req/s); rates below the device's "
                         "capacity measure latency, rates above it throughput")
    ap.add_argument("--no-neko-arrays", action="store_true",
                    help="skip the operator built from Neko's arrays")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the small (8^3) Poisson solve timing")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the "
                          "multi-GPU path (element partitions, RCCL gs halo "
                          "exchange) is designed but not built; run --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from neko_amd import solve as S

    E, lx = args.elems, args.lx
    log(f"[bench] setup {E}^3 elements lx={lx}")
    t0 = time.time()
    p = S.Problem.box(E, E, E, lx, amp=0.1, seed=2024)
    torch.cuda.synchronize()
    t_setup = time.time() - t0
    info = p.info
    n = p.n
    log(f"[bench] setup {t_setup:.1f}s n={n} nlocal={info.nlocal}")

    stream = torch.cuda.current_stream()
    u = p.rhs_normal(7)
    w = torch.empty_like(u)
    tuned = p.autotune(reps=5)
    log(f"[bench] autotuned variants {tuned}")
    # reference copy kernel of known size (8n read + 8n write) for scaling
    # the profiler's FETCH_SIZE / WRITE_SIZE counters
    w.copy_(u)
    torch.cuda.synchronize()

    # ---- Ax+gs steps -------------------------------------------------------
    for _ in range(args.warmup):
        p.ax_gs(w, u)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        p.ax_gs(w, u)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_step = ev0.elapsed_time(ev1) / args.steps
    gdofs = n / (ms_step * 1e-3) / 1e9
    log(f"[bench] Ax+gs {ms_step:.3f} ms/step  {gdofs:.1f} GDOF/s")

    # ---- dominant kernel alone (Ax) for the roofline -----------------------
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        p.ax_local(w, u)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_ax = e0.elapsed_time(e1) / args.steps
    code64 = tuned["ax64"]
    ax_bytes = ax_bytes_per_dof(lx, code64) * n
    ax_gbs = ax_bytes / (ms_ax * 1e-3) / 1e9
    ax_tfs = ax_flops_per_dof(lx, code64) * n / (ms_ax * 1e-3) / 1e12
    # gs alone
    e0.record(stream)
    for _ in range(args.steps):
        p.gs_op(w)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_gs = e0.elapsed_time(e1) / args.steps
    gsb = gs_bytes(info)
    log(f"[bench] Ax {ms_ax:.3f} ms ({ax_gbs:.0f} GB/s, {ax_tfs:.1f} TFLOP/s)  gs {ms_gs:.3f} ms "
        f"({gsb / (ms_gs * 1e-3) / 1e9:.0f} GB/s)")
    axgs_bytes = ax_bytes + gsb

    result = {
        "metric": "GDOF/s Ax+gs & Poisson solves/s, mixed vs FP64 @1e-10 rel.res; p99 latency",
        "value": round(gdofs, 3),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded perturbed box mesh, N(0,1) gs-consistent RHS)",
        "config": {
            "workload": f"Poisson Ax+gs, box {E}x{E}x{E} hex, lx={lx}, "
                        f"+-0.1h perturbed, Dirichlet",
            "elements": E ** 3, "lx": lx, "local_dof": n,
            "gs_entries": info.nlocal, "parallelism": "single GPU",
            "kernel_variants": tuned,
        },
        "axgs": {"ms": round(ms_step, 4), "bytes_per_apply": axgs_bytes,
                 "GBs": round(axgs_bytes / (ms_step * 1e-3) / 1e9, 1),
                 "roofline_frac": round(axgs_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": roofline_entry(code64, lx, ax_gbs, ax_tfs, ax_bytes, ms_ax, n),
        "gs": {"ms": round(ms_gs, 4), "bytes": gsb,
               "GBs": round(gsb / (ms_gs * 1e-3) / 1e9, 1)},
        "setup_s": round(t_setup, 2),
    }

    if not args.no_stream:
        result["hbm_stream"] = hbm_stream()
        log(f"[bench] stream {result['hbm_stream']}")

    # ---- the operator built from Neko's arrays (geometry streamed) -----------
    if not args.no_neko_arrays:
        result["neko_arrays"] = neko_arrays_section(S, p, E, lx, args)

    # ---- solves --------------------------------------------------------------
    if not args.no_solve:
        f = p.rhs_normal(11)
        res_start = None
        solves = {}
        for prec in ("fp64", "mixed"):
            cg = S.CG(p, prec)
            x = torch.zeros_like(f)
            if res_start is None:
                # Neko's norm of the start residual r = f (x0 = 0)
                mult = p.field(S.FIELD_MULT)
                res_start = float(torch.sqrt(torch.sum(f * f * mult)) / np.sqrt(p.volume))
            tol = 1e-10 * res_start
            torch.cuda.synchronize()
            t1 = time.time()
            mon = cg.solve(x, f, tol, max_iter=args.max_iter)
            torch.cuda.synchronize()
            dt = time.time() - t1
            solves[prec] = {"seconds": round(dt, 3), "iterations": mon.iter,
                            "outer": mon.outer, "res_start": mon.res_start,
                            "res_final": mon.res_final, "converged": mon.converged,
                            "solves_per_s": round(1.0 / dt, 4),
                            "it_per_s": round(mon.iter / dt, 2),
                            "roofline": pcg_roofline(prec, mon, dt, info, n, lx,
                                                     stored=tuned["ax64"] < 200)}
            solves[prec]["_x"] = x
            log(f"[bench] {prec}: {mon.iter} it ({mon.outer} outer) {dt:.2f}s "
                f"res {mon.res_start:.3e}->{mon.res_final:.3e}")
            cg.free()
        xd = solves["fp64"].pop("_x")
        xm = solves["mixed"].pop("_x")
        solves["mixed_vs_fp64_rel_l2"] = float(torch.linalg.norm(xm - xd) / torch.linalg.norm(xd))
        solves["speedup_mixed"] = round(solves["fp64"]["seconds"] / solves["mixed"]["seconds"], 3)
        result["solves"] = solves
        if not args.no_c1:
            result["c1_solve"] = c1_solve_section(S)

    # ---- Helmholtz with the Chebyshev/Schwarz p-multigrid ---------------------
    if not args.no_helmholtz:
        result["helmholtz_pmg"] = helmholtz_pmg(S, default_cap=args.pmg_default_cap)

    # ---- batched uneven instances -------------------------------------------
    if not args.no_batch:
        result["batched"] = batched(S, args.batch_size)

    # ---- serving stream ------------------------------------------------------
    if not args.no_serve:
        del w, u
        if not args.no_solve:
            del f, x
        p.free()
        torch.cuda.empty_cache()
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        result["serving"] = serving(S, rates, args.serve_requests)

    # ---- CPU baseline (oracle) on a bounded sample ------------------------
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(lx, args.cpu_seconds)
        except Exception as exc:  # the baseline is reported, never required
            result["cpu_baseline"] = {"error": repr(exc)}

    print(json.dumps(result), flush=True)


def neko_arrays_section(S, p, E, lx, args):
    """Ax+gs and Jacobi-PCG (FP64 and mixed, 1e-10) through an operator
    built from Neko's arrays (nm_op_create), whose Ax streams G11..G23."""
    import numpy as np
    import torch
    t0 = time.time()
    q = neko_array_operator(S, p, E, lx)
    tuned = q.autotune(reps=3)
    info, n = q.info, q.n
    out = {"setup_s": round(time.time() - t0, 2), "kernel_variants": tuned}
    tm = axgs_timing(q, args.steps, args.warmup)
    code = tuned["ax64"]
    axb = ax_bytes_per_dof(lx, code) * n
    gsb = gs_bytes(info)
    out["axgs"] = {
        "ms": round(tm["axgs"], 4), "GDOF_s": round(n / (tm["axgs"] * 1e-3) / 1e9, 3),
        "bytes_per_apply": axb + gsb,
        "bytes_per_dof": round((axb + gsb) / n, 2),
        "roofline_frac": round((axb + gsb) / (tm["axgs"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out["ax"] = {"ms": round(tm["ax"], 4), "kernel":
                 f"ax_kstep_kernel<double,{lx},{code % 100 or 1}>",
                 "GBs": round(axb / (tm["ax"] * 1e-3) / 1e9, 1),
                 "hbm_frac": round(axb / (tm["ax"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out["gs"] = {"ms": round(tm["gs"], 4),
                 "GBs": round(gsb / (tm["gs"] * 1e-3) / 1e9, 1)}
    log(f"[bench] neko arrays: {out}")
    if not args.no_solve:
        f = q.rhs_normal(11)
        mult = q.field(S.FIELD_MULT)
        r0 = float(torch.sqrt(torch.sum(f * f * mult)) / np.sqrt(q.volume))
        xs = {}
        for prec in ("fp64", "mixed"):
            cg = S.CG(q, prec)
            x = torch.zeros_like(f)
            torch.cuda.synchronize()
            t1 = time.time()
            mon = cg.solve(x, f, 1e-10 * r0, max_iter=args.max_iter)
            torch.cuda.synchronize()
            dt = time.time() - t1
            out[prec] = {"seconds": round(dt, 3), "iterations": mon.iter, "outer": mon.outer,
                         "converged": mon.converged, "solves_per_s": round(1.0 / dt, 4),
                         "it_per_s": round(mon.iter / dt, 2),
                         "roofline": pcg_roofline(prec, mon, dt, info, n, lx, stored=True)}
            xs[prec] = x
            log(f"[bench] neko arrays {prec}: {mon.iter} it {dt:.2f}s")
            cg.free()
        out["mixed_vs_fp64_rel_l2"] = float(torch.linalg.norm(xs["mixed"] - xs["fp64"]) /
                                            torch.linalg.norm(xs["fp64"]))
        out["speedup_mixed"] = round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3)
        del xs, f, mult
    q.free()
    torch.cuda.empty_cache()
    return out


def _box_dims(E):
    """Box dimensions (a, b, c) with a*b*c close to E, as cubic as possible."""
    a = max(1, round(E ** (1 / 3)))
    b = max(1, round((E / a) ** 0.5))
    c = max(1, round(E / (a * b)))
    return a, b, c


def serving(S, rates, nreq, pool=12, budget=200e9, rel_tol=1e-8, seed=31):
    """Stream of Poisson solves (lx=8, perturbed boxes with element counts
    log-uniform in [512, 32768], N(0,1) RHS, tolerance 1e-8 relative) with
    Poisson arrivals at each rate, in four configurations: {FP64, mixed}
    x {served: up to 8 in flight under a 200 GB HBM budget, sequential: one
    at a time}.  Meshes are a resident pool built before the stream.  Per
    configuration: p50/p99 latency (arrival to completion), solves/s, and
    the HBM-roofline fraction of the bytes the solves moved (iterations x
    bytes per DOF-iteration of the schedule run, plus the FP64 residual
    replacements of mixed solves).  The error of served results against the
    FP64 CPU oracle is measured on requests of the smallest pool mesh.  At
    rates up to 100 req/s the mixed served configuration also runs the
    stream live: a producer thread submits each request at its arrival
    time while the server runs (nm_server_serve)."""
    import numpy as np
    import torch

    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    Es = np.exp(rng.uniform(np.log(512), np.log(32768), pool))
    dims = [_box_dims(E) for E in Es]
    out = {"pool": [list(d) for d in dims], "budget_GB": budget / 1e9,
           "tolerance": rel_tol, "requests_per_rate": nreq}
    configs = (("mixed_served", "mixed", 8), ("fp64_served", "fp64", 8),
               ("mixed_sequential", "mixed", 1), ("fp64_sequential", "fp64", 1))
    servers = {}
    for label, precision, inflight in configs:
        sv = S.Server(budget, max_inflight=inflight, precision=precision,
                      chunk_iters=8, keep_solutions=label.endswith("served"))
        ids = [sv.add_box_mesh(*d, lx=8, amp=0.1, seed=100 + k,
                               prewarm=min(inflight, 2), tune_reps=3)
               for k, d in enumerate(dims)]
        servers[label] = (sv, ids, precision)
    # bytes per DOF-iteration of each pool mesh's schedule (FP64 / FP32)
    sv0, ids0, _ = servers["fp64_sequential"]
    mesh_info = [sv0.mesh_problem(i).info for i in ids0]
    small = int(np.argmin([d[0] * d[1] * d[2] for d in dims]))

    def moved_bytes(res, mesh_of, precision):
        word = 8 if precision == "fp64" else 4
        tot = 0.0
        for r, m in zip(res, mesh_of):
            inf = mesh_info[m]
            b = fused_pcg_bytes_per_dof(inf, r.n, 8, word) if inf.cg_face_sums else \
                pcg_bytes_per_dof(inf, r.n, word)
            tot += r.iter * b * r.n
            if precision == "mixed":
                tot += r.outer * replacement_bytes_per_dof(inf, r.n) * r.n
        return tot

    oracle_err = {}

    def run(label, arrivals, mesh_of, seed0, live=False):
        sv, ids, precision = servers[label]
        if live:
            # a producer thread submits each request at its arrival time
            # while the server runs (nm_server_serve); the arrival stamp is
            # the submission time on the server's clock
            import threading
            rids = []

            def producer():
                t_start = time.perf_counter()
                try:
                    for q, (t, m) in enumerate(zip(arrivals, mesh_of)):
                        wait = t_start + float(t) - time.perf_counter()
                        if wait > 0:
                            time.sleep(wait)
                        rids.append(sv.submit(ids[m], seed0 + q, rel_tol=rel_tol,
                                              t_arrive=-1.0))
                finally:
                    sv.close()

            th = threading.Thread(target=producer)
            th.start()
            wall = sv.serve()
            th.join()
        else:
            rids = [sv.submit(ids[m], seed0 + q, rel_tol=rel_tol, t_arrive=float(t))
                    for q, (t, m) in enumerate(zip(arrivals, mesh_of))]
            wall = sv.run()
        res = [sv.result(r) for r in rids]
        lat = np.array([r.t_done - r.t_arrive for r in res])
        _, peak = sv.memory()
        byt = moved_bytes(res, mesh_of, precision)
        d = {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
             "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
             "solves_per_s": round(len(res) / wall, 3),
             "wall_s": round(wall, 3),
             "arrival_span_s": round(float(arrivals[-1]), 3),
             "mean_iter": round(float(np.mean([r.iter for r in res])), 1),
             "mean_dof": int(sum(r.n for r in res) / len(res)),
             "all_converged": all(r.state == 2 and r.converged for r in res),
             "peak_GB": round(peak / 1e9, 2),
             "hbm_GBs": round(byt / wall / 1e9, 1),
             "roofline_frac": round(byt / wall / 1e9 / HBM_PEAK_GBS, 4)}
        # error against the FP64 oracle: the first request on the smallest mesh
        if label.endswith("served") and precision not in oracle_err:
            for q, m in enumerate(mesh_of):
                if m != small:
                    continue
                xs = sv.solution(rids[q]).cpu().numpy()
                pr = sv.mesh_problem(ids[m])
                f = pr.rhs_normal_device(seed0 + q).cpu().numpy()
                c = O.OracleCase(*dims[m], 8, amp=0.1, seed=100 + m)
                x_or, m_or = c.cg(f, rel_tol * c.res_norm(f))
                oracle_err[precision] = {
                    "mesh": list(dims[m]), "request_seed": seed0 + q,
                    "iterations": res[q].iter, "oracle_iterations": m_or.iter,
                    "rel_l2_vs_oracle": float(np.linalg.norm(xs - x_or) / np.linalg.norm(x_or))}
                break
        log(f"[bench] serve {label}{' (live arrivals)' if live else ''}: {d}")
        return d

    for k, rate in enumerate(rates):
        arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
        arrivals -= arrivals[0]
        mesh_of = rng.integers(0, pool, nreq)
        out[f"{int(rate)}_req_s"] = {
            label: run(label, arrivals, mesh_of, 1000 * (k + 1)) for label, _, _ in configs}
        if rate <= 100:
            # the same stream with real-time arrivals (submitted while the
            # server runs) on the mixed served configuration
            out[f"{int(rate)}_req_s"]["mixed_served_live"] = run(
                "mixed_served", arrivals, mesh_of, 1000 * (k + 1), live=True)
    out["error_vs_oracle"] = oracle_err
    for sv, _, _ in servers.values():
        sv.free()
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE passes, scaled by the reference copy), or None
    when that summary does not cover the kernel."""
    path = os.path.join(ROOT, "profiles", "axgs_pcg_poisson_64cube_lx8.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                val, name = line.split(":", 1)[1].strip().split(None, 1)
                if name == f"kernel={kernel}":
                    return int(val)
    except OSError:
        pass
    return None


def ax_kernel_name(code, lx):
    """Name of the FP64 Ax kernel a variant code launches (no fused dot or
    CG prologue)."""
    if code >= 300 and lx in (4, 8):
        return f"ax_tri_mfma_kernel<{lx}, false>"  # the Poisson operator: no mass term
    if code >= 200:
        return f"ax_tri_kernel<double,{lx},{code % 100 or 1}>"
    return f"ax_kstep_kernel<double,{lx},{code % 100 or 1}>"


def roofline_entry(code, lx, gbs, tfs, nbytes, ms, n):
    """Roofline object of the dominant kernel (the FP64 Ax launch).  The
    stored-geometry variants stream 64 B/DOF and are HBM-bound.  The
    in-kernel-geometry variants read 16.6 B/DOF and execute ~180 FP64
    flop/DOF (a count modelled from the kernel source, ax_flops_per_dof), so
    they are graded against the FP64 peak, 78.6 TFLOP/s for the vector ALUs
    and for the matrix cores alike on MI355X: "fp64-valu" for the vector
    kernel, "fp64" for the kernel that runs the contractions (12 lx
    flop/DOF) on v_mfma_f64_4x4x4f64 and the geometry on the vector ALUs;
    both fractions are kept."""
    tri = code >= 200
    mfma = code >= 300 and lx in (4, 8)
    name = ax_kernel_name(code, lx)
    hbm_frac = gbs / HBM_PEAK_GBS
    fp_frac = tfs / FP64_PEAK_TFS
    common = {"kernel": name, "traffic": measured_traffic(name),
              "algorithmic_bytes_per_launch": nbytes,
              "flops_per_dof": ax_flops_per_dof(lx, code), "flops_modelled": True,
              "ms_per_launch": round(ms, 4),
              "hbm_GBs": round(gbs, 1), "hbm_frac": round(hbm_frac, 4),
              "fp64_TFs": round(tfs, 2), "fp64_frac": round(fp_frac, 4)}
    if mfma:
        common["matrix_flops_per_dof"] = 12 * lx
    if tri:
        return {"bound": "fp64" if mfma else "fp64-valu", "achieved": round(tfs, 2),
                "peak": FP64_PEAK_TFS, "unit": "TFLOP/s", "frac": round(fp_frac, 4), **common}
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(hbm_frac, 4), **common}


def fourier_rhs(p, seed, kmax=4):
    """b = mask(gs(B f)), f = sum over |k| <= kmax of a_k cos(2 pi k.x)
    + b_k sin(2 pi k.x) with a_k, b_k ~ N(0,1) seeded (k in Z^3, k_x >= 0)."""
    import numpy as np
    import torch
    from neko_amd import solve as S
    rng = np.random.default_rng(seed)
    x, y, z = (p.field(c) for c in (S.FIELD_X, S.FIELD_Y, S.FIELD_Z))
    f = torch.zeros_like(x)
    two_pi = 2.0 * np.pi
    for kx in range(0, kmax + 1):
        for ky in range(-kmax, kmax + 1):
            for kz in range(-kmax, kmax + 1):
                if kx * kx + ky * ky + kz * kz > kmax * kmax:
                    continue
                a, b = rng.standard_normal(2)
                ph = two_pi * (kx * x + ky * y + kz * z)
                f += a * torch.cos(ph) + b * torch.sin(ph)
    f = (f * p.field(S.FIELD_B)).contiguous()
    p.gs_op(f)
    p.bc_apply(f)
    return f


def helmholtz_pmg(S, E=48, lx=8, rel_tol=1e-10, default_cap=300):
    """Velocity-type Helmholtz: 48^3 elements, lx = 8, +-0.08h, periodic in
    x, h1 = 1/1600, h2 = 1.5/dt = 1500 (dt = 1e-3), Fourier-mode RHS
    (|k| <= 4, N(0,1) amplitudes).  GMRES(30) preconditioned by the
    p-multigrid V-cycle on hsmg's levels 8 -> 4 -> 2 with Chebyshev(4)
    smoothing accelerated by overlapping Schwarz and the coarse Jacobi-CG.
    Level operators with the fine h2 on every level (smoothers in FP32 and
    in FP64), and with Neko's coef default h2 = 1 on the coarser levels
    (phmg's own set-up, capped at 300 iterations: its coarse correction is
    mis-scaled for this h2); the coarse solve is hsmg's Jacobi-CG or
    phmg's own tree-AMG (3 levels, one V-cycle, Chebyshev(4))."""
    import numpy as np
    import torch
    t0 = time.time()
    p = S.Problem.box(E, E, E, lx, periodic=(1, 0, 0), amp=0.08, seed=77,
                      h1=1.0 / 1600, h2=1500.0)
    tuned = p.autotune(reps=3)
    f = fourier_rhs(p, 88)
    mult = p.field(S.FIELD_MULT)
    r0 = float(torch.sqrt(torch.sum(f * f * mult)) / np.sqrt(p.volume))
    out = {"workload": f"Helmholtz box {E}^3 lx={lx} periodic x, h1=1/1600, h2=1500",
           "local_dof": p.n, "setup_s": round(time.time() - t0, 2), "kernel_variants": tuned}
    runs = [("fine", "fp32", 2000, "cg", 0), ("fine", "fp32", 2000, "amg", 0),
            ("fine", "fp64", 2000, "cg", 0), ("default", "fp64", default_cap, "cg", 0),
            ("default", "fp64", default_cap, "amg", 0),
            # the Schwarz FDM solves on the vector ALUs instead of the
            # matrix cores, for comparison
            ("fine", "fp32", 2000, "cg", 1), ("fine", "fp64", 2000, "cg", 1)]
    for lh2, smoother, cap, coarse, fdm_valu in runs:
        S.fdm_set_variant(fdm_valu)
        t1 = time.time()
        ph = S.PHMG(p, facet_bc=2, smooth_its=4, level_h2=lh2, smoother=smoother)
        if coarse == "amg":
            ph.set_coarse_amg(3, 1, 4)  # phmg's own tree-AMG (defaults 3, 1, 4)
        torch.cuda.synchronize()
        t_pc = time.time() - t1
        g = S.GMRES(p, pc=ph, lgmres=30)
        x = torch.zeros_like(f)
        torch.cuda.synchronize()
        t2 = time.time()
        m = g.solve(x, f, rel_tol * r0, max_iter=cap)
        torch.cuda.synchronize()
        dt = time.time() - t2
        key = f"gmres_phmg_level_h2_{lh2}_smoother_{smoother}_coarse_{coarse}"
        if fdm_valu:
            key += "_fdm_vector_alu"
        # roofline on a lower-bound byte count of the fine level per GMRES
        # iteration: the outer FP64 Ax+gs (in-kernel geometry), the 9 fine
        # Ax+gs of the V-cycle (4 pre- and 4 post-smoothing Chebyshev steps
        # and the residual; FP32 smoothers count 4-byte words), classical
        # Gram-Schmidt over j basis vectors (dots read j+1 vectors, the
        # update reads j+1 and writes one: 8 (2j+3) B/DOF, j = 15.5 on
        # average over a restart cycle of 30) and the stored z_j (8 B/DOF).
        # The Schwarz/FDM solves and the coarser levels are left out, so the
        # fraction understates the bandwidth the solve uses.
        info = p.info
        ax64 = ax_bytes_per_dof(lx, tuned["ax64"], h2=True) + gs_bytes(info) / p.n
        wsm = 4 if smoother == "fp32" else 8
        axs = ax_bytes_per_dof(lx, tuned["ax32" if wsm == 4 else "ax64"], h2=True) * wsm / 8 \
            + gs_bytes(info, wsm) / p.n
        b_it = ax64 + 9 * axs + 8 * (2 * 15.5 + 3) + 8
        frac = m.iter * b_it * p.n / dt / 1e9 / HBM_PEAK_GBS if m.converged else None
        out[key] = {
            "roofline_lower_bound_bytes_per_dof_iteration": round(b_it, 1),
            "roofline_frac_lower_bound": round(frac, 4) if frac else None,
            "iterations": m.iter, "converged": m.converged, "seconds": round(dt, 3),
            "solves_per_s": round(1.0 / dt, 4) if m.converged else None,
            "ms_per_iteration": round(dt / max(m.iter, 1) * 1e3, 3),
            "res_final_rel": m.res_final / m.res_start, "pc_setup_s": round(t_pc, 2),
            "coarse_levels": ph.coarse_amg_levels() if coarse == "amg" else None,
            "chebyshev_bounds": ph.eigs()}
        log(f"[bench] helmholtz {key}: {out[key]}")
        g.free()
        ph.free()
    S.fdm_set_variant(0)
    # Jacobi-PCG on the same system for reference
    cg = S.CG(p, "fp64")
    x = torch.zeros_like(f)
    torch.cuda.synchronize()
    t3 = time.time()
    m = cg.solve(x, f, rel_tol * r0)
    torch.cuda.synchronize()
    out["cg_jacobi_fp64"] = {"iterations": m.iter, "seconds": round(time.time() - t3, 4)}
    log(f"[bench] helmholtz cg jacobi: {out['cg_jacobi_fp64']}")
    cg.free()
    p.free()
    return out


def batched(S, ninst, seed=47, rel_tol=1e-10):
    """Batched uneven instances: element counts log-uniform in [64, 4096]
    (box dims chosen to match), lx uniform in {6, 8, 10}, h2 = 0 or 1/dt
    with dt ~ U[1e-4, 1e-2] (then h1 = 1/1600), +-0.1h perturbed, RHS
    N(0,1); FP32 CG with FP64 residual replacement to 1e-10 relative, all
    instances in one batch (per-lx autotuned batched Ax, captured graph
    steps) vs the same instances solved one after another by the
    single-problem mixed solver on the same trilinear operator form (fused
    CG schedule)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    t0 = time.time()
    probs, fs, spec = [], [], []
    for k in range(ninst):
        E = float(np.exp(rng.uniform(np.log(64), np.log(4096))))
        lx = int(rng.choice([6, 8, 10]))
        if rng.random() < 0.5:
            h1, h2 = 1.0, 0.0
        else:
            h1, h2 = 1.0 / 1600, 1.0 / rng.uniform(1e-4, 1e-2)
        d = _box_dims(E)
        p = S.Problem.box(*d, lx, amp=0.1, seed=500 + k, h1=h1, h2=h2)
        probs.append(p)
        spec.append((h1, h2))
        fs.append(p.rhs_normal_device(900 + k))
    dof = sum(p.n for p in probs)
    t_setup = time.time() - t0
    b = S.Batch(probs)
    tuned = b.autotune(reps=3)
    xs = [torch.zeros_like(f) for f in fs]
    b.solve(xs, fs, rel_tol=rel_tol)          # warm-up (graph upload, caches)
    torch.cuda.synchronize()
    t1 = time.time()
    mons = b.solve(xs, fs, rel_tol=rel_tol)
    torch.cuda.synchronize()
    t_batch = time.time() - t1
    # sequential: single-problem mixed solver, one instance after another
    t_seq = 0.0
    seq_iters, max_dev = [], 0.0
    for p, f, xb in zip(probs, fs, xs):
        p.set_epb(201, 201)  # in-kernel geometry and the fused CG, as the batch
        p.set_gs_variant(3)  # block sums as packed segments, as the batch
        cg = S.CG(p, "mixed")
        x = torch.zeros_like(f)
        mult = p.field(S.FIELD_MULT)
        r0 = float(torch.sqrt(torch.sum(f * f * mult)) / np.sqrt(p.volume))
        torch.cuda.synchronize()
        t2 = time.time()
        m = cg.solve(x, f, rel_tol * r0)
        torch.cuda.synchronize()
        t_seq += time.time() - t2
        seq_iters.append(m.iter)
        max_dev = max(max_dev, float(torch.linalg.norm(x - xb) / torch.linalg.norm(xb)))
        cg.free()
    steps, launched = b.stats()
    useful = float(sum(m.iter * p.n for m, p in zip(mons, probs)))
    # HBM bytes the solves moved: iterations x the FP32 fused schedule's
    # bytes per DOF-iteration (in-kernel geometry) + the FP64 replacements
    moved = 0.0
    for m, p in zip(mons, probs):
        moved += m.iter * fused_pcg_bytes_per_dof(p.info, p.n, p.lx, 4) * p.n
        moved += m.outer * replacement_bytes_per_dof(p.info, p.n) * p.n
    # error against the FP64 CPU oracle on the smallest problems
    from oracle import oracle as O
    order = np.argsort([p.n for p in probs])[:3]
    errs = []
    for k in order:
        p, k = probs[k], int(k)
        c = O.OracleCase(*p.box_dims, p.lx, amp=0.1, seed=500 + k, h1=spec[k][0], h2=spec[k][1])
        f_or = fs[k].cpu().numpy()
        x_or, _ = c.cg(f_or, rel_tol * c.res_norm(f_or))
        errs.append(float(np.linalg.norm(xs[k].cpu().numpy() - x_or) / np.linalg.norm(x_or)))
    out = {"instances": ninst, "total_dof": dof, "setup_s": round(t_setup, 2),
           "autotuned_epb": tuned,
           "graph_steps": steps,
           "active_dof_per_step": round(useful / max(steps, 1)),
           "launched_dof_per_step": round(launched / max(steps, 1)),
           "active_fraction_of_launched": round(useful / max(launched, 1.0), 4),
           "hbm_GBs": round(moved / t_batch / 1e9, 1),
           "roofline_frac": round(moved / t_batch / 1e9 / HBM_PEAK_GBS, 4),
           "max_rel_l2_vs_oracle_smallest3": max(errs),
           "batched_s": round(t_batch, 3), "batched_solves_per_s": round(ninst / t_batch, 2),
           "batched_GDOF_it_per_s": round(sum(m.iter * p.n for m, p in zip(mons, probs))
                                          / t_batch / 1e9, 2),
           "sequential_s": round(t_seq, 3),
           "sequential_solves_per_s": round(ninst / t_seq, 2),
           "speedup": round(t_seq / t_batch, 2),
           "all_converged": all(m.converged for m in mons),
           "iters_batched_mean": round(float(np.mean([m.iter for m in mons])), 1),
           "iters_sequential_mean": round(float(np.mean(seq_iters)), 1),
           "max_rel_l2_batched_vs_sequential": max_dev}
    log(f"[bench] batched: {out}")
    b.free()
    for p in probs:
        p.free()
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(lx, seconds):
    """The oracle (the C FP64 restatement of Neko's CPU path, OpenMP, one
    process) on the GPU host's cores, timed the way the reference harnesses
    time (tests/bench/ax, tests/bench/nekbone/driver.f90:84-108): on a
    16^3-element lx=8 sample of the same perturbed box, warm-up then 100
    timed Ax+gs applies (mean and standard deviation; fewer if they exceed
    half the time budget), 50 Jacobi-PCG iterations (iterations/s), and
    whole solves of the 8^3 sine problem to 1e-10 (solves/s)."""
    import numpy as np

    from oracle import oracle as O
    Es = 16
    cores = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", cores))
    c = O.OracleCase(Es, Es, Es, lx, amp=0.1, seed=2024)
    u = np.random.default_rng(0).standard_normal(c.n)
    w = np.zeros(c.n)
    L = O.lib()
    args = (w, u, c.d, c.dt, c.h1, *c.g, c.nel, lx, c.h2, c.B, 0)

    def axgs():
        L.or_ax_helm(*args)
        c.gs(w)
    for _ in range(2):
        axgs()
    times = []
    t_all = time.time()
    while len(times) < 100 and time.time() - t_all < seconds / 2:
        t0 = time.perf_counter()
        axgs()
        times.append(time.perf_counter() - t0)
    times = np.array(times)
    gdofs = c.n / times.mean() / 1e9
    # PCG iterations/s on the same sample (absolute tolerance 0: exactly 50)
    f = w.copy()
    c.mask_apply(f)
    t0 = time.time()
    _, mon = c.cg(f, 0.0, max_iter=50)
    it_s = 50 / (time.time() - t0)
    # the 8^3 sine problem solved to 1e-10
    c1 = O.OracleCase(8, 8, 8, lx)
    f1 = c1.rhs_sine()
    tol = 1e-10 * c1.res_norm(f1)
    reps, t0 = 0, time.time()
    while reps < 3 or time.time() - t0 < 2.0:
        _, m1 = c1.cg(f1, tol)
        reps += 1
    c1_s = (time.time() - t0) / reps
    return {"value": round(gdofs, 4), "unit": "GDOF/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle Ax+gs (FP64 C restatement of Neko's CPU path, OpenMP, "
                      f"1 process x {threads} threads) on a {Es}^3-element lx={lx} "
                      f"+-0.1h box ({c.n} DOF), {len(times)} timed applies",
            "cpu_model": _cpu_model(), "affinity_cores": cores,
            "axgs_ms_mean": round(times.mean() * 1e3, 3),
            "axgs_ms_std": round(times.std() * 1e3, 3),
            "pcg_it_per_s": round(it_s, 2),
            "pcg_GDOF_it_per_s": round(it_s * c.n / 1e9, 4),
            "c1_solve": {"workload": "8^3 undeformed box, lx=8, sine RHS, Jacobi-PCG 1e-10",
                         "iterations": m1.iter, "seconds": round(c1_s, 4),
                         "solves_per_s": round(1.0 / c1_s, 3)}}


if __name__ == "__main__":
    main()
