"""GPU tests of the serving loop (nm_server_*): a request solved inside an
interleaved stream gives bit-for-bit the result of the same solve run
alone through nm_cg_solve (iterations equal, x identical), in FP64 and in
mixed precision; a tight HBM budget serialises admission without failing;
latencies are ordered (arrival <= admission <= completion).  The device
RHS generator agrees with the host/oracle generator to 1e-14."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

MESHES = [(4, 4, 4), (5, 3, 4), (3, 3, 6)]


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def test_device_rhs_matches_oracle_rhs():
    from neko_amd.solve import Problem
    c = O.OracleCase(3, 2, 3, 6, amp=0.1, seed=4)
    p = Problem.box(3, 2, 3, 6, amp=0.1, seed=4)
    f = _np(p.rhs_normal_device(17))
    f_or = c.rhs_normal(17)
    assert np.max(np.abs(f - f_or)) <= 1e-14 * np.max(np.abs(f_or))


def _alone(sv, mesh_id, seed, rel, precision):
    from neko_amd.solve import CG
    pr = sv.mesh_problem(mesh_id)
    f = pr.rhs_normal_device(seed)
    r0 = float(torch.sqrt(torch.sum(f * f * pr.field(7))) / np.sqrt(pr.volume))
    cg = CG(pr, precision)
    x = torch.zeros_like(f)
    mon = cg.solve(x, f, rel * r0, max_iter=20000)
    return x, mon


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_serve_interleaved_equals_alone(precision):
    from neko_amd.solve import Server
    print("serve", precision, flush=True)
    sv = Server(8e9, max_inflight=4, precision=precision, chunk_iters=8,
                keep_solutions=True)
    ids = [sv.add_box_mesh(*m, lx=8, amp=0.1, seed=10 + k, prewarm=1,
                           tune_reps=2 if k == 0 else 0)
           for k, m in enumerate(MESHES)]
    reqs = []
    arrivals = [0.0, 0.0, 0.001, 0.002, 0.002, 0.004, 0.02]
    for k, t in enumerate(arrivals):
        reqs.append((sv.submit(ids[k % len(ids)], 100 + k, rel_tol=1e-8,
                               t_arrive=t), ids[k % len(ids)], 100 + k))
    wall = sv.run()
    assert wall > 0
    for rid, mid, seed in reqs:
        r = sv.result(rid)
        assert r.state == 2 and r.converged
        assert r.t_arrive <= r.t_admit + 1e-9 and r.t_admit <= r.t_done
        assert r.res_final < 1e-8 * r.res_start
        x_alone, mon = _alone(sv, mid, seed, 1e-8, precision)
        assert mon.iter == r.iter, (mon.iter, r.iter)
        assert torch.equal(sv.solution(rid), x_alone)
    used, peak = sv.memory()
    assert 0 < used <= peak <= 8e9


def test_serve_tight_budget_serialises():
    from neko_amd.solve import Server
    print("serve budget", flush=True)
    probe = Server(64e9, max_inflight=1, precision="mixed")
    mid = probe.add_box_mesh(4, 4, 4, prewarm=1)
    used_one, _ = probe.memory()
    probe.free()
    # room for the mesh and one CG state only
    sv = Server(used_one * 1.2, max_inflight=4, precision="mixed")
    mid = sv.add_box_mesh(4, 4, 4, prewarm=0)
    rids = [sv.submit(mid, 7 + k, rel_tol=1e-8) for k in range(3)]
    sv.run()
    res = [sv.result(r) for r in rids]
    assert all(r.state == 2 and r.converged for r in res)
    # admissions happen one after another
    res.sort(key=lambda r: r.t_admit)
    for a, b in zip(res, res[1:]):
        assert b.t_admit >= a.t_done - 1e-9
    _, peak = sv.memory()
    assert peak <= used_one * 1.2


def test_serve_reference_meshes():
    """Requests on the reference's hemisphere and lid meshes, interleaved
    with a box mesh: every result equals the same solve run alone."""
    import os
    from neko_amd.solve import Server, read_nmsh
    print("serve meshes", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sv = Server(8e9, max_inflight=3, precision="mixed", chunk_iters=8,
                keep_solutions=True)
    ids = []
    for name, lx in (("hemi.nmsh", 5), ("lid.nmsh", 6)):
        v, pid, _ = read_nmsh(os.path.join(root, "tests", "golden", "nmsh", name))
        ids.append(sv.add_mesh(v, pid, lx=lx, prewarm=1))
    ids.append(sv.add_box_mesh(3, 3, 3, lx=6, prewarm=1))
    reqs = [(sv.submit(ids[k % 3], 300 + k, rel_tol=1e-8), ids[k % 3], 300 + k)
            for k in range(6)]
    sv.run()
    for rid, mid, seed in reqs:
        r = sv.result(rid)
        assert r.state == 2 and r.converged
        x_alone, mon = _alone(sv, mid, seed, 1e-8, "mixed")
        assert mon.iter == r.iter
        assert torch.equal(sv.solution(rid), x_alone)


def test_live_stream_submissions_from_another_thread():
    """nm_server_serve with requests submitted by another thread while it
    runs (arrival stamped at submission): every request completes, its
    times are ordered on the stream clock, and each served solve equals
    the same solve run alone bit for bit."""
    import threading
    import time
    from neko_amd.solve import Server
    print("serve live", flush=True)
    sv = Server(8e9, max_inflight=3, precision="mixed", chunk_iters=8, keep_solutions=True)
    ids = [sv.add_box_mesh(*m, lx=8, amp=0.1, seed=30 + k, prewarm=1,
                           tune_reps=2 if k == 0 else 0)
           for k, m in enumerate(MESHES)]
    submitted = []

    def producer():
        try:
            for k in range(8):
                mid = ids[k % len(ids)]
                rid = sv.submit(mid, 300 + k, rel_tol=1e-8, t_arrive=-1.0)
                submitted.append((rid, mid, 300 + k))
                time.sleep(0.005)
        finally:
            sv.close()

    th = threading.Thread(target=producer)
    th.start()
    wall = sv.serve()
    th.join()
    assert wall > 0 and len(submitted) == 8
    last_arrival = -1.0
    for rid, mid, seed in submitted:
        r = sv.result(rid)
        assert r.state == 2 and r.converged
        assert 0.0 <= r.t_arrive <= r.t_admit <= r.t_done <= wall + 1e-6
        assert r.t_arrive >= last_arrival  # submission order = arrival order
        last_arrival = r.t_arrive
        x_alone, mon = _alone(sv, mid, seed, 1e-8, "mixed")
        assert r.iter == mon.iter
        assert torch.equal(sv.solution(rid), x_alone)
