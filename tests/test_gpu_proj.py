"""GPU parity of the solution-space projection (nm_proj_*) against the
oracle's restatement over a sequence of solves with a slowly drifting RHS
(the pressure solve of a time-stepper): the projected RHS, the basis size,
the logged original residual, the Jacobi-PCG iterations and the final
solutions of every step; projection cuts the iterations as the basis
fills, and the basis restarts after L vectors."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_projection_sequence_matches_oracle(precision):
    from neko_amd.solve import CG, Problem, Projection
    print("projection", precision, flush=True)
    args = (4, 4, 3, 6)
    c = O.OracleCase(*args, amp=0.1, seed=8)
    p = Problem.box(*args, amp=0.1, seed=8)
    p.set_epb(201, 201)
    L, activ = 4, 1
    pj_or = O.OracleProjection(c, L, activ_step=activ)
    pj = Projection(p, L, activ_step=activ)
    cg = CG(p, precision)
    f = [c.rhs_normal(s) for s in (1, 2, 3, 4)]
    iters, iters_or, used, ms = [], [], [], []
    for t in range(1, 10):
        # a 4-dimensional family: once 4 solutions are held the projected
        # RHS is (nearly) zero
        b = (f[0] + np.sin(0.3 * t) * f[1] + 0.1 * np.cos(0.7 * t) * f[2] +
             0.01 * t * f[3])
        tol = 1e-10 * c.res_norm(b)
        b_or = b.copy()
        pj_or.pre(b_or, t)
        bd = torch.from_numpy(b).cuda()
        pj.pre(bd, t)
        # the basis holds solutions, which agree to the solve tolerance
        # (FP64) or to the mixed solver's ~1e-10 deviation from them
        btol = 1e-10 if precision == "fp64" else 1e-8
        assert np.max(np.abs(_np(bd) - b_or)) <= btol * np.max(np.abs(b))
        m, pm, res = pj.info()
        assert m == pj_or.m
        if t > activ and pm > 0:
            assert pm == pj_or.proj_m
            assert abs(res - pj_or.proj_res) <= 1e-12 * pj_or.proj_res
        x_or, mon_or = c.cg(b_or, tol, max_iter=5000)
        x = torch.zeros_like(bd)
        mon = cg.solve(x, bd, tol, max_iter=5000)
        assert mon.converged
        if precision == "fp64":
            assert abs(mon.iter - mon_or.iter) <= 1, (t, mon.iter, mon_or.iter)
        pj_or.post(x_or, t)
        pj.post(x, t)
        assert pj.info()[0] == pj_or.m
        rel = np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or)
        assert rel < 1e-8, (t, rel)
        # the solution of the original (unprojected) system
        r = b - c.ax_full(_np(x))
        assert c.res_norm(r) < 2e-10 * c.res_norm(b)
        iters.append(mon.iter)
        iters_or.append(mon_or.iter)
        used.append(pm if t > activ else 0)
        ms.append(pj.info()[0])
    print("iterations", iters, "oracle", iters_or, "vectors", used, "m", ms,
          flush=True)
    # restart after L vectors, and a full basis nearly solves the system
    assert 4 in ms and ms[ms.index(4) + 1] == 1, ms
    full = [it for it, u in zip(iters, used) if u == 4]
    assert full and max(full) < 0.3 * iters[0], (iters, used)
