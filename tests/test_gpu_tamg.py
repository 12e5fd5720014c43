"""phmg's tree-AMG coarse solve (nm_phmg_set_coarse_amg) against the
oracle's restatement (oracle/neko_oracle_tamg.c): the aggregation (level
sizes and every coarse-dof -> aggregate map) equal bit for bit, one phmg
V-cycle with the AMG coarse solve within 1e-10 (max-norm relative), and
GMRES + phmg(AMG) iterations +-1 with the solution within 1e-9."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    # dims, lx, amp, periodic, h1, h2, facet_bc, level_h2
    ((4, 4, 3), 6, 0.1, (0, 0, 0), 1.0, 0.0, 1, "default"),
    ((5, 4, 4), 8, 0.08, (1, 0, 0), 1.0 / 1600, 1500.0, 2, "fine"),
    ((6, 6, 6), 5, 0.1, (0, 0, 0), 1.0, 0.0, 2, "default"),
]


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _pair(case):
    from neko_amd.solve import PHMG, Problem
    dims, lx, amp, per, h1, h2, fb, lh2 = case
    c = O.OracleCase(*dims, lx, amp=amp, seed=7, periodic=per, h1=h1, h2=h2)
    p = Problem.box(*dims, lx, periodic=per, amp=amp, seed=7, h1=h1, h2=h2)
    P = O.OraclePhmg(c, facet_bc=fb, level_h2=lh2, coarse_amg=(3, 1, 4))
    ph = PHMG(p, facet_bc=fb, smooth_its=4, level_h2=lh2)
    ph.set_coarse_amg(3, 1, 4)
    return c, p, P, ph


@pytest.mark.parametrize("case", CASES)
def test_amg_hierarchy_and_vcycle(case):
    print("tamg", case, flush=True)
    c, p, P, ph = _pair(case)
    nn = ph.coarse_amg_levels()
    assert list(nn) == [int(v) for v in P.amg.nn]
    for l in range(1, len(nn)):
        assert np.array_equal(ph.coarse_amg_map(l), P.amg.map(l)), l
    r = c.rhs_normal(21)
    z_or = P.apply(r)
    z = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    ph.apply(z, _t(r))
    zz = _np(z)
    assert np.max(np.abs(zz - z_or)) < 1e-10 * np.max(np.abs(z_or))


@pytest.mark.parametrize("case", [CASES[0], CASES[1]])
def test_gmres_phmg_amg_matches_oracle(case):
    from neko_amd.solve import GMRES
    print("gmres phmg amg", case, flush=True)
    c, p, P, ph = _pair(case)
    f = c.rhs_normal(5)
    tol = 1e-10 * c.res_norm(f)
    x_or, m_or = P.gmres(f, tol)
    g = GMRES(p, pc=ph, lgmres=30)
    x = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    mon = g.solve(x, _t(f), tol)
    assert mon.converged and m_or.converged
    assert abs(mon.iter - m_or.iter) <= 1, (mon.iter, m_or.iter)
    assert np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or) < 1e-9


def test_amg_on_reference_mesh():
    """The hemisphere mesh's unstructured element graph."""
    from neko_amd.solve import GMRES, PHMG, Problem, read_nmsh
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    v, pid, _ = read_nmsh(os.path.join(root, "tests", "golden", "nmsh", "hemi.nmsh"))
    c = O.OracleCase.from_mesh(v, pid, 5)
    p = Problem.mesh(v, pid, 5)
    P = O.OraclePhmg(c, facet_bc=2, coarse_amg=(3, 1, 4))
    ph = PHMG(p, facet_bc=2, smooth_its=4)
    ph.set_coarse_amg(3, 1, 4)
    nn = ph.coarse_amg_levels()
    assert list(nn) == [int(q) for q in P.amg.nn]
    for l in range(1, len(nn)):
        assert np.array_equal(ph.coarse_amg_map(l), P.amg.map(l))
    f = c.rhs_normal(6)
    tol = 1e-10 * c.res_norm(f)
    x_or, m_or = P.gmres(f, tol)
    g = GMRES(p, pc=ph, lgmres=30)
    x = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    mon = g.solve(x, _t(f), tol)
    assert mon.converged and abs(mon.iter - m_or.iter) <= 1, (mon.iter, m_or.iter)
    assert np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or) < 1e-9


def test_amg_off_restores_cg():
    from neko_amd.solve import PHMG
    c, p, P, ph = _pair(CASES[0])
    ph.set_coarse_amg(0)
    assert ph.coarse_amg_levels() == []
    P2 = O.OraclePhmg(c, facet_bc=1, level_h2="default")
    r = c.rhs_normal(3)
    z = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    ph.apply(z, _t(r))
    zo = P2.apply(r)
    assert np.max(np.abs(_np(z) - zo)) < 1e-9 * np.max(np.abs(zo))
