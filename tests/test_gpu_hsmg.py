"""GPU parity of the hsmg preconditioner (pc_hsmg.f90, schwarz.f90, fdm.f90)
and of GMRES(m) (gmres.f90) against the CPU FP64 oracle.

Bars: one Schwarz smoothing and one hsmg application within 1e-11
relative (max-norm; the coarse level runs 10 CG iterations, whose dot
products sum in a different order than the oracle's); GMRES solves to
1e-10 of the initial residual: iteration count within +-1 and solution
within 1e-9 relative L2 of the oracle's.  All calls go through the C ABI.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", torch.float64)


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _pair(dims, lx, amp, per, h1=1.0, h2=0.0, seed=7):
    from neko_amd.solve import Problem
    c = O.OracleCase(*dims, lx, amp=amp, seed=seed, periodic=per, h1=h1, h2=h2)
    p = Problem.box(*dims, lx, periodic=per, amp=amp, seed=seed, h1=h1, h2=h2)
    return c, p


# A periodic direction has at least 3 elements: with 2, the two elements
# share both faces and the box's vertex-keyed topology identifies distinct
# edges and faces, so the assembled operator is not the Laplacian (Jacobi-
# GMRES does not converge on it either).
HS_CASES = [
    # dims, lx, amp, periodic, facet_bc
    ((3, 3, 3), 8, 0.1, (0, 0, 0), 1),
    ((3, 2, 2), 5, 0.1, (1, 0, 0), 2),
    ((2, 2, 3), 4, 0.05, (0, 0, 1), 1),
    ((2, 2, 2), 6, 0.0, (0, 0, 0), 2),
    ((2, 2, 3), 3, 0.1, (0, 0, 0), 1),
    ((2, 3, 2), 10, 0.1, (0, 1, 0), 1),
]


@pytest.mark.parametrize("case", HS_CASES)
def test_schwarz_fine_matches_oracle(case):
    from neko_amd.solve import HSMG
    dims, lx, amp, per, fb = case
    print("schwarz", case, flush=True)
    c, p = _pair(dims, lx, amp, per)
    H = O.OracleHsmg(c, facet_bc=fb)
    hs = HSMG(p, facet_bc=fb)
    r = c.rhs_normal(11)
    e_or = H.schwarz_fine(r)
    rd = _t(r)
    e = torch.zeros_like(rd)
    hs.schwarz_fine(e, rd)
    assert _relmax(_np(e), e_or) < 1e-11


@pytest.mark.parametrize("case", HS_CASES)
def test_hsmg_apply_matches_oracle(case):
    from neko_amd.solve import HSMG
    dims, lx, amp, per, fb = case
    print("hsmg apply", case, flush=True)
    c, p = _pair(dims, lx, amp, per)
    H = O.OracleHsmg(c, facet_bc=fb)
    hs = HSMG(p, facet_bc=fb)
    r = c.rhs_normal(13)
    z_or = H.apply(r)
    rd = _t(r)
    z = torch.zeros_like(rd)
    hs.apply(z, rd)
    assert _relmax(_np(z), z_or) < 1e-11
    # the input is left untouched and a second application is identical
    assert np.array_equal(_np(rd), r)
    z2 = torch.zeros_like(rd)
    hs.apply(z2, rd)
    assert torch.equal(z, z2)


def _gmres_pair(dims, lx, amp, per, pc, fb=1, h1=1.0, h2=0.0, lgmres=30):
    from neko_amd.solve import GMRES, HSMG
    c, p = _pair(dims, lx, amp, per, h1=h1, h2=h2)
    f = c.rhs_normal(5)
    tol = 1e-10 * c.res_norm(f)
    if pc == "hsmg":
        H = O.OracleHsmg(c, facet_bc=fb)
        x_or, m_or = H.gmres(f, tol, lgmres=lgmres)
        kp = HSMG(p, facet_bc=fb)
    else:
        x_or, m_or = c.gmres(f, tol, lgmres=lgmres,
                             pc=c.jacobi() if pc == "jacobi" else None)
        kp = pc
    g = GMRES(p, pc=kp, lgmres=lgmres)
    x = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    mon = g.solve(x, _t(f), tol)
    err = np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or)
    return mon, m_or, err


def test_gmres_hsmg_poisson():
    """Pressure-type Poisson on a deformed box, GMRES + hsmg (the reference's
    pressure solver pairing)."""
    print("gmres hsmg poisson", flush=True)
    mon, m_or, err = _gmres_pair((3, 3, 3), 8, 0.1, (0, 0, 0), "hsmg", fb=1)
    assert mon.converged and m_or.converged
    assert abs(mon.iter - m_or.iter) <= 1, (mon.iter, m_or.iter)
    assert abs(mon.res_start - m_or.res_start) < 1e-12 * m_or.res_start
    assert err < 1e-9


def test_gmres_hsmg_periodic_lx6():
    print("gmres hsmg periodic", flush=True)
    mon, m_or, err = _gmres_pair((3, 3, 2), 6, 0.1, (1, 0, 0), "hsmg", fb=1)
    assert mon.converged and m_or.converged
    assert abs(mon.iter - m_or.iter) <= 1, (mon.iter, m_or.iter)
    assert err < 1e-9


def test_gmres_jacobi_helmholtz_velocity():
    """Velocity-type Helmholtz (h1 = 1/Re, h2 = rho/dt), Jacobi."""
    print("gmres jacobi helmholtz", flush=True)
    mon, m_or, err = _gmres_pair((3, 3, 3), 6, 0.08, (1, 0, 0), "jacobi",
                                 h1=1.0 / 1600, h2=1500.0, lgmres=3)
    assert mon.converged and m_or.converged
    assert abs(mon.iter - m_or.iter) <= 1, (mon.iter, m_or.iter)
    assert err < 1e-9


def test_gmres_jacobi_restarts():
    """Restart length 5 on a mildly shifted Laplacian: ~50 restart cycles
    (the oracle needs 257 iterations)."""
    print("gmres jacobi restarts", flush=True)
    mon, m_or, err = _gmres_pair((3, 3, 3), 6, 0.08, (1, 0, 0), "jacobi",
                                 h1=1.0, h2=10.0, lgmres=5)
    assert mon.converged and m_or.converged
    assert abs(mon.iter - m_or.iter) <= 1, (mon.iter, m_or.iter)
    assert err < 1e-9


def test_gmres_identity_poisson():
    print("gmres identity", flush=True)
    mon, m_or, err = _gmres_pair((2, 2, 2), 5, 0.1, (0, 0, 0), "identity",
                                 lgmres=20)
    assert mon.converged and m_or.converged
    assert abs(mon.iter - m_or.iter) <= 1, (mon.iter, m_or.iter)
    assert err < 1e-9


def test_hsmg_rejects_unsupported_order():
    from neko_amd.solve import HSMG, Problem
    p = Problem.box(2, 2, 2, 12)
    with pytest.raises(RuntimeError, match="lx=12"):
        HSMG(p)


# --- Chebyshev smoothing in a phmg-style V-cycle ------------------------------
PH_CASES = [
    # dims, lx, amp, periodic, h1, h2, facet_bc, level_h2
    ((3, 3, 3), 8, 0.1, (0, 0, 0), 1.0, 0.0, 1, "default"),
    ((3, 2, 3), 6, 0.08, (1, 0, 0), 1.0 / 1600, 1500.0, 2, "default"),
    ((3, 3, 2), 8, 0.08, (1, 0, 0), 1.0 / 1600, 1500.0, 2, "fine"),
    ((2, 3, 2), 5, 0.1, (0, 0, 0), 1.0, 0.0, 2, "default"),
]


def _phmg_pair(case):
    from neko_amd.solve import PHMG
    dims, lx, amp, per, h1, h2, fb, lh2 = case
    c, p = _pair(dims, lx, amp, per, h1=h1, h2=h2)
    P = O.OraclePhmg(c, facet_bc=fb, level_h2=lh2)
    ph = PHMG(p, facet_bc=fb, smooth_its=4, level_h2=lh2)
    return c, p, P, ph


@pytest.mark.parametrize("case", PH_CASES)
def test_phmg_power_estimates_and_apply(case):
    """Chebyshev eigenvalue bounds (150 power iterations) within 1e-9 of the
    oracle's; one V-cycle application within 1e-9 (max-norm relative)."""
    print("phmg", case, flush=True)
    c, p, P, ph = _phmg_pair(case)
    eg = ph.eigs()
    for l in range(2):
        assert abs(eg[l][0] - P.ch[l].tha) < 1e-9 * P.ch[l].tha
        assert abs(eg[l][1] - P.ch[l].dlt) < 1e-9 * P.ch[l].dlt
    r = c.rhs_normal(21)
    z_or = P.apply(r)
    rd = _t(r)
    z = torch.zeros_like(rd)
    ph.apply(z, rd)
    assert _relmax(_np(z), z_or) < 1e-9
    assert np.array_equal(_np(rd), r)


@pytest.mark.parametrize("case", [PH_CASES[0], PH_CASES[2]])
def test_gmres_phmg_matches_oracle(case):
    """GMRES(30) + Chebyshev(4)/Schwarz V-cycle: iterations +-1, solution
    within 1e-9 of the oracle's."""
    from neko_amd.solve import GMRES
    print("gmres phmg", case, flush=True)
    c, p, P, ph = _phmg_pair(case)
    f = c.rhs_normal(5)
    tol = 1e-10 * c.res_norm(f)
    x_or, m_or = P.gmres(f, tol)
    g = GMRES(p, pc=ph, lgmres=30)
    x = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    mon = g.solve(x, _t(f), tol)
    assert mon.converged and m_or.converged
    assert abs(mon.iter - m_or.iter) <= 1, (mon.iter, m_or.iter)
    err = np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or)
    assert err < 1e-9


@pytest.mark.parametrize("case", [PH_CASES[0], PH_CASES[2]])
def test_phmg_fp32_smoothers(case):
    """FP32 Chebyshev/Schwarz smoothing (FP64 transfers, coarse CG and
    outer GMRES): one V-cycle within 1e-5 of the FP64 oracle's; GMRES to
    1e-10 within 2 iterations of the oracle's FP64-smoother count and
    solution within 1e-9."""
    from neko_amd.solve import GMRES, PHMG
    print("phmg fp32", case, flush=True)
    dims, lx, amp, per, h1, h2, fb, lh2 = case
    c, p = _pair(dims, lx, amp, per, h1=h1, h2=h2)
    P = O.OraclePhmg(c, facet_bc=fb, level_h2=lh2)
    ph = PHMG(p, facet_bc=fb, smooth_its=4, level_h2=lh2, smoother="fp32")
    r = c.rhs_normal(21)
    z = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    ph.apply(z, _t(r))
    assert _relmax(_np(z), P.apply(r)) < 1e-5
    f = c.rhs_normal(5)
    tol = 1e-10 * c.res_norm(f)
    x_or, m_or = P.gmres(f, tol)
    g = GMRES(p, pc=ph, lgmres=30)
    x = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    mon = g.solve(x, _t(f), tol)
    assert mon.converged
    assert abs(mon.iter - m_or.iter) <= 2, (mon.iter, m_or.iter)
    assert np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or) < 1e-9


# --- the same preconditioners on the reference's own meshes -----------------
def _mesh_pair(name, lx, h1=1.0, h2=0.0):
    import os
    from neko_amd.solve import Problem, read_nmsh
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    v, pid, _ = read_nmsh(os.path.join(root, "tests", "golden", "nmsh", name))
    return (O.OracleCase.from_mesh(v, pid, lx, h1=h1, h2=h2),
            Problem.mesh(v, pid, lx, h1=h1, h2=h2))


@pytest.mark.parametrize("name,lx", [("hemi.nmsh", 5), ("lid.nmsh", 6)])
def test_hsmg_on_reference_mesh(name, lx):
    """Schwarz on the unstructured hemisphere mesh (and the periodic lid
    mesh): levels and the extended (lx+2) gs come from the mesh topology;
    one application within 1e-11, GMRES + hsmg iterations +-1 and the
    solution within 1e-9 of the oracle's."""
    from neko_amd.solve import GMRES, HSMG
    print("hsmg mesh", name, lx, flush=True)
    c, p = _mesh_pair(name, lx)
    H = O.OracleHsmg(c, facet_bc=1)
    hs = HSMG(p, facet_bc=1)
    r = c.rhs_normal(13)
    z_or = H.apply(r)
    z = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    hs.apply(z, _t(r))
    assert _relmax(_np(z), z_or) < 1e-11
    f = c.rhs_normal(5)
    tol = 1e-10 * c.res_norm(f)
    x_or, m_or = H.gmres(f, tol, lgmres=30)
    g = GMRES(p, pc=hs, lgmres=30)
    x = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    mon = g.solve(x, _t(f), tol)
    assert mon.converged and m_or.converged
    assert abs(mon.iter - m_or.iter) <= 1, (mon.iter, m_or.iter)
    assert np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or) < 1e-9


def test_phmg_on_reference_mesh():
    from neko_amd.solve import PHMG
    print("phmg mesh", flush=True)
    c, p = _mesh_pair("hemi.nmsh", 5, h1=1.0 / 1600, h2=1500.0)
    P = O.OraclePhmg(c, facet_bc=2, level_h2="fine")
    ph = PHMG(p, facet_bc=2, smooth_its=4, level_h2="fine")
    eg = ph.eigs()
    for l in range(2):
        assert abs(eg[l][0] - P.ch[l].tha) < 1e-9 * P.ch[l].tha
    r = c.rhs_normal(21)
    z_or = P.apply(r)
    z = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    ph.apply(z, _t(r))
    assert _relmax(_np(z), z_or) < 1e-9


@pytest.mark.parametrize("case", [HS_CASES[0], HS_CASES[2], HS_CASES[5]])
def test_fdm_matrix_cores_match_vector_alus(case):
    """The Schwarz FDM solve on the matrix cores (v_mfma_f64_16x16x4f64,
    the default) and on the vector ALUs (nm_fdm_set_variant(1)): the same
    fine-level Schwarz application to 1e-13, each within 1e-11 of the
    oracle (lx 8, 4, 10: extended sizes 10, 6, 12)."""
    from neko_amd.solve import HSMG, fdm_set_variant
    dims, lx, amp, per, fb = case
    print("fdm variants", case, flush=True)
    c, p = _pair(dims, lx, amp, per)
    H = O.OracleHsmg(c, facet_bc=fb)
    hs = HSMG(p, facet_bc=fb)
    r = c.rhs_normal(31)
    e_or = H.schwarz_fine(r)
    out = []
    try:
        for variant in (0, 1):
            fdm_set_variant(variant)
            e = torch.zeros(p.n, dtype=torch.float64, device="cuda")
            hs.schwarz_fine(e, _t(r))
            out.append(_np(e))
    finally:
        fdm_set_variant(0)
    assert _relmax(out[0], out[1]) < 1e-13
    for o in out:
        assert _relmax(o, e_or) < 1e-11


def test_fdm_matrix_cores_fp32_smoothers():
    """FP32 smoothing with the extended size 10 (lx 8) FDM on the matrix
    cores (v_mfma_f32_16x16x4f32) and on the vector ALUs: one V-cycle each
    within 1e-5 of the FP64 oracle's and within 1e-5 of each other."""
    from neko_amd.solve import PHMG, fdm_set_variant
    dims, lx, amp, per, h1, h2, fb, lh2 = PH_CASES[0]
    assert lx == 8
    print("fdm variants fp32", flush=True)
    c, p = _pair(dims, lx, amp, per, h1=h1, h2=h2)
    P = O.OraclePhmg(c, facet_bc=fb, level_h2=lh2)
    r = c.rhs_normal(41)
    z_or = P.apply(r)
    out = []
    try:
        for variant in (0, 1):
            fdm_set_variant(variant)
            ph = PHMG(p, facet_bc=fb, smooth_its=4, level_h2=lh2, smoother="fp32")
            z = torch.zeros(p.n, dtype=torch.float64, device="cuda")
            ph.apply(z, _t(r))
            out.append(_np(z))
            ph.free()
    finally:
        fdm_set_variant(0)
    assert _relmax(out[0], out[1]) < 1e-5
    for o in out:
        assert _relmax(o, z_or) < 1e-5
