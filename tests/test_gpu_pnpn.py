"""GPU parity of the pnpn operator glue against the oracle: the weak
gradient (nm_opgrad) on box and reference meshes, the fused velocity
residual (nm_vel_res), and nm_op_set_coeffs -- new h1/h2 reach Ax, the
Jacobi diagonal and an already created (graph-captured) CG solver."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _box(lx=6):
    from neko_amd.solve import Problem
    args = (3, 4, 3, lx)
    return (O.OracleCase(*args, amp=0.1, seed=6), Problem.box(*args, amp=0.1, seed=6))


@pytest.mark.parametrize("lx", [4, 8])
def test_opgrad_matches_oracle(lx):
    print("opgrad", lx, flush=True)
    c, p = _box(lx)
    u = np.random.default_rng(1).standard_normal(c.n)
    ref = c.opgrad(u)
    got = p.opgrad(torch.from_numpy(u).cuda())
    for g, r in zip(got, ref):
        assert np.abs(_np(g) - r).max() <= 1e-12 * np.abs(r).max()


def test_opgrad_on_reference_mesh():
    from neko_amd.solve import Problem, read_nmsh
    print("opgrad hemi", flush=True)
    v, pid, _ = read_nmsh(os.path.join(ROOT, "tests", "golden", "nmsh", "hemi.nmsh"))
    c = O.OracleCase.from_mesh(v, pid, 5)
    p = Problem.mesh(v, pid, 5)
    u = np.random.default_rng(2).standard_normal(c.n)
    ref = c.opgrad(u)
    got = p.opgrad(torch.from_numpy(u).cuda())
    for g, r in zip(got, ref):
        assert np.abs(_np(g) - r).max() <= 1e-12 * np.abs(r).max()


def test_vel_res_matches_oracle():
    print("vel res", flush=True)
    c, p = _box(8)
    rng = np.random.default_rng(3)
    fields = [rng.standard_normal(c.n) for _ in range(7)]
    u, v, w, pr, *f = fields
    mu, rho, bd, dt = 1.0 / 1600, 1.0, 1.5, 1e-3
    ref = c.vel_res(u, v, w, pr, f, mu, rho, bd, dt)
    d = [torch.from_numpy(a).cuda() for a in fields]
    for code in (0, 201):
        p.set_epb(code, code)
        got = p.vel_res(d[0], d[1], d[2], d[3], d[4:], mu, rho, bd, dt)
        for g, r in zip(got, ref):
            assert np.abs(_np(g) - r).max() <= 1e-12 * np.abs(r).max(), code


def test_set_coeffs_reaches_solver():
    """A CG solver created for h1 = 1, h2 = 0 solves the Helmholtz problem
    after set_coeffs(h1, h2): same iterations (+-1) and solution as the
    oracle built with those coefficients."""
    from neko_amd.solve import CG
    print("set coeffs", flush=True)
    p = _box(6)[1]
    h1, h2 = 0.05, 40.0
    c = O.OracleCase(3, 4, 3, 6, amp=0.1, seed=6, h1=h1, h2=h2)
    for prec in ("fp64", "mixed"):
        p.set_coeffs(1.0, 0.0)
        p.set_epb(201, 201)
        cg = CG(p, prec)
        p.set_coeffs(h1, h2)
        f = c.rhs_normal(7)
        tol = 1e-10 * c.res_norm(f)
        x_or, mon_or = c.cg(f, tol, max_iter=5000)
        x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
        mon = cg.solve(x, torch.from_numpy(f).cuda(), tol, max_iter=5000)
        assert mon.converged
        if prec == "fp64":
            assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
        rel = np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or)
        assert rel < 1e-9, (prec, rel)
    jac = _np(p.field(8))
    assert np.abs(jac - c.jacobi()).max() <= 1e-12 * np.abs(c.jacobi()).max()


def test_curl_and_prs_res_match_oracle():
    print("curl / prs res", flush=True)
    c, p = _box(6)
    rng = np.random.default_rng(9)
    f = [rng.standard_normal(c.n) for _ in range(7)]
    ue, fv, pr = f[:3], f[3:6], f[6]
    d = [torch.from_numpy(a).cuda() for a in f]
    ref = c.curl(*ue)
    got = p.curl(d[0], d[1], d[2])
    for g, r in zip(got, ref):
        assert np.abs(_np(g) - r).max() <= 1e-11 * np.abs(r).max()
    mu, rho = 1.0 / 1600, 1.2
    ref = c.prs_res(pr, ue, fv, mu, rho)
    got = p.prs_res(d[6], d[0:3], d[3:6], mu, rho)
    assert np.abs(_np(got) - ref).max() <= 1e-10 * np.abs(ref).max()


def test_surface_normals_and_prs_res_surf_match_oracle():
    """Facet-normal surfaces (two sides sharing an edge, so edge points get
    two facets' contributions) and the whole pressure residual with a
    velocity-prescribed and a symmetry surface: points bit-equal, normals
    1e-12, residual 1e-10 against the oracle."""
    from neko_amd.solve import Surface
    print("surfaces / prs res surf", flush=True)
    c, p = _box(6)
    dims = (3, 4, 3)
    prs_f = c.box_facets(dims, 2) + c.box_facets(dims, 4)
    sym_f = c.box_facets(dims, 5)
    sp, ss = Surface(p, prs_f), Surface(p, sym_f)
    for sf, fl in ((sp, prs_f), (ss, sym_f)):
        pts, nx, ny, nz = sf.normals()
        rp, rx, ry, rz = c.facet_normals(fl)
        assert np.array_equal(pts.astype(np.int64), rp)
        for g, r in ((nx, rx), (ny, ry), (nz, rz)):
            assert np.abs(g - r).max() <= 1e-12 * np.abs(r).max()
    rng = np.random.default_rng(12)
    f = [rng.standard_normal(c.n) for _ in range(10)]
    ue, uc, fv, pr = f[:3], f[3:6], f[6:9], f[9]
    d = [torch.from_numpy(a).cuda() for a in f]
    mu, rho, bd, dt = 1.0 / 1600, 1.2, 1.5, 2e-3
    ref = c.prs_res_surf(pr, ue, uc, fv, mu, rho, bd, dt,
                         prs=c.facet_normals(prs_f), sym=c.facet_normals(sym_f))
    got = p.prs_res_surf(d[9], d[0:3], d[3:6], d[6:9], mu, rho, bd, dt, prs=sp, sym=ss)
    assert np.abs(_np(got) - ref).max() <= 1e-10 * np.abs(ref).max()
    # without surfaces it is nm_prs_res
    got0 = p.prs_res_surf(d[9], d[0:3], None, d[6:9], mu, rho, bd, dt)
    ref0 = p.prs_res(d[9], d[0:3], d[6:9], mu, rho)
    assert torch.equal(got0, ref0)
