"""GPU parity on the reference's own meshes (tests/golden/nmsh): problems
built by nm_op_create_mesh from .nmsh files -- geometry, Ax+gs+mask, the
Poisson example's fixed-iteration plain CG (examples/poisson/driver.f90),
Jacobi-PCG to 1e-10 in FP64 and mixed precision -- against the oracle on
the same mesh."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "nmsh")


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _pair(name, lx):
    from neko_amd.solve import Problem, read_nmsh
    path = os.path.join(GOLD, name)
    v, p, _ = read_nmsh(path)
    return O.OracleCase.from_mesh(v, p, lx), Problem.mesh(v, p, lx)


@pytest.mark.parametrize("name,lx", [("poisson_128.nmsh", 8), ("hemi.nmsh", 5),
                                     ("lid.nmsh", 6), ("tgv_512.nmsh", 4)])
def test_mesh_geometry_and_ax_match_oracle(name, lx):
    print("mesh ax", name, lx, flush=True)
    c, p = _pair(name, lx)
    assert p.n == c.n and p.nel == c.nel
    assert abs(p.volume - c.volume) <= 1e-12 * c.volume
    for which, ref in [(6, c.B), (7, c.mult)]:
        got = _np(p.field(which))
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    u = np.random.default_rng(7).standard_normal(c.n)
    ref = c.ax_full(u)
    ud = torch.from_numpy(u).cuda()
    for code in (0, 201):  # stored G / G from the element edges in-kernel
        p.set_epb(code, code)
        w = torch.zeros_like(ud)
        p.ax_full(w, ud)
        err = np.max(np.abs(_np(w) - ref)) / np.max(np.abs(ref))
        assert err < 1e-12, (code, err)


def test_poisson_example_fixed_iterations():
    """driver.f90: plain CG (no preconditioner), abstol -1, niter iterations
    from x = 0 on the example's 128-element mesh with its Gaussian RHS."""
    from neko_amd.solve import CG
    print("poisson example", flush=True)
    c, p = _pair("poisson_128.nmsh", 8)
    f = c.rhs_poisson_example()
    niter = 100
    x_or, mon_or = c.cg(f, -1.0, max_iter=niter, pc=np.ones(c.n))
    p.set_pc("identity")
    for code in (0, 201):
        p.set_epb(code, code)
        x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
        mon = CG(p, "fp64").solve(x, torch.from_numpy(f).cuda(), -1.0,
                                  max_iter=niter)
        # Fortran loop semantics: an exhausted run reports niter + 1
        assert mon.iter == mon_or.iter == niter + 1
        assert not mon.converged
        assert abs(mon.res_start - mon_or.res_start) <= 1e-12 * mon_or.res_start
        assert abs(mon.res_final - mon_or.res_final) <= 1e-8 * mon_or.res_final
        rel = np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or)
        assert rel < 1e-10, (code, rel)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hemi_jacobi_pcg_matches_oracle(precision):
    from neko_amd.solve import CG
    print("hemi pcg", precision, flush=True)
    c, p = _pair("hemi.nmsh", 5)
    f = c.rhs_normal(11)
    tol = 1e-10 * c.res_norm(f)
    x_or, mon_or = c.cg(f, tol, max_iter=5000)
    assert mon_or.converged
    p.set_epb(201, 201)
    # the fused update takes the face sums, square symmetries included
    assert p.info.cg_face_sums == 1 and p.info.mult_counts == 1
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    mon = CG(p, precision).solve(x, torch.from_numpy(f).cuda(), tol, max_iter=5000)
    assert mon.converged
    if precision == "fp64":
        assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
    rel = np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or)
    assert rel < 1e-9, rel


def test_pc_switch_is_exact():
    """Jacobi -> identity -> Jacobi gives the first Jacobi solve again, bit
    for bit; identity matches the oracle's plain CG iteration count."""
    from neko_amd.solve import CG
    print("pc switch", flush=True)
    c, p = _pair("lid.nmsh", 6)
    f = c.rhs_normal(5)
    tol = 1e-8 * c.res_norm(f)
    fd = torch.from_numpy(f).cuda()

    def run():
        x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
        mon = CG(p, "fp64").solve(x, fd, tol, max_iter=5000)
        return x, mon

    xj, mj = run()
    p.set_pc("identity")
    xi, mi = run()
    _, mon_or = c.cg(f, tol, max_iter=5000, pc=np.ones(c.n))
    assert abs(mi.iter - mon_or.iter) <= 1, (mi.iter, mon_or.iter)
    p.set_pc("jacobi")
    xj2, mj2 = run()
    assert mj2.iter == mj.iter and torch.equal(xj2, xj)
    assert mi.iter != mj.iter


@pytest.mark.parametrize("name", ["poisson_128.nmsh", "hemi.nmsh", "lid.nmsh",
                                  "tgv_512.nmsh"])
def test_mesh_gs_variants_bit_exact(name):
    """All gs variants, including whole-face descriptors with the square
    symmetries an unstructured mesh needs, give the oracle's bits."""
    print("mesh gs", name, flush=True)
    c, p = _pair(name, 5)
    assert p.info.nrest == 0 and p.info.nfaces * 9 == p.info.npairs
    u = np.random.default_rng(4).standard_normal(c.n)
    ref = c.gs(u.copy())
    for v in (0, 1, 2, 3):
        p.set_gs_variant(v)
        ud = torch.from_numpy(u.copy()).cuda()
        p.gs_op(ud)
        assert np.array_equal(_np(ud), ref), v


def test_surfaces_from_reference_mesh_zones():
    """Facet-normal surfaces built from the hemisphere mesh's own labelled
    zones: each label's normals equal the oracle's; all labels together
    close the boundary, so the area-weighted normals sum to zero and the
    flux of x is 3 x the volume (straight-edged elements, exact GLL)."""
    from neko_amd.solve import Surface, read_nmsh_zones, zone_facets
    print("surfaces hemi", flush=True)
    c, p = _pair("hemi.nmsh", 5)
    zones = read_nmsh_zones(os.path.join(GOLD, "hemi.nmsh"))
    for label in sorted(set(int(l) for l in zones[zones[:, 2] == 7][:, 3])):
        fl = zone_facets(zones, label)
        pts, nx, ny, nz = Surface(p, fl).normals()
        rp, rx, ry, rz = c.facet_normals(fl)
        assert np.array_equal(pts.astype(np.int64), rp)
        scale = max(np.abs(rx).max(), np.abs(ry).max(), np.abs(rz).max())
        for g, r in ((nx, rx), (ny, ry), (nz, rz)):
            assert np.abs(g - r).max() <= 1e-12 * scale
    pts, nx, ny, nz = Surface(p, zone_facets(zones)).normals()
    area = np.sqrt(nx ** 2 + ny ** 2 + nz ** 2).sum()
    for comp in (nx, ny, nz):
        assert abs(comp.sum()) <= 1e-12 * area
    flux = (c.x[pts] * nx + c.y[pts] * ny + c.zc[pts] * nz).sum()
    assert abs(flux - 3.0 * c.B.sum()) <= 1e-11 * abs(flux)
