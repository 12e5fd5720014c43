"""rotate_cyc on the GPU (nm_cyclic_create / nm_rotate_cyc and the drop-in
hip_rotate_cyc) against the oracle (coef_cyclic, src/sem/coef.f90:
1380-1430; opr_cpu_rotate_cyc_r1, src/math/bcknd/cpu/opr_cpu.f90:333-357)
on a wedge whose sides are the cyclic facets, and on a rotated box whose
corner element has two adjacent cyclic facets (points listed twice,
rotated twice in list order): the point list bit for bit, R11/R12 and
the rotated fields to 1e-12 (the normals are derivatives of the device
coordinates), the drop-in on the oracle's arrays within an ulp or two."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O

from test_oracle_cyclic import wedge_case

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _problem(c):
    from neko_amd.solve import Problem
    return Problem.mesh(c.vxyz.reshape(-1, 8, 3), c.pid.reshape(-1, 8), c.lx)


def _check(c, facets):
    from neko_amd.solve import Cyclic
    p = _problem(c)
    cy = Cyclic(p, facets)
    msk, r11, r12 = cy.info()
    setup = c.cyclic_setup(facets)
    assert np.array_equal(msk.astype(np.int64), setup[0])
    # the normals come from the device coordinates' derivatives (the
    # surface normals' bar): 1e-12
    assert np.max(np.abs(r11 - setup[1])) < 1e-12
    assert np.max(np.abs(r12 - setup[2])) < 1e-12
    rng = np.random.default_rng(7)
    vx0, vy0, vz0 = (rng.standard_normal(c.n) for _ in range(3))
    for idir in (1, 0):
        ox, oy = vx0.copy(), vy0.copy()
        O.OracleCase.rotate_cyc(ox, oy, idir, setup)
        gx, gy, gz = _t(vx0), _t(vy0), _t(vz0)
        cy.rotate(gx, gy, gz, idir)
        assert np.max(np.abs(_np(gx) - ox)) < 1e-12 * np.max(np.abs(ox))
        assert np.max(np.abs(_np(gy) - oy)) < 1e-12 * np.max(np.abs(oy))
        assert np.array_equal(_np(gz), vz0)
    cy.free()
    return setup


@pytest.mark.parametrize("alpha,lx", [(0.5, 5), (1.2, 8)])
def test_wedge_matches_oracle(alpha, lx):
    print("cyclic wedge", alpha, lx, flush=True)
    c, facets = wedge_case(2, 3, 2, lx, alpha)
    setup = _check(c, facets)
    # the drop-in with Neko's device arrays (cyc_msk(0) = ncyc + 1)
    from neko_amd._lib import lib
    ncyc = len(setup[0])
    msk = np.concatenate([[ncyc + 1], setup[0]]).astype(np.int32)
    rng = np.random.default_rng(9)
    vx0, vy0 = rng.standard_normal(c.n), rng.standard_normal(c.n)
    gx, gy, gz = _t(vx0), _t(vy0), _t(np.zeros(c.n))
    gm, g11, g12 = _t(msk), _t(setup[1]), _t(setup[2])
    x = _t(c.x)
    torch.cuda.synchronize()
    lib().hip_rotate_cyc(gx.data_ptr(), gy.data_ptr(), gz.data_ptr(), x.data_ptr(),
                         x.data_ptr(), x.data_ptr(), gm.data_ptr(), g11.data_ptr(),
                         g12.data_ptr(), C.byref(C.c_int(ncyc)), C.byref(C.c_int(1)))
    torch.cuda.synchronize()
    ox, oy = vx0.copy(), vy0.copy()
    O.OracleCase.rotate_cyc(ox, oy, 1, setup)
    # the drop-in takes the oracle's R (fused multiply-adds on the device:
    # within an ulp or two)
    assert np.max(np.abs(_np(gx) - ox)) < 1e-15 * 8 * np.max(np.abs(ox))
    assert np.max(np.abs(_np(gy) - oy)) < 1e-15 * 8 * np.max(np.abs(oy))


def test_adjacent_cyclic_facets_rotate_in_list_order():
    print("cyclic corner", flush=True)
    beta = 0.6
    c0 = O.OracleCase(3, 2, 2, 2)
    v = c0.vxyz.reshape(-1, 8, 3).copy()
    x, y = v[:, :, 0] + 0.5, v[:, :, 1] + 1.5
    v[:, :, 0] = np.cos(beta) * x - np.sin(beta) * y
    v[:, :, 1] = np.sin(beta) * x + np.cos(beta) * y
    c = O.OracleCase.from_mesh(v, c0.pid.reshape(-1, 8), 4)
    facets = [(0, 1), (0, 3), (3, 1)]  # element 0: two adjacent cyclic facets
    setup = _check(c, facets)
    assert len(np.unique(setup[0])) < len(setup[0])
