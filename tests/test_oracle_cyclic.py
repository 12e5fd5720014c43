"""The oracle's cyclic rotation (coef_cyclic, src/sem/coef.f90:1380-1430,
and opr_cpu_rotate_cyc_r1, src/math/bcknd/cpu/opr_cpu.f90:333-357) pinned
by closed forms on a wedge -- an annular sector r in [1, 2], theta in [0,
alpha], z in [0, 1] whose two sides theta = 0 and theta = alpha are the
cyclic facets:
* the rotation of each point is the unit azimuthal direction -theta_hat on
  both sides (the sign rule turns the two outward normals into one frame);
* a radial field (x, y) becomes (0, r) and a rigid rotation (-y, x)
  becomes (-r, 0) on both sides;
* the rotation is orthonormal and rotating back restores the field;
and on a box rotated by an angle beta about z (cyclic sides x = const):
R = (cos beta, sin beta) up to the sign rule."""
import numpy as np
import pytest

from oracle import oracle as O


def wedge_case(nr, nt, nz, lx, alpha):
    c0 = O.OracleCase(nr, nt, nz, 2)  # unit-box lattice: only its topology is kept
    v = c0.vxyz.reshape(-1, 8, 3).copy()
    r = 1.0 + v[:, :, 0]
    th = alpha * v[:, :, 1]
    v[:, :, 0] = r * np.cos(th)
    v[:, :, 1] = r * np.sin(th)
    c = O.OracleCase.from_mesh(v, c0.pid.reshape(-1, 8), lx)
    facets = []
    for e in range(c.nel):
        ey = (e // nr) % nt
        if ey == 0:
            facets.append((e, 3))
        if ey == nt - 1:
            facets.append((e, 4))
    return c, facets


@pytest.mark.parametrize("alpha", [0.4, 1.1])
def test_wedge_frames_and_fields(alpha):
    c, facets = wedge_case(2, 3, 2, 5, alpha)
    msk, r11, r12 = c.cyclic_setup(facets)
    j = msk - 1
    x, y = c.x[j], c.y[j]
    r = np.hypot(x, y)
    th = np.arctan2(y, x)
    assert np.all((np.abs(th) < 1e-12) | (np.abs(th - alpha) < 1e-12))
    # R = -theta_hat = (sin th, -cos th)
    assert np.max(np.abs(r11 - np.sin(th))) < 1e-13
    assert np.max(np.abs(r12 + np.cos(th))) < 1e-13
    # fields
    vx, vy = c.x.copy(), c.y.copy()
    O.OracleCase.rotate_cyc(vx, vy, 1, (msk, r11, r12))
    assert np.max(np.abs(vx[j])) < 1e-13 and np.max(np.abs(vy[j] - r)) < 1e-13
    vx, vy = -c.y.copy(), c.x.copy()
    O.OracleCase.rotate_cyc(vx, vy, 1, (msk, r11, r12))
    assert np.max(np.abs(vx[j] + r)) < 1e-13 and np.max(np.abs(vy[j])) < 1e-13
    # points off the cyclic facets are untouched
    off = np.setdiff1d(np.arange(c.n), j)
    assert np.array_equal(vx[off], -c.y[off]) and np.array_equal(vy[off], c.x[off])


def test_orthonormal_and_round_trip():
    c, facets = wedge_case(2, 3, 2, 4, 0.7)
    setup = c.cyclic_setup(facets)
    _, r11, r12 = setup
    assert np.max(np.abs(r11 ** 2 + r12 ** 2 - 1.0)) < 1e-15
    assert len(np.unique(setup[0])) == len(setup[0])  # no point listed twice here
    rng = np.random.default_rng(3)
    vx0, vy0 = rng.standard_normal(c.n), rng.standard_normal(c.n)
    vx, vy = vx0.copy(), vy0.copy()
    O.OracleCase.rotate_cyc(vx, vy, 1, setup)
    assert np.max(np.abs(np.hypot(vx, vy) - np.hypot(vx0, vy0))) < 1e-14
    O.OracleCase.rotate_cyc(vx, vy, 0, setup)
    assert np.max(np.abs(vx - vx0)) < 1e-15 and np.max(np.abs(vy - vy0)) < 1e-15


@pytest.mark.parametrize("beta", [0.3, -0.8])
def test_rotated_box_sides(beta):
    c0 = O.OracleCase(3, 2, 2, 2)
    v = c0.vxyz.reshape(-1, 8, 3).copy()
    x, y = v[:, :, 0] + 0.5, v[:, :, 1] + 1.5  # away from the z axis
    v[:, :, 0] = np.cos(beta) * x - np.sin(beta) * y
    v[:, :, 1] = np.sin(beta) * x + np.cos(beta) * y
    c = O.OracleCase.from_mesh(v, c0.pid.reshape(-1, 8), 4)
    facets = [(e, 1) for e in range(c.nel) if e % 3 == 0] + \
             [(e, 2) for e in range(c.nel) if e % 3 == 2]
    msk, r11, r12 = c.cyclic_setup(facets)
    # normals +-(cos beta, sin beta): R is that direction up to the sign rule
    dot = r11 * np.cos(beta) + r12 * np.sin(beta)
    assert np.max(np.abs(np.abs(dot) - 1.0)) < 1e-13
    # and the sign rule: d = y n_x - x n_y >= 0 after the sign is applied
    j = msk - 1
    assert np.all(c.y[j] * r11 - c.x[j] * r12 > 0)
