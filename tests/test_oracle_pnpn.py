"""The oracle's pnpn operator pieces pinned by the mathematics: a constant
field has zero weak gradient; for a linear field a.x the weak gradient is
B a exactly (the cofactors are jac times the inverse Jacobian, so w3 C^T
grad_r u = w3 jac grad_x u = B grad_x u); the cofactors equal the closed
form cross products of the trilinear map's derivatives; the velocity
residual with u = v = w = 0 and p = 0 is f; its h2 term is h2 B u for a
constant u (the stiffness part annihilates constants); dudxyz of a linear
field is its slope; cdtp is the adjoint of opgrad; the curl of a gradient
vanishes and curl(-y, x, 0) = (0, 0, 2); the pressure residual of a
curl-free velocity with zero forcing is -A p."""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def case():
    return O.OracleCase(3, 2, 3, 6, amp=0.1, seed=5)


def test_constant_has_zero_gradient(case):
    g = case.opgrad(np.full(case.n, 2.5))
    for gc in g:
        assert np.abs(gc).max() < 1e-12


def test_linear_field_gradient_is_mass_times_slope(case):
    c = case
    a = (0.4, -1.1, 0.8)
    lin = a[0] * c.x + a[1] * c.y + a[2] * c.zc + 0.3
    g = c.opgrad(lin)
    for gc, ac in zip(g, a):
        assert np.abs(gc - c.B * ac).max() <= 1e-13 * np.abs(c.B).max()


def test_cofactors_match_closed_form(case):
    c = case
    cof = c.cofactors()
    # closed form from the vertex map: derivatives of the trilinear map
    v = c.vxyz.reshape(c.nel, 8, 3)
    z = c.z
    lx = c.lx
    for e in (0, c.nel - 1):
        for (i, j, k) in [(0, 0, 0), (lx - 1, 2, 1), (3, lx - 1, lx - 1)]:
            r, s, t = z[i], z[j], z[k]
            N = lambda q, xi: (1 - xi) / 2 if q == 0 else (1 + xi) / 2  # noqa: E731
            dN = lambda q: -0.5 if q == 0 else 0.5  # noqa: E731
            Xr = sum(dN(cx) * N(cy, s) * N(cz, t) * v[e, cx + 2 * cy + 4 * cz]
                     for cx in (0, 1) for cy in (0, 1) for cz in (0, 1))
            Xs = sum(N(cx, r) * dN(cy) * N(cz, t) * v[e, cx + 2 * cy + 4 * cz]
                     for cx in (0, 1) for cy in (0, 1) for cz in (0, 1))
            Xt = sum(N(cx, r) * N(cy, s) * dN(cz) * v[e, cx + 2 * cy + 4 * cz]
                     for cx in (0, 1) for cy in (0, 1) for cz in (0, 1))
            q = e * lx ** 3 + i + lx * (j + lx * k)
            ref = np.concatenate([np.cross(Xs, Xt), np.cross(Xt, Xr), np.cross(Xr, Xs)])
            assert np.abs(cof[:, q] - ref).max() < 1e-13


def test_vel_res_special_cases(case):
    c = case
    rng = np.random.default_rng(4)
    f = [rng.standard_normal(c.n) for _ in range(3)]
    zero = np.zeros(c.n)
    res = c.vel_res(zero, zero, zero, zero, f, mu=0.01, rho=1.0, bd=1.5, dt=1e-3)
    for rc, fc in zip(res, f):
        assert np.array_equal(rc, fc)
    one = np.full(c.n, 2.0)
    res = c.vel_res(one, one, one, zero, [zero] * 3, mu=0.01, rho=2.0, bd=1.5, dt=1e-3)
    h2 = 2.0 * 1.5 / 1e-3
    for rc in res:
        assert np.abs(rc + h2 * c.B * 2.0).max() <= 1e-10 * h2 * np.abs(c.B).max()


def test_dudxyz_linear_and_cdtp_adjoint(case):
    c = case
    a = (0.3, -1.2, 0.7)
    lin = a[0] * c.x + a[1] * c.y + a[2] * c.zc
    for k in range(3):
        assert np.abs(c.dudxyz(lin, k) - a[k]).max() < 1e-12
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal(c.n), rng.standard_normal(c.n)
    for k in range(3):
        lhs = np.dot(c.cdtp(x, k), y)
        rhs = np.dot(x, c.opgrad(y)[k])
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1.0)


def test_curl_closed_forms(case):
    c = case
    zero = np.zeros(c.n)
    w = c.curl(-c.y, c.x, zero)
    for q, v in zip(w, (0.0, 0.0, 2.0)):
        assert np.abs(q - v).max() < 1e-12
    g = (2 * c.x + c.y, c.x + 3 * c.zc, 3 * c.y)   # grad(x^2 + x y + 3 y z)
    for q in c.curl(*g):
        assert np.abs(q).max() < 1e-12


def test_prs_res_of_curl_free_field(case):
    c = case
    zero = np.zeros(c.n)
    g = (2 * c.x + c.y, c.x + 3 * c.zc, 3 * c.y)
    p = np.random.default_rng(6).standard_normal(c.n)
    r = c.prs_res(p, g, (zero, zero, zero), mu=0.01, rho=2.0)
    ref = -c.ax_local_coeffs(p, 0.5, 0.0)
    assert np.abs(r - ref).max() < 1e-10 * np.abs(ref).max()
