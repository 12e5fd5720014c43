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


# --- facet normals and the pressure residual's surface terms ----------------
ALL_SIDES = (1, 2, 3, 4, 5, 6)


def _sides(c, dims, sides):
    out = []
    for s in sides:
        out += c.box_facets(dims, s)
    return out


def test_facet_normals_of_undeformed_box_are_axis_aligned():
    """Undeformed unit box: on side s every point's area-weighted normal is
    along the side's axis with the side's sign, the weight is the GLL
    quadrature weight of the face (h_a h_b / 4 w_a w_b), and the weights of
    a side sum to its area 1."""
    dims = (3, 2, 4)
    c = O.OracleCase(*dims, 5, amp=0.0, seed=1)
    h = [1.0 / d for d in dims]
    for s in ALL_SIDES:
        pts, nx, ny, nz = c.facet_normals(c.box_facets(dims, s))
        n = np.stack([nx, ny, nz])
        ax, sign = (s - 1) // 2, (-1.0 if s % 2 == 1 else 1.0)
        others = [a for a in range(3) if a != ax]
        assert np.abs(n[others]).max() < 1e-15
        assert np.all(sign * n[ax] > 0)
        assert abs(sign * n[ax].sum() - 1.0) < 1e-13
        # one facet's corner point: w_0 w_0 h_a h_b / 4
        w0 = c.w[0]
        corner = w0 * w0 * h[others[0]] * h[others[1]] / 4.0
        assert abs(sign * n[ax][0] - corner) < 1e-15


def test_closed_surface_normals_sum_to_zero():
    """Deformed box: the area-weighted normals over all six sides integrate
    n over a closed (bilinear-faced) surface, which vanishes; GLL is exact
    for the faces' quadratic integrands."""
    dims = (3, 2, 3)
    c = O.OracleCase(*dims, 6, amp=0.1, seed=5)
    pts, nx, ny, nz = c.facet_normals(_sides(c, dims, ALL_SIDES))
    for comp in (nx, ny, nz):
        assert abs(comp.sum()) < 1e-13


def test_divergence_theorem_on_deformed_box():
    """Sum over the boundary of (x nx + y ny + z nz) A = integral of div(x)
    = 3 volume, and of y nx A = integral of d(y)/dx = 0 (trilinear
    elements, exact quadrature)."""
    dims = (3, 2, 3)
    c = O.OracleCase(*dims, 6, amp=0.1, seed=5)
    pts, nx, ny, nz = c.facet_normals(_sides(c, dims, ALL_SIDES))
    flux = (c.x[pts] * nx + c.y[pts] * ny + c.zc[pts] * nz).sum()
    assert abs(flux - 3.0 * c.B.sum()) < 1e-12
    assert abs((c.y[pts] * nx).sum()) < 1e-13


def test_prs_res_surface_terms_closed_forms():
    """With no surfaces the whole residual is prs_res; a unit x-velocity
    through the x+ side (area 1) of the velocity-prescribed surface lowers
    the residual's sum by bd/dt; with u_e = 0 and f = rho B e_x the
    symmetry term's ta is e_x, so a symmetry surface on the x+ side lowers
    the sum by 1."""
    dims = (3, 2, 3)
    c = O.OracleCase(*dims, 6, amp=0.1, seed=5)
    rng = np.random.default_rng(3)
    p = rng.standard_normal(c.n)
    ue = [rng.standard_normal(c.n) for _ in range(3)]
    f = [rng.standard_normal(c.n) for _ in range(3)]
    mu, rho, bd, dt = 1e-3, 1.3, 1.5, 1e-2
    base = c.prs_res(p, ue, f, mu, rho)
    assert np.array_equal(c.prs_res_surf(p, ue, None, f, mu, rho, bd, dt), base)
    surf = c.facet_normals(c.box_facets(dims, 2))
    one = [np.ones(c.n), np.zeros(c.n), np.zeros(c.n)]
    d = c.prs_res_surf(p, ue, one, f, mu, rho, bd, dt, prs=surf) - base
    assert abs(d.sum() + bd / dt) < 1e-11 * (bd / dt)
    z = [np.zeros(c.n)] * 3
    fb = [rho * c.B, np.zeros(c.n), np.zeros(c.n)]
    base0 = c.prs_res(p, z, fb, mu, rho)
    d = c.prs_res_surf(p, z, None, fb, mu, rho, bd, dt, sym=surf) - base0
    assert abs(d.sum() + 1.0) < 1e-12
