"""The oracle's solution-space projection (OracleProjection, a restatement
of src/common/projection.f90's CPU path) pinned by what the mathematics
fixes: the kept basis is A-orthonormal in the mult-weighted inner product
with B = A X; project_on leaves a RHS W-orthogonal to the basis with
A xbar = b - b_proj, so xbar + A^-1 b_proj solves the original system; a
vector already in the span is dropped; the basis restarts at one vector
once L are held; pre/post are no-ops up to activ_step and dt_last_change
0 clears."""
import numpy as np
import pytest

from oracle import oracle as O


def _continuous(c, rng):
    x = rng.standard_normal(c.n)
    c.gs(x)
    x *= c.mult
    c.mask_apply(x)
    return x


@pytest.fixture(scope="module")
def case():
    return O.OracleCase(3, 3, 3, 5, amp=0.1, seed=3)


def test_basis_is_a_orthonormal_and_restarts(case):
    c = case
    rng = np.random.default_rng(0)
    pj = O.OracleProjection(c, 4, activ_step=0)
    ms = []
    for t in range(1, 11):
        pj.post(_continuous(c, rng), t)
        ms.append(pj.m)
        X, B = pj.xx[:pj.m], pj.bb[:pj.m]
        G = (X * c.mult) @ B.T
        assert np.abs(G - np.eye(pj.m)).max() < 1e-12
        for k in range(pj.m):
            ref = c.ax_full(X[k])
            assert np.abs(B[k] - ref).max() <= 1e-12 * np.abs(ref).max()
    assert ms == [1, 2, 3, 4, 1, 2, 3, 4, 1, 2]


def test_project_on_invariants(case):
    c = case
    rng = np.random.default_rng(1)
    pj = O.OracleProjection(c, 5, activ_step=0)
    for t in range(1, 4):
        pj.post(_continuous(c, rng), t)
    b0 = c.rhs_normal(9)
    b = b0.copy()
    pj.pre(b, 4)
    assert pj.proj_m == 3
    assert abs(pj.proj_res - c.res_norm(b0)) <= 1e-14 * c.res_norm(b0)
    for k in range(pj.m):
        assert abs(np.dot(pj.xx[k] * c.mult, b)) < 1e-12 * np.linalg.norm(b0)
    axb = c.ax_full(pj.xbar)
    assert np.abs(axb - (b0 - b)).max() < 1e-11 * np.abs(b0).max()
    # xbar + A^-1 b_proj solves A x = b0
    y, mon = c.cg(b, 1e-12 * c.res_norm(b0), max_iter=5000)
    x = y.copy()
    pj.post(x, 4)
    r = b0 - c.ax_full(x)
    assert c.res_norm(r) < 1e-10 * c.res_norm(b0)


def test_dependent_vector_is_dropped(case):
    c = case
    rng = np.random.default_rng(2)
    pj = O.OracleProjection(c, 6, activ_step=0)
    xs = [_continuous(c, rng) for _ in range(3)]
    for t, x in enumerate(xs, 1):
        pj.post(x.copy(), t)
    assert pj.m == 3
    # xbar of an empty projection is unused: post of a vector in the span
    pj.xbar[:] = 0.0
    comb = 0.3 * xs[0] - 1.7 * xs[1] + 0.2 * xs[2]
    pj.post(comb, 4)
    assert pj.m == 3


def test_gating_and_clear(case):
    c = case
    rng = np.random.default_rng(3)
    pj = O.OracleProjection(c, 3, activ_step=2)
    x = _continuous(c, rng)
    pj.post(x.copy(), 2)          # tstep <= activ_step: nothing
    assert pj.m == 0
    pj.post(x.copy(), 3)
    assert pj.m == 1
    b = c.rhs_normal(4)
    b1 = b.copy()
    pj.pre(b1, 3, dt_last_change=1)   # dt changed one step ago: not yet
    assert np.array_equal(b1, b)
    pj.pre(b1, 3, dt_last_change=0)   # dt changed now: basis cleared
    assert pj.m == 0
