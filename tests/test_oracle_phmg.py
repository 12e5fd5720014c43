"""Pins of the oracle's Chebyshev smoother and phmg-style V-cycle
(oracle/neko_oracle_phmg.c) by dense linear algebra on tiny meshes:

* the Chebyshev iteration (cheby_impl) reproduces the closed-form
  Chebyshev residual polynomial e_k = T_k((tha - M^-1 A)/dlt) /
  T_k(tha/dlt) e_0 -- a wrong recurrence coefficient, sign or ordering
  fails it;
* the power estimate (cheby_power) matches the largest eigenvalue of the
  dense preconditioned operator, with Jacobi and with Schwarz;
* the Schwarz smoother maps continuous vectors to continuous vectors (so
  the dense operator on unique dofs is well defined);
* the V-cycle cuts GMRES iterations by an order of magnitude on Poisson.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _unique_basis(c):
    msk = np.zeros(c.n, bool)
    msk[c.msk[: c.nmsk] - 1] = True
    ids = np.unique(c.dof[~msk])
    reps = [np.where(c.dof == g)[0] for g in ids]
    return ids, reps


def _dense(c, op):
    """Matrix of a linear map on continuous masked vectors in the basis of
    the unique unmasked dofs (the result sampled at every copy is checked
    to be continuous)."""
    ids, reps = _unique_basis(c)
    N = len(ids)
    M = np.zeros((N, N))
    for k in range(N):
        u = np.zeros(c.n)
        u[reps[k]] = 1.0
        y = op(u)
        for q, r in enumerate(reps):
            assert np.ptp(y[r]) <= 1e-12 * max(1.0, np.max(np.abs(y))), "not continuous"
            M[q, k] = y[r[0]]
    return M, reps


def _to_local(c, reps, v):
    u = np.zeros(c.n)
    for k, r in enumerate(reps):
        u[r] = v[k]
    return u


def _tk(k, z):
    z = np.asarray(z, complex)
    return np.cosh(k * np.arccosh(z))


@pytest.mark.parametrize("lx,h2", [(4, 0.0), (5, 50.0)])
def test_chebyshev_is_the_chebyshev_polynomial(lx, h2):
    c = O.OracleCase(2, 2, 2, lx, amp=0.1, seed=3, h1=1.0, h2=h2)
    A, reps = _dense(c, c.ax_full)
    assert np.max(np.abs(A - A.T)) < 1e-12 * np.max(np.abs(A))
    dinv = c.jacobi()[[r[0] for r in reps]]
    MA = dinv[:, None] * A
    ch = O.OracleCheby(c, None)
    tha, dlt = ch.tha, ch.dlt
    lam = np.linalg.eigvals(MA).real
    # power estimate: lambda_max of M^-1 A (b = 1.1 lam, a = lam / 30)
    assert abs((tha + dlt) / 1.1 - lam.max()) < 1e-6 * lam.max()
    # lambda_min = lambda_max / 30 (lam_factor)
    assert abs((tha - dlt) * 30 - (tha + dlt) / 1.1) < 1e-9 * tha
    xt = np.random.default_rng(1).standard_normal(len(reps))
    f = A @ xt
    w, V = np.linalg.eig(MA)
    c0 = np.linalg.solve(V, xt)
    for k in (1, 2, 3, 4, 7):
        x = ch.solve(_to_local(c, reps, f), k)
        e = xt - x[[r[0] for r in reps]]
        ek = (V @ (_tk(k, (tha - w) / dlt) / _tk(k, tha / dlt) * c0)).real
        assert np.max(np.abs(e - ek)) < 1e-12 * np.max(np.abs(xt)), k


def test_power_estimate_with_schwarz():
    c = O.OracleCase(2, 2, 2, 5, amp=0.1, seed=4)
    H = O.OracleHsmg(c, facet_bc=2)
    ch = O.OracleCheby(c, H.sw_f)

    def sa(u):
        return H.schwarz_fine(c.ax_full(u))

    MA, _ = _dense(c, sa)
    # weighted overlapping Schwarz is not symmetric: a few eigenvalues come
    # in complex pairs, the dominant one is real
    lam = np.linalg.eigvals(MA)
    top = lam[np.argmax(lam.real)]
    assert abs(top.imag) < 1e-10 * abs(top.real)
    assert abs((ch.tha + ch.dlt) / 1.1 - top.real) < 1e-4 * top.real


def test_phmg_accelerates_gmres():
    c = O.OracleCase(3, 3, 3, 8, amp=0.1, seed=7)
    f = c.rhs_normal(5)
    tol = 1e-10 * c.res_norm(f)
    P = O.OraclePhmg(c, facet_bc=1)
    x_p, m_p = P.gmres(f, tol)
    x_j, m_j = c.gmres(f, tol, pc=c.jacobi())
    assert m_p.converged and m_j.converged
    assert m_p.iter * 10 < m_j.iter
    assert np.linalg.norm(x_p - x_j) < 1e-8 * np.linalg.norm(x_j)
