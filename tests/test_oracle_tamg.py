"""Pins of the oracle's tree-AMG restatement (oracle/neko_oracle_tamg.c),
each independent of the C code's own formulas:

* aggregation: every element in exactly one level-2 aggregate, every
  aggregate connected through shared faces, the finest -> level maps
  composed from the per-level parents, one aggregate on the last level,
  the level count rule (stop when nodes / 8 < 4);
* the level operators (applied through the finest level) equal the
  Galerkin products P^T A P of the assembled, masked finest operator A on
  the unique dofs, with P the averaging prolongation of each level's
  aggregates (dense, numpy);
* the Chebyshev smoother from zero leaves the residual T_m((theta -
  lambda) / delta) / T_m(theta / delta) f in the eigenbasis of the
  (symmetric) coarse level operator;
* the power estimate sits at 1.1 lambda_max / (lambda_max / 30) of the
  dense level operator's spectrum (to the power method's accuracy);
* the V-cycle equals its composition from these pieces;
* as phmg's coarse solve, the AMG keeps GMRES + phmg converging in a few
  iterations, close to the count with hsmg's exact-enough coarse CG.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _case(dims=(4, 4, 3), lx=2, per=(0, 0, 0), amp=0.1):
    return O.OracleCase(*dims, lx, amp=amp, seed=7, periodic=per)


def _unique(c):
    ids, rep, q = np.unique(c.dof, return_index=True, return_inverse=True)
    masked = np.zeros(c.n, dtype=bool)
    masked[c.msk[: c.nmsk] - 1] = True
    mu = np.bincount(q, weights=masked.astype(float), minlength=len(ids)) > 0
    return rep, q, mu


@pytest.mark.parametrize("dims,per", [((4, 4, 3), (0, 0, 0)), ((5, 4, 4), (1, 0, 0)),
                                      ((8, 8, 8), (0, 0, 0))])
def test_aggregation_partitions_and_connects(dims, per):
    c = _case(dims, per=per)
    t = O.OracleTamg(c)
    nel = c.nel
    assert t.nn[0] == c.n and t.nn[1] == nel and t.nn[-1] == 1
    m1 = t.map(1)
    assert np.array_equal(m1, np.arange(c.n) // 8)
    if t.nlvls >= 3:
        m2 = t.map(2)
        # the level-2 aggregate of an element is the same for all its dofs
        parent = m2[::8]
        assert np.array_equal(m2, np.repeat(parent, 8))
        assert set(parent) == set(range(int(t.nn[2])))
        fn = c.fneigh.reshape(nel, 6)
        for a in range(int(t.nn[2])):
            mem = set(np.nonzero(parent == a)[0])
            seen, stack = set(), [min(mem)]
            while stack:
                e = stack.pop()
                if e in seen:
                    continue
                seen.add(e)
                stack += [int(v) - 1 for v in fn[e] if v > 0 and (int(v) - 1) in mem]
            assert seen == mem, a
    # level count: stop once nodes / 8 < 4
    for l in range(2, t.nlvls):
        assert t.nn[l - 1] // 8 >= 4
    assert t.nlvls == 3 or t.nn[t.nlvls - 1] // 8 < 4


def _galerkin(c, t, lvl):
    """P^T A P on the unique dofs: A assembled and masked (zero rows and
    columns at masked dofs), P averaging: P[u, a] = (copies of u in
    aggregate a's elements) / (copies of u), masked rows zero."""
    rep, q, mu = _unique(c)
    nu = len(rep)
    A = np.zeros((nu, nu))
    for u in range(nu):
        if mu[u]:
            continue
        e = (q == u).astype(float)
        A[:, u] = np.bincount(q, weights=c.ax_full(e), minlength=nu) / np.bincount(q, minlength=nu)
    A[mu, :] = 0.0
    m = t.map(lvl)
    na = int(t.nn[lvl])
    cnt = np.bincount(q, minlength=nu).astype(float)
    P = np.zeros((nu, na))
    np.add.at(P, (q, m), 1.0)
    P = P / cnt[:, None]
    P[mu, :] = 0.0
    return P.T @ A @ P


@pytest.mark.parametrize("lvl", [1, 2])
def test_level_operator_is_galerkin_product(lvl):
    c = _case((4, 4, 3))
    t = O.OracleTamg(c)
    G = _galerkin(c, t, lvl)
    na = int(t.nn[lvl])
    M = np.array([t.matvec(np.eye(na)[k], lvl) for k in range(na)]).T
    assert np.max(np.abs(M - G)) < 1e-12 * np.max(np.abs(G))
    assert np.max(np.abs(M - M.T)) < 1e-12 * np.max(np.abs(M))


def test_chebyshev_residual_polynomial_and_power_estimate():
    c = _case((5, 4, 4))
    t = O.OracleTamg(c, cheby_degree=4)
    lvl = 1
    na = int(t.nn[lvl])
    M = np.array([t.matvec(np.eye(na)[k], lvl) for k in range(na)]).T
    lam, V = np.linalg.eigh(0.5 * (M + M.T))
    tha, dlt = t.eigs()
    lmax = lam[-1]
    # power estimate (250 iterations) within 1e-3 of the largest eigenvalue
    b_est = tha[lvl] + dlt[lvl]
    a_est = tha[lvl] - dlt[lvl]
    assert abs(b_est / 1.1 - lmax) < 1e-3 * lmax
    assert abs(a_est * 30.0 - b_est / 1.1) < 1e-12 * lmax
    f = np.random.default_rng(1).standard_normal(na)
    x = t.cheby(f, lvl)
    r = f - M @ x
    m = 4
    th, de = tha[lvl], dlt[lvl]
    def Tm(z):
        z = np.asarray(z, dtype=float)
        out = np.empty_like(z)
        inside = np.abs(z) <= 1
        out[inside] = np.cos(m * np.arccos(z[inside]))
        zo = z[~inside]
        out[~inside] = np.sign(zo) ** m * np.cosh(m * np.arccosh(np.abs(zo)))
        return out
    p = Tm((th - lam) / de) / Tm(np.array([th / de]))[0]
    want = V @ (p * (V.T @ f))
    assert np.max(np.abs(r - want)) < 1e-10 * np.max(np.abs(f))


def test_vcycle_equals_its_composition():
    c = _case((5, 4, 4))
    t = O.OracleTamg(c, nlvls=3, max_iter=1)
    assert t.nlvls == 3
    mult = c.mult
    m1, m2 = t.map(1), t.map(2)
    par2 = m2[::8]  # element -> level-2 aggregate

    def restrict1(r0):
        return np.bincount(m1, weights=r0 * mult, minlength=int(t.nn[1]))

    def restrict2(r1):
        return np.bincount(par2, weights=r1, minlength=int(t.nn[2]))

    def prolong1(x1):
        out = x1[m1].copy()
        c.gs(out)
        out = out * mult
        return c.mask_apply(out)

    b0 = c.rhs_normal(3)
    x0 = t.cheby(b0, 0)
    r0 = b0 - t.matvec(x0, 0)
    b1 = restrict1(r0)
    x1 = t.cheby(b1, 1)
    r1 = b1 - t.matvec(x1, 1)
    b2 = restrict2(r1)
    x2 = t.cheby(b2, 2)
    x1 = x1 + x2[par2]
    x1 = t.cheby(b1, 1, x0=x1)
    x0 = x0 + prolong1(x1)
    x0 = t.cheby(b0, 0, x0=x0)
    z = t.solve(b0)
    assert np.max(np.abs(z - x0)) < 1e-13 * np.max(np.abs(x0))


def test_amg_as_phmg_coarse_solve_in_gmres():
    c = O.OracleCase(4, 4, 4, 6, amp=0.1, seed=9, h1=1.0, h2=0.0)
    f = c.rhs_normal(4)
    tol = 1e-10 * c.res_norm(f)
    ph_amg = O.OraclePhmg(c, coarse_amg=(3, 1, 4))
    x, m = ph_amg.gmres(f, tol)
    ph_cg = O.OraclePhmg(c)
    x2, m2 = ph_cg.gmres(f, tol)
    assert m.converged and m2.converged
    assert m.iter <= m2.iter + 6, (m.iter, m2.iter)
    assert np.linalg.norm(x - x2) < 1e-8 * np.linalg.norm(x2)
