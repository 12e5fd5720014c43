"""Pins of the oracle's hsmg restatement (oracle/neko_oracle_hsmg.c) by
closed forms, generic finite-element assembly and library routines -- each
chosen so that a dropped term, a transposed operand or a wrong index fails:

* semhat: the 1-D stiffness/mass integrate monomials exactly;
* setup_intp: GLL-to-GLL interpolation reproduces polynomials;
* tnsr3d: prolongation of a separable polynomial is exact and restriction
  is its adjoint;
* fdm pencil: equals the assembled 1-D stiffness/mass of the three
  elements restricted to the extended element (with identity rows where a
  facet carries a Dirichlet condition);
* fdm eigenproblem: eigenvalues of scipy.linalg.eigh on the same pencil,
  B-orthonormal eigenvectors;
* plane_space: element lengths of an undeformed box;
* Schwarz/FDM on a single undeformed element with Dirichlet facets is the
  exact inverse of the assembled operator;
* hsmg inside GMRES cuts the iteration count by an order of magnitude.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 9])
def test_semhat_integrates_monomials(n):
    ah, bh = O.semhat(n)
    z = O.zwgll(n + 1)[0]
    for a in range(n + 1):
        for b in range(n + 1):
            p, q = z ** a, z ** b
            m = a + b - 2
            exact = 0.0 if (a == 0 or b == 0 or m % 2) else a * b * 2.0 / (m + 1)
            assert abs(p @ ah @ q - exact) < 1e-11 * max(1.0, abs(exact))
            if a + b <= 2 * n - 1:
                m2 = a + b
                ex2 = 0.0 if m2 % 2 else 2.0 / (m2 + 1)
                assert abs(np.sum(bh * p * q) - ex2) < 1e-13


@pytest.mark.parametrize("nt,nf", [(8, 4), (4, 2), (5, 4), (3, 2), (10, 4), (4, 3)])
def test_setup_intp_reproduces_polynomials(nt, nf):
    zt, zf = O.zwgll(nt)[0], O.zwgll(nf)[0]
    jh, jht = O.setup_intp(zt, zf)
    J = jh.reshape(nf, nt).T
    assert np.array_equal(jht.reshape(nt, nf).T, J.T)
    for k in range(nf):
        assert np.max(np.abs(J @ zf ** k - zt ** k)) < 1e-13


def test_tnsr3d_prolongation_exact_and_restriction_adjoint():
    nf, nm, nel = 8, 4, 3
    zf, zm = O.zwgll(nf)[0], O.zwgll(nm)[0]
    jh, _ = O.setup_intp(zf, zm)
    J = jh.reshape(nm, nf).T                      # nf x nm
    P = lambda x: 1 + 2 * x - x ** 3              # distinct per direction
    Q = lambda y: 0.5 - y ** 2
    R = lambda z: z + 3 * z ** 2
    u = np.concatenate([(k + 1) * np.einsum("i,j,k->kji", P(zm), Q(zm), R(zm)).ravel()
                        for k in range(nel)])
    v = O.tnsr3d(u, nm, nf, J, J.T, J.T, nel)
    for k in range(nel):
        ex = (k + 1) * np.einsum("i,j,k->kji", P(zf), Q(zf), R(zf)).ravel()
        assert np.max(np.abs(v[k * nf ** 3:(k + 1) * nf ** 3] - ex)) < 1e-12
    rng = np.random.default_rng(1)
    a = rng.standard_normal(nel * nf ** 3)
    b = rng.standard_normal(nel * nm ** 3)
    ra = O.tnsr3d(a, nf, nm, J.T, J, J, nel)          # restriction
    pb = O.tnsr3d(b, nm, nf, J, J.T, J.T, nel)        # prolongation
    assert abs(ra @ b - a @ pb) < 1e-12 * np.linalg.norm(a) * np.linalg.norm(b)


def _assembled_1d(n, lengths):
    """Global 1-D stiffness and (diagonal) mass of consecutive elements."""
    ah, bh = O.semhat(n)
    ne = len(lengths)
    m = ne * n + 1
    K = np.zeros((m, m))
    M = np.zeros((m, m))
    for e, L in enumerate(lengths):
        sl = slice(e * n, e * n + n + 1)
        K[sl, sl] += (2.0 / L) * ah
        M[sl, sl] += (0.5 * L) * np.diag(bh)
    return K, M


@pytest.mark.parametrize("n", [2, 3, 5, 7])
@pytest.mark.parametrize("lbc,rbc", [(0, 0), (1, 0), (0, 1), (2, 0), (0, 2), (1, 2), (1, 1)])
def test_fdm_pencil_is_restricted_assembly(n, lbc, rbc):
    ll, lm, lr = 0.3, 0.45, 0.2
    A, B, S, St, lam = O.fdm_1d(lbc, rbc, ll, lm, lr, n)
    # elements present: left neighbour (lbc == 0), the element, right (rbc == 0)
    lengths = ([ll] if lbc == 0 else []) + [lm] + ([lr] if rbc == 0 else [])
    K, M = _assembled_1d(n, lengths)
    off = n if lbc == 0 else 0
    # extended-element points: one point inside each neighbour
    idx = [off - 1] + list(range(off, off + n + 1)) + [off + n + 1]
    Ke = np.zeros((n + 3, n + 3))
    Me = np.zeros((n + 3, n + 3))
    valid = [0 <= q < K.shape[0] for q in idx]
    for a_, qa in enumerate(idx):
        for b_, qb in enumerate(idx):
            if valid[a_] and valid[b_]:
                Ke[a_, b_], Me[a_, b_] = K[qa, qb], M[qa, qb]
    ident = [not v for v in valid]
    if lbc == 1:
        ident[1] = True          # Dirichlet facet: the element's own end point
    if rbc == 1:
        ident[n + 1] = True
    for q, d in enumerate(ident):
        if d:
            Ke[q, :] = Ke[:, q] = 0.0
            Me[q, :] = Me[:, q] = 0.0
            Ke[q, q] = Me[q, q] = 1.0
    assert np.max(np.abs(A - Ke)) < 1e-12 * np.max(np.abs(Ke))
    assert np.max(np.abs(B - Me)) < 1e-15
    # eigenproblem: scipy's generalized symmetric solver on the same pencil
    ev = scipy.linalg.eigh(A, B, eigvals_only=True)
    assert np.max(np.abs(lam - ev)) < 1e-11 * np.max(np.abs(ev))
    zero = [lbc > 0, lbc == 1] + [False] * (n - 1) + [rbc == 1, rbc > 0]
    keep = [not zz for zz in zero]
    assert np.all(S[zero] == 0.0)
    assert np.array_equal(St, S.T)
    k = np.array(keep)
    resid = A[np.ix_(k, k)] @ S[k] - B[np.ix_(k, k)] @ S[k] * lam
    assert np.max(np.abs(resid)) < 1e-10 * np.max(np.abs(A))
    G = S[k].T @ B[np.ix_(k, k)] @ S[k]
    d = np.diag(G)
    assert np.max(np.abs(G - np.diag(d))) < 1e-12
    assert np.all((np.abs(d - 1) < 1e-12) | (np.abs(d) < 1e-12))
    assert int(np.sum(np.abs(d - 1) < 1e-12)) == int(k.sum())


def test_plane_space_undeformed_lengths():
    c = O.OracleCase(3, 2, 4, 6, amp=0.0)
    lr, ls, lt = O.plane_space(c)
    assert np.max(np.abs(lr - 1 / 3)) < 1e-14
    assert np.max(np.abs(ls - 1 / 2)) < 1e-14
    assert np.max(np.abs(lt - 1 / 4)) < 1e-14


@pytest.mark.parametrize("lx", [4, 6, 8])
def test_schwarz_single_element_is_exact_inverse(lx):
    """One undeformed element, every facet Dirichlet (facet_type 1): the FDM
    solve is the exact inverse of the masked assembled operator."""
    c = O.OracleCase(1, 1, 1, lx, amp=0.0)
    H = O.OracleHsmg(c, facet_bc=1)
    r = c.rhs_normal(3)
    e = H.schwarz_fine(r)
    ae = c.ax_full(e)
    assert np.max(np.abs(ae - r)) < 1e-10 * np.max(np.abs(r))


def test_hsmg_accelerates_gmres():
    c = O.OracleCase(3, 3, 3, 8, amp=0.1, seed=7)
    f = c.rhs_normal(5)
    tol = 1e-10 * c.res_norm(f)
    H = O.OracleHsmg(c, facet_bc=1)
    x_h, m_h = H.gmres(f, tol)
    x_j, m_j = c.gmres(f, tol, pc=c.jacobi())
    assert m_h.converged and m_j.converged
    assert m_h.iter * 10 < m_j.iter
    assert np.linalg.norm(x_h - x_j) < 1e-8 * np.linalg.norm(x_j)
