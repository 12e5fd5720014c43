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
* Schwarz overlap weights: 1 / (number of overlap-1 extended subdomains
  holding the point), counted by brute force from the mesh topology, on
  deformed and periodic meshes;
* hsmg composition: with an exact coarse solve (8 coarse unknowns, 10 CG
  iterations) the V-cycle equals S_f + P (S_m + P_c A_c^-1 P_c^T) P^T,
  with numpy Lagrange interpolation for P, P_c and a dense inverse;
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


# ------------------------------------------------------------------------
# Schwarz overlap weights: 1 / (number of overlapping extended subdomains
# holding the point), counted by brute force from the mesh topology.
# ------------------------------------------------------------------------
def _face_points(lx, f, layer):
    """Local offsets of the points of element face f (0..5 = i0, i1, j0,
    j1, k0, k1) at distance `layer` from that face."""
    out = []
    d, side = f // 2, f % 2
    fix = layer if side == 0 else lx - 1 - layer
    for b in range(lx):
        for a in range(lx):
            ijk = [a, b]
            ijk.insert(d, fix)
            i, j, k = ijk
            out.append(i + lx * (j + lx * k))
    return np.array(out)


def _overlap_counts(c):
    """count[dof] = number of elements whose overlap-1 extended subdomain
    (its own points plus, across every shared face, the neighbour's layer
    of points next to that face) contains the point."""
    lx, n3 = c.lx, c.lx ** 3
    dof = c.dof
    count = {}
    face_ids = {}
    for e in range(c.nel):
        for f in range(6):
            face_ids[(e, f)] = frozenset(dof[e * n3 + _face_points(lx, f, 0)])
    for e in range(c.nel):
        members = set(dof[e * n3:(e + 1) * n3])
        for f in range(6):
            nb = int(c.fneigh[e * 6 + f]) - 1
            if nb < 0:
                continue
            shared = [g for g in range(6) if face_ids[(nb, g)] == face_ids[(e, f)]]
            assert len(shared) == 1
            members |= set(dof[nb * n3 + _face_points(lx, shared[0], 1)])
        for d in members:
            count[d] = count.get(d, 0) + 1
    return np.array([count[d] for d in dof], dtype=float)


@pytest.mark.parametrize("dims,lx,per,amp", [((3, 3, 3), 6, (0, 0, 0), 0.15),
                                             ((3, 4, 3), 5, (1, 0, 0), 0.1),
                                             ((3, 3, 4), 4, (0, 1, 1), 0.0)])
def test_schwarz_weights_count_overlapping_subdomains(dims, lx, per, amp):
    c = O.OracleCase(*dims, lx, amp=amp, seed=5, periodic=per)
    h = O.OracleHsmg(c)
    L = O.lib()
    L.or_schwarz_wt3d.argtypes = [O._dp, O._dp, O.C.c_int, O.C.c_int]
    wt = np.ones(c.n)
    L.or_schwarz_wt3d(wt, h.wt_f, lx, c.nel)
    want = 1.0 / _overlap_counts(c)
    assert np.max(np.abs(wt - want)) < 1e-15
    assert len(set(np.round(1 / want).astype(int))) >= 4  # several overlap depths


# ------------------------------------------------------------------------
# hsmg composition: with a coarse space small enough for the coarse CG to
# be exact, the V-cycle equals the additive two-level Schwarz operator
#   M^-1 = S_f + P (S_m + P_c A_c^-1 P_c^T) P^T
# on the free unique dofs, with P, P_c the nodal interpolations between the
# continuous spaces (numpy Lagrange interpolation) and A_c^-1 a dense solve.
# ------------------------------------------------------------------------
def _lagrange_matrix(z_to, z_from):
    n = len(z_from)
    M = np.ones((len(z_to), n))
    for j in range(n):
        for m in range(n):
            if m != j:
                M[:, j] *= (z_to - z_from[m]) / (z_from[j] - z_from[m])
    return M


def _unique(c):
    ids, rep, q = np.unique(c.dof, return_index=True, return_inverse=True)
    masked = np.zeros(c.n, dtype=bool)
    masked[c.msk[: c.nmsk] - 1] = True
    free = ~np.bincount(q, weights=masked.astype(float), minlength=len(ids)).astype(bool)
    return rep, q, free


def _interp_unique(cf, cc):
    """P[u_f, u_c] = value at fine unique node u_f of the coarse nodal basis
    function of unique node u_c (element-wise tensor Lagrange)."""
    J = _lagrange_matrix(cf.z, cc.z)
    E = np.kron(J, np.kron(J, J))            # (k,j,i) ordering on both sides
    rep_f, q_f, _ = _unique(cf)
    rep_c, q_c, _ = _unique(cc)
    nf, nc = len(rep_f), len(rep_c)
    P = np.zeros((nf, nc))
    n3f, n3c = cf.lx ** 3, cc.lx ** 3
    for e in range(cf.nel):
        qf = q_f[e * n3f:(e + 1) * n3f]
        qc = q_c[e * n3c:(e + 1) * n3c]
        for a in range(n3c):
            P[qf, qc[a]] = E[:, a]
    return P


def _dense(apply, c_in, c_out):
    rep_i, q_i, free_i = _unique(c_in)
    rep_o, _, _ = _unique(c_out)
    cols = np.zeros((len(rep_o), len(rep_i)))
    for u in np.nonzero(free_i)[0]:
        r = (q_i == u).astype(float)
        cols[:, u] = apply(r)[rep_o]
    return cols


def test_hsmg_composition_with_exact_coarse_solve():
    c = O.OracleCase(3, 3, 3, 5, amp=0.12, seed=17)
    h = O.OracleHsmg(c, niter=10, crs_tol=1e-300)
    mid, crs = h.mid, h.crs
    _, _, free_f = _unique(c)
    _, q_c, free_c = _unique(crs)
    assert free_c.sum() == 8  # coarse space: the interior vertices
    M = _dense(h.apply, c, c)
    Sf = _dense(h.schwarz_fine, c, c)
    Sm = _dense(h.schwarz_mid, mid, mid)
    P = _interp_unique(c, mid)
    Pc = _interp_unique(mid, crs)
    # assembled coarse operator on the free coarse dofs (sum over copies)
    rep_c, _, _ = _unique(crs)
    Ac = np.zeros((len(rep_c), len(rep_c)))
    for u in np.nonzero(free_c)[0]:
        Ac[:, u] = np.bincount(q_c, weights=crs.ax_local((q_c == u).astype(float)),
                               minlength=len(rep_c))
    fc = np.nonzero(free_c)[0]
    Ainv = np.zeros_like(Ac)
    Ainv[np.ix_(fc, fc)] = np.linalg.inv(Ac[np.ix_(fc, fc)])
    want = Sf + P @ (Sm + Pc @ Ainv @ Pc.T) @ P.T
    ff = np.nonzero(free_f)[0]
    got = M[np.ix_(ff, ff)]
    ref = want[np.ix_(ff, ff)]
    assert np.max(np.abs(got - ref)) < 1e-11 * np.max(np.abs(ref))
    # the coarse term is a visible part of the operator
    coarse = (P @ Pc @ Ainv @ Pc.T @ P.T)[np.ix_(ff, ff)]
    assert np.max(np.abs(coarse)) > 1e-3 * np.max(np.abs(ref))
