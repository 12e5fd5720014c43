"""The velocity solve's three-component entry points against the oracle:
ksp_t%solve_coupled of cg (src/krylov/bcknd/cpu/cg.f90:282-305: three
independent solves, each component with its own bc list blstx / blsty /
blstz, src/fluid/fluid_pnpn.f90:899) through nm_cg_set_mask +
nm_cg_solve_coupled, and ax_t%compute_vector (src/math/ax.f90:98-115)
through nm_op_ax_local_vector.  Bars: iterations +-1 (FP64), the FP64
residual reached (mixed), solution within 1e-9 relative L2 of the oracle's
Jacobi-PCG with the same mask; Ax 1e-12."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DIMS, LX = (3, 3, 4), 6


def _masks(c):
    """1-based masks: x-component every boundary face, y-component the
    faces x = 0 and x = 1, z-component the faces z = 0 and z = 1."""
    on = lambda a: (a == 0.0) | (a == 1.0)  # noqa: E731  (boundary vertices are exact)
    full = np.sort(c.msk[: c.nmsk])
    mx = np.nonzero(on(c.x))[0] + 1
    mz = np.nonzero(on(c.zc))[0] + 1
    return [full, mx.astype(np.int32), mz.astype(np.int32)]


def _oracle_with_mask(c, m):
    c.msk = np.ascontiguousarray(m, dtype=np.int32)
    c.nmsk = len(m)
    c._prob = None
    return c


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_coupled_per_component_masks(precision):
    from neko_amd import solve as S
    h1, h2 = 1.0 / 1600, 1500.0
    base = O.OracleCase(*DIMS, LX, amp=0.1, seed=3, h1=h1, h2=h2)
    masks = _masks(base)
    assert len(set(len(m) for m in masks)) == 3
    p = S.Problem.box(*DIMS, LX, amp=0.1, seed=3, h1=h1, h2=h2)
    rng = np.random.default_rng(8)
    solvers, xs, fs, refs, mons_or = [], [], [], [], []
    for c_idx, m in enumerate(masks):
        c = _oracle_with_mask(O.OracleCase(*DIMS, LX, amp=0.1, seed=3, h1=h1, h2=h2), m)
        f = c.gs(rng.standard_normal(c.n))
        f[m - 1] = 0.0
        tol = 1e-10 * c.res_norm(f)
        x_or, mon_or = c.cg(f, tol)
        refs.append(x_or)
        mons_or.append(mon_or)
        cg = S.CG(p, precision)
        cg.set_mask(m)
        solvers.append(cg)
        xs.append(torch.zeros(c.n, dtype=torch.float64, device="cuda"))
        fs.append(torch.from_numpy(f).cuda())
    mons = S.solve_coupled(solvers, xs, fs, tol)
    for c_idx in range(3):
        m, m_or = mons[c_idx], mons_or[c_idx]
        xr = xs[c_idx].cpu().numpy()
        err = np.linalg.norm(xr - refs[c_idx]) / np.linalg.norm(refs[c_idx])
        assert m.converged, (c_idx, m)
        if precision == "fp64":
            assert abs(m.iter - m_or.iter) <= 1, (c_idx, m.iter, m_or.iter)
        assert err < 1e-9, (c_idx, err)
        # the component's own mask holds on its solution
        assert np.all(xr[masks[c_idx] - 1] == 0.0)
    # the masks matter: the x and y solutions differ where only x is masked
    assert np.linalg.norm(refs[0] - refs[1]) > 1e-3 * np.linalg.norm(refs[0])


def test_set_mask_none_restores_operator_mask():
    from neko_amd import solve as S
    c = O.OracleCase(*DIMS, LX, amp=0.1, seed=4)
    p = S.Problem.box(*DIMS, LX, amp=0.1, seed=4)
    f_or = c.rhs_normal(5)
    tol = 1e-10 * c.res_norm(f_or)
    x_or, mon_or = c.cg(f_or, tol)
    cg = S.CG(p, "fp64")
    cg.set_mask(_masks(c)[1])
    cg.set_mask(None)
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    mon = cg.solve(x, torch.from_numpy(f_or).cuda(), tol)
    assert abs(mon.iter - mon_or.iter) <= 1
    assert np.linalg.norm(x.cpu().numpy() - x_or) < 1e-9 * np.linalg.norm(x_or)


def test_set_mask_rejects_out_of_range():
    from neko_amd import solve as S
    from neko_amd._lib import NekoAmdError
    p = S.Problem.box(2, 2, 2, 4)
    cg = S.CG(p, "fp64")
    with pytest.raises(NekoAmdError):
        cg.set_mask(np.array([1, p.n + 1], dtype=np.int32))


def test_ax_local_vector():
    from neko_amd import solve as S
    c = O.OracleCase(*DIMS, LX, amp=0.1, seed=6, h1=0.3, h2=2.0)
    p = S.Problem.box(*DIMS, LX, amp=0.1, seed=6, h1=0.3, h2=2.0)
    rng = np.random.default_rng(2)
    u = [rng.standard_normal(c.n) for _ in range(3)]
    ud = [torch.from_numpy(a).cuda() for a in u]
    out = [torch.zeros(c.n, dtype=torch.float64, device="cuda") for _ in range(3)]
    p.ax_local_vector(*out, *ud)
    for k in range(3):
        ref = c.ax_local(u[k])
        got = out[k].cpu().numpy()
        assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
