"""GPU tests of batched solves (nm_batch_*): problems of uneven size, order
and Helmholtz shift solved in one batch reach 1e-10 of their initial
residual, agree with the FP64 oracle solution within 1e-9 relative L2,
and a problem's result does not depend on the other problems of the batch
(bit-identical to a batch of one)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

# nelx, nely, nelz, lx, h2, seed
SPECS = [
    (3, 2, 2, 6, 0.0, 1),
    (2, 2, 2, 8, 0.0, 2),
    (4, 3, 2, 4, 150.0, 3),
    (2, 3, 2, 10, 0.0, 4),
    (3, 3, 3, 6, 20.0, 5),
    (1, 2, 3, 8, 0.0, 6),
    (5, 2, 1, 5, 0.0, 7),
]


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _build():
    from neko_amd.solve import Problem
    probs, cases, fs = [], [], []
    for nx, ny, nz, lx, h2, sd in SPECS:
        h1 = 1.0 / 1600 if h2 else 1.0
        c = O.OracleCase(nx, ny, nz, lx, amp=0.1, seed=sd, h1=h1, h2=h2)
        p = Problem.box(nx, ny, nz, lx, amp=0.1, seed=sd, h1=h1, h2=h2)
        f = c.rhs_normal(40 + sd)
        cases.append(c)
        probs.append(p)
        fs.append(torch.from_numpy(f).to("cuda"))
    return probs, cases, fs


@pytest.mark.parametrize("face_sums", [True, False])
def test_batch_matches_oracle(face_sums):
    """face_sums: the update adds the face partners' values (the batch's gs
    runs the blocks only); switched off on the problems, the batch runs the
    facet-pair gs and the chunk-run update instead."""
    from neko_amd.solve import Batch
    print("batch vs oracle", face_sums, flush=True)
    probs, cases, fs = _build()
    for p in probs:
        p.set_cg_face_sums(face_sums)
    b = Batch(probs)
    tuned = b.autotune(reps=2)
    assert set(tuned) == {4, 5, 6, 8, 10}
    xs = [torch.zeros_like(f) for f in fs]
    mons = b.solve(xs, fs, rel_tol=1e-10)
    for c, f, x, m in zip(cases, fs, xs, mons):
        assert m.converged
        assert m.res_final < 1e-10 * m.res_start
        fn = _np(f)
        x_or, mon_or = c.cg(fn, 1e-10 * c.res_norm(fn))
        err = np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or)
        assert err < 1e-9, err
        # residual replacement keeps the Krylov process: the FP32 iteration
        # count stays within 15% of the FP64 oracle's (each replacement
        # perturbs the recurrence slightly; small problems see relatively
        # more replacements)
        assert m.iter <= 1.15 * mon_or.iter + 3, (m.iter, mon_or.iter)


def test_batch_member_independent_of_batch():
    from neko_amd.solve import Batch
    print("batch independence", flush=True)
    probs, cases, fs = _build()
    full = Batch(probs)
    xs = [torch.zeros_like(f) for f in fs]
    mons = full.solve(xs, fs, rel_tol=1e-10)
    for k in (0, 3):
        one = Batch([probs[k]])
        x1 = torch.zeros_like(fs[k])
        m1 = one.solve([x1], [fs[k]], rel_tol=1e-10)[0]
        assert m1.iter == mons[k].iter
        assert torch.equal(x1, xs[k])
        one.free()
    # and solving twice gives the same answer
    xs2 = [torch.zeros_like(f) for f in fs]
    full.solve(xs2, fs, rel_tol=1e-10)
    assert all(torch.equal(a, c) for a, c in zip(xs, xs2))


def test_batch_zero_rhs_member_and_reuse():
    """A member with a zero right-hand side is finished at the start (0
    iterations, x = 0) while the others iterate, replace their residuals
    beside the iteration graphs and drop out of the launches as they
    finish; each of the others is bit-identical to its batch of one, and a
    second solve of the same batch (tables compacted by the first) gives
    the same bits."""
    from neko_amd.solve import Batch
    print("batch zero rhs", flush=True)
    probs, cases, fs = _build()
    fs = list(fs)
    fs[2] = torch.zeros_like(fs[2])
    b = Batch(probs)
    xs = [torch.full_like(f, 3.0) for f in fs]
    mons = b.solve(xs, fs, rel_tol=1e-10)
    assert mons[2].iter == 0 and mons[2].converged
    assert torch.count_nonzero(xs[2]).item() == 0
    for k in (0, 4, 6):
        assert mons[k].converged and mons[k].iter > 0
        one = Batch([probs[k]])
        x1 = torch.zeros_like(fs[k])
        m1 = one.solve([x1], [fs[k]], rel_tol=1e-10)[0]
        assert m1.iter == mons[k].iter and torch.equal(x1, xs[k])
        one.free()
    xs2 = [torch.zeros_like(f) for f in fs]
    mons2 = b.solve(xs2, fs, rel_tol=1e-10)
    assert [m.iter for m in mons2] == [m.iter for m in mons]
    assert all(torch.equal(a, c) for a, c in zip(xs, xs2))


def test_batch_rejects_operator_without_geometry():
    from neko_amd._lib import NekoAmdError
    from neko_amd.solve import Batch, Problem
    c = O.OracleCase(2, 2, 2, 6, amp=0.1, seed=1)
    gs = dict(dg=c.dg, gd=c.gd, blk_len=c.blk_len, blk_off=c.blk_off,
              nlocal=c.nlocal, facet_offset=c.facet_offset, nblks=c.nblks)
    p = Problem.from_arrays(c.nel, c.lx, c.d, c.g, c.B, None, gs, c.msk[: c.nmsk],
                            h1c=1.0, dfrmd=c.dfrmd)
    with pytest.raises(NekoAmdError):
        Batch([p])


def test_batch_with_reference_meshes():
    """Problems on the reference's own meshes (the unstructured hemisphere
    and the periodic lid-driven cavity mesh) batched with box problems:
    each reaches the FP64 oracle's solution within 1e-9."""
    import os
    from neko_amd.solve import Batch, Problem, read_nmsh
    print("batch meshes", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    probs, cases, fs = [], [], []
    for name, lx in (("hemi.nmsh", 5), ("lid.nmsh", 6)):
        v, pid, _ = read_nmsh(os.path.join(root, "tests", "golden", "nmsh", name))
        cases.append(O.OracleCase.from_mesh(v, pid, lx))
        probs.append(Problem.mesh(v, pid, lx))
    cases.append(O.OracleCase(3, 2, 2, 6, amp=0.1, seed=1))
    probs.append(Problem.box(3, 2, 2, 6, amp=0.1, seed=1))
    for k, c in enumerate(cases):
        fs.append(torch.from_numpy(c.rhs_normal(70 + k)).to("cuda"))
    b = Batch(probs)
    xs = [torch.zeros_like(f) for f in fs]
    mons = b.solve(xs, fs, rel_tol=1e-10)
    for c, f, x, m in zip(cases, fs, xs, mons):
        assert m.converged
        fn = _np(f)
        x_or, _ = c.cg(fn, 1e-10 * c.res_norm(fn), max_iter=5000)
        err = np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or)
        assert err < 1e-9, err
