"""GPU parity of the HIP path against the CPU FP64 oracle (oracle/).

Bars (written in each test): integer maps and gs sums bit-exact; one Ax
apply within 1e-12 relative (max-norm); geometric factors / Jacobi within
1e-12; full solves: iteration count within +-1 and solution within 1e-9
relative L2 of the oracle.  All calls go through the C ABI of
libneko_amd.so (via neko_amd.solve / neko_amd._lib).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _t(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _problem_from_oracle(c, with_h1_field=False):
    from neko_amd.solve import Problem
    gs = dict(dg=c.dg, gd=c.gd, blk_len=c.blk_len, blk_off=c.blk_off,
              nlocal=c.nlocal, facet_offset=c.facet_offset, nblks=c.nblks)
    return Problem.from_arrays(
        c.nel, c.lx, c.d, c.g, c.B, None, gs, c.msk[: c.nmsk],
        h1=c.h1 if with_h1_field else None, h1c=float(c.h1[0]),
        h2=c.h2 if c.ifh2 else None, dfrmd=c.dfrmd)


CASES = [
    # nelx, nely, nelz, lx, amp, periodic
    (3, 3, 3, 8, 0.0, (0, 0, 0)),
    (3, 4, 2, 8, 0.1, (0, 0, 0)),
    (3, 3, 3, 2, 0.1, (1, 0, 0)),
    (3, 2, 3, 3, 0.1, (0, 0, 0)),
    (2, 3, 3, 4, 0.1, (0, 1, 0)),
    (3, 2, 2, 6, 0.1, (0, 0, 0)),
    (2, 2, 3, 7, 0.05, (0, 0, 1)),
    (2, 2, 2, 10, 0.1, (0, 0, 0)),
    (2, 2, 2, 12, 0.1, (0, 0, 0)),
    (2, 2, 1, 14, 0.0, (0, 0, 0)),
]


@pytest.mark.parametrize("case", CASES)
def test_ax_local_matches_oracle(case):
    nx, ny, nz, lx, amp, per = case
    c = O.OracleCase(nx, ny, nz, lx, amp=amp, seed=11, periodic=per)
    p = _problem_from_oracle(c)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(c.n)
    w = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    p.ax_local(w, _t(u))
    ref = c.ax_local(u)
    assert _relmax(_np(w), ref) < 1e-12


@pytest.mark.parametrize("lx", [6, 8])
def test_ax_h1_field_and_h2(lx):
    c = O.OracleCase(3, 2, 2, lx, amp=0.1, seed=2, h1=1.0 / 1600, h2=1500.0)
    rng = np.random.default_rng(9)
    c.h1 = c.h1 * (1.0 + 0.5 * rng.random(c.n))  # pointwise h1 field
    p = _problem_from_oracle(c, with_h1_field=True)
    u = rng.standard_normal(c.n)
    w = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    p.ax_local(w, _t(u))
    assert _relmax(_np(w), c.ax_local(u)) < 1e-12


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("case", CASES[:7])
def test_gs_bit_exact(case, variant):
    nx, ny, nz, lx, amp, per = case
    c = O.OracleCase(nx, ny, nz, lx, amp=amp, seed=3, periodic=per)
    p = _problem_from_oracle(c)
    p.set_gs_variant(variant)
    if variant >= 2 and lx > 2:
        # every facet pair of a conforming box lies on a whole shared face
        assert p.info.nrest == 0 and p.info.nfaces * (lx - 2) ** 2 == p.info.npairs
    u = np.random.default_rng(1).standard_normal(c.n)
    # negative zeros: a shared node whose entries are all -0 sums to -0
    u[np.random.default_rng(2).random(c.n) < 0.3] = -0.0
    ud = _t(u)
    p.gs_op(ud)
    ref = c.gs(u.copy())
    assert np.array_equal(_np(ud).view(np.int64), ref.view(np.int64))
    z = np.full(c.n, -0.0)
    zd = _t(z)
    p.gs_op(zd)
    assert np.array_equal(_np(zd).view(np.int64), c.gs(z.copy()).view(np.int64))


@pytest.mark.parametrize("code", [0, 2, 4, 100, 102, 104])
def test_ax_variants_match_oracle(code):
    c = O.OracleCase(3, 3, 2, 8, amp=0.1, seed=12)
    p = _problem_from_oracle(c)
    p.set_epb(code, code)
    u = np.random.default_rng(2).standard_normal(c.n)
    w = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    p.ax_local(w, _t(u))
    assert _relmax(_np(w), c.ax_local(u)) < 1e-12
    # FP32 variant: within single-precision rounding of the FP64 oracle
    w32 = torch.zeros(c.n, dtype=torch.float32, device="cuda")
    p.ax_local(w32, _t(u, torch.float32))
    assert _relmax(_np(w32).astype(np.float64), c.ax_local(u)) < 2e-5


def test_autotune_keeps_results():
    c = O.OracleCase(4, 3, 3, 8, amp=0.1, seed=13)
    p = _problem_from_oracle(c)
    choice = p.autotune(reps=2)
    assert set(choice) == {"ax64", "ax32", "gs"}
    u = np.random.default_rng(3).standard_normal(c.n)
    ud = _t(u)
    w = torch.zeros_like(ud)
    p.ax_full(w, ud)
    assert _relmax(_np(w), c.ax_full(u)) < 1e-12


def test_dropin_gather_scatter_bit_exact():
    """hip_gather_kernel + hip_scatter_kernel with Neko's own 1-based maps
    (gather_scatter.f90:1661-1666 call sequence) against gs_cpu."""
    from neko_amd._lib import lib
    c = O.OracleCase(3, 3, 2, 6, amp=0.1, seed=4, periodic=(1, 0, 0))
    u = np.random.default_rng(2).standard_normal(c.n)
    ud = _t(u)
    v = torch.zeros(c.nlocal, dtype=torch.float64, device="cuda")
    dg, gd = _t(c.dg, torch.int32), _t(c.gd, torch.int32)
    b, bo = _t(c.blk_len, torch.int32), _t(c.blk_off, torch.int32)
    m, o, n, nb, op = (C.c_int(c.nlocal), C.c_int(c.facet_offset), C.c_int(c.n),
                       C.c_int(c.nblks), C.c_int(1))
    s = torch.cuda.current_stream().cuda_stream
    L = lib()
    L.hip_gather_kernel(v.data_ptr(), C.byref(m), C.byref(o), dg.data_ptr(),
                        ud.data_ptr(), C.byref(n), gd.data_ptr(), C.byref(nb),
                        b.data_ptr(), bo.data_ptr(), C.byref(op), s)
    L.hip_scatter_kernel(v.data_ptr(), C.byref(m), dg.data_ptr(), ud.data_ptr(),
                         C.byref(n), gd.data_ptr(), C.byref(nb), b.data_ptr(),
                         bo.data_ptr(), s)
    assert np.array_equal(_np(ud), c.gs(u.copy()))


def test_glsc3_known_answer():
    """tests/unit/device_math/test_device_math_parallel.pf:615-647:
    device_glsc3 of ones with n = 4711 gives 4711 exactly (1 rank)."""
    from neko_amd._lib import lib
    n = 4711
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    nn = C.c_int(n)
    r = lib().hip_glsc3(a.data_ptr(), a.data_ptr(), a.data_ptr(), C.byref(nn),
                        torch.cuda.current_stream().cuda_stream)
    assert r == 4711.0
    x = np.random.default_rng(3).standard_normal((3, 100003))
    t = [_t(q) for q in x]
    nn = C.c_int(100003)
    r = lib().hip_glsc3(t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(),
                        C.byref(nn), torch.cuda.current_stream().cuda_stream)
    ref = O.lib().or_glsc3(x[0].copy(), x[1].copy(), x[2].copy(), 100003)
    assert abs(r - ref) <= 1e-12 * np.sum(np.abs(x[0] * x[1] * x[2]))


@pytest.mark.parametrize("case", [(4, 3, 3, 8, 0.1, (0, 0, 0)),
                                  (3, 3, 4, 5, 0.08, (1, 0, 0))])
def test_box_setup_matches_oracle(case):
    """Product setup (host maps + device coef/mult/Jacobi) vs oracle setup."""
    from neko_amd import solve as S
    nx, ny, nz, lx, amp, per = case
    c = O.OracleCase(nx, ny, nz, lx, amp=amp, seed=7, periodic=per)
    p = S.Problem.box(nx, ny, nz, lx, periodic=per, amp=amp, seed=7)
    m = p.gs_maps()
    assert np.array_equal(m["dg"], c.dg) and np.array_equal(m["gd"], c.gd)
    assert np.array_equal(m["blk_len"], c.blk_len)
    assert np.array_equal(m["blk_off"], c.blk_off)
    for k in range(6):
        assert _relmax(_np(p.field(k)), c.g[k]) < 1e-12, k
    assert _relmax(_np(p.field(S.FIELD_B)), c.B) < 1e-12
    assert np.array_equal(_np(p.field(S.FIELD_MULT)), c.mult)
    assert abs(p.volume - c.volume) < 1e-13 * c.volume
    assert _relmax(_np(p.field(S.FIELD_DINV)), c.jacobi()) < 1e-12


def _solve_pair(nx, ny, nz, lx, amp, per, rhs, precision="fp64", h1=1.0, h2=0.0,
                code=None, face_sums=None, gs_variant=None):
    from neko_amd import solve as S
    c = O.OracleCase(nx, ny, nz, lx, amp=amp, seed=21, periodic=per, h1=h1, h2=h2)
    p = S.Problem.box(nx, ny, nz, lx, periodic=per, amp=amp, seed=21, h1=h1, h2=h2)
    if code is not None:
        p.set_epb(code, code)
    if gs_variant is not None:
        p.set_gs_variant(gs_variant)
    if face_sums is not None:
        p.set_cg_face_sums(face_sums)
        assert p.info.cg_face_sums == int(face_sums)
        assert p.info.mult_counts == 1
    if rhs == "sine":
        f_or = c.rhs_sine()
    else:
        f_or = c.rhs_normal(5)
    f = _t(f_or)  # identical RHS on both sides
    res_start = c.res_norm(f_or)
    tol = 1e-10 * res_start
    x_or, mon_or = c.cg(f_or, tol)
    cg = S.CG(p, precision)
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    mon = cg.solve(x, f, tol)
    xr = _np(x)
    err = np.linalg.norm(xr - x_or) / np.linalg.norm(x_or)
    return mon, mon_or, err


def test_cg_fp64_poisson_sine_c1_shape():
    """Unit cube, undeformed, lx=8, f = 3 pi^2 sin sin sin (the CPU
    reference config at 4^3 elements for speed): iterations +-1, solution
    1e-9 relative L2."""
    mon, mon_or, err = _solve_pair(4, 4, 4, 8, 0.0, (0, 0, 0), "sine")
    assert mon.converged and mon_or.converged
    assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
    assert err < 1e-9
    assert abs(mon.res_start - mon_or.res_start) < 1e-12 * mon_or.res_start


def test_cg_fp64_perturbed_normal_rhs():
    mon, mon_or, err = _solve_pair(4, 3, 3, 8, 0.1, (0, 0, 0), "normal")
    assert mon.converged
    assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
    assert err < 1e-9


def test_cg_fp64_helmholtz_periodic():
    mon, mon_or, err = _solve_pair(3, 3, 3, 6, 0.08, (1, 0, 0), "normal",
                                   h1=1.0 / 1600, h2=1500.0)
    assert mon.converged
    assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
    assert err < 1e-9


def test_cg_mixed_refined_matches_fp64_oracle():
    """FP32 inner CG (FP32 Ax/gs/vectors, FP64 dots) inside FP64 refinement
    reaches the same 1e-10 relative residual; solution within 1e-9."""
    mon, mon_or, err = _solve_pair(4, 3, 3, 8, 0.1, (0, 0, 0), "normal",
                                   precision="mixed")
    assert mon.converged
    assert mon.res_final < 1e-10 * mon.res_start
    assert err < 1e-9


# --- Ax with in-kernel trilinear geometry (box meshes, constant h1) ---------
TRI_CASES = [
    # nelx, nely, nelz, lx, amp, periodic, h2
    (3, 3, 2, 8, 0.1, (0, 0, 0), 0.0),
    (3, 2, 3, 8, 0.1, (1, 0, 0), 1500.0),
    (3, 3, 3, 2, 0.1, (0, 0, 0), 0.0),
    (3, 2, 3, 4, 0.1, (0, 0, 1), 5.0),
    (2, 3, 2, 3, 0.1, (0, 0, 0), 10.0),
    (2, 2, 3, 5, 0.08, (0, 0, 0), 0.0),
    (2, 3, 2, 6, 0.1, (0, 1, 0), 0.0),
    (2, 2, 2, 7, 0.1, (0, 0, 0), 3.0),
    (2, 2, 2, 10, 0.1, (0, 0, 0), 0.0),
    (2, 2, 1, 12, 0.05, (0, 0, 0), 0.0),
    (2, 1, 2, 11, 0.1, (0, 0, 0), 2.0),
]


@pytest.mark.parametrize("case", TRI_CASES)
@pytest.mark.parametrize("code", [201, 202, 204, 300])
def test_ax_trilinear_matches_oracle(case, code):
    """G and B formed in the kernel from the element's edge vectors agree
    with the oracle's coef-based operator (spectral derivatives of the
    trilinear GLL coordinates) within 1e-12; FP32 within 2e-5.  Code 300
    runs the FP64 contractions on the matrix cores at lx 4 and 8 (the
    vector kernel elsewhere and in FP32)."""
    from neko_amd import solve as S
    nx, ny, nz, lx, amp, per, h2 = case
    if lx * lx * (code % 100) > 256:
        pytest.skip("workgroup above 256 threads: variant not built")
    print("trilinear", case, code, flush=True)
    h1 = 1.0 / 1600 if h2 else 1.0
    c = O.OracleCase(nx, ny, nz, lx, amp=amp, seed=17, periodic=per, h1=h1, h2=h2)
    p = S.Problem.box(nx, ny, nz, lx, periodic=per, amp=amp, seed=17, h1=h1, h2=h2)
    p.set_epb(code, code)
    u = np.random.default_rng(6).standard_normal(c.n)
    w = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    p.ax_local(w, _t(u))
    ref = c.ax_local(u)
    assert _relmax(_np(w), ref) < 1e-12
    w32 = torch.zeros(c.n, dtype=torch.float32, device="cuda")
    p.ax_local(w32, _t(u, torch.float32))
    assert _relmax(_np(w32).astype(np.float64), ref) < 2e-5


def test_trilinear_rejected_without_vertices():
    """Operators built from Neko's arrays carry no vertex data: the
    in-kernel geometry variants are refused, not silently wrong."""
    from neko_amd._lib import NekoAmdError
    c = O.OracleCase(2, 2, 2, 6, amp=0.1, seed=1)
    p = _problem_from_oracle(c)
    with pytest.raises(NekoAmdError):
        p.set_epb(201, 201)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cg_fused_pap_trilinear(precision):
    """CG with p.Ap fused into the trilinear Ax kernel: iterations +-1 and
    solution within 1e-9 of the oracle (FP64), 1e-10 relative residual
    reached (mixed)."""
    print("cg fused", precision, flush=True)
    mon, mon_or, err = _solve_pair(4, 3, 3, 8, 0.1, (0, 0, 0), "normal",
                                   precision=precision, code=201)
    assert mon.converged
    if precision == "fp64":
        assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
    else:
        assert mon.res_final < 1e-10 * mon.res_start
    assert err < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("lx,per", [(5, (1, 0, 1)), (8, (0, 1, 0)), (4, (0, 0, 1)),
                                    (12, (1, 0, 0))])
@pytest.mark.parametrize("face_sums", [False, True])
def test_cg_update_face_sums(precision, lx, per, face_sums):
    """The fused CG update adding the descriptor-face partner values itself
    (and with it switched off, the face gs pass): oracle iterations +-1
    (FP64), 1e-10 reached (mixed), solution within 1e-9 either way.  lx 4,
    8 and 12 run the three layouts of the FP32 four-point update (several
    elements per workgroup, one, two passes per thread), lx 5 the
    single-point form."""
    print("cg face sums", precision, lx, per, face_sums, flush=True)
    mon, mon_or, err = _solve_pair(3, 4, 3, lx, 0.1, per, "normal",
                                   precision=precision, code=201, face_sums=face_sums)
    assert mon.converged
    if precision == "fp64":
        assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
    else:
        assert mon.res_final < 1e-10 * mon.res_start
    assert err < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("lx,per,code", [(8, (0, 1, 0), 201), (5, (1, 0, 1), 201),
                                         (6, (0, 0, 1), 102)])
def test_cg_packed_segment_blocks(precision, lx, per, code):
    """The fused CG with the gs blocks as packed segments (gs variant 3),
    in-kernel and stored geometry: oracle iterations +-1 (FP64), 1e-10
    reached (mixed), solution within 1e-9."""
    print("cg packed segments", precision, lx, per, code, flush=True)
    mon, mon_or, err = _solve_pair(3, 4, 3, lx, 0.1, per, "normal", precision=precision,
                                   code=code, face_sums=True, gs_variant=3)
    assert mon.converged
    if precision == "fp64":
        assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
    else:
        assert mon.res_final < 1e-10 * mon.res_start
    assert err < 1e-9


def test_cg_fused_pap_helmholtz_periodic():
    mon, mon_or, err = _solve_pair(3, 3, 3, 6, 0.08, (1, 0, 0), "normal",
                                   h1=1.0 / 1600, h2=1500.0, code=202)
    assert mon.converged
    assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
    assert err < 1e-9


def test_wrong_dtype_or_size_is_rejected_on_the_host():
    """The C ABI takes raw pointers; the Python layer refuses a vector of
    the wrong dtype or length before any kernel sees it (an FP32 buffer
    given to the FP64 gs would be read and written past its end)."""
    from neko_amd.solve import Problem
    p = Problem.box(2, 2, 2, 4)
    with pytest.raises(TypeError):
        p.gs_op(torch.zeros(p.n, dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError):
        p.gs_op(torch.zeros(p.n - 1, dtype=torch.float64, device="cuda"))
    with pytest.raises(TypeError):
        p.ax_local(torch.zeros(p.n, dtype=torch.float64, device="cuda"),
                   torch.zeros(p.n, dtype=torch.float32, device="cuda"))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("code", [1, 102])
@pytest.mark.parametrize("face_sums", [False, True])
def test_cg_fused_stored_geometry(precision, code, face_sums):
    """The CG prologue and p.Ap fused into the stored-geometry Ax kernel
    (the schedule of operators built from Neko's arrays): oracle iterations
    +-1 (FP64), 1e-10 reached (mixed), solution within 1e-9."""
    print("cg fused stored", precision, code, face_sums, flush=True)
    mon, mon_or, err = _solve_pair(3, 4, 3, 8, 0.1, (0, 1, 0), "normal",
                                   precision=precision, code=code, face_sums=face_sums)
    assert mon.converged
    if precision == "fp64":
        assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
    else:
        assert mon.res_final < 1e-10 * mon.res_start
    assert err < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cg_neko_arrays_h1_field_helmholtz(precision):
    """An operator from Neko's own arrays (nm_op_create: G, B, a pointwise
    h1 field, h2 B, the gs maps and mask) runs the fused stored-geometry
    schedule: iterations +-1 (FP64) and solution within 1e-9 of the oracle's
    Jacobi-PCG on the same arrays."""
    from neko_amd import solve as S
    c = O.OracleCase(3, 3, 3, 6, amp=0.1, seed=23, periodic=(1, 0, 0),
                     h1=1.0 / 1600, h2=1500.0)
    c.h1 = c.h1 * (1.0 + 0.5 * np.random.default_rng(2).random(c.n))
    c._prob = None
    p = _problem_from_oracle(c, with_h1_field=True)
    f_or = c.rhs_normal(7)
    tol = 1e-10 * c.res_norm(f_or)
    x_or, mon_or = c.cg(f_or, tol)
    cg = S.CG(p, precision)
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    mon = cg.solve(x, _t(f_or), tol)
    err = np.linalg.norm(_np(x) - x_or) / np.linalg.norm(x_or)
    assert mon.converged
    if precision == "fp64":
        assert abs(mon.iter - mon_or.iter) <= 1, (mon.iter, mon_or.iter)
    assert err < 1e-9, err
