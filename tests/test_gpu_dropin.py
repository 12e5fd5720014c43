"""The drop-in entry points, called exactly as Neko's Fortran binds them
(device pointers by value, integers by reference, FP64 real, the global
command queue), against the oracle or numpy:

* hip_ax_helm (ax_helm.hip:64-67 argument order, pointwise h1 field) vs
  or_ax_helm, 1e-12;
* hip_jacobi_update (pc_jacobi.hip:108) vs the oracle's restatement of the
  reference device kernel, and nm_jacobi_update (per-element flags) vs the
  CPU formula, on deformed and undeformed elements, 1e-12;
* hip_glsc3_many (math.hip:922) vs or_glsc3 per vector, 1e-12;
* every BLAS-1 wrapper vs numpy: exact for the single-operation updates,
  within one rounding for the multiply-adds (the device may fuse them).
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


def _lib():
    from neko_amd._lib import lib
    return lib()


def _ref(v):
    return C.byref(C.c_int(int(v)))


@pytest.mark.parametrize("lx,amp", [(8, 0.1), (5, 0.15), (12, 0.0), (2, 0.1)])
def test_hip_ax_helm_neko_arguments(lx, amp):
    c = O.OracleCase(3, 2, 2, lx, amp=amp, seed=3)
    rng = np.random.default_rng(4)
    c.h1 = 0.5 + rng.random(c.n)           # pointwise h1 field
    u = rng.standard_normal(c.n)
    ref = c.ax_local(u)
    ud, wd = _t(u), torch.full((c.n,), np.nan, dtype=torch.float64, device="cuda")
    d, dt = _t(c.d), _t(c.dt)
    h1 = _t(c.h1)
    g = [_t(a) for a in c.g]
    _lib().hip_ax_helm(wd.data_ptr(), ud.data_ptr(), d.data_ptr(), d.data_ptr(),
                       d.data_ptr(), dt.data_ptr(), dt.data_ptr(), dt.data_ptr(),
                       h1.data_ptr(), *[a.data_ptr() for a in g],
                       _ref(c.nel), _ref(lx))
    assert _relmax(_np(wd), ref) < 1e-12


@pytest.mark.parametrize("lx,amp", [(8, 0.1), (6, 0.0), (4, 0.2), (10, 0.1)])
def test_hip_jacobi_update_device_formula(lx, amp):
    c = O.OracleCase(2, 3, 2, lx, amp=amp, seed=8)
    L = O.lib()
    want_dev = np.zeros(c.n)
    L.or_jacobi_diag_dev(want_dev, c.dt, *c.g, c.nel, lx)
    want_cpu = np.zeros(c.n)
    L.or_jacobi_diag(want_cpu, c.dt, *c.g, c.dfrmd, c.nel, lx)
    dt = _t(c.dt)
    g = [_t(a) for a in c.g]
    d = torch.full((c.n,), np.nan, dtype=torch.float64, device="cuda")
    _lib().hip_jacobi_update(d.data_ptr(), dt.data_ptr(), dt.data_ptr(), dt.data_ptr(),
                             *[a.data_ptr() for a in g], _ref(c.nel), _ref(lx))
    assert _relmax(_np(d), want_dev) < 1e-12
    d2 = torch.full((c.n,), np.nan, dtype=torch.float64, device="cuda")
    fl = _t(c.dfrmd, torch.int32)
    _lib().nm_jacobi_update(d2.data_ptr(), dt.data_ptr(), *[a.data_ptr() for a in g],
                            fl.data_ptr(), _ref(c.nel), _ref(lx),
                            torch.cuda.current_stream().cuda_stream)
    assert _relmax(_np(d2), want_cpu) < 1e-12
    if amp > 0:
        # deformed elements: the two formulas differ on the element edges
        assert _relmax(want_cpu, want_dev) > 1e-8


def test_hip_glsc3_many():
    rng = np.random.default_rng(5)
    n, j = 300007, 9
    w = rng.standard_normal(n)
    mult = 1.0 / rng.integers(1, 9, n)
    vs = [rng.standard_normal(n) for _ in range(j)]
    wd, md = _t(w), _t(mult)
    vd = [_t(v) for v in vs]
    vptr = torch.tensor([t.data_ptr() for t in vd], dtype=torch.int64, device="cuda")
    h = (C.c_double * j)()
    _lib().hip_glsc3_many(h, wd.data_ptr(), vptr.data_ptr(), md.data_ptr(),
                          _ref(j), _ref(n), torch.cuda.current_stream().cuda_stream)
    for l in range(j):
        ref = O.lib().or_glsc3(w, vs[l], mult, n)
        assert abs(h[l] - ref) <= 1e-12 * np.sum(np.abs(w * vs[l] * mult))


# (name, operands, scalar, numpy result, fused multiply-add?)
BLAS1 = [
    ("hip_add2s1", 2, 1.7, lambda a, b, c, d, s: s * a + b, True),
    ("hip_add2s2", 2, -0.3, lambda a, b, c, d, s: a + s * b, True),
    ("hip_cmult2", 2, 2.5, lambda a, b, c, d, s: b * s, False),
    ("hip_copy", 2, None, lambda a, b, c, d, s: b, False),
    ("hip_rzero", 1, None, lambda a, b, c, d, s: np.zeros_like(a), False),
    ("hip_cmult", 1, -1.25, lambda a, b, c, d, s: a * s, False),
    ("hip_cfill", 1, 3.75, lambda a, b, c, d, s: np.full_like(a, s), False),
    ("hip_add2", 2, None, lambda a, b, c, d, s: a + b, False),
    ("hip_col2", 2, None, lambda a, b, c, d, s: a * b, False),
    ("hip_col3", 3, None, lambda a, b, c, d, s: b * c, False),
    ("hip_sub2", 2, None, lambda a, b, c, d, s: a - b, False),
    ("hip_invcol1", 1, None, lambda a, b, c, d, s: 1.0 / a, False),
    ("hip_addcol3", 3, None, lambda a, b, c, d, s: a + b * c, True),
    ("hip_addcol4", 4, None, lambda a, b, c, d, s: a + b * c * d, True),
]


@pytest.mark.parametrize("name,nops,scal,fn,fma", BLAS1, ids=[b[0] for b in BLAS1])
def test_blas1_wrappers(name, nops, scal, fn, fma):
    n = 100003
    rng = np.random.default_rng(sum(name.encode()))
    ops = [rng.standard_normal(n) + 3.0 for _ in range(4)]
    want = fn(*ops, scal)
    dev = [_t(a) for a in ops[:nops]]
    args = [t.data_ptr() for t in dev]
    if scal is not None:
        args.append(C.byref(C.c_double(scal)))
    args += [_ref(n), torch.cuda.current_stream().cuda_stream]
    getattr(_lib(), name)(*args)
    got = _np(dev[0])
    if fma:
        # one rounding apart at most (fused or separate multiply-add)
        assert np.all(np.abs(got - want) <= 4 * np.finfo(float).eps * (np.abs(want) + 10))
    else:
        assert np.array_equal(got, want)
