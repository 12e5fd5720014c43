"""The allocation-free entry point nm_glsc3_many_noalloc (partials in the
library's static device array) against the allocating hip_glsc3_many it
mirrors, through the C ABI: two calls in a row, the second with fewer
vectors and a shorter length, each bit-identical to hip_glsc3_many on the
same data (same kernels, same fixed reduction order); too many vectors are
refused with an error."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bind():
    from neko_amd._lib import lib
    return lib()


def _case(L, j, n, seed):
    rng = np.random.default_rng(seed)
    w = torch.from_numpy(rng.standard_normal(n)).cuda()
    mult = torch.from_numpy(1.0 / rng.integers(1, 9, n)).cuda()
    vs = [torch.from_numpy(rng.standard_normal(n)).cuda() for _ in range(j)]
    vptr = torch.tensor([t.data_ptr() for t in vs], dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    # allocating reference entry point (host result)
    h_ref = (C.c_double * j)()
    jj, nn = C.c_int(j), C.c_int(n)
    L.hip_glsc3_many(h_ref, w.data_ptr(), vptr.data_ptr(), mult.data_ptr(),
                     C.byref(jj), C.byref(nn), s)
    # allocation-free entry point (device result)
    h_dev = torch.full((j,), np.nan, dtype=torch.float64, device="cuda")
    st = L.nm_glsc3_many_noalloc(h_dev.data_ptr(), w.data_ptr(), vptr.data_ptr(),
                                 mult.data_ptr(), j, n, s)
    assert st == 0, L.nm_last_error()
    torch.cuda.synchronize()
    got = h_dev.cpu().numpy()
    ref = np.array(h_ref[:])
    return got, ref, (w.cpu().numpy(), [t.cpu().numpy() for t in vs], mult.cpu().numpy())


def test_noalloc_glsc3_many_matches_allocating_entry_point():
    L = _bind()
    print("noalloc glsc3_many: j=7 n=1000003 then j=3 n=4711", flush=True)
    for j, n, seed in ((7, 1000003, 1), (3, 4711, 2)):
        got, ref, (w, vs, mult) = _case(L, j, n, seed)
        assert np.array_equal(got.view(np.int64), ref.view(np.int64)), (got, ref)
        exact = np.array([np.sum(w * v * mult) for v in vs])
        scale = np.array([np.sum(np.abs(w * v * mult)) for v in vs])
        assert np.all(np.abs(got - exact) <= 1e-13 * scale)


def test_noalloc_rejects_too_many_vectors():
    L = _bind()
    x = torch.ones(16, dtype=torch.float64, device="cuda")
    vptr = torch.tensor([x.data_ptr()] * 65, dtype=torch.int64, device="cuda")
    h = torch.zeros(65, dtype=torch.float64, device="cuda")
    st = L.nm_glsc3_many_noalloc(h.data_ptr(), x.data_ptr(), vptr.data_ptr(),
                                 x.data_ptr(), 65, 16, torch.cuda.current_stream().cuda_stream)
    assert st != 0
    assert b"1..64" in L.nm_last_error()
