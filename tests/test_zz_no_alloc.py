"""bssn_constraint_norms_noalloc: the constraint norms with the reduction's
partial sums in the library's static device scratch instead of a caller
buffer.  Through the C ABI, against bssn_constraint_norms (same kernels, a
caller-allocated partial buffer) on two calls in a row, the second with a
smaller grid: the scratch keeps the first call's partial sums, which the
second must not pick up."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _norms_with_partial(lib, _lib, cf):
    grid = _lib.c_grid(cf.shape[1:])
    partial = torch.empty(int(lib.bssn_norm_partial_size(ctypes.byref(grid))), dtype=torch.float64,
                          device=cf.device)
    out = torch.empty(10, dtype=torch.float64, device=cf.device)
    _lib.check(lib.bssn_constraint_norms(ctypes.byref(grid), ctypes.c_void_p(cf.data_ptr()),
                                         ctypes.c_void_p(partial.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                         None))
    return out


def _norms_noalloc(lib, _lib, cf):
    grid = _lib.c_grid(cf.shape[1:])
    out = torch.empty(10, dtype=torch.float64, device=cf.device)
    _lib.check(lib.bssn_constraint_norms_noalloc(ctypes.byref(grid), ctypes.c_void_p(cf.data_ptr()),
                                                 ctypes.c_void_p(out.data_ptr()), None))
    return out


def test_noalloc_norms_equal_the_allocating_entry_point():
    from jax_bssn_amd import _lib
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    lib = _lib.load()
    rng = np.random.default_rng(21)
    # the large grid fills all 1024 partial blocks; the small one needs 3 and
    # finds the large call's sums in the other slots of the scratch
    for shape in ((9, 96, 96, 96), (9, 5, 7, 11)):
        cf = torch.tensor(rng.standard_normal(shape) * rng.uniform(0.1, 10.0, (9, 1, 1, 1)), device="cuda")
        print("[no-alloc norms] shape", shape, flush=True)
        want = _norms_with_partial(lib, _lib, cf)
        got = _norms_noalloc(lib, _lib, cf)
        torch.cuda.synchronize()
        # same kernels, same block count, same summation order: identical
        assert torch.equal(got, want), (got.tolist(), want.tolist())
        # and the norms themselves (float64 host sums; different order)
        a = cf.cpu().numpy()
        l2 = [np.sqrt(np.mean(a[0] ** 2)), np.sqrt(np.mean(a[1:4] ** 2)), np.sqrt(np.mean(a[4] ** 2)),
              np.sqrt(np.mean(a[5] ** 2)), np.sqrt(np.mean(a[6:9] ** 2))]
        np.testing.assert_allclose(got[:5].cpu().numpy(), l2, rtol=1e-12)
        linf = [np.abs(a[0]).max(), np.abs(a[1:4]).max(), np.abs(a[4]).max(), np.abs(a[5]).max(),
                np.abs(a[6:9]).max()]
        np.testing.assert_array_equal(got[5:].cpu().numpy(), linf)


def test_noalloc_norms_reject_what_the_scratch_cannot_hold():
    from jax_bssn_amd import _lib
    lib = _lib.load()
    out = torch.empty(10, dtype=torch.float64, device="cuda")
    buf = torch.empty(1, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="empty grid"):
        _lib.check(lib.bssn_constraint_norms_noalloc(ctypes.byref(_lib.c_grid((0, 4, 4))),
                                                     ctypes.c_void_p(buf.data_ptr()),
                                                     ctypes.c_void_p(out.data_ptr()), None))
    with pytest.raises(ValueError, match="at most"):
        _lib.check(lib.bssn_constraint_norms_noalloc(ctypes.byref(_lib.c_grid((2048, 2048, 1024))),
                                                     ctypes.c_void_p(buf.data_ptr()),
                                                     ctypes.c_void_p(out.data_ptr()), None))
