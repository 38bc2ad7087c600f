"""The allocation-free smoothing entry point (sp_smooth_noalloc) against
sp_smooth on the device: equal bit for bit, on two calls in a row (the
second on a smaller field, so stale scratch contents from the first call
would show), and sizes beyond its scratch refused with an error before any
device work."""
import ctypes

import numpy as np
import pytest

from seispie_amd import _lib

pytestmark = pytest.mark.gpu


def _pair(ctx, nx, nz, sigma, seed):
    import torch
    rng = np.random.default_rng(seed)
    host = rng.standard_normal(nx * nz).astype(np.float32)
    a = torch.tensor(host, device="cuda:0")
    b = a.clone()
    ctx.smooth(a, nx, nz, sigma)             # sp_smooth, on_device = 1 (allocates its buffers)
    ctx.smooth_noalloc(b, nx, nz, sigma)     # scratch in the fixed device array
    ctx.sync()
    return a.cpu().numpy(), b.cpu().numpy(), host


def test_noalloc_smoothing_equals_sp_smooth_on_two_calls(ctx_factory):
    ctx = ctx_factory()
    print("case: sp_smooth_noalloc 700x900 sigma 5, then 90x130 sigma 3", flush=True)
    for nx, nz, sigma, seed in ((700, 900, 5, 1), (90, 130, 3, 2)):
        want, got, host = _pair(ctx, nx, nz, sigma, seed)
        assert np.abs(want - host).max() > 0          # smoothing did something
        np.testing.assert_array_equal(got, want)
        print("%dx%d sigma %d: equal to sp_smooth" % (nx, nz, sigma), flush=True)


def test_noalloc_smoothing_refuses_fields_beyond_its_scratch(ctx_factory):
    ctx = ctx_factory()
    L = ctx.L
    dummy = ctypes.c_void_p(16)   # never dereferenced: the size check comes first
    for nx, nz in ((8193, 10), (10, 8193), (2049, 2048)):
        rc = L.sp_smooth_noalloc(ctx._p, dummy, nx, nz, 3)
        assert rc != 0
        assert b"exceeds its scratch" in L.sp_last_error(ctx._p)
    assert L.sp_smooth_noalloc(ctx._p, dummy, 10, 10, 0) != 0
