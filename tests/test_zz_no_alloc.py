"""The allocation-free entry point wr_smooth_penalty_factor_noalloc, called
through the C ABI: its factor equals wr_smooth_penalty_factor's on the same
input, on two calls in a row (the second on a smaller volume, so the scratch
array holds the first call's partials beyond the second's rows), and sizes
above its scratch maximum are rejected."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _factor_both(lib, ctx, vol, mask, dtype_code):
    import torch
    nz, ny, nx = vol.shape
    host = ctypes.c_double()
    p = ctypes.c_void_p
    mptr = None if mask is None else p(mask.data_ptr())
    rc = lib.wr_smooth_penalty_factor(ctx, p(vol.data_ptr()), dtype_code, nx, ny, nz, mptr, ctypes.byref(host),
                                      p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    dev = torch.full((1,), -1.0, dtype=torch.float64, device=vol.device)
    rc = lib.wr_smooth_penalty_factor_noalloc(ctx, p(vol.data_ptr()), dtype_code, nx, ny, nz, mptr,
                                              p(dev.data_ptr()), p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    return host.value, float(dev.item())


@pytest.mark.parametrize("dtype", ["float32", "uint16"])
def test_noalloc_equals_allocating_entry_point(cuda, dtype):
    import torch

    from wholistic_registration_amd import _lib
    from wholistic_registration_amd import device as D
    lib = _lib.load()
    ctx = _lib.context(0)
    rng = np.random.default_rng(11)
    for shape in ((24, 300, 410), (5, 64, 97)):            # the second call is smaller
        a = rng.uniform(0, 4000, size=shape)
        vol = torch.from_numpy(a.astype(np.float32) if dtype == "float32"
                               else np.rint(a).astype(np.uint16)).to(cuda)
        mask = torch.from_numpy((rng.uniform(size=shape) < 0.2).astype(np.float32)).to(cuda)
        for m in (None, mask):
            want, got = _factor_both(lib, ctx, vol, m, D.dtype_code(vol))
            assert got == want, (shape, m is not None, got, want)


def test_noalloc_rejects_sizes_above_its_scratch(cuda):
    import torch

    from wholistic_registration_amd import _lib
    lib = _lib.load()
    ctx = _lib.context(0)
    nz, ny, nx = 1, 262145, 2                               # one row above the maximum
    vol = torch.zeros((nz, ny, nx), dtype=torch.float32, device=cuda)
    out = torch.zeros((1,), dtype=torch.float64, device=cuda)
    p = ctypes.c_void_p
    rc = lib.wr_smooth_penalty_factor_noalloc(ctx, p(vol.data_ptr()), _lib.WR_F32, nx, ny, nz, None,
                                              p(out.data_ptr()), p(torch.cuda.current_stream().cuda_stream))
    assert rc == _lib.WR_E_ARG
    assert "exceed the scratch maximum" in _lib.last_error()
    torch.cuda.synchronize()
    assert out.item() == 0.0                                 # nothing was launched
