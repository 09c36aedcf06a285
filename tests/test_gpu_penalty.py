"""Parity of wr_smooth_penalty_factor (getSmPnltNormFctr,
preprocess.py:42-66) with the oracle, called through the C ABI.

The oracle follows numpy's arithmetic for the input dtype; the library sums
the same rounded squares in float64 in another order, so the tolerance is
relative 1e-12 for uint16 and float64 inputs and 1e-6 for float32 (numpy's
mean of a float32 array accumulates in float32)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import lk_oracle as O

pytestmark = pytest.mark.gpu


def _zyx(a, cuda):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0))).to(cuda)


def _ramp_vol(rng, shape, dtype):
    x, y, z = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    v = 3.0 * x + 40.0 * np.sin(y / 3.0) + 2.0 * z + rng.uniform(0, 50, size=shape)
    if dtype == np.uint16:
        v = 30000 - 200 * x + rng.integers(0, 5000, size=shape)   # decreasing in x: wraps
    return np.asarray(v).astype(dtype)


@pytest.mark.parametrize("dtype,rtol", [(np.float32, 1e-6), (np.uint16, 1e-12), (np.float64, 1e-12)])
@pytest.mark.parametrize("masked", [False, True])
def test_smooth_penalty_factor_matches_oracle(cuda, rng, dtype, rtol, masked):
    from wholistic_registration_amd import device as D
    shape = (67, 45, 7)                    # (x, y, z)
    ref = _ramp_vol(rng, shape, dtype)
    mask = rng.uniform(size=shape) < 0.3 if masked else np.zeros(shape, bool)
    want = float(O.smooth_penalty_factor(ref, mask))
    got = D.smooth_penalty_factor(_zyx(ref, cuda), _zyx(mask.astype(np.float32), cuda) if masked else None)
    assert got == pytest.approx(want, rel=rtol)


def test_uint16_differences_wrap(cuda):
    """A uint16 step down by 1 along x gives (65535 / 2)^2, as numpy does."""
    from wholistic_registration_amd import device as D
    ref = np.zeros((5, 3, 2), np.uint16)
    ref[:, :, :] = np.arange(5, 0, -1, dtype=np.uint16)[:, None, None]
    want = float(O.smooth_penalty_factor(ref, np.zeros(ref.shape, bool)))
    assert want > 1e8
    assert D.smooth_penalty_factor(_zyx(ref, cuda)) == pytest.approx(want, rel=1e-12)


def test_all_masked_gives_nan(cuda):
    from wholistic_registration_amd import device as D
    ref = np.ones((6, 5, 2), np.float32)
    got = D.smooth_penalty_factor(_zyx(ref, cuda), _zyx(np.ones(ref.shape, np.float32), cuda))
    assert np.isnan(got)


def test_drop_in_getSmPnltNormFctr(cuda, rng):
    from wholistic_registration_amd.preprocess import getSmPnltNormFctr
    shape = (40, 33, 5)
    ref = _ramp_vol(rng, shape, np.float32)
    mask = rng.uniform(size=shape) < 0.2
    want = float(O.smooth_penalty_factor(ref, mask))
    assert getSmPnltNormFctr(ref, {"mask_ref": mask}) == pytest.approx(want, rel=1e-6)
