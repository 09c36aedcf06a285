"""GPU parity of the fixed-plane projections (the reference's two CUDA splat
kernels and their host arithmetic) with the oracle.

Max projection: bit-exact (a maximum does not depend on the order of the
atomics).  Weighted projection: relative 1e-5 (float32 atomics add in no fixed
order, in the reference too)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import splat_oracle as S

pytestmark = pytest.mark.gpu


def _samples(seed, shape=(20, 14, 5), xref=24, yref=18):
    rng = np.random.default_rng(seed)
    coords = np.stack([rng.uniform(-2, xref + 1, shape), rng.uniform(-2, yref + 1, shape),
                       rng.uniform(0, 10, shape)], axis=-1).astype(np.float32)
    values = rng.normal(-50, 80, shape).astype(np.float32)
    values.flat[::37] = np.nan
    coords.reshape(-1, 3)[5::53, 1] = np.inf
    return coords, values


@pytest.mark.parametrize("mode,radius,ds,order", [("subpixel_footprint", 0, 1, "zyx"),
                                                  ("subpixel_footprint", 1, 3, "zxy"),
                                                  ("nearest", 0, 2, "zyx")])
def test_max_projection_bit_exact(cuda, mode, radius, ds, order):
    from wholistic_registration_amd import calFlowCrossResolution as X
    coords, values = _samples(7 + radius + ds)
    ref = np.zeros((24, 18, 11), np.float32)
    planes = [2.0, 4.5, 7.0]
    got = X.project_coords_to_fixed_planes_gpu(coords, ref, planes, values_xyk=values, z_window=1.2,
                                               downsample_xy=ds, xy_splat_mode=mode, xy_extra_radius=radius,
                                               output_order=order, return_numpy=True, fill_value=-1.0)
    want = S.project_max(coords, values, planes, 24, 18, z_window=1.2, downsample_xy=ds,
                         nearest=mode == "nearest", extra_radius=radius, fill_value=-1.0)
    if order == "zyx":
        want = want.transpose(0, 2, 1)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode,ds", [("hard", 1), ("triangular", 2), ("gaussian", 1)])
def test_weighted_projection(cuda, mode, ds):
    from wholistic_registration_amd import calFlowCrossResolution as X
    coords, values = _samples(11 + ds)
    ref = np.zeros((11, 18, 24), np.float32)          # (Z, Y, X) order
    planes = np.array([3.0, 6.0], np.float32)
    out, w = X.project_coords_to_fixed_planes_weighted_gpu(coords, ref, planes, values_xyk=values,
                                                           ref_volume_order="zyx", z_window=1.5,
                                                           z_weight_mode=mode, downsample_xy=ds,
                                                           return_numpy=True)
    w_out, w_w = S.project_weighted(coords, values, planes, 24, 18, z_window=1.5, mode=mode, downsample_xy=ds)
    assert np.allclose(w, w_w.transpose(0, 2, 1), rtol=1e-5, atol=1e-6)
    assert np.allclose(out, w_out.transpose(0, 2, 1), rtol=1e-5, atol=1e-4)


def test_reference_volume_is_required(cuda):
    """Without values the reference volume is sampled (tests/test_gpu_xres.py);
    the volume itself stays required, as in the reference."""
    from wholistic_registration_amd import calFlowCrossResolution as X
    coords, _ = _samples(1)
    with pytest.raises(ValueError):
        X.project_coords_to_fixed_planes_gpu(coords, None, [1.0], values_xyk=np.zeros(coords.shape[:-1]))
