"""Pin the mask / harness oracle (CPU only) against the reference's own test
vector, numpy's percentile on closed forms, brute-force morphology and
connected components, and golden vectors written by running the reference's
pick_initial_reference."""

from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import harness_oracle, mask_oracle
from tests.golden.harness_cases import PICK_CASES, pick_frames

HERE = os.path.dirname(os.path.abspath(__file__))


def test_robust_mean_std_reference_vector():
    """tests/unit/test_preprocess.py:52-57 of the reference: the outlier above
    the 95th percentile is excluded."""
    data = np.concatenate([np.ones(99, dtype=np.float32), np.array([1e6], dtype=np.float32)])
    mean, std = mask_oracle.robust_mean_std(data, percentile=95)
    assert mean == pytest.approx(1.0, abs=1e-6)
    assert std == pytest.approx(0.0, abs=1e-6)


@pytest.mark.parametrize("n", [1, 2, 7, 100, 1001])
def test_percentile_closed_form(n):
    """For 0..n-1 the linear 95th percentile is 0.95 (n - 1) (float32 index),
    and the selected values are 0..floor(thr)."""
    a = np.arange(n, dtype=np.float32)
    thr = mask_oracle.percentile_threshold(a)
    vi = np.float32(n - 1) * (np.float32(95) / np.float32(100))
    assert thr == pytest.approx(float(vi), rel=1e-6)
    k = int(np.floor(thr)) + 1
    mean, std = mask_oracle.robust_mean_std(a)
    assert mean == pytest.approx((k - 1) / 2.0, rel=1e-6, abs=1e-6)
    assert std == pytest.approx(np.sqrt((k * k - 1) / 12.0), rel=1e-5, abs=1e-6)


def _tiny_masks(rng, n=6):
    for _ in range(n):
        shape = tuple(int(s) for s in rng.integers(1, 7, size=3))
        yield rng.random(shape) < rng.uniform(0.2, 0.8)


def test_morphology_matches_brute_force(rng):
    """Erosion treats outside as set, dilation as clear; opening then dilation."""
    for m in _tiny_masks(rng, 10):
        e = mask_oracle.brute_morph(m, erode=True, border=True)
        d1 = mask_oracle.brute_morph(e, erode=False, border=False)
        d2 = mask_oracle.brute_morph(d1, erode=False, border=False)
        from scipy import ndimage as ndi
        fp = mask_oracle.FOOTPRINT
        assert np.array_equal(ndi.binary_erosion(m, structure=fp, border_value=True), e)
        got = ndi.binary_dilation(ndi.binary_dilation(e, structure=fp), structure=fp)
        assert np.array_equal(got, d2)


def test_get_mask_border_rule():
    """A set voxel at the volume's edge survives the opening (outside counts as
    set for the erosion), an interior isolated voxel does not."""
    vol = np.full((7, 7, 1), 10.0, dtype=np.float32)
    vol[0:3, 0:3, 0] = 1000.0          # corner 3x3 block
    vol[5, 5, 0] = 1000.0              # isolated voxel
    m = mask_oracle.get_mask(vol, 3, mu_sigma=(10.0, 1.0))
    assert m[0, 0, 0] and m[3, 3, 0]   # block kept and dilated by one
    assert not m[5, 5, 0]


def test_components_match_flood_fill(rng):
    for m in _tiny_masks(rng, 12):
        comps = mask_oracle.brute_components(m)
        assert mask_oracle.n_components(m) == len(comps)
        lo, hi = 2, 5
        want = np.zeros_like(m)
        for c in comps:
            if lo <= len(c) <= hi:
                for p in c:
                    want[p] = True
        assert np.array_equal(mask_oracle.bwareafilt3(m, (lo, hi)), want)


def test_components_join_through_corners_only():
    """Voxels touching only at a corner are one 26-connected component."""
    m = np.zeros((4, 4, 4), dtype=bool)
    m[0, 0, 0] = m[1, 1, 1] = m[2, 2, 2] = True
    m[0, 3, 3] = True
    assert mask_oracle.n_components(m) == 2
    out = mask_oracle.bwareafilt3(m, (3, 3))       # closed range
    assert out[0, 0, 0] and out[2, 2, 2] and not out[0, 3, 3]
    assert not mask_oracle.bwareafilt3(m, (4, 10)).any()


def test_mean_sq_diff_uint16_wraps():
    a = np.array([0, 5, 65535], dtype=np.uint16)
    b = np.array([1, 2, 0], dtype=np.uint16)
    d = (a.astype(np.int64) - b) % 65536
    want = np.mean((d * d) % 65536)
    assert mask_oracle.mean_sq_diff(a, b) == want


def test_pick_reference_matches_reference_vectors():
    z = np.load(os.path.join(HERE, "golden", "pick_reference.npz"))
    for i, case in enumerate(PICK_CASES):
        frames = pick_frames(i)
        img, top = harness_oracle.pick_initial_reference(frames, max_corr_frames=case["max_corr"])
        assert np.array_equal(top, z[f"top_{i}"]), i
        assert img.dtype == z[f"ref_{i}"].dtype
        assert np.array_equal(img, z[f"ref_{i}"]), i


def test_correlation_of_scaled_copies_is_one():
    rng = np.random.default_rng(3)
    base = rng.uniform(0, 100, size=(3, 16, 16)).astype(np.float32)
    frames = np.stack([base, 2 * base + 5, base])
    cc = harness_oracle.correlation_matrix(frames, spatial_ds=1)
    assert np.allclose(cc, 1.0, atol=1e-5)
