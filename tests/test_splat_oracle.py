"""Pin the splat oracle (CPU) with hand-computed projections."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import splat_oracle as S


def test_max_footprint_shift_and_fill():
    coords = np.array([[2.3, 1.6, 4.9], [2.9, 1.1, 5.2], [0.2, 0.2, 9.0]], dtype=np.float32)
    values = np.array([-5.0, 7.0, 100.0], dtype=np.float32)
    out = S.project_max(coords, values, [5.0], xref=6, yref=4, z_window=0.5)
    # samples 0 and 1 cover the 2x2 footprints {2,3}x{1,2}; sample 2 is off-plane
    want = np.zeros((1, 6, 4), dtype=np.float32)
    want[0, 2:4, 1:3] = 7.0
    assert np.array_equal(out, want)
    out = S.project_max(coords, values, [5.0], xref=6, yref=4, z_window=0.5, fill_value=None)
    assert np.isneginf(out[0, 0, 0]) and out[0, 2, 1] == 7.0


def test_max_nearest_downsample_radius():
    coords = np.array([[4.6, 3.4, 0.0]], dtype=np.float32)
    values = np.array([3.0], dtype=np.float32)
    out = S.project_max(coords, values, [0.0], xref=8, yref=8, nearest=True, downsample_xy=2)
    assert out.shape == (1, 4, 4) and out[0, 2, 1] == 3.0 and np.count_nonzero(out) == 1
    out = S.project_max(coords, values, [0.0], xref=8, yref=8, extra_radius=1)
    assert np.count_nonzero(out) == 16 and out[0, 3:7, 2:6].min() == 3.0


def test_weighted_bilinear_and_z_modes():
    coords = np.array([[1.25, 2.5, 3.0]], dtype=np.float32)
    values = np.array([8.0], dtype=np.float32)
    out, w = S.project_weighted(coords, values, [3.0, 3.5], xref=4, yref=4, z_window=1.0, mode="hard")
    assert w[0, 1, 2] == pytest.approx(0.75 * 0.5) and w[0, 2, 3] == pytest.approx(0.25 * 0.5)
    assert np.allclose(out[0][w[0] > 0], 8.0) and np.allclose(out[1][w[1] > 0], 8.0)
    _, wt = S.project_weighted(coords, values, [3.5], xref=4, yref=4, z_window=1.0, mode="triangular")
    assert wt[0, 1, 2] == pytest.approx(0.375 * (1 - 0.5 / (1 + 1e-6)), rel=1e-6)
    _, wg = S.project_weighted(coords, values, [3.5], xref=4, yref=4, z_window=1.0, mode="gaussian")
    assert wg[0, 1, 2] == pytest.approx(0.375 * np.exp(-0.5 * 0.25 / (0.25 + 1e-6)), rel=1e-6)
    # outside [0, xref - 1] is dropped
    out, w = S.project_weighted(np.array([[3.5, 1.0, 3.0]], np.float32), values, [3.0], xref=4, yref=4)
    assert not w.any() and not out.any()
