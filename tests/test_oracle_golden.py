"""The oracle against the golden vectors written by running the reference
(tests/golden/make_golden.py).  CPU only."""

from __future__ import annotations

import hashlib
import os

import numpy as np
import pytest

from oracle import lk_oracle as O
from tests.golden.cases import (BLOBS_64x128x128, IMRESIZE_CASES, R16_CASES, SMALL_CASES,
                                inputs_64x128x128, imresize_input, small_inputs)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name))


@pytest.mark.parametrize("i", range(len(IMRESIZE_CASES)))
def test_imresize_bit_exact(i):
    g = _load("imresize_cases.npz")
    a = imresize_input(i)
    np.testing.assert_array_equal(a, g[f"in_{i}"])
    _, shape_out, method, _ = IMRESIZE_CASES[i]
    out = O.imresize(a, shape_out, method)
    np.testing.assert_array_equal(out, g[f"out_{i}"])


def _check_log(log, g, levels, exact):
    for L in range(levels, -1, -1):
        e = log[f"layer_{L}"]
        for key in ("diffError", "penaltyError", "currentError", "max_diff_motion"):
            ref = g[f"L{L}_{key}"]
            got = np.asarray(e[key], dtype=np.float64)
            assert got.shape == ref.shape, (L, key, got.shape, ref.shape)
            if exact:
                np.testing.assert_array_equal(got, ref)
            else:
                np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("name", sorted(SMALL_CASES))
@pytest.mark.parametrize("sums", ["correlate", "window"])
def test_get_motion_small_cases(name, sums):
    g = _load(f"getmotion_{name}.npz")
    mov, ref, opt = small_inputs(name)
    np.testing.assert_array_equal(mov, g["mov"])
    np.testing.assert_array_equal(ref, g["ref"])
    motion, err, coords, log = O.get_motion(mov, ref, dict(opt), sums=sums)
    exact = sums == "correlate"
    if exact:
        np.testing.assert_array_equal(motion, g["motion"])
        assert err == float(g["current_error"])
    else:
        np.testing.assert_allclose(motion, g["motion"], rtol=0, atol=1e-5)
        assert err == pytest.approx(float(g["current_error"]), rel=1e-9)
    _check_log(log, g, opt["layer"], exact)
    # coords is motion + grid in float32
    nx, ny, nz = mov.shape
    assert coords.dtype == np.float32 and coords.shape == (nx, ny, nz, 3)
    np.testing.assert_array_equal(coords[..., 0],
                                  motion[..., 0] + np.arange(nx, dtype=np.float32)[:, None, None])


def test_get_motion_64x128x128():
    """The 64x128x128 blob pair (, layer 3, r 8, 10 iterations) against the
    reference's run on the same bytes."""
    g = _load("getmotion_64x128x128.npz")
    mov, ref, opt = inputs_64x128x128()
    assert hashlib.sha256(np.ascontiguousarray(mov).tobytes()).hexdigest() == str(g["mov_sha"])
    assert hashlib.sha256(np.ascontiguousarray(ref).tobytes()).hexdigest() == str(g["ref_sha"])
    assert opt["smoothPenalty"] == float(g["smooth_penalty"])
    motion, err, _, log = O.get_motion(mov, ref, dict(opt), keep_images=False)
    r = BLOBS_64x128x128["r"]
    nx, ny, nz = mov.shape
    xg = np.arange(r, nx - 1, 2 * r + 1)
    yg = np.arange(r, ny - 1, 2 * r + 1)
    np.testing.assert_allclose(motion[np.ix_(xg, yg, np.arange(nz))], g["motion_cp"], rtol=0, atol=1e-5)
    idx = g["sample_idx"]
    np.testing.assert_allclose(motion[idx[:, 0], idx[:, 1], idx[:, 2]], g["sample_motion"], rtol=0,
                               atol=1e-5)
    assert err == pytest.approx(float(g["current_error"]), rel=1e-9)
    for L in range(BLOBS_64x128x128["layer"], -1, -1):
        assert len(log[f"layer_{L}"]["currentError"]) == len(g[f"L{L}_currentError"])


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(R16_CASES))
def test_get_motion_r16_cases(name):
    """The whole-body workload's parameter set (33x33x1 patches, 20
    iterations per level, a mask_ref ellipsoid) against the reference's runs
    on the same bytes: schedule, error logs, and the level-0 field at the
    control points and at sampled voxels."""
    g = _load(f"getmotion_{name}.npz")
    mov, ref, opt = small_inputs(name)
    assert _sha(mov) == str(g["mov_sha"]) and _sha(ref) == str(g["ref_sha"])
    assert _sha(np.asarray(opt["mask_ref"])) == str(g["mask_ref_sha"])
    assert opt["smoothPenalty"] == float(g["smooth_penalty"])
    motion, err, _, log = O.get_motion(mov, ref, dict(opt), keep_images=False)
    _check_log(log, g, opt["layer"], exact=False)
    r = opt["r"]
    nx, ny, nz = mov.shape
    xg = np.arange(r, nx - 1, 2 * r + 1)
    yg = np.arange(r, ny - 1, 2 * r + 1)
    np.testing.assert_allclose(motion[np.ix_(xg, yg, np.arange(nz))], g["motion_cp"], rtol=0, atol=1e-5)
    idx = g["sample_idx"]
    np.testing.assert_allclose(motion[idx[:, 0], idx[:, 1], idx[:, 2]], g["sample_motion"], rtol=0,
                               atol=1e-5)
    assert err == pytest.approx(float(g["current_error"]), rel=1e-9)
