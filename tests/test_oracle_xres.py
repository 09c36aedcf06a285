"""CPU checks of the cross-resolution oracle (oracle/xres_oracle.py) and of
the drop-in's host-side pieces.

* The oracle reproduces the reference's own getMotion_v2 outputs (fixtures
  written by tests/golden/make_golden_xres.py) bit for bit, schedules and
  wrong-region decisions included (from the printed logs).
* The structuring element and the trap mask against the known answers of the
  reference's regression tests (tests/regressions/
  test_cross_resolution_geometry.py), restated from the geometry.
* The outlier detector and the box marking against brute force.
* The spline prefilter restated from its recursion against scipy.
"""

from __future__ import annotations

import os
import re

import numpy as np
import pytest
from scipy import ndimage as ndi

from oracle import xres_oracle as xo
from tests.golden.xres_cases import XRES_CASES, xres_inputs

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _log_schedule(lines):
    """(records, stop) per optimisation and the wrong-region lines of a
    reference log."""
    calls, cur, regions, decisions = [], 0, [], []
    for line in lines:
        if re.search(r"iter=\d+ error=", line):
            k = int(re.search(r"iter=(\d+)", line).group(1))
            if k == 0 and cur:
                calls.append(cur)
            cur = k + 1
        elif "detected wrong regions" in line:
            regions.append(int(line.rsplit(":", 1)[1]))
        elif "accept corrected" in line or "keep original" in line:
            decisions.append("accept" in line)
    if cur:
        calls.append(cur)
    return calls, regions, decisions


@pytest.mark.parametrize("name", list(XRES_CASES))
def test_oracle_reproduces_reference_get_motion_v2(name):
    g = np.load(os.path.join(HERE, f"xres_{name}.npz"))
    mov, ref, opt = xres_inputs(name)
    trace, info = [], {}
    phase, motion, sampled = xo.get_motion_v2(mov, ref, opt, trace=trace, info=info)
    assert np.array_equal(phase, g["phase"])
    assert np.array_equal(motion, g["motion"])
    assert np.array_equal(sampled, g["sampled"])
    calls, regions, decisions = _log_schedule(list(g["log"]))
    assert [t["records"] for t in trace] == calls
    got_regions = [info[L]["regions"] for L in sorted(info, reverse=True) if info[L]]
    assert got_regions == regions
    got_dec = [info[L]["accepted"] for L in sorted(info, reverse=True)
               if info[L] and info[L]["regions"] and info[L]["ratio"] is not None and info[L]["ratio"] >= 0.4]
    assert got_dec == decisions
    # the iteration errors printed by the reference, to its 4 decimals
    errs = [float(re.search(r"error=([\d.]+)", l).group(1)) for l in g["log"] if re.search(r"iter=\d+ error=", l)]
    mine = [e[2] for t in trace for e in t["errors"]]
    assert np.allclose(mine, errs, atol=6e-5, rtol=0)


def test_oracle_input_case_mutation_free():
    mov, ref, opt = xres_inputs("plain")
    before = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in opt.items()}
    xo.get_motion_v2(mov, ref, opt)
    for k, v in before.items():
        assert np.array_equal(opt[k], v) if isinstance(v, np.ndarray) else opt[k] == v


def test_make_ball_isotropic_at_unit_ratio():
    """z_ratio 1 gives the plain ball (reference regression B-090 anchor)."""
    from wholistic_registration_amd.calFlowCrossResolution import _make_ball
    for r in (1, 2, 3, 8):
        s = np.arange(-r, r + 1)
        X, Y, Z = np.meshgrid(s, s, s, indexing="ij")
        want = (X ** 2 + Y ** 2 + Z ** 2) <= r ** 2 + 1e-6
        assert np.array_equal(xo.make_ball(r, 1.0), want)
        assert np.array_equal(_make_ball(r, 1.0), want)


@pytest.mark.parametrize("z_ratio,extent", [(3.0, 1), (2.0, 3), (1.0, 5), (0.5, 9), (0.25, 17), (0.125, 33)])
def test_make_ball_z_extent_is_physical(z_ratio, extent):
    """A radius-2 ball reaches 2 / z_ratio planes (rounded up, at least 1)."""
    from wholistic_registration_amd.calFlowCrossResolution import _make_ball
    for f in (xo.make_ball, _make_ball):
        ball = f(2, z_ratio)
        assert int(np.count_nonzero(ball.any(axis=(0, 1)))) == extent


def test_make_ball_rejects_degenerate_ratio_and_caps_depth():
    from wholistic_registration_amd.calFlowCrossResolution import _make_ball
    for f in (xo.make_ball, _make_ball):
        for bad in (0.0, -1.0, float("nan"), float("inf")):
            with pytest.raises(ValueError, match="positive, finite"):
                f(2, bad)
        ball = f(2, 1e-12)
        assert ball.shape[2] == 2 * 64 + 1 and ball.any(axis=(0, 1)).all()


@pytest.mark.parametrize("z_ratio,reach", [(0.25, 8), (0.125, 16)])
def test_trap_mask_reach_follows_physical_z(z_ratio, reach):
    """The reference's phantom: a 5x5x3 seed inside a bright slab; the trap
    grows 2 voxels in x, y and 2 / z_ratio planes in z (two growth steps of a
    physical-radius-1 ball)."""
    shape = (64, 64, 64)
    bad = np.zeros(shape, bool)
    bad[30:35, 30:35, 30:33] = True
    ph = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).astype(np.float32)
    rng = np.random.default_rng(0)
    data = np.full(shape, 100.0, np.float32)
    data[16:49, 16:49, :] = 900.0
    data = (data + rng.normal(0.0, 10.0, shape)).astype(np.float32)
    trap = xo.trap_mask(bad, ph, data, z_ratio, 2, 1.0, 2.5)
    seed = np.argwhere(bad)
    lo, hi = np.argwhere(trap).min(0), np.argwhere(trap).max(0)
    assert (seed.min(0) - lo).tolist() == [2, 2, reach]
    assert (hi - seed.max(0)).tolist() == [2, 2, reach]
    assert int(trap.sum()) < 0.05 * trap.size


def test_robust_significant_brute_force():
    rng = np.random.default_rng(4)
    err = rng.gamma(2.0, 1.0, (15, 12, 3))
    err[5, 6, 1] = 40.0
    err[:, :, 2] = 1.0                                   # MAD 0: nothing flagged
    sig = xo.robust_significant(err, 3.0)
    for z in range(3):
        inner = err[1:-1, 1:-1, z].ravel()
        med = np.median(inner)
        mad = np.median(np.abs(inner - med))
        for i in range(1, 14):
            for j in range(1, 11):
                want = mad >= 1e-12 and abs(err[i, j, z] - med) / (1.4826 * mad + 1e-12) > 3.0
                assert sig[i, j, z] == want
        assert not sig[0, :, z].any() and not sig[:, 0, z].any()
    assert sig[5, 6, 1] and not sig[..., 2].any()


def test_bad_region_boxes_brute_force():
    """Components of >= 2 significant control points, 4-connected within a
    plane; each marks its control-point box widened by ceil(1.5 r)."""
    r = 3
    nx, ny, nz = 80, 70, 2
    xg, yg, cps = xo.control_grid(nx, ny, nz, r)
    err = np.ones((len(xg), len(yg), nz)) + np.random.default_rng(1).uniform(0, 0.1, (len(xg), len(yg), nz))
    err[3, 4, 0] = err[4, 4, 0] = 50.0                    # a 2-point component
    err[7, 7, 0] = 50.0                                   # a single point: dropped
    err[5, 2, 1] = err[5, 3, 1] = err[6, 3, 1] = 50.0     # an L of 3
    mask, n = xo.bad_region_mask(err, r, xg, yg, nx, ny)
    assert n == 2
    dx = int(np.ceil(1.5 * r))
    want = np.zeros((nx, ny, nz), bool)
    want[xg[3] - dx:xg[4] + dx + 1, yg[4] - dx:yg[4] + dx + 1, 0] = True
    want[xg[5] - dx:xg[6] + dx + 1, yg[2] - dx:yg[3] + dx + 1, 1] = True
    assert np.array_equal(mask, want)


def _prefilter_line(c, z=np.sqrt(3.0) - 2.0):
    """Cubic B-spline prefilter of one line: gain, causal pass from the
    half-sample symmetric initial value, anticausal pass (Unser, Aldroubi and
    Eden, IEEE TSP 41(2), 1993)."""
    c = c.astype(np.float64) * ((1 - z) * (1 - 1 / z))
    n = len(c)
    zn = z ** n
    s = c[0] + zn * c[-1]
    zi = z
    for i in range(1, n):
        s += zi * (c[i] + zn * c[n - 1 - i])
        zi *= z
    c[0] = s * z / (1 - zn * zn) + c[0]
    for i in range(1, n):
        c[i] += z * c[i - 1]
    c[-1] *= z / (z - 1)
    for i in range(n - 2, -1, -1):
        c[i] = z * (c[i + 1] - c[i])
    return c


def test_spline_prefilter_restatement_matches_scipy():
    """The prefilter the HIP kernel runs (padding by 12 edge voxels, then this
    recursion per axis) against scipy's spline_filter1d(mode='nearest') and
    the full interpolant against map_coordinates(order=3, mode='nearest')."""
    rng = np.random.default_rng(2)
    a = rng.uniform(0, 1000, 41)
    assert np.abs(_prefilter_line(a) - ndi.spline_filter1d(a, 3, mode="nearest")).max() < 1e-9
    vol = rng.uniform(0, 1000, (8, 7, 5))
    c = np.pad(vol, 12, mode="edge")
    for ax in range(3):
        c = np.apply_along_axis(_prefilter_line, ax, c)
    pts = np.concatenate([rng.uniform(-3, 10, (60, 3)), rng.uniform(-25, 30, (20, 3))])
    got = []
    for p in pts:
        w, idx = [], []
        for d in range(3):
            x = p[d] + 12
            f = np.floor(x)
            t = x - f
            w.append([(1 - t) ** 3 / 6, (4 - 6 * t * t + 3 * t ** 3) / 6, (1 + 3 * t + 3 * t * t - 3 * t ** 3) / 6,
                      t ** 3 / 6])
            idx.append(np.clip(np.arange(int(f) - 1, int(f) + 3), 0, c.shape[d] - 1))
        got.append(sum(w[0][i] * w[1][j] * w[2][k] * c[idx[0][i], idx[1][j], idx[2][k]]
                       for i in range(4) for j in range(4) for k in range(4)))
    want = ndi.map_coordinates(vol, pts.T, order=3, mode="nearest")
    assert np.abs(np.array(got) - want).max() < 1e-8


@pytest.mark.parametrize("name", ["grad_hann", "raw_flat", "backward"])
def test_oracle_find_init_z_against_reference_scores_and_planted_start(name):
    """The slice features and ZNCC scores equal the reference's own
    (grad_mag_2d, _global_zncc_scores) bit for bit; the search recovers the
    planted start, and brute force over all starts agrees."""
    from tests.golden.xres_cases import FINDZ_CASES, findz_inputs
    mov, ref, delta, direction, grad, hann = findz_inputs(name)
    g = np.load(os.path.join(HERE, f"findz_{name}.npz"))
    z, dbg = xo.find_init_z(mov, ref, delta, grad, hann, direction, return_debug=True)
    assert np.array_equal(dbg["scores"], g["scores"])
    if grad:
        assert np.array_equal(xo.grad_mag_2d(mov[:, :, 0]), g["mov_feat"])
        assert np.array_equal(xo.grad_mag_2d(ref[:, :, 0]), g["ref_feat"])
    z0 = FINDZ_CASES[name][4]
    assert dbg["best_z0"] == z0
    k = np.arange(mov.shape[2])
    offs = (k.astype(np.float32) * np.float32(delta) * np.float32(direction))
    assert np.array_equal(z, (z0 + offs).astype(np.uint16))
    starts = [s for s in range(ref.shape[2]) if 0 <= s + offs.min() and s + offs.max() <= ref.shape[2] - 1]
    sums = [np.sum(dbg["scores"][k, np.rint(s + offs).astype(int)]) for s in starts]
    assert starts[int(np.argmax(sums))] == z0
