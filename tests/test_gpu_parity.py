"""Parity of the HIP path (called through the C ABI) with the oracle and the
reference's golden vectors.

Tolerances (FP32 path, stated by the workload definition):
  * displacement field: max |d_gpu - d_oracle| <= 1e-3 voxel;
  * warped volume: relative L2 <= 1e-5;
  * resampling (float64 tables and accumulation): relative 1e-12 against the
    reference's float64 imresize output, i.e. equal up to the last bits.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import lk_oracle as O
from oracle.metrics import epe_stats, rel_l2
from tests.golden.cases import (BLOBS_64x128x128, IMRESIZE_CASES, R16_CASES, SMALL_CASES,
                                inputs_64x128x128, imresize_input, small_inputs)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIELD_TOL = 1e-3
WARP_TOL = 1e-5


def _dev():
    from wholistic_registration_amd import device
    return device


def _to_zyx(t_xyz, cuda):
    import torch
    a = np.ascontiguousarray(np.asarray(t_xyz).transpose(2, 1, 0))
    return torch.from_numpy(a).to(cuda)


@pytest.mark.parametrize("i", range(len(IMRESIZE_CASES)))
def test_imresize_matches_reference_vectors(cuda, i):
    g = np.load(os.path.join(GOLDEN, "imresize_cases.npz"))
    shape_in, shape_out, method, _ = IMRESIZE_CASES[i]
    a = imresize_input(i)
    out = _dev().imresize_xy(_to_zyx(a, cuda), shape_out[0], shape_out[1], method)
    got = out.cpu().numpy().transpose(2, 1, 0)
    ref = g[f"out_{i}"]
    assert got.shape == ref.shape
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-12 * scale


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_pyramid_level_matches_oracle(cuda, level):
    mov, _, _ = small_inputs("plain")
    nx, ny, nz = mov.shape
    shp = O.level_shape(mov.shape, level)
    want = O.imresize(mov, shp).astype(np.float32)
    got = _dev().pyramid_level(_to_zyx(mov, cuda), level).cpu().numpy().transpose(2, 1, 0)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dtype", [np.float32, np.uint16])
def test_correct_motion_matches_oracle(cuda, dtype, rng):
    vol = rng.uniform(0, 1000, size=(61, 47, 9)).astype(dtype)
    m = rng.normal(scale=3.0, size=vol.shape + (3,)).astype(np.float32)
    m[:3, :, :, 0] -= 10.0                       # coordinates beyond the volume
    m[-2:, :, :, 1] += 60.0
    import torch
    got = _dev().correct_motion(_to_zyx(vol, cuda), torch.from_numpy(
        np.ascontiguousarray(m.transpose(2, 1, 0, 3))).to(cuda)).cpu().numpy().transpose(2, 1, 0)
    want = O.correct_motion(vol, m)
    assert rel_l2(got, want) <= WARP_TOL
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-3)


def test_correct_motion_identity_is_exact(cuda, rng):
    import torch
    vol = rng.uniform(0, 1000, size=(33, 20, 5)).astype(np.float32)
    z = torch.zeros((5, 20, 33, 3), dtype=torch.float32, device=cuda)
    got = _dev().correct_motion(_to_zyx(vol, cuda), z).cpu().numpy().transpose(2, 1, 0)
    np.testing.assert_array_equal(got, vol)


def _run_gpu(mov, ref, opt, cuda, schedule=None, stop_kernels=False):
    def mask(key):
        m = np.asarray(opt[key], dtype=np.float32)
        return _to_zyx(m, cuda) if np.any(m) else None
    init = opt.get("motion")
    init_t = None
    if init is not None and np.any(init):
        import torch
        init_t = torch.from_numpy(np.ascontiguousarray(
            np.asarray(init, np.float32).transpose(2, 1, 0, 3))).to(cuda)
    res = _dev().get_motion(_to_zyx(mov, cuda), _to_zyx(ref, cuda), layer=opt["layer"],
                            iter=opt["iter"], r=opt["r"], zRatio=opt["zRatio"],
                            movRange=opt["movRange"], smoothPenalty=opt["smoothPenalty"],
                            tol=opt["tol"], mask_ref=mask("mask_ref"), mask_mov=mask("mask_mov"),
                            motion_init=init_t, schedule=schedule, stop_kernels=stop_kernels)
    return res, res.motion.cpu().numpy().transpose(2, 1, 0, 3)


@pytest.mark.parametrize("name", sorted(SMALL_CASES))
def test_get_motion_forced_schedule_matches_oracle(cuda, name):
    """Same iteration schedule as the oracle: field within 1e-3 voxel and the
    error logs within float32 round-off."""
    mov, ref, opt = small_inputs(name)
    ex = {}
    want, err, _, log = O.get_motion(mov, ref, dict(opt), extras=ex, keep_images=False)
    res, got = _run_gpu(mov, ref, opt, cuda, schedule=ex["schedule"])
    assert res.schedule() == ex["schedule"]
    assert np.abs(got - want).max() <= FIELD_TOL
    for L in range(opt["layer"], -1, -1):
        n = res.records[L]
        np.testing.assert_allclose(res.current_error[L, :n], log[f"layer_{L}"]["currentError"],
                                   rtol=2e-5)
    assert res.final_error == pytest.approx(err, rel=2e-5)


@pytest.mark.parametrize("stop_kernels", [False, True])
@pytest.mark.parametrize("name", sorted(SMALL_CASES))
def test_get_motion_free_running_matches_reference(cuda, name, stop_kernels):
    """Stop rules evaluated on the device reproduce the reference's schedule
    and field (golden vectors of the reference run), whether they run inside
    the launches that follow them or in their own kernels."""
    g = np.load(os.path.join(GOLDEN, f"getmotion_{name}.npz"))
    mov, ref, opt = small_inputs(name)
    res, got = _run_gpu(mov, ref, opt, cuda, stop_kernels=stop_kernels)
    for L in range(opt["layer"], -1, -1):
        assert res.records[L] == len(g[f"L{L}_currentError"]), L
        assert res.updates[L] == len(g[f"L{L}_max_diff_motion"]), L
    assert np.abs(got - g["motion"]).max() <= FIELD_TOL


@pytest.mark.parametrize("stop_kernels", [False, True])
def test_get_motion_64x128x128_matches_reference(cuda, stop_kernels):
    g = np.load(os.path.join(GOLDEN, "getmotion_64x128x128.npz"))
    mov, ref, opt = inputs_64x128x128()
    res, got = _run_gpu(mov, ref, opt, cuda, stop_kernels=stop_kernels)
    for L in range(BLOBS_64x128x128["layer"], -1, -1):
        assert res.records[L] == len(g[f"L{L}_currentError"]), L
    r = BLOBS_64x128x128["r"]
    nx, ny, nz = mov.shape
    xg = np.arange(r, nx - 1, 2 * r + 1)
    yg = np.arange(r, ny - 1, 2 * r + 1)
    assert np.abs(got[np.ix_(xg, yg, np.arange(nz))] - g["motion_cp"]).max() <= FIELD_TOL
    idx = g["sample_idx"]
    assert np.abs(got[idx[:, 0], idx[:, 1], idx[:, 2]] - g["sample_motion"]).max() <= FIELD_TOL
    assert res.final_error == pytest.approx(float(g["current_error"]), rel=1e-5)


@pytest.mark.parametrize("stop_kernels", [False, True])
@pytest.mark.parametrize("name", sorted(R16_CASES))
def test_get_motion_r16_matches_reference(cuda, name, stop_kernels):
    """The whole-body workload's parameter set (r 16, 20 iterations, mask_ref)
    on the reference's own runs: free-running schedule equal, level-0 field
    within 1e-3 voxel at the control points and sampled voxels."""
    g = np.load(os.path.join(GOLDEN, f"getmotion_{name}.npz"))
    mov, ref, opt = small_inputs(name)
    res, got = _run_gpu(mov, ref, opt, cuda, stop_kernels=stop_kernels)
    for L in range(opt["layer"], -1, -1):
        assert res.records[L] == len(g[f"L{L}_currentError"]), L
        assert res.updates[L] == len(g[f"L{L}_max_diff_motion"]), L
        np.testing.assert_allclose(res.current_error[L, :res.records[L]], g[f"L{L}_currentError"], rtol=2e-5)
    r = opt["r"]
    nx, ny, nz = mov.shape
    xg = np.arange(r, nx - 1, 2 * r + 1)
    yg = np.arange(r, ny - 1, 2 * r + 1)
    assert np.abs(got[np.ix_(xg, yg, np.arange(nz))] - g["motion_cp"]).max() <= FIELD_TOL
    idx = g["sample_idx"]
    assert np.abs(got[idx[:, 0], idx[:, 1], idx[:, 2]] - g["sample_motion"]).max() <= FIELD_TOL
    assert res.final_error == pytest.approx(float(g["current_error"]), rel=1e-5)


@pytest.mark.parametrize("name", sorted(R16_CASES))
def test_get_motion_r16_forced_schedule_matches_oracle(cuda, name):
    """Same cases with the oracle's schedule forced: the whole field within
    1e-3 voxel of the oracle, its warp within relative L2 1e-5."""
    mov, ref, opt = small_inputs(name)
    ex = {}
    want, err, _, _ = O.get_motion(mov, ref, dict(opt), extras=ex, keep_images=False)
    res, got = _run_gpu(mov, ref, opt, cuda, schedule=ex["schedule"])
    assert res.schedule() == ex["schedule"]
    assert np.abs(got - want).max() <= FIELD_TOL
    import torch
    warped = _dev().correct_motion(_to_zyx(mov, cuda), torch.from_numpy(
        np.ascontiguousarray(got.transpose(2, 1, 0, 3))).to(cuda)).cpu().numpy().transpose(2, 1, 0)
    assert rel_l2(warped, O.correct_motion(mov, got)) <= WARP_TOL


def test_drop_in_module_round_trip(cuda):
    """getMotion/correctMotion of the drop-in module on host arrays, as
    wbi_registration_3d calls them (registration.py:320-326)."""
    from wholistic_registration_amd import calFlow3d_Wei_v1 as wr
    mov, ref, opt = small_inputs("masked")
    o = dict(opt)
    motion, err, coords, log = wr.getMotion(mov, ref, o)
    assert motion.shape == mov.shape + (3,) and motion.dtype == np.float32
    assert o["mask_ref"].dtype == np.float32
    g = np.load(os.path.join(GOLDEN, "getmotion_masked.npz"))
    assert np.abs(motion - g["motion"]).max() <= FIELD_TOL
    assert sorted(log) == ["layer_0", "layer_1"]
    nx = mov.shape[0]
    np.testing.assert_allclose(coords[..., 0], motion[..., 0] + np.arange(nx)[:, None, None])
    warped = wr.correctMotion(mov, motion)
    assert rel_l2(warped, O.correct_motion(mov, motion)) <= WARP_TOL


def test_identity_pair_gives_exact_zero(cuda):
    import torch
    from wholistic_registration_amd import synth
    pair = synth.make_pair_np(3, (8, 96, 128), 300, 2, 2.0)
    vol = torch.from_numpy(pair.template).to(cuda)
    res = _dev().get_motion(vol, vol, layer=2, iter=6, r=8, zRatio=1.0, movRange=5.0,
                            smoothPenalty=10.0, tol=1e-3)
    assert torch.count_nonzero(res.motion).item() == 0
    assert res.final_error == 0.0


def test_deterministic(cuda):
    import torch
    mov, ref, opt = inputs_64x128x128()
    a, ma = _run_gpu(mov, ref, opt, cuda)
    b, mb = _run_gpu(mov, ref, opt, cuda)
    np.testing.assert_array_equal(ma, mb)
    np.testing.assert_array_equal(a.current_error, b.current_error)
    del torch


def test_epe_against_fp64_oracle_small(cuda):
    """The workload's accuracy gate, on the 'plain' case: mean endpoint error
    of the FP32 GPU field against the all-FP64 oracle <= 0.1 voxel."""
    mov, ref, opt = small_inputs("plain")
    want, _, _, _ = O.get_motion(mov, ref, dict(opt), mode="fp64", keep_images=False)
    _, got = _run_gpu(mov, ref, opt, cuda)
    assert epe_stats(got, want)["epe_mean"] <= 0.1


def test_invalid_arguments_raise(cuda):
    import torch
    v = torch.zeros((4, 8, 8), dtype=torch.float32, device=cuda)
    with pytest.raises(ValueError):      # level 3 of an 8-wide volume is 1 voxel: no patch fits
        _dev().get_motion(v, v, layer=3, iter=2, r=2, zRatio=1.0, movRange=5.0,
                          smoothPenalty=1.0, tol=1e-3)
    with pytest.raises(ValueError):
        _dev().get_motion(v, v, layer=0, iter=0, r=1, zRatio=1.0, movRange=5.0,
                          smoothPenalty=1.0, tol=1e-3)


def test_resampling_bit_exact(cuda):
    """Unfused float64 products and adds in tap order, with the reference's
    tables: the resampled volumes equal the reference's (and the oracle's)
    bit for bit."""
    g = np.load(os.path.join(GOLDEN, "imresize_cases.npz"))
    for i, (shape_in, shape_out, method, _) in enumerate(IMRESIZE_CASES):
        out = _dev().imresize_xy(_to_zyx(imresize_input(i), cuda), shape_out[0], shape_out[1], method)
        got = out.cpu().numpy().transpose(2, 1, 0)
        ref = g[f"out_{i}"]
        assert np.array_equal(got, ref), f"case {i}: max |diff| {np.abs(got - ref).max():.3g}"
    mov, _, _ = small_inputs("plain")
    for level in (1, 2, 3):
        want = O.imresize(mov, O.level_shape(mov.shape, level)).astype(np.float32)
        got = _dev().pyramid_level(_to_zyx(mov, cuda), level).cpu().numpy().transpose(2, 1, 0)
        assert np.array_equal(got, want), f"level {level}: max |diff| {np.abs(got - want).max():.3g}"
