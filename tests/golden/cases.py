"""Inputs of the golden fixtures, regenerated bit-identically from seeds.

Shared by ``make_golden.py`` (which runs the reference on them) and by the
tests (which run the oracle and the HIP path on them).  Volumes are handed to
``getMotion`` in the reference's internal ``(x, y, z)`` layout, i.e. as the
``transpose(2, 1, 0)`` view of a C-order ``(Z, Y, X)`` volume, exactly as the
reference harness does (registration.py:251,313).
"""

from __future__ import annotations

import numpy as np

from wholistic_registration_amd import synth

# (shape_in, shape_out, method, dtype) -- imresize on (x, y, z) arrays
IMRESIZE_CASES = [
    ((37, 29, 5), (18, 14, 5), "bicubic", np.float32),
    ((64, 48, 4), (8, 6, 4), "bicubic", np.float32),
    ((32, 32, 3), (32, 32, 3), "bicubic", np.float32),
    ((16, 12, 3), (32, 24, 3), "bilinear", np.float32),
    ((15, 11, 3), (31, 22, 3), "bilinear", np.float32),
    ((40, 40, 6), (20, 20, 6), "bicubic", np.float64),
    ((33, 70, 2), (4, 8, 2), "bicubic", np.float32),
    ((24, 20, 4), (12, 10, 4), "bicubic", np.uint16),
    ((9, 7, 2), (18, 14, 2), "bicubic", np.float32),
    ((50, 30, 3), (25, 15, 3), "bilinear", np.float32),
]


def imresize_input(i):
    shape_in, _, _, dtype = IMRESIZE_CASES[i]
    rng = np.random.default_rng(500 + i)
    a = rng.uniform(0.0, 1000.0, size=shape_in)
    if dtype == np.uint16:
        return np.rint(a).astype(np.uint16)
    return a.astype(dtype)


def _base_option(layer, r, n_iter, z_ratio, tol=1e-3):
    return {"layer": layer, "iter": n_iter, "r": r, "zRatio": z_ratio, "movRange": 5.0,
            "tol": tol, "save_ite": 100}


def _smooth_penalty(ref_xyz, mask_xyz, raw):
    # getSmPnltNormFctr(dat_ref, option) * smoothPenalty_raw
    # (registration.py:292-297, preprocess.py:42-66), in float32 input dtype
    gx = (ref_xyz[2:] - ref_xyz[:-2]) / 2
    gy = (ref_xyz[:, 2:] - ref_xyz[:, :-2]) / 2
    mx = np.mean(gx[~mask_xyz[1:-1]] ** 2)
    my = np.mean(gy[~mask_xyz[:, 1:-1]] ** 2)
    return ((mx + my) / 2) * raw


def _smooth_init(shape_xyz, amp, seed):
    """A smooth (x, y, z, 3) float32 field used as the warm-start motion."""
    rng = np.random.default_rng(seed)
    modes = synth.draw_modes(rng, 2)
    zyx = (shape_xyz[2], shape_xyz[1], shape_xyz[0])
    d = synth.field_np(zyx, modes, amp)
    return np.ascontiguousarray(d.transpose(2, 1, 0, 3))


# name -> (seed, (Z, Y, X), n_blobs, n_modes, dmax, layer, r, iter, zRatio, flags)
SMALL_CASES = {
    "plain": (101, (6, 40, 48), 40, 2, 2.0, 2, 3, 8, 1.0, {}),
    "masked": (102, (5, 36, 44), 40, 2, 2.0, 1, 4, 7, 3.0,
               {"mask_ref": 0.15, "mask_mov": True, "init": 1.0}),
    "pyfloat": (103, (4, 34, 40), 30, 2, 1.5, 1, 5, 10, 1.0,
                {"pyfloat": True, "tol": 1e-4}),
    "iter1": (104, (4, 32, 32), 20, 2, 1.5, 2, 2, 1, 1.0, {"init": 2.0}),
    "layer0": (105, (3, 30, 26), 20, 2, 1.0, 0, 3, 6, 2.0, {}),
    "odd": (106, (3, 37, 45), 30, 2, 2.0, 2, 2, 9, 1.0, {"init": 0.5}),
}


# The parameter set of the 96x1024x2048 whole-body workload (33x33x1 patches,
# r = 16; 20 iterations per level; a mask_ref ellipsoid over 10 % of the
# voxels; zRatio 1; blob density 2e-3 per voxel) on volumes small enough for
# the reference's full-grid correlates, with every level still holding a
# patch centre (x_L - 1 > r).  The fixtures store input hashes, the level-0
# control-point field, sampled voxels and the per-level error logs.
R16_CASES = {
    "r16_6x80x96": (107, (6, 80, 96), 90, 2, 3.0, 2, 16, 20, 1.0, {"mask_ref": 0.10}),
    "r16_4x288x320": (108, (4, 288, 320), 740, 3, 4.0, 4, 16, 20, 1.0, {"mask_ref": 0.10}),
}


def small_inputs(name):
    spec = SMALL_CASES[name] if name in SMALL_CASES else R16_CASES[name]
    seed, zyx, nb, nm, dmax, layer, r, n_iter, zr, flags = spec
    pair = synth.make_pair_np(seed, zyx, nb, nm, dmax,
                              mask_fraction=flags.get("mask_ref", 0.0))
    mov = pair.moving.transpose(2, 1, 0)
    ref = pair.template.transpose(2, 1, 0)
    opt = _base_option(layer, r, n_iter, zr, flags.get("tol", 1e-3))
    shape_xyz = mov.shape
    mask_ref = (pair.mask.transpose(2, 1, 0) if pair.mask is not None
                else np.zeros(shape_xyz, dtype=bool))
    if flags.get("mask_mov"):
        mm = np.zeros(shape_xyz, dtype=bool)
        mm[shape_xyz[0] // 3: shape_xyz[0] // 3 + 6, 4:12, :] = True
    else:
        mm = np.zeros(shape_xyz, dtype=bool)
    opt["mask_ref"] = mask_ref
    opt["mask_mov"] = mm
    sp = _smooth_penalty(ref, mask_ref, 0.08)
    opt["smoothPenalty"] = float(sp) if flags.get("pyfloat") else sp
    if "init" in flags:
        opt["motion"] = _smooth_init(shape_xyz, flags["init"], seed + 7)
    else:
        opt["motion"] = np.zeros(shape_xyz + (3,))
    return mov, ref, opt


# the small blob pair: 64x128x128 (z, y, x) float32, 2000 blobs,
# 4 modes with max |d| = 3, layer 3, 17x17x1 patches (r = 8), 10 iterations
BLOBS_64x128x128 = dict(seed=0, zyx=(64, 128, 128), n_blobs=2000, n_modes=4, dmax=3.0,
                        layer=3, r=8, iter=10, zRatio=1.0, raw_penalty=0.08)


def pair_64x128x128():
    c = BLOBS_64x128x128
    return synth.make_pair_np(c["seed"], c["zyx"], c["n_blobs"], c["n_modes"], c["dmax"])


def inputs_64x128x128(pair=None):
    c = BLOBS_64x128x128
    pair = pair if pair is not None else pair_64x128x128()
    mov = pair.moving.transpose(2, 1, 0)
    ref = pair.template.transpose(2, 1, 0)
    opt = _base_option(c["layer"], c["r"], c["iter"], c["zRatio"])
    mask = np.zeros(mov.shape, dtype=bool)
    opt["mask_ref"] = mask
    opt["mask_mov"] = mask
    opt["smoothPenalty"] = _smooth_penalty(ref, mask, c["raw_penalty"])
    opt["motion"] = np.zeros(mov.shape + (3,))
    return mov, ref, opt
