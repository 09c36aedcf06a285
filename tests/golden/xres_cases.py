"""Inputs of the cross-resolution (``getMotion_v2``) fixtures, regenerated
bit-identically from seeds.

A cross-resolution pair is a moving volume on a coarse-z grid and a reference
volume on a finer-z grid (calFlowCrossResolution.py:2470-2534): moving voxel
``(x, y, k)`` sits at reference coordinates ``(x, y, k * zRatio / zRatio_HR)``
before motion.  The moving volume here is the noise-free reference sampled
there plus a smooth displacement (in reference index units), with Poisson
noise; optional planted artefacts make the wrong-region detector fire.

Arrays are in the solver's ``(x, y, z)`` layout (the ``transpose(2, 1, 0)``
view of ``(Z, Y, X)`` C-order volumes).
"""

from __future__ import annotations

import numpy as np
from wholistic_registration_amd import synth

# name -> dict of the case
XRES_CASES = {
    # no wrong-region pass, no masks, two levels
    "plain": dict(seed=301, mov_xyz=(64, 48, 8), ref_xyz=(64, 48, 24), z_ratio=3.0, z_ratio_hr=1.0,
                  n_blobs=90, dmax=1.5, layer=1, r=3, iter=8, wrong=False),
    # wrong-region correction on, a bright artefact in the moving volume that the
    # reference does not hold (detected, trap mask built, corrected run kept or not)
    "wrong": dict(seed=302, mov_xyz=(72, 56, 6), ref_xyz=(72, 56, 18), z_ratio=3.0, z_ratio_hr=1.0,
                  n_blobs=110, dmax=1.5, layer=1, r=3, iter=8, wrong=True, artefacts=3),
    # masks on both sides, a warm start, three levels, reference wider in y
    "masked": dict(seed=303, mov_xyz=(60, 52, 6), ref_xyz=(60, 60, 12), z_ratio=2.0, z_ratio_hr=1.0,
                   n_blobs=100, dmax=1.2, layer=2, r=2, iter=7, wrong=True, artefacts=2,
                   mask_ref=0.12, mask_mov=True, init=0.6),
    # the 'highresidual' exclusion mode, non-unit zRatio_HR
    "highres": dict(seed=304, mov_xyz=(64, 64, 5), ref_xyz=(64, 64, 10), z_ratio=4.0, z_ratio_hr=2.0,
                    n_blobs=120, dmax=1.5, layer=1, r=3, iter=6, wrong=True, artefacts=3,
                    exclude_mode="highresidual"),
}


def _option(c):
    opt = {
        "layer": c["layer"], "iter": c["iter"], "r": c["r"],
        "zRatio": c["z_ratio"], "zRatio_HR": c["z_ratio_hr"],
        "smoothPenalty": 0.05, "movRange": 3.0, "tol": 1e-3,
        "wrong_region_enable": bool(c["wrong"]),
    }
    if "exclude_mode" in c:
        opt["wrong_region_exclude_mode"] = c["exclude_mode"]
    return opt


def xres_inputs(name):
    """``(mov_xyz, ref_xyz, option)`` of a case; ``option`` holds numpy masks."""
    c = XRES_CASES[name]
    rx, ry, rz = c["ref_xyz"]
    mx, my, mz = c["mov_xyz"]
    ref_zyx = (rz, ry, rx)
    mov_xyz, ref_xyz, rng = synth.make_xres_pair_np(c["seed"], c["mov_xyz"], c["ref_xyz"], c["z_ratio"],
                                                    c["z_ratio_hr"], c["n_blobs"], c["dmax"],
                                                    c.get("artefacts", 0))
    opt = _option(c)
    mref = np.zeros(c["ref_xyz"], dtype=bool)
    if c.get("mask_ref"):
        mref = np.ascontiguousarray(synth.ellipsoid_mask_np(rng, ref_zyx, c["mask_ref"]).transpose(2, 1, 0))
    mmov = np.zeros(c["mov_xyz"], dtype=bool)
    if c.get("mask_mov"):
        mmov[mx // 4: mx // 4 + 8, 5:13, :] = True
    opt["mask_ref"] = mref
    opt["mask_mov"] = mmov
    if c.get("init"):
        f = synth.field_np((mz, my, mx), synth.draw_modes(rng, 2), c["init"])
        opt["motion"] = np.ascontiguousarray(f.transpose(2, 1, 0, 3))
    return mov_xyz, ref_xyz, opt


# z-offset search (FindInitZ_stack_global_fixed_spacing): name -> (seed, (H, W),
# K moving slices, Z reference slices, true z0, spacing, direction, gradient, hann)
FINDZ_CASES = {
    "grad_hann": (311, (96, 80), 6, 40, 9, 3.0, 1, True, True),
    "raw_flat": (312, (64, 72), 5, 30, 17, 2.5, 1, False, False),
    "backward": (313, (80, 64), 4, 36, 28, 4.0, -1, True, True),
}


def findz_inputs(name):
    """``(mov (H, W, K), ref (H, W, Z), delta, direction, use_gradient,
    use_hann)``: moving slices are reference slices z0 + k * spacing *
    direction (rounded) with fresh Poisson noise."""
    seed, (h, w), k_n, z_n, z0, delta, direction, grad, hann = FINDZ_CASES[name]
    rng = np.random.default_rng(seed)
    blobs = synth.draw_blobs(rng, (z_n, w, h), 6 * z_n)
    clean = synth.render_template_np((z_n, w, h), blobs).astype(np.float64)      # (Z, W, H)
    ref = rng.poisson(clean).astype(np.float32).transpose(2, 1, 0)
    zs = [int(np.rint(z0 + k * delta * direction)) for k in range(k_n)]
    mov = np.stack([rng.poisson(clean[z]).astype(np.float32).T for z in zs], axis=2)
    return np.ascontiguousarray(mov), np.ascontiguousarray(ref), delta, direction, grad, hann
