"""Inputs of the harness fixtures, regenerated bit-identically from seeds.

Frame stacks for ``pick_initial_reference``: a smooth random scene plus
per-frame shifts and noise, so that the correlations differ clearly between
frames.
"""

from __future__ import annotations

import numpy as np

# (T, Z, Y, X), dtype, max_corr_frames
PICK_CASES = [
    dict(shape=(12, 5, 40, 36), dtype=np.uint16, max_corr=4),
    dict(shape=(9, 7, 33, 50), dtype=np.float32, max_corr=20),
    dict(shape=(6, 2, 24, 24), dtype=np.uint16, max_corr=3),
    dict(shape=(10, 24, 28), dtype=np.float32, max_corr=5),
    dict(shape=(1, 4, 16, 16), dtype=np.uint16, max_corr=3),
]


def pick_frames(i):
    case = PICK_CASES[i]
    shape = case["shape"]
    rng = np.random.default_rng(700 + i)
    T = shape[0]
    spatial = shape[1:]
    base = rng.uniform(50.0, 1000.0, size=tuple(s + 8 for s in spatial))
    frames = []
    for t in range(T):
        off = rng.integers(0, 5, size=len(spatial))
        sl = tuple(slice(o, o + s) for o, s in zip(off, spatial))
        noise = rng.normal(0.0, 20.0 + 30.0 * t, size=spatial)
        frames.append(base[sl] + noise)
    a = np.clip(np.asarray(frames), 0, 65535)
    if case["dtype"] == np.uint16:
        return np.rint(a).astype(np.uint16)
    return a.astype(np.float32)


def harness_stack(seed=900, shape_zyx=(6, 48, 56), n_blobs=60, frames=2, dmax=1.5, uint16=False):
    """(template (Z, Y, X), moving frames (T, Z, Y, X)) for the harness tests:
    one blob scene, each frame pull-warped by its own smooth field, Poisson
    noise everywhere, and a few very bright spots so that the automatic masks
    hold components of several sizes."""
    from wholistic_registration_amd import synth
    rng = np.random.default_rng(seed)
    blobs = synth.draw_blobs(rng, shape_zyx, n_blobs)
    clean = synth.render_template_np(shape_zyx, blobs)
    spots = rng.integers(0, shape_zyx, size=(4, 3))
    for k, (z, y, x) in enumerate(spots):
        clean[z, max(y - k, 0):y + 1, max(x - k, 0):x + 1] += 20000.0
    tmpl = rng.poisson(clean.astype(np.float64)).astype(np.float32)
    movs = []
    for _ in range(frames):
        d = synth.field_np(shape_zyx, synth.draw_modes(rng, 2), dmax)
        warped = synth.pull_warp_np(clean, -d)
        movs.append(rng.poisson(np.maximum(warped, 0).astype(np.float64)).astype(np.float32))
    movs = np.asarray(movs)
    if uint16:
        tmpl = np.clip(np.rint(tmpl), 0, 65535).astype(np.uint16)
        movs = np.clip(np.rint(movs), 0, 65535).astype(np.uint16)
    return tmpl, movs
