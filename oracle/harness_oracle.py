"""numpy restatement of the registration harness around the solver.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
harness path.  The product path never calls it.

Functions and the reference code they follow (paths under
``src/wholistic_registration/utils/``):

``transform``  registration.py:26-36 (same numpy expressions).
``normalize_std``  preprocess.py:288-293 (same numpy expressions).
``pick_initial_reference``  reference.py:19-99 with ``cp`` = numpy (the
    reference's CPU fallback, utils/__init__.py:12-65): the middle three z
    slices every ``spatial_ds`` voxels, per-frame demeaning, the Gram matrix,
    normalised correlations, the frame with the best mean of its
    ``max_corr_frames`` top correlations, and the mean of the frames most
    correlated with it.  Pinned by golden vectors written by running the
    reference's function (tests/golden/make_golden_harness.py).
``wbi_registration_3d``  registration.py:192-352 for single-channel configs
    with a given reference image, composed of the mask oracle, the solver
    oracle (lk_oracle.get_motion / correct_motion) and the error metric; the
    caller can force the solver's iteration schedule per frame.
"""

from __future__ import annotations

import numpy as np

from . import lk_oracle
from . import mask_oracle


def transform(image, k=1, method="raw"):
    """registration.py:26-36."""
    if method == "raw":
        return k * image
    if method == "sqrt":
        return k * np.sqrt(image)
    if method == "log2":
        return k * np.log2(1 + image)
    if method == "log10":
        return k * np.log10(1 + image)
    raise ValueError(f"Unknown method to process the image:{method}")


def normalize_std(mean, std, image):
    """preprocess.py:288-293."""
    mean_prev = np.mean(image)
    std_prev = np.std(image)
    image_normalized = (image - mean_prev) / std_prev
    return image_normalized * std + mean


def correlation_matrix(frames, spatial_ds=4):
    """reference.py:31-78: normalised correlations of the sampled frames."""
    if frames.ndim == 4:
        Z = frames.shape[1]
        zs = slice(Z // 2 - 1, Z // 2 + 2) if Z >= 3 else slice(0, Z)
        used = frames[:, zs, :, :]
    elif frames.ndim == 3:
        used = frames[:, None, :, :]
    else:
        raise ValueError("frames must be [T,Y,X] or [T,Z,Y,X]")
    small = np.ascontiguousarray(used[:, :, ::spatial_ds, ::spatial_ds], dtype=np.float32)
    flat = small.reshape(small.shape[0], -1)
    demeaned = flat - flat.mean(axis=1, keepdims=True)
    cc = demeaned @ demeaned.T
    diag = np.sqrt(np.diag(cc)) + 1e-12
    return cc / (diag[:, None] * diag[None, :])


def pick_initial_reference(frames, max_corr_frames=20, downsample=1, spatial_ds=4):
    """reference.py:19-99 on numpy arrays: ``(refImg, top_indices)``."""
    T = frames.shape[0]
    cc = correlation_matrix(frames, spatial_ds)
    ncorr = min(max_corr_frames, T - 1)
    if ncorr < 1:
        return frames.mean(axis=0), np.arange(T)
    cc_sort = -np.sort(-cc, axis=1)
    best = cc_sort[:, 1:ncorr + 1].mean(axis=1)
    imax = int(np.argmax(best))
    indsort = np.argsort(-cc[imax])
    top = indsort[:ncorr]
    return frames[top].mean(axis=0), top


def wbi_registration_3d(moving, config, reference_image, schedules=None, mode="ref"):
    """registration.py:192-352, single channel, reference image given.

    ``moving`` (T, Z, Y, X); ``config`` a dict with the reference's keys;
    ``schedules[i]`` forces frame i's solver schedule ({level: (records,
    updates)}).  Returns ``(mem_out, motion_out, errors, masks)`` where masks
    lists ``mask_ref`` then every frame's ``mask_mov`` (bool, (x, y, z)).
    """
    dat_ref = reference_image.transpose(2, 1, 0)
    mcfg = config["mask"]
    mask_ref = mask_oracle.bwareafilt3(mask_oracle.get_mask(dat_ref, mcfg["thresFactor"]), mcfg["maskRange"])
    pyr = config["pyramid"]
    option = {
        "layer": pyr["layer"], "iter": pyr["iter"], "r": pyr["r"], "movRange": 5.0,
        "tol": pyr["tolerance"], "zRatio": config.get("MetaData", {}).get("zRatio", 27.693),
        "save_ite": 100, "mask_ref": mask_ref,
    }
    factor = lk_oracle.smooth_penalty_factor(dat_ref, mask_ref)
    option["smoothPenalty"] = factor * pyr["smoothPenalty"]
    mem_out, motions, errors, masks = [], [], [], [mask_ref]
    for i in range(moving.shape[0]):
        dat_mov = moving[i].transpose(2, 1, 0)
        mask_mov = mask_oracle.bwareafilt3(mask_oracle.get_mask(dat_mov, mcfg["thresFactor"]),
                                           mcfg["maskRange"])
        masks.append(mask_mov)
        opt = dict(option, mask_mov=mask_mov)
        sched = None if schedules is None else schedules[i]
        motion, _, _, _ = lk_oracle.get_motion(dat_mov, dat_ref, opt, mode=mode, schedule=sched)
        corrected = lk_oracle.correct_motion(dat_mov, motion)
        errors.append({"initial_error": mask_oracle.mean_sq_diff(dat_mov, dat_ref),
                       "eventual_error": mask_oracle.mean_sq_diff(corrected, dat_ref)})
        mem_out.append(np.asarray(corrected).transpose(2, 1, 0))
        motions.append(np.asarray(motion).transpose(2, 1, 0, 3))
    return np.asarray(mem_out), np.asarray(motions), errors, masks
