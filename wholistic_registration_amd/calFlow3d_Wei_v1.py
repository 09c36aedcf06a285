"""Drop-in replacement of the reference solver module on MI355X.

Same names, arguments and return values as
``src/wholistic_registration/utils/calFlow3d_Wei_v1.py``:

* ``getMotion(dat_mov, dat_ref, option, verbose=False)`` (:295-673)
* ``correctMotion(data_raw, motion_field)`` (:676-706)

so ``wbi_registration_3d`` (registration.py:320-326) can use this module by
assignment, the seam the reference's own test patches
(tests/regressions/test_registration_option_assembly.py:31-36).  All compute
runs in libwrhip.so on the current HIP device; this module only converts
layouts and copies host arrays to and from the device.

Inputs follow the reference: ``(x, y, z)`` arrays (numpy, or torch device
tensors), typically the ``transpose(2, 1, 0)`` view of a C-order
``(Z, Y, X)`` volume, for which the conversion to the library's ``(Z, Y, X)``
layout is free.  float32 and uint16 volumes go to the device as they are;
other dtypes are converted to float32.

Differences from the reference, by design:

* ``error_log`` carries the error lists of every level; the periodic host
  snapshots (``motion_current``, ``data_mov_corrected``) and the level images
  (``data_ref``, ``data_mov``) are recorded only when
  ``option["save_snapshots"]`` is true, because copying full volumes to the
  host each level defeats a device-resident solver.  The reference also keeps
  device level images alive in the log (its audit item B-102).
* ``coords_new`` is returned as a host array for host inputs.
* Error values are computed from float32 residuals accumulated in float64.
"""

from __future__ import annotations

import numpy as np

from . import device as _dev
from ._util import torch_module as _torch, is_tensor as _is_tensor, hip_device as _device

REQUIRED_KEYS = ("layer", "iter", "r", "zRatio", "movRange", "smoothPenalty", "tol",
                 "save_ite", "mask_ref", "mask_mov")


def _volume_to_device(a, dev, name):
    """(x, y, z) host array or tensor -> contiguous (Z, Y, X) device tensor."""
    torch = _torch()
    if _is_tensor(a):
        if a.dim() != 3:
            raise ValueError(f"{name} must be 3-D (x, y, z), got {tuple(a.shape)}")
        t = a.permute(2, 1, 0)
        if t.dtype not in (torch.float32, torch.uint16):
            t = t.to(torch.float32)
        return t.to(dev).contiguous()
    arr = np.asarray(a)
    if arr.ndim != 3:
        raise ValueError(f"{name} must be 3-D (x, y, z), got {arr.shape}")
    zyx = arr.transpose(2, 1, 0)
    if zyx.dtype not in (np.float32, np.uint16):
        zyx = zyx.astype(np.float32)
    zyx = np.ascontiguousarray(zyx)
    return torch.from_numpy(zyx).to(dev)


def _mask_to_device(m, shape_xyz, dev, name):
    torch = _torch()
    if _is_tensor(m):
        if tuple(m.shape) != tuple(shape_xyz):
            raise ValueError(f"option[{name!r}] must have the volume's shape {tuple(shape_xyz)}, "
                             f"got {tuple(m.shape)}")
        return m.permute(2, 1, 0).to(dev, torch.float32).contiguous()
    arr = np.asarray(m, dtype=np.float32)
    if arr.shape != tuple(shape_xyz):
        raise ValueError(f"option[{name!r}] must have the volume's shape {tuple(shape_xyz)}, "
                         f"got {arr.shape}")
    if not np.any(arr > 0) and not np.any(arr < 0):
        return None                 # an all-zero mask excludes nothing
    return torch.from_numpy(np.ascontiguousarray(arr.transpose(2, 1, 0))).to(dev)


def _field_to_device(f, shape_xyz, dev):
    torch = _torch()
    if f is None:
        return None
    if _is_tensor(f):
        if tuple(f.shape) != tuple(shape_xyz) + (3,):
            raise ValueError(f"option['motion'] must have shape {tuple(shape_xyz) + (3,)}, "
                             f"got {tuple(f.shape)}")
        return f.permute(2, 1, 0, 3).to(dev, torch.float32).contiguous()
    arr = np.asarray(f, dtype=np.float32)
    if arr.shape != tuple(shape_xyz) + (3,):
        raise ValueError(f"option['motion'] must have shape {tuple(shape_xyz) + (3,)}, got {arr.shape}")
    if not np.any(arr):
        return None                 # resampling zeros gives zeros
    return torch.from_numpy(np.ascontiguousarray(arr.transpose(2, 1, 0, 3))).to(dev)


def getMotion(dat_mov, dat_ref, option, verbose=False):
    """Coarse-to-fine patch-wise Lucas-Kanade registration of ``dat_mov`` onto
    ``dat_ref`` (calFlow3d_Wei_v1.py:295-673).

    Returns ``(motion (x, y, z, 3) float32, currentError, coords_new,
    error_log)``.  Like the reference, ``option["mask_ref"]`` and
    ``option["mask_mov"]`` are replaced by float32 arrays.
    """
    for key in REQUIRED_KEYS:
        option[key]                 # KeyError, as the reference raises
    for key in ("mask_ref", "mask_mov"):
        if not _is_tensor(option[key]):
            option[key] = np.asarray(option[key], dtype=np.float32)
        elif option[key].dtype != _torch().float32:
            option[key] = option[key].to(_torch().float32)
    dev = _device()
    torch_in = _is_tensor(dat_mov)
    shape_xyz = tuple(dat_mov.shape)
    if tuple(dat_ref.shape) != shape_xyz:
        raise ValueError("dat_mov and dat_ref must have the same shape")
    mov = _volume_to_device(dat_mov, dev, "dat_mov")
    ref = _volume_to_device(dat_ref, dev, "dat_ref")
    if mov.dtype != ref.dtype:
        mov, ref = mov.float(), ref.float()
    mref = _mask_to_device(option["mask_ref"], shape_xyz, dev, "mask_ref")
    mmov = _mask_to_device(option["mask_mov"], shape_xyz, dev, "mask_mov")
    init = _field_to_device(option.get("motion"), shape_xyz, dev)
    res = _dev.get_motion(mov, ref, layer=option["layer"], iter=option["iter"], r=option["r"],
                          zRatio=option["zRatio"], movRange=option["movRange"],
                          smoothPenalty=option["smoothPenalty"], tol=option["tol"],
                          mask_ref=mref, mask_mov=mmov, motion_init=init, want_coords=True,
                          schedule=option.get("forced_schedule"),
                          precision=option.get("precision", "fp32"))
    log = res.error_log()
    for L in range(res.layer, -1, -1):
        if verbose:
            print(f"starting layer {L} out of {res.layer}")
            e = log[f"layer_{L}"]
            for k in range(len(e["currentError"])):
                print(f"Downsample layer: {L}\tIter: {k}\tError: {e['currentError'][k]:.3f}, "
                      f"Diff Error: {e['diffError'][k]:.3f}, "
                      f"Penalty Error: {e['penaltyError'][k]:.3f}")
        if res.nans[L]:
            print(f"number of nans: {res.nans[L]}")
    if option.get("save_snapshots"):
        _attach_level_images(log, mov, ref, res.layer)
    if torch_in:
        motion = res.motion.permute(2, 1, 0, 3)
        coords = res.coords.permute(2, 1, 0, 3)
    else:
        motion = res.motion.cpu().numpy().transpose(2, 1, 0, 3)
        coords = res.coords.cpu().numpy().transpose(2, 1, 0, 3)
    return motion, res.final_error, coords, log


def _attach_level_images(log, mov, ref, layer):
    for L in range(layer, -1, -1):
        log[f"layer_{L}"]["data_mov"] = _dev.pyramid_level(mov, L).cpu().numpy().transpose(2, 1, 0)
        log[f"layer_{L}"]["data_ref"] = _dev.pyramid_level(ref, L).cpu().numpy().transpose(2, 1, 0)


def correctMotion(data_raw, motion_field):
    """Warp ``data_raw`` by ``motion_field`` (calFlow3d_Wei_v1.py:676-706):
    ``out(p) = trilinear(data_raw, p + motion(p))``, float32."""
    torch = _torch()
    dev = _device()
    torch_in = _is_tensor(data_raw)
    shape_xyz = tuple(data_raw.shape)
    vol = _volume_to_device(data_raw, dev, "data_raw")
    if _is_tensor(motion_field):
        mot = motion_field.permute(2, 1, 0, 3).to(dev, torch.float32).contiguous()
    else:
        m = np.asarray(motion_field)
        if m.shape != shape_xyz + (3,):
            raise ValueError(f"motion_field must have shape {shape_xyz + (3,)}, got {m.shape}")
        mot = torch.from_numpy(np.ascontiguousarray(m.transpose(2, 1, 0, 3), dtype=np.float32)).to(dev)
    out = _dev.correct_motion(vol, mot)
    if torch_in:
        return out.permute(2, 1, 0)
    return out.cpu().numpy().transpose(2, 1, 0)
