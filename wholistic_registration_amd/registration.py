"""Drop-in for the 3-D registration harness ``wbi_registration_3d``
(``src/wholistic_registration/utils/registration.py:192-352``) on MI355X.

Same signature, config keys, ``option`` side effects, printed lines and return
tuple as the reference.  Every frame stays on the device from its upload to
its outputs: the automatic masks (``mask.getMask`` + ``mask.bwareafilt3_wei``),
the smoothness normaliser, ``getMotion``, the ``correctMotion`` warps and the
initial / eventual errors all run in libwrhip.so.  The solver functions are
looked up as attributes of ``calFlow3d_Wei_v1`` (and the mask functions of
``mask``) at call time, the seam the reference's own test patches
(tests/regressions/test_registration_option_assembly.py:31-36).

Differences from the reference, by design:

* The config is read with ``tomli`` (the reference uses ``toml``, which is not
  a dependency here); ``config_file`` may also be a dict.
* Frames of dtypes other than uint16 / float32 / float64 are converted to
  float32 on upload.
* The calcium channel (dual-channel configs) is transformed, normalised and
  added in float32; numpy keeps float64 where a float64 operand enters (a
  float ``k`` with ``"raw"``, the float64 mean of uint16 frames).
* ``option["motion"]`` is None instead of a zero field when no initial
  motion is given (the solver treats both alike).
"""

from __future__ import annotations

import os

import numpy as np

from . import calFlow3d_Wei_v1, mask, option, reference
from . import device as _dev
from . import preprocess as prep
from ._util import torch_module as _torch, is_tensor as _is_tensor, hip_device as _device

# The module default of zRatio, snapshot at import (registration.py:19-23):
# the fallback for configs without MetaData.zRatio.
_DEFAULT_ZRATIO = option["zRatio"]


def load_config(config_file):
    """The parsed TOML config (``toml.load`` in the reference)."""
    if isinstance(config_file, dict):
        return config_file
    import tomli
    if hasattr(config_file, "read"):
        data = config_file.read()
        return tomli.loads(data.decode() if isinstance(data, bytes) else data)
    with open(os.fspath(config_file), "rb") as f:
        return tomli.load(f)


def _upload_zyx(a, dev):
    """(Z, Y, X) host array or tensor -> contiguous device tensor (uint16,
    float32 or float64 kept; other dtypes as float32)."""
    torch = _torch()
    if _is_tensor(a):
        t = a
        if t.dtype not in (torch.uint16, torch.float32, torch.float64):
            t = t.to(torch.float32)
        return t.to(dev).contiguous()
    arr = np.asarray(a)
    if arr.dtype not in (np.uint16, np.float32, np.float64):
        arr = arr.astype(np.float32)
    return torch.from_numpy(np.ascontiguousarray(arr)).to(dev)


def _xyz_to_zyx_device(a, dev):
    """An (x, y, z) array (host or device) as a contiguous (Z, Y, X) device tensor."""
    if _is_tensor(a):
        return _upload_zyx(a.permute(2, 1, 0), dev)
    return _upload_zyx(np.asarray(a).transpose(2, 1, 0), dev)


def _host(a):
    if _is_tensor(a):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def transform(image, k=1, method="raw"):
    """Calcium-channel transform (registration.py:26-36): ``k * f(image)`` with
    f = identity, sqrt, log2(1 + .), log10(1 + .), in numpy's arithmetic on
    host arrays; on device tensors it runs in libwrhip.so (float32)."""
    if method not in ("raw", "sqrt", "log2", "log10"):
        raise ValueError(f"Unknown method to process the image:{method}")
    if _is_tensor(image):
        return _dev.channel_transform(image.contiguous(), k, method)
    if method == "raw":
        return k * image
    if method == "sqrt":
        return k * np.sqrt(image)
    if method == "log2":
        return k * np.log2(1 + image)
    return k * np.log10(1 + image)


def _mse(a_xyz, b_xyz, dev):
    """np.mean((a - b) ** 2) (registration.py:328-329) on the device."""
    return _dev.mean_sq_diff(_xyz_to_zyx_device(a_xyz, dev), _xyz_to_zyx_device(b_xyz, dev))


def wbi_registration_3d(
    moving_membrane_image,
    config_file,
    reference_image=None,
    motion_init=None,
    verbose=True,
    frame=None,
    direction="forward",
    moving_Ca_image=None,
):
    """Register every frame of ``moving_membrane_image`` ((T, Z, Y, X) or
    (Z, Y, X)) onto one reference volume (registration.py:192-352).

    Returns ``(mem_out (T, Z, Y, X) float32, ca_out, dat_ref (x, y, z),
    errors [{"initial_error", "eventual_error"}], motion_out (T, Z, Y, X, 3))``
    as host arrays.
    """
    config = load_config(config_file)
    if moving_membrane_image.ndim == 3:
        moving_membrane_image = moving_membrane_image[None]
        if moving_Ca_image is not None:
            moving_Ca_image = moving_Ca_image[None]
    if moving_membrane_image.ndim != 4:
        raise ValueError("moving_membrane_image must be 3D or 4D array")
    nimg = moving_membrane_image.shape[0]
    frames = range(nimg)
    dev = _device()
    torch = _torch()

    # reference volume (registration.py:222-250)
    refer = config["reference"]
    channels = config["channels"]
    if refer["pick_reference_auto"]:
        membrane_ref, indsort = reference.pick_initial_reference(
            moving_membrane_image, max_corr_frames=refer["chunk_size"])
    else:
        membrane_ref = reference_image
    if channels["dual_channel"] and moving_Ca_image is None:
        raise ValueError("[ERROR] Calcium image stack shouldn't be empty")
    ref_norm = None
    if channels["dual_channel"] and refer["pick_reference_auto"] and channels["k"] != 0:
        # the calcium mean of the chosen frames, transformed, and its
        # statistics for normalising every moving calcium frame (float32)
        ca_ref = reference.mean_of_frames(moving_Ca_image, indsort).to(torch.float32)
        ca_ref_t = transform(ca_ref, channels["k"], channels["function"])
        ref_norm = _dev.mean_std(ca_ref_t)
        ref_zyx = _dev.add_volumes(_upload_zyx(membrane_ref, dev).to(torch.float32), ca_ref_t)
    else:
        ref_zyx = _upload_zyx(membrane_ref, dev)
    dat_ref = ref_zyx.permute(2, 1, 0)

    # automatic reference mask (registration.py:252-260)
    mask_cfg = config["mask"]
    mask_range = mask_cfg["maskRange"]
    thres_factor = mask_cfg["thresFactor"]
    option["mask_ref"] = mask.getMask(dat_ref, thres_factor)
    option["mask_ref"] = mask.bwareafilt3_wei(option["mask_ref"], mask_range)

    mem_channel, ca_channel, errors, motions = [], [], [], []

    # initial motion (registration.py:268-279): the caller's (Z, Y, X, 3)
    # field back in the solver's (x, y, z, 3) layout
    if motion_init is None:
        option["motion"] = None
    elif _is_tensor(motion_init):
        option["motion"] = motion_init.permute(2, 1, 0, 3)
    else:
        option["motion"] = np.asarray(motion_init).transpose(2, 1, 0, 3)

    # pyramid parameters (registration.py:280-296)
    pyramid = config["pyramid"]
    option["r"] = pyramid["r"]
    option["layer"] = pyramid["layer"]
    option["iter"] = pyramid["iter"]
    option["movRange"] = 5.0
    option["tol"] = pyramid["tolerance"]
    option["zRatio"] = config.get("MetaData", {}).get("zRatio", _DEFAULT_ZRATIO)
    smooth_raw = pyramid["smoothPenalty"]
    pnlt_factor = prep.getSmPnltNormFctr(dat_ref, option)
    option["smoothPenalty"] = pnlt_factor * smooth_raw

    for i in frames:
        mem_zyx = _upload_zyx(moving_membrane_image[i], dev)
        dat_mem = mem_zyx.permute(2, 1, 0)
        if channels["dual_channel"]:
            ca_zyx = _upload_zyx(moving_Ca_image[i], dev)
            ca_t = transform(ca_zyx, channels["k"], channels["function"])
            if refer["pick_reference_auto"]:
                if ref_norm is None:
                    raise ValueError("normalising the calcium channel needs k != 0 "
                                     "(registration.py:229-245 leaves its statistics undefined)")
                ca_t = _dev.normalize_std(ref_norm[0], ref_norm[1], ca_t)
            mov_zyx = _dev.add_volumes(mem_zyx, ca_t)
        else:
            mov_zyx = mem_zyx
        dat_mov = mov_zyx.permute(2, 1, 0)

        # moving mask (registration.py:311-312)
        option["mask_mov"] = mask.getMask(dat_mov, thres_factor)
        option["mask_mov"] = mask.bwareafilt3_wei(option["mask_mov"], mask_range)

        motion_current, _, new_coords, error_logs = calFlow3d_Wei_v1.getMotion(dat_mov, dat_ref, option)
        if channels["dual_channel"]:
            corrected_ca = calFlow3d_Wei_v1.correctMotion(ca_zyx.permute(2, 1, 0), motion_current)
        corrected_mem = calFlow3d_Wei_v1.correctMotion(dat_mem, motion_current)
        # without a calcium channel the moving volume is the membrane volume:
        # its warp is the same array (the warp is deterministic)
        corrected_mov = (corrected_mem if mov_zyx is mem_zyx
                         else calFlow3d_Wei_v1.correctMotion(dat_mov, motion_current))
        initial_error = _mse(dat_mov, dat_ref, dev)
        eventual_error = _mse(corrected_mov, dat_ref, dev)
        if verbose:
            label = i + 1 if frame is None else frame[i]
            print(f"        Frame: {label}\tInitial Error is:{initial_error:.4f}"
                  f"\tEventual Error: {eventual_error:.4f}")
        errors.append({"initial_error": initial_error, "eventual_error": eventual_error})
        mem_channel.append(_host(corrected_mem).transpose(2, 1, 0))
        if channels["dual_channel"]:
            ca_channel.append(_host(corrected_ca).transpose(2, 1, 0))
        motions.append(_host(motion_current).transpose(2, 1, 0, 3))

    mem_out = np.asarray(mem_channel)
    ca_out = np.asarray(ca_channel) if ca_channel else np.array([])
    motion_out = np.asarray(motions)
    return mem_out, ca_out, _host(dat_ref), errors, motion_out
