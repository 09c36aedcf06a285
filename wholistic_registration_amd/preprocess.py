"""Drop-in for the registration-path part of ``src/wholistic_registration/utils/preprocess.py``.

``getSmPnltNormFctr(dat_ref, option)`` (preprocess.py:42-66) is the factor the
harness multiplies the raw smoothness weight by before ``getMotion``
(registration.py:295).  ``robust_mean_std(data, percentile)``
(preprocess.py:69-84) gives the statistics of ``getMask``.  Both run in
libwrhip.so (``wr_smooth_penalty_factor``, ``wr_robust_mean_std``) on the
current HIP device.
"""

from __future__ import annotations

import numpy as np

from . import device as _dev
from ._util import torch_module as _torch


def _to_zyx(a, dev, name, dtype=None):
    torch = _torch()
    if isinstance(a, torch.Tensor):
        if a.dim() != 3:
            raise ValueError(f"{name} must be 3-D (x, y, z), got {tuple(a.shape)}")
        t = a.permute(2, 1, 0)
        if dtype is not None:
            t = t.to(dtype)
        return t.to(dev).contiguous()
    arr = np.asarray(a)
    if arr.ndim != 3:
        raise ValueError(f"{name} must be 3-D (x, y, z), got {arr.shape}")
    zyx = np.ascontiguousarray(arr.transpose(2, 1, 0))
    t = torch.from_numpy(zyx)
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev)


def getSmPnltNormFctr(dat_ref, option):
    """Smoothness-penalty factor of the reference volume ``dat_ref``
    ``(x, y, z)``: half the sum of the mean squared central differences along
    x and y over the voxels where ``option["mask_ref"]`` is False.

    The arithmetic is the reference's for the input dtype: uint16 differences
    wrap modulo 2^16, float32 stays float32 up to the sums (accumulated in
    float64 here), float64 is float64.  Signed integer volumes are exact in
    float64 and are converted; other dtypes raise TypeError.
    """
    torch = _torch()
    mask = option["mask_ref"]
    if not torch.cuda.is_available():
        raise RuntimeError("wholistic_registration_amd needs a HIP device (none is visible)")
    dev = torch.device("cuda", torch.cuda.current_device())
    kind = dat_ref.dtype
    if isinstance(dat_ref, torch.Tensor):
        ok = kind in (torch.uint16, torch.float32, torch.float64)
        signed = kind in (torch.int8, torch.int16, torch.int32, torch.int64)
    else:
        kind = np.asarray(dat_ref).dtype
        ok = kind in (np.uint16, np.float32, np.float64)
        signed = np.issubdtype(kind, np.signedinteger)
    if not ok and not signed:
        raise TypeError(f"getSmPnltNormFctr: unsupported volume dtype {kind}")
    ref = _to_zyx(dat_ref, dev, "dat_ref", None if ok else torch.float64)
    m = None
    if mask is not None:
        if isinstance(mask, torch.Tensor):
            m = _to_zyx(mask, dev, "mask_ref", None if mask.dtype == torch.float32 else torch.float32)
        else:
            m = _to_zyx(np.asarray(mask, dtype=np.float32), dev, "mask_ref")
        if tuple(m.shape) != tuple(ref.shape):
            raise ValueError("option['mask_ref'] must have the volume's shape")
    return _dev.smooth_penalty_factor(ref, m)


def robust_mean_std(data, percentile=95):
    """Mean and std of the values at or below the ``percentile``-th
    percentile of ``data`` (preprocess.py:69-84), in float32 arithmetic:
    float32 data as the reference computes it, uint16 data converted to
    float32 first (as ``getMask`` does before calling it).  numpy inputs are
    copied to the device; the percentile selection and the moments run in
    libwrhip.so.  Returns numpy float32 scalars."""
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("wholistic_registration_amd needs a HIP device (none is visible)")
    dev = torch.device("cuda", torch.cuda.current_device())
    if isinstance(data, torch.Tensor):
        t = data.reshape(-1)
        if t.dtype not in (torch.float32, torch.uint16):
            t = t.to(torch.float32)
        t = t.to(dev).contiguous()
    else:
        a = np.asarray(data).reshape(-1)
        if a.dtype not in (np.float32, np.uint16):
            a = a.astype(np.float32)
        t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    _, mean, std = _dev.robust_mean_std(t, percentile)
    return mean, std
