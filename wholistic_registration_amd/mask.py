"""Drop-in for ``src/wholistic_registration/utils/mask.py`` on MI355X.

* ``getMask(dat_mov, thres_factor)`` (mask.py:59-111)
* ``bwareafilt3_wei(vol, size_range)`` (mask.py:8-56)

Same names and arguments as the reference.  Inputs are ``(x, y, z)`` arrays
(numpy, or torch device tensors, typically the ``permute(2, 1, 0)`` view of a
C-order ``(Z, Y, X)`` volume); all compute runs in libwrhip.so on the current
HIP device.  numpy inputs give numpy bool outputs, like the reference.  Device
inputs give device outputs in the same ``(x, y, z)`` view: ``getMask`` a uint8
0/1 tensor, ``bwareafilt3_wei`` a float32 0/1 tensor, the form ``getMotion``
converts its masks to (calFlow3d_Wei_v1.py:314-315), so the harness hands it
over without a conversion.
"""

from __future__ import annotations

import numpy as np

from . import device as _dev
from ._util import torch_module as _torch, is_tensor as _is_tensor, hip_device as _device


def _volume_zyx(a, name):
    """(x, y, z) array or tensor -> contiguous (Z, Y, X) float32/uint16 device tensor."""
    torch = _torch()
    dev = _device()
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
        zyx = zyx.astype(np.float32)   # getMask converts to float32 first (mask.py:83)
    return torch.from_numpy(np.ascontiguousarray(zyx)).to(dev)


def _mask_zyx(a, name):
    """(x, y, z) array or tensor (nonzero = set) -> (Z, Y, X) uint8 device tensor."""
    torch = _torch()
    dev = _device()
    if _is_tensor(a):
        if a.dim() != 3:
            raise ValueError(f"{name} must be 3-D (x, y, z), got {tuple(a.shape)}")
        t = a.permute(2, 1, 0)
        if t.dtype != torch.uint8:
            t = t != 0
        return t.to(dev, torch.uint8).contiguous()
    arr = np.asarray(a)
    if arr.ndim != 3:
        raise ValueError(f"{name} must be 3-D (x, y, z), got {arr.shape}")
    zyx = np.ascontiguousarray((arr != 0).transpose(2, 1, 0), dtype=np.uint8)
    return torch.from_numpy(zyx).to(dev)


def getMask(dat_mov, thres_factor):
    """Outlier mask of ``dat_mov`` (mask.py:59-111): float32 z-score
    ``|(v - mu) / sigma|`` above ``thres_factor`` with the robust statistics
    of the values at or below the 95th percentile, then a 3x3x1 binary
    opening and a 3x3x1 binary dilation (in-plane, x-y)."""
    vol = _volume_zyx(dat_mov, "dat_mov")
    out = _dev.get_mask(vol, thres_factor)
    if _is_tensor(dat_mov):
        return out.permute(2, 1, 0)
    return out.cpu().numpy().astype(bool).transpose(2, 1, 0)


def bwareafilt3_wei(vol, size_range):
    """Keep the 26-connected components of ``vol != 0`` whose voxel count lies
    in ``[size_range[0], size_range[1]]`` (mask.py:8-56)."""
    m = _mask_zyx(vol, "vol")
    if _is_tensor(vol):
        return _dev.bwareafilt3(m, size_range, out_dtype="float32").permute(2, 1, 0)
    out = _dev.bwareafilt3(m, size_range, out_dtype="uint8")
    return out.cpu().numpy().astype(bool).transpose(2, 1, 0)
