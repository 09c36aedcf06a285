"""Drop-in for ``src/wholistic_registration/utils/reference.py`` on MI355X.

* ``pick_initial_reference(frames, max_corr_frames=20, downsample=1,
  spatial_ds=4)`` (reference.py:19-99): the frame whose ``max_corr_frames``
  best correlations (over the middle three z slices sampled every
  ``spatial_ds`` voxels) are highest, and the mean of the frames most
  correlated with it.
* ``compute_reference_from_block(mem_block, config, ca_block=None)``
  (reference.py:102-124).

The sampled frame matrix, its per-frame means and its demeaned Gram matrix
(the O(T^2 D) part) and the mean of the chosen frames run in libwrhip.so;
the T x T normalisation, sort and arg-max are host numpy, the same calls as
the reference.  numpy inputs give numpy outputs; device tensors give device
tensors.
"""

from __future__ import annotations

import numpy as np

from . import device as _dev
from ._util import torch_module as _torch, is_tensor as _is_tensor, hip_device as _device


def _to_device(a, dev):
    """Host array or tensor -> contiguous uint16 / float32 device tensor."""
    torch = _torch()
    if _is_tensor(a):
        t = a if a.dtype in (torch.uint16, torch.float32) else a.to(torch.float32)
        return t.to(dev).contiguous()
    arr = np.asarray(a)
    if arr.dtype not in (np.uint16, np.float32):
        arr = arr.astype(np.float32)
    return torch.from_numpy(np.ascontiguousarray(arr)).to(dev)


def mean_of_frames(frames, indices):
    """``frames[indices].mean(axis=0)`` summed in the order of ``indices`` on
    the device: float64 for uint16 frames, float32 for float32 frames (numpy's
    result types).  Returns a device tensor."""
    dev = _device()
    vols = [_to_device(frames[int(i)], dev) for i in np.asarray(indices).ravel()]
    return _dev.frames_mean(vols)


def _pick(frames, max_corr_frames, spatial_ds):
    """(reference volume as a device tensor, top indices)."""
    ndim = frames.ndim
    if ndim == 4:
        T, Z = frames.shape[0], frames.shape[1]
        zs = slice(Z // 2 - 1, Z // 2 + 2) if Z >= 3 else slice(0, Z)
        used = frames[:, zs]
    elif ndim == 3:
        T = frames.shape[0]
        used = frames[:, None]
    else:
        raise ValueError("frames must be [T,Y,X] or [T,Z,Y,X]")
    dev = _device()
    slab = _to_device(used, dev)
    cc = _dev.frame_gram(slab, spatial_ds).astype(np.float32)
    del slab
    # reference.py:80-99, the same numpy calls on the T x T matrix
    diag = np.sqrt(np.diag(cc)) + 1e-12
    cc = cc / (diag[:, None] * diag[None, :])
    ncorr = min(max_corr_frames, T - 1)
    if ncorr < 1:
        return mean_of_frames(frames, np.arange(T)), np.arange(T)
    cc_sort = -np.sort(-cc, axis=1)
    best_cc = cc_sort[:, 1:ncorr + 1].mean(axis=1)
    imax = int(np.argmax(best_cc))
    indsort = np.argsort(-cc[imax])
    top_indices = indsort[:ncorr]
    return mean_of_frames(frames, top_indices), top_indices


def pick_initial_reference(frames, max_corr_frames=20, downsample=1, spatial_ds=4):
    """Reference volume of a block of frames (reference.py:19-99).
    ``frames`` is ``[T, Z, Y, X]`` or ``[T, Y, X]``; returns
    ``(refImg, top_indices)``."""
    ref, top = _pick(frames, max_corr_frames, spatial_ds)
    if _is_tensor(frames):
        return ref, top
    return ref.cpu().numpy(), top


def compute_reference_from_block(mem_block, config, ca_block=None):
    """Reference image of a block of frames (reference.py:102-124):
    ``pick_initial_reference`` over ``min(len // 2, 50)`` correlations, plus
    the transformed calcium mean of the same frames for dual-channel configs
    with ``k != 0``.  Returns host arrays ``(result, indsort)``."""
    from .registration import transform
    k = config["channels"]["k"]
    function = config["channels"]["function"]
    dual = config["channels"]["dual_channel"]
    frames = min(len(mem_block) // 2, 50)
    mem_ref, indsort = _pick(mem_block, frames, 4)
    if dual and k != 0:
        ca_ref = mean_of_frames(ca_block, indsort)
        ca_t = transform(ca_ref.to(_torch().float32), k, function)
        result = _dev.add_volumes(mem_ref.to(_torch().float32), ca_t)
    else:
        result = mem_ref
    return result.cpu().numpy(), np.asarray(indsort)
