"""Device-resident entry points over torch tensors (the hot path without
host copies).

Volumes are contiguous ``(Z, Y, X)`` CUDA/HIP tensors (float32 or uint16);
motion fields are contiguous ``(Z, Y, X, 3)`` float32 tensors with components
ordered ``(x, y, z)``.  Each call enqueues on torch's current stream.  PyTorch
only provides the device memory and the stream; the work is done by
libwrhip.so.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def dtype_code(t):
    torch = _torch()
    if t.dtype == torch.float32:
        return _lib.WR_F32
    if t.dtype == torch.uint16:
        return _lib.WR_U16
    if t.dtype == torch.float64:
        return _lib.WR_F64
    raise ValueError(f"unsupported volume dtype {t.dtype}")


def _check_vol(t, name, ndim=3, last=None):
    if t is None:
        return
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dim() != ndim or (last is not None and t.shape[-1] != last):
        raise ValueError(f"{name} has shape {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous (Z, Y, X) memory")


@dataclass
class MotionResult:
    motion: object                 # (Z, Y, X, 3) float32 tensor
    coords: object                 # (Z, Y, X, 3) float32 tensor or None
    records: list                  # per level, indexed by level
    updates: list
    nans: list
    diff_error: np.ndarray         # (layer+1, iter)
    penalty_error: np.ndarray
    current_error: np.ndarray
    max_diff_motion: np.ndarray
    final_error: float
    layer: int = 0
    extras: dict = field(default_factory=dict)

    def schedule(self):
        """``{level: (records, updates)}``, the form the oracle accepts."""
        return {L: (self.records[L], self.updates[L]) for L in range(self.layer + 1)}

    def error_log(self):
        """The lists of the reference's ``error_log`` (calFlow3d:329-335)."""
        out = {}
        for L in range(self.layer, -1, -1):
            n, u = self.records[L], self.updates[L]
            out[f"layer_{L}"] = {
                "diffError": [float(v) for v in self.diff_error[L, :n]],
                "penaltyError": [float(v) for v in self.penalty_error[L, :n]],
                "currentError": [float(v) for v in self.current_error[L, :n]],
                "max_diff_motion": [np.float32(v) for v in self.max_diff_motion[L, :u]],
                "motion_current": [],
                "data_mov_corrected": [],
            }
        return out


def get_motion(mov, ref, *, layer, iter, r, zRatio, movRange, smoothPenalty, tol,
               mask_ref=None, mask_mov=None, motion_init=None, want_coords=False,
               schedule=None, tile_patches_x=0, motion_out=None, device=None, timing=False,
               stop_kernels=False):
    """Pyramid LK registration of ``mov`` onto ``ref`` on the GPU.

    Replaces ``getMotion`` (calFlow3d_Wei_v1.py:295-673); the stop rules run on
    the device.  ``schedule`` ({level: (records, updates)}) forces another
    run's iteration schedule.  ``stop_kernels`` runs the stop rules in their
    own single-workgroup kernels on every level (by default that is done only
    where the launches that follow cannot afford to evaluate them).
    """
    torch = _torch()
    _check_vol(mov, "mov")
    _check_vol(ref, "ref")
    if mov.shape != ref.shape or mov.dtype != ref.dtype:
        raise ValueError("mov and ref must have the same shape and dtype")
    for m, name in ((mask_ref, "mask_ref"), (mask_mov, "mask_mov")):
        _check_vol(m, name)
        if m is not None and (m.shape != mov.shape or m.dtype != torch.float32):
            raise ValueError(f"{name} must be float32 with the volume's shape")
    _check_vol(motion_init, "motion_init", 4, 3)
    nz, ny, nx = mov.shape
    dev = mov.device if device is None else device
    if motion_out is None:
        motion_out = torch.empty((nz, ny, nx, 3), dtype=torch.float32, device=dev)
    coords = torch.empty_like(motion_out) if want_coords else None

    prm = _lib.WrParams()
    prm.layer, prm.iter, prm.r = int(layer), int(iter), int(r)
    prm.zRatio, prm.movRange = float(zRatio), float(movRange)
    prm.smoothPenalty, prm.tol = float(smoothPenalty), float(tol)
    prm.tile_patches_x = int(tile_patches_x)
    prm.stop_kernels = 1 if stop_kernels else 0
    prm.timing = 1 if timing else 0
    keep = []
    if schedule is not None:
        rec = (ctypes.c_int * (layer + 1))(*[schedule[L][0] for L in range(layer + 1)])
        upd = (ctypes.c_int * (layer + 1))(*[schedule[L][1] for L in range(layer + 1)])
        prm.forced_records = ctypes.cast(rec, ctypes.POINTER(ctypes.c_int))
        prm.forced_updates = ctypes.cast(upd, ctypes.POINTER(ctypes.c_int))
        keep += [rec, upd]
    nlev = layer + 1
    bufs = {k: np.zeros((nlev, max(iter, 1)), dtype=np.float64)
            for k in ("diff", "pen", "cur", "md")}
    log = _lib.WrLog()
    dp = ctypes.POINTER(ctypes.c_double)
    log.diff_error = bufs["diff"].ctypes.data_as(dp)
    log.penalty_error = bufs["pen"].ctypes.data_as(dp)
    log.current_error = bufs["cur"].ctypes.data_as(dp)
    log.max_diff_motion = bufs["md"].ctypes.data_as(dp)

    lib = _lib.load()
    ctx = _lib.context(dev.index if dev.index is not None else 0)
    rc = lib.wr_get_motion(ctx, _ptr(mov), _ptr(ref), dtype_code(mov), nx, ny, nz,
                           _ptr(mask_ref), _ptr(mask_mov), _ptr(motion_init), ctypes.byref(prm),
                           _ptr(motion_out), _ptr(coords), ctypes.byref(log), _stream())
    _lib.check(rc, "wr_get_motion")
    return MotionResult(
        motion=motion_out, coords=coords,
        records=[log.records[L] for L in range(nlev)],
        updates=[log.updates[L] for L in range(nlev)],
        nans=[log.nans[L] for L in range(nlev)],
        diff_error=bufs["diff"], penalty_error=bufs["pen"], current_error=bufs["cur"],
        max_diff_motion=bufs["md"], final_error=float(log.final_error), layer=layer)


def stats(device=0):
    """Launch statistics of the last get_motion on ``device`` (HIP-event
    durations of the LK passes when it ran with ``timing=True``)."""
    st = _lib.WrStats()
    _lib.check(_lib.load().wr_get_stats(_lib.context(device), ctypes.byref(st)), "wr_get_stats")
    return st


def correct_motion(vol, motion, out=None):
    """``correctMotion`` (calFlow3d_Wei_v1.py:676-706) on device tensors."""
    torch = _torch()
    _check_vol(vol, "volume")
    _check_vol(motion, "motion", 4, 3)
    if tuple(motion.shape[:3]) != tuple(vol.shape) or motion.dtype != torch.float32:
        raise ValueError("motion must be float32 (Z, Y, X, 3) matching the volume")
    if out is None:
        out = torch.empty(vol.shape, dtype=torch.float32, device=vol.device)
    nz, ny, nx = vol.shape
    lib = _lib.load()
    ctx = _lib.context(vol.device.index or 0)
    rc = lib.wr_correct_motion(ctx, _ptr(vol), dtype_code(vol), nx, ny, nz, _ptr(motion),
                               _ptr(out), _stream())
    _lib.check(rc, "wr_correct_motion")
    return out


def pyramid_level(vol, level):
    """Level image ``imresize(vol, (x/2^L, y/2^L, z))`` as float32 (Z, Y_L, X_L)."""
    torch = _torch()
    _check_vol(vol, "volume")
    nz, ny, nx = vol.shape
    lx, ly = int(nx / 2 ** level), int(ny / 2 ** level)
    out = torch.empty((nz, max(ly, 1), max(lx, 1)), dtype=torch.float32, device=vol.device)
    lib = _lib.load()
    ctx = _lib.context(vol.device.index or 0)
    rc = lib.wr_pyramid_level(ctx, _ptr(vol), dtype_code(vol), nx, ny, nz, int(level), _ptr(out),
                              _stream())
    _lib.check(rc, "wr_pyramid_level")
    return out


def imresize_xy(vol, nx_out, ny_out, method="bicubic"):
    """``imresize(vol, output_shape=(nx_out, ny_out, nz), method)`` as float64."""
    torch = _torch()
    _check_vol(vol, "volume")
    nz, ny, nx = vol.shape
    if method not in ("bicubic", "bilinear"):
        raise ValueError("unidentified kernel method supplied")
    out = torch.empty((nz, ny_out, nx_out), dtype=torch.float64, device=vol.device)
    lib = _lib.load()
    ctx = _lib.context(vol.device.index or 0)
    rc = lib.wr_imresize(ctx, _ptr(vol), dtype_code(vol), nx, ny, nz, int(nx_out), int(ny_out),
                         0 if method == "bicubic" else 1, _ptr(out), _stream())
    _lib.check(rc, "wr_imresize")
    return out


def smooth_penalty_factor(ref, mask_ref=None):
    """``getSmPnltNormFctr`` (preprocess.py:42-66) of a (Z, Y, X) device
    volume: half the sum of the mean squared central differences along x and
    y over the voxels whose ``mask_ref`` is zero, in the arithmetic of the
    volume's dtype (uint16 differences wrap).  ``mask_ref``: (Z, Y, X) float32
    or bool device tensor (nonzero = excluded) or None."""
    torch = _torch()
    _check_vol(ref, "ref")
    if mask_ref is not None:
        _check_vol(mask_ref, "mask_ref")
        if tuple(mask_ref.shape) != tuple(ref.shape):
            raise ValueError("mask_ref must match the volume's shape")
        if mask_ref.dtype != torch.float32:
            mask_ref = mask_ref.to(torch.float32)
    nz, ny, nx = ref.shape
    out = ctypes.c_double()
    lib = _lib.load()
    ctx = _lib.context(ref.device.index or 0)
    rc = lib.wr_smooth_penalty_factor(ctx, _ptr(ref), dtype_code(ref), nx, ny, nz, _ptr(mask_ref),
                                      ctypes.byref(out), _stream())
    _lib.check(rc, "wr_smooth_penalty_factor")
    return float(out.value)
