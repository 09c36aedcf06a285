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
from ._util import is_tensor as _is_tensor
from ._util import torch_module as _torch


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


PRECISIONS = {"fp32": 0, "bf16": 1, "fp16": 2}


def _params(layer, iter, r, zRatio, movRange, smoothPenalty, tol, tile_patches_x=0, stop_kernels=False,
            precision="fp32", reuse_template=False, timing=False, schedule=None, resample_start=False,
            maxdiff_gather=False):
    """wr_params of a getMotion call, and the ctypes arrays it points to."""
    prm = _lib.WrParams()
    prm.layer, prm.iter, prm.r = int(layer), int(iter), int(r)
    prm.zRatio, prm.movRange = float(zRatio), float(movRange)
    prm.smoothPenalty, prm.tol = float(smoothPenalty), float(tol)
    keep = []
    if isinstance(tile_patches_x, (list, tuple)):
        if len(tile_patches_x) != layer + 1:
            raise ValueError(f"tile_patches_x needs one entry per level (layer + 1 = {layer + 1})")
        arr = (ctypes.c_int * (layer + 1))(*[int(v) for v in tile_patches_x])
        prm.tile_patches_level = ctypes.cast(arr, ctypes.POINTER(ctypes.c_int))
        keep.append(arr)
    else:
        prm.tile_patches_x = int(tile_patches_x)
    prm.stop_kernels = 1 if stop_kernels else 0
    prm.resample_start = 1 if resample_start else 0
    prm.maxdiff_gather = 1 if maxdiff_gather else 0
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}")
    prm.precision = PRECISIONS[precision]
    prm.reuse_template = 1 if reuse_template else 0
    prm.timing = 1 if timing else 0
    if schedule is not None:
        rec = (ctypes.c_int * (layer + 1))(*[schedule[L][0] for L in range(layer + 1)])
        upd = (ctypes.c_int * (layer + 1))(*[schedule[L][1] for L in range(layer + 1)])
        prm.forced_records = ctypes.cast(rec, ctypes.POINTER(ctypes.c_int))
        prm.forced_updates = ctypes.cast(upd, ctypes.POINTER(ctypes.c_int))
        keep += [rec, upd]
    return prm, keep


def _log_struct(nlev, iter):
    """A wr_log whose arrays point into fresh (nlev, iter) float64 buffers."""
    bufs = {k: np.zeros((nlev, max(iter, 1)), dtype=np.float64) for k in ("diff", "pen", "cur", "md")}
    log = _lib.WrLog()
    dp = ctypes.POINTER(ctypes.c_double)
    log.diff_error = bufs["diff"].ctypes.data_as(dp)
    log.penalty_error = bufs["pen"].ctypes.data_as(dp)
    log.current_error = bufs["cur"].ctypes.data_as(dp)
    log.max_diff_motion = bufs["md"].ctypes.data_as(dp)
    return log, bufs


def _result(motion, coords, log, bufs, layer):
    nlev = layer + 1
    return MotionResult(
        motion=motion, coords=coords,
        records=[log.records[L] for L in range(nlev)],
        updates=[log.updates[L] for L in range(nlev)],
        nans=[log.nans[L] for L in range(nlev)],
        diff_error=bufs["diff"], penalty_error=bufs["pen"], current_error=bufs["cur"],
        max_diff_motion=bufs["md"], final_error=float(log.final_error), layer=layer)


def get_motion(mov, ref, *, layer, iter, r, zRatio, movRange, smoothPenalty, tol,
               mask_ref=None, mask_mov=None, motion_init=None, want_coords=False,
               schedule=None, tile_patches_x=0, motion_out=None, device=None, timing=False,
               stop_kernels=False, precision="fp32", reuse_template=False, context=None, stream=None,
               motion_init_cp=None, resample_start=False, maxdiff_gather=False, warp_src=None,
               warped_out=None, init_event=None):
    """Pyramid LK registration of ``mov`` onto ``ref`` on the GPU.

    Replaces ``getMotion`` (calFlow3d_Wei_v1.py:295-673); the stop rules run on
    the device.  ``schedule`` ({level: (records, updates)}) forces another
    run's iteration schedule.  ``stop_kernels`` runs the stop rules in their
    own single-workgroup kernels on every level (by default that is done only
    where the launches that follow cannot afford to evaluate them).
    ``precision`` stores the pyramid level images the iterations read as
    ``"fp32"`` (the reference's values), ``"bf16"`` or ``"fp16"``; every
    interpolation, sum and solve stays float32 / float64.  ``reuse_template``
    skips rebuilding the template's pyramid when the previous call on the same
    context had the same ``ref`` tensor (unchanged), size, layer and precision.
    ``context`` (a handle from ``_lib``) and ``stream`` (a torch stream)
    default to the device's shared context and torch's current stream.
    ``tile_patches_x``: patches per LK workgroup along x, one value for every
    level or a list of ``layer + 1`` values (level 0 first; 0 = the built-in
    rule), as ``autotune.tune_tiles`` picks them.  ``warp_src`` (a (Z, Y, X)
    float32 or uint16 volume, e.g. ``mov``) is also warped by the returned
    field (``correctMotion``, calFlow3d_Wei_v1.py:676-706) into ``warped_out``
    (float32, allocated when None; ``result.extras["warped"]``) in the kernel
    that writes the field (wr_get_motion_warp).  ``motion_init`` may have
    another x-y size than the volume (full-resolution voxel units; the
    reference's warm start resamples it to the coarsest level).
    ``init_event`` (a ``torch.cuda.Event``): where ``motion_init`` /
    ``motion_init_cp`` is written on another stream; the iterations wait for
    it, the level images are built without waiting (``wr_params.init_event``).
    ``motion_init_cp`` (instead of ``motion_init``): the coarsest level's
    control-point field itself, (3, Z, ncy, ncx) float32 in that level's
    voxel units (``cp_grid``); the level starts from its interpolant.
    ``resample_start`` and ``maxdiff_gather`` select the reference-shaped
    kernels that the automatic path replaces (a materialised dense field per
    level end, upsampled by the resampler; the gather form of the first
    update's |delta motion| maximum), for tests that compare the two.
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
    if motion_init_cp is not None:
        if motion_init is not None:
            raise ValueError("give motion_init or motion_init_cp, not both")
        want = (3, nz) + cp_grid((nz, ny, nx), r, layer)[::-1]
        if tuple(motion_init_cp.shape) != want or motion_init_cp.dtype != torch.float32 or \
                not motion_init_cp.is_contiguous():
            raise ValueError(f"motion_init_cp must be contiguous float32 {want}")
    if motion_init is not None and (motion_init.shape[0] != nz or motion_init.dtype != torch.float32):
        raise ValueError("motion_init must be float32 (Z, Y', X', 3) with the volume's Z")
    dev = mov.device if device is None else device
    if motion_out is None:
        motion_out = torch.empty((nz, ny, nx, 3), dtype=torch.float32, device=dev)
    coords = torch.empty_like(motion_out) if want_coords else None

    prm, keep = _params(layer, iter, r, zRatio, movRange, smoothPenalty, tol, tile_patches_x, stop_kernels,
                        precision, reuse_template, timing, schedule, resample_start, maxdiff_gather)
    if init_event is not None:     # motion_init / motion_init_cp written on another stream
        prm.init_event = ctypes.c_void_p(init_event.cuda_event)
    if motion_init_cp is not None:
        prm.init_cp = 1
        motion_init = motion_init_cp
    elif motion_init is not None and tuple(motion_init.shape[1:3]) != (ny, nx):
        # a warm start of another x-y size (full-resolution units)
        prm.init_ny, prm.init_nx = int(motion_init.shape[1]), int(motion_init.shape[2])
    log, bufs = _log_struct(layer + 1, iter)
    lib = _lib.load()
    ctx = context if context is not None else _lib.context(dev.index if dev.index is not None else 0)
    strm = _stream() if stream is None else ctypes.c_void_p(stream.cuda_stream)
    if warp_src is not None:
        _check_vol(warp_src, "warp_src")
        if tuple(warp_src.shape) != (nz, ny, nx):
            raise ValueError("warp_src must have the volume's shape")
        if want_coords:
            raise ValueError("want_coords and warp_src together are not supported")
        if warped_out is None:
            warped_out = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
        _check_vol(warped_out, "warped_out")
        if tuple(warped_out.shape) != (nz, ny, nx) or warped_out.dtype != torch.float32:
            raise ValueError("warped_out must be float32 with the volume's shape")
        rc = lib.wr_get_motion_warp(ctx, _ptr(mov), _ptr(ref), dtype_code(mov), nx, ny, nz,
                                    _ptr(mask_ref), _ptr(mask_mov), _ptr(motion_init), ctypes.byref(prm),
                                    _ptr(motion_out), _ptr(warp_src), dtype_code(warp_src), _ptr(warped_out),
                                    ctypes.byref(log), strm)
        _lib.check(rc, "wr_get_motion_warp")
        res = _result(motion_out, coords, log, bufs, layer)
        res.extras["warped"] = warped_out
        return res
    rc = lib.wr_get_motion(ctx, _ptr(mov), _ptr(ref), dtype_code(mov), nx, ny, nz,
                           _ptr(mask_ref), _ptr(mask_mov), _ptr(motion_init), ctypes.byref(prm),
                           _ptr(motion_out), _ptr(coords), ctypes.byref(log), strm)
    _lib.check(rc, "wr_get_motion")
    return _result(motion_out, coords, log, bufs, layer)


def get_motion_batch(movs, ref, *, layer, iter, r, zRatio, movRange, smoothPenalty, tol,
                     mask_ref=None, mask_mov=None, motion_init=None, motion_out=None, tile_patches_x=0,
                     timing=False, stop_kernels=False, precision="fp32", reuse_template=False, context=None,
                     stream=None, schedule=None, logs=True, warp_src=None, warped_out=None, motion_init_cp=None,
                     init_event=None):
    """getMotion of every frame of ``movs`` (F, Z, Y, X) against ``ref`` in one
    launch sequence (wr_get_motion_batch): one launch per kernel and
    level-iteration over all frames, each frame with its own stop rules.
    Every frame's field equals ``get_motion(movs[f], ref, ...)`` bit for bit.
    ``mask_mov`` (F, Z, Y, X) float32 and ``motion_init`` (F, Z, Y, X, 3)
    are per frame (or ``motion_init_cp`` (F, 3, Z, ncy, ncx): each frame's
    coarsest-level control-point field, as in ``get_motion``); ``motion_out``
    is (F, Z, Y, X, 3).  Returns one
    MotionResult per frame (its ``motion`` a view of ``motion_out``).  With
    ``warp_src`` (F, Z, Y, X) float32 or uint16 (e.g. ``movs``) is warped by
    each frame's field into ``warped_out`` (F, Z, Y, X) float32 in the kernel
    that writes the field.  With
    ``logs=False`` the call only enqueues the work on ``stream`` and returns
    ``motion_out`` without waiting: repeated with the same tensors after a
    first call, it allocates nothing and can be captured in a CUDA/HIP graph
    (``torch.cuda.graph``).  ``init_event`` (a ``torch.cuda.Event``) marks
    where ``motion_init`` / ``motion_init_cp`` is written on another stream:
    the iterations wait for it, the level images are built without waiting
    (``wr_params.init_event``)."""
    torch = _torch()
    if not _is_tensor(movs) or movs.dim() != 4 or not movs.is_cuda or not movs.is_contiguous():
        raise ValueError("movs must be a contiguous (F, Z, Y, X) device tensor")
    nf, nz, ny, nx = (int(v) for v in movs.shape)
    _check_vol(ref, "ref")
    if tuple(ref.shape) != (nz, ny, nx) or ref.dtype != movs.dtype:
        raise ValueError("ref must be (Z, Y, X) with the frames' shape and dtype")
    _check_vol(mask_ref, "mask_ref")
    if mask_ref is not None and (tuple(mask_ref.shape) != (nz, ny, nx) or mask_ref.dtype != torch.float32):
        raise ValueError("mask_ref must be float32 with the volume's shape")
    if mask_mov is not None and (tuple(mask_mov.shape) != (nf, nz, ny, nx) or mask_mov.dtype != torch.float32
                                 or not mask_mov.is_contiguous()):
        raise ValueError("mask_mov must be contiguous float32 (F, Z, Y, X)")
    if motion_init is not None and (tuple(motion_init.shape) != (nf, nz, ny, nx, 3)
                                    or motion_init.dtype != torch.float32 or not motion_init.is_contiguous()):
        raise ValueError("motion_init must be contiguous float32 (F, Z, Y, X, 3)")
    init_stride = 3 * nz * ny * nx
    if motion_init_cp is not None:
        if motion_init is not None:
            raise ValueError("give motion_init or motion_init_cp, not both")
        ncx, ncy = cp_grid((nz, ny, nx), r, layer)
        if (tuple(motion_init_cp.shape) != (nf, 3, nz, ncy, ncx) or motion_init_cp.dtype != torch.float32
                or not motion_init_cp.is_contiguous()):
            raise ValueError(f"motion_init_cp must be contiguous float32 {(nf, 3, nz, ncy, ncx)}")
        init_stride = 3 * nz * ncy * ncx
    if motion_out is None:
        motion_out = torch.empty((nf, nz, ny, nx, 3), dtype=torch.float32, device=movs.device)
    elif tuple(motion_out.shape) != (nf, nz, ny, nx, 3) or not motion_out.is_contiguous():
        raise ValueError("motion_out must be contiguous float32 (F, Z, Y, X, 3)")
    prm, keep = _params(layer, iter, r, zRatio, movRange, smoothPenalty, tol, tile_patches_x, stop_kernels,
                        precision, reuse_template, timing, schedule)
    if motion_init_cp is not None:
        prm.init_cp = 1
        motion_init = motion_init_cp
    if init_event is not None:
        prm.init_event = ctypes.c_void_p(init_event.cuda_event)
    if warp_src is not None:
        if (not _is_tensor(warp_src) or tuple(warp_src.shape) != (nf, nz, ny, nx) or not warp_src.is_contiguous()
                or warp_src.dtype not in (torch.float32, torch.uint16)):
            raise ValueError("warp_src must be contiguous float32 or uint16 (F, Z, Y, X)")
        if warped_out is None:
            warped_out = torch.empty((nf, nz, ny, nx), dtype=torch.float32, device=movs.device)
        elif (tuple(warped_out.shape) != (nf, nz, ny, nx) or warped_out.dtype != torch.float32
              or not warped_out.is_contiguous()):
            raise ValueError("warped_out must be contiguous float32 (F, Z, Y, X)")
    bufs = []
    if logs:
        logs = (_lib.WrLog * nf)()
        for f in range(nf):
            lg, b = _log_struct(layer + 1, iter)
            logs[f] = lg
            bufs.append(b)
    else:
        logs = None
    lib = _lib.load()
    ctx = context if context is not None else _lib.context(movs.device.index or 0)
    strm = _stream() if stream is None else ctypes.c_void_p(stream.cuda_stream)
    nvox = nz * ny * nx
    rc = lib.wr_get_motion_batch(ctx, _ptr(movs), nvox, _ptr(ref), dtype_code(movs), nx, ny, nz, nf,
                                 _ptr(mask_ref), _ptr(mask_mov), nvox, _ptr(motion_init), init_stride,
                                 ctypes.byref(prm), _ptr(motion_out), _ptr(warp_src),
                                 dtype_code(warp_src) if warp_src is not None else 0, nvox, _ptr(warped_out),
                                 logs, strm)
    _lib.check(rc, "wr_get_motion_batch")
    del keep
    if logs is None:
        return motion_out
    out = [_result(motion_out[f], None, logs[f], bufs[f], layer) for f in range(nf)]
    if warp_src is not None:
        for f, res in enumerate(out):
            res.extras["warped"] = warped_out[f]
    return out


def cp_grid(shape_zyx, r, level=0):
    """(ncx, ncy): control points per row and column of pyramid level
    ``level`` (calFlow3d_Wei_v1.py:434-436: range(r, n_L - 1, 2r + 1))."""
    _, ny, nx = shape_zyx
    lx, ly = int(nx / 2 ** level), int(ny / 2 ** level)
    return len(range(r, lx - 1, 2 * r + 1)), len(range(r, ly - 1, 2 * r + 1))


def stats(device=0, context=None):
    """Launch statistics of the last get_motion on ``device``'s shared
    context, or on ``context`` (HIP-event durations of the LK passes when it
    ran with ``timing=True``)."""
    st = _lib.WrStats()
    ctx = context if context is not None else _lib.context(device)
    _lib.check(_lib.load().wr_get_stats(ctx, ctypes.byref(st)), "wr_get_stats")
    return st


def correct_motion(vol, motion, out=None, context=None, stream=None):
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
    ctx = context if context is not None else _lib.context(vol.device.index or 0)
    strm = _stream() if stream is None else ctypes.c_void_p(stream.cuda_stream)
    rc = lib.wr_correct_motion(ctx, _ptr(vol), dtype_code(vol), nx, ny, nz, _ptr(motion),
                               _ptr(out), strm)
    _lib.check(rc, "wr_correct_motion")
    return out


def pyramid_level(vol, level, context=None, stream=None):
    """Level image ``imresize(vol, (x/2^L, y/2^L, z))`` as float32 (Z, Y_L, X_L)."""
    torch = _torch()
    _check_vol(vol, "volume")
    nz, ny, nx = vol.shape
    lx, ly = int(nx / 2 ** level), int(ny / 2 ** level)
    out = torch.empty((nz, max(ly, 1), max(lx, 1)), dtype=torch.float32, device=vol.device)
    lib = _lib.load()
    ctx = context if context is not None else _lib.context(vol.device.index or 0)
    strm = _stream() if stream is None else ctypes.c_void_p(stream.cuda_stream)
    rc = lib.wr_pyramid_level(ctx, _ptr(vol), dtype_code(vol), nx, ny, nz, int(level), _ptr(out), strm)
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


def robust_mean_std(data, percentile=95):
    """``robust_mean_std`` (preprocess.py:69-84) of a float32 or uint16 device
    tensor, in float32 arithmetic: ``(threshold, mean, std)`` as numpy
    float32 scalars, where threshold is ``numpy.percentile(data, percentile)``
    and mean / std are over the values at or below it."""
    _check_dense(data, "data")
    out = (ctypes.c_float * 3)()
    lib = _lib.load()
    ctx = _lib.context(data.device.index or 0)
    rc = lib.wr_robust_mean_std(ctx, _ptr(data), dtype_code(data), int(data.numel()), float(percentile),
                                out, _stream())
    _lib.check(rc, "wr_robust_mean_std")
    return np.float32(out[0]), np.float32(out[1]), np.float32(out[2])


def get_mask(vol, thres_factor, mu_sigma=None, out=None):
    """``getMask`` (mask.py:59-111) of a (Z, Y, X) float32 or uint16 device
    volume: uint8 (Z, Y, X) tensor, 1 where the float32 z-score
    ``|(v - mu) / sigma|`` exceeds ``thres_factor`` after a 3x3 in-plane
    opening and dilation.  ``mu_sigma`` overrides the robust statistics."""
    torch = _torch()
    _check_vol(vol, "volume")
    nz, ny, nx = vol.shape
    if out is None:
        out = torch.empty(vol.shape, dtype=torch.uint8, device=vol.device)
    ms = None if mu_sigma is None else (ctypes.c_float * 2)(float(mu_sigma[0]), float(mu_sigma[1]))
    lib = _lib.load()
    ctx = _lib.context(vol.device.index or 0)
    rc = lib.wr_get_mask(ctx, _ptr(vol), dtype_code(vol), nx, ny, nz, float(thres_factor), ms, _ptr(out),
                         _stream())
    _lib.check(rc, "wr_get_mask")
    return out


def bwareafilt3(mask, size_range, out_dtype="uint8", count=False):
    """``bwareafilt3_wei`` (mask.py:8-56) of a (Z, Y, X) uint8 device mask:
    the 26-connected components whose voxel count lies in
    ``[size_range[0], size_range[1]]``, as uint8 or float32 (0 / 1).  With
    ``count=True`` returns ``(mask, number_of_components)``."""
    torch = _torch()
    _check_vol(mask, "mask")
    if mask.dtype != torch.uint8:
        raise ValueError("mask must be a uint8 tensor")
    if out_dtype not in ("uint8", "float32"):
        raise ValueError("out_dtype must be 'uint8' or 'float32'")
    nz, ny, nx = mask.shape
    out = torch.empty(mask.shape, dtype=getattr(torch, out_dtype), device=mask.device)
    n = ctypes.c_longlong(-1)
    lib = _lib.load()
    ctx = _lib.context(mask.device.index or 0)
    rc = lib.wr_bwareafilt3(ctx, _ptr(mask), nx, ny, nz, float(size_range[0]), float(size_range[1]),
                            _ptr(out) if out_dtype == "uint8" else None,
                            _ptr(out) if out_dtype == "float32" else None,
                            ctypes.byref(n) if count else None, _stream())
    _lib.check(rc, "wr_bwareafilt3")
    return (out, int(n.value)) if count else out


def mean_sq_diff(a, b):
    """``np.mean((a - b) ** 2)`` (registration.py:328-329) of two device
    tensors of equal size in numpy's dtypes (float32, uint16, float64):
    uint16 differences and squares wrap modulo 2^16; returned as numpy
    float32 when both operands are float32 / uint16 with one float32, else as
    float64."""
    torch = _torch()
    _check_dense(a, "a")
    _check_dense(b, "b")
    if a.numel() != b.numel():
        raise ValueError("a and b must have the same number of elements")
    out = ctypes.c_double()
    lib = _lib.load()
    ctx = _lib.context(a.device.index or 0)
    rc = lib.wr_mean_sq_diff(ctx, _ptr(a), dtype_code(a), _ptr(b), dtype_code(b), int(a.numel()),
                             ctypes.byref(out), _stream())
    _lib.check(rc, "wr_mean_sq_diff")
    kinds = {a.dtype, b.dtype}
    if torch.float64 in kinds or kinds == {torch.uint16}:
        return np.float64(out.value)
    return np.float32(out.value)


def _check_dense(t, name):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


_METHODS = {"raw": 0, "sqrt": 1, "log2": 2, "log10": 3}


def channel_transform(image, k, method):
    """``transform(image, k, method)`` (registration.py:26-36) of a float32 or
    uint16 device tensor, as a float32 tensor of the same shape."""
    torch = _torch()
    _check_dense(image, "image")
    if method not in _METHODS:
        raise ValueError(f"Unknown method to process the image:{method}")
    k_int = isinstance(k, (int, np.integer)) and not isinstance(k, bool)
    if k_int and image.dtype == torch.uint16 and method == "raw" and not 0 <= int(k) <= 65535:
        raise OverflowError(f"Python integer {k} out of bounds for uint16")
    out = torch.empty(image.shape, dtype=torch.float32, device=image.device)
    lib = _lib.load()
    ctx = _lib.context(image.device.index or 0)
    rc = lib.wr_channel_transform(ctx, _ptr(image), dtype_code(image), int(image.numel()), float(k),
                                  _METHODS[method], 1 if k_int else 0, _ptr(out), _stream())
    _lib.check(rc, "wr_channel_transform")
    return out


def mean_std(image):
    """``(np.mean(image), np.std(image))`` of a float32 device tensor, as
    numpy float32 scalars (float64 sums)."""
    torch = _torch()
    _check_dense(image, "image")
    if image.dtype != torch.float32:
        raise ValueError("image must be float32")
    out = (ctypes.c_float * 2)()
    lib = _lib.load()
    ctx = _lib.context(image.device.index or 0)
    _lib.check(lib.wr_mean_std(ctx, _ptr(image), int(image.numel()), out, _stream()), "wr_mean_std")
    return np.float32(out[0]), np.float32(out[1])


def normalize_std(mean, std, image):
    """``normalize_std(mean, std, image)`` (preprocess.py:288-293) of a
    float32 device tensor: rescaled to the given mean and std."""
    torch = _torch()
    _check_dense(image, "image")
    if image.dtype != torch.float32:
        raise ValueError("image must be float32")
    out = torch.empty_like(image)
    lib = _lib.load()
    ctx = _lib.context(image.device.index or 0)
    rc = lib.wr_normalize_std(ctx, _ptr(image), int(image.numel()), float(mean), float(std), _ptr(out),
                              _stream())
    _lib.check(rc, "wr_normalize_std")
    return out


def add_volumes(a, b):
    """``a + b`` for a float32/uint16 and a float32 device tensor of equal
    size, as float32 (numpy's type for these operands)."""
    torch = _torch()
    _check_dense(a, "a")
    _check_dense(b, "b")
    if b.dtype != torch.float32 or a.numel() != b.numel():
        raise ValueError("b must be float32 with a's size")
    out = torch.empty(a.shape, dtype=torch.float32, device=a.device)
    lib = _lib.load()
    ctx = _lib.context(a.device.index or 0)
    rc = lib.wr_add_volumes(ctx, _ptr(a), dtype_code(a), _ptr(b), int(a.numel()), _ptr(out), _stream())
    _lib.check(rc, "wr_add_volumes")
    return out


def frames_mean(frames):
    """Mean of a sequence of equally shaped float32 or uint16 device tensors,
    summed in the given order: float64 for uint16 frames, float32 for
    float32 (``frames[idx].mean(axis=0)`` in numpy's types)."""
    torch = _torch()
    frames = list(frames)
    if not frames:
        raise ValueError("no frames")
    f0 = frames[0]
    for f in frames:
        _check_dense(f, "frame")
        if f.shape != f0.shape or f.dtype != f0.dtype:
            raise ValueError("frames must share shape and dtype")
    out_dtype = torch.float64 if f0.dtype == torch.uint16 else torch.float32
    out = torch.empty(f0.shape, dtype=out_dtype, device=f0.device)
    ptrs = (ctypes.c_void_p * len(frames))(*[f.data_ptr() for f in frames])
    lib = _lib.load()
    ctx = _lib.context(f0.device.index or 0)
    rc = lib.wr_frames_mean(ctx, ptrs, len(frames), dtype_code(f0), int(f0.numel()), _ptr(out), _stream())
    _lib.check(rc, "wr_frames_mean")
    return out


def frame_gram(slab, ds):
    """Gram matrix (float64 numpy (T, T)) of the demeaned float32 rows of a
    (T, Zs, Y, X) device tensor sampled at ``[:, :, ::ds, ::ds]``."""
    _check_dense(slab, "slab")
    if slab.dim() != 4:
        raise ValueError("slab must be (T, Zs, Y, X)")
    nt, nzs, ny, nx = slab.shape
    cc = np.zeros((nt, nt), dtype=np.float64)
    lib = _lib.load()
    ctx = _lib.context(slab.device.index or 0)
    rc = lib.wr_frame_gram(ctx, _ptr(slab), dtype_code(slab), nt, nzs, ny, nx, int(ds),
                           cc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _stream())
    _lib.check(rc, "wr_frame_gram")
    return cc
