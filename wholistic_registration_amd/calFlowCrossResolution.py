"""Drop-in for ``src/wholistic_registration/utils/calFlowCrossResolution.py``
on MI355X: the cross-resolution solver and the fixed-plane projections.

* ``getMotion_v2`` (:2470-2663) with ``optimize_layer_cross_resolution``
  (:2031-2237), ``correct_wrong_regions_one_layer`` (:2240-2467),
  ``build_bad_region_mask_from_cp_error`` (:1570-1787),
  ``build_reference_trap_mask_from_bad_moving_fast_roi`` (:1413-1567),
  ``compute_valid_mask_on_moving_grid`` (:2006-2028), ``_make_ball``
  (:1363-1396) and the spline interpolant ``generate_continuous_H_gpu`` /
  ``apply_H_to_matrix_gpu`` (:463-477, :1144-1153);
* ``project_coords_to_fixed_planes_gpu`` (:878-1141, kernel :480-568) and
  ``project_coords_to_fixed_planes_weighted_gpu`` (:700-875, kernel
  :571-669), including the sampling of the reference volume when
  ``values_xyk`` is None.

Same names, arguments, option keys, defaults, printed logs, validation
messages and outputs as the reference; the arithmetic runs in libwrhip.so
(``wr_spline_*``, ``wr_xres_*``, ``wr_project_planes_*``) on device buffers
that torch allocates.  Volumes cross the boundary in the reference's
``(x, y, z)`` layout (numpy or torch) and live on the device as ``(z, y, x)``
arrays; fields as 3 planes of ``(z, y, x)``.  ``getMotion_v2`` returns numpy
arrays like the reference; the projections return torch device tensors (the
reference returns CuPy arrays) or numpy arrays with ``return_numpy=True``.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import device as _dev
from ._util import torch_module as _torch, is_tensor as _is_tensor, hip_device as _device


def _f32_device(a, dev):
    torch = _torch()
    if _is_tensor(a):
        return a.to(dev, torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float32))).to(dev)


def _common(coords_ref_xyk_xyz, ref_volume, target_z_planes, values_xyk, ref_volume_order, downsample_xy):
    dev = _device()
    coords = _f32_device(coords_ref_xyk_xyz, dev)
    if coords.dim() != 4 or coords.shape[-1] != 3:
        raise ValueError(f"coords_ref_xyk_xyz should have shape (X,Y,K,3), got {tuple(coords.shape)}")
    shape = tuple(ref_volume.shape)
    if len(shape) != 3:
        raise ValueError(f"ref_volume should be 3D, got {shape}")
    if ref_volume_order == "xyz":
        xref, yref, _ = shape
    elif ref_volume_order == "zyx":
        _, yref, xref = shape
    else:
        raise ValueError("ref_volume_order should be 'xyz' or 'zyx'.")
    downsample_xy = int(downsample_xy)
    if downsample_xy < 1:
        raise ValueError("downsample_xy should be >= 1.")
    xout = (xref + downsample_xy - 1) // downsample_xy
    yout = (yref + downsample_xy - 1) // downsample_xy
    tz = _f32_device(np.atleast_1d(np.asarray(target_z_planes, dtype=np.float32)) if not _is_tensor(target_z_planes)
                     else target_z_planes.reshape(-1), dev)
    if values_xyk is None:
        # the reference volume's cubic spline at the coordinates (:803-805, :1033-1037)
        ref = _f32_device(ref_volume, dev)
        ref_zyx = ref.permute(2, 1, 0).contiguous() if ref_volume_order == "xyz" else ref.contiguous()
        h = generate_continuous_H_gpu(ref_zyx, 1, layout="zyx")
        values = h(coords).to(_torch().float32)
    else:
        values = _f32_device(values_xyk, dev)
        if tuple(values.shape) != tuple(coords.shape[:-1]):
            raise ValueError(f"values_xyk shape {tuple(values.shape)} does not match coords shape "
                             f"{tuple(coords.shape[:-1])}")
    return dev, coords.reshape(-1, 3).contiguous(), values.reshape(-1).contiguous(), tz, xref, yref, xout, yout, \
        downsample_xy


def _finish(out, output_order, return_numpy):
    if output_order == "zyx":
        out = out.transpose(1, 2)
    elif output_order != "zxy":
        raise ValueError("output_order should be 'zyx' or 'zxy'.")
    return out.cpu().numpy() if return_numpy else out


def project_coords_to_fixed_planes_gpu(coords_ref_xyk_xyz, ref_volume=None, target_z_planes=None, values_xyk=None,
                                       ref_volume_order="xyz", z_window=1.5, downsample_xy=1, fill_value=0.0,
                                       return_numpy=False, output_order="zyx", xy_splat_mode="subpixel_footprint",
                                       xy_extra_radius=0):
    """Max projection of the samples onto fixed reference z planes
    (calFlowCrossResolution.py:878-1141): ``(Nplanes, Yout, Xout)`` for
    ``output_order="zyx"``, ``(Nplanes, Xout, Yout)`` for ``"zxy"``."""
    torch = _torch()
    if target_z_planes is None:
        raise ValueError("target_z_planes must be provided.")
    if ref_volume is None:
        raise ValueError("ref_volume must be provided to infer reference shape.")
    if xy_splat_mode == "nearest":
        nearest, radius = 1, 0
    elif xy_splat_mode == "subpixel_footprint":
        nearest, radius = 0, int(xy_extra_radius)
        if radius < 0:
            raise ValueError("xy_extra_radius should be >= 0.")
    else:
        raise ValueError("xy_splat_mode should be 'nearest' or 'subpixel_footprint'.")
    dev, coords, values, tz, xref, yref, xout, yout, ds = _common(
        coords_ref_xyk_xyz, ref_volume, target_z_planes, values_xyk, ref_volume_order, downsample_xy)
    out = torch.empty((tz.numel(), xout, yout), dtype=torch.float32, device=dev)
    lib = _lib.load()
    ctx = _lib.context(dev.index or 0)
    rc = lib.wr_project_planes_max(ctx, ctypes.c_void_p(values.data_ptr()), ctypes.c_void_p(coords.data_ptr()),
                                   values.numel(), ctypes.c_void_p(tz.data_ptr()), tz.numel(), xref, yref, ds,
                                   float(z_window), nearest, radius, 0 if fill_value is None else 1,
                                   0.0 if fill_value is None else float(fill_value),
                                   ctypes.c_void_p(out.data_ptr()), _dev._stream())
    _lib.check(rc, "wr_project_planes_max")
    return _finish(out, output_order, return_numpy)


def project_coords_to_fixed_planes_weighted_gpu(coords_ref_xyk_xyz, ref_volume, target_z_planes, values_xyk=None,
                                                ref_volume_order="xyz", z_window=1.5, z_sigma=None,
                                                z_weight_mode="gaussian", downsample_xy=1, fill_value=0.0,
                                                return_numpy=False, output_order="zyx", eps=1e-6):
    """Weighted-average projection onto fixed reference z planes
    (calFlowCrossResolution.py:700-875): ``(out, weights)``, each
    ``(Nplanes, Yout, Xout)`` for ``output_order="zyx"``."""
    torch = _torch()
    modes = {"hard": 0, "triangular": 1, "gaussian": 2}
    if z_weight_mode not in modes:
        raise ValueError("z_weight_mode should be 'hard', 'triangular', or 'gaussian'.")
    dev, coords, values, tz, xref, yref, xout, yout, ds = _common(
        coords_ref_xyk_xyz, ref_volume, target_z_planes, values_xyk, ref_volume_order, downsample_xy)
    if z_sigma is None:
        z_sigma = float(z_window) * 0.5
    out = torch.empty((tz.numel(), xout, yout), dtype=torch.float32, device=dev)
    weights = torch.empty_like(out)
    lib = _lib.load()
    ctx = _lib.context(dev.index or 0)
    rc = lib.wr_project_planes_weighted(ctx, ctypes.c_void_p(values.data_ptr()), ctypes.c_void_p(coords.data_ptr()),
                                        values.numel(), ctypes.c_void_p(tz.data_ptr()), tz.numel(), xref, yref, ds,
                                        float(z_window), float(z_sigma), modes[z_weight_mode], float(eps),
                                        float(fill_value), ctypes.c_void_p(out.data_ptr()),
                                        ctypes.c_void_p(weights.data_ptr()), _dev._stream())
    _lib.check(rc, "wr_project_planes_weighted")
    if output_order not in ("zyx", "zxy"):
        raise ValueError("output_order should be 'zyx' or 'zxy'.")
    return _finish(out, output_order, return_numpy), _finish(weights, output_order, return_numpy)


# ============================================================================
# the spline interpolant (:463-477, :1144-1153)
# ============================================================================

SPLINE_PAD = 12


def _ctx_stream(dev):
    return _lib.context(dev.index or 0), _dev._stream()


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _spline_coef(vol_zyx):
    """Cubic B-spline coefficients (float64, 12-voxel padded) of a (z, y, x)
    device volume of float64, float32 or uint8."""
    torch = _torch()
    codes = {torch.float64: _lib.WR_F64, torch.float32: _lib.WR_F32, torch.uint8: _lib.WR_U8}
    if vol_zyx.dtype not in codes:
        raise ValueError(f"unsupported volume dtype {vol_zyx.dtype}")
    nz, ny, nx = vol_zyx.shape
    coef = torch.empty((nz + 2 * SPLINE_PAD) * (ny + 2 * SPLINE_PAD) * (nx + 2 * SPLINE_PAD),
                       dtype=torch.float64, device=vol_zyx.device)
    ctx, st = _ctx_stream(vol_zyx.device)
    _lib.check(_lib.load().wr_spline_coef(ctx, _vp(vol_zyx), codes[vol_zyx.dtype], nx, ny, nz, _vp(coef), st),
               "wr_spline_coef")
    return coef


class _Spline:
    """``H`` of ``generate_continuous_H_gpu``: callable on ``(..., 3)`` point
    arrays in (x, y, z) index units; values in the volume's float type
    (float64 for float64 volumes, float32 otherwise, as scipy's spline
    interpolation returns them)."""

    def __init__(self, vol_zyx, z_ratio):
        self.shape_zyx = tuple(vol_zyx.shape)
        self.coef = _spline_coef(vol_zyx.contiguous())
        self.z_ratio = float(z_ratio)
        torch = _torch()
        self.out_dtype = torch.float64 if vol_zyx.dtype == torch.float64 else torch.float32

    def __call__(self, coords_phys):
        torch = _torch()
        pts = _f32_device(coords_phys, self.coef.device)
        shp = tuple(pts.shape[:-1])
        flat = pts.reshape(-1, 3).contiguous()
        out = torch.empty(flat.shape[0], dtype=torch.float64, device=self.coef.device)
        nz, ny, nx = self.shape_zyx
        ctx, st = _ctx_stream(self.coef.device)
        _lib.check(_lib.load().wr_spline_sample(ctx, _vp(self.coef), nx, ny, nz, _vp(flat), flat.shape[0],
                                                self.z_ratio, _vp(out), st), "wr_spline_sample")
        return out.to(self.out_dtype).reshape(shp)


def generate_continuous_H_gpu(stack, zRatio, layout="xyz"):
    """The cubic-spline interpolant of ``stack`` (``(X, Y, Z)`` like the
    reference, or a ``(z, y, x)`` device volume with ``layout="zyx"``)."""
    torch = _torch()
    dev = _device()
    if _is_tensor(stack):
        v = stack.to(dev)
    else:
        a = np.asarray(stack)
        if a.dtype not in (np.float64, np.float32, np.uint8):
            a = a.astype(np.float64 if a.dtype.kind == "f" else np.float32)
        v = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    if v.dtype == torch.bool:
        v = v.to(torch.uint8)
    if v.dim() != 3:
        raise ValueError(f"stack should be 3D, got {tuple(v.shape)}")
    if layout == "xyz":
        v = v.permute(2, 1, 0)
    return _Spline(v.contiguous(), zRatio)


def apply_H_to_matrix_gpu(A, H):
    """``H`` at the points of ``A`` (``(..., 3)``)."""
    return H(A)


# ============================================================================
# the cross-resolution solver (getMotion_v2)
# ============================================================================

_MAX_BALL_RADIUS_Z = 64


def _make_ball(radius_xy, z_ratio):
    """The anisotropic ball structuring element (:1363-1396), as a numpy bool
    array ``(2R+1, 2R+1, 2Rz+1)`` over (x, y, z).  The device kernels build
    the same element from the same rule."""
    radius_xy = int(max(1, radius_xy))
    z_ratio = float(z_ratio)
    if not np.isfinite(z_ratio) or z_ratio <= 0:
        raise ValueError(
            f"_make_ball: z_ratio must be a positive, finite physical-units-per-z-index "
            f"ratio, got {z_ratio!r}.")
    radius_z = int(max(1, min(np.ceil(radius_xy / z_ratio), _MAX_BALL_RADIUS_Z)))
    r = np.arange(-radius_xy, radius_xy + 1, dtype=np.float32)
    rz = np.arange(-radius_z, radius_z + 1, dtype=np.float32)
    gx, gy, gz = np.meshgrid(r, r, rz, indexing="ij")
    return (gx ** 2 + gy ** 2 + (gz * z_ratio) ** 2) <= radius_xy ** 2 + 1e-6


def _to_zyx(a, dev, dtype):
    """A reference-layout (x, y, z) array as a contiguous (z, y, x) device tensor."""
    torch = _torch()
    if _is_tensor(a):
        t = a.to(dev)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a))).to(dev)
    if t.dtype == torch.bool:
        t = t.to(torch.uint8)
    return t.permute(2, 1, 0).to(dtype).contiguous()


def _field_to_planes(a, dev):
    """An (x, y, z, 3) field as (3, z, y, x) float32 device planes."""
    torch = _torch()
    t = a.to(dev) if _is_tensor(a) else torch.from_numpy(np.ascontiguousarray(np.asarray(a))).to(dev)
    return t.to(torch.float32).permute(3, 2, 1, 0).contiguous()


def _planes_to_field(p):
    """(3, z, y, x) planes as an (x, y, z, 3) host array."""
    return p.permute(3, 2, 1, 0).cpu().numpy()


def _imresize_zyx(vol, nx_out, ny_out, method=0):
    """imresize of a (z, y, x) device volume in x and y (float64 out)."""
    torch = _torch()
    nz, ny, nx = vol.shape
    out = torch.empty((nz, ny_out, nx_out), dtype=torch.float64, device=vol.device)
    ctx, st = _ctx_stream(vol.device)
    code = _dev.dtype_code(vol)
    _lib.check(_lib.load().wr_imresize(ctx, _vp(vol), code, nx, ny, nz, nx_out, ny_out, method, _vp(out), st),
               "wr_imresize")
    return out


def _scale(vol64, mul=1.0, div=1.0, out=None):
    torch = _torch()
    if out is None:
        out = torch.empty(vol64.shape, dtype=torch.float32, device=vol64.device)
    ctx, st = _ctx_stream(vol64.device)
    _lib.check(_lib.load().wr_xres_scale(ctx, _vp(vol64), vol64.numel(), float(mul), float(div), _vp(out), st),
               "wr_xres_scale")
    return out


def resample_exclude_mask(mask_zyx, nx_out, ny_out, threshold=0.5):
    """``resample_exclude_mask`` (:1917-1938) of a (z, y, x) float32 device
    mask: bicubic imresize > threshold, as bytes."""
    torch = _torch()
    r = _imresize_zyx(mask_zyx, nx_out, ny_out, 0)
    out = torch.empty(r.shape, dtype=torch.uint8, device=r.device)
    ctx, st = _ctx_stream(r.device)
    _lib.check(_lib.load().wr_xres_threshold(ctx, _vp(r), r.numel(), float(threshold), _vp(out), st),
               "wr_xres_threshold")
    return out


def initialize_motion_for_layer(option, layer, layer_num, SZ, x, y, z, prev_motion=None, dev=None):
    """``initialize_motion_for_layer`` (:1941-1975): (3, z, y, x) planes."""
    torch = _torch()
    out = torch.zeros((3, z, y, x), dtype=torch.float32, device=dev)
    if layer == layer_num:
        init = option.get("motion")
        if init is not None:
            m = _field_to_planes(init, dev)
            for c, dim in enumerate((x, y, z)):
                _scale(_imresize_zyx(m[c], x, y, 0), div=SZ[c] / dim, out=out[c])
        return out
    for c in range(3):
        _scale(_imresize_zyx(prev_motion[c], x, y, 1), mul=2.0 if c < 2 else 1.0, out=out[c])
    return out


def initialize_phase_for_layer(option, SZ, layer, x, y, z, zRatio, zRatio_hr, dev=None):
    """``initialize_phase_for_layer`` (:1978-1992): (3, z, y, x) planes."""
    torch = _torch()
    out = torch.empty((3, z, y, x), dtype=torch.float32, device=dev)
    ph = option.get("phase")
    if ph is not None:
        p = _field_to_planes(ph, dev)
        for c, dim in enumerate((x, y, z)):
            _scale(_imresize_zyx(p[c], x, y, 0), div=SZ[c] / dim, out=out[c])
        return out
    ctx, st = _ctx_stream(dev)
    _lib.check(_lib.load().wr_xres_phase_identity(ctx, x, y, z, float(zRatio), float(zRatio_hr), _vp(out), st),
               "wr_xres_phase_identity")
    return out


def _count(mask):
    n = ctypes.c_longlong(0)
    ctx, st = _ctx_stream(mask.device)
    _lib.check(_lib.load().wr_count_u8(ctx, _vp(mask), mask.numel(), ctypes.byref(n), st), "wr_count_u8")
    return int(n.value)


class _Layer:
    """Device images of one pyramid layer.  An empty exclusion mask has an
    all-zero interpolant, so its coefficients are left out (None: nothing
    excluded) and the kernels skip its reads."""

    def __init__(self, mov_l, ref_l, mmov_l, mref_l):
        self.mov_l, self.ref_l, self.mmov_l, self.mref_l = mov_l, ref_l, mmov_l, mref_l
        self.nz, self.ny, self.nx = mov_l.shape
        self.rz, self.ry, self.rx = ref_l.shape
        self.ref_coef = _spline_coef(ref_l)
        self.mask_coef = _spline_coef(mref_l) if _count(mref_l) else None


def compute_valid_mask_on_moving_grid(phase_new, H_mask_ref_layer, mask_mov_layer, lay=None):
    """``compute_valid_mask_on_moving_grid`` (:2006-2028) with the reference
    mask's spline coefficients ``H_mask_ref_layer`` (None = nothing
    excluded): ``(valid bytes, count)``."""
    torch = _torch()
    _, nz, ny, nx = phase_new.shape
    valid = torch.empty((nz, ny, nx), dtype=torch.uint8, device=phase_new.device)
    n = ctypes.c_longlong(0)
    ctx, st = _ctx_stream(phase_new.device)
    _lib.check(_lib.load().wr_xres_valid(ctx, _vp(phase_new), nx, ny, nz, _vp(H_mask_ref_layer), lay.rx, lay.ry,
                                         lay.rz, _vp(mask_mov_layer), _vp(valid), ctypes.byref(n), st),
               "wr_xres_valid")
    return valid, int(n.value)


_REASONS = {1: "error increased repeatedly", 2: "error change below tol", 3: "motion update below tol"}


def optimize_layer_cross_resolution(lay, mask_coef, phase_current, motion_init, r, iterNum, movRange,
                                    smoothPenaltySum, zRatio_hr, tol=1e-3, verbose=False, layer=None,
                                    keep_first=False, schedule=None, trace=None):
    """``optimize_layer_cross_resolution`` (:2031-2237) on the device: returns
    the reference's result dict (motion and phase as (3, z, y, x) planes,
    sampled / residual float64, valid bytes, current / diff / penalty
    error).  ``schedule`` = (records, updates) forces the iteration
    schedule; ``trace`` (list) receives the iteration record."""
    torch = _torch()
    dev = lay.mov_l.device
    shape3 = (3, lay.nz, lay.ny, lay.nx)
    motion = torch.empty(shape3, dtype=torch.float32, device=dev)
    phase = torch.empty(shape3, dtype=torch.float32, device=dev)
    sampled = torch.empty((lay.nz, lay.ny, lay.nx), dtype=torch.float64, device=dev)
    residual = torch.empty_like(sampled)
    valid = torch.empty((lay.nz, lay.ny, lay.nx), dtype=torch.uint8, device=dev)
    first = torch.empty_like(sampled) if keep_first else None
    L = _lib.WrXresLayer(lay.nx, lay.ny, lay.nz, lay.rx, lay.ry, lay.rz, int(r), int(iterNum),
                         float(smoothPenaltySum), float(movRange), float(zRatio_hr), float(tol))
    it = int(iterNum)
    arrs = [(ctypes.c_double * it)() for _ in range(3)]
    tr = _lib.WrXresTrace()
    tr.diff_error = ctypes.cast(arrs[0], ctypes.POINTER(ctypes.c_double))
    tr.penalty_error = ctypes.cast(arrs[1], ctypes.POINTER(ctypes.c_double))
    tr.current_error = ctypes.cast(arrs[2], ctypes.POINTER(ctypes.c_double))
    forced = (ctypes.c_int * 2)(*schedule) if schedule is not None else None
    ctx, st = _ctx_stream(dev)
    _lib.check(_lib.load().wr_xres_optimize(
        ctx, ctypes.byref(L), _vp(lay.mov_l), _vp(lay.ref_l), _vp(lay.ref_coef), _vp(mask_coef), _vp(lay.mmov_l),
        _vp(phase_current), _vp(motion_init.contiguous()), forced, _vp(motion), _vp(phase), _vp(sampled),
        _vp(residual), _vp(valid), _vp(first), ctypes.byref(tr), st), "wr_xres_optimize")
    errs = [(arrs[0][k], arrs[1][k], arrs[2][k]) for k in range(tr.records)]
    if verbose:
        for k, (d, p, e) in enumerate(errs):
            print(f"[layer={layer}] iter={k:02d} error={e:.4f} diff={d:.4f} penalty={p:.4f}")
            if k == len(errs) - 1 and tr.reason in _REASONS and (tr.reason != 3 or tr.updates == tr.records):
                print(f"[layer={layer}] stop: {_REASONS[tr.reason]}")
    if trace is not None:
        reason = {0: "last", 1: "increase", 2: "flat", 3: "motion", 4: "forced"}[tr.reason]
        trace.append(dict(records=tr.records, updates=tr.updates, reason=reason, errors=errs,
                          final=(tr.final_diff, tr.final_penalty, tr.final_error)))
    return {"motion": motion, "phase_new": phase, "data_ref_sampled": sampled, "residual": residual,
            "first_residual": first, "valid_mask": valid, "current_error": tr.final_error,
            "diff_error": tr.final_diff, "penalty_error": tr.final_penalty}


def build_bad_region_mask_from_cp_error(residual, valid, r, mad_threshold=3.0, min_component_size=2,
                                        metric="mse", exclude_mode="direct", first_residual=None,
                                        quantile_threshold_q=0.8):
    """Local control-point errors, ``build_bad_region_mask_from_cp_error``
    (:1570-1787), the 'highresidual' selection and ``& valid``
    (:2306-2351): ``(bad bytes, number of regions)``."""
    torch = _torch()
    if metric.lower() not in ("mse", "mae"):
        raise ValueError("metric must be 'mse' or 'mae'")
    if exclude_mode not in ("direct", "highresidual"):
        raise ValueError(f"unknown wrong_region_exclude_mode {exclude_mode!r}")
    nz, ny, nx = residual.shape
    bad = torch.empty((nz, ny, nx), dtype=torch.uint8, device=residual.device)
    n = ctypes.c_int(0)
    ctx, st = _ctx_stream(residual.device)
    _lib.check(_lib.load().wr_xres_bad_regions(
        ctx, _vp(residual), _vp(valid), _vp(first_residual), nx, ny, nz, int(r),
        0 if metric.lower() == "mse" else 1, float(mad_threshold), int(min_component_size),
        0 if exclude_mode == "direct" else 1, float(quantile_threshold_q), _vp(bad), ctypes.byref(n), st),
        "wr_xres_bad_regions")
    return bad, int(n.value)


def build_reference_trap_mask_from_bad_moving_fast_roi(bad_mask, phase_new, lay, z_ratio_ref, expand_radius_xy=2,
                                                       sigma_grad=1.0, intensity_k=2.5):
    """``mask_ref_layer | build_reference_trap_mask_from_bad_moving_fast_roi``
    (:1413-1567, :2364): ``(corrected reference mask bytes, trap voxels)``."""
    torch = _torch()
    corrected = torch.empty((lay.rz, lay.ry, lay.rx), dtype=torch.uint8, device=lay.ref_l.device)
    n = ctypes.c_longlong(0)
    ctx, st = _ctx_stream(lay.ref_l.device)
    _lib.check(_lib.load().wr_xres_trap_mask(
        ctx, _vp(bad_mask), _vp(phase_new), lay.nx, lay.ny, lay.nz, _vp(lay.ref_l), lay.rx, lay.ry, lay.rz,
        _vp(lay.mref_l), float(z_ratio_ref), float(expand_radius_xy), float(sigma_grad), float(intensity_k),
        _vp(corrected), ctypes.byref(n), st), "wr_xres_trap_mask")
    return corrected, int(n.value)


def correct_wrong_regions_one_layer(lay, phase_current, motion_init_layer, r, iterNum, movRange, smoothPenaltySum,
                                    zRatio_hr, tol=1e-3, error_metric="mse", mad_threshold=3.0,
                                    min_component_size=2, bad_region_exclude_mode="direct", verbose=False,
                                    layer=None, expand_radius_xy=2.0, sigma_grad=1.0, intensity_k=2.0,
                                    quantile_threshold_q=0.8, schedules=None, trace=None, info=None):
    """``correct_wrong_regions_one_layer`` (:2240-2467): plain optimisation,
    bad-region detection, a rerun with the reference trap mask excluded and
    one from that field with the original mask; the rerun is kept only when
    its final objective is lower."""
    high = bad_region_exclude_mode == "highresidual"

    def run(mcoef, m0):
        return optimize_layer_cross_resolution(lay, mcoef, phase_current, m0, r, iterNum, movRange,
                                               smoothPenaltySum, zRatio_hr, tol, verbose, layer, keep_first=high,
                                               schedule=next(schedules) if schedules else None, trace=trace)

    res0 = run(lay.mask_coef, motion_init_layer)
    bad, num_regions = build_bad_region_mask_from_cp_error(
        res0["residual"], res0["valid_mask"], r, mad_threshold, min_component_size, error_metric,
        bad_region_exclude_mode, res0["first_residual"], quantile_threshold_q)
    if info is not None:
        info.update(regions=num_regions, ratio=None, accepted=False)
    if verbose:
        print(f"[layer={layer}] detected wrong regions: {num_regions}")
    if num_regions == 0:
        return res0
    corrected, n_trap = build_reference_trap_mask_from_bad_moving_fast_roi(
        bad, res0["phase_new"], lay, zRatio_hr, expand_radius_xy / (2.0 ** layer), sigma_grad, intensity_k)
    mask_coef1 = _spline_coef(corrected)
    n_valid_old = _count(res0["valid_mask"])
    _, n_valid_new = compute_valid_mask_on_moving_grid(res0["phase_new"], mask_coef1, lay.mmov_l, lay)
    valid_ratio = n_valid_new / max(n_valid_old, 1)
    if info is not None:
        info.update(ratio=valid_ratio, trap=n_trap, bad=_count(bad))
    if verbose:
        print(f"[layer={layer}] mask kept ratio after correction: {valid_ratio:.3f}")
    if valid_ratio < 0.4:
        if verbose:
            print(f"[layer={layer}] skip correction: corrected mask too small")
        return res0
    res1 = run(mask_coef1, motion_init_layer)
    res2 = run(lay.mask_coef, res1["motion"])
    if verbose:
        print(f"[layer={layer}] compare original vs corrected: "
              f"{float(res0['current_error']):.4f} -> {float(res2['current_error']):.4f}")
    if res2["current_error"] < res0["current_error"]:
        if info is not None:
            info["accepted"] = True
        if verbose:
            print(f"[layer={layer}] accept corrected result")
        return res2
    if verbose:
        print(f"[layer={layer}] keep original result")
    return res0


def getMotion_v2(data_mov, data_ref, option, verbose=False, *, schedules=None, trace=None, info=None,
                 output="numpy"):
    """Cross-resolution registration with wrong-region correction
    (calFlowCrossResolution.py:2470-2663).

    ``data_mov`` (X, Y, Z_mov) and ``data_ref`` (X_hr, Y_hr, Z_ref), numpy or
    torch; ``option`` with the reference's keys and defaults.  Returns
    ``(phase_new, motion_current, data_ref_sampled)`` as numpy arrays
    ((x, y, z, 3) float32, (x, y, z, 3) float32, (x, y, z) float64).

    Beyond the reference: ``output="device"`` returns the device tensors
    instead (phase and motion as (3, z, y, x) float32 planes, the sampled
    reference (z, y, x) float64), for callers that keep working on the GPU.
    Testing: ``schedules`` (list of (records, updates) per optimisation, in
    call order) forces the iteration schedules; ``trace`` (list) and
    ``info`` (dict) receive every optimisation's record and each layer's
    wrong-region outcome.
    """
    if output not in ("numpy", "device"):
        raise ValueError("output must be 'numpy' or 'device'")
    torch = _torch()
    dev = _device()
    option["mask_ref"] = np.asarray(option["mask_ref"], dtype=np.float32) if not _is_tensor(option["mask_ref"]) \
        else option["mask_ref"].to(torch.float32)
    option["mask_mov"] = np.asarray(option["mask_mov"], dtype=np.float32) if not _is_tensor(option["mask_mov"]) \
        else option["mask_mov"].to(torch.float32)
    layer_num = int(option["layer"])
    iterNum = int(option["iter"])
    r = int(option["r"])
    zRatio_raw = float(option.get("zRatio", 1.0))
    zRatio_HR = float(option["zRatio_HR"])
    smoothPenalty = float(option["smoothPenalty"])
    movRange = float(option.get("movRange", 5.0))
    tol = float(option.get("tol", 1e-3))
    wrong_region_enable = bool(option.get("wrong_region_enable", True))
    wrong_region_metric = option.get("wrong_region_metric", "mse")
    wrong_region_mad_threshold = float(option.get("wrong_region_mad_threshold", 3.0))
    wrong_region_min_component_size = int(option.get("wrong_region_min_component_size", 2))
    wrong_region_exclude_mode = option.get("wrong_region_exclude_mode", "direct")
    expand_radius_xy = float(option.get("expand_radius_xy", 2.0))
    sigma_grad = float(option.get("sigma_grad", 1.0))
    intensity_k = float(option.get("intensity_k", 2.0))
    quantile_threshold_q = float(option.get("quantile_threshold_q", 0.8))

    mov = _to_zyx(data_mov, dev, torch.float32)
    ref = _to_zyx(data_ref, dev, torch.float32)
    mask_mov = _to_zyx(option["mask_mov"], dev, torch.float32)
    mask_ref = _to_zyx(option["mask_ref"], dev, torch.float32)
    SZ = tuple(int(v) for v in reversed(mov.shape))
    SZ_HR = tuple(int(v) for v in reversed(ref.shape))
    sched = iter(schedules) if schedules is not None else None
    patch = (2 * r + 1) ** 2
    smoothPenaltySum = smoothPenalty * patch
    res = None
    for layer in range(layer_num, -1, -1):
        if verbose:
            print(f"\n========== start layer {layer}/{layer_num} ==========")
        x, y, z = int(SZ[0] / (2 ** layer)), int(SZ[1] / (2 ** layer)), SZ[2]
        x_hr, y_hr = int(SZ_HR[0] / (2 ** layer)), int(SZ_HR[1] / (2 ** layer))
        lay = _Layer(_imresize_zyx(mov, x, y), _imresize_zyx(ref, x_hr, y_hr),
                     resample_exclude_mask(mask_mov, x, y), resample_exclude_mask(mask_ref, x_hr, y_hr))
        zRatio = zRatio_raw / (2 ** layer)
        zRatio_hr = zRatio_HR / (2 ** layer)
        m0 = initialize_motion_for_layer(option, layer, layer_num, SZ, x, y, z,
                                         None if res is None else res["motion"], dev)
        ph0 = initialize_phase_for_layer(option, SZ, layer, x, y, z, zRatio, zRatio_hr, dev)
        lvl = {} if info is not None else None
        if wrong_region_enable:
            res = correct_wrong_regions_one_layer(
                lay, ph0, m0, r, iterNum, movRange, smoothPenaltySum, zRatio_hr, tol, wrong_region_metric,
                wrong_region_mad_threshold, wrong_region_min_component_size, wrong_region_exclude_mode, verbose,
                layer, expand_radius_xy, sigma_grad, intensity_k, quantile_threshold_q, schedules=sched,
                trace=trace, info=lvl)
        else:
            res = optimize_layer_cross_resolution(lay, lay.mask_coef, ph0, m0, r, iterNum, movRange,
                                                  smoothPenaltySum, zRatio_hr, tol, verbose, layer,
                                                  schedule=next(sched) if sched else None, trace=trace)
        if info is not None:
            info[layer] = lvl
    if output == "device":
        return res["phase_new"], res["motion"], res["data_ref_sampled"]
    phase_np = _planes_to_field(res["phase_new"])
    motion_np = _planes_to_field(res["motion"])
    sampled_np = res["data_ref_sampled"].permute(2, 1, 0).cpu().numpy()
    return phase_np, motion_np, sampled_np


# ============================================================================
# z-offset search (:370-460)
# ============================================================================


def _stack_3d(a, what, dev):
    torch = _torch()
    t = a.to(dev, torch.float32) if _is_tensor(a) else torch.from_numpy(np.asarray(a, dtype=np.float32)).to(dev)
    if t.dim() != 3:
        raise ValueError(f"{what} must be 3D (H,W,Z), got shape {tuple(t.shape)}")
    return t


def FindInitZ_stack_global_fixed_spacing(data_mov, data_ref, delta_ref_idx, use_gradient=True, use_hann_weight=True,
                                         direction=1, weight_eps=1e-6, return_debug=False):
    """The reference z of each moving slice (slices ``delta_ref_idx *
    direction`` apart), chosen by the summed Hann-weighted ZNCC of the
    slices' Sobel gradient magnitudes (calFlowCrossResolution.py:370-460).
    ``data_mov`` (H, W, K), ``data_ref`` (H, W, Z); returns the z indices as
    uint16 (and the reference's debug dict with ``return_debug``)."""
    dev = _device()
    mov = _stack_3d(data_mov, "data_ref", dev)
    ref = _stack_3d(data_ref, "data_ref", dev)
    h, w, k_n = mov.shape
    hr, wr_, z_n = ref.shape
    if (hr, wr_) != (h, w):
        raise ValueError(f"XY size mismatch: data_mov has {(h, w)}, data_ref has {(hr, wr_)}")
    if k_n < 1:
        raise ValueError("Moving input must have at least one slice.")
    if z_n < 1:
        raise ValueError("Reference input must have at least one slice.")
    m = mov.permute(2, 1, 0).contiguous()
    r = ref.permute(2, 1, 0).contiguous()
    z_init = np.zeros(k_n, dtype=np.uint16)
    scores = np.zeros((k_n, z_n), dtype=np.float32)
    best = ctypes.c_double(0.0)
    z0 = ctypes.c_int(0)
    ctx, st = _ctx_stream(dev)
    _lib.check(_lib.load().wr_find_init_z(
        ctx, _vp(m), k_n, _vp(r), z_n, h, w, float(delta_ref_idx), int(bool(use_gradient)),
        int(bool(use_hann_weight)), int(direction), float(weight_eps), z_init.ctypes.data_as(ctypes.c_void_p),
        scores.ctypes.data_as(ctypes.c_void_p), ctypes.byref(best), ctypes.byref(z0), st), "FindInitZ")
    if return_debug:
        return z_init, {"scores": scores, "best_score": float(best.value), "best_z0": int(z0.value),
                        "delta_ref_idx": float(delta_ref_idx), "direction": int(direction)}
    return z_init
