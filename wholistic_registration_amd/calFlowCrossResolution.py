"""Drop-in for the fixed-plane projections of
``src/wholistic_registration/utils/calFlowCrossResolution.py`` on MI355X.

* ``project_coords_to_fixed_planes_gpu`` (:878-1141, kernel :480-568)
* ``project_coords_to_fixed_planes_weighted_gpu`` (:700-875, kernel :571-669)

Same names, arguments, validation messages and outputs as the reference.  The
splats, the value shift of the max projection and the normalisation of the
weighted one run in libwrhip.so (``wr_project_planes_max``,
``wr_project_planes_weighted``).  Outputs are torch device tensors (the
reference returns CuPy arrays), or numpy arrays with ``return_numpy=True``.

Not built: the rest of the cross-resolution solver (``getMotion_v2``), and
sampling the reference volume at the coordinates when ``values_xyk`` is None
(the reference evaluates a prefiltered cubic spline of order 3 there): pass
the values to project.
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
        raise NotImplementedError(
            "sampling ref_volume at the coordinates (cubic spline, order 3) is not part of this "
            "package: pass values_xyk")
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
