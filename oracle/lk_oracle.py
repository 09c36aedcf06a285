"""numpy/scipy restatement of the reference registration hot path (the oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of bench.py.  The product path never calls it.

Layout follows the reference solver: arrays are indexed ``(x, y, z)`` and a
motion field is ``(x, y, z, 3)`` with component ``c`` displacing axis ``c``.
All file:line citations are into the reference tree
``src/wholistic_registration/utils/``.

Modes
-----
``mode="ref"``
    Reproduces the reference's dtype flow: float64 level images, float32
    warp input, float32 coordinates and motion, float64 structure tensor and
    solve, float32 neighbour differences (calFlow3d_Wei_v1.py:37-706).
``mode="fp64"``
    Everything that the reference keeps in float32 is carried in float64
    instead.  This is the FP64 oracle for endpoint-error reporting.

Patch sums
----------
``sums="correlate"`` runs the reference's nine full-grid
``scipy.ndimage.correlate`` calls and samples them at the control points
(calFlow3d_Wei_v1.py:546-588).  ``sums="window"`` (default) sums each
(2r+1)x(2r+1)x1 window at the control points directly, with the same clamped
indices; it is the same sum in a different float64 order and is thousands of
times cheaper at large r.
"""

from __future__ import annotations

from math import ceil

import numpy as np
from scipy import ndimage as ndi

# ----------------------------------------------------------------------------
# imresize (MATLAB-compatible antialiased resampling), imresize.py:41-191
# ----------------------------------------------------------------------------


def keys_cubic(t):
    """Keys cubic kernel, a = -0.5 (imresize.py:49-57)."""
    a = np.abs(np.asarray(t).astype(np.float64))
    a2 = np.multiply(a, a)
    a3 = np.multiply(a2, a)
    inner = np.multiply(1.5 * a3 - 2.5 * a2 + 1, a <= 1)
    outer = np.multiply(-0.5 * a3 + 2.5 * a2 - 4 * a + 2, (a > 1) & (a <= 2))
    return inner + outer


def tent(t):
    """Triangle kernel (imresize.py:41-46)."""
    t = np.asarray(t).astype(np.float64)
    neg = np.logical_and(t >= -1, t < 0)
    pos = np.logical_and(t <= 1, t >= 0)
    return np.multiply(t + 1, neg) + np.multiply(1 - t, pos)


def resample_tables(in_len, out_len, scale, kernel, support=4.0):
    """Weight and source-index tables of one axis (imresize.py:60-82).

    Returns ``(weights, indices)`` of shape ``(out_len, K)``: output ``j`` is
    ``sum_k weights[j, k] * src[indices[j, k]]``.  For ``scale < 1`` the kernel
    is stretched to ``scale * kernel(scale * t)`` with width ``support/scale``
    (antialiasing); source taps outside ``[0, in_len)`` are mirrored
    half-sample symmetrically; tap columns that are zero for every output are
    dropped.
    """
    if scale < 1:
        def h(t):
            return scale * kernel(scale * t)
        width = 1.0 * support / scale
    else:
        h = kernel
        width = support
    j = np.arange(1, out_len + 1).astype(np.float64)
    u = j / scale + 0.5 * (1 - 1 / scale)
    left = np.floor(u - width / 2)
    ntap = int(ceil(width)) + 2
    taps = (np.expand_dims(left, axis=1) + np.arange(ntap) - 1).astype(np.int32)
    w = h(np.expand_dims(u, axis=1) - taps - 1)
    w = np.divide(w, np.expand_dims(np.sum(w, axis=1), axis=1))
    mirror = np.concatenate(
        (np.arange(in_len), np.arange(in_len - 1, -1, step=-1))).astype(np.int32)
    taps = mirror[np.mod(taps, mirror.size)]
    live = np.flatnonzero(np.any(w, axis=0))
    return w[:, live], taps[:, live]


def _resample_axis(img, axis, w, idx):
    """Gather-and-weight along one axis in float64 (imresize.py:121-137)."""
    nout, k = w.shape
    if axis == 0:
        g = img[idx].astype(np.float64)
        return np.sum(w.reshape((nout, k, 1, 1)) * g, axis=1)
    if axis == 1:
        g = img[:, idx].astype(np.float64)
        return np.sum(w.reshape((1, nout, k, 1)) * g, axis=2)
    g = img[:, :, idx].astype(np.float64)
    return np.sum(w.reshape((1, 1, nout, k)) * g, axis=3)


def imresize(img, output_shape, method="bicubic"):
    """Separable MATLAB-style resize of a 2-D or 3-D array to ``output_shape``.

    Restates ``imresize(I, output_shape=...)`` (imresize.py:148-191): per-axis
    scale ``out/in``, axes processed in ascending-scale order, float64 out.
    """
    if method == "bicubic":
        kernel = keys_cubic
    elif method == "bilinear":
        kernel = tent
    else:
        raise ValueError("unidentified kernel method supplied")
    shape_in = img.shape
    scales = [1.0 * output_shape[k] / shape_in[k] for k in range(3)]
    order = np.argsort(np.array(scales), kind="stable")
    tables = [resample_tables(shape_in[k], output_shape[k], scales[k], kernel)
              for k in range(3)]
    out = np.copy(img)
    flat = out.ndim == 2
    if flat:
        out = np.expand_dims(out, axis=2)
    for k in range(3):
        ax = int(order[k])
        out = _resample_axis(out, ax, *tables[ax])
    if flat:
        out = np.squeeze(out, axis=2)
    return out


# ----------------------------------------------------------------------------
# interpolation and filters (interp.py:28-45,96-100; calculate.py:30-92)
# ----------------------------------------------------------------------------


def warp_linear(img, coords):
    """Trilinear pull-sample with clamp-to-edge: ``map_coordinates(order=1,
    mode='nearest')`` (interp.py:28-45).  ``coords`` is ``(3, ...)``."""
    return ndi.map_coordinates(img, coords, order=1, mode="nearest", cval=0)


def add_grid(motion, grid):
    """``coords = motion + grid`` in motion's dtype (interp.py:96-100)."""
    out = motion.copy()
    for c, g in enumerate(grid):
        out[..., c] += g
    return out


def warp_by_coords(img, coords_xyz3, dtype=np.float32):
    """``correctMotionGrid`` (calFlow3d_Wei_v1.py:37-66): warp ``img`` cast to
    ``dtype`` at per-voxel coordinates ``(x, y, z, 3)``."""
    src = np.asarray(img, dtype=dtype)
    return warp_linear(src, np.transpose(coords_xyz3, (3, 0, 1, 2)))


def box3(a):
    """3x3x1 mean filter with replicated edges (calFlow3d_Wei_v1.py:480-482)."""
    return ndi.correlate(a, np.ones((3, 3, 1)) / 9, mode="nearest")


def nei_diff(cp_field, dtype=np.float32):
    """Neighbour difference on the control-point field
    (calFlow3d_Wei_v1.py:69-98): mean of the 8 x-y neighbours minus the
    centre, replicated edges, per z and per component."""
    f = np.ones((3, 3, 1, 1), dtype=dtype) / 8
    f[1, 1] = -1
    return ndi.correlate(cp_field, f, mode="nearest")


def det3(xx, xy, xz, yy, yz, zz):
    """calculate.py:67-79."""
    return xx * (yy * zz - yz ** 2) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz)


def det2(a, b, c, d):
    """calculate.py:82-92."""
    return a * d - b * c


def spatial_gradient(img, coords_xyz3):
    """Central differences of the trilinear image at deformed coordinates
    (calFlow3d_Wei_v1.py:134-206).  Only the perturbed coordinate is clipped
    to ``[0, n-1]``; the difference is still divided by 2 at the faces."""
    dims = img.shape
    base = [coords_xyz3[..., a] for a in range(3)]
    grads = []
    for a in range(3):
        hi = list(base)
        lo = list(base)
        hi[a] = np.clip(base[a] + 1.0, 0, dims[a] - 1)
        lo[a] = np.clip(base[a] - 1.0, 0, dims[a] - 1)
        up = warp_linear(img, np.asarray(hi))
        dn = warp_linear(img, np.asarray(lo))
        grads.append((up - dn) / (2 * 1.0))
    return grads


def window_index(centres, r, n):
    """Clamped source indices of each (2r+1) window: ``(len(centres), 2r+1)``."""
    return np.clip(centres[:, None] + np.arange(-r, r + 1)[None, :], 0, n - 1)


def patch_sums(prod, xg, yg, r):
    """Sum of ``prod`` over the (2r+1)x(2r+1)x1 window of every control point,
    clamped indices (= full-grid ``correlate(mode='nearest')`` sampled at the
    control points, calFlow3d_Wei_v1.py:546-588)."""
    ix = window_index(xg, r, prod.shape[0])
    iy = window_index(yg, r, prod.shape[1])
    sx = prod[ix].sum(axis=1)                 # (ncx, y, z)
    return sx[:, iy].sum(axis=2)              # (ncx, ncy, z)


def patch_sums_correlate(prod, xg, yg, r):
    """Reference form: full-grid correlate with a ones((2r+1,2r+1,1)) kernel,
    then sampled at the control points (calFlow3d_Wei_v1.py:546-588)."""
    k = np.ones((2 * r + 1, 2 * r + 1, 1))
    full = ndi.correlate(prod, k, mode="nearest")
    return full[np.ix_(xg, yg, np.arange(prod.shape[2]))]


def solve_patches(sxx, sxy, sxz, syy, syz, szz, sxt, syt, szt, s_pen, nei_sum):
    """Regularised 3x3 adjugate solve per control point
    (calFlow3d_Wei_v1.py:209-262).  No minus sign; NaN -> 0.

    Returns ``(v, n_nan)`` with ``v`` of shape ``(..., 3)``.
    """
    sxx = sxx + s_pen
    syy = syy + s_pen
    szz = szz + s_pen
    sxt = sxt + nei_sum[..., 0]
    syt = syt + nei_sum[..., 1]
    szt = szt + nei_sum[..., 2]
    det = det3(sxx, sxy, sxz, syy, syz, szz)
    m11 = det2(syy, syz, syz, szz)
    m12 = -det2(sxy, syz, sxz, szz)
    m13 = det2(sxy, syy, sxz, syz)
    m22 = det2(sxx, sxz, sxz, szz)
    m23 = -det2(sxx, sxy, sxz, syz)
    m33 = det2(sxx, sxy, sxy, syy)
    with np.errstate(divide="ignore", invalid="ignore"):
        vx = (m11 * sxt + m12 * syt + m13 * szt) / det
        vy = (m12 * sxt + m22 * syt + m23 * szt) / det
        vz = (m13 * sxt + m23 * syt + m33 * szt) / det
    v = np.stack((vx, vy, vz), axis=-1)
    bad = np.isnan(v)
    n_nan = int(bad.sum())
    if n_nan:
        v[bad] = 0
    return v, n_nan


def cp_coords(grid, r, cp_shape):
    """Dense voxel -> control-point-grid coordinates (calFlow3d_Wei_v1.py:
    265-292): ``clamp((x-r)/(2r+1), 0, ncx)``, same for y, z unchanged."""
    gx, gy, gz = grid
    nx = (gx - r) / (2 * r + 1)
    ny = (gy - r) / (2 * r + 1)
    nx = np.minimum(np.maximum(nx, 0.0), cp_shape[0])
    ny = np.minimum(np.maximum(ny, 0.0), cp_shape[1])
    return np.stack([nx, ny, gz], axis=0)


def control_points(nx, ny, r):
    """Control-point centres (calFlow3d_Wei_v1.py:434-436)."""
    return np.arange(r, nx - 1, step=2 * r + 1), np.arange(r, ny - 1, step=2 * r + 1)


def level_shape(shape, level):
    """x and y halve per level (truncating), z is kept (calFlow3d:341-343)."""
    return int(shape[0] / (2 ** level)), int(shape[1] / (2 ** level)), shape[2]


# ----------------------------------------------------------------------------
# the driver: getMotion / correctMotion (calFlow3d_Wei_v1.py:295-706)
# ----------------------------------------------------------------------------


def _stop_rule(k, n_iter, old_err, err, tol):
    """Error-based stop rule before the update (calFlow3d:515-530).
    Mutates ``old_err`` when the level continues."""
    if k == n_iter - 1:
        return True
    if np.sum(old_err <= err) > 1:
        return True
    if np.abs(old_err[-1] - err) < tol:
        return True
    old_err[:-1] = old_err[1:]
    old_err[-1] = err
    return False


def get_motion(dat_mov, dat_ref, option, verbose=False, mode="ref", sums="window",
               schedule=None, extras=None, keep_images=True):
    """Restatement of ``getMotion`` (calFlow3d_Wei_v1.py:295-673).

    ``option`` needs ``layer, iter, r, zRatio, movRange, smoothPenalty, tol,
    save_ite, mask_ref, mask_mov`` and may hold ``motion`` (full-resolution
    ``(x, y, z, 3)`` initial field).  Unlike the reference, ``option`` is not
    modified.

    ``schedule``: optional ``{level: (records, updates)}`` that replaces the
    two data-dependent break tests, so that a run can follow another run's
    iteration schedule exactly.  ``extras``: optional dict filled with the
    per-level schedule (``"schedule"``), the final float64 control-point field
    of each level (``"cp"``) and the NaN counts.

    Returns ``(motion, current_error, coords, error_log)`` like the reference.
    """
    if mode not in ("ref", "fp64"):
        raise ValueError(f"unknown oracle mode {mode!r}")
    ft = np.float32 if mode == "ref" else np.float64
    psum = patch_sums if sums == "window" else patch_sums_correlate

    mask_ref_full = np.asarray(option["mask_ref"], dtype=np.float32)
    mask_mov_full = np.asarray(option["mask_mov"], dtype=np.float32)
    n_levels = option["layer"]
    n_iter = option["iter"]
    r = option["r"]
    z_ratio_raw = option["zRatio"]
    mov_range = option["movRange"]
    size = dat_mov.shape
    log = {}
    sched = {}
    cps = {}
    nans = {}
    motion = None
    current_error = None
    mov_l = None

    for level in range(n_levels, -1, -1):
        entry = {k: [] for k in ("diffError", "penaltyError", "currentError",
                                 "motion_current", "data_mov_corrected",
                                 "max_diff_motion")}
        log[f"layer_{level}"] = entry
        if verbose:
            print(f"starting layer {level} out of {n_levels}")
        shp = level_shape(size, level)
        mov_l = imresize(dat_mov, shp)
        ref_l = imresize(dat_ref, shp)
        if keep_images:
            entry["data_ref"] = ref_l
            entry["data_mov"] = mov_l
        nx, ny, nz = mov_l.shape
        z_ratio = z_ratio_raw / (2 ** level)

        # initial field of this level (calFlow3d:362-410)
        if level == n_levels:
            fresh = np.zeros((nx, ny, nz, 3), dtype=ft)
            init = option.get("motion")
            if init is not None:
                init = np.array(init, dtype=np.float32)
                for c in range(3):
                    fresh[..., c] = imresize(init[..., c], shp) / (size[c] / shp[c])
            motion = fresh
        else:
            prev = np.asarray(motion)
            up = np.zeros((nx, ny, nz, 3), dtype=ft)
            for c in range(3):
                comp = imresize(prev[..., c], shp, method="bilinear")
                up[..., c] = comp * 2 if c < 2 else comp
            motion = np.asarray(up, dtype=ft)

        grid = np.meshgrid(*[np.arange(n, dtype=ft) for n in mov_l.shape],
                           indexing="ij", sparse=False)
        mref = imresize(mask_ref_full, shp) > 0
        mmov = imresize(mask_mov_full, shp)
        old_err = np.inf * np.ones(3)
        s_pen = option["smoothPenalty"] * (r * 2 + 1) ** 2
        xg, yg = control_points(nx, ny, r)
        zg = np.arange(0, nz)
        gx, gy, gz = np.meshgrid(xg, yg, zg, indexing="ij")
        n_rec = 0
        n_upd = 0
        n_nan = 0
        cp_new = None

        for k in range(n_iter):
            old_motion = motion.copy()
            coords = add_grid(motion, grid)
            warped = warp_by_coords(mov_l, coords, dtype=ft)
            if k % option["save_ite"] == 0:
                entry["data_mov_corrected"].append(np.asarray(warped).copy())
                entry["motion_current"].append(np.asarray(motion).copy())
            excl = (warp_by_coords(mmov, coords, dtype=ft) > 0) | mref
            it = box3(ref_l - warped)
            it[excl] = 0

            diff = nei_diff(motion[gx, gy, gz, :], dtype=ft)
            diff[..., 2] = diff[..., 2] * z_ratio
            nei_sum = s_pen * diff
            diff_err = float(np.mean(it ** 2))
            pen_err = float(np.sum(np.sum(diff ** 2, axis=3) * s_pen) / (nx * ny * nz))
            err = diff_err + pen_err
            entry["diffError"].append(diff_err)
            entry["penaltyError"].append(pen_err)
            entry["currentError"].append(err)
            current_error = err
            n_rec += 1
            if verbose:
                print(f"Downsample layer: {level}\tIter: {k}\tError: {err:.3f}, "
                      f"Diff Error: {diff_err:.3f}, Penalty Error: {pen_err:.3f}")

            if schedule is not None:
                rec, upd = schedule[level]
                stop = (k == rec - 1) and (upd == rec - 1)
            else:
                stop = _stop_rule(k, n_iter, old_err, err, option["tol"])
            if stop:
                break

            ix, iy, iz = spatial_gradient(mov_l, coords)
            ix[excl] = 0
            iy[excl] = 0
            iz[excl] = 0
            iz = iz / z_ratio
            terms = [ix ** 2, ix * iy, ix * iz, iy ** 2, iy * iz, iz ** 2,
                     ix * it, iy * it, iz * it]
            s = [psum(t, xg, yg, r) for t in terms]
            v, bad = solve_patches(*s, s_pen, nei_sum)
            n_nan += bad
            if bad:
                print(f"number of nans: {bad}")
            dist = np.sqrt(np.sum(v ** 2, axis=3))
            dist = np.maximum(dist / mov_range, 1.0)
            v = v / dist[..., np.newaxis]
            v[..., 2] = v[..., 2] / z_ratio
            cp_new = motion[gx, gy, gz, :] + v
            cc = cp_coords(grid, r, cp_new.shape)
            for c in range(3):
                motion[..., c] = warp_linear(np.asarray(cp_new[..., c]), cc).reshape(nx, ny, nz)
            n_upd += 1
            max_diff = np.max(np.abs(motion - old_motion))
            entry["max_diff_motion"].append(max_diff)
            if verbose:
                print(f"Downsample layer: {level}\tIter: {k}\tMax motion: "
                      f"{np.max(np.abs(motion)):.2f}\tMax diff. old vs new motion: {max_diff:.4f}")
            if schedule is not None:
                rec, upd = schedule[level]
                if k == rec - 1:
                    break
            elif max_diff < 1e-3:
                break

        sched[level] = (n_rec, n_upd)
        cps[level] = cp_new
        nans[level] = n_nan

    grid = np.meshgrid(*[np.arange(n, dtype=ft) for n in mov_l.shape],
                       indexing="ij", sparse=False)
    coords = add_grid(motion, grid)
    if extras is not None:
        extras["schedule"] = sched
        extras["cp"] = cps
        extras["nans"] = nans
    return np.asarray(motion), current_error, coords, log


def correct_motion(data_raw, motion_field, dtype=np.float32):
    """Restatement of ``correctMotion`` (calFlow3d_Wei_v1.py:676-706)."""
    grid = np.meshgrid(*[np.arange(n, dtype=dtype) for n in data_raw.shape],
                       indexing="ij", sparse=False)
    motion_field = np.asarray(motion_field)
    if dtype == np.float64:
        motion_field = motion_field.astype(np.float64)
    coords = add_grid(motion_field, grid)
    return np.asarray(warp_by_coords(data_raw, coords, dtype=dtype))


def smooth_penalty_factor(dat_ref, mask_ref):
    """``getSmPnltNormFctr`` (preprocess.py:42-66): half the sum of the
    unmasked mean squared central differences along x and y, in the input
    dtype.  ``mask_ref`` True means excluded."""
    mask_ref = np.asarray(mask_ref, dtype=bool)
    gx = (dat_ref[2:, :, :] - dat_ref[:-2, :, :]) / 2
    gy = (dat_ref[:, 2:, :] - dat_ref[:, :-2, :]) / 2
    mx = np.mean(gx[~mask_ref[1:-1, :, :]] ** 2)
    my = np.mean(gy[~mask_ref[:, 1:-1, :]] ** 2)
    return (mx + my) / 2
