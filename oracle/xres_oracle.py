"""numpy/scipy restatement of the cross-resolution solver ``getMotion_v2``
(the oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's checks, never by
the product package.

Reference: ``src/wholistic_registration/utils/calFlowCrossResolution.py``
(cited below by line).  The reference's array namespace falls back to numpy
and its ndimage calls to ``scipy.ndimage`` without CuPy
(``utils/__init__.py``); this restatement makes the same library calls with
the same float32 / float64 cast points, so it reproduces the reference's
outputs bit for bit on the golden cases (tests/test_oracle_xres.py).

Layout: arrays ``(x, y, z)``, fields ``(x, y, z, 3)`` with component ``c``
along axis ``c`` (the reference solver's layout).

Cross-resolution geometry (:1978-1992, :1904-1914): moving voxel ``p`` maps to
reference coordinates ``phase(p) = phase0(p) + motion(p)``, ``phase0 =
(x, y, z * zRatio / zRatio_HR)`` per level; the reference level image is read
there by its cubic B-spline interpolant (``map_coordinates(order=3,
mode="nearest")``, :463-477, :1144-1153) and its gradient by central
differences of the trilinear interpolant (:106-178).
"""

from __future__ import annotations

import numpy as np
from scipy import ndimage as ndi

from . import lk_oracle as lk

BOX3 = np.ones((3, 3, 1), dtype=np.float32) / 9      # the residual's mean filter, float32 weights


# ---------------------------------------------------------------- primitives


def spline_sample(vol, pts):
    """Cubic B-spline interpolant of ``vol`` at points ``(..., 3)``
    (``generate_continuous_H_gpu`` with zRatio 1 + ``apply_H_to_matrix_gpu``,
    :463-477, :1144-1153).  Output dtype = ``vol``'s."""
    shp = pts.shape[:-1]
    flat = pts.reshape(-1, 3).T
    return ndi.map_coordinates(vol, flat, order=3, mode="nearest").reshape(shp)


def exclude_mask(mask, shape):
    """``resample_exclude_mask`` (:1917-1938): bicubic imresize > 0.5."""
    return lk.imresize(np.asarray(mask, dtype=np.float32), shape) > 0.5


def control_grid(nx, ny, nz, r):
    """``make_control_point_grid`` (:1995-2003)."""
    xg = np.arange(r, nx - 1, step=2 * r + 1)
    yg = np.arange(r, ny - 1, step=2 * r + 1)
    gx, gy, gz = np.meshgrid(xg, yg, np.arange(nz), indexing="ij")
    return xg, yg, (gx, gy, gz)


def initial_motion(option, level, n_levels, size, shp, prev):
    """``initialize_motion_for_layer`` (:1941-1975)."""
    out = np.zeros(shp + (3,), dtype=np.float32)
    if level == n_levels:
        init = option.get("motion")
        if init is not None:
            init = np.asarray(init, dtype=np.float32)
            for c in range(3):
                out[..., c] = lk.imresize(init[..., c], shp) / (size[c] / shp[c])
        return out
    for c in range(3):
        up = lk.imresize(prev[..., c], shp, method="bilinear")
        out[..., c] = up * 2 if c < 2 else up
    return out


def initial_phase(option, size, shp, z_ratio, z_ratio_hr):
    """``initialize_phase_for_layer`` (:1978-1992)."""
    ph = option.get("phase")
    if ph is not None:
        ph = np.asarray(ph, dtype=np.float32)
        out = np.zeros(shp + (3,), dtype=np.float32)
        for c in range(3):
            out[..., c] = lk.imresize(ph[..., c], shp) / (size[c] / shp[c])
        return out
    gx, gy, gz = np.indices(shp)
    return np.stack([gx, gy, gz * z_ratio / z_ratio_hr], axis=-1).astype(np.float32)


def valid_on_moving_grid(phase, mask_ref_coef_vol, mask_mov_l):
    """``compute_valid_mask_on_moving_grid`` (:2006-2028): a voxel counts
    unless the moving mask or the reference mask (sampled by its spline at the
    phase, > 0.5) excludes it."""
    ref_ex = spline_sample(mask_ref_coef_vol, phase) > 0.5
    return ~(mask_mov_l > 0) & ~ref_ex


def masked_residual(mov_l, sampled, valid):
    """Residual, 3x3x1 mean with replicated edges, zero outside ``valid``
    (:2117-2122)."""
    res = ndi.correlate(mov_l - sampled, BOX3, mode="nearest")
    res[~valid] = 0
    return res


def cp_nei_diff(motion, cps, z_ratio_hr):
    """Neighbour difference of the field at the control points, z component
    scaled to physical units (:2130-2131)."""
    d = lk.nei_diff(motion[cps[0], cps[1], cps[2], :], dtype=np.float32)
    d[..., 2] *= z_ratio_hr
    return d


def error_terms(res, nei, s_pen):
    """``calError`` (:73-103): mean squared residual, and the penalty summed in
    float32 over the control points divided by the voxel count."""
    x, y, z = res.shape
    diff = float(np.mean(res ** 2))
    pen = float(np.sum(np.sum(nei ** 2, axis=3) * s_pen) / (x * y * z))
    return diff, pen


def ref_gradient(ref_l, phase):
    """``getSpatialGradientInOrgGrid`` (:106-178): central differences of the
    trilinear reference at the phase, the stepped coordinate clipped to the
    reference's extent."""
    return lk.spatial_gradient(ref_l, phase)


def stop_before_update(k, n_iter, old_err, err, tol):
    """The error stop rules (:2146-2157); ``old_err`` is the reference's float32
    history and is shifted when the iteration goes on."""
    if k == n_iter - 1:
        return "last"
    if np.sum(old_err <= err) > 1:
        return "increase"
    if np.abs(old_err[-1] - err) < tol:
        return "flat"
    old_err[:-1] = old_err[1:]
    old_err[-1] = err
    return None


# ---------------------------------------------------------------- one layer


def optimize_layer(mov_l, ref_l, ref_coef, mask_coef, mask_mov_l, phase0, motion_init, cps, r,
                   n_iter, mov_range, s_pen, z_ratio_hr, tol, schedule=None, trace=None):
    """``optimize_layer_cross_resolution`` (:2031-2237).

    ``ref_coef`` / ``mask_coef`` are the volumes the spline interpolants are
    built from (the reference level image, the exclusion mask as float32).
    ``schedule``: optional ``(records, updates)`` replacing the data-dependent
    break tests.  ``trace``: optional list that receives this call's
    ``(records, updates, reason, errors)``.
    """
    x, y, z = mov_l.shape
    motion = motion_init.copy()
    old_err = np.inf * np.ones(3, dtype=np.float32)
    err_first = None
    errs = []
    reason = "end"
    n_upd = 0
    gx, gy, gz = cps
    for k in range(n_iter):
        old_motion = motion.copy()
        phase = phase0 + motion
        sampled = spline_sample(ref_coef, phase)
        valid = valid_on_moving_grid(phase, mask_coef, mask_mov_l)
        res = masked_residual(mov_l, sampled, valid)
        if k == 0:
            err_first = res ** 2
        nei = cp_nei_diff(motion, cps, z_ratio_hr)
        nei_sum = s_pen * nei
        d_err, p_err = error_terms(res, nei, s_pen)
        err = d_err + p_err
        errs.append((d_err, p_err, err))
        if schedule is not None:
            stop = "forced" if (k == schedule[0] - 1 and schedule[1] == schedule[0] - 1) else None
        else:
            stop = stop_before_update(k, n_iter, old_err, err, tol)
        if stop:
            reason = stop
            break
        ix, iy, iz = ref_gradient(ref_l, phase)
        ix[~valid] = 0
        iy[~valid] = 0
        iz[~valid] = 0
        iz /= z_ratio_hr
        terms = [ix ** 2, ix * iy, ix * iz, iy ** 2, iy * iz, iz ** 2, ix * res, iy * res, iz * res]
        sums = [ndi.correlate(t, np.ones((2 * r + 1, 2 * r + 1, 1), dtype=np.float32),
                              mode="nearest")[gx, gy, gz] for t in terms]
        v, _ = lk.solve_patches(*sums, s_pen, nei_sum)
        scale = np.maximum(np.sqrt(np.sum(v ** 2, axis=3)) / mov_range, 1.0)
        v = v / scale[..., np.newaxis]
        v[..., 2] = v[..., 2] / z_ratio_hr
        cp_new = motion[gx, gy, gz, :] + v
        grid = np.meshgrid(*[np.arange(n, dtype=np.float32) for n in mov_l.shape], indexing="ij")
        cc = lk.cp_coords(grid, r, cp_new.shape)
        for c in range(3):
            motion[..., c] = lk.warp_linear(cp_new[..., c], cc).reshape(x, y, z)
        n_upd += 1
        if schedule is not None:
            if k == schedule[0] - 1:
                reason = "forced"
                break
        elif np.max(np.abs(motion - old_motion)) < tol:
            reason = "motion"
            break
    records = len(errs)
    phase = phase0 + motion
    sampled = spline_sample(ref_coef, phase)
    valid = valid_on_moving_grid(phase, mask_coef, mask_mov_l)
    res = masked_residual(mov_l, sampled, valid)
    err_last = res ** 2
    nei = lk.nei_diff(motion[gx, gy, gz, :], dtype=np.float32)     # no z scaling here (:2214)
    d_err, p_err = error_terms(res, nei, s_pen)
    if trace is not None:
        trace.append(dict(records=records, updates=n_upd, reason=reason, errors=errs,
                          final=(d_err, p_err, d_err + p_err)))
    return dict(motion=motion, phase=phase, sampled=sampled, residual=res,
                err_first=err_first if err_first is not None else err_last.copy(),
                err_last=err_last, current_error=d_err + p_err, valid=valid)


# ---------------------------------------------------------------- wrong regions


def cp_local_error(res, cps, k, metric="mse"):
    """``get_local_error_on_control_points`` (:1190-1222): k x k x 1 mean of
    the squared ('mse') or absolute ('mae') residual at the control points
    (float32 weights)."""
    if metric.lower() == "mse":
        dense = res ** 2
    elif metric.lower() == "mae":
        dense = np.abs(res)
    else:
        raise ValueError("metric must be 'mse' or 'mae'")
    kern = np.ones((k, k, 1), dtype=np.float32) / (k ** 2)
    return ndi.correlate(dense, kern, mode="nearest")[cps[0], cps[1], cps[2]]


def robust_significant(err_cp, mad_threshold, large=144, eps=1e-12):
    """Per z plane: |e - median| / (1.4826 MAD + eps) > threshold, over the
    inner control points when the plane holds more than ``large`` of them
    (:1627-1649).  Planes with MAD < eps flag nothing."""
    nx, ny, nz = err_cp.shape
    sig = np.zeros((nx, ny, nz), dtype=bool)
    inner = nx * ny > large and nx > 2 and ny > 2
    sub = err_cp[1:-1, 1:-1, :] if inner else err_cp
    flat = sub.reshape(-1, nz)
    med = np.median(flat, axis=0)
    mad = np.median(np.abs(flat - med[None, :]), axis=0)
    z = np.abs(flat - med[None, :]) / (1.4826 * mad[None, :] + eps)
    s = z > mad_threshold
    s[:, mad < eps] = False
    if inner:
        sig[1:-1, 1:-1, :] = s.reshape(nx - 2, ny - 2, nz)
    else:
        sig = s.reshape(nx, ny, nz)
    return sig


def bad_region_mask(err_cp, r, xg, yg, x_size, y_size, mad_threshold=3.0, min_size=2):
    """``build_bad_region_mask_from_cp_error`` (:1570-1787): significant
    control points, 4-connected components within each z plane, components of
    at least ``min_size`` points; each kept component's control-point
    bounding box widened by ceil(1.5 r) voxels marks the dense mask.
    Returns ``(mask, n_regions)``."""
    nz = err_cp.shape[2]
    out = np.zeros((x_size, y_size, nz), dtype=bool)
    if err_cp.size == 0:
        return out, 0
    sig = robust_significant(err_cp, mad_threshold)
    if not sig.any():
        return out, 0
    st = np.zeros((3, 3, 3), dtype=bool)
    st[:, :, 1] = [[0, 1, 0], [1, 1, 1], [0, 1, 0]]
    lab, n = ndi.label(sig, structure=st)
    if n == 0:
        return out, 0
    sizes = np.bincount(lab.ravel(), minlength=n + 1)
    dx = int(np.ceil(1.5 * r))
    kept = 0
    for l in range(1, n + 1):
        if sizes[l] < min_size:
            continue
        kept += 1
        ii, jj, kk = np.nonzero(lab == l)
        x0 = max(0, int(xg[ii.min()]) - dx)
        x1 = min(x_size - 1, int(xg[ii.max()]) + dx)
        y0 = max(0, int(yg[jj.min()]) - dx)
        y1 = min(y_size - 1, int(yg[jj.max()]) + dx)
        out[x0:x1 + 1, y0:y1 + 1, kk.min():kk.max() + 1] = True
    return out, kept


def quantile_threshold(vals, q):
    """``get_quantile_threshold`` (:1790-1799): the floor((n-1) q)-th smallest."""
    flat = np.asarray(vals).ravel()
    if flat.size == 0:
        return np.float32(0)
    k = int(np.clip(np.floor((flat.size - 1) * q), 0, flat.size - 1))
    return np.partition(flat, k)[k]


MAX_BALL_RADIUS_Z = 64


def make_ball(radius_xy, z_ratio):
    """``_make_ball`` (:1363-1396): voxels within ``radius_xy`` physical units,
    z index spacing ``z_ratio`` physical units, z reach capped at 64."""
    radius_xy = int(max(1, radius_xy))
    z_ratio = float(z_ratio)
    if not np.isfinite(z_ratio) or z_ratio <= 0:
        raise ValueError("make_ball: z_ratio must be a positive, finite physical-units-per-z-index ratio")
    rz = int(max(1, min(np.ceil(radius_xy / z_ratio), MAX_BALL_RADIUS_Z)))
    a = np.arange(-radius_xy, radius_xy + 1, dtype=np.float32)
    c = np.arange(-rz, rz + 1, dtype=np.float32)
    X, Y, Z = np.meshgrid(a, a, c, indexing="ij")
    return (X ** 2 + Y ** 2 + (Z * z_ratio) ** 2) <= radius_xy ** 2 + 1e-6


def trap_mask(bad, phase, ref_l, z_ratio_ref, expand=2.0, sigma=1.0, k_int=2.5):
    """``build_reference_trap_mask_from_bad_moving_fast_roi`` (:1413-1567): the
    reference voxels the bad moving voxels map to (rounded phase), closed,
    then grown inside a box around them through reference voxels of similar
    smoothed intensity (median +- k 1.4826 MAD of the seeds)."""
    xr, yr, zr = ref_l.shape
    trap = np.zeros((xr, yr, zr), dtype=bool)
    if not bad.any():
        return trap
    seeds = np.rint(phase[bad]).astype(np.int32)
    sx, sy, sz = seeds[:, 0], seeds[:, 1], seeds[:, 2]
    keep = (sx >= 0) & (sx < xr) & (sy >= 0) & (sy < yr) & (sz >= 0) & (sz < zr)
    sx, sy, sz = sx[keep], sy[keep], sz[keep]
    if sx.size == 0:
        return trap
    mxy = int(max(1, expand + 2))
    mz = int(max(1, np.ceil((expand + 2) / max(float(z_ratio_ref), 1e-6))))
    x0, x1 = max(0, sx.min() - mxy), min(xr, sx.max() + mxy + 1)
    y0, y1 = max(0, sy.min() - mxy), min(yr, sy.max() + mxy + 1)
    z0, z1 = max(0, sz.min() - mz), min(zr, sz.max() + mz + 1)
    seed = np.zeros((x1 - x0, y1 - y0, z1 - z0), dtype=bool)
    seed[sx - x0, sy - y0, sz - z0] = True
    seed = ndi.binary_closing(seed, structure=make_ball(1, z_ratio_ref))
    if not seed.any():
        return trap
    crop = ref_l[x0:x1, y0:y1, z0:z1].astype(np.float32)
    sm = ndi.gaussian_filter(crop, sigma=(sigma, sigma, sigma)) if sigma > 0 else crop
    vals = sm[seed]
    if vals.size == 0:
        return trap
    med = float(np.median(vals))
    mad = float(np.median(np.abs(vals - med)))
    tol = max(k_int * 1.4826 * mad, 1e-3)
    near = ndi.binary_dilation(seed, structure=make_ball(max(1, int(round(expand))), z_ratio_ref), iterations=1)
    allowed = (near & (np.abs(sm - med) <= tol)) | seed
    grow_st = make_ball(1, z_ratio_ref)
    region = seed.copy()
    for _ in range(int(expand)):
        g = ndi.binary_dilation(region, structure=grow_st) & allowed
        if np.array_equal(g, region):
            break
        region = g
    trap[x0:x1, y0:y1, z0:z1] = region | seed
    return trap


def correct_wrong_regions(mov_l, ref_l, ref_coef, mask_coef, mask_mov_l, mask_ref_l, phase0, motion_init,
                          xg, yg, cps, r, n_iter, mov_range, s_pen, z_ratio_hr, tol, level, opts,
                          schedules=None, trace=None, info=None):
    """``correct_wrong_regions_one_layer`` (:2240-2467): plain optimisation; if
    bad regions are found, rerun with their reference trap mask excluded, then
    again from that field with the original mask; keep the rerun only when
    its final objective is lower."""
    def run(mcoef, m0):
        return optimize_layer(mov_l, ref_l, ref_coef, mcoef, mask_mov_l, phase0, m0, cps, r, n_iter, mov_range,
                              s_pen, z_ratio_hr, tol, schedule=next(schedules) if schedules else None,
                              trace=trace)

    res0 = run(mask_coef, motion_init)
    err_cp = cp_local_error(res0["residual"], cps, 2 * r + 1, opts["metric"])
    bad, n_reg = bad_region_mask(err_cp, r, xg, yg, mov_l.shape[0], mov_l.shape[1],
                                 opts["mad_threshold"], opts["min_component_size"])
    if info is not None:
        info.update(regions=n_reg, ratio=None, accepted=False)
    if n_reg == 0:
        return res0
    if opts["exclude_mode"] == "highresidual":
        dense = ndi.correlate(res0["err_first"] - res0["err_last"], BOX3, mode="nearest")
        dense[~bad] = 0
        th = quantile_threshold(dense[bad], opts["quantile_q"]) if bad.any() else 0
        bad = bad & (dense >= th)
    bad = bad & res0["valid"]
    trap = trap_mask(bad, res0["phase"], ref_l, z_ratio_hr, opts["expand_radius_xy"] / (2.0 ** level),
                     opts["sigma_grad"], opts["intensity_k"])
    mcoef1 = (mask_ref_l | trap).astype(np.float32)
    n_old = int(res0["valid"].sum())
    n_new = int(valid_on_moving_grid(res0["phase"], mcoef1, mask_mov_l).sum())
    ratio = n_new / max(n_old, 1)
    if info is not None:
        info.update(ratio=ratio, trap=int(trap.sum()), bad=int(bad.sum()))
    if ratio < 0.4:
        return res0
    res1 = run(mcoef1, motion_init)
    res2 = run(mask_coef, res1["motion"])
    if res2["current_error"] < res0["current_error"]:
        if info is not None:
            info["accepted"] = True
        return res2
    return res0


# ---------------------------------------------------------------- driver


def options(option):
    """The option keys ``getMotion_v2`` reads, with its defaults (:2501-2524)."""
    return dict(
        layer=int(option["layer"]), iter=int(option["iter"]), r=int(option["r"]),
        zRatio=float(option.get("zRatio", 1.0)), zRatio_HR=float(option["zRatio_HR"]),
        smoothPenalty=float(option["smoothPenalty"]), movRange=float(option.get("movRange", 5.0)),
        tol=float(option.get("tol", 1e-3)), wrong=bool(option.get("wrong_region_enable", True)),
        metric=option.get("wrong_region_metric", "mse"),
        mad_threshold=float(option.get("wrong_region_mad_threshold", 3.0)),
        min_component_size=int(option.get("wrong_region_min_component_size", 2)),
        exclude_mode=option.get("wrong_region_exclude_mode", "direct"),
        expand_radius_xy=float(option.get("expand_radius_xy", 2.0)),
        sigma_grad=float(option.get("sigma_grad", 1.0)), intensity_k=float(option.get("intensity_k", 2.0)),
        quantile_q=float(option.get("quantile_threshold_q", 0.8)),
    )


def get_motion_v2(data_mov, data_ref, option, schedules=None, trace=None, info=None):
    """Restatement of ``getMotion_v2`` (:2470-2663).

    Returns ``(phase, motion, sampled)`` like the reference.  ``option`` is
    not modified.  ``schedules``: optional list, in call order, of
    ``(records, updates)`` per layer optimisation, forcing the iteration
    schedule.  ``trace`` / ``info``: optional list / dict filled with every
    optimisation's schedule and errors and each layer's wrong-region outcome.
    """
    o = options(option)
    mask_ref = np.asarray(option["mask_ref"], dtype=np.float32)
    mask_mov = np.asarray(option["mask_mov"], dtype=np.float32)
    mov = np.asarray(data_mov, dtype=np.float32)
    ref = np.asarray(data_ref, dtype=np.float32)
    size, size_hr = mov.shape, ref.shape
    r = o["r"]
    s_pen = o["smoothPenalty"] * (2 * r + 1) ** 2
    sched = iter(schedules) if schedules is not None else None
    motion = phase = sampled = None
    for level in range(o["layer"], -1, -1):
        shp = (int(size[0] / 2 ** level), int(size[1] / 2 ** level), size[2])
        shp_hr = (int(size_hr[0] / 2 ** level), int(size_hr[1] / 2 ** level), size_hr[2])
        mov_l = lk.imresize(mov, shp)
        ref_l = lk.imresize(ref, shp_hr)
        mmov_l = exclude_mask(mask_mov, shp)
        mref_l = exclude_mask(mask_ref, shp_hr)
        z_ratio = o["zRatio"] / 2 ** level
        z_ratio_hr = o["zRatio_HR"] / 2 ** level
        m0 = initial_motion(option, level, o["layer"], size, shp, motion)
        ph0 = initial_phase(option, size, shp, z_ratio, z_ratio_hr)
        xg, yg, cps = control_grid(*shp, r)
        mcoef = mref_l.astype(np.float32)
        lvl_info = {} if info is not None else None
        if o["wrong"]:
            res = correct_wrong_regions(mov_l, ref_l, ref_l, mcoef, mmov_l, mref_l, ph0, m0, xg, yg, cps, r,
                                        o["iter"], o["movRange"], s_pen, z_ratio_hr, o["tol"], level, o,
                                        schedules=sched, trace=trace, info=lvl_info)
        else:
            res = optimize_layer(mov_l, ref_l, ref_l, mcoef, mmov_l, ph0, m0, cps, r, o["iter"], o["movRange"],
                                 s_pen, z_ratio_hr, o["tol"], schedule=next(sched) if sched else None,
                                 trace=trace)
        if info is not None:
            info[level] = lvl_info
        motion, phase, sampled = res["motion"], res["phase"], res["sampled"]
    return np.asarray(phase), np.asarray(motion), np.asarray(sampled)


# ---------------------------------------------------------------- z-offset search


def grad_mag_2d(img):
    """``calculate.grad_mag_2d`` (calculate.py:109-112): Sobel magnitude with
    replicated edges, float32."""
    gx = ndi.sobel(img, axis=1, mode="nearest")
    gy = ndi.sobel(img, axis=0, mode="nearest")
    return np.sqrt(gx * gx + gy * gy).astype(np.float32)


def hann2d(h, w):
    """``calculate.hann2d`` (calculate.py:115-121)."""
    wy = np.hanning(h) if h > 1 else np.ones(1, dtype=np.float32)
    wx = np.hanning(w) if w > 1 else np.ones(1, dtype=np.float32)
    return np.maximum(np.outer(wy, wx).astype(np.float32), 1e-3)


def zncc_scores(mov_feat, ref_feat, use_hann_weight=True, eps=1e-6):
    """``_global_zncc_scores`` (:332-367): (K, Z) weighted ZNCC of every moving
    slice with every reference slice, clipped to [-1, 1]."""
    h, w, k_n = mov_feat.shape
    mov_feat = mov_feat.astype(np.float32, copy=False)
    ref_feat = ref_feat.astype(np.float32, copy=False)
    weight = hann2d(h, w).astype(np.float32) if use_hann_weight else np.ones((h, w), dtype=np.float32)
    weight = weight / (np.sum(weight) + eps)
    scores = np.zeros((k_n, ref_feat.shape[2]), dtype=np.float32)
    ref_mean = np.sum(weight[:, :, None] * ref_feat, axis=(0, 1))
    ref_c = ref_feat - ref_mean[None, None, :]
    ref_var = np.sum(weight[:, :, None] * ref_c * ref_c, axis=(0, 1))
    for k in range(k_n):
        m = mov_feat[:, :, k]
        mm = np.sum(weight * m)
        mc = m - mm
        mv = np.sum(weight * mc * mc)
        num = np.sum(weight[:, :, None] * mc[:, :, None] * ref_c, axis=(0, 1))
        scores[k, :] = num / (np.sqrt(mv * ref_var) + eps)
    return np.clip(scores, -1.0, 1.0)


def find_init_z(data_mov, data_ref, delta_ref_idx, use_gradient=True, use_hann_weight=True, direction=1,
                weight_eps=1e-6, return_debug=False):
    """``FindInitZ_stack_global_fixed_spacing`` (:370-460): the reference z of
    each moving slice, slices a fixed spacing apart, chosen by the summed
    ZNCC of their gradient magnitudes."""
    mov = np.asarray(data_mov, dtype=np.float32)
    ref = np.asarray(data_ref, dtype=np.float32)
    if mov.ndim != 3 or ref.ndim != 3:
        raise ValueError("data_ref must be 3D (H,W,Z)")
    h, w, k_n = mov.shape
    if ref.shape[:2] != (h, w):
        raise ValueError(f"XY size mismatch: data_mov has {(h, w)}, data_ref has {ref.shape[:2]}")
    z_n = ref.shape[2]
    if use_gradient:
        mf = np.stack([grad_mag_2d(mov[:, :, k]) for k in range(k_n)], axis=2)
        rf = np.stack([grad_mag_2d(ref[:, :, z]) for z in range(z_n)], axis=2)
    else:
        mf, rf = mov, ref
    scores = zncc_scores(mf, rf, use_hann_weight, weight_eps)
    offsets = np.arange(k_n, dtype=np.float32) * float(delta_ref_idx) * float(direction)
    z0_min = int(np.ceil(-np.min(offsets)).item())
    z0_max = int(np.floor((z_n - 1) - np.max(offsets)).item())
    if z0_max < z0_min:
        raise ValueError(f"No valid z0 exists. Z={z_n}, K={k_n}, delta_ref_idx={delta_ref_idx}. "
                         f"Need reference z range >= {(k_n - 1) * abs(delta_ref_idx) + 1}.")
    best, best_init, best_z0 = -np.inf, None, None
    for z0 in range(z0_min, z0_max + 1):
        z_init = z0 + offsets
        idx = np.rint(z_init).astype(np.int32)
        sc = np.sum(scores[np.arange(k_n), idx])
        if sc > best:
            best, best_init, best_z0 = sc, z_init.copy(), z0
    out = best_init.astype(np.uint16)
    if return_debug:
        return out, {"scores": scores, "best_score": float(best), "best_z0": int(best_z0),
                     "delta_ref_idx": float(delta_ref_idx), "direction": int(direction)}
    return out
