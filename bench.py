#!/usr/bin/env python
"""Registered voxels/s of one full registration (getMotion + correctMotion)
on one MI355X.

Workload (default): a uint16 pair of 96 x 1024 x 2048 (z, y, x) voxels,
400k Gaussian blobs, a smooth 8-mode displacement field with max |d| = 8
voxels, a random ellipsoid mask covering 10 % of the voxels as mask_ref;
layer 4 (five scales), 33x33x1 patches (r = 16), 20 iterations per level,
float32 throughout.  One step = getMotion + correctMotion of the moving
volume with the inputs already resident in HBM.  Data are synthetic,
generated on the device from fixed seeds.

Prints ONE JSON line (rank 0).  The roofline object is for the dominant
kernel (the fused LK pass at level 0): algorithmic bytes per launch
((2 * 4 + 1) B per level-0 voxel: moving and template float32 level images
plus the mask byte) divided by its average duration, measured with HIP
events recorded on the launch stream inside the timed steps.  cpu_baseline
times the numpy/scipy oracle (the reference's CPU path restated) on a crop of
the same pair, one thread.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (z, y, x), blobs, modes, max|d|, mask fraction, layer, r, iter, seed
    "plain_96x1024x2048_u16": dict(zyx=(96, 1024, 2048), n_blobs=400_000, n_modes=8, dmax=8.0,
                                   mask_fraction=0.10, layer=4, r=16, iter=20, seed=1, u16=True),
    "plain_64x128x128_f32": dict(zyx=(64, 128, 128), n_blobs=2000, n_modes=4, dmax=3.0,
                                 mask_fraction=0.0, layer=3, r=8, iter=10, seed=0, u16=False),
}
DEFAULT = "plain_96x1024x2048_u16"
HBM_PEAK_GBS = 8000.0
RAW_PENALTY = 0.08


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_pair_device(w, dev):
    import torch

    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import synth
    rng = np.random.default_rng(w["seed"])
    zyx = w["zyx"]
    blobs = synth.draw_blobs(rng, zyx, w["n_blobs"])
    modes = synth.draw_modes(rng, w["n_modes"])
    clean = synth.render_template_torch(zyx, blobs, dev)
    d = synth.field_torch(zyx, modes, w["dmax"], dev)
    warped = D.correct_motion(clean, (-d).contiguous())
    g = torch.Generator(device=dev)
    g.manual_seed(w["seed"])
    tmpl = torch.poisson(clean, generator=g)
    mov = torch.poisson(warped.clamp_min(0), generator=g)
    del clean, warped
    if w["u16"]:
        tmpl = tmpl.round_().clamp_(0, 65535).to(torch.uint16)
        mov = mov.round_().clamp_(0, 65535).to(torch.uint16)
    mask = None
    if w["mask_fraction"] > 0:
        mask = synth.ellipsoid_mask_torch(rng, zyx, w["mask_fraction"], dev).float().contiguous()
    torch.cuda.synchronize()
    return tmpl.contiguous(), mov.contiguous(), mask, d


def smooth_penalty(tmpl, mask):
    """getSmPnltNormFctr (preprocess.py:42-66) of the template on the device
    (wr_smooth_penalty_factor), after the harness's uint16 -> float32
    conversion."""
    from wholistic_registration_amd import device as D
    return D.smooth_penalty_factor(tmpl.float().contiguous(), mask)


def measured_traffic(workload):
    """HBM bytes per launch of the level-0 LK pass from the committed counter
    runs (FETCH_SIZE and WRITE_SIZE, scaled by a copy of known size), or None
    when the workload has no such record."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(workload)
    except (OSError, ValueError):
        return None


def algorithmic_bytes(w, records, s_in, s_p=4, s_m=1):
    """Bytes per registration: LK passes, pyramid, level-0 conversion, final
    warp with motion write."""
    nz, ny, nx = w["zyx"]
    n0 = nx * ny * nz
    tot = 0.0
    for L in range(w["layer"] + 1):
        nl = int(nx / 2 ** L) * int(ny / 2 ** L) * nz
        tot += nl * records[L] * (2 * s_p + s_m)
        if L >= 1:
            tot += 2 * s_in * n0 + nl * 2 * s_p
    if s_p != s_in:
        tot += 2 * (s_in + s_p) * n0
    tot += (s_in + 4 + 12) * n0
    return tot


def _cpu_iteration_seconds(args):
    """Seconds of one level-0 iteration (error record, gradients, full-grid
    window sums, solve, update) of the reference's CPU path on a crop: the
    oracle in the reference's own form (scipy.ndimage.correlate with a
    ones((2r+1, 2r+1, 1)) kernel and map_coordinates, its dtype flow) run with
    the schedule of one record per level, then with one more level-0
    iteration; the difference isolates that iteration.  Runs in a worker
    process of the all-cores leg as well."""
    mov, ref, opt, layer = args
    from oracle import lk_oracle as O
    one = {L: (1, 0) for L in range(layer + 1)}
    two = dict(one)
    two[0] = (2, 1)
    t = []
    for sched in (one, two):
        t0 = time.perf_counter()
        O.get_motion(mov, ref, dict(opt), sums="correlate", schedule=sched, keep_images=False)
        t.append(time.perf_counter() - t0)
    return t[1] - t[0], t[0] + t[1]


def cpu_baseline_measure(tmpl, mov, mask, w, records, updates, crop=(2, 320, 640), workers=0):
    """The reference's CPU path on the host cores, extrapolated.  A full
    registration of the workload takes hours on a CPU (at r = 16 every
    iteration correlates 9 full-resolution products with a 33 x 33 kernel), so
    one level-0 iteration is timed on a z-slab crop of the same pair
    (_cpu_iteration_seconds) and the registration time is extrapolated as
    seconds per voxel-iteration x sum over levels of (updates executed x level
    voxels), with the schedule the GPU run executed.  Error-only records, the
    pyramid and the final warp are left out, so the estimate favours the CPU.
    One thread (scipy.ndimage and numpy elementwise work are single-threaded),
    then `workers` processes at once (the all-cores leg: independent
    registrations, aggregate throughput)."""
    import multiprocessing as mp
    import platform

    from oracle import lk_oracle as O
    nz, ny, nx = w["zyx"]
    cz, cy, cx = crop
    z0, y0, x0 = (nz - cz) // 2, (ny - cy) // 2, (nx - cx) // 2
    sl = (slice(z0, z0 + cz), slice(y0, y0 + cy), slice(x0, x0 + cx))
    ref = np.ascontiguousarray(tmpl[sl].float().cpu().numpy().transpose(2, 1, 0))
    mv = np.ascontiguousarray(mov[sl].float().cpu().numpy().transpose(2, 1, 0))
    mk = (mask[sl].cpu().numpy() > 0).transpose(2, 1, 0) if mask is not None else np.zeros(ref.shape, bool)
    sp = O.smooth_penalty_factor(ref, mk) * RAW_PENALTY
    opt = dict(layer=w["layer"], iter=w["iter"], r=w["r"], zRatio=1.0, movRange=5.0, tol=1e-3,
               save_ite=10 ** 9, mask_ref=mk, mask_mov=np.zeros(ref.shape, bool), smoothPenalty=sp)
    n0_crop = cx * cy * cz
    vox_iters = sum(updates[L] * int(nx / 2 ** L) * int(ny / 2 ** L) * nz for L in range(w["layer"] + 1))
    job = (mv, ref, opt, w["layer"])
    t_iter, t_all = _cpu_iteration_seconds(job)
    per_vox_iter = t_iter / n0_crop
    t_reg = per_vox_iter * vox_iters
    n0 = nx * ny * nz
    try:
        model = next((ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name")),
                     platform.processor())
    except OSError:
        model = platform.processor()
    nproc = os.cpu_count() or 1
    out = {"value": n0 / t_reg, "unit": "registered voxels/s", "cores": 1, "kind": "port",
           "extrapolated": True, "seconds_per_registration": t_reg, "seconds_per_voxel_iteration": per_vox_iter,
           "cpu_model": model, "nproc": nproc,
           "sample": f"reference CPU path (the oracle in the reference's form: full-grid scipy correlate window "
                     f"sums, map_coordinates warps, its dtype flow), one level-0 iteration isolated on the centre "
                     f"crop {cz}x{cy}x{cx} (z,y,x) of the same pair ({t_all:.1f} s of CPU work), extrapolated "
                     f"to {vox_iters / 1e6:.0f}M voxel-iterations (the GPU run's updates per level "
                     f"{dict((f'L{L}', updates[L]) for L in range(w['layer'], -1, -1))}); one thread"}
    workers = workers or min(nproc, 16)
    if workers > 1:
        ctx = mp.get_context("spawn")   # fresh interpreters: the parent holds a GPU context
        t0 = time.perf_counter()
        with ctx.Pool(workers) as pool:
            res = pool.map(_cpu_iteration_seconds, [job] * workers)
        wall = time.perf_counter() - t0
        agg = sum(n0 / (ti / n0_crop * vox_iters) for ti, _ in res)
        out["all_cores"] = {"value": agg, "unit": "registered voxels/s", "cores": workers, "processes": workers,
                            "extrapolated": True, "wall_s": wall,
                            "sample": f"{workers} independent processes running the same timed sample at once "
                                      f"(aggregate of independent registrations)"}
    return out


def crop_checks(tmpl, mov, mask, w, crop, want_cpu, want_epe, precisions=()):
    """On the centre crop of the pair: time the oracle in the reference's
    dtype flow with direct per-patch window sums (one thread), and compare
    the GPU field with the all-float64 oracle on the same crop (the EPE <= 0.1
    voxel gate)."""
    import torch

    from oracle import lk_oracle as O
    from oracle.metrics import epe_stats
    from wholistic_registration_amd import device as D
    cz, cy, cx = crop
    nz, ny, nx = w["zyx"]
    z0, y0, x0 = (nz - cz) // 2, (ny - cy) // 2, (nx - cx) // 2
    sl = (slice(z0, z0 + cz), slice(y0, y0 + cy), slice(x0, x0 + cx))
    ref = tmpl[sl].float().cpu().numpy().transpose(2, 1, 0)
    mv = mov[sl].float().cpu().numpy().transpose(2, 1, 0)
    mk = (mask[sl].cpu().numpy() > 0).transpose(2, 1, 0) if mask is not None else \
        np.zeros(ref.shape, bool)
    sp = O.smooth_penalty_factor(ref, mk) * RAW_PENALTY
    opt = dict(layer=w["layer"], iter=w["iter"], r=w["r"], zRatio=1.0, movRange=5.0, tol=1e-3,
               save_ite=10 ** 9, mask_ref=mk, mask_mov=np.zeros(ref.shape, bool), smoothPenalty=sp)
    out = {}
    nvox = cx * cy * cz
    if want_cpu:
        t0 = time.perf_counter()
        m, _, _, _ = O.get_motion(mv, ref, opt, keep_images=False)
        O.correct_motion(mv, m)
        dt = time.perf_counter() - t0
        out["cpu_oracle_window_sums"] = {
            "value": nvox / dt, "unit": "registered voxels/s", "cores": 1, "kind": "port",
            "sample": f"oracle with direct per-patch window sums instead of the reference's full-grid "
                      f"correlates (faster than the reference's form) on the centre crop {cz}x{cy}x{cx} "
                      f"(z,y,x) of the same pair, same parameters, one thread, {dt:.1f} s"}
    if want_epe:
        m64, _, _, _ = O.get_motion(mv, ref, opt, mode="fp64", keep_images=False)
        dev = tmpl.device
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).to(dev)
        mref_t = t(mk.astype(np.float32)) if mk.any() else None
        for prec in ("fp32",) + tuple(precisions):
            res = D.get_motion(t(mv), t(ref), mask_ref=mref_t, layer=w["layer"], iter=w["iter"],
                               r=w["r"], zRatio=1.0, movRange=5.0, smoothPenalty=sp, tol=1e-3,
                               precision=prec)
            got = res.motion.cpu().numpy().transpose(2, 1, 0, 3)
            key = "epe_vs_fp64" if prec == "fp32" else f"epe_vs_fp64_{prec}"
            out[key] = dict(epe_stats(got, m64),
                            sample=f"GPU {prec} level images vs all-float64 oracle on the centre "
                                   f"crop {cz}x{cy}x{cx} (z,y,x), both free-running")
    return out


def precision_measure(w, n0, step, motion_fp32, field, mask, steps, prec):
    """The getMotion + correctMotion step with bfloat16 / float16 level
    images: throughput, the level-0 LK pass roofline (2-byte template and
    moving image plus the mask byte per voxel), and the endpoint error of its
    field against the float32 field and the generating field."""
    import torch

    from wholistic_registration_amd import device as D
    res = step(False, prec)          # warm-up (and workspace) for this storage
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = step(False, prec)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    lk_ms = lk_n = 0.0
    for _ in range(steps):           # with HIP events around the LK passes (not timed)
        step(True, prec)
        st = D.stats()
        lk_ms += st.lk_ms[0]
        lk_n += st.lk_exec[0]
    motion = res.motion
    e = torch.sqrt(((motion - motion_fp32) ** 2).sum(-1))
    keep = mask <= 0 if mask is not None else torch.ones_like(e, dtype=torch.bool)
    ek = e[keep].float()
    bytes_per_launch = (2 * 2 + (1 if mask is not None else 0)) * n0
    avg_ms = lk_ms / max(lk_n, 1)
    roof = {"kernel": "k_lk_pass (level 0)", "achieved": bytes_per_launch / (avg_ms * 1e-3) / 1e9,
            "unit": "GB/s", "frac": bytes_per_launch / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "avg_launch_ms": avg_ms, "bytes_per_launch": bytes_per_launch}
    name = next(k for k, v in WORKLOADS.items() if v is w)
    counted = (measured_traffic(name) or {}).get("bf16") if prec == "bf16" else None
    if counted:
        roof.update(traffic=counted["bytes_per_launch"], ta_busy=counted["ta_busy"],
                    traffic_note="records of 8 B per voxel (the image and its central differences)")
    return {
        "value": n0 / dt, "ms_per_step": 1e3 * dt,
        "records_per_level": {f"L{L}": res.records[L] for L in range(w["layer"], -1, -1)},
        "roofline": roof,
        "epe_vs_fp32_field": {"epe_mean": float(ek.mean()), "epe_max": float(ek.max()),
                              "sample": "full volume outside the mask"},
        "epe_vs_truth": truth_epe(motion, field, mask),
    }


def batch_measure(n_frames, worker_counts, seed=2, batched=((8, 1), (16, 1), (32, 1), (16, 2))):
    """Frames of a time series against one template (the time-series frame
    geometry): a 64 x 512 x 1024 uint16 template with 100k blobs, frames
    pull-warped by their own 6-mode fields (max |d| = 6) with fresh Poisson
    noise; layer 4, r = 8, 15 iterations, no mask.  Registered voxels/s with
    one context (frames in sequence), with several concurrent contexts
    (FrameBatch.register), and with the batched launch over chunks of frames
    (FrameBatch.register_batched, `batched` = (chunk, contexts) pairs), on the
    same frames."""
    import torch

    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import synth
    from wholistic_registration_amd.batch import FrameBatch
    dev = torch.device("cuda", 0)
    zyx = (64, 512, 1024)
    rng = np.random.default_rng(seed)
    clean = synth.render_template_torch(zyx, synth.draw_blobs(rng, zyx, 100_000), dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    tmpl = torch.poisson(clean, generator=g).round_().clamp_(0, 65535).to(torch.uint16).contiguous()
    frames = []
    for i in range(n_frames):
        d = synth.field_torch(zyx, synth.draw_modes(np.random.default_rng(1000 + i), 6), 6.0, dev)
        warped = D.correct_motion(clean, (-d).contiguous())
        frames.append(torch.poisson(warped.clamp_min(0), generator=g).round_().clamp_(0, 65535)
                      .to(torch.uint16).contiguous())
        del d, warped
    sp = float(D.smooth_penalty_factor(tmpl.float().contiguous())) * RAW_PENALTY
    params = dict(layer=4, iter=15, r=8, zRatio=1.0, movRange=5.0, smoothPenalty=sp, tol=1e-3)
    n_vox = n_frames * zyx[0] * zyx[1] * zyx[2]
    out = {"workload": f"{n_frames} frames of {zyx[0]}x{zyx[1]}x{zyx[2]} (z,y,x) uint16 against one "
                       f"template, layer 4, r 8, 15 iterations, getMotion + correctMotion per frame",
           "frames": n_frames}
    warped_out = torch.empty((n_frames,) + zyx, dtype=torch.float32, device=dev)
    for wk in worker_counts:
        with FrameBatch(workers=wk) as fb:
            fb.register(frames[: min(n_frames, 2 * wk)], tmpl, **params)   # workspaces, warm caches
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fb.register(frames, tmpl, warped_out=warped_out, **params)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        out[f"workers_{wk}"] = {"value": n_vox / dt, "ms_per_frame": 1e3 * dt / n_frames}
    # the batched launch: every kernel of a level-iteration one launch over a chunk of frames
    stack = torch.stack(frames).contiguous()
    del frames
    for ch, wk in batched:
        with FrameBatch(workers=wk) as fb:
            fb.register_batched(stack[: min(n_frames, ch * wk)], tmpl, chunk=ch, **params)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fb.register_batched(stack, tmpl, chunk=ch, warped_out=warped_out, **params)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        out[f"batched_chunk{ch}_workers_{wk}"] = {"value": n_vox / dt, "ms_per_frame": 1e3 * dt / n_frames}
    out["unit"] = "registered voxels/s"
    return out


def splat_measure(reps=3):
    """The cross-resolution path's fixed-plane projections on one
    64 x 512 x 1024 frame of warped coordinates (the voxel grid plus a smooth
    6-mode field, max |d| = 6) onto 8 reference planes: max splat with the
    2x2 footprint, Gaussian-weighted splat.  Kernel time per call with events
    on the launch stream; algorithmic bytes = the samples read (16 B each)
    plus the output planes written."""
    import torch

    from wholistic_registration_amd import calFlowCrossResolution as X
    from wholistic_registration_amd import synth
    dev = torch.device("cuda", 0)
    zyx = (64, 512, 1024)
    d = synth.field_torch(zyx, synth.draw_modes(np.random.default_rng(5), 6), 6.0, dev)
    z, y, x = torch.meshgrid(*[torch.arange(n, device=dev, dtype=torch.float32) for n in zyx], indexing="ij")
    coords = torch.stack([x + d[..., 0], y + d[..., 1], z + d[..., 2]], dim=-1).permute(2, 1, 0, 3).contiguous()
    values = (x * 0.01 + y * 0.02).permute(2, 1, 0).contiguous()
    ref = torch.empty(zyx[::-1], device=dev)          # (x, y, z) shape only
    planes = torch.linspace(4, 60, 8, device=dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t_max, t_w = [], []
    for _ in range(reps + 1):
        ev[0].record()
        X.project_coords_to_fixed_planes_gpu(coords, ref, planes, values_xyk=values, z_window=1.5)
        ev[1].record()
        X.project_coords_to_fixed_planes_weighted_gpu(coords, ref, planes, values_xyk=values, z_window=1.5)
        ev[2].record()
        torch.cuda.synchronize()
        t_max.append(ev[0].elapsed_time(ev[1]))
        t_w.append(ev[1].elapsed_time(ev[2]))
    n = values.numel()
    out_bytes = 8 * zyx[1] * zyx[2] * 4
    t_max, t_w = float(np.mean(t_max[1:])), float(np.mean(t_w[1:]))
    return {"samples": n, "planes": 8, "max_ms": t_max, "weighted_ms": t_w,
            "max_gbs": (16 * n + out_bytes) / (t_max * 1e-3) / 1e9,
            "weighted_gbs": (16 * n + 2 * out_bytes) / (t_w * 1e-3) / 1e9}


XRES = dict(seed=41, mov_xyz=(512, 512, 12), ref_xyz=(512, 512, 48), z_ratio=4.0, z_ratio_hr=1.0, n_blobs=6000,
            dmax=2.0, artefacts=24, layer=2, r=8, iter=10)


def xres_option(c):
    return {"layer": c["layer"], "iter": c["iter"], "r": c["r"], "zRatio": c["z_ratio"],
            "zRatio_HR": c["z_ratio_hr"], "smoothPenalty": 0.05, "movRange": 3.0, "tol": 1e-3,
            "mask_ref": np.zeros(c["ref_xyz"], np.float32), "mask_mov": np.zeros(c["mov_xyz"], np.float32)}


def xres_measure(reps=3, cpu=True, cpu_xy=128):
    """Cross-resolution registration (getMotion_v2, wrong-region correction
    on): moving 12 x 512 x 512 (z, y, x) at zRatio 4 against a 48 x 512 x 512
    reference at zRatio_HR 1, 6000 blobs, a 3-mode field of max |d| = 2, 24
    planted artefacts, layer 2, r 8, 10 iterations.  Inputs resident on the
    device; wall time of the whole call (its host waits per iteration
    included).  CPU: the oracle (bit-identical to the reference's numpy path)
    on the same generator at 12 x 128 x 128 / 48 x 128 x 128, one thread."""
    import time

    import torch

    from wholistic_registration_amd import calFlowCrossResolution as X
    from wholistic_registration_amd import synth
    c = XRES
    mov, ref, _ = synth.make_xres_pair_np(c["seed"], c["mov_xyz"], c["ref_xyz"], c["z_ratio"], c["z_ratio_hr"],
                                          c["n_blobs"], c["dmax"], c["artefacts"])
    dev = torch.device("cuda", 0)
    mov_t = torch.from_numpy(np.ascontiguousarray(mov.transpose(2, 1, 0))).to(dev).permute(2, 1, 0)
    ref_t = torch.from_numpy(np.ascontiguousarray(ref.transpose(2, 1, 0))).to(dev).permute(2, 1, 0)
    times, trace, info = [], [], {}
    for k in range(reps + 1):
        tr, inf = [], {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        X.getMotion_v2(mov_t, ref_t, xres_option(c), trace=tr, info=inf, output="device")
        torch.cuda.synchronize()
        if k:
            times.append(time.perf_counter() - t0)
        trace, info = tr, inf
    n = int(np.prod(c["mov_xyz"]))
    ms = 1e3 * float(np.median(times))
    out = {"moving_xyz": list(c["mov_xyz"]), "reference_xyz": list(c["ref_xyz"]), "layer": c["layer"],
           "r": c["r"], "iter": c["iter"], "ms": ms, "moving_voxels_per_s": n / (ms * 1e-3),
           "optimisations": [(t["records"], t["updates"]) for t in trace],
           "wrong_regions": {f"L{L}": (v or {}).get("regions") for L, v in info.items()},
           "accepted": {f"L{L}": (v or {}).get("accepted") for L, v in info.items()}}
    if cpu:
        from oracle import xres_oracle as xo
        small_m, small_r, _ = synth.make_xres_pair_np(c["seed"], (cpu_xy, cpu_xy, c["mov_xyz"][2]),
                                                      (cpu_xy, cpu_xy, c["ref_xyz"][2]), c["z_ratio"],
                                                      c["z_ratio_hr"], c["n_blobs"] * cpu_xy * cpu_xy // (512 * 512),
                                                      c["dmax"], c["artefacts"] * cpu_xy * cpu_xy // (512 * 512))
        o = xres_option(c)
        o["mask_ref"] = np.zeros(small_r.shape, np.float32)
        o["mask_mov"] = np.zeros(small_m.shape, np.float32)
        t0 = time.perf_counter()
        xo.get_motion_v2(small_m, small_r, o)
        sec = time.perf_counter() - t0
        nn = int(np.prod(small_m.shape))
        out["cpu_reference_path"] = {"moving_xyz": list(small_m.shape), "seconds": sec,
                                     "moving_voxels_per_s": nn / sec, "cores": 1,
                                     "kind": "port (bit-identical to the reference's numpy path)"}
    return out


def mask_measure(vol, crop, reps=3):
    """The harness's automatic mask of one moving volume (getMask +
    bwareafilt3_wei with the default config's thresFactor 5 and maskRange
    [5, 4000]) on the device, timed with events on the launch stream, and the
    oracle's mask of the centre crop on one host thread."""
    import torch

    from oracle import mask_oracle as MO
    from wholistic_registration_amd import device as D
    nz, ny, nx = vol.shape
    n = vol.numel()
    s_in = vol.element_size()
    m8 = torch.empty(vol.shape, dtype=torch.uint8, device=vol.device)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t_mask, t_cc = [], []
    for _ in range(reps + 1):
        ev[0].record()
        D.get_mask(vol, 5, out=m8)
        ev[1].record()
        f = D.bwareafilt3(m8, (5, 4000), out_dtype="float32")
        ev[2].record()
        torch.cuda.synchronize()
        t_mask.append(ev[0].elapsed_time(ev[1]))
        t_cc.append(ev[1].elapsed_time(ev[2]))
    t_mask, t_cc = float(np.mean(t_mask[1:])), float(np.mean(t_cc[1:]))
    frac = float((f > 0).float().mean())
    cz, cy, cx = crop
    z0, y0, x0 = (nz - cz) // 2, (ny - cy) // 2, (nx - cx) // 2
    sub = vol[z0:z0 + cz, y0:y0 + cy, x0:x0 + cx].cpu().numpy().transpose(2, 1, 0)
    t0 = time.perf_counter()
    MO.bwareafilt3(MO.get_mask(sub, 5), (5, 4000))
    dt = time.perf_counter() - t0
    b_mask = (s_in + 1) * n          # read the volume, write the byte mask
    b_cc = (1 + 4) * n               # read the byte mask, write the float32 mask
    return {
        "get_mask_ms": t_mask, "bwareafilt3_ms": t_cc, "masked_fraction": frac,
        "get_mask_gbs": b_mask / (t_mask * 1e-3) / 1e9, "bwareafilt3_gbs": b_cc / (t_cc * 1e-3) / 1e9,
        "algorithmic_bytes": {"get_mask": b_mask, "bwareafilt3": b_cc},
        "voxels_per_s": n / ((t_mask + t_cc) * 1e-3),
        "cpu_baseline": {"value": sub.size / dt, "unit": "masked voxels/s", "cores": 1, "kind": "port",
                         "sample": f"mask oracle (numpy percentile/z-score, scipy.ndimage morphology "
                                   f"and 26-connected labelling) on the centre crop {cz}x{cy}x{cx}, "
                                   f"one thread, {dt:.2f} s"},
        "sample": "moving volume of the workload, thresFactor 5, maskRange [5, 4000]",
    }


# Held-out pairs of the learned start, declared in advance: templates 90000 +
# 1000 k (k < 16) of the 64 x 512 x 1024 uint16 frame geometry, frames
# template + 1 .. template + 16.  Training used templates 10000 + 1000 i
# (i < 64) and validation 20000, 21000: none of these.
HELD_OUT_TEMPLATES = tuple(90000 + 1000 * k for k in range(16))
HELD_OUT_FRAMES = 16


def learned_measure(n_templates=16, per_template=HELD_OUT_FRAMES, reps=3, weights=None):
    """Plain FP32 pyramid LK against the learned start on the declared
    held-out pairs (up to 16 x 16 = 256).  Per pair: median wall time of
    getMotion + correctMotion over ``reps`` runs for plain FP32 and the
    learned start with FP32 and BF16 level images (the learned path also
    computes the moving frame's level-3 image and runs the network; the
    template's level-3 image once per template); one timed run of the learned
    BF16 start with the fallback as specified (the plain pyramid method also
    runs and its field is kept when its final residual is lower), which is
    what the learned start costs when every frame is checked; endpoint error
    of every field against the plain FP32 field (layer 4, r 8, 15 iterations:
    the gate's yardstick) and against the generating field; the same
    schedule from zero (the method without the network)."""
    import torch

    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    dev = torch.device("cuda", 0)
    net = Lr.InitNet.load(weights) if weights else Lr.InitNet.load()
    nz, ny, nx = T.FRAME
    n0 = nz * ny * nx
    warped = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
    motion = torch.empty((nz, ny, nx, 3), dtype=torch.float32, device=dev)

    def timed(fn, n):
        fn()
        ts = []
        for _ in range(n):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return r, 1e3 * float(np.median(ts))

    names = ("plain_fp32", "learned_fp32", "learned_bf16", "learned_bf16_fallback", "plain_fp32_same_schedule")
    rows = {k: {"ms": [], "epe": [], "epe_truth": []} for k in names}
    pairs, used_plain, frames_first = [], 0, []
    for tseed in HELD_OUT_TEMPLATES[:n_templates]:
        clean, tmpl = T.make_template(tseed, dev)
        base = dict(r=8, zRatio=1.0, movRange=5.0, smoothPenalty=T.penalty(tmpl), tol=1e-3)
        ref_l = D.pyramid_level(tmpl, Lr.LEVEL)

        # getMotion + correctMotion of the frame (the warp formed with the field)
        def plain(mov, layer, it):
            return D.get_motion(mov, tmpl, layer=layer, iter=it, motion_out=motion, warp_src=mov,
                                warped_out=warped, **base)

        def learned(mov, precision, fallback=None):
            return Lr.register_learned(mov, tmpl, net, ref_level=ref_l, precision=precision, motion_out=motion,
                                       fallback=fallback, warp_src=mov, warped_out=warped, **Lr.SCHEDULE, **base)

        for i in range(1, per_template + 1):
            mov, d = T.make_moving(clean, tseed + i, dev)
            if tseed == HELD_OUT_TEMPLATES[0] and len(frames_first) < 8:
                frames_first.append((mov, tmpl, base))
            truth = T.truth_field(d)
            _, ms = timed(lambda: plain(mov, 4, 15), reps)
            pf = motion.clone()
            res_t = {"plain_fp32": ms}
            _, res_t["learned_fp32"] = timed(lambda: learned(mov, "fp32"), reps)
            fields = {"learned_fp32": motion.clone()}
            _, res_t["learned_bf16"] = timed(lambda: learned(mov, "bf16"), reps)
            fields["learned_bf16"] = motion.clone()
            (_, info), res_t["learned_bf16_fallback"] = timed(lambda: learned(mov, "bf16", "plain"), 1)
            used_plain += info["used"] == "plain"
            fields["learned_bf16_fallback"] = motion.clone()
            _, res_t["plain_fp32_same_schedule"] = timed(lambda: plain(mov, 0, Lr.SCHEDULE["iter"]), 1)
            fields["plain_fp32_same_schedule"] = motion.clone()
            fields["plain_fp32"] = pf
            for k in names:
                rows[k]["ms"].append(res_t[k])
                rows[k]["epe"].append(float(torch.sqrt(((fields[k] - pf) ** 2).sum(-1)).mean()))
                rows[k]["epe_truth"].append(float(torch.sqrt(((fields[k] - truth) ** 2).sum(-1)).mean()))
            pairs.append(tseed + i)
            del d, truth, fields, pf
    out = {"workload": f"{len(pairs)} declared held-out pairs of {nz}x{ny}x{nx} (z,y,x) uint16 frames (templates "
                       f"{HELD_OUT_TEMPLATES[0]}..{HELD_OUT_TEMPLATES[n_templates - 1]} step 1000, frames +1..+"
                       f"{per_template}; 6-mode fields, max |d| 6 voxels), getMotion + correctMotion per frame",
           "epe_against": "plain FP32 field, layer 4, r 8, 15 iterations", "gate": "mean EPE per frame <= 0.1 voxel",
           "learned_schedule": dict(Lr.SCHEDULE), "fallback_kept_plain": used_plain}
    for k, row in rows.items():
        ms = float(np.mean(row["ms"]))
        e = np.asarray(row["epe"])
        miss = [(pairs[j], round(float(e[j]), 4)) for j in np.flatnonzero(e > 0.1)]
        out[k] = {"value": n0 / (ms * 1e-3), "unit": "registered voxels/s", "ms_per_frame": ms,
                  "epe_vs_plain_mean": float(e.mean()), "epe_vs_plain_max": float(e.max()),
                  "frames_within_gate": int((e <= 0.1).sum()), "frames": len(e), "misses": miss[:32],
                  "epe_vs_truth_mean": float(np.mean(row["epe_truth"]))}
    for k in ("learned_fp32", "learned_bf16", "learned_bf16_fallback"):
        out[f"speedup_{k}"] = out["plain_fp32"]["ms_per_frame"] / out[k]["ms_per_frame"]
    # the first template's frames on 4 concurrent contexts
    from wholistic_registration_amd.batch import FrameBatch
    batch = [f[0] for f in frames_first] * 2
    tmpl, base = frames_first[0][1], frames_first[0][2]
    conc = {}
    with FrameBatch(workers=4) as fb:
        wo = torch.empty((len(batch), nz, ny, nx), dtype=torch.float32, device=dev)
        mo = torch.empty((len(batch), nz, ny, nx, 3), dtype=torch.float32, device=dev)
        for name, kw in (("plain_fp32_4ctx", dict(layer=4, iter=15)),
                         ("learned_bf16_4ctx", dict(net=net, precision="bf16", **Lr.SCHEDULE))):
            fb.register(batch[:4], tmpl, motion_out=mo[:4], warped_out=wo[:4], **kw, **base)
            ts = []
            for _ in range(3):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                fb.register(batch, tmpl, motion_out=mo, warped_out=wo, **kw, **base)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            ms = 1e3 * float(np.median(ts)) / len(batch)
            conc[name] = {"value": n0 / (ms * 1e-3), "unit": "registered voxels/s", "ms_per_frame": ms,
                          "frames": len(batch)}
    out.update(conc)
    out["speedup_learned_bf16_4ctx"] = conc["plain_fp32_4ctx"]["ms_per_frame"] / conc["learned_bf16_4ctx"]["ms_per_frame"]
    del frames_first, batch, wo, mo
    # the first template's 16 held-out frames as one batch: the plain method's
    # batched launch sequence against the learned start's (one network pass of
    # batch 16, then one batched sequence warm-started by the control points)
    clean0, tmpl0 = T.make_template(HELD_OUT_TEMPLATES[0], dev)
    base0 = dict(r=8, zRatio=1.0, movRange=5.0, smoothPenalty=T.penalty(tmpl0), tol=1e-3)
    ref_l0 = D.pyramid_level(tmpl0, Lr.LEVEL)
    movs = torch.stack([T.make_moving(clean0, HELD_OUT_TEMPLATES[0] + i, dev)[0] for i in range(1, 17)])
    mo = torch.empty(tuple(movs.shape) + (3,), dtype=torch.float32, device=dev)
    mo_l = torch.empty_like(mo)
    wo = torch.empty(tuple(movs.shape), dtype=torch.float32, device=dev)
    runs = {"plain_fp32_batched16": lambda: D.get_motion_batch(movs, tmpl0, layer=4, iter=15, motion_out=mo,
                                                                warp_src=movs, warped_out=wo, **base0),
            "learned_bf16_batched16": lambda: Lr.register_learned_batch(
                movs, tmpl0, net, ref_level=ref_l0, precision="bf16", chunk=16, motion_out=mo_l, warp_src=movs,
                warped_out=wo, **Lr.SCHEDULE, **base0)}
    for name, fn in runs.items():
        _, ms = timed(fn, 3)
        ms /= movs.shape[0]
        out[name] = {"value": n0 / (ms * 1e-3), "unit": "registered voxels/s", "ms_per_frame": ms,
                     "frames": int(movs.shape[0])}
    e = torch.sqrt(((mo_l - mo) ** 2).sum(-1)).mean(dim=(1, 2, 3)).cpu().numpy()
    out["learned_bf16_batched16"]["frames_within_gate"] = int((e <= 0.1).sum())
    out["learned_bf16_batched16"]["epe_vs_plain_mean"] = float(e.mean())
    out["speedup_learned_bf16_batched16"] = (out["plain_fp32_batched16"]["ms_per_frame"] /
                                             out["learned_bf16_batched16"]["ms_per_frame"])
    out["speedup_learned_bf16_batched16_vs_plain_single"] = (out["plain_fp32"]["ms_per_frame"] /
                                                             out["learned_bf16_batched16"]["ms_per_frame"])
    del movs, mo, mo_l, wo
    m_l = D.pyramid_level(T.make_moving(T.make_template(HELD_OUT_TEMPLATES[0], dev)[0], HELD_OUT_TEMPLATES[0] + 1,
                                        dev)[0], Lr.LEVEL)
    ref_l = D.pyramid_level(T.make_template(HELD_OUT_TEMPLATES[0], dev)[1], Lr.LEVEL)
    for _ in range(3):
        net.predict(m_l, ref_l)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        net.predict(m_l, ref_l)
    torch.cuda.synchronize()
    net_ms = 1e3 * (time.perf_counter() - t0) / 10
    flops = Lr.forward_flops(*m_l.shape, channels=net.channels)
    n_par = sum(w.size + b.size for w, b in net.params().values())
    out["network"] = {"channels": list(net.channels), "parameters": int(n_par),
                      "forward_ms": net_ms, "flops": flops, "achieved_tflops": flops / (net_ms * 1e-3) / 1e12,
                      "peak_tflops": 2500.0, "frac": flops / (net_ms * 1e-3) / 2.5e15,
                      "input": list(m_l.shape) + [8], "note": "bf16 MFMA, dense peak; wall time of predict()"}
    return out


def crops_measure(n_crops=16, seed=3, reps=3):
    """Crops of uneven sizes (x in 128..1024, y in 128..768, z in 32..96,
    multiples of 8, random positions; seed 3) of one 96 x 768 x 1024 uint16
    frame pair (225k blobs, 6-mode field, max |d| = 6), r 8, 15 iterations,
    up to 4 levels (fewer where a crop's coarsest level would hold no patch),
    getMotion + correctMotion per crop.  Total time of all crops: on one
    context with the built-in tile rule, with autotuned per-level tiles, and
    with autotuned tiles on 4 concurrent contexts (FrameBatch); the tuning
    itself (once per geometry) is timed separately.  Then the learned start (the frame
    network, control-point start, layer 0, 5 iterations) on the same crops,
    on one and on 4 contexts, with each crop's endpoint error against its
    plain field."""
    import torch

    from wholistic_registration_amd import autotune
    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import synth
    from wholistic_registration_amd.batch import FrameBatch, max_layer
    dev = torch.device("cuda", 0)
    zyx = (96, 768, 1024)
    rng = np.random.default_rng(seed)
    clean = synth.render_template_torch(zyx, synth.draw_blobs(rng, zyx, 225_000), dev)
    d = synth.field_torch(zyx, synth.draw_modes(rng, 6), 6.0, dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    tmpl = torch.poisson(clean, generator=g).round_().clamp_(0, 65535).to(torch.uint16)
    mov = torch.poisson(D.correct_motion(clean, (-d).contiguous()).clamp_min(0), generator=g)
    mov = mov.round_().clamp_(0, 65535).to(torch.uint16)
    del clean, d
    sp = float(D.smooth_penalty_factor(tmpl.float().contiguous())) * RAW_PENALTY
    params = dict(r=8, zRatio=1.0, movRange=5.0, smoothPenalty=sp, tol=1e-3)
    pairs, shapes = [], []
    for _ in range(n_crops):
        cz, cy, cx = (int(rng.integers(lo // 8, hi // 8 + 1)) * 8 for lo, hi in ((32, 96), (128, 768), (128, 1024)))
        z0, y0, x0 = (int(rng.integers(0, n - c + 1)) for n, c in zip(zyx, (cz, cy, cx)))
        sl = (slice(z0, z0 + cz), slice(y0, y0 + cy), slice(x0, x0 + cx))
        pairs.append((mov[sl].contiguous(), tmpl[sl].contiguous()))
        shapes.append((cz, cy, cx))
    del mov, tmpl
    n_vox = sum(int(np.prod(s)) for s in shapes)
    cache = autotune.TileCache()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tiles = [cache.tiles(m, t, layer=max_layer(m.shape, 8), **params) for m, t in pairs]
    torch.cuda.synchronize()
    tune_s = time.perf_counter() - t0

    def sequential(tuned):
        for (m, t), tl in zip(pairs, tiles):
            res = D.get_motion(m, t, layer=max_layer(m.shape, 8), iter=15, tile_patches_x=tl if tuned else 0, **params)
            D.correct_motion(m, res.motion)

    def timed(fn):
        fn()
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    out = {"workload": f"{n_crops} crops of a 96x768x1024 uint16 pair, sizes (z,y,x) {shapes}, r 8, "
                       f"15 iterations, up to 4 levels, getMotion + correctMotion each",
           "voxels": n_vox, "tiles_per_level": [list(t) for t in tiles], "tuning_s": tune_s}
    for name, fn in (("rule_1ctx", lambda: sequential(False)), ("tuned_1ctx", lambda: sequential(True))):
        dt = timed(fn)
        out[name] = {"value": n_vox / dt, "ms": 1e3 * dt}
    with FrameBatch(workers=4, autotune=True) as fb:
        fb.tiles = cache
        dt = timed(lambda: fb.register_crops(pairs, iter=15, warp=True, **params))
        plain = [r.motion.clone() for r in fb.register_crops(pairs, iter=15, **params)[0]]
    out["tuned_4ctx"] = {"value": n_vox / dt, "ms": 1e3 * dt}
    # the learned start on the same crops (network trained on 64 x 512 x 1024 frames)
    from wholistic_registration_amd import learned as Lr
    net = Lr.InitNet.load()
    for workers, name in ((1, "learned_1ctx"), (4, "learned_4ctx")):
        with FrameBatch(workers=workers) as fb:
            dt = timed(lambda: fb.register_crops(pairs, iter=Lr.SCHEDULE["iter"], net=net, warp=True, **params))
            res = fb.register_crops(pairs, iter=Lr.SCHEDULE["iter"], net=net, **params)[0]
        epe = [float(torch.sqrt(((r.motion - pm) ** 2).sum(-1)).mean()) for r, pm in zip(res, plain)]
        out[name] = {"value": n_vox / dt, "ms": 1e3 * dt, "epe_vs_plain": [round(e, 4) for e in epe],
                     "crops_within_gate": int(sum(e <= 0.1 for e in epe))}
    out["unit"] = "registered voxels/s"
    return out


def net_layer_table(d, h, w, reps=20, warmup=3, device=0):
    """Per-layer timing of the learned start's convolutions: each 3x3x3
    convolution of the shipped network (forward, bf16 channels-last, fused
    bias and LeakyReLU) at the grid it has inside the network for a (D, H, W)
    input, on random data; the mean time of ``reps`` launches timed with HIP
    events on the launch stream.  Returns [(name, cin, cout, (d, h, w), us,
    flops, algorithmic bytes)] (activations in and out once, weights once)."""
    import ctypes

    from wholistic_registration_amd import _lib
    from wholistic_registration_amd.learned import SLOPE, InitNet, unet_layers
    import torch
    lib = _lib.load()
    ctx = _lib.context(device)
    dev = torch.device("cuda", device)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    s = ctypes.c_void_p(stream.cuda_stream)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    layers, depth = unet_layers(InitNet.load().channels)
    rows = []
    for name, cin, cout, act in layers:
        k = 2 ** depth[name]
        dd, hh, ww = d // k, h // k, w // k
        x = torch.randn((1, dd, hh, ww, cin), device=dev, generator=gen).to(torch.bfloat16)
        wt = (0.1 * torch.randn((cout, 27, cin), device=dev, generator=gen)).to(torch.bfloat16)
        b = torch.zeros(cout, device=dev)
        y = torch.empty((1, dd, hh, ww, cout), dtype=torch.bfloat16, device=dev)
        args = [ctypes.c_void_p(t.data_ptr()) for t in (x, wt, b, y)]

        def launch():
            _lib.check(lib.wr_net_conv3d(ctx, *args, 1, dd, hh, ww, cin, cout, int(act), SLOPE, s),
                       "wr_net_conv3d")

        for _ in range(warmup):
            launch()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            launch()
        e1.record(stream)
        torch.cuda.synchronize(dev)
        us = e0.elapsed_time(e1) * 1e3 / reps
        vox = dd * hh * ww
        flops = 2 * 27 * cin * cout * vox
        nbytes = 2 * (cin + cout) * vox + 2 * 27 * cin * cout + 4 * cout
        rows.append((name, cin, cout, (dd, hh, ww), us, flops, nbytes))
    return rows



def truth_epe(motion, d, mask):
    """Endpoint error of the GPU field against the synthetic registration
    field m = d(p + m) (fixed point, evaluated with the library's warp)."""
    import torch

    from wholistic_registration_amd import device as D
    comps = [d[..., c].contiguous() for c in range(3)]
    m = d.clone()
    for _ in range(25):
        m = torch.stack([D.correct_motion(comps[c], m) for c in range(3)], dim=-1).contiguous()
    e = torch.sqrt(((motion - m) ** 2).sum(-1))
    keep = mask <= 0 if mask is not None else torch.ones_like(e, dtype=torch.bool)
    ek = e[keep].float()
    q = torch.quantile(ek[:: max(1, ek.numel() // 4_000_000)], torch.tensor([0.5, 0.99], device=e.device))
    return {"epe_mean": float(ek.mean()), "epe_median": float(q[0]), "epe_p99": float(q[1]),
            "epe_max": float(ek.max()),
            "sample": "full volume outside the mask, against the generating field"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-epe-check", action="store_true",
                    help="skip the float64-oracle endpoint-error check on the crop")
    ap.add_argument("--cpu-crop", default="8,512,1024", help="z,y,x of the EPE-check (and window-sum oracle) crop")
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="processes of the all-cores CPU leg (0: the host's cores, at most 16)")
    ap.add_argument("--tile-patches-x", type=int, default=0)
    ap.add_argument("--precisions", default="bf16,fp16",
                    help="reduced-precision level-image variants to time beside the float32 path "
                         "(comma list of bf16, fp16; 'none' to skip)")
    ap.add_argument("--batch-frames", type=int, default=32,
                    help="frames of the time-series measurement (64x512x1024 frames; 0 to skip)")
    ap.add_argument("--no-mask-measure", action="store_true",
                    help="skip timing the harness's automatic mask kernels")
    ap.add_argument("--learned-templates", type=int, default=16,
                    help="held-out templates (16 frames each) of the plain vs learned-start comparison (0 to skip)")
    ap.add_argument("--crops", type=int, default=16,
                    help="crops of uneven sizes for the tile-autotuner measurement (0 to skip)")
    ap.add_argument("--crops-only", action="store_true",
                    help="only run the crops measurement (prints its JSON)")
    ap.add_argument("--learned-weights", default=None, metavar="PATH",
                    help="network weights of the learned-start comparison (default: the shipped file)")
    ap.add_argument("--learned-only", action="store_true",
                    help="only run the plain vs learned-start comparison (prints its JSON)")
    ap.add_argument("--net-layers", default=None, metavar="D,H,W",
                    help="only time each convolution of the learned start's network for a (D, H, W) input "
                         "(64,64,128: the level-3 image of a 64x512x1024 frame) and print one line per layer")
    ap.add_argument("--no-xres", action="store_true", help="skip the cross-resolution (getMotion_v2) measurement")
    ap.add_argument("--xres-only", action="store_true",
                    help="only run the cross-resolution measurement (prints its JSON)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}))
        return
    if args.net_layers:
        d, h, w = (int(v) for v in args.net_layers.split(","))
        tot_us = tot_f = 0.0
        for name, cin, cout, (dd, hh, ww), us, fl, nb in net_layer_table(d, h, w):
            tot_us += us
            tot_f += fl
            print(f"{name:6s} {cin:3d}->{cout:3d} @ {dd}x{hh}x{ww}: {us:7.1f} us {fl / us / 1e6:7.1f} TFLOP/s "
                  f"{nb / 1e6:6.1f} MB {nb / us / 1e3:6.0f} GB/s", flush=True)
        print(f"total {tot_us:.1f} us {tot_f / tot_us / 1e6:.1f} TFLOP/s", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (frames or z-slabs "
              f"shard without communication, see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch

    from wholistic_registration_amd import _lib
    from wholistic_registration_amd import device as D
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.xres_only:
        print(json.dumps(xres_measure(cpu=not args.no_cpu_baseline)), flush=True)
        return
    if args.learned_only:
        print(json.dumps(learned_measure(max(args.learned_templates, 1), weights=args.learned_weights)),
              flush=True)
        return
    if args.crops_only:
        print(json.dumps(crops_measure(max(args.crops, 1))), flush=True)
        return
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    w = WORKLOADS[args.workload]
    nz, ny, nx = w["zyx"]
    n0 = nx * ny * nz

    log(f"[bench] generating {args.workload} on the device")
    tmpl, mov, mask, field = make_pair_device(w, dev)
    sp = smooth_penalty(tmpl, mask) * RAW_PENALTY
    s_in = 2 if tmpl.dtype == torch.uint16 else 4
    motion = torch.empty((nz, ny, nx, 3), dtype=torch.float32, device=dev)
    warped = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
    params = dict(layer=w["layer"], iter=w["iter"], r=w["r"], zRatio=1.0, movRange=5.0,
                  smoothPenalty=sp, tol=1e-3, tile_patches_x=args.tile_patches_x)

    def step(timing, precision="fp32", motion_out=motion):
        # getMotion + correctMotion of the moving volume, the warp formed in the
        # kernel that writes the field (wr_get_motion_warp)
        return D.get_motion(mov, tmpl, mask_ref=mask, motion_out=motion_out, timing=timing,
                            precision=precision, warp_src=mov, warped_out=warped, **params)

    for i in range(args.warmup):
        res = step(False)
        torch.cuda.synchronize()
        log(f"[bench] warmup {i}: records {res.records[::-1]} (levels {w['layer']}..0)")
    torch.cuda.synchronize()
    times, recs = [], None
    t_all = time.perf_counter()
    for i in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = step(False)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        recs = res.records
        upds = res.updates
        log(f"[bench] step {i}: {times[-1] * 1e3:.1f} ms, records {res.records[::-1]}")
    total = time.perf_counter() - t_all
    ms_per_step = 1e3 * total / args.steps
    value = n0 / (total / args.steps)
    # the same registration with HIP events around every LK pass (outside the
    # timed region): per-level LK time and the level-0 pass duration
    layers = range(w["layer"], -1, -1)
    lk_lvl = {L: [0.0, 0] for L in layers}
    for i in range(args.steps):
        step(True)
        torch.cuda.synchronize()
        st = D.stats()
        for L in layers:
            lk_lvl[L][0] += st.lk_ms[L]
            lk_lvl[L][1] += st.lk_exec[L]
    lk_ms, lk_n = lk_lvl[0]

    # roofline of the level-0 LK pass
    bytes_per_launch = (2 * 4 + (1 if mask is not None else 0)) * n0
    avg_ms = lk_ms / max(lk_n, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    alg = algorithmic_bytes(w, recs, s_in, s_m=1 if mask is not None else 0)
    out = {
        "metric": "registered voxels/s at EPE<=0.1 vox vs FP64 oracle (plain FP32 path)",
        "value": value, "unit": "registered voxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "zyx": list(w["zyx"]), "input_dtype": str(tmpl.dtype),
                   "layer": w["layer"], "r": w["r"], "iter": w["iter"], "mask": mask is not None,
                   "records_per_level": {f"L{L}": recs[L] for L in range(w["layer"], -1, -1)},
                   "tile_cols": D.stats().tile_cols,
                   "lk_ms_per_level": {f"L{L}": {"ms": lk_lvl[L][0] / args.steps,
                                                "passes": lk_lvl[L][1] / args.steps} for L in layers}},
        "roofline": {"bound": "hbm", "kernel": "k_lk_pass (level 0)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "avg_launch_ms": avg_ms,
                     "bytes_per_launch": bytes_per_launch,
                     "bound_note": "the pass is bound by the vector-memory issue path (texture-address unit "
                                   "busy 0.83 of the cycles, the bfloat16 record layout with half the "
                                   "vector-memory instructions runs 1.6x faster at the same VALU count); HBM "
                                   "is the roofline it is priced against"},
        "pipeline_gbs": alg / (total / args.steps) / 1e9,
        "algorithmic_bytes_per_step": alg,
    }
    traffic = measured_traffic(args.workload)
    if traffic is not None:
        out["roofline"]["traffic"] = traffic["bytes_per_launch"]
        out["roofline"]["traffic_source"] = traffic["source"]
        # issue fractions of the same kernel from the committed counter runs
        out["roofline"]["issue"] = {"ta_busy": traffic["ta_busy"],
                                    "valu_per_wave": traffic["valu_instructions"] / traffic["waves"],
                                    "vmem_read_per_wave": traffic["vmem_read_instructions"] / traffic["waves"],
                                    "source": traffic["source"]}
    out["epe_vs_truth"] = truth_epe(motion, field, mask)
    precisions = [] if args.precisions == "none" else args.precisions.split(",")
    if precisions:
        variants = {}
        m2 = torch.empty_like(motion)
        for prec in precisions:
            log(f"[bench] {prec} level images")
            variants[prec] = precision_measure(w, n0, lambda t, p: step(t, p, m2), motion, field, mask,
                                               args.steps, prec)
        out["precision_variants"] = dict(variants, note=(
            "same workload and step with the pyramid level images stored as bfloat16 / float16; "
            "interpolation, sums and solves stay float32 / float64; value = registered voxels/s"))
    if args.batch_frames > 0:
        log(f"[bench] {args.batch_frames} frames against one template")
        out["batched_frames"] = batch_measure(args.batch_frames, (1, 2, 4, 8))
    if args.crops > 0:
        log(f"[bench] {args.crops} crops of uneven sizes")
        out["crops"] = crops_measure(args.crops)
    if args.learned_templates > 0:
        log(f"[bench] plain vs learned start on {16 * args.learned_templates} held-out frames")
        out["learned_init"] = learned_measure(args.learned_templates, weights=args.learned_weights)
    if not args.no_xres:
        log("[bench] cross-resolution registration (getMotion_v2)")
        out["cross_resolution"] = xres_measure(cpu=not args.no_cpu_baseline)
    if not args.no_mask_measure:
        log("[bench] fixed-plane projections")
        out["plane_projections"] = splat_measure()
        log("[bench] harness mask of the moving volume")
        out["harness_masks"] = mask_measure(mov, tuple(int(v) for v in args.cpu_crop.split(",")))
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (the reference's CPU path on a crop, extrapolated)")
        out["cpu_baseline"] = cpu_baseline_measure(tmpl, mov, mask, w, recs, upds,
                                                   workers=args.cpu_workers)
    if not (args.no_cpu_baseline and args.no_epe_check):
        crop = tuple(int(v) for v in args.cpu_crop.split(","))
        log(f"[bench] oracle runs on crop {crop}")
        out.update(crop_checks(tmpl, mov, mask, w, crop, not args.no_cpu_baseline,
                               not args.no_epe_check, precisions))
    print(json.dumps(out), flush=True)
    _ = _lib, math


if __name__ == "__main__":
    main()
