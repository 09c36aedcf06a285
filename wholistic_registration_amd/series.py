"""Time-series scheduler of the registration pipeline on one MI355X.

Restates the frame schedule of ``Registration_v3``
(``src/wholistic_registration/core/main_function.py:512-964``) and its
helpers ``build_chunks`` (:1443-1475) and ``process_directional_chunks``
(:1478-1697) around this package's harness
(``registration.wbi_registration_3d``):

1. the middle window of the recording gives the initial reference
   (``reference.compute_reference_from_block`` over every ``mid_stride``-th
   frame of it);
2. the middle block is registered in batches against it, keeping its first
   and last ``reference_chunk`` registered frames;
3. the frames before it are registered backward and the frames after it
   forward, in chunks of ``window_size`` frames (a residual chunk shorter than
   half a window joins the previous one); each chunk's reference is computed
   from the rolling window of the last ``reference_chunk`` registered frames,
   and each batch starts from the previous batch's last motion field (warm
   start).

Frame I/O is the caller's: ``read(frame_ids, channel)`` returns a
``(len(frame_ids), Z, Y, X)`` array (channel 1 membrane, 0 calcium, the
reference's convention), and ``sink(kind, frame_id, array)`` receives every
output (``"membrane"``, ``"calcium"``, ``"motion"``, ``"reference"``).  The
reference reads ND2 files and writes OME-TIFF / HDF5 there
(``IO.readND2Frame``, ``IO.write_volume_as_ome_tiff``); those formats are not
part of this package.  With one GPU both directions run in series, as the
reference's serial mode does.
"""

from __future__ import annotations

from collections import deque

import numpy as np

from . import reference, registration


def build_chunks(start_frame, end_frame, window_size, direction):
    """Frame-index chunks of one direction (main_function.py:1443-1475):
    forward ``start..end-1``, backward ``start-1`` down to ``end``; a last
    chunk shorter than ``window_size // 2`` is merged into the one before."""
    assert direction in ("forward", "backward")
    if direction == "forward":
        frames = list(range(start_frame, end_frame))
    else:
        frames = list(range(start_frame - 1, end_frame - 1, -1))
    chunks = [frames[i:i + window_size] for i in range(0, len(frames), window_size)]
    if len(chunks) >= 2 and len(chunks[-1]) < window_size // 2:
        chunks[-2].extend(chunks[-1])
        chunks.pop(-1)
    return chunks


def _frames_of(config, key, value):
    """Minutes or frames -> frames (main_function.py:573-581, 626-637)."""
    ref = config["reference"]
    fd = config["frames"]["frame_downsample"]
    if ref["time_measurement"] == "minute":
        return int(value * 60 * config["MetaData"]["fps"] / fd)
    if ref["time_measurement"] == "frame":
        return int(value / fd)
    raise ValueError("[ERROR]Invalid time_measurement in reference config.")


def process_directional_chunks(direction, frame_list, start_frame, end_frame, init_mem_windows,
                               init_ca_windows, config, read, sink, window_size_frames, reference_chunk,
                               batch_size=50, motion_init=None, save_ref=False, save_motion=False,
                               verbose=False):
    """One direction of the schedule (main_function.py:1478-1697)."""
    dual = config["channels"]["dual_channel"]
    win_mem = deque(init_mem_windows, maxlen=reference_chunk)
    win_ca = deque(init_ca_windows, maxlen=reference_chunk)
    for frames in build_chunks(start_frame, end_frame, window_size_frames, direction):
        ref_img, _ = reference.compute_reference_from_block(
            np.asarray(list(win_mem)), config, np.asarray(list(win_ca)) if dual else None)
        win_mem.clear()
        win_ca.clear()
        if save_ref:
            ends = (frames[0], frames[-1]) if direction == "forward" else (frames[-1], frames[0])
            sink("reference", (frame_list[ends[0]], frame_list[ends[1]]), ref_img)
        for i in range(0, len(frames), batch_size):
            batch = [frame_list[j] for j in frames[i:i + batch_size]]
            mem_batch = read(batch, 1)
            ca_batch = read(batch, 0) if dual else None
            mem_reg, ca_reg, _, _, motion = registration.wbi_registration_3d(
                mem_batch, config, ref_img, frame=batch, direction=direction, motion_init=motion_init,
                moving_Ca_image=ca_batch, verbose=verbose)
            motion_init = motion[-1]
            for k, fid in enumerate(batch):
                sink("membrane", fid, mem_reg[k])
                if dual:
                    sink("calcium", fid, ca_reg[k])
                if save_motion:
                    sink("motion", fid, motion[k])
                win_mem.append(mem_reg[k])
                if dual:
                    win_ca.append(ca_reg[k])
    return motion_init


def register_series(config, read, sink, verbose=False):
    """The schedule of ``Registration_v3`` (main_function.py:512-964) for a
    3-D recording of ``len(config["frames"]["frames"])`` frames."""
    process_frames = list(config["frames"]["frames"])
    total = len(process_frames)
    dual = config["channels"]["dual_channel"]
    batch_size = config["processing"]["batch_size"]
    window_frames = _frames_of(config, "window_size", config["reference"]["window_size"])
    mid_frames = _frames_of(config, "mid_window_size", config["reference"]["mid_window_size"])
    reference_chunk = _frames_of(config, "reference_chunk", config["reference"]["reference_chunk"])
    save_ref = bool(config["save_config"]["save_ref"])
    save_motion = bool(config["save_config"]["save_motion"])

    # middle window, clamped to the recording (main_function.py:638-648)
    mid_start = max(0, total // 2 - mid_frames // 2)
    mid_end = min(mid_start + mid_frames, total)
    frames_mid = list(range(mid_start, min(mid_end + 1, total)))
    stride = int(config["reference"]["mid_stride"])
    sampled = [process_frames[i] for i in range(mid_start, mid_end, stride)]
    mem_mid = read(sampled, 1)
    ca_mid = read(sampled, 0) if dual else None
    ref_img, indsort = reference.compute_reference_from_block(mem_mid, config, ca_mid)
    if save_ref:
        sink("reference", (process_frames[frames_mid[0]], process_frames[frames_mid[-1]]), ref_img)

    # middle block in batches, keeping its head and tail (main_function.py:705-804)
    head_mem, head_ca = deque(maxlen=reference_chunk), deque(maxlen=reference_chunk)
    tail_mem, tail_ca = deque(maxlen=reference_chunk), deque(maxlen=reference_chunk)
    n_mid = len(frames_mid)
    motion_start = motion_end = None
    for i in range(0, n_mid, batch_size):
        batch = [process_frames[j] for j in range(mid_start + i, mid_start + min(n_mid, i + batch_size))]
        mem_batch = read(batch, 1)
        ca_batch = read(batch, 0) if dual else None
        mem_reg, ca_reg, _, _, motion = registration.wbi_registration_3d(
            mem_batch, config, ref_img, frame=batch, moving_Ca_image=ca_batch, verbose=verbose)
        if i == 0:
            motion_start = motion[0].copy()
        if i + batch_size >= n_mid:
            motion_end = motion[-1].copy()
        for k, fid in enumerate(batch):
            sink("membrane", fid, mem_reg[k])
            if dual:
                sink("calcium", fid, ca_reg[k])
            if save_motion:
                sink("motion", fid, motion[k])
        for k in range(len(mem_reg)):
            if len(head_mem) < reference_chunk:
                head_mem.append(mem_reg[k])
                if dual:
                    head_ca.append(ca_reg[k])
            tail_mem.append(mem_reg[k])
            if dual:
                tail_ca.append(ca_reg[k])

    # backward, then forward (the reference's serial mode, main_function.py:836-902)
    common = dict(frame_list=process_frames, config=config, read=read, sink=sink,
                  window_size_frames=window_frames, reference_chunk=reference_chunk, batch_size=batch_size,
                  save_ref=save_ref, save_motion=save_motion, verbose=verbose)
    process_directional_chunks("backward", start_frame=mid_start, end_frame=0, init_mem_windows=head_mem,
                               init_ca_windows=head_ca, motion_init=motion_start, **common)
    process_directional_chunks("forward", start_frame=mid_end + 1, end_frame=total, init_mem_windows=tail_mem,
                               init_ca_windows=tail_ca, motion_init=motion_end, **common)
    return indsort
