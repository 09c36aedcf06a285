"""The time-series schedule (main_function.py:512-964, 1443-1697 of the
reference) with the per-batch registration stubbed out: CPU only."""

from __future__ import annotations

import numpy as np
import pytest

from wholistic_registration_amd import series


def test_build_chunks_forward_and_backward():
    assert series.build_chunks(15, 23, 5, "forward") == [[15, 16, 17, 18, 19], [20, 21, 22]]
    assert series.build_chunks(8, 0, 5, "backward") == [[7, 6, 5, 4, 3], [2, 1, 0]]
    # a residual shorter than half a window joins the previous chunk
    assert series.build_chunks(0, 11, 5, "forward") == [[0, 1, 2, 3, 4], [5, 6, 7, 8, 9, 10]]
    assert series.build_chunks(5, 5, 5, "forward") == []
    with pytest.raises(AssertionError):
        series.build_chunks(0, 3, 2, "sideways")


def _config(total, dual=False):
    return {
        "MetaData": {"zRatio": 1.0, "Dim": 3, "fps": 1.0},
        "frames": {"frames": list(range(100, 100 + total)), "frame_downsample": 1},
        "reference": {"pick_reference_auto": False, "time_measurement": "frame", "mid_window_size": 6,
                      "window_size": 5, "mid_stride": 1, "reference_chunk": 3},
        "processing": {"batch_size": 2},
        "channels": {"dual_channel": dual, "k": 0, "function": "raw"},
        "save_config": {"save_ref": True, "save_motion": True},
        "mask": {"maskRange": [5, 4000], "thresFactor": 5},
        "pyramid": {"r": 2, "layer": 1, "iter": 2, "tolerance": 1e-3, "smoothPenalty": 0.08},
    }


def test_schedule_order_references_and_warm_start(monkeypatch):
    total = 23
    cfg = _config(total)
    calls, refs, out = [], [], []

    def read(ids, channel):
        return np.asarray([np.full((2, 3, 4), fid, dtype=np.float32) for fid in ids])

    def fake_ref(block, config, ca_block=None):
        ids = [int(b.flat[0]) for b in np.asarray(block)]
        refs.append(ids)
        return np.full((2, 3, 4), -1.0, dtype=np.float32), np.arange(len(ids))

    def fake_reg(mem_batch, config, ref_img, frame=None, direction="forward", motion_init=None,
                 moving_Ca_image=None, verbose=True):
        init = None if motion_init is None else int(motion_init.flat[0])
        calls.append((list(frame), direction, init))
        motion = np.asarray([np.full((2, 3, 4, 3), f, dtype=np.float32) for f in frame])
        return mem_batch.copy(), np.array([]), None, [], motion

    monkeypatch.setattr(series.reference, "compute_reference_from_block", fake_ref)
    monkeypatch.setattr(series.registration, "wbi_registration_3d", fake_reg)
    series.register_series(cfg, read, lambda kind, fid, a: out.append((kind, fid)))

    f = lambda *idx: [100 + i for i in idx]   # noqa: E731  frame ids of indices
    # middle window [8, 14): reference from frames 8..13, block 8..14 in batches of 2
    assert refs[0] == f(8, 9, 10, 11, 12, 13)
    assert [c[0] for c in calls[:4]] == [f(8, 9), f(10, 11), f(12, 13), f(14)]
    assert all(c[2] is None for c in calls[:4])
    # backward: chunks [7..3], [2, 1, 0]; first reference from the block's head 8, 9, 10
    assert refs[1] == f(8, 9, 10)
    bwd = [c for c in calls if c[1] == "backward"]
    assert [c[0] for c in bwd] == [f(7, 6), f(5, 4), f(3), f(2, 1), f(0)]
    assert [c[2] for c in bwd] == [108, 106, 104, 103, 101]      # warm start chain
    assert refs[2] == f(5, 4, 3)                                 # rolling window of the chunk
    # forward: chunks [15..19], [20..22]; reference from the block's tail 12, 13, 14
    assert refs[3] == f(12, 13, 14)
    fwd = [c for c in calls if c[1] == "forward" and c[2] is not None]
    assert [c[0] for c in fwd] == [f(15, 16), f(17, 18), f(19), f(20, 21), f(22)]
    assert [c[2] for c in fwd] == [114, 116, 118, 119, 121]
    assert refs[4] == f(17, 18, 19)
    # every frame is written exactly once, with its motion
    mem = [fid for kind, fid in out if kind == "membrane"]
    assert sorted(mem) == f(*range(total)) and len(mem) == total
    assert sum(kind == "motion" for kind, _ in out) == total
    assert sum(kind == "reference" for kind, _ in out) == 5
