"""A batched registration enqueued without logs (wr_get_motion_batch with
logs = NULL) is captured in a HIP graph and replayed: every replay, with
other work on the device between replays, gives the eager call's fields bit
for bit (the stop rules run on the device, so a replay takes each frame's
own schedule again)."""

from __future__ import annotations

import numpy as np
import pytest

from tests.golden.harness_cases import harness_stack

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_batched_registration_replays_in_a_graph(cuda, precision):
    import torch

    from wholistic_registration_amd import _lib
    from wholistic_registration_amd import device as D
    tmpl, movs = harness_stack(seed=920, shape_zyx=(6, 64, 96), frames=3, uint16=True)
    ref = torch.from_numpy(tmpl).to(cuda)
    stack = torch.from_numpy(np.ascontiguousarray(movs)).to(cuda)
    mref = torch.from_numpy((tmpl > 2000).astype(np.float32)).to(cuda)
    prm = dict(layer=2, iter=8, r=3, zRatio=1.0, movRange=5.0, smoothPenalty=50.0, tol=1e-3,
               precision=precision, mask_ref=mref)
    want = [r.motion.clone() for r in D.get_motion_batch(stack, ref, **prm)]
    h = _lib.load().wr_create(cuda.index or 0)
    try:
        s = torch.cuda.Stream(device=cuda)
        out = torch.empty((3,) + tuple(stack.shape[1:]) + (3,), dtype=torch.float32, device=cuda)
        with torch.cuda.stream(s):      # first call: workspace and tables
            D.get_motion_batch(stack, ref, motion_out=out, context=h, stream=s, logs=False, **prm)
        s.synchronize()
        for f in range(3):
            assert torch.equal(out[f], want[f]), f
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            D.get_motion_batch(stack, ref, motion_out=out, context=h, stream=s, logs=False, **prm)
        for rep in range(3):
            out.fill_(float("nan"))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            junk = torch.randn(1 << 20, device=cuda).sum()     # other work between replays
            for f in range(3):
                assert torch.equal(out[f], want[f]), (rep, f)
            del junk
    finally:
        _lib.load().wr_destroy(h)


def test_capture_with_new_tables_is_refused(cuda):
    """A geometry the context has not run before needs a table upload, which
    cannot be captured: the call raises instead."""
    import torch

    from wholistic_registration_amd import _lib
    from wholistic_registration_amd import device as D
    tmpl, movs = harness_stack(seed=921, shape_zyx=(4, 48, 64), frames=2, uint16=True)
    ref = torch.from_numpy(tmpl).to(cuda)
    stack = torch.from_numpy(np.ascontiguousarray(movs)).to(cuda)
    prm = dict(layer=1, iter=4, r=3, zRatio=1.0, movRange=5.0, smoothPenalty=50.0, tol=1e-3)
    h = _lib.load().wr_create(cuda.index or 0)
    try:
        s = torch.cuda.Stream(device=cuda)
        out = torch.empty((2,) + tuple(stack.shape[1:]) + (3,), dtype=torch.float32, device=cuda)
        with torch.cuda.stream(s):
            D.get_motion_batch(stack, ref, motion_out=out, context=h, stream=s, logs=False, **prm)
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with pytest.raises((ValueError, RuntimeError)):
            with torch.cuda.graph(g, stream=s):
                D.get_motion_batch(stack, ref, motion_out=out, context=h, stream=s, logs=False,
                                   **dict(prm, layer=0))
    finally:
        torch.cuda.synchronize()
        _lib.load().wr_destroy(h)
