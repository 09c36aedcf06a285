"""Frames of a time series registered concurrently against one template
(FrameBatch, wr_params.reuse_template): every frame's field and warp equal a
single getMotion / correctMotion call on it, bit for bit."""

from __future__ import annotations

import numpy as np
import pytest

from tests.golden.harness_cases import harness_stack

pytestmark = pytest.mark.gpu

PARAMS = dict(layer=1, iter=6, r=3, zRatio=1.0, movRange=5.0, smoothPenalty=50.0, tol=1e-3)


@pytest.mark.parametrize("workers,precision", [(3, "fp32"), (2, "bf16"), (1, "fp32")])
def test_batch_equals_single_calls(cuda, workers, precision):
    import torch

    from wholistic_registration_amd import device as D
    from wholistic_registration_amd.batch import FrameBatch
    tmpl, movs = harness_stack(seed=904, frames=5, uint16=True)
    ref = torch.from_numpy(tmpl).to(cuda)
    frames = [torch.from_numpy(m).to(cuda) for m in movs]
    mask = torch.from_numpy((tmpl > 2000).astype(np.float32)).to(cuda)
    warped = torch.empty((len(frames),) + tuple(ref.shape), dtype=torch.float32, device=cuda)
    with FrameBatch(workers=workers) as fb:
        motion, res = fb.register(frames, ref, warped_out=warped, mask_ref=mask, precision=precision, **PARAMS)
        again, _ = fb.register(frames, ref, mask_ref=mask, precision=precision, **PARAMS)   # template reused
    torch.cuda.synchronize()
    for i, f in enumerate(frames):
        one = D.get_motion(f, ref, mask_ref=mask, precision=precision, **PARAMS)
        w1 = D.correct_motion(f, one.motion)
        torch.cuda.synchronize()
        assert torch.equal(motion[i], one.motion), i
        assert torch.equal(again[i], one.motion), i
        assert torch.equal(warped[i], w1), i
        assert res[i].records == one.records


def test_reuse_template_needs_a_previous_call(cuda):
    import torch

    from wholistic_registration_amd import _lib
    from wholistic_registration_amd import device as D
    tmpl, movs = harness_stack(seed=905, frames=1)
    ref = torch.from_numpy(tmpl).to(cuda)
    mov = torch.from_numpy(movs[0]).to(cuda)
    h = _lib.load().wr_create(0)
    try:
        with pytest.raises(ValueError):
            D.get_motion(mov, ref, context=h, reuse_template=True, **PARAMS)
        D.get_motion(mov, ref, context=h, **PARAMS)
        other = ref.clone()
        with pytest.raises(ValueError):     # another template tensor
            D.get_motion(mov, other, context=h, reuse_template=True, **PARAMS)
    finally:
        _lib.load().wr_destroy(h)
