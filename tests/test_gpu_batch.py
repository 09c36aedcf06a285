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


def test_reuse_template_with_another_mask_ref_is_an_error(cuda):
    """The reuse key holds mask_ref's pointer: a new mask_ref with
    reuse_template raises instead of reusing the old mask pyramid."""
    import torch

    from wholistic_registration_amd import _lib
    from wholistic_registration_amd import device as D
    tmpl, movs = harness_stack(seed=906, frames=1)
    ref = torch.from_numpy(tmpl).to(cuda)
    mov = torch.from_numpy(movs[0]).to(cuda)
    m1 = torch.from_numpy((tmpl > 2000).astype(np.float32)).to(cuda)
    m2 = torch.from_numpy((tmpl > 1500).astype(np.float32)).to(cuda)
    h = _lib.load().wr_create(0)
    try:
        D.get_motion(mov, ref, mask_ref=m1, context=h, **PARAMS)
        D.get_motion(mov, ref, mask_ref=m1, context=h, reuse_template=True, **PARAMS)
        with pytest.raises(ValueError):
            D.get_motion(mov, ref, mask_ref=m2, context=h, reuse_template=True, **PARAMS)
        with pytest.raises(ValueError):     # the mask dropped
            D.get_motion(mov, ref, context=h, reuse_template=True, **PARAMS)
    finally:
        _lib.load().wr_destroy(h)


def test_second_register_on_a_new_template_at_the_same_address(cuda):
    """A second FrameBatch.register whose template occupies the freed
    address of the first call's template gets a fresh pyramid: its fields
    equal single calls on the new template."""
    import torch

    from wholistic_registration_amd import device as D
    from wholistic_registration_amd.batch import FrameBatch
    tmpl_a, movs = harness_stack(seed=907, frames=3, uint16=True)
    tmpl_b, _ = harness_stack(seed=908, frames=1, uint16=True)
    frames = [torch.from_numpy(m).to(cuda) for m in movs]
    buf = torch.from_numpy(tmpl_a).to(cuda)
    with FrameBatch(workers=2) as fb:
        fb.register(frames, buf, **PARAMS)
        buf.copy_(torch.from_numpy(tmpl_b))          # same address, new template
        torch.cuda.synchronize()
        motion, _ = fb.register(frames, buf, **PARAMS)
    torch.cuda.synchronize()
    for i, f in enumerate(frames):
        one = D.get_motion(f, buf, **PARAMS)
        torch.cuda.synchronize()
        assert torch.equal(motion[i], one.motion), i
