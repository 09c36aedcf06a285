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


def _frames_case(cuda, seed, frames, shape=(6, 48, 56), uint16=True):
    import torch
    tmpl, movs = harness_stack(seed=seed, shape_zyx=shape, frames=frames, uint16=uint16)
    ref = torch.from_numpy(tmpl).to(cuda)
    stack = torch.from_numpy(np.ascontiguousarray(movs)).to(cuda)
    return tmpl, ref, stack, movs


def _same_result(a, b):
    import torch
    assert torch.equal(a.motion, b.motion)
    assert a.records == b.records and a.updates == b.updates and a.nans == b.nans
    for k in ("diff_error", "penalty_error", "current_error", "max_diff_motion"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    assert a.final_error == b.final_error or (np.isnan(a.final_error) and np.isnan(b.final_error))


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("masks", [False, True])
def test_batched_launch_equals_single_calls(cuda, precision, masks):
    """wr_get_motion_batch: one launch per kernel and level-iteration over all
    frames, each frame's field, schedule and logs equal to wr_get_motion on
    that frame alone, bit for bit (frames stop at different iterations)."""
    import torch

    from wholistic_registration_amd import device as D
    tmpl, ref, stack, movs = _frames_case(cuda, 910 + masks, frames=5)
    mref = torch.from_numpy((tmpl > 2000).astype(np.float32)).to(cuda) if masks else None
    mmov = torch.from_numpy((movs > 2500).astype(np.float32)).to(cuda) if masks else None
    prm = dict(layer=2, iter=8, r=3, zRatio=1.0, movRange=5.0, smoothPenalty=50.0, tol=1e-3, precision=precision)
    res = D.get_motion_batch(stack, ref, mask_ref=mref, mask_mov=mmov, **prm)
    scheds = set()
    for f in range(stack.shape[0]):
        one = D.get_motion(stack[f], ref, mask_ref=mref, mask_mov=None if mmov is None else mmov[f], **prm)
        _same_result(res[f], one)
        scheds.add(tuple(sorted(one.schedule().items())))
    print(f"{len(scheds)} distinct schedules over {stack.shape[0]} frames")


def test_batched_launch_with_warm_start_and_forced_schedule(cuda):
    import torch

    from wholistic_registration_amd import device as D
    _, ref, stack, _ = _frames_case(cuda, 912, frames=3, shape=(4, 64, 80), uint16=False)
    g = torch.Generator(device="cpu").manual_seed(3)
    init = (0.3 * torch.randn((3,) + tuple(stack.shape[1:]) + (3,), generator=g)).to(cuda).contiguous()
    prm = dict(layer=1, iter=6, r=4, zRatio=2.0, movRange=5.0, smoothPenalty=30.0, tol=1e-3)
    sched = {1: (3, 2), 0: (6, 5)}
    for kw in (dict(), dict(schedule=sched)):
        res = D.get_motion_batch(stack, ref, motion_init=init, **prm, **kw)
        for f in range(stack.shape[0]):
            one = D.get_motion(stack[f], ref, motion_init=init[f], **prm, **kw)
            _same_result(res[f], one)


def test_batched_launch_single_frame_and_argument_errors(cuda):
    import torch

    from wholistic_registration_amd import device as D
    _, ref, stack, _ = _frames_case(cuda, 913, frames=2)
    prm = dict(layer=1, iter=5, r=3, zRatio=1.0, movRange=5.0, smoothPenalty=50.0, tol=1e-3)
    res = D.get_motion_batch(stack[:1].contiguous(), ref, **prm)
    _same_result(res[0], D.get_motion(stack[0], ref, **prm))
    with pytest.raises(ValueError):
        D.get_motion_batch(stack, ref[:, :, :-1].contiguous(), **prm)
    with pytest.raises(ValueError):
        D.get_motion_batch(stack, ref, mask_mov=torch.zeros_like(stack[0], dtype=torch.float32), **prm)


def test_register_batched_equals_register(cuda):
    """FrameBatch.register_batched (chunks of frames per batched launch,
    chunks spread over two contexts, warps included) equals register."""
    import torch

    from wholistic_registration_amd.batch import FrameBatch
    _, ref, stack, _ = _frames_case(cuda, 914, frames=5)
    prm = dict(layer=1, iter=6, r=3, zRatio=1.0, movRange=5.0, smoothPenalty=50.0, tol=1e-3)
    w1 = torch.empty(stack.shape, dtype=torch.float32, device=cuda)
    w2 = torch.empty_like(w1)
    with FrameBatch(workers=2) as fb:
        m1, r1 = fb.register(list(stack), ref, warped_out=w1, **prm)
        m2, r2 = fb.register_batched(stack, ref, chunk=2, warped_out=w2, **prm)
    torch.cuda.synchronize()
    assert torch.equal(m1, m2) and torch.equal(w1, w2)
    assert [r.records for r in r1] == [r.records for r in r2]


@pytest.mark.parametrize("src_u16", [True, False])
def test_fused_warp_equals_separate_calls(cuda, src_u16):
    """wr_get_motion_warp / the batched warp: the field and the warp of a
    volume by it (formed in the kernel that writes the field) equal
    getMotion then correctMotion, bit for bit; the warped volume may be
    another channel than the registered one."""
    import torch

    from wholistic_registration_amd import device as D
    tmpl, ref, stack, movs = _frames_case(cuda, 915, frames=3, uint16=True)
    other = torch.from_numpy(np.ascontiguousarray(movs[::-1])).to(cuda)      # another channel
    src = other if src_u16 else other.float().contiguous()
    mref = torch.from_numpy((tmpl > 2000).astype(np.float32)).to(cuda)
    prm = dict(layer=2, iter=8, r=3, zRatio=1.0, movRange=5.0, smoothPenalty=50.0, tol=1e-3, mask_ref=mref)
    res = D.get_motion_batch(stack, ref, warp_src=src, **prm)
    warped = torch.empty(stack.shape, dtype=torch.float32, device=cuda)
    res2 = D.get_motion_batch(stack, ref, warp_src=src, warped_out=warped, **prm)
    for f in range(stack.shape[0]):
        one = D.get_motion(stack[f], ref, warp_src=src[f], **prm)
        plain = D.get_motion(stack[f], ref, **prm)
        sep = D.correct_motion(src[f], plain.motion)
        torch.cuda.synchronize()
        _same_result(one, plain)
        assert torch.equal(one.extras["warped"], sep)
        assert torch.equal(res[f].motion, plain.motion) and torch.equal(res2[f].motion, plain.motion)
        assert torch.equal(warped[f], sep)
