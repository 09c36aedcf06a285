"""What the trained network gains over the method without it, on held-out
64 x 512 x 1024 uint16 frame pairs (template seeds 90000 and 92000 and their
moving frames; training used template seeds 10000 + 1000 i, i < 64, with
frames +1..+16, and validation 20000, 21000).

The plain pyramid LK (layer 4, r 8, 15 iterations, float32) is the yardstick:
endpoint errors are measured against its field.  The network's field starts
getMotion at full resolution (layer 0) with 5 iterations and ends within the
0.1-voxel gate; the same schedule started from zero (the method without the
network, at the same cost) ends outside it with about twice the error.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEEDS = (90001, 90002)


@pytest.fixture(scope="module")
def setup(cuda):
    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    if not os.path.exists(Lr.WEIGHTS):
        pytest.fail(f"trained weights missing: {Lr.WEIGHTS}")
    net = Lr.InitNet.load()
    clean, tmpl = T.make_template(90000, cuda)
    params = dict(r=8, zRatio=1.0, movRange=5.0, smoothPenalty=T.penalty(tmpl), tol=1e-3)
    ref_l = D.pyramid_level(tmpl, Lr.LEVEL)
    yield dict(net=net, clean=clean, tmpl=tmpl, params=params, ref_l=ref_l)


def _epe(a, b):
    import torch
    return float(torch.sqrt(((a - b) ** 2).sum(-1)).mean())


def _other_template(seed, cuda):
    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    clean, tmpl = T.make_template(seed, cuda)
    params = dict(r=8, zRatio=1.0, movRange=5.0, smoothPenalty=T.penalty(tmpl), tol=1e-3)
    return dict(clean=clean, tmpl=tmpl, params=params, ref_l=D.pyramid_level(tmpl, Lr.LEVEL))


@pytest.mark.parametrize("template,seed", [(90000, SEEDS[0]), (90000, SEEDS[1]), (92000, 92001)])
def test_learned_start_beats_method_without_it(setup, cuda, template, seed):
    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    s = setup if template == 90000 else dict(setup, **_other_template(template, cuda))
    mov, _ = T.make_moving(s["clean"], seed, s["tmpl"].device)
    p = s["params"]
    plain_field = D.get_motion(mov, s["tmpl"], layer=4, iter=15, **p).motion.clone()
    learned, info = Lr.register_learned(mov, s["tmpl"], s["net"], ref_level=s["ref_l"], **Lr.SCHEDULE, **p)
    assert info["used"] == "learned"
    e_learned = _epe(learned.motion, plain_field)
    zero = D.get_motion(mov, s["tmpl"], **Lr.SCHEDULE, **p)
    e_zero = _epe(zero.motion, plain_field)
    assert e_learned <= 0.1, e_learned               # the gate
    assert e_zero > 0.1 and e_zero > 1.5 * e_learned, (e_zero, e_learned)
    assert learned.final_error < zero.final_error


def test_fallback_keeps_the_lower_residual(setup):
    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    s = setup
    mov, _ = T.make_moving(s["clean"], SEEDS[0], s["tmpl"].device)
    p = s["params"]
    res, info = Lr.register_learned(mov, s["tmpl"], s["net"], ref_level=s["ref_l"], fallback="plain", **Lr.SCHEDULE,
                                    **p)
    assert set(info) >= {"used", "learned_error", "plain_error"}
    assert res.final_error == min(info["learned_error"], info["plain_error"])
    assert info["used"] == ("plain" if info["plain_error"] < info["learned_error"] else "learned")
    # the plain side is the full pyramid method (4 levels, 15 iterations from zero)
    from wholistic_registration_amd import device as D
    plain = D.get_motion(mov, s["tmpl"], layer=4, iter=15, **p)
    assert info["plain_error"] == plain.final_error


def test_fallback_fills_the_callers_buffer(setup):
    """When the plain field is kept it is the one in motion_out (forced here
    with a plain run of more iterations than the learned one can match)."""
    import torch

    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    s = setup
    mov, _ = T.make_moving(s["clean"], SEEDS[0], s["tmpl"].device)
    out = torch.full(tuple(mov.shape) + (3,), float("nan"), device=mov.device)
    res, info = Lr.register_learned(mov, s["tmpl"], s["net"], ref_level=s["ref_l"], fallback="plain", layer=0, iter=1,
                                    motion_out=out, **s["params"])
    assert info["used"] == "plain", info
    assert res.motion.data_ptr() == out.data_ptr()
    assert bool(torch.isfinite(out).all())


def test_prediction_is_deterministic_and_finite(setup):
    import torch

    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    s = setup
    mov, _ = T.make_moving(s["clean"], SEEDS[1], s["tmpl"].device)
    m_l = D.pyramid_level(mov, Lr.LEVEL)
    a = s["net"].predict(m_l, s["ref_l"]).clone()
    b = s["net"].predict(m_l, s["ref_l"])
    assert torch.equal(a, b)
    assert torch.isfinite(a).all()
    assert tuple(a.shape) == (64, 64, 128, 3)
    assert float(a.abs().max()) < 20.0


def test_learned_frames_on_concurrent_contexts_equal_single_runs(setup):
    import torch

    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    from wholistic_registration_amd.batch import FrameBatch
    s = setup
    frames = [T.make_moving(s["clean"], seed, s["tmpl"].device)[0] for seed in (90003, 90004, 90005)]
    with FrameBatch(workers=2) as fb:
        motion, results = fb.register(frames, s["tmpl"], net=s["net"], **Lr.SCHEDULE, **s["params"])
    for f, m, res in zip(frames, motion, results):
        single, _ = Lr.register_learned(f, s["tmpl"], s["net"], ref_level=s["ref_l"], **Lr.SCHEDULE, **s["params"])
        assert torch.equal(single.motion, m)
        assert single.records == res.records


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_batched_learned_frames_equal_single_runs(setup, precision):
    """register_learned_batch (one network pass of batch 3, one batched
    launch sequence warm-started by each frame's control points, the warp
    formed with the field) gives every frame the field, schedule and warp of
    register_learned on the frame alone."""
    import torch

    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    s = setup
    movs = torch.stack([T.make_moving(s["clean"], seed, s["tmpl"].device)[0] for seed in (90006, 90007, 90008)])
    mo = torch.empty(tuple(movs.shape) + (3,), dtype=torch.float32, device=movs.device)
    wo = torch.empty(tuple(movs.shape), dtype=torch.float32, device=movs.device)
    res = Lr.register_learned_batch(movs, s["tmpl"], s["net"], ref_level=s["ref_l"], precision=precision, chunk=3,
                                    motion_out=mo, warp_src=movs, warped_out=wo, **Lr.SCHEDULE, **s["params"])
    for f in range(movs.shape[0]):
        w1 = torch.empty_like(wo[f])
        single, _ = Lr.register_learned(movs[f], s["tmpl"], s["net"], ref_level=s["ref_l"], precision=precision,
                                        warp_src=movs[f], warped_out=w1, **Lr.SCHEDULE, **s["params"])
        assert torch.equal(single.motion, mo[f])
        assert single.records == res[f].records and single.updates == res[f].updates
        assert torch.equal(w1, wo[f])


def test_batched_learned_on_the_registration_context_equals_single_runs(setup):
    """With the network forked onto the registration's own context (one
    workspace), register_learned_batch runs the network on the registration's
    stream instead of a second one; the fields still equal the single runs."""
    import torch

    from wholistic_registration_amd import _lib
    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    s = setup
    dev = s["tmpl"].device
    ctx = _lib.context(dev.index or 0)
    net = s["net"].fork(ctx)
    movs = torch.stack([T.make_moving(s["clean"], seed, dev)[0] for seed in (90009, 90010)])
    mo = torch.empty(tuple(movs.shape) + (3,), dtype=torch.float32, device=dev)
    res = Lr.register_learned_batch(movs, s["tmpl"], net, ref_level=s["ref_l"], precision="bf16", chunk=2,
                                    motion_out=mo, context=ctx, **Lr.SCHEDULE, **s["params"])
    for f in range(movs.shape[0]):
        single, _ = Lr.register_learned(movs[f], s["tmpl"], s["net"], ref_level=s["ref_l"], precision="bf16",
                                        **Lr.SCHEDULE, **s["params"])
        assert torch.equal(single.motion, mo[f])
        assert single.records == res[f].records


def test_learned_start_on_uneven_crops(setup):
    """Crops whose level-3 image is not a multiple of the network's (4, 4, 32)
    grid run on an edge-padded grid; crops on concurrent contexts equal their
    single runs, the learned start beats the zero start on a deep crop, and by
    default a crop shallower than 48 planes runs the plain pyramid method."""
    import torch

    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    from wholistic_registration_amd.batch import FrameBatch
    s = setup
    mov, _ = T.make_moving(s["clean"], SEEDS[0], s["tmpl"].device)
    boxes = [(slice(4, 60), slice(40, 424), slice(100, 388)),      # 56 x 384 x 288 (level 3: 56 x 48 x 36)
             (slice(0, 40), slice(300, 500), slice(17, 377))]      # 40 x 200 x 360 (level 3: 40 x 25 x 45)
    pairs = [(mov[b].contiguous(), s["tmpl"][b].contiguous()) for b in boxes]
    p = s["params"]
    with FrameBatch(workers=2) as fb:
        results, _ = fb.register_crops(pairs, net=s["net"], iter=Lr.SCHEDULE["iter"], learned_min_depth=0, **p)
        policy, _ = fb.register_crops(pairs, net=s["net"], iter=Lr.SCHEDULE["iter"], **p)
    for (m, t), res in zip(pairs, results):
        single, _ = Lr.register_learned(m, t, s["net"], **Lr.SCHEDULE, **p)
        assert torch.equal(single.motion, res.motion)
    # default policy: the 40-plane crop (< 48) runs the plain pyramid method
    from wholistic_registration_amd.batch import max_layer
    assert torch.equal(policy[0].motion, results[0].motion)
    m, t = pairs[1]
    plain40 = D.get_motion(m, t, layer=max_layer(m.shape, p["r"]), iter=15, **p)
    assert torch.equal(policy[1].motion, plain40.motion) and policy[1].records == plain40.records
    m, t = pairs[0]
    plain = D.get_motion(m, t, layer=4, iter=15, **p).motion
    zero = D.get_motion(m, t, **Lr.SCHEDULE, **p).motion
    e_learned = float(torch.sqrt(((results[0].motion - plain) ** 2).sum(-1)).mean())
    e_zero = float(torch.sqrt(((zero - plain) ** 2).sum(-1)).mean())
    assert e_learned < e_zero, (e_learned, e_zero)


def test_learned_frames_with_fallback_on_concurrent_contexts(setup):
    """FrameBatch(net=..., fallback="plain"): the plain fallback run rebuilds a
    context's template levels for its own layer, so the next frame on that
    context must not reuse them; every frame equals its single run."""
    import torch

    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    from wholistic_registration_amd.batch import FrameBatch
    s = setup
    frames = [T.make_moving(s["clean"], seed, s["tmpl"].device)[0] for seed in (90003, 90004, 90005, 90006)]
    with FrameBatch(workers=2) as fb:
        motion, results = fb.register(frames, s["tmpl"], net=s["net"], fallback="plain", **Lr.SCHEDULE,
                                      **s["params"])
    for f, m, res in zip(frames, motion, results):
        single, _ = Lr.register_learned(f, s["tmpl"], s["net"], ref_level=s["ref_l"], fallback="plain",
                                        **Lr.SCHEDULE, **s["params"])
        assert torch.equal(single.motion, m)
        assert single.records == res.records


# The held-out set of the gate, declared before measuring: templates 90000,
# 91000, 92000, 93000 and their frames +1 .. +10 (the bench's pairs 90001 ..
# 90006 included).  None was seen in training or validation.
GATE_TEMPLATES = (90000, 91000, 92000, 93000)
GATE_FRAMES = 10
GATE_MIN_FRACTION = 0.9


def test_learned_gate_on_declared_held_out_set(setup, cuda):
    """The learned start (network field, layer 0, 5 iterations) on every
    pair of the declared held-out set: at least 90 % of the frames end within
    the 0.1-voxel gate of the plain FP32 field, the mean over frames is
    within it, and the same schedule from zero is worse on average.  Misses
    are printed, not hidden."""
    import torch

    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import learned as Lr
    from wholistic_registration_amd import train_init as T
    net = setup["net"]
    epe, epe0, names = [], [], []
    for tseed in GATE_TEMPLATES:
        s = setup if tseed == 90000 else dict(setup, **_other_template(tseed, cuda))
        p = s["params"]
        for i in range(1, GATE_FRAMES + 1):
            mov, _ = T.make_moving(s["clean"], tseed + i, cuda)
            plain = D.get_motion(mov, s["tmpl"], layer=4, iter=15, **p).motion
            res, _ = Lr.register_learned(mov, s["tmpl"], net, ref_level=s["ref_l"], **Lr.SCHEDULE, **p)
            zero = D.get_motion(mov, s["tmpl"], **Lr.SCHEDULE, **p).motion
            epe.append(_epe(res.motion, plain))
            epe0.append(_epe(zero, plain))
            names.append(tseed + i)
            del mov, plain, zero, res
        torch.cuda.empty_cache()
    epe, epe0 = np.asarray(epe), np.asarray(epe0)
    misses = [(n, round(float(e), 4)) for n, e in zip(names, epe) if e > 0.1]
    print(f"learned start: {int((epe <= 0.1).sum())}/{len(epe)} frames within 0.1 voxel, mean {epe.mean():.4f}, "
          f"max {epe.max():.4f}, misses {misses}; same schedule from zero: mean {epe0.mean():.4f}")
    assert (epe <= 0.1).mean() >= GATE_MIN_FRACTION, misses
    assert epe.mean() <= 0.1
    assert epe.mean() < epe0.mean()
