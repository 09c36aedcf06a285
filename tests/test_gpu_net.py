"""The learned initialisation's kernels (called through the C ABI) against the
network oracle, the whole network's gradient, training, and the warm start
at the network's x-y size.

Tolerances: inputs are bfloat16 values; a convolution output is the float32
sum of exact products rounded once to bfloat16, so it lies within one
bfloat16 step (2^-8 relative) of the oracle's float64 sum rounded the same
way, plus the float32 summation error (1e-6 of the sum of |terms|); float32 weight gradients within 1e-5 of the sum of |terms|; pooling,
upsampling, concatenation, split and the LeakyReLU backward are bit-exact.
The whole network (11 convolutions, bfloat16 storage between layers) is
compared by relative Frobenius norm: 2e-2 on every weight gradient.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import net_oracle as N

pytestmark = pytest.mark.gpu


def _lib():
    from wholistic_registration_amd import _lib as L
    return L


def _ctx():
    return _lib().context(0)


def _s():
    from wholistic_registration_amd import device as D
    return D._stream()


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _bf(a, cuda):
    import torch
    return torch.from_numpy(np.ascontiguousarray(N.bf16(a))).to(cuda).to(torch.bfloat16)


def _np(t):
    import torch
    return t.to(torch.float32).cpu().numpy()


def _within_bf16_step(got, want, mag=None):
    """|got - bf16(want)| within one bfloat16 step of want, plus float32
    accumulation error (1e-6 of ``mag``, the sum of |terms|)."""
    want_r = N.bf16(want)
    step = np.abs(want_r) * 2.0 ** -7 + 1e-30 + (0.0 if mag is None else 1e-6 * mag)
    bad = np.abs(got - want_r) > step
    assert not bad.any(), (np.argwhere(bad)[:5], got[bad][:5], want_r[bad][:5])


@pytest.mark.parametrize("shape,cin,cout,leaky", [
    ((2, 4, 5, 12), 8, 16, True),       # 240 voxels: partial 64-voxel tiles
    ((1, 3, 4, 16), 48, 32, True),      # 27 * 48 = 1296: partial 32-wide k step
    ((2, 2, 3, 8), 96, 64, False),
    ((1, 2, 2, 8), 128, 128, True),     # the bottleneck's channels on a tiny volume (one-tile waves)
])
def test_conv_forward_matches_oracle(cuda, shape, cin, cout, leaky):
    import torch
    rng = np.random.default_rng(cin * cout)
    B, D, H, W = shape
    x = N.bf16(rng.standard_normal((B, D, H, W, cin)))
    w = N.bf16(rng.standard_normal((cout, 27, cin)) * 0.2)
    b = rng.standard_normal(cout).astype(np.float32)
    y = torch.empty((B, D, H, W, cout), dtype=torch.bfloat16, device=cuda)
    bt = torch.from_numpy(b).to(cuda)
    xt, wt = _bf(x, cuda), _bf(w, cuda)          # held: the kernel runs after these lines
    rc = _lib().load().wr_net_conv3d(_ctx(), _p(xt), _p(wt), _p(bt), _p(y), B, D, H, W, cin,
                                     cout, int(leaky), N.SLOPE, _s())
    _lib().check(rc, "wr_net_conv3d")
    want = N.conv3d(x, w, b, leaky, N.SLOPE)
    _within_bf16_step(_np(y), want, N.conv3d(np.abs(x), np.abs(w), np.abs(b)))


@pytest.mark.parametrize("shape,cin,cout", [((2, 3, 4, 8), 16, 32), ((1, 4, 4, 8), 48, 16)])
def test_conv_backward_data_matches_oracle(cuda, shape, cin, cout):
    import torch
    rng = np.random.default_rng(cin + cout)
    B, D, H, W = shape
    dy = N.bf16(rng.standard_normal((B, D, H, W, cout)))
    w = N.bf16(rng.standard_normal((cout, 27, cin)) * 0.2)
    dx = torch.empty((B, D, H, W, cin), dtype=torch.bfloat16, device=cuda)
    dyt, wt = _bf(dy, cuda), _bf(w, cuda)
    rc = _lib().load().wr_net_conv3d_bwd_data(_ctx(), _p(dyt), _p(wt), _p(dx), B, D, H, W,
                                              cin, cout, _s())
    _lib().check(rc, "wr_net_conv3d_bwd_data")
    _within_bf16_step(_np(dx), N.conv3d_bwd_data(dy, w), N.conv3d_bwd_data(np.abs(dy), np.abs(w)))


@pytest.mark.parametrize("shape,cin,cout", [((2, 4, 5, 16), 8, 16), ((1, 3, 4, 24), 48, 32),
                                            ((3, 2, 2, 8), 16, 64)])
def test_conv_backward_weight_matches_oracle(cuda, shape, cin, cout):
    import torch
    rng = np.random.default_rng(7 * cin + cout)
    B, D, H, W = shape
    dy = N.bf16(rng.standard_normal((B, D, H, W, cout)))
    x = N.bf16(rng.standard_normal((B, D, H, W, cin)))
    dw = torch.full((cout, 27, cin), 0.5, dtype=torch.float32, device=cuda)      # accumulates
    db = torch.full((cout,), -1.0, dtype=torch.float32, device=cuda)
    dyt, xt = _bf(dy, cuda), _bf(x, cuda)
    rc = _lib().load().wr_net_conv3d_bwd_weight(_ctx(), _p(dyt), _p(xt), _p(dw), _p(db), B, D,
                                                H, W, cin, cout, _s())
    _lib().check(rc, "wr_net_conv3d_bwd_weight")
    want_w, want_b = N.conv3d_bwd_weight(dy, x)
    mag_w, _ = N.conv3d_bwd_weight(np.abs(dy), np.abs(x))
    assert np.all(np.abs(dw.cpu().numpy() - 0.5 - want_w) <= 1e-5 * mag_w + 1e-6)
    np.testing.assert_allclose(db.cpu().numpy() + 1.0, want_b, rtol=1e-5, atol=1e-4)


def test_pool_upsample_concat_split_leaky_bit_exact(cuda):
    import torch
    L = _lib().load()
    rng = np.random.default_rng(11)
    B, D, H, W = 2, 4, 6, 8
    x = N.bf16(rng.standard_normal((B, D, H, W, 16)))
    xt = _bf(x, cuda)
    y = torch.empty((B, D // 2, H // 2, W // 2, 16), dtype=torch.bfloat16, device=cuda)
    _lib().check(L.wr_net_down2(_ctx(), _p(xt), _p(y), B, D, H, W, 16, 0.125, _s()), "down2")
    assert np.array_equal(_np(y), N.bf16(N.down2(x, 0.125)))
    acc = N.bf16(rng.standard_normal((B, D, H, W, 16)))
    at = _bf(acc, cuda)
    _lib().check(L.wr_net_up2(_ctx(), _p(y), _p(at), B, D, H, W, 16, 0.125, 1, _s()), "up2")
    assert np.array_equal(_np(at), N.bf16(acc + N.up2(_np(y), 0.125)))
    u = torch.empty_like(xt)
    _lib().check(L.wr_net_up2(_ctx(), _p(y), _p(u), B, D, H, W, 16, 1.0, 0, _s()), "up2")
    assert np.array_equal(_np(u), N.up2(_np(y), 1.0))
    b8 = _bf(rng.standard_normal((B, D, H, W, 8)), cuda)
    cat = torch.empty((B, D, H, W, 24), dtype=torch.bfloat16, device=cuda)
    n = B * D * H * W
    _lib().check(L.wr_net_concat(_ctx(), _p(xt), 16, _p(b8), 8, n, _p(cat), _s()), "concat")
    assert torch.equal(cat, torch.cat([xt, b8], dim=-1))
    da, db = torch.empty_like(xt), torch.empty_like(b8)
    _lib().check(L.wr_net_split(_ctx(), _p(cat), 16, 8, n, _p(da), _p(db), _s()), "split")
    assert torch.equal(da, xt) and torch.equal(db, b8)
    g = _bf(rng.standard_normal((B, D, H, W, 16)), cuda)
    want = N.bf16(np.where(x > 0, _np(g), _np(g) * np.float32(N.SLOPE)))
    _lib().check(L.wr_net_leaky_bwd(_ctx(), _p(g), _p(xt), g.numel(), N.SLOPE, _s()), "leaky")
    assert np.array_equal(_np(g), want)


def test_mse_and_adam_match_oracle(cuda):
    import torch
    L = _lib().load()
    rng = np.random.default_rng(12)
    p = N.bf16(rng.standard_normal((2, 3, 4, 8, 16)))
    t = rng.standard_normal((2, 3, 4, 8, 3)).astype(np.float32)
    n = 2 * 3 * 4 * 8
    dp = torch.empty((2, 3, 4, 8, 16), dtype=torch.bfloat16, device=cuda)
    loss = torch.zeros(1, dtype=torch.float32, device=cuda)
    pt, tt = _bf(p, cuda), torch.from_numpy(t).to(cuda)
    _lib().check(L.wr_net_mse(_ctx(), _p(pt), 16, _p(tt), 3, n, N.PSCALE,
                              1.0 / (3 * n), _p(dp), _p(loss), _s()), "mse")
    want_loss, want_g = N.mse(p, t)
    assert loss.item() == pytest.approx(want_loss, rel=1e-5)
    _within_bf16_step(_np(dp), want_g)
    w = rng.standard_normal(1000).astype(np.float32)
    m = np.zeros(1000, np.float32)
    v = np.zeros(1000, np.float32)
    wt, mt, vt = (torch.from_numpy(a.copy()).to(cuda) for a in (w, m, v))
    w16 = torch.empty(1000, dtype=torch.bfloat16, device=cuda)
    for step in range(1, 4):
        g = rng.standard_normal(1000).astype(np.float32)
        gt = torch.from_numpy(g).to(cuda)
        _lib().check(L.wr_net_adam(_ctx(), _p(wt), _p(gt), _p(mt), _p(vt), _p(w16), 1000,
                                   1e-2, 0.9, 0.999, 1e-8, step, _s()), "adam")
        w, m, v = N.adam(w.astype(np.float64), g, m, v, step, 1e-2)
    np.testing.assert_allclose(wt.cpu().numpy(), w, rtol=1e-5, atol=1e-6)
    assert np.array_equal(_np(w16), N.bf16(wt.cpu().numpy()))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("channels,dhw", [((16, 32, 64), (4, 8, 32)), ((16, 32, 64, 128), (16, 16, 64))])
def test_network_gradient_matches_oracle(cuda, channels, dhw):
    import torch

    from wholistic_registration_amd.learned import InitNet
    params = N.init_params(seed=3, channels=channels)
    rng = np.random.default_rng(4)
    for name in params:                       # nonzero biases exercise every term
        w, b = params[name]
        params[name] = (w, (0.05 * rng.standard_normal(b.shape)).astype(np.float32))
    B, (D, H, W) = 2, dhw
    x = np.zeros((B, D, H, W, 8), np.float32)
    x[..., :2] = rng.standard_normal((B, D, H, W, 2))
    x = N.bf16(x)
    t = rng.standard_normal((B, D, H, W, 3)).astype(np.float32)
    net = InitNet(params)
    out = net.forward(_bf(x, cuda))
    ref = N.UNet(params, round_bf16=True)
    want_out = ref.forward(x)
    assert _rel(_np(out), want_out) <= 1e-2
    net.zero_grad()
    loss, dout = net.loss(out, torch.from_numpy(t).to(cuda))
    want_loss, want_dout = N.mse(want_out, t)
    assert loss.item() == pytest.approx(want_loss, rel=1e-2)
    net.backward(dout)
    grads = ref.backward(want_dout)
    errs = {}
    for name, _, _, _ in N.unet_layers(channels):
        gw, gb = net.dw[name].cpu().numpy(), net.db[name].cpu().numpy()
        ww, wb = grads[name]
        if name == "enc1a":
            gw, ww = gw[..., :2], ww[..., :2]
        errs[name] = (round(_rel(gw, ww), 4), round(_rel(gb, wb), 4))
    print(f"channels {channels}: relative Frobenius errors (weight, bias) {errs}")
    # bfloat16 storage points are the same in both; the float32 accumulation
    # order of the matrix cores differs from the oracle's float64 sums, and
    # the difference grows toward the bottleneck: measured 0.2 % (3 levels)
    # and up to 3 % (4 levels, 2 x 2 x 8 bottleneck grid)
    tol = 2e-2 if len(channels) == 3 else 4e-2
    assert all(max(e) <= tol for e in errs.values()), errs


def test_training_reduces_loss(cuda):
    import torch

    from wholistic_registration_amd.learned import InitNet
    rng = np.random.default_rng(5)
    B, D, H, W = 2, 8, 8, 32
    x = np.zeros((B, D, H, W, 8), np.float32)
    x[..., :2] = rng.standard_normal((B, D, H, W, 2))
    xt = _bf(x, cuda)
    # target: a smooth function of the inputs
    t = np.stack([x[..., 0] - x[..., 1], 0.5 * x[..., 0], np.full(x.shape[:4], 0.25)], -1).astype(np.float32)
    tt = torch.from_numpy(t).to(cuda)
    net = InitNet(seed=6, lr=2e-3, channels=(16, 32, 64))
    losses = []
    for _ in range(100):
        net.zero_grad()
        out = net.forward(xt)
        loss, dout = net.loss(out, tt)
        net.backward(dout)
        net.step()
        losses.append(loss.item())
    assert all(np.isfinite(losses))
    assert losses[-1] < 0.35 * losses[0], losses[::10]


def test_low_resolution_warm_start_matches_oracle(cuda):
    """motion_init at another x-y size (init_nx, init_ny): the reference's
    warm start imresizes it to the coarsest level (calFlow3d_Wei_v1.py:362-377)."""
    import torch

    from oracle import lk_oracle as O
    from tests.golden.cases import small_inputs
    from wholistic_registration_amd import device as D
    mov, ref, opt = small_inputs("plain")
    nx, ny, nz = mov.shape
    rng = np.random.default_rng(9)
    init = rng.normal(scale=0.7, size=(nx // 4, ny // 4, nz, 3)).astype(np.float32)
    opt = dict(opt)
    opt["motion"] = init
    ex = {}
    want, err, _, _ = O.get_motion(mov, ref, dict(opt), extras=ex, keep_images=False)
    init_t = torch.from_numpy(np.ascontiguousarray(init.transpose(2, 1, 0, 3))).to(cuda)
    zyx = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0))).to(cuda)
    res = D.get_motion(zyx(mov), zyx(ref), layer=opt["layer"], iter=opt["iter"], r=opt["r"], zRatio=opt["zRatio"],
                       movRange=opt["movRange"], smoothPenalty=opt["smoothPenalty"], tol=opt["tol"],
                       motion_init=init_t, schedule=ex["schedule"])
    got = res.motion.cpu().numpy().transpose(2, 1, 0, 3)
    assert np.abs(got - want).max() <= 1e-3
    assert res.final_error == pytest.approx(err, rel=2e-5)


def test_entry_points_reject_bad_shapes(cuda):
    """Shape and alignment rules are errors, not faults."""
    import torch
    L = _lib().load()
    x = torch.zeros((1, 4, 4, 8, 16), dtype=torch.bfloat16, device=cuda)
    w = torch.zeros((16, 27, 16), dtype=torch.bfloat16, device=cuda)
    y = torch.zeros_like(x)
    cases = [
        L.wr_net_conv3d(_ctx(), _p(x), _p(w), None, _p(y), 1, 4, 4, 8, 12, 16, 1, 0.1, _s()),   # cin % 8
        L.wr_net_conv3d(_ctx(), _p(x), _p(w), None, _p(y), 1, 4, 4, 8, 16, 24, 1, 0.1, _s()),   # cout % 16
        L.wr_net_conv3d(_ctx(), ctypes.c_void_p(x.data_ptr() + 2), _p(w), None, _p(y), 1, 4, 4, 8, 16, 16, 1,
                        0.1, _s()),                                                          # alignment
        L.wr_net_conv3d(_ctx(), _p(x), _p(w), None, _p(y), 0, 4, 4, 8, 16, 16, 1, 0.1, _s()),    # empty batch
        L.wr_net_conv3d_bwd_data(_ctx(), _p(x), _p(w), _p(y), 1, 4, 4, 8, 8, 16, _s()),          # cin % 16
        L.wr_net_conv3d_bwd_weight(_ctx(), _p(x), _p(x), _p(y), None, 1, 4, 4, 12, 16, 16, _s()),  # w % 8
        L.wr_net_down2(_ctx(), _p(x), _p(y), 1, 4, 4, 7, 16, 0.125, _s()),                        # odd extent
        L.wr_net_up2(_ctx(), _p(x), _p(y), 1, 3, 4, 8, 16, 1.0, 0, _s()),                          # odd extent
        L.wr_net_leaky_bwd(_ctx(), _p(x), _p(y), 12, 0.1, _s()),                                   # n % 8
        L.wr_net_mse(_ctx(), _p(x), 2, _p(x), 3, 16, 4.0, 1.0, _p(y), _p(y), _s()),                # ct > cp
        L.wr_net_adam(_ctx(), _p(y), _p(y), _p(y), _p(y), None, 16, 1e-3, 0.9, 0.999, 1e-8, 0, _s()),  # step 0
    ]
    assert all(rc == _lib().WR_E_ARG for rc in cases), cases


def test_cp_start_equals_dense_start_of_its_interpolant(cuda):
    """motion_init_cp (wr_params.init_cp): a level started from a CP field
    behaves as the reference's dense warm start with that field's bilinear
    interpolant; the interpolant equals the CP values at the control points."""
    import torch

    from tests.golden.cases import small_inputs
    from wholistic_registration_amd import device as D
    mov, ref, opt = small_inputs("plain")
    zyx = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).transpose(2, 1, 0))).to(cuda)
    mt, rt = zyx(mov), zyx(ref)
    nz, ny, nx = mt.shape
    r = opt["r"]
    P = dict(r=r, zRatio=opt["zRatio"], movRange=opt["movRange"], smoothPenalty=opt["smoothPenalty"], tol=opt["tol"])
    ncx, ncy = D.cp_grid((nz, ny, nx), r)
    rng = np.random.default_rng(13)
    cp = torch.from_numpy((0.6 * rng.standard_normal((3, nz, ncy, ncx))).astype(np.float32)).to(cuda)
    start = D.get_motion(mt, rt, layer=0, iter=1, motion_init_cp=cp, **P)
    assert start.updates[0] == 0
    p = 2 * r + 1
    at_cp = start.motion[:, r:r + p * ncy:p, r:r + p * ncx:p, :].permute(3, 0, 1, 2)
    assert torch.equal(at_cp, cp)
    dense = D.get_motion(mt, rt, layer=0, iter=4, motion_init=start.motion.clone(), **P)
    via_cp = D.get_motion(mt, rt, layer=0, iter=4, motion_init_cp=cp, **P)
    assert dense.records == via_cp.records
    assert float((dense.motion - via_cp.motion).abs().max()) <= 1e-4
    with pytest.raises(ValueError):
        D.get_motion(mt, rt, layer=0, iter=2, motion_init_cp=cp[:, :, :-1].contiguous(), **P)
