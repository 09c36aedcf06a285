"""Pins the network oracle (oracle/net_oracle.py) on the CPU with checks that
do not restate its formulas: scipy's correlate, adjoint identities, central
finite differences and Adam's closed-form first step."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import net_oracle as N


def _rand(rng, *shape):
    return rng.standard_normal(shape)


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0, np.inf], dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 0.0, np.inf], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins;
    # 1 + 3 * 2^-8 ties between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert np.array_equal(N.bf16(x), want)
    assert np.isnan(N.bf16(np.array([np.nan], dtype=np.float32)))[0]


def test_conv3d_matches_scipy_correlate():
    from scipy import ndimage as ndi
    rng = np.random.default_rng(0)
    x = _rand(rng, 2, 5, 6, 7, 8)
    w = _rand(rng, 16, 27, 8)
    b = _rand(rng, 16)
    y = N.conv3d(x, w, b)
    for bi in range(2):
        for co in (0, 5, 15):
            want = np.full((5, 6, 7), b[co])
            for ci in range(8):
                k = w[co, :, ci].reshape(3, 3, 3)
                want += ndi.correlate(x[bi, ..., ci], k, mode="constant", cval=0.0)
            np.testing.assert_allclose(y[bi, ..., co], want, rtol=1e-12, atol=1e-12)


def test_conv3d_leaky():
    rng = np.random.default_rng(1)
    x = _rand(rng, 1, 3, 3, 4, 8)
    w = _rand(rng, 16, 27, 8)
    lin = N.conv3d(x, w)
    act = N.conv3d(x, w, leaky=True, slope=0.2)
    assert np.allclose(act, np.where(lin < 0, 0.2 * lin, lin))


def test_conv3d_backward_adjoint_identities():
    rng = np.random.default_rng(2)
    x = _rand(rng, 2, 4, 5, 6, 8)
    w = _rand(rng, 16, 27, 8)
    dy = _rand(rng, 2, 4, 5, 6, 16)
    lhs = float((N.conv3d(x, w) * dy).sum())
    dx = N.conv3d_bwd_data(dy, w)
    dw, db = N.conv3d_bwd_weight(dy, x)
    assert lhs == pytest.approx(float((x * dx).sum()), rel=1e-12)
    assert lhs == pytest.approx(float((w * dw).sum()), rel=1e-12)
    assert np.allclose(db, dy.reshape(-1, 16).sum(0))


def test_conv3d_bwd_data_is_flipped_conv():
    """The GPU computes dX as a forward convolution with w'[ci][t][co] =
    w[co][26 - t][ci]; the oracle's scatter form must agree with it."""
    rng = np.random.default_rng(3)
    dy = _rand(rng, 1, 4, 4, 5, 16)
    w = _rand(rng, 16, 27, 32)
    wf = np.transpose(w[:, ::-1, :], (2, 1, 0))
    assert np.allclose(N.conv3d_bwd_data(dy, w), N.conv3d(dy, wf), atol=1e-10)


def test_pool_up_adjoint():
    rng = np.random.default_rng(4)
    x = _rand(rng, 2, 4, 6, 8, 8).astype(np.float32)
    y = _rand(rng, 2, 2, 3, 4, 8).astype(np.float32)
    lhs = float((N.down2(x, 0.125).astype(np.float64) * y).sum())
    rhs = float((x.astype(np.float64) * N.up2(y, 0.125)).sum())
    assert lhs == pytest.approx(rhs, rel=1e-5)
    assert np.allclose(N.down2(np.ones((1, 2, 2, 2, 8), np.float32), 0.125), 1.0)


def test_mse_gradient_finite_difference():
    rng = np.random.default_rng(5)
    p = _rand(rng, 1, 2, 3, 4, 16)
    t = _rand(rng, 1, 2, 3, 4, 3)
    loss, g = N.mse(p, t)
    assert np.all(g[..., 3:] == 0)
    for idx in [(0, 0, 0, 0, 0), (0, 1, 2, 3, 2), (0, 1, 1, 1, 1)]:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[idx] += e
        pm[idx] -= e
        fd = (N.mse(pp, t)[0] - N.mse(pm, t)[0]) / (2 * e)
        assert g[idx] == pytest.approx(fd, rel=1e-6, abs=1e-10)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(6)
    w = _rand(rng, 100)
    g = _rand(rng, 100)
    w1, m, v = N.adam(w, g, np.zeros(100), np.zeros(100), 1, lr=1e-2, eps=1e-8)
    assert np.allclose(w1, w - 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-10, atol=1e-14)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


@pytest.mark.parametrize("channels,shape", [((16, 32, 64), (1, 4, 4, 8, 8)), ((16, 32, 64, 128), (1, 8, 8, 16, 8))])
def test_unet_gradient_finite_difference(channels, shape):
    """Whole-network backward (float64, no rounding) against central
    differences of the loss, on a few weights of every layer, for the
    3-level and the 4-level network."""
    params = N.init_params(seed=7, channels=channels)
    net = N.UNet(params, round_bf16=False)
    rng = np.random.default_rng(8)
    x = rng.standard_normal(shape)
    x[..., 2:] = 0.0
    t = rng.standard_normal(shape[:4] + (3,))
    loss, dout = N.mse(net.forward(x), t)
    grads = net.backward(dout)
    e = 1e-6          # small against the distance of pre-activations to LeakyReLU's kink
    for name, _, _, _ in N.unet_layers(channels):
        w, b = params[name]
        dw, db = grads[name]
        for j in range(2):
            idx = tuple(int(rng.integers(0, s)) for s in w.shape)
            if name == "enc1a":
                idx = idx[:2] + (int(rng.integers(0, 2)),)
            w64 = w.astype(np.float64)
            wp, wm = w64.copy(), w64.copy()
            wp[idx] += e
            wm[idx] -= e
            params[name] = (wp, b)
            lp = N.mse(net.forward(x), t)[0]
            params[name] = (wm, b)
            lm = N.mse(net.forward(x), t)[0]
            params[name] = (w, b)
            fd = (lp - lm) / (2 * e)
            assert dw[idx] == pytest.approx(fd, rel=1e-3, abs=1e-9), (name, idx)
        co = int(rng.integers(0, b.shape[0]))
        bp, bm = b.astype(np.float64).copy(), b.astype(np.float64).copy()
        bp[co] += e
        bm[co] -= e
        params[name] = (w, bp)
        lp = N.mse(net.forward(x), t)[0]
        params[name] = (w, bm)
        lm = N.mse(net.forward(x), t)[0]
        params[name] = (w, b)
        assert db[co] == pytest.approx((lp - lm) / (2 * e), rel=1e-3, abs=1e-9), (name, "bias")
