"""CPU oracle of the learned initialisation network (test infrastructure only).

The reference has no learned component: this restates, in numpy, the layers
the HIP kernels implement (3x3x3 convolution with padding 1, LeakyReLU,
2x2x2 average pooling, nearest upsampling, channel concatenation, mean
squared error, Adam), each from its textbook definition, plus the U-Net that
composes them.  Its warm-start output enters getMotion through option["motion"]
(src/wholistic_registration/utils/calFlow3d_Wei_v1.py:362-377).

Layout: channels-last (B, D, H, W, C); weights (Cout, 27, Cin) with tap
t = (kd * 3 + kh) * 3 + kw at offset (kd - 1, kh - 1, kw - 1).

``bf16`` rounds float32 to bfloat16 (round to nearest even), the storage of
the GPU's activations, weights and gradients; the network oracle can apply it
at the same points as the kernels (``round_bf16=True``) or run exactly in
float64 (for finite-difference checks).

Pinning (tests/test_net_oracle.py): the convolution against
scipy.ndimage.correlate per channel pair; both backward passes by the adjoint
identities <conv(x, w), dy> = <x, dX(dy)> = <w, dW(dy, x)>; pooling and
upsampling by their adjoint identity; the network's gradient by central
finite differences in float64; Adam's first step by its closed form
(w - lr * g / (|g| + eps)).
"""

from __future__ import annotations

import numpy as np

SLOPE = 0.1
PSCALE = 4.0

CHANNELS = (16, 32, 64)


def unet_layers(channels=CHANNELS, cin=8):
    """(name, cin, cout, leaky) of the convolutions of the U-Net with the
    given channels per level (the bottleneck last), in forward order."""
    n = len(channels)
    layers, c = [], cin
    for i in range(n - 1):
        layers += [(f"enc{i + 1}a", c, channels[i], True), (f"enc{i + 1}b", channels[i], channels[i], True)]
        c = channels[i]
    layers += [("bota", c, channels[-1], True), ("botb", channels[-1], channels[-1], True)]
    c = channels[-1]
    for i in range(n - 2, -1, -1):
        layers += [(f"dec{i + 1}a", c + channels[i], channels[i], True), (f"dec{i + 1}b", channels[i], channels[i], True)]
        c = channels[i]
    layers.append(("out", channels[0], 16, False))
    return tuple(layers)


# (name, cin, cout, leaky) of the default network's convolutions, in forward order
LAYERS = unet_layers()


def bf16(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out


def taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))


def conv3d(x, w, bias=None, leaky=False, slope=SLOPE):
    """y[b, v, co] = sum_{t, ci} w[co, t, ci] x[b, v + off(t), ci] (+ bias), float64."""
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    B, D, H, W, _ = x.shape
    xp = _pad(x)
    y = np.zeros((B, D, H, W, w.shape[0]))
    for t, (dd, dh, dw) in enumerate(taps()):
        xs = xp[:, 1 + dd:1 + dd + D, 1 + dh:1 + dh + H, 1 + dw:1 + dw + W, :]
        y += xs @ w[:, t, :].T
    if bias is not None:
        y += np.asarray(bias, dtype=np.float64)
    if leaky:
        y = np.where(y < 0, y * slope, y)
    return y


def conv3d_bwd_data(dy, w):
    """dx[b, u, ci] = sum_{t, co} w[co, t, ci] dy[b, u - off(t), co] (scatter form)."""
    dy = np.asarray(dy, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    B, D, H, W, _ = dy.shape
    dxp = np.zeros((B, D + 2, H + 2, W + 2, w.shape[2]))
    for t, (dd, dh, dw) in enumerate(taps()):
        dxp[:, 1 + dd:1 + dd + D, 1 + dh:1 + dh + H, 1 + dw:1 + dw + W, :] += dy @ w[:, t, :]
    return dxp[:, 1:-1, 1:-1, 1:-1, :]


def conv3d_bwd_weight(dy, x):
    """dw[co, t, ci] = sum_{b, v} dy[b, v, co] x[b, v + off(t), ci]; db = sum dy."""
    dy = np.asarray(dy, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    B, D, H, W, cin = x.shape
    cout = dy.shape[-1]
    xp = _pad(x)
    dw = np.zeros((cout, 27, cin))
    dyf = dy.reshape(-1, cout)
    for t, (dd, dh, dw_) in enumerate(taps()):
        xs = xp[:, 1 + dd:1 + dd + D, 1 + dh:1 + dh + H, 1 + dw_:1 + dw_ + W, :].reshape(-1, cin)
        dw[:, t, :] = dyf.T @ xs
    return dw, dyf.sum(axis=0)


def leaky_bwd(g, y, slope=SLOPE):
    return np.where(np.asarray(y) > 0, g, np.asarray(g) * slope)


def down2(x, scale, dtype=np.float32):
    """scale * sum of the 2x2x2 children, summed in (dz, dy, dx) order in
    ``dtype`` (float32: the kernel's arithmetic)."""
    x = np.asarray(x, dtype=dtype)
    s = np.zeros((x.shape[0], x.shape[1] // 2, x.shape[2] // 2, x.shape[3] // 2, x.shape[4]), dtype=dtype)
    for k in range(8):
        s = s + x[:, (k >> 2)::2, ((k >> 1) & 1)::2, (k & 1)::2, :]
    return s * dtype(scale)


def up2(x, scale, dtype=np.float32):
    x = np.asarray(x, dtype=dtype)
    return np.repeat(np.repeat(np.repeat(x, 2, axis=1), 2, axis=2), 2, axis=3) * dtype(scale)


def mse(p, t, pscale=PSCALE):
    """loss = mean over voxels and the first t.shape[-1] channels of
    (pscale p - t)^2 and its gradient with respect to p (all channels)."""
    p = np.asarray(p, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    ct = t.shape[-1]
    n = t.size
    e = p[..., :ct] * pscale - t
    g = np.zeros_like(p)
    g[..., :ct] = 2.0 * e * pscale / n
    return float((e * e).sum() / n), g


def adam(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0, channels=CHANNELS):
    """He-normal weights (LeakyReLU gain) and zero biases, float32."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, cin, cout, leaky in unet_layers(channels):
        fan_in = 27 * (2 if name == "enc1a" else cin)      # enc1a: 2 live input channels of 8
        std = np.sqrt(2.0 / ((1 + SLOPE ** 2) * fan_in)) if leaky else np.sqrt(1.0 / fan_in)
        w = (rng.standard_normal((cout, 27, cin)) * std).astype(np.float32)
        if name == "out":
            w *= np.float32(0.1)
        params[name] = (w, np.zeros(cout, dtype=np.float32))
    return params


class UNet:
    """Forward and backward of the network, float64 arithmetic; with
    ``round_bf16`` every stored activation, weight and gradient is rounded to
    bfloat16 where the GPU stores it."""

    def __init__(self, params, round_bf16=True, slope=SLOPE):
        self.p = params
        self.r = round_bf16
        self.slope = slope
        ch, i = [], 1
        while f"enc{i}a" in params:
            ch.append(np.asarray(params[f"enc{i}a"][0]).shape[0])
            i += 1
        ch.append(np.asarray(params["bota"][0]).shape[0])
        self.channels = tuple(ch)
        self.leaky = {n: l for n, _, _, l in unet_layers(self.channels)}

    def _q(self, a):
        return bf16(a).astype(np.float64) if self.r else np.asarray(a, dtype=np.float64)

    def _w(self, name):
        w, b = self.p[name]
        return self._q(w), np.asarray(b, dtype=np.float64)

    def _conv(self, name, x):
        w, b = self._w(name)
        return self._q(conv3d(x, w, b, self.leaky[name], self.slope))

    def _pool(self, x):
        return self._q(down2(x, 0.125)) if self.r else down2(x, 0.125, np.float64)

    def _up(self, x):
        return up2(x, 1.0).astype(np.float64) if self.r else up2(x, 1.0, np.float64)

    def forward(self, x):
        n = len(self.channels)
        a = {"x": self._q(x)}
        h = a["x"]
        for i in range(n - 1):
            a[f"enc{i + 1}a"] = self._conv(f"enc{i + 1}a", h)
            a[f"enc{i + 1}b"] = self._conv(f"enc{i + 1}b", a[f"enc{i + 1}a"])
            h = a[f"p{i + 1}"] = self._pool(a[f"enc{i + 1}b"])
        a["bota"] = self._conv("bota", h)
        h = a["botb"] = self._conv("botb", a["bota"])
        for i in range(n - 2, -1, -1):
            a[f"cat{i + 1}"] = np.concatenate([self._up(h), a[f"enc{i + 1}b"]], axis=-1)
            a[f"dec{i + 1}a"] = self._conv(f"dec{i + 1}a", a[f"cat{i + 1}"])
            h = a[f"dec{i + 1}b"] = self._conv(f"dec{i + 1}b", a[f"dec{i + 1}a"])
        a["out"] = self._conv("out", h)
        self.a = a
        return a["out"]

    def backward(self, dout):
        """Gradients {name: (dw, db)} for d loss / d out = dout."""
        a, q, s = self.a, self._q, self.slope
        n = len(self.channels)
        grads = {}

        def conv_back(name, d, x_in, data=True):
            w, _ = self._w(name)
            grads[name] = conv3d_bwd_weight(d, x_in)
            return q(conv3d_bwd_data(d, w)) if data else None

        def down1(g):
            return q(down2(g, 1.0)) if self.r else down2(g, 1.0, np.float64)

        def up_acc(g, add):
            return q(add + up2(g, 0.125)) if self.r else add + up2(g, 0.125, np.float64)

        d = q(dout)
        d = conv_back("out", d, a["dec1b"])
        d = q(leaky_bwd(d, a["dec1b"], s))
        skips = {}
        for i in range(n - 1):
            cu = self.channels[i + 1]
            d = conv_back(f"dec{i + 1}b", d, a[f"dec{i + 1}a"])
            d = q(leaky_bwd(d, a[f"dec{i + 1}a"], s))
            dcat = conv_back(f"dec{i + 1}a", d, a[f"cat{i + 1}"])
            du, skips[i] = dcat[..., :cu], dcat[..., cu:]
            d = down1(du)
            d = q(leaky_bwd(d, a["botb" if i == n - 2 else f"dec{i + 2}b"], s))
        d = conv_back("botb", d, a["bota"])
        d = q(leaky_bwd(d, a["bota"], s))
        dp = conv_back("bota", d, a[f"p{n - 1}"])
        for i in range(n - 2, -1, -1):
            ds = up_acc(dp, skips[i])
            d = q(leaky_bwd(ds, a[f"enc{i + 1}b"], s))
            d = conv_back(f"enc{i + 1}b", d, a[f"enc{i + 1}a"])
            d = q(leaky_bwd(d, a[f"enc{i + 1}a"], s))
            if i > 0:
                dp = conv_back(f"enc{i + 1}a", d, a[f"p{i}"])
            else:
                conv_back("enc1a", d, a["x"], data=False)
        return grads
