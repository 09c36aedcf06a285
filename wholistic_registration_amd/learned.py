"""Learned initialisation of getMotion on MI355X.

A small 3-D U-Net reads the moving and template volumes at pyramid level 3
and predicts the registration field there (in full-resolution voxel units).
That field is the warm start the reference already accepts,
``option["motion"]`` (calFlow3d_Wei_v1.py:362-377): getMotion resamples it
to the coarsest level and the pyramid LK refines it, with fewer iterations
per level than from zero.  The plain pyramid LK stays the yardstick of its
endpoint error and the fallback: ``register_learned(..., fallback="plain")``
also runs the plain method and keeps whichever field leaves the lower final
residual.

Every layer runs in libwrhip.so (``wr_net_*``): 3x3x3 convolutions as
bfloat16 MFMA implicit GEMMs with float32 accumulation (forward, data and
weight gradients), LeakyReLU, 2x2x2 average pooling, nearest upsampling,
concatenation, the squared-error loss and Adam.  Torch tensors only hold
device memory.

Network (channels-last (B, D, H, W, C), D = Z, H = Y / 8, W = X / 8): a 3-D
U-Net of n levels with channels c_0 .. c_{n-1} (CHANNELS; the weights file
fixes them):
  enc_i c_{i-1}->c_i->c_i | pool | ... | bottleneck c_{n-2}->c_{n-1}->c_{n-1} |
  up, concat enc_i | dec_i (c_{i+1}+c_i)->c_i->c_i | ... | out c_0->16
(channels 0..2 of the output: the field / 4).  D, H must be multiples of
2^(n-1) and W of 8 * 2^(n-1).
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from . import device as _dev
from ._util import torch_module as _torch

SLOPE = 0.1
PSCALE = 4.0          # field = PSCALE * network output
LEVEL = 3             # pyramid level the network reads
CHANNELS = (16, 32, 64, 128)     # levels of the shipped network (the bottleneck last)


def unet_layers(channels=CHANNELS, cin=8):
    """(name, cin, cout, leaky) of every 3x3x3 convolution of the U-Net with
    the given channels per level, in forward order, and the level (number of
    halvings of the input grid) each runs at."""
    n = len(channels)
    layers, depth, c = [], {}, cin
    for i in range(n - 1):
        layers += [(f"enc{i + 1}a", c, channels[i], True), (f"enc{i + 1}b", channels[i], channels[i], True)]
        depth[f"enc{i + 1}a"] = depth[f"enc{i + 1}b"] = i
        c = channels[i]
    layers += [("bota", c, channels[-1], True), ("botb", channels[-1], channels[-1], True)]
    depth["bota"] = depth["botb"] = n - 1
    c = channels[-1]
    for i in range(n - 2, -1, -1):
        layers += [(f"dec{i + 1}a", c + channels[i], channels[i], True), (f"dec{i + 1}b", channels[i], channels[i], True)]
        depth[f"dec{i + 1}a"] = depth[f"dec{i + 1}b"] = i
        c = channels[i]
    layers.append(("out", channels[0], 16, False))
    depth["out"] = 0
    return tuple(layers), depth


LAYERS, LAYER_DEPTH = unet_layers(CHANNELS)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "init_net.safetensors")
# getMotion schedule after the network's start: the full-resolution level
# only, 5 iterations (a start at the coarsest level is pulled back to that
# level's own optimum within its first iterations and gains nothing)
SCHEDULE = dict(layer=0, iter=5)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def init_params(seed=0, channels=CHANNELS):
    """He-normal weights (LeakyReLU gain), zero biases; the output layer
    scaled by 0.1 (float32 numpy)."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, cin, cout, leaky in unet_layers(channels)[0]:
        fan_in = 27 * (2 if name == "enc1a" else cin)
        std = np.sqrt(2.0 / ((1 + SLOPE ** 2) * fan_in)) if leaky else np.sqrt(1.0 / fan_in)
        w = (rng.standard_normal((cout, 27, cin)) * std).astype(np.float32)
        if name == "out":
            w *= np.float32(0.1)
        params[name] = (w, np.zeros(cout, dtype=np.float32))
    return params


def channels_of(params):
    """The per-level channels of a parameter set (enc1a, enc2a, ..., bota)."""
    ch, i = [], 1
    while f"enc{i}a" in params:
        ch.append(int(np.asarray(params[f"enc{i}a"][0]).shape[0]))
        i += 1
    ch.append(int(np.asarray(params["bota"][0]).shape[0]))
    return tuple(ch)


class InitNet:
    """Parameters, optimizer state and activations of the network on one device."""

    def __init__(self, params=None, device=0, seed=0, lr=1e-3, channels=CHANNELS):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("wholistic_registration_amd needs a HIP device (none is visible)")
        self.dev = torch.device("cuda", device)
        self.lib = _lib.load()
        # a context of its own: the batched start runs the network on a second
        # stream while the registration's context builds the level images
        self.ctx = _lib.context(device, slot=1)
        if params is None:
            params = init_params(seed, channels)
        self.channels = channels_of(params)
        self.layers, self.depth = unet_layers(self.channels)
        self.spec = {name: (cin, cout, leaky) for name, cin, cout, leaky in self.layers}
        self.w32, self.b32, self.w16, self.dw, self.db = {}, {}, {}, {}, {}
        self.adam_m, self.adam_v = {}, {}
        for name, cin, cout, _ in self.layers:
            w, b = params[name]
            w = np.asarray(w, dtype=np.float32)
            if w.shape != (cout, 27, cin):
                raise ValueError(f"{name}: weight shape {w.shape}, want {(cout, 27, cin)}")
            self.w32[name] = torch.from_numpy(np.ascontiguousarray(w)).to(self.dev)
            self.b32[name] = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float32)).to(self.dev)
            self.w16[name] = torch.empty((cout, 27, cin), dtype=torch.bfloat16, device=self.dev)
            self._check(self.lib.wr_net_to_bf16(self.ctx, _p(self.w32[name]), self.w32[name].numel(),
                                                _p(self.w16[name]), self._s()), "wr_net_to_bf16")
            self.dw[name] = torch.zeros_like(self.w32[name])
            self.db[name] = torch.zeros_like(self.b32[name])
            for st in (self.adam_m, self.adam_v):
                st[name] = (torch.zeros_like(self.w32[name]), torch.zeros_like(self.b32[name]))
        self.lr = lr
        self.step_count = 0
        self.act = {}
        self.shape = None

    # ---------------------------------------------------------------- plumbing

    @staticmethod
    def _s():
        return _dev._stream()

    @staticmethod
    def _check(rc, what):
        _lib.check(rc, what)

    def _buf(self, key, shape):
        torch = _torch()
        t = self.act.get(key)
        if t is None or tuple(t.shape) != tuple(shape):
            t = torch.empty(shape, dtype=torch.bfloat16, device=self.dev)
            self.act[key] = t
        return t

    def _conv(self, name, x, y, dims, batch):
        cin, cout, leaky = self.spec[name]
        d, h, w = dims
        self._check(self.lib.wr_net_conv3d(self.ctx, _p(x), _p(self.w16[name]), _p(self.b32[name]), _p(y), batch,
                                           d, h, w, cin, cout, 1 if leaky else 0, SLOPE, self._s()),
                    "wr_net_conv3d")
        return y

    def grid(self):
        """(dz, dy, dx): the input's D, H and W must be multiples of these."""
        k = 2 ** (len(self.channels) - 1)
        return k, k, 8 * k

    # ---------------------------------------------------------------- forward

    def forward(self, x):
        """x: (B, D, H, W, 8) bfloat16 device tensor -> (B, D, H, W, 16)."""
        B, D, H, W, C = x.shape
        gd, gh, gw = self.grid()
        if C != 8 or D % gd or H % gh or W % gw:
            raise ValueError(f"network input must be (B, D, H, W, 8) with D % {gd}, H % {gh}, W % {gw} == 0, "
                             f"got {tuple(x.shape)}")
        self.shape = (B, D, H, W)
        n = len(self.channels)
        dims = [(D >> i, H >> i, W >> i) for i in range(n)]
        a, L = self.act, self.lib
        a["x"] = x
        h = x
        for i in range(n - 1):               # encoder level i, then 2x2x2 average pooling
            c = self.channels[i]
            h = self._conv(f"enc{i + 1}a", h, self._buf(f"enc{i + 1}a", (B,) + dims[i] + (c,)), dims[i], B)
            h = self._conv(f"enc{i + 1}b", h, self._buf(f"enc{i + 1}b", (B,) + dims[i] + (c,)), dims[i], B)
            p = self._buf(f"p{i + 1}", (B,) + dims[i + 1] + (c,))
            self._check(L.wr_net_down2(self.ctx, _p(h), _p(p), B, *dims[i], c, 0.125, self._s()), "wr_net_down2")
            h = p
        cb = self.channels[-1]
        h = self._conv("bota", h, self._buf("bota", (B,) + dims[-1] + (cb,)), dims[-1], B)
        h = self._conv("botb", h, self._buf("botb", (B,) + dims[-1] + (cb,)), dims[-1], B)
        cu = cb
        for i in range(n - 2, -1, -1):       # decoder level i: upsample, concatenate the skip, two convs
            c = self.channels[i]
            # nearest upsampling of h and the skip connection, concatenated
            cat = self._buf(f"cat{i + 1}", (B,) + dims[i] + (cu + c,))
            self._check(L.wr_net_upcat(self.ctx, _p(h), cu, _p(a[f"enc{i + 1}b"]), c, B, *dims[i], _p(cat),
                                       self._s()), "wr_net_upcat")
            h = self._conv(f"dec{i + 1}a", cat, self._buf(f"dec{i + 1}a", (B,) + dims[i] + (c,)), dims[i], B)
            h = self._conv(f"dec{i + 1}b", h, self._buf(f"dec{i + 1}b", (B,) + dims[i] + (c,)), dims[i], B)
            cu = c
        return self._conv("out", h, self._buf("out", (B,) + dims[0] + (16,)), dims[0], B)

    # ---------------------------------------------------------------- backward

    def zero_grad(self):
        for name, _, _, _ in self.layers:
            self.dw[name].zero_()
            self.db[name].zero_()

    def _wgrad(self, name, d, x_in, dims, batch):
        cin, cout, _ = self.spec[name]
        self._check(self.lib.wr_net_conv3d_bwd_weight(self.ctx, _p(d), _p(x_in), _p(self.dw[name]),
                                                      _p(self.db[name]), batch, *dims, cin, cout, self._s()),
                    "wr_net_conv3d_bwd_weight")

    def _dgrad(self, name, d, out_key, dims, batch):
        cin, cout, _ = self.spec[name]
        dx = self._buf(out_key, (batch,) + tuple(dims) + (cin,))
        self._check(self.lib.wr_net_conv3d_bwd_data(self.ctx, _p(d), _p(self.w16[name]), _p(dx), batch, *dims, cin,
                                                    cout, self._s()), "wr_net_conv3d_bwd_data")
        return dx

    def _leaky(self, g, y):
        self._check(self.lib.wr_net_leaky_bwd(self.ctx, _p(g), _p(y), g.numel(), SLOPE, self._s()),
                    "wr_net_leaky_bwd")
        return g

    def backward(self, dout):
        """Accumulates the weight and bias gradients for d loss / d out = dout
        (the activations of the last forward)."""
        B, D, H, W = self.shape
        n = len(self.channels)
        dims = [(D >> i, H >> i, W >> i) for i in range(n)]
        a, L, s = self.act, self.lib, self._s

        def layer(name, d, x_in, dm, out_key, y_in):
            self._wgrad(name, d, x_in, dm, B)
            g = self._dgrad(name, d, out_key, dm, B)
            return self._leaky(g, y_in) if y_in is not None else g

        d = layer("out", dout, a["dec1b"], dims[0], "g_dec1b", a["dec1b"])
        skips = {}
        for i in range(n - 1):               # decoders, finest first
            c = self.channels[i]
            cu = self.channels[i + 1]
            d = layer(f"dec{i + 1}b", d, a[f"dec{i + 1}a"], dims[i], f"g_dec{i + 1}a", a[f"dec{i + 1}a"])
            dcat = layer(f"dec{i + 1}a", d, a[f"cat{i + 1}"], dims[i], f"g_cat{i + 1}", None)
            nv = B * dims[i][0] * dims[i][1] * dims[i][2]
            du, ds = self._buf(f"g_u{i + 1}", (B,) + dims[i] + (cu,)), self._buf(f"g_s{i + 1}", (B,) + dims[i] + (c,))
            self._check(L.wr_net_split(self.ctx, _p(dcat), cu, c, nv, _p(du), _p(ds), s()), "wr_net_split")
            skips[i] = ds
            up_in = "botb" if i == n - 2 else f"dec{i + 2}b"
            d = self._buf(f"g_{up_in}", (B,) + dims[i + 1] + (cu,))
            self._check(L.wr_net_down2(self.ctx, _p(du), _p(d), B, *dims[i], cu, 1.0, s()), "wr_net_down2")
            self._leaky(d, a[up_in])
        d = layer("botb", d, a["bota"], dims[-1], "g_bota", a["bota"])
        dp = layer("bota", d, a[f"p{n - 1}"], dims[-1], f"g_p{n - 1}", None)
        for i in range(n - 2, -1, -1):       # encoders, coarsest first
            c = self.channels[i]
            ds = skips[i]
            self._check(L.wr_net_up2(self.ctx, _p(dp), _p(ds), B, *dims[i], c, 0.125, 1, s()), "wr_net_up2")
            self._leaky(ds, a[f"enc{i + 1}b"])
            d = layer(f"enc{i + 1}b", ds, a[f"enc{i + 1}a"], dims[i], f"g_enc{i + 1}a", a[f"enc{i + 1}a"])
            if i > 0:
                dp = layer(f"enc{i + 1}a", d, a[f"p{i}"], dims[i], f"g_p{i}", None)
            else:
                self._wgrad("enc1a", d, a["x"], dims[0], B)

    def loss(self, out, target):
        """Mean squared error of PSCALE * out[..., :3] against target (B, D, H,
        W, 3) float32; returns (loss device scalar, d loss / d out)."""
        torch = _torch()
        n = out.numel() // 16
        dout = self._buf("g_out", tuple(out.shape))
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self._check(self.lib.wr_net_mse(self.ctx, _p(out), 16, _p(target), 3, n, PSCALE, 1.0 / (3 * n), _p(dout),
                                        _p(loss), self._s()), "wr_net_mse")
        return loss, dout

    def step(self, lr=None, beta1=0.9, beta2=0.999, eps=1e-8):
        """One Adam step on every parameter; refreshes the bfloat16 weights."""
        self.step_count += 1
        lr = self.lr if lr is None else lr
        for name, _, _, _ in self.layers:
            m, v = self.adam_m[name], self.adam_v[name]
            self._check(self.lib.wr_net_adam(self.ctx, _p(self.w32[name]), _p(self.dw[name]), _p(m[0]), _p(v[0]),
                                             _p(self.w16[name]), self.w32[name].numel(), lr, beta1, beta2, eps,
                                             self.step_count, self._s()), "wr_net_adam")
            self._check(self.lib.wr_net_adam(self.ctx, _p(self.b32[name]), _p(self.db[name]), _p(m[1]), _p(v[1]),
                                             None, self.b32[name].numel(), lr, beta1, beta2, eps, self.step_count,
                                             self._s()), "wr_net_adam")

    # ---------------------------------------------------------------- inference

    def fork(self, context=None):
        """Another handle on the same weights with its own activations (and
        library context): one per host thread that runs the network."""
        other = object.__new__(InitNet)
        other.__dict__.update(self.__dict__)
        other.act, other.shape = {}, None
        if context is not None:
            other.ctx = context
        return other

    def predict(self, mov_l, ref_l, field=None):
        """Warm-start field (Z, Y_L, X_L, 3) float32, full-resolution voxel
        units, from the level-3 images of the moving and template volumes
        (on torch's current stream)."""
        torch = _torch()
        out = self.forward(net_input(mov_l, ref_l, context=self.ctx))
        n = mov_l.numel()
        if field is None:
            field = torch.empty(tuple(mov_l.shape) + (3,), dtype=torch.float32, device=self.dev)
        self._check(self.lib.wr_net_output(self.ctx, _p(out), 16, n, PSCALE, _p(field), self._s()), "wr_net_output")
        return field

    def predict_cp(self, mov_l, ref_l, full_shape, r, level=0, out=None):
        """The predicted field at the control points of pyramid level
        ``level`` of a ``full_shape`` (Z, Y, X) volume, in that level's units:
        (3, Z, ncy, ncx) float32, the ``motion_init_cp`` of getMotion."""
        torch = _torch()
        d, h, w = mov_l.shape
        nz, ny, nx = full_shape
        if d != nz or int(nx / 2 ** LEVEL) != w or int(ny / 2 ** LEVEL) != h:
            raise ValueError(f"level images must be the full volume's level {LEVEL} (x, y / 2^{LEVEL}, same z)")
        if mov_l.shape != ref_l.shape or mov_l.dtype != torch.float32 or ref_l.dtype != torch.float32:
            raise ValueError("level images must be float32 volumes of one shape")
        ncx, ncy = _dev.cp_grid(full_shape, r, level)
        # the network's grid: edge-padded up to multiples of grid()
        gd, gh, gw = self.grid()
        dp, hp, wp = -(-d // gd) * gd, -(-h // gh) * gh, -(-w // gw) * gw
        x = torch.empty((1, dp, hp, wp, 8), dtype=torch.bfloat16, device=self.dev)
        mov_c, ref_c = mov_l.contiguous(), ref_l.contiguous()
        self._check(self.lib.wr_net_input_padded(self.ctx, _p(mov_c), _p(ref_c), d, h, w, dp, hp, wp, _p(x),
                                                 self._s()), "wr_net_input_padded")
        p = self.forward(x)
        if out is None:
            out = torch.empty((3, nz, ncy, ncx), dtype=torch.float32, device=self.dev)
        self._check(self.lib.wr_net_output_cp(self.ctx, _p(p), 16, d, h, w, hp, wp, LEVEL, ncx, ncy, int(r),
                                              int(level), PSCALE, _p(out), self._s()), "wr_net_output_cp")
        return out

    def predict_cp_batch(self, mov_ls, ref_l, full_shape, r, level=0, out=None):
        """``predict_cp`` of B frames in one forward pass of batch B:
        ``mov_ls`` (B, D, H, W) level images of the frames, ``ref_l`` the
        template's; returns (B, 3, Z, ncy, ncx).  Each frame's field equals
        its own ``predict_cp`` (a convolution's sums do not depend on the
        batch)."""
        torch = _torch()
        nb, d, h, w = mov_ls.shape
        nz, ny, nx = full_shape
        if d != nz or int(nx / 2 ** LEVEL) != w or int(ny / 2 ** LEVEL) != h:
            raise ValueError(f"level images must be the full volume's level {LEVEL} (x, y / 2^{LEVEL}, same z)")
        if tuple(ref_l.shape) != (d, h, w) or mov_ls.dtype != torch.float32 or ref_l.dtype != torch.float32:
            raise ValueError("level images must be float32 volumes of one shape")
        ncx, ncy = _dev.cp_grid(full_shape, r, level)
        gd, gh, gw = self.grid()
        dp, hp, wp = -(-d // gd) * gd, -(-h // gh) * gh, -(-w // gw) * gw
        x = torch.empty((nb, dp, hp, wp, 8), dtype=torch.bfloat16, device=self.dev)
        ref_c = ref_l.contiguous()
        mov_c = mov_ls.contiguous()
        self._check(self.lib.wr_net_input_padded_batch(self.ctx, _p(mov_c), nb, _p(ref_c), d, h, w, dp, hp, wp,
                                                       _p(x), self._s()), "wr_net_input_padded_batch")
        p = self.forward(x)
        if out is None:
            out = torch.empty((nb, 3, nz, ncy, ncx), dtype=torch.float32, device=self.dev)
        elif not out.is_contiguous() or tuple(out.shape) != (nb, 3, nz, ncy, ncx):
            raise ValueError("out must be a contiguous (B, 3, Z, ncy, ncx) float32 tensor")
        self._check(self.lib.wr_net_output_cp_batch(self.ctx, _p(p), nb, 16, d, dp, h, w, hp, wp, LEVEL, ncx, ncy,
                                                    int(r), int(level), PSCALE, _p(out), self._s()),
                    "wr_net_output_cp_batch")
        return out

    # ---------------------------------------------------------------- storage

    def params(self):
        return {name: (self.w32[name].cpu().numpy(), self.b32[name].cpu().numpy()) for name, _, _, _ in self.layers}

    def save(self, path):
        from safetensors.numpy import save_file
        tensors = {}
        for name, (w, b) in self.params().items():
            tensors[f"{name}.weight"] = w
            tensors[f"{name}.bias"] = b
        save_file(tensors, path)

    @classmethod
    def load(cls, path=WEIGHTS, device=0):
        from safetensors.numpy import load_file
        t = load_file(path)
        names = sorted({k.rsplit(".", 1)[0] for k in t})
        params = {name: (t[f"{name}.weight"], t[f"{name}.bias"]) for name in names}
        return cls(params, device=device)


def forward_flops(d, h, w, batch=1, channels=CHANNELS):
    """Multiply-add work of one forward pass (2 flops per MAC, the 8 padded
    input channels included): what the matrix cores execute."""
    layers, depth = unet_layers(channels)
    total = 0
    for name, cin, cout, _ in layers:
        k = 2 ** depth[name]
        total += 2 * 27 * cin * cout * (d // k) * (h // k) * (w // k)
    return batch * total


def net_input(mov_l, ref_l, out=None, context=None):
    """(Z, Y, X) float32 level images -> (1, Z, Y, X, 8) bfloat16 network
    input: both normalised by the template's mean and standard deviation."""
    torch = _torch()
    if mov_l.shape != ref_l.shape or mov_l.dtype != torch.float32 or ref_l.dtype != torch.float32:
        raise ValueError("level images must be float32 volumes of one shape")
    if out is None:
        out = torch.empty((1,) + tuple(mov_l.shape) + (8,), dtype=torch.bfloat16, device=mov_l.device)
    lib = _lib.load()
    mov_c, ref_c = mov_l.contiguous(), ref_l.contiguous()      # held until the launch
    ctx = context if context is not None else _lib.context(mov_l.device.index or 0)
    _lib.check(lib.wr_net_input(ctx, _p(mov_c), _p(ref_c), mov_l.numel(), _p(out), _dev._stream()),
               "wr_net_input")
    return out


def level_images(mov, ref, level=LEVEL, ref_level=None, context=None, stream=None):
    """The network's inputs: pyramid level images of mov and ref (float32)."""
    m = _dev.pyramid_level(mov, level, context=context, stream=stream)
    r = ref_level if ref_level is not None else _dev.pyramid_level(ref, level, context=context, stream=stream)
    return m, r


def register_learned(mov, ref, net, *, layer, iter, r, zRatio, movRange, smoothPenalty, tol, fallback=None,
                     plain_layer=4, plain_iter=15, ref_level=None, context=None, stream=None, start="cp", **kw):
    """getMotion warm-started by the network's field.

    Returns ``(result, info)``: ``result`` is a ``device.MotionResult``;
    ``info`` holds ``used`` ("learned" or "plain") and, with
    ``fallback="plain"``, both runs' final residuals.  With ``fallback="plain"``
    the plain pyramid method (from zero, ``plain_layer`` levels and
    ``plain_iter`` iterations; default the 4 levels and 15 iterations the
    network's targets were made with) also runs and its field is kept when
    the learned run's final residual (``currentError`` of level 0) is higher.  ``context`` / ``stream``: the
    library context (``net`` must run on it: ``net.fork(context)``) and torch
    stream to run on (default: the shared context, torch's current stream).
    ``start``: ``"cp"`` hands the network's field to getMotion as the
    coarsest level's control-point field (``motion_init_cp``: no dense start
    field is formed); ``"dense"`` as the reference's warm start
    ``option["motion"]`` at the network's x-y size."""
    torch = _torch()
    params = dict(layer=layer, r=r, zRatio=zRatio, movRange=movRange, smoothPenalty=smoothPenalty, tol=tol,
                  context=context, stream=stream, **kw)
    if start not in ("cp", "dense"):
        raise ValueError("start must be 'cp' or 'dense'")
    main = stream if stream is not None else torch.cuda.current_stream(mov.device)
    reg_ctx = context if context is not None else _lib.context(mov.device.index or 0)
    if start == "cp" and _ctx_value(net.ctx) != _ctx_value(reg_ctx):
        # the network (on a context of its own) runs on a second stream while
        # getMotion builds its level images; the iterations wait for it
        side = _net_stream(mov.device)
        side.wait_stream(main)
        with torch.cuda.stream(side):
            m_l, r_l = level_images(mov, ref, ref_level=ref_level, context=net.ctx)
            cp_init = net.predict_cp(m_l, r_l, tuple(mov.shape), r, level=layer)
            ready = torch.cuda.Event()
            ready.record(side)
        cp_init.record_stream(main)
        init = {"motion_init_cp": cp_init, "init_event": ready}
    else:
        with torch.cuda.stream(main):
            m_l, r_l = level_images(mov, ref, ref_level=ref_level, context=context, stream=stream)
            if start == "cp":
                init = {"motion_init_cp": net.predict_cp(m_l, r_l, tuple(mov.shape), r, level=layer)}
            else:
                init = {"motion_init": net.predict(m_l, r_l)}
    res = _dev.get_motion(mov, ref, iter=iter, **init, **params)
    info = {"used": "learned", "learned_error": res.final_error}
    if fallback == "plain":
        kw_plain = {k: v for k, v in params.items()
                    if k not in ("motion_out", "layer", "reuse_template", "warp_src", "warped_out")}
        if isinstance(kw_plain.get("tile_patches_x"), (list, tuple)):    # per-level widths of the learned run
            kw_plain.pop("tile_patches_x")
        plain = _dev.get_motion(mov, ref, layer=plain_layer, iter=plain_iter, **kw_plain)
        info["plain_error"] = plain.final_error
        if plain.final_error < res.final_error:
            out = params.get("motion_out")
            if out is not None:                      # the caller's buffer holds the kept field
                with torch.cuda.stream(stream if stream is not None else torch.cuda.current_stream(mov.device)):
                    out.copy_(plain.motion)
                plain.motion = out
            if params.get("warp_src") is not None:   # the warp of the kept field
                plain.extras["warped"] = _dev.correct_motion(params["warp_src"], plain.motion,
                                                             out=res.extras["warped"], context=context,
                                                             stream=stream)
            res, info["used"] = plain, "plain"
    elif fallback is not None:
        raise ValueError("fallback must be None or 'plain'")
    return res, info


def register_learned_batch(movs, ref, net, *, layer, iter, r, zRatio, movRange, smoothPenalty, tol, chunk=16,
                           ref_level=None, motion_out=None, warp_src=None, warped_out=None, context=None,
                           stream=None, **kw):
    """``register_learned`` (control-point start, no fallback) of every frame
    of ``movs`` (F, Z, Y, X) against ``ref``, ``chunk`` frames at a time: the
    level images of a chunk, one network pass of batch ``chunk``, then one
    batched launch sequence (``device.get_motion_batch``) warm-started by
    each frame's control-point field.  Each frame's field equals
    ``register_learned(movs[f], ...)``'s.  ``motion_out`` (F, Z, Y, X, 3) and,
    with ``warp_src`` (F, Z, Y, X), ``warped_out`` are filled; returns the
    per-frame ``MotionResult`` list.

    Per chunk, the level images and the network run on a second stream
    (with the network's own library context) while the registration's
    stream builds the level-0 images of the chunk; the iterations wait for
    the network through ``init_event``."""
    torch = _torch()
    nf = int(movs.shape[0])
    full = tuple(int(v) for v in movs.shape[1:])
    if motion_out is None:
        motion_out = torch.empty((nf,) + full + (3,), dtype=torch.float32, device=movs.device)
    if warp_src is not None and warped_out is None:
        warped_out = torch.empty((nf,) + full, dtype=torch.float32, device=movs.device)
    out = []
    main = stream if stream is not None else torch.cuda.current_stream(movs.device)
    # the network's work overlaps the registration's only on a context of its
    # own (two streams must not share one context's workspace); a network
    # forked onto the registration's context runs on the registration's stream
    reg_ctx = context if context is not None else _lib.context(movs.device.index or 0)
    own = _ctx_value(net.ctx) != _ctx_value(reg_ctx)
    side = _net_stream(movs.device) if own else main
    with torch.cuda.stream(main):
        r_l = ref_level if ref_level is not None else _dev.pyramid_level(ref, LEVEL, context=context)
        for f0 in range(0, nf, chunk):
            f1 = min(nf, f0 + chunk)
            side.wait_stream(main)
            with torch.cuda.stream(side):
                # the level images of the chunk in one call: the resampling is
                # x-y only, so F frames of Z planes are one volume of F*Z planes
                chunk_vol = movs[f0:f1].reshape((f1 - f0) * full[0], full[1], full[2])
                m_ls = _dev.pyramid_level(chunk_vol, LEVEL, context=net.ctx)
                m_ls = m_ls.view(f1 - f0, full[0], m_ls.shape[1], m_ls.shape[2])
                cps = net.predict_cp_batch(m_ls, r_l, full, r, level=layer)
                ready = None
                if own:
                    ready = torch.cuda.Event()
                    ready.record(side)
            if own:
                cps.record_stream(main)
            out += _dev.get_motion_batch(movs[f0:f1], ref, layer=layer, iter=iter, r=r, zRatio=zRatio,
                                         movRange=movRange, smoothPenalty=smoothPenalty, tol=tol,
                                         motion_init_cp=cps, init_event=ready, motion_out=motion_out[f0:f1],
                                         warp_src=None if warp_src is None else warp_src[f0:f1],
                                         warped_out=None if warp_src is None else warped_out[f0:f1],
                                         context=context, stream=main, **kw)
    return out


_NET_STREAMS = {}


def _ctx_value(h):
    """The address of a library context handle (an int or a ctypes pointer)."""
    return getattr(h, "value", h)


def _net_stream(device):
    """The stream the batched start runs its network on (default priority)."""
    torch = _torch()
    key = device.index or 0
    if key not in _NET_STREAMS:
        _NET_STREAMS[key] = torch.cuda.Stream(device=device)
    return _NET_STREAMS[key]
