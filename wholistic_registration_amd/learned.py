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

Network (channels-last (B, D, H, W, C), D = Z, H = Y / 8, W = X / 8):
  enc1 8->16->16 | pool | enc2 16->32->32 | pool | bottleneck 32->64->64 |
  up, concat enc2 | dec2 96->32->32 | up, concat enc1 | dec1 48->16->16 |
  out 16->16 (channels 0..2: the field / 4)
D, H must be multiples of 4 and W of 32.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from . import device as _dev

SLOPE = 0.1
PSCALE = 4.0          # field = PSCALE * network output
LEVEL = 3             # pyramid level the network reads
LAYERS = (
    ("enc1a", 8, 16, True), ("enc1b", 16, 16, True),
    ("enc2a", 16, 32, True), ("enc2b", 32, 32, True),
    ("bota", 32, 64, True), ("botb", 64, 64, True),
    ("dec2a", 96, 32, True), ("dec2b", 32, 32, True),
    ("dec1a", 48, 16, True), ("dec1b", 16, 16, True),
    ("out", 16, 16, False),
)
# halvings of the input grid each layer runs at (0 = the input's own size)
LAYER_DEPTH = {"enc1a": 0, "enc1b": 0, "enc2a": 1, "enc2b": 1, "bota": 2, "botb": 2, "dec2a": 1, "dec2b": 1,
               "dec1a": 0, "dec1b": 0, "out": 0}
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "init_net.safetensors")
# getMotion schedule after the network's start: the full-resolution level
# only, 5 iterations (a start at the coarsest level is pulled back to that
# level's own optimum within its first iterations and gains nothing)
SCHEDULE = dict(layer=0, iter=5)


def _torch():
    import torch
    return torch


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def init_params(seed=0):
    """He-normal weights (LeakyReLU gain), zero biases; the output layer
    scaled by 0.1 (float32 numpy)."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, cin, cout, leaky in LAYERS:
        fan_in = 27 * (2 if name == "enc1a" else cin)
        std = np.sqrt(2.0 / ((1 + SLOPE ** 2) * fan_in)) if leaky else np.sqrt(1.0 / fan_in)
        w = (rng.standard_normal((cout, 27, cin)) * std).astype(np.float32)
        if name == "out":
            w *= np.float32(0.1)
        params[name] = (w, np.zeros(cout, dtype=np.float32))
    return params


class InitNet:
    """Parameters, optimizer state and activations of the network on one device."""

    def __init__(self, params=None, device=0, seed=0, lr=1e-3):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("wholistic_registration_amd needs a HIP device (none is visible)")
        self.dev = torch.device("cuda", device)
        self.lib = _lib.load()
        self.ctx = _lib.context(device)
        if params is None:
            params = init_params(seed)
        self.w32, self.b32, self.w16, self.dw, self.db = {}, {}, {}, {}, {}
        self.adam_m, self.adam_v = {}, {}
        for name, cin, cout, _ in LAYERS:
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
        _, cin, cout, leaky = next(layer for layer in LAYERS if layer[0] == name)
        d, h, w = dims
        self._check(self.lib.wr_net_conv3d(self.ctx, _p(x), _p(self.w16[name]), _p(self.b32[name]), _p(y), batch,
                                           d, h, w, cin, cout, 1 if leaky else 0, SLOPE, self._s()),
                    "wr_net_conv3d")
        return y

    # ---------------------------------------------------------------- forward

    def forward(self, x):
        """x: (B, D, H, W, 8) bfloat16 device tensor -> (B, D, H, W, 16)."""
        B, D, H, W, C = x.shape
        if C != 8 or D % 4 or H % 4 or W % 32:
            raise ValueError(f"network input must be (B, D, H, W, 8) with D, H % 4 == 0 and W % 32 == 0, "
                             f"got {tuple(x.shape)}")
        self.shape = (B, D, H, W)
        d0, d1, d2 = (D, H, W), (D // 2, H // 2, W // 2), (D // 4, H // 4, W // 4)
        a, L = self.act, self.lib
        a["x"] = x
        self._conv("enc1a", x, self._buf("enc1a", (B,) + d0 + (16,)), d0, B)
        self._conv("enc1b", a["enc1a"], self._buf("enc1b", (B,) + d0 + (16,)), d0, B)
        self._check(L.wr_net_down2(self.ctx, _p(a["enc1b"]), _p(self._buf("p1", (B,) + d1 + (16,))), B, *d0, 16,
                                   0.125, self._s()), "wr_net_down2")
        self._conv("enc2a", a["p1"], self._buf("enc2a", (B,) + d1 + (32,)), d1, B)
        self._conv("enc2b", a["enc2a"], self._buf("enc2b", (B,) + d1 + (32,)), d1, B)
        self._check(L.wr_net_down2(self.ctx, _p(a["enc2b"]), _p(self._buf("p2", (B,) + d2 + (32,))), B, *d1, 32,
                                   0.125, self._s()), "wr_net_down2")
        self._conv("bota", a["p2"], self._buf("bota", (B,) + d2 + (64,)), d2, B)
        self._conv("botb", a["bota"], self._buf("botb", (B,) + d2 + (64,)), d2, B)
        u2 = self._buf("u2", (B,) + d1 + (64,))
        self._check(L.wr_net_up2(self.ctx, _p(a["botb"]), _p(u2), B, *d1, 64, 1.0, 0, self._s()), "wr_net_up2")
        n1 = B * d1[0] * d1[1] * d1[2]
        self._check(L.wr_net_concat(self.ctx, _p(u2), 64, _p(a["enc2b"]), 32, n1,
                                    _p(self._buf("cat2", (B,) + d1 + (96,))), self._s()), "wr_net_concat")
        self._conv("dec2a", a["cat2"], self._buf("dec2a", (B,) + d1 + (32,)), d1, B)
        self._conv("dec2b", a["dec2a"], self._buf("dec2b", (B,) + d1 + (32,)), d1, B)
        u1 = self._buf("u1", (B,) + d0 + (32,))
        self._check(L.wr_net_up2(self.ctx, _p(a["dec2b"]), _p(u1), B, *d0, 32, 1.0, 0, self._s()), "wr_net_up2")
        n0 = B * D * H * W
        self._check(L.wr_net_concat(self.ctx, _p(u1), 32, _p(a["enc1b"]), 16, n0,
                                    _p(self._buf("cat1", (B,) + d0 + (48,))), self._s()), "wr_net_concat")
        self._conv("dec1a", a["cat1"], self._buf("dec1a", (B,) + d0 + (16,)), d0, B)
        self._conv("dec1b", a["dec1a"], self._buf("dec1b", (B,) + d0 + (16,)), d0, B)
        return self._conv("out", a["dec1b"], self._buf("out", (B,) + d0 + (16,)), d0, B)

    # ---------------------------------------------------------------- backward

    def zero_grad(self):
        for name, _, _, _ in LAYERS:
            self.dw[name].zero_()
            self.db[name].zero_()

    def _wgrad(self, name, d, x_in, dims, batch):
        _, cin, cout, _ = next(layer for layer in LAYERS if layer[0] == name)
        self._check(self.lib.wr_net_conv3d_bwd_weight(self.ctx, _p(d), _p(x_in), _p(self.dw[name]),
                                                      _p(self.db[name]), batch, *dims, cin, cout, self._s()),
                    "wr_net_conv3d_bwd_weight")

    def _dgrad(self, name, d, out_key, dims, batch):
        _, cin, cout, _ = next(layer for layer in LAYERS if layer[0] == name)
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
        d0, d1, d2 = (D, H, W), (D // 2, H // 2, W // 2), (D // 4, H // 4, W // 4)
        a, L, s = self.act, self.lib, self._s
        n0, n1 = B * D * H * W, B * d1[0] * d1[1] * d1[2]

        def layer(name, d, x_in, dims, out_key, y_in):
            self._wgrad(name, d, x_in, dims, B)
            g = self._dgrad(name, d, out_key, dims, B)
            return self._leaky(g, y_in) if y_in is not None else g

        d = layer("out", dout, a["dec1b"], d0, "g_dec1b", a["dec1b"])
        d = layer("dec1b", d, a["dec1a"], d0, "g_dec1a", a["dec1a"])
        dcat1 = layer("dec1a", d, a["cat1"], d0, "g_cat1", None)
        du1, ds1 = self._buf("g_u1", (B,) + d0 + (32,)), self._buf("g_s1", (B,) + d0 + (16,))
        self._check(L.wr_net_split(self.ctx, _p(dcat1), 32, 16, n0, _p(du1), _p(ds1), s()), "wr_net_split")
        d = self._buf("g_dec2b", (B,) + d1 + (32,))
        self._check(L.wr_net_down2(self.ctx, _p(du1), _p(d), B, *d0, 32, 1.0, s()), "wr_net_down2")
        self._leaky(d, a["dec2b"])
        d = layer("dec2b", d, a["dec2a"], d1, "g_dec2a", a["dec2a"])
        dcat2 = layer("dec2a", d, a["cat2"], d1, "g_cat2", None)
        du2, ds2 = self._buf("g_u2", (B,) + d1 + (64,)), self._buf("g_s2", (B,) + d1 + (32,))
        self._check(L.wr_net_split(self.ctx, _p(dcat2), 64, 32, n1, _p(du2), _p(ds2), s()), "wr_net_split")
        d = self._buf("g_botb", (B,) + d2 + (64,))
        self._check(L.wr_net_down2(self.ctx, _p(du2), _p(d), B, *d1, 64, 1.0, s()), "wr_net_down2")
        self._leaky(d, a["botb"])
        d = layer("botb", d, a["bota"], d2, "g_bota", a["bota"])
        dp2 = layer("bota", d, a["p2"], d2, "g_p2", None)
        self._check(L.wr_net_up2(self.ctx, _p(dp2), _p(ds2), B, *d1, 32, 0.125, 1, s()), "wr_net_up2")
        self._leaky(ds2, a["enc2b"])
        d = layer("enc2b", ds2, a["enc2a"], d1, "g_enc2a", a["enc2a"])
        dp1 = layer("enc2a", d, a["p1"], d1, "g_p1", None)
        self._check(L.wr_net_up2(self.ctx, _p(dp1), _p(ds1), B, *d0, 16, 0.125, 1, s()), "wr_net_up2")
        self._leaky(ds1, a["enc1b"])
        d = layer("enc1b", ds1, a["enc1a"], d0, "g_enc1a", a["enc1a"])
        self._wgrad("enc1a", d, a["x"], d0, B)

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
        for name, _, _, _ in LAYERS:
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
        # the network's grid: edge-padded up to multiples of (4, 4, 32)
        dp, hp, wp = -(-d // 4) * 4, -(-h // 4) * 4, -(-w // 32) * 32
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

    # ---------------------------------------------------------------- storage

    def params(self):
        return {name: (self.w32[name].cpu().numpy(), self.b32[name].cpu().numpy()) for name, _, _, _ in LAYERS}

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
        params = {name: (t[f"{name}.weight"], t[f"{name}.bias"]) for name, _, _, _ in LAYERS}
        return cls(params, device=device)


def forward_flops(d, h, w, batch=1):
    """Multiply-add work of one forward pass (2 flops per MAC, the 8 padded
    input channels included): what the matrix cores execute."""
    total = 0
    for name, cin, cout, _ in LAYERS:
        k = 2 ** LAYER_DEPTH[name]
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
    with torch.cuda.stream(stream if stream is not None else torch.cuda.current_stream(mov.device)):
        m_l, r_l = level_images(mov, ref, ref_level=ref_level, context=context, stream=stream)
        if start == "cp":
            init = {"motion_init_cp": net.predict_cp(m_l, r_l, tuple(mov.shape), r, level=layer)}
        else:
            init = {"motion_init": net.predict(m_l, r_l)}
    res = _dev.get_motion(mov, ref, iter=iter, **init, **params)
    info = {"used": "learned", "learned_error": res.final_error}
    if fallback == "plain":
        kw_plain = {k: v for k, v in params.items() if k not in ("motion_out", "layer", "reuse_template")}
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
            res, info["used"] = plain, "plain"
    elif fallback is not None:
        raise ValueError("fallback must be None or 'plain'")
    return res, info
