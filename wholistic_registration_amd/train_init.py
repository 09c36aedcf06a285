"""Training of the learned initialisation on generated frame pairs.

A pair is a 64 x 512 x 1024 (z, y, x) uint16 frame and its template: a
template of 100k Gaussian blobs over a background of 100 with Poisson noise,
and a moving frame pull-warped from it by a 6-mode sinusoidal field (max
|d| = 6 voxels) with fresh Poisson noise.  The network reads the pyramid
level 3 images of both (64 x 64 x 128) and learns the registration field at
that level, in full-resolution voxel units:

* ``target="lk"``: the converged field of the plain pyramid LK on the pair
  (layer 4, r = 8, 15 iterations), resampled to level 3 -- the method's own
  result, which the learned start has to reproduce with fewer iterations;
* ``target="truth"``: the generating registration field m = d(p + m).

``python -m wholistic_registration_amd.train_init --out weights.safetensors``
reproduces the shipped weights' recipe on one GPU: the 4-level 16/32/64/128
network (1.47 M parameters) on 20000 pairs (1250 templates x 16 frames; about
8 minutes to generate, the LK targets included), 15k Adam steps of batch 8
(about 3.5 minutes).
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import device as D
from . import learned as Lr
from . import synth
from ._util import torch_module as _torch

FRAME = (64, 512, 1024)
LK_PARAMS = dict(layer=4, iter=15, r=8, zRatio=1.0, movRange=5.0, tol=1e-3)
RAW_PENALTY = 0.08

# Pair recipes: the frame geometry of the time-series frames (the default,
# 100k blobs, 6-mode fields up to 6 voxels, r 8, 15 iterations) and of the
# large single volume (400k blobs, 8-mode fields up to 8 voxels, a 10 %
# ellipsoid mask_ref, r 16, 20 iterations).
RECIPES = {
    "frame_64x512x1024": dict(zyx=(64, 512, 1024), blobs=100_000, modes=6, dmax=6.0, mask_fraction=0.0,
                              lk=dict(LK_PARAMS)),
    "volume_96x1024x2048": dict(zyx=(96, 1024, 2048), blobs=400_000, modes=8, dmax=8.0, mask_fraction=0.10,
                                lk=dict(layer=4, iter=20, r=16, zRatio=1.0, movRange=5.0, tol=1e-3)),
}


def make_template(seed, dev, zyx=FRAME, n_blobs=100_000):
    """Noise-free template and its Poisson-noise uint16 frame."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    clean = synth.render_template_torch(zyx, synth.draw_blobs(rng, zyx, n_blobs), dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    tmpl = torch.poisson(clean, generator=g).round_().clamp_(0, 65535).to(torch.uint16).contiguous()
    return clean, tmpl


def make_mask(seed, zyx, fraction, dev):
    """A random ellipsoid covering ``fraction`` of the voxels (float32, 1 = excluded)."""
    return synth.ellipsoid_mask_torch(np.random.default_rng(seed), zyx, fraction, dev).float().contiguous()


def make_moving(clean, seed, dev, n_modes=6, dmax=6.0):
    """Moving frame (uint16) pull-warped by -d, and d."""
    torch = _torch()
    zyx = tuple(clean.shape)
    d = synth.field_torch(zyx, synth.draw_modes(np.random.default_rng(seed), n_modes), dmax, dev)
    warped = D.correct_motion(clean, (-d).contiguous())
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    mov = torch.poisson(warped.clamp_min(0), generator=g).round_().clamp_(0, 65535).to(torch.uint16).contiguous()
    return mov, d


def truth_field(d, n_iter=25):
    """m with m(p) = d(p + m(p)) by fixed-point iteration (library warp)."""
    torch = _torch()
    comps = [d[..., c].contiguous() for c in range(3)]
    m = d.clone()
    for _ in range(n_iter):
        m = torch.stack([D.correct_motion(comps[c], m) for c in range(3)], dim=-1).contiguous()
    return m


def to_level(field, level=Lr.LEVEL):
    """A full-resolution (Z, Y, X, 3) field resampled (imresize, bicubic) to
    level ``level``'s x-y size, values unchanged (float32)."""
    torch = _torch()
    nz, ny, nx, _ = field.shape
    lx, ly = int(nx / 2 ** level), int(ny / 2 ** level)
    comps = [D.imresize_xy(field[..., c].contiguous(), lx, ly).float() for c in range(3)]
    return torch.stack(comps, dim=-1).contiguous()


def penalty(tmpl):
    return float(D.smooth_penalty_factor(tmpl.float().contiguous())) * RAW_PENALTY


class PairSet:
    """Generated pairs resident on the device in compact form: per pair the
    moving frame's normalised level-3 channel (bfloat16) and the target field
    at level 3 (float16); per template its normalised level-3 channel.  A
    batch is assembled into the network's 8-channel input on the fly, so 20k
    pairs of the frame geometry take about 85 GB."""

    def __init__(self, dev, templates, fields_per_template, seed0, target="lk", log=None, recipe="frame_64x512x1024",
                 depths=None, budget_s=None):
        torch = _torch()
        rc0 = RECIPES[recipe]
        self.dev = dev
        self.mov, self.t, self.tid, self.meta, self.ref = [], [], [], [], []
        # frames of several depths (z extents): each template is made at every depth, blobs
        # in proportion; a batch is drawn from pairs of one shape
        depths = [rc0["zyx"][0]] if not depths else list(depths)
        jobs = [(ti, dz) for ti in range(templates) for dz in depths]
        t_start = time.time()
        for ti, dz in jobs:
            if budget_s is not None and time.time() - t_start > budget_s:
                if log:
                    log(f"[pairs] time budget of {budget_s:.0f} s used: {len(self.mov)} pairs")
                break
            rc = dict(rc0, zyx=(dz,) + tuple(rc0["zyx"][1:]), blobs=int(rc0["blobs"] * dz / rc0["zyx"][0]))
            clean, tmpl = make_template(seed0 + 1000 * ti + 100 * depths.index(dz), dev, zyx=rc["zyx"],
                                        n_blobs=rc["blobs"])
            ref_l = D.pyramid_level(tmpl, Lr.LEVEL)
            mask = make_mask(seed0 + 1000 * ti, rc["zyx"], rc["mask_fraction"], dev) if rc["mask_fraction"] else None
            sp = penalty(tmpl) if mask is None else float(D.smooth_penalty_factor(
                tmpl.float().contiguous(), mask)) * RAW_PENALTY
            k = len(self.ref)
            for fi in range(fields_per_template):
                seed = seed0 + 1000 * ti + 100 * depths.index(dz) + 1 + fi
                mov, d = make_moving(clean, seed, dev, rc["modes"], rc["dmax"])
                if target == "lk":
                    res = D.get_motion(mov, tmpl, smoothPenalty=sp, mask_ref=mask, **rc["lk"])
                    full = res.motion
                elif target == "truth":
                    full = truth_field(d)
                else:
                    raise ValueError("target must be 'lk' or 'truth'")
                x = Lr.net_input(D.pyramid_level(mov, Lr.LEVEL), ref_l)
                if fi == 0:
                    self.ref.append(x[0, ..., 1].clone())
                self.mov.append(x[0, ..., 0].clone())
                self.t.append(to_level(full).to(torch.float16))
                self.tid.append(k)
                self.meta.append(seed)
                del mov, d, full, x
            del clean, tmpl
            if log and (ti + 1) % max(1, templates // 20) == 0:
                log(f"[pairs] template {ti + 1}/{templates} depth {dz}: {len(self.mov)} pairs, "
                    f"{time.time() - t_start:.0f} s")
        torch.cuda.empty_cache()
        self.groups = {}
        for i, m in enumerate(self.mov):
            self.groups.setdefault(tuple(m.shape), []).append(i)

    def __len__(self):
        return len(self.mov)

    def batch(self, idx):
        """(x (B, D, H, W, 8) bfloat16 network input, t (B, D, H, W, 3) float32)."""
        torch = _torch()
        d, h, w = self.mov[idx[0]].shape
        x = torch.zeros((len(idx), d, h, w, 8), dtype=torch.bfloat16, device=self.dev)
        for j, i in enumerate(idx):
            x[j, ..., 0] = self.mov[i]
            x[j, ..., 1] = self.ref[self.tid[i]]
        return x, torch.stack([self.t[i] for i in idx]).float()

    def sample(self, rng, batch):
        """Indices of ``batch`` pairs of one shape (shapes drawn in proportion
        to their pairs)."""
        keys = list(self.groups)
        sizes = np.array([len(self.groups[k]) for k in keys], dtype=np.float64)
        g = self.groups[keys[int(rng.choice(len(keys), p=sizes / sizes.sum()))]]
        return [g[int(i)] for i in rng.choice(len(g), size=min(batch, len(g)), replace=False)]


def train(net, pairs, steps, batch, lr, log=None, log_every=50, warmup=200):
    """Adam over ``steps`` steps of ``batch`` random pairs, the learning rate
    rising linearly over ``warmup`` steps and then decaying on a cosine;
    returns the loss curve (mean per logging window)."""
    rng = np.random.default_rng(0)
    curve, window = [], []
    for k in range(steps):
        x, t = pairs.batch(pairs.sample(rng, batch))
        net.zero_grad()
        out = net.forward(x)
        loss, dout = net.loss(out, t)
        net.backward(dout)
        scale = (k + 1) / warmup if k < warmup else 0.5 * (1.0 + math.cos(math.pi * (k - warmup) / (steps - warmup)))
        net.step(lr=lr * scale)
        window.append(loss)
        if (k + 1) % log_every == 0 or k + 1 == steps:
            v = float(sum(float(w.item()) for w in window) / len(window))
            curve.append((k + 1, v))
            window = []
            if log:
                log(f"[train] step {k + 1}/{steps} loss {v:.5f}")
    return curve


def evaluate(net, pairs):
    """Mean squared error of the network on ``pairs`` (and of the zero field)."""
    torch = _torch()
    tot, zero = 0.0, 0.0
    for i in range(len(pairs)):
        x, t = pairs.batch([i])
        out = net.forward(x)
        loss, _ = net.loss(out, t)
        tot += float(loss.item())
        zero += float((t ** 2).mean().item())
    torch.cuda.synchronize()
    return tot / len(pairs), zero / len(pairs)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True, help="safetensors file for the trained weights")
    # defaults: the shipped weights (weights/init_net.safetensors)
    ap.add_argument("--templates", type=int, default=1250)
    ap.add_argument("--fields", type=int, default=16, help="moving frames per template")
    ap.add_argument("--val-templates", type=int, default=2)
    ap.add_argument("--steps", type=int, default=15000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--target", default="lk", choices=("lk", "truth"))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--recipe", default="frame_64x512x1024", choices=sorted(RECIPES))
    ap.add_argument("--init", default=None, help="safetensors weights to start from (fine-tuning)")
    ap.add_argument("--depths", default=None, help="comma list of frame depths (z extents) to mix")
    ap.add_argument("--report", default=None, help="JSON file for the loss curve and validation")
    ap.add_argument("--channels", default="16,32,64,128", help="channels per U-Net level, the bottleneck last")
    ap.add_argument("--data-budget-s", type=float, default=None,
                    help="stop generating training templates after this many seconds")
    args = ap.parse_args(argv)
    torch = _torch()

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    dev = torch.device("cuda", 0)
    t0 = time.time()
    depths = [int(v) for v in args.depths.split(",")] if args.depths else None
    pairs = PairSet(dev, args.templates, args.fields, 10000, args.target, log, args.recipe, depths,
                    budget_s=args.data_budget_s)
    val = PairSet(dev, args.val_templates, 4, 20000, args.target, log, args.recipe, depths)
    t1 = time.time()
    channels = tuple(int(v) for v in args.channels.split(","))
    net = Lr.InitNet.load(args.init) if args.init else Lr.InitNet(seed=args.seed, lr=args.lr, channels=channels)
    torch.cuda.synchronize()
    t2 = time.time()
    curve = train(net, pairs, args.steps, args.batch, args.lr, log, warmup=args.warmup)
    torch.cuda.synchronize()
    t3 = time.time()
    v_loss, v_zero = evaluate(net, val)
    net.save(args.out)
    rep = {"recipe": args.recipe, "init": args.init, "depths": depths, "pairs": len(pairs), "val_pairs": len(val),
           "target": args.target, "channels": list(net.channels),
           "steps": args.steps,
           "batch": args.batch, "lr": args.lr, "data_s": t1 - t0, "train_s": t3 - t2,
           "ms_per_step": 1e3 * (t3 - t2) / args.steps, "curve": curve,
           "val_mse": v_loss, "val_mse_zero_field": v_zero}
    log(json.dumps({k: v for k, v in rep.items() if k != "curve"}))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()
