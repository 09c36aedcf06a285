"""Per-layer timing of the learned start's convolutions on the device.

``python -m wholistic_registration_amd.net_layers [D H W]`` runs each 3x3x3
convolution of the network (forward, bf16 channels-last, fused bias and
LeakyReLU) at the grid it has inside the network for a (D, H, W) input
(default 64x64x128: the level-3 image of a 64x512x1024 frame) on random
data, and prints per layer the mean time of ``reps`` launches timed with HIP
events on the launch stream, the matrix-core rate and the algorithmic HBM
bytes (input and output activations once, weights once) with the rate they
imply.  A counter run of the same command (``rocprofv3 --pmc FETCH_SIZE``
or ``WRITE_SIZE``) gives the bytes each launch really moves.
"""

from __future__ import annotations

import argparse
import ctypes

from . import _lib
from .learned import LAYER_DEPTH, LAYERS, SLOPE


def layer_table(d, h, w, reps=20, warmup=3, device=0):
    """[(name, cin, cout, (d, h, w), us, flops, bytes)] for every conv layer."""
    import torch
    lib = _lib.load()
    ctx = _lib.context(device)
    dev = torch.device("cuda", device)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    s = ctypes.c_void_p(stream.cuda_stream)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    rows = []
    for name, cin, cout, act in LAYERS:
        k = 2 ** LAYER_DEPTH[name]
        dd, hh, ww = d // k, h // k, w // k
        x = torch.randn((1, dd, hh, ww, cin), device=dev, generator=gen).to(torch.bfloat16)
        wt = (0.1 * torch.randn((cout, 27, cin), device=dev, generator=gen)).to(torch.bfloat16)
        b = torch.zeros(cout, device=dev)
        y = torch.empty((1, dd, hh, ww, cout), dtype=torch.bfloat16, device=dev)
        args = [ctypes.c_void_p(t.data_ptr()) for t in (x, wt, b, y)]

        def launch():
            _lib.check(lib.wr_net_conv3d(ctx, *args, 1, dd, hh, ww, cin, cout, int(act), SLOPE, s),
                       "wr_net_conv3d")

        for _ in range(warmup):
            launch()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            launch()
        e1.record(stream)
        torch.cuda.synchronize(dev)
        us = e0.elapsed_time(e1) * 1e3 / reps
        vox = dd * hh * ww
        flops = 2 * 27 * cin * cout * vox
        nbytes = 2 * (cin + cout) * vox + 2 * 27 * cin * cout + 4 * cout
        rows.append((name, cin, cout, (dd, hh, ww), us, flops, nbytes))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("shape", nargs="*", type=int, default=[64, 64, 128])
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args(argv)
    if len(a.shape) != 3 or any(v % 4 for v in a.shape[:2]) or a.shape[2] % 32:
        raise SystemExit("shape: D H W with D, H multiples of 4 and W a multiple of 32")
    rows = layer_table(*a.shape, reps=a.reps)
    tot_us = tot_f = 0.0
    for name, cin, cout, (dd, hh, ww), us, fl, nb in rows:
        tot_us += us
        tot_f += fl
        print(f"{name:6s} {cin:3d}->{cout:3d} @ {dd}x{hh}x{ww}: {us:7.1f} us {fl / us / 1e6:7.1f} TFLOP/s "
              f"{nb / 1e6:6.1f} MB {nb / us / 1e3:6.0f} GB/s", flush=True)
    print(f"total {tot_us:.1f} us {tot_f / tot_us / 1e6:.1f} TFLOP/s", flush=True)


if __name__ == "__main__":
    main()
