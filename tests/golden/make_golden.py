"""Write the golden fixtures of the registration path by running the reference.

Run in the development container only (the reference tree is not present on
GPU machines):

    python tests/golden/make_golden.py [--reference /root/reference] [--skip-64x128x128] [--only-r16]

It imports the reference's hot-path modules
(``src/wholistic_registration/utils/{calFlow3d_Wei_v1,imresize,interp,
calculate}.py``) through an empty parent package, because the package's own
``__init__`` pulls in I/O dependencies that the hot path does not need.  Only
the two functions whose exact behaviour nothing else pins are exercised:

* ``imresize`` (edge mirroring, zero-column trimming, axis order and float64
  table arithmetic in combination, imresize.py:60-191);
* ``getMotion`` end to end (iteration schedule, stop rule and final field,
  calFlow3d_Wei_v1.py:295-673).

Outputs are small ``.npz`` files next to this script: inputs (or the seed and a
hash that regenerates them) and the reference's outputs.  No reference source
is copied.
"""

from __future__ import annotations

import argparse
import hashlib
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from wholistic_registration_amd import synth  # noqa: E402

from tests.golden.cases import (  # noqa: E402
    BLOBS_64x128x128, IMRESIZE_CASES, R16_CASES, SMALL_CASES, inputs_64x128x128, imresize_input, small_inputs,
)


def import_reference(root):
    src = os.path.join(root, "src", "wholistic_registration")
    if "wholistic_registration" not in sys.modules:
        pkg = types.ModuleType("wholistic_registration")
        pkg.__path__ = [src]
        sys.modules["wholistic_registration"] = pkg
    import importlib
    cf = importlib.import_module("wholistic_registration.utils.calFlow3d_Wei_v1")
    im = importlib.import_module("wholistic_registration.utils.imresize")
    return cf, im


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def log_arrays(error_log, levels):
    out = {}
    for L in range(levels, -1, -1):
        e = error_log[f"layer_{L}"]
        for key in ("diffError", "penaltyError", "currentError", "max_diff_motion"):
            out[f"L{L}_{key}"] = np.asarray(e[key], dtype=np.float64)
    return out


def write_imresize(im):
    out = {}
    for i, (shape_in, shape_out, method, dtype) in enumerate(IMRESIZE_CASES):
        a = imresize_input(i)
        b = im.imresize(a, output_shape=shape_out, method=method)
        out[f"in_{i}"] = a
        out[f"out_{i}"] = np.asarray(b)
    np.savez_compressed(os.path.join(HERE, "imresize_cases.npz"), **out)
    print("imresize cases:", len(IMRESIZE_CASES))


def run_reference(cf, mov, ref, option):
    opt = dict(option)
    motion, err, _coords, log = cf.getMotion(mov, ref, opt)
    return np.asarray(motion, dtype=np.float32), float(err), log


def write_small(cf):
    for name, spec in SMALL_CASES.items():
        mov, ref, option = small_inputs(name)
        motion, err, log = run_reference(cf, mov, ref, option)
        payload = dict(mov=mov, ref=ref, motion=motion, current_error=np.float64(err),
                       **log_arrays(log, option["layer"]))
        if option.get("motion") is not None:
            payload["motion_init"] = option["motion"]
        mr = np.asarray(option["mask_ref"])
        mm = np.asarray(option["mask_mov"])
        payload["mask_ref"] = mr
        payload["mask_mov"] = mm
        np.savez_compressed(os.path.join(HERE, f"getmotion_{name}.npz"), **payload)
        print(name, "levels", option["layer"], "err", err,
              "records", [len(log[f"layer_{L}"]["currentError"])
                          for L in range(option["layer"], -1, -1)])


def write_64x128x128(cf):
    mov, ref, option = inputs_64x128x128()
    motion, err, log = run_reference(cf, mov, ref, option)
    r = option["r"]
    nx, ny, nz = mov.shape
    xg = np.arange(r, nx - 1, 2 * r + 1)
    yg = np.arange(r, ny - 1, 2 * r + 1)
    rng = np.random.default_rng(12345)
    pick = rng.integers(0, [nx, ny, nz], size=(4096, 3))
    payload = dict(
        mov_sha=sha(mov), ref_sha=sha(ref),
        smooth_penalty=np.float64(option["smoothPenalty"]),
        motion_cp=motion[np.ix_(xg, yg, np.arange(nz))],
        sample_idx=pick, sample_motion=motion[pick[:, 0], pick[:, 1], pick[:, 2]],
        motion_absmax=np.float32(np.abs(motion).max()),
        current_error=np.float64(err), **log_arrays(log, option["layer"]))
    np.savez_compressed(os.path.join(HERE, "getmotion_64x128x128.npz"), **payload)
    print("64x128x128 err", err, "records",
          [len(log[f"layer_{L}"]["currentError"]) for L in range(option["layer"], -1, -1)])


def write_r16(cf):
    """The whole-body workload's parameter set (r 16, 20 iterations, mask_ref)
    on small volumes: input hashes, the level-0 field at the control points
    and at 4096 sampled voxels, the per-level error logs."""
    for name in R16_CASES:
        mov, ref, option = small_inputs(name)
        motion, err, log = run_reference(cf, mov, ref, option)
        r = option["r"]
        nx, ny, nz = mov.shape
        xg = np.arange(r, nx - 1, 2 * r + 1)
        yg = np.arange(r, ny - 1, 2 * r + 1)
        pick = np.random.default_rng(54321).integers(0, [nx, ny, nz], size=(4096, 3))
        payload = dict(
            mov_sha=sha(mov), ref_sha=sha(ref), mask_ref_sha=sha(np.asarray(option["mask_ref"])),
            smooth_penalty=np.float64(option["smoothPenalty"]),
            motion_cp=motion[np.ix_(xg, yg, np.arange(nz))],
            sample_idx=pick, sample_motion=motion[pick[:, 0], pick[:, 1], pick[:, 2]],
            motion_absmax=np.float32(np.abs(motion).max()),
            current_error=np.float64(err), **log_arrays(log, option["layer"]))
        np.savez_compressed(os.path.join(HERE, f"getmotion_{name}.npz"), **payload)
        print(name, "err", err, "records",
              [len(log[f"layer_{L}"]["currentError"]) for L in range(option["layer"], -1, -1)])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--skip-64x128x128", action="store_true")
    ap.add_argument("--only-r16", action="store_true",
                    help="write only the fixtures of the r = 16 parameter set")
    args = ap.parse_args()
    cf, im = import_reference(args.reference)
    if args.only_r16:
        write_r16(cf)
        return
    write_imresize(im)
    write_small(cf)
    write_r16(cf)
    if not args.skip_64x128x128:
        write_64x128x128(cf)
    _ = synth  # the inputs come from the package's own generator


if __name__ == "__main__":
    main()
