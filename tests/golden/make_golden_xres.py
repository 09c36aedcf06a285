"""Write the cross-resolution (``getMotion_v2``) fixtures by running the reference.

Run in the development container only (the reference tree is not present on
GPU machines):

    python tests/golden/make_golden_xres.py [--reference /root/reference]

It imports ``src/wholistic_registration/utils/calFlowCrossResolution.py``
through an empty parent package (the package ``__init__`` needs I/O
libraries the solver does not use).  Without CuPy the module's own fallback
binds its array namespace to numpy and its ndimage calls to
``scipy.ndimage`` (utils/__init__.py), so ``getMotion_v2`` runs on the CPU as
written.  Each case of ``xres_cases.py`` is run with ``verbose=True``; the
fixture holds the three returned arrays and the printed log (per-iteration
errors, stop reasons, wrong-region counts and decisions).  For the z-offset search
``FindInitZ_stack_global_fixed_spacing`` (which ends in a CuPy-only call), the
reference's slice features (``calculate.grad_mag_2d``) and ZNCC score matrix
(``_global_zncc_scores``) are written for the cases of ``FINDZ_CASES``.  No reference
source is copied.
"""

from __future__ import annotations

import argparse
import contextlib
import io
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from tests.golden.xres_cases import FINDZ_CASES, XRES_CASES, findz_inputs, xres_inputs  # noqa: E402


def import_reference(root):
    src = os.path.join(root, "src", "wholistic_registration")
    if "wholistic_registration" not in sys.modules:
        pkg = types.ModuleType("wholistic_registration")
        pkg.__path__ = [src]
        sys.modules["wholistic_registration"] = pkg
    import importlib
    return importlib.import_module("wholistic_registration.utils.calFlowCrossResolution")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--only", default=None, help="comma-separated case names")
    args = ap.parse_args()
    cf = import_reference(args.reference)
    names = args.only.split(",") if args.only else list(XRES_CASES) + list(FINDZ_CASES)
    for name in names:
        if name in FINDZ_CASES:
            # FindInitZ_stack_global_fixed_spacing itself ends in cp.asnumpy, which the
            # numpy fallback lacks (AttributeError): its two numeric steps are run instead,
            # the reference's own grad_mag_2d per slice and _global_zncc_scores; the search
            # over the score matrix is pinned by the planted offsets of the case
            mov, ref, delta, direction, grad, hann = findz_inputs(name)
            calc = cf.calculate
            if grad:
                mf = np.stack([calc.grad_mag_2d(mov[:, :, k]) for k in range(mov.shape[2])], axis=2)
                rf = np.stack([calc.grad_mag_2d(ref[:, :, z]) for z in range(ref.shape[2])], axis=2)
            else:
                mf, rf = mov, ref
            scores = cf._global_zncc_scores(mf, rf, use_hann_weight=hann, eps=1e-6)
            np.savez_compressed(os.path.join(HERE, f"findz_{name}.npz"), scores=np.asarray(scores),
                                mov_feat=np.asarray(mf)[:, :, 0], ref_feat=np.asarray(rf)[:, :, 0])
            print(name, "scores", np.asarray(scores).shape)
            continue
        mov, ref, opt = xres_inputs(name)
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf):
            phase, motion, sampled = cf.getMotion_v2(mov.copy(), ref.copy(), dict(opt), verbose=True)
        log = np.array(buf.getvalue().splitlines())
        np.savez_compressed(os.path.join(HERE, f"xres_{name}.npz"),
                            phase=np.asarray(phase, dtype=np.float32),
                            motion=np.asarray(motion, dtype=np.float32),
                            sampled=np.asarray(sampled, dtype=np.float64), log=log)
        print(name, "phase", phase.shape, "log lines", len(log))
        for line in log:
            if "stop" in line or "region" in line or "accept" in line or "keep" in line or "skip" in line:
                print("   ", line)


if __name__ == "__main__":
    main()
