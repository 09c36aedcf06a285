"""Write the golden fixtures of the harness's reference picking by running the reference.

Run in the development container only (the reference tree is not present on
GPU machines):

    python tests/golden/make_golden_harness.py [--reference /root/reference]

It imports ``src/wholistic_registration/utils/reference.py`` through an empty
parent package (the package ``__init__`` needs I/O libraries this function
does not use; ``utils/__init__`` falls back to numpy when CuPy is absent) and
records ``pick_initial_reference`` (reference.py:19-99) on small seeded frame
stacks: the chosen frame indices and the reference volume.  Nothing else pins
its selection rule (top-k correlation means, arg-max, arg-sort) in
combination.  Output: ``pick_reference.npz`` next to this script (inputs are
regenerated from the seeds in ``harness_cases.py``).
"""

from __future__ import annotations

import argparse
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from tests.golden.harness_cases import PICK_CASES, pick_frames  # noqa: E402


def import_reference(root):
    src = os.path.join(root, "src", "wholistic_registration")
    if "wholistic_registration" not in sys.modules:
        pkg = types.ModuleType("wholistic_registration")
        pkg.__path__ = [src]
        sys.modules["wholistic_registration"] = pkg
    import importlib
    return importlib.import_module("wholistic_registration.utils.reference")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    ref = import_reference(args.reference)
    out = {}
    for i, case in enumerate(PICK_CASES):
        frames = pick_frames(i)
        img, top = ref.pick_initial_reference(frames, max_corr_frames=case["max_corr"])
        out[f"top_{i}"] = np.asarray(top)
        out[f"ref_{i}"] = np.asarray(img)
        print(i, case, "top", np.asarray(top)[:8], "ref dtype", np.asarray(img).dtype)
    np.savez_compressed(os.path.join(HERE, "pick_reference.npz"), **out)


if __name__ == "__main__":
    main()
