"""Build libwrhip.so (the HIP kernels and their C ABI) in-tree for gfx950.

    python -m wholistic_registration_amd.build_lib

hipcc cross-compiles without a GPU.  Packed-FP32 instruction selection is
disabled for every device compilation (-target-feature -packed-fp32-ops).
Each source compiles to its own object in parallel (objects under
wholistic_registration_amd/build/, kept out of git); one hipcc call links them.
"""

from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libwrhip.so")
SOURCES = ["api.hip", "resample.hip", "lk.hip", "lk_pass.hip", "warp.hip", "penalty.hip", "mask.hip", "harness.hip", "splat.hip", "net.hip", "xres.hip"]
ARCH = os.environ.get("WR_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), "hipcc"):
        if os.path.sep not in cand or os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
OBJ = os.path.join(PKG, "build")


def compile_command(src, obj):
    return [hipcc(), f"--offload-arch={ARCH}", "-c"] + FLAGS + ["-o", obj, os.path.join(CSRC, src)]


def link_command(objs, out=LIB):
    return [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out] + objs


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}: {' '.join(cmd[-1:])}")
    noise = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if verbose and noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build(verbose=False, jobs=None):
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    jobs = jobs or min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(_run, compile_command(s, o), verbose) for s, o in zip(SOURCES, objs)]:
            f.result()
    tmp = LIB + ".tmp"
    _run(link_command(objs, tmp), verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
