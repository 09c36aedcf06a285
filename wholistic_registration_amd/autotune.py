"""Autotuner of the LK tile shape on the device.

The level-0..layer LK passes split each level into workgroup tiles of
``np_x`` patches along x (one patch row along y, every z plane its own
tiles).  The built-in rule (the widest tile that still gives 2048
workgroups) is right for large levels and wrong by a few per cent where
the x extent leaves a ragged last tile or a level has few patches.
``tune_tiles`` times every admissible width on the caller's own data,
with every candidate running the same forced iteration schedule, reads
the per-level pass times the library records with HIP events
(``wr_params.timing``), and picks the fastest width per level.  The field
does not depend on the tile shape beyond float32 summation order
(per-patch sums are reduced across the waves of a workgroup), so a tuned
run stays within the parity tolerance of an untuned one.

``TileCache`` keeps one choice per geometry (volume shape, layer, r,
precision, mask) so that every later frame or crop of that geometry uses
it without tuning again.
"""

from __future__ import annotations

import threading

import numpy as np

from . import device as _dev

# widest LK tile in voxel columns (4 waves x 62 owned columns)
MAX_TILE_COLS = 248


def candidates(r):
    """Admissible tile widths (patches per workgroup along x) for patch half-width r."""
    return list(range(1, max(1, MAX_TILE_COLS // (2 * r + 1)) + 1))


def tune_tiles(mov, ref, *, layer, r, iters=3, reps=2, context=None, stream=None, **params):
    """Per-level tile widths (list of ``layer + 1`` ints, level 0 first) that
    minimise the measured LK pass time on this pair, and the table of
    measurements ``{width: [ms per pass of level 0..layer]}`` (width 0 = the
    built-in rule).  ``params`` are the remaining ``device.get_motion``
    arguments (zRatio, movRange, smoothPenalty, tol, mask_ref, precision...)."""
    sched = {L: (iters, iters - 1) for L in range(layer + 1)}
    table = {}
    for width in [0] + candidates(r):
        best = None
        for _ in range(reps):
            _dev.get_motion(mov, ref, layer=layer, iter=iters, r=r, tile_patches_x=width, timing=True,
                            schedule=sched, context=context, stream=stream, **params)
            st = _dev.stats(context=context)
            ms = np.array([st.lk_ms[L] / max(st.lk_exec[L], 1) for L in range(layer + 1)])
            best = ms if best is None else np.minimum(best, ms)
        table[width] = best.tolist()
    widths = [c for c in table if c > 0]
    choice = []
    for L in range(layer + 1):
        c = min(widths, key=lambda w: table[w][L])
        # keep the rule unless a width is measurably (2 %) faster
        choice.append(c if table[c][L] < 0.98 * table[0][L] else 0)
    return choice, table


class TileCache:
    """Tuned tile widths per geometry, shared by the threads of one process."""

    def __init__(self):
        self._lock = threading.Lock()
        self._tiles = {}

    @staticmethod
    def key(mov, layer, r, precision="fp32", mask_ref=None):
        return (tuple(mov.shape), str(mov.dtype), int(layer), int(r), precision, mask_ref is not None)

    def get(self, key):
        with self._lock:
            return self._tiles.get(key)

    def tiles(self, mov, ref, *, layer, r, context=None, stream=None, **params):
        """The tuned widths for this pair's geometry (tuned on first use)."""
        k = self.key(mov, layer, r, params.get("precision", "fp32"), params.get("mask_ref"))
        with self._lock:
            hit = self._tiles.get(k)
        if hit is not None:
            return hit
        tune_params = {key: v for key, v in params.items() if key not in ("iter", "motion_out", "tile_patches_x")}
        choice, _ = tune_tiles(mov, ref, layer=layer, r=r, context=context, stream=stream, **tune_params)
        with self._lock:
            self._tiles.setdefault(k, choice)
            return self._tiles[k]
