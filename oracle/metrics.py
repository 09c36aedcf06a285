"""Endpoint-error statistics (TEST INFRASTRUCTURE ONLY).

Restates ``epe_3d`` of ``compute_motion_metrics``
(src/wholistic_registration/utils/simulation.py:303-339): the per-voxel
endpoint error is the Euclidean norm of the difference of two displacement
fields in index units.  Unlike the reference helper, which fixes the channel
order to (y, x, z), the three components enter symmetrically, so the norm is
the same for any consistent channel order.
"""

from __future__ import annotations

import numpy as np


def epe_stats(pred, gt):
    """Mean, median, p99 and max endpoint error between two ``(..., 3)`` fields."""
    pred = np.asarray(pred, dtype=np.float64)
    gt = np.asarray(gt, dtype=np.float64)
    e = np.sqrt(np.sum((pred - gt) ** 2, axis=-1)).ravel()
    return {
        "epe_mean": float(e.mean()),
        "epe_median": float(np.median(e)),
        "epe_p99": float(np.percentile(e, 99)),
        "epe_max": float(e.max()),
    }


def rel_l2(a, b):
    """Relative L2 difference ``||a - b|| / ||b||``."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
