"""numpy/scipy restatement of the registration harness's mask and error helpers.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP mask
kernels.  The product path never calls it.

Functions and the reference code they follow (paths under
``src/wholistic_registration/utils/``):

``robust_mean_std``  preprocess.py:69-84 -- 95th percentile (numpy's linear
    method, in the data's float dtype), then mean and std of the values at or
    below it.  Same numpy calls as the reference.
``get_mask``  mask.py:59-111 -- float32 z-score ``|(x - mu) / sigma|`` above
    ``thres_factor``, then scikit-image ``binary_opening`` and
    ``binary_dilation`` with a 3x3x1 footprint.  scikit-image is not installed
    here (``scikit-image>=0.21`` in pyproject.toml:22, unpinned); its binary
    morphology is ``scipy.ndimage.binary_erosion(..., border_value=True)``
    for erosion and ``scipy.ndimage.binary_dilation`` (border value False) for
    dilation, opening = erosion then dilation.  That is what this function
    calls.
``bwareafilt3``  mask.py:8-56 -- keep the 26-connected components
    (``skimage.measure.label(connectivity=3)``) whose voxel count lies in
    ``[size_range[0], size_range[1]]``.  Restated with
    ``scipy.ndimage.label`` and the full 3x3x3 structuring element, which
    defines the same components.
``mean_sq_diff``  registration.py:328-329 -- ``np.mean((a - b) ** 2)`` with
    numpy's dtype rules (uint16 differences and squares wrap modulo 2^16).

Pinning (tests/test_mask_oracle.py): the reference's own robust_mean_std test
vector (tests/unit/test_preprocess.py:52-57), numpy's percentile on closed
forms, brute-force loops for the morphology with its border rules, a
breadth-first flood fill over the 26-neighbourhood for the components, and
hand-made shapes that only 26-connectivity joins.
"""

from __future__ import annotations

import numpy as np
from scipy import ndimage as ndi

FOOTPRINT = np.ones((3, 3, 1), dtype=bool)     # mask.py:101 (x, y, z) layout
STRUCT26 = np.ones((3, 3, 3), dtype=bool)


def robust_mean_std(data, percentile=95):
    """preprocess.py:69-84: ``(mean, std)`` of the values <= the percentile."""
    data_flat = np.asarray(data).flatten()
    threshold = np.percentile(data_flat, percentile)
    selected = data_flat[data_flat <= threshold]
    return np.mean(selected), np.std(selected)


def percentile_threshold(data, percentile=95):
    """The threshold robust_mean_std compares against (np.percentile)."""
    return np.percentile(np.asarray(data).flatten(), percentile)


def zscore_mask(dat, thres_factor, mu_sigma=None):
    """mask.py:87-98: float32 z-score above the threshold.  ``mu_sigma``
    overrides the robust statistics (both cast to float32)."""
    d = np.asarray(dat).astype(np.float32)
    if mu_sigma is None:
        mu, sigma = robust_mean_std(d)
    else:
        mu, sigma = np.float32(mu_sigma[0]), np.float32(mu_sigma[1])
    normalized = np.abs((d - mu) / sigma)
    return normalized > thres_factor


def get_mask(dat, thres_factor, mu_sigma=None):
    """mask.py:59-111 on an ``(x, y, z)`` volume; returns a bool volume."""
    mask = zscore_mask(dat, thres_factor, mu_sigma)
    eroded = ndi.binary_erosion(mask, structure=FOOTPRINT, border_value=True)
    opened = ndi.binary_dilation(eroded, structure=FOOTPRINT)
    return ndi.binary_dilation(opened, structure=FOOTPRINT)


def bwareafilt3(vol, size_range):
    """mask.py:8-56: 26-connected components with area in the closed range."""
    fg = np.asarray(vol).astype(bool)
    labels, n = ndi.label(fg, structure=STRUCT26)
    area = np.bincount(labels.ravel(), minlength=n + 1)
    keep = np.zeros(n + 1, dtype=bool)
    keep[1:] = (size_range[0] <= area[1:]) & (area[1:] <= size_range[1])
    return keep[labels] & fg


def n_components(vol):
    """Number of 26-connected foreground components."""
    return int(ndi.label(np.asarray(vol).astype(bool), structure=STRUCT26)[1])


def mean_sq_diff(a, b):
    """registration.py:328-329: ``np.mean((a - b) ** 2)`` in numpy's dtypes."""
    return np.mean((np.asarray(a) - np.asarray(b)) ** 2)


# ----------------------------------------------------------------------------
# Brute-force restatements on tiny volumes (they pin the functions above)
# ----------------------------------------------------------------------------


def brute_morph(mask, erode, border):
    """3x3 in-plane erosion (min) or dilation (max) over axes 0, 1 with the
    given value outside the volume, by explicit loops."""
    m = np.asarray(mask, dtype=bool)
    nx, ny, nz = m.shape
    out = np.zeros_like(m)
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                vals = []
                for dx in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        i, j = x + dx, y + dy
                        inside = 0 <= i < nx and 0 <= j < ny
                        vals.append(bool(m[i, j, z]) if inside else border)
                out[x, y, z] = all(vals) if erode else any(vals)
    return out


def brute_components(vol):
    """Breadth-first flood fill over the 26-neighbourhood: list of voxel sets."""
    fg = np.asarray(vol, dtype=bool)
    seen = np.zeros_like(fg)
    shape = fg.shape
    comps = []
    for start in zip(*np.nonzero(fg)):
        if seen[start]:
            continue
        seen[start] = True
        queue, comp = [start], []
        while queue:
            p = queue.pop()
            comp.append(p)
            for d in np.ndindex(3, 3, 3):
                q = tuple(p[k] + d[k] - 1 for k in range(3))
                if all(0 <= q[k] < shape[k] for k in range(3)) and fg[q] and not seen[q]:
                    seen[q] = True
                    queue.append(q)
        comps.append(comp)
    return comps
