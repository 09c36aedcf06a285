"""numpy restatement of the fixed-plane projections of the cross-resolution path.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP splat
kernels.  The product path never calls it.

``project_max``  calFlowCrossResolution.py:480-568 (kernel) and :878-1141
    (host side: value shift ``max(1 - min(finite values), 0)`` in float32,
    -inf as the empty marker, shift taken off, fill, axis order).
``project_weighted``  calFlowCrossResolution.py:571-669 (kernel) and :700-875
    (host side): z weights (hard / triangular / Gaussian), bilinear weights
    at ``(x / ds, y / ds)``, ``sum(v * w) / sum(w)`` where ``sum(w) > eps``.
    Sums in float64 (the reference's float atomics add in no fixed order).

Both loop over samples in Python; they are for the small inputs of the tests.
Pinned by hand-computed cases (tests/test_splat_oracle.py).
"""

from __future__ import annotations

import math

import numpy as np


def _out_shape(xref, yref, ds):
    return (xref + ds - 1) // ds, (yref + ds - 1) // ds


def project_max(coords, values, target_z, xref, yref, z_window=1.5, downsample_xy=1, nearest=False,
                extra_radius=0, fill_value=0.0):
    """Returns ``(Nplanes, Xout, Yout)`` float32."""
    c = np.asarray(coords, dtype=np.float32).reshape(-1, 3)
    v = np.asarray(values, dtype=np.float32).reshape(-1)
    tz = np.atleast_1d(np.asarray(target_z, dtype=np.float32))
    ds = int(downsample_xy)
    xout, yout = _out_shape(xref, yref, ds)
    fin = v[np.isfinite(v)]
    shift = np.maximum(-fin.min() + np.float32(1.0), np.float32(0.0)) if fin.size else np.float32(0.0)
    vs = v + shift
    out = np.full((tz.size, xout, yout), -np.inf, dtype=np.float32)
    for i in range(v.size):
        val = vs[i]
        x, y, z = c[i]
        if not (np.isfinite(val) and np.isfinite(x) and np.isfinite(y) and np.isfinite(z)):
            continue
        if nearest:
            xs = xe = int(math.floor(np.float32(x + np.float32(0.5))))
            ys = ye = int(math.floor(np.float32(y + np.float32(0.5))))
        else:
            xf, yf = int(math.floor(x)), int(math.floor(y))
            xs, xe, ys, ye = xf - extra_radius, xf + 1 + extra_radius, yf - extra_radius, yf + 1 + extra_radius
        for p, z0 in enumerate(tz):
            if abs(np.float32(z - z0)) > np.float32(z_window):
                continue
            for xi in range(xs, xe + 1):
                if xi < 0 or xi >= xref:
                    continue
                for yi in range(ys, ye + 1):
                    if yi < 0 or yi >= yref:
                        continue
                    o = (p, xi // ds, yi // ds)
                    out[o] = max(out[o], val)
    done = np.isfinite(out)
    out[done] = out[done] - shift
    if fill_value is not None:
        out[~np.isfinite(out)] = np.float32(fill_value)
    return out


def project_weighted(coords, values, target_z, xref, yref, z_window=1.5, z_sigma=None, mode="gaussian",
                     downsample_xy=1, fill_value=0.0, eps=1e-6):
    """Returns ``(out, sum_w)``, each ``(Nplanes, Xout, Yout)``."""
    c = np.asarray(coords, dtype=np.float32).reshape(-1, 3)
    v = np.asarray(values, dtype=np.float32).reshape(-1)
    tz = np.atleast_1d(np.asarray(target_z, dtype=np.float32))
    ds = int(downsample_xy)
    xout, yout = _out_shape(xref, yref, ds)
    sigma = np.float32(z_window * 0.5 if z_sigma is None else z_sigma)
    if sigma <= 0:
        sigma = np.float32(z_window * 0.5)
    zw = np.float32(z_window)
    sv = np.zeros((tz.size, xout, yout), dtype=np.float64)
    sw = np.zeros_like(sv)
    for i in range(v.size):
        val = v[i]
        x, y, z = c[i]
        if not (np.isfinite(val) and np.isfinite(x) and np.isfinite(y) and np.isfinite(z)):
            continue
        if x < 0 or x > xref - 1 or y < 0 or y > yref - 1:
            continue
        xo_f, yo_f = np.float32(x / np.float32(ds)), np.float32(y / np.float32(ds))
        xo0, yo0 = int(math.floor(xo_f)), int(math.floor(yo_f))
        dx, dy = np.float32(xo_f - np.float32(xo0)), np.float32(yo_f - np.float32(yo0))
        for p, z0 in enumerate(tz):
            dz = np.float32(abs(np.float32(z - z0)))
            if dz > zw:
                continue
            if mode == "hard":
                wz = np.float32(1.0)
            elif mode == "triangular":
                wz = max(np.float32(1.0) - dz / (zw + np.float32(1e-6)), np.float32(0.0))
            else:
                wz = np.float32(math.exp(-0.5 * float(dz) * float(dz) / (float(sigma) * float(sigma) + 1e-6)))
            if wz <= 0:
                continue
            for ax in (0, 1):
                xo = xo0 + ax
                if xo < 0 or xo >= xout:
                    continue
                wx = np.float32(1.0) - dx if ax == 0 else dx
                for ay in (0, 1):
                    yo = yo0 + ay
                    if yo < 0 or yo >= yout:
                        continue
                    wy = np.float32(1.0) - dy if ay == 0 else dy
                    w = float(wx) * float(wy) * float(wz)
                    if w <= 0:
                        continue
                    sv[p, xo, yo] += float(val) * w
                    sw[p, xo, yo] += w
    out = np.where(sw > eps, sv / np.maximum(sw, eps), fill_value).astype(np.float32)
    return out, sw.astype(np.float32)
