"""Seeded synthetic registration workloads.

A template is a background of 100 plus isotropic Gaussian blobs
(sigma ~ U[1.5, 3] voxels, amplitude ~ LogNormal(6, 0.5), centres uniform),
followed by Poisson noise.  A displacement field is a sum of 3-D sinusoidal
modes ``a_k * sin(2*pi*<w_k, p>/lambda_k + phi_k)`` (unit ``w_k``,
``lambda_k ~ U[32, 96]``) rescaled so that ``max_p |d(p)| = dmax``.  The
moving volume is the noise-free template pull-warped by ``-d`` (trilinear,
clamp to edge), the convention of
``src/wholistic_registration/utils/simulation.py:244-245``, plus independent
Poisson noise.  The registration field that undoes it solves
``m(p) = d(p + m(p))`` (:func:`truth_field`).

Volumes are ``(Z, Y, X)`` C-order arrays; fields are ``(Z, Y, X, 3)`` with
components ordered ``(x, y, z)``, the output layout of the reference harness
(``registration.py:343-346``).

Two renderers share one parameter stream: numpy (small volumes, host) and
torch (large volumes, on a device).  Blob and mode parameters are drawn from
numpy's PCG64 with the given seed in both, so the structure is identical; the
Poisson draws come from each backend's own generator.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

BLOB_RADIUS = 9          # 3 * max sigma


@dataclass
class Blobs:
    centre: np.ndarray   # (n, 3) in (z, y, x) voxel units
    sigma: np.ndarray    # (n,)
    amp: np.ndarray      # (n,)


@dataclass
class Modes:
    omega: np.ndarray    # (k, 3) unit vectors over (x, y, z)
    lam: np.ndarray      # (k,)
    phase: np.ndarray    # (k,)
    amp: np.ndarray      # (k, 3) per-axis amplitude before the global rescale


def draw_blobs(rng, shape_zyx, n):
    c = rng.uniform(0.0, 1.0, size=(n, 3)) * (np.asarray(shape_zyx, dtype=np.float64) - 1.0)
    sigma = rng.uniform(1.5, 3.0, size=n)
    amp = rng.lognormal(mean=6.0, sigma=0.5, size=n)
    return Blobs(c, sigma, amp)


def draw_modes(rng, k):
    w = rng.normal(size=(k, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    lam = rng.uniform(32.0, 96.0, size=k)
    phase = rng.uniform(0.0, 2.0 * math.pi, size=k)
    amp = rng.uniform(-1.0, 1.0, size=(k, 3))
    return Modes(w, lam, phase, amp)


# ---------------------------------------------------------------- numpy side


def render_template_np(shape_zyx, blobs, background=100.0):
    vol = np.full(shape_zyx, background, dtype=np.float64)
    nz, ny, nx = shape_zyx
    for (cz, cy, cx), s, a in zip(blobs.centre, blobs.sigma, blobs.amp):
        r = int(math.ceil(3.0 * s))
        z0, z1 = max(int(math.floor(cz)) - r, 0), min(int(math.floor(cz)) + r + 1, nz)
        y0, y1 = max(int(math.floor(cy)) - r, 0), min(int(math.floor(cy)) + r + 1, ny)
        x0, x1 = max(int(math.floor(cx)) - r, 0), min(int(math.floor(cx)) + r + 1, nx)
        dz = (np.arange(z0, z1) - cz)[:, None, None]
        dy = (np.arange(y0, y1) - cy)[None, :, None]
        dx = (np.arange(x0, x1) - cx)[None, None, :]
        vol[z0:z1, y0:y1, x0:x1] += a * np.exp(-(dz * dz + dy * dy + dx * dx) / (2 * s * s))
    return vol.astype(np.float32)


def field_raw_np(shape_zyx, modes):
    nz, ny, nx = shape_zyx
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d = np.zeros(shape_zyx + (3,), dtype=np.float64)
    for w, lam, ph, a in zip(modes.omega, modes.lam, modes.phase, modes.amp):
        s = np.sin(2 * math.pi * (w[0] * x + w[1] * y + w[2] * z) / lam + ph)
        d += s[..., None] * a[None, None, None, :]
    return d


def field_np(shape_zyx, modes, dmax):
    d = field_raw_np(shape_zyx, modes)
    peak = np.sqrt((d ** 2).sum(-1)).max()
    return (d * (dmax / peak)).astype(np.float32)


def pull_warp_np(vol_zyx, field_zyx3):
    """moving(p) = vol(p + field(p)), trilinear, clamp to edge (host)."""
    from scipy import ndimage as ndi
    nz, ny, nx = vol_zyx.shape
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.float32), np.arange(ny, dtype=np.float32),
                          np.arange(nx, dtype=np.float32), indexing="ij")
    coords = np.stack([z + field_zyx3[..., 2], y + field_zyx3[..., 1], x + field_zyx3[..., 0]])
    return ndi.map_coordinates(vol_zyx.astype(np.float32), coords, order=1, mode="nearest")


def truth_field(field_zyx3, n_iter=30):
    """Registration field ``m`` with ``m(p) = d(p + m(p))`` for a moving volume
    made by pull-warping with ``-d`` (fixed-point iteration, host)."""
    from scipy import ndimage as ndi
    nz, ny, nx, _ = field_zyx3.shape
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                          np.arange(nx, dtype=np.float64), indexing="ij")
    d = field_zyx3.astype(np.float64)
    m = d.copy()
    for _ in range(n_iter):
        coords = np.stack([z + m[..., 2], y + m[..., 1], x + m[..., 0]])
        m = np.stack([ndi.map_coordinates(d[..., c], coords, order=1, mode="nearest")
                      for c in range(3)], axis=-1)
    return m.astype(np.float32)


def ellipsoid_mask_np(rng, shape_zyx, fraction):
    nz, ny, nx = shape_zyx
    centre = np.array([rng.uniform(0.3, 0.7) * (n - 1) for n in shape_zyx])
    ratio = rng.uniform(0.6, 1.4, size=3) * np.asarray(shape_zyx, dtype=np.float64)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    q = (((z - centre[0]) / ratio[0]) ** 2 + ((y - centre[1]) / ratio[1]) ** 2
         + ((x - centre[2]) / ratio[2]) ** 2)
    thr = np.quantile(q.ravel(), fraction)
    return q <= thr


@dataclass
class Pair:
    template: np.ndarray      # (Z, Y, X) float32 (or uint16-valued float32)
    moving: np.ndarray        # (Z, Y, X)
    field: np.ndarray         # (Z, Y, X, 3) displacement d used to make moving
    mask: np.ndarray | None   # (Z, Y, X) bool, True = excluded


def make_pair_np(seed, shape_zyx, n_blobs, n_modes, dmax, uint16=False, mask_fraction=0.0):
    """Host-side pair (small volumes).  Returns float32 volumes; with
    ``uint16=True`` they are rounded and clipped to uint16 first."""
    rng = np.random.default_rng(seed)
    blobs = draw_blobs(rng, shape_zyx, n_blobs)
    modes = draw_modes(rng, n_modes)
    clean = render_template_np(shape_zyx, blobs)
    d = field_np(shape_zyx, modes, dmax)
    warped = pull_warp_np(clean, -d)
    tmpl = rng.poisson(clean.astype(np.float64)).astype(np.float32)
    mov = rng.poisson(np.maximum(warped, 0).astype(np.float64)).astype(np.float32)
    if uint16:
        tmpl = np.clip(np.rint(tmpl), 0, 65535).astype(np.uint16).astype(np.float32)
        mov = np.clip(np.rint(mov), 0, 65535).astype(np.uint16).astype(np.float32)
    mask = ellipsoid_mask_np(rng, shape_zyx, mask_fraction) if mask_fraction > 0 else None
    return Pair(tmpl, mov, d, mask)


def make_xres_pair_np(seed, mov_xyz, ref_xyz, z_ratio, z_ratio_hr, n_blobs, dmax, artefacts=0,
                      artefact_amp=900.0):
    """Host-side cross-resolution pair in the solver's ``(x, y, z)`` layout:
    a blob reference on its own z grid, and a moving volume on a coarser z
    grid that samples the noise-free reference at ``(x, y, k zRatio /
    zRatio_HR) + d`` (trilinear) for a smooth 3-mode field ``d`` of max
    |d| = dmax (reference index units), plus ``artefacts`` bright flat blobs
    that the reference does not hold, then Poisson noise on both.  Returns
    ``(mov_xyz, ref_xyz, rng)`` (float32; the generator continues for masks
    and warm starts)."""
    from scipy import ndimage as ndi
    rng = np.random.default_rng(seed)
    rx, ry, rz = ref_xyz
    mx, my, mz = mov_xyz
    ref_zyx = (rz, ry, rx)
    blobs = draw_blobs(rng, ref_zyx, n_blobs)
    clean = render_template_np(ref_zyx, blobs).astype(np.float64)
    modes = draw_modes(rng, 3)
    d = field_np((mz, my, mx), modes, dmax).astype(np.float64)
    k, yy, xx = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    zs = k * (z_ratio / z_ratio_hr)
    coords = np.stack([zs + d[..., 2], yy + d[..., 1], xx + d[..., 0]])
    mov_clean = ndi.map_coordinates(clean, coords, order=1, mode="nearest")
    for _ in range(artefacts):
        cz, cy, cx = rng.integers(0, mz), rng.integers(6, my - 6), rng.integers(6, mx - 6)
        rr = ((yy - cy) ** 2 + (xx - cx) ** 2) / 9.0 + (k - cz) ** 2
        mov_clean = mov_clean + artefact_amp * np.exp(-rr)
    ref = rng.poisson(clean).astype(np.float32)
    mov = rng.poisson(np.maximum(mov_clean, 0)).astype(np.float32)
    return np.ascontiguousarray(mov.transpose(2, 1, 0)), np.ascontiguousarray(ref.transpose(2, 1, 0)), rng


# ---------------------------------------------------------------- torch side


def render_template_torch(shape_zyx, blobs, device, background=100.0, chunk=16384):
    import torch
    nz, ny, nx = shape_zyx
    vol = torch.full((nz * ny * nx,), background, dtype=torch.float32, device=device)
    r = BLOB_RADIUS
    off = torch.arange(-r, r + 1, device=device)
    oz, oy, ox = torch.meshgrid(off, off, off, indexing="ij")
    oz, oy, ox = oz.reshape(1, -1), oy.reshape(1, -1), ox.reshape(1, -1)
    cen = torch.as_tensor(blobs.centre, dtype=torch.float64, device=device)
    sig = torch.as_tensor(blobs.sigma, dtype=torch.float64, device=device)
    amp = torch.as_tensor(blobs.amp, dtype=torch.float64, device=device)
    for s in range(0, cen.shape[0], chunk):
        c = cen[s:s + chunk]
        base = torch.floor(c).long()
        pz, py, px = base[:, 0:1] + oz, base[:, 1:2] + oy, base[:, 2:3] + ox
        dz, dy, dx = pz - c[:, 0:1], py - c[:, 1:2], px - c[:, 2:3]
        sg = sig[s:s + chunk, None]
        half = torch.ceil(3 * sg).long()
        # same cube window as the host renderer: |p - floor(c)| <= ceil(3 sigma)
        keep = (oz.abs() <= half) & (oy.abs() <= half) & (ox.abs() <= half) & \
               (pz >= 0) & (pz < nz) & (py >= 0) & (py < ny) & (px >= 0) & (px < nx)
        w = amp[s:s + chunk, None] * torch.exp(-(dz * dz + dy * dy + dx * dx) / (2 * sg * sg))
        idx = (pz * ny + py) * nx + px
        vol.index_add_(0, idx[keep], w[keep].float())
    return vol.view(nz, ny, nx)


def field_torch(shape_zyx, modes, dmax, device):
    import torch
    nz, ny, nx = shape_zyx
    d = torch.zeros((nz, ny, nx, 3), dtype=torch.float32, device=device)
    x = torch.arange(nx, device=device, dtype=torch.float64).view(1, 1, nx)
    y = torch.arange(ny, device=device, dtype=torch.float64).view(1, ny, 1)
    for w, lam, ph, a in zip(modes.omega, modes.lam, modes.phase, modes.amp):
        for zs in range(0, nz, 8):
            z = torch.arange(zs, min(zs + 8, nz), device=device, dtype=torch.float64).view(-1, 1, 1)
            s = torch.sin(2 * math.pi * (w[0] * x + w[1] * y + w[2] * z) / lam + ph).float()
            d[zs:zs + z.shape[0]] += s[..., None] * torch.as_tensor(a, dtype=torch.float32,
                                                                    device=device)
    peak = torch.sqrt((d * d).sum(-1)).max()
    d *= float(dmax) / float(peak)
    return d


def ellipsoid_mask_torch(rng, shape_zyx, fraction, device):
    import torch
    nz, ny, nx = shape_zyx
    centre = [rng.uniform(0.3, 0.7) * (n - 1) for n in shape_zyx]
    ratio = rng.uniform(0.6, 1.4, size=3) * np.asarray(shape_zyx, dtype=np.float64)
    z = torch.arange(nz, device=device, dtype=torch.float32).view(-1, 1, 1)
    y = torch.arange(ny, device=device, dtype=torch.float32).view(1, -1, 1)
    x = torch.arange(nx, device=device, dtype=torch.float32).view(1, 1, -1)
    q = (((z - centre[0]) / ratio[0]) ** 2 + ((y - centre[1]) / ratio[1]) ** 2
         + ((x - centre[2]) / ratio[2]) ** 2)
    # bisection on the threshold so that the excluded fraction is `fraction`
    lo, hi = 0.0, float(q.max())
    total = q.numel()
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if float((q <= mid).sum()) / total < fraction:
            lo = mid
        else:
            hi = mid
    return q <= hi
