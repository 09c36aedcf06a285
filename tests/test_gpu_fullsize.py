"""The workload's full size (96 x 1024 x 2048 uint16, 400k blobs, 8-mode field,
10 % ellipsoid mask), checked through properties that do not depend on size.

The oracle cannot run the whole registration at this size, so the checks use
what the mathematics makes size-independent:
  * pyramid level images and getMask act plane by plane (x-y only): sampled
    z planes of the full-size GPU result equal the oracle on those planes
    alone, bit for bit (getMask with the kernel's robust statistics, which
    are themselves checked against numpy on all 201M values);
  * the 26-connected components of the full-size mask equal scipy's;
  * an identity pair registers to exactly zero motion; two runs give the same
    field bit for bit; integer shifts warp exactly;
  * the bfloat16 / float16 level images stay within the endpoint-error gate of
    the float32 field.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import lk_oracle as O
from oracle import mask_oracle

pytestmark = pytest.mark.gpu

PLANES = (0, 47, 95)


@pytest.fixture(scope="module")
def full(cuda):
    import torch

    import bench
    w = bench.WORKLOADS[bench.DEFAULT]
    tmpl, mov, mask, _ = bench.make_pair_device(w, cuda)
    sp = bench.smooth_penalty(tmpl, mask) * bench.RAW_PENALTY
    params = dict(layer=w["layer"], iter=w["iter"], r=w["r"], zRatio=1.0, movRange=5.0, smoothPenalty=sp, tol=1e-3)
    yield dict(tmpl=tmpl, mov=mov, mask=mask, params=params)
    del tmpl, mov, mask
    torch.cuda.empty_cache()


@pytest.mark.parametrize("level", [1, 4])
def test_fullsize_pyramid_planes_bit_exact(full, level):
    from wholistic_registration_amd import device as D
    got = D.pyramid_level(full["tmpl"], level)
    for z in PLANES:
        plane = full["tmpl"][z].cpu().numpy().T[:, :, None]           # (x, y, 1) uint16
        shp = O.level_shape(plane.shape, level)
        want = O.imresize(plane, shp).astype(np.float32)[:, :, 0].T
        assert np.array_equal(got[z].cpu().numpy(), want), z


def test_fullsize_robust_stats_and_mask_planes(full):
    import torch

    from wholistic_registration_amd import device as D
    mov = full["mov"]
    thr, mu, sigma = D.robust_mean_std(mov.reshape(-1))
    f = mov.reshape(-1).cpu().numpy().astype(np.float32)
    assert thr == np.percentile(f, 95)                                # numpy's float32 index arithmetic
    sel = f[f <= thr].astype(np.float64)
    assert mu == pytest.approx(sel.mean(), rel=1e-6)
    assert sigma == pytest.approx(sel.std(), rel=1e-5)
    m = D.get_mask(mov, 5)
    for z in PLANES:
        plane = mov[z].cpu().numpy().T[:, :, None]
        want = mask_oracle.get_mask(plane, 5, mu_sigma=(mu, sigma))[:, :, 0].T
        assert np.array_equal(m[z].cpu().numpy().astype(bool), want), z
    del m
    torch.cuda.empty_cache()


def test_fullsize_components(full):
    from wholistic_registration_amd import device as D
    from scipy import ndimage as ndi
    m = D.get_mask(full["mov"], 5)
    kept, n = D.bwareafilt3(m, (5, 4000), count=True)
    labels, n_ref = ndi.label(m.cpu().numpy().astype(bool), structure=mask_oracle.STRUCT26)
    assert n == n_ref
    area = np.bincount(labels.ravel())
    keep = (area >= 5) & (area <= 4000)
    keep[0] = False
    assert np.array_equal(kept.cpu().numpy().astype(bool), keep[labels])


def test_fullsize_identity_and_determinism(full):
    import torch

    from wholistic_registration_amd import device as D
    p = full["params"]
    ident = D.get_motion(full["tmpl"], full["tmpl"], mask_ref=full["mask"], **p)
    assert torch.count_nonzero(ident.motion).item() == 0
    del ident
    a = D.get_motion(full["mov"], full["tmpl"], mask_ref=full["mask"], **p)
    b = D.get_motion(full["mov"], full["tmpl"], mask_ref=full["mask"], **p)
    assert a.records == b.records
    assert torch.equal(a.motion, b.motion)


@pytest.mark.parametrize("shift", [(1.0, 0.0, 0.0), (0.0, -2.0, 1.0)])
def test_fullsize_integer_shift_warp_is_exact(full, shift):
    import torch

    from wholistic_registration_amd import device as D
    vol = full["mov"]
    nz, ny, nx = vol.shape
    motion = torch.empty((nz, ny, nx, 3), dtype=torch.float32, device=vol.device)
    for c in range(3):
        motion[..., c] = shift[c]
    out = D.correct_motion(vol, motion)
    z = torch.clamp(torch.arange(nz, device=vol.device) + int(shift[2]), 0, nz - 1)
    y = torch.clamp(torch.arange(ny, device=vol.device) + int(shift[1]), 0, ny - 1)
    x = torch.clamp(torch.arange(nx, device=vol.device) + int(shift[0]), 0, nx - 1)
    want = vol.to(torch.int32)[z][:, y][:, :, x].to(torch.float32)
    assert torch.equal(out, want)


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_fullsize_reduced_precision_epe(full, precision):
    import torch

    from wholistic_registration_amd import device as D
    p = full["params"]
    a = D.get_motion(full["mov"], full["tmpl"], mask_ref=full["mask"], **p).motion
    b = D.get_motion(full["mov"], full["tmpl"], mask_ref=full["mask"], precision=precision, **p).motion
    e = torch.sqrt(((a - b) ** 2).sum(-1)).mean().item()
    assert e <= 0.1


CROP = (12, 320, 640)    # (z, y, x): level 4 is 40 x 20, one patch centre


def test_fullsize_parameter_set_on_masked_crop_matches_oracle(full):
    """The workload's own parameters (layer 4, r 16, 20 iterations, mask_ref)
    on a crop of the workload's pair across the mask ellipsoid's edge, against
    the oracle on the same bytes (calFlow3d_Wei_v1.py:430-439,515-530,546-588,
    641-654): free-running schedule equal to the oracle's; with that schedule
    forced, field max |d| <= 1e-3 voxel and the warp by that field within
    relative L2 1e-5 of the oracle's warp; mean
    endpoint error against the all-float64 oracle <= 0.1 voxel."""
    import torch

    from oracle.metrics import epe_stats, rel_l2
    from wholistic_registration_amd import device as D
    mask = full["mask"]
    nz, ny, nx = mask.shape
    # centred on the ellipsoid's edge: the row through its centroid, where
    # the mask begins along x
    cen = [int(round(c)) for c in torch.nonzero(mask > 0).float().mean(0).tolist()]
    row = torch.nonzero(mask[cen[0], cen[1]] > 0).flatten()
    cen[2] = int(row[0]) if row.numel() else cen[2]
    lo = [int(min(max(c - s // 2, 0), n - s)) for c, s, n in zip(cen, CROP, (nz, ny, nx))]
    sl = tuple(slice(a, a + s) for a, s in zip(lo, CROP))
    tmpl = full["tmpl"][sl].contiguous()
    mov = full["mov"][sl].contiguous()
    mk = mask[sl].contiguous()
    frac = float((mk > 0).float().mean())
    print(f"crop origin (z,y,x) {lo}, size {CROP}, masked fraction {frac:.3f}", flush=True)
    assert 0.0 < frac < 1.0
    ref_x = tmpl.float().cpu().numpy().transpose(2, 1, 0)
    mov_x = mov.float().cpu().numpy().transpose(2, 1, 0)
    mk_x = (mk.cpu().numpy() > 0).transpose(2, 1, 0)
    p = dict(full["params"])
    p["smoothPenalty"] = O.smooth_penalty_factor(ref_x, mk_x) * 0.08
    opt = dict(layer=p["layer"], iter=p["iter"], r=p["r"], zRatio=1.0, movRange=5.0, tol=1e-3,
               save_ite=10 ** 9, mask_ref=mk_x, mask_mov=np.zeros(mk_x.shape, bool),
               smoothPenalty=p["smoothPenalty"])
    ex = {}
    want, _, _, _ = O.get_motion(mov_x, ref_x, dict(opt), extras=ex, keep_images=False)
    free = D.get_motion(mov, tmpl, mask_ref=mk, **p)
    print(f"schedule oracle {ex['schedule']}, GPU {free.schedule()}", flush=True)
    assert free.schedule() == ex["schedule"]
    forced = D.get_motion(mov, tmpl, mask_ref=mk, schedule=ex["schedule"], **p)
    got = forced.motion.cpu().numpy().transpose(2, 1, 0, 3)
    assert np.abs(got - want).max() <= 1e-3
    warped = D.correct_motion(mov, forced.motion).cpu().numpy().transpose(2, 1, 0)
    assert rel_l2(warped, O.correct_motion(mov_x, got)) <= 1e-5
    m64, _, _, _ = O.get_motion(mov_x, ref_x, dict(opt), mode="fp64", keep_images=False)
    e = epe_stats(free.motion.cpu().numpy().transpose(2, 1, 0, 3), m64)
    print(f"EPE vs float64 oracle: {e}", flush=True)
    assert e["epe_mean"] <= 0.1
