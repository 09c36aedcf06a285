"""GPU parity of the harness's mask, statistics and channel kernels against the
oracle, through the C ABI (libwrhip.so)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import harness_oracle, mask_oracle

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _robust(data):
    from wholistic_registration_amd import device as D
    return D.robust_mean_std(_dev(data.ravel()))


@pytest.mark.parametrize("n", [1, 2, 100, 1001, 65537, 1_000_003])
def test_robust_stats_random(cuda, n):
    rng = np.random.default_rng(n)
    data = rng.normal(100.0, 15.0, size=n).astype(np.float32)
    data[rng.random(n) < 0.02] = 5e4                      # bright outliers
    thr, mu, sigma = _robust(data)
    assert thr == mask_oracle.percentile_threshold(data)  # exact order statistics + numpy lerp
    m_ref, s_ref = mask_oracle.robust_mean_std(data)
    sel = data[data <= thr].astype(np.float64)
    assert mu == pytest.approx(sel.mean(), rel=1e-6, abs=1e-6)
    assert mu == pytest.approx(float(m_ref), rel=2e-6, abs=1e-6)
    assert sigma == pytest.approx(float(s_ref), rel=2e-5, abs=1e-5)


def test_robust_stats_reference_vector(cuda):
    """tests/unit/test_preprocess.py:52-57 of the reference."""
    data = np.concatenate([np.ones(99, dtype=np.float32), np.array([1e6], dtype=np.float32)])
    thr, mu, sigma = _robust(data)
    assert mu == pytest.approx(1.0, abs=1e-6) and sigma == pytest.approx(0.0, abs=1e-6)
    assert thr == mask_oracle.percentile_threshold(data)


def test_robust_stats_uint16_ties_and_large_index(cuda):
    """uint16 data (many ties, converted to float32) and n above 2^24, where
    numpy's float32 index arithmetic rounds."""
    from wholistic_registration_amd import device as D
    rng = np.random.default_rng(7)
    n = (1 << 24) + 12345
    data = rng.poisson(120.0, size=n).astype(np.uint16)
    data[:5000] = 60000
    thr, mu, sigma = D.robust_mean_std(_dev(data))
    f = data.astype(np.float32)
    assert thr == mask_oracle.percentile_threshold(f)
    m_ref, s_ref = mask_oracle.robust_mean_std(f)
    assert mu == pytest.approx(float(m_ref), rel=2e-6)
    assert sigma == pytest.approx(float(s_ref), rel=2e-5)


def _scene(shape_xyz, seed, u16=False):
    rng = np.random.default_rng(seed)
    vol = rng.normal(100.0, 10.0, size=shape_xyz)
    for _ in range(max(1, vol.size // 400)):
        x, y, z = (int(rng.integers(0, s)) for s in shape_xyz)
        w = int(rng.integers(0, 4))
        vol[max(x - w, 0):x + w + 1, max(y - w, 0):y + w + 1, z] += rng.uniform(50, 500)
    if u16:
        return np.clip(np.rint(vol), 0, 65535).astype(np.uint16)
    return vol.astype(np.float32)


@pytest.mark.parametrize("shape_xyz,thres,u16", [((70, 45, 3), 5, False), ((128, 130, 5), 3, True),
                                                  ((1, 1, 1), 5, False), ((65, 17, 2), 4.7, False),
                                                  ((200, 3, 4), 2.5, True)])
def test_get_mask_matches_oracle(cuda, shape_xyz, thres, u16):
    from wholistic_registration_amd import mask as M
    vol = _scene(shape_xyz, 11 + shape_xyz[0], u16)
    got = M.getMask(vol, thres)
    _, mu, sigma = _robust(vol.transpose(2, 1, 0).astype(np.float32))   # the kernel's (Z, Y, X) order
    want = mask_oracle.get_mask(vol, thres, mu_sigma=(mu, sigma))
    assert got.dtype == bool and got.shape == vol.shape
    assert np.array_equal(got, want)
    # with the oracle's own (numpy float32) statistics: the same mask here
    assert np.array_equal(got, mask_oracle.get_mask(vol, thres))


@pytest.mark.parametrize("shape_xyz,density,rng_lo_hi", [
    ((64, 50, 7), 0.05, (1, np.inf)), ((64, 50, 7), 0.3, (5, 4000)), ((33, 65, 17), 0.6, (2, 10)),
    ((300, 7, 3), 0.35, (3, 3)), ((1, 1, 1), 1.0, (1, 1)), ((40, 40, 40), 0.25, (0, 0))])
def test_bwareafilt3_matches_oracle(cuda, shape_xyz, density, rng_lo_hi):
    from wholistic_registration_amd import device as D
    from wholistic_registration_amd import mask as M
    rng = np.random.default_rng(int(density * 1000) + shape_xyz[0])
    m = rng.random(shape_xyz) < density
    got = M.bwareafilt3_wei(m, rng_lo_hi)
    want = mask_oracle.bwareafilt3(m, rng_lo_hi)
    assert np.array_equal(got, want)
    _, n = D.bwareafilt3(_dev(m.transpose(2, 1, 0).astype(np.uint8)), rng_lo_hi, count=True)
    assert n == mask_oracle.n_components(m)


def test_bwareafilt3_one_large_component(cuda):
    """A solid ellipsoid of ~10^5 voxels is one component (many concurrent unions)."""
    from wholistic_registration_amd import mask as M
    z, y, x = np.mgrid[0:30, 0:90, 0:110]
    m = ((x - 55) / 50.0) ** 2 + ((y - 45) / 40.0) ** 2 + ((z - 15) / 14.0) ** 2 <= 1.0
    m = m.transpose(2, 1, 0)
    area = int(m.sum())
    assert np.array_equal(M.bwareafilt3_wei(m, (area, area)), m)
    assert not M.bwareafilt3_wei(m, (area + 1, np.inf)).any()


def test_mask_functions_on_device_tensors(cuda):
    """Device (x, y, z) views in, device views out (float32 0/1 from the filter)."""
    import torch

    from wholistic_registration_amd import mask as M
    vol = _scene((48, 40, 3), 5)
    t = _dev(vol.transpose(2, 1, 0)).permute(2, 1, 0)
    g = M.getMask(t, 4)
    f = M.bwareafilt3_wei(g, (2, 50))
    assert f.dtype == torch.float32 and tuple(f.shape) == vol.shape
    want = mask_oracle.bwareafilt3(M.getMask(vol, 4), (2, 50))
    assert np.array_equal(f.cpu().numpy() > 0, want)


@pytest.mark.parametrize("ta,tb", [("u16", "u16"), ("f32", "u16"), ("f32", "f32"), ("u16", "f64")])
def test_mean_sq_diff(cuda, ta, tb):
    from wholistic_registration_amd import device as D
    rng = np.random.default_rng(5)
    kinds = {"u16": np.uint16, "f32": np.float32, "f64": np.float64}
    a = rng.integers(0, 3000, size=(5, 33, 41)).astype(kinds[ta])
    b = rng.integers(0, 3000, size=(5, 33, 41)).astype(kinds[tb])
    got = D.mean_sq_diff(_dev(a), _dev(b))
    want = mask_oracle.mean_sq_diff(a, b)
    assert type(got) is type(want)
    if ta == tb == "u16":
        assert got == want                       # exact integer sums
    else:
        assert got == pytest.approx(want, rel=2e-6)


@pytest.mark.parametrize("method", ["raw", "sqrt", "log2", "log10"])
@pytest.mark.parametrize("kind,k", [("u16", 3), ("u16", 250), ("f32", 0.3), ("f32", 250)])
def test_channel_transform(cuda, method, kind, k):
    from wholistic_registration_amd import device as D
    rng = np.random.default_rng(9)
    a = rng.integers(0, 4000, size=(3, 17, 19))
    a.flat[:3] = [0, 65535, 1]
    a = a.astype(np.uint16) if kind == "u16" else a.astype(np.float32)
    got = D.channel_transform(_dev(a), k, method).cpu().numpy()
    want = harness_oracle.transform(a, k, method)
    if method in ("raw", "sqrt"):
        assert np.array_equal(got, want.astype(np.float32))
    else:
        fin = np.isfinite(want)
        assert np.array_equal(fin, np.isfinite(got))
        assert np.allclose(got[fin], want[fin], rtol=4e-7, atol=0)


def test_normalize_std_and_frames_mean(cuda):
    from wholistic_registration_amd import device as D
    rng = np.random.default_rng(4)
    img = rng.normal(7.0, 3.0, size=(4, 20, 30)).astype(np.float32)
    mean, std = D.mean_std(_dev(img))
    assert mean == pytest.approx(float(np.mean(img)), rel=1e-6)
    assert std == pytest.approx(float(np.std(img)), rel=1e-5)
    got = D.normalize_std(np.float32(50.0), np.float32(4.0), _dev(img)).cpu().numpy()
    want = harness_oracle.normalize_std(np.float32(50.0), np.float32(4.0), img)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-4)
    for kind in (np.uint16, np.float32):
        fr = rng.integers(0, 60000, size=(5, 3, 11, 13)).astype(kind)
        idx = [3, 0, 4]
        got = D.frames_mean([_dev(fr[i]) for i in idx]).cpu().numpy()
        want = fr[idx].mean(axis=0)
        assert got.dtype == want.dtype
        assert np.array_equal(got, want)


def test_pick_initial_reference_matches_reference_vectors(cuda):
    import os

    from tests.golden.harness_cases import PICK_CASES, pick_frames
    from wholistic_registration_amd import reference as R
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "pick_reference.npz"))
    for i, case in enumerate(PICK_CASES):
        frames = pick_frames(i)
        img, top = R.pick_initial_reference(frames, max_corr_frames=case["max_corr"])
        assert np.array_equal(top, z[f"top_{i}"]), i
        assert img.dtype == z[f"ref_{i}"].dtype, i
        assert np.array_equal(img, z[f"ref_{i}"]), i
