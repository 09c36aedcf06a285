"""GPU tests of the registration harness (wbi_registration_3d) and the
reference-block scheduler helper, through the C ABI."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import harness_oracle, metrics
from tests.golden.harness_cases import harness_stack

pytestmark = pytest.mark.gpu


def _config(meta=None, layer=1, r=3, n_iter=6, thres=5, mask_range=(5, 4000), dual=False, auto=False):
    return {
        "MetaData": meta if meta is not None else {"zRatio": 1.0, "Dim": 3},
        "reference": {"pick_reference_auto": auto, "chunk_size": 2},
        "channels": {"dual_channel": dual, "k": 250 if dual else 0, "function": "log10"},
        "mask": {"maskRange": list(mask_range), "thresFactor": thres},
        "pyramid": {"r": r, "layer": layer, "iter": n_iter, "tolerance": 1e-3, "smoothPenalty": 0.08},
    }


def _write_toml(tmp_path, cfg):
    """A TOML file with the config's tables (no toml writer is installed)."""
    def val(v):
        if isinstance(v, bool):
            return "true" if v else "false"
        if isinstance(v, str):
            return f'"{v}"'
        if isinstance(v, (list, tuple)):
            return "[" + ", ".join(val(x) for x in v) + "]"
        return repr(v)
    lines = []
    for table, body in cfg.items():
        lines.append(f"[{table}]")
        lines += [f"{k} = {val(v)}" for k, v in body.items()]
    path = tmp_path / "config.toml"
    path.write_text("\n".join(lines) + "\n")
    return str(path)


def _run_capturing_option(monkeypatch, config_path, preset_zratio):
    """The reference's own harness test, run against this package: the solver
    stubbed out, the option dict handed to getMotion captured
    (tests/regressions/test_registration_option_assembly.py:22-45)."""
    from wholistic_registration_amd import calFlow3d_Wei_v1, option, registration
    monkeypatch.setitem(option, "zRatio", preset_zratio)
    seen = {}

    def fake_getMotion(dat_mov, dat_ref, opt):
        seen.update(opt)
        return np.zeros(tuple(dat_mov.shape) + (3,), dtype=np.float32), 0.0, None, {}

    monkeypatch.setattr(calFlow3d_Wei_v1, "getMotion", fake_getMotion)
    monkeypatch.setattr(calFlow3d_Wei_v1, "correctMotion", lambda data, motion: data)
    rng = np.random.default_rng(0)
    moving = rng.uniform(0.0, 1000.0, size=(1, 4, 12, 12)).astype(np.float32)
    registration.wbi_registration_3d(moving, config_path, reference_image=moving[0], verbose=False)
    return seen


def test_zratio_reaches_the_flow_solver(cuda, tmp_path, monkeypatch):
    """Reference regression B-071: MetaData.zRatio reaches option."""
    path = _write_toml(tmp_path, {k: v for k, v in _config({"zRatio": 5.5, "Dim": 3}).items()})
    seen = _run_capturing_option(monkeypatch, path, preset_zratio=27.693)
    assert seen["zRatio"] == 5.5


def test_zratio_falls_back_to_the_module_default(cuda, tmp_path, monkeypatch):
    """A config without MetaData.zRatio keeps the import-time default."""
    path = _write_toml(tmp_path, _config({"Dim": 3}))
    seen = _run_capturing_option(monkeypatch, path, preset_zratio=27.693)
    assert seen["zRatio"] == 27.693


@pytest.mark.parametrize("u16", [False, True])
def test_harness_matches_oracle(cuda, monkeypatch, u16):
    """Masks bit-exact, motion within 1e-3 voxel (schedules forced from the
    GPU run), corrected volumes within 1e-5 relative L2, errors within 1e-5."""
    from wholistic_registration_amd import calFlow3d_Wei_v1, option, registration
    tmpl, movs = harness_stack(uint16=u16)
    cfg = _config(layer=1, r=3, n_iter=6)
    real = calFlow3d_Wei_v1.getMotion
    scheds, masks = [], []

    def recording_getMotion(dat_mov, dat_ref, opt, verbose=False):
        masks.append(np.asarray(opt["mask_mov"].cpu()) > 0)
        out = real(dat_mov, dat_ref, opt, verbose)
        log = out[3]
        scheds.append({L: (len(log[f"layer_{L}"]["currentError"]), len(log[f"layer_{L}"]["max_diff_motion"]))
                       for L in range(opt["layer"] + 1)})
        return out

    monkeypatch.setattr(calFlow3d_Wei_v1, "getMotion", recording_getMotion)
    mem, ca, dat_ref, errors, motion = registration.wbi_registration_3d(
        movs, cfg, reference_image=tmpl, verbose=False)
    mask_ref = np.asarray(option["mask_ref"].cpu()) > 0
    assert mem.shape == movs.shape and mem.dtype == np.float32
    assert motion.shape == movs.shape + (3,)
    assert ca.size == 0 and dat_ref.shape == tmpl.transpose(2, 1, 0).shape
    w_mem, w_mot, w_err, w_masks = harness_oracle.wbi_registration_3d(movs, cfg, tmpl, schedules=scheds)
    assert np.array_equal(mask_ref, w_masks[0])
    assert w_masks[0].any()                           # the bright spots are masked
    for i in range(len(movs)):
        assert np.array_equal(masks[i], w_masks[1 + i]), i
        assert np.abs(motion[i] - w_mot[i]).max() <= 1e-3
        assert metrics.rel_l2(mem[i], w_mem[i]) <= 1e-5
        for key in ("initial_error", "eventual_error"):
            assert float(errors[i][key]) == pytest.approx(float(w_err[i][key]), rel=1e-5)
        if not u16:   # uint16 - uint16 wraps modulo 2^16 in the reference's initial error
            assert errors[i]["eventual_error"] < errors[i]["initial_error"]


def test_harness_dual_channel_auto_reference(cuda):
    """Dual channel with the automatic reference: the calcium channel enters the
    moving volume through transform + normalize_std, the reference through
    pick_initial_reference; the harness runs end to end and registers."""
    from wholistic_registration_amd import registration
    tmpl, movs = harness_stack(seed=901, frames=4)
    ca = (movs * 0.5).astype(np.float32)
    cfg = _config(dual=True, auto=True)
    mem, ca_out, dat_ref, errors, motion = registration.wbi_registration_3d(
        movs, cfg, moving_Ca_image=ca, verbose=False)
    assert mem.shape == movs.shape and ca_out.shape == movs.shape
    assert all(np.isfinite(float(e["eventual_error"])) for e in errors)
    # reference = membrane mean of the chosen frames + transformed calcium mean
    ref_mem, top = harness_oracle.pick_initial_reference(movs, max_corr_frames=2)
    ca_ref = harness_oracle.transform(np.mean(ca.reshape(len(movs), -1)[top], axis=0).reshape(movs.shape[1:]),
                                      250, "log10")
    want = (ref_mem + ca_ref).transpose(2, 1, 0)
    assert np.allclose(dat_ref, want, rtol=1e-5, atol=1e-3)


def test_compute_reference_from_block(cuda):
    from wholistic_registration_amd import reference as R
    tmpl, movs = harness_stack(seed=902, frames=6)
    cfg = _config()
    ref, idx = R.compute_reference_from_block(movs, cfg)
    w_ref, w_idx = harness_oracle.pick_initial_reference(movs, max_corr_frames=3)
    assert np.array_equal(idx, w_idx)
    assert np.array_equal(ref, w_ref)


def test_register_series_end_to_end(cuda):
    """The whole time-series schedule on the GPU: every frame registered once,
    and the middle block's outputs equal the harness run directly on its
    batches against the block's reference (7 frames: middle window 2..5)."""
    from wholistic_registration_amd import reference as R
    from wholistic_registration_amd import registration, series
    tmpl, movs = harness_stack(seed=903, frames=7)
    cfg = _config(layer=1, r=3, n_iter=5)
    cfg.update({
        "frames": {"frames": list(range(7)), "frame_downsample": 1},
        "processing": {"batch_size": 2},
        "save_config": {"save_ref": True, "save_motion": True},
    })
    cfg["MetaData"]["fps"] = 1.0
    cfg["reference"].update({"time_measurement": "frame", "mid_window_size": 3, "window_size": 2,
                             "mid_stride": 1, "reference_chunk": 2})
    out = {}
    series.register_series(cfg, lambda ids, ch: movs[list(ids)],
                           lambda kind, fid, a: out.setdefault(kind, {}).update({fid: a}))
    assert sorted(out["membrane"]) == list(range(7)) and sorted(out["motion"]) == list(range(7))
    assert all(np.isfinite(a).all() for a in out["membrane"].values())
    ref, _ = R.compute_reference_from_block(movs[[2, 3, 4]], cfg)
    for batch in ([2, 3], [4, 5]):
        mem, _, _, _, motion = registration.wbi_registration_3d(movs[batch], cfg, ref, verbose=False)
        for k, fid in enumerate(batch):
            assert np.array_equal(out["membrane"][fid], mem[k])
            assert np.array_equal(out["motion"][fid], motion[k])
