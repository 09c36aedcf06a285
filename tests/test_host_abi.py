"""Host-side checks that need no GPU: the C ABI library loads and exports
what the header declares, the ctypes structures match the header, and the
drop-in module's argument handling mirrors the reference."""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from wholistic_registration_amd import _lib
from wholistic_registration_amd import calFlow3d_Wei_v1 as drop_in


def test_library_exports_every_header_symbol():
    names = _lib.header_symbols()
    assert {"wr_get_motion", "wr_correct_motion", "wr_create", "wr_destroy"} <= set(names)
    lib = _lib.load()
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib._SIGNATURES), "ctypes signatures out of sync with the header"


def test_integration_guide_names_every_header_symbol():
    guide = open(os.path.join(os.path.dirname(_lib.PKG), "INTEGRATION.md")).read()
    missing = [n for n in _lib.header_symbols() if f"`{n}`" not in guide]
    assert not missing, missing


def _c_struct_fields(struct_name):
    text = open(_lib.HEADER).read()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct_name, struct_name), text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return re.findall(r"(\w+)(?:\[\w+\])?;", body)


@pytest.mark.parametrize("struct,py", [("wr_params", _lib.WrParams), ("wr_log", _lib.WrLog),
                                       ("wr_stats", _lib.WrStats), ("wr_xres_layer", _lib.WrXresLayer),
                                       ("wr_xres_trace", _lib.WrXresTrace)])
def test_ctypes_structs_match_header(struct, py):
    assert _c_struct_fields(struct) == [f[0] for f in py._fields_]


def test_wr_create_without_device_reports_error():
    """No HIP device in this container: wr_create must fail cleanly."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    lib = _lib.load()
    assert not lib.wr_create(0)
    assert "no HIP device" in _lib.last_error()


def test_null_arguments_are_rejected():
    lib = _lib.load()
    rc = lib.wr_correct_motion(None, None, 0, 1, 1, 1, None, None, None)
    assert rc == _lib.WR_E_ARG
    with pytest.raises(ValueError):
        _lib.check(rc, "wr_correct_motion")


def _option(shape):
    return {"layer": 1, "iter": 3, "r": 2, "zRatio": 1.0, "movRange": 5.0, "smoothPenalty": 1.0,
            "tol": 1e-3, "save_ite": 100, "mask_ref": np.zeros(shape, bool),
            "mask_mov": np.zeros(shape, bool)}


@pytest.mark.parametrize("key", drop_in.REQUIRED_KEYS)
def test_missing_option_key_raises_keyerror(key):
    vol = np.zeros((8, 8, 2), np.float32)
    opt = _option(vol.shape)
    del opt[key]
    with pytest.raises(KeyError):
        drop_in.getMotion(vol, vol, opt)


def test_masks_become_float32_arrays():
    """Side effect of the reference (calFlow3d:314-315): option masks are
    replaced by float32 arrays before anything else happens."""
    import torch
    vol = np.zeros((8, 8, 2), np.float32)
    opt = _option(vol.shape)
    if torch.cuda.is_available():
        pytest.skip("checks the host path before the device call")
    with pytest.raises(RuntimeError):
        drop_in.getMotion(vol, vol, opt)
    assert opt["mask_ref"].dtype == np.float32 and opt["mask_mov"].dtype == np.float32


def test_product_path_has_no_cpu_fallback():
    """The package never imports the oracle or scipy's interpolation."""
    pkg = os.path.dirname(_lib.__file__)
    for fn in os.listdir(pkg):
        if fn.endswith(".py"):
            src = open(os.path.join(pkg, fn)).read()
            assert "from oracle" not in src and "import oracle" not in src, fn
            if fn not in ("synth.py",):
                assert "map_coordinates" not in src, fn


def test_smooth_penalty_null_arguments_are_rejected():
    lib = _lib.load()
    out = ctypes.c_double()
    rc = lib.wr_smooth_penalty_factor(None, None, 0, 4, 4, 4, None, ctypes.byref(out), None)
    assert rc == _lib.WR_E_ARG


@pytest.mark.parametrize("name,args", [
    ("wr_net_conv3d", (None, None, None, None, None, 1, 4, 4, 8, 8, 16, 1, 0.1, None)),
    ("wr_net_conv3d_bwd_data", (None, None, None, None, 1, 4, 4, 8, 16, 16, None)),
    ("wr_net_conv3d_bwd_weight", (None, None, None, None, None, 1, 4, 4, 8, 8, 16, None)),
    ("wr_net_leaky_bwd", (None, None, None, 8, 0.1, None)),
    ("wr_net_down2", (None, None, None, 1, 4, 4, 8, 8, 0.125, None)),
    ("wr_net_up2", (None, None, None, 1, 4, 4, 8, 8, 1.0, 0, None)),
    ("wr_net_concat", (None, None, 8, None, 8, 16, None, None)),
    ("wr_net_upcat", (None, None, 8, None, 8, 1, 4, 4, 8, None, None)),
    ("wr_net_split", (None, None, 8, 8, 16, None, None, None)),
    ("wr_net_mse", (None, None, 16, None, 3, 16, 4.0, 1.0, None, None, None)),
    ("wr_net_adam", (None, None, None, None, None, None, 10, 1e-3, 0.9, 0.999, 1e-8, 1, None)),
    ("wr_net_to_bf16", (None, None, 10, None, None)),
    ("wr_net_input", (None, None, None, 10, None, None)),
    ("wr_net_output", (None, None, 16, 10, 4.0, None, None)),
    ("wr_net_output_cp", (None, None, 16, 4, 8, 32, 8, 32, 3, 4, 4, 8, 0, 4.0, None, None)),
    ("wr_net_input_padded", (None, None, None, 4, 8, 32, 4, 8, 32, None, None)),
    ("wr_net_input_padded_batch", (None, None, 2, None, 4, 8, 32, 4, 8, 32, None, None)),
    ("wr_net_output_cp_batch", (None, None, 2, 16, 4, 4, 8, 32, 8, 32, 3, 4, 4, 8, 0, 4.0, None, None)),
])
def test_network_entry_points_reject_null_arguments(name, args):
    """The learned initialisation's entry points validate before touching a
    device: a null context or buffer is WR_E_ARG with the function named."""
    lib = _lib.load()
    rc = getattr(lib, name)(*args)
    assert rc == _lib.WR_E_ARG
    assert name in _lib.last_error()
