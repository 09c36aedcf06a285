"""ctypes binding of libwrhip.so (declared in include/wr_hip.h).

The library is required: every entry point raises when it is not built or
cannot be loaded.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
import re
import threading

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libwrhip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "wr_hip.h")

WR_OK, WR_E_ARG, WR_E_KEY, WR_E_HIP, WR_E_MEM = 0, -1, -2, -3, -4
WR_F32, WR_U16, WR_F64, WR_U8 = 0, 1, 2, 3
WR_MAX_LEVELS = 16


class WrParams(ctypes.Structure):
    _fields_ = [
        ("layer", ctypes.c_int),
        ("iter", ctypes.c_int),
        ("r", ctypes.c_int),
        ("timing", ctypes.c_int),
        ("zRatio", ctypes.c_double),
        ("movRange", ctypes.c_double),
        ("smoothPenalty", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("forced_records", ctypes.POINTER(ctypes.c_int)),
        ("forced_updates", ctypes.POINTER(ctypes.c_int)),
        ("tile_patches_x", ctypes.c_int),
        ("stop_kernels", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("reuse_template", ctypes.c_int),
        ("init_nx", ctypes.c_int),
        ("init_ny", ctypes.c_int),
        ("tile_patches_level", ctypes.POINTER(ctypes.c_int)),
        ("init_cp", ctypes.c_int),
        ("resample_start", ctypes.c_int),
        ("maxdiff_gather", ctypes.c_int),
        ("init_event", ctypes.c_void_p),
    ]


class WrLog(ctypes.Structure):
    _fields_ = [
        ("records", ctypes.c_int * WR_MAX_LEVELS),
        ("updates", ctypes.c_int * WR_MAX_LEVELS),
        ("nans", ctypes.c_int * WR_MAX_LEVELS),
        ("diff_error", ctypes.POINTER(ctypes.c_double)),
        ("penalty_error", ctypes.POINTER(ctypes.c_double)),
        ("current_error", ctypes.POINTER(ctypes.c_double)),
        ("max_diff_motion", ctypes.POINTER(ctypes.c_double)),
        ("final_error", ctypes.c_double),
    ]


class WrStats(ctypes.Structure):
    _fields_ = [
        ("lk_launches", ctypes.c_int * WR_MAX_LEVELS),
        ("tile_cols", ctypes.c_int),
        ("tile_rows", ctypes.c_int),
        ("lk_blocks", ctypes.c_int * WR_MAX_LEVELS),
        ("lk_exec", ctypes.c_int * WR_MAX_LEVELS),
        ("lk_ms", ctypes.c_double * WR_MAX_LEVELS),
        ("solve_fp64", ctypes.c_int * WR_MAX_LEVELS),
        ("start_from_cp", ctypes.c_int * WR_MAX_LEVELS),
    ]


class WrXresLayer(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("rx", ctypes.c_int), ("ry", ctypes.c_int), ("rz", ctypes.c_int),
        ("r", ctypes.c_int),
        ("iters", ctypes.c_int),
        ("smooth_penalty_sum", ctypes.c_double),
        ("mov_range", ctypes.c_double),
        ("z_ratio_hr", ctypes.c_double),
        ("tol", ctypes.c_double),
    ]


class WrXresTrace(ctypes.Structure):
    _fields_ = [
        ("records", ctypes.c_int), ("updates", ctypes.c_int), ("reason", ctypes.c_int),
        ("diff_error", ctypes.POINTER(ctypes.c_double)),
        ("penalty_error", ctypes.POINTER(ctypes.c_double)),
        ("current_error", ctypes.POINTER(ctypes.c_double)),
        ("final_diff", ctypes.c_double),
        ("final_penalty", ctypes.c_double),
        ("final_error", ctypes.c_double),
    ]


_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_LL = ctypes.c_longlong
_SIGNATURES = {
    "wr_create": ([_I], _VP),
    "wr_destroy": ([_VP], None),
    "wr_last_error": ([], ctypes.c_char_p),
    "wr_get_motion": ([_VP, _VP, _VP, _I, _I, _I, _I, _VP, _VP, _VP, ctypes.POINTER(WrParams),
                       _VP, _VP, ctypes.POINTER(WrLog), _VP], _I),
    "wr_get_motion_batch": ([_VP, _VP, ctypes.c_longlong, _VP, _I, _I, _I, _I, _I, _VP, _VP, ctypes.c_longlong,
                             _VP, ctypes.c_longlong, ctypes.POINTER(WrParams), _VP, _VP, _I, ctypes.c_longlong, _VP,
                             ctypes.POINTER(WrLog), _VP], _I),
    "wr_get_motion_warp": ([_VP, _VP, _VP, _I, _I, _I, _I, _VP, _VP, _VP, ctypes.POINTER(WrParams),
                            _VP, _VP, _I, _VP, ctypes.POINTER(WrLog), _VP], _I),
    "wr_correct_motion": ([_VP, _VP, _I, _I, _I, _I, _VP, _VP, _VP], _I),
    "wr_pyramid_level": ([_VP, _VP, _I, _I, _I, _I, _I, _VP, _VP], _I),
    "wr_imresize": ([_VP, _VP, _I, _I, _I, _I, _I, _I, _I, _VP, _VP], _I),
    "wr_get_stats": ([_VP, ctypes.POINTER(WrStats)], _I),
    "wr_smooth_penalty_factor": ([_VP, _VP, _I, _I, _I, _I, _VP, ctypes.POINTER(ctypes.c_double), _VP], _I),
    "wr_smooth_penalty_factor_noalloc": ([_VP, _VP, _I, _I, _I, _I, _VP, _VP, _VP], _I),
    "wr_robust_mean_std": ([_VP, _VP, _I, ctypes.c_longlong, ctypes.c_double, ctypes.POINTER(ctypes.c_float),
                            _VP], _I),
    "wr_get_mask": ([_VP, _VP, _I, _I, _I, _I, ctypes.c_double, ctypes.POINTER(ctypes.c_float), _VP, _VP], _I),
    "wr_bwareafilt3": ([_VP, _VP, _I, _I, _I, ctypes.c_double, ctypes.c_double, _VP, _VP,
                        ctypes.POINTER(ctypes.c_longlong), _VP], _I),
    "wr_mean_sq_diff": ([_VP, _VP, _I, _VP, _I, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double), _VP], _I),
    "wr_channel_transform": ([_VP, _VP, _I, ctypes.c_longlong, ctypes.c_double, _I, _I, _VP, _VP], _I),
    "wr_mean_std": ([_VP, _VP, ctypes.c_longlong, ctypes.POINTER(ctypes.c_float), _VP], _I),
    "wr_normalize_std": ([_VP, _VP, ctypes.c_longlong, ctypes.c_double, ctypes.c_double, _VP, _VP], _I),
    "wr_add_volumes": ([_VP, _VP, _I, _VP, ctypes.c_longlong, _VP, _VP], _I),
    "wr_frames_mean": ([_VP, ctypes.POINTER(_VP), _I, _I, ctypes.c_longlong, _VP, _VP], _I),
    "wr_frame_gram": ([_VP, _VP, _I, _I, _I, _I, _I, _I, ctypes.POINTER(ctypes.c_double), _VP], _I),
    "wr_project_planes_max": ([_VP, _VP, _VP, ctypes.c_longlong, _VP, _I, _I, _I, _I, ctypes.c_double, _I, _I, _I,
                               ctypes.c_double, _VP, _VP], _I),
    "wr_project_planes_weighted": ([_VP, _VP, _VP, ctypes.c_longlong, _VP, _I, _I, _I, _I, ctypes.c_double,
                                    ctypes.c_double, _I, ctypes.c_double, ctypes.c_double, _VP, _VP, _VP], _I),
    "wr_net_conv3d": ([_VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _I, ctypes.c_double, _VP], _I),
    "wr_net_conv3d_bwd_data": ([_VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _VP], _I),
    "wr_net_conv3d_bwd_weight": ([_VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _VP], _I),
    "wr_net_leaky_bwd": ([_VP, _VP, _VP, ctypes.c_longlong, ctypes.c_double, _VP], _I),
    "wr_net_down2": ([_VP, _VP, _VP, _I, _I, _I, _I, _I, ctypes.c_double, _VP], _I),
    "wr_net_up2": ([_VP, _VP, _VP, _I, _I, _I, _I, _I, ctypes.c_double, _I, _VP], _I),
    "wr_net_concat": ([_VP, _VP, _I, _VP, _I, ctypes.c_longlong, _VP, _VP], _I),
    "wr_net_upcat": ([_VP, _VP, _I, _VP, _I, _I, _I, _I, _I, _VP, _VP], _I),
    "wr_net_split": ([_VP, _VP, _I, _I, ctypes.c_longlong, _VP, _VP, _VP], _I),
    "wr_net_mse": ([_VP, _VP, _I, _VP, _I, ctypes.c_longlong, ctypes.c_double, ctypes.c_double, _VP, _VP, _VP],
                   _I),
    "wr_net_adam": ([_VP, _VP, _VP, _VP, _VP, _VP, ctypes.c_longlong, ctypes.c_double, ctypes.c_double,
                     ctypes.c_double, ctypes.c_double, _I, _VP], _I),
    "wr_net_to_bf16": ([_VP, _VP, ctypes.c_longlong, _VP, _VP], _I),
    "wr_net_input": ([_VP, _VP, _VP, ctypes.c_longlong, _VP, _VP], _I),
    "wr_net_input_padded": ([_VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _VP, _VP], _I),
    "wr_net_output": ([_VP, _VP, _I, ctypes.c_longlong, ctypes.c_double, _VP, _VP], _I),
    "wr_net_output_cp": ([_VP, _VP, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, ctypes.c_double, _VP, _VP], _I),
    "wr_net_input_padded_batch": ([_VP, _VP, _I, _VP, _I, _I, _I, _I, _I, _I, _VP, _VP], _I),
    "wr_net_output_cp_batch": ([_VP, _VP, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, ctypes.c_double,
                                _VP, _VP], _I),
    "wr_spline_coef": ([_VP, _VP, _I, _I, _I, _I, _VP, _VP], _I),
    "wr_spline_sample": ([_VP, _VP, _I, _I, _I, _VP, _LL, _D, _VP, _VP], _I),
    "wr_xres_optimize": ([_VP, ctypes.POINTER(WrXresLayer), _VP, _VP, _VP, _VP, _VP, _VP, _VP,
                          ctypes.POINTER(ctypes.c_int), _VP, _VP, _VP, _VP, _VP, _VP,
                          ctypes.POINTER(WrXresTrace), _VP], _I),
    "wr_xres_valid": ([_VP, _VP, _I, _I, _I, _VP, _I, _I, _I, _VP, _VP, ctypes.POINTER(_LL), _VP], _I),
    "wr_xres_bad_regions": ([_VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _D, _I, _I, _D, _VP,
                             ctypes.POINTER(ctypes.c_int), _VP], _I),
    "wr_xres_trap_mask": ([_VP, _VP, _VP, _I, _I, _I, _VP, _I, _I, _I, _VP, _D, _D, _D, _D, _VP,
                           ctypes.POINTER(_LL), _VP], _I),
    "wr_count_u8": ([_VP, _VP, _LL, ctypes.POINTER(_LL), _VP], _I),
    "wr_find_init_z": ([_VP, _VP, _I, _VP, _I, _I, _I, _D, _I, _I, _I, _D, _VP, _VP, ctypes.POINTER(_D),
                        ctypes.POINTER(_I), _VP], _I),
    "wr_xres_scale": ([_VP, _VP, _LL, _D, _D, _VP, _VP], _I),
    "wr_xres_threshold": ([_VP, _VP, _LL, _D, _VP, _VP], _I),
    "wr_xres_phase_identity": ([_VP, _I, _I, _I, _D, _D, _VP, _VP], _I),
}

_lib = None
_lock = threading.Lock()
_ctx = {}


def header_symbols(path=HEADER):
    """Function names declared in the C header."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wr_[a-z0-9_]+)\s*\(", text)))


def load():
    """Load libwrhip.so (raises if it is missing: there is no fallback)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m wholistic_registration_amd.build_lib`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
        return lib


def last_error():
    msg = load().wr_last_error()
    return msg.decode() if msg else ""


def check(rc, what):
    if rc == WR_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc == WR_E_ARG:
        raise ValueError(msg)
    if rc == WR_E_KEY:
        raise KeyError(msg)
    if rc == WR_E_MEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def context(device=0, slot=0):
    """The library context of one HIP device (created on first use).  Slot 1
    is a second context of the device for work enqueued on another stream at
    the same time as slot 0's (each context has its own workspace)."""
    lib = load()
    with _lock:
        h = _ctx.get((device, slot))
        if h is None:
            h = lib.wr_create(device)
            if not h:
                raise RuntimeError(f"wr_create({device}): {last_error()}")
            _ctx[(device, slot)] = h
        return h
