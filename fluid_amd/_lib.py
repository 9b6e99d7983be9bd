"""ctypes binding of ``libpbf_hip.so`` (the C ABI of ``include/pbf.h``).

The library is built in-tree by ``fluid_amd.build``.  There is no fallback:
if the library is missing or cannot be loaded, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libpbf_hip.so")


class PbfParams(C.Structure):
    """Mirror of ``pbf_params`` (include/pbf.h)."""

    _fields_ = [
        ("h", C.c_float),
        ("rest_density", C.c_float),
        ("iterations", C.c_int32),
        ("epsilon", C.c_float),
        ("k_tensile", C.c_float),
        ("n_tensile", C.c_int32),
        ("dq", C.c_float),
        ("xsph_c", C.c_float),
        ("vort_eps", C.c_float),
        ("gravity", C.c_float),
        ("dt", C.c_float),
        ("box_lo", C.c_float * 3),
        ("box_hi", C.c_float * 3),
        ("wall_stop_mask", C.c_uint32),
        ("cell_scale", C.c_float),
        ("max_neighbors", C.c_int32),
        ("poly6_coef", C.c_float),
        ("spiky_coef", C.c_float),
        ("eps_d", C.c_float),
    ]


# Every symbol declared in include/pbf.h with its (restype, argtypes).
_P = C.c_void_p
_F = C.POINTER(C.c_float)
_U = C.POINTER(C.c_uint32)
SIGNATURES = {
    "pbf_params_default": (None, [C.POINTER(PbfParams)]),
    "pbf_create": (C.c_int, [C.POINTER(PbfParams), C.POINTER(_P)]),
    "pbf_create_batch": (
        C.c_int,
        [C.POINTER(PbfParams), C.c_int32, C.POINTER(C.c_int64), _F, _F, C.POINTER(_P)],
    ),
    "pbf_destroy": (C.c_int, [_P]),
    "pbf_load": (C.c_int, [_P, _F, _F, C.c_int64]),
    "pbf_step_dt": (C.c_int, [_P, C.c_float, C.c_int32]),
    "pbf_step": (C.c_int, [_P, C.c_int32]),
    "pbf_read": (C.c_int, [_P, _F, _F, _F]),
    "pbf_estimate_densities": (C.c_int, [_P, _F]),
    "pbf_density_at": (C.c_int, [_P, C.c_int32, _F, C.c_int64, _F]),
    "pbf_save_snapshot": (C.c_int, [_P, C.c_char_p]),
    "pbf_set_mesh": (C.c_int, [_P, _F, _F, C.c_int64, C.POINTER(C.c_int32), C.c_int64,
                               C.c_float]),
    "pbf_load_snapshot": (C.c_int, [_P, C.c_char_p]),
    "pbf_density_lattice": (C.c_int, [_P, C.c_int32, _F, _F, C.c_float,
                                      C.POINTER(C.c_int32), _F]),
    "pbf_step_device_noalloc": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_int64, C.c_float,
                                          C.c_int32]),
    "pbf_simulate_time": (C.c_double, [_P]),
    "pbf_count": (C.c_int64, [_P]),
    "pbf_last_error": (C.c_char_p, [_P]),
    "pbf_neighbors": (C.c_int, [_P, _U, _U, C.c_int32]),
    "pbf_cells": (C.c_int, [_P, _U]),
    "pbf_neighbor_stats": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "pbf_grid_dims": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "pbf_step_debug": (C.c_int, [_P, C.c_float, C.c_int32]),
    "pbf_adj_create": (C.c_int, [C.POINTER(PbfParams), C.c_int64, C.c_int32, C.c_int32,
                                 C.POINTER(_P)]),
    "pbf_adj_destroy": (C.c_int, [_P]),
    "pbf_adj_forward": (C.c_int, [_P, _F, _F, C.c_float, _F, _F]),
    "pbf_adj_backward": (C.c_int, [_P, _F, _F, _F, _F, C.POINTER(C.c_double)]),
    "pbf_adj_stats": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "pbf_adj_replay_check": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "pbf_adj_last_error": (C.c_char_p, [_P]),
    "pbf_set_schedule": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "pbf_step_async": (C.c_int, [_P, C.c_float, C.c_int32]),
    "pbf_sync": (C.c_int, [_P]),
    "pbf_step_profiled": (C.c_int, [_P, C.c_float, C.c_int32, _F, C.POINTER(C.c_int32), C.c_int32]),
    "pbf_debug_read": (C.c_int, [_P, C.c_int32, _F]),
    "pbf_debug_derived": (C.c_int, [_P, _F, C.c_int32]),
}

ERRORS = {
    0: "PBF_OK",
    -1: "PBF_ERR_INVALID",
    -2: "PBF_ERR_HIP",
    -3: "PBF_ERR_OVERFLOW",
    -4: "PBF_ERR_STATE",
    -5: "PBF_ERR_OUT_OF_BOX",
    -6: "PBF_ERR_NOMEM",
}

STAGE = {"predict": 1, "neighbors": 2, "lambda0": 3, "deltap0": 4, "solve": 5, "xsph": 6}
BUF = {"pred": 1, "vel": 2, "vxsph": 3, "omega": 4, "pos": 5}
PASSES = ["predict", "scan", "sort", "neighbors", "lambda", "deltap", "xsph", "vorticity"]
SCHEDULE = {"flat": 0, "graph": 1, "persistent": 2}

_lib = None


class PbfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m fluid_amd.build` "
                "(the PBF step has no CPU fallback)"
            )
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, ctx=None) -> int:
    if rc < 0:
        msg = lib().pbf_last_error(ctx)
        raise PbfError(rc, msg.decode() if msg else "")
    return rc
