"""ctypes declarations of include/geostat_hip.h and library loading.

The library must exist in-tree (geostat-rs_amd/build/libgeostat_hip.so);
there is no fallback path: importing the predictor API without it raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent
# GEOSTAT_HIP_LIB selects another build of the same library (the
# host-sanitized one in build/asan/, used by the sanitizer tests)
LIB_PATH = Path(os.environ.get("GEOSTAT_HIP_LIB") or PKG_ROOT / "build" / "libgeostat_hip.so")
SYNTH_PATH = PKG_ROOT / "build" / "libgs_synth.so"

GS_NONE = -1
GS_MAX_STRUCTURES = 8

# gs_status
OK, INSUFFICIENT_DATA, DIMENSION_MISMATCH, SINGULAR_SYSTEM = 0, 1, 2, 3
INVALID_PARAMETER, NO_NEIGHBORS, DUPLICATE_POINTS = 4, 5, 6
DEVICE_ERROR, OUT_OF_MEMORY = 100, 101

PREC_FP64, PREC_FP32_REFINED = 0, 1


class gs_anisotropy(C.Structure):
    _fields_ = [("azimuth_deg", C.c_double), ("ratio", C.c_double), ("ratio_z", C.c_double),
                ("dip_deg", C.c_double), ("rake_deg", C.c_double)]


class gs_structure(C.Structure):
    _fields_ = [("kind", C.c_int32), ("param", C.c_double), ("sill", C.c_double),
                ("range", C.c_double), ("has_anis", C.c_int32), ("anis", gs_anisotropy)]


class gs_model(C.Structure):
    _fields_ = [("nugget", C.c_double), ("n_structures", C.c_int32),
                ("structures", gs_structure * GS_MAX_STRUCTURES)]


class gs_kconfig(C.Structure):
    _fields_ = [("method", C.c_int32), ("simple_mean", C.c_double), ("degree", C.c_int32),
                ("max_neighbors", C.c_int64), ("min_neighbors", C.c_int64),
                ("max_per_octant", C.c_int64), ("has_search_radius", C.c_int32),
                ("search_radius", C.c_double), ("has_anisotropic_search", C.c_int32),
                ("anisotropic_search", gs_anisotropy)]


_P = C.c_void_p
_DP = C.POINTER(C.c_double)
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every function declared in include/geostat_hip.h
SIGNATURES = {
    "gs_ctx_create": (C.c_int32, [C.c_int32, C.POINTER(_P)]),
    "gs_ctx_destroy": (None, [_P]),
    "gs_ctx_synchronize": (C.c_int32, [_P]),
    "gs_last_error": (C.c_char_p, []),
    "gs_last_solve_kernel": (C.c_char_p, []),
    "gs_abi_version": (C.c_int32, []),
    "gs_kriging_new": (C.c_int32, [_P, C.c_int32, _DP, _DP, C.c_int64, C.POINTER(gs_model),
                                   C.POINTER(gs_kconfig), C.POINTER(_P), _I64P]),
    "gs_kriging_free": (None, [_P]),
    "gs_kriging_predict_many": (C.c_int32, [_P, _DP, C.c_int64, C.c_int32, _DP, _DP, _DP, _I32P]),
    "gs_kriging_predict_many_device": (C.c_int32, [_P, _P, C.c_int64, C.c_int32, _P, _P, _P, _P, _P]),
    "gs_kriging_predict_many_noalloc": (C.c_int32, [_P, _P, C.c_int64, C.c_int32, _P, _P, _P, _P, _P]),
    "gs_kriging_graph_create": (C.c_int32, [_P, _P, C.c_int64, C.c_int32, _P, _P, _P, _P,
                                            C.POINTER(_P)]),
    "gs_graph_launch": (C.c_int32, [_P, _P]),
    "gs_graph_free": (None, [_P]),
    "gs_kriging_neighbors": (C.c_int32, [_P, _DP, C.c_int64, C.c_int64, _I32P, _I32P]),
    "gs_kriging_max_neighbors": (C.c_int64, [_P]),
    "gs_kriging_set_timing": (C.c_int32, [_P, C.c_int32]),
    "gs_kriging_timing_summary": (C.c_int32, [_P, _I64P, _DP, _DP]),
    "gs_maxmin_order": (C.c_int32, [C.c_int32, _DP, C.c_int64, _I64P]),
    "gs_maxmin_order_gpu": (C.c_int32, [C.c_void_p, C.c_int32, _DP, C.c_int64, _I64P, _I64P]),
    "gs_vecchia_plan_new": (C.c_int32, [_P, C.c_int32, _DP, C.c_int64, C.c_int64, _I64P,
                                        C.POINTER(_P)]),
    "gs_vecchia_plan_free": (None, [_P]),
    "gs_vecchia_plan_get": (C.c_int32, [_P, _I64P, _I32P, _I32P]),
    "gs_vecchia_loglik_batch": (C.c_int32, [_P, _DP, C.POINTER(gs_model), C.c_int32, C.c_int32,
                                            _DP, _I32P]),
    "gs_vecchia_loglik_grad_batch": (C.c_int32, [_P, _DP, C.POINTER(gs_model), C.c_int32,
                                                 C.c_int32, _DP, _DP, _I32P]),
    "gs_vecchia_loglik": (C.c_int32, [_P, C.c_int32, _DP, _DP, C.c_int64, C.POINTER(gs_model),
                                      C.c_int64, _I64P, _DP]),
    "gs_kriging_new_with_measurement_error": (C.c_int32, [_P, C.c_int32, _DP, _DP, C.c_int64,
                                                           C.POINTER(gs_model),
                                                           C.POINTER(gs_kconfig), _DP,
                                                           C.POINTER(_P), _I64P]),
    "gs_kriging_new_with_external_drift": (C.c_int32, [_P, C.c_int32, _DP, _DP, C.c_int64,
                                                        C.POINTER(gs_model),
                                                        C.POINTER(gs_kconfig), C.c_int32, _DP,
                                                        C.POINTER(_P), _I64P]),
    "gs_kriging_predict_many_with_drift": (C.c_int32, [_P, _DP, C.c_int64, _DP, C.c_int32, _DP,
                                                        _DP, _DP, _I32P]),
    "gs_kriging_predict_block_many": (C.c_int32, [_P, _DP, C.c_int64, _DP, C.c_int32, C.c_int32,
                                                   _DP, _DP, _DP, _I32P]),
    "gs_cv_leave_one_out": (C.c_int32, [_P, C.c_int32, _DP, _DP, C.c_int64, C.POINTER(gs_model),
                                        C.POINTER(gs_kconfig), _DP, _DP]),
    "gs_cv_leave_one_out_with_drift": (C.c_int32, [_P, C.c_int32, _DP, _DP, C.c_int64,
                                                   C.POINTER(gs_model), C.POINTER(gs_kconfig),
                                                   C.c_int32, _DP, _DP, _DP]),
    "gs_cv_k_fold": (C.c_int32, [_P, C.c_int32, _DP, _DP, C.c_int64, C.POINTER(gs_model),
                                 C.POINTER(gs_kconfig), C.c_int64, C.c_uint64, _DP, _DP]),
    "gs_cv_block": (C.c_int32, [_P, C.c_int32, _DP, _DP, C.c_int64, C.POINTER(gs_model),
                                C.POINTER(gs_kconfig), _I64P, _DP, _DP]),
    "gs_server_create": (C.c_int32, [_P, C.c_int32, C.c_int64, C.POINTER(_P)]),
    "gs_server_destroy": (None, [_P]),
    "gs_server_submit_krige": (C.c_int32, [_P, _P, _DP, C.c_int64, C.c_int32, _DP, _DP, _DP,
                                           _I32P, C.POINTER(C.c_int64)]),
    "gs_server_submit_vecchia": (C.c_int32, [_P, _P, _DP, C.POINTER(gs_model), C.c_int32,
                                             C.c_int32, _DP, _I32P, C.POINTER(C.c_int64)]),
    "gs_server_wait": (C.c_int32, [_P, C.c_int64, _DP, _DP]),
    "gs_model_validate": (C.c_int32, [C.POINTER(gs_model)]),
    "gs_vecchia_mle": (C.c_int32, [_P, _DP, C.c_int32, C.c_double, C.POINTER(gs_model), _DP]),
}

_lib = None
_synth = None


def lib() -> C.CDLL:
    """The loaded libgeostat_hip.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python geostat-rs_amd/build_lib.py` "
                "(there is no CPU fallback)")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def synth() -> C.CDLL:
    global _synth
    if _synth is None:
        if not SYNTH_PATH.exists():
            raise RuntimeError(f"{SYNTH_PATH} is missing: build with geostat-rs_amd/build_lib.py")
        h = C.CDLL(str(SYNTH_PATH))
        h.gs_synth_raw.argtypes = [C.c_uint64, C.c_int64, C.POINTER(C.c_uint64)]
        h.gs_synth_uniform.argtypes = [C.c_uint64, C.c_int64, _DP]
        h.gs_synth_points.argtypes = [C.c_uint64, C.c_int64, C.c_int, _DP, C.c_int, C.c_double,
                                      _DP, _DP]
        h.gs_synth_vecchia_field.argtypes = [C.c_uint64, C.c_int64, C.c_int, _DP, _I64P, _I32P,
                                             _I32P, C.c_int64, C.c_double, C.c_double,
                                             C.c_double, C.c_int, _DP]
        _synth = h
    return _synth


def last_error() -> str:
    return lib().gs_last_error().decode(errors="replace")


def last_solve_kernel() -> str:
    """Kernel family of the last solve launched on this thread."""
    return lib().gs_last_solve_kernel().decode(errors="replace")


def dptr(a):
    return a.ctypes.data_as(_DP)


def i32ptr(a):
    return a.ctypes.data_as(_I32P)


def i64ptr(a):
    return a.ctypes.data_as(_I64P)
