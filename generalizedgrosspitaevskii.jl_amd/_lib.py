"""ctypes binding of libgpe_hip.so (the C ABI declared in include/gpe_hip.h).

The library is loaded from this package directory.  There is no fallback: if
the shared library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpe_hip.so")

GPE_OK = 0
GPE_ERR_INVALID = -1
GPE_ERR_UNSUPPORTED = -2
GPE_ERR_DEVICE = -3
GPE_ERR_STATE = -4
GPE_ERR_NOMEM = -5

GPE_C64 = 0
GPE_C128 = 1

GPE_FORM_NONE = 0
GPE_FORM_SCALAR = 1
GPE_FORM_DIAG = 2
GPE_FORM_FULL = 3

GPE_PUMP_NONE = 0
GPE_PUMP_TABLE = 1
GPE_PUMP_SEPARABLE = 2
GPE_PUMP_SAMPLES = 3

GPE_NOISE_NONE = 0
GPE_NOISE_COMPLEX = 1
GPE_NOISE_REAL = 2

GPE_SCHEDULE_AUTO = 0
GPE_SCHEDULE_LAUNCHES = 1
GPE_SCHEDULE_GRAPH = 2

GPE_PLAN_ESTIMATE = 0
GPE_PLAN_MEASURE = 1
GPE_PLAN_AUTO = 2


class ProblemDesc(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32),
        ("ncomp", ctypes.c_int32),
        ("n", ctypes.c_int64 * 3),
        ("L", ctypes.c_double * 3),
        ("batch", ctypes.c_int64),
        ("dtype", ctypes.c_int32),
        ("device", ctypes.c_int32),
    ]


class GpeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libgpe_hip error {code}: {msg}")
        self.code = code


class GpeUnsupported(GpeError, NotImplementedError):
    pass


P = ctypes.c_void_p
D = ctypes.POINTER(ctypes.c_double)
I64 = ctypes.c_int64
I32 = ctypes.c_int32

# name -> (restype, argtypes); every function declared in include/gpe_hip.h
SIGNATURES = {
    "gpe_version": (ctypes.c_int, []),
    "gpe_last_error": (ctypes.c_char_p, []),
    "gpe_create": (ctypes.c_int, [ctypes.POINTER(ProblemDesc), ctypes.POINTER(P)]),
    "gpe_destroy": (ctypes.c_int, [P]),
    "gpe_set_dispersion": (ctypes.c_int, [P, I32, P]),
    "gpe_set_potential": (ctypes.c_int, [P, I32, P]),
    "gpe_set_nonlinearity": (ctypes.c_int, [P, P, P]),
    "gpe_set_nonlinearity_matrix": (ctypes.c_int, [P, P, P]),
    "gpe_set_pump": (ctypes.c_int, [P, I32, P, P, I64]),
    "gpe_set_noise": (ctypes.c_int, [P, I32, P, ctypes.c_uint64, ctypes.c_uint64]),
    "gpe_set_observables": (ctypes.c_int, [P, P, P]),
    "gpe_resolve_timestepping": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, I64, D,
                                                 ctypes.POINTER(I64)]),
    "gpe_pump_times": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, I64, P, I64]),
    "gpe_solve": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.c_double, ctypes.c_double, ctypes.c_double, I64, I32,
                                  ctypes.POINTER(P), P]),
    "gpe_set_schedule": (ctypes.c_int, [P, I32]),
    "gpe_get_schedule": (ctypes.c_int, [P, ctypes.POINTER(I32)]),
    "gpe_set_planner": (ctypes.c_int, [P, I32]),
    "gpe_set_plan": (ctypes.c_int, [P, I32, I32]),
    "gpe_get_plan": (ctypes.c_int, [P, ctypes.POINTER(I32), ctypes.POINTER(I32), ctypes.POINTER(I32), D]),
    "gpe_kspace_g2": (ctypes.c_int, [P, P, ctypes.c_double, P, P]),
    "gpe_windowed_g1": (ctypes.c_int, [P, P, P, P, P]),
    "gpe_density_g2_x": (ctypes.c_int, [P, P, ctypes.c_double, P, P]),
    "gpe_get_field": (ctypes.c_int, [P, ctypes.POINTER(P)]),
    "gpe_last_solve_ms": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_float)]),
    "gpe_advance": (ctypes.c_int, [P, I64, ctypes.POINTER(ctypes.c_float)]),
    "gpe_profile_kernels": (ctypes.c_int, [P, I64, D, D, D]),
    "gpe_steady_state": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.c_double, ctypes.c_double, I64, I64,
                                         ctypes.POINTER(I64), D, P, I64, ctypes.POINTER(I64),
                                         ctypes.POINTER(ctypes.c_float)]),
    "gpe_fno_gemm": (ctypes.c_int, [P, P, P, I64, I64, I64, P, I64, ctypes.c_float, ctypes.c_float, P]),
    "gpe_fno_gemm_f32": (ctypes.c_int, [P, P, P, I64, I64, I64, P, I64, ctypes.c_float, ctypes.c_float, P]),
    "gpe_fno_bias_act": (ctypes.c_int, [P, P, P, P, I64, I64, I64, I32, I32, P]),
    "gpe_fno_act_backward": (ctypes.c_int, [P, P, P, I64, P]),
    "gpe_fno_axpby": (ctypes.c_int, [P, P, ctypes.c_float, ctypes.c_float, I64, P]),
    "gpe_fno_adam": (ctypes.c_int, [P, P, P, P, I64, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     I64, P]),
    "gpe_sweep_create": (ctypes.c_int, [I64, P, P, I32, I32, ctypes.POINTER(P)]),
    "gpe_sweep_destroy": (ctypes.c_int, [P]),
    "gpe_sweep_set_dispersion": (ctypes.c_int, [P, P]),
    "gpe_sweep_set_nonlinearity": (ctypes.c_int, [P, P, P]),
    "gpe_sweep_set_pump": (ctypes.c_int, [P, P, P, I64]),
    "gpe_sweep_set_noise": (ctypes.c_int, [P, I32, P, ctypes.c_uint64, ctypes.c_uint64]),
    "gpe_sweep_solve": (ctypes.c_int, [P, P, ctypes.c_double, ctypes.c_double, ctypes.c_double, I64, I32, P, P]),
    "gpe_sweep_last_ms": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_float)]),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load libgpe_hip.so from the package directory (raises if absent)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python generalizedgrosspitaevskii.jl_amd/build.py` "
                "(or __graft_entry__.build()) first")
        # torch ships its own libamdhip64.so.  Loaded after this library, it
        # would start a second HIP runtime in the process (this library's from
        # /opt/rocm), which then finds no devices and cannot use the other's
        # pointers.  Loading torch first makes its runtime the one this
        # library binds to (same soname).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(code):
    if code != GPE_OK:
        msg = load().gpe_last_error().decode(errors="replace")
        if code == GPE_ERR_UNSUPPORTED:
            raise GpeUnsupported(code, msg)
        if code == GPE_ERR_INVALID:
            err = GpeError(code, msg)
            raise err
        raise GpeError(code, msg)
    return code


def ptr_array(ptrs):
    arr = (P * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr
