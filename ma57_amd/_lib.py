"""ctypes binding of libsymla_hip.so (declared in include/symla_hip.h).

The library is loaded from the package directory, where
``python -m ma57_amd.build`` puts it.  Loading fails loudly when it is
missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libsymla_hip.so"

SYMLA_OK = 0
SYMLA_ERR_INVALID = 1
SYMLA_ERR_LOGIC = 2
SYMLA_ERR_RUNTIME = 3
SYMLA_ERR_DEVICE = 4

ORDER_AMD, ORDER_METIS, ORDER_NATURAL, ORDER_USER = 0, 1, 2, 3
MODE_THRESHOLD, MODE_STATIC = 0, 1
PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2

i32p = C.POINTER(C.c_int)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)


class SymbolicInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("nsn", C.c_int), ("nnz_l", C.c_int64), ("pattern_total", C.c_int64),
                ("max_front", C.c_int)]


class FactorOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("pivot_threshold", C.c_double), ("zero_tolerance", C.c_double),
                ("relative_pivot_floor", C.c_double), ("expected_sign", i32p), ("initial_delta", C.c_double),
                ("max_retries", C.c_int), ("growth", C.c_double), ("static_rel_floor", C.c_double),
                ("static_abs_floor", C.c_double), ("precision", C.c_int)]


class FactorInfo(C.Structure):
    _fields_ = [("n_pos", C.c_int), ("n_neg", C.c_int), ("n_zero", C.c_int), ("singular", C.c_int),
                ("n_singular", C.c_int), ("total_perturbation", C.c_double), ("num_perturbed", C.c_int),
                ("retries_used", C.c_int), ("last_delta", C.c_double), ("delayed_fronts", C.c_int),
                ("repaired", C.c_int), ("precision_used", C.c_int)]


class RefineOpts(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("tolerance", C.c_double)]


class RefineInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("initial_relative_residual", C.c_double),
                ("final_relative_residual", C.c_double), ("backward_error", C.c_double)]


class DenseOpts(C.Structure):
    _fields_ = [("zero_tolerance", C.c_double), ("relative_pivot_floor", C.c_double), ("static_mode", C.c_int),
                ("static_delta", C.c_double), ("static_rel_floor", C.c_double), ("static_abs_floor", C.c_double),
                ("expected_sign", i32p), ("exact_panel", C.c_int), ("deferred_cb", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("small_front_max", C.c_int), ("assemble_tile", C.c_int), ("speculative_panel", C.c_int),
                ("speculative_min_rows", C.c_int), ("deferred_cb_min_rows", C.c_int)]


class TuningTrial(C.Structure):
    _fields_ = [("tuning", Tuning), ("ms", C.c_double)]


class Profile(C.Structure):
    _fields_ = [("syrk_ms", C.c_double), ("syrk_launches", C.c_longlong), ("schur_flops", C.c_double),
                ("panel_ms", C.c_double), ("panel_launches", C.c_longlong), ("panel_rows_ms", C.c_double),
                ("panel_rows_launches", C.c_longlong), ("small_ms", C.c_double), ("small_launches", C.c_longlong),
                ("assemble_ms", C.c_double), ("assemble_launches", C.c_longlong), ("fwd_ms", C.c_double),
                ("fwd_launches", C.c_longlong), ("bwd_ms", C.c_double), ("bwd_launches", C.c_longlong),
                ("cb_ms", C.c_double), ("cb_launches", C.c_longlong), ("cb_flops", C.c_double)]


# name -> (restype, argtypes); every symbol declared in include/symla_hip.h
SIGNATURES = {
    "symla_last_error": (C.c_char_p, []),
    "symla_analyze": (C.c_int, [C.c_int, i64p, i32p, C.c_int, i32p, C.c_int, C.c_double, C.POINTER(C.c_void_p)]),
    "symla_symbolic_info": (C.c_int, [C.c_void_p, C.POINTER(SymbolicInfo)]),
    "symla_symbolic_get": (C.c_int, [C.c_void_p, i32p, i32p, i32p, i32p, i64p, i32p]),
    "symla_symbolic_colcounts": (C.c_int, [C.c_void_p, i32p]),
    "symla_batch_analyze": (C.c_int, [C.c_int, i32p, C.POINTER(i64p), C.POINTER(i32p), C.c_int, C.c_int, C.c_double,
                                      C.POINTER(C.c_void_p)]),
    "symla_batch_plan_create": (C.c_int, [C.c_void_p, C.c_int, i32p, C.POINTER(i64p), C.POINTER(i32p),
                                          C.POINTER(C.c_void_p)]),
    "symla_batch_solve_refined": (C.c_int, [C.c_void_p, C.c_int, i32p, f64p, f64p, C.c_int64, C.c_int, f64p,
                                            C.c_int64, C.POINTER(RefineOpts), C.POINTER(RefineInfo)]),
    "symla_batch_solve_refined_device": (C.c_int, [C.c_void_p, C.c_int, i32p, C.c_void_p, C.c_void_p, C.c_int64,
                                                   C.c_int, C.c_void_p, C.c_int64, C.POINTER(RefineOpts),
                                                   C.POINTER(RefineInfo)]),
    "symla_read_matrix_market": (C.c_int, [C.c_char_p, i32p, i32p, i64p, i32p, i64p, i32p, f64p]),
    "symla_solve_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_int64]),
    "symla_symbolic_free": (None, [C.c_void_p]),
    "symla_plan_create": (C.c_int, [C.c_void_p, i64p, i32p, C.POINTER(C.c_void_p)]),
    "symla_plan_destroy": (None, [C.c_void_p]),
    "symla_plan_last_error": (C.c_char_p, [C.c_void_p]),
    "symla_factorize": (C.c_int, [C.c_void_p, f64p, C.POINTER(FactorOpts), C.POINTER(FactorInfo)]),
    "symla_factorize_device": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FactorOpts), C.POINTER(FactorInfo)]),
    "symla_singular_cols": (C.c_int, [C.c_void_p, i32p, C.c_int]),
    "symla_solve": (C.c_int, [C.c_void_p, f64p, C.c_int64, C.c_int, f64p, C.c_int64]),
    "symla_group_inertia": (C.c_int, [C.c_void_p, C.c_int, i32p, i32p, i32p]),
    "symla_solve_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_int64]),
    "symla_solve_refined": (C.c_int, [C.c_void_p, f64p, f64p, C.c_int64, C.c_int, f64p, C.c_int64,
                                      C.POINTER(RefineOpts), C.POINTER(RefineInfo)]),
    "symla_solve_refined_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p,
                                             C.c_int64, C.POINTER(RefineOpts), C.POINTER(RefineInfo)]),
    "symla_graph_capture": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_int64,
                                      C.POINTER(FactorOpts)]),
    "symla_graph_capture_refined": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p,
                                              C.c_int64, C.POINTER(FactorOpts), C.POINTER(RefineOpts)]),
    "symla_graph_replay_refined": (C.c_int, [C.c_void_p, C.POINTER(FactorInfo), C.POINTER(RefineInfo)]),
    "symla_graph_replay": (C.c_int, [C.c_void_p, C.POINTER(FactorInfo)]),
    "symla_graph_release": (C.c_int, [C.c_void_p]),
    "symla_plan_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "symla_plan_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "symla_plan_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(Tuning),
                                      C.POINTER(TuningTrial), C.c_int, i32p]),
    "symla_front_count": (C.c_int, [C.c_void_p]),
    "symla_front_view": (C.c_int, [C.c_void_p, C.c_int, i32p, i32p, i32p, f64p, f64p, f64p, i32p, i32p, i32p]),
    "symla_dense_factor": (C.c_int, [C.c_int, f64p, C.c_int, C.c_int, C.POINTER(DenseOpts), f64p, f64p, i32p,
                                     i32p, i32p, i32p, i32p, i32p, f64p, i32p]),
    "symla_plan_timing": (C.c_int, [C.c_void_p, f64p, f64p]),
    "symla_plan_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "symla_plan_get_profile": (C.c_int, [C.c_void_p, C.POINTER(Profile)]),
    "symla_plan_autotune_cached": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                             C.POINTER(Tuning), C.POINTER(TuningTrial), C.c_int,
                                             C.POINTER(C.c_int), C.POINTER(C.c_int)]),
}

_lib = None


def load() -> C.CDLL:
    """Load libsymla_hip.so (RTLD_GLOBAL not needed).  Raises if absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m ma57_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class SymlaError(Exception):
    pass


def raise_for(code: int, msg: str) -> None:
    """Map a C-ABI status to the exception the reference throws."""
    if code == SYMLA_OK:
        return
    if code == SYMLA_ERR_INVALID:
        raise ValueError(msg)
    if code == SYMLA_ERR_LOGIC:
        raise RuntimeError("logic_error: " + msg)
    if code == SYMLA_ERR_RUNTIME:
        raise RuntimeError(msg)
    raise SymlaError(msg)


def ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def as_i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def as_i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


def as_f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)
