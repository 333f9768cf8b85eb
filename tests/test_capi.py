"""The C-ABI library loads and exports every entry point include/symla_hip.h
declares, and its host-only entry points report errors the way the
reference's exceptions are mapped (no device calls here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from ma57_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "symla_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(symla_\w+)\s*\(", text)))


def test_header_declares_the_reference_roles():
    syms = declared_symbols()
    for s in ("symla_analyze", "symla_plan_create", "symla_factorize", "symla_factorize_device", "symla_solve",
              "symla_solve_device", "symla_solve_refined", "symla_singular_cols", "symla_front_view",
              "symla_dense_factor", "symla_plan_destroy"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert missing == []
    # the ctypes binding covers every declared symbol too
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_host_only_error_paths():
    lib = _lib.load()
    out = C.c_void_p()
    cp = np.array([0, 1, 2], dtype=np.int64)
    ri = np.array([0, 1], dtype=np.int32)
    rc = lib.symla_analyze(2, _lib.ptr(cp, _lib.i64p), _lib.ptr(ri, _lib.i32p), 9, None, 64, 0.25, C.byref(out))
    assert rc == _lib.SYMLA_ERR_INVALID
    assert b"OrderingType" in lib.symla_last_error()
    bad = np.array([0, 5], dtype=np.int32)
    rc = lib.symla_analyze(2, _lib.ptr(cp, _lib.i64p), _lib.ptr(bad, _lib.i32p), 2, None, 64, 0.25, C.byref(out))
    assert rc == _lib.SYMLA_ERR_INVALID
    rc = lib.symla_analyze(2, _lib.ptr(cp, _lib.i64p), _lib.ptr(ri, _lib.i32p), 2, None, 64, 0.25, C.byref(out))
    assert rc == _lib.SYMLA_OK and out.value
    info = _lib.SymbolicInfo()
    assert lib.symla_symbolic_info(out, C.byref(info)) == 0
    assert (info.n, info.nsn, info.nnz_l) == (2, 2, 2)
    lib.symla_symbolic_free(out)
    rc = lib.symla_dense_factor(-1, None, 1, -1, None, None, None, None, None, None, None, None, None, None, None)
    assert rc == _lib.SYMLA_ERR_INVALID


def test_null_plan_arguments_are_rejected():
    lib = _lib.load()
    t = _lib.Tuning(64, 64)
    assert lib.symla_plan_set_tuning(None, C.byref(t)) == _lib.SYMLA_ERR_INVALID
    assert lib.symla_plan_get_tuning(None, C.byref(t)) == _lib.SYMLA_ERR_INVALID
    assert lib.symla_graph_replay(None, None) == _lib.SYMLA_ERR_INVALID
    starts = np.array([0, 1], dtype=np.int32)
    inr = np.zeros(3, dtype=np.int32)
    assert lib.symla_group_inertia(None, 1, _lib.ptr(starts, _lib.i32p), _lib.ptr(inr, _lib.i32p),
                                   None) == _lib.SYMLA_ERR_INVALID


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr
