"""ctypes binding of the CPU oracle (oracle/symla_oracle.cpp).

TEST INFRASTRUCTURE ONLY: the checker of the HIP product path and the CPU
baseline of the benchmark.  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "build" / "libsymla_oracle.so"

i32p = C.POINTER(C.c_int)
i64p = C.POINTER(C.c_longlong)
f64p = C.POINTER(C.c_double)


class DenseOpts(C.Structure):
    _fields_ = [("pivot_threshold", C.c_double), ("zero_tolerance", C.c_double),
                ("relative_pivot_floor", C.c_double), ("panel_size", C.c_int)]


class StaticOpts(C.Structure):
    _fields_ = [("delta", C.c_double), ("relative_pivot_floor", C.c_double), ("absolute_floor", C.c_double),
                ("panel_size", C.c_int)]


class DenseResult(C.Structure):
    _fields_ = [("n_factored", C.c_int), ("n_delayed", C.c_int), ("n_pivots", C.c_int), ("n_pos", C.c_int),
                ("n_neg", C.c_int), ("n_zero", C.c_int), ("max_growth", C.c_double),
                ("total_perturbation", C.c_double), ("num_perturbed", C.c_int)]


class MfOpts(C.Structure):
    _fields_ = [("parallel", C.c_int), ("num_threads", C.c_int), ("task_cutoff", C.c_longlong),
                ("static_pivoting", C.c_int), ("dense", DenseOpts), ("stat", StaticOpts), ("sn_limit", C.c_int)]


class MfInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("n_pos", C.c_int), ("n_neg", C.c_int), ("n_zero", C.c_int), ("singular", C.c_int),
                ("n_singular", C.c_int), ("total_perturbation", C.c_double), ("num_perturbed", C.c_int),
                ("nsn", C.c_int), ("total_pivots", C.c_longlong), ("n_two_by_two", C.c_int), ("max_front", C.c_int)]


class RefineResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("initial_relative_residual", C.c_double),
                ("final_relative_residual", C.c_double)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is not built; run `python -m ma57_amd.build`")
        lib = C.CDLL(str(LIB))
        lib.oracle_dense_factor.restype = C.c_int
        lib.oracle_dense_factor.argtypes = [C.c_int, f64p, C.c_int, C.c_int, C.POINTER(DenseOpts), f64p, f64p, i32p,
                                            i32p, i32p, i32p, C.POINTER(DenseResult)]
        lib.oracle_dense_factor_static.restype = C.c_int
        lib.oracle_dense_factor_static.argtypes = [C.c_int, f64p, C.c_int, C.c_int, C.POINTER(StaticOpts), i32p,
                                                   C.c_int, f64p, f64p, i32p, i32p, i32p, C.POINTER(DenseResult)]
        lib.oracle_mf_factorize.restype = C.c_void_p
        lib.oracle_mf_factorize.argtypes = [C.c_int, i32p, i32p, C.c_int, i32p, i32p, i64p, i32p, i64p, i32p, f64p,
                                            i32p, C.POINTER(MfOpts)]
        lib.oracle_mf_free.argtypes = [C.c_void_p]
        lib.oracle_mf_get_info.argtypes = [C.c_void_p, C.POINTER(MfInfo)]
        lib.oracle_mf_singular_cols.restype = C.c_int
        lib.oracle_mf_singular_cols.argtypes = [C.c_void_p, i32p, C.c_int]
        lib.oracle_mf_front_m.restype = C.c_int
        lib.oracle_mf_front_m.argtypes = [C.c_void_p, C.c_int]
        lib.oracle_mf_front_npiv.restype = C.c_int
        lib.oracle_mf_front_npiv.argtypes = [C.c_void_p, C.c_int]
        lib.oracle_mf_front_get.restype = C.c_int
        lib.oracle_mf_front_get.argtypes = [C.c_void_p, C.c_int, i32p, f64p, f64p, f64p, i32p, i32p]
        lib.oracle_mf_solve.restype = C.c_int
        lib.oracle_mf_solve.argtypes = [C.c_void_p, f64p, C.c_int, f64p, C.c_int, C.c_int]
        lib.oracle_refine.restype = C.c_int
        lib.oracle_refine.argtypes = [C.c_void_p, C.c_int, i64p, i32p, f64p, f64p, C.c_int, f64p, C.c_int, C.c_double,
                                      C.c_int, C.c_int, C.POINTER(RefineResult)]
        lib.oracle_static_regularized.restype = C.c_void_p
        lib.oracle_static_regularized.argtypes = [C.c_int, i32p, i32p, C.c_int, i32p, i32p, i64p, i32p, i64p, i32p,
                                                  f64p, i32p, C.c_double, C.c_int, C.c_double, C.c_double,
                                                  C.c_double, C.c_int, C.c_int, f64p, i32p, i32p, C.c_char_p,
                                                  C.c_int]
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def dense_opts(zero_tolerance=1e-300, relative_pivot_floor=1e-12, panel_size=64):
    return DenseOpts(0.01, zero_tolerance, relative_pivot_floor, panel_size)


def dense_factor(A, n_eligible=-1, zero_tolerance=1e-300, relative_pivot_floor=1e-12, panel_size=64):
    """DenseLDLT::factor (dense_kernel.hpp:393-837) on a copy of A (column-major)."""
    lib = load()
    F = np.array(A, dtype=np.float64, order="F", copy=True)
    n = F.shape[0]
    o = dense_opts(zero_tolerance, relative_pivot_floor, panel_size)
    dd = np.zeros(max(n, 1)); do = np.zeros(max(n, 1))
    perm = np.zeros(max(n, 1), np.int32); ps = np.zeros(max(n, 1), np.int32); pk = np.zeros(max(n, 1), np.int32)
    dl = np.zeros(max(n, 1), np.int32)
    r = DenseResult()
    rc = lib.oracle_dense_factor(n, _p(F, f64p), max(n, 1), n_eligible, C.byref(o), _p(dd, f64p), _p(do, f64p),
                                 _p(perm, i32p), _p(ps, i32p), _p(pk, i32p), _p(dl, i32p), C.byref(r))
    if rc != 0:
        raise ValueError("oracle dense factor failed")
    return dict(A=F, d_diag=dd[:n], d_off=do[:n], perm=perm[:n],
                pivots=[(int(ps[i]), int(pk[i])) for i in range(r.n_pivots)], n_factored=r.n_factored,
                delayed=[int(v) for v in dl[:r.n_delayed]], inertia=(r.n_pos, r.n_neg, r.n_zero),
                max_growth=r.max_growth)


def dense_factor_static(A, expected_sign=None, n_eligible=-1, delta=0.0, relative_pivot_floor=1e-8,
                        absolute_floor=1e-300, panel_size=64):
    """DenseLDLT::factorStatic (dense_kernel.hpp:919-1059)."""
    lib = load()
    F = np.array(A, dtype=np.float64, order="F", copy=True)
    n = F.shape[0]
    o = StaticOpts(delta, relative_pivot_floor, absolute_floor, panel_size)
    sg = np.ascontiguousarray(expected_sign if expected_sign is not None else np.zeros(n), np.int32)
    dd = np.zeros(max(n, 1)); do = np.zeros(max(n, 1))
    perm = np.zeros(max(n, 1), np.int32); ps = np.zeros(max(n, 1), np.int32); pk = np.zeros(max(n, 1), np.int32)
    r = DenseResult()
    rc = lib.oracle_dense_factor_static(n, _p(F, f64p), max(n, 1), n_eligible, C.byref(o), _p(sg, i32p), len(sg),
                                        _p(dd, f64p), _p(do, f64p), _p(perm, i32p), _p(ps, i32p), _p(pk, i32p),
                                        C.byref(r))
    if rc != 0:
        raise ValueError("oracle static factor failed")
    return dict(A=F, d_diag=dd[:n], d_off=do[:n], perm=perm[:n], n_factored=r.n_factored,
                inertia=(r.n_pos, r.n_neg, r.n_zero), total_perturbation=r.total_perturbation,
                num_perturbed=r.num_perturbed, max_growth=r.max_growth)


class Symbolic:
    """Arrays of a symbolic analysis (perm, etree parent, supernodes)."""

    def __init__(self, perm, etree_parent, sn_first, sn_ncols, sn_rowptr, sn_rows):
        self.perm = np.ascontiguousarray(perm, np.int32)
        self.parent = np.ascontiguousarray(etree_parent, np.int32)
        self.first = np.ascontiguousarray(sn_first, np.int32)
        self.ncols = np.ascontiguousarray(sn_ncols, np.int32)
        self.rowptr = np.ascontiguousarray(sn_rowptr, np.int64)
        self.rows = np.ascontiguousarray(sn_rows if len(sn_rows) else np.zeros(1), np.int32)

    @classmethod
    def from_product(cls, sf):
        """From ma57_amd.solver.SymbolicFactor (the analysis both paths share)."""
        rowptr = np.zeros(len(sf.supernodes) + 1, np.int64)
        for i, s in enumerate(sf.supernodes):
            rowptr[i + 1] = rowptr[i] + len(s.rowPattern)
        rows = np.concatenate([s.rowPattern for s in sf.supernodes]) if sf.supernodes else np.zeros(0)
        return cls(sf.perm, sf.etree_parent, [s.firstCol for s in sf.supernodes],
                   [s.ncols for s in sf.supernodes], rowptr, rows)

    def args(self):
        return (len(self.perm), _p(self.perm, i32p), _p(self.parent, i32p), len(self.first), _p(self.first, i32p),
                _p(self.ncols, i32p), _p(self.rowptr, i64p), _p(self.rows, i32p))


def _csc(A):
    A = A.tocsc()
    return (np.ascontiguousarray(A.indptr, np.int64), np.ascontiguousarray(A.indices, np.int32),
            np.ascontiguousarray(A.data, np.float64))


class Factor:
    """A NumericFactor held by the oracle."""

    def __init__(self, handle, A_arrays, n):
        self.h = handle
        self.keep = A_arrays
        self.n = n

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_mf_free(self.h)
            self.h = None

    def info(self):
        i = MfInfo()
        load().oracle_mf_get_info(self.h, C.byref(i))
        return i

    def inertia(self):
        i = self.info()
        return (i.n_pos, i.n_neg, i.n_zero)

    def singular_cols(self):
        i = self.info()
        out = np.zeros(max(i.n_singular, 1), np.int32)
        k = load().oracle_mf_singular_cols(self.h, _p(out, i32p), len(out))
        return [int(v) for v in out[:k]]

    def front(self, si):
        lib = load()
        m = lib.oracle_mf_front_m(self.h, si)
        k = lib.oracle_mf_front_npiv(self.h, si)
        rows = np.zeros(max(m, 1), np.int32)
        L = np.zeros((max(m, 1), max(k, 1)), order="F")
        dd = np.zeros(max(k, 1)); do = np.zeros(max(k, 1))
        ps = np.zeros(max(k, 1), np.int32); pk = np.zeros(max(k, 1), np.int32)
        nb = lib.oracle_mf_front_get(self.h, si, _p(rows, i32p), _p(L, f64p), _p(dd, f64p), _p(do, f64p),
                                     _p(ps, i32p), _p(pk, i32p))
        return dict(rows=rows[:m], npiv=k, L=L[:m, :k], d_diag=dd[:k], d_off=do[:k],
                    pivots=[(int(ps[i]), int(pk[i])) for i in range(nb)])

    def solve(self, B, parallel=False, threads=0):
        B = np.asfortranarray(np.asarray(B, dtype=np.float64).reshape(self.n, -1))
        X = np.zeros_like(B, order="F")
        rc = load().oracle_mf_solve(self.h, _p(B, f64p), B.shape[1], _p(X, f64p), int(parallel), threads)
        if rc != 0:
            raise RuntimeError("oracle solve failed")
        return X

    def refine(self, A, B, X, max_iterations=10, tolerance=1e-14, parallel=False, threads=0):
        cp, ri, vals = _csc(A)
        B = np.asfortranarray(np.asarray(B, dtype=np.float64).reshape(self.n, -1))
        X = np.asfortranarray(np.array(X, dtype=np.float64).reshape(self.n, -1))
        r = RefineResult()
        rc = load().oracle_refine(self.h, self.n, _p(cp, i64p), _p(ri, i32p), _p(vals, f64p), _p(B, f64p),
                                  B.shape[1], _p(X, f64p), max_iterations, tolerance, int(parallel), threads,
                                  C.byref(r))
        if rc != 0:
            raise RuntimeError("oracle refine failed")
        return X, r


def mf_factorize(sym: Symbolic, A, parallel=False, threads=0, static_pivoting=False, expected_sign_final=None,
                 delta=0.0, static_rel_floor=1e-8, static_abs_floor=1e-300, sn_limit=0):
    """MultifrontalFactorizer::factorize (multifrontal.hpp:299-805).  sn_limit > 0 factors only the first
    sn_limit supernodes in postorder (complete subtrees): a CPU timing sample, not a factorization."""
    lib = load()
    cp, ri, vals = _csc(A)
    o = MfOpts(int(parallel), threads, -1, int(static_pivoting), dense_opts(),
               StaticOpts(delta, static_rel_floor, static_abs_floor, 64), int(sn_limit))
    sg = None if expected_sign_final is None else np.ascontiguousarray(expected_sign_final, np.int32)
    h = lib.oracle_mf_factorize(*sym.args(), _p(cp, i64p), _p(ri, i32p), _p(vals, f64p),
                                _p(sg, i32p) if sg is not None else None, C.byref(o))
    if not h:
        raise RuntimeError("oracle factorize failed")
    return Factor(h, (cp, ri, vals), len(sym.perm))


def static_regularized(sym: Symbolic, A, signs_orig=None, initial_delta=0.0, max_retries=10, growth=10.0,
                       rel_floor=1e-8, abs_floor=1e-300, parallel=False, threads=0):
    """factorizeStaticRegularized (solver.hpp:328-417).  Returns (Factor or None, last_delta, retries, error)."""
    lib = load()
    cp, ri, vals = _csc(A)
    sg = None if signs_orig is None else np.ascontiguousarray(signs_orig, np.int32)
    ld = C.c_double()
    ru = C.c_int()
    threw = C.c_int()
    msg = C.create_string_buffer(1024)
    h = lib.oracle_static_regularized(*sym.args(), _p(cp, i64p), _p(ri, i32p), _p(vals, f64p),
                                      _p(sg, i32p) if sg is not None else None, initial_delta, max_retries, growth,
                                      rel_floor, abs_floor, int(parallel), threads, C.byref(ld), C.byref(ru),
                                      C.byref(threw), msg, 1024)
    if threw.value:
        return None, ld.value, ru.value, msg.value.decode()
    return Factor(h, (cp, ri, vals), len(sym.perm)), ld.value, ru.value, None
