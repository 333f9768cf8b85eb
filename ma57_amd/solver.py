"""Python mirror of ``symla::SymLDLT<double>`` over the C-ABI.

Same method names, argument meaning and error behaviour as the reference
facade (include/symla/solver.hpp:29-418); the C++ facade with the same
surface is include/symla/solver.hpp in this repository.  Matrices are
``scipy.sparse`` CSC (any stored triangle; the lower triangle is
authoritative, multifrontal.hpp:190-198) and dense ``numpy`` arrays.

Exceptions: ``InvalidArgument`` (std::invalid_argument), ``LogicError``
(std::logic_error), ``RuntimeFailure`` (std::runtime_error).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import as_f64, as_i32, as_i64, f64p, i32p, i64p, ptr


class InvalidArgument(ValueError):
    pass


class LogicError(RuntimeError):
    pass


class RuntimeFailure(RuntimeError):
    pass


class DeviceFailure(RuntimeError):
    pass


def _raise(code: int, msg: str) -> None:
    if code == _lib.SYMLA_OK:
        return
    if code == _lib.SYMLA_ERR_INVALID:
        raise InvalidArgument(msg)
    if code == _lib.SYMLA_ERR_LOGIC:
        raise LogicError(msg)
    if code == _lib.SYMLA_ERR_RUNTIME:
        raise RuntimeFailure(msg)
    raise DeviceFailure(msg)


class Mode:
    ThresholdPivot = _lib.MODE_THRESHOLD
    StaticRegularized = _lib.MODE_STATIC


class OrderingType:
    AMD = _lib.ORDER_AMD
    Metis = _lib.ORDER_METIS
    Natural = _lib.ORDER_NATURAL


class Precision:
    FP64 = _lib.PREC_FP64
    FP32 = _lib.PREC_FP32
    BF16X3 = _lib.PREC_BF16X3  # FP32 fronts; deferred contribution-block updates as BF16x3 split products


@dataclass
class Inertia:
    n_pos: int = 0
    n_neg: int = 0
    n_zero: int = 0


@dataclass
class SymbolicFactorOptions:
    max_relax_size: int = 64
    max_relax_fill_fraction: float = 0.25


@dataclass
class Supernode:
    firstCol: int
    ncols: int
    rowPattern: np.ndarray


@dataclass
class SymbolicFactor:
    perm: np.ndarray
    etree_parent: np.ndarray
    supernodes: List[Supernode]
    nnz_l: int


@dataclass
class SupernodeFactor:
    rowIndices: np.ndarray
    nPivots: int
    L: np.ndarray
    D_diag: np.ndarray
    D_off: np.ndarray
    pivotBlocks: List[tuple]  # (start, kind) with kind 1 or 2


@dataclass
class RefinementOptions:
    max_iterations: int = 10
    tolerance: float = 1e-14


@dataclass
class RefinementResult:
    iterations: int = 0
    initial_relative_residual: float = 0.0
    final_relative_residual: float = 0.0
    backward_error: float = 0.0


@dataclass
class RefinedSolve:
    x: np.ndarray
    diagnostics: RefinementResult = field(default_factory=RefinementResult)


def csc_arrays(A):
    """(n, colptr int64, rowidx int32, values float64) of a scipy sparse matrix, stored order kept."""
    import scipy.sparse as sp

    if not sp.issparse(A):
        raise InvalidArgument("expected a scipy.sparse matrix")
    A = A.tocsc(copy=False)
    if A.shape[0] != A.shape[1]:
        raise InvalidArgument("symla::computeOrdering: matrix must be square")
    return A.shape[0], as_i64(A.indptr), as_i32(A.indices), as_f64(A.data)


def analyze(A, ordering: int = OrderingType.AMD, options: Optional[SymbolicFactorOptions] = None,
            user_perm=None):
    """Run the CPU symbolic analysis; returns (handle, SymbolicFactor)."""
    lib = _lib.load()
    opts = options or SymbolicFactorOptions()
    n, cp, ri, _ = csc_arrays(A) if not isinstance(A, tuple) else A
    h = C.c_void_p()
    up = None
    if user_perm is not None:
        up = as_i32(user_perm)
        ordering = _lib.ORDER_USER
    rc = lib.symla_analyze(n, ptr(cp, i64p), ptr(ri, i32p), ordering, ptr(up, i32p) if up is not None else None,
                           opts.max_relax_size, opts.max_relax_fill_fraction, C.byref(h))
    _raise(rc, lib.symla_last_error().decode())
    return h, symbolic_factor(h)


def symbolic_factor(h) -> "SymbolicFactor":
    """The SymbolicFactor fields of an analysis handle (symla_symbolic_get)."""
    lib = _lib.load()
    info = _lib.SymbolicInfo()
    _raise(lib.symla_symbolic_info(h, C.byref(info)), lib.symla_last_error().decode())
    n = info.n
    perm = np.empty(n, np.int32)
    par = np.empty(n, np.int32)
    first = np.empty(info.nsn, np.int32)
    nc = np.empty(info.nsn, np.int32)
    rp = np.empty(info.nsn + 1, np.int64)
    rows = np.empty(max(info.pattern_total, 1), np.int32)
    _raise(lib.symla_symbolic_get(h, ptr(perm, i32p), ptr(par, i32p), ptr(first, i32p), ptr(nc, i32p),
                                  ptr(rp, i64p), ptr(rows, i32p)), lib.symla_last_error().decode())
    sns = [Supernode(int(first[i]), int(nc[i]), rows[rp[i]:rp[i + 1]].copy()) for i in range(info.nsn)]
    return SymbolicFactor(perm, par, sns, int(info.nnz_l))


class SymLDLT:
    """Mirror of symla::SymLDLT<double> (solver.hpp:29-418) on the MI355X path."""

    def __init__(self):
        self._lib = _lib.load()
        self.mode = Mode.ThresholdPivot
        self.pivot_threshold = 0.01
        self.ordering = OrderingType.AMD
        self._sym_handle = None
        self._plan = None
        self._symbolic: Optional[SymbolicFactor] = None
        self._n = 0
        self._nnz = 0
        self._pattern_analyzed = False
        self._factorized = False
        self._inertia = Inertia()
        self._singular = False
        self._use_kkt = False
        self._kkt = (0, 0)
        self._use_signs = False
        self._signs = None
        self._init_delta = 0.0
        self._max_retries = 10
        self._growth = 10.0
        self._static_rel = 1e-8
        self._static_abs = 1e-300
        self._last_delta = 0.0
        self._retries = 0
        self._total_pert = 0.0
        self._num_pert = 0
        self.precision = Precision.FP64
        self.last_info = None

    # --- setters (solver.hpp:36-126) ---
    def setPrecision(self, precision):
        """Front storage precision (no counterpart in the reference): Precision.FP64, or Precision.FP32
        (FP32 fronts and Schur updates, FP64 pivots/D/inertia; use solveWithRefinement for FP64 accuracy)."""
        if precision not in (Precision.FP64, Precision.FP32, Precision.BF16X3):
            raise InvalidArgument("symla: unknown precision")
        self.precision = precision

    def setMode(self, mode):
        self.mode = mode

    def setPivotThreshold(self, u):
        self.pivot_threshold = float(u)

    def setOrdering(self, ordering):
        self.ordering = ordering

    def setKKTBlockSizes(self, n1, n2):
        if n1 < 0 or n2 < 0:
            raise InvalidArgument("symla::SymLDLT::setKKTBlockSizes: sizes must be nonnegative")
        self._kkt = (int(n1), int(n2))
        self._use_kkt = True
        self._use_signs = False

    def setExpectedSignPattern(self, signs):
        self._signs = as_i32(signs)
        self._use_signs = True
        self._use_kkt = False

    def setInitialRegularizationDelta(self, d):
        self._init_delta = float(d)

    def setMaxRegularizationRetries(self, r):
        self._max_retries = int(r)

    def setRegularizationGrowth(self, g):
        self._growth = float(g)

    def setStaticPivotFloor(self, rel, absolute=1e-300):
        self._static_rel = float(rel)
        self._static_abs = float(absolute)

    def setParallel(self, parallel):  # the device path has no thread setting
        pass

    def setNumThreads(self, n):
        pass

    # --- diagnostics ---
    def lastRegularizationDelta(self):
        return self._last_delta

    def regularizationRetriesUsed(self):
        return self._retries

    def totalPerturbation(self):
        return self._total_pert

    def numPerturbedPivots(self):
        return self._num_pert

    def __del__(self):
        self._release()

    def _release(self):
        lib = getattr(self, "_lib", None)
        if lib is None:
            return
        if self._plan is not None:
            lib.symla_plan_destroy(self._plan)
            self._plan = None
        if self._sym_handle is not None:
            lib.symla_symbolic_free(self._sym_handle)
            self._sym_handle = None

    # --- pipeline ---
    def analyzePattern(self, A, options: Optional[SymbolicFactorOptions] = None):
        self._release()
        n, cp, ri, _ = csc_arrays(A)
        h, sf = analyze((n, cp, ri, None), self.ordering, options)
        self._sym_handle = h
        self._symbolic = sf
        plan = C.c_void_p()
        rc = self._lib.symla_plan_create(h, ptr(cp, i64p), ptr(ri, i32p), C.byref(plan))
        _raise(rc, self._lib.symla_last_error().decode())
        self._plan = plan
        self._n = n
        self._nnz = int(cp[-1])
        self._pattern_analyzed = True
        self._factorized = False

    def _expected_signs(self):
        n = self._n
        if self._use_signs:
            if len(self._signs) != n:
                raise InvalidArgument("symla::SymLDLT::factorize: setExpectedSignPattern's vector size does not "
                                      "match the matrix size")
            return self._signs
        if self._use_kkt:
            n1, n2 = self._kkt
            if n1 + n2 != n:
                raise InvalidArgument("symla::SymLDLT::factorize: setKKTBlockSizes(n1, n2) sizes do not sum to the "
                                      "matrix size")
            s = np.ones(n, np.int32)
            s[:n1] = -1
            return s
        return None

    def _opts(self):
        o = _lib.FactorOpts()
        o.mode = self.mode
        o.pivot_threshold = self.pivot_threshold
        o.zero_tolerance = 1e-300
        o.relative_pivot_floor = 1e-12
        o.expected_sign = None
        o.initial_delta = self._init_delta
        o.max_retries = self._max_retries
        o.growth = self._growth
        o.static_rel_floor = self._static_rel
        o.static_abs_floor = self._static_abs
        o.precision = self.precision
        return o

    def factorize(self, A):
        if not self._pattern_analyzed:
            raise LogicError("symla::SymLDLT::factorize: analyzePattern must be called first")
        n, cp, ri, vals = csc_arrays(A)
        if n != self._n or int(cp[-1]) != self._nnz:
            raise InvalidArgument("symla: matrix pattern differs from the analyzed pattern")
        o = self._opts()
        signs = None
        if self.mode == Mode.StaticRegularized:
            signs = self._expected_signs()
            if signs is not None:
                o.expected_sign = ptr(signs, i32p)
        info = _lib.FactorInfo()
        self._factorized = False
        rc = self._lib.symla_factorize(self._plan, ptr(vals, f64p), C.byref(o), C.byref(info))
        _raise(rc, self._lib.symla_plan_last_error(self._plan).decode())
        self.last_info = info
        self._inertia = Inertia(info.n_pos, info.n_neg, info.n_zero)
        self._singular = bool(info.singular)
        self._factorized = True
        if self.mode == Mode.StaticRegularized:
            self._last_delta = info.last_delta
            self._retries = info.retries_used
        else:
            self._last_delta = 0.0
            self._retries = 0
        self._total_pert = info.total_perturbation
        self._num_pert = info.num_perturbed

    def compute(self, A):
        self.analyzePattern(A)
        self.factorize(A)

    def refactorize(self, A):
        self.factorize(A)

    def solve(self, B):
        if not self._factorized:
            raise LogicError("symla::SymLDLT::solve: factorize() must be called before solve()")
        B = np.asarray(B, dtype=np.float64)
        vec = B.ndim == 1
        B2 = B.reshape(-1, 1) if vec else B
        if B2.shape[0] != self._n:
            raise InvalidArgument("symla::SymLDLT::solve: B.rows() must match the factorized matrix size")
        Bf = np.asfortranarray(B2)
        X = np.empty_like(Bf, order="F")
        nrhs = Bf.shape[1]
        rc = self._lib.symla_solve(self._plan, ptr(Bf, f64p), self._n, nrhs, ptr(X, f64p), self._n)
        _raise(rc, self._lib.symla_plan_last_error(self._plan).decode())
        return X.reshape(-1) if vec else X

    def solveWithRefinement(self, A, B, options: Optional[RefinementOptions] = None) -> RefinedSolve:
        if not self._factorized:
            raise LogicError("symla::SymLDLT::solve: factorize() must be called before solve()")
        opts = options or RefinementOptions()
        n, cp, ri, vals = csc_arrays(A)
        B = np.asarray(B, dtype=np.float64)
        vec = B.ndim == 1
        B2 = B.reshape(-1, 1) if vec else B
        if B2.shape[0] != self._n:
            raise InvalidArgument("symla::SymLDLT::solve: B.rows() must match the factorized matrix size")
        Bf = np.asfortranarray(B2)
        X = np.empty_like(Bf, order="F")
        ro = _lib.RefineOpts(opts.max_iterations, opts.tolerance)
        ri_ = _lib.RefineInfo()
        rc = self._lib.symla_solve_refined(self._plan, ptr(vals, f64p), ptr(Bf, f64p), self._n, Bf.shape[1],
                                           ptr(X, f64p), self._n, C.byref(ro), C.byref(ri_))
        _raise(rc, self._lib.symla_plan_last_error(self._plan).decode())
        diag = RefinementResult(ri_.iterations, ri_.initial_relative_residual, ri_.final_relative_residual,
                                ri_.backward_error)
        return RefinedSolve(X.reshape(-1) if vec else X, diag)

    # --- scheduling choices (no reference counterpart) ---
    def tuning(self):
        """(small_front_max, assemble_tile) of the plan's schedule."""
        t = self._get_tuning()
        return (t.small_front_max, t.assemble_tile)

    def _get_tuning(self):
        t = _lib.Tuning()
        rc = self._lib.symla_plan_get_tuning(self._plan, C.byref(t))
        _raise(rc, "symla_plan_get_tuning failed")
        return t

    def setTuning(self, small_front_max: int, assemble_tile: int, speculative_panel=None, speculative_min_rows=None,
                  deferred_cb_min_rows=None):
        """Scheduling choices; the optional ones left as None keep their current values."""
        cur = self._get_tuning()
        spec = cur.speculative_panel if speculative_panel is None else (1 if speculative_panel else 0)
        rows = cur.speculative_min_rows if speculative_min_rows is None else int(speculative_min_rows)
        cb = cur.deferred_cb_min_rows if deferred_cb_min_rows is None else int(deferred_cb_min_rows)
        t = _lib.Tuning(small_front_max, assemble_tile, spec, rows, cb)
        rc = self._lib.symla_plan_set_tuning(self._plan, C.byref(t))
        _raise(rc, self._lib.symla_plan_last_error(self._plan).decode())

    def speculativePanel(self) -> bool:
        """Whether panels of large fronts start with the speculative blocked step."""
        return bool(self._get_tuning().speculative_panel)

    def deferredCbMinRows(self) -> int:
        """Contribution blocks of at least this many rows (large fronts, >= 256 eligible columns) are updated
        once after the front's last panel; -1: every block panel by panel."""
        return int(self._get_tuning().deferred_cb_min_rows)

    def setSpeculativePanel(self, on: bool, min_rows=None):
        """Switch the speculative panel step; min_rows: fronts with more eligible rows than this take it."""
        small, tile = self.tuning()
        self.setTuning(small, tile, on, min_rows)

    def autotune(self, d_vals: int, nrhs: int = 1, reps: int = 2, cache: Optional[str] = None):
        """Time factorize + solve by coordinate descent over the scheduling choices on device values; keep the
        fastest.  Returns (chosen, [(tuning, ms), ...], from_cache) with each tuning a tuple (small_front_max,
        assemble_tile, speculative_panel, speculative_min_rows, deferred_cb_min_rows).  With `cache` (a file
        path), a stored result for the same GPU, precision, nrhs and front-order histogram is applied without
        timing."""
        if not self._pattern_analyzed:
            raise LogicError("symla::SymLDLT::factorize: analyzePattern must be called first")
        chosen = _lib.Tuning()
        trials = (_lib.TuningTrial * 128)()
        nt = C.c_int(0)
        fc = C.c_int(0)
        if cache is None:
            rc = self._lib.symla_plan_autotune(self._plan, C.c_void_p(d_vals), nrhs, self.precision, reps,
                                               C.byref(chosen), trials, 128, C.byref(nt))
        else:
            rc = self._lib.symla_plan_autotune_cached(self._plan, C.c_void_p(d_vals), nrhs, self.precision, reps,
                                                      str(cache).encode(), C.byref(chosen), trials, 128,
                                                      C.byref(nt), C.byref(fc))
        _raise(rc, self._lib.symla_plan_last_error(self._plan).decode())
        self._factorized = False  # the trials factored the tuning values: factorize() before solving
        tup = lambda t: (t.small_front_max, t.assemble_tile, t.speculative_panel, t.speculative_min_rows,
                         t.deferred_cb_min_rows)
        return tup(chosen), [(tup(trials[i].tuning), trials[i].ms) for i in range(min(nt.value, 128))], bool(fc.value)

    # --- recorded refactorize + solve (no reference counterpart) ---
    def captureGraph(self, d_vals: int, d_B: int, nrhs: int, d_X: int, refine: Optional[RefinementOptions] = None):
        """Record factorize(values at device address d_vals) + solve(d_B -> d_X), n x nrhs column-major; with
        `refine`, the solve is refined in FP64 (solveWithRefinement semantics) by recorded refinement steps."""
        if not self._pattern_analyzed:
            raise LogicError("symla::SymLDLT::factorize: analyzePattern must be called first")
        o = self._opts()
        if refine is None:
            rc = self._lib.symla_graph_capture(self._plan, C.c_void_p(d_vals), C.c_void_p(d_B), self._n, nrhs,
                                               C.c_void_p(d_X), self._n, C.byref(o))
        else:
            ro = _lib.RefineOpts(refine.max_iterations, refine.tolerance)
            rc = self._lib.symla_graph_capture_refined(self._plan, C.c_void_p(d_vals), C.c_void_p(d_B), self._n,
                                                       nrhs, C.c_void_p(d_X), self._n, C.byref(o), C.byref(ro))
        self._graph_refined = refine is not None
        _raise(rc, self._lib.symla_plan_last_error(self._plan).decode())

    def replayGraph(self):
        """Re-run the recorded factorize + solve on the current buffer contents; returns the factor info (and,
        for a graph recorded with refinement, `last_refinement` holds the RefinementResult)."""
        info = _lib.FactorInfo()
        self._factorized = False
        if getattr(self, "_graph_refined", False):
            ri = _lib.RefineInfo()
            rc = self._lib.symla_graph_replay_refined(self._plan, C.byref(info), C.byref(ri))
            self.last_refinement = RefinementResult(ri.iterations, ri.initial_relative_residual,
                                                    ri.final_relative_residual, ri.backward_error)
        else:
            rc = self._lib.symla_graph_replay(self._plan, C.byref(info))
        _raise(rc, self._lib.symla_plan_last_error(self._plan).decode())
        self.last_info = info
        self._inertia = Inertia(info.n_pos, info.n_neg, info.n_zero)
        self._singular = bool(info.singular)
        self._factorized = True
        return info

    def releaseGraph(self):
        rc = self._lib.symla_graph_release(self._plan)
        _raise(rc, self._lib.symla_plan_last_error(self._plan).decode())

    def inertia(self) -> Inertia:
        return self._inertia

    def patternAnalyzed(self):
        return self._pattern_analyzed

    def factorized(self):
        return self._factorized

    def isSingular(self):
        return self._singular

    def singularColumns(self):
        cnt = self._lib.symla_singular_cols(self._plan, None, 0)
        out = np.empty(max(cnt, 1), np.int32)
        self._lib.symla_singular_cols(self._plan, ptr(out, i32p), cnt)
        return [int(v) for v in out[:cnt]]

    def symbolicFactor(self) -> SymbolicFactor:
        return self._symbolic

    def numericFactor(self) -> List[SupernodeFactor]:
        """Host copy of NumericFactor::fronts (multifrontal.hpp:148-186)."""
        lib = self._lib
        nsn = lib.symla_front_count(self._plan)
        fronts = []
        for si in range(nsn):
            m = C.c_int()
            k = C.c_int()
            _raise(lib.symla_front_view(self._plan, si, C.byref(m), C.byref(k), None, None, None, None, None, None,
                                        None), lib.symla_plan_last_error(self._plan).decode())
            mm, kk = m.value, k.value
            rows = np.empty(max(mm, 1), np.int32)
            L = np.zeros((mm, max(kk, 1)), order="F")
            dd = np.empty(max(kk, 1))
            do = np.empty(max(kk, 1))
            ps = np.empty(max(kk, 1), np.int32)
            pk = np.empty(max(kk, 1), np.int32)
            nb = C.c_int()
            _raise(lib.symla_front_view(self._plan, si, C.byref(m), C.byref(k), ptr(rows, i32p), ptr(L, f64p),
                                        ptr(dd, f64p), ptr(do, f64p), ptr(ps, i32p), ptr(pk, i32p), C.byref(nb)),
                   lib.symla_plan_last_error(self._plan).decode())
            fronts.append(SupernodeFactor(rows[:mm].copy(), kk, L[:, :kk].copy(), dd[:kk].copy(), do[:kk].copy(),
                                          [(int(ps[i]), int(pk[i])) for i in range(nb.value)]))
        return fronts


def dense_factor(A: np.ndarray, n_eligible: int = -1, zero_tolerance: float = 1e-300,
                 relative_pivot_floor: float = 1e-12, static_mode: bool = False, static_delta: float = 0.0,
                 static_rel_floor: float = 1e-8, static_abs_floor: float = 1e-300, expected_sign=None,
                 exact_panel: bool = False, deferred_cb: bool = False):
    """DenseLDLT::factor / factorStatic (dense_kernel.hpp:393-837, :919-1059) on the device front kernels.

    Returns a dict with the factored matrix (lower triangle meaningful), D, perm, pivot blocks, n_factored,
    inertia, and the static-mode perturbation statistics."""
    lib = _lib.load()
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    n = A.shape[0]
    o = _lib.DenseOpts(zero_tolerance, relative_pivot_floor, 1 if static_mode else 0, static_delta,
                       static_rel_floor, static_abs_floor, None, 1 if exact_panel else 0, 1 if deferred_cb else 0)
    sg = None
    if expected_sign is not None:
        sg = as_i32(expected_sign)
        o.expected_sign = ptr(sg, i32p)
    dd = np.zeros(max(n, 1))
    do = np.zeros(max(n, 1))
    perm = np.zeros(max(n, 1), np.int32)
    ps = np.zeros(max(n, 1), np.int32)
    pk = np.zeros(max(n, 1), np.int32)
    nb = C.c_int()
    nf = C.c_int()
    inert = np.zeros(3, np.int32)
    tp = C.c_double()
    npert = C.c_int()
    rc = lib.symla_dense_factor(n, ptr(A, f64p), max(n, 1), n_eligible, C.byref(o), ptr(dd, f64p), ptr(do, f64p),
                                ptr(perm, i32p), ptr(ps, i32p), ptr(pk, i32p), C.byref(nb), C.byref(nf),
                                ptr(inert, i32p), C.byref(tp), C.byref(npert))
    _raise(rc, lib.symla_last_error().decode())
    k = nf.value
    return dict(A=A, d_diag=dd[:n], d_off=do[:n], perm=perm[:n], pivots=[(int(ps[i]), int(pk[i]))
                                                                        for i in range(nb.value)],
                n_factored=k, delayed=list(range(k, n if n_eligible < 0 else n_eligible)),
                inertia=(int(inert[0]), int(inert[1]), int(inert[2])), total_perturbation=tp.value,
                num_perturbed=npert.value)
