"""Many independent symmetric indefinite systems factored and solved together.

No counterpart in the reference, which factors one system per SymLDLT
(include/symla/solver.hpp:29-418).  symla_batch_analyze analyzes every
system on its own (on host threads) and joins the analyses as that of the
block-diagonal matrix of the systems: a forest with one tree per system.  The
device plan then schedules the fronts of every system by tree level, so each
kernel launch of the numeric phase (assembly, small-front factorization,
panel steps, Schur updates, solve sweeps) covers the fronts of all systems at
that level.  Per-system results come back through symla_group_inertia
(inertia, singular columns), symla_batch_solve_refined (refine.hpp's loop run
for each system on its own: iterations and backward error per system) and row
slices of the solution.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np
import scipy.sparse as sp

from . import _lib
from .solver import (InvalidArgument, OrderingType, Precision, RefinementOptions, RefinementResult, SymLDLT,
                     SymbolicFactorOptions, _raise, csc_arrays, ptr, symbolic_factor)


class BatchedLDLT:
    """SymLDLT over a batch of systems A_0 .. A_{s-1} (each lower triangle or both stored)."""

    def __init__(self, ordering=OrderingType.AMD, precision=Precision.FP64):
        self._solver = SymLDLT()
        self._solver.setOrdering(ordering)
        self._solver.setPrecision(precision)
        self.sizes: List[int] = []
        self.starts = np.zeros(1, np.int32)
        self._A: Optional[sp.csc_matrix] = None
        self._patterns = None

    @staticmethod
    def _arrays(systems: Sequence):
        out = []
        for A in systems:
            n, cp, ri, v = csc_arrays(sp.csc_matrix(A))
            out.append((n, cp, ri, v))
        return out

    def _join(self, arrays) -> sp.csc_matrix:
        """The block-diagonal matrix, entries in the order the plan's pattern has them (system by system)."""
        n = int(self.starts[-1])
        cps, ris, vs = [np.zeros(1, np.int64)], [], []
        nnz = 0
        for (ns, cp, ri, v), off in zip(arrays, self.starts[:-1]):
            cps.append(cp[1:] + nnz)
            ris.append(ri + off)
            vs.append(v)
            nnz += int(cp[-1])
        cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)
        return sp.csc_matrix((cat(vs, np.float64), cat(ris, np.int32), cat(cps, np.int64)), shape=(n, n))

    def setPrecision(self, precision):
        self._solver.setPrecision(precision)

    def analyzePattern(self, systems: Sequence, options: SymbolicFactorOptions = SymbolicFactorOptions()):
        if len(systems) == 0:
            raise InvalidArgument("symla: empty batch")
        arrays = self._arrays(systems)
        self.sizes = [a[0] for a in arrays]
        self.starts = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int32)
        lib = _lib.load()
        ns = len(arrays)
        n_arr = np.array(self.sizes, np.int32)
        self._patterns = arrays  # keeps the pointer targets alive
        cps = (_lib.i64p * ns)(*[ptr(a[1], _lib.i64p) for a in arrays])
        ris = (_lib.i32p * ns)(*[ptr(a[2], _lib.i32p) for a in arrays])
        h = C.c_void_p()
        rc = lib.symla_batch_analyze(ns, ptr(n_arr, _lib.i32p), cps, ris, self._solver.ordering,
                                     options.max_relax_size, options.max_relax_fill_fraction, C.byref(h))
        _raise(rc, lib.symla_last_error().decode())
        plan = C.c_void_p()
        rc = lib.symla_batch_plan_create(h, ns, ptr(n_arr, _lib.i32p), cps, ris, C.byref(plan))
        if rc:
            lib.symla_symbolic_free(h)
        _raise(rc, lib.symla_last_error().decode())
        s = self._solver
        s._release()
        s._sym_handle = h
        s._symbolic = symbolic_factor(h)
        s._plan = plan
        self._A = self._join(arrays)
        s._n = int(self.starts[-1])
        s._nnz = int(self._A.indptr[-1])
        s._pattern_analyzed = True
        s._factorized = False

    def factorize(self, systems: Sequence):
        """Factor new values on the analyzed patterns (the first call after analyzePattern, or a refactorize)."""
        arrays = self._arrays(systems)
        if [a[0] for a in arrays] != self.sizes:
            raise InvalidArgument("symla: batch sizes differ from the analyzed batch")
        self._A = self._join(arrays)
        self._solver.factorize(self._A)

    def compute(self, systems: Sequence):
        self.analyzePattern(systems)
        self.factorize(systems)

    def inertia(self) -> List[tuple]:
        """(n_pos, n_neg, n_zero) of every system."""
        lib = _lib.load()
        ng = len(self.sizes)
        out = np.zeros(3 * ng, np.int32)
        sing = np.zeros(ng, np.int32)
        rc = lib.symla_group_inertia(self._solver._plan, ng, ptr(self.starts, _lib.i32p), ptr(out, _lib.i32p),
                                     ptr(sing, _lib.i32p))
        _raise(rc, lib.symla_plan_last_error(self._solver._plan).decode())
        self._singular = sing
        return [tuple(int(v) for v in out[3 * g:3 * g + 3]) for g in range(ng)]

    def singularCounts(self) -> List[int]:
        self.inertia()
        return [int(v) for v in self._singular]

    def _stack_rhs(self, Bs: Sequence) -> np.ndarray:
        if len(Bs) != len(self.sizes):
            raise InvalidArgument("symla: one right-hand side block per system")
        cols = {np.asarray(b).reshape(np.asarray(b).shape[0], -1).shape[1] for b in Bs}
        if len(cols) != 1:
            raise InvalidArgument("symla: every system needs the same number of right-hand sides")
        return np.vstack([np.asarray(b, dtype=np.float64).reshape(n, -1) for b, n in zip(Bs, self.sizes)])

    def _split(self, X: np.ndarray) -> List[np.ndarray]:
        return [X[self.starts[g]:self.starts[g + 1]] for g in range(len(self.sizes))]

    def solve(self, Bs: Sequence) -> List[np.ndarray]:
        return self._split(self._solver.solve(self._stack_rhs(Bs)))

    def solveWithRefinement(self, Bs: Sequence, options: Optional[RefinementOptions] = None):
        """FP64 iterative refinement of every system on its own (refine.hpp:97-144 per system); returns
        (per-system solutions, per-system RefinementResult)."""
        s = self._solver
        if not s._factorized:
            raise RuntimeError("logic_error: symla::SymLDLT::solve: factorize() must be called before solve()")
        opts = options or RefinementOptions()
        B = np.asfortranarray(self._stack_rhs(Bs))
        n, nrhs = B.shape
        X = np.zeros((n, nrhs), order="F")
        ng = len(self.sizes)
        infos = (_lib.RefineInfo * ng)()
        vals = np.ascontiguousarray(self._A.data, np.float64)
        lib = _lib.load()
        rc = lib.symla_batch_solve_refined(s._plan, ng, ptr(self.starts, _lib.i32p), ptr(vals, _lib.f64p),
                                           ptr(B, _lib.f64p), n, nrhs, ptr(X, _lib.f64p), n,
                                           C.byref(_lib.RefineOpts(opts.max_iterations, opts.tolerance)), infos)
        _raise(rc, lib.symla_plan_last_error(s._plan).decode())
        diags = [RefinementResult(i.iterations, i.initial_relative_residual, i.final_relative_residual,
                                  i.backward_error) for i in infos]
        return self._split(X), diags

    @property
    def matrix(self) -> sp.csc_matrix:
        return self._A

    @property
    def solver(self) -> SymLDLT:
        return self._solver
