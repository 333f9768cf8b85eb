"""Many independent symmetric indefinite systems factored and solved together.

No counterpart in the reference, which factors one system per SymLDLT
(include/symla/solver.hpp:29-418).  The systems are laid out as one
block-diagonal matrix: the symbolic analysis then yields a forest with one
tree per system, and the device plan schedules the fronts of every system by
tree level, so each kernel launch of the numeric phase (assembly, small-front
factorization, panel steps, Schur updates, solve sweeps) covers the fronts of
all systems at that level.  Per-system results come back through
symla_group_inertia (inertia, singular columns) and column slices of the
solution.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np
import scipy.sparse as sp

from . import _lib
from .solver import (InvalidArgument, OrderingType, Precision, RefinedSolve, RefinementOptions, SymLDLT,
                     SymbolicFactorOptions, _raise, csc_arrays, ptr)


class BatchedLDLT:
    """SymLDLT over a batch of systems A_0 .. A_{s-1} (each lower triangle or both stored)."""

    def __init__(self, ordering=OrderingType.AMD, precision=Precision.FP64):
        self._solver = SymLDLT()
        self._solver.setOrdering(ordering)
        self._solver.setPrecision(precision)
        self.sizes: List[int] = []
        self.starts = np.zeros(1, np.int32)
        self._A: Optional[sp.csc_matrix] = None

    @staticmethod
    def _stack(systems: Sequence) -> sp.csc_matrix:
        blocks = [sp.csc_matrix(A) for A in systems]
        for A in blocks:
            if A.shape[0] != A.shape[1]:
                raise InvalidArgument("symla: every system of a batch must be square")
        return sp.block_diag(blocks, format="csc")

    def setPrecision(self, precision):
        self._solver.setPrecision(precision)

    def analyzePattern(self, systems: Sequence, options: SymbolicFactorOptions = SymbolicFactorOptions()):
        if len(systems) == 0:
            raise InvalidArgument("symla: empty batch")
        self.sizes = [int(A.shape[0]) for A in systems]
        self.starts = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int32)
        self._A = self._stack(systems)
        self._solver.analyzePattern(self._A, options)

    def factorize(self, systems: Sequence):
        """Factor new values on the analyzed patterns (the first call after analyzePattern, or a refactorize)."""
        if [int(A.shape[0]) for A in systems] != self.sizes:
            raise InvalidArgument("symla: batch sizes differ from the analyzed batch")
        self._A = self._stack(systems)
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
        """FP64 iterative refinement of the whole batch; returns (per-system solutions, diagnostics)."""
        r: RefinedSolve = self._solver.solveWithRefinement(self._A, self._stack_rhs(Bs), options)
        return self._split(r.x), r.diagnostics

    @property
    def matrix(self) -> sp.csc_matrix:
        return self._A

    @property
    def solver(self) -> SymLDLT:
        return self._solver
