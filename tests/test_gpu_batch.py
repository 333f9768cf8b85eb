"""Batched independent systems (the 2D-grid-plus-random-couplings batch) on
the HIP path: one block-diagonal plan whose level schedule covers every
system, checked system by system.

Bars: the inertia of every system equals the inertia of the same system
factored alone, and Sylvester's inertia from numpy.linalg.eigvalsh where the
system is small enough; backward error <= 1e-12 per system (FP64 solve, and
FP32 fronts after FP64 refinement).
"""
import numpy as np
import pytest

from ma57_amd import problems as P
from ma57_amd.batch import BatchedLDLT
from ma57_amd.solver import SymLDLT, OrderingType, Precision, RefinementOptions

pytestmark = pytest.mark.gpu

BE_TOL = 1e-12
SEED = 0x53594D4C41 + 4


def _systems(count):
    return [P.batched_system(SEED + 1000 * i) for i in range(count)]


def _eig_inertia(A):
    ev = np.linalg.eigvalsh(P.sym_dense(A))
    return (int((ev > 0).sum()), int((ev < 0).sum()), int((ev == 0).sum()))


def test_batch_fp64_per_system():
    systems = _systems(24)
    b = BatchedLDLT(OrderingType.AMD)
    b.compute(systems)
    inr = b.inertia()
    assert b.singularCounts() == [0] * len(systems)
    checked_eig = 0
    for A, got in zip(systems, inr):
        s = SymLDLT()
        s.analyzePattern(A)
        s.factorize(A)
        i = s.inertia()
        assert got == (i.n_pos, i.n_neg, i.n_zero)
        if A.shape[0] <= 1200 and checked_eig < 3:
            assert got == _eig_inertia(A)
            checked_eig += 1
    rng = np.random.default_rng(SEED)
    Bs = [rng.standard_normal((A.shape[0], 2)) for A in systems]
    Xs = b.solve(Bs)
    for A, X, B in zip(systems, Xs, Bs):
        assert P.backward_error(A, X, B) <= BE_TOL


def test_batch_fp32_refined_and_refactorize():
    systems = _systems(16)
    b64 = BatchedLDLT(OrderingType.AMD)
    b64.compute(systems)
    b = BatchedLDLT(OrderingType.AMD, Precision.FP32)
    b.compute(systems)
    assert b.inertia() == b64.inertia()
    rng = np.random.default_rng(SEED + 1)
    Bs = [rng.standard_normal(A.shape[0]) for A in systems]
    Xs, diags = b.solveWithRefinement(Bs, RefinementOptions(10, 1e-14))
    assert len(diags) == len(systems)
    for A, X, B, d in zip(systems, Xs, Bs, diags):
        assert d.iterations >= 1
        be = P.backward_error(A, X, B)
        assert be <= BE_TOL
        # the library's per-system backward error is the same formula (its residual is summed in another
        # order: near 1e-15 that moves the last digits)
        assert d.backward_error <= BE_TOL and abs(d.backward_error - be) <= 0.05 * be + 1e-16
    # new values on the same patterns: the shift moves, the inertia follows the individual factorizations
    shifted = [_shift_diagonal(A, -0.05) for A in systems]
    b64.factorize(shifted)
    for A, got in zip(shifted, b64.inertia()):
        s = SymLDLT()
        s.analyzePattern(A)
        s.factorize(A)
        i = s.inertia()
        assert got == (i.n_pos, i.n_neg, i.n_zero)


def _shift_diagonal(A, delta):
    """A + delta*I with the stored pattern unchanged (every diagonal entry is stored)."""
    B = A.tocsc(copy=True)
    for j in range(B.shape[1]):
        for t in range(B.indptr[j], B.indptr[j + 1]):
            if B.indices[t] == j:
                B.data[t] += delta
    return B
