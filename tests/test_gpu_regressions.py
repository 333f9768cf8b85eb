"""Regression tests for fixed GPU bugs.

First solve after a fresh factorization: the backward sweep's last own-row
block of a front with no ancestors' rows (every root) reaches its tail with
no barrier between staging its row indices in LDS and reading them from
other waves; without the barrier after the staging, the first solve of a
fresh plan could read stale LDS (a repeat solve happened to find the same
values there).  Several fresh batched plans, each with many roots, are
solved once each and checked.
"""
import numpy as np
import pytest

from ma57_amd import problems as P
from ma57_amd.batch import BatchedLDLT
from ma57_amd.solver import OrderingType

pytestmark = pytest.mark.gpu


def test_first_solve_of_fresh_plans_with_many_roots():
    systems = [P.batched_system(0x5EED + 101 * i) for i in range(12)]
    rng = np.random.default_rng(17)
    Bs = [rng.standard_normal((A.shape[0], 3)) for A in systems]
    for trial in range(5):
        b = BatchedLDLT(OrderingType.AMD)
        b.compute(systems)
        Xs = b.solve(Bs)
        for i, (A, X, B) in enumerate(zip(systems, Xs, Bs)):
            be = P.backward_error(A, X, B)
            assert be <= 1e-12, (trial, i, be)
        del b


def test_refinement_reports_non_finite_right_hand_side():
    """A NaN in the right-hand side makes the solution and the residual NaN.
    The device max reductions map NaN to +inf (fmax alone would drop it), so
    the refinement reports an infinite backward error instead of converging
    on garbage, while the other columns stay unaffected in the solution."""
    from ma57_amd.solver import SymLDLT
    A = P.laplace2d(30, 30, 0.3)
    s = SymLDLT()
    s.setOrdering(OrderingType.AMD)
    s.analyzePattern(A)
    s.factorize(A)
    B = np.random.default_rng(3).standard_normal((A.shape[0], 2))
    good = s.solveWithRefinement(A, B)
    assert good.diagnostics.backward_error <= 1e-12
    B[7, 1] = np.nan
    bad = s.solveWithRefinement(A, B)
    assert not np.isfinite(bad.diagnostics.backward_error)
    assert not np.isfinite(bad.diagnostics.final_relative_residual)
    assert np.isnan(bad.x[:, 1]).any()
    # the healthy column's solution is the same as from the clean solve up to refinement steps
    assert P.backward_error(A, bad.x[:, :1], B[:, :1]) <= 1e-10
