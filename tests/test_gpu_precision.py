"""FP32 front storage with FP64 iterative refinement, and the larger
configurations (SQD KKT in static-regularized mode, 27-point stencil) on the
HIP path, through the C-ABI.

Bars: exact inertia (closed form, Sylvester's law of inertia, or the FP64
path on the same matrix); backward error <= 1e-12 after FP64 refinement,
with the refinement iteration count reported; the unrefined FP32 solve is
only required to be a usable first iterate.
"""
import numpy as np
import pytest

from ma57_amd import problems as P
from ma57_amd.solver import SymLDLT, Mode, OrderingType, Precision, RefinementOptions

pytestmark = pytest.mark.gpu

BE_TOL = 1e-12


def _factor(A, precision, ordering=OrderingType.Metis, kkt=None):
    s = SymLDLT()
    s.setOrdering(ordering)
    s.setPrecision(precision)
    if kkt is not None:
        s.setMode(Mode.StaticRegularized)
        s.setKKTBlockSizes(*kkt)
    s.analyzePattern(A)
    s.factorize(A)
    return s


def _inertia(s):
    i = s.inertia()
    return (i.n_pos, i.n_neg, i.n_zero)


def test_c2_fp32_refined():
    A = P.laplace3d(40, 40, 40, 0.5)
    s = _factor(A, Precision.FP32)
    assert _inertia(s) == (63671, 329, 0)
    rng = np.random.default_rng(0x53594D4C41 + 2)
    for nrhs in (1, 16):
        B = rng.standard_normal((A.shape[0], nrhs))
        raw = P.backward_error(A, s.solve(B), B)
        assert raw < 1e-4
        r = s.solveWithRefinement(A, B, RefinementOptions(10, 1e-14))
        assert P.backward_error(A, r.x, B) <= BE_TOL
        assert r.diagnostics.backward_error <= BE_TOL
        assert r.diagnostics.iterations >= 1


@pytest.mark.parametrize("n,seed", [(500, 3), (2000, 4)])
def test_random_indefinite_fp32_matches_fp64(n, seed):
    A = P.random_sparse_indefinite(n, 5, seed)
    s64 = _factor(A, Precision.FP64, OrderingType.AMD)
    s32 = _factor(A, Precision.FP32, OrderingType.AMD)
    assert _inertia(s32) == _inertia(s64)
    b = np.random.default_rng(seed + 7).standard_normal((n, 3))
    x64 = s64.solve(b)
    r = s32.solveWithRefinement(A, b, RefinementOptions(10, 1e-14))
    assert P.backward_error(A, r.x, b) <= BE_TOL
    assert np.linalg.norm(r.x - x64) <= 1e-8 * np.linalg.norm(x64)


def test_precision_switch_on_one_plan():
    A = P.laplace3d(16, 16, 16, 0.5)
    s = SymLDLT()
    s.setOrdering(OrderingType.Metis)
    s.analyzePattern(A)
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    s.factorize(A)
    x64 = s.solve(b)
    s.setPrecision(Precision.FP32)
    s.refactorize(A)
    x32 = s.solveWithRefinement(A, b).x
    s.setPrecision(Precision.FP64)
    s.refactorize(A)
    assert np.array_equal(s.solve(b), x64)
    assert np.linalg.norm(x32 - x64) <= 1e-10 * np.linalg.norm(x64)


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32])
def test_c3_sqd_kkt_static(precision):
    """SQD KKT, H = 7-point 30^3 (n1 = 27,000), J 9,000 x 27,000 with 5 entries per row:
    inertia (n2, n1, 0) by Sylvester; FP64 refinement to <= 1e-12, nrhs = 4."""
    K, n1, n2 = P.sqd_kkt(30, 9000)
    s = _factor(K, precision, OrderingType.AMD, kkt=(n1, n2))
    assert _inertia(s) == (n2, n1, 0)
    assert s.regularizationRetriesUsed() <= 1
    # FP32 fronts either hold the expected inertia, or the factorization reports the FP64 fallback
    assert s.last_info.precision_used in ((Precision.FP64,) if precision == Precision.FP64
                                          else (Precision.FP32, Precision.FP64))
    B = np.random.default_rng(0x53594D4C41 + 3).standard_normal((n1 + n2, 4))
    r = s.solveWithRefinement(K, B, RefinementOptions(10, 1e-14))
    assert P.backward_error(K, r.x, B) <= BE_TOL


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32])
def test_stencil27_fp64_and_fp32_agree(precision):
    """27-point stencil with random edge weights on 30^3, sigma = 0.2, 64 right-hand sides: the inertia of
    the FP32 factorization equals the FP64 one, both refine to <= 1e-12."""
    A = P.stencil27(30, 0.2)
    s64 = _factor(A, Precision.FP64)
    s = s64 if precision == Precision.FP64 else _factor(A, precision)
    assert _inertia(s) == _inertia(s64)
    assert sum(_inertia(s)) == A.shape[0]
    B = np.random.default_rng(0x53594D4C41 + 5).standard_normal((A.shape[0], 64))
    if precision == Precision.FP64:
        assert P.backward_error(A, s.solve(B), B) <= BE_TOL
    r = s.solveWithRefinement(A, B, RefinementOptions(10, 1e-14))
    assert P.backward_error(A, r.x, B) <= BE_TOL
