"""The speculative panel step against the exact pivot-by-pivot panel.

Each panel of a large front is first factored as 1x1 pivots in order, the
threshold Bunch-Kaufman tests (dense_kernel.hpp:601-618) are checked for all
its columns at once, and the exact kernel continues from the first column
that fails.  The pivot decisions are the reference's; only the rounding of
the values they are taken on differs.  So: the same pivot sequence,
permutation and inertia as the exact path, factors equal to rounding, and
on sparse problems the same inertia with the plain FP64 backward-error bar
met either way.
"""
import numpy as np
import pytest

from ma57_amd import problems as P
from ma57_amd.solver import SymLDLT, Mode, OrderingType, Precision, dense_factor

pytestmark = pytest.mark.gpu

BE_TOL = 1e-12


def _dominant_indefinite(n, seed, n_pairs):
    """Diagonally weighted indefinite front whose panels mostly pass the 1x1 test, with n_pairs planted
    2x2 blocks (zero diagonals, large coupling) that make the test fail inside a panel."""
    rng = np.random.default_rng(seed)
    R = rng.standard_normal((n, n)) / np.sqrt(n)
    A = (R + R.T) / 2.0
    A += np.diag(np.where(rng.random(n) < 0.5, -1.0, 1.0) * rng.uniform(2.0, 4.0, n))
    for c in rng.choice(n - 1, size=n_pairs, replace=False):
        A[c, c] = A[c + 1, c + 1] = 0.0
        A[c, c + 1] = A[c + 1, c] = 3.0
    return A


@pytest.mark.parametrize("n,ne,pairs", [(300, 300, 0), (700, 600, 5), (2000, 2000, 12), (5000, 4800, 20)])
def test_dense_speculative_matches_exact(n, ne, pairs):
    A = _dominant_indefinite(n, n + pairs, pairs)
    g = dense_factor(A, n_eligible=ne)
    e = dense_factor(A, n_eligible=ne, exact_panel=True)
    assert g["n_factored"] == e["n_factored"] == ne
    assert g["pivots"] == e["pivots"]
    assert list(g["perm"]) == list(e["perm"])
    assert g["inertia"] == e["inertia"]
    if pairs:
        assert any(k == 2 for _, k in g["pivots"])
    scale = np.abs(A).max()
    assert np.abs(np.tril(g["A"]) - np.tril(e["A"])).max() <= 1e-11 * scale
    assert np.allclose(g["d_diag"][:ne], e["d_diag"][:ne], rtol=1e-11, atol=1e-13 * scale)


def test_dense_speculative_first_column_fails():
    """A zero leading diagonal fails the test at the first column: the exact kernel takes the whole first
    panel, starting with an interchange."""
    n = 400
    A = _dominant_indefinite(n, 5, 0)
    A[0, 0] = 0.0
    g = dense_factor(A)
    e = dense_factor(A, exact_panel=True)
    assert g["pivots"] == e["pivots"]
    assert list(g["perm"]) == list(e["perm"]) and g["perm"][0] != 0
    assert g["inertia"] == e["inertia"]
    assert np.abs(np.tril(g["A"]) - np.tril(e["A"])).max() <= 1e-11 * np.abs(A).max()


def _factor(A, spec, ordering=OrderingType.Metis, precision=Precision.FP64):
    s = SymLDLT()
    s.setOrdering(ordering)
    s.setPrecision(precision)
    s.analyzePattern(A)
    small, tile = s.tuning()
    s.setTuning(small, tile, spec, 0)  # every front's panels take the step when spec is on
    assert s.speculativePanel() == spec
    s.factorize(A)
    return s


@pytest.mark.parametrize("name", ["laplace3d_24", "kkt_threshold"])
def test_sparse_speculative_matches_exact(name):
    """The SQD KKT matrix in threshold mode has a 1e-8 (2,2) block: its unrefined backward error is about
    4e-10 on either path, so that case is compared after refinement."""
    if name == "laplace3d_24":
        A = P.laplace3d(24, 24, 24, 0.5)
    else:
        A, _, _ = P.sqd_kkt(nh=16, m2=1200, seed=77)
    b = np.random.default_rng(3).standard_normal((A.shape[0], 4))
    res = {}
    for spec in (True, False):
        s = _factor(A, spec)
        inr = s.inertia()
        x = s.solve(b) if name == "laplace3d_24" else s.solveWithRefinement(A, b).x
        res[spec] = ((inr.n_pos, inr.n_neg, inr.n_zero), P.backward_error(A, x, b))
    assert res[True][0] == res[False][0]
    assert res[True][1] <= BE_TOL and res[False][1] <= BE_TOL


def test_static_speculative_matches_exact():
    """Static pivots always pass: the whole factorization runs through the speculative step."""
    K, n1, n2 = P.sqd_kkt(nh=16, m2=1200, seed=78)
    b = np.random.default_rng(4).standard_normal((n1 + n2, 2))
    out = {}
    for spec in (True, False):
        s = SymLDLT()
        s.setOrdering(OrderingType.AMD)
        s.setMode(Mode.StaticRegularized)
        s.setKKTBlockSizes(n1, n2)
        s.analyzePattern(K)
        small, tile = s.tuning()
        s.setTuning(small, tile, spec, 0)
        s.factorize(K)
        inr = s.inertia()
        r = s.solveWithRefinement(K, b)
        out[spec] = ((inr.n_pos, inr.n_neg, inr.n_zero), P.backward_error(K, r.x, b))
    assert out[True][0] == out[False][0] == (n2, n1, 0)
    assert out[True][1] <= BE_TOL and out[False][1] <= BE_TOL
