"""Deferred contribution-block update against the panel-by-panel one.

A large front whose eligible columns are factored by 64-column panels can
leave its contribution block (rows and columns past the eligible ones) out of
the panels' trailing updates and take it as one rank-nE product after the
last panel (dense_kernel.hpp:793-823 forms the same block one pivot block at
a time).  The pivot tests never read that block, so the pivot sequence,
permutation and inertia are the same and every entry of the factored front,
contribution block included, agrees to rounding.
"""
import numpy as np
import pytest

from ma57_amd import problems as P
from ma57_amd.solver import SymLDLT, OrderingType, Precision, dense_factor

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _front(n, ne, seed, pairs=0, delays=0):
    """Indefinite front with planted 2x2 pivots; `delays` eligible columns are zero inside the eligible block
    and coupled only to contribution-block rows, so no pivot can be taken on them."""
    rng = np.random.default_rng(seed)
    R = rng.standard_normal((n, n)) / np.sqrt(n)
    A = (R + R.T) / 2.0
    A += np.diag(np.where(rng.random(n) < 0.5, -1.0, 1.0) * rng.uniform(2.0, 4.0, n))
    for c in rng.choice(ne - 1 - delays, size=pairs, replace=False):
        A[c, c] = A[c + 1, c + 1] = 0.0
        A[c, c + 1] = A[c + 1, c] = 3.0
    for c in range(ne - delays, ne):
        A[c, :ne] = 0.0
        A[:ne, c] = 0.0
        A[c, ne:] = A[ne:, c] = 5.0 * rng.standard_normal(n - ne)
    return A


@pytest.mark.parametrize("n,ne,pairs,delays", [(600, 300, 3, 0), (3000, 1200, 8, 0), (2500, 900, 4, 7),
                                                 (5000, 2600, 6, 0)])
def test_dense_deferred_cb_matches_panel_updates(n, ne, pairs, delays):
    A = _front(n, ne, n + ne, pairs, delays)
    g = dense_factor(A, n_eligible=ne, deferred_cb=True)
    e = dense_factor(A, n_eligible=ne)
    assert g["n_factored"] == e["n_factored"] == ne - delays
    assert g["pivots"] == e["pivots"]
    assert list(g["perm"]) == list(e["perm"])
    assert g["inertia"] == e["inertia"]
    scale = np.abs(A).max()
    # the whole factored front: L, the delayed block and the updated contribution block
    assert np.abs(np.tril(g["A"]) - np.tril(e["A"])).max() <= 1e-11 * scale
    # the contribution block against its closed form: A22 - L21 D L21^T over the pivots taken
    k = ne - delays
    Ap = A[np.ix_(g["perm"], g["perm"])]
    L = np.tril(g["A"][:, :k], -1)
    L[np.arange(k), np.arange(k)] = 1.0
    D = np.diag(g["d_diag"][:k]) + np.diag(g["d_off"][:k - 1], -1) + np.diag(g["d_off"][:k - 1], 1)
    S = Ap[k:, k:] - L[k:] @ D @ L[k:].T
    assert np.abs(np.tril(g["A"][ne:, ne:]) - np.tril(S[ne - k:, ne - k:])).max() <= 1e-10 * scale


def _sparse(A, cb_rows, precision=Precision.FP64):
    s = SymLDLT()
    s.setOrdering(OrderingType.Metis)
    s.setPrecision(precision)
    s.analyzePattern(A)
    small, tile = s.tuning()
    s.setTuning(small, tile, deferred_cb_min_rows=cb_rows)
    assert s.deferredCbMinRows() == cb_rows
    s.factorize(A)
    return s


@pytest.mark.parametrize("name", ["laplace3d_32", "kkt_threshold"])
def test_sparse_deferred_cb_matches_panel_updates(name):
    """Every large front with at least 256 eligible columns and 64 contribution rows defers its block.  The
    SQD KKT matrix in threshold mode delays pivots, so its factorization goes through the repair path."""
    if name == "laplace3d_32":
        A = P.laplace3d(32, 32, 32, 0.5)
    else:
        A, _, _ = P.sqd_kkt(nh=16, m2=1200, seed=77)
    b = np.random.default_rng(5).standard_normal((A.shape[0], 3))
    out = {}
    for cb in (64, -1):
        s = _sparse(A, cb)
        inr = s.inertia()
        x = s.solve(b) if name == "laplace3d_32" else s.solveWithRefinement(A, b).x
        out[cb] = ((inr.n_pos, inr.n_neg, inr.n_zero), P.backward_error(A, x, b))
    assert out[64][0] == out[-1][0]
    assert out[64][1] <= 1e-12 and out[-1][1] <= 1e-12


def test_fp32_deferred_cb_refines():
    A = P.laplace3d(32, 32, 32, 0.5)
    b = np.random.default_rng(6).standard_normal((A.shape[0], 2))
    s = _sparse(A, 64, Precision.FP32)
    r = s.solveWithRefinement(A, b)
    inr = s.inertia()
    expect = tuple(P.closed_form_inertia(P.laplace_eigs_3d(32, 32, 32), 0.5)[:3])
    assert (inr.n_pos, inr.n_neg, inr.n_zero) == expect
    assert P.backward_error(A, r.x, b) <= 1e-12


def test_graph_replays_with_deferred_cb_match_eager():
    from ma57_amd.solver import csc_arrays
    G = 32
    shifts = [0.5, 0.41]
    mats = [P.laplace3d(G, G, G, sh) for sh in shifts]
    n = mats[0].shape[0]
    dev = torch.device("cuda:0")
    s = _sparse(mats[0], 64)
    eager = _sparse(mats[0], 64)
    d_vals = torch.from_numpy(csc_arrays(mats[0])[3].copy()).to(dev)
    d_B = torch.zeros((2, n), dtype=torch.float64, device=dev)
    d_X = torch.zeros_like(d_B)
    torch.cuda.synchronize()
    s.captureGraph(d_vals.data_ptr(), d_B.data_ptr(), 2, d_X.data_ptr())
    rng = np.random.default_rng(12)
    for rep in range(2):
        for A, sh in zip(mats, shifts):
            B = rng.standard_normal((n, 2))
            d_vals.copy_(torch.from_numpy(csc_arrays(A)[3]))
            d_B.copy_(torch.from_numpy(np.ascontiguousarray(B.T)))
            torch.cuda.synchronize()
            info = s.replayGraph()
            X = d_X.cpu().numpy().T
            eager.factorize(A)
            Xe = eager.solve(B)
            expect = tuple(P.closed_form_inertia(P.laplace_eigs_3d(G, G, G), sh)[:3])
            assert (info.n_pos, info.n_neg, info.n_zero) == expect
            assert info.repaired == 0
            assert np.array_equal(X, Xe)
            assert P.backward_error(A, X, B) <= 1e-12
            junk = torch.randn(4096, device=dev, dtype=torch.float64)
            assert torch.isfinite(junk.sum()).item()
            torch.cuda.synchronize()


def test_bf16x3_deferred_cb_refines_to_fp64_accuracy():
    """BF16x3 split products for the deferred contribution-block updates (FP32 fronts): the same inertia as
    FP64, FP64 refinement reaches the 1e-12 bar, and the unrefined solution sits between the FP32 one and a
    plain-BF16 update's (the dropped lo*lo term is below 2^-16 of each product, against 2^-24 for FP32 and
    2^-8 for one bf16 product)."""
    A = P.laplace3d(32, 32, 32, 0.5)
    b = np.random.default_rng(7).standard_normal((A.shape[0], 2))
    expect = tuple(P.closed_form_inertia(P.laplace_eigs_3d(32, 32, 32), 0.5)[:3])
    unrefined = {}
    for prec in (Precision.FP32, Precision.BF16X3):
        s = _sparse(A, 64, prec)
        inr = s.inertia()
        assert (inr.n_pos, inr.n_neg, inr.n_zero) == expect
        assert s.last_info.precision_used == prec
        unrefined[prec] = P.backward_error(A, s.solve(b), b)
        r = s.solveWithRefinement(A, b)
        assert P.backward_error(A, r.x, b) <= 1e-12
        assert r.diagnostics.iterations >= 1
    assert unrefined[Precision.BF16X3] < 1e-3
    assert unrefined[Precision.BF16X3] <= 200.0 * unrefined[Precision.FP32], unrefined
