"""The two configurations without a closed-form inertia, at their full sizes.

27-point stencil on 80^3 (n = 512,000) with random weights, shifted by
-0.2 I, 64 right-hand sides:
  * plain FP64 and FP32 fronts + FP64 refinement both reach backward error
    <= 1e-12 (the refinement count is reported);
  * the inertia is pinned by a second FP64 factorization of the same matrix
    under another ordering (minimum degree instead of nested dissection):
    Sylvester's law of inertia makes it ordering-invariant, and the two
    factorizations share no fronts (test/correctness/
    multifrontal_correctness_test.cpp:65-75 pins inertia by an independent
    factorization the same way);
  * the number of negative eigenvalues does not decrease as the shift grows
    (A - sigma I, sigma = 0.1, 0.2, 0.3);
  * ten replays of the recorded FP32 refactorize + refined solve on new
    seeded values, each at backward error <= 1e-12 and nonsingular; the
    first and last replays' inertia equals a minimum-degree FP64
    factorization of their own values.

2,048 independent systems (n_i ~ U[200, 5000], 5.4M unknowns) in one
batched plan:
  * every system's inertia equals that of a second batched FP64
    factorization in which each system's unknowns were relabeled by a random
    permutation before ordering (a different elimination order, same
    inertia by Sylvester), and numpy.linalg.eigvalsh on sampled systems;
  * FP64 and FP32 + per-system FP64 refinement reach per-system backward
    error <= 1e-12, with per-system refinement counts."""
import numpy as np
import pytest
import scipy.sparse as sp

from ma57_amd import problems as P
from ma57_amd.batch import BatchedLDLT
from ma57_amd.solver import SymLDLT, OrderingType, Precision, RefinementOptions, csc_arrays

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SEED = 0x53594D4C41
BAR = 1e-12


def _device_be(A_lower, d_X, d_B):
    """Per-column backward error with the symmetric infinity norm, max over columns, on the GPU (the formula
    of problems.backward_error; X, B are (nrhs, n) row-major device tensors)."""
    L = sp.tril(A_lower.tocsc(), format="csc")
    S = (L + L.T - sp.diags(L.diagonal())).tocsr()
    normA = float(np.abs(S).sum(axis=1).max())
    St = torch.sparse_csr_tensor(torch.from_numpy(S.indptr.astype(np.int64)),
                                 torch.from_numpy(S.indices.astype(np.int64)), torch.from_numpy(S.data),
                                 size=S.shape, dtype=torch.float64).to(d_X.device)
    R = d_B.T - St @ d_X.T.contiguous()
    den = normA * d_X.abs().amax(dim=1) + d_B.abs().amax(dim=1)
    return float((R.abs().amax(dim=0) / den).max().item())


def test_stencil27_80_full_size():
    N, nrhs = 80, 64
    A = P.stencil27(N, 0.2)
    n = A.shape[0]
    dev = torch.device("cuda:0")
    B = np.random.default_rng(SEED + 5).standard_normal((n, nrhs))
    d_B = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)
    d_X = torch.zeros_like(d_B)

    # second ordering first: the minimum-degree FP64 factorization pins the inertia
    s2 = SymLDLT()
    s2.setOrdering(OrderingType.AMD)
    s2.analyzePattern(A)
    s2.factorize(A)
    i2 = s2.inertia()
    pinned = (i2.n_pos, i2.n_neg, i2.n_zero)
    assert sum(pinned) == n and pinned[2] == 0
    del s2

    s = SymLDLT()
    s.setOrdering(OrderingType.Metis)
    s.analyzePattern(A)
    nd_front = max(len(x.rowPattern) for x in s.symbolicFactor().supernodes)
    # a front past 4,096 eligible rows: every panel front takes the speculative step by default
    assert s._get_tuning().speculative_min_rows == 0
    for prec in (Precision.FP64, Precision.FP32):
        s.setPrecision(prec)
        s.factorize(A)
        i = s.inertia()
        assert (i.n_pos, i.n_neg, i.n_zero) == pinned, prec
        r = s.solveWithRefinement(A, B, RefinementOptions(10, BAR))
        d_X.copy_(torch.from_numpy(np.ascontiguousarray(r.x.T)))
        assert _device_be(A, d_X, d_B) <= BAR, prec
        assert r.diagnostics.backward_error <= BAR
        if prec == Precision.FP32:
            assert 1 <= r.diagnostics.iterations <= 4
    assert nd_front < 12000  # the nested dissection keeps the root front near a grid plane (80^2 = 6,400)

    # shift sweep on the same pattern: n_neg(sigma) is non-decreasing
    base = csc_arrays(A)[3]
    diag = A.indices == np.repeat(np.arange(n), np.diff(A.indptr))
    negs = []
    s.setPrecision(Precision.FP64)
    for sigma in (0.1, 0.2, 0.3):
        As = A.copy()
        As.data = base + np.where(diag, 0.2 - sigma, 0.0)
        s.factorize(As)
        negs.append(s.inertia().n_neg)
    assert negs[1] == pinned[1] and negs[0] <= negs[1] <= negs[2], negs

    # ten recorded FP32 refactorize + refined solves on new seeded values
    d_vals = torch.from_numpy(base.copy()).to(dev)
    torch.cuda.synchronize()
    s.setPrecision(Precision.FP32)
    s.captureGraph(d_vals.data_ptr(), d_B.data_ptr(), nrhs, d_X.data_ptr(), refine=RefinementOptions(10, BAR))
    for r in range(10):
        Ar = P.stencil27(N, 0.2, seed=SEED + 5 + r)
        d_vals.copy_(torch.from_numpy(csc_arrays(Ar)[3]))
        torch.cuda.synchronize()
        info = s.replayGraph()
        rr = s.last_refinement
        assert info.repaired == 0 and info.singular == 0 and info.n_zero == 0, r
        assert rr.backward_error <= BAR and rr.iterations >= 1, (r, rr)
        assert _device_be(Ar, d_X, d_B) <= BAR, r
        if r in (0, 9):  # the replay's inertia against an FP64 factorization of the same values
            e = SymLDLT()
            e.setOrdering(OrderingType.AMD)
            e.analyzePattern(Ar)
            e.factorize(Ar)
            ie = e.inertia()
            assert (info.n_pos, info.n_neg, info.n_zero) == (ie.n_pos, ie.n_neg, ie.n_zero), r
            del e
    s.releaseGraph()


def _permuted(A, seed):
    """P A P^T (lower triangle) for a random permutation: the same system with its unknowns relabeled."""
    n = A.shape[0]
    p = np.random.default_rng(seed).permutation(n)
    inv = np.empty(n, np.int64)
    inv[p] = np.arange(n)
    C = A.tocoo()
    r, c = inv[C.row], inv[C.col]
    return sp.csc_matrix((C.data, (np.maximum(r, c), np.minimum(r, c))), shape=A.shape), p


def test_batch_2048_full_size():
    systems = [P.batched_system(SEED + 4 + 1000 * i) for i in range(2048)]
    b = BatchedLDLT(OrderingType.AMD)
    b.compute(systems)
    inr = b.inertia()
    assert b.singularCounts() == [0] * len(systems)
    rng = np.random.default_rng(SEED + 44)
    Bs = [rng.standard_normal(A.shape[0]) for A in systems]
    Xs = b.solve(Bs)
    worst = max(P.backward_error(A, X, Bv) for A, X, Bv in zip(systems, Xs, Bs))
    assert worst <= BAR
    del b

    # the same systems with relabeled unknowns: another elimination order, the same inertia
    relabeled = [_permuted(A, 7 + g)[0] for g, A in enumerate(systems)]
    b2 = BatchedLDLT(OrderingType.AMD)
    b2.compute(relabeled)
    assert b2.inertia() == inr
    del b2

    # Sylvester's inertia from eigvalsh on sampled systems (the three smallest and the median size)
    order = np.argsort([A.shape[0] for A in systems])
    for g in list(order[:3]) + [int(order[len(order) // 2])]:
        ev = np.linalg.eigvalsh(P.sym_dense(systems[g]))
        assert inr[g] == (int((ev > 0).sum()), int((ev < 0).sum()), int((ev == 0).sum())), g

    # FP32 fronts + per-system FP64 refinement
    b32 = BatchedLDLT(OrderingType.AMD, Precision.FP32)
    b32.compute(systems)
    assert b32.inertia() == inr
    Xs, diags = b32.solveWithRefinement(Bs, RefinementOptions(10, BAR))
    its = [d.iterations for d in diags]
    assert min(its) >= 1 and max(its) <= 10
    assert max(d.backward_error for d in diags) <= BAR
    assert max(P.backward_error(A, X, Bv) for A, X, Bv in zip(systems, Xs, Bs)) <= BAR
