"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Bars: exact inertia; backward error <= 1e-12 for the plain FP64 path;
pivot structure equal to the oracle on the reference's hand cases; factor
reconstruction P A P^T = L D L^T to 1e-10 on random dense fronts.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.linalg import lapack

from ma57_amd import problems as P
from ma57_amd.solver import (SymLDLT, Mode, OrderingType, SymbolicFactorOptions, RefinementOptions, LogicError,
                             RuntimeFailure, InvalidArgument, dense_factor)

pytestmark = pytest.mark.gpu

BE_TOL = 1e-12  # backward error bar of the plain FP64 path


def _reconstruct(res, n_f):
    """L D L^T of the leading n_f x n_f block (d21 slots excluded from L)."""
    F = res["A"]
    L = np.tril(F[:n_f, :n_f], -1) + np.eye(n_f)
    D = np.diag(res["d_diag"][:n_f])
    for s, k in res["pivots"]:
        if k == 2 and s + 1 < n_f:
            L[s + 1, s] = 0.0
            D[s + 1, s] = D[s, s + 1] = res["d_off"][s]
    return L @ D @ L.T


def _sym_rand(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    e = rng.uniform(0.5, 5.0, n) * np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return (Q * e) @ Q.T


def _lapack_inertia(A):
    lu, piv, info = lapack.dsytrf(A, lower=1)
    pos = neg = 0
    k = 0
    n = A.shape[0]
    while k < n:
        if piv[k] > 0:
            d = lu[k, k]
            pos += d > 0
            neg += d < 0
            k += 1
        else:
            d11, d21, d22 = lu[k, k], lu[k + 1, k], lu[k + 1, k + 1]
            det = d11 * d22 - d21 * d21
            if det < 0:
                pos += 1
                neg += 1
            elif d11 + d22 > 0:
                pos += 2
            else:
                neg += 2
            k += 2
    return pos, neg


HAND = {
    "diag3": (np.array([[10, 1, 1], [1, 8, 1], [1, 1, 6.0]]), {}),
    "forces2x2": (np.array([[0, 1, 0.2], [1, 0, 0.3], [0.2, 0.3, 5.0]]), {}),
    "mixed5": (np.array([[20, 1, 0.5, 0.1, 0.1], [1, 15, 0.2, 0.1, 0.1], [0.5, 0.2, 0, 2, 0.1],
                         [0.1, 0.1, 2, 0, 0.2], [0.1, 0.1, 0.1, 0.2, 12]]), {}),
    "rankdef": (np.array([[100, 0, 0], [0, 0, 1e-20], [0, 1e-20, 2e-20]]), {"zero_tolerance": 1e-8}),
}


@pytest.mark.parametrize("name", list(HAND))
def test_dense_hand_cases_match_oracle(name, oracle):
    A, kw = HAND[name]
    g = dense_factor(A, **kw)
    o = oracle.dense_factor(A, **kw)
    assert g["n_factored"] == o["n_factored"]
    assert g["pivots"] == o["pivots"]
    assert g["inertia"] == o["inertia"]
    assert list(g["perm"]) == list(o["perm"])
    nf = g["n_factored"]
    Ap = A[np.ix_(g["perm"][:nf], g["perm"][:nf])]
    assert np.linalg.norm(_reconstruct(g, nf) - Ap) / np.linalg.norm(A) < 1e-12


def test_dense_delayed_trailing_block(oracle):
    A = np.zeros((5, 5))
    A[0, 0], A[1, 1], A[2, 2] = 10, 8, 6
    A[0, 1] = A[1, 0] = 1
    A[0, 2] = A[2, 0] = 0.5
    A[1, 2] = A[2, 1] = 0.2
    g = dense_factor(A)
    assert g["n_factored"] == 3 and g["delayed"] == [3, 4]


@pytest.mark.parametrize("n,seed", [(10, 1), (50, 2), (150, 3), (200, 4), (300, 5), (520, 6)])
def test_dense_random_reconstruction_and_lapack_inertia(n, seed, oracle):
    A = _sym_rand(n, seed)
    g = dense_factor(A)
    assert g["n_factored"] == n
    Ap = A[np.ix_(g["perm"], g["perm"])]
    assert np.linalg.norm(_reconstruct(g, n) - Ap) / np.linalg.norm(A) < 1e-10
    pos, neg = _lapack_inertia(A)
    assert g["inertia"] == (pos, neg, 0)
    assert g["inertia"] == oracle.dense_factor(A)["inertia"]


@pytest.mark.parametrize("n,ne,seed", [(40, 20, 7), (260, 100, 8), (400, 300, 9)])
def test_dense_front_schur_complement(n, ne, seed, oracle):
    """A front with n_eligible < n: the trailing block is the Schur complement handed to the parent."""
    rng = np.random.default_rng(seed)
    A = _sym_rand(n, seed) + np.diag(rng.uniform(2, 3, n))
    g = dense_factor(A, n_eligible=ne)
    o = oracle.dense_factor(A, n_eligible=ne)
    assert g["n_factored"] == o["n_factored"] == ne
    assert g["inertia"] == o["inertia"]
    Sg = np.tril(g["A"][ne:, ne:])
    So = np.tril(o["A"][ne:, ne:])
    assert np.abs(Sg - So).max() <= 1e-11 * max(1.0, np.abs(So).max())


@pytest.mark.parametrize("n,ne", [(5000, 5000), (5200, 4800)])
def test_dense_cooperative_panel(n, ne):
    """Fronts with more than 4096 eligible rows take the cooperative (multi-workgroup) panel: an indefinite
    random front needs 2x2 pivots and swaps (the general step); checked against LAPACK's inertia
    (Sylvester), the reconstruction of the eligible block, and the Schur complement of the rest."""
    rng = np.random.default_rng(n)
    R = rng.standard_normal((n, n))
    A = (R + R.T) / 2.0
    g = dense_factor(A, n_eligible=ne)
    assert g["n_factored"] == ne
    assert any(k == 2 for _, k in g["pivots"])
    A11 = A[:ne, :ne]
    pos, neg = _lapack_inertia(A11)
    assert g["inertia"] == (pos, neg, 0)
    p = g["perm"][:ne]
    assert np.linalg.norm(_reconstruct(g, ne) - A11[np.ix_(p, p)]) / np.linalg.norm(A11) < 1e-10
    if ne < n:
        S = A[ne:, ne:] - A[ne:, :ne] @ np.linalg.solve(A11, A[:ne, ne:])
        Sg = np.tril(g["A"][ne:, ne:])
        assert np.abs(Sg - np.tril(S)).max() <= 1e-8 * max(1.0, np.abs(S).max())


@pytest.mark.parametrize("n", [30, 250, 400])
def test_dense_static_matches_oracle(n, oracle):
    """Quasi-definite front (diagonal dominant with the expected signs) plus a few wrong-signed and tiny
    diagonals that must be perturbed; static pivoting is unpivoted, so a well-conditioned front keeps the
    comparison about the kernel and not about error growth."""
    rng = np.random.default_rng(n)
    sign = np.where(rng.random(n) < 0.5, -1, 1).astype(np.int32)
    A = 0.2 * _sym_rand(n, n + 1) / np.sqrt(n) + np.diag(sign * rng.uniform(2.0, 3.0, n))
    bad = rng.choice(n, size=max(4, n // 20), replace=False)
    wrong = bad[0::2]
    A[wrong, wrong] = -sign[wrong] * 2.0  # wrong-signed: must be perturbed by +sign*delta
    A[bad[1::2], bad[1::2]] = 1e-12
    g = dense_factor(A, static_mode=True, static_delta=5.0, expected_sign=sign)
    o = oracle.dense_factor_static(A, expected_sign=sign, delta=5.0, panel_size=1 if n < 256 else 64)
    assert g["n_factored"] == o["n_factored"] == n
    assert g["inertia"] == o["inertia"]
    assert g["num_perturbed"] == o["num_perturbed"] >= len(wrong)
    assert abs(g["total_perturbation"] - o["total_perturbation"]) <= 1e-12 * o["total_perturbation"]
    assert np.allclose(g["d_diag"], o["d_diag"], rtol=1e-11, atol=1e-14)


def _solver(A, ordering=OrderingType.AMD, options=None):
    s = SymLDLT()
    s.setOrdering(ordering)
    s.analyzePattern(A, options)
    s.factorize(A)
    return s


def test_c1_laplacian_inertia_and_backward_error():
    A = P.laplace2d(100, 100, 0.3)
    pos, neg, zero, margin = P.closed_form_inertia(P.laplace_eigs_2d(100, 100), 0.3)
    s = _solver(A)
    inr = s.inertia()
    assert (inr.n_pos, inr.n_neg, inr.n_zero) == (pos, neg, zero) == (9770, 230, 0)
    b = np.random.default_rng(0x53594D4C41 + 1).standard_normal(A.shape[0])
    x = s.solve(b)
    assert P.backward_error(A, x, b) <= BE_TOL


def test_arrow_delayed_pivots():
    n = 5
    A = np.zeros((n, n))
    A[2, 2], A[3, 3], A[4, 4] = 5.0, 7.0, 20.0
    A[2, 3] = A[3, 2] = 1.1
    A[3, 4] = A[4, 3] = 1.3
    A[2, 4] = A[4, 2] = 0.9
    A[0, 4] = A[4, 0] = 2.0
    A[1, 3] = A[3, 1] = 1.7
    S = sp.csc_matrix(A)
    s = _solver(S, OrderingType.Natural, SymbolicFactorOptions(1, 0.0))
    sf = s.symbolicFactor()
    assert [(x.firstCol, x.ncols) for x in sf.supernodes] == [(0, 1), (1, 1), (2, 3)]
    fronts = s.numericFactor()
    assert [f.nPivots for f in fronts] == [0, 0, 5]
    assert not s.isSingular()
    inr = s.inertia()
    assert inr.n_pos + inr.n_neg == 5 and inr.n_zero == 0
    b = np.array([1, 2, 3, 4, 5.0])
    x = s.solve(b)
    assert np.linalg.norm(A @ x - b) / (np.abs(A).max() * np.linalg.norm(x) + np.linalg.norm(b)) < 1e-10
    assert s.last_info.repaired == 1


def test_genuinely_singular():
    A = np.zeros((3, 3))
    A[0, 0], A[2, 2] = 4.0, 3.0
    A[0, 2] = A[2, 0] = 0.5
    s = SymLDLT()
    s.compute(sp.csc_matrix(A))
    assert s.factorized() and s.isSingular()
    assert s.singularColumns() == [1]
    assert sum(f.nPivots for f in s.numericFactor()) == 2
    with pytest.raises(LogicError):
        s.solve(np.ones(3))


def test_solve_before_factorize_raises():
    s = SymLDLT()
    with pytest.raises(LogicError):
        s.solve(np.zeros(2))
    with pytest.raises(LogicError):
        s.factorize(sp.identity(2, format="csc"))


def test_smoke_diag():
    A = sp.csc_matrix(np.diag([2.0, 3.0]))
    s = SymLDLT()
    s.setPivotThreshold(0.01)
    s.analyzePattern(A)
    s.factorize(A)
    assert (s.inertia().n_pos, s.inertia().n_neg) == (2, 0)
    x = s.solve(np.array([4.0, 9.0]))
    assert abs(x[0] - 2) < 1e-12 and abs(x[1] - 3) < 1e-12
    with pytest.raises(InvalidArgument):
        s.solve(np.ones(3))


def _hinge(nh, seed):
    rng = np.random.default_rng(seed)
    n = 2 * nh
    r, c, v = [], [], []
    for h in range(nh):
        a, b = 2 * h, 2 * h + 1
        cc = rng.uniform(0.5, 2.0)
        r += [a, b, b]; c += [a, b, a]; v += [0.0, 0.0, cc]
    for h in range(nh - 1):
        r.append(2 * (h + 1)); c.append(2 * h); v.append(rng.uniform(-0.05, 0.05))
    return sp.csc_matrix((v, (r, c)), shape=(n, n))


@pytest.mark.parametrize("nh,ordering", [(2, OrderingType.AMD), (10, OrderingType.Natural), (50, OrderingType.AMD),
                                         (200, OrderingType.Natural)])
def test_hinge_chain_two_by_two(nh, ordering):
    A = _hinge(nh, 1000 + nh)
    s = _solver(A, ordering)
    inr = s.inertia()
    assert (inr.n_pos, inr.n_neg, inr.n_zero) == (nh, nh, 0)
    n2 = sum(1 for f in s.numericFactor() for (_, k) in f.pivotBlocks if k == 2)
    assert n2 >= nh
    Ad = P.sym_dense(A)
    xt = np.random.default_rng(nh).standard_normal(2 * nh)
    b = Ad @ xt
    x = s.solve(b)
    assert P.backward_error(A, x, b) < 1e-13


def _kkt(n1, n2, seed):
    E = P.random_sparse_spd(n1, 4, seed)
    F = P.random_sparse_spd(n2, 4, seed + 17)
    rng = np.random.default_rng(seed + 31)
    r, c, v = [], [], []
    for i in range(n2):
        for _ in range(1 + int(rng.integers(0, 3))):
            r.append(n1 + i); c.append(int(rng.integers(0, n1))); v.append(rng.uniform(-0.5, 0.5))
    Ec, Fc = sp.tril(E).tocoo(), sp.tril(F).tocoo()
    rows = np.concatenate([Ec.row, n1 + Fc.row, r])
    cols = np.concatenate([Ec.col, n1 + Fc.col, c])
    vals = np.concatenate([-Ec.data, Fc.data, v])
    return sp.csc_matrix((vals, (rows, cols)), shape=(n1 + n2, n1 + n2))


@pytest.mark.parametrize("n1,n2,seed", [(10, 6, 17), (80, 40, 19), (50, 5, 4242), (5, 50, 4243)])
def test_kkt_threshold_inertia(n1, n2, seed):
    K = _kkt(n1, n2, seed)
    s = _solver(K)
    inr = s.inertia()
    assert (inr.n_pos, inr.n_neg, inr.n_zero) == (n2, n1, 0)
    xt = np.random.default_rng(seed).standard_normal(n1 + n2)
    b = P.sym_matvec(K, xt)
    assert P.backward_error(K, s.solve(b), b) < 1e-12


def test_static_regularized_paths():
    K = _kkt(40, 20, 11)
    s = SymLDLT()
    s.setMode(Mode.StaticRegularized)
    s.setKKTBlockSizes(40, 20)
    s.analyzePattern(K)
    s.factorize(K)
    assert (s.inertia().n_pos, s.inertia().n_neg, s.inertia().n_zero) == (20, 40, 0)
    assert s.regularizationRetriesUsed() <= 1
    # near-SQD: wrong-signed diagonal forces escalation; budget 0 throws
    Kl = K.tolil()
    Kl[0, 0] = 0.05
    Kn = Kl.tocsc()
    s2 = SymLDLT()
    s2.setMode(Mode.StaticRegularized)
    s2.setKKTBlockSizes(40, 20)
    s2.setInitialRegularizationDelta(1e-6)
    s2.analyzePattern(Kn)
    s2.factorize(Kn)
    assert (s2.inertia().n_pos, s2.inertia().n_neg) == (20, 40)
    assert s2.regularizationRetriesUsed() >= 1
    s3 = SymLDLT()
    s3.setMode(Mode.StaticRegularized)
    s3.setKKTBlockSizes(40, 20)
    s3.setInitialRegularizationDelta(1e-6)
    s3.setMaxRegularizationRetries(0)
    s3.analyzePattern(Kn)
    with pytest.raises(RuntimeFailure):
        s3.factorize(Kn)
    assert not s3.factorized()


def test_refinement_recovers_regularization_error():
    K = _kkt(40, 20, 3034)
    s = SymLDLT()
    s.setMode(Mode.StaticRegularized)
    s.setKKTBlockSizes(40, 20)
    s.setStaticPivotFloor(100.0)
    s.setInitialRegularizationDelta(0.005)
    s.analyzePattern(K)
    s.factorize(K)
    xt = np.random.default_rng(3034 + 555).standard_normal(60)
    b = P.sym_matvec(K, xt)
    raw = P.backward_error(K, s.solve(b), b)
    ref = s.solveWithRefinement(K, b, RefinementOptions(10, 1e-14))
    fin = P.backward_error(K, ref.x, b)
    if raw > 1e-10:
        assert fin <= raw * 1e-2
    assert fin < 1e-12
    assert ref.diagnostics.iterations >= 1


def test_multi_rhs_matches_single(oracle):
    A = P.random_sparse_indefinite(400, 5, 77)
    s = _solver(A)
    B = np.random.default_rng(5).standard_normal((400, 5))
    X = s.solve(B)
    for j in range(5):
        xj = s.solve(B[:, j])
        assert np.linalg.norm(xj - X[:, j]) <= 1e-12 * max(1.0, np.linalg.norm(X[:, j]))


@pytest.mark.parametrize("n,seed,ordering", [(20, 1, OrderingType.AMD), (100, 2, OrderingType.Natural),
                                             (500, 3, OrderingType.AMD), (2000, 4, OrderingType.Metis)])
def test_random_indefinite_matches_oracle(n, seed, ordering, oracle):
    A = P.random_sparse_indefinite(n, 5, seed)
    s = _solver(A, ordering)
    sym = oracle.Symbolic.from_product(s.symbolicFactor())
    o = oracle.mf_factorize(sym, A)
    inr = s.inertia()
    assert (inr.n_pos, inr.n_neg, inr.n_zero) == o.inertia()
    b = np.random.default_rng(seed + 555).standard_normal(n)
    xg = s.solve(b)
    xo = o.solve(b)[:, 0]
    assert P.backward_error(A, xg, b) <= BE_TOL
    assert np.linalg.norm(xg - xo) <= 1e-9 * np.linalg.norm(xo)


def test_determinism_bitwise():
    A = P.laplace3d(16, 16, 16, 0.5)
    s = SymLDLT()
    s.setOrdering(OrderingType.Metis)
    s.analyzePattern(A)
    b = np.random.default_rng(3).standard_normal((A.shape[0], 2))
    s.factorize(A)
    x1 = s.solve(b)
    s.refactorize(A)
    x2 = s.solve(b)
    assert np.array_equal(x1, x2)


def test_c2_laplacian_40cubed():
    A = P.laplace3d(40, 40, 40, 0.5)
    pos, neg, zero, _ = P.closed_form_inertia(P.laplace_eigs_3d(40, 40, 40), 0.5)
    s = SymLDLT()
    s.setOrdering(OrderingType.Metis)
    s.analyzePattern(A)
    s.factorize(A)
    inr = s.inertia()
    assert (inr.n_pos, inr.n_neg, inr.n_zero) == (pos, neg, zero) == (63671, 329, 0)
    rng = np.random.default_rng(0x53594D4C41 + 2)
    for nrhs in (1, 16):
        B = rng.standard_normal((A.shape[0], nrhs))
        X = s.solve(B)
        assert P.backward_error(A, X, B) <= BE_TOL
