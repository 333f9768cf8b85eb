"""Pins the CPU oracle (oracle/symla_oracle.cpp) before it is trusted as the checker.

Checks, per oracle function:
* DenseLDLT::factor restatement -- the reference's known-answer cases
  (tests/golden/reference_cases.json), LAPACK dsytrf inertia (scipy), Sylvester
  inertia from eigvalsh, reconstruction P A P^T = L D L^T, panel-size invariance,
  and planted mistakes that each check must catch.
* DenseLDLT::factorStatic -- perturbation rule against a direct evaluation of
  the rule on a diagonal matrix (closed form), SQD inertia, reconstruction.
* Multifrontal factorize/solve -- reference cases (arrow delayed pivots,
  singular column, smoke solve), closed-form Laplacian spectra, KKT/SQD
  Sylvester inertia, flattened-factor reconstruction, solve residuals,
  multi-RHS consistency, serial vs OpenMP task-DAG equivalence.
* Refinement and the static retry loop -- the reference's own assertions.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.linalg import lapack

from ma57_amd import problems as P
from ma57_amd.solver import analyze, SymbolicFactorOptions

CASES = json.loads((Path(__file__).parent / "golden" / "reference_cases.json").read_text())


def recon(res, nf):
    F = res["A"]
    L = np.tril(F[:nf, :nf], -1) + np.eye(nf)
    D = np.diag(res["d_diag"][:nf]).astype(float)
    for s, k in res["pivots"]:
        if k == 2 and s + 1 < nf:
            L[s + 1, s] = 0.0
            D[s + 1, s] = D[s, s + 1] = res["d_off"][s]
    return L, D


def recon_err(A, res, nf):
    L, D = recon(res, nf)
    p = res["perm"][:nf]
    return np.linalg.norm(L @ D @ L.T - A[np.ix_(p, p)]) / np.linalg.norm(A)


def sym_rand(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    e = rng.uniform(0.5, 5.0, n) * np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return (Q * e) @ Q.T


def lapack_inertia(A):
    lu, piv, info = lapack.dsytrf(A, lower=1)
    pos = neg = 0
    k, n = 0, A.shape[0]
    while k < n:
        if piv[k] > 0:
            pos += lu[k, k] > 0
            neg += lu[k, k] < 0
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
    return int(pos), int(neg)


def eig_inertia(A):
    e = np.linalg.eigvalsh(A)
    return int((e > 0).sum()), int((e < 0).sum()), int((e == 0).sum())


# ---------------------------------------------------------------- dense factor

@pytest.mark.parametrize("case", CASES["dense"], ids=lambda c: c["name"])
def test_dense_reference_cases(case, oracle):
    A = np.array(case["A"], dtype=float)
    r = oracle.dense_factor(A, **case["opts"])
    e = case["expect"]
    assert r["n_factored"] == e["n_factored"]
    if "pivot_kinds" in e:
        assert [k for _, k in r["pivots"]] == e["pivot_kinds"]
    if "perm" in e:
        assert list(r["perm"]) == e["perm"]
    if "first_pivot" in e:
        assert list(r["pivots"][0]) == e["first_pivot"]
    if "inertia" in e:
        assert list(r["inertia"]) == e["inertia"]
    if "n_zero" in e:
        assert r["inertia"][2] == e["n_zero"]
    if "delayed" in e:
        assert r["delayed"] == e["delayed"]
    if "recon_tol" in e:
        assert recon_err(A, r, r["n_factored"]) < e["recon_tol"]


@pytest.mark.parametrize("n", [5, 10, 25, 50, 100])
@pytest.mark.parametrize("seed", [11, 22, 33])
def test_dense_lapack_crosscheck(n, seed, oracle):
    """dense_kernel_test.cpp:142-205, :305-333: inertia equal to LAPACK dsytrf's."""
    A = sym_rand(n, seed * 1000 + n)
    r = oracle.dense_factor(A)
    assert r["n_factored"] == n
    assert r["inertia"][:2] == lapack_inertia(A)
    assert r["inertia"] == eig_inertia(A)
    assert recon_err(A, r, n) < 1e-10
    assert r["max_growth"] < 1000.0


@pytest.mark.parametrize("panel", [1, 2, 4, 64, 10_000])
def test_dense_panel_size_invariance(panel, oracle):
    """The blocked schedule is a reassociation of the unblocked one (dense_kernel.hpp:386-392)."""
    A = sym_rand(90, 5)
    ref = oracle.dense_factor(A, panel_size=1)
    r = oracle.dense_factor(A, panel_size=panel)
    assert r["inertia"] == ref["inertia"]
    assert recon_err(A, r, 90) < 1e-12


def test_dense_front_schur_complement_closed_form(oracle):
    """With n_eligible < n the trailing block is the Schur complement S = A22 - A21 A11^{-1} A12."""
    rng = np.random.default_rng(3)
    n, ne = 40, 25
    A = sym_rand(n, 9) + np.diag(rng.uniform(2, 3, n))
    r = oracle.dense_factor(A, n_eligible=ne)
    assert r["n_factored"] == ne
    p = r["perm"]
    assert list(p[ne:]) == list(range(ne, n))
    A11 = A[np.ix_(p[:ne], p[:ne])]
    A21 = A[np.ix_(p[ne:], p[:ne])]
    S = A[ne:, ne:] - A21 @ np.linalg.solve(A11, A21.T)
    assert np.abs(np.tril(r["A"][ne:, ne:]) - np.tril(S)).max() < 1e-11 * np.abs(S).max()


def test_dense_planted_mistakes_are_caught(oracle):
    """Each check above fails on a plausible corruption of the oracle's output."""
    A = sym_rand(30, 4)
    r = oracle.dense_factor(A)
    assert recon_err(A, r, 30) < 1e-10
    bad = dict(r, A=r["A"].copy())
    bad["A"][20, 7] = -bad["A"][20, 7]  # sign flip of one multiplier
    assert recon_err(A, bad, 30) > 1e-6
    bad = dict(r, perm=r["perm"].copy())
    bad["perm"][[3, 17]] = bad["perm"][[17, 3]]  # transposed permutation entries
    assert recon_err(A, bad, 30) > 1e-6
    bad = dict(r, d_diag=r["d_diag"].copy())
    bad["d_diag"][5] *= -1  # dropped sign of one pivot -> inertia and reconstruction both off
    assert recon_err(A, bad, 30) > 1e-6


def test_static_rule_on_diagonal_closed_form(oracle):
    """On a diagonal matrix factorStatic's pivots are the perturbation rule itself (dense_kernel.hpp:982-1011)."""
    d = np.array([3.0, -2.0, 1e-12, -1e-12, 0.5, -0.5, 0.0, 4.0])
    sign = np.array([1, -1, 1, -1, -1, 1, 0, -1], dtype=np.int32)
    delta = 0.25
    r = oracle.dense_factor_static(np.diag(d), expected_sign=sign, delta=delta)
    floor = 1e-8 * np.abs(d).max()
    exp = []
    for a, s in zip(d, sign):
        need = (a < floor) if s > 0 else (a > -floor) if s < 0 else (abs(a) < floor)
        if need:
            ss = s if s != 0 else (1.0 if a >= 0 else -1.0)
            exp.append(a + ss * delta)
        else:
            exp.append(a)
    assert np.allclose(r["d_diag"], exp, rtol=0, atol=0)
    assert r["num_perturbed"] == 6


def test_static_sqd_inertia_and_reconstruction(oracle):
    rng = np.random.default_rng(8)
    n1, n2 = 30, 20
    E = sym_rand(n1, 1) @ sym_rand(n1, 1) + np.eye(n1)
    F = np.eye(n2) * 2
    J = rng.standard_normal((n2, n1)) * 0.3
    K = np.block([[-E, J.T], [J, F]])
    sign = np.r_[-np.ones(n1), np.ones(n2)].astype(np.int32)
    r = oracle.dense_factor_static(K, expected_sign=sign)
    assert r["inertia"] == (n2, n1, 0) == eig_inertia(K)[:2] + (0,)
    assert r["num_perturbed"] == 0
    L = np.tril(r["A"], -1) + np.eye(n1 + n2)
    assert np.linalg.norm(L @ np.diag(r["d_diag"]) @ L.T - K) / np.linalg.norm(K) < 1e-12


# ------------------------------------------------------------- multifrontal

def _analyze(A, ordering="amd", relax=None):
    order = {"amd": 0, "metis": 1, "natural": 2}[ordering]
    opts = SymbolicFactorOptions(*relax) if relax else None
    h, sf = analyze(sp.csc_matrix(A), order, opts)
    return sf


def flatten(sf, f, n):
    """test/unit/multifrontal_test_helpers.hpp:40-88: dense L, D in actual elimination order."""
    fronts = [f.front(i) for i in range(len(sf.supernodes))]
    order = [int(fr["rows"][p]) for fr in fronts for p in range(fr["npiv"])]
    pos = {g: t for t, g in enumerate(order)}
    L = np.eye(n)
    D = np.zeros((n, n))
    for fr in fronts:
        k = fr["npiv"]
        two = {s + 1 for s, kk in fr["pivots"] if kk == 2}
        for p in range(k):
            tp = pos[int(fr["rows"][p])]
            for i in range(p + 1, len(fr["rows"])):
                if i < k and i in two and i == p + 1:
                    continue
                L[pos[int(fr["rows"][i])], tp] = fr["L"][i, p]
        for s, kk in fr["pivots"]:
            D[pos[int(fr["rows"][s])], pos[int(fr["rows"][s])]] = fr["d_diag"][s]
            if kk == 2:
                a, b = pos[int(fr["rows"][s])], pos[int(fr["rows"][s + 1])]
                D[b, b] = fr["d_diag"][s + 1]
                D[a, b] = D[b, a] = fr["d_off"][s]
    orig = np.array([sf.perm[g] for g in order])
    return L, D, orig


@pytest.mark.parametrize("case", CASES["multifrontal"], ids=lambda c: c["name"])
def test_multifrontal_reference_cases(case, oracle):
    A = np.array(case["A"], dtype=float)
    n = A.shape[0]
    S = sp.csc_matrix(A)
    sf = _analyze(A, case["ordering"], case.get("relax"))
    e = case["expect"]
    if "supernodes" in e:
        assert [[s.firstCol, s.ncols] for s in sf.supernodes] == e["supernodes"]
    f = oracle.mf_factorize(oracle.Symbolic.from_product(sf), S)
    info = f.info()
    assert bool(info.singular) == e["singular"] if "singular" in e else True
    if "npivots" in e:
        assert [f.front(i)["npiv"] for i in range(info.nsn)] == e["npivots"]
    if "inertia" in e:
        assert list(f.inertia()) == e["inertia"]
    if "singular_cols" in e:
        assert [int(sf.perm[c]) for c in f.singular_cols()] == e["singular_cols"]
    if "total_pivots" in e:
        assert info.total_pivots == e["total_pivots"]
    if not info.singular:
        L, D, orig = flatten(sf, f, n)
        Ap = A[np.ix_(orig, orig)]
        assert np.linalg.norm(L @ D @ L.T - Ap) / max(1.0, np.abs(Ap).max()) < 1e-10
        if "b" in case:
            b = np.array(case["b"], dtype=float)
            x = f.solve(b)[:, 0]
            if "x" in e:
                assert np.abs(x - np.array(e["x"])).max() < e["x_tol"]
            if "resid_tol" in e:
                assert np.linalg.norm(A @ x - b) / (np.abs(A).max() * np.linalg.norm(x) + np.linalg.norm(b)) < \
                    e["resid_tol"]


def test_c1_closed_form_inertia(oracle):
    A = P.laplace2d(100, 100, 0.3)
    pos, neg, zero, margin = P.closed_form_inertia(P.laplace_eigs_2d(100, 100), 0.3)
    assert (pos, neg, zero) == (9770, 230, 0) and margin > 9e-4
    sf = _analyze_sparse(A, 0)
    f = oracle.mf_factorize(oracle.Symbolic.from_product(sf), A)
    assert f.inertia() == (pos, neg, zero)
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    x = f.solve(b)
    assert P.backward_error(A, x, b) <= 1e-12


def _analyze_sparse(A, order):
    h, sf = analyze(A, order)
    return sf


def test_laplacian_closed_form_planted_shift(oracle):
    """A dropped shift term (planted in the input, not the code) must change the inertia the check expects."""
    A = P.laplace2d(30, 30, 0.3)
    sf = _analyze_sparse(A, 0)
    f = oracle.mf_factorize(oracle.Symbolic.from_product(sf), A)
    right = P.closed_form_inertia(P.laplace_eigs_2d(30, 30), 0.3)[:3]
    wrong = P.closed_form_inertia(P.laplace_eigs_2d(30, 30), 0.0)[:3]
    assert f.inertia() == right and f.inertia() != wrong


def test_3d_laplacian_closed_form_nd(oracle):
    A = P.laplace3d(12, 12, 12, 0.5)
    sf = _analyze_sparse(A, 1)
    f = oracle.mf_factorize(oracle.Symbolic.from_product(sf), A, parallel=True, threads=4)
    assert f.inertia() == P.closed_form_inertia(P.laplace_eigs_3d(12, 12, 12), 0.5)[:3]



@pytest.mark.parametrize("parallel", [False, True])
def test_timing_sample_fronts_equal_full_factorization(parallel, oracle):
    """The CPU baseline's timing sample (the first K supernodes in postorder, complete subtrees) factors those
    fronts exactly as the full factorization does: same rows, pivots, D and L."""
    A = P.laplace3d(12, 12, 12, 0.5)
    sym = oracle.Symbolic.from_product(_analyze_sparse(A, 1))
    full = oracle.mf_factorize(sym, A)
    K = len(sym.first) // 3
    samp = oracle.mf_factorize(sym, A, parallel=parallel, threads=4, sn_limit=K)
    assert samp.info().nsn == K
    for si in range(K):
        a, b = full.front(si), samp.front(si)
        assert np.array_equal(a["rows"], b["rows"]) and a["pivots"] == b["pivots"]
        assert np.array_equal(a["d_diag"], b["d_diag"]) and np.array_equal(a["L"], b["L"])

def _kkt(n1, n2, seed):
    E = P.random_sparse_spd(n1, 4, seed)
    F = P.random_sparse_spd(n2, 4, seed + 17)
    rng = np.random.default_rng(seed + 31)
    r, c, v = [], [], []
    for i in range(n2):
        for _ in range(1 + int(rng.integers(0, 3))):
            r.append(n1 + i); c.append(int(rng.integers(0, n1))); v.append(rng.uniform(-0.5, 0.5))
    Ec, Fc = sp.tril(E).tocoo(), sp.tril(F).tocoo()
    return sp.csc_matrix((np.concatenate([-Ec.data, Fc.data, v]),
                          (np.concatenate([Ec.row, n1 + Fc.row, r]), np.concatenate([Ec.col, n1 + Fc.col, c]))),
                         shape=(n1 + n2, n1 + n2))


@pytest.mark.parametrize("n1,n2,seed", [(10, 6, 17), (80, 40, 19), (50, 5, 4242), (5, 50, 4243)])
def test_kkt_sylvester_inertia(n1, n2, seed, oracle):
    """kkt_test.cpp:104-137: inertia (n2, n1, 0) regardless of pivot order."""
    K = _kkt(n1, n2, seed)
    sf = _analyze_sparse(K, 0)
    f = oracle.mf_factorize(oracle.Symbolic.from_product(sf), K)
    assert f.inertia() == (n2, n1, 0)
    assert eig_inertia(P.sym_dense(K)) == (n2, n1, 0)
    xt = np.random.default_rng(seed).standard_normal(n1 + n2)
    b = P.sym_matvec(K, xt)
    assert P.backward_error(K, f.solve(b), b) < 1e-13


@pytest.mark.parametrize("n,seed", [(20, 1), (100, 2), (500, 3)])
def test_random_indefinite_against_eigvalsh(n, seed, oracle):
    """multifrontal_correctness_test.cpp:39-76 with eigvalsh as the independent inertia oracle."""
    A = P.random_sparse_indefinite(n, 5, seed)
    for order in (0, 2):
        sf = _analyze_sparse(A, order)
        f = oracle.mf_factorize(oracle.Symbolic.from_product(sf), A)
        assert f.inertia() == eig_inertia(A.toarray())
        L, D, orig = flatten(sf, f, n)
        Ad = A.toarray()
        Ap = Ad[np.ix_(orig, orig)]
        assert np.linalg.norm(L @ D @ L.T - Ap) / max(1.0, np.abs(Ap).max()) < 1e-10


def test_multi_rhs_equals_single_and_parallel_equals_serial(oracle):
    """solve_test.cpp:69-75 and thread_equivalence_test.cpp:72-107."""
    A = P.laplace3d(14, 14, 14, 0.5)
    sf = _analyze_sparse(A, 1)
    sym = oracle.Symbolic.from_product(sf)
    fs = oracle.mf_factorize(sym, A)
    fp = oracle.mf_factorize(sym, A, parallel=True, threads=4)
    assert fs.inertia() == fp.inertia()
    B = np.random.default_rng(2).standard_normal((A.shape[0], 5))
    Xs = fs.solve(B)
    Xp = fp.solve(B, parallel=True, threads=4)
    assert np.linalg.norm(Xs - Xp) <= 1e-10 * np.linalg.norm(Xs)
    for j in range(5):
        xj = fs.solve(B[:, j])[:, 0]
        assert np.linalg.norm(xj - Xs[:, j]) <= 1e-12 * max(1.0, np.linalg.norm(Xs[:, j]))
    fp2 = oracle.mf_factorize(sym, A, parallel=True, threads=4)
    assert np.array_equal(fp2.solve(B), fs.solve(B))  # fixed child order -> repeatable


def test_static_regularized_loop(oracle):
    """kkt_static_pivoting_test.cpp: genuine SQD needs <= 1 retry; near-SQD escalates; budget 0 throws."""
    K = _kkt(40, 20, 11)
    sf = _analyze_sparse(K, 0)
    sym = oracle.Symbolic.from_product(sf)
    signs = np.r_[-np.ones(40), np.ones(20)].astype(np.int32)
    f, delta, retries, err = oracle.static_regularized(sym, K, signs)
    assert err is None and f.inertia() == (20, 40, 0) and retries <= 1
    Kl = K.tolil()
    Kl[0, 0] = 0.05
    Kn = Kl.tocsc()
    f, delta, retries, err = oracle.static_regularized(sym, Kn, signs, initial_delta=1e-6)
    assert err is None and f.inertia() == (20, 40, 0) and retries >= 1
    f, delta, retries, err = oracle.static_regularized(sym, Kn, signs, initial_delta=1e-6, max_retries=0)
    assert f is None and "exhausted" in err


def test_refinement_recovers_regularization(oracle):
    """kkt_static_pivoting_test.cpp:247-293: refinement improves the residual >= 100x, final < 1e-8."""
    K = _kkt(40, 20, 3034)
    sf = _analyze_sparse(K, 0)
    sym = oracle.Symbolic.from_product(sf)
    signs = np.r_[-np.ones(40), np.ones(20)].astype(np.int32)
    f, *_ = oracle.static_regularized(sym, K, signs, initial_delta=0.005, rel_floor=100.0)
    xt = np.random.default_rng(3589).standard_normal(60)
    b = P.sym_matvec(K, xt)
    x0 = f.solve(b)
    raw = P.backward_error(K, x0, b)
    x, rr = f.refine(K, b, x0)
    fin = P.backward_error(K, x, b)
    assert raw > 1e-10 and fin <= raw * 1e-2 and fin < 1e-8
    assert rr.iterations >= 1 and rr.final_relative_residual < rr.initial_relative_residual
