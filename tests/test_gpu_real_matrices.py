"""Real-matrix regression in the manner of the reference's
test/correctness/real_matrix_test.cpp:62-190: Matrix Market inputs read by
the library's reader, analyzePattern + factorize (AMD ordering) + solve with
b = A x_true, inertia self-consistency n_pos + n_neg + n_zero = n - #singular
and the reference's residual ||A x - b||2 / (max|A| max(1, ||x||2) + ||b||2).

The SuiteSparse matrices the reference curates are fetched from the network
there (bench/fetch_matrices.py), which this build cannot do: they are used
when present under $SYMLA_MATRICES/<group>/<name>/<name>.mtx and skipped
otherwise.  Three generated indefinite matrices of the same families (a
2D Stokes-like saddle point, a KKT system, a shifted Laplacian) are always
written to .mtx files and run through the same path."""
import os
from pathlib import Path

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

from ma57_amd import problems as P
from ma57_amd.mmio import read_matrix_market
from ma57_amd.solver import SymLDLT, OrderingType

pytestmark = pytest.mark.gpu

CURATED = [("GHS_indef", m) for m in ("sit100", "tuma2", "ncvxqp1", "tuma1", "qpband", "bratu3d", "c-55",
                                       "aug3dcqp", "stokes128", "dawson5", "blockqp1", "cont-201", "turon_m",
                                       "helm2d03")] + \
          [("Schenk_IBMNA", m) for m in ("c-18", "c-22", "c-26", "c-30", "c-62", "c-67")] + \
          [("HB", "bcsstk14"), ("HB", "bcsstk16")]


def _residual(A_lower, X, B):
    """real_matrix_test.cpp:51-59 (2-norms, max |A| over the stored entries)."""
    L = sp.tril(A_lower, format="csc")
    S = L + L.T - sp.diags(L.diagonal())
    amax = np.abs(A_lower.data).max() if A_lower.nnz else 1.0
    return np.linalg.norm(S @ X - B) / (amax * max(1.0, np.linalg.norm(X)) + np.linalg.norm(B))


def _stokes2d(N, seed=3):
    """[[K, D^T], [D, 0]] with K the 5-point vector Laplacian on N x N (two velocity components) and D a
    forward-difference divergence to an (N-1)^2 pressure grid: a saddle point like GHS_indef/stokes*."""
    K1 = P.laplace2d(N, N, 0.0)
    K1 = (K1 + sp.tril(K1, -1).T).tocsr()
    K = sp.block_diag([K1, K1])
    idx = np.arange(N * N).reshape(N, N)
    rows, cols, vals = [], [], []
    for i in range(N - 1):
        for j in range(N - 1):
            p = i * (N - 1) + j
            for (a, b, s) in ((idx[i + 1, j], 0, 1.0), (idx[i, j], 0, -1.0), (idx[i, j + 1], 1, 1.0),
                              (idx[i, j], 1, -1.0)):
                rows.append(p)
                cols.append(a + b * N * N)
                vals.append(s)
    D = sp.csr_matrix((vals, (rows, cols)), shape=((N - 1) ** 2, 2 * N * N))
    M = sp.bmat([[K, D.T], [D, None]], format="csc")
    return sp.tril(M, format="csc")


def _cases(tmp_path):
    out = []
    gen = {"stokes2d_24": _stokes2d(24), "kkt_threshold": P.sqd_kkt(nh=12, m2=600, seed=5)[0],
           "laplace2d_60": P.laplace2d(60, 60, 0.3)}
    for name, A in gen.items():
        path = tmp_path / f"{name}.mtx"
        sio.mmwrite(str(path), A, symmetry="general", precision=17)
        out.append((name, path, True))
    root = os.environ.get("SYMLA_MATRICES")
    if root:
        for group, name in CURATED:
            path = Path(root) / group / name / f"{name}.mtx"
            if path.exists():
                out.append((f"{group}/{name}", path, False))
    return out


def test_matrix_market_inputs(tmp_path):
    tested = 0
    for name, path, generated in _cases(tmp_path):
        A, _ = read_matrix_market(path)
        if A.shape[0] != A.shape[1] or A.shape[0] > 25000:  # the reference's ctest size gate (:112)
            continue
        A = sp.tril(A, format="csc")
        n = A.shape[0]
        s = SymLDLT()
        s.setOrdering(OrderingType.AMD)
        s.compute(A)
        inr = s.inertia()
        nsing = len(s.singularColumns())
        assert inr.n_pos + inr.n_neg + inr.n_zero == n - nsing, name
        if s.isSingular():
            assert not generated, name
            continue
        xtrue = np.random.default_rng(1234567 + n).standard_normal(n)
        L = sp.tril(A, format="csc")
        b = (L + L.T - sp.diags(L.diagonal())) @ xtrue
        x = s.solve(b)
        r = _residual(A, x.reshape(n, -1), b.reshape(n, -1))
        # generated matrices: the reference's KKT residual bar (kkt_test.cpp:134; the threshold-mode SQD
        # system reaches ~6e-9 unrefined, as the CPU oracle does); curated ones: real_matrix_test.cpp:189
        assert r < (1e-8 if generated else 0.05), (name, r)
        tested += 1
    assert tested >= 3
