"""Synthetic inputs of the benchmark configurations and of the tests.

All matrices are returned as scipy CSC holding the LOWER triangle (row >= col),
the reference's authoritative triangle (multifrontal.hpp:190-198).  Seeds are
fixed; "N(0, s)" means standard deviation s.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

SEED_BASE = 0x53594D4C41  # "SYMLA"


def _lower(rows, cols, vals, n):
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    vals = np.asarray(vals, np.float64)
    lo = rows >= cols
    A = sp.coo_matrix((vals[lo], (rows[lo], cols[lo])), shape=(n, n)).tocsc()
    A.sum_duplicates()
    A.sort_indices()
    return A


def laplace2d(nx: int, ny: int, shift: float = 0.0):
    """5-point Dirichlet Laplacian on an nx x ny grid (diag 4, off -1) minus shift*I."""
    n = nx * ny
    idx = np.arange(n).reshape(ny, nx)
    r = [idx.ravel()]
    c = [idx.ravel()]
    v = [np.full(n, 4.0 - shift)]
    # neighbours (lower triangle: larger index first)
    r.append(idx[:, 1:].ravel()); c.append(idx[:, :-1].ravel()); v.append(np.full(ny * (nx - 1), -1.0))
    r.append(idx[1:, :].ravel()); c.append(idx[:-1, :].ravel()); v.append(np.full((ny - 1) * nx, -1.0))
    return _lower(np.concatenate(r), np.concatenate(c), np.concatenate(v), n)


def laplace3d(nx: int, ny: int, nz: int, shift: float = 0.0):
    """7-point Dirichlet Laplacian on nx x ny x nz (diag 6, off -1) minus shift*I."""
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    r = [idx.ravel()]
    c = [idx.ravel()]
    v = [np.full(n, 6.0 - shift)]
    for a, b in ((idx[:, :, 1:], idx[:, :, :-1]), (idx[:, 1:, :], idx[:, :-1, :]), (idx[1:, :, :], idx[:-1, :, :])):
        r.append(a.ravel()); c.append(b.ravel()); v.append(np.full(a.size, -1.0))
    return _lower(np.concatenate(r), np.concatenate(c), np.concatenate(v), n)


def laplace_eigs_2d(nx, ny):
    i = np.arange(1, nx + 1)
    j = np.arange(1, ny + 1)
    return (4 - 2 * np.cos(i * np.pi / (nx + 1)))[None, :] - 2 * np.cos(j * np.pi / (ny + 1))[:, None]


def closed_form_inertia(eigs: np.ndarray, shift: float):
    """Inertia of (L - shift I) from the closed-form spectrum of L."""
    e = eigs.ravel() - shift
    return int((e > 0).sum()), int((e < 0).sum()), int((e == 0).sum()), float(np.abs(e).min())


def laplace_eigs_3d(nx, ny, nz):
    f = lambda k: 2 - 2 * np.cos(np.arange(1, k + 1) * np.pi / (k + 1))
    return f(nx)[None, None, :] + f(ny)[None, :, None] + f(nz)[:, None, None]


def sqd_kkt(nh: int = 30, m2: int = 9000, per_row: int = 5, seed: int = SEED_BASE + 3, eps_h: float = 1e-4,
            eps_f: float = 1e-8):
    """K = [[-(H + eps_h I), J^T], [J, eps_f I]]; H = 7-point Laplacian on nh^3; J is m2 x nh^3 with
    `per_row` N(0,1) entries per row at uniform columns (duplicates summed).  Returns (K_lower, n1, n2)."""
    H = laplace3d(nh, nh, nh, 0.0)
    n1 = H.shape[0]
    rng = np.random.default_rng(seed)
    jr = np.repeat(np.arange(m2), per_row)
    jc = rng.integers(0, n1, size=m2 * per_row)
    jv = rng.standard_normal(m2 * per_row)
    Hc = H.tocoo()
    rows = np.concatenate([Hc.row, np.arange(n1), n1 + jr, n1 + np.arange(m2)])
    cols = np.concatenate([Hc.col, np.arange(n1), jc, n1 + np.arange(m2)])
    vals = np.concatenate([-Hc.data, np.full(n1, -eps_h), jv, np.full(m2, eps_f)])
    return _lower(rows, cols, vals, n1 + m2), n1, m2


def stencil27(N: int, sigma: float, seed: int = SEED_BASE + 5):
    """27-point stencil on N^3: a_ij = -w_ij with w ~ U(0.5, 1.5) per undirected edge, a_ii = sum of w over
    all 26 positions (out-of-domain ones drawn too), minus sigma*I."""
    rng = np.random.default_rng(seed)
    n = N ** 3
    idx = np.arange(n).reshape(N, N, N)
    diag = np.zeros(n)
    rows, cols, vals = [], [], []
    offsets = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
               if (dz, dy, dx) > (0, 0, 0)]
    for dz, dy, dx in offsets:
        # edge between p and p + d for every p whose neighbour lies in the padded domain
        zs = slice(max(0, -dz), N - max(0, dz))
        ys = slice(max(0, -dy), N - max(0, dy))
        xs = slice(max(0, -dx), N - max(0, dx))
        a = idx[zs, ys, xs]
        b = idx[(slice(zs.start + dz, zs.stop + dz)), slice(ys.start + dy, ys.stop + dy),
                slice(xs.start + dx, xs.stop + dx)]
        w = rng.uniform(0.5, 1.5, size=a.size)
        rows.append(b.ravel()); cols.append(a.ravel()); vals.append(-w)
        np.add.at(diag, a.ravel(), w)
        np.add.at(diag, b.ravel(), w)
    # out-of-domain neighbours (Dirichlet): each vertex has 26 - deg of them
    deg = np.zeros(n)
    for r_ in rows:
        np.add.at(deg, r_, 1)
    for c_ in cols:
        np.add.at(deg, c_, 1)
    missing = (26 - deg).astype(np.int64)
    diag += _sum_uniform(rng, missing)
    rows.append(np.arange(n)); cols.append(np.arange(n)); vals.append(diag - sigma)
    return _lower(np.concatenate(rows), np.concatenate(cols), np.concatenate(vals), n)


def _sum_uniform(rng, counts):
    total = int(counts.sum())
    draws = rng.uniform(0.5, 1.5, size=total)
    cs = np.concatenate([[0], np.cumsum(counts)])
    return np.add.reduceat(np.concatenate([draws, [0.0]]), cs[:-1]) * (counts > 0)


def random_sparse_spd(n: int, avg_nnz: int, seed: int):
    """Diagonally dominant random SPD matrix (the shape of the reference's randomSparseSPD,
    test/unit/test_helpers.hpp:38-77): 1 + rng%avg off-diagonal draws per row, values U(0.1, 1),
    diagonal = row |sum| + 1 + 0.1*(i%5).  Full symmetric storage."""
    rng = np.random.default_rng(seed)
    r, c, v = [], [], []
    for i in range(n):
        k = 1 + int(rng.integers(0, avg_nnz)) if avg_nnz > 0 else 0
        for _ in range(k):
            j = int(rng.integers(0, n))
            if j <= i:
                continue
            val = rng.uniform(0.1, 1.0)
            r += [i, j]; c += [j, i]; v += [val, val]
    A = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    rs = np.asarray(abs(A).sum(axis=1)).ravel()
    d = rs + 1.0 + 0.1 * (np.arange(n) % 5)
    A = (A + sp.diags(d)).tocsc()
    A.sort_indices()
    return A


def random_sparse_indefinite(n: int, avg_nnz: int, seed: int):
    """random_sparse_spd with a random half of the diagonal negated (test_helpers.hpp:137-150)."""
    A = random_sparse_spd(n, avg_nnz, seed).tolil()
    rng = np.random.default_rng(seed ^ 0x9E3779B9)
    flip = rng.integers(0, 2, size=n) == 0
    for i in np.nonzero(flip)[0]:
        A[i, i] = -A[i, i]
    A = A.tocsc()
    A.sort_indices()
    return A


def batched_system(seed: int):
    """One system of the batched configuration: n ~ U{200..5000}, grid nx ~ U{ceil(sqrt(n/4))..floor(sqrt(4n))},
    ny = ceil(n/nx), 5-point Laplacian plus 2 symmetric entries per row at random columns with N(0, 0.1)
    values, minus 0.3*I."""
    rng = np.random.default_rng(seed)
    n_req = int(rng.integers(200, 5001))
    lo = int(np.ceil(np.sqrt(n_req / 4)))
    hi = int(np.floor(np.sqrt(4 * n_req)))
    nx = int(rng.integers(lo, hi + 1))
    ny = int(np.ceil(n_req / nx))
    L = laplace2d(nx, ny, 0.3).tocoo()
    n = nx * ny
    rr = np.repeat(np.arange(n), 2)
    cc = rng.integers(0, n - 1, size=2 * n)
    cc = cc + (cc >= rr)
    vv = rng.normal(0.0, 0.1, size=2 * n)
    hi_ = np.maximum(rr, cc)
    lo_ = np.minimum(rr, cc)
    rows = np.concatenate([L.row, hi_])
    cols = np.concatenate([L.col, lo_])
    vals = np.concatenate([L.data, vv])
    return _lower(rows, cols, vals, n)


def sym_dense(A_lower) -> np.ndarray:
    """Dense symmetric matrix from a lower-stored CSC (lower triangle authoritative)."""
    A = A_lower.tocoo()
    D = np.zeros(A.shape)
    lo = A.row >= A.col
    np.add.at(D, (A.row[lo], A.col[lo]), A.data[lo])
    off = lo & (A.row != A.col)
    np.add.at(D, (A.col[off], A.row[off]), A.data[off])
    return D


def sym_matvec(A_lower, X):
    """A_sym @ X with A given by its lower triangle."""
    A = A_lower.tocsc()
    L = sp.tril(A, format="csc")
    D = sp.diags(L.diagonal())
    S = L + L.T - D
    return S @ X


def backward_error(A_lower, X, B) -> float:
    """max over columns of ||A x - b||inf / (||A||inf ||x||inf + ||b||inf) with the symmetric ||A||inf."""
    X = np.asarray(X).reshape(A_lower.shape[0], -1)
    B = np.asarray(B).reshape(A_lower.shape[0], -1)
    A = A_lower.tocsc()
    L = sp.tril(A, format="csc")
    S = L + L.T - sp.diags(L.diagonal())
    normA = np.abs(S).sum(axis=1).max()
    R = B - S @ X
    out = 0.0
    for c in range(X.shape[1]):
        den = normA * np.abs(X[:, c]).max() + np.abs(B[:, c]).max()
        out = max(out, np.abs(R[:, c]).max() / den if den > 0 else np.abs(R[:, c]).max())
    return float(out)
