"""The product's CPU symbolic analysis (ma57_amd/csrc/symbolic.cpp) against
the Python restatement of the reference's supernode construction
(oracle/symbolic_ref.py) and a brute-force symbolic elimination.

Orderings: the reference's AMD and METIS are third-party libraries that are
absent here; the product has its own minimum degree and nested dissection.
Given the SAME first-stage permutation (OrderingType Natural, or the product's
own order handed back as a user permutation), the supernode partition must
equal the restatement's exactly.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from ma57_amd import problems as P
from ma57_amd.solver import analyze, SymbolicFactorOptions
from oracle import symbolic_ref as R


def _arrays(A):
    A = sp.csc_matrix(A)
    return A.shape[0], A.indptr.astype(np.int64), A.indices.astype(np.int32)


def _product(A, order=2, relax=(64, 0.25), user_perm=None):
    h, sf = analyze(sp.csc_matrix(A), order, SymbolicFactorOptions(*relax), user_perm=user_perm)
    return sf


def _same(sf, ref):
    perm, parent, sns = ref
    assert list(sf.perm) == list(perm)
    assert list(sf.etree_parent) == list(parent)
    assert [(s.firstCol, s.ncols, list(s.rowPattern)) for s in sf.supernodes] == [tuple(x) for x in sns]


def dense_colcounts(A_perm):
    """test/unit/test_helpers.hpp:86-125: literal symbolic Gaussian elimination."""
    n = A_perm.shape[0]
    nz = (A_perm != 0) | (A_perm.T != 0) | np.eye(n, dtype=bool)
    for j in range(n):
        below = np.nonzero(nz[j + 1:, j])[0] + j + 1
        nz[np.ix_(below, below)] = True
    return [int(nz[j:, j].sum()) for j in range(n)]


@pytest.mark.parametrize("n,seed", [(8, 1), (30, 2), (60, 3), (120, 4)])
@pytest.mark.parametrize("relax", [(64, 0.25), (1, 0.0), (8, 0.5)])
def test_natural_order_matches_restatement(n, seed, relax):
    A = P.random_sparse_spd(n, 3, seed)
    sf = _product(A, 2, relax)
    nn, cp, ri = _arrays(A)
    _same(sf, R.analyze(nn, cp, ri, list(range(n)), *relax))


@pytest.mark.parametrize("order", [0, 1])
def test_product_orders_fed_back_match_restatement(order):
    """Same first-stage permutation -> same postorder, etree, supernodes and patterns."""
    A = P.laplace2d(9, 11, 0.3)
    sf0 = _product(A, order)
    nn, cp, ri = _arrays(A)
    # the product's final perm is itself a valid first-stage order; feed it to both
    sf = _product(A, 3, user_perm=sf0.perm)
    _same(sf, R.analyze(nn, cp, ri, list(sf0.perm)))


def test_arrow_supernodes_reference_case():
    """multifrontal_test.cpp:128-139: Natural order, no amalgamation -> {0}, {1}, {2,3,4}."""
    A = np.zeros((5, 5))
    A[2, 2], A[3, 3], A[4, 4] = 5.0, 7.0, 20.0
    A[2, 3] = A[3, 2] = 1.1
    A[3, 4] = A[4, 3] = 1.3
    A[2, 4] = A[4, 2] = 0.9
    A[0, 4] = A[4, 0] = 2.0
    A[1, 3] = A[3, 1] = 1.7
    sf = _product(A, 2, (1, 0.0))
    assert [(s.firstCol, s.ncols) for s in sf.supernodes] == [(0, 1), (1, 1), (2, 3)]


@pytest.mark.parametrize("order", [0, 1, 2])
def test_column_counts_and_structure(order):
    A = P.random_sparse_indefinite(70, 4, 9)
    sf = _product(A, order, (1, 0.0))
    n = A.shape[0]
    perm = np.asarray(sf.perm)
    assert sorted(perm) == list(range(n))
    Ad = A.toarray()
    cc = dense_colcounts(Ad[np.ix_(perm, perm)])
    assert sf.nnz_l == sum(cc)
    # supernode invariants: own columns contiguous and leading the sorted pattern,
    # index order a postorder (parent index > child index)
    for s in sf.supernodes:
        pat = list(s.rowPattern)
        assert pat == sorted(pat)
        assert pat[:s.ncols] == list(range(s.firstCol, s.firstCol + s.ncols))
    parent = np.asarray(sf.etree_parent)
    assert all(parent[j] == -1 or parent[j] > j for j in range(n))


def test_orderings_reduce_fill_on_grids():
    """AMD-like and ND orders beat the natural order on a 2D grid (ordering_test.cpp:85-132 in spirit)."""
    A = P.laplace2d(30, 30, 0.0)
    nat = _product(A, 2).nnz_l
    md = _product(A, 0).nnz_l
    nd = _product(A, 1).nnz_l
    assert md < 0.7 * nat and nd < 0.7 * nat


def test_invalid_inputs_raise():
    from ma57_amd.solver import InvalidArgument
    A = sp.csc_matrix(np.eye(3))
    with pytest.raises(InvalidArgument):
        analyze(A, 7)
    with pytest.raises(InvalidArgument):
        analyze(A, 3, user_perm=[0, 0, 1])


@pytest.mark.parametrize("case", ["stencil27_7", "laplace3d_9", "batched", "random400"])
@pytest.mark.parametrize("relax", [(64, 0.25), (16, 0.1), (128, 0.6)])
def test_larger_partitions_match_restatement(case, relax):
    """Bushier trees (3D stencils, random couplings) with the product's own
    orders fed back: the amalgamation decisions and the relabeling must equal
    the restatement's on many merges, not only on small examples."""
    if case == "stencil27_7":
        A, order = P.stencil27(7, 0.2), 1
    elif case == "laplace3d_9":
        A, order = P.laplace3d(9, 9, 9, 0.5), 1
    elif case == "batched":
        A, order = P.batched_system(1234), 0
    else:
        A, order = P.random_sparse_indefinite(400, 5, 11), 0
    sf0 = _product(A, order, relax)
    nn, cp, ri = _arrays(A)
    sf = _product(A, 3, relax, user_perm=sf0.perm)
    ref = R.analyze(nn, cp, ri, list(sf0.perm), *relax)
    _same(sf, ref)
    # the case really amalgamates: fewer supernodes than the fundamental ones
    assert len(ref[2]) < len(R.analyze(nn, cp, ri, list(sf0.perm), 1, 0.0)[2])
