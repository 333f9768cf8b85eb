"""Matrix Market reader (symla_read_matrix_market; the reference reads its
real-matrix inputs with Eigen::loadMarket, test/correctness/
real_matrix_test.cpp:84) against scipy.io's independent reader/writer."""
import io

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

from ma57_amd import problems as P
from ma57_amd.mmio import read_matrix_market


def _same(A, B):
    A, B = sp.csc_matrix(A), sp.csc_matrix(B)
    A.sum_duplicates()
    B.sum_duplicates()
    A.sort_indices()
    B.sort_indices()
    return A.shape == B.shape and np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices) \
        and np.array_equal(A.data, B.data)


@pytest.mark.parametrize("symmetry", ["general", "symmetric"])
def test_roundtrip_against_scipy(tmp_path, symmetry):
    A = P.laplace2d(13, 11, 0.3).tocsc()  # lower triangle stored
    full = (A + sp.tril(A, -1).T).tocsc()
    path = tmp_path / f"m_{symmetry}.mtx"
    sio.mmwrite(str(path), full, symmetry=symmetry, precision=17)
    got, sym = read_matrix_market(path)
    assert sym == symmetry
    # a symmetric file stores the lower triangle, exactly what scipy writes and reads back as stored
    expect = A if symmetry == "symmetric" else full
    assert _same(got, expect)
    assert _same(got, sp.tril(sio.mmread(str(path))) if symmetry == "symmetric" else sio.mmread(str(path)))


def test_pattern_integer_duplicates_and_comments(tmp_path):
    path = tmp_path / "d.mtx"
    path.write_text("%%MatrixMarket matrix coordinate integer general\n% comment\n%\n3 4 5\n1 1 2\n3 2 -7\n"
                    "1 1 3\n2 4 1\n3 2 1\n")
    A, sym = read_matrix_market(path)
    assert sym == "general" and A.shape == (3, 4)
    dense = np.zeros((3, 4))
    dense[0, 0], dense[2, 1], dense[1, 3] = 5, -6, 1
    assert np.array_equal(A.toarray(), dense)
    path2 = tmp_path / "p.mtx"
    path2.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n3 3 3\n1 1\n2 1\n3 3\n")
    B, sym = read_matrix_market(path2)
    assert sym == "symmetric" and np.array_equal(B.toarray(), [[1, 0, 0], [1, 0, 0], [0, 0, 1]])


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",
    "%%MatrixMarket matrix array real general\n1 1\n1\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
    "not a matrix market file\n",
])
def test_rejects_unsupported_or_malformed(tmp_path, text):
    path = tmp_path / "bad.mtx"
    path.write_text(text)
    with pytest.raises(ValueError):
        read_matrix_market(path)
