"""Matrix Market input through the library's reader (symla_read_matrix_market).

The reference loads its real-matrix regression inputs with Eigen::loadMarket
(test/correctness/real_matrix_test.cpp:84); this is the same role for the
Python tests and benches: entries as stored (a symmetric file's one
triangle), duplicates summed.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

SYMMETRY = {0: "general", 1: "symmetric", 2: "skew-symmetric"}


def read_matrix_market(path):
    """(scipy.sparse.csc_matrix, symmetry name) of a coordinate Matrix Market file."""
    import scipy.sparse as sp

    lib = _lib.load()
    p = str(path).encode()
    nr, nc, sym = C.c_int(), C.c_int(), C.c_int()
    nnz = C.c_int64()
    rc = lib.symla_read_matrix_market(p, C.byref(nr), C.byref(nc), C.byref(nnz), C.byref(sym), None, None, None)
    if rc:
        _lib.raise_for(rc, lib.symla_last_error().decode())
    colptr = np.zeros(nc.value + 1, np.int64)
    rowidx = np.zeros(max(nnz.value, 1), np.int32)
    vals = np.zeros(max(nnz.value, 1), np.float64)
    rc = lib.symla_read_matrix_market(p, None, None, None, None, colptr.ctypes.data_as(_lib.i64p),
                                      rowidx.ctypes.data_as(_lib.i32p), vals.ctypes.data_as(_lib.f64p))
    if rc:
        _lib.raise_for(rc, lib.symla_last_error().decode())
    A = sp.csc_matrix((vals[:nnz.value], rowidx[:nnz.value], colptr), shape=(nr.value, nc.value))
    return A, SYMMETRY[sym.value]
