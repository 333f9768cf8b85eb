"""Writes reference_cases.json: the known-answer cases of the reference's own
test suite for the numeric path, transcribed as data (inputs and the
expectations the reference asserts), with the file:line each comes from.

The reference is header-only C++ on Eigen and SuiteSparse AMD, neither of
which is installed here, so its tests cannot be run; these values are the
ones its test sources state.  Eigenvalue-derived inertias (computed here with
numpy.linalg.eigvalsh, Sylvester's law of inertia) are added where the
reference only asserts a weaker property.
"""
import json
from pathlib import Path

import numpy as np


def inertia_of(A):
    e = np.linalg.eigvalsh(np.array(A, dtype=float))
    return [int((e > 0).sum()), int((e < 0).sum()), int((e == 0).sum())]


def main():
    dense = [
        {"name": "diag3", "src": "test/unit/dense_kernel_test.cpp:207-227",
         "A": [[10, 1, 1], [1, 8, 1], [1, 1, 6]], "opts": {},
         "expect": {"n_factored": 3, "pivot_kinds": [1, 1, 1], "perm": [0, 1, 2], "inertia": [3, 0, 0],
                    "recon_tol": 1e-12}},
        {"name": "forces2x2", "src": "test/unit/dense_kernel_test.cpp:229-255",
         "A": [[0, 1, 0.2], [1, 0, 0.3], [0.2, 0.3, 5.0]], "opts": {},
         "expect": {"n_factored": 3, "first_pivot": [0, 2], "inertia": [2, 1, 0], "recon_tol": 1e-10}},
        {"name": "mixed5", "src": "test/unit/dense_kernel_test.cpp:257-278",
         "A": [[20, 1, 0.5, 0.1, 0.1], [1, 15, 0.2, 0.1, 0.1], [0.5, 0.2, 0, 2, 0.1], [0.1, 0.1, 2, 0, 0.2],
               [0.1, 0.1, 0.1, 0.2, 12]], "opts": {},
         "expect": {"n_factored": 5, "n_zero": 0, "recon_tol": 1e-10}},
        {"name": "delayed_zero_trailing", "src": "test/unit/dense_kernel_test.cpp:335-362",
         "A": [[10, 1, 0.5, 0, 0], [1, 8, 0.2, 0, 0], [0.5, 0.2, 6, 0, 0], [0, 0, 0, 0, 0], [0, 0, 0, 0, 0]],
         "opts": {}, "expect": {"n_factored": 3, "delayed": [3, 4], "recon_tol": 1e-10}},
        {"name": "rankdef2x2", "src": "test/unit/dense_kernel_test.cpp:364-387",
         "A": [[100, 0, 0], [0, 0, 1e-20], [0, 1e-20, 2e-20]], "opts": {"zero_tolerance": 1e-8},
         "expect": {"n_factored": 1, "delayed": [1, 2]}},
    ]
    for c in dense:
        if c["name"] == "mixed5":
            c["expect"]["inertia"] = inertia_of(c["A"])
    arrow = np.zeros((5, 5))
    arrow[2, 2], arrow[3, 3], arrow[4, 4] = 5.0, 7.0, 20.0
    arrow[2, 3] = arrow[3, 2] = 1.1
    arrow[3, 4] = arrow[4, 3] = 1.3
    arrow[2, 4] = arrow[4, 2] = 0.9
    arrow[0, 4] = arrow[4, 0] = 2.0
    arrow[1, 3] = arrow[3, 1] = 1.7
    sing = np.zeros((3, 3))
    sing[0, 0], sing[2, 2] = 4.0, 3.0
    sing[0, 2] = sing[2, 0] = 0.5
    multifrontal = [
        {"name": "arrow", "src": "test/unit/multifrontal_test.cpp:102-164", "A": arrow.tolist(),
         "ordering": "natural", "relax": [1, 0.0], "b": [1, 2, 3, 4, 5],
         "expect": {"supernodes": [[0, 1], [1, 1], [2, 3]], "npivots": [0, 0, 5], "singular": False,
                    "inertia": inertia_of(arrow), "resid_tol": 1e-10}},
        {"name": "singular3", "src": "test/unit/multifrontal_test.cpp:171-196", "A": sing.tolist(),
         "ordering": "amd", "expect": {"singular": True, "singular_cols": [1], "total_pivots": 2}},
        {"name": "smoke_diag", "src": "test/unit/smoke_test.cpp:24-55", "A": [[2, 0], [0, 3]], "ordering": "amd",
         "b": [4, 9], "expect": {"x": [2, 3], "inertia": [2, 0, 0], "x_tol": 1e-12}},
        {"name": "spd4", "src": "test/unit/multifrontal_test.cpp:58-62",
         "A": [[10, 1, 0, 1], [1, 8, 1, 0], [0, 1, 6, 1], [1, 0, 1, 9]], "ordering": "amd",
         "expect": {"inertia": [4, 0, 0], "singular": False}},
        {"name": "indef5", "src": "test/unit/multifrontal_test.cpp:64-77",
         "A": [[10, 1, 0, 0, 1], [1, -8, 1, 0, 0], [0, 1, 6, 1, 0], [0, 0, 1, -9, 1], [1, 0, 0, 1, 7]],
         "ordering": "amd", "expect": {"singular": False}},
    ]
    for c in multifrontal:
        if c["name"] in ("indef5",):
            c["expect"]["inertia"] = inertia_of(c["A"])
    out = {"about": __doc__.strip().splitlines()[0], "dense": dense, "multifrontal": multifrontal}
    Path(__file__).with_name("reference_cases.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
