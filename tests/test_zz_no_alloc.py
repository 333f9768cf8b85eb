"""The allocation-free solve entry point (symla_solve_device_noalloc, no
counterpart in the reference) against symla_solve_device: the same kernels
on the same factorization, so the solutions agree bit for bit.  Two calls
in a row share the library's static scratch, the second on a smaller
system, after the first left its values there."""
import ctypes as C

import numpy as np
import pytest

from ma57_amd import _lib, problems as P
from ma57_amd.solver import SymLDLT, OrderingType, Precision

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _factored(A, ordering, precision):
    s = SymLDLT()
    s.setOrdering(ordering)
    s.setPrecision(precision)
    s.analyzePattern(A)
    s.factorize(A)
    return s


def _out(n, nrhs):
    return torch.full((nrhs, n), float("nan"), dtype=torch.float64, device="cuda:0")


def _call(lib, plan, fn, dB, n, nrhs, dX=None):
    dX = _out(n, nrhs) if dX is None else dX
    torch.cuda.synchronize()
    rc = fn(plan, C.c_void_p(dB.data_ptr()), n, nrhs, C.c_void_p(dX.data_ptr()), n)
    torch.cuda.synchronize()
    return rc, dX


def test_noalloc_solve_equals_device_solve_twice():
    lib = _lib.load()
    # the runtime loads the solve kernels' code object on their first launch: do one solve elsewhere first
    w = _factored(P.laplace2d(5, 5, 0.3), OrderingType.AMD, Precision.FP64)
    w.solve(np.ones(25))
    # the second call is on a smaller system (fewer unknowns and right-hand sides)
    runs = [(P.laplace3d(18, 18, 18, 0.5), OrderingType.Metis, Precision.FP64, 12),
            (P.laplace2d(30, 30, 0.3), OrderingType.AMD, Precision.FP32, 3)]
    for A, ordering, precision, nrhs in runs:
        n = A.shape[0]
        s = _factored(A, ordering, precision)
        B = np.random.default_rng(n + nrhs).standard_normal((n, nrhs))
        dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda:0")
        dXn = _out(n, nrhs)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc, dXn = _call(lib, s._plan, lib.symla_solve_device_noalloc, dB, n, nrhs, dXn)
        assert rc == 0, lib.symla_plan_last_error(s._plan).decode()
        free1 = torch.cuda.mem_get_info()[0]
        rc, dXr = _call(lib, s._plan, lib.symla_solve_device, dB, n, nrhs)
        assert rc == 0, lib.symla_plan_last_error(s._plan).decode()
        Xn, Xr = dXn.cpu().numpy().T, dXr.cpu().numpy().T
        assert np.array_equal(Xn, Xr)
        assert P.backward_error(A, Xn, B) <= (1e-12 if precision == Precision.FP64 else 1e-4)
        # the no-alloc call on a plan that never solved before took no device memory (the static scratch
        # was set up with the plan; symla_solve_device would have grown the plan's workspace here)
        assert free1 == free0


def test_noalloc_rejects_a_workspace_above_the_scratch():
    lib = _lib.load()
    A = P.laplace2d(100, 100, 0.3)
    n = A.shape[0]
    s = _factored(A, OrderingType.AMD, Precision.FP64)
    nrhs = 700  # 3 n nrhs doubles alone exceed 128 MiB
    dB = torch.ones((nrhs, n), dtype=torch.float64, device="cuda:0")
    rc, _ = _call(lib, s._plan, lib.symla_solve_device_noalloc, dB, n, nrhs)
    assert rc == _lib.SYMLA_ERR_INVALID
    assert "MiB" in lib.symla_plan_last_error(s._plan).decode()
    # a call within the limit still works afterwards
    rc, dX = _call(lib, s._plan, lib.symla_solve_device_noalloc, dB[:4].contiguous(), n, 4)
    assert rc == 0
    X = dX.cpu().numpy().T
    assert P.backward_error(A, X, np.ones((n, 4))) <= 1e-12
