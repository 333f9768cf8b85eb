"""Recorded refactorize + solve (one hipGraph on the plan's stream) against
the eager path, through the C-ABI.

Each test replays the recorded graph several times with new matrix values
and right-hand sides written into the same device buffers, with eager device
work and host synchronizations between the replays, and checks every replay
against an eager factorize + solve of the same data on another plan (bitwise:
same kernels, same order) and against the closed-form inertia.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from ma57_amd import problems as P
from ma57_amd.solver import SymLDLT, OrderingType, Precision, SymbolicFactorOptions, csc_arrays

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _solver(A, ordering, precision, opts=None):
    s = SymLDLT()
    s.setOrdering(ordering)
    s.setPrecision(precision)
    if opts is None:
        s.analyzePattern(A)
    else:
        s.analyzePattern(A, opts)
    return s


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32])
def test_graph_replays_match_eager(precision):
    G = 16
    shifts = [0.5, 0.37, 0.61, 0.45]
    mats = [P.laplace3d(G, G, G, sh) for sh in shifts]
    n = mats[0].shape[0]
    nrhs = 3
    dev = torch.device("cuda:0")
    s = _solver(mats[0], OrderingType.Metis, precision)
    eager = _solver(mats[0], OrderingType.Metis, precision)
    d_vals = torch.from_numpy(csc_arrays(mats[0])[3].copy()).to(dev)
    d_B = torch.zeros((nrhs, n), dtype=torch.float64, device=dev)
    d_X = torch.zeros_like(d_B)
    torch.cuda.synchronize()
    s.captureGraph(d_vals.data_ptr(), d_B.data_ptr(), nrhs, d_X.data_ptr())
    rng = np.random.default_rng(11)
    for rep in range(2):
        for A, sh in zip(mats, shifts):
            B = rng.standard_normal((n, nrhs))
            d_vals.copy_(torch.from_numpy(csc_arrays(A)[3]))
            d_B.copy_(torch.from_numpy(np.ascontiguousarray(B.T)))
            torch.cuda.synchronize()
            info = s.replayGraph()
            X = d_X.cpu().numpy().T
            eager.factorize(A)
            Xe = eager.solve(B)
            expect = tuple(P.closed_form_inertia(P.laplace_eigs_3d(G, G, G), sh)[:3])
            assert (info.n_pos, info.n_neg, info.n_zero) == expect
            assert info.repaired == 0
            assert np.array_equal(X, Xe)
            if precision == Precision.FP64:
                assert P.backward_error(A, X, B) <= 1e-12
            # eager device work between replays
            junk = torch.randn(4096, device=dev, dtype=torch.float64)
            assert torch.isfinite(junk.sum()).item()
            torch.cuda.synchronize()


def test_graph_replay_with_delays_is_repaired():
    A = np.zeros((5, 5))
    A[2, 2], A[3, 3], A[4, 4] = 5.0, 7.0, 20.0
    A[2, 3] = A[3, 2] = 1.1
    A[3, 4] = A[4, 3] = 1.3
    A[2, 4] = A[4, 2] = 0.9
    A[0, 4] = A[4, 0] = 2.0
    A[1, 3] = A[3, 1] = 1.7
    S = sp.csc_matrix(np.tril(A))
    dev = torch.device("cuda:0")
    s = _solver(S, OrderingType.Natural, Precision.FP64, SymbolicFactorOptions(1, 0.0))
    d_vals = torch.from_numpy(S.data.copy()).to(dev)
    b = np.array([1, 2, 3, 4, 5.0])
    d_B = torch.from_numpy(b.copy()).to(dev).reshape(1, 5)
    d_X = torch.zeros_like(d_B)
    torch.cuda.synchronize()
    s.captureGraph(d_vals.data_ptr(), d_B.data_ptr(), 1, d_X.data_ptr())
    for _ in range(3):  # the second and third replays re-record after the repair moved the layout
        info = s.replayGraph()
        assert info.repaired == 1 and info.singular == 0
        assert info.n_pos + info.n_neg == 5 and info.n_zero == 0
        x = d_X.cpu().numpy().reshape(5)
        assert np.linalg.norm(A @ x - b) / (np.abs(A).max() * np.linalg.norm(x) + np.linalg.norm(b)) < 1e-10
        torch.cuda.synchronize()


def test_graph_rejects_static_mode():
    from ma57_amd.solver import Mode, InvalidArgument
    A = P.laplace2d(6, 6, 0.3)
    s = _solver(A, OrderingType.AMD, Precision.FP64)
    s.setMode(Mode.StaticRegularized)
    dev = torch.device("cuda:0")
    d_vals = torch.from_numpy(csc_arrays(A)[3].copy()).to(dev)
    d_B = torch.zeros((1, 36), dtype=torch.float64, device=dev)
    with pytest.raises(InvalidArgument):
        s.captureGraph(d_vals.data_ptr(), d_B.data_ptr(), 1, d_B.data_ptr())


def test_graph_replay_with_delays_on_large_fronts_interleaved_with_eager_work():
    """Threshold-mode SQD KKT (fronts well above 64 rows delay pivots): every
    replay is repaired.  Between replays the same plan runs an eager
    factorize + solve (which lays the plan out again) and a repaired replay
    reallocates per-front storage; the next replay must notice that the
    recorded graph refers to moved buffers and record it again rather than
    run on freed memory.  Inertia is Sylvester's (n2, n1, 0) for every
    positive scaling of the matrix.  Without refinement the reference's
    algorithm reaches backward error 1.4e-9 / 9.1e-10 on this matrix at scales
    1.0 / 0.7 (the CPU oracle, same ordering); the bar is the reference's own
    KKT residual bar, 1e-8 (test/correctness/kkt_test.cpp:134)."""
    K, n1, n2 = P.sqd_kkt(nh=16, m2=1200, seed=77)
    n = K.shape[0]
    base = csc_arrays(K)[3].copy()
    dev = torch.device("cuda:0")
    s = _solver(K, OrderingType.Metis, Precision.FP64)
    d_vals = torch.from_numpy(base.copy()).to(dev)
    d_B = torch.zeros((2, n), dtype=torch.float64, device=dev)
    d_X = torch.zeros_like(d_B)
    torch.cuda.synchronize()
    s.captureGraph(d_vals.data_ptr(), d_B.data_ptr(), 2, d_X.data_ptr())
    rng = np.random.default_rng(21)
    for step, scale in enumerate([1.0, 0.7, 1.3, 0.9, 1.1]):
        Ks = K * scale
        B = rng.standard_normal((n, 2))
        d_vals.copy_(torch.from_numpy(base * scale))
        d_B.copy_(torch.from_numpy(np.ascontiguousarray(B.T)))
        torch.cuda.synchronize()
        info = s.replayGraph()
        assert info.repaired == 1 and info.singular == 0, step
        assert (info.n_pos, info.n_neg, info.n_zero) == (n2, n1, 0), step
        X = d_X.cpu().numpy().T
        assert P.backward_error(Ks, X, B) <= 1e-8, step
        if step % 2 == 0:
            # eager work on the same plan between replays
            s.factorize(Ks)
            inr = s.inertia()
            assert (inr.n_pos, inr.n_neg, inr.n_zero) == (n2, n1, 0)
            Xe = s.solve(B)
            assert P.backward_error(Ks, Xe, B) <= 1e-8
        torch.cuda.synchronize()


def test_refined_graph_replays_match_eager_refinement():
    """Recorded factorize + FP64-refined solve (FP32 fronts): every replay on new values reaches the 1e-12
    backward-error bar with the closed-form inertia and equals the eager factorize + solveWithRefinement of
    the same values bit for bit (same kernels, same order), with eager work between the replays."""
    from ma57_amd.solver import RefinementOptions, csc_arrays
    G = 20
    shifts = [0.5, 0.37, 0.61]
    n = G ** 3
    dev = torch.device("cuda:0")
    s = _solver(P.laplace3d(G, G, G, shifts[0]), OrderingType.Metis, Precision.FP32)
    eager = _solver(P.laplace3d(G, G, G, shifts[0]), OrderingType.Metis, Precision.FP32)
    d_vals = torch.from_numpy(csc_arrays(P.laplace3d(G, G, G, shifts[0]))[3].copy()).to(dev)
    d_B = torch.zeros((3, n), dtype=torch.float64, device=dev)
    d_X = torch.zeros_like(d_B)
    torch.cuda.synchronize()
    opts = RefinementOptions(max_iterations=10, tolerance=1e-12)
    s.captureGraph(d_vals.data_ptr(), d_B.data_ptr(), 3, d_X.data_ptr(), refine=opts)
    rng = np.random.default_rng(33)
    for rep in range(2):
        for sh in shifts:
            A = P.laplace3d(G, G, G, sh)
            B = rng.standard_normal((n, 3))
            d_vals.copy_(torch.from_numpy(csc_arrays(A)[3]))
            d_B.copy_(torch.from_numpy(np.ascontiguousarray(B.T)))
            torch.cuda.synchronize()
            info = s.replayGraph()
            r = s.last_refinement
            X = d_X.cpu().numpy().T
            expect = tuple(P.closed_form_inertia(P.laplace_eigs_3d(G, G, G), sh)[:3])
            assert (info.n_pos, info.n_neg, info.n_zero) == expect and info.repaired == 0
            assert r.iterations >= 1 and r.backward_error <= 1e-12
            assert P.backward_error(A, X, B) <= 1e-12
            eager.factorize(A)
            re = eager.solveWithRefinement(A, B, opts)
            assert re.diagnostics.iterations == r.iterations
            assert np.array_equal(re.x, X)
            junk = torch.randn(4096, device=dev, dtype=torch.float64)
            assert torch.isfinite(junk.sum()).item()
            torch.cuda.synchronize()


def test_refined_graph_replay_with_delays_is_repaired_and_refined():
    """Threshold-mode SQD KKT: every replay delays pivots, is repaired on the regular path and refined there;
    refinement brings the backward error from the unrefined ~1e-9 to the 1e-12 bar."""
    from ma57_amd.solver import RefinementOptions, csc_arrays
    K, n1, n2 = P.sqd_kkt(nh=16, m2=1200, seed=77)
    n = K.shape[0]
    base = csc_arrays(K)[3].copy()
    dev = torch.device("cuda:0")
    s = _solver(K, OrderingType.Metis, Precision.FP64)
    d_vals = torch.from_numpy(base.copy()).to(dev)
    d_B = torch.zeros((2, n), dtype=torch.float64, device=dev)
    d_X = torch.zeros_like(d_B)
    torch.cuda.synchronize()
    s.captureGraph(d_vals.data_ptr(), d_B.data_ptr(), 2, d_X.data_ptr(),
                   refine=RefinementOptions(max_iterations=10, tolerance=1e-14))
    rng = np.random.default_rng(8)
    for scale in (1.0, 0.8):
        B = rng.standard_normal((n, 2))
        d_vals.copy_(torch.from_numpy(base * scale))
        d_B.copy_(torch.from_numpy(np.ascontiguousarray(B.T)))
        torch.cuda.synchronize()
        info = s.replayGraph()
        assert info.repaired == 1 and (info.n_pos, info.n_neg, info.n_zero) == (n2, n1, 0)
        assert s.last_refinement.iterations >= 1
        assert P.backward_error(K * scale, d_X.cpu().numpy().T, B) <= 1e-12
