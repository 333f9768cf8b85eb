"""Scheduling choices (small-front bound, assembly tile) and the autotuner on
the HIP path: every choice must give the same inertia and the FP64 backward
error bar, since it only moves fronts between the LDS kernel and the panel
path and changes the assembly tiling."""
import numpy as np
import pytest

from ma57_amd import problems as P
from ma57_amd.solver import SymLDLT, OrderingType, InvalidArgument, csc_arrays

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _expected(G, shift):
    return tuple(P.closed_form_inertia(P.laplace_eigs_3d(G, G, G), shift)[:3])


@pytest.mark.parametrize("small,tile", [(0, 16), (64, 32), (128, 128), (192, 1024)])
def test_every_tuning_is_correct(small, tile):
    G, shift = 20, 0.5
    A = P.laplace3d(G, G, G, shift)
    s = SymLDLT()
    s.setOrdering(OrderingType.Metis)
    s.analyzePattern(A)
    s.setTuning(small, tile)
    assert s.tuning() == (small, tile)
    s.factorize(A)
    i = s.inertia()
    assert (i.n_pos, i.n_neg, i.n_zero) == _expected(G, shift)
    B = np.random.default_rng(9).standard_normal((A.shape[0], 4))
    assert P.backward_error(A, s.solve(B), B) <= 1e-12


def test_bad_tuning_rejected():
    A = P.laplace2d(8, 8, 0.3)
    s = SymLDLT()
    s.analyzePattern(A)
    for small, tile in ((193, 64), (-1, 64), (64, 48), (64, 8), (64, 2048)):
        with pytest.raises(InvalidArgument):
            s.setTuning(small, tile)


def test_autotune_picks_fastest_and_stays_correct():
    G, shift = 24, 0.5
    A = P.laplace3d(G, G, G, shift)
    s = SymLDLT()
    s.setOrdering(OrderingType.Metis)
    s.analyzePattern(A)
    d_vals = torch.from_numpy(csc_arrays(A)[3].copy()).to("cuda:0")
    torch.cuda.synchronize()
    chosen, trials = s.autotune(d_vals.data_ptr(), nrhs=4, reps=1)
    assert len(trials) == 15
    best = min(trials, key=lambda t: t[1])
    assert chosen == best[0] == s.tuning()
    s.factorize(A)
    i = s.inertia()
    assert (i.n_pos, i.n_neg, i.n_zero) == _expected(G, shift)
    B = np.random.default_rng(10).standard_normal((A.shape[0], 2))
    assert P.backward_error(A, s.solve(B), B) <= 1e-12
