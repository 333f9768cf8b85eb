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


def test_speculative_default_for_mid_size_fronts():
    """Plans whose fronts stay at or below 4,096 eligible rows keep the speculative step for fronts above
    2,048 rows (the 7-point Laplacians' panels fail the first test too often for it to pay below that)."""
    A = P.laplace3d(20, 20, 20, 0.5)
    s = SymLDLT()
    s.setOrdering(OrderingType.Metis)
    s.analyzePattern(A)
    assert max(x.ncols for x in s.symbolicFactor().supernodes) <= 4096
    t = s._get_tuning()
    assert t.speculative_panel == 1 and t.speculative_min_rows == 2048


def test_bad_tuning_rejected():
    A = P.laplace2d(8, 8, 0.3)
    s = SymLDLT()
    s.analyzePattern(A)
    for small, tile in ((193, 64), (-1, 64), (64, 48), (64, 8), (64, 2048)):
        with pytest.raises(InvalidArgument):
            s.setTuning(small, tile)


def _tuple(s):
    t = s._get_tuning()
    return (t.small_front_max, t.assemble_tile, t.speculative_panel, t.speculative_min_rows,
            t.deferred_cb_min_rows)


def test_autotune_coordinate_descent_picks_fastest_and_stays_correct(tmp_path):
    """The autotuner times every candidate of each scheduling choice with the others fixed (no tuning twice),
    keeps the fastest trial, leaves a correct factorization behind, and a cached result for the same key is
    applied without timing; another precision is another key."""
    from ma57_amd.solver import Precision
    G, shift = 24, 0.5
    A = P.laplace3d(G, G, G, shift)
    s = SymLDLT()
    s.setOrdering(OrderingType.Metis)
    s.analyzePattern(A)
    d_vals = torch.from_numpy(csc_arrays(A)[3].copy()).to("cuda:0")
    torch.cuda.synchronize()
    cache = tmp_path / "tune.txt"
    chosen, trials, cached = s.autotune(d_vals.data_ptr(), nrhs=4, reps=1, cache=cache)
    assert not cached
    tunings = [t for t, _ in trials]
    assert len(tunings) == len(set(tunings)) and len(tunings) >= 15  # 6 + 3 + 5 + 4 candidates, the start shared
    assert {t[0] for t in tunings} >= {0, 32, 64, 96, 128, 192} and {t[4] for t in tunings} >= {-1, 1024, 2048, 4096}
    best = min(trials, key=lambda t: t[1])
    assert chosen == best[0] == _tuple(s)
    s.factorize(A)
    i = s.inertia()
    assert (i.n_pos, i.n_neg, i.n_zero) == _expected(G, shift)
    B = np.random.default_rng(10).standard_normal((A.shape[0], 2))
    assert P.backward_error(A, s.solve(B), B) <= 1e-12
    # the same key: applied from the cache
    s2 = SymLDLT()
    s2.setOrdering(OrderingType.Metis)
    s2.analyzePattern(A)
    chosen2, trials2, cached2 = s2.autotune(d_vals.data_ptr(), nrhs=4, reps=1, cache=cache)
    assert cached2 and trials2 == [] and chosen2 == chosen == _tuple(s2)
    assert len(cache.read_text().splitlines()) == 1
    # FP32 fronts are another key
    s2.setPrecision(Precision.FP32)
    _, trials3, cached3 = s2.autotune(d_vals.data_ptr(), nrhs=4, reps=1, cache=cache)
    assert not cached3 and len(trials3) >= 15
    assert len(cache.read_text().splitlines()) == 2
