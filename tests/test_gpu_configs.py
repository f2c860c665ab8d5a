"""The benchmark configurations at full size on the GPU, checked against the FP64 oracle.

Each synthetic survey is built at scale 1.0 (the whole store goes to the device); the
correlation runs over every neighbour pair of a seeded sample of delta1 forests (about 2e4
forest pairs, 1e10 pixel pairs: seconds for the oracle), so that the full-size store, the
pruning and the schedule choice of the pair kernel are all in play.  Bar: num_pairs
bit-exact, edge and overflow counts equal, weights and the normalised xi / r_par / r_trans /
z within 1e-9 relative (xi against the scale of its row, as its sums cancel).

    config 0  survey_small       1000 forests, log10 lambda grid: every pair
    config 1  survey_100k        100k forests, 0.8 A grid, L ~ U{50..1500}
    config 2  survey_300k        300k forests, L log-uniform in [50, 1500]; and the
                                 distortion matrix of survey_100k, rej 0.99, 50x50 -> 100x50
    config 3  survey_100k_heavy  delta = 0.3 t(3), w log-uniform in [1e-3, 1e2]
"""
import numpy as np
import pytest

import walk_cases as W
from oracle import cf_oracle as O
from picca_amd import _native, batch, synth
from picca_amd.constants import Cosmo
from picca_amd.engine import get_engine, params_from_globals
from picca_amd.neighbours import find_neighbours
from picca_amd.utils import compute_ang_max

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def sampled_pairs(survey, n_sources, seed, r_max=200.0):
    """Every neighbour pair of n_sources random delta1 forests, in picca_cf's order (rows =
    healpix ascending, delta1 in store order, neighbours healpix ascending then file
    order), as xi_pairs builds them for the whole catalogue."""
    store, hp = survey.store, np.asarray(survey.healpix)
    ang_max = compute_ang_max(Cosmo(0.315, verbose=False), r_max, survey.z_min)
    rng = np.random.default_rng(seed)
    src = np.sort(rng.choice(store.num_forests, n_sources, replace=False))
    src = src[np.argsort(hp[src], kind="stable")].astype(np.int64)
    tgt = np.argsort(hp, kind="stable").astype(np.int64)
    nl = find_neighbours(store, src, tgt, ang_max, auto=True)
    s_of = nl.source_index()
    uniq, row_of_src = np.unique(hp[src], return_inverse=True)
    return batch.XiPairs(src[s_of], tgt[nl.target], nl.angle,
                         np.zeros(len(s_of), dtype=np.uint8),
                         row_of_src[s_of].astype(np.int32), uniq)


def run_and_check(store, pairs, g=None, schedule=None):
    g = g or W.Globals()
    res = batch.run_xi(store, pairs, params_from_globals(g), schedule=schedule)
    ref, ost = O.xi_rows(O.params_from_module(g), store, pairs.pair_a, pairs.pair_b,
                         pairs.pair_ang, pairs.pair_flags, pairs.pair_row, pairs.num_rows,
                         num_threads=16)
    rep = res.report
    W.assert_matches([res.weights, res.xi, res.r_par, res.r_trans, res.z, res.num_pairs],
                     rep["edge_pairs"], rep["overflow_pairs"], ref, ost, RTOL)
    assert rep["accepted_pairs"] == ost[0] > 0
    return res, ref, ost


def test_config0_full_catalogue():
    s = synth.make_survey("survey_small", seed=0, scale=1.0)
    ang_max = compute_ang_max(Cosmo(0.315, verbose=False), 200.0, s.z_min)
    pairs = batch.xi_pairs(s.store, s.healpix, ang_max)
    assert s.store.num_forests == 1000 and pairs.num_pairs > 100_000
    res, _, ost = run_and_check(s.store, pairs)
    # about two pixels per bin on the log grid: the automatic schedule keeps the pair loop
    assert res.report["walk_tasks"] == 0
    assert ost[0] > 1e9


@pytest.mark.parametrize("name", ["survey_100k", "survey_300k", "survey_100k_heavy"])
def test_full_size_surveys(name):
    """Both schedules of the pair kernel: the default (automatic) and the run walk, which
    takes every task whose pruned window is not empty on this grid."""
    s = synth.make_survey(name, seed=0, scale=1.0)
    pairs = sampled_pairs(s, 15, seed=11)
    assert 10_000 < pairs.num_pairs < 60_000
    run_and_check(s.store, pairs)
    res, _, _ = run_and_check(s.store, pairs, schedule=_native.XI_SCHEDULE_RUNS)
    rep = res.report
    assert rep["walk_tasks"] > 0.9 * (rep["tasks"] - rep["empty_tasks"])


def test_config2_distortion_matrix_linear_grid():
    """compute_dmat on survey_100k with rej 0.99, 50x50 data and 100x50 model bins,
    redshift evolution on, over 16 kept pairs of sampled forests."""
    s = synth.make_survey("survey_100k", seed=0, scale=1.0)
    pairs = sampled_pairs(s, 20, seed=5)
    keep = np.random.RandomState(5).rand(pairs.num_pairs) > 0.99
    idx = np.nonzero(keep)[0][:16]
    assert len(idx) == 16
    a, b, ang = pairs.pair_a[idx], pairs.pair_b[idx], pairs.pair_ang[idx]
    flags = np.zeros(len(idx), dtype=np.uint8)
    g = W.Globals(num_model_bins_r_par=100, num_model_bins_r_trans=50)
    eng = get_engine(0)
    eng.set_forests(s.store)
    mine = eng.dmat(params_from_globals(g), a, b, ang, flags)
    ref, ref_stats = O.dmat_rows(O.params_from_module(g), s.store, a, b, ang, flags)
    assert list(mine[6]) == list(ref_stats)
    assert ref_stats[0] > 1e6 and ref_stats[1] > 0
    names = ("weights_dmat", "dmat", "r_par_eff", "r_trans_eff", "z_eff", "weight_eff")
    for k, name in enumerate(names):
        m, r = np.asarray(mine[k]), np.asarray(ref[k])
        scale = float(np.abs(r).max()) if r.size else 0.0
        np.testing.assert_allclose(m, r, rtol=1e-9, atol=1e-12 * max(scale, 1e-300),
                                   err_msg=name)
