"""The run walk of the pair kernel on the GPU: the same inputs as the CPU harness tests.

Each case runs with the three schedules (runs, pairs, automatic) and each result is checked
against the FP64 oracle (num_pairs bit-exact, edge and overflow counts equal, sums within
1e-9 relative after normalisation).  The GPU's run and probe counts must equal the CPU
harness's, which runs the same lane program under bounds checks: the walk took the same
path on both.
"""
import numpy as np
import pytest

import walk_cases as W
import walk_harness as H
from picca_amd import _native, batch
from picca_amd.engine import params_from_globals

pytestmark = pytest.mark.gpu

SCHEDULES = {"runs": _native.XI_SCHEDULE_RUNS, "pairs": _native.XI_SCHEDULE_PAIRS,
             "auto": _native.XI_SCHEDULE_AUTO}


def gpu_rows(res):
    return [res.weights, res.xi, res.r_par, res.r_trans, res.z, res.num_pairs]


def check_case(case, rtol=1e-9):
    ref, ost = W.oracle(case)
    params = params_from_globals(case.g)
    p = case.pairs
    _, hst = H.xi_walk(params, case.store, p.pair_a, p.pair_b, p.pair_ang, p.pair_row,
                       p.num_rows, schedule=H.SCHEDULE_RUNS)
    reports = {}
    for name, schedule in SCHEDULES.items():
        res = batch.run_xi(case.store, case.pairs, params, schedule=schedule)
        rep = res.report
        W.assert_matches(gpu_rows(res), rep["edge_pairs"], rep["overflow_pairs"], ref, ost,
                         rtol)
        assert rep["accepted_pairs"] == ost[0]
        reports[name] = rep
    runs = reports["runs"]
    assert runs["walk_tasks"] == hst["walk_tasks"]
    assert runs["runs"] == hst["runs"]
    assert runs["probes"] == hst["probes"]
    assert reports["pairs"]["walk_tasks"] == 0
    return reports


def test_log_grid_survey():
    reports = check_case(W.survey_case("survey_small", 0.05))
    assert reports["runs"]["walk_tasks"] > 1000
    assert reports["auto"]["walk_tasks"] == 0


def test_linear_grid_survey_slice():
    reports = check_case(W.survey_case("survey_100k", 0.001))
    assert reports["runs"]["walk_tasks"] > 1000


@pytest.mark.parametrize("case", ["cf", "cf_image"])
def test_golden_delta_files(case):
    check_case(W.golden_case(case))


def test_exact_bin_edges_on_a_lattice():
    reports = check_case(W.lattice_case())
    assert reports["runs"]["edge_pairs"] > 1000


def test_zero_weights_short_and_irregular_forests():
    check_case(W.irregular_case())


@pytest.mark.parametrize("seed", range(12))
def test_random_geometries(seed):
    check_case(W.random_case(seed), rtol=1e-8)
