"""The xi launch sequence replayed from its captured HIP graph against direct launches.

picca_xi_execute captures its clearing kernels and the pair kernel in a graph and replays it
while the launch arguments stay the same.  Each replay, with eager device work, other xi runs
(some captured in a graph of their own) between replays, must give what the
direct launch gives: num_pairs and every counter equal; the FP64 sums within 1e-12 relative
(the LDS histogram adds pairs with atomics, so the last bits follow the order in which the
waves reach a bin, run to run).  The first replay is also checked against the oracle.
"""
import numpy as np
import pytest

import walk_cases as W
from picca_amd import batch
from picca_amd.engine import get_engine, params_from_globals

pytestmark = pytest.mark.gpu

COUNTERS = ("accepted_pairs", "edge_pairs", "overflow_pairs", "tasks", "empty_tasks",
            "walk_tasks", "runs", "probes")


def rows(res):
    return [res.weights, res.xi, res.r_par, res.r_trans, res.z]


def same_result(a, b):
    np.testing.assert_array_equal(a.num_pairs, b.num_pairs)
    for x, y in zip(rows(a), rows(b)):
        scale = float(np.abs(y).max()) or 1.0
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-12 * scale)
    for k in COUNTERS:
        assert a.report[k] == b.report[k], k


@pytest.mark.parametrize("name", ["golden_cf", "survey_small"])
def test_graph_replays_equal_direct_launch(name):
    case = W.golden_case("cf") if name == "golden_cf" else W.survey_case("survey_small", 0.05)
    other = W.random_case(3)
    eng = get_engine(0)
    params = params_from_globals(case.g)
    try:
        eng.use_xi_graph(False)
        direct = batch.run_xi(case.store, case.pairs, params)
        eng.use_xi_graph(True)
        ref, ost = W.oracle(case)
        for replay in range(4):
            res = batch.run_xi(case.store, case.pairs, params)
            same_result(res, direct)
            if replay == 0:
                W.assert_matches(rows(res) + [res.num_pairs], res.report["edge_pairs"],
                                 res.report["overflow_pairs"], ref, ost, 1e-9)
            # the engine's own work between replays: a direct launch on other forests, and
            # every other replay a captured run of those (a new graph)
            eng.use_xi_graph(replay % 2 == 0)
            o = batch.run_xi(other.store, other.pairs, params_from_globals(other.g))
            assert o.report["accepted_pairs"] > 0
            eng.use_xi_graph(True)
        # the same forests and pairs with another r_par range: another graph
        g2 = W.Globals(**{**case.g.__dict__, "r_par_max": 100.0})
        res2 = batch.run_xi(case.store, case.pairs, params_from_globals(g2))
        eng.use_xi_graph(False)
        direct2 = batch.run_xi(case.store, case.pairs, params_from_globals(g2))
        same_result(res2, direct2)
    finally:
        eng.use_xi_graph(True)
