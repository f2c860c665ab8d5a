"""The run walk of the pair kernel, run on the CPU against the FP64 oracle.

tests/native/xi_walk_host.hip compiles the kernel's lane program (picca_amd/csrc/xi_walk.h)
for the host and runs it with the kernel's task split under bounds checks.  Bar, as for the
GPU: num_pairs bit-exact, edge and overflow counts equal to the oracle's, the five sums
within 1e-9 relative after normalisation (the walk sums runs from prefix sums, so only the
FP64 summation order differs), and no access outside a forest.
"""
import pytest

import walk_cases as W
import walk_harness as H
from picca_amd.engine import params_from_globals

RTOL = 1e-9


def run(case, schedule=H.SCHEDULE_RUNS, rtol=RTOL):
    p = case.pairs
    out, st = H.xi_walk(params_from_globals(case.g), case.store, p.pair_a, p.pair_b,
                        p.pair_ang, p.pair_row, p.num_rows, schedule=schedule)
    ref, ost = W.oracle(case)
    assert st["oob"] == 0
    W.assert_matches(out, st["edge"], st["overflow"], ref, ost, rtol)
    return st, ost


def test_log_grid_survey():
    """Config-0 geometry (log10 lambda grid, about two pixels per r_par bin)."""
    st, ost = run(W.survey_case("survey_small", 0.05))
    assert st["pair_tasks"] == 0 and st["walk_tasks"] > 1000
    assert st["runs"] > 0 and ost[0] > 1e7


def test_linear_grid_survey_slice():
    """Config-1 geometry (0.8 A linear grid, forests of 50-1500 pixels), a 0.1 % slice."""
    st, _ = run(W.survey_case("survey_100k", 0.001))
    assert st["pair_tasks"] == 0
    # about eight pixel pairs per run on this grid, about two probes per run
    assert st["cover"] > 6 * st["runs"]
    assert st["probes"] < 3 * st["runs"]


def test_auto_schedule_takes_pairs_on_the_log_grid():
    """The default schedule walks only tasks with enough pixels per bin: on the log grid
    (about two per bin) it keeps the pair loop."""
    st, _ = run(W.survey_case("survey_small", 0.05), schedule=H.SCHEDULE_AUTO)
    assert st["walk_tasks"] == 0 and st["pair_tasks"] > 1000


def test_pair_schedule_of_harness_matches_oracle():
    st, _ = run(W.survey_case("survey_small", 0.05), schedule=H.SCHEDULE_PAIRS)
    assert st["walk_tasks"] == 0 and st["runs"] == 0


@pytest.mark.parametrize("case", ["cf", "cf_image"])
def test_golden_delta_files(case):
    """picca's own test deltas (Delta_LYA, 936 forests; Delta_LYA_image, 55 forests)."""
    st, ost = run(W.golden_case(case))
    assert ost[0] > 1e6
    assert st["walk_tasks"] > 0


def test_exact_bin_edges_on_a_lattice():
    st, ost = run(W.lattice_case())
    assert ost[1] > 1000          # many exact edge pairs
    assert st["pair_tasks"] == 0


def test_zero_weights_short_and_irregular_forests():
    st, _ = run(W.irregular_case())
    assert st["walk_tasks"] > 0 and st["pair_tasks"] > 0


@pytest.mark.parametrize("seed", range(12))
def test_random_geometries(seed):
    run(W.random_case(seed), rtol=1e-8)


def test_reciprocal_guess_of_the_run_end():
    """walk_rcp_guess: the seed (bit pattern of 2^1023 minus that of x) is within 12.5 % of
    1 / x, and each Newton step squares the relative error: (1/8)^4 = 2.4e-4 at most.  The
    worst case is the mantissa 1.5 (seed 0.75 / 2^k, x * seed = 9/8); powers of two are
    exact."""
    import numpy as np
    rng = np.random.default_rng(7)
    xs = np.concatenate([10.0 ** rng.uniform(-12, 12, 4000), 2.0 ** np.arange(-40, 41),
                         1.5 * 2.0 ** np.arange(-40, 41), [0.05, 0.25, 4.0 / 3.0]])
    err = np.array([abs(x * H.rcp_guess(x) - 1.0) for x in xs])
    assert err.max() < 2.5e-4
    assert err[4000:4081].max() == 0.0                     # powers of two
    np.testing.assert_allclose(err[4081:4162], 0.125 ** 4, rtol=1e-9)
