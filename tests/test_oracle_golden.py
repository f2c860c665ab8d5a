"""Pins the CPU oracle against picca's own golden correlation files.

Each case replays one of picca's regression commands with the oracle (brute-force
neighbours, scalar per-pair angles, the C restatement of the pair kernel, the per-healpix
loop and the parent reduction) and checks the result with picca's comparator: NB and
HEALPID exactly, every other column within rtol 1e-5 / atol 1e-8
(py/picca/tests/test_helpers.py:93-101).
"""
import numpy as np
import pytest

import cor_cases
from oracle import cf_oracle as O

_NEIGHBOURS = {}


def oracle_rows(setup, ang_correlation=False):
    params = O.params_from_module(setup)
    key = (id(setup.data), id(setup.data2), setup.ang_max)
    neigh = _NEIGHBOURS.get(key)
    if neigh is None:
        neigh = O.brute_force_neighbours(setup.data, setup.data2, sorted(setup.data),
                                         setup.ang_max)
        _NEIGHBOURS[key] = neigh
    rows = []
    stats = np.zeros(3, dtype=np.int64)
    for hp in sorted(setup.data):
        res, st = O.compute_xi(params, setup.data, [hp], neigh, ang_correlation)
        rows.append(res)
        stats += st
    return rows, stats


@pytest.mark.parametrize("case", ["cf", "cf_image", "cf_zcuts", "cf_image_zcuts"])
def test_oracle_reproduces_golden_auto(case):
    setup = cor_cases.Setup(cor_cases.CF_CASES[case])
    rows, stats = oracle_rows(setup)
    result = cor_cases.reduce_rows(rows)
    result["HEALPID"] = np.array(sorted(setup.data))
    golden = cor_cases.golden_columns(case)
    cor_cases.compare_to_golden(result, golden)
    assert stats[0] == golden["NB"].sum()
    assert stats[2] == 0


@pytest.mark.parametrize("case", ["cf_cross", "cf_cross_zcuts"])
def test_oracle_reproduces_golden_cross(case):
    setup = cor_cases.Setup(cor_cases.CF_CASES[case])
    rows, stats = oracle_rows(setup)
    result = cor_cases.reduce_rows(rows)
    golden = cor_cases.golden_columns(case)
    cor_cases.compare_to_golden(result, golden)


def test_known_answers():
    """Known answers read from the golden headers/columns."""
    assert cor_cases.golden_columns("cf")["NB"].sum() == 36543950
    assert cor_cases.golden_columns("cf_cross")["NB"].sum() == 73087900
    assert cor_cases.golden_columns("cf_zcuts")["NB"].sum() == 153260
    assert cor_cases.golden_columns("cf_image")["NB"].sum() == 21661136
    assert list(cor_cases.golden_columns("cf_image")["HEALPID"]) == [16, 69, 94, 711]


def test_oracle_reproduces_golden_angular():
    setup = cor_cases.AngularSetup()
    rows, stats = oracle_rows(setup, ang_correlation=True)
    result = cor_cases.reduce_rows(rows)
    result["HEALPID"] = np.array(sorted(setup.data))
    golden = cor_cases.golden_columns("cf_angl")
    cor_cases.compare_to_golden(result, golden)
    assert stats[0] == golden["NB"].sum() > 0


@pytest.mark.parametrize("case", ["xcf", "xcf_zcuts"])
def test_oracle_reproduces_golden_xcf(case):
    setup = cor_cases.XcfSetup(cor_cases.XCF_CASES[case])
    params = O.params_from_module(setup)
    healpixs = sorted(setup.data)
    neigh = O.xcf_neighbours(setup.data, setup.objs, healpixs, setup.ang_max,
                             setup.r_par_min, setup.r_par_max, setup.r_trans_max)
    rows = []
    stats = np.zeros(3, dtype=np.int64)
    for hp in healpixs:
        res, st = O.compute_xcf(params, setup.data, [hp], neigh)
        rows.append(res)
        stats += st
    result = cor_cases.reduce_rows(rows)
    result["HEALPID"] = np.array(healpixs)
    golden = cor_cases.golden_columns(case)
    cor_cases.compare_to_golden(result, golden)
    assert stats[0] == golden["NB"].sum() > 0
    assert stats[2] == 0
