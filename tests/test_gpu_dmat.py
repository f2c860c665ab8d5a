"""Parity of the HIP distortion-matrix kernels with the CPU oracle and picca's golden files.

The GPU regroups the reference's fill loop (py/picca/cf.py:846-887) into per-pixel sums
and outer products, so the FP64 sums are added in another order.  Bar: the counts (NPALL,
NPUSED, in-range / selected / beyond-count-pass / overflow pairs) identical; weights_dmat,
the effective coordinates and every dmat entry within 1e-9 relative or 1e-12 of the
largest entry (whichever is larger); normalised columns within picca's own test tolerance
of the golden files.
"""
import numpy as np
import pytest

import cor_cases
from oracle import cf_oracle as O
from picca_amd import batch, cf, synth
from picca_amd.constants import Cosmo
from picca_amd.engine import get_engine, params_from_globals
from picca_amd.utils import compute_ang_max
from test_gpu_xi import set_globals
from test_oracle_dmat_golden import dmat_columns, oracle_dmat

pytestmark = pytest.mark.gpu

RTOL = 1e-9
ATOL_REL = 1e-12


def assert_close(mine, ref, name):
    mine = np.asarray(mine)
    ref = np.asarray(ref)
    assert mine.shape == ref.shape, name
    scale = float(np.abs(ref).max()) if ref.size else 0.0
    np.testing.assert_allclose(mine, ref, rtol=RTOL, atol=ATOL_REL * max(scale, 1e-300),
                               err_msg=name)


def assert_outputs_close(mine, ref):
    names = ("weights_dmat", "dmat", "r_par_eff", "r_trans_eff", "z_eff", "weight_eff")
    for k, name in enumerate(names):
        assert_close(mine[k], ref[k], name)


@pytest.mark.parametrize("case", sorted(cor_cases.DMAT_CASES))
def test_fixture_dmat_matches_oracle_and_golden(case):
    setup = cor_cases.Setup(cor_cases.DMAT_CASES[case], dmat=True)
    set_globals(setup)
    healpixs = sorted(setup.data)
    ref, ref_stats = oracle_dmat(setup)
    # picca_dmat seeds the legacy stream with the first healpix of the process list
    # (py/picca/bin/picca_dmat.py:36), then fills the neighbours and computes
    np.random.seed(healpixs[0])
    cf.fill_neighs(healpixs)
    mine = cf.compute_dmat(healpixs)
    assert mine[6] == ref[6]     # num_pairs
    assert mine[7] == ref[7]     # num_pairs_used
    assert_outputs_close(mine[:6], ref[:6])
    rep = cf.last_report
    assert rep["in_range_pairs"] == ref_stats[0]
    assert rep["selected_pairs"] == ref_stats[1]
    assert rep["beyond_count_pass"] == ref_stats[2]
    cols, npall, npused = dmat_columns(mine, setup.redshift_evolution_in_distortion_matrix)
    golden = cor_cases.FITS(cor_cases.golden_path(case))
    header = golden["DMAT"].read_header()
    assert npall == header["NPALL"]
    assert npused == header["NPUSED"]
    cor_cases.compare_to_golden(cols, cor_cases.golden_columns(case))


@pytest.mark.parametrize("case", sorted(cor_cases.DMAT_CASES))
def test_dmat_driver_reproduces_golden(case, tmp_path):
    from picca_amd.bin import picca_dmat
    out = str(tmp_path / (case + ".fits.gz"))
    picca_dmat.main(cor_cases.DMAT_CASES[case] + ["--out", out])
    mine = cor_cases.FITS(out)
    golden = cor_cases.FITS(cor_cases.golden_path(case))
    for key in ("NPALL", "NPUSED"):
        assert mine["DMAT"].read_header()[key] == golden["DMAT"].read_header()[key]
    result = {}
    for hdu in mine.hdus[1:]:
        assert hdu.get_colnames() == golden[hdu.get_extname()].get_colnames()
        for col in hdu.get_colnames():
            result[col] = hdu[col][:]
    cor_cases.compare_to_golden(result, cor_cases.golden_columns(case))


class Globals:
    num_bins_r_par = 15
    num_bins_r_trans = 15
    num_model_bins_r_par = 30
    num_model_bins_r_trans = 15
    r_par_min = 0.0
    r_par_max = 60.0
    r_trans_max = 60.0
    z_min_pairs = None
    z_max_pairs = None
    zerr_cut_deg = None
    zerr_cut_kms = None
    z_ref = 2.25
    alpha = 2.9
    alpha2 = 2.9
    x_correlation = False
    rmu_binning = False
    ang_correlation = False
    remove_same_half_plate_close_pairs = False
    redshift_evolution_in_distortion_matrix = True


def make_globals(**overrides):
    g = Globals()
    for k, v in overrides.items():
        setattr(g, k, v)
    return g


@pytest.fixture(scope="module")
def dmat_survey():
    """~130 forests of the 1000-forest log-grid survey and 60 of their neighbour pairs."""
    s = synth.make_survey("survey_small", seed=3, scale=0.13)
    ang_max = compute_ang_max(Cosmo(0.315, verbose=False), 60.0, s.z_min)
    pairs = batch.xi_pairs(s.store, s.healpix, ang_max)
    rng = np.random.default_rng(5)
    sel = np.sort(rng.choice(pairs.num_pairs, min(60, pairs.num_pairs), replace=False))
    return s.store, pairs.pair_a[sel], pairs.pair_b[sel], pairs.pair_ang[sel]


MODES = {
    "auto": {},
    "no_evolution": {"redshift_evolution_in_distortion_matrix": False, "alpha2": 1.5},
    "cross": {"x_correlation": True, "r_par_min": -60.0, "num_bins_r_par": 30,
              "num_model_bins_r_par": 60},
    "zcuts": {"z_min_pairs": 2.3, "z_max_pairs": 2.6},
    "zerr": {"zerr_cut_deg": 5.0, "zerr_cut_kms": 25000.0},
    "same_half_plate": {"remove_same_half_plate_close_pairs": True},
    "rmu": {"rmu_binning": True, "r_par_min": 0.0, "r_par_max": 1.0},
    "config_binning": {"num_bins_r_par": 50, "num_bins_r_trans": 50,
                       "num_model_bins_r_par": 100, "num_model_bins_r_trans": 50,
                       "r_par_max": 200.0, "r_trans_max": 200.0},
    "chunked": {"num_bins_r_par": 80, "num_bins_r_trans": 40, "num_model_bins_r_par": 150,
                "num_model_bins_r_trans": 60},
}


@pytest.mark.parametrize("mode", sorted(MODES))
def test_synthetic_modes_match_oracle(dmat_survey, mode):
    store, a, b, ang = dmat_survey
    g = make_globals(**MODES[mode])
    # every other pair on one half plate (only read with the removal switched on)
    flags = (np.arange(len(a)) % 2).astype(np.uint8)
    if mode == "no_evolution":
        store.order[::3] = 0     # some forests fitted with order 0
    try:
        eng = get_engine(0)
        eng.set_forests(store)
        mine = eng.dmat(params_from_globals(g), a, b, ang, flags)
        ref, ref_stats = O.dmat_rows(O.params_from_module(g), store, a, b, ang, flags)
    finally:
        store.order[:] = 1
    assert ref_stats[0] > 0 and ref_stats[1] > 0
    assert_outputs_close(mine[:6], ref)
    assert list(mine[6]) == list(ref_stats)
    if mode == "same_half_plate":
        assert ref_stats[2] > 0   # the count-pass truncation of U is exercised


def test_empty_pair_list_and_zero_weight_forest(dmat_survey):
    store, a, b, ang = dmat_survey
    g = make_globals()
    eng = get_engine(0)
    eng.set_forests(store)
    out = eng.dmat(params_from_globals(g), [], [], [], [])
    assert all(np.count_nonzero(x) == 0 for x in out[:6])
    # a forest whose weights are all zero adds nothing and does not break the others
    w = store.weights.copy()
    k = int(a[0])
    store.weights[store.offsets[k]:store.offsets[k + 1]] = 0.0
    try:
        eng.set_forests(store)
        mine = eng.dmat(params_from_globals(g), a[:10], b[:10], ang[:10], np.zeros(10, np.uint8))
        ref, ref_stats = O.dmat_rows(O.params_from_module(g), store, a[:10], b[:10], ang[:10],
                                     np.zeros(10, np.uint8))
    finally:
        store.weights[:] = w
    assert_outputs_close(mine[:6], ref)
    assert list(mine[6]) == list(ref_stats)
