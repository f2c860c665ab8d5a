"""CPU tests of the command-line layer (picca_amd.driver, picca_amd.bin) and of the error
conditions picca raises on its correlation path, each forced from host inputs before any
device work:

* ValueError for --zerr-cut-deg without --zerr-cut-kms (py/picca/bin/picca_cf.py:362-365);
* AssertionError when no forest is read (py/picca/io.py:489-490);
* RuntimeError for a combined-reobservation fiber id under same-half-plate removal
  (py/picca/cf.py:171-179);
* RuntimeError for a distortion matrix of deltas without a continuum-fit order
  (py/picca/cf.py:433-438);
* IndexError for a negative data bin (py/picca/cf.py:855-856), which the library reports as
  PICCA_E_INDEX.
"""
import argparse

import numpy as np
import pytest

import cor_cases
from picca_amd import _native, cf, constants, driver, io, utils
from picca_amd.bin import picca_cf, picca_dmat, picca_metal_dmat

# (program, option, default) of the reference's command lines, the interface kept
# (py/picca/bin/picca_cf.py:42-338, picca_xcf.py, picca_dmat.py, picca_cf_angl.py,
# picca_metal_xdmat.py, picca_wick.py, picca_export.py)
KNOWN_DEFAULTS = [
    ("picca_cf", "rp_min", 0.0), ("picca_cf", "np", 50), ("picca_cf", "nside", 16),
    ("picca_cf", "z_evol2", 2.9), ("picca_cf", "lambda_abs", "LYA"),
    ("picca_xcf", "rp_min", -200.0), ("picca_xcf", "np", 100), ("picca_xcf", "z_evol_obj", 1.44),
    ("picca_xcf", "mode", "sdss"), ("picca_dmat", "rej", 1.0),
    ("picca_dmat", "coef_binning_model", 1), ("picca_cf_angl", "wr_max", 1.1),
    ("picca_cf_angl", "ang_max", 0.02), ("picca_metal_xdmat", "abs_igm", None),
    ("picca_metal_dmat", "abs_igm", []), ("picca_wick", "max_diagram", 3),
    ("picca_export", "num_boot_cov", 0),
]


def required_argv(program):
    args = {"picca_export": ["--data", "x", "--out", "y"],
            "picca_wick": ["--out", "y", "--in-dir", "x", "--cf1d", "z"]}
    if program in args:
        return args[program]
    argv = ["--out", "y", "--in-dir", "x"]
    if "--drq" in driver.PROGRAMS[program][1]:
        argv += ["--drq", "q"]
    return argv


@pytest.mark.parametrize("program,option,value", KNOWN_DEFAULTS)
def test_flag_defaults(program, option, value):
    opts = driver.parse(program, required_argv(program))
    assert getattr(opts, option) == value


def test_every_program_parses_and_requires_its_inputs():
    for program in driver.PROGRAMS:
        driver.parse(program, required_argv(program))
        with pytest.raises(SystemExit):
            driver.parse(program, [])


def test_choices_are_enforced():
    argv = required_argv("picca_xcf") + ["--mode", "boss"]
    with pytest.raises(SystemExit):
        driver.parse("picca_xcf", argv)


def test_rmu_binning_forces_the_r_par_range():
    opts = driver.parse("picca_cf", required_argv("picca_cf") + ["--rmu-binning",
                                                                  "--rp-min", "-5"])
    s = driver.binning_settings(opts)
    assert (s["r_par_min"], s["r_par_max"], s["rmu_binning"]) == (0.0, 1.0, True)
    opts = driver.parse("picca_xcf", required_argv("picca_xcf") + ["--rmu-binning"])
    s = driver.binning_settings(opts, signed=True)
    assert (s["r_par_min"], s["r_par_max"]) == (-1.0, 1.0)


def test_combine_rows_weighted_means():
    # two pixels, two bins; rows already normalised per pixel
    rows = np.zeros((2, 6, 2))
    rows[0, 0] = [1.0, 0.0]      # weights
    rows[1, 0] = [3.0, 2.0]
    rows[0, 2] = [10.0, 99.0]    # r_par (bin 1 of pixel 0 has no weight)
    rows[1, 2] = [14.0, 5.0]
    rows[:, 5] = [[4, 0], [6, 7]]
    out = driver.combine_rows(rows)
    assert out["RP"][0] == pytest.approx((1 * 10 + 3 * 14) / 4)
    assert out["RP"][1] == pytest.approx(5.0)
    assert list(out["NB"]) == [10, 7]
    assert out["WE"] is not None and out["WE"].shape == (2, 2)


def test_combine_rows_empty_bin_stays_zero():
    rows = np.zeros((3, 6, 4))
    out = driver.combine_rows(rows)
    assert not np.any(out["RP"]) and not np.any(out["NB"])


def test_process_lists_round_robin():
    data = {7: [], 3: [], 12: [], 5: [], 9: []}
    assert driver.process_lists(data, 2) == [[3, 7, 12], [5, 9]]
    assert driver.process_lists(data, 1) == [[3, 5, 7, 9, 12]]
    assert driver.process_lists(data, 8) == [[3], [5], [7], [9], [12]]


def test_normalise_matrix_and_rescale():
    wdm = np.array([2.0, 0.0])
    dmat = np.array([[2.0, 4.0], [1.0, 1.0]])
    r_par, r_trans, z = np.array([4.0, 0.0]), np.array([6.0, 0.0]), np.array([5.0, 0.0])
    weights = np.array([2.0, 0.0])
    mean_z = driver.normalise_matrix(wdm, dmat, r_par, r_trans, z, weights)
    assert mean_z == 2.5
    assert list(dmat[0]) == [1.0, 2.0] and list(dmat[1]) == [1.0, 1.0]
    driver.redshift_rescale(dmat, z, mean_z, 2.25, 2.9, 2.9)
    factor = (3.25 / 3.5) ** 3.8
    assert dmat[0, 1] == pytest.approx(2.0 * factor)
    assert list(z) == [2.5, 2.5]


def test_absorber_pairs_of_one_field():
    pairs = picca_metal_dmat.absorber_pairs(["LYA", "SiII(1260)", "SiIII(1207)"],
                                            ["LYA", "SiII(1260)", "SiIII(1207)"], False)
    assert pairs == [("LYA", "SiII(1260)"), ("LYA", "SiIII(1207)"),
                     ("SiII(1260)", "SiII(1260)"), ("SiII(1260)", "SiIII(1207)"),
                     ("SiIII(1207)", "SiIII(1207)")]


def test_absorber_pairs_of_two_fields_and_fast_mode():
    pairs = picca_metal_dmat.absorber_pairs(["LYA", "SiII(1260)"], ["LYB", "SiII(1260)"],
                                            False)
    assert pairs == [("LYA", "SiII(1260)"), ("SiII(1260)", "LYB"),
                     ("SiII(1260)", "SiII(1260)")]
    fast = picca_metal_dmat.absorber_pairs(["LYA", "SiII(1260)", "CIV(eff)"],
                                           ["LYA", "SiII(1260)", "CIV(eff)"], True)
    assert fast == [("LYA", "SiII(1260)"), ("CIV(eff)", "CIV(eff)")]


def test_settings_reset_module_globals():
    cf.reject = 0.5
    cf.x_correlation = True
    driver.Settings(nside=8).apply(cf)
    assert cf.reject is None and cf.x_correlation is False and cf.nside == 8
    driver.reset_module(cf)
    assert cf.nside is None


# --------------------------------------------------------------------------- errors


def test_zerr_options_must_come_together():
    argv = cor_cases.CF_CASES["cf"] + ["--out", "unused.fits", "--zerr-cut-deg", "1.0"]
    with pytest.raises(ValueError):
        picca_cf.main(argv)
    argv = cor_cases.DMAT_CASES["dmat"] + ["--out", "unused.fits", "--zerr-cut-kms", "100"]
    with pytest.raises(ValueError):
        picca_dmat.main(argv)


def test_no_data_is_an_assertion_error():
    cosmo = constants.Cosmo(Om=0.315, verbose=False)
    with pytest.raises(AssertionError):
        io.read_deltas(cor_cases.DELTA_LYA, 16, constants.ABSORBER_IGM["LYA"], 2.9, 2.25,
                       cosmo, z_min_qso=100.0, z_max_qso=200.0,
                       delta_attributes=cor_cases.ATTRIBUTES)


def _cf_globals(setup):
    driver.reset_module(cf)
    for name in ("num_bins_r_par", "num_bins_r_trans", "num_model_bins_r_par",
                 "num_model_bins_r_trans", "r_par_min", "r_par_max", "r_trans_max",
                 "z_min_pairs", "z_max_pairs", "zerr_cut_deg", "zerr_cut_kms", "z_ref",
                 "alpha", "alpha2", "x_correlation", "rmu_binning", "ang_correlation",
                 "remove_same_half_plate_close_pairs", "ang_max", "data", "data2", "reject",
                 "redshift_evolution_in_distortion_matrix"):
        setattr(cf, name, getattr(setup, name))
    cf.nside = 16


def test_string_fiberid_under_same_half_plate_removal():
    setup = cor_cases.Setup(cor_cases.CF_CASES["cf"])
    _cf_globals(setup)
    healpixs = sorted(setup.data)
    # one forest of the second pixel carries a combined-reobservation id: it is a delta2
    # of the first pixel's pairs and a delta1 of its own
    setup.data[healpixs[1]][0].fiberid = "combined"
    with pytest.raises(RuntimeError):
        cf.build_xi_pairs(healpixs[:2])
    cf.remove_same_half_plate_close_pairs = False
    cf.build_xi_pairs(healpixs[:2])


def test_distortion_matrix_needs_the_fit_order():
    setup = cor_cases.Setup(cor_cases.DMAT_CASES["dmat"], dmat=True)
    _cf_globals(setup)
    hp = sorted(setup.data)[0]
    for delta in setup.data[hp]:
        delta.order = None
    cf.reject = -1.0
    cf.fill_neighs([hp])
    with pytest.raises(RuntimeError):
        cf.build_dmat_pairs([hp])


def test_negative_bin_status_is_an_index_error():
    with pytest.raises(IndexError):
        _native.check(_native.PICCA_E_INDEX, "dmat")
    with pytest.raises(ValueError):
        _native.check(_native.PICCA_E_VALUE)
    with pytest.raises(RuntimeError):
        _native.check(_native.PICCA_E_RUNTIME)


def test_parser_is_argparse():
    assert isinstance(driver.parse("picca_cf", required_argv("picca_cf")), argparse.Namespace)


def test_angle_limit_uses_lowest_redshift():
    cosmo = constants.Cosmo(Om=0.315, verbose=False)
    a = utils.compute_ang_max(cosmo, 200.0, 2.0)
    b = utils.compute_ang_max(cosmo, 200.0, 2.0, 1.5)
    assert b > a
