"""Shared set-up of the correlation test cases.

The cases replay the commands of picca's own regression tests
(``py/picca/tests/test_3_cor.py``: test_cf :229-257, test_cf_image_data :259-286,
test_cf_cross :458-488, test_cf_zcuts :791-823, test_cf_image_data_zcuts :825-856,
test_cf_cross_zcuts :895-929, and the dmat tests :318-349, :490-524, :858-893,
:931-968) against the golden files copied into tests/golden.
"""
import os

import numpy as np

from picca_amd import constants, io, utils
from picca_amd.fitsio_lite import FITS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DELTA_LYA = os.path.join(GOLDEN, "Delta_LYA")
DELTA_LYA_IMAGE = os.path.join(GOLDEN, "Delta_LYA_image")
ATTRIBUTES = os.path.join(GOLDEN, "delta_attributes.fits.gz")

COMMON = ["--rp-max", "+60.0", "--rt-max", "+60.0", "--nt", "15", "--nproc", "1",
          "--in-attributes", ATTRIBUTES]
ZCUTS = ["--z-min-pairs", "2.25", "--z-max-pairs", "2.3", "--z-min-sources", "2.3",
         "--z-max-sources", "2.5"]

CF_CASES = {
    "cf": ["--in-dir", DELTA_LYA, "--rp-min", "+0.0", "--np", "15",
           "--remove-same-half-plate-close-pairs"] + COMMON,
    "cf_image": ["--in-dir", DELTA_LYA_IMAGE, "--rp-min", "+0.0", "--np", "15"] + COMMON,
    "cf_cross": ["--in-dir", DELTA_LYA, "--in-dir2", DELTA_LYA, "--rp-min", "-60.0", "--np",
                 "30", "--remove-same-half-plate-close-pairs", "--unfold-cf"] + COMMON,
    "cf_zcuts": ["--in-dir", DELTA_LYA, "--rp-min", "+0.0", "--np", "15",
                 "--remove-same-half-plate-close-pairs"] + ZCUTS + COMMON,
    "cf_image_zcuts": ["--in-dir", DELTA_LYA_IMAGE, "--rp-min", "+0.0", "--np", "15"]
                      + ZCUTS + COMMON,
    "cf_cross_zcuts": ["--in-dir", DELTA_LYA, "--in-dir2", DELTA_LYA, "--rp-min", "-60.0",
                       "--np", "30", "--remove-same-half-plate-close-pairs", "--unfold-cf"]
                      + ZCUTS + COMMON,
}

DMAT_CASES = {
    "dmat": ["--in-dir", DELTA_LYA, "--rp-min", "+0.0", "--np", "15", "--rej", "0.99",
             "--remove-same-half-plate-close-pairs", "--no-redshift-evolution"] + COMMON,
    "dmat_cross": ["--in-dir", DELTA_LYA, "--in-dir2", DELTA_LYA, "--rp-min", "-60.0",
                   "--np", "30", "--rej", "0.99", "--remove-same-half-plate-close-pairs",
                   "--unfold-cf", "--no-redshift-evolution"] + COMMON,
    "dmat_zcuts": ["--in-dir", DELTA_LYA, "--rp-min", "+0.0", "--np", "15", "--rej", "0.99",
                   "--remove-same-half-plate-close-pairs", "--no-redshift-evolution"]
                  + ZCUTS + COMMON,
    "dmat_cross_zcuts": ["--in-dir", DELTA_LYA, "--in-dir2", DELTA_LYA, "--rp-min", "-60.0",
                         "--np", "30", "--rej", "0.99", "--remove-same-half-plate-close-pairs",
                         "--unfold-cf", "--no-redshift-evolution"] + ZCUTS + COMMON,
}


def golden_path(name):
    return os.path.join(GOLDEN, "cor", name + ".fits.gz")


class Setup:
    """Globals + data of one case, as the picca_cf / picca_dmat drivers set them."""

    def __init__(self, argv, dmat=False):
        args = {}
        flags = set()
        k = 0
        while k < len(argv):
            if k + 1 < len(argv) and not argv[k + 1].startswith("--"):
                args[argv[k]] = argv[k + 1]
                k += 2
            else:
                flags.add(argv[k])
                k += 1
        self.r_par_min = float(args.get("--rp-min", 0.0))
        self.r_par_max = float(args.get("--rp-max", 200.0))
        self.r_trans_max = float(args.get("--rt-max", 200.0))
        self.num_bins_r_par = int(args.get("--np", 50))
        self.num_bins_r_trans = int(args.get("--nt", 50))
        self.coef = int(args.get("--coef-binning-model", 1))
        self.num_model_bins_r_par = self.num_bins_r_par * self.coef if dmat else None
        self.num_model_bins_r_trans = self.num_bins_r_trans * self.coef if dmat else None
        self.z_min_pairs = float(args["--z-min-pairs"]) if "--z-min-pairs" in args else None
        self.z_max_pairs = float(args["--z-max-pairs"]) if "--z-max-pairs" in args else None
        self.zerr_cut_deg = None
        self.zerr_cut_kms = None
        self.z_ref = 2.25
        self.alpha = 2.9
        self.alpha2 = 2.9 if dmat else None
        self.reject = float(args.get("--rej", 1.0))
        self.x_correlation = False
        self.rmu_binning = False
        self.ang_correlation = False
        self.remove_same_half_plate_close_pairs = "--remove-same-half-plate-close-pairs" in flags
        self.redshift_evolution_in_distortion_matrix = "--no-redshift-evolution" not in flags
        z_min_qso = float(args.get("--z-min-sources", 0.0))
        z_max_qso = float(args.get("--z-max-sources", 10.0))
        cosmo = constants.Cosmo(Om=0.315, verbose=False)
        rebin = int(args["--rebin-factor"]) if "--rebin-factor" in args else None
        self.data, self.num_data, z_min, _ = io.read_deltas(
            args["--in-dir"], 16, constants.ABSORBER_IGM["LYA"], 2.9, 2.25, cosmo,
            z_min_qso=z_min_qso, z_max_qso=z_max_qso, delta_attributes=ATTRIBUTES,
            rebin_factor=rebin)
        self.ang_max = utils.compute_ang_max(cosmo, self.r_trans_max, z_min)
        self.data2 = None
        if "--in-dir2" in args:
            if "--unfold-cf" in flags:
                self.x_correlation = True
            self.alpha2 = 2.9
            self.data2, _, z_min2, _ = io.read_deltas(
                args["--in-dir2"], 16, constants.ABSORBER_IGM["LYA"], 2.9, 2.25, cosmo,
                z_min_qso=z_min_qso, z_max_qso=z_max_qso, delta_attributes=ATTRIBUTES)
            self.ang_max = utils.compute_ang_max(cosmo, self.r_trans_max, z_min, z_min2)


CF_ANGL_CASE = ["--in-dir", DELTA_LYA, "--nproc", "1", "--in-attributes", ATTRIBUTES]


class AngularSetup:
    """Globals + data of picca's test_cf_angl (py/picca/tests/test_3_cor.py:206-227) as
    picca_cf_angl sets them (py/picca/bin/picca_cf_angl.py:172-233): wavelength ratio in
    [1, 1.1) for r_par, angle below 0.02 rad for r_trans, deltas read without cosmology."""

    def __init__(self):
        self.r_par_min, self.r_par_max = 1.0, 1.1
        self.r_trans_max = 0.02
        self.ang_max = 0.02
        self.num_bins_r_par = 50
        self.num_bins_r_trans = 50
        self.num_model_bins_r_par = None
        self.num_model_bins_r_trans = None
        self.z_min_pairs = self.z_max_pairs = None
        self.zerr_cut_deg = self.zerr_cut_kms = None
        self.z_ref = 2.25
        self.alpha = 2.9
        self.alpha2 = None
        self.reject = 1.0
        self.x_correlation = False
        self.rmu_binning = False
        self.ang_correlation = True
        self.remove_same_half_plate_close_pairs = False
        self.redshift_evolution_in_distortion_matrix = True
        self.data, self.num_data, _, _ = io.read_deltas(
            DELTA_LYA, 16, constants.ABSORBER_IGM["LYA"], 2.9, 2.25, None,
            delta_attributes=ATTRIBUTES)
        self.data2 = None


CAT = os.path.join(GOLDEN, "cat.fits")
XCF_COMMON = ["--in-dir", DELTA_LYA, "--drq", CAT, "--rp-min", "-60.0", "--rp-max", "+60.0",
              "--rt-max", "+60.0", "--np", "30", "--nt", "15", "--nproc", "1",
              "--z-evol-obj", "1.", "--in-attributes", ATTRIBUTES]
XCF_CASES = {
    "xcf": list(XCF_COMMON),
    "xcf_zcuts": XCF_COMMON + ["--z-min-pairs", "2.25", "--z-max-pairs", "2.3",
                               "--z-min-sources", "2.3", "--z-max-sources", "2.5"],
}


def _remove_mean_lambda_obs(data, z_min, z_max, lambda_abs):
    """py/picca/bin/picca_xcf.py:436-468 with py/picca/prep_del.py:13-60, restated."""
    step = None
    for hp in data:
        for d in data[hp]:
            s = np.asarray([d.log_lambda[k] - d.log_lambda[k - 1]
                            for k in range(1, d.log_lambda.size)]).min()
            step = s if step is None else min(s, step)
    ll_min = np.log10((z_min + 1.0) * lambda_abs) - step / 2.0
    ll_max = np.log10((z_max + 1.0) * lambda_abs) + step / 2.0
    nbins = int((ll_max - ll_min) / step) + 1
    stack = np.zeros(nbins)
    weight = np.zeros(nbins)
    for hp in sorted(data):
        for d in data[hp]:
            bins = ((d.log_lambda - ll_min) / step + 0.5).astype(int)
            r = np.bincount(bins, weights=d.delta * d.weights)
            stack[:len(r)] += r
            r = np.bincount(bins, weights=d.weights)
            weight[:len(r)] += r
    w = weight > 0
    stack[w] /= weight[w]
    for hp in data:
        for d in data[hp]:
            bins = ((d.log_lambda - ll_min) / step + 0.5).astype(int)
            d.delta -= stack[bins]


class XcfSetup:
    """Globals, deltas and quasars of one picca_xcf test command, read as picca_xcf
    does (py/picca/bin/picca_xcf.py:337-490)."""

    def __init__(self, argv):
        args = {}
        k = 0
        while k < len(argv):
            args[argv[k]] = argv[k + 1]
            k += 2
        self.r_par_min = float(args["--rp-min"])
        self.r_par_max = float(args["--rp-max"])
        self.r_trans_max = float(args["--rt-max"])
        self.num_bins_r_par = int(args["--np"])
        self.num_bins_r_trans = int(args["--nt"])
        self.num_model_bins_r_par = None
        self.num_model_bins_r_trans = None
        self.z_min_pairs = float(args["--z-min-pairs"]) if "--z-min-pairs" in args else None
        self.z_max_pairs = float(args["--z-max-pairs"]) if "--z-max-pairs" in args else None
        self.zerr_cut_deg = self.zerr_cut_kms = None
        self.z_ref = 2.25
        self.alpha = 2.9
        self.alpha2 = 1.0
        self.x_correlation = True
        self.rmu_binning = False
        self.ang_correlation = False
        self.remove_same_half_plate_close_pairs = False
        self.redshift_evolution_in_distortion_matrix = True
        cosmo = constants.Cosmo(Om=0.315, verbose=False)
        self.objs, z_min2 = io.read_objects(CAT, 16, 0, 10, float(args["--z-evol-obj"]), 2.25,
                                            cosmo)
        self.data, self.num_data, z_min, z_max = io.read_deltas(
            DELTA_LYA, 16, constants.ABSORBER_IGM["LYA"], 2.9, 2.25, cosmo,
            z_min_qso=float(args.get("--z-min-sources", 0.0)),
            z_max_qso=float(args.get("--z-max-sources", 10.0)), delta_attributes=ATTRIBUTES)
        _remove_mean_lambda_obs(self.data, z_min, z_max, constants.ABSORBER_IGM["LYA"])
        self.ang_max = utils.compute_ang_max(cosmo, self.r_trans_max, z_min, z_min2)


XDMAT_CASES = {
    "xdmat": XCF_COMMON + ["--rej", "0.99", "--no-redshift-evolution"],
    "xdmat_zcuts": XCF_COMMON + ["--rej", "0.99", "--z-min-pairs", "2.25",
                                 "--z-max-pairs", "2.3", "--z-min-sources", "2.3",
                                 "--z-max-sources", "2.5", "--no-redshift-evolution"],
}


class XdmatSetup(XcfSetup):
    """Globals, deltas and quasars of one picca_xdmat test command
    (py/picca/bin/picca_xdmat.py:338-459; the mean-delta removal of XcfSetup does not
    change a distortion matrix, which reads no delta values)."""

    def __init__(self, argv):
        argv = list(argv)
        self.redshift_evolution = "--no-redshift-evolution" not in argv
        argv = [a for a in argv if a != "--no-redshift-evolution"]
        opts = dict(zip(argv[0::2], argv[1::2]))
        super().__init__(argv)
        self.reject = float(opts["--rej"])
        self.num_model_bins_r_par = self.num_bins_r_par
        self.num_model_bins_r_trans = self.num_bins_r_trans
        self.alpha2 = float(opts["--z-evol-obj"])
        self.redshift_evolution_in_distortion_matrix = self.redshift_evolution


def reduce_rows(rows):
    """The parent-process reduction of py/picca/bin/picca_cf.py:466-484."""
    rows = np.array(rows)
    we = rows[:, 0, :]
    sw = we.sum(axis=0)
    w = sw > 0
    out = {}
    for name, k in (("RP", 2), ("RT", 3), ("Z", 4)):
        v = (rows[:, k, :] * we).sum(axis=0)
        v[w] /= sw[w]
        out[name] = v
    out["NB"] = rows[:, 5, :].astype(np.int64).sum(axis=0)
    out["WE"] = we
    out["DA"] = rows[:, 1, :]
    return out


def golden_columns(name):
    g = FITS(golden_path(name))
    out = {}
    for hdu in g.hdus[1:]:
        for col in hdu.get_colnames():
            out[col] = hdu[col][:]
    return out


def compare_to_golden(result, golden, rtol=1e-5, atol=1e-8, exact=("NB", "HEALPID")):
    """compare_fits of py/picca/tests/test_helpers.py:45-112, column by column.

    Returns {column: max relative difference}; raises AssertionError on failure.
    """
    report = {}
    for col, ref in golden.items():
        if col not in result:
            continue
        mine = np.asarray(result[col])
        assert mine.shape == ref.shape, (col, mine.shape, ref.shape)
        if col in exact:
            assert np.array_equal(mine, ref), col
            report[col] = 0.0
            continue
        assert np.allclose(mine, ref, rtol=rtol, atol=atol), col
        denom = np.where(ref != 0, np.abs(ref), 1.0)
        report[col] = float(np.max(np.abs(mine - ref) / denom)) if ref.size else 0.0
    return report
