"""Shared layer of the command-line programs in ``picca_amd.bin``.

The programs keep picca's command lines (flag names and defaults) and its FITS products
(extension names, column names, header keywords), because other tools read them.  What
they do in between is assembled here from a few tables and steps:

* ``FLAGS``: every flag of every program, with its value kind and default; ``PROGRAMS``
  lists which flags a program takes and the defaults it changes.  ``parse`` builds the
  argparse parser from the two.
* ``Settings``: the module globals of ``picca_amd.cf`` / ``picca_amd.xcf`` (the
  reference's ``py/picca/cf.py:28-79`` and ``py/picca/xcf.py:30-66``) derived from the
  parsed options, applied in one call.
* ``load_deltas`` / ``load_objects``: the reading of the delta fields and quasar
  catalogues with their cosmology, angle limit and optional shuffling.
* ``healpix_rows`` and ``seeded_lists``: the two ways the reference splits a run over
  healpix pixels (one task per pixel for the correlations, ``py/picca/bin/picca_cf.py:
  452-458``; round-robin lists seeding numpy's legacy stream with their first pixel for
  the matrices, ``py/picca/bin/picca_dmat.py:21-38,469-485``).  Here the correlations run
  every pixel as one row of a single GPU launch; the matrices run list after list.
* ``combine_rows``, ``add_parts``, ``normalise_matrix``: the reductions
  (``py/picca/bin/picca_cf.py:466-484``, ``py/picca/bin/picca_dmat.py:493-519``).
* ``HEADER_KEYS`` and ``Product``: the FITS layouts, written by ``write_products``.
"""
import argparse
import sys
import time
from multiprocessing import cpu_count

import numpy as np

from picca_amd import constants, io, utils
from picca_amd.fitsio_lite import FITSWriter

# ------------------------------------------------------------------------------------------
# flags
# ------------------------------------------------------------------------------------------

REQUIRED = "required"
SWITCH = "switch"

# option -> (value kind, default, what it sets).  Kinds: float, int, str, SWITCH (a
# store_true flag) and "strs" (zero or more strings).
FLAGS = {
    "--out": (str, REQUIRED, "FITS file to write"),
    "--in-dir": (str, REQUIRED, "directory of the delta files of the first field"),
    "--in-dir2": (str, None, "directory of the delta files of a second field"),
    "--drq": (str, REQUIRED, "quasar catalogue (DRQ FITS)"),
    "--mode": (str, "sdss", "layout of the quasar catalogue"),
    "--rp-min": (float, 0.0, "lower r_par limit (Mpc/h)"),
    "--rp-max": (float, 200.0, "upper r_par limit (Mpc/h)"),
    "--rt-max": (float, 200.0, "upper r_trans limit (Mpc/h)"),
    "--np": (int, 50, "r_par bins"),
    "--nt": (int, 50, "r_trans bins"),
    "--rmu-binning": (SWITCH, False, "bin in (mu, r) instead of (r_par, r_trans)"),
    "--coef-binning-model": (int, 1, "model bins per data bin along each axis"),
    "--zerr-cut-deg": (float, None, "angle (deg) below which the redshift-error cut acts"),
    "--zerr-cut-kms": (float, None, "velocity (km/s) of the redshift-error cut"),
    "--z-min-sources": (float, 0.0, "lowest quasar redshift of the forests read"),
    "--z-max-sources": (float, 10.0, "highest quasar redshift of the forests read"),
    "--z-min-pairs": (float, None, "lowest mean redshift of a pixel pair"),
    "--z-max-pairs": (float, None, "highest mean redshift of a pixel pair"),
    "--z-min-obj": (float, 0, "lowest redshift of the catalogue objects"),
    "--z-max-obj": (float, 10, "highest redshift of the catalogue objects"),
    "--lambda-abs": (str, "LYA", "absorber of the first field (constants.ABSORBER_IGM)"),
    "--lambda-abs2": (str, None, "absorber of the second field"),
    "--z-ref": (float, 2.25, "reference redshift of the weight evolution"),
    "--z-evol": (float, 2.9, "bias evolution exponent of the first field"),
    "--z-evol2": (float, 2.9, "bias evolution exponent of the second field"),
    "--z-evol-del": (float, 2.9, "bias evolution exponent of the deltas"),
    "--z-evol-obj": (float, 1.44, "bias evolution exponent of the objects"),
    "--metal-alpha": (float, 1.0, "bias evolution exponent of the metals"),
    "--fid-Om": (float, 0.315, "fiducial Omega_m"),
    "--fid-Or": (float, 0.0, "fiducial Omega_r"),
    "--fid-Ok": (float, 0.0, "fiducial Omega_k"),
    "--fid-wl": (float, -1.0, "fiducial dark-energy equation of state"),
    "--distance-template": (str, None, "camb distance template (refused by Cosmo)"),
    "--no-project": (SWITCH, False, "keep the continuum-fit modes in the deltas"),
    "--no-remove-mean-lambda-obs": (SWITCH, False,
                                    "keep the mean delta per observed wavelength"),
    "--remove-same-half-plate-close-pairs": (SWITCH, False,
                                             "drop first-r_par-bin pairs of one half plate"),
    "--rej": (float, 1.0, "fraction of neighbour pairs left out at random"),
    "--nside": (int, 16, "healpix resolution of the sky split"),
    "--nproc": (int, None, "worker count (readers; seeds the matrix draws)"),
    "--nspec": (int, None, "read at most this many forests"),
    "--unfold-cf": (SWITCH, False, "signed r_par for two fields"),
    "--shuffle-distrib-forest-seed": (int, None, "seed of a shuffle of the forests on the sky"),
    "--shuffle-distrib-obj-seed": (int, None, "seed of a shuffle of the objects on the sky"),
    "--rebin-factor": (int, None, "merge this many pixels of every forest"),
    "--in-attributes": (str, None, "delta attributes file (continuum-fit order)"),
    "--no-redshift-evolution": (SWITCH, False, "no redshift evolution in the matrix"),
    "--wr-min": (float, 1.0, "lowest wavelength ratio"),
    "--wr-max": (float, 1.1, "highest wavelength ratio"),
    "--ang-max": (float, 0.02, "largest angle (rad)"),
    "--abs-igm": ("strs", [], "metal absorbers of the first field"),
    "--abs-igm2": ("strs", [], "metal absorbers of the second field"),
    "--fast-metals": (SWITCH, False, "only LYA x Si metals and CIV(eff) x CIV(eff)"),
    "--obj-name": (str, "QSO", "name of the objects in the column names"),
    "--max-diagram": (int, 3, "highest Wick diagram computed"),
    "--cf1d": (str, REQUIRED, "1-D correlation file of the first field"),
    "--cf1d2": (str, None, "1-D correlation file of the second field"),
    "--cf": (str, None, "3-D correlation file of the first field"),
    "--cf2": (str, None, "3-D correlation file of the second field"),
    "--cf12": (str, None, "3-D cross-correlation file of the two fields"),
    "--data": (str, REQUIRED, "correlation file to export"),
    "--dmat": (str, None, "distortion-matrix file (identity when absent)"),
    "--multipoles": (SWITCH, False, "export multipoles"),
    "--lmax-model": (int, 6, "highest model multipole"),
    "--lmax-data": (int, 4, "highest data multipole"),
    "--cov": (str, None, "covariance file to use instead of the subsampling one"),
    "--cor": (str, None, "correlation-matrix file scaled by the subsampling variances"),
    "--remove-shuffled-correlation": (str, None, "correlation of shuffled forests to subtract"),
    "--num-boot-cov": (int, 0, "bootstrap realisations (0: subsampling covariance)"),
    "--do-not-smooth-cov": (SWITCH, False, "keep the raw covariance"),
    "--smooth-per-r-par": (SWITCH, False, "smooth the covariance per r_par bin"),
    "--blind-corr-type": (str, None, "correlation type of DESI's blinding"),
    "--device": (int, 0, "HIP device of the engine"),
}

CATALOGUE_MODES = ("sdss", "desi", "desi_mocks", "desi_healpix")
BLIND_TYPES = ("lyaxlya", "lyaxlyb", "qsoxlya", "qsoxlyb", "lyaxqso", "lybxqso")

_FIELD = ("--z-min-sources --z-max-sources --lambda-abs --nside --nproc --nspec "
          "--in-attributes --device").split()
_COSMO = "--fid-Om --fid-Or --fid-Ok --fid-wl".split()
_RANGE = "--rp-min --rp-max --rt-max --np --nt".split()
_PAIR_CUTS = "--zerr-cut-deg --zerr-cut-kms --z-min-pairs --z-max-pairs".split()

# program -> (summary, flags, changed defaults)
PROGRAMS = {
    "picca_cf": (
        "3D auto- or cross-correlation of delta fields",
        ["--out", "--in-dir", "--in-dir2"] + _RANGE + ["--rmu-binning"] + _PAIR_CUTS
        + _FIELD + ["--lambda-abs2", "--z-ref", "--z-evol", "--z-evol2"] + _COSMO
        + ["--distance-template", "--no-project", "--remove-same-half-plate-close-pairs",
           "--unfold-cf", "--shuffle-distrib-forest-seed", "--rebin-factor"], {}),
    "picca_dmat": (
        "distortion matrix of the correlation of delta fields",
        ["--out", "--in-dir", "--in-dir2"] + _RANGE + ["--rmu-binning", "--coef-binning-model"]
        + _PAIR_CUTS + _FIELD + ["--lambda-abs2", "--z-evol", "--z-evol2"] + _COSMO
        + ["--distance-template", "--remove-same-half-plate-close-pairs", "--rej",
           "--unfold-cf", "--rebin-factor", "--no-redshift-evolution"], {}),
    "picca_xcf": (
        "cross-correlation of a delta field with an object catalogue",
        ["--out", "--in-dir", "--drq", "--mode"] + _RANGE + ["--rmu-binning", "--z-min-obj",
                                                            "--z-max-obj"]
        + _PAIR_CUTS + _FIELD + ["--z-ref", "--z-evol-del", "--z-evol-obj"] + _COSMO
        + ["--distance-template", "--no-project", "--no-remove-mean-lambda-obs",
           "--shuffle-distrib-obj-seed", "--shuffle-distrib-forest-seed", "--rebin-factor"],
        {"--rp-min": -200.0, "--np": 100}),
    "picca_xdmat": (
        "distortion matrix of the delta x object cross-correlation",
        ["--out", "--in-dir", "--drq", "--mode"] + _RANGE
        + ["--rmu-binning", "--coef-binning-model", "--z-min-obj", "--z-max-obj"]
        + _PAIR_CUTS + _FIELD + ["--z-evol-del", "--z-evol-obj"] + _COSMO
        + ["--distance-template", "--rej", "--rebin-factor", "--no-redshift-evolution"],
        {"--rp-min": -200.0, "--np": 100}),
    "picca_cf_angl": (
        "correlation of delta fields in wavelength ratio and angle",
        ["--out", "--in-dir", "--in-dir2", "--wr-min", "--wr-max", "--ang-max", "--np",
         "--nt"] + _FIELD + ["--lambda-abs2", "--z-ref", "--z-evol", "--z-evol2",
                             "--no-project", "--remove-same-half-plate-close-pairs"], {}),
    "picca_metal_dmat": (
        "metal distortion matrices of the correlation of delta fields",
        ["--out", "--in-dir", "--in-dir2"] + _RANGE + ["--coef-binning-model"] + _FIELD
        + ["--lambda-abs2", "--abs-igm", "--abs-igm2", "--z-ref", "--z-evol", "--z-evol2",
           "--metal-alpha"] + _COSMO
        + ["--remove-same-half-plate-close-pairs", "--rej", "--unfold-cf", "--rebin-factor",
           "--fast-metals"], {}),
    "picca_metal_xdmat": (
        "metal distortion matrices of the delta x object cross-correlation",
        ["--out", "--in-dir", "--drq", "--mode"] + _RANGE
        + ["--coef-binning-model", "--z-min-obj", "--z-max-obj"] + _FIELD
        + ["--obj-name", "--abs-igm", "--z-ref", "--z-evol-del", "--z-evol-obj",
           "--metal-alpha"] + _COSMO + ["--rej", "--rebin-factor"],
        {"--rp-min": -200.0, "--np": 100, "--abs-igm": None}),
    "picca_wick": (
        "Wick covariance (diagrams T1-T3) of the auto-correlation",
        ["--out", "--in-dir", "--in-dir2"] + _RANGE + _FIELD
        + ["--lambda-abs2", "--z-ref", "--z-evol", "--z-evol2"] + _COSMO
        + ["--no-project", "--max-diagram", "--cf1d", "--cf1d2", "--cf", "--cf2", "--cf12",
           "--unfold-cf", "--rej", "--rebin-factor"], {}),
    "picca_export": (
        "export of a correlation with its covariance and distortion matrix",
        ["--data", "--out", "--dmat", "--multipoles", "--lmax-model", "--lmax-data", "--cov",
         "--cor", "--remove-shuffled-correlation", "--num-boot-cov", "--do-not-smooth-cov",
         "--smooth-per-r-par", "--blind-corr-type"], {}),
}


def parse(program, argv):
    """Parsed options of `program` from `argv` (the flag table and the program's list)."""
    summary, flags, changed = PROGRAMS[program]
    parser = argparse.ArgumentParser(prog=program, description=summary,
                                     formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    for option in flags:
        kind, default, note = FLAGS[option]
        default = changed.get(option, default)
        if kind == SWITCH:
            parser.add_argument(option, action="store_true", help=note)
            continue
        extra = {}
        if option == "--mode":
            extra["choices"] = list(CATALOGUE_MODES)
        if option == "--blind-corr-type":
            extra["choices"] = list(BLIND_TYPES)
        if kind == "strs":
            extra["nargs"] = "*"
            kind = str
        if default is REQUIRED:
            parser.add_argument(option, type=kind, default=None, required=True, help=note,
                                **extra)
        else:
            parser.add_argument(option, type=None if option == "--blind-corr-type" else kind,
                                default=default, required=False, help=note, **extra)
    return parser.parse_args(argv)


# ------------------------------------------------------------------------------------------
# progress
# ------------------------------------------------------------------------------------------

class Clock:
    """Phase timer printing `[program] phase: seconds` lines to stdout."""

    def __init__(self, program):
        self.program = program
        self.start = self.mark = time.time()

    def lap(self, phase):
        now = time.time()
        print("[%s] %s: %.2f s" % (self.program, phase, now - self.mark), flush=True)
        self.mark = now

    def total(self):
        print("[%s] total: %.2f s" % (self.program, time.time() - self.start), flush=True)


def say(*parts):
    print(*parts, flush=True)


# ------------------------------------------------------------------------------------------
# module settings
# ------------------------------------------------------------------------------------------

# the module globals a run sets, with their values before any run
# (py/picca/tests/test_3_cor.py:39-85 resets the same names between tests)
_NEUTRAL = dict(
    num_bins_r_par=None, num_bins_r_trans=None, num_model_bins_r_trans=None,
    num_model_bins_r_par=None, r_par_max=None, r_par_min=None, z_min_pairs=None,
    z_max_pairs=None, r_trans_max=None, ang_max=None, nside=None, zerr_cut_deg=None,
    zerr_cut_kms=None, counter=None, num_data=None, num_data2=None, z_ref=None, alpha=None,
    alpha2=None, alpha_abs=None, lambda_abs=None, lambda_abs2=None, data=None, data2=None,
    cosmo=None, reject=None, lock=None, x_correlation=False, rmu_binning=False,
    ang_correlation=False, remove_same_half_plate_close_pairs=False,
    redshift_evolution_in_distortion_matrix=True)


class Settings(dict):
    """Module globals of one run, applied with `apply` after the neutral values."""

    def apply(self, module, reset=True):
        if reset:
            for name, value in _NEUTRAL.items():
                if hasattr(module, name):
                    setattr(module, name, value)
        for name, value in self.items():
            setattr(module, name, value)
        return module


def reset_module(module):
    """Puts every run setting of `module` back to its value before any run."""
    Settings().apply(module)


def pair_cut_settings(opts):
    """The redshift-pair and redshift-error cuts; the two zerr options go together
    (py/picca/bin/picca_cf.py:362-365 raises the same ValueError)."""
    if (opts.zerr_cut_deg is None) != (opts.zerr_cut_kms is None):
        raise ValueError("--zerr-cut-deg and --zerr-cut-kms must be given together")
    return dict(z_min_pairs=opts.z_min_pairs, z_max_pairs=opts.z_max_pairs,
                zerr_cut_deg=opts.zerr_cut_deg, zerr_cut_kms=opts.zerr_cut_kms)


def binning_settings(opts, model_coef=None, signed=False, data_coef=1):
    """Binning globals.  In (mu, r) binning r_par runs over [0, 1] ([-1, 1] when signed,
    for the cross-correlation with objects); `model_coef` adds the model binning of the
    matrices, `data_coef` multiplies the data binning (the metal matrices bin their data
    as finely as their model)."""
    r_min, r_max = opts.rp_min, opts.rp_max
    if getattr(opts, "rmu_binning", False):
        r_min, r_max = (-1.0 if signed else 0.0), 1.0
        opts.rp_min, opts.rp_max = r_min, r_max
    out = dict(r_par_min=r_min, r_par_max=r_max, r_trans_max=opts.rt_max,
               num_bins_r_par=opts.np * data_coef, num_bins_r_trans=opts.nt * data_coef,
               rmu_binning=bool(getattr(opts, "rmu_binning", False)), nside=opts.nside)
    if model_coef is not None:
        out.update(num_model_bins_r_par=opts.np * model_coef,
                   num_model_bins_r_trans=opts.nt * model_coef)
    return out


def process_count(opts):
    """--nproc, or the reference's default of half the host's cores."""
    if opts.nproc is None:
        opts.nproc = max(1, cpu_count() // 2)
    return opts.nproc


# ------------------------------------------------------------------------------------------
# reading
# ------------------------------------------------------------------------------------------

def cosmology(opts, blinding=None, template=True):
    kwargs = dict(Om=opts.fid_Om, Or=opts.fid_Or, Ok=opts.fid_Ok, wl=opts.fid_wl)
    if blinding is not None:
        kwargs["blinding"] = blinding
    if template:
        kwargs["template"] = getattr(opts, "distance_template", None)
    return constants.Cosmo(**kwargs)


def load_deltas(opts, folder, absorber, alpha, cosmo, project=True, nproc=True,
                rebin=True):
    """One delta field as a {healpix: [Delta]} dict, with its forest count and the
    smallest and largest pixel redshifts (``io.read_deltas``)."""
    data, count, z_low, z_high = io.read_deltas(
        folder, opts.nside, constants.ABSORBER_IGM[absorber], alpha, opts_z_ref(opts), cosmo,
        max_num_spec=opts.nspec, no_project=not project,
        nproc=opts.nproc if nproc else None,
        rebin_factor=getattr(opts, "rebin_factor", None) if rebin else None,
        z_min_qso=opts.z_min_sources, z_max_qso=opts.z_max_sources,
        delta_attributes=opts.in_attributes)
    say("%s: %d forests in %d healpix pixels" % (folder, count, len(data)))
    return data, count, z_low, z_high


def opts_z_ref(opts):
    return getattr(opts, "z_ref", None) or 2.25


class Fields:
    """The delta fields of an auto/cross run and their angle limit."""

    def __init__(self, opts, cosmo, alpha, alpha2, project=True, nproc=True, rebin=True,
                 angular=False):
        self.opts = opts
        self.first = load_deltas(opts, opts.in_dir, opts.lambda_abs, alpha, cosmo, project,
                                 nproc, rebin)
        self.second = None
        self.absorber2 = None
        self.cross = False
        if opts.in_dir2 or opts.lambda_abs2:
            # a second field, the first one again under another absorber, or both
            self.absorber2 = opts.lambda_abs2 or opts.lambda_abs
            self.cross = bool(opts.lambda_abs2) or angular or bool(
                getattr(opts, "unfold_cf", False))
            self.second = load_deltas(opts, opts.in_dir2 or opts.in_dir, self.absorber2,
                                      alpha2, cosmo, project, nproc, rebin)
        self.ang_max = None
        if cosmo is not None:
            lows = [self.first[2]] + ([self.second[2]] if self.second else [])
            self.ang_max = utils.compute_ang_max(cosmo, opts.rt_max, *lows)

    def settings(self, alpha2):
        out = dict(data=self.first[0], num_data=self.first[1], ang_max=self.ang_max,
                   x_correlation=self.cross)
        if self.second is not None:
            out.update(data2=self.second[0], num_data2=self.second[1], alpha2=alpha2,
                       lambda_abs2=constants.ABSORBER_IGM[self.absorber2])
        return out


def object_redshift_limits(opts, cosmo, z_low, z_high):
    """Catalogue redshift range; an unset end follows the deltas' range widened by the
    r_par limits."""
    if opts.z_min_obj is None:
        r_low = max(0.0, float(cosmo.get_r_comov(z_low)) + opts.rp_min)
        opts.z_min_obj = float(cosmo.distance_to_redshift(r_low))
    if opts.z_max_obj is None:
        r_high = max(0.0, float(cosmo.get_r_comov(z_high)) + opts.rp_max)
        opts.z_max_obj = float(cosmo.distance_to_redshift(r_high))
    return opts.z_min_obj, opts.z_max_obj


def load_objects(opts, cosmo):
    objs, z_low = io.read_objects(opts.drq, opts.nside, opts.z_min_obj, opts.z_max_obj,
                                  opts.z_evol_obj, opts_z_ref(opts), cosmo, mode=opts.mode)
    say("%s: %d objects" % (opts.drq, sum(len(v) for v in objs.values())))
    return objs, z_low


# ------------------------------------------------------------------------------------------
# running over healpix pixels
# ------------------------------------------------------------------------------------------

def healpix_rows(module, data):
    """All pixels of `data` as rows of one launch: (sorted healpix ids, [rows, 6, nb])."""
    ids = np.array(sorted(data))
    return ids, np.asarray(module.compute_xi_rows(ids))


def process_lists(data, count):
    """The reference's split of the sorted pixels into `count` round-robin lists, in the
    order its pool receives them (py/picca/bin/picca_dmat.py:471-476)."""
    lists = {}
    for position, healpix in enumerate(sorted(data)):
        lists.setdefault(position % count, []).append(healpix)
    return sorted(lists.values())


def seeded_lists(module, lists, compute):
    """Runs `compute(list)` per list after filling the neighbours and seeding numpy's legacy
    generator with the list's first pixel: the rejection draws of the matrices then come
    out as in the reference (py/picca/bin/picca_dmat.py:21-38)."""
    parts = []
    for healpixs in lists:
        module.fill_neighs(healpixs)
        np.random.seed(healpixs[0])
        parts.append(compute(healpixs))
    return parts


# ------------------------------------------------------------------------------------------
# reductions
# ------------------------------------------------------------------------------------------

def combine_rows(rows):
    """Per-pixel rows [n, 6, nb] (weights, xi, r_par, r_trans, z, counts, each row already
    normalised by its weights) -> the weight-averaged RP, RT, Z and the summed NB of the
    ATTRI table, plus the WE and DA columns of COR (py/picca/bin/picca_cf.py:466-484)."""
    weights = rows[:, 0, :]
    total = weights.sum(axis=0)
    filled = total > 0
    means = []
    for k in (2, 3, 4):
        column = (rows[:, k, :] * weights).sum(axis=0)
        column[filled] /= total[filled]
        means.append(column)
    counts = rows[:, 5, :].astype(np.int64).sum(axis=0)
    return dict(RP=means[0], RT=means[1], Z=means[2], NB=counts, WE=weights,
                DA=rows[:, 1, :])


def add_parts(parts):
    """Element-wise sums of the per-list result tuples."""
    return [np.array(column).sum(axis=0) for column in zip(*parts)]


def normalise_matrix(weights_dmat, dmat, r_par, r_trans, z, weights):
    """Effective coordinates over their weights, matrix rows over the data weights; returns
    the mean effective redshift of the filled model bins."""
    filled = weights > 0
    for column in (r_par, r_trans, z):
        column[filled] /= weights[filled]
    mean_z = float(np.mean(z[filled])) if filled.any() else float("nan")
    rows = weights_dmat > 0
    dmat[rows] /= weights_dmat[rows, None]
    return mean_z


def redshift_rescale(dmat, z, mean_z, z_ref, alpha, alpha2):
    """The matrix at the mean redshift: every model bin takes the mean effective redshift
    and the matrix is scaled by ((1 + z_ref) / (1 + mean_z))^(alpha + alpha2 - 2)."""
    z[:] = mean_z
    dmat *= ((1 + z_ref) / (1 + mean_z)) ** ((alpha - 1) + (alpha2 - 1))


# ------------------------------------------------------------------------------------------
# FITS products
# ------------------------------------------------------------------------------------------

# header keyword -> card comment (the keywords are the products' interface)
HEADER_KEYS = {
    "RPMIN": "lower r_par limit",
    "RPMAX": "upper r_par limit",
    "RTMAX": "upper r_trans limit",
    "NP": "r_par bins",
    "NT": "r_trans bins",
    "COEFMOD": "model bins per data bin",
    "ZMIN": "lowest pair redshift",
    "ZMAX": "highest pair redshift",
    "REJ": "fraction of pairs left out",
    "NPALL": "neighbour pairs",
    "NPUSED": "pairs kept",
    "NSIDE": "healpix nside",
    "OMEGAM": "fiducial Omega_m",
    "OMEGAR": "fiducial Omega_r",
    "OMEGAK": "fiducial Omega_k",
    "WL": "fiducial dark-energy w",
    "BLINDING": "blinding strategy",
    "RMU_BIN": "binned in (mu, r)",
    "ALPHAMET": "metal bias evolution exponent",
    "HLPXSCHM": "healpix scheme",
    "NSAMPLES": "subsamples of the covariance",
    "MLTPOLE": "multipoles",
}

# column -> (unit, comment)
COLUMNS = {
    "RP": ("h^-1 Mpc", "mean r_par"),
    "RT": ("h^-1 Mpc", "mean r_trans"),
    "Z": ("", "mean pair redshift"),
    "NB": ("", "pair count"),
    "HEALPID": ("", "healpix pixel"),
    "WE": ("", "weight sum"),
    "DA": ("", "correlation"),
    "WDM": ("", "data weight sum"),
    "DM": ("", "distortion matrix"),
    "CO": ("", "covariance"),
    "NPALL": ("", "neighbour pairs"),
    "NPUSED": ("", "pairs kept"),
    "ABS_IGM": ("", "absorber pair"),
    "WALL": ("", "weight sum"),
    "DMRP": ("h^-1 Mpc", "model r_par"),
    "DMRT": ("h^-1 Mpc", "model r_trans"),
    "DMZ": ("", "model redshift"),
}


def header(**values):
    """Header cards from keyword=value pairs, with the comments of HEADER_KEYS."""
    return [{"name": key, "value": value, "comment": HEADER_KEYS.get(key, "")}
            for key, value in values.items()]


class Product:
    """One BINTABLE extension: ordered columns and header cards."""

    def __init__(self, extname, cards=None):
        self.extname = extname
        self.cards = cards or []
        self.names, self.values, self.units, self.notes = [], [], [], []

    def add(self, name, value, unit=None, note=None, like=None):
        base = COLUMNS.get(like or name, ("", ""))
        self.names.append(name)
        self.values.append(value)
        self.units.append(base[0] if unit is None else unit)
        self.notes.append(base[1] if note is None else note)
        return self


def write_products(path, products):
    out = FITSWriter(path)
    for prod in products:
        out.write(prod.values, names=prod.names, header=prod.cards, extname=prod.extname,
                  comment=prod.notes, units=prod.units)
    out.close()
    say("wrote %s" % path)


def blinded(name, blinding):
    """Column names of blinded products carry a _BLIND suffix (DA_BLIND, DM_BLIND)."""
    return name if blinding == "none" else name + "_BLIND"


def cosmo_cards(opts):
    return header(OMEGAM=opts.fid_Om, OMEGAR=opts.fid_Or, OMEGAK=opts.fid_Ok, WL=opts.fid_wl)


def correlation_products(path, table, healpix, cards, blinding, rmu=False, units=None):
    """ATTRI (RP, RT, Z, NB) and COR (HEALPID, WE, DA or DA_BLIND) of a correlation."""
    attri = Product("ATTRI", cards)
    if rmu:
        attri.add("RP", table["RP"], unit="", note="mean mu")
        attri.add("RT", table["RT"], note="mean r")
    else:
        attri.add("RP", table["RP"], unit=units[0] if units else None)
        attri.add("RT", table["RT"], unit=units[1] if units else None)
    attri.add("Z", table["Z"]).add("NB", table["NB"])
    cor = Product("COR", header(HLPXSCHM="RING"))
    cor.add("HEALPID", healpix).add("WE", table["WE"])
    cor.add(blinded("DA", blinding), table["DA"], like="DA")
    write_products(path, [attri, cor])


def matrix_products(path, parts, cards, blinding):
    """DMAT (WDM, DM or DM_BLIND) and ATTRI (RP, RT, Z) of a distortion matrix."""
    weights_dmat, dmat, r_par, r_trans, z = parts
    mat = Product("DMAT", cards).add("WDM", weights_dmat)
    mat.add(blinded("DM", blinding), dmat, like="DM")
    attri = Product("ATTRI").add("RP", r_par).add("RT", r_trans).add("Z", z)
    write_products(path, [mat, attri])


def metal_products(path, cards, matrices, blinding, prefix=""):
    """ATTRI (NPALL, NPUSED, ABS_IGM per absorber pair) and MDMAT (RP_, RT_, Z_, DM_ or
    DM_BLIND_, WDM_ columns per absorber pair, named after the pair and `prefix`)."""
    names = [m["name"] for m in matrices]
    width = max(len(n) for n in names)
    attri = Product("ATTRI", cards)
    attri.add("NPALL", np.array([m["npall"] for m in matrices]))
    attri.add("NPUSED", np.array([m["npused"] for m in matrices]))
    attri.add("ABS_IGM", np.array(names, dtype="S%d" % width))
    mdmat = Product("MDMAT")
    dm = "DM_" if blinding == "none" else "DM_BLIND_"
    for m in matrices:
        tag = prefix + m["name"]
        mdmat.add("RP_" + tag, m["r_par"], like="RP").add("RT_" + tag, m["r_trans"], like="RT")
        mdmat.add("Z_" + tag, m["z"], like="Z")
        mdmat.add(dm + tag, m["dmat"], like="DM").add("WDM_" + tag, m["weights_dmat"],
                                                       like="WDM")
    write_products(path, [attri, mdmat])


def main_entry(main):
    """`python -m picca_amd.bin.X args...`"""
    main(sys.argv[1:])
