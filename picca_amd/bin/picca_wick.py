#!/usr/bin/env python3
"""Wick covariance of the auto-correlation of forests on the MI355X engine.

Command-line mirror of ``py/picca/bin/picca_wick.py``: same flags and defaults, module
globals (:325-336), delta reading (:352-387), the nearest-neighbour interpolations of the
1-D correlation and variance of the ``--cf1d`` files (:392-419), the 3-D correlations of
``--cf``/``--cf2``/``--cf12`` (read, used by diagrams 4-5 only), the split of the healpix list
into ``--nproc`` process lists, each of which fills its neighbours and seeds numpy's legacy
stream with its first healpix (``calc_wick_terms``, :20-37), the merge and normalisation
(:523-560) and the ``COV`` output (:562-628).  Diagrams 1-3 run on the GPU
(``cf.compute_wick_terms``); ``--max-diagram`` above 3 is refused.
"""
import argparse
import sys
import time
from multiprocessing import cpu_count

import numpy as np
from scipy.interpolate import interp1d

from picca_amd import cf, constants, io, utils
from picca_amd.bin.picca_cf import reset_globals
from picca_amd.fitsio_lite import FITS, FITSWriter
from picca_amd.utils import userprint


def parse_args(cmdargs=None):
    parser = argparse.ArgumentParser(
        formatter_class=argparse.ArgumentDefaultsHelpFormatter,
        description="Compute the wick covariance for the auto-correlation of forests")
    parser.add_argument("--out", type=str, default=None, required=True, help="Output file name")
    parser.add_argument("--in-dir", type=str, default=None, required=True,
                        help="Directory to delta files")
    parser.add_argument("--in-dir2", type=str, default=None, required=False,
                        help="Directory to 2nd delta files")
    parser.add_argument("--rp-min", type=float, default=0.0, required=False,
                        help="Min r-parallel [h^-1 Mpc]")
    parser.add_argument("--rp-max", type=float, default=200.0, required=False,
                        help="Max r-parallel [h^-1 Mpc]")
    parser.add_argument("--rt-max", type=float, default=200.0, required=False,
                        help="Max r-transverse [h^-1 Mpc]")
    parser.add_argument("--np", type=int, default=50, required=False,
                        help="Number of r-parallel bins")
    parser.add_argument("--nt", type=int, default=50, required=False,
                        help="Number of r-transverse bins")
    parser.add_argument("--z-min-sources", type=float, default=0.0, required=False,
                        help="Limit the minimum redshift of the quasars used as sources for "
                             "spectra")
    parser.add_argument("--z-max-sources", type=float, default=10.0, required=False,
                        help="Limit the maximum redshift of the quasars used as sources for "
                             "spectra")
    parser.add_argument("--lambda-abs", type=str, default="LYA", required=False,
                        help="Name of the absorption in picca.constants defining the redshift "
                             "of the delta")
    parser.add_argument("--lambda-abs2", type=str, default=None, required=False,
                        help="Name of the absorption in picca.constants defining the redshift "
                             "of the 2nd delta")
    parser.add_argument("--z-ref", type=float, default=2.25, required=False,
                        help="Reference redshift")
    parser.add_argument("--z-evol", type=float, default=2.9, required=False,
                        help="Exponent of the redshift evolution of the delta field")
    parser.add_argument("--z-evol2", type=float, default=2.9, required=False,
                        help="Exponent of the redshift evolution of the 2nd delta field")
    parser.add_argument("--fid-Om", type=float, default=0.315, required=False,
                        help="Omega_matter(z=0) of fiducial LambdaCDM cosmology")
    parser.add_argument("--fid-Or", type=float, default=0.0, required=False,
                        help="Omega_radiation(z=0) of fiducial LambdaCDM cosmology")
    parser.add_argument("--fid-Ok", type=float, default=0.0, required=False,
                        help="Omega_k(z=0) of fiducial LambdaCDM cosmology")
    parser.add_argument("--fid-wl", type=float, default=-1.0, required=False,
                        help="Equation of state of dark energy of fiducial LambdaCDM cosmology")
    parser.add_argument("--no-project", action="store_true", required=False,
                        help="Do not project out continuum fitting modes")
    parser.add_argument("--max-diagram", type=int, default=3, required=False,
                        help="Maximum diagram to compute")
    parser.add_argument("--cf1d", type=str, required=True,
                        help="1D auto-correlation of pixels from the same forest file: "
                             "picca_cf1d.py")
    parser.add_argument("--cf1d2", type=str, default=None, required=False,
                        help="1D auto-correlation of pixels from the same forest file of the "
                             "2nd delta field: picca_cf1d.py")
    parser.add_argument("--cf", type=str, default=None, required=False,
                        help="3D auto-correlation of pixels from different forests: "
                             "picca_cf.py")
    parser.add_argument("--cf2", type=str, default=None, required=False,
                        help="3D auto-correlation of pixels from different forests for 2nd "
                             "catalog: picca_cf.py")
    parser.add_argument("--cf12", type=str, default=None, required=False,
                        help="3D auto-correlation of pixels from different forests for cross "
                             "1st and 2nd catalog: picca_cf.py")
    parser.add_argument("--unfold-cf", action="store_true", required=False,
                        help="rp can be positive or negative depending on the relative "
                             "position between absorber1 and absorber2")
    parser.add_argument("--rej", type=float, default=1.0, required=False,
                        help="Fraction of rejected pairs: -1=no rejection, 1=all rejection")
    parser.add_argument("--nside", type=int, default=16, required=False, help="Healpix nside")
    parser.add_argument("--nproc", type=int, default=None, required=False,
                        help="Number of process lists (fixes the rejection draws)")
    parser.add_argument("--nspec", type=int, default=None, required=False,
                        help="Maximum number of spectra to read")
    parser.add_argument("--rebin-factor", type=int, default=None, required=False,
                        help="Rebin factor for deltas. If not None, deltas will be rebinned "
                             "by that factor")
    parser.add_argument("--in-attributes", type=str, default=None, required=False,
                        help="Delta attributes file holding the continuum-fit order")
    parser.add_argument("--device", type=int, default=0, required=False,
                        help="HIP device (engine option)")
    return parser.parse_args(cmdargs)


def read_cf1d(filename):
    """Nearest-neighbour interpolations of the 1-D variance and correlation of a picca_cf1d
    file (py/picca/bin/picca_wick.py:396-419)."""
    hdu = FITS(filename)[1]
    header = hdu.read_header()
    log_lambda_min = header["LLMIN"]
    delta_log_lambda = header["DLL"]
    num_pairs_variance_1d = hdu["nv1d"][:]
    variance_1d = hdu["v1d"][:]
    log_lambda = log_lambda_min + delta_log_lambda * np.arange(len(variance_1d))
    variance = interp1d(log_lambda[num_pairs_variance_1d > 0],
                        variance_1d[num_pairs_variance_1d > 0], kind="nearest",
                        fill_value="extrapolate")
    num_pairs1d = hdu["nb1d"][:]
    xi_1d = hdu["c1d"][:]
    xi = interp1d((log_lambda - log_lambda_min)[num_pairs1d > 0], xi_1d[num_pairs1d > 0],
                  kind="nearest", fill_value="extrapolate")
    return variance, xi


def calc_wick_terms(healpixs):
    """One process list (py/picca/bin/picca_wick.py:20-37)."""
    cf.fill_neighs(healpixs)
    np.random.seed(healpixs[0])
    return cf.compute_wick_terms(healpixs)


def _set_fname(data, fname):
    for deltas in data.values():
        for delta in deltas:
            delta.fname = fname


def main(cmdargs=None):
    args = parse_args(cmdargs)
    reset_globals()
    cf.device = args.device
    if args.nproc is None:
        args.nproc = max(1, cpu_count() // 2)
    userprint("nproc", args.nproc)
    cf.r_par_max = args.rp_max
    cf.r_trans_max = args.rt_max
    cf.r_par_min = args.rp_min
    cf.num_bins_r_par = args.np
    cf.num_bins_r_trans = args.nt
    cf.nside = args.nside
    cf.z_ref = args.z_ref
    cf.alpha = args.z_evol
    cf.alpha2 = args.z_evol
    cf.lambda_abs = constants.ABSORBER_IGM[args.lambda_abs]
    cf.reject = args.rej
    cf.max_diagram = args.max_diagram
    if args.max_diagram > 3:
        raise NotImplementedError("diagrams 4 and 5 (--max-diagram > 3) are not part of this "
                                  "engine")

    blinding = io.read_blinding(args.in_dir)
    if (args.fid_Or != 0.0) or (args.fid_Ok != 0.0) or (args.fid_wl != -1.0):
        userprint("ERROR: Cosmology with other than Omega_m set are not yet implemented")
        sys.exit()
    cosmo = constants.Cosmo(Om=args.fid_Om, Or=args.fid_Or, Ok=args.fid_Ok, wl=args.fid_wl)
    t0 = time.time()
    data, num_data, z_min, z_max = io.read_deltas(
        args.in_dir, cf.nside, cf.lambda_abs, cf.alpha, cf.z_ref, cosmo,
        max_num_spec=args.nspec, nproc=args.nproc, rebin_factor=args.rebin_factor,
        z_min_qso=args.z_min_sources, z_max_qso=args.z_max_sources,
        delta_attributes=args.in_attributes)
    _set_fname(data, "D1")
    del z_max
    cf.data = data
    cf.num_data = num_data
    cf.ang_max = utils.compute_ang_max(cosmo, cf.r_trans_max, z_min)
    userprint("done, npix = {}".format(len(data)))

    cf.xi_1d = {}
    cf.get_variance_1d = {}
    for fname, filename in {"D1": args.cf1d, "D2": args.cf1d2}.items():
        if filename is None:
            continue
        cf.get_variance_1d[fname], cf.xi_1d[fname] = read_cf1d(filename)

    cf.xi_wick = {}
    for fname, filename in {"D1_D1": args.cf, "D2_D2": args.cf2, "D1_D2": args.cf12,
                            "D2_D1": args.cf12}.items():
        if filename is None:
            continue
        hdul = FITS(filename)
        header = hdul[1].read_header()
        assert cf.num_bins_r_par == header["NP"]
        assert cf.num_bins_r_trans == header["NT"]
        assert cf.r_par_min == header["RPMIN"]
        assert cf.r_par_max == header["RPMAX"]
        assert cf.r_trans_max == header["RTMAX"]
        xi = hdul[2]["DA"][:]
        weights = hdul[2]["WE"][:]
        xi = (xi * weights).sum(axis=0)
        weights = weights.sum(axis=0)
        w = weights > 0.0
        xi[w] /= weights[w]
        cf.xi_wick[fname] = xi.copy()

    if args.in_dir2 or args.lambda_abs2:
        if args.lambda_abs2 or args.unfold_cf:
            cf.x_correlation = True
        cf.alpha2 = args.z_evol2
        if args.in_dir2 is None:
            args.in_dir2 = args.in_dir
        if args.lambda_abs2:
            cf.lambda_abs2 = constants.ABSORBER_IGM[args.lambda_abs2]
        else:
            cf.lambda_abs2 = cf.lambda_abs
        data2, num_data2, z_min2, z_max2 = io.read_deltas(
            args.in_dir2, cf.nside, cf.lambda_abs2, cf.alpha2, cf.z_ref, cosmo,
            max_num_spec=args.nspec, nproc=args.nproc, rebin_factor=args.rebin_factor,
            z_min_qso=args.z_min_sources, z_max_qso=args.z_max_sources,
            delta_attributes=args.in_attributes)
        # picca_wick renames the first field's deltas here (py/picca/bin/picca_wick.py:474)
        _set_fname(data, "D2")
        del z_max2
        cf.data2 = data2
        cf.num_data2 = num_data2
        cf.ang_max = utils.compute_ang_max(cosmo, cf.r_trans_max, z_min, z_min2)
        userprint("done, npix = {}".format(len(data2)))

    cpu_data = {}
    for index, healpix in enumerate(sorted(data)):
        cpu_data.setdefault(index % args.nproc, []).append(healpix)
    wick_data = [calc_wick_terms(lst) for lst in sorted(cpu_data.values())]
    reports = [cf.last_report]

    weights_wick = np.array([item[0] for item in wick_data]).sum(axis=0)
    num_pairs_wick = np.array([item[1] for item in wick_data]).sum(axis=0)
    num_pairs = np.array([item[2] for item in wick_data]).sum(axis=0)
    num_pairs_used = np.array([item[3] for item in wick_data]).sum(axis=0)
    t = [np.array([item[k] for item in wick_data]).sum(axis=0) for k in range(4, 10)]
    weights = weights_wick * weights_wick[:, None]
    w = weights > 0.0
    for x in t:
        x[w] /= weights[w]
    for x in t:
        x *= 1.0 * num_pairs_used / num_pairs
    t_tot = t[0] + t[1] + t[2] + t[3] + t[4] + t[5]

    header = [
        {"name": "RPMIN", "value": cf.r_par_min, "comment": "Minimum r-parallel [h^-1 Mpc]"},
        {"name": "RPMAX", "value": cf.r_par_max, "comment": "Maximum r-parallel [h^-1 Mpc]"},
        {"name": "RTMAX", "value": cf.r_trans_max, "comment": "Maximum r-transverse [h^-1 Mpc]"},
        {"name": "NP", "value": cf.num_bins_r_par, "comment": "Number of bins in r-parallel"},
        {"name": "NT", "value": cf.num_bins_r_trans,
         "comment": "Number of bins in r-transverse"},
        {"name": "REJ", "value": cf.reject, "comment": "Rejection factor"},
        {"name": "NPALL", "value": int(num_pairs), "comment": "Number of pairs"},
        {"name": "NPUSED", "value": int(num_pairs_used), "comment": "Number of used pairs"},
        {"name": "OMEGAM", "value": args.fid_Om,
         "comment": "Omega_matter(z=0) of fiducial LambdaCDM cosmology"},
        {"name": "OMEGAR", "value": args.fid_Or,
         "comment": "Omega_radiation(z=0) of fiducial LambdaCDM cosmology"},
        {"name": "OMEGAK", "value": args.fid_Ok,
         "comment": "Omega_k(z=0) of fiducial LambdaCDM cosmology"},
        {"name": "WL", "value": args.fid_wl,
         "comment": "Equation of state of dark energy of fiducial LambdaCDM cosmology"},
        {"name": "BLINDING", "value": blinding,
         "comment": "String specifying the blinding strategy"},
    ]
    out = FITSWriter(args.out)
    out.write([t_tot, weights_wick, num_pairs_wick] + t,
              names=["CO", "WALL", "NB", "T1", "T2", "T3", "T4", "T5", "T6"],
              comment=["Sum of weight", "Covariance", "Nomber of pairs", "T1", "T2", "T3",
                       "T4", "T5", "T6"], header=header, extname="COV")
    out.close()
    userprint(f"picca_wick.py - Time total : {(time.time() - t0) / 60:.3f} minutes")
    return reports


if __name__ == "__main__":
    main(sys.argv[1:])
