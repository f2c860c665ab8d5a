#!/usr/bin/env python3
"""Auto- and cross-correlation of delta fields on the MI355X engine.

Command-line mirror of ``py/picca/bin/picca_cf.py``: same flags and defaults
(:42-338), same module-global setup (:343-368), same reading (:370-447) and the same
``ATTRI`` / ``COR`` FITS output (:486-580).  Instead of a fork pool with one task per
healpix (:452-458), every healpix is one output row of a single GPU launch
(``cf.compute_xi_rows``); the per-row normalisation and the reduction of
``py/picca/bin/picca_cf.py:466-484`` are then applied unchanged.
"""
import argparse
import sys
import time

import numpy as np

from picca_amd import cf, constants, io, utils
from picca_amd.fitsio_lite import FITSWriter
from picca_amd.utils import userprint


def parse_args(cmdargs=None):
    parser = argparse.ArgumentParser(
        formatter_class=argparse.ArgumentDefaultsHelpFormatter,
        description="Compute the auto and cross-correlation of delta fields")
    parser.add_argument("--out", type=str, default=None, required=True,
                        help="Output file name")
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
    parser.add_argument("--rmu-binning", action="store_true",
                        help="Estimate in r,mu binning (np: mu bins, nt: r bins)")
    parser.add_argument("--zerr-cut-deg", type=float, default=None, required=False,
                        help="Angular separation below which the redshift-error cut applies")
    parser.add_argument("--zerr-cut-kms", type=float, default=None, required=False,
                        help="Velocity separation of the redshift-error cut [km/s]")
    parser.add_argument("--z-min-sources", type=float, default=0.0, required=False,
                        help="Limit the minimum redshift of the quasars")
    parser.add_argument("--z-max-sources", type=float, default=10.0, required=False,
                        help="Limit the maximum redshift of the quasars")
    parser.add_argument("--z-min-pairs", type=float, default=None, required=False,
                        help="Limit the minimum redshift of the pairs")
    parser.add_argument("--z-max-pairs", type=float, default=None, required=False,
                        help="Limit the maximum redshift of the pairs")
    parser.add_argument("--lambda-abs", type=str, default="LYA", required=False,
                        help="Name of the absorption in picca.constants defining the redshift")
    parser.add_argument("--lambda-abs2", type=str, default=None, required=False,
                        help="Name of the absorption of the 2nd delta field")
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
    parser.add_argument("--distance-template", type=str, default=None, required=False,
                        help="camb ini file for a distance template (not supported)")
    parser.add_argument("--no-project", action="store_true", required=False,
                        help="Do not project out continuum fitting modes")
    parser.add_argument("--remove-same-half-plate-close-pairs", action="store_true",
                        required=False,
                        help="Reject pairs in the first bin in r-parallel from same half plate")
    parser.add_argument("--nside", type=int, default=16, required=False, help="Healpix nside")
    parser.add_argument("--nproc", type=int, default=None, required=False,
                        help="Number of processors (reading only; the pairs run on the GPU)")
    parser.add_argument("--nspec", type=int, default=None, required=False,
                        help="Maximum number of spectra to read")
    parser.add_argument("--unfold-cf", action="store_true", required=False,
                        help="rp can be positive or negative depending on the relative "
                             "position between absorber1 and absorber2")
    parser.add_argument("--shuffle-distrib-forest-seed", type=int, default=None,
                        required=False, help="Shuffle the distribution of forests on the sky")
    parser.add_argument("--rebin-factor", type=int, default=None, required=False,
                        help="Rebin factor for deltas")
    parser.add_argument("--in-attributes", type=str, default=None, required=False,
                        help="Delta attributes file holding the continuum-fit order")
    parser.add_argument("--device", type=int, default=0, required=False,
                        help="HIP device (engine option)")
    return parser.parse_args(cmdargs)


def reset_globals():
    """Reset the module globals the way picca's tests do between runs."""
    for name in ("num_bins_r_par", "num_bins_r_trans", "num_model_bins_r_trans",
                 "num_model_bins_r_par", "r_par_max", "r_par_min", "z_min_pairs",
                 "z_max_pairs", "r_trans_max", "ang_max", "nside", "zerr_cut_deg",
                 "zerr_cut_kms", "counter", "num_data", "num_data2", "z_ref", "alpha",
                 "alpha2", "alpha_abs", "lambda_abs", "lambda_abs2", "data", "data2", "cosmo",
                 "reject", "lock"):
        setattr(cf, name, None)
    cf.x_correlation = False
    cf.rmu_binning = False
    cf.ang_correlation = False
    cf.remove_same_half_plate_close_pairs = False
    cf.redshift_evolution_in_distortion_matrix = True


def main(cmdargs=None):
    args = parse_args(cmdargs)
    reset_globals()
    cf.device = args.device

    cf.rmu_binning = args.rmu_binning
    if cf.rmu_binning:
        args.rp_min = 0
        args.rp_max = 1
    cf.r_par_max = args.rp_max
    cf.r_trans_max = args.rt_max
    cf.r_par_min = args.rp_min
    cf.z_min_pairs = args.z_min_pairs
    cf.z_max_pairs = args.z_max_pairs
    cf.num_bins_r_par = args.np
    cf.num_bins_r_trans = args.nt
    cf.nside = args.nside
    cf.z_ref = args.z_ref
    cf.alpha = args.z_evol
    cf.lambda_abs = constants.ABSORBER_IGM[args.lambda_abs]
    cf.remove_same_half_plate_close_pairs = args.remove_same_half_plate_close_pairs
    if (args.zerr_cut_deg is None) != (args.zerr_cut_kms is None):
        raise ValueError("Options --zerr-cut-deg and --zerr-cut-kms must be specified together")
    cf.zerr_cut_deg = args.zerr_cut_deg
    cf.zerr_cut_kms = args.zerr_cut_kms

    blinding = io.read_blinding(args.in_dir)
    cosmo = constants.Cosmo(Om=args.fid_Om, Or=args.fid_Or, Ok=args.fid_Ok, wl=args.fid_wl,
                            blinding=blinding, template=args.distance_template)

    t0 = time.time()
    data, num_data, z_min, z_max = io.read_deltas(
        args.in_dir, cf.nside, cf.lambda_abs, cf.alpha, cf.z_ref, cosmo,
        max_num_spec=args.nspec, no_project=args.no_project, nproc=args.nproc,
        rebin_factor=args.rebin_factor, z_min_qso=args.z_min_sources,
        z_max_qso=args.z_max_sources, delta_attributes=args.in_attributes)
    del z_max
    cf.data = data
    cf.num_data = num_data
    cf.ang_max = utils.compute_ang_max(cosmo, cf.r_trans_max, z_min)
    userprint("")
    userprint("done, npix = {}".format(len(data)))

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
            max_num_spec=args.nspec, no_project=args.no_project, nproc=args.nproc,
            rebin_factor=args.rebin_factor, z_min_qso=args.z_min_sources,
            z_max_qso=args.z_max_sources, delta_attributes=args.in_attributes)
        del z_max2
        cf.data2 = data2
        cf.num_data2 = num_data2
        cf.ang_max = utils.compute_ang_max(cosmo, cf.r_trans_max, z_min, z_min2)
        userprint("")
        userprint("done, npix = {}".format(len(data2)))

    if args.shuffle_distrib_forest_seed is not None:
        cf.data = utils.shuffle_distrib_forests(cf.data, args.shuffle_distrib_forest_seed)

    t1 = time.time()
    userprint(f"picca_cf.py - Time reading data: {(t1 - t0) / 60:.3f} minutes")

    # one row per healpix of the data as read (py/picca/bin/picca_cf.py:452,472)
    healpix_list = np.array(sorted(data))
    rows = cf.compute_xi_rows(healpix_list)
    t2 = time.time()
    userprint(f"picca_cf.py - Time computing correlation function: {(t2 - t1) / 60:.3f} minutes")
    if cf.last_report is not None:
        userprint("engine report: {}".format(cf.last_report))

    rows = np.array(rows)
    weights_list = rows[:, 0, :]
    xi_list = rows[:, 1, :]
    r_par_list = rows[:, 2, :]
    r_trans_list = rows[:, 3, :]
    z_list = rows[:, 4, :]
    num_pairs_list = rows[:, 5, :].astype(np.int64)

    sum_weights = weights_list.sum(axis=0)
    w = sum_weights > 0
    r_par = (r_par_list * weights_list).sum(axis=0)
    r_par[w] /= sum_weights[w]
    r_trans = (r_trans_list * weights_list).sum(axis=0)
    r_trans[w] /= sum_weights[w]
    z = (z_list * weights_list).sum(axis=0)
    z[w] /= sum_weights[w]
    num_pairs = num_pairs_list.sum(axis=0)

    header = [
        {"name": "RPMIN", "value": cf.r_par_min, "comment": "Minimum r-parallel [h^-1 Mpc]"},
        {"name": "RPMAX", "value": cf.r_par_max, "comment": "Maximum r-parallel [h^-1 Mpc]"},
        {"name": "RTMAX", "value": cf.r_trans_max, "comment": "Maximum r-transverse [h^-1 Mpc]"},
        {"name": "NP", "value": cf.num_bins_r_par, "comment": "Number of bins in r-parallel"},
        {"name": "NT", "value": cf.num_bins_r_trans, "comment": "Number of bins in r-transverse"},
        {"name": "ZMIN", "value": cf.z_min_pairs, "comment": "Minimum redshift of pairs"},
        {"name": "ZMAX", "value": cf.z_max_pairs, "comment": "Maximum redshift of pairs"},
        {"name": "NSIDE", "value": cf.nside, "comment": "Healpix nside"},
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
        {"name": "RMU_BIN", "value": cf.rmu_binning, "comment": "True if binned in r, mu"},
    ]
    results = FITSWriter(args.out)
    results.write([r_par, r_trans, z, num_pairs], names=["RP", "RT", "Z", "NB"],
                  comment=["R-parallel" if not cf.rmu_binning else "Mu",
                           "R-transverse" if not cf.rmu_binning else "Radial separation",
                           "Redshift", "Number of pairs"],
                  units=["h^-1 Mpc" if not cf.rmu_binning else "", "h^-1 Mpc", "", ""],
                  header=header, extname="ATTRI")
    xi_list_name = "DA" if blinding == "none" else "DA_BLIND"
    results.write([healpix_list, weights_list, xi_list], names=["HEALPID", "WE", xi_list_name],
                  comment=["Healpix index", "Sum of weight", "Correlation"],
                  header=[{"name": "HLPXSCHM", "value": "RING", "comment": "Healpix scheme"}],
                  extname="COR")
    results.close()
    t3 = time.time()
    userprint(f"picca_cf.py - Time total : {(t3 - t0) / 60:.3f} minutes")
    return cf.last_report


if __name__ == "__main__":
    main(sys.argv[1:])
