#!/usr/bin/env python3
"""Cross-correlation of a delta field with a quasar catalogue on the MI355X engine.

Command-line mirror of ``py/picca/bin/picca_xcf.py``: same flags and defaults, module
globals (:337-367), quasar and delta reading (:369-434), removal of the mean delta
against observed wavelength (:436-468), maximum angle (:478-479) and the ``ATTRI`` /
``COR`` output (:506-620).  Every healpix of the deltas is one output row of a single GPU
launch instead of one pool task (:489-493); the reduction is unchanged.
"""
import argparse
import sys
import time

import numpy as np

from picca_amd import constants, io, prep_del, utils, xcf
from picca_amd.fitsio_lite import FITSWriter
from picca_amd.utils import userprint


def parse_args(cmdargs=None):
    parser = argparse.ArgumentParser(
        formatter_class=argparse.ArgumentDefaultsHelpFormatter,
        description="Compute the cross-correlation between a catalog of objects and a "
                    "delta field.")
    parser.add_argument("--out", type=str, default=None, required=True, help="Output file name")
    parser.add_argument("--in-dir", type=str, default=None, required=True,
                        help="Directory to delta files")
    parser.add_argument("--drq", type=str, default=None, required=True,
                        help="Catalog of objects in DRQ format")
    parser.add_argument("--mode", type=str, default="sdss", required=False,
                        choices=["desi", "sdss", "desi_mocks", "desi_healpix"],
                        help="type of catalog supplied")
    parser.add_argument("--rp-min", type=float, default=-200.0, required=False,
                        help="Min r-parallel [h^-1 Mpc]")
    parser.add_argument("--rp-max", type=float, default=200.0, required=False,
                        help="Max r-parallel [h^-1 Mpc]")
    parser.add_argument("--rt-max", type=float, default=200.0, required=False,
                        help="Max r-transverse [h^-1 Mpc]")
    parser.add_argument("--np", type=int, default=100, required=False,
                        help="Number of r-parallel bins")
    parser.add_argument("--nt", type=int, default=50, required=False,
                        help="Number of r-transverse bins")
    parser.add_argument("--rmu-binning", action="store_true",
                        help="Estimate in r,mu binning (np: mu bins, nt: r bins)")
    parser.add_argument("--z-min-obj", type=float, default=0, required=False,
                        help="Min redshift for object field")
    parser.add_argument("--z-max-obj", type=float, default=10, required=False,
                        help="Max redshift for object field")
    parser.add_argument("--zerr-cut-deg", type=float, default=None, required=False,
                        help="Angular separation below which the redshift-error cut applies")
    parser.add_argument("--zerr-cut-kms", type=float, default=None, required=False,
                        help="Velocity separation of the redshift-error cut [km/s]")
    parser.add_argument("--z-min-sources", type=float, default=0.0, required=False,
                        help="Limit the minimum redshift of the quasars used as sources "
                             "for spectra")
    parser.add_argument("--z-max-sources", type=float, default=10.0, required=False,
                        help="Limit the maximum redshift of the quasars used as sources "
                             "for spectra")
    parser.add_argument("--z-min-pairs", type=float, default=None, required=False,
                        help="Limit the minimum redshift of the pairs")
    parser.add_argument("--z-max-pairs", type=float, default=None, required=False,
                        help="Limit the maximum redshift of the pairs")
    parser.add_argument("--lambda-abs", type=str, default="LYA", required=False,
                        help="Name of the absorption in picca.constants defining the redshift")
    parser.add_argument("--z-ref", type=float, default=2.25, required=False,
                        help="Reference redshift")
    parser.add_argument("--z-evol-del", type=float, default=2.9, required=False,
                        help="Exponent of the redshift evolution of the delta field")
    parser.add_argument("--z-evol-obj", type=float, default=1.44, required=False,
                        help="Exponent of the redshift evolution of the object field")
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
    parser.add_argument("--no-remove-mean-lambda-obs", action="store_true", required=False,
                        help="Do not remove mean delta versus lambda_obs")
    parser.add_argument("--nside", type=int, default=16, required=False, help="Healpix nside")
    parser.add_argument("--nproc", type=int, default=None, required=False,
                        help="Number of processors (reading only; the pairs run on the GPU)")
    parser.add_argument("--nspec", type=int, default=None, required=False,
                        help="Maximum number of spectra to read")
    parser.add_argument("--shuffle-distrib-obj-seed", type=int, default=None, required=False,
                        help="Shuffle the distribution of objects on the sky")
    parser.add_argument("--shuffle-distrib-forest-seed", type=int, default=None,
                        required=False, help="Shuffle the distribution of forests on the sky")
    parser.add_argument("--rebin-factor", type=int, default=None, required=False,
                        help="Rebin factor for deltas")
    parser.add_argument("--in-attributes", type=str, default=None, required=False,
                        help="Delta attributes file holding the continuum-fit order")
    parser.add_argument("--device", type=int, default=0, required=False,
                        help="HIP device (engine option)")
    return parser.parse_args(cmdargs)


def main(cmdargs=None):
    args = parse_args(cmdargs)
    xcf.device = args.device
    xcf.rmu_binning = args.rmu_binning
    if xcf.rmu_binning:
        args.rp_min = -1
        args.rp_max = 1
    xcf.r_par_max = args.rp_max
    xcf.r_par_min = args.rp_min
    xcf.z_min_pairs = args.z_min_pairs
    xcf.z_max_pairs = args.z_max_pairs
    xcf.r_trans_max = args.rt_max
    xcf.num_bins_r_par = args.np
    xcf.num_bins_r_trans = args.nt
    xcf.nside = args.nside
    xcf.lambda_abs = constants.ABSORBER_IGM[args.lambda_abs]
    xcf.z_ref = args.z_ref
    xcf.ang_correlation = False
    if (args.zerr_cut_deg is None) != (args.zerr_cut_kms is None):
        raise ValueError("Options --zerr-cut-deg and --zerr-cut-kms must be specified together")
    xcf.zerr_cut_deg = args.zerr_cut_deg
    xcf.zerr_cut_kms = args.zerr_cut_kms

    blinding = io.read_blinding(args.in_dir)
    cosmo = constants.Cosmo(Om=args.fid_Om, Or=args.fid_Or, Ok=args.fid_Ok, wl=args.fid_wl,
                            blinding=blinding, template=args.distance_template)
    t0 = time.time()
    objs, z_min2 = io.read_objects(args.drq, args.nside, args.z_min_obj, args.z_max_obj,
                                   args.z_evol_obj, args.z_ref, cosmo, mode=args.mode)
    xcf.objs = objs
    data, num_data, z_min, z_max = io.read_deltas(
        args.in_dir, args.nside, xcf.lambda_abs, args.z_evol_del, args.z_ref, cosmo,
        max_num_spec=args.nspec, no_project=args.no_project, nproc=args.nproc,
        rebin_factor=args.rebin_factor, z_min_qso=args.z_min_sources,
        z_max_qso=args.z_max_sources, delta_attributes=args.in_attributes)
    xcf.data = data
    xcf.num_data = num_data
    userprint("")
    userprint("done, npix = {}\n".format(len(data)))
    if not args.no_remove_mean_lambda_obs:
        prep_del.remove_mean_lambda_obs(xcf.data, z_min, z_max, xcf.lambda_abs)
    if args.shuffle_distrib_obj_seed is not None:
        xcf.objs = utils.shuffle_distrib_forests(objs, args.shuffle_distrib_obj_seed)
    if args.shuffle_distrib_forest_seed is not None:
        xcf.data = utils.shuffle_distrib_forests(xcf.data, args.shuffle_distrib_forest_seed)
    xcf.ang_max = utils.compute_ang_max(cosmo, xcf.r_trans_max, z_min, z_min2)
    t1 = time.time()
    userprint(f"picca_xcf.py - Time reading data: {(t1 - t0) / 60:.3f} minutes")

    healpix_list = np.array(sorted(xcf.data))
    rows = np.array(xcf.compute_xi_rows(healpix_list))
    t2 = time.time()
    userprint(f"picca_xcf.py - Time computing cross-correlation function: "
              f"{(t2 - t1) / 60:.3f} minutes")
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
        {"name": "RPMIN", "value": xcf.r_par_min, "comment": "Minimum r-parallel [h^-1 Mpc]"},
        {"name": "RPMAX", "value": xcf.r_par_max, "comment": "Maximum r-parallel [h^-1 Mpc]"},
        {"name": "RTMAX", "value": xcf.r_trans_max, "comment": "Maximum r-transverse [h^-1 Mpc]"},
        {"name": "NP", "value": xcf.num_bins_r_par, "comment": "Number of bins in r-parallel"},
        {"name": "NT", "value": xcf.num_bins_r_trans, "comment": "Number of bins in r-transverse"},
        {"name": "ZMIN", "value": xcf.z_min_pairs, "comment": "Minimum redshift of pairs"},
        {"name": "ZMAX", "value": xcf.z_max_pairs, "comment": "Maximum redshift of pairs"},
        {"name": "NSIDE", "value": xcf.nside, "comment": "Healpix nside"},
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
        {"name": "RMU_BIN", "value": xcf.rmu_binning, "comment": "True if binned in r, mu"},
    ]
    results = FITSWriter(args.out)
    results.write([r_par, r_trans, z, num_pairs], names=["RP", "RT", "Z", "NB"],
                  comment=["R-parallel" if not xcf.rmu_binning else "Mu",
                           "R-transverse" if not xcf.rmu_binning else "Radial separation",
                           "Redshift", "Number of pairs"],
                  units=["h^-1 Mpc" if not xcf.rmu_binning else "", "h^-1 Mpc", "", ""],
                  header=header, extname="ATTRI")
    da_name = "DA" if blinding == "none" else "DA_BLIND"
    results.write([healpix_list, weights_list, xi_list], names=["HEALPID", "WE", da_name],
                  comment=["Healpix index", "Sum of weight", "Correlation"],
                  header=[{"name": "HLPXSCHM", "value": "RING", "comment": "Healpix scheme"}],
                  extname="COR")
    results.close()
    userprint(f"picca_xcf.py - Time total: {(time.time() - t0) / 60:.3f} minutes")
    return xcf.last_report


if __name__ == "__main__":
    main(sys.argv[1:])
