#!/usr/bin/env python3
"""Distortion matrix of the delta x quasar cross-correlation on the MI355X engine.

Command-line mirror of ``py/picca/bin/picca_xdmat.py``: same flags and defaults, module
globals (:346-385), delta then quasar reading with the default quasar redshift range
(:398-446), maximum angle (:450), the split of the healpix list into ``--nproc`` process
lists (:455-461) -- each list fills its neighbours and seeds numpy's legacy stream with its
first healpix before its rejection draws (``calc_dmat``, :22-38), so the kept pairs depend
on ``--nproc`` exactly as in picca -- the merge and normalisation (:476-506) and the
``DMAT`` / ``ATTRI`` output (:508-617).  The lists run one after the other, each as one
batched GPU computation (``xcf.compute_dmat``).
"""
import argparse
import sys
import time
from multiprocessing import cpu_count

import numpy as np

from picca_amd import constants, io, utils, xcf
from picca_amd.fitsio_lite import FITSWriter
from picca_amd.utils import userprint


def parse_args(cmdargs=None):
    parser = argparse.ArgumentParser(
        formatter_class=argparse.ArgumentDefaultsHelpFormatter,
        description="Compute the distortion matrix of the cross-correlation delta x "
                    "object.")
    parser.add_argument("--out", type=str, default=None, required=True, help="Output file name")
    parser.add_argument("--in-dir", type=str, default=None, required=True,
                        help="Directory to delta files")
    parser.add_argument("--drq", type=str, default=None, required=True,
                        help="Catalog of objects in DRQ format")
    parser.add_argument("--mode", type=str, default="sdss", required=False,
                        choices=["desi", "sdss", "desi_mocks", "desi_healpix"],
                        help="type of catalog supplied, default sdss")
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
    parser.add_argument("--coef-binning-model", type=int, default=1, required=False,
                        help="Coefficient multiplying np and nt to get finner binning for "
                             "the model")
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
    parser.add_argument("--rej", type=float, default=1.0, required=False,
                        help="Fraction of rejected object-forests pairs: -1=no rejection, "
                             "1=all rejection")
    parser.add_argument("--nside", type=int, default=16, required=False, help="Healpix nside")
    parser.add_argument("--nproc", type=int, default=None, required=False,
                        help="Number of process lists (fixes the rejection draws)")
    parser.add_argument("--nspec", type=int, default=None, required=False,
                        help="Maximum number of spectra to read")
    parser.add_argument("--rebin-factor", type=int, default=None, required=False,
                        help="Rebin factor for deltas")
    parser.add_argument("--in-attributes", type=str, default=None, required=False,
                        help="Delta attributes file holding the continuum-fit order")
    parser.add_argument("--no-redshift-evolution", action="store_true",
                        help="Ignore redshift evolution when computing distortion matrix")
    parser.add_argument("--device", type=int, default=0, required=False,
                        help="HIP device (engine option)")
    return parser.parse_args(cmdargs)


def calc_dmat(healpixs):
    """One process list of picca_xdmat (py/picca/bin/picca_xdmat.py:22-38)."""
    xcf.fill_neighs(healpixs)
    np.random.seed(healpixs[0])
    return xcf.compute_dmat(healpixs)


def main(cmdargs=None):
    args = parse_args(cmdargs)
    if args.nproc is None:
        args.nproc = max(1, cpu_count() // 2)
    xcf.device = args.device
    xcf.rmu_binning = args.rmu_binning
    if xcf.rmu_binning:
        args.rp_min = -1
        args.rp_max = 1
    xcf.r_par_max = args.rp_max
    xcf.r_par_min = args.rp_min
    xcf.r_trans_max = args.rt_max
    xcf.z_min_pairs = args.z_min_pairs
    xcf.z_max_pairs = args.z_max_pairs
    xcf.num_bins_r_par = args.np
    xcf.num_bins_r_trans = args.nt
    xcf.num_model_bins_r_par = args.np * args.coef_binning_model
    xcf.num_model_bins_r_trans = args.nt * args.coef_binning_model
    xcf.nside = args.nside
    xcf.redshift_evolution_in_distortion_matrix = not args.no_redshift_evolution
    if args.no_redshift_evolution:
        userprint("ignore redshift evolution in the distortion matrix")
    xcf.z_ref = 2.25
    xcf.alpha = args.z_evol_del
    xcf.alpha_obj = args.z_evol_obj
    xcf.lambda_abs = constants.ABSORBER_IGM[args.lambda_abs]
    xcf.reject = args.rej
    xcf.ang_correlation = False
    if (args.zerr_cut_deg is None) != (args.zerr_cut_kms is None):
        raise ValueError("Options --zerr-cut-deg and --zerr-cut-kms must be specified together")
    xcf.zerr_cut_deg = args.zerr_cut_deg
    xcf.zerr_cut_kms = args.zerr_cut_kms

    blinding = io.read_blinding(args.in_dir)
    cosmo = constants.Cosmo(Om=args.fid_Om, Or=args.fid_Or, Ok=args.fid_Ok, wl=args.fid_wl,
                            blinding=blinding, template=args.distance_template)
    t0 = time.time()
    data, num_data, z_min, z_max = io.read_deltas(
        args.in_dir, args.nside, xcf.lambda_abs, args.z_evol_del, xcf.z_ref, cosmo,
        max_num_spec=args.nspec, nproc=args.nproc, rebin_factor=args.rebin_factor,
        z_min_qso=args.z_min_sources, z_max_qso=args.z_max_sources,
        delta_attributes=args.in_attributes)
    xcf.data = data
    xcf.num_data = num_data
    userprint("")
    userprint("done, npix = {}\n".format(len(data)))
    if args.z_min_obj is None:
        r_comov_min = max(0.0, float(cosmo.get_r_comov(z_min)) + xcf.r_par_min)
        args.z_min_obj = float(cosmo.distance_to_redshift(r_comov_min))
        userprint("z_min_obj = {}".format(args.z_min_obj), end="")
    if args.z_max_obj is None:
        r_comov_max = max(0.0, float(cosmo.get_r_comov(z_max)) + xcf.r_par_max)
        args.z_max_obj = float(cosmo.distance_to_redshift(r_comov_max))
        userprint("z_max_obj = {}".format(args.z_max_obj), end="")
    objs, z_min2 = io.read_objects(args.drq, args.nside, args.z_min_obj, args.z_max_obj,
                                   args.z_evol_obj, xcf.z_ref, cosmo, mode=args.mode)
    userprint("")
    xcf.objs = objs
    xcf.ang_max = utils.compute_ang_max(cosmo, xcf.r_trans_max, z_min, z_min2)
    t1 = time.time()
    userprint(f"picca_xdmat.py - Time reading data: {(t1 - t0) / 60:.3f} minutes")

    cpu_data = {}
    for index, healpix in enumerate(sorted(data)):
        cpu_data.setdefault(index % args.nproc, []).append(healpix)
    dmat_data = [calc_dmat(lst) for lst in sorted(cpu_data.values())]
    t2 = time.time()
    userprint(f"picca_xdmat.py - Time computing distortion matrix: {(t2 - t1) / 60:.3f} "
              f"minutes")

    weights_dmat = np.array([item[0] for item in dmat_data]).sum(axis=0)
    dmat = np.array([item[1] for item in dmat_data]).sum(axis=0)
    r_par = np.array([item[2] for item in dmat_data]).sum(axis=0)
    r_trans = np.array([item[3] for item in dmat_data]).sum(axis=0)
    zeff = np.array([item[4] for item in dmat_data]).sum(axis=0)
    weights = np.array([item[5] for item in dmat_data]).sum(axis=0)
    num_pairs = np.array([item[6] for item in dmat_data]).sum(axis=0)
    num_pairs_used = np.array([item[7] for item in dmat_data]).sum(axis=0)

    w = weights > 0.0
    r_par[w] /= weights[w]
    r_trans[w] /= weights[w]
    zeff[w] /= weights[w]
    mean_zeff = np.mean(zeff[w])
    w = weights_dmat > 0.0
    dmat[w, :] /= weights_dmat[w, None]
    if xcf.redshift_evolution_in_distortion_matrix:
        zeff[:] = mean_zeff
        zfac = ((1 + xcf.z_ref) / (1 + mean_zeff)) ** ((xcf.alpha - 1) + (xcf.alpha_obj - 1))
        dmat *= zfac

    # the header of py/picca/bin/picca_xdmat.py:509-583
    header = [
        {"name": "RPMIN", "value": xcf.r_par_min, "comment": "Minimum r-parallel [h^-1 Mpc]"},
        {"name": "RPMAX", "value": xcf.r_par_max, "comment": "Maximum r-parallel [h^-1 Mpc]"},
        {"name": "RTMAX", "value": xcf.r_trans_max, "comment": "Maximum r-transverse [h^-1 Mpc]"},
        {"name": "NP", "value": xcf.num_bins_r_par, "comment": "Number of bins in r-parallel"},
        {"name": "NT", "value": xcf.num_bins_r_trans, "comment": "Number of bins in r-transverse"},
        {"name": "COEFMOD", "value": args.coef_binning_model,
         "comment": "Coefficient for model binning"},
        {"name": "ZMIN", "value": xcf.z_min_pairs, "comment": "Minimum redshift of pairs"},
        {"name": "ZMAX", "value": xcf.z_max_pairs, "comment": "Maximum redshift of pairs"},
        {"name": "REJ", "value": xcf.reject, "comment": "Rejection factor"},
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
        {"name": "RMU_BIN", "value": xcf.rmu_binning, "comment": "True if binned in r, mu"},
    ]
    dmat_name = "DM" if blinding == "none" else "DM_BLIND"
    results = FITSWriter(args.out)
    results.write([weights_dmat, dmat], names=["WDM", dmat_name],
                  comment=["Sum of weight", "Distortion matrix"], units=["", ""],
                  header=header, extname="DMAT")
    results.write([r_par, r_trans, zeff], names=["RP", "RT", "Z"],
                  comment=["R-parallel", "R-transverse", "Redshift"],
                  units=["h^-1 Mpc", "h^-1 Mpc", ""], extname="ATTRI")
    results.close()
    userprint(f"picca_xdmat.py - Time total: {(time.time() - t0) / 60:.3f} minutes")
    return xcf.last_report


if __name__ == "__main__":
    main(sys.argv[1:])
