#!/usr/bin/env python3
"""Metal distortion matrices of the delta x object cross-correlation on the MI355X engine.

Command-line mirror of ``py/picca/bin/picca_metal_xdmat.py``: same flags and defaults, module
globals (:318-335), delta then object reading (:344-395), maximum angle, the split of the
healpix list into ``--nproc`` process lists, each of which fills its neighbours and seeds
numpy's legacy stream with its first healpix before its rejection draws
(``calc_metal_dmat``, :20-44), the loop over the metals, the merge and normalisation and
the ``ATTRI`` / ``MDMAT`` output (columns ``RP_<obj-name>_<metal>`` ...).  Each process
list runs as one batched GPU pass (``xcf.compute_metal_dmat``).
"""
import argparse
import sys
import time
from multiprocessing import cpu_count

import numpy as np

from picca_amd import constants, io, utils, xcf
from picca_amd.bin.picca_metal_dmat import write_metal_file
from picca_amd.utils import userprint


def parse_args(cmdargs=None):
    parser = argparse.ArgumentParser(
        formatter_class=argparse.ArgumentDefaultsHelpFormatter,
        description="Compute the metal distortion matrices of the cross-correlation delta x "
                    "object for a list of IGM absorption.")
    parser.add_argument("--out", type=str, default=None, required=True, help="Output file name")
    parser.add_argument("--in-dir", type=str, default=None, required=True,
                        help="Directory to delta files")
    parser.add_argument("--drq", type=str, default=None, required=True,
                        help="Catalog of objects in DRQ format")
    parser.add_argument("--mode", type=str, default="sdss",
                        choices=["sdss", "desi", "desi_mocks", "desi_healpix"], required=False,
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
    parser.add_argument("--coef-binning-model", type=int, default=1, required=False,
                        help="Coefficient multiplying np and nt to get finner binning for the "
                             "model")
    parser.add_argument("--z-min-obj", type=float, default=0, required=False,
                        help="Min redshift for object field")
    parser.add_argument("--z-max-obj", type=float, default=10, required=False,
                        help="Max redshift for object field")
    parser.add_argument("--z-min-sources", type=float, default=0.0, required=False,
                        help="Limit the minimum redshift of the quasars used as sources for "
                             "spectra")
    parser.add_argument("--z-max-sources", type=float, default=10.0, required=False,
                        help="Limit the maximum redshift of the quasars used as sources for "
                             "spectra")
    parser.add_argument("--lambda-abs", type=str, default="LYA", required=False,
                        help="Name of the absorption in picca.constants defining the redshift "
                             "of the delta")
    parser.add_argument("--obj-name", type=str, default="QSO", required=False,
                        help="Name of the object tracer")
    parser.add_argument("--abs-igm", type=str, default=None, required=False, nargs="*",
                        help="List of names of metal absorption in picca.constants")
    parser.add_argument("--z-ref", type=float, default=2.25, required=False,
                        help="Reference redshift")
    parser.add_argument("--z-evol-del", type=float, default=2.9, required=False,
                        help="Exponent of the redshift evolution of the delta field")
    parser.add_argument("--z-evol-obj", type=float, default=1.44, required=False,
                        help="Exponent of the redshift evolution of the object field")
    parser.add_argument("--metal-alpha", type=float, default=1.0, required=False,
                        help="Exponent of the redshift evolution of the metal delta field")
    parser.add_argument("--fid-Om", type=float, default=0.315, required=False,
                        help="Omega_matter(z=0) of fiducial LambdaCDM cosmology")
    parser.add_argument("--fid-Or", type=float, default=0.0, required=False,
                        help="Omega_radiation(z=0) of fiducial LambdaCDM cosmology")
    parser.add_argument("--fid-Ok", type=float, default=0.0, required=False,
                        help="Omega_k(z=0) of fiducial LambdaCDM cosmology")
    parser.add_argument("--fid-wl", type=float, default=-1.0, required=False,
                        help="Equation of state of dark energy of fiducial LambdaCDM cosmology")
    parser.add_argument("--rej", type=float, default=1.0, required=False,
                        help="Fraction of rejected object-forests pairs: -1=no rejection, "
                             "1=all rejection")
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


def calc_metal_dmat(abs_igm, healpixs):
    """One process list (py/picca/bin/picca_metal_xdmat.py:20-44)."""
    xcf.fill_neighs(healpixs)
    np.random.seed(healpixs[0])
    return xcf.compute_metal_dmat(healpixs, abs_igm=abs_igm)


def main(cmdargs=None):
    args = parse_args(cmdargs)
    if args.nproc is None:
        args.nproc = max(1, cpu_count() // 2)
    xcf.device = args.device
    xcf.r_par_max = args.rp_max
    xcf.r_par_min = args.rp_min
    xcf.r_trans_max = args.rt_max
    xcf.z_min_pairs = None
    xcf.z_max_pairs = None
    xcf.zerr_cut_deg = None
    xcf.zerr_cut_kms = None
    xcf.rmu_binning = False
    xcf.ang_correlation = False
    xcf.num_bins_r_par = args.np * args.coef_binning_model
    xcf.num_bins_r_trans = args.nt * args.coef_binning_model
    xcf.num_model_bins_r_par = args.np * args.coef_binning_model
    xcf.num_model_bins_r_trans = args.nt * args.coef_binning_model
    xcf.nside = args.nside
    xcf.z_ref = args.z_ref
    xcf.lambda_abs = constants.ABSORBER_IGM[args.lambda_abs]
    xcf.reject = args.rej
    xcf.alpha = args.z_evol_del
    xcf.alpha_obj = args.z_evol_obj
    xcf.alpha_abs = {args.lambda_abs: args.z_evol_del}
    abs_igm = args.abs_igm or []
    for metal in abs_igm:
        xcf.alpha_abs[metal] = args.metal_alpha

    blinding = io.read_blinding(args.in_dir)
    cosmo = constants.Cosmo(Om=args.fid_Om, Or=args.fid_Or, Ok=args.fid_Ok, wl=args.fid_wl)
    xcf.cosmo = cosmo
    t0 = time.time()
    data, num_data, z_min, z_max = io.read_deltas(
        args.in_dir, args.nside, xcf.lambda_abs, args.z_evol_del, args.z_ref, cosmo,
        max_num_spec=args.nspec, nproc=args.nproc, rebin_factor=args.rebin_factor,
        z_min_qso=args.z_min_sources, z_max_qso=args.z_max_sources,
        delta_attributes=args.in_attributes)
    xcf.data = data
    xcf.num_data = num_data
    if args.z_min_obj is None:
        r_comov_min = max(0.0, float(cosmo.get_r_comov(z_min)) + xcf.r_par_min)
        args.z_min_obj = float(cosmo.distance_to_redshift(r_comov_min))
        userprint("z_min_obj = {}".format(args.z_min_obj), end="")
    if args.z_max_obj is None:
        r_comov_max = max(0.0, float(cosmo.get_r_comov(z_max)) + xcf.r_par_max)
        args.z_max_obj = float(cosmo.distance_to_redshift(r_comov_max))
        userprint("z_max_obj = {}".format(args.z_max_obj), end="")
    objs, z_min2 = io.read_objects(args.drq, args.nside, args.z_min_obj, args.z_max_obj,
                                   args.z_evol_obj, args.z_ref, cosmo, mode=args.mode)
    xcf.objs = objs
    xcf.ang_max = utils.compute_ang_max(cosmo, xcf.r_trans_max, z_min, z_min2)
    t1 = time.time()
    userprint(f"picca_metal_xdmat.py - Time reading data: {(t1 - t0) / 60:.3f} minutes")

    cpu_data = {}
    for index, healpix in enumerate(sorted(data)):
        cpu_data.setdefault(index % args.nproc, []).append(healpix)

    results = {"dmat": [], "weights_dmat": [], "r_par": [], "r_trans": [], "z": [],
               "names": [], "num_pairs": [], "num_pairs_used": []}
    reports = []
    for metal in abs_igm:
        dmat_data = [calc_metal_dmat(metal, lst) for lst in sorted(cpu_data.values())]
        reports.append(xcf.last_report)
        weights_dmat = np.array([item[0] for item in dmat_data]).sum(axis=0)
        dmat = np.array([item[1] for item in dmat_data]).sum(axis=0)
        r_par = np.array([item[2] for item in dmat_data]).sum(axis=0)
        r_trans = np.array([item[3] for item in dmat_data]).sum(axis=0)
        z = np.array([item[4] for item in dmat_data]).sum(axis=0)
        weights = np.array([item[5] for item in dmat_data]).sum(axis=0)
        w = weights > 0
        r_par[w] /= weights[w]
        r_trans[w] /= weights[w]
        z[w] /= weights[w]
        num_pairs = np.array([item[6] for item in dmat_data]).sum(axis=0)
        num_pairs_used = np.array([item[7] for item in dmat_data]).sum(axis=0)
        w = weights_dmat > 0
        dmat[w, :] /= weights_dmat[w, None]
        results["dmat"].append(dmat)
        results["weights_dmat"].append(weights_dmat)
        results["r_par"].append(r_par)
        results["r_trans"].append(r_trans)
        results["z"].append(z)
        results["names"].append(metal)
        results["num_pairs"].append(num_pairs)
        results["num_pairs_used"].append(num_pairs_used)
    t2 = time.time()
    userprint(f"picca_metal_xdmat.py - Time computing all metal matrix: "
              f"{(t2 - t1) / 60:.3f} minutes")

    header = [
        {"name": "RPMIN", "value": xcf.r_par_min, "comment": "Minimum r-parallel [h^-1 Mpc]"},
        {"name": "RPMAX", "value": xcf.r_par_max, "comment": "Maximum r-parallel [h^-1 Mpc]"},
        {"name": "RTMAX", "value": xcf.r_trans_max,
         "comment": "Maximum r-transverse [h^-1 Mpc]"},
        {"name": "NP", "value": xcf.num_bins_r_par, "comment": "Number of bins in r-parallel"},
        {"name": "NT", "value": xcf.num_bins_r_trans,
         "comment": "Number of bins in r-transverse"},
        {"name": "COEFMOD", "value": args.coef_binning_model,
         "comment": "Coefficient for model binning"},
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
    write_metal_file(args.out, header, results, blinding, prefix=args.obj_name + "_")
    t3 = time.time()
    userprint(f"picca_metal_xdmat.py - Time total : {(t3 - t0) / 60:.3f} minutes")
    return reports


if __name__ == "__main__":
    main(sys.argv[1:])
