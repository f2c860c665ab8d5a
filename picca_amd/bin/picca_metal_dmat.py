#!/usr/bin/env python3
"""Metal distortion matrices of the auto- and cross-correlation of delta fields on the
MI355X engine.

Command-line mirror of ``py/picca/bin/picca_metal_dmat.py``: same flags and defaults, module
globals (:333-350), reading (:352-421), the split of the healpix list into ``--nproc``
process lists, each of which fills its neighbours and seeds numpy's legacy stream with its
first healpix before its rejection draws (``calc_metal_dmat``, :20-41), the loop over the
absorber pairs (:445-462, with ``--fast-metals``), the merge and normalisation (:470-500)
and the ``ATTRI`` / ``MDMAT`` output (:510-641).  Each process list runs as one batched
GPU computation per absorber pass (``cf.compute_metal_dmat``).
"""
import argparse
import sys
import time
from multiprocessing import cpu_count

import numpy as np

from picca_amd import cf, constants, io, utils
from picca_amd.bin.picca_cf import reset_globals
from picca_amd.fitsio_lite import FITSWriter
from picca_amd.utils import userprint

# the silicon metals modelled in the CF/XCF (py/picca/bin/picca_metal_dmat.py:18)
DEFAULT_SI_METALS = ["SiIII(1207)", "SiII(1190)", "SiII(1193)", "SiII(1260)"]


def parse_args(cmdargs=None):
    parser = argparse.ArgumentParser(
        formatter_class=argparse.ArgumentDefaultsHelpFormatter,
        description="Compute the auto and cross-correlation of delta fields for a list of "
                    "IGM absorption.")
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
    parser.add_argument("--coef-binning-model", type=int, default=1, required=False,
                        help="Coefficient multiplying np and nt to get finner binning for the "
                             "model")
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
    parser.add_argument("--abs-igm", type=str, default=[], required=False, nargs="*",
                        help="List of names of metal absorption in picca.constants present in "
                             "forest")
    parser.add_argument("--abs-igm2", type=str, default=[], required=False, nargs="*",
                        help="List of names of metal absorption in picca.constants present in "
                             "2nd forest")
    parser.add_argument("--z-ref", type=float, default=2.25, required=False,
                        help="Reference redshift")
    parser.add_argument("--z-evol", type=float, default=2.9, required=False,
                        help="Exponent of the redshift evolution of the delta field")
    parser.add_argument("--z-evol2", type=float, default=2.9, required=False,
                        help="Exponent of the redshift evolution of the 2nd delta field")
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
    parser.add_argument("--remove-same-half-plate-close-pairs", action="store_true",
                        required=False,
                        help="Reject pairs in the first bin in r-parallel from same half plate")
    parser.add_argument("--rej", type=float, default=1.0, required=False,
                        help="Fraction of rejected forest-forest pairs: -1=no rejection, "
                             "1=all rejection")
    parser.add_argument("--nside", type=int, default=16, required=False, help="Healpix nside")
    parser.add_argument("--nproc", type=int, default=None, required=False,
                        help="Number of process lists (fixes the rejection draws)")
    parser.add_argument("--nspec", type=int, default=None, required=False,
                        help="Maximum number of spectra to read")
    parser.add_argument("--unfold-cf", action="store_true", required=False,
                        help="rp can be positive or negative depending on the relative "
                             "position between absorber1 and absorber2")
    parser.add_argument("--rebin-factor", type=int, default=None, required=False,
                        help="Rebin factor for deltas. If not None, deltas will be rebinned "
                             "by that factor")
    parser.add_argument("--fast-metals", action="store_true", required=False,
                        help="compute only the metal correlations used by Vega, i.e. 4 LyaxSi "
                             "matrices and CIVxCIV")
    parser.add_argument("--in-attributes", type=str, default=None, required=False,
                        help="Delta attributes file holding the continuum-fit order")
    parser.add_argument("--device", type=int, default=0, required=False,
                        help="HIP device (engine option)")
    return parser.parse_args(cmdargs)


def calc_metal_dmat(abs_igm1, abs_igm2, healpixs):
    """One process list (py/picca/bin/picca_metal_dmat.py:20-41)."""
    cf.fill_neighs(healpixs)
    np.random.seed(healpixs[0])
    return cf.compute_metal_dmat(healpixs, abs_igm1=abs_igm1, abs_igm2=abs_igm2)


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
    cf.num_bins_r_par = args.np * args.coef_binning_model
    cf.num_bins_r_trans = args.nt * args.coef_binning_model
    cf.num_model_bins_r_par = args.np * args.coef_binning_model
    cf.num_model_bins_r_trans = args.nt * args.coef_binning_model
    cf.nside = args.nside
    cf.z_ref = args.z_ref
    cf.alpha = args.z_evol
    cf.reject = args.rej
    cf.lambda_abs = constants.ABSORBER_IGM[args.lambda_abs]
    cf.remove_same_half_plate_close_pairs = args.remove_same_half_plate_close_pairs
    cf.alpha_abs = {args.lambda_abs: cf.alpha}
    for metal in args.abs_igm:
        cf.alpha_abs[metal] = args.metal_alpha

    blinding = io.read_blinding(args.in_dir)
    cf.cosmo = constants.Cosmo(Om=args.fid_Om, Or=args.fid_Or, Ok=args.fid_Ok, wl=args.fid_wl)
    t0 = time.time()
    data, num_data, z_min, z_max = io.read_deltas(
        args.in_dir, cf.nside, cf.lambda_abs, cf.alpha, cf.z_ref, cf.cosmo,
        max_num_spec=args.nspec, nproc=args.nproc, rebin_factor=args.rebin_factor,
        z_min_qso=args.z_min_sources, z_max_qso=args.z_max_sources,
        delta_attributes=args.in_attributes)
    del z_max
    cf.data = data
    cf.num_data = num_data
    cf.ang_max = utils.compute_ang_max(cf.cosmo, cf.r_trans_max, z_min)
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
        cf.alpha_abs[args.lambda_abs2] = cf.alpha2
        for metal in args.abs_igm2:
            cf.alpha_abs[metal] = args.metal_alpha
        data2, num_data2, z_min2, z_max2 = io.read_deltas(
            args.in_dir2, cf.nside, cf.lambda_abs2, cf.alpha2, cf.z_ref, cf.cosmo,
            max_num_spec=args.nspec, nproc=args.nproc, rebin_factor=args.rebin_factor,
            z_min_qso=args.z_min_sources, z_max_qso=args.z_max_sources,
            delta_attributes=args.in_attributes)
        del z_max2
        cf.data2 = data2
        cf.num_data2 = num_data2
        cf.ang_max = utils.compute_ang_max(cf.cosmo, cf.r_trans_max, z_min, z_min2)
        userprint("")
        userprint("done, npix = {}".format(len(data2)))

    t1 = time.time()
    userprint(f"picca_metal_dmat.py - Time reading data: {(t1 - t0) / 60:.3f} minutes")

    cpu_data = {}
    for index, healpix in enumerate(sorted(data)):
        cpu_data.setdefault(index % args.nproc, []).append(healpix)

    abs_igm = [args.lambda_abs] + args.abs_igm
    userprint("abs_igm = {}".format(abs_igm))
    if args.lambda_abs2 is None:
        args.lambda_abs2 = args.lambda_abs
        args.abs_igm2 = args.abs_igm
    abs_igm_2 = [args.lambda_abs2] + args.abs_igm2
    if cf.x_correlation:
        userprint("abs_igm2 = {}".format(abs_igm_2))

    results = {"dmat": [], "weights_dmat": [], "r_par": [], "r_trans": [], "z": [],
               "names": [], "num_pairs": [], "num_pairs_used": []}
    reports = []
    for index1, abs_igm1 in enumerate(abs_igm):
        index0 = index1
        if args.lambda_abs != args.lambda_abs2:
            index0 = 0
        for index2, abs_igm2 in enumerate(abs_igm_2[index0:]):
            if index1 == 0 and index2 == 0:
                continue
            if args.fast_metals:
                if not (abs_igm1 == "LYA" and abs_igm2 in DEFAULT_SI_METALS) and not (
                        abs_igm1 == "CIV(eff)" and abs_igm1 == abs_igm2):
                    continue
            cf.counter = None
            dmat_data = [calc_metal_dmat(abs_igm1, abs_igm2, lst)
                         for lst in sorted(cpu_data.values())]
            reports.append(cf.last_report)
            weights_dmat = np.array([item[0] for item in dmat_data]).sum(axis=0)
            dmat = np.array([item[1] for item in dmat_data]).sum(axis=0)
            r_par = np.array([item[2] for item in dmat_data]).sum(axis=0)
            r_trans = np.array([item[3] for item in dmat_data]).sum(axis=0)
            z = np.array([item[4] for item in dmat_data]).sum(axis=0)
            weights = np.array([item[5] for item in dmat_data]).sum(axis=0)
            num_pairs = np.array([item[6] for item in dmat_data]).sum(axis=0)
            num_pairs_used = np.array([item[7] for item in dmat_data]).sum(axis=0)
            w = weights > 0
            r_par[w] /= weights[w]
            r_trans[w] /= weights[w]
            z[w] /= weights[w]
            w = weights_dmat > 0
            dmat[w, :] /= weights_dmat[w, None]
            results["dmat"].append(dmat)
            results["weights_dmat"].append(weights_dmat)
            results["r_par"].append(r_par)
            results["r_trans"].append(r_trans)
            results["z"].append(z)
            results["names"].append(abs_igm1 + "_" + abs_igm2)
            results["num_pairs"].append(num_pairs)
            results["num_pairs_used"].append(num_pairs_used)

    t2 = time.time()
    userprint(f"picca_metal_dmat.py - Time computing all metal matrices : "
              f"{(t2 - t1) / 60:.3f} minutes")
    header = [
        {"name": "RPMIN", "value": cf.r_par_min, "comment": "Minimum r-parallel [h^-1 Mpc]"},
        {"name": "RPMAX", "value": cf.r_par_max, "comment": "Maximum r-parallel [h^-1 Mpc]"},
        {"name": "RTMAX", "value": cf.r_trans_max, "comment": "Maximum r-transverse [h^-1 Mpc]"},
        {"name": "NP", "value": cf.num_bins_r_par, "comment": "Number of bins in r-parallel"},
        {"name": "NT", "value": cf.num_bins_r_trans,
         "comment": " Number of bins in r-transverse"},
        {"name": "COEFMOD", "value": args.coef_binning_model,
         "comment": "Coefficient for model binning"},
        {"name": "REJ", "value": cf.reject, "comment": "Rejection factor"},
        {"name": "ALPHAMET", "value": args.metal_alpha, "comment": "Evolution of metal bias"},
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
    write_metal_file(args.out, header, results, blinding)
    t3 = time.time()
    userprint(f"picca_metal_dmat.py - Time total : {(t3 - t0) / 60:.3f} minutes")
    return reports


def write_metal_file(path, header, results, blinding, prefix=""):
    """ATTRI (NPALL, NPUSED, ABS_IGM) and MDMAT (RP_, RT_, Z_, DM_, WDM_ per absorber
    pair) of py/picca/bin/picca_metal_dmat.py:580-636; ``prefix`` ("QSO_") names the
    columns of the forest x quasar matrices."""
    names = results["names"]
    len_names = max(len(name) for name in names)
    out = FITSWriter(path)
    out.write([np.array(results["num_pairs"]), np.array(results["num_pairs_used"]),
               np.array(names, dtype="S" + str(len_names))],
              names=["NPALL", "NPUSED", "ABS_IGM"], header=header,
              comment=["Number of pairs", "Number of used pairs", "Absorption name"],
              extname="ATTRI")
    dmat_name = "DM_" if blinding == "none" else "DM_BLIND_"
    out_list, out_names, out_comment, out_units = [], [], [], []
    for index, name in enumerate(names):
        tag = prefix + name
        out_names += ["RP_" + tag, "RT_" + tag, "Z_" + tag, dmat_name + tag, "WDM_" + tag]
        out_list += [results["r_par"][index], results["r_trans"][index], results["z"][index],
                     results["dmat"][index], results["weights_dmat"][index]]
        out_comment += ["R-parallel", "R-transverse", "Redshift", "Distortion matrix",
                        "Sum of weight"]
        out_units += ["h^-1 Mpc", "h^-1 Mpc", "", "", ""]
    out.write(out_list, names=out_names, comment=out_comment, units=out_units,
              extname="MDMAT")
    out.close()


if __name__ == "__main__":
    main(sys.argv[1:])
