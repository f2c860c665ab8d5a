#!/usr/bin/env python3
"""Export of a correlation function with its covariance and distortion matrix.

Command-line mirror of ``py/picca/bin/picca_export.py`` for the correlation path: reads the
COR rows (``DA`` or ``DA_BLIND``, ``WE``) and attributes of a picca_cf / picca_xcf file,
computes the subsampling covariance on the GPU (``utils.compute_cov``,
``picca_cov_subsample``) or the bootstrap one, smooths it unless ``--do-not-smooth-cov``
(:270-294), takes the weighted mean of the rows (:296-300), reads the distortion matrix
(:307-341) and writes ``COR`` and ``DMATTRI`` (:404-577).  ``--cov``/``--cor`` and
``--remove-shuffled-correlation`` are supported; the multipole conversion and DESI's
blinding templates (site files at NERSC) are not part of this engine and are refused.
"""
import argparse
import sys

import numpy as np
import scipy.linalg

from picca_amd.fitsio_lite import FITS, FITSWriter
from picca_amd.utils import compute_cov, compute_cov_boot, smooth_cov, userprint

# blinding strategies whose correlations are not blinded
# (py/picca/bin/picca_export.py:26)
UNBLINDABLE_STRATEGIES = ["none", "desi_m2", "desi_y1", "desi_y3"]


def parse_args(cmdargs=None):
    parser = argparse.ArgumentParser(
        formatter_class=argparse.ArgumentDefaultsHelpFormatter,
        description="Export auto and cross-correlation for the fitter.")
    parser.add_argument("--data", type=str, default=None, required=True,
                        help="Correlation produced via picca_cf.py, picca_xcf.py, ...")
    parser.add_argument("--out", type=str, default=None, required=True, help="Output file name")
    parser.add_argument("--dmat", type=str, default=None, required=False,
                        help="Distortion matrix produced via picca_dmat.py, picca_xdmat.py... "
                             "(if not provided will be identity)")
    parser.add_argument("--multipoles", action="store_true", help="Use multipole extension")
    parser.add_argument("--lmax-model", type=int, default=6, help="Maximum model multipole")
    parser.add_argument("--lmax-data", type=int, default=4, help="Maximum data multipole")
    parser.add_argument("--cov", type=str, default=None, required=False,
                        help="Covariance matrix (if not provided will be calculated by "
                             "subsampling)")
    parser.add_argument("--cor", type=str, default=None, required=False,
                        help="Correlation matrix (if not provided will be calculated by "
                             "subsampling)")
    parser.add_argument("--remove-shuffled-correlation", type=str, default=None,
                        required=False,
                        help="Remove a correlation from shuffling the distribution of los")
    parser.add_argument("--num-boot-cov", type=int, default=0,
                        help="Number of bootstrap realizations. <=0 turns it off.")
    parser.add_argument("--do-not-smooth-cov", action="store_true", default=False,
                        help="Do not smooth the covariance matrix")
    parser.add_argument("--smooth-per-r-par", action="store_true", default=False,
                        help="Consider different correlation coefficients per r_par")
    parser.add_argument("--blind-corr-type", default=None,
                        choices=["lyaxlya", "lyaxlyb", "qsoxlya", "qsoxlyb", "lyaxqso",
                                 "lybxqso"],
                        help="Type of correlation. Required to apply blinding in DESI")
    return parser.parse_args(cmdargs)


def main(cmdargs=None):
    args = parse_args(cmdargs)
    if args.multipoles:
        raise NotImplementedError("--multipoles is not part of this engine")
    hdul = FITS(args.data)
    r_par = np.array(hdul[1]["RP"][:])
    r_trans = np.array(hdul[1]["RT"][:])
    z = np.array(hdul[1]["Z"][:])
    num_pairs = np.array(hdul[1]["NB"][:])
    weights = np.array(hdul[2]["WE"][:])
    if "DA_BLIND" in hdul[2].get_colnames():
        xi = np.array(hdul[2]["DA_BLIND"][:])
        data_name = "DA_BLIND"
    else:
        xi = np.array(hdul[2]["DA"][:])
        data_name = "DA"
    head = hdul[1].read_header()
    num_bins_r_par = head["NP"]
    num_bins_r_trans = head["NT"]
    r_trans_max = head["RTMAX"]
    r_par_min = head["RPMIN"]
    r_par_max = head["RPMAX"]
    nsamples = xi.shape[0]
    blinding = head["BLINDING"] if "BLINDING" in head else "none"
    if blinding == "minimal":
        blinding = "corr_yshift"
        userprint("The minimal strategy is no longer supported."
                  "Automatically switch to corr_yshift.")

    if args.remove_shuffled_correlation is not None:
        sh = FITS(args.remove_shuffled_correlation)
        xi_shuffled = sh["COR"][data_name][:]
        weight_shuffled = sh["COR"]["WE"][:]
        xi_shuffled = (xi_shuffled * weight_shuffled).sum(axis=1)
        weight_shuffled = weight_shuffled.sum(axis=1)
        w = weight_shuffled > 0.0
        xi_shuffled[w] /= weight_shuffled[w]
        xi -= xi_shuffled[:, None]

    if args.cov is not None:
        userprint("INFO: The covariance-matrix will be read from file: {}".format(args.cov))
        covariance = FITS(args.cov)[1]["CO"][:]
    elif args.cor is not None:
        userprint("INFO: The correlation-matrix will be read from file: {}".format(args.cor))
        correlation = FITS(args.cor)[1]["CO"][:]
        if ((correlation.min() < -1.0) or (correlation.min() > 1.0)
                or (correlation.max() < -1.0) or (correlation.max() > 1.0)
                or np.any(np.diag(correlation) != 1.0)):
            userprint("WARNING: The correlation-matrix has some incorrect values")
        var = np.diagonal(correlation)
        correlation = correlation / np.sqrt(var * var[:, None])
        covariance = compute_cov(xi, weights)
        var = np.diagonal(covariance)
        covariance = correlation * np.sqrt(var * var[:, None])
    else:
        delta_r_par = (r_par_max - r_par_min) / num_bins_r_par
        delta_r_trans = (r_trans_max - 0.0) / num_bins_r_trans
        if args.num_boot_cov > 0:
            userprint(f"INFO: Covariance with {args.num_boot_cov} bootstrap realizations.")
            covariance = compute_cov_boot(xi, weights, nboots=args.num_boot_cov)
        else:
            covariance = compute_cov(xi, weights)
        if args.do_not_smooth_cov:
            userprint("INFO: The covariance will not be smoothed")
        else:
            userprint("INFO: The covariance will be smoothed")
            covariance = smooth_cov(None, None, r_par, r_trans, delta_r_trans=delta_r_trans,
                                    delta_r_par=delta_r_par, covariance=covariance,
                                    per_r_par=args.smooth_per_r_par)

    xi = (xi * weights).sum(axis=0)
    weights = weights.sum(axis=0)
    w = weights > 0
    xi[w] /= weights[w]
    try:
        scipy.linalg.cholesky(covariance)
    except scipy.linalg.LinAlgError:
        userprint("WARNING: Covariance matrix is not positive definite")

    if args.dmat is not None:
        dm = FITS(args.dmat)
        cols = dm[1].get_colnames()
        if data_name == "DA_BLIND" and "DM_BLIND" in cols:
            dmat = np.array(dm[1]["DM_BLIND"][:])
            dmat_name = "DM_BLIND"
        elif data_name == "DA_BLIND":
            userprint("Blinded correlations were given but distortion matrix is unblinded. "
                      "These files should not mix. Exiting...")
            sys.exit(1)
        elif "DM_BLIND" in cols:
            userprint("Non-blinded correlations were given but distortion matrix is blinded. "
                      "These files should not mix. Exiting...")
            sys.exit(1)
        else:
            dmat = np.array(dm[1]["DM"][:])
            dmat_name = "DM"
        try:
            r_par_dmat = np.array(dm[2]["RP"][:])
            r_trans_dmat = np.array(dm[2]["RT"][:])
            z_dmat = np.array(dm[2]["Z"][:])
        except (IndexError, KeyError):
            r_par_dmat, r_trans_dmat, z_dmat = r_par.copy(), r_trans.copy(), z.copy()
        if dmat.shape == (xi.size, xi.size):
            r_par_dmat, r_trans_dmat, z_dmat = r_par.copy(), r_trans.copy(), z.copy()
    else:
        dmat = np.eye(len(xi))
        r_par_dmat, r_trans_dmat, z_dmat = r_par.copy(), r_trans.copy(), z.copy()
        dmat_name = "DM_EMPTY"

    if blinding in UNBLINDABLE_STRATEGIES:
        if blinding != "none":
            userprint(f"'{blinding}' correlations are not blinded.")
        blinding = "none"
        data_name = "DA"
        dmat_name = "DM"
    else:
        raise NotImplementedError("blinding '%s' needs DESI's blinding templates, which are "
                                  "not part of this engine" % blinding)

    header = [
        {"name": "BLINDING", "value": blinding,
         "comment": "String specifying the blinding strategy"},
        {"name": "RPMIN", "value": r_par_min, "comment": "Minimum r-parallel"},
        {"name": "RPMAX", "value": r_par_max, "comment": "Maximum r-parallel"},
        {"name": "RTMAX", "value": r_trans_max, "comment": "Maximum r-transverse"},
        {"name": "NP", "value": num_bins_r_par, "comment": "Number of bins in r-parallel"},
        {"name": "NT", "value": num_bins_r_trans, "comment": "Number of bins in r-transverse"},
    ]
    for key, comment in (("ZMIN", "Minimum redshift of pairs"),
                         ("ZMAX", "Maximum redshift of pairs"),
                         ("OMEGAM", "Omega_matter(z=0) of fiducial LambdaCDM cosmology"),
                         ("OMEGAR", "Omega_radiation(z=0) of fiducial LambdaCDM cosmology"),
                         ("OMEGAK", "Omega_k(z=0) of fiducial LambdaCDM cosmology"),
                         ("WL", "Equation of state of dark energy of fiducial LambdaCDM "
                                "cosmology")):
        if key in head:
            header.append({"name": key, "value": head[key], "comment": comment})
    header += [
        {"name": "RMU_BIN", "value": bool("RMU_BIN" in head and head["RMU_BIN"]),
         "comment": "True if binned in r, mu"},
        {"name": "NSAMPLES", "value": nsamples, "comment": "Number of samples"},
        {"name": "MLTPOLE", "value": False, "comment": "Specifying if in multipoles."},
    ]
    out = FITSWriter(args.out)
    out.write([xi, r_par, r_trans, z, covariance, dmat, num_pairs],
              names=[data_name, "RP", "RT", "Z", "CO", dmat_name, "NB"],
              comment=["Correlation", "R-parallel", "R-transverse", "Redshift",
                       "Covariance matrix", "Distortion matrix", "Number of pairs"],
              header=header, extname="COR")
    out.write([r_par_dmat, r_trans_dmat, z_dmat], names=["DMRP", "DMRT", "DMZ"],
              comment=["R-parallel model", "R-transverse model", "Redshift model"],
              extname="DMATTRI")
    out.close()


if __name__ == "__main__":
    main(sys.argv[1:])
