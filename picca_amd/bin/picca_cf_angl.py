#!/usr/bin/env python3
"""Auto- and cross-correlation of delta fields in wavelength ratio and angle, on the
MI355X engine.

Command-line mirror of ``py/picca/bin/picca_cf_angl.py``: same flags and defaults
(:35-205), module globals (:172-187: wavelength-ratio range for r_par, the maximum angle
for r_trans, ``ang_correlation``), deltas read without a cosmology (:188-233) and the same
``ATTRI`` / ``COR`` output (:263-296).  All healpix rows run in one GPU launch.
"""
import argparse
import sys
import time

import numpy as np

from picca_amd import cf, constants, io
from picca_amd.bin.picca_cf import reset_globals
from picca_amd.fitsio_lite import FITSWriter
from picca_amd.utils import userprint


def parse_args(cmdargs=None):
    parser = argparse.ArgumentParser(
        formatter_class=argparse.ArgumentDefaultsHelpFormatter,
        description="Compute the auto and cross-correlation of delta fields as a function "
                    "of angle and wavelength ratio")
    parser.add_argument("--out", type=str, default=None, required=True, help="Output file name")
    parser.add_argument("--in-dir", type=str, default=None, required=True,
                        help="Directory to delta files")
    parser.add_argument("--in-dir2", type=str, default=None, required=False,
                        help="Directory to 2nd delta files")
    parser.add_argument("--wr-min", type=float, default=1.0, required=False,
                        help="Min of wavelength ratio")
    parser.add_argument("--wr-max", type=float, default=1.1, required=False,
                        help="Max of wavelength ratio")
    parser.add_argument("--ang-max", type=float, default=0.02, required=False,
                        help="Max angle (rad)")
    parser.add_argument("--np", type=int, default=50, required=False,
                        help="Number of wavelength ratio bins")
    parser.add_argument("--nt", type=int, default=50, required=False,
                        help="Number of angular bins")
    parser.add_argument("--z-min-sources", type=float, default=0.0, required=False,
                        help="Limit the minimum redshift of the quasars")
    parser.add_argument("--z-max-sources", type=float, default=10.0, required=False,
                        help="Limit the maximum redshift of the quasars")
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
    parser.add_argument("--in-attributes", type=str, default=None, required=False,
                        help="Delta attributes file holding the continuum-fit order")
    parser.add_argument("--device", type=int, default=0, required=False,
                        help="HIP device (engine option)")
    return parser.parse_args(cmdargs)


def main(cmdargs=None):
    args = parse_args(cmdargs)
    reset_globals()
    cf.device = args.device
    cf.r_par_min = args.wr_min
    cf.r_par_max = args.wr_max
    cf.r_trans_max = args.ang_max
    cf.num_bins_r_par = args.np
    cf.num_bins_r_trans = args.nt
    cf.nside = args.nside
    cf.z_ref = args.z_ref
    cf.alpha = args.z_evol
    cf.x_correlation = False
    cf.ang_correlation = True
    cf.ang_max = args.ang_max
    cf.lambda_abs = constants.ABSORBER_IGM[args.lambda_abs]
    cf.remove_same_half_plate_close_pairs = args.remove_same_half_plate_close_pairs

    t0 = time.time()
    data, num_data, z_min, z_max = io.read_deltas(
        args.in_dir, cf.nside, cf.lambda_abs, cf.alpha, cf.z_ref, None,
        max_num_spec=args.nspec, no_project=args.no_project, z_min_qso=args.z_min_sources,
        z_max_qso=args.z_max_sources, delta_attributes=args.in_attributes)
    cf.data = data
    cf.num_data = num_data
    del z_min, z_max
    userprint("")
    userprint("done, npix = {}".format(len(data)))

    if args.in_dir2 or args.lambda_abs2:
        cf.x_correlation = True
        cf.alpha2 = args.z_evol2
        if args.in_dir2 is None:
            args.in_dir2 = args.in_dir
        if args.lambda_abs2:
            cf.lambda_abs2 = constants.ABSORBER_IGM[args.lambda_abs2]
        else:
            cf.lambda_abs2 = cf.lambda_abs
        data2, num_data2, z_min2, z_max2 = io.read_deltas(
            args.in_dir2, cf.nside, cf.lambda_abs2, cf.alpha2, cf.z_ref, None,
            max_num_spec=args.nspec, no_project=args.no_project, z_min_qso=args.z_min_sources,
            z_max_qso=args.z_max_sources, delta_attributes=args.in_attributes)
        cf.data2 = data2
        cf.num_data2 = num_data2
        del z_min2, z_max2
        userprint("")
        userprint("done, npix = {}".format(len(data2)))

    healpix_list = np.array(sorted(data))
    rows = np.array(cf.compute_xi_rows(healpix_list))
    t1 = time.time()
    userprint(f"picca_cf_angl.py - Time computing correlation function: {(t1 - t0) / 60:.3f} "
              "minutes")
    weights_list = rows[:, 0, :]
    xi_list = rows[:, 1, :]
    r_par_list = rows[:, 2, :]
    r_trans_list = rows[:, 3, :]
    z_list = rows[:, 4, :]
    num_pairs_list = rows[:, 5, :].astype(np.int64)

    w = weights_list.sum(axis=0) > 0.0
    r_par = (r_par_list * weights_list).sum(axis=0)
    r_par[w] /= weights_list.sum(axis=0)[w]
    r_trans = (r_trans_list * weights_list).sum(axis=0)
    r_trans[w] /= weights_list.sum(axis=0)[w]
    z = (z_list * weights_list).sum(axis=0)
    z[w] /= weights_list.sum(axis=0)[w]
    num_pairs = num_pairs_list.sum(axis=0)

    header = [
        {"name": "RPMIN", "value": cf.r_par_min, "comment": "Minimum wavelength ratio"},
        {"name": "RPMAX", "value": cf.r_par_max, "comment": "Maximum wavelength ratio"},
        {"name": "RTMAX", "value": cf.r_trans_max, "comment": "Maximum angle [rad]"},
        {"name": "NP", "value": cf.num_bins_r_par,
         "comment": "Number of bins in wavelength ratio"},
        {"name": "NT", "value": cf.num_bins_r_trans, "comment": "Number of bins in angle"},
        {"name": "NSIDE", "value": cf.nside, "comment": "Healpix nside"},
    ]
    results = FITSWriter(args.out)
    results.write([r_par, r_trans, z, num_pairs], names=["RP", "RT", "Z", "NB"],
                  comment=["Wavelength ratio", "Angle", "Redshift", "Number of pairs"],
                  units=["", "rad", "", ""], header=header, extname="ATTRI")
    results.write([healpix_list, weights_list, xi_list], names=["HEALPID", "WE", "DA"],
                  comment=["Healpix index", "Sum of weight", "Correlation"],
                  header=[{"name": "HLPXSCHM", "value": "RING", "comment": "Healpix scheme"}],
                  extname="COR")
    results.close()
    return cf.last_report


if __name__ == "__main__":
    main(sys.argv[1:])
