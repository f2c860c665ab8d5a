#!/usr/bin/env python3
"""picca_xdmat on the MI355X engine: the distortion matrix of the delta x object
cross-correlation, written as picca's DMAT + ATTRI file.

Command line and products as ``py/picca/bin/picca_xdmat.py``; the healpix pixels run as
seeded round-robin lists like picca_dmat (``xcf.compute_dmat`` per list).
"""
from picca_amd import constants, io, utils, xcf
from picca_amd import driver as D

MATRIX_Z_REF = 2.25


def main(cmdargs=None):
    opts = D.parse("picca_xdmat", cmdargs)
    opts.z_ref = MATRIX_Z_REF
    clock = D.Clock("picca_xdmat")
    xcf.device = opts.device
    D.process_count(opts)
    s = D.Settings(D.binning_settings(opts, model_coef=opts.coef_binning_model, signed=True))
    s.update(D.pair_cut_settings(opts))
    s.update(redshift_evolution_in_distortion_matrix=not opts.no_redshift_evolution,
             z_ref=MATRIX_Z_REF, alpha=opts.z_evol_del, alpha_obj=opts.z_evol_obj,
             lambda_abs=constants.ABSORBER_IGM[opts.lambda_abs], reject=opts.rej,
             ang_correlation=False)

    blinding = io.read_blinding(opts.in_dir)
    cosmo = D.cosmology(opts, blinding=blinding)
    data, count, z_low, z_high = D.load_deltas(opts, opts.in_dir, opts.lambda_abs,
                                               opts.z_evol_del, cosmo)
    D.object_redshift_limits(opts, cosmo, z_low, z_high)
    objs, z_low_obj = D.load_objects(opts, cosmo)
    s.update(data=data, num_data=count, objs=objs,
             ang_max=utils.compute_ang_max(cosmo, opts.rt_max, z_low, z_low_obj))
    s.apply(xcf, reset=False)
    clock.lap("read")

    parts = D.seeded_lists(xcf, D.process_lists(xcf.data, opts.nproc), xcf.compute_dmat)
    clock.lap("matrix")
    weights_dmat, dmat, r_par, r_trans, z, weights, npall, npused = D.add_parts(parts)
    mean_z = D.normalise_matrix(weights_dmat, dmat, r_par, r_trans, z, weights)
    if xcf.redshift_evolution_in_distortion_matrix:
        D.redshift_rescale(dmat, z, mean_z, xcf.z_ref, xcf.alpha, xcf.alpha_obj)

    cards = D.header(RPMIN=xcf.r_par_min, RPMAX=xcf.r_par_max, RTMAX=xcf.r_trans_max,
                     NP=xcf.num_bins_r_par, NT=xcf.num_bins_r_trans,
                     COEFMOD=opts.coef_binning_model, ZMIN=xcf.z_min_pairs,
                     ZMAX=xcf.z_max_pairs, REJ=xcf.reject, NPALL=int(npall),
                     NPUSED=int(npused))
    cards += D.cosmo_cards(opts) + D.header(BLINDING=blinding, RMU_BIN=xcf.rmu_binning)
    D.matrix_products(opts.out, (weights_dmat, dmat, r_par, r_trans, z), cards, blinding)
    clock.total()
    return xcf.last_report


if __name__ == "__main__":
    D.main_entry(main)
