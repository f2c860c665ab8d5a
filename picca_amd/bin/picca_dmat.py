#!/usr/bin/env python3
"""picca_dmat on the MI355X engine: the distortion matrix of the correlation of delta fields,
written as picca's DMAT + ATTRI file.

Command line and products as ``py/picca/bin/picca_dmat.py``.  The healpix pixels are split
into ``--nproc`` round-robin lists, each seeding numpy's legacy generator with its first
pixel before its rejection draws, so the kept pairs are the reference's for the same
``--nproc``; each list is one batched GPU computation (``cf.compute_dmat``).
"""
from picca_amd import cf, constants, io
from picca_amd import driver as D

# the matrix is always referred to z = 2.25 (py/picca/bin/picca_dmat.py:377)
MATRIX_Z_REF = 2.25


def settings(opts, fields):
    s = D.Settings(D.binning_settings(opts, model_coef=opts.coef_binning_model))
    s.update(D.pair_cut_settings(opts))
    s.update(z_ref=MATRIX_Z_REF, alpha=opts.z_evol, alpha2=opts.z_evol, reject=opts.rej,
             lambda_abs=constants.ABSORBER_IGM[opts.lambda_abs],
             remove_same_half_plate_close_pairs=opts.remove_same_half_plate_close_pairs,
             redshift_evolution_in_distortion_matrix=not opts.no_redshift_evolution)
    s.update(fields.settings(opts.z_evol2))
    return s


def main(cmdargs=None):
    opts = D.parse("picca_dmat", cmdargs)
    opts.z_ref = MATRIX_Z_REF
    clock = D.Clock("picca_dmat")
    cf.device = opts.device
    D.process_count(opts)
    D.pair_cut_settings(opts)
    blinding = io.read_blinding(opts.in_dir)
    fields = D.Fields(opts, D.cosmology(opts), opts.z_evol, opts.z_evol2)
    settings(opts, fields).apply(cf)
    clock.lap("read")

    lists = D.process_lists(cf.data, opts.nproc)
    parts = D.seeded_lists(cf, lists, cf.compute_dmat)
    clock.lap("matrix")
    weights_dmat, dmat, r_par, r_trans, z, weights, npall, npused = D.add_parts(parts)
    mean_z = D.normalise_matrix(weights_dmat, dmat, r_par, r_trans, z, weights)
    if cf.redshift_evolution_in_distortion_matrix:
        D.redshift_rescale(dmat, z, mean_z, cf.z_ref, cf.alpha, cf.alpha2)

    cards = D.header(RPMIN=cf.r_par_min, RPMAX=cf.r_par_max, RTMAX=cf.r_trans_max,
                     NP=cf.num_bins_r_par, NT=cf.num_bins_r_trans,
                     COEFMOD=opts.coef_binning_model, ZMIN=cf.z_min_pairs,
                     ZMAX=cf.z_max_pairs, REJ=cf.reject, NPALL=int(npall),
                     NPUSED=int(npused))
    cards += D.cosmo_cards(opts) + D.header(BLINDING=blinding, RMU_BIN=cf.rmu_binning)
    D.matrix_products(opts.out, (weights_dmat, dmat, r_par, r_trans, z), cards, blinding)
    clock.total()
    return cf.last_report


if __name__ == "__main__":
    D.main_entry(main)
