#!/usr/bin/env python3
"""picca_cf on the MI355X engine: the 3D correlation of one delta field with itself or with
a second field, written as picca's ATTRI + COR file.

Command line and products as ``py/picca/bin/picca_cf.py``; every healpix pixel of the first
field is one output row of a single pair-kernel launch (``cf.compute_xi_rows``).
"""
from picca_amd import cf, constants, io, utils
from picca_amd import driver as D


def settings(opts, fields, blinding):
    s = D.Settings(D.binning_settings(opts))
    s.update(D.pair_cut_settings(opts))
    s.update(z_ref=opts.z_ref, alpha=opts.z_evol,
             lambda_abs=constants.ABSORBER_IGM[opts.lambda_abs],
             remove_same_half_plate_close_pairs=opts.remove_same_half_plate_close_pairs)
    s.update(fields.settings(opts.z_evol2))
    return s


def main(cmdargs=None):
    opts = D.parse("picca_cf", cmdargs)
    clock = D.Clock("picca_cf")
    cf.device = opts.device
    D.pair_cut_settings(opts)  # option errors before any reading
    blinding = io.read_blinding(opts.in_dir)
    cosmo = D.cosmology(opts, blinding=blinding)
    fields = D.Fields(opts, cosmo, opts.z_evol, opts.z_evol2, project=not opts.no_project)
    settings(opts, fields, blinding).apply(cf)
    if opts.shuffle_distrib_forest_seed is not None:
        cf.data = utils.shuffle_distrib_forests(cf.data, opts.shuffle_distrib_forest_seed)
    clock.lap("read")

    healpix, rows = D.healpix_rows(cf, cf.data)
    clock.lap("pairs")
    if cf.last_report is not None:
        D.say("engine:", cf.last_report)

    cards = D.header(RPMIN=cf.r_par_min, RPMAX=cf.r_par_max, RTMAX=cf.r_trans_max,
                     NP=cf.num_bins_r_par, NT=cf.num_bins_r_trans, ZMIN=cf.z_min_pairs,
                     ZMAX=cf.z_max_pairs, NSIDE=cf.nside)
    cards += D.cosmo_cards(opts) + D.header(BLINDING=blinding, RMU_BIN=cf.rmu_binning)
    D.correlation_products(opts.out, D.combine_rows(rows), healpix, cards, blinding,
                           rmu=cf.rmu_binning)
    clock.total()
    return cf.last_report


if __name__ == "__main__":
    D.main_entry(main)
