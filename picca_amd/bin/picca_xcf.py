#!/usr/bin/env python3
"""picca_xcf on the MI355X engine: the cross-correlation of a delta field with an object
catalogue, written as picca's ATTRI + COR file.

Command line and products as ``py/picca/bin/picca_xcf.py``; every healpix pixel of the
deltas is one output row of a single launch (``xcf.compute_xi_rows``).
"""
from picca_amd import constants, io, prep_del, utils, xcf
from picca_amd import driver as D


def main(cmdargs=None):
    opts = D.parse("picca_xcf", cmdargs)
    clock = D.Clock("picca_xcf")
    xcf.device = opts.device
    s = D.Settings(D.binning_settings(opts, signed=True))
    s.update(D.pair_cut_settings(opts))
    absorber = constants.ABSORBER_IGM[opts.lambda_abs]
    s.update(lambda_abs=absorber, z_ref=opts.z_ref, ang_correlation=False)

    blinding = io.read_blinding(opts.in_dir)
    cosmo = D.cosmology(opts, blinding=blinding)
    objs, z_low_obj = D.load_objects(opts, cosmo)
    data, count, z_low, z_high = D.load_deltas(opts, opts.in_dir, opts.lambda_abs,
                                               opts.z_evol_del, cosmo,
                                               project=not opts.no_project)
    if not opts.no_remove_mean_lambda_obs:
        prep_del.remove_mean_lambda_obs(data, z_low, z_high, absorber)
    if opts.shuffle_distrib_obj_seed is not None:
        objs = utils.shuffle_distrib_forests(objs, opts.shuffle_distrib_obj_seed)
    if opts.shuffle_distrib_forest_seed is not None:
        data = utils.shuffle_distrib_forests(data, opts.shuffle_distrib_forest_seed)
    s.update(objs=objs, data=data, num_data=count,
             ang_max=utils.compute_ang_max(cosmo, opts.rt_max, z_low, z_low_obj))
    s.apply(xcf, reset=False)
    clock.lap("read")

    healpix, rows = D.healpix_rows(xcf, xcf.data)
    clock.lap("pairs")
    cards = D.header(RPMIN=xcf.r_par_min, RPMAX=xcf.r_par_max, RTMAX=xcf.r_trans_max,
                     NP=xcf.num_bins_r_par, NT=xcf.num_bins_r_trans, ZMIN=xcf.z_min_pairs,
                     ZMAX=xcf.z_max_pairs, NSIDE=xcf.nside)
    cards += D.cosmo_cards(opts) + D.header(BLINDING=blinding, RMU_BIN=xcf.rmu_binning)
    D.correlation_products(opts.out, D.combine_rows(rows), healpix, cards, blinding,
                           rmu=xcf.rmu_binning)
    clock.total()
    return xcf.last_report


if __name__ == "__main__":
    D.main_entry(main)
