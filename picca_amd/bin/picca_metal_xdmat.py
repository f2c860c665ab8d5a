#!/usr/bin/env python3
"""picca_metal_xdmat on the MI355X engine: the metal distortion matrices of the delta x
object cross-correlation, one per metal, written as picca's ATTRI + MDMAT file.

Command line and products as ``py/picca/bin/picca_metal_xdmat.py``; every metal runs over the
seeded round-robin healpix lists (``xcf.compute_metal_dmat`` per list).
"""
from picca_amd import constants, io, utils, xcf
from picca_amd import driver as D
from picca_amd.bin.picca_metal_dmat import matrix_of


def main(cmdargs=None):
    opts = D.parse("picca_metal_xdmat", cmdargs)
    clock = D.Clock("picca_metal_xdmat")
    xcf.device = opts.device
    D.process_count(opts)
    metals = list(opts.abs_igm or [])
    coef = opts.coef_binning_model
    cosmo = D.cosmology(opts, template=False)
    s = D.Settings(D.binning_settings(opts, model_coef=coef, data_coef=coef))
    s.update(z_min_pairs=None, z_max_pairs=None, zerr_cut_deg=None, zerr_cut_kms=None,
             rmu_binning=False, ang_correlation=False, z_ref=opts.z_ref,
             lambda_abs=constants.ABSORBER_IGM[opts.lambda_abs], reject=opts.rej,
             alpha=opts.z_evol_del, alpha_obj=opts.z_evol_obj, cosmo=cosmo,
             alpha_abs=dict({opts.lambda_abs: opts.z_evol_del},
                            **{m: opts.metal_alpha for m in metals}))

    blinding = io.read_blinding(opts.in_dir)
    data, count, z_low, z_high = D.load_deltas(opts, opts.in_dir, opts.lambda_abs,
                                               opts.z_evol_del, cosmo)
    D.object_redshift_limits(opts, cosmo, z_low, z_high)
    objs, z_low_obj = D.load_objects(opts, cosmo)
    s.update(data=data, num_data=count, objs=objs,
             ang_max=utils.compute_ang_max(cosmo, opts.rt_max, z_low, z_low_obj))
    s.apply(xcf, reset=False)
    clock.lap("read")

    lists = D.process_lists(xcf.data, opts.nproc)
    matrices, reports = [], []
    for metal in metals:
        parts = D.seeded_lists(xcf, lists, lambda hp: xcf.compute_metal_dmat(hp, abs_igm=metal))
        matrices.append(matrix_of(parts, metal))
        reports.append(xcf.last_report)
        clock.lap(metal)

    cards = D.header(RPMIN=xcf.r_par_min, RPMAX=xcf.r_par_max, RTMAX=xcf.r_trans_max,
                     NP=xcf.num_bins_r_par, NT=xcf.num_bins_r_trans, COEFMOD=coef)
    cards += D.cosmo_cards(opts) + D.header(BLINDING=blinding)
    D.metal_products(opts.out, cards, matrices, blinding, prefix=opts.obj_name + "_")
    clock.total()
    return reports


if __name__ == "__main__":
    D.main_entry(main)
