#!/usr/bin/env python3
"""picca_metal_dmat on the MI355X engine: the metal distortion matrices of the correlation of
delta fields, one per absorber pair, written as picca's ATTRI + MDMAT file.

Command line and products as ``py/picca/bin/picca_metal_dmat.py``.  Every absorber pair is
computed over the seeded round-robin healpix lists (``cf.compute_metal_dmat`` per list).
"""
from picca_amd import cf, constants, io
from picca_amd import driver as D

# the silicon lines the fast mode keeps next to LYA (py/picca/bin/picca_metal_dmat.py:18)
SILICON_LINES = ("SiIII(1207)", "SiII(1190)", "SiII(1193)", "SiII(1260)")


def absorber_pairs(first, second, fast):
    """Absorber pairs of the matrices: every (first-field absorber, second-field absorber)
    except the main pair, the second index starting at the first one's when both fields
    share their main absorber (each unordered pair once); `fast` keeps LYA with a silicon
    line and CIV(eff) with itself."""
    same_main = first[0] == second[0]
    out = []
    for k, a1 in enumerate(first):
        for k2 in range(k if same_main else 0, len(second)):
            a2 = second[k2]
            if k == 0 and k2 == (k if same_main else 0):
                continue
            if fast and not ((a1 == "LYA" and a2 in SILICON_LINES)
                             or (a1 == "CIV(eff)" and a2 == "CIV(eff)")):
                continue
            out.append((a1, a2))
    return out


def matrix_of(parts, name):
    weights_dmat, dmat, r_par, r_trans, z, weights, npall, npused = D.add_parts(parts)
    D.normalise_matrix(weights_dmat, dmat, r_par, r_trans, z, weights)
    return dict(name=name, weights_dmat=weights_dmat, dmat=dmat, r_par=r_par,
                r_trans=r_trans, z=z, npall=npall, npused=npused)


def main(cmdargs=None):
    opts = D.parse("picca_metal_dmat", cmdargs)
    clock = D.Clock("picca_metal_dmat")
    cf.device = opts.device
    D.process_count(opts)
    cosmo = D.cosmology(opts, template=False)
    blinding = io.read_blinding(opts.in_dir)
    fields = D.Fields(opts, cosmo, opts.z_evol, opts.z_evol2)
    evolution = {opts.lambda_abs: opts.z_evol}
    evolution.update({m: opts.metal_alpha for m in opts.abs_igm})
    if fields.second is not None:
        # keyed by the option as given (None when the second field has no absorber of its
        # own), as py/picca/bin/picca_metal_dmat.py:398 does
        evolution[opts.lambda_abs2] = opts.z_evol2
        evolution.update({m: opts.metal_alpha for m in opts.abs_igm2})
    coef = opts.coef_binning_model
    s = D.Settings(D.binning_settings(opts, model_coef=coef, data_coef=coef))
    s.update(z_ref=opts.z_ref, alpha=opts.z_evol, reject=opts.rej, cosmo=cosmo,
             lambda_abs=constants.ABSORBER_IGM[opts.lambda_abs], alpha_abs=evolution,
             remove_same_half_plate_close_pairs=opts.remove_same_half_plate_close_pairs)
    s.update(fields.settings(opts.z_evol2))
    s.apply(cf)
    clock.lap("read")

    first = [opts.lambda_abs] + list(opts.abs_igm)
    if opts.lambda_abs2 is None:
        second = [opts.lambda_abs] + list(opts.abs_igm)
    else:
        second = [opts.lambda_abs2] + list(opts.abs_igm2)
    lists = D.process_lists(cf.data, opts.nproc)
    matrices, reports = [], []
    for a1, a2 in absorber_pairs(first, second, opts.fast_metals):
        cf.counter = None
        parts = D.seeded_lists(cf, lists, lambda hp: cf.compute_metal_dmat(
            hp, abs_igm1=a1, abs_igm2=a2))
        matrices.append(matrix_of(parts, a1 + "_" + a2))
        reports.append(cf.last_report)
        clock.lap("%s x %s" % (a1, a2))

    cards = D.header(RPMIN=cf.r_par_min, RPMAX=cf.r_par_max, RTMAX=cf.r_trans_max,
                     NP=cf.num_bins_r_par, NT=cf.num_bins_r_trans, COEFMOD=coef,
                     REJ=cf.reject, ALPHAMET=opts.metal_alpha)
    cards += D.cosmo_cards(opts) + D.header(BLINDING=blinding)
    D.metal_products(opts.out, cards, matrices, blinding)
    clock.total()
    return reports


if __name__ == "__main__":
    D.main_entry(main)
