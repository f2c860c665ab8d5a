#!/usr/bin/env python3
"""picca_cf_angl on the MI355X engine: the correlation of delta fields in wavelength ratio
(r_par) and angle (r_trans), written as picca's ATTRI + COR file.

Command line and products as ``py/picca/bin/picca_cf_angl.py``.  No cosmology is involved:
the deltas are read without distances and the pair kernel runs its angular form.
"""
from picca_amd import cf, constants
from picca_amd import driver as D


def main(cmdargs=None):
    opts = D.parse("picca_cf_angl", cmdargs)
    clock = D.Clock("picca_cf_angl")
    cf.device = opts.device
    fields = D.Fields(opts, None, opts.z_evol, opts.z_evol2, project=not opts.no_project,
                      nproc=False, rebin=False, angular=True)
    s = D.Settings(r_par_min=opts.wr_min, r_par_max=opts.wr_max, r_trans_max=opts.ang_max,
                   num_bins_r_par=opts.np, num_bins_r_trans=opts.nt, nside=opts.nside,
                   z_ref=opts.z_ref, alpha=opts.z_evol, ang_correlation=True,
                   lambda_abs=constants.ABSORBER_IGM[opts.lambda_abs],
                   remove_same_half_plate_close_pairs=opts.remove_same_half_plate_close_pairs)
    s.update(fields.settings(opts.z_evol2))
    s["ang_max"] = opts.ang_max
    s.apply(cf)
    clock.lap("read")

    healpix, rows = D.healpix_rows(cf, cf.data)
    clock.lap("pairs")
    cards = D.header(RPMIN=cf.r_par_min, RPMAX=cf.r_par_max, RTMAX=cf.r_trans_max,
                     NP=cf.num_bins_r_par, NT=cf.num_bins_r_trans, NSIDE=cf.nside)
    table = D.combine_rows(rows)
    attri = D.Product("ATTRI", cards)
    attri.add("RP", table["RP"], unit="", note="mean wavelength ratio")
    attri.add("RT", table["RT"], unit="rad", note="mean angle")
    attri.add("Z", table["Z"]).add("NB", table["NB"])
    cor = D.Product("COR", D.header(HLPXSCHM="RING"))
    cor.add("HEALPID", healpix).add("WE", table["WE"]).add("DA", table["DA"])
    D.write_products(opts.out, [attri, cor])
    clock.total()
    return cf.last_report


if __name__ == "__main__":
    D.main_entry(main)
