#!/usr/bin/env python3
"""picca_wick on the MI355X engine: the Wick covariance (diagrams T1-T3) of the
auto-correlation of forests, written as picca's COV file.

Command line and product as ``py/picca/bin/picca_wick.py``.  The 1-D correlation and
variance come from ``--cf1d`` files as nearest-neighbour look-ups; diagrams 1-3 run on the
GPU over the seeded round-robin healpix lists (``cf.compute_wick_terms``); diagrams 4-5
(``--max-diagram`` above 3) are refused.
"""
import numpy as np
from scipy.interpolate import interp1d

from picca_amd import cf, constants, io
from picca_amd import driver as D
from picca_amd.fitsio_lite import FITS


def one_d_lookups(path):
    """(variance(log_lambda), xi_1d(delta log_lambda)) of a picca_cf1d file: each a
    nearest-neighbour look-up over the grid points that hold pairs, extrapolated flat."""
    table = FITS(path)[1]
    head = table.read_header()
    grid = head["LLMIN"] + head["DLL"] * np.arange(len(table["v1d"][:]))
    has_var = table["nv1d"][:] > 0
    has_xi = table["nb1d"][:] > 0
    variance = interp1d(grid[has_var], table["v1d"][:][has_var], kind="nearest",
                        fill_value="extrapolate")
    xi = interp1d((grid - head["LLMIN"])[has_xi], table["c1d"][:][has_xi], kind="nearest",
                  fill_value="extrapolate")
    return variance, xi


def mean_correlation(path, opts):
    """Weight-averaged DA of a picca_cf file, whose binning must be this run's."""
    cor = FITS(path)
    head = cor[1].read_header()
    for key, value in (("NP", opts.np), ("NT", opts.nt), ("RPMIN", opts.rp_min),
                       ("RPMAX", opts.rp_max), ("RTMAX", opts.rt_max)):
        assert head[key] == value, "%s of %s differs from this run" % (key, path)
    weights = cor[2]["WE"][:]
    xi = (cor[2]["DA"][:] * weights).sum(axis=0)
    total = weights.sum(axis=0)
    xi[total > 0] /= total[total > 0]
    return xi


def label(data, name):
    for deltas in data.values():
        for delta in deltas:
            delta.fname = name


def main(cmdargs=None):
    opts = D.parse("picca_wick", cmdargs)
    clock = D.Clock("picca_wick")
    cf.device = opts.device
    D.process_count(opts)
    if opts.max_diagram > 3:
        raise NotImplementedError("Wick diagrams 4 and 5 (--max-diagram > 3) are not part of "
                                  "this engine")
    if (opts.fid_Or, opts.fid_Ok, opts.fid_wl) != (0.0, 0.0, -1.0):
        raise SystemExit("picca_wick: only the fiducial Omega_m may differ from the "
                         "default cosmology")
    blinding = io.read_blinding(opts.in_dir)
    cosmo = D.cosmology(opts, template=False)
    fields = D.Fields(opts, cosmo, opts.z_evol, opts.z_evol2)
    # the first field's deltas are "D1", or "D2" once a second field is read
    # (py/picca/bin/picca_wick.py:386,474 label them so)
    label(fields.first[0], "D1" if fields.second is None else "D2")

    s = D.Settings(D.binning_settings(opts))
    s.update(z_ref=opts.z_ref, alpha=opts.z_evol, alpha2=opts.z_evol, reject=opts.rej,
             lambda_abs=constants.ABSORBER_IGM[opts.lambda_abs],
             max_diagram=opts.max_diagram)
    s.update(fields.settings(opts.z_evol2))
    s.update(xi_1d={}, get_variance_1d={}, xi_wick={})
    for name, path in (("D1", opts.cf1d), ("D2", opts.cf1d2)):
        if path is not None:
            s["get_variance_1d"][name], s["xi_1d"][name] = one_d_lookups(path)
    for name, path in (("D1_D1", opts.cf), ("D2_D2", opts.cf2), ("D1_D2", opts.cf12),
                       ("D2_D1", opts.cf12)):
        if path is not None:
            s["xi_wick"][name] = mean_correlation(path, opts)
    s.apply(cf)
    clock.lap("read")

    parts = D.seeded_lists(cf, D.process_lists(cf.data, opts.nproc), cf.compute_wick_terms)
    clock.lap("wick")
    sums = D.add_parts(parts)
    weights_wick, num_pairs_wick, npall, npused = sums[:4]
    terms = sums[4:10]
    norm = weights_wick * weights_wick[:, None]
    filled = norm > 0
    for t in terms:
        t[filled] /= norm[filled]
        t *= 1.0 * npused / npall
    total = terms[0] + terms[1] + terms[2] + terms[3] + terms[4] + terms[5]

    cards = D.header(RPMIN=cf.r_par_min, RPMAX=cf.r_par_max, RTMAX=cf.r_trans_max,
                     NP=cf.num_bins_r_par, NT=cf.num_bins_r_trans, REJ=cf.reject,
                     NPALL=int(npall), NPUSED=int(npused))
    cards += D.cosmo_cards(opts) + D.header(BLINDING=blinding)
    cov = D.Product("COV", cards).add("CO", total).add("WALL", weights_wick)
    cov.add("NB", num_pairs_wick)
    for k, t in enumerate(terms):
        cov.add("T%d" % (k + 1), t, unit="", note="Wick term %d" % (k + 1))
    D.write_products(opts.out, [cov])
    clock.total()
    return [cf.last_report]


if __name__ == "__main__":
    D.main_entry(main)
