"""Parity of the HIP cross-correlation kernels (forests x quasars) with the CPU oracle and
with picca's golden xcf and xdmat files.

Bar for xi: num_pairs bit-exact per healpix row; weights / xi / r_par / r_trans / z within
1e-6 relative after normalisation; edge and overflow counts equal to the oracle's.  Bar for
the distortion matrix (the GPU regroups the fill loop, so FP64 sums are added in another
order): NPALL, NPUSED and the pair counts identical; every entry within 1e-9 relative or
1e-12 of the largest entry.  The drivers' outputs within picca's own test tolerance of the
golden files.
"""
import numpy as np
import pytest

import cor_cases
from oracle import cf_oracle as O
from picca_amd import xcf
from picca_amd.fitsio_lite import FITS

pytestmark = pytest.mark.gpu

RTOL = 1e-6


def set_xcf_globals(setup, **overrides):
    for name in ("num_bins_r_par", "num_bins_r_trans", "r_par_min", "r_par_max", "r_trans_max",
                 "z_min_pairs", "z_max_pairs", "zerr_cut_deg", "zerr_cut_kms", "z_ref",
                 "rmu_binning", "ang_correlation", "ang_max"):
        setattr(xcf, name, getattr(setup, name))
    for name, value in overrides.items():
        setattr(xcf, name, value)
    xcf.data = setup.data
    xcf.objs = setup.objs


@pytest.mark.parametrize("case", sorted(cor_cases.XCF_CASES))
def test_xcf_rows_match_oracle(case):
    setup = cor_cases.XcfSetup(cor_cases.XCF_CASES[case])
    set_xcf_globals(setup)
    healpixs = sorted(setup.data)
    rows = xcf.compute_xi_rows(healpixs)
    report = xcf.last_report
    params = O.params_from_module(setup)
    neigh = O.xcf_neighbours(setup.data, setup.objs, healpixs, setup.ang_max, setup.r_par_min,
                             setup.r_par_max, setup.r_trans_max)
    stats = np.zeros(3, dtype=np.int64)
    for hp, mine in zip(healpixs, rows):
        ref, st = O.compute_xcf(params, setup.data, [hp], neigh)
        stats += st
        assert np.array_equal(mine[5], ref[5])
        for k in range(5):
            np.testing.assert_allclose(mine[k], ref[k], rtol=RTOL, atol=1e-12)
    assert report["accepted_pairs"] == stats[0] > 0
    assert report["edge_pairs"] == stats[1]
    assert report["overflow_pairs"] == stats[2]


@pytest.mark.parametrize("mode", ["rmu", "narrow", "windows"])
def test_xcf_modes_match_oracle(mode):
    setup = cor_cases.XcfSetup(cor_cases.XCF_CASES["xcf"])
    if mode == "rmu":
        setup.rmu_binning = True
        setup.r_par_min, setup.r_par_max = -1.0, 1.0
    elif mode == "narrow":   # asymmetric r_par range with the r_comov neighbour cut active
        setup.r_par_min, setup.r_par_max = -20.0, 45.0
        setup.num_bins_r_par = 13
    else:   # 100 x 50 bins: the LDS histogram is split into windows of r_trans rows
        setup.num_bins_r_par, setup.num_bins_r_trans = 100, 50
    set_xcf_globals(setup)
    healpixs = sorted(setup.data)
    xcf.fill_neighs(healpixs)
    mine = xcf.compute_xi(healpixs)
    params = O.params_from_module(setup)
    neigh = O.xcf_neighbours(setup.data, setup.objs, healpixs, setup.ang_max, setup.r_par_min,
                             setup.r_par_max, setup.r_trans_max,
                             rmu_binning=setup.rmu_binning)
    ref, st = O.compute_xcf(params, setup.data, healpixs, neigh)
    assert st[0] > 0
    assert np.array_equal(mine[5], ref[5])
    for k in range(5):
        np.testing.assert_allclose(mine[k], ref[k], rtol=RTOL, atol=1e-12)
    report = xcf.last_report
    assert report["accepted_pairs"] == st[0]
    assert report["edge_pairs"] == st[1]
    assert report["overflow_pairs"] == st[2]


@pytest.mark.parametrize("case", sorted(cor_cases.XCF_CASES))
def test_xcf_driver_reproduces_golden(case, tmp_path):
    from picca_amd.bin import picca_xcf
    out = str(tmp_path / (case + ".fits.gz"))
    picca_xcf.main(cor_cases.XCF_CASES[case] + ["--out", out])
    mine = FITS(out)
    golden = FITS(cor_cases.golden_path(case))
    result = {}
    for hdu in mine.hdus[1:]:
        assert hdu.get_colnames() == golden[hdu.get_extname()].get_colnames()
        for col in hdu.get_colnames():
            result[col] = hdu[col][:]
    cor_cases.compare_to_golden(result, cor_cases.golden_columns(case))


# ---------------------------------------------------------------------------------------
# distortion matrix of the cross-correlation (xcf.compute_dmat, py/picca/xcf.py:324-658)
# ---------------------------------------------------------------------------------------

DM_RTOL = 1e-9
DM_ATOL_REL = 1e-12


def assert_dmat_close(mine, ref):
    names = ("weights_dmat", "dmat", "r_par_eff", "r_trans_eff", "z_eff", "weight_eff")
    for k, name in enumerate(names):
        a = np.asarray(mine[k])
        b = np.asarray(ref[k])
        assert a.shape == b.shape, name
        scale = float(np.abs(b).max()) if b.size else 0.0
        np.testing.assert_allclose(a, b, rtol=DM_RTOL, atol=DM_ATOL_REL * max(scale, 1e-300),
                                   err_msg=name)
    assert mine[6] == ref[6]   # num_pairs
    assert mine[7] == ref[7]   # num_pairs_used


def run_xdmat(setup):
    """The GPU and oracle distortion matrices of one setup, each from a fresh draw."""
    setup.redshift_evolution_in_distortion_matrix = setup.redshift_evolution
    set_xcf_globals(setup)
    for name in ("num_model_bins_r_par", "num_model_bins_r_trans", "reject"):
        setattr(xcf, name, getattr(setup, name))
    xcf.alpha = setup.alpha
    xcf.alpha_obj = setup.alpha2
    xcf.redshift_evolution_in_distortion_matrix = setup.redshift_evolution
    healpixs = sorted(setup.data)
    xcf.fill_neighs(healpixs)
    np.random.seed(healpixs[0])
    mine = xcf.compute_dmat(healpixs)
    report = xcf.last_report
    params = O.params_from_module(setup)
    params.alpha2 = setup.alpha2
    neigh = O.xcf_neighbours(setup.data, setup.objs, healpixs, setup.ang_max,
                             setup.r_par_min, setup.r_par_max, setup.r_trans_max,
                             rmu_binning=setup.rmu_binning)
    np.random.seed(healpixs[0])
    ref, stats = O.compute_xdmat(params, setup.data, healpixs, neigh, setup.reject)
    return mine, report, ref, stats


@pytest.mark.parametrize("case", sorted(cor_cases.XDMAT_CASES))
def test_xdmat_matches_oracle_and_golden(case):
    from test_oracle_dmat_golden import dmat_columns
    setup = cor_cases.XdmatSetup(cor_cases.XDMAT_CASES[case])
    mine, report, ref, stats = run_xdmat(setup)
    assert_dmat_close(mine, ref)
    assert report["in_range_pairs"] == stats[0] > 0
    assert report["selected_pairs"] == stats[1]
    assert report["overflow_pairs"] == stats[3]
    cols, npall, npused = dmat_columns(mine, setup.redshift_evolution, alpha2=setup.alpha2)
    header = FITS(cor_cases.golden_path(case))["DMAT"].read_header()
    assert npall == header["NPALL"] and npused == header["NPUSED"]
    cor_cases.compare_to_golden(cols, cor_cases.golden_columns(case))


XDMAT_MODES = {
    # more kept pairs, redshift evolution with the quasars' own exponent
    "evolution": {"reject": 0.8, "redshift_evolution": True, "alpha2": 1.44},
    # model binning finer than the data binning
    "coef2": {"reject": 0.8, "num_model_bins_r_par": 60, "num_model_bins_r_trans": 30},
    "rmu": {"reject": 0.8, "rmu_binning": True, "r_par_min": -1.0, "r_par_max": 1.0},
    "zcuts": {"reject": 0.8, "z_min_pairs": 2.2, "z_max_pairs": 2.5},
    "order0": {"reject": 0.8},
}


@pytest.mark.parametrize("mode", sorted(XDMAT_MODES))
def test_xdmat_modes_match_oracle(mode):
    setup = cor_cases.XdmatSetup(cor_cases.XDMAT_CASES["xdmat"])
    for name, value in XDMAT_MODES[mode].items():
        setattr(setup, name, value)
    if mode == "order0":
        for hp in setup.data:
            for d in setup.data[hp]:
                d.order = 0
    mine, report, ref, stats = run_xdmat(setup)
    assert stats[1] > 0
    assert_dmat_close(mine, ref)
    assert report["in_range_pairs"] == stats[0]
    assert report["selected_pairs"] == stats[1]


def test_xdmat_forest_kernel_matches_oracle():
    """compute_dmat_forest_pairs_fast of one delta against a few quasars, in place."""
    setup = cor_cases.XdmatSetup(cor_cases.XDMAT_CASES["xdmat"])
    set_xcf_globals(setup)
    xcf.num_model_bins_r_par, xcf.num_model_bins_r_trans = 30, 15
    xcf.alpha, xcf.alpha_obj = 2.9, 1.44
    xcf.redshift_evolution_in_distortion_matrix = True
    healpixs = sorted(setup.data)
    neigh = O.xcf_neighbours(setup.data, setup.objs, healpixs, setup.ang_max,
                             setup.r_par_min, setup.r_par_max, setup.r_trans_max)
    d1 = max((d for hp in healpixs for d in setup.data[hp]), key=lambda d: len(neigh[id(d)][0]))
    objs_n, ang = neigh[id(d1)]
    assert len(objs_n) > 2
    q = [np.array([getattr(o, k) for o in objs_n]) for k in ("r_comov", "dist_m", "z_qso",
                                                            "weights")]
    nb, nm = 30 * 15, 30 * 15
    mine = [np.zeros(nb), np.zeros(nb * nm)] + [np.zeros(nm) for _ in range(4)]
    xcf.compute_dmat_forest_pairs_fast(d1.log_lambda, d1.r_comov, d1.dist_m, d1.z, d1.weights,
                                       *q, ang, *mine, d1.order)
    params = O.params_from_module(setup)
    params.num_model_bins_r_par, params.num_model_bins_r_trans = 30, 15
    params.alpha, params.alpha2, params.redshift_evolution = 2.9, 1.44, 1
    ref = [np.zeros(nb), np.zeros(nb * nm)] + [np.zeros(nm) for _ in range(4)]
    stats = np.zeros(4, dtype=np.int64)
    O.xdmat_forest(params, d1.log_lambda, d1.r_comov, d1.dist_m, d1.z, d1.weights, *q, ang, ref,
                   d1.order, stats)
    assert stats[1] > 0
    for k in range(6):
        scale = float(np.abs(ref[k]).max())
        np.testing.assert_allclose(mine[k], ref[k], rtol=DM_RTOL,
                                   atol=DM_ATOL_REL * max(scale, 1e-300))


@pytest.mark.parametrize("case", sorted(cor_cases.XDMAT_CASES))
def test_xdmat_driver_reproduces_golden(case, tmp_path):
    from picca_amd.bin import picca_xdmat
    out = str(tmp_path / (case + ".fits.gz"))
    picca_xdmat.main(cor_cases.XDMAT_CASES[case] + ["--out", out])
    mine = FITS(out)
    golden = FITS(cor_cases.golden_path(case))
    for key in ("NPALL", "NPUSED"):
        assert mine["DMAT"].read_header()[key] == golden["DMAT"].read_header()[key]
    result = {}
    for hdu in mine.hdus[1:]:
        assert hdu.get_colnames() == golden[hdu.get_extname()].get_colnames()
        for col in hdu.get_colnames():
            result[col] = hdu[col][:]
    cor_cases.compare_to_golden(result, cor_cases.golden_columns(case))


def test_xcf_many_rows_match_oracle():
    """A synthetic survey crossed with its own quasars at 100 x 50 bins: many healpix rows
    per workgroup, so each workgroup moves its LDS histogram to HBM many times; every row
    must equal the oracle's (a wave adding to a bin while the histogram is being moved
    would lose pairs)."""
    import types
    from picca_amd import synth
    from picca_amd.constants import Cosmo
    from picca_amd.engine import get_engine, params_from_globals
    from picca_amd.neighbours import find_neighbours

    s = synth.make_survey("survey_small", seed=4, scale=1.0)
    store = s.store
    cosmo = Cosmo(0.315, verbose=False)

    class G:
        num_bins_r_par, num_bins_r_trans = 100, 50
        num_model_bins_r_par = num_model_bins_r_trans = None
        r_par_min, r_par_max, r_trans_max = -100.0, 100.0, 100.0
        z_min_pairs = z_max_pairs = zerr_cut_deg = zerr_cut_kms = None
        z_ref, alpha, alpha2 = 2.25, 2.9, 1.44
        x_correlation, rmu_binning, ang_correlation = True, False, False
        remove_same_half_plate_close_pairs = False

    qso = types.SimpleNamespace(x_cart=store.x_cart, y_cart=store.y_cart, z_cart=store.z_cart,
                                ra=store.ra, dec=store.dec, cos_dec=store.cos_dec,
                                thingid=store.thingid, z_qso=store.z_qso)
    qso.r_comov = np.asarray(cosmo.get_r_comov(store.z_qso), dtype=np.float64)
    qso.dist_m = np.asarray(cosmo.get_dist_m(store.z_qso), dtype=np.float64)
    qso.weights = ((1 + store.z_qso) / 3.25) ** 0.44
    ang_max = 2 * np.arcsin(G.r_trans_max / (2 * float(cosmo.get_dist_m(s.z_min))))
    # rows: healpix pixels at nside 256 (a few forests each), so that the workgroups
    # change rows often
    from picca_amd.healpix import ang2pix
    hp = ang2pix(256, np.pi / 2 - store.dec, store.ra)
    order = np.argsort(hp, kind="stable")
    uniq, counts = np.unique(hp[order], return_counts=True)
    row_of_src = np.repeat(np.arange(len(uniq), dtype=np.int32), counts)
    nl = find_neighbours(store, order, order, ang_max, auto=False, target_store=qso)
    src = nl.source_index()
    forest, quasar, ang, row = order[src], order[nl.target], nl.angle, row_of_src[src]
    assert len(uniq) > 100
    eng = get_engine(0)
    eng.set_forests(store)
    eng.xcf_set_pairs(forest, qso.r_comov[quasar], qso.dist_m[quasar], qso.z_qso[quasar],
                      qso.weights[quasar], ang, row, len(uniq))
    res = eng.xcf(params_from_globals(G))
    params = O.params_from_module(G)
    nb = G.num_bins_r_par * G.num_bins_r_trans
    total = 0
    for r in range(len(uniq)):
        out = [np.zeros(nb) for _ in range(5)] + [np.zeros(nb, dtype=np.int64)]
        stats = np.zeros(3, dtype=np.int64)
        for f in np.unique(forest[row == r]):
            idx = np.nonzero(forest == f)[0]
            q = quasar[idx]
            o0, o1 = store.offsets[f], store.offsets[f + 1]
            O.xcf_pair(params, store.z[o0:o1], store.r_comov[o0:o1], store.dist_m[o0:o1],
                       store.weights[o0:o1], store.delta[o0:o1], qso.z_qso[q], qso.r_comov[q],
                       qso.dist_m[q], qso.weights[q], ang[idx], out, stats)
        assert np.array_equal(res.num_pairs[r], out[5]), r
        np.testing.assert_allclose(res.weights[r], out[0], rtol=RTOL, atol=1e-12)
        np.testing.assert_allclose(res.xi[r], out[1], rtol=RTOL, atol=1e-9)
        total += int(out[5].sum())
    assert total > 0 and res.report["accepted_pairs"] == total
