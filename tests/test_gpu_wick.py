"""Parity of the HIP Wick terms (diagrams 1-3) with the CPU oracle and picca's golden wick
file.

The GPU sums each covariance row over the pixel pairs of one bin in another order than the
reference's double loop, so T1-T3 agree to 1e-9 relative (or 1e-12 of the largest entry);
weights_wick within 1e-12 relative; num_pairs_wick, NPALL and NPUSED exact.
"""
import numpy as np
import pytest

import cor_cases
from oracle import cf_oracle as O
from picca_amd import cf, synth
from picca_amd.engine import get_engine, params_from_globals
from test_gpu_xi import set_globals
from test_oracle_wick_golden import WICK_CASE, normalise, oracle_wick

pytestmark = pytest.mark.gpu


def assert_close(a, b, name, rtol=1e-9):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = float(np.abs(b).max()) if b.size else 0.0
    np.testing.assert_allclose(a, b, rtol=rtol, atol=1e-12 * max(scale, 1e-300), err_msg=name)


def test_fixture_wick_matches_oracle_and_golden():
    setup, ref = oracle_wick()
    set_globals(setup)
    cf.alpha2 = setup.alpha2
    cf.max_diagram = 3
    variance, xi = O.read_cf1d(cor_cases.golden_path("cf1d"))
    cf.xi_1d = {"D1": xi}
    cf.get_variance_1d = {"D1": variance}
    for deltas in setup.data.values():
        for d in deltas:
            d.fname = "D1"
    healpixs = sorted(setup.data)
    cf.fill_neighs(healpixs)
    np.random.seed(healpixs[0])
    mine = cf.compute_wick_terms(healpixs)
    assert mine[2] == ref[2] and mine[3] == ref[3]
    assert np.array_equal(mine[1], ref[1])
    assert_close(mine[0], ref[0], "weights_wick", rtol=1e-12)
    for k, name in ((4, "t1"), (5, "t2"), (6, "t3")):
        assert_close(mine[k], ref[k], name)
    assert not any(x.any() for x in mine[7:])
    assert cf.last_report["overflow_pairs"] == 0
    cols = normalise(*mine[:7])
    cor_cases.compare_to_golden(cols, cor_cases.golden_columns("wick"))


def test_wick_driver_reproduces_golden(tmp_path):
    from picca_amd.bin import picca_wick
    out = str(tmp_path / "wick.fits.gz")
    picca_wick.main(WICK_CASE + ["--cf1d", cor_cases.golden_path("cf1d"), "--out", out])
    mine = cor_cases.FITS(out)
    golden = cor_cases.FITS(cor_cases.golden_path("wick"))
    for key in ("NPALL", "NPUSED"):
        assert mine["COV"].read_header()[key] == golden["COV"].read_header()[key]
    assert mine["COV"].get_colnames() == golden["COV"].get_colnames()
    result = {c: mine["COV"][c][:] for c in mine["COV"].get_colnames()}
    cor_cases.compare_to_golden(result, cor_cases.golden_columns("wick"))


class _G:
    num_bins_r_par = 50
    num_bins_r_trans = 50
    num_model_bins_r_par = 0
    num_model_bins_r_trans = 0
    r_par_min = 0.0
    r_par_max = 200.0
    r_trans_max = 200.0
    z_min_pairs = None
    z_max_pairs = None
    zerr_cut_deg = None
    zerr_cut_kms = None
    z_ref = 2.25
    alpha = 2.9
    alpha2 = 2.9
    x_correlation = False
    rmu_binning = False
    ang_correlation = False
    remove_same_half_plate_close_pairs = False
    redshift_evolution_in_distortion_matrix = True


@pytest.mark.parametrize("xcorr", [False, True])
def test_synthetic_pairs_match_oracle(xcorr):
    """Random pairs of a synthetic survey with 50 x 50 bins to 200 Mpc/h and smooth random
    weighted 1-D correlations; the signed form with r_par in [-200, 200)."""
    g = _G()
    if xcorr:
        g.x_correlation = True
        g.r_par_min = -200.0
        g.num_bins_r_par = 100
    survey = synth.make_survey("survey_small", seed=11, scale=0.1)
    store = survey.store
    rng = np.random.default_rng(5)
    nf = store.num_forests
    sel = rng.choice(nf, min(nf, 40), replace=False)
    a = rng.choice(sel, 60)
    b = rng.choice(sel, 60)
    ang = rng.uniform(0.0, 0.03, 60)
    lengths = np.diff(store.offsets)
    wx_off = np.full(nf, -1, dtype=np.int64)
    mats = []
    used = 0
    for f in np.unique(np.concatenate([a, b])):
        n = lengths[f]
        x = np.arange(n)
        m = 0.01 * np.exp(-np.abs(x[:, None] - x[None, :]) / 7.0) * np.outer(
            store.weights[store.offsets[f]:store.offsets[f + 1]],
            store.weights[store.offsets[f]:store.offsets[f + 1]])
        wx_off[f] = used
        used += n * n
        mats.append(m.ravel())
    wx = np.concatenate(mats)
    zev1 = ((1 + store.z) / 3.25) ** 1.9
    zev2 = ((1 + store.z) / 3.25) ** 1.9
    eng = get_engine(0)
    eng.set_forests(store, key=("wick_synth", xcorr))
    mine = eng.wick_t123(params_from_globals(g), a, b, ang, zev1, zev2, wx_off, wx)
    params = O.params_from_module(g)
    nb = g.num_bins_r_par * g.num_bins_r_trans
    ww = np.zeros(nb)
    nw = np.zeros(nb, dtype=np.int64)
    t = [np.zeros(nb * nb) for _ in range(3)]
    a32, b32 = a.astype(np.int32), b.astype(np.int32)   # kept alive across the call
    O.lib().oracle_wick_t123(O.ctypes.byref(params), len(a), O._p(a32), O._p(b32),
                             O._p(ang), O._p(store.offsets),
                             O._p(store.r_comov), O._p(store.weights), O._p(zev1), O._p(zev2),
                             O._p(wx_off), O._p(wx), O._p(ww), O._p(nw), *[O._p(x) for x in t])
    assert nw.sum() > 0 and mine[5]["in_range_pairs"] == nw.sum()
    assert np.array_equal(mine[1], nw)
    assert_close(mine[0], ww, "weights_wick", rtol=1e-12)
    for k in range(3):
        assert_close(mine[2 + k], t[k].reshape(nb, nb), "t%d" % (k + 1))
