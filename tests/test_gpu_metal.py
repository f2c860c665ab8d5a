"""Parity of the HIP metal distortion-matrix pass with the CPU oracle and picca's golden
metal_dmat / metal_dmat_cross / metal_xdmat files.

The GPU merges consecutive pairs of one (data bin, model bin) and adds in another order
than the reference's bincount, and it multiplies by 1/(1+z_ref)^(a1+a2-2) where the
reference divides.  Bar: the pair counts (NPALL, NPUSED, pairs added to weights_dmat and
to dmat) identical; weights_dmat, every dmat entry and the effective coordinates within
1e-9 relative or 1e-12 of the largest entry; normalised columns within picca's own test
tolerance of the golden files.
"""
import numpy as np
import pytest

import cor_cases
from oracle import cf_oracle as O
from picca_amd import cf, constants, metal, synth, xcf
from picca_amd import _native
from picca_amd.engine import get_engine, params_from_globals
from test_gpu_xcf import set_xcf_globals
from test_gpu_xi import set_globals
from test_oracle_metal_golden import METAL_CASES, normalise

pytestmark = pytest.mark.gpu

RTOL = 1e-9
ATOL_REL = 1e-12
ALPHA_ABS = {"LYA": 2.9, "SiIII(1207)": 1.0}


def assert_close(mine, ref):
    names = ("weights_dmat", "dmat", "r_par_eff", "r_trans_eff", "z_eff", "weight_eff")
    for k, name in enumerate(names):
        a = np.asarray(mine[k]).ravel()
        b = np.asarray(ref[k]).ravel()
        assert a.shape == b.shape, name
        scale = float(np.abs(b).max()) if b.size else 0.0
        np.testing.assert_allclose(a, b, rtol=RTOL, atol=ATOL_REL * max(scale, 1e-300),
                                   err_msg=name)


def set_metal_globals(setup, cosmo):
    set_globals(setup)
    cf.alpha_abs = dict(ALPHA_ABS)
    cf.cosmo = cosmo
    cf.lambda_abs = constants.ABSORBER_IGM["LYA"]
    cf.lambda_abs2 = constants.ABSORBER_IGM["LYA"] if setup.data2 is not None else None


@pytest.mark.parametrize("case", sorted(METAL_CASES))
def test_fixture_metal_dmat_matches_oracle_and_golden(case):
    setup = cor_cases.Setup(METAL_CASES[case], dmat=True)
    cosmo = constants.Cosmo(Om=0.315, verbose=False)
    set_metal_globals(setup, cosmo)
    healpixs = sorted(setup.data)
    params = O.params_from_module(setup)
    neigh = O.brute_force_neighbours(setup.data, setup.data2, healpixs, setup.ang_max)
    golden = cor_cases.golden_columns(case)
    attri = cor_cases.FITS(cor_cases.golden_path(case))["ATTRI"]
    for k, (abs1, abs2) in enumerate([("LYA", "SiIII(1207)"),
                                      ("SiIII(1207)", "SiIII(1207)")]):
        np.random.seed(healpixs[0])
        ref, ref_stats = O.compute_metal_dmat(params, setup.data, healpixs, neigh,
                                              setup.reject, cosmo, abs1, abs2,
                                              constants.ABSORBER_IGM, ALPHA_ABS,
                                              constants.ABSORBER_IGM["LYA"])
        cf.fill_neighs(healpixs)
        np.random.seed(healpixs[0])
        mine = cf.compute_metal_dmat(healpixs, abs1, abs2)
        assert mine[6] == ref[6] == attri["NPALL"][:][k]
        assert mine[7] == ref[7] == attri["NPUSED"][:][k]
        assert_close(mine[:6], ref[:6])
        passes = cf.last_report["passes"]
        assert sum(p["data_pairs"] for p in passes) == ref_stats[0]
        assert sum(p["model_pairs"] for p in passes) == ref_stats[1]
        assert all(p["overflow_pairs"] == 0 for p in passes)
        cols, _, _ = normalise(mine)
        name = abs1 + "_" + abs2
        mine_cols = {c + "_" + name: v for c, v in cols.items()}
        ref_cols = {c: v for c, v in golden.items() if c.endswith("_" + name)}
        cor_cases.compare_to_golden(mine_cols, ref_cols)


def test_fixture_metal_xdmat_matches_oracle_and_golden():
    argv = cor_cases.XCF_COMMON + ["--rej", "0.99"]
    setup = cor_cases.XcfSetup(argv)
    setup.reject = 0.99
    setup.num_model_bins_r_par = setup.num_bins_r_par
    setup.num_model_bins_r_trans = setup.num_bins_r_trans
    cosmo = constants.Cosmo(Om=0.315, verbose=False)
    set_xcf_globals(setup)
    xcf.num_model_bins_r_par = setup.num_bins_r_par
    xcf.num_model_bins_r_trans = setup.num_bins_r_trans
    xcf.reject = 0.99
    xcf.alpha_abs = dict(ALPHA_ABS)
    xcf.cosmo = cosmo
    healpixs = sorted(setup.data)
    params = O.params_from_module(setup)
    neigh = O.xcf_neighbours(setup.data, setup.objs, healpixs, setup.ang_max,
                             setup.r_par_min, setup.r_par_max, setup.r_trans_max)
    np.random.seed(healpixs[0])
    ref, ref_stats = O.compute_metal_xdmat(params, setup.data, healpixs, neigh, setup.reject,
                                           cosmo, "SiIII(1207)", constants.ABSORBER_IGM,
                                           ALPHA_ABS)
    xcf.fill_neighs(healpixs)
    np.random.seed(healpixs[0])
    mine = xcf.compute_metal_dmat(healpixs, "SiIII(1207)")
    attri = cor_cases.FITS(cor_cases.golden_path("metal_xdmat"))["ATTRI"]
    assert mine[6] == ref[6] == attri["NPALL"][:][0]
    assert mine[7] == ref[7] == attri["NPUSED"][:][0]
    assert_close(mine[:6], ref[:6])
    rep = xcf.last_report["passes"][0]
    assert rep["data_pairs"] == ref_stats[0]
    assert rep["model_pairs"] == ref_stats[1]
    cols, _, _ = normalise(mine)
    mine_cols = {c + "_QSO_SiIII(1207)": v for c, v in cols.items()}
    cor_cases.compare_to_golden(mine_cols, cor_cases.golden_columns("metal_xdmat"))


class _Globals:
    num_bins_r_par = 50
    num_bins_r_trans = 50
    num_model_bins_r_par = 50
    num_model_bins_r_trans = 50
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
    remove_same_half_plate_close_pairs = True
    redshift_evolution_in_distortion_matrix = True


def synthetic_pairs(seed, count):
    survey = synth.make_survey("survey_small", seed=seed, scale=0.3)
    store = survey.store
    rng = np.random.default_rng(seed)
    nf = store.num_forests
    a = rng.integers(0, nf, count)
    b = rng.integers(0, nf, count)
    ang = rng.uniform(0.0, 0.04, count)
    flags = rng.integers(0, 2, count).astype(np.uint8)
    return store, a, b, ang, flags


@pytest.mark.parametrize("mode,npm,ntm", [(_native.PICCA_METAL_AUTO, 50, 50),
                                         (_native.PICCA_METAL_CROSS, 50, 50),
                                         (_native.PICCA_METAL_AUTO, 100, 60)])
def test_synthetic_pass_matches_oracle(mode, npm, ntm):
    """Random pairs of the synthetic survey, both absorber passes of LYA x SiIII(1207);
    the 100 x 60 model grid does not fit the LDS histogram (global-atomic form)."""
    g = _Globals()
    g.num_model_bins_r_par, g.num_model_bins_r_trans = npm, ntm
    g.x_correlation = mode == _native.PICCA_METAL_CROSS
    if g.x_correlation:
        g.r_par_min = -200.0
    store, a, b, ang, flags = synthetic_pairs(3, 400)
    cosmo = constants.Cosmo(Om=0.315, verbose=False)
    lya = metal.absorber_table(store, constants.ABSORBER_IGM["LYA"], 2.9, cosmo)
    si = metal.absorber_table(store, constants.ABSORBER_IGM["SiIII(1207)"], 1.0, cosmo)
    den = (1 + g.z_ref) ** (2.9 + 1.0 - 2)
    eng = get_engine(0)
    eng.set_forests(store, key=("metal_synth", mode))
    eng.metal_set_absorber(0, *lya)
    eng.metal_set_absorber(1, *si)
    params = params_from_globals(g)
    nb, nm = 50 * 50, npm * ntm
    mine = metal.outputs(nb, nm)
    reps = [eng.metal_pass(params, mode, 0, 1, den, a, b, ang, flags, mine),
            eng.metal_pass(params, mode, 1, 0, den, a, b, ang, flags, mine)]
    oparams = O.params_from_module(g)
    ref = metal.outputs(nb, nm)
    stats = np.zeros(3, dtype=np.int64)
    off = store.offsets
    for k in range(len(a)):
        sa = slice(off[a[k]], off[a[k] + 1])
        sb = slice(off[b[k]], off[b[k] + 1])
        for ta, tb in ((lya, si), (si, lya)):
            O.metal_pair(oparams, mode, den, (store.r_comov[sa], store.dist_m[sa]),
                         [x[sa] for x in ta], (store.r_comov[sb], store.dist_m[sb]),
                         [x[sb] for x in tb], ang[k], bool(flags[k]), ref, stats)
    assert sum(r["data_pairs"] for r in reps) == stats[0] > 0
    assert sum(r["model_pairs"] for r in reps) == stats[1] > 0
    assert_close(mine, ref)


def test_empty_and_invalid_inputs():
    g = _Globals()
    store, a, b, ang, flags = synthetic_pairs(5, 4)
    cosmo = constants.Cosmo(Om=0.315, verbose=False)
    eng = get_engine(0)
    eng.set_forests(store, key=("metal_empty",))
    eng.metal_set_absorber(2, *metal.absorber_table(store, 1215.67, 2.9, cosmo))
    params = params_from_globals(g)
    out = metal.outputs(2500, 2500)
    rep = eng.metal_pass(params, _native.PICCA_METAL_AUTO, 2, 2, 1.0, [], [], [], [], out)
    assert rep["data_pairs"] == 0 and not any(x.any() for x in out)
    with pytest.raises(ValueError):
        eng.metal_pass(params, 7, 2, 2, 1.0, a, b, ang, flags, out)
    with pytest.raises(ValueError):
        eng.metal_pass(params, 0, 2, 2, 1.0, a, b + store.num_forests, ang, flags, out)
    with pytest.raises(RuntimeError):
        eng.metal_pass(params, 0, 3, 2, 1.0, a, b, ang, flags, out)


def compare_metal_files(mine_path, golden_name):
    mine = cor_cases.FITS(mine_path)
    golden = cor_cases.FITS(cor_cases.golden_path(golden_name))
    for key in ("NPALL", "NPUSED"):
        assert np.array_equal(mine["ATTRI"][key][:], golden["ATTRI"][key][:]), key
    assert [n.strip() for n in mine["ATTRI"]["ABS_IGM"][:]] == \
        [n.strip() for n in golden["ATTRI"]["ABS_IGM"][:]]
    assert mine["MDMAT"].get_colnames() == golden["MDMAT"].get_colnames()
    result = {c: mine["MDMAT"][c][:] for c in mine["MDMAT"].get_colnames()}
    cor_cases.compare_to_golden(result, cor_cases.golden_columns(golden_name))


METAL_DRIVER_CASES = {
    "metal_dmat": METAL_CASES["metal_dmat"] + ["--abs-igm", "SiIII(1207)"],
    "metal_dmat_cross": METAL_CASES["metal_dmat_cross"] + ["--abs-igm", "SiIII(1207)",
                                                           "--abs-igm2", "SiIII(1207)"],
}


@pytest.mark.parametrize("case", sorted(METAL_DRIVER_CASES))
def test_metal_dmat_driver_reproduces_golden(case, tmp_path):
    from picca_amd.bin import picca_metal_dmat
    out = str(tmp_path / (case + ".fits.gz"))
    picca_metal_dmat.main(METAL_DRIVER_CASES[case] + ["--out", out])
    compare_metal_files(out, case)


def test_metal_xdmat_driver_reproduces_golden(tmp_path):
    from picca_amd.bin import picca_metal_xdmat
    out = str(tmp_path / "metal_xdmat.fits.gz")
    argv = ["--in-dir", cor_cases.DELTA_LYA, "--drq", cor_cases.CAT, "--abs-igm",
            "SiIII(1207)", "--rp-min", "-60.0", "--rp-max", "+60.0", "--rt-max", "+60.0",
            "--np", "30", "--nt", "15", "--rej", "0.99", "--nproc", "1", "--z-evol-obj", "1.",
            "--in-attributes", cor_cases.ATTRIBUTES, "--out", out]
    picca_metal_xdmat.main(argv)
    compare_metal_files(out, "metal_xdmat")
