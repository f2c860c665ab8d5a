"""Pins the metal distortion-matrix oracle against picca's golden metal_dmat files.

Replays test_metal_dmat (py/picca/tests/test_3_cor.py:351-381), test_metal_dmat_cross
(:526-560) and test_metal_xdmat (:700-729) with the oracle: brute-force neighbours, the
legacy numpy stream seeded with the first healpix before every matrix
(py/picca/bin/picca_metal_dmat.py:36-38, py/picca/bin/picca_metal_xdmat.py), the C pass
oracle_metal_pair and the merge/normalisation of py/picca/bin/picca_metal_dmat.py:470-500.
"""
import numpy as np
import pytest

import cor_cases
from oracle import cf_oracle as O
from picca_amd import constants

METAL_CASES = {
    "metal_dmat": ["--in-dir", cor_cases.DELTA_LYA, "--rp-min", "+0.0", "--rp-max", "+60.0",
                   "--rt-max", "+60.0", "--np", "15", "--nt", "15", "--rej", "0.99",
                   "--nproc", "1", "--in-attributes", cor_cases.ATTRIBUTES],
    "metal_dmat_cross": ["--in-dir", cor_cases.DELTA_LYA, "--in-dir2", cor_cases.DELTA_LYA,
                         "--rp-min", "-60.0", "--rp-max", "+60.0", "--rt-max", "+60.0",
                         "--np", "30", "--nt", "15", "--rej", "0.99", "--nproc", "1",
                         "--remove-same-half-plate-close-pairs", "--unfold-cf",
                         "--in-attributes", cor_cases.ATTRIBUTES],
}


def normalise(out):
    """py/picca/bin/picca_metal_dmat.py:470-500 (no redshift rescaling)."""
    weights_dmat, dmat, r_par, r_trans, z, weights, num_pairs, num_pairs_used = out
    r_par, r_trans, z, dmat = r_par.copy(), r_trans.copy(), z.copy(), dmat.copy()
    w = weights > 0
    r_par[w] /= weights[w]
    r_trans[w] /= weights[w]
    z[w] /= weights[w]
    w = weights_dmat > 0
    dmat[w, :] /= weights_dmat[w, None]
    return {"RP": r_par, "RT": r_trans, "Z": z, "DM": dmat, "WDM": weights_dmat}, \
        num_pairs, num_pairs_used


def metal_setup(case):
    setup = cor_cases.Setup(METAL_CASES[case], dmat=True)
    cosmo = constants.Cosmo(Om=0.315, verbose=False)
    alpha_abs = {"LYA": 2.9, "SiIII(1207)": 1.0}
    return setup, cosmo, alpha_abs


@pytest.mark.parametrize("case", ["metal_dmat", "metal_dmat_cross"])
def test_oracle_reproduces_golden_metal_dmat(case):
    setup, cosmo, alpha_abs = metal_setup(case)
    params = O.params_from_module(setup)
    healpixs = sorted(setup.data)
    neigh = O.brute_force_neighbours(setup.data, setup.data2, healpixs, setup.ang_max)
    golden = cor_cases.golden_columns(case)
    attri = cor_cases.FITS(cor_cases.golden_path(case))["ATTRI"]
    names = [n.strip() for n in attri["ABS_IGM"][:]]
    assert names == ["LYA_SiIII(1207)", "SiIII(1207)_SiIII(1207)"]
    for k, (abs1, abs2) in enumerate([("LYA", "SiIII(1207)"),
                                      ("SiIII(1207)", "SiIII(1207)")]):
        np.random.seed(healpixs[0])
        out, stats = O.compute_metal_dmat(params, setup.data, healpixs, neigh, setup.reject,
                                          cosmo, abs1, abs2, constants.ABSORBER_IGM,
                                          alpha_abs, constants.ABSORBER_IGM["LYA"])
        assert stats[2] == 0
        cols, npall, npused = normalise(out)
        assert npall == attri["NPALL"][:][k]
        assert npused == attri["NPUSED"][:][k]
        name = abs1 + "_" + abs2
        mine = {c + "_" + name: v for c, v in cols.items()}
        ref = {c: v for c, v in golden.items() if c.endswith("_" + name)}
        assert len(ref) == 5
        cor_cases.compare_to_golden(mine, ref)


def test_oracle_reproduces_golden_metal_xdmat():
    argv = cor_cases.XCF_COMMON + ["--rej", "0.99"]
    setup = cor_cases.XcfSetup(argv)
    setup.reject = 0.99
    setup.num_model_bins_r_par = setup.num_bins_r_par
    setup.num_model_bins_r_trans = setup.num_bins_r_trans
    cosmo = constants.Cosmo(Om=0.315, verbose=False)
    params = O.params_from_module(setup)
    healpixs = sorted(setup.data)
    neigh = O.xcf_neighbours(setup.data, setup.objs, healpixs, setup.ang_max,
                             setup.r_par_min, setup.r_par_max, setup.r_trans_max)
    np.random.seed(healpixs[0])
    out, stats = O.compute_metal_xdmat(params, setup.data, healpixs, neigh, setup.reject,
                                       cosmo, "SiIII(1207)", constants.ABSORBER_IGM,
                                       {"LYA": 2.9, "SiIII(1207)": 1.0})
    assert stats[2] == 0
    cols, npall, npused = normalise(out)
    attri = cor_cases.FITS(cor_cases.golden_path("metal_xdmat"))["ATTRI"]
    assert npall == attri["NPALL"][:][0]
    assert npused == attri["NPUSED"][:][0]
    mine = {c + "_QSO_SiIII(1207)": v for c, v in cols.items()}
    cor_cases.compare_to_golden(mine, cor_cases.golden_columns("metal_xdmat"))
