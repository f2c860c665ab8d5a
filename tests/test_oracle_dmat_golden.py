"""Pins the distortion-matrix oracles against picca's golden dmat and xdmat files.

Replays test_dmat / test_dmat_cross / test_dmat_zcuts / test_dmat_cross_zcuts
(py/picca/tests/test_3_cor.py:318-349, :490-524, :858-893, :931-968) with the oracle:
brute-force neighbours, the legacy numpy stream seeded with the first healpix
(py/picca/bin/picca_dmat.py:36), the C restatement of compute_dmat_forest_pairs_fast and
the merge/normalisation of py/picca/bin/picca_dmat.py:493-519.  The same for the
cross-correlation: test_xdmat / test_xdmat_zcuts (py/picca/tests/test_3_cor.py:666-698,
:1004-1039) through the C restatement of xcf's compute_dmat_forest_pairs_fast.
"""
import numpy as np
import pytest

import cor_cases
from oracle import cf_oracle as O


def dmat_columns(out, redshift_evolution, z_ref=2.25, alpha=2.9, alpha2=2.9):
    """Normalisation of py/picca/bin/picca_dmat.py:503-519."""
    weights_dmat, dmat, r_par, r_trans, zeff, weights, num_pairs, num_pairs_used = out
    r_par, r_trans, zeff, dmat = r_par.copy(), r_trans.copy(), zeff.copy(), dmat.copy()
    w = weights > 0.0
    r_par[w] /= weights[w]
    r_trans[w] /= weights[w]
    zeff[w] /= weights[w]
    mean_zeff = np.mean(zeff[w])
    w = weights_dmat > 0
    dmat[w] /= weights_dmat[w, None]
    if redshift_evolution:
        zeff[:] = mean_zeff
        dmat *= ((1 + z_ref) / (1 + mean_zeff)) ** ((alpha - 1) + (alpha2 - 1))
    return {"WDM": weights_dmat, "DM": dmat, "RP": r_par, "RT": r_trans, "Z": zeff}, \
        num_pairs, num_pairs_used


def oracle_dmat(setup):
    params = O.params_from_module(setup)
    healpixs = sorted(setup.data)
    neigh = O.brute_force_neighbours(setup.data, setup.data2, healpixs, setup.ang_max)
    np.random.seed(healpixs[0])
    out, stats = O.compute_dmat(params, setup.data, healpixs, neigh, setup.reject)
    return out, stats


@pytest.mark.parametrize("case", ["dmat", "dmat_cross", "dmat_zcuts", "dmat_cross_zcuts"])
def test_oracle_reproduces_golden_dmat(case):
    setup = cor_cases.Setup(cor_cases.DMAT_CASES[case], dmat=True)
    out, stats = oracle_dmat(setup)
    cols, npall, npused = dmat_columns(out, setup.redshift_evolution_in_distortion_matrix)
    golden = cor_cases.FITS(cor_cases.golden_path(case))
    header = golden["DMAT"].read_header()
    assert npall == header["NPALL"]
    assert npused == header["NPUSED"]
    cor_cases.compare_to_golden(cols, cor_cases.golden_columns(case))


def oracle_xdmat(setup):
    params = O.params_from_module(setup)
    params.alpha2 = setup.alpha2
    healpixs = sorted(setup.data)
    neigh = O.xcf_neighbours(setup.data, setup.objs, healpixs, setup.ang_max,
                             setup.r_par_min, setup.r_par_max, setup.r_trans_max)
    # picca_xdmat fills the neighbours, then seeds the legacy stream with the first
    # healpix of the process list (py/picca/bin/picca_xdmat.py:35-37)
    np.random.seed(healpixs[0])
    return O.compute_xdmat(params, setup.data, healpixs, neigh, setup.reject)


@pytest.mark.parametrize("case", ["xdmat", "xdmat_zcuts"])
def test_oracle_reproduces_golden_xdmat(case):
    setup = cor_cases.XdmatSetup(cor_cases.XDMAT_CASES[case])
    out, stats = oracle_xdmat(setup)
    assert stats[0] > 0 and stats[1] > 0
    cols, npall, npused = dmat_columns(out, setup.redshift_evolution, alpha2=setup.alpha2)
    golden = cor_cases.FITS(cor_cases.golden_path(case))
    header = golden["DMAT"].read_header()
    assert npall == header["NPALL"]
    assert npused == header["NPUSED"]
    cor_cases.compare_to_golden(cols, cor_cases.golden_columns(case))
