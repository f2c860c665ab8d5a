"""Parity of the FP64 matrix-core subsampling covariance with the oracle and picca's
golden exported_cf file.

The GPU sums the samples of Y^T Y in the matrix cores' order where numpy's BLAS uses its
own, so entries agree to 1e-12 of the largest entry; S and the means are sequential sums
as numpy's, so a bin with zero total weight keeps the same raw value.  The GPU matrix is
exactly symmetric.
"""
import numpy as np
import pytest

import cor_cases
from oracle import cov_oracle as C
from picca_amd.engine import get_engine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ns,nb", [(1, 1), (14, 225), (100, 65), (37, 700), (1196, 130),
                                   (5, 64), (300, 2500)])
def test_cov_matches_oracle(ns, nb):
    rng = np.random.default_rng(ns * 1000 + nb)
    xi = rng.normal(0.0, 1e-2, size=(ns, nb))
    we = rng.uniform(1e2, 1e4, size=(ns, nb))
    we[rng.random(size=we.shape) < 0.05] = 0.0
    if nb > 3:
        we[:, 3] = 0.0           # a bin without weight: S = 0, left unnormalised
    mine, ms, ms_syrk = get_engine(0).cov_subsample(xi, we)
    ref = C.compute_cov(xi, we)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(mine, ref, rtol=1e-10, atol=1e-12 * scale)
    assert np.array_equal(mine, mine.T)
    assert ms >= ms_syrk >= 0


def test_cov_empty_and_invalid():
    eng = get_engine(0)
    cov, _, _ = eng.cov_subsample(np.zeros((0, 4)), np.zeros((0, 4)))
    assert cov.shape == (4, 4) and not cov.any()
    with pytest.raises(ValueError):
        eng.cov_subsample(np.full((2, 3), np.nan), np.ones((2, 3)))
    with pytest.raises(ValueError):
        eng.cov_subsample(np.zeros((2, 3)), np.zeros((3, 2)))


@pytest.mark.parametrize("smooth", [True, False])
def test_export_driver_matches_golden_and_oracle(smooth, tmp_path):
    from picca_amd.bin import picca_export
    out = str(tmp_path / "exported_cf.fits.gz")
    argv = ["--data", cor_cases.golden_path("cf"), "--dmat", cor_cases.golden_path("dmat"),
            "--out", out]
    if not smooth:
        argv.append("--do-not-smooth-cov")
    picca_export.main(argv)
    mine = cor_cases.FITS(out)
    cols = {c: mine["COR"][c][:] for c in mine["COR"].get_colnames()}
    cols.update({c: mine["DMATTRI"][c][:] for c in mine["DMATTRI"].get_colnames()})
    f = cor_cases.FITS(cor_cases.golden_path("cf"))
    d = cor_cases.FITS(cor_cases.golden_path("dmat"))
    hd = [{c: f[k][c][:] for c in f[k].get_colnames()} for k in (1, 2)]
    dd = [{c: d[k][c][:] for c in d[k].get_colnames()} for k in (1, 2)]
    ref = C.export_columns((hd[0], hd[1], f[1].read_header()), (dd[0], dd[1]), smooth=smooth)
    for name in ("DA", "RP", "RT", "Z", "DM", "NB", "DMRP", "DMRT", "DMZ"):
        np.testing.assert_array_equal(cols[name], ref[name], err_msg=name)
    scale = np.abs(ref["CO"]).max()
    np.testing.assert_allclose(cols["CO"], ref["CO"], rtol=1e-9, atol=1e-12 * scale)
    if smooth:
        cor_cases.compare_to_golden(cols, cor_cases.golden_columns("exported_cf"))
