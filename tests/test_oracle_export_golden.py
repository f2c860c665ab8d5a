"""Pins the covariance-export oracle against picca's golden exported_cf file.

Replays test_export_cf (py/picca/tests/test_3_cor.py:443-456: picca_export on the golden
cf and dmat files, default smoothing) with oracle/cov_oracle.py, and checks compute_cov
against closed forms: with equal weights it is the population covariance of the rows
divided by their number; a row that repeats the mean adds nothing.
"""
import numpy as np

import cor_cases
from oracle import cov_oracle as C


def read_hdus(name):
    f = cor_cases.FITS(cor_cases.golden_path(name))
    return [{c: f[k][c][:] for c in f[k].get_colnames()} for k in (1, 2)], \
        f[1].read_header()


def test_oracle_reproduces_golden_export():
    (attri, cor), head = read_hdus("cf")
    (dm, dm_attri), _ = read_hdus("dmat")
    out = C.export_columns((attri, cor, head), (dm, dm_attri))
    golden = cor_cases.golden_columns("exported_cf")
    report = cor_cases.compare_to_golden(out, golden)
    assert {"DA", "CO", "DM", "NB", "DMRP"} <= set(report)
    # beyond picca's own tolerance: every covariance entry and xi within 1e-6 of the
    # largest (observed 4e-8 and 2e-8: the golden export was made from a cf file that
    # differs from the golden cf at that level)
    for name in ("CO", "DA"):
        scale = np.abs(golden[name]).max()
        assert np.abs(out[name] - golden[name]).max() < 1e-6 * scale, name


def test_compute_cov_closed_forms():
    rng = np.random.default_rng(4)
    xi = rng.normal(size=(9, 6))
    ones = np.ones_like(xi)
    cov = C.compute_cov(xi, ones)
    np.testing.assert_allclose(cov, np.cov(xi, rowvar=False, bias=True) / 9, rtol=1e-12)
    we = rng.uniform(0.5, 2.0, size=xi.shape)
    mean = (xi * we).sum(axis=0) / we.sum(axis=0)
    xi2 = np.vstack([xi, mean])
    we2 = np.vstack([we, rng.uniform(0.5, 2.0, size=(1, 6))])
    cov2 = C.compute_cov(xi2, we2)
    s1, s2 = we.sum(axis=0), we2.sum(axis=0)
    np.testing.assert_allclose(cov2 * np.outer(s2, s2), C.compute_cov(xi, we) * np.outer(s1, s1),
                               rtol=1e-10, atol=1e-14)
