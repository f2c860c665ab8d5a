"""Host side of the metal distortion matrix (no GPU): the absorber tables handed to the
engine agree with the oracle's per-delta view of the same pixels on picca's test deltas,
and the engine library exports the metal entry points."""
import numpy as np

import cor_cases
from oracle import cf_oracle as O
from picca_amd import _native, constants, metal
from picca_amd.store import ForestStore


def test_absorber_table_matches_oracle_view():
    setup = cor_cases.Setup(cor_cases.DMAT_CASES["dmat"], dmat=True)
    cosmo = constants.Cosmo(Om=0.315, verbose=False)
    deltas = [d for hp in sorted(setup.data) for d in setup.data[hp]]
    store = ForestStore.from_deltas(deltas)
    lam = constants.ABSORBER_IGM["SiIII(1207)"]
    table = metal.absorber_table(store, lam, 1.0, cosmo)
    xtable = metal.absorber_table(store, lam, 1.7, cosmo, z_ref=2.25)
    lam_masked = constants.ABSORBER_IGM["SiII(1190)"]
    table_masked = metal.absorber_table(store, lam_masked, 1.0, cosmo)
    masked = 0
    for k, d in enumerate(deltas):
        sl = slice(store.offsets[k], store.offsets[k + 1])
        ref = O.absorber_pixels(d, lam, 1.0, cosmo)
        for mine, want in zip(table, ref):
            np.testing.assert_array_equal(mine[sl], want)
        xref = O.absorber_pixels(d, lam, 1.7, cosmo, z_ref=2.25)
        np.testing.assert_array_equal(xtable[4][sl], xref[4])
        mref = O.absorber_pixels(d, lam_masked, 1.0, cosmo)
        np.testing.assert_array_equal(table_masked[3][sl], mref[3])
        masked += int(np.sum((mref[3] == 0) & (d.weights != 0)))
    # the forest reaches 1200 A rest frame: as SiII(1190) its reddest pixels lie beyond
    # z_qso and are dropped
    assert masked > 0


def test_library_exports_metal_entry_points():
    names = _native.exported_symbols()
    assert "picca_metal_set_absorber" in names and "picca_metal_pass" in names
    lib = _native.load()
    for name in ("picca_metal_set_absorber", "picca_metal_pass"):
        assert hasattr(lib, name)


def test_smooth_cov_matches_oracle_loops():
    """The vectorised smoothing against the oracle's restatement of the reference's loops,
    plain and per r_par, on a random positive-definite covariance of a 6 x 5 grid."""
    from oracle import cov_oracle as C
    from picca_amd.utils import smooth_cov
    rng = np.random.default_rng(7)
    rp = np.repeat(np.arange(6) * 4.0 + 2.0, 5) + rng.normal(0, 0.3, 30)
    rt = np.tile(np.arange(5) * 4.0 + 2.0, 6) + rng.normal(0, 0.3, 30)
    m = rng.normal(size=(40, 30))
    cov = m.T @ m / 40
    for per in (False, True):
        mine = smooth_cov(None, None, rp, rt, 4.0, 4.0, covariance=cov, per_r_par=per)
        ref = C.smooth_cov(cov, rp, rt, 4.0, 4.0, per_r_par=per)
        np.testing.assert_array_equal(mine, ref)
