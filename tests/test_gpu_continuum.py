"""Parity of the batched GPU continuum fit with the CPU oracle.

Bar: for every forest the GPU minimum chi2 is no higher than the oracle's and within
1e-9 relative of it (both sit at the minimum of the same FP64 objective), (a, b) within
1e-5 relative of the oracle's (the valley of chi2 fixes them only to about sqrt(tol)),
ndata and the validity verdict identical;
use_constant_weight fits equal the closed-form least squares to 1e-9.
"""
import numpy as np
import pytest

from oracle import cont_oracle as C
from picca_amd import continuum, synth_spectra as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def inputs():
    spectra = S.make_spectra(48, seed=21)
    return S.forest_inputs(spectra)


def specs_of(inputs):
    return [S.forest_spec(inputs, k) for k in range(len(inputs["offsets"]) - 1)]


@pytest.mark.parametrize("order", [0, 1])
def test_fit_matches_oracle(inputs, order):
    res = continuum.fit_arrays(inputs, order=order)
    assert np.all(res["status"] == 0)
    for k, spec in enumerate(specs_of(inputs)):
        a, b, chi, ndata, valid = C.fit(spec, order=order)
        assert valid
        assert res["ndata"][k] == ndata
        assert res["chi2"][k] <= chi + 1e-12 * abs(chi)
        assert res["chi2"][k] == pytest.approx(chi, rel=1e-9)
        assert res["a"][k] == pytest.approx(a, rel=1e-5)
        scale = max(abs(a), 1e-12)
        assert abs(res["b"][k] - b) <= 1e-5 * scale
        if order == 0:
            assert res["b"][k] == 0.0


def test_constant_weight_is_least_squares(inputs):
    res = continuum.fit_arrays(inputs, order=1, use_constant_weight=True)
    for k, spec in enumerate(specs_of(inputs)):
        a, b = C.ols(spec, order=1)
        assert res["a"][k] == pytest.approx(a, rel=1e-9)
        assert abs(res["b"][k] - b) <= 1e-9 * abs(a)
        assert res["chi2"][k] == pytest.approx(C.chi2(a, b, spec, True), rel=1e-12)


def test_starts_in_the_basin_reach_the_same_minimum(inputs):
    """chi2 is not convex (the log D term), so only starts near the minimum are bound to
    find it: perturb the plain fit's solution by up to 20 % and refit."""
    base = continuum.fit_arrays(inputs, order=1)
    n = len(inputs["offsets"]) - 1
    rng = np.random.default_rng(3)
    a0 = base["a"] * rng.uniform(0.8, 1.2, n)
    b0 = base["b"] + 0.2 * np.abs(base["a"]) * rng.uniform(-1, 1, n)
    res = continuum.fit_arrays(inputs, order=1, a0=a0, b0=b0)
    np.testing.assert_allclose(res["chi2"], base["chi2"], rtol=1e-10)
    # chi2 pins the minimum; along its flattest direction (a, b) are fixed only to about
    # sqrt(tol) of their scale
    np.testing.assert_allclose(res["a"], base["a"], rtol=1e-5)


def test_negative_and_zero_continuum_are_flagged():
    n = 200
    ll = np.log10(np.linspace(3900.0, 4200.0, n))
    x = (ll - ll[0]) / (ll[-1] - ll[0])
    flux_neg = 1.0 - 2.0 * x                # best line crosses zero: negative continuum
    flux_zero = np.ones(n)                  # mean continuum 0: the continuum is 0
    mean = np.concatenate([np.ones(n), np.zeros(n)])
    inputs = {"offsets": np.array([0, n, 2 * n]), "flux": np.concatenate([flux_neg, flux_zero]),
              "ivar": np.ones(2 * n), "mean_cont": mean, "x": np.concatenate([x, x]),
              "eta": np.ones(2 * n), "var_lss": np.zeros(2 * n), "fudge": np.zeros(2 * n)}
    res = continuum.fit_arrays(inputs, order=1, use_constant_weight=True,
                               a0=np.array([0.5, 0.1]), b0=np.array([0.0, 0.0]))
    assert res["status"][0] == 2
    assert res["status"][1] == 3
    spec = {"log_lambda": ll, "flux": flux_neg, "ivar": np.ones(n), "mean_cont": np.ones(n),
            "eta": np.ones(n), "var_lss": np.zeros(n), "fudge": np.zeros(n),
            "log_lambda_min": ll[0], "log_lambda_max": ll[-1]}
    assert not C.fit(spec, order=1, use_constant_weight=True)[4]


class _Forest:
    def __init__(self, spec, z):
        self.log_lambda = spec["log_lambda"]
        self.flux = spec["flux"]
        self.ivar = spec["ivar"]
        self.z = z
        self.transmission_correction = np.ones_like(self.flux)


def test_compute_continuum_api(inputs):
    grid = np.arange(np.log10(S.FOREST_MIN), np.log10(S.FOREST_MAX) + 1e-4 / 2, 1e-4)
    spec = S.forest_spec(inputs, 5)
    forest = _Forest(spec, inputs["z_qso"][5])
    one = lambda ll: np.ones_like(ll)
    cont, reason, params = continuum.compute_continuum(
        forest, one, one, lambda ll: np.full_like(ll, 0.1), lambda ll: np.zeros_like(ll),
        False, 1, grid)
    assert reason is None
    a, b, chi, ndata, valid = C.fit(spec, order=1)
    assert params[0] == pytest.approx(a, rel=1e-6)
    assert params[2] == pytest.approx(chi, rel=1e-9)
    assert params[3] == ndata
    expected = C.continuum_model(params[0], params[1], spec["log_lambda"],
                                 grid[0] + np.log10(1 + forest.z),
                                 grid[-1] + np.log10(1 + forest.z), np.ones_like(spec["flux"]))
    np.testing.assert_allclose(cont, expected, rtol=1e-12)
