"""Pins the continuum-fit oracle (oracle/cont_oracle.py) without the reference.

iminuit is not installed, so picca's own fit cannot run here and no golden (a, b) exist for
synthetic spectra.  What fixes the oracle instead:
  * use_constant_weight turns the objective into a plain sum of squares: the fit must be
    the closed-form linear least-squares solution;
  * the weights of get_continuum_weights equal 1 / (eta / ivar + var_lss C^2 + fudge ivar
    C^4) (the form the GPU kernel differentiates);
  * at the minimum the finite-difference gradient vanishes, the Hessian is positive
    definite, and no point of a dense grid around it is lower;
  * order 0 keeps the slope at 0;
  * stacking noise-free spectra whose flux is exactly proportional to their continuum
    leaves the mean continuum unchanged (mean of F/C = 1 after normalisation).
"""
import numpy as np
import pytest

from oracle import cont_oracle as C
from picca_amd import synth_spectra as S


@pytest.fixture(scope="module")
def forests():
    spectra = S.make_spectra(24, seed=11)
    inputs = S.forest_inputs(spectra)
    return [S.forest_spec(inputs, k) for k in range(spectra.num_spectra)]


def test_weights_closed_form(forests):
    spec = forests[0]
    rng = np.random.default_rng(0)
    cont = np.abs(rng.normal(2.0, 0.5, spec["flux"].size))
    eta = rng.uniform(0.5, 1.5, cont.size)
    lss = rng.uniform(0.0, 0.3, cont.size)
    fudge = rng.uniform(0.0, 1e-3, cont.size)
    ivar = spec["ivar"].copy()
    ivar[::7] = 0
    w = C.continuum_weights(ivar, cont, eta, lss, fudge, False)
    m = ivar > 0
    d = eta[m] / ivar[m] + lss[m] * cont[m] ** 2 + fudge[m] * ivar[m] * cont[m] ** 4
    np.testing.assert_allclose(w[m], 1.0 / d, rtol=1e-12)
    assert np.all(w[~m] == 0)


@pytest.mark.parametrize("order", [0, 1])
def test_constant_weight_is_least_squares(forests, order):
    for spec in forests[:8]:
        a, b, chi, ndata, valid = C.fit(spec, order=order, use_constant_weight=True)
        a_ls, b_ls = C.ols(spec, order=order)
        np.testing.assert_allclose([a, b], [a_ls, b_ls], rtol=1e-7, atol=1e-9 * abs(a_ls))
        assert chi == pytest.approx(C.chi2(a_ls, b_ls, spec, True), rel=1e-12)
        assert ndata == int((spec["ivar"] > 0).sum())


def test_minimum_is_stationary_and_lowest(forests):
    for spec in forests[:10]:
        a, b, chi, ndata, valid = C.fit(spec, order=1)
        f = lambda p, q: C.chi2(p, q, spec)
        h = 1e-5 * max(abs(a), 1e-3)
        ga = (f(a + h, b) - f(a - h, b)) / (2 * h)
        gb = (f(a, b + h) - f(a, b - h)) / (2 * h)
        haa = (f(a + h, b) - 2 * chi + f(a - h, b)) / h ** 2
        hbb = (f(a, b + h) - 2 * chi + f(a, b - h)) / h ** 2
        hab = (f(a + h, b + h) - f(a + h, b - h) - f(a - h, b + h) + f(a - h, b - h)) / (4 * h * h)
        assert abs(ga) < 1e-4 * haa * abs(a) + 1e-6
        assert abs(gb) < 1e-4 * hbb * abs(a) + 1e-6
        assert haa > 0 and haa * hbb - hab * hab > 0
        # a dense grid around the minimum has no lower point
        da = 0.2 * abs(a) + 1e-3
        grid = [f(a + s * da, b + t * da) for s in np.linspace(-1, 1, 21)
                for t in np.linspace(-1, 1, 21)]
        assert min(grid) >= chi - 1e-9 * max(1.0, abs(chi))


def test_order_zero_keeps_slope(forests):
    a, b, chi, ndata, valid = C.fit(forests[3], order=0)
    assert b == 0.0
    assert chi <= C.chi2(C.start_point(forests[3])[0], 0.0, forests[3])


def test_stack_of_proportional_spectra_is_unchanged(forests):
    grid = np.arange(np.log10(1040.0), np.log10(1200.0) + 1e-4 / 2, 3e-4)
    previous = np.ones_like(grid)
    specs, conts = [], []
    rng = np.random.default_rng(4)
    for spec in forests[:6]:
        s = dict(spec)
        cont = rng.uniform(1.0, 3.0) * (1 + 0.3 * (s["log_lambda"] - s["log_lambda"][0]))
        s["flux"] = 0.8 * cont
        specs.append(s)
        conts.append(cont)
    kept, new, weight = C.stack_mean_cont(specs, conts, grid, C.find_bins_log, previous)
    np.testing.assert_allclose(new, 1.0, rtol=1e-12)
    assert np.all(weight > 0)
