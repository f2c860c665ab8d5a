"""GPU kernels of the continuum seed network against the FP64 oracle, and what the trained
network gains the continuum fit.

Bar: forward outputs within 1e-4 of the oracle's (relative to the output scale; the
oracle rounds the matrix-core operands to bfloat16 like the kernels and accumulates in
FP64, the kernels in FP32); gradients within 2e-3 of the largest gradient of each
parameter; the seeded fit ends at the plain fit's minimum (chi2 within 1e-9 relative)
with fewer objective evaluations.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from picca_amd import continuum, seed_mlp, synth_spectra as S

pytestmark = pytest.mark.gpu


def random_params(rng, n_in, n_hidden):
    return [rng.normal(0, np.sqrt(2.0 / n_in), (n_in, n_hidden)).astype(np.float32),
            rng.normal(0, 0.1, n_hidden).astype(np.float32),
            rng.normal(0, 0.3, (n_hidden, 2)).astype(np.float32),
            rng.normal(0, 0.1, 2).astype(np.float32)]


@pytest.mark.parametrize("shape", [(256, 128, 100), (64, 32, 37)])
def test_forward_matches_oracle(shape):
    n_in, n_hidden, n = shape
    rng = np.random.default_rng(1)
    params = random_params(rng, n_in, n_hidden)
    net = seed_mlp.SeedNetwork(n_in, n_hidden, params=params)
    x = rng.normal(size=(n, n_in)).astype(np.float32)
    y = net.predict(x)
    _, ref = M.forward(params, x)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-4 * scale)
    got = net.params()
    for p, q in zip(got, params):
        np.testing.assert_array_equal(p, q)


@pytest.mark.parametrize("shape", [(256, 128, 300), (64, 32, 50)])
def test_gradients_match_oracle(shape):
    n_in, n_hidden, n = shape
    rng = np.random.default_rng(2)
    params = random_params(rng, n_in, n_hidden)
    net = seed_mlp.SeedNetwork(n_in, n_hidden, params=params)
    x = rng.normal(size=(n, n_in)).astype(np.float32)
    t = rng.normal(size=(n, 2)).astype(np.float32)
    *g, loss = net.gradients(x, t)
    *ref, ref_loss = M.gradients(params, x, t)
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    for mine, expect, name in zip(g, ref, ("w1", "b1", "w2", "b2")):
        scale = np.abs(expect).max()
        np.testing.assert_allclose(mine, expect, rtol=0, atol=2e-3 * scale, err_msg=name)


def test_training_lowers_the_loss():
    rng = np.random.default_rng(3)
    n_in, n = 64, 4096
    x = rng.normal(size=(n, n_in)).astype(np.float32)
    wt = rng.normal(size=(n_in, 2)) / np.sqrt(n_in)
    t = (np.tanh(x @ wt) + 0.5).astype(np.float32)
    net = seed_mlp.SeedNetwork(n_in, 64, seed=0)
    losses = net.train(x, t, epochs=30, batch=256, lr=2e-3)
    assert np.all(np.isfinite(losses))
    assert losses[-10:].mean() < 0.1 * losses[:10].mean()


def test_seeded_fit_reaches_the_plain_minimum_faster():
    spectra = S.make_spectra(6000, seed=7)
    inputs = S.forest_inputs(spectra)
    x, scale = S.binned_inputs(spectra)
    plain = continuum.fit_arrays(inputs, order=1)
    ok = plain["status"] == 0
    train = np.arange(6000) < 5000
    net = seed_mlp.SeedNetwork(seed=0)
    sel = train & ok
    net.train(x[sel], seed_mlp.targets(plain["a"][sel], plain["b"][sel], scale[sel]),
              epochs=40, batch=256, lr=1e-3)
    test = ~train
    a0, b0 = seed_mlp.seeds(net.predict(x), scale)
    # seeds for the held-out spectra only; the training spectra keep their plain start
    start_a, start_b = continuum.start_values(inputs["offsets"], inputs["flux"], inputs["ivar"])
    a0 = np.where(test, a0, start_a)
    b0 = np.where(test, b0, start_b)
    seeded = continuum.fit_arrays(inputs, order=1, a0=a0, b0=b0)
    both = test & ok & (seeded["status"] == 0)
    assert both.sum() > 0.95 * test.sum()
    np.testing.assert_allclose(seeded["chi2"][both], plain["chi2"][both], rtol=1e-9)
    assert seeded["evaluations"][both].mean() < plain["evaluations"][both].mean()
