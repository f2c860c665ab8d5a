"""Learned start values for the continuum fit.

A two-layer perceptron (256 binned-flux inputs -> 128 relu units -> 2 outputs) predicts,
from a quasar's rest-frame spectrum binned by ``synth_spectra.binned_inputs``, the zero
point and slope (a, b) of picca's continuum fit divided by the spectrum's mean flux.  The
batched GPU fit (``continuum.fit_arrays``) then starts from these values instead of
Migrad's start (mean flux, slope 0); the plain fit is the judge: the seeded fit must end
at the same minimum chi2.

Training and inference run on the GPU (``picca_mlp_*`` in the C-ABI): BF16 matrix-core
products with FP32 accumulation, Adam in FP32.  Targets are the (a, b) of the last
iteration of picca's DR16 loop (``continuum.compute_expected_flux``) on a training set of
spectra.
"""
import ctypes

import numpy as np

from . import _native
from . import engine as _engine
from ._native import as_f64, check, ptr


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class SeedNetwork:
    def __init__(self, n_in=256, n_hidden=128, seed=0, device=0, params=None):
        self.n_in = n_in
        self.n_hidden = n_hidden
        self._eng = _engine.get_engine(device)
        self._lib = self._eng._lib
        if params is None:
            rng = np.random.default_rng(seed)
            params = (rng.normal(0.0, np.sqrt(2.0 / n_in), (n_in, n_hidden)),
                      np.zeros(n_hidden),
                      rng.normal(0.0, np.sqrt(1.0 / n_hidden), (n_hidden, 2)) * 0.1,
                      np.zeros(2))
        self.set_params(params)

    def set_params(self, params):
        w1, b1, w2, b2 = (_f32(p) for p in params)
        assert w1.shape == (self.n_in, self.n_hidden) and w2.shape == (self.n_hidden, 2)
        check(self._lib.picca_mlp_init(self._eng._ctx, self.n_in, self.n_hidden, ptr(w1),
                                       ptr(b1), ptr(w2), ptr(b2)), "picca_mlp_init")

    def params(self):
        out = (np.zeros((self.n_in, self.n_hidden), np.float32), np.zeros(self.n_hidden, np.float32),
               np.zeros((self.n_hidden, 2), np.float32), np.zeros(2, np.float32))
        check(self._lib.picca_mlp_get(self._eng._ctx, *(ptr(p) for p in out)), "picca_mlp_get")
        return out

    def predict(self, x, timing=False):
        """Network outputs [n, 2]; with ``timing`` also the forward kernel's device ms."""
        x = _f32(x)
        y = np.zeros((x.shape[0], 2), np.float32)
        ms = ctypes.c_double()
        check(self._lib.picca_mlp_forward(self._eng._ctx, x.shape[0], ptr(x), ptr(y),
                                          ctypes.byref(ms)), "picca_mlp_forward")
        return (y, ms.value) if timing else y

    def gradients(self, x, t):
        x, t = _f32(x), _f32(t)
        g = (np.zeros((self.n_in, self.n_hidden), np.float32), np.zeros(self.n_hidden, np.float32),
             np.zeros((self.n_hidden, 2), np.float32), np.zeros(2, np.float32))
        loss = np.zeros(1, np.float32)
        check(self._lib.picca_mlp_gradients(self._eng._ctx, x.shape[0], ptr(x), ptr(t),
                                            *(ptr(a) for a in g), ptr(loss)),
              "picca_mlp_gradients")
        return g + (float(loss[0]),)

    def train(self, x, t, epochs=20, batch=512, lr=1e-3, seed=0):
        """Adam over minibatches; returns the loss of every step."""
        x, t = _f32(x), _f32(t)
        check(self._lib.picca_mlp_set_data(self._eng._ctx, x.shape[0], ptr(x), ptr(t)),
              "picca_mlp_set_data")
        steps = max(1, epochs * (x.shape[0] // batch))
        losses = np.zeros(steps, np.float32)
        check(self._lib.picca_mlp_train(self._eng._ctx, steps, batch, ctypes.c_float(lr),
                                        ctypes.c_uint64(seed), ptr(losses)), "picca_mlp_train")
        return losses


def targets(a, b, scale):
    """Training targets: the fit's (a, b) over the spectrum's mean flux."""
    return np.column_stack([as_f64(a) / scale, as_f64(b) / scale]).astype(np.float32)


def seeds(y, scale):
    """Start values of the continuum fit from the network's outputs."""
    y = np.asarray(y, dtype=np.float64)
    return y[:, 0] * scale, y[:, 1] * scale
