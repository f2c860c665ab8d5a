"""Pins the seed-network oracle (oracle/mlp_oracle.py): its analytic gradients equal
central finite differences of the loss in FP64 (bf16 rounding off), and its bfloat16
rounding is round-to-nearest-even on the float32 bit pattern."""
import numpy as np

from oracle import mlp_oracle as M


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    n, n_in, n_hidden = 7, 32, 16
    x = rng.normal(size=(n, n_in))
    t = rng.normal(size=(n, 2))
    params = [rng.normal(0, 0.3, (n_in, n_hidden)), rng.normal(0, 0.1, n_hidden),
              rng.normal(0, 0.3, (n_hidden, 2)), rng.normal(0, 0.1, 2)]
    grads = M.gradients(params, x, t, bf16=False)
    for k, p in enumerate(params):
        flat = p.reshape(-1)
        for idx in rng.choice(flat.size, min(12, flat.size), replace=False):
            old = flat[idx]
            h = 1e-6
            flat[idx] = old + h
            lp = M.loss(params, x, t)
            flat[idx] = old - h
            lm = M.loss(params, x, t)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(grads[k].reshape(-1)[idx] - fd) < 1e-6 * max(1.0, abs(fd))


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.0e-5, 65504.0, 0.0,
                     1.0 + 2 ** -7 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8 + 2 ** -10], dtype=np.float32)
    out = M.to_bf16(vals)
    # exactly representable values stay; ties go to even; other values to the nearest
    assert out[0] == 1.0
    assert out[1] == 1.0                       # tie between 1 and 1 + 2^-7 -> even (1.0)
    assert out[2] == 1.0 + 2 ** -7             # below the tie -> rounds down
    assert out[6] == 1.0 + 2 ** -6             # tie between odd and even -> even (up)
    assert out[7] == 1.0 + 2 ** -6             # above the tie -> rounds up
    assert abs(out[3] - (-3.0e-5)) <= 2 ** -8 * 3.0e-5
    assert out[5] == 0.0
