"""CPU oracle of the continuum seed network -- TEST INFRASTRUCTURE ONLY.

Only tests use this module, as the checker of the GPU kernels of picca_amd.seed_mlp.

FP64 numpy restatement of the two-layer perceptron y = relu(x W1 + b1) W2 + b2 and of the
gradients of loss = sum (y - t)^2 / (2 n) by the chain rule:
    dy = (y - t) / n,  dW2 = h^T dy,  db2 = sum dy,  dh = (dy W2^T) * (h > 0),
    db1 = sum dh,      dW1 = x^T dh.
With ``bf16=True`` the operands the GPU multiplies on its matrix cores (x and W1 in the
forward product, x and dh in the dW1 product) are first rounded to bfloat16 (round to
nearest even), as the kernels do; everything else stays FP64.
"""
import numpy as np


def to_bf16(a):
    """Round float32 values to bfloat16 (nearest even) and return them as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward(params, x, bf16=True):
    w1, b1, w2, b2 = (np.asarray(p, dtype=np.float64) for p in params)
    xs = to_bf16(x) if bf16 else np.asarray(x, dtype=np.float64)
    w1s = to_bf16(w1.astype(np.float32)) if bf16 else w1
    h = np.maximum(xs @ w1s + b1, 0.0)
    return h, h @ w2 + b2


def gradients(params, x, t, bf16=True):
    """Returns (g_w1, g_b1, g_w2, g_b2, loss)."""
    w1, b1, w2, b2 = (np.asarray(p, dtype=np.float64) for p in params)
    n = x.shape[0]
    h, y = forward(params, x, bf16)
    e = y - np.asarray(t, dtype=np.float64)
    loss = 0.5 * (e ** 2).sum() / n
    dy = e / n
    g_w2 = h.T @ dy
    g_b2 = dy.sum(axis=0)
    dh = (dy @ w2.T) * (h > 0)
    g_b1 = dh.sum(axis=0)
    xs = to_bf16(x) if bf16 else np.asarray(x, dtype=np.float64)
    dhs = to_bf16(dh.astype(np.float32)) if bf16 else dh
    g_w1 = xs.T @ dhs
    return g_w1, g_b1, g_w2, g_b2, loss


def loss(params, x, t):
    _, y = forward(params, x, bf16=False)
    return 0.5 * ((y - t) ** 2).sum() / x.shape[0]
