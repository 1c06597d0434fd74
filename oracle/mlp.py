"""ORACLE (test infrastructure only): NumPy restatement of the q/rank predictor.

The learned path of the library (csrc/learned.hip) serves and trains a small
multilayer perceptron with its own HIP kernels.  This module computes the same
network, loss and gradients in float64 NumPy so the tests can check the HIP
forward pass and gradients against it; its own gradient is pinned against
central finite differences (tests/test_oracle_mlp.py).  Only tests/ import it.

Network (all weights [out][in] row-major, packed in this order, float32 on
the device):
    h1 = gelu(W1 x + b1)      W1: 128 x 64
    h2 = gelu(W2 h1 + b2)     W2: 128 x 128
    h3 = gelu(W3 h2 + b3)     W3: 128 x 128
    o  = W4 h3 + b4           W4:   6 x 128
o[0:5] are the logits of q in {0, 1, 2, 3, 4}; o[5] regresses the rank as a
fraction of min(m, n).  gelu(z) = z Phi(z) (exact, erf form).  Per-sample loss:
    L = logsumexp(o[0:5]) - o[q] + lam * (o[5] - r)^2,
averaged over the batch.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

SIZES = [(128, 64), (128, 128), (128, 128), (6, 128)]


def n_params() -> int:
    return sum(o * i + o for o, i in SIZES)


def unpack(theta: np.ndarray):
    out, off = [], 0
    for o, i in SIZES:
        W = theta[off:off + o * i].reshape(o, i)
        off += o * i
        b = theta[off:off + o]
        off += o
        out.append((W, b))
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_prime(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(theta: np.ndarray, X: np.ndarray):
    """X: batch x 64 -> outputs batch x 6 (and the cache for backward)."""
    layers = unpack(np.asarray(theta, dtype=np.float64))
    h = np.asarray(X, dtype=np.float64)
    cache = [h]
    for li, (W, b) in enumerate(layers):
        z = h @ W.T + b
        if li < 3:
            cache.append(z)
            h = gelu(z)
            cache.append(h)
        else:
            h = z
    return h, cache


def loss_and_grad(theta, X, q, r, lam):
    """Mean loss over the batch and its gradient w.r.t. the packed parameters."""
    theta = np.asarray(theta, dtype=np.float64)
    layers = unpack(theta)
    o, cache = forward(theta, X)
    B = X.shape[0]
    lg = o[:, :5]
    mx = lg.max(axis=1, keepdims=True)
    e = np.exp(lg - mx)
    lse = mx[:, 0] + np.log(e.sum(axis=1))
    q = np.asarray(q, dtype=np.int64)
    r = np.asarray(r, dtype=np.float64)
    loss = np.mean(lse - lg[np.arange(B), q] + lam * (o[:, 5] - r) ** 2)
    do = np.zeros_like(o)
    do[:, :5] = e / e.sum(axis=1, keepdims=True)
    do[np.arange(B), q] -= 1.0
    do[:, 5] = 2.0 * lam * (o[:, 5] - r)
    do /= B
    grads = [None] * 4
    delta = do
    for li in range(3, -1, -1):
        W, b = layers[li]
        h_in = cache[0] if li == 0 else cache[2 * li]
        grads[li] = (delta.T @ h_in, delta.sum(axis=0))
        if li > 0:
            dh = delta @ W
            z = cache[2 * li - 1]
            delta = dh * gelu_prime(z)
    g = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, g


def init_params(seed: int) -> np.ndarray:
    """He-style initialisation (the device uses its own; tests pass explicit
    parameters to both sides)."""
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in SIZES:
        parts.append(rng.standard_normal(o * i) * np.sqrt(2.0 / i))
        parts.append(np.zeros(o))
    return np.concatenate(parts)
