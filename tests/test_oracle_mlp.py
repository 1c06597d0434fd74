"""Pins the NumPy q/rank-predictor oracle (oracle/mlp.py) before it is used to
check the HIP training kernels: its analytic gradient against central finite
differences, its GELU against the definition, and a hand-computed loss."""
import numpy as np
from scipy.special import erf

from oracle import mlp


def test_gelu_definition():
    z = np.linspace(-6, 6, 101)
    phi = 0.5 * (1 + erf(z / np.sqrt(2)))
    assert np.allclose(mlp.gelu(z), z * phi)
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    assert np.abs(fd - mlp.gelu_prime(z)).max() < 1e-8


def test_param_count_and_layout():
    assert mlp.n_params() == 128 * 64 + 128 + 2 * (128 * 128 + 128) + 6 * 128 + 6
    theta = np.arange(mlp.n_params(), dtype=np.float64)
    (W1, b1), (W2, b2), (W3, b3), (W4, b4) = mlp.unpack(theta)
    assert W1[0, 1] == 1 and b1[0] == 128 * 64 and W4.shape == (6, 128) and b4[-1] == theta[-1]


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = mlp.init_params(1) + 0.01 * rng.standard_normal(mlp.n_params())
    X = rng.standard_normal((7, 64))
    q = rng.integers(0, 5, 7)
    r = rng.random(7)
    loss, g = mlp.loss_and_grad(theta, X, q, r, 0.7)
    idx = rng.choice(mlp.n_params(), 60, replace=False)
    idx = np.concatenate([idx, [0, 8192, 8320, mlp.n_params() - 1, mlp.n_params() - 7]])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mlp.loss_and_grad(tp, X, q, r, 0.7)[0] - mlp.loss_and_grad(tm, X, q, r, 0.7)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_loss_of_a_known_case():
    # all weights zero except b4: o = b4 for every sample
    theta = np.zeros(mlp.n_params())
    b4 = np.array([1.0, 2.0, 0.0, -1.0, 0.5, 0.3])
    theta[-6:] = b4
    X = np.ones((2, 64))
    loss, _ = mlp.loss_and_grad(theta, X, np.array([1, 3]), np.array([0.3, 0.5]), 2.0)
    lse = np.log(np.exp(b4[:5]).sum())
    want = np.mean([lse - 2.0 + 0.0, lse + 1.0 + 2.0 * 0.2 ** 2])
    assert abs(loss - want) < 1e-12
