"""Pins the MLP oracle: backpropagated gradients equal central finite
differences of the loss (closed form check), and Adam matches its textbook
update on a scalar."""
import numpy as np

from oracle import mlp_oracle as MO


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    F, H, K, n = 6, 8, 3, 17
    p = rng.normal(0, 0.5, H * F + H + K * H + K)
    X = rng.normal(0, 1, (n, F))
    T = rng.normal(0, 1, (n, K))
    M = (rng.random((n, K)) < 0.7).astype(float)
    _, g = MO.grad(p, X, T, M, F, H, K, 0.1)
    fd = MO.fd_grad(p, X, T, M, F, H, K, 0.1)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-7)


def test_adam_textbook_scalar():
    p, m, v = MO.adam(np.array([1.0]), np.array([0.5]), np.zeros(1), np.zeros(1), 0.1, 0.9, 0.999, 1e-8, 1)
    # first step: mhat = g, vhat = g^2 -> p - lr * sign(g)
    assert np.allclose(p, [1.0 - 0.1 * 0.5 / (0.5 + 1e-8)])
