"""Pins the binary64 MLP oracle on the CPU: analytic gradients against
central finite differences, and GELU against its closed form values."""
import math

import numpy as np

from oracle import mlp_oracle as MO


def test_gelu_known_values():
    assert MO.gelu(np.array([0.0]))[0] == 0.0
    # gelu(1) = Phi(1) = 0.8413447460685429...
    assert abs(MO.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    z = np.linspace(-4, 4, 81)
    h = 1e-6
    fd = (MO.gelu(z + h) - MO.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(MO.gelu_grad(z), fd, atol=1e-8)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    n_in, h, n_out, rows = 5, 7, 4, 6
    npar = n_in * h + h + h * h + h + h * n_out + n_out
    p = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((rows, n_in))
    T = rng.standard_normal((rows, n_out))
    loss, g = MO.loss_grad(p, X, T, n_in, h, n_out)
    eps = 1e-6
    for i in rng.choice(npar, size=40, replace=False):
        e = np.zeros(npar)
        e[i] = eps
        fd = (MO.loss_grad(p + e, X, T, n_in, h, n_out)[0] - MO.loss_grad(p - e, X, T, n_in, h, n_out)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 * (1 + abs(fd)), (i, fd, g[i])
