"""Binary64 numpy restatement of the warm-start MLP (TEST INFRASTRUCTURE).

Same architecture, parameter layout and loss as the library's kernels
(pdos_amd/csrc/mlp.hip): W1 [n_in x h], b1, W2 [h x h], b2, W3 [h x n_out], b3;
Z1 = X W1 + b1, H1 = gelu(Z1), Z2 = H1 W2 + b2, H2 = gelu(Z2), Y = H2 W3 + b3,
loss = mean((Y - T)^2) over rows and outputs; gelu(z) = z Phi(z).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def unpack(p, n_in, h, n_out):
    sizes = [(n_in, h), (h,), (h, h), (h,), (h, n_out), (n_out,)]
    out, o = [], 0
    for s in sizes:
        k = int(np.prod(s))
        out.append(np.asarray(p[o:o + k], dtype=np.float64).reshape(s))
        o += k
    assert o == len(p)
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, X, n_in, h, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, h, n_out)
    Z1 = X @ W1 + b1
    H1 = gelu(Z1)
    Z2 = H1 @ W2 + b2
    H2 = gelu(Z2)
    return H2 @ W3 + b3, (Z1, H1, Z2, H2)


def loss_grad(p, X, T, n_in, h, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, h, n_out)
    X = np.asarray(X, np.float64)
    Y, (Z1, H1, Z2, H2) = forward(p, X, n_in, h, n_out)
    R = Y - T
    loss = float(np.mean(R * R))
    dY = 2.0 * R / R.size
    gW3 = H2.T @ dY
    gb3 = dY.sum(0)
    dZ2 = (dY @ W3.T) * gelu_grad(Z2)
    gW2 = H1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * gelu_grad(Z1)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(0)
    return loss, np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])
