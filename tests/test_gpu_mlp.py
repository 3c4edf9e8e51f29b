"""The MLP's BF16 MFMA kernels against the binary64 oracle (oracle/mlp_oracle.py).

Tolerance: the kernels round weights, inputs and hidden activations to
bfloat16 (8-bit mantissa) and accumulate in binary32, so gradients agree with
binary64 to about 1e-2 relative (norm-wise) -- the test uses 3e-2."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from pdos_amd.mlp import Mlp

pytestmark = pytest.mark.gpu


def _bf16(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16  # round to nearest even
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("shape", [(70, 96, 130, 200), (501, 512, 2545, 256)])
def test_forward_and_gradients_vs_oracle(shape):
    n_in, h, n_out, rows = shape
    print("mlp", shape, flush=True)
    rng = np.random.default_rng(1)
    net = Mlp(n_in, h, n_out, seed=3)
    p = net.get_params()
    p = p + 0.01 * rng.standard_normal(p.size).astype(np.float32)  # nonzero biases
    net.set_params(p)
    X = rng.standard_normal((rows, n_in)).astype(np.float32)
    T = rng.standard_normal((rows, n_out)).astype(np.float32)
    loss, g = net.loss_grad(X, T)
    pr = _bf16(p).astype(np.float64)
    b_off = [n_in * h, n_in * h + h + h * h, n_in * h + h + h * h + h + h * n_out]
    # biases stay binary32 in the kernels
    for o, k in zip(b_off, (h, h, n_out)):
        pr[o:o + k] = p[o:o + k]
    loss_ref, g_ref = MO.loss_grad(pr, _bf16(X), T, n_in, h, n_out)
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref
    sizes = [n_in * h, h, h * h, h, h * n_out, n_out]
    o = 0
    for name, k in zip(["W1", "b1", "W2", "b2", "W3", "b3"], sizes):
        a, b = g[o:o + k], g_ref[o:o + k]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert rel < 3e-2, (name, rel)
        o += k
    # forward through predict with identity affine maps
    net.set_affine(np.zeros(n_in), np.ones(n_in), np.ones(n_out), np.zeros(n_out))
    Y = net.predict(X)
    Yr, _ = MO.forward(pr, _bf16(X), n_in, h, n_out)
    assert np.linalg.norm(Y - Yr) <= 2e-2 * np.linalg.norm(Yr)


def test_training_reduces_loss():
    rng = np.random.default_rng(2)
    n_in, h, n_out, rows = 64, 128, 64, 4096
    A = rng.standard_normal((n_in, n_out)) / np.sqrt(n_in)
    X = rng.standard_normal((rows, n_in)).astype(np.float32)
    T = np.tanh(X @ A).astype(np.float32)
    net = Mlp(n_in, h, n_out, seed=5)
    l0, _ = net.loss_grad(X[:1024], T[:1024])
    net.train(X, T, epochs=20, batch=256, lr=2e-3, seed=1)
    l1, _ = net.loss_grad(X[:1024], T[:1024])
    assert l1 < 0.2 * l0, (l0, l1)
