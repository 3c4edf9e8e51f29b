"""pdos_mlp_predict_dev_noalloc (the entry point that allocates nothing)
against pdos_mlp_predict_dev on the same network and inputs: equal bit for
bit, on two calls in a row (the second with fewer rows), through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import pdos_amd
from pdos_amd import DeviceArray
from pdos_amd.mlp import Mlp

pytestmark = pytest.mark.gpu


def _net(n_in, n_h, n_out, seed):
    net = Mlp(n_in, n_h, n_out, seed=seed)
    rng = np.random.default_rng(seed)
    net.set_affine(rng.normal(size=n_in), 0.5 + rng.random(n_in), 0.5 + rng.random(n_out),
                   rng.normal(size=n_out))
    return net


def _both(net, X, ldy):
    rows = X.shape[0]
    dX = DeviceArray(X)
    ref = DeviceArray(np.full((rows, ldy), np.nan, dtype=np.float32))
    out = DeviceArray(np.full((rows, ldy), np.nan, dtype=np.float32))
    L = pdos_amd.lib("indirect")
    assert L.pdos_mlp_predict_dev(net.h, dX.ptr, rows, ref.ptr, ldy) == 0, L.pdos_mlp_last_error()
    assert L.pdos_mlp_predict_dev_noalloc(net.h, dX.ptr, rows, out.ptr, ldy) == 0, L.pdos_mlp_last_error()
    return ref.numpy(), out.numpy()


@pytest.mark.parametrize("shape,rows", [((501, 512, 2545), (3000, 700)), ((37, 96, 70), (130, 5))])
def test_noalloc_predict_equals_predict_dev(shape, rows):
    print(f"case noalloc {shape} rows {rows}", flush=True)
    net = _net(*shape, seed=11)
    rng = np.random.default_rng(5)
    ldy = shape[2] + 3
    for r in rows:  # the second call is smaller and reuses the scratch the first one filled
        X = rng.normal(size=(r, shape[0])).astype(np.float32)
        ref, out = _both(net, X, ldy)
        assert np.isfinite(ref[:, :shape[2]]).all()
        assert np.array_equal(ref[:, :shape[2]], out[:, :shape[2]])
        assert np.isnan(out[:, shape[2]:]).all()  # padding columns of dY untouched


def test_noalloc_refuses_a_network_wider_than_the_scratch():
    n_out = 140_000  # 64 rows of binary32 output alone: 35.8 MB > 32 MiB
    net = Mlp(16, 64, n_out, seed=1)
    dX = DeviceArray(np.zeros((4, 16), dtype=np.float32))
    dY = DeviceArray(shape=(4, n_out))
    L = pdos_amd.lib("indirect")
    assert L.pdos_mlp_predict_dev_noalloc(net.h, dX.ptr, 4, dY.ptr, n_out) == -1
    assert b"too wide" in L.pdos_mlp_last_error()
