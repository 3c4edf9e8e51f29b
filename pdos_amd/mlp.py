"""Python handle on the library's MLP (pdos_mlp_* in include/pdos_ext.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib

_fp = C.POINTER(C.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Mlp:
    """n_in -> n_hidden -> n_hidden -> n_out, GELU; BF16 MFMA kernels."""

    def __init__(self, n_in: int, n_hidden: int, n_out: int, seed: int = 0, device: int = 0):
        self.L = lib("indirect")
        self.n_in, self.n_hidden, self.n_out = n_in, n_hidden, n_out
        self.h = self.L.pdos_mlp_create(n_in, n_hidden, n_out, seed, device)
        if not self.h:
            raise RuntimeError(self.L.pdos_mlp_last_error().decode())

    @property
    def num_params(self) -> int:
        return self.L.pdos_mlp_num_params(self.h)

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, dtype=np.float32)
        self.L.pdos_mlp_get_params(self.h, out.ctypes.data_as(_fp))
        return out

    def set_params(self, p) -> None:
        p = _f32(p)
        assert p.size == self.num_params
        self.L.pdos_mlp_set_params(self.h, p.ctypes.data_as(_fp))

    def set_affine(self, in_mean, in_std, out_std, out_mean) -> None:
        a, b, c, d = (_f32(v) for v in (in_mean, in_std, out_std, out_mean))
        self._aff = (a, b, c, d)
        self.L.pdos_mlp_set_affine(self.h, *(v.ctypes.data_as(_fp) for v in (a, b, c, d)))

    def loss_grad(self, X, T):
        X, T = _f32(X), _f32(T)
        loss = C.c_double()
        g = np.zeros(self.num_params, dtype=np.float32)
        rc = self.L.pdos_mlp_loss_grad(self.h, X.ctypes.data_as(_fp), T.ctypes.data_as(_fp), X.shape[0],
                                       C.byref(loss), g.ctypes.data_as(_fp))
        if rc != 0:
            raise RuntimeError(self.L.pdos_mlp_last_error().decode())
        return loss.value, g

    def train(self, X, T, epochs: int, batch: int, lr: float, seed: int = 0) -> float:
        X, T = _f32(X), _f32(T)
        loss = C.c_double()
        rc = self.L.pdos_mlp_train(self.h, X.ctypes.data_as(_fp), T.ctypes.data_as(_fp), X.shape[0],
                                   epochs, batch, lr, seed, C.byref(loss))
        if rc != 0:
            raise RuntimeError(self.L.pdos_mlp_last_error().decode())
        return loss.value

    def predict(self, X) -> np.ndarray:
        X = _f32(X)
        Y = np.zeros((X.shape[0], self.n_out), dtype=np.float32)
        ms = C.c_double()
        rc = self.L.pdos_mlp_predict(self.h, X.ctypes.data_as(_fp), X.shape[0], Y.ctypes.data_as(_fp),
                                     C.byref(ms))
        if rc != 0:
            raise RuntimeError(self.L.pdos_mlp_last_error().decode())
        self.last_predict_ms = ms.value
        return Y

    def predict_dev(self, dX, rows: int, dY, ldy: int) -> None:
        rc = self.L.pdos_mlp_predict_dev(self.h, dX.ptr, rows, dY.ptr, ldy)
        if rc != 0:
            raise RuntimeError(self.L.pdos_mlp_last_error().decode())

    def predict_dev_noalloc(self, dX, rows: int, dY, ldy: int) -> None:
        """predict_dev without allocation (pdos_mlp_predict_dev_noalloc)."""
        rc = self.L.pdos_mlp_predict_dev_noalloc(self.h, dX.ptr, rows, dY.ptr, ldy)
        if rc != 0:
            raise RuntimeError(self.L.pdos_mlp_last_error().decode())

    def __del__(self):
        try:
            if self.h:
                self.L.pdos_mlp_free(self.h)
                self.h = None
        except Exception:
            pass
