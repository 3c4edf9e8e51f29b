"""Python handle on the shared-A family engine (pdos_family_* in include/pdos_ext.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import PdosOptions, lib, options
from .abi import Problem

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)


def _bind(L):
    if getattr(L, "_fam_bound", False):
        return
    L.pdos_family_create.restype = C.c_void_p
    L.pdos_family_create.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(PdosOptions)]
    L.pdos_family_load.restype = C.c_int
    L.pdos_family_load.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    L.pdos_family_run.restype = C.c_int
    L.pdos_family_run.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_int, _dp]
    L.pdos_family_fetch.restype = C.c_int
    L.pdos_family_fetch.argtypes = [C.c_void_p, _dp, _dp, _dp, _lp, _lp, _lp, _dp]
    L.pdos_family_free.argtypes = [C.c_void_p]
    L.pdos_family_last_error.restype = C.c_char_p
    L.pdos_family_timing.restype = C.c_int
    L.pdos_family_timing.argtypes = [C.c_void_p, C.c_int, _dp, _lp, _dp]
    L.pdos_family_info.restype = C.c_int
    L.pdos_family_info.argtypes = [C.c_void_p, _lp]
    L.pdos_family_engines.restype = C.c_int
    L.pdos_family_engines.argtypes = [C.c_void_p, C.c_int]
    L.pdos_family_scales.restype = C.c_int
    L.pdos_family_scales.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
    L._fam_bound = True


class Family:
    """Instances sharing the template's A, cones and parameters; each has its
    own c (and optionally b)."""

    def __init__(self, template: Problem, precision: str = "fp64", device: int = 0, bf16_products: int = 0,
                 equil: str = "reference"):
        self.L = lib("direct")
        _bind(self.L)
        self.n, self.m = template.n, template.m
        self.h = None
        d, k, self._keep = template.to_c()
        o = options("direct", precision, "auto", device, bf16_products, equil)
        self.h = self.L.pdos_family_create(C.byref(d), C.byref(k), C.byref(o))
        if not self.h:
            raise RuntimeError(self.L.pdos_family_last_error().decode())
        self.count = 0

    def load(self, Cm, Bm=None) -> None:
        Cm = np.ascontiguousarray(Cm, dtype=np.float64)
        self._C = Cm
        self._B = None if Bm is None else np.ascontiguousarray(Bm, dtype=np.float64)
        self.count = Cm.shape[0]
        rc = self.L.pdos_family_load(self.h, self.count, Cm.ctypes.data_as(_dp),
                                     self._B.ctypes.data_as(_dp) if self._B is not None else None)
        if rc != 0:
            raise RuntimeError(self.L.pdos_family_last_error().decode())

    def run(self, warm=None, ld: int = 0, judge: bool = True) -> float:
        """Solves the loaded batch; warm is a DeviceArray [count x ld] of
        (x, y, s) in original units.  Returns the device time in ms."""
        ms = C.c_double()
        rc = self.L.pdos_family_run(self.h, warm.ptr if warm is not None else None, ld, int(judge),
                                    C.byref(ms))
        if rc != 0:
            raise RuntimeError(self.L.pdos_family_last_error().decode())
        return ms.value

    def scales(self) -> dict:
        """Per loaded instance: lambda and the stop test's scales 1 + |b|,
        1 + |c|, as computed on the device at load."""
        lam, ps, ds = np.zeros(self.count), np.zeros(self.count), np.zeros(self.count)
        rc = self.L.pdos_family_scales(self.h, lam.ctypes.data_as(_dp), ps.ctypes.data_as(_dp),
                                       ds.ctypes.data_as(_dp))
        if rc != 0:
            raise RuntimeError(self.L.pdos_family_last_error().decode())
        return dict(lam=lam, pscale=ps, dscale=ds)

    def fetch(self) -> dict:
        c = self.count
        X = np.zeros((c, self.n))
        Y = np.zeros((c, self.m))
        S = np.zeros((c, self.m))
        it = np.zeros(c, dtype=np.int64)
        st = np.zeros(c, dtype=np.int64)
        fb = np.zeros(c, dtype=np.int64)
        rs = np.zeros((c, 5))
        rc = self.L.pdos_family_fetch(self.h, X.ctypes.data_as(_dp), Y.ctypes.data_as(_dp),
                                      S.ctypes.data_as(_dp), it.ctypes.data_as(_lp),
                                      st.ctypes.data_as(_lp), fb.ctypes.data_as(_lp),
                                      rs.ctypes.data_as(_dp))
        if rc != 0:
            raise RuntimeError(self.L.pdos_family_last_error().decode())
        return dict(x=X, y=Y, s=S, iters=it, status=np.where(st == 0, "Solved", "Indeterminate"),
                    fallback=fb, resid=rs)

    KINDS = ["phase_a", "At_product", "phase_b", "xupdate_gemm", "A_product", "phase_c"]

    def timing(self, enable: bool) -> dict:
        """Returns per-kernel {kind: (total_ms, calls, columns)} since the last
        call and switches event timing on or off for later runs."""
        ms = np.zeros(6)
        calls = np.zeros(6, dtype=np.int64)
        cols = np.zeros(6)
        self.L.pdos_family_timing(self.h, int(enable), ms.ctypes.data_as(_dp), calls.ctypes.data_as(_lp),
                                  cols.ctypes.data_as(_dp))
        return {k: (float(ms[i]), int(calls[i]), float(cols[i])) for i, k in enumerate(self.KINDS)}

    def engines(self, n: int = 0) -> int:
        """Engines the next load splits a large mixed batch over (1 or 2; 0
        only queries)."""
        return int(self.L.pdos_family_engines(self.h, int(n)))

    def info(self) -> dict:
        out = np.zeros(9, dtype=np.int64)
        self.L.pdos_family_info(self.h, out.ctypes.data_as(_lp))
        return dict(n=int(out[0]), m=int(out[1]), np=int(out[2]), compactions=int(out[3]),
                    dense=bool(out[4]), fp64=bool(out[5]), products=int(out[6]),
                    dense_rows=int(out[7]), engines=int(out[8]))

    def __del__(self):
        try:
            if self.h:
                self.L.pdos_family_free(self.h)
                self.h = None
        except Exception:
            pass


def polish(template: Problem, Cm, mixed: dict, Bm=None, eps_abs: float = 1e-4, max_iters: int = 300,
           device: int = 0) -> dict:
    """Binary64 polish of mixed-precision family results: the same instances
    (template's A, cones and parameters; rows of Cm, and Bm if given) solved
    by the FP64 family engine from the mixed x, y, s as warm start (original
    units, passed through binary32 as every family warm start is), without
    the warm-start judge, to eps_abs within max_iters iterations.  Returns
    fetch()'s dict of the polished results (iters = polish iterations)."""
    from . import DeviceArray
    W = np.concatenate([mixed["x"], mixed["y"], mixed["s"]], axis=1).astype(np.float32)
    dW = DeviceArray(W)
    P = Family(template.with_params(EPS_ABS=eps_abs, MAX_ITERS=max_iters), precision="fp64", device=device)
    P.load(Cm, Bm)
    P.run(dW, W.shape[1], judge=False)
    return P.fetch()
