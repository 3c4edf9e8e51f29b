"""ctypes front-end to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and the
benchmark's CPU-baseline leg, as the checker -- never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from pdos_amd.abi import Cone, Data, Problem, Sol, sol_to_numpy

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OracleInfo(C.Structure):
    _fields_ = [
        ("iters", C.c_long), ("status", C.c_long), ("cg_iters", C.c_long), ("lnz", C.c_long),
        ("lam", C.c_double), ("p_res", C.c_double), ("d_res", C.c_double),
        ("p_obj", C.c_double), ("d_obj", C.c_double), ("eta", C.c_double),
        ("hist", C.POINTER(C.c_double)), ("hist_cap", C.c_long), ("hist_len", C.c_long),
        ("setup_s", C.c_double), ("solve_s", C.c_double),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_pdos.restype = C.POINTER(Sol)
        L.oracle_pdos.argtypes = [C.POINTER(Data), C.POINTER(Cone), C.c_int,
                                  C.POINTER(C.c_long), C.POINTER(OracleInfo)]
        L.oracle_free_sol.argtypes = [C.POINTER(Sol)]
        L.oracle_proj_cone.argtypes = [C.POINTER(C.c_double), C.POINTER(Cone)]
        L.oracle_setup.argtypes = [C.POINTER(Data), C.POINTER(Cone), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oracle_kkt_order.argtypes = [C.POINTER(Data), C.POINTER(Cone), C.POINTER(C.c_long)]
        L.oracle_set_equil.argtypes = [C.c_int]
        L.oracle_pdos_many.restype = C.c_long
        L.oracle_pdos_many.argtypes = [C.POINTER(C.POINTER(Data)), C.POINTER(C.POINTER(Cone)),
                                       C.c_long, C.c_int, C.POINTER(C.c_long)]
        _lib = L
    return _lib


EQUIL = {"reference": 0, "ruiz": 1}


def solve(p: Problem, plugin: str = "indirect", perm=None, hist: int = 0, equil: str = "reference") -> dict:
    """Runs the oracle; returns x, s, y, status, iters, cg_iters, residuals.
    equil "ruiz": the reference's alternative scaling (ruiz_equil_common.h)."""
    lib().oracle_set_equil(EQUIL[equil])
    try:
        return _solve(p, plugin, perm, hist)
    finally:
        lib().oracle_set_equil(0)


def _solve(p, plugin, perm, hist):
    d, k, keep = p.to_c()
    info = OracleInfo()
    hbuf = None
    if hist:
        hbuf = np.zeros(4 * hist)
        info.hist = hbuf.ctypes.data_as(C.POINTER(C.c_double))
        info.hist_cap = hist
    pp = None
    if perm is not None:
        pp = np.ascontiguousarray(perm, dtype=np.int64)
    res = lib().oracle_pdos(C.byref(d), C.byref(k), 1 if plugin == "direct" else 0,
                            pp.ctypes.data_as(C.POINTER(C.c_long)) if pp is not None else None,
                            C.byref(info))
    if not res:
        raise RuntimeError("oracle_pdos returned NULL")
    out = sol_to_numpy(res)
    lib().oracle_free_sol(res)
    out.update(iters=info.iters, cg_iters=info.cg_iters, lnz=info.lnz, lam=info.lam,
               p_res=info.p_res, d_res=info.d_res, p_obj=info.p_obj, d_obj=info.d_obj,
               eta=info.eta, setup_s=info.setup_s, solve_s=info.solve_s)
    if hist:
        out["hist"] = hbuf[:4 * info.hist_len].reshape(-1, 4).copy()
    return out


def proj_cone(z: np.ndarray, f: int, l: int, q) -> np.ndarray:
    x = np.array(z, dtype=np.float64, copy=True)
    qa = np.array(q if len(q) else [0], dtype=np.int64)
    k = Cone(f, l, qa.ctypes.data_as(C.POINTER(C.c_long)), len(q))
    lib().oracle_proj_cone(x.ctypes.data_as(C.POINTER(C.c_double)), C.byref(k))
    return x


def setup(p: Problem, equil: str = "reference"):
    d, k, keep = p.to_c()
    D = np.zeros(p.m)
    E = np.zeros(p.n)
    lam = C.c_double()
    lib().oracle_set_equil(EQUIL[equil])
    try:
        lib().oracle_setup(C.byref(d), C.byref(k), D.ctypes.data_as(C.POINTER(C.c_double)),
                           E.ctypes.data_as(C.POINTER(C.c_double)), C.byref(lam))
    finally:
        lib().oracle_set_equil(0)
    return D, E, lam.value


def solve_many(problems, plugin: str = "direct"):
    """Solves problems back to back on this thread. Returns (solved, iters)."""
    keeps = [p.to_c() for p in problems]
    n = len(problems)
    DA = (C.POINTER(Data) * n)(*[C.pointer(kk[0]) for kk in keeps])
    KA = (C.POINTER(Cone) * n)(*[C.pointer(kk[1]) for kk in keeps])
    iters = np.zeros(n, dtype=np.int64)
    solved = lib().oracle_pdos_many(DA, KA, n, 1 if plugin == "direct" else 0,
                                    iters.ctypes.data_as(C.POINTER(C.c_long)))
    return int(solved), iters
