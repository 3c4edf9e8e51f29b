"""MI355X PDOS: Python front-end over the C ABI of libpdos{indir,dir}_hip.so.

`solve(c, A, b, dims, opts, x0, y0, s0)` keeps the contract of the reference's
Python binding (python/pdosmodule.c:27-344): A is a sparse m x n matrix (any
scipy sparse format; converted to CSC), dims = {'f': int, 'l': int,
'q': [ints]}, opts may set MAX_ITERS, CG_MAX_ITS, EPS_ABS, ALPHA, CG_TOL,
VERBOSE, NORMALIZE (defaults pdosmodule.c:63-74, NORMALIZE 0), and the result
is a dict with 'x', 'y', 's' and 'status'.  Extra keys: 'info'.

The library is the HIP build; there is no CPU fallback.  If the shared
library is missing, `lib()` raises.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np

from .abi import Cone, Data, Params, Problem, Sol, sol_to_numpy, DEFAULT_PARAMS

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {"indirect": os.path.join(HERE, "libpdosindir_hip.so"),
        "direct": os.path.join(HERE, "libpdosdir_hip.so")}

PLUGIN = {"indirect": 0, "direct": 1}
PRECISION = {"fp64": 0, "mixed": 1}
ENGINE = {"auto": 0, "wg": 1, "grid": 2}


class PdosOptions(C.Structure):
    _fields_ = [("plugin", C.c_int), ("precision", C.c_int), ("engine", C.c_int),
                ("device", C.c_int), ("block", C.c_int), ("bf16_products", C.c_int),
                ("equil", C.c_int), ("tune", C.c_int)]


class PdosInfo(C.Structure):
    _fields_ = [("iters", C.c_long), ("cg_iters", C.c_long), ("status", C.c_long),
                ("engine", C.c_long), ("fallback", C.c_long), ("p_res", C.c_double), ("d_res", C.c_double),
                ("p_obj", C.c_double), ("d_obj", C.c_double), ("eta", C.c_double),
                ("lam", C.c_double), ("setup_ms", C.c_double), ("solve_ms", C.c_double),
                ("total_ms", C.c_double), ("lnz", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every symbol include/pdos.h and include/pdos_ext.h declare.
EXPORTS = ["pdos", "freeData", "freeSol", "tic", "toc", "tocq", "pdos_default_options",
           "pdos_ex", "pdos_last_info", "pdos_batch", "pdos_batch_prepare", "pdos_batch_run",
           "pdos_batch_fetch", "pdos_batch_free", "pdos_batch_device_bytes", "pdos_batch_tuning", "pdos_batch_prepare_times",
           "pdos_batch_set_warm_dev", "pdos_device_alloc",
           "pdos_device_free", "pdos_device_upload", "pdos_device_download", "pdos_device_copy",
           "pdos_mlp_create",
           "pdos_mlp_free", "pdos_mlp_num_params", "pdos_mlp_get_params", "pdos_mlp_set_params",
           "pdos_mlp_set_affine", "pdos_mlp_loss_grad", "pdos_mlp_train", "pdos_mlp_predict",
           "pdos_mlp_predict_dev", "pdos_mlp_predict_dev_noalloc", "pdos_mlp_last_error", "pdos_family_create", "pdos_family_load",
           "pdos_family_run", "pdos_family_fetch", "pdos_family_free", "pdos_family_last_error",
           "pdos_family_timing", "pdos_family_info", "pdos_family_engines", "pdos_family_scales",
           "pdos_batch_scales",
           "pdos_kkt_order", "pdos_kkt_solve_host", "pdos_setup_host", "pdos_last_error",
           "pdos_version"]

_libs = {}


def lib(plugin: str = "indirect"):
    """Loads the HIP library (raises if it was not built)."""
    if plugin not in _libs:
        path = LIBS[plugin]
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `make -C pdos_amd/csrc`")
        L = C.CDLL(path)
        L.pdos.restype = C.POINTER(Sol)
        L.pdos.argtypes = [C.POINTER(Data), C.POINTER(Cone)]
        L.pdos_ex.restype = C.POINTER(Sol)
        L.pdos_ex.argtypes = [C.POINTER(Data), C.POINTER(Cone), C.POINTER(PdosOptions),
                              C.POINTER(PdosInfo)]
        L.freeSol.argtypes = [C.POINTER(C.POINTER(Sol))]
        L.freeData.argtypes = [C.POINTER(C.POINTER(Data)), C.POINTER(C.POINTER(Cone))]
        L.pdos_last_error.restype = C.c_char_p
        L.pdos_version.restype = C.c_char_p
        L.pdos_last_info.argtypes = [C.POINTER(PdosInfo)]
        L.pdos_default_options.argtypes = [C.POINTER(PdosOptions)]
        L.pdos_batch_prepare.restype = C.c_void_p
        L.pdos_batch_prepare.argtypes = [C.POINTER(C.POINTER(Data)), C.POINTER(C.POINTER(Cone)),
                                         C.c_int, C.POINTER(PdosOptions)]
        L.pdos_batch_run.restype = C.c_int
        L.pdos_batch_run.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.pdos_batch_fetch.restype = C.c_int
        L.pdos_batch_fetch.argtypes = [C.c_void_p, C.POINTER(C.POINTER(Sol)), C.POINTER(PdosInfo)]
        L.pdos_batch_free.argtypes = [C.c_void_p]
        L.pdos_batch_device_bytes.restype = C.c_size_t
        L.pdos_batch_device_bytes.argtypes = [C.c_void_p]
        L.pdos_batch_tuning.restype = C.c_char_p
        L.pdos_batch_tuning.argtypes = [C.c_void_p]
        L.pdos_batch_scales.restype = C.c_int
        L.pdos_batch_scales.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
        L.pdos_batch_prepare_times.restype = C.c_int
        L.pdos_batch_prepare_times.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.pdos_setup_host.restype = C.c_int
        L.pdos_setup_host.argtypes = [C.POINTER(Data), C.POINTER(Cone), C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.pdos_kkt_order.restype = C.c_int
        L.pdos_kkt_order.argtypes = [C.POINTER(Data), C.POINTER(Cone), C.POINTER(C.c_long)]
        L.pdos_kkt_solve_host.restype = C.c_int
        L.pdos_kkt_solve_host.argtypes = [C.POINTER(Data), C.POINTER(Cone), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.POINTER(C.c_long)]
        L.pdos_batch_set_warm_dev.restype = C.c_int
        L.pdos_batch_set_warm_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_int]
        L.pdos_device_alloc.restype = C.c_void_p
        L.pdos_device_alloc.argtypes = [C.c_size_t]
        L.pdos_device_free.argtypes = [C.c_void_p]
        L.pdos_device_upload.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.pdos_device_download.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.pdos_device_copy.restype = C.c_int
        L.pdos_device_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        fp = C.POINTER(C.c_float)
        L.pdos_mlp_create.restype = C.c_void_p
        L.pdos_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_ulong, C.c_int]
        L.pdos_mlp_free.argtypes = [C.c_void_p]
        L.pdos_mlp_num_params.argtypes = [C.c_void_p]
        L.pdos_mlp_get_params.argtypes = [C.c_void_p, fp]
        L.pdos_mlp_set_params.argtypes = [C.c_void_p, fp]
        L.pdos_mlp_set_affine.argtypes = [C.c_void_p, fp, fp, fp, fp]
        L.pdos_mlp_loss_grad.argtypes = [C.c_void_p, fp, fp, C.c_int, C.POINTER(C.c_double), fp]
        L.pdos_mlp_train.argtypes = [C.c_void_p, fp, fp, C.c_int, C.c_int, C.c_int, C.c_float,
                                     C.c_ulong, C.POINTER(C.c_double)]
        L.pdos_mlp_predict.argtypes = [C.c_void_p, fp, C.c_int, fp, C.POINTER(C.c_double)]
        L.pdos_mlp_predict_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        L.pdos_mlp_predict_dev_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        L.pdos_mlp_last_error.restype = C.c_char_p
        L.pdos_batch.restype = C.c_int
        L.pdos_batch.argtypes = [C.POINTER(C.POINTER(Data)), C.POINTER(C.POINTER(Cone)), C.c_int,
                                 C.POINTER(PdosOptions), C.POINTER(C.POINTER(Sol)),
                                 C.POINTER(PdosInfo)]
        _libs[plugin] = L
    return _libs[plugin]


EQUIL = {"reference": 0, "ruiz": 1}


def options(plugin="indirect", precision="fp64", engine="auto", device=0, bf16_products=0,
            equil="reference") -> PdosOptions:
    """equil "ruiz": the scaling of the reference's alternative header
    ruiz_equil_common.h instead of common.h's (not for families)."""
    o = PdosOptions()
    o.bf16_products = bf16_products
    o.equil = EQUIL[equil]
    o.plugin = PLUGIN[plugin]
    o.precision = PRECISION[precision]
    o.engine = ENGINE[engine]
    o.device = device
    return o


def _take_sol(L, sp):
    out = sol_to_numpy(sp)
    L.freeSol(C.byref(sp))
    return out


def solve_problem(p: Problem, plugin="indirect", precision="fp64", engine="auto", equil="reference") -> dict:
    """Solves one Problem with the HIP library; returns x, s, y, status, info."""
    L = lib(plugin)
    d, k, keep = p.to_c()
    o = options(plugin, precision, engine, equil=equil)
    info = PdosInfo()
    sp = L.pdos_ex(C.byref(d), C.byref(k), C.byref(o), C.byref(info))
    if not sp:
        raise RuntimeError(f"pdos_ex failed: {L.pdos_last_error().decode()}")
    out = _take_sol(L, sp)
    out["info"] = info.as_dict()
    return out


def solve_batch(problems, plugin="indirect", precision="fp64", equil="reference", tune=False) -> list:
    """Solves independent problems in one launch (one workgroup each); tune:
    the launch-shape tuner (PdosOptions.tune)."""
    L = lib(plugin)
    n = len(problems)
    keeps = [p.to_c() for p in problems]
    DA = (C.POINTER(Data) * n)(*[C.pointer(kk[0]) for kk in keeps])
    KA = (C.POINTER(Cone) * n)(*[C.pointer(kk[1]) for kk in keeps])
    sols = (C.POINTER(Sol) * n)()
    infos = (PdosInfo * n)()
    o = options(plugin, precision, "wg", equil=equil)
    o.tune = int(tune)
    rc = L.pdos_batch(DA, KA, n, C.byref(o), sols, infos)
    if rc != 0:
        raise RuntimeError(f"pdos_batch failed: {L.pdos_last_error().decode()}")
    out = []
    for i in range(n):
        r = _take_sol(L, sols[i])
        r["info"] = infos[i].as_dict()
        out.append(r)
    return out


def kkt_solve_host(p: Problem, rhs):
    """Solves the scaled KKT system with the library's host factorization;
    returns (z, stats = [N, nnz(L), forward levels, backward levels])."""
    L = lib("direct")
    d, k, keep = p.to_c()
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    z = np.zeros(p.n + p.m)
    st = np.zeros(4, dtype=np.int64)
    rc = L.pdos_kkt_solve_host(C.byref(d), C.byref(k), rhs.ctypes.data_as(C.POINTER(C.c_double)),
                               z.ctypes.data_as(C.POINTER(C.c_double)),
                               st.ctypes.data_as(C.POINTER(C.c_long)))
    if rc != 0:
        raise RuntimeError(L.pdos_last_error().decode())
    return z, st


def kkt_order(p: Problem) -> np.ndarray:
    """The direct plugin's ordering of p's scaled KKT matrix (pivot order)."""
    L = lib("direct")
    d, k, keep = p.to_c()
    perm = np.zeros(p.n + p.m, dtype=np.int64)
    if L.pdos_kkt_order(C.byref(d), C.byref(k), perm.ctypes.data_as(C.POINTER(C.c_long))) != 0:
        raise RuntimeError(L.pdos_last_error().decode())
    return perm


def setup_host(p: Problem, equil="reference", plugin="indirect"):
    """The library's scaling of p on the host: (D, E with the x1e3, lambda)."""
    L = lib(plugin)
    d, k, keep = p.to_c()
    D, E, lam = np.zeros(p.m), np.zeros(p.n), C.c_double()
    rc = L.pdos_setup_host(C.byref(d), C.byref(k), EQUIL[equil], D.ctypes.data_as(C.POINTER(C.c_double)),
                           E.ctypes.data_as(C.POINTER(C.c_double)), C.byref(lam))
    if rc != 0:
        raise RuntimeError(L.pdos_last_error().decode())
    return D, E, lam.value


class BatchHandle:
    """A prepared batch resident on the device (pdos_batch_prepare/run/fetch)."""

    def __init__(self, problems, plugin="indirect", precision="fp64", tune=False):
        t0 = time.perf_counter()  # the Python-side conversion is timed apart from the library's prepare
        self.L = lib(plugin)
        self.n = len(problems)
        self.dims = [(p.n, p.m) for p in problems]
        keeps = [p.to_c() for p in problems]
        DA = (C.POINTER(Data) * self.n)(*[C.pointer(kk[0]) for kk in keeps])
        KA = (C.POINTER(Cone) * self.n)(*[C.pointer(kk[1]) for kk in keeps])
        o = options(plugin, precision, "wg")
        o.tune = int(tune)
        self.convert_s = time.perf_counter() - t0
        self.h = self.L.pdos_batch_prepare(DA, KA, self.n, C.byref(o))
        if not self.h:
            raise RuntimeError(f"pdos_batch_prepare failed: {self.L.pdos_last_error().decode()}")

    def run(self) -> float:
        """Solves on the device; returns the launch time in ms (HIP events)."""
        ms = C.c_double()
        if self.L.pdos_batch_run(self.h, C.byref(ms)) != 0:
            raise RuntimeError(f"pdos_batch_run failed: {self.L.pdos_last_error().decode()}")
        return ms.value

    def fetch(self) -> list:
        sols = (C.POINTER(Sol) * self.n)()
        infos = (PdosInfo * self.n)()
        if self.L.pdos_batch_fetch(self.h, sols, infos) != 0:
            raise RuntimeError(f"pdos_batch_fetch failed: {self.L.pdos_last_error().decode()}")
        out = []
        for i in range(self.n):
            r = _take_sol(self.L, sols[i])
            r["info"] = infos[i].as_dict()
            out.append(r)
        return out

    def set_warm_dev(self, dev_ptr, ld: int, check: bool = True):
        """Warm start from device memory (see pdos_batch_set_warm_dev)."""
        if self.L.pdos_batch_set_warm_dev(self.h, dev_ptr, ld, int(check)) != 0:
            raise RuntimeError(self.L.pdos_last_error().decode())

    def device_bytes(self) -> int:
        return int(self.L.pdos_batch_device_bytes(self.h))

    def prepare_times(self) -> dict:
        """Host ms of the library's prepare: setup, grouping and factorization,
        tuning, packing and upload."""
        v = (C.c_double * 4)()
        self.L.pdos_batch_prepare_times(self.h, v)
        return dict(zip(("setup", "group_factor", "tune", "pack_upload"), (round(x, 2) for x in v)))

    def scales(self, i: int):
        """D, E (with the x1e3), lambda of instance i and whether the device
        formed them (pdos_batch_scales)."""
        n, m = self.dims[i]
        D = np.zeros(m)
        E = np.zeros(n)
        lam = C.c_double()
        r = self.L.pdos_batch_scales(self.h, i, D.ctypes.data_as(C.POINTER(C.c_double)),
                                     E.ctypes.data_as(C.POINTER(C.c_double)), C.byref(lam))
        if r < 0:
            raise RuntimeError("pdos_batch_scales: bad arguments")
        return D, E, lam.value, bool(r)

    def tuning(self) -> list:
        """The tuner's decisions per size bucket ([] without tuning)."""
        t = self.L.pdos_batch_tuning(self.h).decode()
        return json.loads(t) if t else []

    def close(self):
        if self.h:
            self.L.pdos_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A float32 buffer in device memory (plain hipMalloc via the library)."""

    def __init__(self, arr=None, shape=None):
        self.L = lib("indirect")
        if arr is not None:
            arr = np.ascontiguousarray(arr, dtype=np.float32)
            shape = arr.shape
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        self.ptr = self.L.pdos_device_alloc(self.nbytes)
        if not self.ptr:
            raise RuntimeError("device allocation failed")
        if arr is not None:
            self.L.pdos_device_upload(self.ptr, arr.ctypes.data, self.nbytes)

    def upload(self, arr) -> None:
        """Copies a host array of this array's size into it."""
        arr = np.ascontiguousarray(arr, dtype=np.float32)
        if arr.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {arr.nbytes} bytes into {self.nbytes}")
        if self.L.pdos_device_upload(self.ptr, arr.ctypes.data, self.nbytes) != 0:
            raise RuntimeError("device upload failed")

    def numpy(self):
        out = np.empty(self.shape, dtype=np.float32)
        self.L.pdos_device_download(out.ctypes.data, self.ptr, self.nbytes)
        return out

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy by the library's copy kernel."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        if self.L.pdos_device_copy(self.ptr, other.ptr, self.nbytes) != 0:
            raise RuntimeError("device copy failed")

    def free(self):
        if self.ptr:
            self.L.pdos_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def binding_problem(c, A, b, dims, opts=None, x0=None, y0=None, s0=None) -> Problem:
    """The Problem the reference binding builds from its arguments
    (python/pdosmodule.c:63-311): defaults MAX_ITERS 2000, EPS_ABS 5e-3,
    ALPHA 1.0, VERBOSE 1, CG_MAX_ITS 20, CG_TOL 1e-4 (:67-74), NORMALIZE 0
    (README.md:89; the binding leaves it unset); an unknown option raises."""
    import scipy.sparse as sp
    A = sp.csc_matrix(A)
    m, n = A.shape
    params = dict(DEFAULT_PARAMS)
    params["VERBOSE"] = 1
    if opts:
        for key, val in opts.items():
            if key not in params:
                raise KeyError(f"unknown option {key}")
            params[key] = val
    q = list(dims.get("q", []))
    p = Problem(n, m, A.indptr, A.indices, A.data, np.asarray(b, float).ravel(),
                np.asarray(c, float).ravel(), int(dims.get("f", 0)), int(dims.get("l", 0)), q,
                params, x0, y0, s0)
    return p


def solve(c, A, b, dims, opts=None, x0=None, y0=None, s0=None, plugin="indirect",
          precision="fp64", equil="reference"):
    """The reference binding's solve() (python/pdosmodule.c:27-344): returns
    a dict with 'x', 's', 'y', 'status' (and 'info'); equil "ruiz" selects
    the scaling of the reference's ruiz_equil_common.h."""
    p = binding_problem(c, A, b, dims, opts, x0, y0, s0)
    return solve_problem(p, plugin, precision, equil=equil)
