"""ctypes mirror of the C ABI in include/pdos.h.

Struct layouts follow the reference built with -DDLONG (idxint = long):
Params pdos.h:15-19, Data pdos.h:22-35, Sol pdos.h:37-41, Cone cones.h:8-13.
`Problem` is the Python-side owner of one cone program; `Problem.to_c()`
builds the C structs over numpy buffers that the Problem keeps alive.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

c_long_p = C.POINTER(C.c_long)
c_double_p = C.POINTER(C.c_double)


class Params(C.Structure):
    _fields_ = [
        ("MAX_ITERS", C.c_long),
        ("CG_MAX_ITS", C.c_long),
        ("EPS_ABS", C.c_double),
        ("ALPHA", C.c_double),
        ("CG_TOL", C.c_double),
        ("VERBOSE", C.c_long),
        ("NORMALIZE", C.c_long),
    ]


class Data(C.Structure):
    _fields_ = [
        ("n", C.c_long),
        ("m", C.c_long),
        ("Ax", c_double_p),
        ("Ai", c_long_p),
        ("Ap", c_long_p),
        ("b", c_double_p),
        ("c", c_double_p),
        ("x", c_double_p),
        ("y", c_double_p),
        ("s", c_double_p),
        ("p", C.POINTER(Params)),
    ]


class Sol(C.Structure):
    _fields_ = [
        ("n", C.c_long),
        ("m", C.c_long),
        ("x", c_double_p),
        ("s", c_double_p),
        ("y", c_double_p),
        ("status", C.c_char * 16),
    ]


class Cone(C.Structure):
    _fields_ = [
        ("f", C.c_long),
        ("l", C.c_long),
        ("q", c_long_p),
        ("qsize", C.c_long),
    ]


assert C.sizeof(Params) == 56 and C.sizeof(Data) == 88
assert C.sizeof(Sol) == 56 and C.sizeof(Cone) == 32

# Defaults of the reference's Python binding (python/pdosmodule.c:63-74),
# with NORMALIZE = 0 as README.md:89 documents.
DEFAULT_PARAMS = dict(MAX_ITERS=2000, CG_MAX_ITS=20, EPS_ABS=5e-3, ALPHA=1.0,
                      CG_TOL=1e-4, VERBOSE=0, NORMALIZE=0)


def _dptr(a):
    return a.ctypes.data_as(c_double_p) if a is not None else None


def _lptr(a):
    return a.ctypes.data_as(c_long_p) if a is not None else None


@dataclass
class Problem:
    """min c'x s.t. Ax + s = b, s in K; A is m x n in CSC (Ap, Ai, Ax)."""
    n: int
    m: int
    Ap: np.ndarray
    Ai: np.ndarray
    Ax: np.ndarray
    b: np.ndarray
    c: np.ndarray
    f: int
    l: int
    q: list
    params: dict = field(default_factory=lambda: dict(DEFAULT_PARAMS))
    x0: np.ndarray | None = None
    y0: np.ndarray | None = None
    s0: np.ndarray | None = None
    name: str = ""

    def __post_init__(self):
        self.Ap = np.ascontiguousarray(self.Ap, dtype=np.int64)
        self.Ai = np.ascontiguousarray(self.Ai, dtype=np.int64)
        self.Ax = np.ascontiguousarray(self.Ax, dtype=np.float64)
        self.b = np.ascontiguousarray(self.b, dtype=np.float64)
        self.c = np.ascontiguousarray(self.c, dtype=np.float64)
        self.q = [int(v) for v in self.q]
        for k in ("x0", "y0", "s0"):
            v = getattr(self, k)
            if v is not None:
                setattr(self, k, np.ascontiguousarray(v, dtype=np.float64))

    @property
    def nnz(self) -> int:
        return int(self.Ap[-1])

    def with_params(self, **kw) -> "Problem":
        p = dict(self.params)
        p.update(kw)
        return Problem(self.n, self.m, self.Ap, self.Ai, self.Ax, self.b, self.c, self.f,
                       self.l, list(self.q), p, self.x0, self.y0, self.s0, self.name)

    def with_warm(self, x0=None, y0=None, s0=None) -> "Problem":
        return Problem(self.n, self.m, self.Ap, self.Ai, self.Ax, self.b, self.c, self.f,
                       self.l, list(self.q), dict(self.params), x0, y0, s0, self.name)

    def dense_A(self) -> np.ndarray:
        A = np.zeros((self.m, self.n))
        for j in range(self.n):
            for p in range(self.Ap[j], self.Ap[j + 1]):
                A[self.Ai[p], j] += self.Ax[p]
        return A

    def to_c(self):
        """Returns (Data, Cone, keepalive). Buffers are copies, so a callee
        that writes into them cannot disturb this Problem."""
        keep = {}
        par = Params(**{k: self.params.get(k, DEFAULT_PARAMS[k]) for k in DEFAULT_PARAMS})
        keep["par"] = par
        arrs = {}
        for k in ("Ax", "Ai", "Ap", "b", "c"):
            arrs[k] = np.array(getattr(self, k), copy=True)
        for k, kk in (("x0", "x"), ("y0", "y"), ("s0", "s")):
            v = getattr(self, k)
            arrs[kk] = None if v is None else np.array(v, copy=True)
        qarr = np.array(self.q if self.q else [0], dtype=np.int64)
        keep["arrs"] = arrs
        keep["q"] = qarr
        d = Data(self.n, self.m, _dptr(arrs["Ax"]), _lptr(arrs["Ai"]), _lptr(arrs["Ap"]),
                 _dptr(arrs["b"]), _dptr(arrs["c"]), _dptr(arrs["x"]), _dptr(arrs["y"]),
                 _dptr(arrs["s"]), C.pointer(par))
        k = Cone(self.f, self.l, _lptr(qarr), len(self.q))
        keep["d"] = d
        keep["k"] = k
        return d, k, keep


def sol_to_numpy(sol_p) -> dict:
    s = sol_p.contents
    n, m = s.n, s.m
    return dict(
        x=np.ctypeslib.as_array(s.x, shape=(n,)).copy(),
        s=np.ctypeslib.as_array(s.s, shape=(m,)).copy(),
        y=np.ctypeslib.as_array(s.y, shape=(m,)).copy(),
        status=s.status.decode(),
    )


# ---------------------------------------------------------------------------
# Binary problem-file format of the demo CLI (run_pdos.c:31-94), little-endian
# i64 / f64: n m f l qsize | q[qsize] | b[m] c[n] | MAX_ITERS CG_MAX_ITS |
# ALPHA EPS_ABS CG_TOL | VERBOSE NORMALIZE | Anz | Ai[Anz] Ap[n+1] Ax[Anz]
# ---------------------------------------------------------------------------

def read_problem_file(path: str) -> Problem:
    raw = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    n, m, f, l, qsize = (int(v) for v in take("<i8", 5))
    q = [int(v) for v in take("<i8", qsize)]
    b = take("<f8", m)
    c = take("<f8", n)
    max_iters, cg_max = (int(v) for v in take("<i8", 2))
    alpha, eps, cg_tol = (float(v) for v in take("<f8", 3))
    verbose, normalize = (int(v) for v in take("<i8", 2))
    (anz,) = (int(v) for v in take("<i8", 1))
    Ai = take("<i8", anz)
    Ap = take("<i8", n + 1)
    Ax = take("<f8", anz)
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    params = dict(MAX_ITERS=max_iters, CG_MAX_ITS=cg_max, EPS_ABS=eps, ALPHA=alpha,
                  CG_TOL=cg_tol, VERBOSE=verbose, NORMALIZE=normalize)
    return Problem(n, m, Ap, Ai, Ax, b, c, f, l, q, params, name=path.rsplit("/", 1)[-1])


def write_problem_file(path: str, p: Problem) -> None:
    par = dict(DEFAULT_PARAMS)
    par.update(p.params)
    with open(path, "wb") as fh:
        fh.write(np.array([p.n, p.m, p.f, p.l, len(p.q)], "<i8").tobytes())
        fh.write(np.array(p.q, "<i8").tobytes())
        fh.write(np.asarray(p.b, "<f8").tobytes())
        fh.write(np.asarray(p.c, "<f8").tobytes())
        fh.write(np.array([par["MAX_ITERS"], par["CG_MAX_ITS"]], "<i8").tobytes())
        fh.write(np.array([par["ALPHA"], par["EPS_ABS"], par["CG_TOL"]], "<f8").tobytes())
        fh.write(np.array([par["VERBOSE"], par["NORMALIZE"]], "<i8").tobytes())
        fh.write(np.array([p.nnz], "<i8").tobytes())
        fh.write(np.asarray(p.Ai, "<i8").tobytes())
        fh.write(np.asarray(p.Ap, "<i8").tobytes())
        fh.write(np.asarray(p.Ax, "<f8").tobytes())
