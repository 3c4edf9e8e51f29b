"""Writes the golden parity fixtures from the compiled reference C library.

Runs only in the container that holds the reference tree: it builds
oracle/_ref/ (make -C oracle ref) from the reference sources in place, calls
the reference's exported `pdos`, `initWork` and `projCone` through ctypes, and
stores inputs and outputs as data under tests/golden/:

  inputs/*.bin        problem files in the demo format (run_pdos.c:31-94)
  data/portfolio_*    the reference's own fixtures (copied data files)
  golden.npz          outputs (x, s, y; D, E, lambda; AMD permutation; projCone)
  golden.json         manifest: status, iteration count, last residual row

Iteration counts come from the reference's own VERBOSE summary line
(pdos.c:108,126 -> printSummary(i, ...)), captured from fd 1.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from pdos_amd.abi import Cone, Data, Params, Problem, Sol, read_problem_file, sol_to_numpy, write_problem_file  # noqa: E402
from pdos_amd import problems  # noqa: E402

REF = "/root/reference"
REF_LIBS = os.path.join(ROOT, "oracle", "_ref")


class Work(C.Structure):  # pdos.h:46-74 (reference internal workspace)
    _fields_ = [("n", C.c_long), ("m", C.c_long), ("Ax", C.POINTER(C.c_double)),
                ("Ai", C.POINTER(C.c_long)), ("Ap", C.POINTER(C.c_long)),
                ("Atx", C.POINTER(C.c_double)), ("Ati", C.POINTER(C.c_long)),
                ("Atp", C.POINTER(C.c_long)), ("b", C.POINTER(C.c_double)),
                ("c", C.POINTER(C.c_double)), ("x", C.POINTER(C.c_double)),
                ("s", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
                ("stilde", C.POINTER(C.c_double)), ("D", C.POINTER(C.c_double)),
                ("E", C.POINTER(C.c_double)), ("lam", C.c_double),
                ("params", C.POINTER(Params)), ("p", C.c_void_p)]


class PrivDirect(C.Structure):  # direct/private.h:8-12
    _fields_ = [("L", C.c_void_p), ("D", C.POINTER(C.c_double)), ("P", C.POINTER(C.c_long))]


def load_ref(plugin, variant=""):
    """The compiled reference; variant "_ruiz" is the build with
    ruiz_equil_common.h in place of common.h (oracle/Makefile)."""
    L = C.CDLL(os.path.join(REF_LIBS, f"libpdos{'dir' if plugin == 'direct' else 'indir'}{variant}_ref.so"))
    L.pdos.restype = C.POINTER(Sol)
    L.pdos.argtypes = [C.POINTER(Data), C.POINTER(Cone)]
    L.initWork.restype = C.POINTER(Work)
    L.initWork.argtypes = [C.POINTER(Data), C.POINTER(Cone)]
    L.projCone.argtypes = [C.POINTER(C.c_double), C.POINTER(Cone)]
    return L


libc = C.CDLL(None)
SUMMARY = re.compile(r"^\s*(\d+) \|\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s*$")


def run_ref(lib, prob):
    """Reference pdos() with VERBOSE=1 captured; returns outputs + iterations."""
    p = prob.with_params(VERBOSE=1)
    d, k, keep = p.to_c()
    with tempfile.TemporaryFile() as tmp:
        libc.fflush(None)
        saved = os.dup(1)
        os.dup2(tmp.fileno(), 1)
        try:
            sol = lib.pdos(C.byref(d), C.byref(k))
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        text = tmp.read().decode()
    out = sol_to_numpy(sol)
    rows = [SUMMARY.match(line) for line in text.splitlines()]
    rows = [r for r in rows if r]
    last = rows[-1]
    out["iters"] = int(last.group(1))
    out["last_row"] = [float(last.group(i)) for i in range(2, 7)]
    return out


def ref_setup(lib, prob, direct):
    d, k, keep = prob.to_c()
    w = lib.initWork(C.byref(d), C.byref(k)).contents
    D = np.ctypeslib.as_array(w.D, shape=(prob.m,)).copy()
    E = np.ctypeslib.as_array(w.E, shape=(prob.n,)).copy()
    P = None
    if direct:
        pr = C.cast(w.p, C.POINTER(PrivDirect)).contents
        P = np.ctypeslib.as_array(pr.P, shape=(prob.n + prob.m,)).copy()
    return D, E, w.lam, P


def golden_cases(arrays):
    """The parity cases as (tag, input path, problem): the reference's data
    files (copied under data/) and generated problems (written to inputs/ in
    the demo format; warm starts, which the format lacks, go into `arrays`)."""
    os.makedirs(os.path.join(HERE, "inputs"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "data"), exist_ok=True)
    cases = []
    for i in range(1, 11):
        name = f"portfolio_test_{i:02d}_10x100"
        src = os.path.join(REF, "data", name)
        dst = os.path.join(HERE, "data", name)
        shutil.copyfile(src, dst)
        prob = read_problem_file(dst)
        prob.params["VERBOSE"] = 0
        cases.append((f"fixture{i:02d}", f"data/{name}", prob))

    def add_generated(tag, prob):
        path = f"inputs/{tag}.bin"
        write_problem_file(os.path.join(HERE, path), prob)
        back = read_problem_file(os.path.join(HERE, path))
        if prob.x0 is not None:  # the demo format has no warm start: keep it as arrays
            arrays[f"{tag}__x0"], arrays[f"{tag}__y0"], arrays[f"{tag}__s0"] = prob.x0, prob.y0, prob.s0
            back = back.with_warm(prob.x0, prob.y0, prob.s0)
        cases.append((tag, path, back))

    base = problems.portfolio(seed=0)
    add_generated("portfolio_n101_norm0", base.with_params(NORMALIZE=0))
    add_generated("portfolio_n101_norm1", base.with_params(NORMALIZE=1))
    add_generated("portfolio_n101_alpha18", base.with_params(ALPHA=1.8, NORMALIZE=1))
    for its in (1, 2, 10):
        add_generated(f"portfolio_n101_steps{its}", base.with_params(EPS_ABS=0.0, MAX_ITERS=its))
    add_generated("portfolio_n101_maxit0", base.with_params(MAX_ITERS=0))
    add_generated("portfolio_n101_cg1", base.with_params(CG_MAX_ITS=1))
    add_generated("portfolio_n101_warm", base.with_warm(
        x0=np.full(base.n, 0.01), y0=np.full(base.m, -0.1), s0=np.full(base.m, 0.2)
    ).with_params(NORMALIZE=1))
    small = problems.random_feasible(seed=11, n=300, m=900, nnz=6000, f=10, l=440, nsoc=30,
                                     soc_lo=3, soc_hi=25, EPS_ABS=1e-3, MAX_ITERS=3000)
    add_generated("random_n300_m900", small)
    add_generated("random_n300_m900_norm1", small.with_params(NORMALIZE=1))
    for j, u in enumerate(problems.uneven_portfolios(seed=3, count=3, na_range=(50, 200),
                                                     k_range=(5, 20), EPS_ABS=1e-3)):
        add_generated(f"uneven_{j}", u)
    # structural edge cases of the input format: CSC columns in random row
    # order with duplicated entries, an empty column and empty rows (one of
    # them a zero-cone row); SOCs only; no SOC; b = 0 (lambda = 1e-15 / (1e-15 + |c|))
    edge = problems.random_feasible(seed=21, n=120, m=300, nnz=1500, f=5, l=150, nsoc=10, soc_lo=3,
                                    soc_hi=30, empty_rows=2, empty_cols=1, EPS_ABS=1e-3, MAX_ITERS=3000)
    add_generated("jumbled_n120_m300", problems.jumbled(edge, seed=1))
    add_generated("jumbled_n120_m300_norm1", problems.jumbled(edge, seed=1).with_params(NORMALIZE=1))
    add_generated("soc_only_n80_m200", problems.random_feasible(
        seed=22, n=80, m=200, nnz=1200, f=0, l=0, nsoc=8, soc_lo=5, soc_hi=40, EPS_ABS=1e-3, MAX_ITERS=3000))
    add_generated("lp_only_n80_m200", problems.random_feasible(
        seed=23, n=80, m=200, nnz=1200, f=10, l=190, nsoc=0, EPS_ABS=1e-3, MAX_ITERS=3000))
    bz = problems.random_feasible(seed=24, n=60, m=150, nnz=800, f=0, l=100, nsoc=5, soc_lo=3, soc_hi=20,
                                  EPS_ABS=1e-3, MAX_ITERS=400)
    add_generated("b_zero_n60_m150", Problem(bz.n, bz.m, bz.Ap, bz.Ai, bz.Ax, np.zeros(bz.m), bz.c, bz.f,
                                             bz.l, bz.q, bz.params, name="b_zero"))
    dense = problems.DenseFamily(seed=2, m=96, n=24, l=64, nsoc=4, qdim=8, EPS_ABS=1e-3)
    B, Cm, _ = dense.sample(5, 2)
    add_generated("dense_m96_n24", dense.instance(B[0], Cm[0]))
    fam = problems.PortfolioFamily(seed=4, na=60, k=6, EPS_ABS=1e-3)
    _, Cf = fam.sample(6, 2)
    add_generated("family_na60", fam.instance(Cf[0]))
    return cases


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    libs = {pl: load_ref(pl) for pl in ("indirect", "direct")}
    arrays = {}
    manifest = {"cases": {}, "projcone": {}}
    cases = golden_cases(arrays)

    for tag, path, prob in cases:
        entry = {"input": path, "n": prob.n, "m": prob.m, "nnz": prob.nnz}
        for pl in ("indirect", "direct"):
            out = run_ref(libs[pl], prob)
            for key in ("x", "s", "y"):
                arrays[f"{tag}__{pl}__{key}"] = out[key]
            entry[pl] = {"status": out["status"], "iters": out["iters"], "last_row": out["last_row"]}
        D, E, lam, P = ref_setup(libs["direct"], prob, True)
        arrays[f"{tag}__D"] = D
        arrays[f"{tag}__E"] = E
        arrays[f"{tag}__lambda"] = np.array([lam])
        arrays[f"{tag}__amd_perm"] = P
        manifest["cases"][tag] = entry
        print(tag, {pl: entry[pl]["iters"] for pl in ("indirect", "direct")},
              entry["indirect"]["status"], flush=True)

    # projCone on seeded vectors plus edge cases (cones.c:5-57)
    rng = np.random.default_rng(123)
    pc_cases = {
        "mixed": (3, 5, [4, 1, 2, 7], rng.standard_normal(3 + 5 + 14) * 2),
        "edge_eq": (0, 0, [3, 3, 3], np.array([5.0, 3.0, 4.0, -5.0, 3.0, 4.0, 0.0, -0.0, 0.0])),
        "edge_neg0": (1, 3, [2], np.array([-0.0, -0.0, 0.0, -1e-300, 1.0, -1.0])),
        "dim1_soc": (0, 0, [1, 1], np.array([-2.0, 3.0])),
        "big": (10, 100, [50, 200, 2], rng.standard_normal(362)),
    }
    for tag, (f, l, q, z) in pc_cases.items():
        qa = np.array(q, dtype=np.int64)
        k = Cone(f, l, qa.ctypes.data_as(C.POINTER(C.c_long)), len(q))
        x = np.array(z, dtype=np.float64, copy=True)
        libs["direct"].projCone(x.ctypes.data_as(C.POINTER(C.c_double)), C.byref(k))
        arrays[f"projcone__{tag}__in"] = np.array(z, dtype=np.float64)
        arrays[f"projcone__{tag}__out"] = x
        manifest["projcone"][tag] = {"f": f, "l": l, "q": q}

    np.savez_compressed(os.path.join(HERE, "golden.npz"), **arrays)
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)
    print("wrote", len(arrays), "arrays")


if __name__ == "__main__":
    main()
