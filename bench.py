"""Benchmark: SOCPs solved/sec at EPS_ABS=1e-3 on one MI355X.

  python bench.py [--steps K] [--warmup W] [--workload family|uneven|random|portfolio]

A step is one pass of the hot path over one batch of synthetic input, with
every input resident in device memory when the timed region starts (setup and
upload happen once, before warm-up).  The step is a synchronous library call
(the C API blocks until its HIP stream is idle), timed with the host clock;
the dominant kernel is also timed with HIP events on its own stream.

Workloads (seeded, see pdos_amd/problems.py):
  family     16,384 held-out instances of the parametric portfolio family
             (n_a=500, k=20; A and b shared; mu, gamma vary), batched
  uneven     4,096 portfolios with n_a in [50,1000], k in [5,50], batched
  random     one random feasible SOCP, n=100,000, m=300,000, nnz~5M (grid engine)
  portfolio  the n=101, m=212 portfolio, one instance

Printed: one JSON line (metric, value, ..., roofline, cpu_baseline, plain_gpu,
accuracy).  --gpus N > 1 exits: this build targets one GPU.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import pdos_amd  # noqa: E402
from pdos_amd import problems  # noqa: E402

METRIC = "SOCPs solved/sec at EPS_ABS=1e-3: plain GPU vs batched+learned warm start"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ----------------------------------------------------------------- workloads

def make_workload(name, count):
    if name == "family":
        fam = problems.PortfolioFamily(seed=4, na=500, k=20, EPS_ABS=1e-3, MAX_ITERS=2000)
        _, Cm = fam.sample(seed=1004, count=count or 16384)
        probs = [fam.instance(c) for c in Cm]
        return probs, {"workload": "portfolio family, held-out batch", "n_a": 500, "k": 20,
                       "n": fam.n, "m": fam.m, "nnz": int(fam.Ap[-1]), "count": len(probs),
                       "eps_abs": 1e-3}
    if name == "uneven":
        probs = problems.uneven_portfolios(seed=3, count=count or 4096, EPS_ABS=1e-3,
                                           MAX_ITERS=2000)
        return probs, {"workload": "uneven portfolio batch", "count": len(probs),
                       "n_a": [50, 1000], "k": [5, 50], "eps_abs": 1e-3,
                       "nnz_total": int(sum(p.nnz for p in probs))}
    if name == "random":
        p = problems.random_feasible(seed=1, EPS_ABS=1e-3, MAX_ITERS=2000)
        return [p], {"workload": "random feasible SOCP", "n": p.n, "m": p.m, "nnz": p.nnz,
                     "count": 1, "eps_abs": 1e-3}
    if name == "portfolio":
        p = problems.portfolio(seed=0, EPS_ABS=1e-3)
        return [p], {"workload": "portfolio n=101 m=212", "n": p.n, "m": p.m, "nnz": p.nnz,
                     "count": 1, "eps_abs": 1e-3}
    raise ValueError(name)


# ---------------------------------------------------------- FP64 verification

def verify_fp64(probs, outs):
    """Recomputes the stop test of every returned solution in binary64 on the
    host (pdos.c:95-107 in original units); returns (verified, solved)."""
    ok = solved = 0
    for p, o in zip(probs, outs):
        if o["status"] != "Solved":
            continue
        solved += 1
        A = sp.csc_matrix((p.Ax, p.Ai, p.Ap), shape=(p.m, p.n))
        eps = p.params["EPS_ABS"]
        pres = np.linalg.norm(A @ o["x"] + o["s"] - p.b)
        dres = np.linalg.norm(p.c + A.T @ o["y"])
        pobj, dobj = p.c @ o["x"], -p.b @ o["y"]
        slack = 1.001
        if (pres < slack * eps * (1 + np.linalg.norm(p.b)) and
                dres < slack * eps * (1 + np.linalg.norm(p.c)) and
                abs(pobj - dobj) < slack * eps * (1 + abs(pobj) + abs(dobj))):
            ok += 1
    return ok, solved


# ------------------------------------------------------------------ roofline

def algorithmic_bytes(probs, outs, plugin, sv):
    """SpMV bytes the launch executed (B_A / B_At per SpMV with s_v = value
    size, s_i = 4), plus the direct plugin's two triangular sweeps per
    iteration.  Counts: initial A x, then per update k+1 SpMVs with A and k+2
    with A' (indirect) or 1 and 1 (direct), plus the stop check's A' SpMV."""
    si = 4
    tot = 0.0
    for p, o in zip(probs, outs):
        it, cg = o["info"]["iters"], o["info"]["cg_iters"]
        solved = 1 if o["status"] == "Solved" else 0
        nnz, n, m = p.nnz, p.n, p.m
        BA = nnz * (sv + si) + (m + 1) * si + n * sv + m * sv
        BAt = nnz * (sv + si) + (n + 1) * si + m * sv + n * sv
        if plugin == "indirect":
            NA, NAt = 1 + it + cg, 2 * it + cg + solved
            tot += NA * BA + NAt * BAt
        else:
            NA, NAt = 1 + it, it + solved
            N, lnz = n + m, o["info"]["lnz"]
            BL = lnz * (8 + si) + (N + 1) * si + 2 * N * 8
            tot += NA * BA + NAt * BAt + it * 2 * BL
    return tot


# -------------------------------------------------------------- CPU baseline

def _cpu_worker(args):
    kind, plugin, blobs = args
    import ctypes as C
    sys.path.insert(0, ROOT)
    from pdos_amd.abi import Cone, Data, Problem, Sol
    if kind == "reference":
        L = C.CDLL(os.path.join(ROOT, "oracle", "_ref",
                                f"libpdos{'dir' if plugin == 'direct' else 'indir'}_ref.so"))
        L.pdos.restype = C.POINTER(Sol)
        L.pdos.argtypes = [C.POINTER(Data), C.POINTER(Cone)]
        L.freeSol.argtypes = [C.POINTER(C.POINTER(Sol))]
        solve = lambda d, k: L.freeSol(C.byref(L.pdos(C.byref(d), C.byref(k))))  # noqa: E731
    else:
        from oracle import oracle as O
        O.lib()
        solve = lambda d, k: O.lib().oracle_free_sol(  # noqa: E731
            O.lib().oracle_pdos(C.byref(d), C.byref(k), int(plugin == "direct"), None, None))
    probs = [Problem(**b) for b in blobs]
    cs = [p.to_c() for p in probs]
    t0 = time.perf_counter()
    for d, k, _ in cs:
        solve(d, k)
    return time.perf_counter() - t0


def cpu_baseline(probs, plugin, sample, procs):
    """Times the reference C library (oracle/_ref, built from the reference
    sources) -- or the oracle port if it is absent -- on `sample` instances,
    split over `procs` worker processes, one per core."""
    kind = "reference" if os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libpdosdir_ref.so")) else "port"
    sel = probs[:sample]
    blobs = [dict(n=p.n, m=p.m, Ap=p.Ap, Ai=p.Ai, Ax=p.Ax, b=p.b, c=p.c, f=p.f, l=p.l, q=p.q,
                  params=dict(p.params, VERBOSE=0)) for p in sel]
    shards = [blobs[i::procs] for i in range(procs)]
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        pool.map(_cpu_worker, [(kind, plugin, s) for s in shards if s])
    wall = time.perf_counter() - t0
    return {"value": len(sel) / wall, "unit": "SOCPs/s", "cores": procs, "kind": kind,
            "plugin": plugin,
            "sample": f"first {len(sel)} instances of the same batch, {procs} processes, "
                      f"{wall:.1f} s wall (factorization included)"}


# --------------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="family")
    ap.add_argument("--count", type=int, default=0)
    ap.add_argument("--plugin", default="direct")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--plain-sample", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    probs, config = make_workload(args.workload, args.count)
    plugin, precision = args.plugin, args.precision
    if args.workload == "random":
        plugin = "indirect"
    config.update(plugin=plugin, precision=precision)
    log(f"workload {args.workload}: {len(probs)} instances, plugin {plugin}, {precision}")

    sv = 8 if precision == "fp64" else 4
    if args.workload in ("random", "portfolio"):
        # single instance: the step is one full pdos_ex call on prepared data
        run = lambda: pdos_amd.solve_problem(probs[0], plugin=plugin, precision=precision)  # noqa
        for _ in range(args.warmup):
            out = run()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = run()
        wall = (time.perf_counter() - t0) / args.steps
        outs = [out]
        kernel_ms = out["info"]["solve_ms"]
    else:
        t0 = time.perf_counter()
        h = pdos_amd.BatchHandle(probs, plugin=plugin, precision=precision)
        log(f"prepare (host setup + factor + upload): {time.perf_counter() - t0:.2f} s")
        for _ in range(args.warmup):
            h.run()
        kms = []
        t0 = time.perf_counter()
        for i in range(args.steps):
            kms.append(h.run())
            log(f"step {i}: {kms[-1]:.2f} ms")
        wall = (time.perf_counter() - t0) / args.steps
        kernel_ms = float(np.mean(kms))
        outs = h.fetch()
        h.close()

    solved = sum(o["status"] == "Solved" for o in outs)
    verified, _ = verify_fp64(probs, outs)
    iters = np.array([o["info"]["iters"] for o in outs])
    value = solved / wall
    log(f"{solved}/{len(outs)} solved, {verified} verified in FP64, {value:.1f} SOCPs/s, "
        f"mean iters {iters.mean():.1f}")

    abytes = algorithmic_bytes(probs, outs, plugin, sv)
    achieved = abytes / (kernel_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_wg_solve" if len(outs) > 1 or args.workload == "portfolio" else "grid",
                "algorithmic_bytes_per_launch": abytes, "kernel_ms": kernel_ms}

    plain = None
    if len(probs) > 1 and args.plain_sample > 0:
        sel = probs[:args.plain_sample]
        pdos_amd.solve_problem(sel[0], plugin=plugin, precision=precision)
        t0 = time.perf_counter()
        ns = sum(pdos_amd.solve_problem(p, plugin=plugin, precision=precision)["status"] == "Solved"
                 for p in sel)
        dt = time.perf_counter() - t0
        plain = {"value": ns / dt, "unit": "SOCPs/s",
                 "sample": f"first {len(sel)} instances, one pdos_ex call each (setup, upload, "
                           f"solve, download)"}
        log(f"plain single-instance GPU path: {plain['value']:.1f} SOCPs/s")

    cpu = None
    if not args.no_cpu:
        procs = max(1, min(16, os.cpu_count() or 1))
        cpu = cpu_baseline(probs, "direct" if args.workload != "random" else "indirect",
                           min(args.cpu_sample, len(probs)), procs)
        log(f"cpu baseline: {cpu['value']:.1f} SOCPs/s on {cpu['cores']} cores")

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "SOCPs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32", "data": "synthetic",
        "config": config, "roofline": roofline, "cpu_baseline": cpu, "plain_gpu": plain,
        "accuracy": {"solved": int(solved), "count": len(outs), "verified_fp64": int(verified),
                     "mean_iters": float(iters.mean()), "max_iters": int(iters.max())},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
