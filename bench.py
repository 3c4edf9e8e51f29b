"""Benchmark: SOCPs solved/sec at EPS_ABS=1e-3 on one MI355X.

  python bench.py [--steps K] [--warmup W] [--workload family|dense|uneven|random|portfolio]

A step is one pass of the hot path over one batch of synthetic input with
every input resident in device memory when the timed region starts (setup,
factorization, training and upload happen before warm-up).  The library calls
are synchronous (they return once their HIP stream is idle), so the host
clock around them brackets the device work; the dominant kernel is also
timed live with HIP events on the stream it is launched on.

Workloads (seeded, pdos_amd/problems.py):
  family     (default, the headline) portfolio family n_a=500, k=20: A and b
             shared, mu and gamma vary.  65,536 training instances are solved
             by the batched GPU solver, the MLP (501 -> 512 -> 512 -> 2545,
             GELU, BF16 MFMA kernels) is trained on them on the GPU, and the
             timed step is: MLP inference for 16,384 held-out instances +
             batched warm-started ADMM with the in-solver fallback judge.
             Also reported: the same batch cold, and the plain single-instance
             GPU path (pdos_ex, one call per instance).
  dense      1,024 instances sharing one dense 8,192 x 2,048 A (per-instance b, c)
  uneven     4,096 portfolios, n_a in [50,1000], k in [5,50] (one workgroup each)
  random     one random feasible SOCP, n=100,000, m=300,000, nnz~5M (grid engine)
  portfolio  the n=101, m=212 portfolio, one instance

Prints one JSON line.  --gpus N > 1 exits: this build targets one GPU.
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
BF16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------- FP64 verification

def verify_batch(A, B, Cm, X, Y, S, status, eps):
    """Recomputes the stop test (pdos.c:95-107, original units) of every
    Solved instance in binary64 on the host.  A is shared (sparse or dense),
    B is m-vector (shared) or count x m.  Returns the number that pass."""
    Bm = np.broadcast_to(B, (X.shape[0], A.shape[0])) if B.ndim == 1 else B
    P = (A @ X.T).T + S - Bm
    Dd = (A.T @ Y.T).T + Cm
    pres = np.linalg.norm(P, axis=1)
    dres = np.linalg.norm(Dd, axis=1)
    po = np.einsum("ij,ij->i", Cm, X)
    do = -np.einsum("ij,ij->i", Bm, Y)
    ok = ((pres < eps * (1 + np.linalg.norm(Bm, axis=1))) &
          (dres < eps * (1 + np.linalg.norm(Cm, axis=1))) &
          (np.abs(po - do) < eps * (1 + np.abs(po) + np.abs(do))) & (status == "Solved"))
    return int(ok.sum())


def verify_list(probs, outs):
    ok = 0
    for p, o in zip(probs, outs):
        if o["status"] != "Solved":
            continue
        A = sp.csc_matrix((p.Ax, p.Ai, p.Ap), shape=(p.m, p.n))
        ok += verify_batch(A, p.b, p.c[None, :], o["x"][None, :], o["y"][None, :], o["s"][None, :],
                           np.array([o["status"]]), p.params["EPS_ABS"])
    return ok


# -------------------------------------------------------------- CPU baseline

def _cpu_worker(args):
    kind, plugin, blobs, core = args
    if core is not None:
        os.sched_setaffinity(0, {core})  # one worker per core
    import ctypes as C
    sys.path.insert(0, ROOT)
    from pdos_amd.abi import Cone, Data, Problem, Sol
    if kind == "reference":
        L = C.CDLL(os.path.join(ROOT, "oracle", "_ref",
                                f"libpdos{'dir' if plugin == 'direct' else 'indir'}_ref.so"))
        L.pdos.restype = C.POINTER(Sol)
        L.pdos.argtypes = [C.POINTER(Data), C.POINTER(Cone)]
        L.freeSol.argtypes = [C.POINTER(C.POINTER(Sol))]

        def solve(d, k):
            sol = L.pdos(C.byref(d), C.byref(k))
            ok = int(sol.contents.status.split(b"\0")[0] == b"Solved")
            L.freeSol(C.byref(sol))
            return ok
    else:
        from oracle import oracle as O

        def solve(d, k):
            sol = O.lib().oracle_pdos(C.byref(d), C.byref(k), int(plugin == "direct"), None, None)
            ok = int(bytes(sol.contents.status).split(b"\0")[0] == b"Solved")
            O.lib().oracle_free_sol(sol)
            return ok
    probs = [Problem(**b) for b in blobs]
    cs = [p.to_c() for p in probs]
    solved = 0
    t0 = time.perf_counter()
    for d, k, _ in cs:
        solved += solve(d, k)
    return time.perf_counter() - t0, solved


# The GPU box allots each GPU a share of its host cores (16 per GPU): the
# baseline runs one pinned worker process per core of that share; nproc and
# the CPU model are reported beside it.
HOST_CORES_PER_GPU = 16


def host_cpu_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)), "cpu_model": model,
            "cores_per_gpu_share": HOST_CORES_PER_GPU}


def cpu_baseline(probs, plugin, procs, cap=0, gpu_mean_iters=0.0, gpu_solved_frac=1.0, inner=False):
    """The reference C library (oracle/_ref, built from the reference sources)
    -- or the oracle port if absent -- on the given instances, split over
    `procs` worker processes (one per core), factorization included; value =
    Solved instances per second.  With cap > 0 (instances too slow to solve
    on the CPU inside the bench) every instance runs exactly `cap` ADMM
    iterations (EPS_ABS 0) and the rate is extrapolated to the GPU run's mean
    iteration count and solved fraction (same FP64 algorithm, same counts).
    inner: the time is the workers' own solve loop (the longest), without the
    pool's dispatch (for samples of sub-millisecond solves)."""
    kind = "reference" if os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libpdosdir_ref.so")) else "port"
    over = dict(VERBOSE=0) if not cap else dict(VERBOSE=0, MAX_ITERS=cap, EPS_ABS=0.0)
    blobs = [dict(n=p.n, m=p.m, Ap=p.Ap, Ai=p.Ai, Ax=p.Ax, b=p.b, c=p.c, f=p.f, l=p.l, q=p.q,
                  params=dict(p.params, **over)) for p in probs]
    procs = max(1, min(procs, len(blobs)))
    cores = sorted(os.sched_getaffinity(0))
    pin = [cores[i % len(cores)] if len(cores) >= procs else None for i in range(procs)]
    shards = [blobs[i::procs] for i in range(procs)]
    ctx = mp.get_context("fork")
    with ctx.Pool(procs) as pool:
        if not cap:
            pool.map(_cpu_worker, [(kind, plugin, s[:1], pin[i]) for i, s in enumerate(shards) if s])  # warm
        t0 = time.perf_counter()
        res = pool.map(_cpu_worker, [(kind, plugin, s, pin[i]) for i, s in enumerate(shards) if s])
        wall = time.perf_counter() - t0
    if inner:
        wall = max(r[0] for r in res)
    out = {"unit": "SOCPs/s", "cores": procs, "kind": kind, "plugin": plugin, "pinned": all(c is not None for c in pin)}
    out.update(host_cpu_info())
    if not cap:
        solved = sum(r[1] for r in res)
        out["value"] = round(solved / wall, 3)
        out["sample"] = (f"{len(probs)} instances of the same workload ({solved} Solved), {procs} processes, "
                         f"{wall:.3f} s {'in the solve loop' if inner else 'wall'}, factorization included, "
                         f"cold start")
    else:
        per_inst = wall / (len(probs) / procs)  # seconds per instance for cap iterations on one core
        t_solve = per_inst * gpu_mean_iters / cap
        out["value"] = round(procs * gpu_solved_frac / t_solve, 4)
        out["extrapolated"] = True
        out["sample"] = (f"{len(probs)} instances x {cap} ADMM iterations (EPS_ABS 0), {procs} processes, "
                         f"{wall:.1f} s wall incl. setup, extrapolated to {gpu_mean_iters:.1f} iterations per "
                         f"instance and the solved fraction {gpu_solved_frac:.3f} of the GPU run")
    return out


def plain_gpu(probs, plugin, precision):
    """The plain single-instance GPU path: one pdos_ex call per instance
    (host setup, upload, solve, download), cold start."""
    pdos_amd.solve_problem(probs[0], plugin=plugin, precision=precision)
    t0 = time.perf_counter()
    ns = sum(pdos_amd.solve_problem(p, plugin=plugin, precision=precision)["status"] == "Solved"
             for p in probs)
    dt = time.perf_counter() - t0
    return {"value": round(ns / dt, 3), "unit": "SOCPs/s", "plugin": plugin, "precision": precision,
            "sample": f"{len(probs)} instances, one pdos_ex call each (setup, upload, solve, download)"}


# ------------------------------------------------------------------ roofline

# kernel names behind each timed kind of the family engine (sparse A)
# (sparse A; with dense rows in A the A' product is k_fam_at_hyb, the A
# product and the stop test ride in the x-update and phase C)
# kernels of each timed kind, in passes: one launch of each pass per
# iteration; a pass has a kernel for wide batches and one for narrow ones
FAMILY_KERNELS = {"At_product": [["k_fam_at_b<{t}>", "k_fam_at_hyb<{t}>"]],
                  "A_product": [["k_fam_ax_decide<{t}>"]], "xupdate_gemm": [["k_fam_xupdate"]],
                  "phase_b": [["k_fam_decide<{t}>"]],
                  "phase_c": [["k_fam_c1<{t}>", "k_fam_c1_ll<{t}>"], ["k_fam_c2<{t}>", "k_fam_c2_ll<{t}>"]],
                  "phase_a": [["k_fam_a<{t}>"]]}


def measured_traffic(profile, kind, fp64):
    """HBM bytes per launch of `kind` from a committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by a calibration copy; the JSON
    block of profiles/<profile>.md), or None when it does not cover the kind."""
    path = os.path.join(ROOT, "profiles", profile + ".md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    if "```json" not in text:
        return None
    data = json.loads(text.split("```json", 1)[1].split("```", 1)[0])["kernels"]
    t = "double" if fp64 else "float"
    passes = [[k.format(t=t) for k in ks if k.format(t=t) in data] for ks in FAMILY_KERNELS.get(kind, [])]
    if not passes or not passes[0]:
        return None
    # per-launch averages in the summary: total bytes over every launch of
    # the kind's kernels, per launch of its first pass (one per iteration)
    total = sum((data[k]["read_bytes"] + data[k]["write_bytes"]) * data[k]["launches"] for ks in passes for k in ks)
    return total / sum(data[k]["launches"] for k in passes[0])


def family_roofline(kt, fam_info, nnz, fp64, lin_rows, b_inst, traffic_profile=None, dense_cone=None):
    """Roofline of the dominant kernel of a family run, from per-kernel HIP
    event totals {kind: (ms, calls, columns)}.  Algorithmic bytes per column
    (instance), values of sv bytes:
      A' product  sparse A, fused with phase B: Y, W read (2m), c, X read
                  (2n), the x-update right-hand side written (bf16 planes
                  or sv); dense A: Y, W read, U, V written (2m + 2n)
                  (+ the shared matrix once per launch: nnz (sv + 4))
      phase B     dense A only: c, X, U, V read (4n), R written; sparse A:
                  the stop test, 16 bytes of partials per chunk
      A product   n read, m written (+ the matrix)
      phase A     S, Y, AX (+ b) read, W written: 4m (5m)
      phase C     zero/nonneg rows: AX, S, Y (+ b) read, S, Y, W written
                  (6, 7); cone rows: pass 1 AX, S, Y (+ b) read, W, S
                  written; pass 2 S, W, Y (+ AX, + b) read, Y, S, W written
                  (11 in mixed mode, where s~ stands in for b - AX; 12 in
                  fp64; + 2 with per-instance b).  With dense rows in A
                  phase C forms AX and does not store it: zero / nonneg rows
                  6 values, cone rows 9 (S, Y, XN read in each pass; Y, S, W
                  written), and each dense row reads the x-update's np / 64
                  planes in every pass over it (dense_cone: those in cones)
    x-update GEMM: 2 n np flops per product (bf16) or 2 n^2 (fp64), with
                  dense rows in A + 2 * 32 np (their A_h XN planes)."""
    n, m, npad = fam_info["n"], fam_info["m"], fam_info["np"]
    dense = bool(fam_info.get("dense", 0))
    nprod = fam_info.get("products", 1)
    sv = 8 if fp64 else 4
    kind = max(kt, key=lambda k: kt[k][0])
    ms, calls, cols = kt[kind]
    cone_rows = m - lin_rows
    rhs = n * sv if fp64 else n * 2 * nprod
    if dense and kind in ("At_product", "A_product"):
        # dense A: GEMMs.  A' products: two GEMMs (Y and W) of 2 m n flops per
        # column; the A product one.  fp64: rocBLAS dgemm against the fp64
        # peak; mixed: the hand-written kernels run each product as three bf16
        # products over split operands (hi hi + hi lo + lo hi), the flops
        # executed on the bf16 MFMA
        if kind == "At_product":
            flops, peak = (2 * 2.0 * m * n * cols, FP64_PEAK_TFLOPS) if fp64 else \
                (3 * 2 * 2.0 * m * n * cols, BF16_PEAK_TFLOPS)
        else:
            flops, peak = (2.0 * m * n * cols, FP64_PEAK_TFLOPS) if fp64 else (3 * 2.0 * m * n * cols, BF16_PEAK_TFLOPS)
        ach = flops / (ms * 1e-3) / 1e12
        return {"bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 4), "traffic": None, "kernel": kind,
                "algorithmic_flops_per_launch": flops / max(calls, 1), "avg_launch_us": round(1e3 * ms / max(calls, 1), 2)}
    nh = fam_info.get("dense_rows", 0)
    if kind == "xupdate_gemm":
        flops = (2.0 * n * n if fp64 else 2.0 * npad * nprod * npad + (2.0 * 32 * npad if nh else 0.0)) * cols
        ach = flops / (ms * 1e-3) / 1e12
        peak = FP64_PEAK_TFLOPS if fp64 else BF16_PEAK_TFLOPS
        return {"bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 4), "traffic": None, "kernel": kind,
                "algorithmic_flops_per_launch": flops / max(calls, 1), "avg_launch_us": round(1e3 * ms / max(calls, 1), 2)}
    if kind == "At_product":
        per_col = 2 * (m + n) * sv + (0 if dense else rhs)
        byt = cols * per_col + calls * nnz * (sv + 4)
    elif kind == "A_product":
        byt = cols * (m + n) * sv + calls * nnz * (sv + 4)
    elif kind == "phase_a":
        byt = cols * m * sv * (5 if b_inst else 4)
    elif kind == "phase_b":
        byt = cols * (n * sv * 4 + rhs) if dense else cols * 16 * ((n + 63) // 64 + (m + 63) // 64)
    elif nh:
        # dense rows in A (mixed): AX is formed from XN (or the planes) and
        # not stored; a zero / nonneg row reads S, Y, XN and writes S, Y, W;
        # a cone row is read twice (S, Y, XN: once for the norm, once for
        # the projection) and written once (Y, S, W); a dense row reads the
        # np / 64 planes in each pass that touches it
        nh_cone = dense_cone if dense_cone is not None else 0
        rows = (7 if b_inst else 6) * lin_rows + (9 + (2 if b_inst else 0)) * cone_rows \
            + ((nh - nh_cone) + 2 * nh_cone) * (npad // 64)
        byt = cols * rows * sv
    else:
        cone = (12 if fp64 else 11) + (2 if b_inst else 0)
        rows = (7 if b_inst else 6) * lin_rows + cone * cone_rows
        byt = cols * rows * sv
    ach = byt / (ms * 1e-3) / 1e9
    traffic = measured_traffic(traffic_profile, kind, fp64) if traffic_profile and not dense else None
    out = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kind,
           "algorithmic_bytes_per_launch": byt / max(calls, 1), "avg_launch_us": round(1e3 * ms / max(calls, 1), 2)}
    if traffic is not None:
        out["traffic_unit"] = "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/%s.md)" % traffic_profile
    return out


def wg_roofline(probs, outs, plugin, sv, kernel_ms):
    """SpMV bytes one k_wg launch executed (B_A / B_At per SpMV, s_i = 4)
    plus the direct plugin's two triangular sweeps per iteration."""
    si, tot = 4, 0.0
    for p, o in zip(probs, outs):
        it, cg = o["info"]["iters"], o["info"]["cg_iters"]
        solved = 1 if o["status"] == "Solved" else 0
        nnz, n, m = p.nnz, p.n, p.m
        BA = nnz * (sv + si) + (m + 1) * si + n * sv + m * sv
        BAt = nnz * (sv + si) + (n + 1) * si + m * sv + n * sv
        if plugin == "indirect":
            tot += (1 + it + cg) * BA + (2 * it + cg + solved) * BAt
        else:
            N, lnz = n + m, o["info"]["lnz"]
            tot += (1 + it) * BA + (it + solved) * BAt + it * 2 * (lnz * (8 + si) + (N + 1) * si + 2 * N * 8)
    ach = tot / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_wg",
            "algorithmic_bytes_per_launch": tot, "avg_launch_us": round(kernel_ms * 1e3, 1)}


# ------------------------------------------------------------------ workloads

def run_family(args):
    from pdos_amd import DeviceArray
    from pdos_amd.family import Family
    from pdos_amd.warmstart import LearnedWarmStart, portfolio_features

    fam = problems.PortfolioFamily(seed=4, na=500, k=20, EPS_ABS=1e-3, MAX_ITERS=2000)
    n_tr, n_te = args.train or 65536, args.count or 16384
    th_tr, C_tr = fam.sample(seed=1001, count=n_tr)
    th_te, C_te = fam.sample(seed=1004, count=n_te)
    prec = args.precision or "mixed"
    F = Family(fam.instance(C_te[0]), precision=prec, bf16_products=args.products)
    if args.engines:
        F.engines(args.engines)
    info = F.info()
    # 1. training data: the solver's own solutions
    t0 = time.perf_counter()
    F.load(C_tr)
    F.run()
    tr = F.fetch()
    t_data = time.perf_counter() - t0
    keep = tr["status"] == "Solved"
    T = np.concatenate([tr["x"], tr["y"], tr["s"]], axis=1)[keep]
    log(f"training data: {n_tr} solved on the GPU in {t_data:.1f} s, {keep.sum()} Solved, "
        f"mean iters {tr['iters'].mean():.1f}")
    # 2. training on the GPU (own BF16 kernels)
    t0 = time.perf_counter()
    ws = LearnedWarmStart(th_tr.shape[1], T.shape[1], hidden=512, seed=0)
    loss = ws.fit(portfolio_features(th_tr[keep]), T, epochs=args.epochs, batch=args.mlp_batch, lr=args.mlp_lr)
    t_train = time.perf_counter() - t0
    log(f"MLP trained: {args.epochs} epochs of batch {args.mlp_batch}, loss {loss:.4f}, {t_train:.1f} s")
    # 3. held-out batch resident on the device
    F.load(C_te)
    dF = DeviceArray(portfolio_features(th_te))
    ld = T.shape[1]
    dW = DeviceArray(shape=(n_te, ld))
    # cold batched
    for _ in range(args.warmup):
        F.run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        F.run()
    cold_wall = (time.perf_counter() - t0) / args.steps
    cold = F.fetch()

    def step():
        ws.predict_dev(dF, n_te, dW, ld)
        F.run(dW, ld, judge=True)

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    wall = (time.perf_counter() - t0) / args.steps
    warm = F.fetch()
    info = F.info()
    comp = info["compactions"]
    # one more step with per-kernel events for the roofline, on one engine
    # over the whole batch (the timed steps split it over two running at once)
    engines = F.engines()
    F.engines(1)
    F.load(C_te)
    F.timing(True)
    step()
    kt = F.timing(False)
    F.engines(engines)
    # end to end: the same step with the per-instance data uploaded (c, and
    # the features the network reads) and the solutions downloaded and
    # unscaled, as a caller holding host arrays sees it
    feats = portfolio_features(th_te)
    e2e_steps = max(1, min(args.steps, 5))
    dF2 = DeviceArray(shape=feats.shape)  # a caller's input buffer, reused from call to call
    t0 = time.perf_counter()
    for _ in range(e2e_steps):
        F.load(C_te)
        dF2.upload(feats)
        ws.predict_dev(dF2, n_te, dW, ld)
        F.run(dW, ld, judge=True)
        e2e = F.fetch()
    e2e_wall = (time.perf_counter() - t0) / e2e_steps
    A = sp.csc_matrix((fam.Ax, fam.Ai, fam.Ap), shape=(fam.m, fam.n))
    v_e2e = verify_batch(A, fam.b, C_te, e2e["x"], e2e["y"], e2e["s"], e2e["status"], 1e-3)
    v_warm = verify_batch(A, fam.b, C_te, warm["x"], warm["y"], warm["s"], warm["status"], 1e-3)
    v_cold = verify_batch(A, fam.b, C_te, cold["x"], cold["y"], cold["s"], cold["status"], 1e-3)
    solved = int((warm["status"] == "Solved").sum())
    solved_cold = int((cold["status"] == "Solved").sum())
    value = v_warm / wall  # FP64-verified Solved instances per second
    log(f"warm: {solved}/{n_te} Solved ({v_warm} verified), {value:.0f} SOCPs/s, mean iters "
        f"{warm['iters'].mean():.1f}, fallbacks {int(warm['fallback'].sum())}; cold: {solved_cold / cold_wall:.0f} "
        f"SOCPs/s, mean iters {cold['iters'].mean():.1f}")
    # error against binary64: every result against the same batch solved by
    # the family engine in FP64 to EPS_ABS 1e-6; then FP64 polish iterations
    # warm-started from the mixed results (to EPS_ABS 1e-4, at most 300)
    fp64_ref = fp64_error_report(fam, C_te, warm, args) if args.fp64_check else None
    probs = [fam.instance(C_te[i]) for i in range(max(args.plain_sample, args.cpu_sample))]
    plain = None
    if args.plain_sample:
        plain = {prec_: plain_gpu(probs[:args.plain_sample], "direct", prec_) for prec_ in ("fp64", "mixed")}
        log(f"plain GPU: fp64 {plain['fp64']['value']} / mixed {plain['mixed']['value']} SOCPs/s")
    config = {"workload": "portfolio family, learned warm start, held-out batch", "n_a": 500, "k": 20,
              "n": fam.n, "m": fam.m, "nnz": int(fam.Ap[-1]), "count": n_te, "train_count": n_tr,
              "eps_abs": 1e-3, "max_iters": 2000, "precision": prec, "engine": "family (shared A)",
              "engines": info["engines"], "dense_rows": info["dense_rows"],
              "mlp": "501-512-512-2545 GELU bf16"}
    extra = {
        "batched_cold": {"value": round(solved_cold / cold_wall, 3), "unit": "SOCPs/s",
                         "ms_per_step": round(cold_wall * 1e3, 3), "mean_iters": float(cold["iters"].mean()),
                         "solved": solved_cold, "verified_fp64": v_cold},
        "warm_start": {"mean_iters": float(warm["iters"].mean()), "fallbacks": int(warm["fallback"].sum()),
                       "train_loss": loss, "train_s": round(t_train, 2), "data_s": round(t_data, 2),
                       "compactions": comp},
        "kernels": {k: {"ms": round(v[0], 3), "calls": v[1]} for k, v in kt.items()},
        "end_to_end": {"value": round(v_e2e / e2e_wall, 3), "unit": "SOCPs/s", "ms_per_step": round(e2e_wall * 1e3, 3),
                       "verified_fp64": v_e2e,
                       "includes": "host-to-device copy of c and the features, inference, solve, download and "
                                   "unscale of x, y, s"},
        "plain_gpu": plain,
        # against the faster of the two plain single-instance paths
        "ratio_vs_plain_gpu": round(value / max(v["value"] for v in plain.values()), 1) if plain else None,
        "fp64_error": fp64_ref,
    }
    prof = "family_portfolio_16384_pmc_traffic" if (args.count or 16384) == 16384 else None
    row_nnz = np.bincount(fam.Ai, minlength=fam.m)
    dense_cone = int(((row_nnz >= 64) & (np.arange(fam.m) >= fam.f + fam.l)).sum())
    roof = family_roofline(kt, info, int(fam.Ap[-1]), prec == "fp64", fam.f + fam.l, False, prof, dense_cone)
    acc = {"solved": solved, "count": n_te, "verified_fp64": v_warm,
           "mean_iters": float(warm["iters"].mean()), "max_iters": int(warm["iters"].max())}
    return value, wall, "f32+bf16" if prec == "mixed" else "f64", config, roof, acc, extra, probs[:args.cpu_sample], "direct"


def _errors(C, out, ref):
    """Per instance: relative objective gap and relative distance of (x, y, s)
    to the reference solution."""
    po = np.einsum("ij,ij->i", C, out["x"])
    pr = np.einsum("ij,ij->i", C, ref["x"])
    gap = np.abs(po - pr) / (1 + np.abs(pr))
    num = np.sqrt(sum(np.sum((out[k] - ref[k]) ** 2, axis=1) for k in ("x", "y", "s")))
    den = np.sqrt(sum(np.sum(ref[k] ** 2, axis=1) for k in ("x", "y", "s")))
    return gap, num / np.maximum(den, 1e-300)


def _stats(v):
    return {"max": float(v.max()), "p99": float(np.quantile(v, 0.99)), "mean": float(v.mean()),
            "median": float(np.median(v))}


def fp64_error_report(fam, C_te, warm, args):
    from pdos_amd.family import Family, polish
    t0 = time.perf_counter()
    R = Family(fam.instance(C_te[0]).with_params(EPS_ABS=1e-6, MAX_ITERS=20000), precision="fp64")
    R.load(C_te)
    R.run()
    ref = R.fetch()
    t_ref = time.perf_counter() - t0
    keep = (warm["status"] == "Solved") & (ref["status"] == "Solved")
    gap, dist = _errors(C_te[keep], {k: warm[k][keep] for k in "xys"}, {k: ref[k][keep] for k in "xys"})
    out = {"reference": "the same batch, family engine in binary64 to EPS_ABS 1e-6 (MAX_ITERS 20,000)",
           "reference_solved": int((ref["status"] == "Solved").sum()), "compared": int(keep.sum()),
           "reference_s": round(t_ref, 2), "obj_rel_gap": _stats(gap), "dist_rel": _stats(dist)}
    # FP64 polish: iterations in binary64 from the mixed results (pdos_amd.family.polish)
    polish(fam.instance(C_te[0]), C_te, warm)  # first use: kernels loaded, device buffers sized
    t0 = time.perf_counter()
    pol = polish(fam.instance(C_te[0]), C_te, warm)
    t_pol = time.perf_counter() - t0
    gap2, dist2 = _errors(C_te[keep], {k: pol[k][keep] for k in "xys"}, {k: ref[k][keep] for k in "xys"})
    out["polish"] = {"eps_abs": 1e-4, "max_iters": 300, "ms": round(t_pol * 1e3, 2),
                     "ms_includes": "family creation (G formed), load, upload of the warm start, run, fetch",
                     "mean_iters": float(pol["iters"].mean()),
                     "solved_at_1e-4": int((pol["status"] == "Solved").sum()),
                     "obj_rel_gap": _stats(gap2), "dist_rel": _stats(dist2)}
    if args.errors_out:
        np.savez_compressed(args.errors_out, keep=keep, obj_gap=gap, dist=dist, obj_gap_polished=gap2,
                            dist_polished=dist2)
    log(f"fp64 error: obj gap max {gap.max():.2e} p99 {np.quantile(gap, .99):.2e}; dist max {dist.max():.2e}; "
        f"after polish ({t_pol * 1e3:.0f} ms, mean {pol['iters'].mean():.1f} iters): obj gap max {gap2.max():.2e}, "
        f"dist max {dist2.max():.2e}")
    return out


def run_dense(args):
    from pdos_amd.family import Family
    dense = problems.DenseFamily(seed=2, EPS_ABS=1e-3, MAX_ITERS=2000)
    cnt = args.count or 1024
    B, Cm, opt = dense.sample(seed=1002, count=cnt)
    prec = args.precision or "mixed"
    F = Family(dense.instance(B[0], Cm[0]), precision=prec)
    F.load(Cm, B)
    for _ in range(args.warmup):
        F.run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        F.run()
    wall = (time.perf_counter() - t0) / args.steps
    out = F.fetch()
    F.timing(True)
    F.run()
    kt = F.timing(False)
    solved = int((out["status"] == "Solved").sum())
    ver = verify_batch(dense.A, B, Cm, out["x"], out["y"], out["s"], out["status"], 1e-3)
    gap = np.abs(np.einsum("ij,ij->i", Cm, out["x"]) - opt) / (1 + np.abs(opt))
    config = {"workload": "dense shared-A family", "m": dense.m, "n": dense.n, "count": cnt, "eps_abs": 1e-3,
              "precision": prec, "engine": "family (shared dense A)"}
    roof = family_roofline(kt, F.info(), dense.m * dense.n, prec == "fp64", dense.f + dense.l, True)
    acc = {"solved": solved, "count": cnt, "verified_fp64": ver, "mean_iters": float(out["iters"].mean()),
           "max_iters": int(out["iters"].max()), "max_rel_obj_err_vs_planted": float(gap.max())}
    extra = {"kernels": {k: {"ms": round(v[0], 3), "calls": v[1]} for k, v in kt.items()}}
    probs = [dense.instance(B[i], Cm[i]) for i in range(min(args.cpu_sample, 16))]
    return ver / wall, wall, "f32+bf16" if prec == "mixed" else "f64", config, roof, acc, extra, probs, "direct"


def run_wg(args, probs, config, plugin):
    prec = args.precision or "fp64"
    t0 = time.perf_counter()
    h = pdos_amd.BatchHandle(probs, plugin=plugin, precision=prec, tune=bool(args.tune))
    prep = time.perf_counter() - t0
    tuning = h.tuning()
    parts = h.prepare_times()
    log(f"prepare: {prep:.2f} s (Python conversion {h.convert_s:.2f} s; library ms {parts})")
    for r in tuning:
        log(f"  tuned n+m in {r['n_plus_m']}: {r['count']} instances -> {r['variant']}, persistent "
            f"{r['persistent']}, {r['entries_per_lane']} entries per lane, {r['threads']} threads; candidates "
            f"(variant, persistent, entries per lane, threads, ms): {r['candidates']}")
    for _ in range(args.warmup):
        h.run()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kms.append(h.run())
    wall = (time.perf_counter() - t0) / args.steps
    outs = h.fetch()
    dev_bytes = h.device_bytes()
    h.close()
    solved = sum(o["status"] == "Solved" for o in outs)
    iters = np.array([o["info"]["iters"] for o in outs])
    roof = wg_roofline(probs, outs, plugin, 8 if prec == "fp64" else 4, float(np.mean(kms)))
    acc = {"solved": int(solved), "count": len(outs), "verified_fp64": verify_list(probs, outs),
           "mean_iters": float(iters.mean()), "max_iters": int(iters.max())}
    extra = {"plain_gpu": plain_gpu(probs[:args.plain_sample], plugin, prec) if args.plain_sample else None,
             "prepare_s": round(prep, 3), "prepare_parts_ms": parts, "python_convert_s": round(h.convert_s, 3),
             "device_bytes": dev_bytes,
             "host_threads": int(os.environ.get("PDOS_HOST_THREADS", os.cpu_count())),
             "tuning": [{k: r[k] for k in ("n_plus_m", "count", "variant", "persistent", "entries_per_lane",
                                          "threads")} for r in tuning]}
    return acc["verified_fp64"] / wall, wall, "f64" if prec == "fp64" else "f32", config, roof, acc, extra, \
        probs[:args.cpu_sample], plugin


def run_single(args, p, config, plugin):
    prec = args.precision or "fp64"
    for _ in range(args.warmup):
        pdos_amd.solve_problem(p, plugin=plugin, precision=prec)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = pdos_amd.solve_problem(p, plugin=plugin, precision=prec)
    wall = (time.perf_counter() - t0) / args.steps
    solved = int(out["status"] == "Solved")
    if plugin == "indirect" and out["info"]["engine"] == 2:
        it, cg = out["info"]["iters"], out["info"]["cg_iters"]
        sv, si = (8 if prec == "fp64" else 4), 4
        BA = p.nnz * (sv + si) + (p.m + 1) * si + (p.n + p.m) * sv
        BAt = p.nnz * (sv + si) + (p.n + 1) * si + (p.n + p.m) * sv
        byt = (1 + it + cg) * BA + (2 * it + cg + solved) * BAt
        ach = byt / (out["info"]["solve_ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "grid engine (all kernels)",
                "algorithmic_bytes_per_solve": byt, "spmv_count": 1 + it + cg + 2 * it + cg + solved}
    else:
        roof = wg_roofline([p], [out], plugin, 8 if prec == "fp64" else 4, out["info"]["solve_ms"])
    acc = {"solved": solved, "count": 1, "verified_fp64": verify_list([p], [out]),
           "mean_iters": float(out["info"]["iters"]), "max_iters": int(out["info"]["iters"])}
    return acc["verified_fp64"] / wall, wall, "f64" if prec == "fp64" else "f32", config, roof, acc, {}, [p], plugin


def calibration_copy(mib=256):
    """One plain copy of `mib` MiB between two device buffers by the library's
    copy kernel (k_copy): under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE
    per dispatch, against the known mib * 1024 KiB, gives the counters'
    scale."""
    src = pdos_amd.DeviceArray(np.ones(mib << 18, dtype=np.float32))
    dst = pdos_amd.DeviceArray(shape=(mib << 18,))
    dst.copy_from(src)
    log(f"calibration copy: {mib} MiB device to device (k_copy)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="family")
    ap.add_argument("--count", type=int, default=0)
    ap.add_argument("--train", type=int, default=0)
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--mlp-batch", type=int, default=512, help="family: MLP training mini-batch")
    ap.add_argument("--mlp-lr", type=float, default=1e-3, help="family: MLP peak learning rate (Adam)")
    ap.add_argument("--plugin", default="direct")
    ap.add_argument("--precision", default="")
    ap.add_argument("--cpu-sample", type=int, default=1024)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-check", type=int, default=1,
                    help="family: error of every result against an FP64 solve of the batch, and FP64 polish")
    ap.add_argument("--errors-out", default="", help="family: write per-instance errors to this .npz")
    ap.add_argument("--products", type=int, default=0, help="family, mixed: bf16 products (1-3)")
    ap.add_argument("--engines", type=int, default=0,
                    help="family: engines the batch is split over (0: the library's default)")
    ap.add_argument("--tune", type=int, default=1,
                    help="uneven: time candidate launch shapes per size bucket at prepare (PdosOptions.tune)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with one 256 MiB device-to-device copy (the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters in profiled runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    w = args.workload
    if w == "family":
        res = run_family(args)
    elif w == "dense":
        res = run_dense(args)
    elif w == "uneven":
        probs = problems.uneven_portfolios(seed=3, count=args.count or 4096, EPS_ABS=1e-3, MAX_ITERS=2000)
        cfg = {"workload": "uneven portfolio batch", "count": len(probs), "n_a": [50, 1000], "k": [5, 50],
               "eps_abs": 1e-3, "nnz_total": int(sum(p.nnz for p in probs)), "plugin": args.plugin,
               "engine": "workgroup per instance"}
        res = run_wg(args, probs, cfg, args.plugin)
    elif w == "random":
        p = problems.random_feasible(seed=1, EPS_ABS=1e-3, MAX_ITERS=2000)
        cfg = {"workload": "random feasible SOCP", "n": p.n, "m": p.m, "nnz": p.nnz, "count": 1,
               "eps_abs": 1e-3, "plugin": "indirect", "engine": "grid"}
        res = run_single(args, p, cfg, "indirect")
    elif w == "portfolio":
        p = problems.portfolio(seed=0)  # the n=101 portfolio's own options: EPS_ABS 5e-3, ALPHA 1, NORMALIZE 0
        cfg = {"workload": "portfolio n=101 m=212", "n": p.n, "m": p.m, "nnz": p.nnz, "count": 1,
               "eps_abs": p.params["EPS_ABS"], "plugin": args.plugin, "engine": "workgroup"}
        res = run_single(args, p, cfg, args.plugin)
    else:
        raise SystemExit(f"unknown workload {w}")
    value, wall, dtype, config, roof, acc, extra, cpu_probs, cpu_plugin = res

    cpu = None
    if not args.no_cpu and cpu_probs:
        procs = max(1, min(HOST_CORES_PER_GPU, len(os.sched_getaffinity(0))))
        if w in ("random", "dense"):  # bounded samples: a full CPU solve takes minutes (random) or hours (dense)
            cpu = cpu_baseline(cpu_probs, "indirect", procs, cap=10 if w == "random" else 3,
                               gpu_mean_iters=acc["mean_iters"], gpu_solved_frac=acc["solved"] / acc["count"])
        elif w == "portfolio":  # one instance: the latency of one core, over 200 back-to-back calls
            cpu = cpu_baseline(cpu_probs * 200, cpu_plugin, 1, inner=True)
        else:
            cpu = cpu_baseline(cpu_probs, cpu_plugin, procs)
        log(f"cpu baseline: {cpu['value']} SOCPs/s on {cpu['cores']} cores ({cpu['kind']})")

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "SOCPs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic", "config": config, "roofline": roof, "cpu_baseline": cpu,
        "accuracy": acc,
    }
    line.update(extra)
    if cpu:
        line["ratio_vs_cpu_baseline"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    if args.calibrate_copy:
        calibration_copy()


if __name__ == "__main__":
    main()
