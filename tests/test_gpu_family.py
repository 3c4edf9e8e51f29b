"""Shared-A family engine against the oracle, instance by instance.

FP64 mode: the x-update is G (x - lam c + A'(b - s - lam y)) with
G = (I + A'A)^-1 -- the direct plugin's solve up to roundoff -- so status,
iteration count and x, s, y (1e-6 relative) match the oracle's direct plugin.
Mixed mode: every Solved instance meets the stop test in binary64.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import golden_data as G
from oracle import oracle
from pdos_amd import DeviceArray, problems
from pdos_amd.abi import Problem
from pdos_amd.family import Family

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)


def _fam(na=60, k=6, **kw):
    return problems.PortfolioFamily(seed=4, na=na, k=k, EPS_ABS=1e-3, **kw)


@pytest.mark.parametrize("normalize", [0, 1])
def test_family_fp64_matches_oracle(normalize):
    fam = _fam(NORMALIZE=normalize)
    _, Cm = fam.sample(seed=7, count=40)
    print("family fp64", normalize, flush=True)
    F = Family(fam.instance(Cm[0]), precision="fp64")
    F.load(Cm)
    F.run()
    out = F.fetch()
    for t in range(len(Cm)):
        ref = oracle.solve(fam.instance(Cm[t]), "direct")
        assert out["status"][t] == ref["status"], t
        assert out["iters"][t] == ref["iters"], (t, out["iters"][t], ref["iters"])
        for key in ("x", "s", "y"):
            assert _rel(out[key][t], ref[key]) <= 1e-6, (t, key)


def test_dense_family_fp64_matches_oracle():
    dense = problems.DenseFamily(seed=2, m=96, n=24, l=64, nsoc=4, qdim=8, EPS_ABS=1e-3)
    B, Cm, _ = dense.sample(5, 10)
    print("dense family", flush=True)
    F = Family(dense.instance(B[0], Cm[0]), precision="fp64")
    F.load(Cm, B)
    F.run()
    out = F.fetch()
    for t in range(len(Cm)):
        ref = oracle.solve(dense.instance(B[t], Cm[t]), "direct")
        assert out["status"][t] == ref["status"]
        assert out["iters"][t] == ref["iters"]
        for key in ("x", "s", "y"):
            assert _rel(out[key][t], ref[key]) <= 1e-6
    # golden case of the same family: the reference's own output
    exp = G.expected("dense_m96_n24", "direct")
    gp = G.problem("dense_m96_n24")
    F2 = Family(gp, precision="fp64")
    F2.load(gp.c[None, :], gp.b[None, :])
    F2.run()
    o2 = F2.fetch()
    assert o2["iters"][0] == exp["iters"] and o2["status"][0] == exp["status"]
    for key in ("x", "s", "y"):
        assert _rel(o2[key][0], exp[key]) <= 1e-6


def _verify(fam, Cm, out, slack):
    A = sp.csc_matrix((fam.Ax, fam.Ai, fam.Ap), shape=(fam.m, fam.n))
    eps = fam.params["EPS_ABS"]
    nb = np.linalg.norm(fam.b)
    ok = 0
    for t in range(len(Cm)):
        if out["status"][t] != "Solved":
            continue
        x, s, y, c = out["x"][t], out["s"][t], out["y"][t], Cm[t]
        pres = np.linalg.norm(A @ x + s - fam.b)
        dres = np.linalg.norm(c + A.T @ y)
        po, do = c @ x, -fam.b @ y
        ok += (pres < slack * eps * (1 + nb) and dres < slack * eps * (1 + np.linalg.norm(c))
               and abs(po - do) < slack * eps * (1 + abs(po) + abs(do)))
    return ok


def test_family_mixed_solutions_verify_in_fp64():
    fam = _fam(na=200, k=15)
    _, Cm = fam.sample(seed=8, count=300)
    print("family mixed", flush=True)
    F = Family(fam.instance(Cm[0]), precision="mixed")
    F.load(Cm)
    F.run()
    out = F.fetch()
    solved = int((out["status"] == "Solved").sum())
    assert solved >= 0.8 * len(Cm)
    assert _verify(fam, Cm, out, 1.05) == solved
    # objectives agree with the FP64 engine within the stop tolerance
    F64 = Family(fam.instance(Cm[0]), precision="fp64")
    F64.load(Cm)
    F64.run()
    o64 = F64.fetch()
    both = (out["status"] == "Solved") & (o64["status"] == "Solved")
    p32 = np.einsum("ij,ij->i", Cm, out["x"])[both]
    p64 = np.einsum("ij,ij->i", Cm, o64["x"])[both]
    assert np.all(np.abs(p32 - p64) <= 20 * 1e-3 * (1 + np.abs(p64)))


def test_family_split_batch_matches_single_engine_runs():
    """A mixed batch of 8,192 or more runs as two halves on two engines at
    once.  With dense rows in A no kernel's sums depend on the batch width,
    so every instance must come out bit for bit as in separate runs of 2,048."""
    fam = _fam(na=200, k=15, MAX_ITERS=300)
    _, Cm = fam.sample(seed=12, count=8192)
    print("family split batch", flush=True)
    F = Family(fam.instance(Cm[0]), precision="mixed")
    F.load(Cm)
    F.run()
    whole = F.fetch()
    parts = []
    for q in range(4):
        F.load(Cm[q * 2048:(q + 1) * 2048])
        F.run()
        parts.append(F.fetch())
    for key in ("iters", "status", "x", "y", "s"):
        np.testing.assert_array_equal(whole[key], np.concatenate([p[key] for p in parts]))
    assert (whole["status"] == "Solved").sum() >= 0.5 * len(Cm)


def test_warm_start_judge():
    fam = _fam()
    _, Cm = fam.sample(seed=9, count=16)
    print("judge", flush=True)
    F = Family(fam.instance(Cm[0]), precision="fp64")
    F.load(Cm)
    F.run()
    cold = F.fetch()
    n, m = fam.n, fam.m
    # the converged solution as warm start: accepted, stops at once
    W = np.concatenate([cold["x"], cold["y"], cold["s"]], axis=1).astype(np.float32)
    dW = DeviceArray(W)
    F.run(dW, W.shape[1], judge=True)
    hot = F.fetch()
    assert np.all(hot["fallback"] == 0)
    solved = cold["status"] == "Solved"
    assert np.all(hot["iters"][solved] <= 2)
    # a wild warm start: rejected, identical to the cold solve
    Wb = (1e6 * np.random.default_rng(0).standard_normal(W.shape)).astype(np.float32)
    dWb = DeviceArray(Wb)
    F.run(dWb, W.shape[1], judge=True)
    bad = F.fetch()
    assert np.all(bad["fallback"] == 1)
    np.testing.assert_array_equal(bad["iters"], cold["iters"])
    for key in ("x", "s", "y"):
        np.testing.assert_array_equal(bad[key], cold[key])


def test_max_iters_indeterminate():
    fam = _fam(MAX_ITERS=3)
    _, Cm = fam.sample(seed=10, count=5)
    F = Family(fam.instance(Cm[0]), precision="fp64")
    F.load(Cm)
    F.run()
    out = F.fetch()
    for t in range(5):
        ref = oracle.solve(fam.instance(Cm[t]), "direct")
        assert out["status"][t] == ref["status"] == "Indeterminate"
        assert out["iters"][t] == 3
        for key in ("x", "s", "y"):
            assert _rel(out[key][t], ref[key]) <= 1e-6


def test_dense_family_mixed_verifies_in_fp64():
    """Mixed precision with a dense A: the A and A' products are split-bf16
    GEMMs (three products, fp32 accumulation); every Solved instance meets the
    stop test in binary64 and the objective matches the planted optimum."""
    dense = problems.DenseFamily(seed=2, m=256, n=64, l=192, nsoc=8, qdim=8, EPS_ABS=1e-3)
    B, Cm, opt = dense.sample(6, 64)
    print("dense family mixed", flush=True)
    F = Family(dense.instance(B[0], Cm[0]), precision="mixed")
    F.load(Cm, B)
    F.run()
    out = F.fetch()
    solved = out["status"] == "Solved"
    assert solved.sum() >= 60
    A = dense.A
    for t in np.nonzero(solved)[0]:
        x, s, y, b, c = out["x"][t], out["s"][t], out["y"][t], B[t], Cm[t]
        assert np.linalg.norm(A @ x + s - b) < 1.01e-3 * (1 + np.linalg.norm(b))
        assert np.linalg.norm(c + A.T @ y) < 1.01e-3 * (1 + np.linalg.norm(c))
        assert abs(c @ x - opt[t]) <= 2e-2 * (1 + abs(opt[t]))


def test_family_wide_batch_uses_both_phase_c_variants():
    """9,216 instances on one engine: phase C runs its wide kernels while
    more than 128 blocks of 64 instances are live and its narrow,
    latency-first kernels after compaction; the same instances in runs of
    3,072 use only the narrow ones.  Both do the same arithmetic in the same
    order, so statuses and iteration counts agree, and every Solved point
    meets the stop test in binary64."""
    fam = _fam(na=200, k=15, MAX_ITERS=400)
    _, Cm = fam.sample(seed=13, count=9216)
    print("family wide batch", flush=True)
    F = Family(fam.instance(Cm[0]), precision="mixed")
    F.engines(1)
    F.load(Cm)
    F.run()
    wide = F.fetch()
    parts = []
    for q in range(3):
        F.load(Cm[q * 3072:(q + 1) * 3072])
        F.run()
        parts.append(F.fetch())
    for key in ("iters", "status"):
        np.testing.assert_array_equal(wide[key], np.concatenate([p[key] for p in parts]))
    solved = int((wide["status"] == "Solved").sum())
    assert solved >= 0.5 * len(Cm)
    assert _verify(fam, Cm, wide, 1.05) == solved


def test_family_mixed_relaxed_normalized_verifies_in_fp64():
    """Over-relaxation (ALPHA 1.8) and NORMALIZE 1 in mixed precision, with
    dense rows in A: phase C's relaxed branches (W from b, AX, s, y rather
    than from s~) in both its wide and narrow kernels (9,000 instances on
    one engine).  Every Solved point meets the stop test in binary64, and the
    solved fraction matches the FP64 engine's on a subset."""
    fam = _fam(na=200, k=15, ALPHA=1.8, NORMALIZE=1, MAX_ITERS=600)
    _, Cm = fam.sample(seed=14, count=9000)
    print("family mixed relaxed", flush=True)
    F = Family(fam.instance(Cm[0]), precision="mixed")
    F.engines(1)
    F.load(Cm)
    F.run()
    out = F.fetch()
    solved = int((out["status"] == "Solved").sum())
    assert solved >= 0.5 * len(Cm)
    assert _verify(fam, Cm, out, 1.05) == solved
    F64 = Family(fam.instance(Cm[0]), precision="fp64")
    F64.load(Cm[:200])
    F64.run()
    o64 = F64.fetch()
    assert abs(int((o64["status"] == "Solved").sum()) - int((out["status"][:200] == "Solved").sum())) <= 6


def test_family_jumbled_csc_with_duplicates():
    """A stored as the input format allows: each column's rows in random order
    and 30% of the entries split into two of the same row, including the
    dense (budget and factor) rows that the mixed engine runs as MFMA tiles.
    FP64: status, iterations and x, s, y (1e-6) as the oracle's direct plugin
    on the same stored A; mixed: every Solved point meets the stop test in
    binary64 against the summed A."""
    fam = _fam(na=200, k=15, MAX_ITERS=2000)
    _, Cm = fam.sample(seed=15, count=512)
    t0 = problems.jumbled(fam.instance(Cm[0]), seed=3)
    assert np.bincount(t0.Ai, minlength=t0.m).max() >= 64
    print("family jumbled fp64", flush=True)
    F = Family(t0, precision="fp64")
    F.load(Cm[:12])
    F.run()
    out = F.fetch()
    for t in range(12):
        pt = Problem(t0.n, t0.m, t0.Ap, t0.Ai, t0.Ax, t0.b, Cm[t], t0.f, t0.l, t0.q, t0.params)
        ref = oracle.solve(pt, "direct")
        assert out["status"][t] == ref["status"], t
        assert out["iters"][t] == ref["iters"], (t, out["iters"][t], ref["iters"])
        for key in ("x", "s", "y"):
            assert _rel(out[key][t], ref[key]) <= 1e-6, (t, key)
    print("family jumbled mixed", flush=True)
    F = Family(t0, precision="mixed")
    assert F.info()["dense_rows"] > 0
    F.load(Cm)
    F.run()
    mix = F.fetch()
    solved = int((mix["status"] == "Solved").sum())
    assert solved >= 0.8 * len(Cm)
    A = sp.csc_matrix((t0.Ax, t0.Ai, t0.Ap), shape=(t0.m, t0.n))
    nb = np.linalg.norm(t0.b)
    for t in np.nonzero(mix["status"] == "Solved")[0]:
        x, s, y, c = mix["x"][t], mix["s"][t], mix["y"][t], Cm[t]
        assert np.linalg.norm(A @ x + s - t0.b) < 1.05e-3 * (1 + nb)
        assert np.linalg.norm(c + A.T @ y) < 1.05e-3 * (1 + np.linalg.norm(c))


def _seq_norm1p(v):
    """1 + |v| with the sum of squares in index order, one rounding per
    operation (the reference's loop, pdos.c:76-77)."""
    acc = 0.0
    for x in v:
        acc = acc + float(x) * float(x)
    return 1.0 + float(np.sqrt(acc))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_family_load_setup_on_device_bitwise(precision):
    """pdos_family_load scales each instance's c (and b) on the device: lambda
    equals the reference's initWork lambda (the oracle's setup of the same
    instance) bit for bit, and the stop-test scales equal the reference's
    sequential norms; shared b (portfolio family) and per-instance b (dense
    family); a second load of the same width reuses the device buffers."""
    fam = _fam(NORMALIZE=1)
    _, Cm = fam.sample(seed=9, count=70)
    print("family load scales, shared b", precision, flush=True)
    F = Family(fam.instance(Cm[0]), precision=precision)
    for rep in range(2):  # the second load reuses the batch's buffers
        F.load(Cm)
        sc = F.scales()
        for t in (0, 1, 33, 69):
            p = fam.instance(Cm[t])
            _, _, lam = oracle.setup(p)
            assert sc["lam"][t] == lam, (rep, t, sc["lam"][t], lam)
            assert sc["pscale"][t] == _seq_norm1p(p.b), t
            assert sc["dscale"][t] == _seq_norm1p(Cm[t]), t
    dense = problems.DenseFamily(seed=3, m=96, n=24, l=64, nsoc=4, qdim=8, EPS_ABS=1e-3, NORMALIZE=1)
    B, C2, _ = dense.sample(seed=5, count=10)
    print("family load scales, per-instance b", precision, flush=True)
    F2 = Family(dense.instance(B[0], C2[0]), precision=precision)
    F2.load(C2, B)
    sc = F2.scales()
    for t in range(10):
        p = dense.instance(B[t], C2[t])
        _, _, lam = oracle.setup(p)
        assert sc["lam"][t] == lam, (t, sc["lam"][t], lam)
        assert sc["pscale"][t] == _seq_norm1p(B[t]), t
        assert sc["dscale"][t] == _seq_norm1p(C2[t]), t
    # and the solve after the device-side setup still matches the oracle (fp64)
    if precision == "fp64":
        F2.run()
        out = F2.fetch()
        for t in range(10):
            ref = oracle.solve(dense.instance(B[t], C2[t]), "direct")
            assert out["status"][t] == ref["status"] and out["iters"][t] == ref["iters"], t
            for k in "xsy":
                assert _rel(out[k][t], ref[k]) <= 1e-6, (t, k)


def test_fp64_polish_moves_mixed_results_to_the_fp64_solution():
    """pdos_amd.family.polish: the FP64 family engine warm-started from the
    mixed results reaches its tighter tolerance in fewer iterations than a
    cold FP64 solve, and brings every result closer to the FP64 solution."""
    from pdos_amd.family import polish
    fam = _fam(na=60, k=6)
    _, Cm = fam.sample(seed=17, count=256)
    print("fp64 polish", flush=True)
    F = Family(fam.instance(Cm[0]), precision="mixed")
    F.load(Cm)
    F.run()
    mixed = F.fetch()
    R = Family(fam.instance(Cm[0]).with_params(EPS_ABS=1e-7, MAX_ITERS=20000), precision="fp64")
    R.load(Cm)
    R.run()
    ref = R.fetch()
    pol = polish(fam.instance(Cm[0]), Cm, mixed, eps_abs=1e-5, max_iters=2000)
    C5 = Family(fam.instance(Cm[0]).with_params(EPS_ABS=1e-5, MAX_ITERS=2000), precision="fp64")
    C5.load(Cm)
    C5.run()
    cold = C5.fetch()
    ok = (mixed["status"] == "Solved") & (ref["status"] == "Solved") & (pol["status"] == "Solved")
    assert ok.sum() >= 0.9 * len(Cm)
    assert pol["iters"][ok].mean() < cold["iters"][ok].mean()

    def dist(out):
        num = np.sqrt(sum(np.sum((out[k] - ref[k]) ** 2, axis=1) for k in ("x", "y", "s")))
        den = np.sqrt(sum(np.sum(ref[k] ** 2, axis=1) for k in ("x", "y", "s")))
        return (num / den)[ok]

    d0, d1 = dist(mixed), dist(pol)
    assert np.median(d1) < 0.2 * np.median(d0), (np.median(d0), np.median(d1))
    assert d1.max() < d0.max()


def test_fetch_in_chunks_with_outputs_left_out():
    """pdos_family_fetch copies x, y, s down through pinned staging chunks
    (32 MB) into the caller's arrays: a batch whose outputs span several
    chunks gives the same x and iterations when fetched again with y, s and
    the residuals left out (NULL), and every y differs from instance to
    instance (no chunk copied twice or left stale)."""
    import ctypes as C
    fam = _fam(na=500, k=20, MAX_ITERS=50)
    _, Cm = fam.sample(seed=41, count=4096)  # x, y, s: 4096 x (501 + 2 x 1022) doubles = 83 MB
    print("chunked fetch", flush=True)
    F = Family(fam.instance(Cm[0]), precision="mixed")
    F.load(Cm)
    F.run()
    full = F.fetch()
    X = np.zeros((len(Cm), fam.n))
    it = np.zeros(len(Cm), dtype=np.int64)
    st = np.zeros(len(Cm), dtype=np.int64)
    fb = np.zeros(len(Cm), dtype=np.int64)
    dp, lp = C.POINTER(C.c_double), C.POINTER(C.c_long)
    rc = F.L.pdos_family_fetch(F.h, X.ctypes.data_as(dp), None, None, it.ctypes.data_as(lp),
                               st.ctypes.data_as(lp), fb.ctypes.data_as(lp), None)
    assert rc == 0
    np.testing.assert_array_equal(X, full["x"])
    np.testing.assert_array_equal(it, full["iters"])
    assert np.all(np.isfinite(full["y"])) and np.all(np.isfinite(full["s"]))
    # y and s differ between instances (their own c), so no chunk was repeated
    assert len(np.unique(full["y"][:, 0])) > len(Cm) // 2
