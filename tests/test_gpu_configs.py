"""Parity at the benchmark configurations' own sizes (the n=101 portfolio is
covered by the golden cases):

  random SOCP, n=100,000, m=300,000, nnz ~5M (grid engine, FP64 and FP32
      bulk): Solved, the stop test recomputed in binary64, and the optimum
      against the planted one;
  dense shared A, 8,192 x 2,048, mixed precision (bf16 / fp32 products):
      64 instances, each Solved, the stop test in binary64, the objective
      against the planted optimum; and the dense mixed engine instance by
      instance against the oracle on a reduced dense family;
  uneven portfolio batch with n_a in [800, 1000]: 16 instances in one
      pdos_batch, FP64, both plugins against the oracle (status, iterations,
      x, s, y within 1e-6), and the FP32 bulk mode against the stop test;
  portfolio family n_a=500, k=20: held-out instances, the FP64 family engine
      against the oracle's direct plugin; mixed against the stop test; the
      learned warm start's gain on this family.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import pdos_amd
from oracle import oracle
from pdos_amd import problems

pytestmark = pytest.mark.gpu

RTOL = 1e-6


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)


def _stop_ok(A, b, c, x, y, s, eps, slack=1.0):
    """The reference's stop test (pdos.c:102-104) in binary64, original units."""
    pres = np.linalg.norm(A @ x + s - b)
    dres = np.linalg.norm(A.T @ y + c)
    po, do = c @ x, -b @ y
    return (pres < slack * eps * (1 + np.linalg.norm(b)) and dres < slack * eps * (1 + np.linalg.norm(c))
            and abs(po - do) < slack * eps * (1 + abs(po) + abs(do)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_random_full_size_planted_optimum(precision):
    p = problems.random_feasible(seed=1, EPS_ABS=1e-3, MAX_ITERS=2000)
    print(f"case random n=100000 m=300000 {precision}", flush=True)
    out = pdos_amd.solve_problem(p, plugin="indirect", precision=precision)
    assert out["info"]["engine"] == 2
    assert out["status"] == "Solved"
    A = sp.csc_matrix((p.Ax, p.Ai, p.Ap), shape=(p.m, p.n))
    # the returned point meets the stop test in binary64 (1% slack for the
    # binary32 iterates' rounding in mixed mode)
    assert _stop_ok(A, p.b, p.c, out["x"], out["y"], out["s"], 1e-3, 1.0 if precision == "fp64" else 1.01)
    # the planted optimal value c'x* = -b'y* (problems.random_feasible)
    po = p.c @ out["x"]
    assert abs(po - p.planted_opt) <= 20 * 1e-3 * (1 + abs(p.planted_opt)), (po, p.planted_opt)


def test_dense_full_size_mixed():
    from pdos_amd.family import Family
    dense = problems.DenseFamily(seed=2, EPS_ABS=1e-3, MAX_ITERS=2000)
    B, Cm, opt = dense.sample(seed=1002, count=64)
    print("case dense 8192x2048 mixed, 64 instances", flush=True)
    F = Family(dense.instance(B[0], Cm[0]), precision="mixed")
    F.load(Cm, B)
    F.run()
    out = F.fetch()
    assert (out["status"] == "Solved").all()
    for i in range(64):
        assert _stop_ok(dense.A, B[i], Cm[i], out["x"][i], out["y"][i], out["s"][i], 1e-3, 1.01), i
    po = np.einsum("ij,ij->i", Cm, out["x"])
    assert np.all(np.abs(po - opt) <= 20 * 1e-3 * (1 + np.abs(opt)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dense_family_vs_oracle_reduced(precision):
    """The dense-A family engine on a reduced dense family (512 x 128, 8
    instances): FP64 instance by instance against the oracle (status,
    iterations, x, s, y within 1e-6); mixed against the stop test and the
    oracle's objective."""
    from pdos_amd.family import Family
    dense = problems.DenseFamily(seed=12, m=512, n=128, l=384, nsoc=4, qdim=32, EPS_ABS=1e-3, MAX_ITERS=2000)
    B, Cm, opt = dense.sample(seed=3, count=8)
    print(f"case dense 512x128 family {precision}", flush=True)
    F = Family(dense.instance(B[0], Cm[0]), precision=precision)
    assert F.info()["dense"]
    F.load(Cm, B)
    F.run()
    out = F.fetch()
    for i in range(8):
        exp = oracle.solve(dense.instance(B[i], Cm[i]), "direct")
        got = {k: out[k][i] for k in "xys"}
        if precision == "fp64":
            assert out["status"][i] == exp["status"] and out["iters"][i] == exp["iters"], i
            for k in "xys":
                assert _rel(got[k], exp[k]) <= RTOL, (i, k)
        else:
            assert out["status"][i] == "Solved"
            assert _stop_ok(dense.A, B[i], Cm[i], got["x"], got["y"], got["s"], 1e-3, 1.01), i
            pe = Cm[i] @ exp["x"]
            assert abs(Cm[i] @ got["x"] - pe) <= 20 * 1e-3 * (1 + abs(pe)), i


@pytest.mark.parametrize("plugin", ["direct", "indirect"])
def test_uneven_large_members_vs_oracle(plugin):
    probs = problems.uneven_portfolios(seed=31, count=16, na_range=(800, 1000), EPS_ABS=1e-3, MAX_ITERS=2000)
    print(f"case uneven n_a 800-1000 x16 {plugin} fp64", flush=True)
    outs = pdos_amd.solve_batch(probs, plugin=plugin)
    for i, (p, o) in enumerate(zip(probs, outs)):
        exp = oracle.solve(p, plugin, perm=pdos_amd.kkt_order(p) if plugin == "direct" else None)
        assert o["status"] == exp["status"] and o["info"]["iters"] == exp["iters"], (i, o["info"]["iters"],
                                                                                     exp["iters"])
        for k in "xsy":
            assert _rel(o[k], exp[k]) <= RTOL, (i, k, _rel(o[k], exp[k]))


def test_uneven_large_members_fp32():
    probs = problems.uneven_portfolios(seed=31, count=16, na_range=(800, 1000), EPS_ABS=1e-3, MAX_ITERS=2000)
    print("case uneven n_a 800-1000 x16 indirect fp32", flush=True)
    outs = pdos_amd.solve_batch(probs, plugin="indirect", precision="mixed")
    solved = 0
    for p, o in zip(probs, outs):
        if o["status"] != "Solved":
            continue
        solved += 1
        A = sp.csc_matrix((p.Ax, p.Ai, p.Ap), shape=(p.m, p.n))
        assert _stop_ok(A, p.b, p.c, o["x"], o["y"], o["s"], 1e-3, 1.01), p.name
    assert solved >= 10


def test_portfolio_family_full_size_vs_oracle():
    """n_a = 500, k = 20, held-out instances: the FP64 family engine (shared
    A, exact (I + A'A)^-1 x-update) against the oracle's direct plugin; the
    mixed engine against the stop test in binary64."""
    from pdos_amd.family import Family
    fam = problems.PortfolioFamily(seed=4, na=500, k=20, EPS_ABS=1e-3, MAX_ITERS=2000)
    th, Cm = fam.sample(seed=1004, count=24)
    A = sp.csc_matrix((fam.Ax, fam.Ai, fam.Ap), shape=(fam.m, fam.n))
    for precision in ("fp64", "mixed"):
        print(f"case portfolio family n_a=500 {precision}", flush=True)
        F = Family(fam.instance(Cm[0]), precision=precision)
        F.load(Cm)
        F.run()
        out = F.fetch()
        for i in range(24):
            got = {k: out[k][i] for k in "xys"}
            if precision == "fp64":
                exp = oracle.solve(fam.instance(Cm[i]), "direct")
                assert out["status"][i] == exp["status"] and out["iters"][i] == exp["iters"], i
                for k in "xys":
                    assert _rel(got[k], exp[k]) <= RTOL, (i, k)
            elif out["status"][i] == "Solved":
                assert _stop_ok(A, fam.b, Cm[i], got["x"], got["y"], got["s"], 1e-3, 1.01), i
