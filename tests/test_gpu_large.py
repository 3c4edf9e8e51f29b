"""Instances past one workgroup's LDS: the whole-GPU engine with both plugins
(the direct plugin's level-scheduled triangular sweeps, direct/private.c:
16-56, ldl.c:326-428), through pdos_ex, the reference's own pdos() entry
point and pdos_batch's per-instance route, against the oracle.

The oracle's direct plugin is given the library's own ordering
(pdos_kkt_order), so both factor the same permuted matrix; FP64 bar: same
status and iteration count, x, s, y within 1e-6 relative."""
import ctypes as C

import numpy as np
import pytest

import golden_data as G
import pdos_amd
from oracle import oracle
from pdos_amd import problems

pytestmark = pytest.mark.gpu

RTOL = 1e-6
_CASES = {}


def _case(name):
    if name not in _CASES:
        if name == "portfolio_na3000":
            p = problems.portfolio(seed=5, na=3000, k=20, EPS_ABS=1e-3)
        elif name == "portfolio_na5000":  # n + m = 15,023: z does not fit the sweep workgroup's LDS
            p = problems.portfolio(seed=6, na=5000, k=30, EPS_ABS=1e-3)
        else:
            p = problems.random_feasible(seed=7, n=2000, m=6000, nnz=14000, f=50, l=2950, nsoc=60, soc_lo=10,
                                         soc_hi=90, EPS_ABS=1e-3, MAX_ITERS=2000)
        _CASES[name] = p
    return _CASES[name]


def _expected(p, plugin):
    return oracle.solve(p, plugin, perm=pdos_amd.kkt_order(p) if plugin == "direct" else None)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)


def _check(out, exp, tag):
    assert out["status"] == exp["status"], tag
    assert out["info"]["iters"] == exp["iters"], (tag, out["info"]["iters"], exp["iters"])
    for k in ("x", "s", "y"):
        assert _rel(out[k], exp[k]) <= RTOL, (tag, k, _rel(out[k], exp[k]))


@pytest.mark.parametrize("name", ["portfolio_na3000", "random_n2000_m6000"])
@pytest.mark.parametrize("plugin", ["direct", "indirect"])
def test_large_vs_oracle(name, plugin):
    p = _case(name)
    print(f"case {name} {plugin} grid", flush=True)
    exp = _expected(p, plugin)
    out = pdos_amd.solve_problem(p, plugin=plugin, precision="fp64")
    assert out["info"]["engine"] == 2  # past the workgroup engine's LDS: the whole-GPU engine
    _check(out, exp, f"{name}/{plugin}")
    if plugin == "indirect":
        # a CG stop test within roundoff of CG_TOL (indirect/private.c:86) may
        # take one step more or less under a different summation order; the
        # iterates stay within the bar above
        assert abs(out["info"]["cg_iters"] - exp["cg_iters"]) <= max(2, exp["cg_iters"] // 1000)


def test_large_direct_z_in_global_memory():
    """n + m = 15,023: the sweep workgroup keeps z in global memory (it does
    not fit the LDS beside the staged rows)."""
    p = _case("portfolio_na5000")
    print("case portfolio_na5000 direct grid", flush=True)
    _check(pdos_amd.solve_problem(p, plugin="direct", precision="fp64"), _expected(p, "direct"), "na5000")


def test_large_direct_wide_levels(monkeypatch):
    """Levels of 64 rows or more as whole-GPU launches (k_sweep_wide) between
    the staged runs: the same solution."""
    monkeypatch.setenv("PDOS_SWEEP_WIDE_ROWS", "64")
    for name in ("portfolio_na3000", "random_n2000_m6000"):
        p = _case(name)
        print(f"case {name} direct grid wide levels", flush=True)
        _check(pdos_amd.solve_problem(p, plugin="direct", precision="fp64"), _expected(p, "direct"),
               f"{name}/wide")


def test_large_direct_through_pdos_abi():
    """libpdosdir_hip.so's pdos() (pdos.h:85) solves it instead of returning
    NULL."""
    p = _case("portfolio_na3000")
    L = pdos_amd.lib("direct")
    d, k, keep = p.to_c()
    print("case portfolio_na3000 pdos()", flush=True)
    sp = L.pdos(C.byref(d), C.byref(k))
    assert sp, L.pdos_last_error().decode()
    out = pdos_amd._take_sol(L, sp)
    exp = _expected(p, "direct")
    assert out["status"] == exp["status"]
    for key in ("x", "s", "y"):
        assert _rel(out[key], exp[key]) <= RTOL, key


@pytest.mark.parametrize("plugin", ["direct", "indirect"])
def test_batch_routes_oversized_member(plugin):
    """A pdos_batch with one member past the LDS: that member runs on the
    whole-GPU engine, the rest in the workgroup launch; every member matches."""
    big = problems.portfolio(seed=8, na=2000, k=10, EPS_ABS=1e-3)
    probs = [G.problem("fixture01"), big, G.problem("portfolio_n101_norm0"), G.problem("random_n300_m900")]
    print(f"case batch with an oversized member {plugin}", flush=True)
    outs = pdos_amd.solve_batch(probs, plugin=plugin)
    assert [o["info"]["engine"] for o in outs] == [1, 2, 1, 1]
    for tag, o in zip(["fixture01", None, "portfolio_n101_norm0", "random_n300_m900"], outs):
        exp = G.expected(tag, plugin) if tag else _expected(big, plugin)
        _check(o, exp, f"batch/{tag or 'big'}/{plugin}")


def test_large_direct_mixed_precision():
    """Mixed mode (binary32 iterates, binary64 sweeps and norms): Solved, and
    the point meets the stop test recomputed in binary64."""
    import scipy.sparse as sp
    p = _case("portfolio_na3000")
    print("case portfolio_na3000 direct mixed grid", flush=True)
    out = pdos_amd.solve_problem(p, plugin="direct", precision="mixed")
    assert out["status"] == "Solved"
    A = sp.csc_matrix((p.Ax, p.Ai, p.Ap), shape=(p.m, p.n))
    eps = p.params["EPS_ABS"]
    pres = np.linalg.norm(A @ out["x"] + out["s"] - p.b)
    dres = np.linalg.norm(A.T @ out["y"] + p.c)
    po, do = p.c @ out["x"], -p.b @ out["y"]
    assert pres < eps * (1 + np.linalg.norm(p.b)) * 1.01
    assert dres < eps * (1 + np.linalg.norm(p.c)) * 1.01
    assert abs(po - do) < eps * (1 + abs(po) + abs(do)) * 1.01
