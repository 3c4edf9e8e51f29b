"""Pins the CPU oracle (oracle/pdos_oracle.c).

1. Bitwise against the compiled reference's outputs (tests/golden/), for both
   plugins; the direct plugin is given the reference's AMD permutation.
2. Against mathematics that does not go through the reference: cone-projection
   identities and brute force, exact linear-system solves, the DIMACS stop
   test, planted optima, ordering independence.
3. Planted mistakes (compiled mutants of the oracle) must fail the pins.
"""
import os
import subprocess

import numpy as np
import pytest

import golden_data as G
from oracle import oracle
from pdos_amd import problems

CASES = G.case_names()


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


@pytest.mark.parametrize("tag", CASES)
@pytest.mark.parametrize("plugin", ["indirect", "direct"])
def test_oracle_bitwise_vs_reference(tag, plugin):
    prob = G.problem(tag)
    exp = G.expected(tag, plugin)
    perm = G.setup_expected(tag)[3] if plugin == "direct" else None
    out = oracle.solve(prob, plugin, perm=perm)
    assert out["status"] == exp["status"]
    assert out["iters"] == exp["iters"]
    for k in ("x", "s", "y"):
        np.testing.assert_array_equal(out[k], exp[k], err_msg=f"{tag}/{plugin}/{k}")
    if exp["iters"] > 0:
        _check_last_row(out, exp)


def _check_last_row(out, exp):
    """The reference prints the last check's residuals with %.3e (pdos.c:306-317)."""
    got = [out["p_res"], out["d_res"], out["p_obj"], out["d_obj"], out["eta"]]
    for g, e in zip(got, exp["last_row"]):
        assert abs(g - e) <= 5.01e-4 * abs(e) + 1e-300, (got, exp["last_row"])


@pytest.mark.parametrize("tag", CASES)
def test_oracle_setup_bitwise(tag):
    prob = G.problem(tag)
    D, E, lam, _ = G.setup_expected(tag)
    d2, e2, l2 = oracle.setup(prob)
    np.testing.assert_array_equal(d2, D)
    np.testing.assert_array_equal(e2, E)
    assert l2 == lam


RUIZ_CASES = G.case_names("ruiz")


@pytest.mark.parametrize("tag", RUIZ_CASES)
@pytest.mark.parametrize("plugin", ["indirect", "direct"])
def test_oracle_ruiz_bitwise_vs_reference(tag, plugin):
    """The Ruiz scaling option against the reference built with its
    ruiz_equil_common.h in place of common.h."""
    prob = G.problem(tag, "ruiz")
    exp = G.expected(tag, plugin, "ruiz")
    perm = G.setup_expected(tag, "ruiz")[3] if plugin == "direct" else None
    out = oracle.solve(prob, plugin, perm=perm, equil="ruiz")
    assert out["status"] == exp["status"]
    assert out["iters"] == exp["iters"]
    for k in ("x", "s", "y"):
        np.testing.assert_array_equal(out[k], exp[k], err_msg=f"{tag}/{plugin}/{k}")
    if exp["iters"] > 0:
        _check_last_row(out, exp)


@pytest.mark.parametrize("tag", RUIZ_CASES)
def test_oracle_ruiz_setup_bitwise(tag):
    prob = G.problem(tag, "ruiz")
    D, E, lam, _ = G.setup_expected(tag, "ruiz")
    d2, e2, l2 = oracle.setup(prob, "ruiz")
    np.testing.assert_array_equal(d2, D)
    np.testing.assert_array_equal(e2, E)
    assert l2 == lam


def test_ruiz_closed_form_on_a_diagonal():
    """A diagonal, no cones, b = c = 0: a sweep maps D <- sqrt(D / (|a| E)) and
    then E <- sqrt(E / (|a| D)), so from D = E = 1 three sweeps give
    D = |a|^(-21/32), E = |a|^(-21/64) (E then x 1e3), and lambda = 1e-6 / 1e-6."""
    from pdos_amd.abi import Problem
    a = np.array([0.01, -0.5, 1.0, 3.0, -250.0])
    n = len(a)
    p = Problem(n, n, np.arange(n + 1), np.arange(n), a, np.zeros(n), np.zeros(n), 0, n, []).with_params(NORMALIZE=1)
    D, E, lam = oracle.setup(p, "ruiz")
    np.testing.assert_allclose(D, np.abs(a) ** (-21 / 32), rtol=1e-14)
    np.testing.assert_allclose(E, 1e3 * np.abs(a) ** (-21 / 64), rtol=1e-14)
    assert lam == 1.0


@pytest.mark.parametrize("tag", [t for t in RUIZ_CASES if G.problem(t, "ruiz").q])
def test_ruiz_cone_rows_share_one_scale(tag):
    """The scaling must not change a cone's shape: one D for all of a cone's rows."""
    prob = G.problem(tag, "ruiz").with_params(NORMALIZE=1)
    D = oracle.setup(prob, "ruiz")[0]
    r0 = prob.f + prob.l
    for q in prob.q:
        assert np.all(D[r0:r0 + q] == D[r0]), (tag, r0, q)
        r0 += q


def test_projcone_bitwise():
    a = G.arrays()
    for tag, meta in G.manifest()["projcone"].items():
        z = a[f"projcone__{tag}__in"]
        out = oracle.proj_cone(z, meta["f"], meta["l"], meta["q"])
        np.testing.assert_array_equal(out, a[f"projcone__{tag}__out"], err_msg=tag)
        # sign of zero too (memset -> +0.0; clamp keeps -0.0)
        assert np.array_equal(np.signbit(out), np.signbit(a[f"projcone__{tag}__out"]))


# ---------------------------------------------------------------- mathematics

def _soc_brute(t, v):
    """Nearest point of the 3-dim SOC {(a, u): ||u|| <= a} by dense search over
    the point itself (if inside), the apex and a fine grid of the boundary."""
    cands = [np.zeros(3)]
    if np.linalg.norm(v) <= t:
        cands.append(np.array([t, *v]))
    for a in np.linspace(0, 8, 1601):
        for th in np.linspace(0, 2 * np.pi, 721):
            cands.append(np.array([a, a * np.cos(th), a * np.sin(th)]))
    cands = np.array(cands)
    d = ((cands - np.array([t, *v])) ** 2).sum(1)
    return cands[np.argmin(d)]


def test_projcone_brute_force_dim3():
    rng = np.random.default_rng(3)
    for _ in range(4):
        z = rng.uniform(-2, 2, size=3)
        p = oracle.proj_cone(z, 0, 0, [3])
        np.testing.assert_allclose(p, _soc_brute(z[0], z[1:]), atol=0.01)


def _kkt_step_exact(prob, x, s, y, lam, c, b, A):
    """(I + A'A) x+ = x - lam c + A'(b - s - lam y) in scaled space."""
    n = A.shape[1]
    rhs = x - lam * c + A.T @ (b - s - lam * y)
    return np.linalg.solve(np.eye(n) + A.T @ A, rhs)


def test_direct_one_step_is_exact_linear_solve():
    """MAX_ITERS=1, EPS=0: x after one update equals the dense KKT solve of
    pdos.c's first step (computed here in numpy from the scaled data)."""
    prob = G.problem("portfolio_n101_steps1")
    out = oracle.solve(prob, "direct")
    D, E, lam = oracle.setup(prob)
    A = prob.dense_A() * D[:, None] * E[None, :]  # E includes the x1e3
    b = prob.b * D
    c = prob.c * E
    x1 = _kkt_step_exact(prob, np.zeros(prob.n), np.zeros(prob.m), np.zeros(prob.m), lam, c, b, A)
    st = b - A @ x1
    alpha = prob.params["ALPHA"]
    st = alpha * st  # s = 0 before the step
    s1 = problems.proj_cone_np(st, prob.f, prob.l, prob.q)
    y1 = (s1 - st) / lam
    np.testing.assert_allclose(out["x"], E * x1, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(out["s"], s1 / D, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(out["y"], D * y1, rtol=1e-9, atol=1e-12)


def test_indirect_converged_cg_matches_direct():
    """With CG run to a tiny tolerance the indirect x-update is the exact
    linear solve, so both plugins follow the same ADMM path."""
    prob = G.problem("portfolio_n101_steps10").with_params(CG_MAX_ITS=500, CG_TOL=1e-12)
    a = oracle.solve(prob, "indirect")
    b = oracle.solve(prob, "direct")
    for k in ("x", "s", "y"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("tag", ["fixture01", "portfolio_n101_norm1", "random_n300_m900", "uneven_1"])
def test_direct_ordering_independent(tag):
    prob = G.problem(tag)
    exp = G.expected(tag, "direct")
    out = oracle.solve(prob, "direct")  # oracle's own minimum-degree ordering
    assert out["status"] == exp["status"]
    assert abs(out["iters"] - exp["iters"]) <= 1
    for k in ("x", "s", "y"):
        np.testing.assert_allclose(out[k], exp[k], rtol=1e-6, atol=1e-8 * (1 + np.abs(exp[k]).max()))


@pytest.mark.parametrize("tag", ["fixture03", "portfolio_n101_norm0", "random_n300_m900"])
def test_solved_meets_dimacs(tag):
    """Recompute the stop test from the returned (x, s, y) in original units."""
    prob = G.problem(tag)
    out = oracle.solve(prob, "direct")
    assert out["status"] == "Solved"
    A = prob.dense_A()
    eps = prob.params["EPS_ABS"]
    # the scaled residuals divided by D / E are the original-unit residuals
    pres = np.linalg.norm(A @ out["x"] + out["s"] - prob.b)
    dres = np.linalg.norm(prob.c + A.T @ out["y"])
    pobj = prob.c @ out["x"]
    dobj = -prob.b @ out["y"]
    slack = 1.0 + 1e-6  # roundoff between the solver's sums and numpy's
    assert pres < eps * (1 + np.linalg.norm(prob.b)) * slack
    assert dres < eps * (1 + np.linalg.norm(prob.c)) * slack
    assert abs(pobj - dobj) < eps * (1 + abs(pobj) + abs(dobj)) * slack
    # iterates lie in the cone
    sp = problems.proj_cone_np(out["s"], prob.f, prob.l, prob.q)
    np.testing.assert_allclose(sp, out["s"], atol=1e-12)


def test_planted_optimum():
    prob = problems.random_feasible(seed=21, n=200, m=600, nnz=4000, f=5, l=295, nsoc=20,
                                    soc_lo=3, soc_hi=25, EPS_ABS=1e-5, MAX_ITERS=20000)
    out = oracle.solve(prob, "direct")
    assert out["status"] == "Solved"
    # strong duality at the planted pair: c'x* = -b'y*; recompute it
    pobj = prob.c @ out["x"]
    dobj = -prob.b @ out["y"]
    assert abs(pobj - dobj) <= 1e-4 * (1 + abs(pobj))


# ---------------------------------------------------------------- planted mistakes

MUTANTS = {
    "CG_SIGN": "the sign of A'(A x + r) in the CG residual",
    "RELAX": "the (1 - alpha) term of the relaxation",
    "SOC_HALF": "the /2 in the SOC projection",
    "DUAL_LAMBDA": "the 1/lambda in the dual update",
    "RES_D": "the D scaling of the primal residual",
    "LDL_ORDER": "the order of the forward solve",
}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistake_is_caught(mutant, tmp_path):
    src = os.path.join(os.path.dirname(oracle.HERE), "oracle", "pdos_oracle.c")
    so = tmp_path / f"mut_{mutant}.so"
    subprocess.run(["gcc", "-O1", "-std=gnu99", "-fPIC", "-shared", "-w", f"-DORACLE_MUTANT_{mutant}",
                    "-o", str(so), src, "-lm"], check=True)
    import ctypes as C
    from pdos_amd.abi import Sol, Data, Cone, sol_to_numpy
    L = C.CDLL(str(so))
    L.oracle_pdos.restype = C.POINTER(Sol)
    L.oracle_pdos.argtypes = [C.POINTER(Data), C.POINTER(Cone), C.c_int, C.POINTER(C.c_long),
                              C.POINTER(oracle.OracleInfo)]
    caught = False
    for tag in ("fixture01", "portfolio_n101_steps2"):
        prob = G.problem(tag)
        for plugin in ("indirect", "direct"):
            exp = G.expected(tag, plugin)
            perm = G.setup_expected(tag)[3] if plugin == "direct" else None
            pp = perm.ctypes.data_as(C.POINTER(C.c_long)) if perm is not None else None
            d, k, keep = prob.to_c()
            info = oracle.OracleInfo()
            r = sol_to_numpy(L.oracle_pdos(C.byref(d), C.byref(k), int(plugin == "direct"), pp,
                                           C.byref(info)))
            r.update(p_res=info.p_res, d_res=info.d_res, p_obj=info.p_obj, d_obj=info.d_obj,
                     eta=info.eta)
            if r["status"] != exp["status"] or any(not np.array_equal(r[v], exp[v]) for v in "xsy"):
                caught = True
            try:
                _check_last_row(r, exp)
            except AssertionError:
                caught = True
    assert caught, f"planted mistake in {MUTANTS[mutant]} went unnoticed"


RUIZ_MUTANTS = {
    "RUIZ_CONE": "giving every row of a cone its first row's maximum",
    "RUIZ_SCAN": "the cone-collapse scan that runs on into the next column",
}


@pytest.mark.parametrize("mutant", sorted(RUIZ_MUTANTS))
def test_planted_ruiz_mistake_is_caught(mutant, tmp_path):
    src = os.path.join(os.path.dirname(oracle.HERE), "oracle", "pdos_oracle.c")
    so = tmp_path / f"mut_{mutant}.so"
    subprocess.run(["gcc", "-O1", "-std=gnu99", "-fPIC", "-shared", "-w", f"-DORACLE_MUTANT_{mutant}",
                    "-o", str(so), src, "-lm"], check=True)
    import ctypes as C
    from pdos_amd.abi import Data, Cone
    L = C.CDLL(str(so))
    L.oracle_setup.argtypes = [C.POINTER(Data), C.POINTER(Cone), C.POINTER(C.c_double),
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.oracle_set_equil(1)
    caught = []
    for tag in RUIZ_CASES:
        prob = G.problem(tag, "ruiz")
        D, E, lam, _ = G.setup_expected(tag, "ruiz")
        d, k, keep = prob.to_c()
        d2, e2, l2 = np.zeros(prob.m), np.zeros(prob.n), C.c_double()
        L.oracle_setup(C.byref(d), C.byref(k), d2.ctypes.data_as(C.POINTER(C.c_double)),
                       e2.ctypes.data_as(C.POINTER(C.c_double)), C.byref(l2))
        if not (np.array_equal(d2, D) and np.array_equal(e2, E) and l2.value == lam):
            caught.append(tag)
    assert caught, f"planted mistake in {RUIZ_MUTANTS[mutant]} went unnoticed"
