"""GPU parity: the HIP path against the compiled reference's golden outputs.

FP64 mode: same status and iteration count as the reference, and x, s, y
within 1e-6 relative (norm-wise; the bar the solver's FP64 mode promises).
Mixed mode (binary32 bulk, binary64 norms): status Solved and the returned
point within EPS_ABS-scaled distance of the FP64 optimum.
"""
import ctypes as C
import json
import os
import re
import numpy as np
import pytest

import golden_data as G
import pdos_amd
from pdos_amd import problems
from oracle import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-6


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)


def _check_fp64(out, exp, tag):
    assert out["status"] == exp["status"], tag
    assert out["info"]["iters"] == exp["iters"], (tag, out["info"]["iters"], exp["iters"])
    for k in ("x", "s", "y"):
        assert _rel(out[k], exp[k]) <= RTOL, (tag, k, _rel(out[k], exp[k]))


@pytest.mark.parametrize("tag", G.case_names())
@pytest.mark.parametrize("plugin", ["indirect", "direct"])
def test_wg_fp64_vs_reference(tag, plugin):
    print(f"case {tag} {plugin}", flush=True)
    out = pdos_amd.solve_problem(G.problem(tag), plugin=plugin, precision="fp64", engine="wg")
    _check_fp64(out, G.expected(tag, plugin), f"{tag}/{plugin}")


@pytest.mark.parametrize("tag", ["random_n300_m900", "portfolio_n101_norm1", "fixture03",
                                 "portfolio_n101_steps10", "portfolio_n101_cg1", "jumbled_n120_m300",
                                 "soc_only_n80_m200", "b_zero_n60_m150"])
def test_grid_fp64_vs_reference(tag):
    print(f"case {tag} grid", flush=True)
    out = pdos_amd.solve_problem(G.problem(tag), plugin="indirect", precision="fp64", engine="grid")
    _check_fp64(out, G.expected(tag, "indirect"), f"{tag}/grid")


@pytest.mark.parametrize("tag", G.case_names("ruiz"))
@pytest.mark.parametrize("plugin", ["indirect", "direct"])
def test_wg_fp64_ruiz_vs_reference(tag, plugin):
    """The Ruiz scaling option against the reference built with its
    ruiz_equil_common.h in place of common.h."""
    print(f"case {tag} {plugin} ruiz", flush=True)
    out = pdos_amd.solve_problem(G.problem(tag, "ruiz"), plugin=plugin, precision="fp64", engine="wg",
                                 equil="ruiz")
    _check_fp64(out, G.expected(tag, plugin, "ruiz"), f"{tag}/{plugin}/ruiz")


@pytest.mark.parametrize("tag", ["random_n300_m900_norm1", "portfolio_n101_norm1"])
def test_grid_fp64_ruiz_vs_reference(tag):
    print(f"case {tag} grid ruiz", flush=True)
    out = pdos_amd.solve_problem(G.problem(tag, "ruiz"), plugin="indirect", precision="fp64", engine="grid",
                                 equil="ruiz")
    _check_fp64(out, G.expected(tag, "indirect", "ruiz"), f"{tag}/grid/ruiz")


def test_pdos_abi_takes_ruiz_from_environment(monkeypatch):
    """pdos() (the reference's own entry point) reads PDOS_EQUIL at each call."""
    tag = "portfolio_n101_norm1"
    prob = G.problem(tag, "ruiz")
    L = pdos_amd.lib("indirect")
    print(f"case {tag} pdos() ruiz", flush=True)
    for env, variant in (("ruiz", "ruiz"), ("", "")):
        monkeypatch.setenv("PDOS_EQUIL", env)
        d, k, keep = prob.to_c()
        sp = L.pdos(C.byref(d), C.byref(k))
        assert sp, L.pdos_last_error().decode()
        out = pdos_amd._take_sol(L, sp)
        exp = G.expected(tag, "indirect", variant)
        assert out["status"] == exp["status"]
        for key in ("x", "s", "y"):
            assert _rel(out[key], exp[key]) <= RTOL, (variant, key)


def test_grid_long_rows_vs_oracle():
    """A full row (6,000 entries) and a full column (9,000 entries): both
    exceed one SpMV row block and take the whole-workgroup row path.  This
    instance is well conditioned for the bar: solving it in the oracle with
    the columns permuted (other summation orders) moves x by ~2e-8 and keeps
    the iteration count."""
    prob = problems.random_feasible(seed=6, n=6000, m=9000, nnz=30000, f=100, l=4900, nsoc=40,
                                    soc_lo=10, soc_hi=190, dense_rows=1, dense_cols=1,
                                    EPS_ABS=1e-3, MAX_ITERS=2000)
    print("case random_long_rows grid", flush=True)
    exp = oracle.solve(prob, "indirect")
    out = pdos_amd.solve_problem(prob, plugin="indirect", precision="fp64", engine="grid")
    _check_fp64(out, exp, "long_rows/grid")
    assert out["info"]["cg_iters"] == exp["cg_iters"]


@pytest.mark.parametrize("engine", ["wg", "grid"])
@pytest.mark.parametrize("tag", ["fixture01", "portfolio_n101_norm0", "random_n300_m900", "uneven_1",
                                 "jumbled_n120_m300"])
def test_mixed_precision_within_eps(tag, engine):
    print(f"case {tag} mixed {engine}", flush=True)
    prob = G.problem(tag)
    exp = G.expected(tag, "indirect")
    out = pdos_amd.solve_problem(prob, plugin="indirect", precision="mixed", engine=engine)
    assert out["status"] == "Solved"
    eps = prob.params["EPS_ABS"]
    # objective agrees to the stop test's relative gap tolerance
    pobj, pobj_ref = prob.c @ out["x"], prob.c @ exp["x"]
    assert abs(pobj - pobj_ref) <= 10 * eps * (1 + abs(pobj_ref))
    # and the returned point meets the stop test in binary64 arithmetic
    A = prob.dense_A()
    assert np.linalg.norm(A @ out["x"] + out["s"] - prob.b) < 1.01 * eps * (1 + np.linalg.norm(prob.b))
    assert np.linalg.norm(prob.c + A.T @ out["y"]) < 1.01 * eps * (1 + np.linalg.norm(prob.c))


def test_batch_matches_single():
    tags = ["fixture01", "fixture02", "uneven_1", "portfolio_n101_norm1", "random_n300_m900"]
    probs = [G.problem(t) for t in tags]
    print("batch of", len(probs), flush=True)
    outs = pdos_amd.solve_batch(probs, plugin="indirect", precision="fp64")
    for t, o in zip(tags, outs):
        _check_fp64(o, G.expected(t, "indirect"), f"batch/{t}")
    outs = pdos_amd.solve_batch(probs, plugin="direct", precision="fp64")
    for t, o in zip(tags, outs):
        _check_fp64(o, G.expected(t, "direct"), f"batch-direct/{t}")


@pytest.mark.parametrize("variant", ["", "ruiz"])
@pytest.mark.parametrize("plugin", ["indirect", "direct"])
def test_batch_of_every_golden_case(plugin, variant):
    """Every golden case in one pdos_batch call: members with their own
    parameters (MAX_ITERS 0, EPS 0, ALPHA 1.8, NORMALIZE, CG_MAX_ITS 1), warm
    starts, cone mixes and stored-A quirks each match the reference's output
    for that case alone (status, iterations, x, s, y at 1e-6)."""
    tags = G.case_names(variant)
    probs = [G.problem(t, variant) for t in tags]
    print(f"batch of every golden case {plugin} {variant or 'reference'}", flush=True)
    outs = pdos_amd.solve_batch(probs, plugin=plugin, precision="fp64", equil=variant or "reference")
    for t, o in zip(tags, outs):
        _check_fp64(o, G.expected(t, plugin, variant), f"batch-all/{plugin}/{t}")


def test_batch_uneven_vs_oracle():
    probs = problems.uneven_portfolios(seed=31, count=24, na_range=(20, 300), k_range=(3, 30),
                                       EPS_ABS=1e-3, MAX_ITERS=4000)
    print("uneven batch", flush=True)
    outs = pdos_amd.solve_batch(probs, plugin="indirect", precision="fp64")
    for p, o in zip(probs, outs):
        ref = oracle.solve(p, "indirect")
        assert o["status"] == ref["status"]
        assert o["info"]["iters"] == ref["iters"]
        for k in ("x", "s", "y"):
            assert _rel(o[k], ref[k]) <= RTOL


def test_maxiters_zero_returns_warm_start():
    base = G.problem("portfolio_n101_norm1")
    rng = np.random.default_rng(5)
    prob = base.with_warm(x0=rng.standard_normal(base.n), y0=rng.standard_normal(base.m),
                          s0=rng.standard_normal(base.m)).with_params(MAX_ITERS=0)
    out = pdos_amd.solve_problem(prob, engine="wg")
    assert out["status"] == "Indeterminate"
    np.testing.assert_allclose(out["x"], prob.x0, rtol=1e-15)
    np.testing.assert_allclose(out["y"], prob.y0, rtol=1e-15)
    np.testing.assert_allclose(out["s"], prob.s0, rtol=1e-15)


@pytest.mark.parametrize("plugin", ["indirect", "direct"])
def test_demo_cli_on_reference_fixture(plugin):
    """The demo CLI (run_pdos.c semantics) on the reference's own data file."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(pdos_amd.__file__), f"demo_{plugin}_hip")
    path = os.path.join(G.GOLDEN, "data", "portfolio_test_01_10x100")
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "status: Solved" in r.stdout
    # VERBOSE is 1 in the file: the summary row carries the iteration count
    assert any(line.strip().startswith("20 |") for line in r.stdout.splitlines()), r.stdout


def test_device_copy_kernel():
    """pdos_device_copy (the counter-calibration copy): bulk 16-byte path and
    a 12-byte tail."""
    src = np.random.default_rng(3).standard_normal(1003).astype(np.float32)
    a = pdos_amd.DeviceArray(src)
    b = pdos_amd.DeviceArray(shape=(1003,))
    b.copy_from(a)
    np.testing.assert_array_equal(b.numpy(), src)


_VERBOSE_LOGS = os.path.join(G.GOLDEN, "verbose_logs.json")
_NUMBER = re.compile(r"[-+]?\d+\.\d+(?:e[+-]\d+)?|\d+")


def _capture_fd1(fn):
    import ctypes
    import tempfile
    libc = ctypes.CDLL(None)
    with tempfile.TemporaryFile() as tmp:
        libc.fflush(None)
        saved = os.dup(1)
        os.dup2(tmp.fileno(), 1)
        try:
            fn()
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return tmp.read().decode()


@pytest.mark.parametrize("key", sorted(json.load(open(_VERBOSE_LOGS))))
def test_verbose_log_matches_reference(key):
    """pdos() with VERBOSE 1 prints the reference's log (pdos.c:79-127,
    306-335; direct/private.c:157; indirect/private.c:100): the same lines in
    the same layout -- banner, factorization time (direct), lambda, header and
    rule, a row for every tenth check, a CG line per update (indirect), the
    final row, the solve time.  Fields are fixed width, so every line has the
    reference's length; numbers agree to the printed digits (one unit of the
    last digit), counts exactly.  Times and the banner's credit line differ."""
    tag, plugin = key.split("/")
    ref = json.load(open(_VERBOSE_LOGS))[key].splitlines()
    prob = G.problem(tag).with_params(VERBOSE=1)
    print(f"verbose log {key}", flush=True)
    got = _capture_fd1(lambda: pdos_amd.solve_problem(prob, plugin=plugin, precision="fp64", engine="wg"))
    got = got.splitlines()
    assert len(got) == len(ref), (len(got), len(ref), got[:12])
    for k, (g, r) in enumerate(zip(got, ref)):
        if k == 2:  # the banner's second line: the build's own
            continue
        if r.startswith(("Factorization time:", "Total solve time is")):
            assert g.split(":")[0].split(" is")[0] == r.split(":")[0].split(" is")[0], (k, g, r)
            continue
        assert len(g) == len(r), (k, g, r)
        assert _NUMBER.sub("#", g).split() == _NUMBER.sub("#", r).split(), (k, g, r)
        for a, b in zip(_NUMBER.findall(g), _NUMBER.findall(r)):
            if "." not in b:
                assert a == b, (k, g, r)
            else:
                digits = 1e-3 if "e" in b else 1e-6  # %.3e: 4 significant digits; %4f: 6 decimals
                scale = abs(float(b)) if "e" in b else 1.0
                assert abs(float(a) - float(b)) <= 1.01 * digits * scale + 1e-300, (k, g, r)


@pytest.mark.parametrize("plugin", ["indirect", "direct"])
def test_python_binding_solve(plugin):
    """pdos_amd.solve (the reference binding's solve(c, A, b, dims, opts),
    python/pdosmodule.c:27-344): the keys it returns, the fixture's own
    options against the reference's golden outputs, and the binding's
    defaults (:67-74, NORMALIZE 0) against the oracle."""
    import scipy.sparse as sp
    tag = "fixture01"
    prob = G.problem(tag)
    A = sp.csc_matrix((prob.Ax, prob.Ai, prob.Ap), shape=(prob.m, prob.n))
    dims = {"f": prob.f, "l": prob.l, "q": list(prob.q)}
    print(f"case {tag} binding {plugin}", flush=True)
    out = pdos_amd.solve(prob.c, A, prob.b, dims, opts=dict(prob.params, VERBOSE=0), plugin=plugin)
    assert {"x", "s", "y", "status"} <= set(out)
    _check_fp64(out, G.expected(tag, plugin), f"{tag}/binding/{plugin}")
    out = pdos_amd.solve(prob.c, A, prob.b, dims, plugin=plugin)
    exp = oracle.solve(prob.with_params(MAX_ITERS=2000, CG_MAX_ITS=20, EPS_ABS=5e-3, ALPHA=1.0, CG_TOL=1e-4,
                                        VERBOSE=0, NORMALIZE=0), plugin)
    _check_fp64(out, exp, f"{tag}/binding-defaults/{plugin}")


def test_batch_verbose_keeps_no_history():
    """VERBOSE instances in a batch: the batch path prints nothing, so it keeps
    no residual history -- the device memory of a prepared batch does not
    depend on MAX_ITERS."""
    base = problems.portfolio(seed=0, VERBOSE=1)
    sizes = []
    for mi in (2000, 2_000_000):
        probs = [base.with_params(MAX_ITERS=mi, VERBOSE=1) for _ in range(8)]
        print(f"case verbose batch MAX_ITERS {mi}", flush=True)
        h = pdos_amd.BatchHandle(probs, plugin="indirect")
        sizes.append(h.device_bytes())
        h.close()
    assert sizes[0] == sizes[1] > 0, sizes


@pytest.mark.parametrize("variant", ["stream", "resident", "image"])
@pytest.mark.parametrize("plugin", ["indirect", "direct"])
def test_wg_streaming_and_resident_variants(monkeypatch, variant, plugin):
    """The workgroup engine's three variants -- matrices streamed from global
    memory (the streaming body), the wave-batched body with its image in LDS
    (the default where it fits) and the same with the image in
    device memory -- each against the reference on cases covering long rows,
    jumbled columns, empty rows and warm starts (PDOS_WG_VARIANT forces one)."""
    monkeypatch.setenv("PDOS_WG_VARIANT", variant)
    resident = variant
    for tag in ["fixture01", "portfolio_n101_norm1", "portfolio_n101_warm", "jumbled_n120_m300",
                "random_n300_m900", "dense_m96_n24", "uneven_1", "b_zero_n60_m150"]:
        print(f"case {tag} {plugin} resident={resident}", flush=True)
        out = pdos_amd.solve_problem(G.problem(tag), plugin=plugin, precision="fp64", engine="wg")
        _check_fp64(out, G.expected(tag, plugin), f"{tag}/{plugin}/res{resident}")
