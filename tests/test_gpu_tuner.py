"""The workgroup engine's launch shapes on uneven batches: the persistent
work-queue launch against the one-workgroup-per-instance launch (the same
kernels and per-instance arithmetic, so the results are identical bit for
bit), and batches prepared with the launch-shape tuner against the oracle
(the tuner may change the wave batches' entries per lane, hence the
summation order, so the comparison is the usual status, iterations and
x, s, y within 1e-6)."""
import numpy as np
import pytest

import pdos_amd
from oracle import oracle
from pdos_amd import problems

pytestmark = pytest.mark.gpu

RTOL = 1e-6


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)


@pytest.mark.parametrize("plugin", ["indirect", "direct"])
@pytest.mark.parametrize("variant", ["stream", "resident", "image"])
def test_persistent_launch_identical(monkeypatch, plugin, variant):
    # more instances than CUs, so that workgroups take several instances each
    probs = problems.uneven_portfolios(seed=7, count=300, na_range=(10, 60), k_range=(2, 8), EPS_ABS=1e-4,
                                       MAX_ITERS=400)
    monkeypatch.setenv("PDOS_WG_VARIANT", variant)
    print(f"case uneven x300 {plugin} {variant} one workgroup per instance", flush=True)
    monkeypatch.setenv("PDOS_WG_PERSISTENT", "0")
    a = pdos_amd.solve_batch(probs, plugin=plugin)
    print(f"case uneven x300 {plugin} {variant} persistent", flush=True)
    monkeypatch.setenv("PDOS_WG_PERSISTENT", "1")
    b = pdos_amd.solve_batch(probs, plugin=plugin)
    for i, (u, v) in enumerate(zip(a, b)):
        assert u["status"] == v["status"] and u["info"]["iters"] == v["info"]["iters"], i
        for k in "xsy":
            assert np.array_equal(u[k], v[k]), (i, k)
    # and a few against the oracle
    for i in (0, 150, 299):
        exp = oracle.solve(probs[i], plugin, perm=pdos_amd.kkt_order(probs[i]) if plugin == "direct" else None)
        assert b[i]["status"] == exp["status"] and b[i]["info"]["iters"] == exp["iters"], i
        for k in "xsy":
            assert _rel(b[i][k], exp[k]) <= RTOL, (i, k)


@pytest.mark.parametrize("plugin", ["indirect", "direct"])
def test_tuned_batch_vs_oracle(plugin):
    probs = problems.uneven_portfolios(seed=11, count=400, na_range=(20, 200), k_range=(2, 12), EPS_ABS=1e-4,
                                       MAX_ITERS=1000)
    print(f"case uneven x400 {plugin} tuned", flush=True)
    h = pdos_amd.BatchHandle(probs, plugin=plugin, tune=True)
    h.run()
    outs = h.fetch()
    rep = h.tuning()
    h.close()
    assert rep and sum(r["count"] for r in rep) == len(probs)
    for r in rep:
        assert r["variant"] in ("stream", "resident", "image") and r["persistent"] in (0, 1)
        assert r["entries_per_lane"] in (4, 8) and r["threads"] in (256, 512, 1024) and r["candidates"]
        # candidates [variant, persistent, entries per lane, threads, ms]: the
        # fastest, the work-queue launch and wider workgroups only by a 2%
        # margin each
        best = min(c[4] for c in r["candidates"])
        chosen = [c[4] for c in r["candidates"] if c[:4] == [r["variant"], r["persistent"], r["entries_per_lane"],
                                                              r["threads"]]]
        assert chosen and chosen[0] <= best * 1.02 ** 2 * (1 + 1e-9), r
    for i in range(0, len(probs), 50):
        p = probs[i]
        exp = oracle.solve(p, plugin, perm=pdos_amd.kkt_order(p) if plugin == "direct" else None)
        assert outs[i]["status"] == exp["status"] and outs[i]["info"]["iters"] == exp["iters"], i
        for k in "xsy":
            assert _rel(outs[i][k], exp[k]) <= RTOL, (i, k)
    # a second batch of the same sizes takes the cached decisions
    h2 = pdos_amd.BatchHandle(probs[:300], plugin=plugin, tune=True)
    assert all(r["cached"] for r in h2.tuning())
    h2.close()


@pytest.mark.parametrize("plugin", ["indirect", "direct"])
@pytest.mark.parametrize("variant", ["resident", "image"])
@pytest.mark.parametrize("threads", ["512", "1024"])
def test_wide_workgroups_vs_oracle(monkeypatch, plugin, variant, threads):
    """The image variants with 8 or 16 waves per workgroup (k_wg_wide): the
    same rows, batches and levels dealt over more waves, the block sums over
    more wave partials -- against the oracle (status, iterations, x, s, y
    within 1e-6), with the work-queue launch on."""
    probs = problems.uneven_portfolios(seed=13, count=40, na_range=(20, 300), k_range=(2, 12), EPS_ABS=1e-4,
                                       MAX_ITERS=1000)
    monkeypatch.setenv("PDOS_WG_VARIANT", variant)
    monkeypatch.setenv("PDOS_WG_THREADS", threads)
    monkeypatch.setenv("PDOS_WG_PERSISTENT", "1")
    print(f"case uneven x40 {plugin} {variant} {threads} threads", flush=True)
    outs = pdos_amd.solve_batch(probs, plugin=plugin)
    for i in range(0, 40, 5):
        exp = oracle.solve(probs[i], plugin, perm=pdos_amd.kkt_order(probs[i]) if plugin == "direct" else None)
        assert outs[i]["status"] == exp["status"] and outs[i]["info"]["iters"] == exp["iters"], i
        for k in "xsy":
            assert _rel(outs[i][k], exp[k]) <= RTOL, (i, k)
