"""Setup of a batch's instances on the device (batch_setup.hip: NORMALIZE
equilibration, lambda, the x1e3, the scaled values of A in row and column
order; common.h:69-215, cs.c:146-168) against the reference's initWork
scalings stored with the golden cases, and against the same batch set up on
the host (PDOS_HOST_SETUP=1): equal bit for bit, so the solves are too."""
import dataclasses

import numpy as np
import pytest

import golden_data as G
import pdos_amd
from pdos_amd import problems

pytestmark = pytest.mark.gpu


def _batch_probs():
    probs = [G.problem(t) for t in G.case_names()]
    for norm in (0, 1):
        probs += problems.uneven_portfolios(seed=40 + norm, count=24, na_range=(20, 400), k_range=(3, 40),
                                            EPS_ABS=1e-3, MAX_ITERS=400, NORMALIZE=norm)
    return probs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_setup_bitwise_vs_reference(precision):
    """Every golden case in one indirect batch: D, E (with the x1e3) and lambda
    formed on the device equal the reference's bit for bit (duplicate and
    unsorted CSC entries, empty rows and columns, NORMALIZE 0 and 1, b = 0)."""
    tags = G.case_names()
    probs = [G.problem(t) for t in tags]
    print(f"device setup vs reference scalings, {len(tags)} cases, {precision}", flush=True)
    h = pdos_amd.BatchHandle(probs, plugin="indirect", precision=precision)
    try:
        for i, t in enumerate(tags):
            D, E, lam, dev = h.scales(i)
            D0, E0, lam0, _ = G.setup_expected(t)
            assert dev, t
            np.testing.assert_array_equal(D, D0, err_msg=f"{t} D")
            np.testing.assert_array_equal(E, E0, err_msg=f"{t} E")
            assert lam == lam0, t
    finally:
        h.close()


@pytest.mark.parametrize("variant", ["", "stream"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_setup_matches_host_setup(monkeypatch, precision, variant):
    """The golden cases and 48 uneven portfolios (NORMALIZE 0 and 1) solved in
    one batch with the setup on the device and again with it on the host:
    status, iterations, CG steps, lambda and x, s, y equal bit for bit, with
    the automatic kernel variants and with the streaming variant (whose
    matrix copies the device also fills)."""
    probs = _batch_probs()
    if variant:
        monkeypatch.setenv("PDOS_WG_VARIANT", variant)
    print(f"device vs host setup, {len(probs)} instances, {precision} {variant or 'auto'}", flush=True)
    dev = pdos_amd.solve_batch(probs, plugin="indirect", precision=precision)
    monkeypatch.setenv("PDOS_HOST_SETUP", "1")
    host = pdos_amd.solve_batch(probs, plugin="indirect", precision=precision)
    for i, (a, b) in enumerate(zip(dev, host)):
        assert a["status"] == b["status"], i
        for k in ("iters", "cg_iters", "lam"):
            assert a["info"][k] == b["info"][k], (i, k)
        for k in ("x", "s", "y"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"instance {i} {k}")


def test_device_setup_shared_matrix_groups(monkeypatch):
    """Instances with the same A (a group set up once) but their own c and
    warm starts: per-instance lambda and stop-test scales, formed in place on
    the device, equal the host setup's, and the solves equal a host-setup
    batch's bit for bit."""
    base = problems.portfolio(seed=3, na=60, k=6, NORMALIZE=1, EPS_ABS=1e-4)
    rng = np.random.default_rng(0)
    probs = []
    for t in range(12):
        p = dataclasses.replace(base, c=base.c + 0.1 * rng.standard_normal(base.n))
        if t % 3 == 0:
            p = p.with_warm(rng.standard_normal(p.n), rng.standard_normal(p.m), np.abs(rng.standard_normal(p.m)))
        probs.append(p)
    print("shared-matrix batch, device setup", flush=True)
    h = pdos_amd.BatchHandle(probs, plugin="indirect", precision="fp64")
    try:
        for i, p in enumerate(probs):
            D, E, lam, dev = h.scales(i)
            d2, e2, l2 = pdos_amd.setup_host(p, "reference")
            assert dev
            np.testing.assert_array_equal(D, d2)
            np.testing.assert_array_equal(E, e2)
            assert lam == l2, i
        h.run()
        outs = h.fetch()
    finally:
        h.close()
    monkeypatch.setenv("PDOS_HOST_SETUP", "1")
    host = pdos_amd.solve_batch(probs, plugin="indirect", precision="fp64")
    for o, r in zip(outs, host):
        assert o["info"]["iters"] == r["info"]["iters"]
        for k in ("x", "s", "y"):
            np.testing.assert_array_equal(o[k], r[k])
