"""The learned warm start end to end on a small family: trained on the
solver's own solutions, it must cut the mean iteration count of held-out
instances against the cold start, and every returned solution must still be
one the stop test accepts (ADMM judges the prediction)."""
import numpy as np
import pytest

from pdos_amd import DeviceArray, problems
from pdos_amd.family import Family
from pdos_amd.warmstart import LearnedWarmStart, portfolio_features

pytestmark = pytest.mark.gpu


def test_learned_warm_start_gains_over_cold():
    fam = problems.PortfolioFamily(seed=4, na=60, k=6, EPS_ABS=1e-3)
    th_tr, C_tr = fam.sample(seed=11, count=4096)
    th_te, C_te = fam.sample(seed=12, count=512)
    print("train set", flush=True)
    F = Family(fam.instance(C_tr[0]), precision="mixed")
    F.load(C_tr)
    F.run()
    tr = F.fetch()
    keep = tr["status"] == "Solved"
    T = np.concatenate([tr["x"], tr["y"], tr["s"]], axis=1)[keep]
    ws = LearnedWarmStart(th_tr.shape[1], T.shape[1], hidden=256, seed=1)
    loss = ws.fit(portfolio_features(th_tr[keep]), T, epochs=40, batch=256, lr=1e-3)
    print("train loss", loss, flush=True)
    F.load(C_te)
    F.run()
    cold = F.fetch()
    dF = DeviceArray(portfolio_features(th_te))
    dW = DeviceArray(shape=(len(C_te), T.shape[1]))
    ws.predict_dev(dF, len(C_te), dW, T.shape[1])
    F.run(dW, T.shape[1], judge=True)
    warm = F.fetch()
    ci, wi = cold["iters"].mean(), warm["iters"].mean()
    print("mean iters cold", ci, "warm", wi, "fallbacks", int(warm["fallback"].sum()), flush=True)
    assert wi < 0.8 * ci
    assert (warm["status"] == "Solved").sum() >= (cold["status"] == "Solved").sum()
