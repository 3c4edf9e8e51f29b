"""Learned warm start for a parametric family (the MLP of pdos_amd.mlp).

The network maps an instance's varying data (features) to a full ADMM
starting point (x0, y0, s0) in original units.  It is trained on solutions
the solver itself produced; at solve time ADMM judges every prediction
against the cold start (pdos_batch_set_warm_dev / pdos_family_run with
judge on) and falls back to the cold start when the prediction is worse.
"""
from __future__ import annotations

import numpy as np

from .mlp import Mlp


def portfolio_features(theta: np.ndarray) -> np.ndarray:
    """(mu, gamma) -> (mu, log gamma): the portfolio family's varying data."""
    theta = np.asarray(theta, dtype=np.float64)
    return np.concatenate([theta[:, :-1], np.log(theta[:, -1:])], axis=1).astype(np.float32)


class LearnedWarmStart:
    def __init__(self, n_in: int, n_out: int, hidden: int = 512, seed: int = 0):
        self.net = Mlp(n_in, hidden, n_out, seed=seed)
        self.n_in, self.n_out = n_in, n_out

    # (fraction of the epochs, learning-rate factor): a step decay; Adam's
    # moments carry over from one stage to the next
    SCHEDULE = ((0.6, 1.0), (0.25, 1.0 / 3.0), (0.15, 0.1))

    def fit(self, F, T, epochs: int = 30, batch: int = 512, lr: float = 1e-3, seed: int = 0) -> float:
        """Trains on features F and targets T (rows); returns the last epoch's
        mean loss in the standardized target space."""
        F = np.asarray(F, np.float64)
        T = np.asarray(T, np.float64)
        self.fmu, self.fsd = F.mean(0), F.std(0)
        self.fsd[self.fsd < 1e-8] = 1.0
        self.tmu, self.tsd = T.mean(0), T.std(0)
        self.tsd[self.tsd < 1e-8] = 1.0
        self.net.set_affine(self.fmu, self.fsd, self.tsd, self.tmu)
        Tn = ((T - self.tmu) / self.tsd).astype(np.float32)
        Ff = F.astype(np.float32)
        loss, done = float("nan"), 0
        for k, (frac, f) in enumerate(self.SCHEDULE):
            e = epochs - done if k == len(self.SCHEDULE) - 1 else int(round(frac * epochs))
            if e > 0:
                loss = self.net.train(Ff, Tn, e, batch, lr * f, seed + k)
                done += e
        return loss

    def predict(self, F) -> np.ndarray:
        return self.net.predict(np.asarray(F, np.float32))

    def predict_dev(self, dF, rows: int, dW, ld: int) -> None:
        self.net.predict_dev(dF, rows, dW, ld)
