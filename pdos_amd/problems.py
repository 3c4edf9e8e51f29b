"""Seeded synthetic cone programs for the five workloads.

All generators use numpy's PCG64 (`np.random.default_rng(seed)`), so the same
seed gives the same bits on every machine with this numpy.

  portfolio        long-only factor-risk portfolio SOCP (the shape of the
                   reference's data/ fixtures, decoded in tests/golden):
                   x in R^na, t;  min -mu'x + gamma*t
                   s.t. 1'x = 1 (zero cone), x >= 0 (nonneg),
                        ||(F'x, D^1/2 x)|| <= t (one SOC of dim 1+k+na)
  random_feasible  random sparse A with a planted primal-dual pair
                   (x*, s*, y*) so that the optimum c'x* = -b'y* is known
  dense_family     one dense A shared by a batch, per-instance b, c
  portfolio_family shared F, D; per-instance mu, gamma (A and b shared)
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .abi import DEFAULT_PARAMS, Problem


def proj_cone_np(z: np.ndarray, f: int, l: int, q) -> np.ndarray:
    """Projection onto {0}^f x R_+^l x SOC(q_1) x ... (cones.c:5-57 semantics)."""
    s = np.array(z, dtype=np.float64, copy=True)
    s[:f] = 0.0
    lin = s[f:f + l]
    lin[lin < 0] = 0.0
    r = f + l
    for qi in q:
        t = s[r]
        v = s[r + 1:r + qi]
        ns = float(np.sqrt(np.dot(v, v)))
        if ns <= t:
            pass
        elif ns <= -t:
            s[r:r + qi] = 0.0
        else:
            a = (ns + t) / 2.0
            s[r] = a
            s[r + 1:r + qi] = a * v / ns
        r += qi
    return s


def _params(**kw):
    p = dict(DEFAULT_PARAMS)
    p.update(kw)
    return p


def portfolio_data(rng, na: int, k: int, density: float = 0.5):
    """F (na x k, N(0,1) at `density`), D ~ U[0,1]."""
    F = np.zeros((na, k))
    cnt = int(round(density * na * k))
    pos = rng.choice(na * k, size=cnt, replace=False)
    F.flat[pos] = rng.standard_normal(cnt)
    D = rng.random(na)
    return F, D


def portfolio_matrix(F: np.ndarray, D: np.ndarray):
    """CSC A and b of the portfolio cone program; rows are
    [budget (1)] [-x <= 0 (na)] [SOC: t (1), F'x (k), sqrt(D) x (na)]."""
    na, k = F.shape
    n = na + 1
    m = 2 * na + k + 2
    soc0 = 1 + na
    Ap = [0]
    Ai = []
    Ax = []
    sqD = np.sqrt(D)
    for j in range(na):
        rows = [0, 1 + j]
        vals = [1.0, -1.0]
        nzf = np.nonzero(F[j])[0]
        rows += [soc0 + 1 + r for r in nzf]
        vals += [-F[j, r] for r in nzf]
        rows.append(soc0 + 1 + k + j)
        vals.append(-sqD[j])
        Ai += rows
        Ax += vals
        Ap.append(len(Ai))
    Ai.append(soc0)
    Ax.append(-1.0)
    Ap.append(len(Ai))
    b = np.zeros(m)
    b[0] = 1.0
    return n, m, np.array(Ap), np.array(Ai), np.array(Ax), b, (1, na, [1 + k + na])


def portfolio(seed: int = 0, na: int = 100, k: int = 10, gamma: float = 1.0,
              **params) -> Problem:
    """Single portfolio instance.  Defaults are the n=101, m=212, nnz=801 case:
    MAX_ITERS 2000, EPS_ABS 5e-3, ALPHA 1.0, CG 20 / 1e-4, NORMALIZE 0."""
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal(na)
    F, D = portfolio_data(rng, na, k)
    n, m, Ap, Ai, Ax, b, (f, l, q) = portfolio_matrix(F, D)
    c = np.concatenate([-mu, [gamma]])
    return Problem(n, m, Ap, Ai, Ax, b, c, f, l, q, _params(**params),
                   name=f"portfolio_s{seed}_na{na}_k{k}")


def uneven_portfolios(seed: int, count: int, na_range=(50, 1000), k_range=(5, 50),
                      **params) -> list:
    """Batch of portfolio instances with na ~ U{lo..hi}, k ~ U{lo..hi},
    gamma log-uniform in [0.1, 10]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        na = int(rng.integers(na_range[0], na_range[1] + 1))
        k = int(rng.integers(k_range[0], k_range[1] + 1))
        gamma = float(np.exp(rng.uniform(np.log(0.1), np.log(10.0))))
        mu = rng.standard_normal(na)
        F, D = portfolio_data(rng, na, k)
        n, m, Ap, Ai, Ax, b, (f, l, q) = portfolio_matrix(F, D)
        c = np.concatenate([-mu, [gamma]])
        out.append(Problem(n, m, Ap, Ai, Ax, b, c, f, l, q, _params(**params),
                           name=f"uneven_s{seed}_{i}"))
    return out


class PortfolioFamily:
    """Parametric family: F, D (hence A and b) fixed by the seed; each
    instance draws mu = mu0 + 0.1 N(0,1) and gamma log-uniform [0.1, 10]."""

    def __init__(self, seed: int = 4, na: int = 500, k: int = 20, **params):
        rng = np.random.default_rng(seed)
        self.na, self.k = na, k
        self.mu0 = rng.standard_normal(na)
        self.F, self.D = portfolio_data(rng, na, k)
        (self.n, self.m, self.Ap, self.Ai, self.Ax, self.b,
         (self.f, self.l, self.q)) = portfolio_matrix(self.F, self.D)
        self.params = _params(**params)

    def sample(self, seed: int, count: int):
        """Returns (theta [count, na+1] = (mu, gamma), C [count, n])."""
        rng = np.random.default_rng(seed)
        mu = self.mu0[None, :] + 0.1 * rng.standard_normal((count, self.na))
        gamma = np.exp(rng.uniform(np.log(0.1), np.log(10.0), size=count))
        theta = np.concatenate([mu, gamma[:, None]], axis=1)
        C = np.concatenate([-mu, gamma[:, None]], axis=1)
        return theta, C

    def instance(self, c: np.ndarray, **params) -> Problem:
        p = dict(self.params)
        p.update(params)
        return Problem(self.n, self.m, self.Ap, self.Ai, self.Ax, self.b, c, self.f,
                       self.l, list(self.q), p, name="family")


def soc_sizes(rng, total: int, count: int, lo: int, hi: int) -> list:
    """`count` SOC sizes ~ U{lo..hi}, trimmed deterministically so they sum to
    `total` (each at least 2)."""
    q = rng.integers(lo, hi + 1, size=count).astype(np.int64)
    diff = int(q.sum()) - total
    i = 0
    while diff != 0:
        j = i % count
        if diff > 0 and q[j] > 2:
            step = min(diff, int(q[j]) - 2)
            q[j] -= step
            diff -= step
        elif diff < 0:
            q[j] += -diff
            diff = 0
        i += 1
    return [int(v) for v in q]


def random_feasible(seed: int = 1, n: int = 100_000, m: int = 300_000, nnz: int = 5_000_000,
                    f: int = 1_000, l: int = 149_000, nsoc: int = 2_000, soc_lo: int = 10,
                    soc_hi: int = 140, dense_rows: int = 0, dense_cols: int = 0, empty_rows: int = 0,
                    empty_cols: int = 0, **params) -> Problem:
    """Random sparse SOCP with a planted optimal pair.  Entries N(0,1) at
    uniform random (i, j) (duplicates summed), plus `dense_rows` full rows
    (entries N(0,1)/sqrt(n)) and `dense_cols` full columns (N(0,1)/sqrt(m)) at
    random positions.  `empty_rows` rows (0, m/e, 2m/e, ...) and the last
    `empty_cols` columns then lose all their entries (a row of A that is
    empty leaves s_i = b_i; an empty column leaves x_j free with c_j = 0).
    x* ~ N(0,1), z ~ N(0,1), s* = Pi_K(z), y* = Pi_K(z) - z, b = A x* + s*,
    c = -A' y*; then (x*, s*, y*) is primal-dual optimal with c'x* = -b'y*."""
    rng = np.random.default_rng(seed)
    q = soc_sizes(rng, m - f - l, nsoc, soc_lo, soc_hi) if nsoc > 0 else []
    rows = rng.integers(0, m, size=nnz)
    cols = rng.integers(0, n, size=nnz)
    vals = rng.standard_normal(nnz)
    if dense_rows or dense_cols:
        rd = np.random.default_rng(seed + 7919)
        dr = rd.choice(m, size=dense_rows, replace=False)
        dc = rd.choice(n, size=dense_cols, replace=False)
        rows = np.concatenate([rows, np.repeat(dr, n), np.tile(np.arange(m), dense_cols)])
        cols = np.concatenate([cols, np.tile(np.arange(n), dense_rows), np.repeat(dc, m)])
        vals = np.concatenate([vals, rd.standard_normal(dense_rows * n) / np.sqrt(n),
                               rd.standard_normal(dense_cols * m) / np.sqrt(m)])
    if empty_rows or empty_cols:
        gone = np.isin(rows, np.arange(empty_rows) * (m // max(empty_rows, 1))) | (cols >= n - empty_cols)
        rows, cols, vals = rows[~gone], cols[~gone], vals[~gone]
    A = sp.csc_matrix((vals, (rows, cols)), shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    xs = rng.standard_normal(n)
    z = rng.standard_normal(m)
    ss = proj_cone_np(z, f, l, q)
    ys = ss - z
    b = A @ xs + ss
    c = -(A.T @ ys)
    p = Problem(n, m, A.indptr, A.indices, A.data, b, c, f, l, q, _params(**params),
                name=f"random_s{seed}_n{n}_m{m}")
    p.planted_opt = float(c @ xs)  # the optimal value c'x* = -b'y*
    return p


def jumbled(p: Problem, seed: int, dup_frac: float = 0.3) -> Problem:
    """The same problem with A stored as the reference's CSC input allows but
    a canonical matrix never is: each column's entries in random row order,
    and a fraction `dup_frac` of them split into two entries of the same row
    (a - w and w, w ~ N(0,1)), which every product sums.  The split changes A
    by the rounding of a - w + w, so a planted optimum stays optimal to
    roundoff."""
    rng = np.random.default_rng(seed)
    Ap, Ai, Ax = [0], [], []
    for j in range(p.n):
        r = p.Ai[p.Ap[j]:p.Ap[j + 1]]
        v = p.Ax[p.Ap[j]:p.Ap[j + 1]]
        split = rng.random(len(r)) < dup_frac
        w = rng.standard_normal(len(r))
        rr = np.concatenate([r, r[split]])
        vv = np.concatenate([np.where(split, v - w, v), w[split]])
        order = rng.permutation(len(rr))
        Ai.append(rr[order])
        Ax.append(vv[order])
        Ap.append(Ap[-1] + len(rr))
    return Problem(p.n, p.m, np.array(Ap), np.concatenate(Ai), np.concatenate(Ax), p.b.copy(), p.c.copy(),
                   p.f, p.l, list(p.q), dict(p.params), name=p.name + "_jumbled")


class DenseFamily:
    """One dense A (m x n, N(0,1)/sqrt(n)) shared by a batch; per-instance
    (b, c) from planted primal-dual pairs.  Cones: l rows then `nsoc` SOCs of
    size `qdim`."""

    def __init__(self, seed: int = 2, m: int = 8192, n: int = 2048, l: int = 6144,
                 nsoc: int = 64, qdim: int = 32, **params):
        assert l + nsoc * qdim == m
        rng = np.random.default_rng(seed)
        self.m, self.n, self.f, self.l = m, n, 0, l
        self.q = [qdim] * nsoc
        self.A = (rng.standard_normal((m, n)) / np.sqrt(n))
        self.params = _params(**params)

    def sample(self, seed: int, count: int):
        """Returns B [count, m], C [count, n], and the planted optimal values."""
        rng = np.random.default_rng(seed)
        X = rng.standard_normal((count, self.n))
        Z = rng.standard_normal((count, self.m))
        S = np.stack([proj_cone_np(z, self.f, self.l, self.q) for z in Z])
        Y = S - Z
        B = X @ self.A.T + S
        Cm = -(Y @ self.A)
        opt = np.einsum("ij,ij->i", Cm, X)
        return B, Cm, opt

    def instance(self, b: np.ndarray, c: np.ndarray, **params) -> Problem:
        A = sp.csc_matrix(self.A)
        p = dict(self.params)
        p.update(params)
        return Problem(self.n, self.m, A.indptr, A.indices, A.data, b, c, self.f, self.l,
                       list(self.q), p, name="dense_family")
