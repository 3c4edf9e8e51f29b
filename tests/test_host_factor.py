"""The direct plugin's host factorization (ordering, LDL', level sets),
checked on the CPU against a dense solve of the scaled KKT system."""
import numpy as np
import pytest

import golden_data as G
import pdos_amd
from oracle import oracle
from pdos_amd import problems


def _cases():
    yield "portfolio", problems.portfolio(0)
    yield "portfolio_norm", problems.portfolio(1).with_params(NORMALIZE=1)
    yield "fixture07", G.problem("fixture07")
    yield "random", G.problem("random_n300_m900")
    yield "uneven", problems.uneven_portfolios(3, 1, na_range=(300, 300), k_range=(30, 30))[0]
    yield "dense", G.problem("dense_m96_n24")


@pytest.mark.parametrize("name,p", list(_cases()), ids=[c[0] for c in _cases()])
def test_kkt_solve_matches_dense(name, p):
    N = p.n + p.m
    rhs = np.random.default_rng(1).standard_normal(N)
    z, st = pdos_amd.kkt_solve_host(p, rhs)
    D, E, lam = oracle.setup(p)  # scalings the solver iterates in (x1e3 included)
    A = p.dense_A() * D[:, None] * E[None, :]
    K = np.block([[-np.eye(p.n), A.T], [A, np.eye(p.m)]])
    zd = np.linalg.solve(K, rhs)
    assert np.linalg.norm(z - zd) <= 1e-9 * np.linalg.norm(zd)
    assert st[0] == N and st[2] >= 1 and st[3] >= 1 and st[1] >= 0
