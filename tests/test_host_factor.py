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


@pytest.mark.parametrize("variant", ["", "ruiz"])
def test_library_setup_bitwise_vs_reference(variant):
    """The library's host setup (pdos_setup_host) against the reference's
    initWork on every golden case: D, E (with the x1e3) and lambda bit for bit,
    for the reference build and for its ruiz_equil_common.h build."""
    equil = "ruiz" if variant else "reference"
    for tag in G.case_names(variant):
        prob = G.problem(tag, variant)
        D, E, lam, _ = G.setup_expected(tag, variant)
        d2, e2, l2 = pdos_amd.setup_host(prob, equil)
        np.testing.assert_array_equal(d2, D, err_msg=f"{tag} D")
        np.testing.assert_array_equal(e2, E, err_msg=f"{tag} E")
        assert l2 == lam, tag


def test_family_rejects_ruiz():
    """Ruiz scaling depends on each instance's b and c; a family shares one
    scaling of A, so the family engine refuses the option."""
    from pdos_amd.family import Family
    fam = problems.PortfolioFamily(seed=4, na=20, k=3)
    _, Cm = fam.sample(seed=1, count=2)
    with pytest.raises(RuntimeError, match="PDOS_EQUIL_REFERENCE"):
        Family(fam.instance(Cm[0]), precision="mixed", equil="ruiz")


def test_ruiz_unsorted_columns_library_matches_oracle():
    """Columns whose row indices do not ascend, with NORMALIZE 1: the
    reference's cone collapse (ruiz_equil_common.h:29-66) then indexes k->q
    past qsize, so there is no reference output.  The library and the oracle
    both stop the walk of a column at its last cone; their scalings agree bit
    for bit, and with sorted columns they equal the reference's (golden)."""
    prob = G.problem("jumbled_n120_m300_norm1")
    assert prob.params["NORMALIZE"] == 1
    assert any(np.any(np.diff(prob.Ai[prob.Ap[j]:prob.Ap[j + 1]]) < 0) for j in range(prob.n))
    d2, e2, l2 = pdos_amd.setup_host(prob, "ruiz")
    d3, e3, l3 = oracle.setup(prob, equil="ruiz")
    np.testing.assert_array_equal(d2, d3)
    np.testing.assert_array_equal(e2, e3)
    assert l2 == l3
    assert np.all(np.isfinite(d2)) and np.all(np.isfinite(e2))


def test_library_ordering_keeps_reference_iterations():
    """The library's own ordering (approximate minimum degree) in place of the
    reference's AMD: the oracle, which equals the reference bit for bit given
    a permutation, run with the library's permutation reaches the reference's
    status and iteration count on every golden case, x, s, y within 1e-6."""
    for tag in G.case_names():
        prob = G.problem(tag)
        perm = pdos_amd.kkt_order(prob)
        assert sorted(perm) == list(range(prob.n + prob.m)), tag
        out = oracle.solve(prob, "direct", perm=perm)
        exp = G.expected(tag, "direct")
        assert out["status"] == exp["status"] and out["iters"] == exp["iters"], tag
        for k in ("x", "s", "y"):
            assert np.linalg.norm(out[k] - exp[k]) <= 1e-6 * max(np.linalg.norm(exp[k]), 1e-12), (tag, k)


@pytest.mark.parametrize("name,p", [
    ("portfolio_na3000", problems.portfolio(seed=5, na=3000, k=20)),
    ("random_n2000_m6000", problems.random_feasible(seed=7, n=2000, m=6000, nnz=14000, f=50, l=2950, nsoc=60,
                                                    soc_lo=10, soc_hi=90)),
], ids=["portfolio_na3000", "random_n2000_m6000"])
def test_large_kkt_factor(name, p):
    """Beyond the old exact-degree cap's comfort: the ordering's fill stays
    within 1.5x of exact minimum degree (the oracle's), and the factor solves
    the KKT system to 1e-9 against a sparse LU solve."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla
    N = p.n + p.m
    rhs = np.random.default_rng(2).standard_normal(N)
    z, st = pdos_amd.kkt_solve_host(p, rhs)
    D, E, lam = oracle.setup(p)
    A = sps.csc_matrix((p.Ax, p.Ai, p.Ap), shape=(p.m, p.n))
    As = sps.diags(D) @ A @ sps.diags(E)
    K = sps.bmat([[-sps.eye(p.n), As.T], [As, sps.eye(p.m)]], format="csc")
    zd = spla.spsolve(K, rhs)
    assert np.linalg.norm(z - zd) <= 1e-9 * np.linalg.norm(zd)
    exact = oracle.solve(p.with_params(MAX_ITERS=0), "direct")["lnz"]
    assert st[1] <= 1.5 * exact, (st[1], exact)
