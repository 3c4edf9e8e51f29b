"""The Python binding's argument handling (python/pdosmodule.c:27-344) on the
host: defaults, option overrides, the cone dims and the CSC conversion."""
import numpy as np
import pytest
import scipy.sparse as sp

import golden_data as G
import pdos_amd


def _args(tag):
    p = G.problem(tag)
    A = sp.csc_matrix((p.Ax, p.Ai, p.Ap), shape=(p.m, p.n))
    return p, A, {"f": p.f, "l": p.l, "q": list(p.q)}


def test_binding_defaults_follow_the_reference_binding():
    p, A, dims = _args("fixture01")
    q = pdos_amd.binding_problem(p.c, A, p.b, dims)
    # pdosmodule.c:67-74 (indirect build), NORMALIZE 0 (README.md:89)
    assert q.params == dict(MAX_ITERS=2000, CG_MAX_ITS=20, EPS_ABS=5e-3, ALPHA=1.0, CG_TOL=1e-4, VERBOSE=1,
                            NORMALIZE=0)
    assert (q.n, q.m, q.f, q.l, list(q.q)) == (p.n, p.m, p.f, p.l, list(p.q))
    assert np.array_equal(q.Ap, p.Ap) and np.array_equal(q.Ai, p.Ai) and np.array_equal(q.Ax, p.Ax)


def test_binding_options_override_and_unknown_option_raises():
    p, A, dims = _args("fixture01")
    q = pdos_amd.binding_problem(p.c, A.tocoo(), p.b, dims, opts={"ALPHA": 1.8, "NORMALIZE": 1, "VERBOSE": 0})
    assert q.params["ALPHA"] == 1.8 and q.params["NORMALIZE"] == 1 and q.params["VERBOSE"] == 0
    assert q.params["MAX_ITERS"] == 2000
    assert np.array_equal(q.Ap, p.Ap)  # COO input converted to the same CSC
    with pytest.raises(KeyError):
        pdos_amd.binding_problem(p.c, A, p.b, dims, opts={"RHO": 1.0})
