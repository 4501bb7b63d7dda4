"""Oracle basis tables against the reference's literal tables
(src/cartesian/dg/master.h:150-214, fixture tests/golden/master_tables.json)
and against what the mathematics fixes."""
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "master_tables.json")))["tables"]


@pytest.mark.parametrize("P", range(1, 7))
def test_tables_match_reference_literals(oracle, P):
    t = oracle.basis_tables(P)
    for mine, ref in [("nodes", "gl_nodes"), ("denom", "denom"), ("mass", "mass"), ("invmass", "invmass"), ("diff", "diff")]:
        r = np.array(GOLD[ref][str(P)])
        v = np.ravel(t[mine])
        assert np.abs(v - r).max() <= 1e-14 * max(1.0, np.abs(r).max()), (mine, P)


@pytest.mark.parametrize("Q", range(1, 13))
def test_gauss_rules_match_reference_literals(oracle, Q):
    t = oracle.basis_tables(2, Q)
    assert np.abs(t["qnodes"] - GOLD["quad_nodes"][str(Q)]).max() < 1e-15
    assert np.abs(t["qweights"] - GOLD["quad_weights"][str(Q)]).max() < 1e-15


@pytest.mark.parametrize("P", range(2, 7))
def test_basis_math_invariants(oracle, P):
    t = oracle.basis_tables(P)
    x = t["nodes"]
    # nodal property of the Lagrange evaluation (master.h:78-95)
    import ctypes as C
    L = oracle.lib()
    for i, xi in enumerate(x):
        out = np.zeros(P)
        L.orc_lagrange_eval(P, float(xi), out.ctypes.data_as(C.POINTER(C.c_double)))
        assert np.allclose(out, np.eye(P)[i], atol=1e-14)
    # mass is the Gram matrix: 1^T M 1 = |[0,1]| = 1 ; x^T M 1 = 1/2
    M = t["mass"]
    assert abs(np.ones(P) @ M @ np.ones(P) - 1) < 1e-14
    assert abs(x @ M @ np.ones(P) - 0.5) < 1e-14
    assert np.allclose(M @ t["invmass"], np.eye(P), atol=1e-12)
    # differentiation is exact on polynomials of degree P-1
    for k in range(P):
        assert np.allclose(t["diff"] @ x ** k, k * x ** max(k - 1, 0) * (k > 0), atol=1e-11)
    # Gauss rule with Q points integrates degree 2Q-1 exactly
    Q = t["qnodes"].size
    for k in range(2 * Q):
        assert abs(t["qweights"] @ t["qnodes"] ** k - 1 / (k + 1)) < 1e-14
