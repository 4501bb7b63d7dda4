"""The oracle's stationary V-cycle solver (the driver's `--solver vcycles`,
tests/cartesian/dg_tests.cpp:309-332), pinned by identities of the V-cycle:
one cycle from zero is the preconditioner's V(b); every step of a V-cycle
is affine in (x, b), so a cycle from x0 equals x0 + V(b - A x0); the residual
history starts at ||b - A x0|| and contracts.  Dirichlet, Neumann and
periodic problems (the last two with the right-hand side made mean-zero, as
the driver does, dg_tests.cpp:282-286)."""
import numpy as np
import pytest

from conftest import sine_f

CASES = {"dirichlet": ({}, 0), "neumann": ({}, 1), "periodic": ({"periodic": True}, 1)}


def setup(oracle, case, relaxation="gs"):
    kw, bct = CASES[case]
    qt = oracle.Quadtree(2, lambda x: 1 / 16, **kw)
    pr = oracle.Problem(qt, 1, bctype=[bct] * 4)
    mg = oracle.Multigrid(pr, relaxation, npre=2, npost=2)
    b = pr.rhs(sine_f(pr.nodes()), np.zeros(pr.nqb))
    if bct:
        b = b - b.mean()        # Function::meanZero (function.h:156-170)
    return pr, mg, b


@pytest.mark.parametrize("case", list(CASES))
def test_one_cycle_from_zero_is_the_preconditioner(oracle, case):
    pr, mg, b = setup(oracle, case)
    x, rn = mg.vcycles(b, np.zeros_like(b), 1)
    assert np.array_equal(x, mg.vcycle(b))
    assert rn[0] == pytest.approx(np.linalg.norm(b), rel=1e-14)


@pytest.mark.parametrize("relaxation", ["gs", "mcgs", "jacobi"])
@pytest.mark.parametrize("case", list(CASES))
def test_cycle_is_affine_in_the_iterate(oracle, case, relaxation):
    pr, mg, b = setup(oracle, case, relaxation)
    x0 = np.random.default_rng(1).standard_normal(len(b))
    x, rn = mg.vcycles(b, x0, 1)
    expect = x0 + mg.vcycle(b - mg.spmv(0, x0))
    d = x - expect
    if CASES[case][1]:
        # Neumann / periodic: A has the constants as its kernel and the
        # coarsest block factorisation leaves their component to rounding,
        # so iterates agree up to a constant (residuals are unaffected)
        d = d - d.mean()
    assert np.abs(d).max() < 1e-12 * np.abs(expect).max()
    assert rn[0] == pytest.approx(np.linalg.norm(b - mg.spmv(0, x0)), rel=1e-13)
    assert rn[1] == pytest.approx(np.linalg.norm(b - mg.spmv(0, x)), rel=1e-13)


@pytest.mark.parametrize("case", list(CASES))
def test_residuals_contract(oracle, case):
    pr, mg, b = setup(oracle, case)
    x, rn = mg.vcycles(b, np.zeros_like(b), 10)
    ratios = rn[1:] / rn[:-1]
    assert np.all(ratios < 0.85) and np.all(ratios[1:] < 0.5)
    assert rn[-1] < 1e-3 * rn[0]
    # a plant: a cycle that skipped the coarse-grid correction would contract far less
    y, rs = mg.vcycles(b, x, 0)
    assert len(rs) == 1 and np.array_equal(y, x)
