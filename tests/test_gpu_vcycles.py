"""GPU: Neumann and periodic problems, the driver's stationary V-cycle
solver (`--solver vcycles`, tests/cartesian/dg_tests.cpp:309-332) and
Function::meanZero (function.h:156-170), through the C ABI, against the
oracle.

Neumann and periodic operators have the constants as their kernel; the
driver makes the right-hand side mean-zero (dg_tests.cpp:282-286) and the
coarsest block factorisation leaves the constant component of an iterate to
rounding, so iterates of those problems are compared up to a constant.
Bars: patterns bit-exact, operators 1e-13, FP64 iterates and residual
histories 1e-10, MG-PCG iteration counts equal to the oracle's."""
import numpy as np
import pytest

from conftest import sine_f

pytestmark = pytest.mark.gpu

# case: (quadtree kwargs, library bc, oracle bc type)
CASES = {"dirichlet": ({}, "dirichlet", 0), "neumann": ({}, "neumann", 1), "periodic": ({"periodic": True}, "dirichlet", 1)}
RELAX = {"gs-lex": "gs", "gs": "mcgs", "jacobi": "jacobi"}


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rel_mod_const(a, b, singular):
    d = a - b
    if singular:
        d = d - d.mean()
        b = b - b.mean()
    return np.abs(d).max() / max(np.abs(b).max(), 1e-300)


def build(gctx, oracle, case, relax="gs", collapse=0, p=1, n=16):
    import mgldg
    kw, bc, bct = CASES[case]
    qt = mgldg.Quadtree(gctx, 2, lambda x: 1 / n, **kw)
    pr = mgldg.LDGPoisson(qt, p, bc=bc)
    mg = mgldg.Multigrid(pr, relax, npre=2, npost=2, precision="mixed", coarse_dense_max=collapse)
    oqt = oracle.Quadtree(2, lambda x: 1 / n, **kw)
    opr = oracle.Problem(oqt, p, bctype=[bct] * 4)
    omg = oracle.Multigrid(opr, RELAX[relax], npre=2, npost=2)
    b = pr.rhs(sine_f(pr.nodes()))
    if bct:
        b = b - b.mean()
    return pr, mg, opr, omg, b


@pytest.mark.parametrize("case", ["neumann", "periodic"])
def test_operators_match_oracle(gctx, oracle, case):
    pr, mg, opr, omg, b = build(gctx, oracle, case, p=2, n=8)
    assert (pr.ne, pr.nf, pr.nbf, pr.ndof) == (opr.ne, opr.nf, opr.nbf, opr.ndof)
    for w in ["M", "T", "A", "G0", "G1"]:
        a, o = pr.export(w), opr.export(w)
        assert np.array_equal(a[0], o[0]) and np.array_equal(a[1], o[1]), w
        assert rel(a[2], o[2]) < 1e-13, w
    from oracle import bsr_to_dense
    rp, cols, blk = pr.export("A")
    A = bsr_to_dense(rp, cols, blk, pr.ne, pr.ne)
    assert np.abs(A @ np.ones(pr.ndof)).max() < 1e-10 * np.abs(A).max()   # constants in the kernel


@pytest.mark.parametrize("collapse", [0, 4096])
@pytest.mark.parametrize("relax", list(RELAX))
@pytest.mark.parametrize("case", list(CASES))
def test_vcycles_match_oracle(gctx, oracle, case, relax, collapse):
    import mgldg
    pr, mg, opr, omg, b = build(gctx, oracle, case, relax, collapse)
    x0 = np.random.default_rng(5).standard_normal(pr.ndof)
    xo, rno = omg.vcycles(b, x0, 6)
    x, rn = mgldg.vcycles(mg, b, x0, ncycle=6)
    assert len(rn) == 7
    assert rel(rn, rno) < 1e-10
    assert rel_mod_const(x, xo, CASES[case][2]) < 1e-10


@pytest.mark.parametrize("case", list(CASES))
def test_vcycles_mixed_and_tolerance(gctx, oracle, case):
    import mgldg
    import torch
    pr, mg, opr, omg, b = build(gctx, oracle, case)
    _, rno = omg.vcycles(b, np.zeros(pr.ndof), 4)
    bd = torch.tensor(b, device="cuda")
    xd = torch.zeros_like(bd)
    x, rn = mgldg.vcycles(mg, bd, xd, ncycle=4, precision="mixed")
    assert x is xd and len(rn) == 5
    assert rel(rn, rno) < 1e-4            # FP32 V-cycle, FP64 residual and update
    tol = 1e-6
    x2, rn2 = mgldg.vcycles(mg, b, None, ncycle=60, tol=tol)
    assert len(rn2) < 61 and rn2[-1] <= tol * np.linalg.norm(b) < rn2[-2]


def test_mean_zero_matches_definition(gctx):
    import mgldg
    import torch
    rng = np.random.default_rng(2)
    for shape in [(1000,), (777, 3)]:
        v = rng.standard_normal(shape) + 3.0
        t = torch.tensor(v, device="cuda")
        mgldg.mean_zero(gctx, t)
        assert np.abs(t.cpu().numpy() - (v - v.mean(axis=0))).max() < 1e-13


@pytest.mark.parametrize("case", ["neumann", "periodic"])
def test_pcg_matches_oracle(gctx, oracle, case):
    import mgldg
    import torch
    pr, mg, opr, omg, b = build(gctx, oracle, case, "gs", p=2, n=16)
    xo, ito, _, _ = oracle.pcg(opr, omg, b, tol=1e-10, maxit=200, criterion="reference")
    x, st = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion="reference")
    assert st["status"] == "converged" and st["iterations"] == ito
    assert rel_mod_const(x, xo, True) < 1e-8
    # the right-hand side made mean-zero on the device gives the same solve
    bd = torch.tensor(pr.rhs(sine_f(pr.nodes())), device="cuda")
    mgldg.mean_zero(pr.ctx, bd)
    assert rel(bd.cpu().numpy(), b) < 1e-13
