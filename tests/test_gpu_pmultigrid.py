"""GPU: p- and hp-multigrid (Multigrid<N,P1,P2,...> with
InterpolationHierarchy<N,P1,P2,...>, multigrid.h:435-639,
interpolationhierarchy.h:98-154) through the C ABI, against the oracle.

Bars: patterns bit-exact, level operators 1e-12, interpolation 1e-15,
MG-PCG iteration counts equal to the oracle's (FP64; mixed within one),
solutions 1e-9.  The StopCoarsening case is a Neumann problem (its last
p-level's coarse solve subtracts the mean of its right-hand side,
multigrid.h:596-599), compared up to a constant."""
import numpy as np
import pytest

from conftest import sine_f

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def h_adaptive(x):
    return 1 / 8 if x[0] > 0.5 else 1 / 16


@pytest.mark.parametrize("restriction", ["ratrat", "rat"])
@pytest.mark.parametrize("p,degrees", [(3, [2, 1]), (3, [1]), (2, [0]), (5, [3, 1])])
def test_hp_hierarchy_matches_oracle(gctx, oracle, p, degrees, restriction):
    import mgldg
    qt = mgldg.Quadtree(gctx, 2, h_adaptive)
    pr = mgldg.LDGPoisson(qt, p, tau0=1.0)
    oqt = oracle.Quadtree(2, h_adaptive)
    opr = oracle.Problem(oqt, p, tau0=1.0)
    mg = mgldg.Multigrid(pr, "gs", restriction=restriction, pdegrees=degrees)
    omg = oracle.Multigrid(opr, "mcgs", restriction=restriction, pdegrees=degrees)
    h = oracle.Multigrid(oracle.Problem(oqt, degrees[-1]), "mcgs")
    assert mg.nlevels == omg.nlevels == len(degrees) + h.nlevels
    assert mg.level_B[:len(degrees) + 1] == [(d + 1) ** 2 for d in [p] + degrees]
    for l in range(mg.nlevels):
        a, b = mg.export(l, "A"), omg.export(l, "A")
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), l
        assert rel(a[2], b[2]) < 1e-12, l
        if l + 1 < mg.nlevels:
            ti, to = mg.export(l, "interp"), omg.export(l, "interp")
            assert np.array_equal(ti[1], to[1]) and rel(ti[2], to[2]) < 1e-15, l


@pytest.mark.parametrize("relax", ["gs", "jacobi"])
@pytest.mark.parametrize("p,degrees", [(3, [2, 1]), (3, [1])])
def test_hp_pcg_matches_oracle(gctx, oracle, p, degrees, relax):
    import mgldg
    qt = mgldg.Quadtree(gctx, 2, h_adaptive)
    pr = mgldg.LDGPoisson(qt, p)
    oqt = oracle.Quadtree(2, h_adaptive)
    opr = oracle.Problem(oqt, p)
    mg = mgldg.Multigrid(pr, relax, npre=2, npost=2, precision="mixed", pdegrees=degrees)
    omg = oracle.Multigrid(opr, "mcgs" if relax == "gs" else relax, npre=2, npost=2, pdegrees=degrees)
    b = pr.rhs(sine_f(pr.nodes()))
    xo, ito, _, _ = oracle.pcg(opr, omg, b, tol=1e-10, maxit=200, criterion="true")
    x, st = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion="true")
    assert st["status"] == "converged" and st["iterations"] == ito
    assert rel(x, xo) < 1e-9
    xm, stm = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion="true", precision="mixed", flexible=True)
    assert stm["status"] == "converged" and abs(stm["iterations"] - ito) <= 1 and rel(xm, xo) < 1e-8
    # batched columns equal single solves
    import torch
    B = torch.tensor(np.stack([b, 2 * b, -b], 1), device="cuda")
    X, sts = mgldg.pcg(mg, B, torch.zeros_like(B), tol=1e-10, maxit=200, criterion="true")
    assert all(s["iterations"] == ito for s in sts)
    assert rel(X[:, 1].cpu().numpy() / 2, xo) < 1e-9


def test_p_only_stop_coarsening_neumann(gctx, oracle):
    import mgldg
    qt = mgldg.Quadtree(gctx, 2, lambda x: 1 / 8)
    pr = mgldg.LDGPoisson(qt, 3, bc="neumann")
    opr = oracle.Problem(oracle.Quadtree(2, lambda x: 1 / 8), 3, bctype=[1] * 4)
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, pdegrees=[1], hcoarsen=False)
    omg = oracle.Multigrid(opr, "mcgs", npre=2, npost=2, pdegrees=[1], hcoarsen=False)
    assert mg.nlevels == omg.nlevels == 2
    b = pr.rhs(sine_f(pr.nodes()))
    b = b - b.mean()
    xo, ito, _, _ = oracle.pcg(opr, omg, b, tol=1e-10, maxit=100, criterion="true")
    x, st = mgldg.pcg(mg, b, tol=1e-10, maxit=100, criterion="true")
    assert st["status"] == "converged" and st["iterations"] == ito
    d = x - xo
    assert np.abs(d - d.mean()).max() < 1e-9 * np.abs(xo - xo.mean()).max()


def test_invalid_degrees_raise(gctx):
    import mgldg
    qt = mgldg.Quadtree(gctx, 2, lambda x: 1 / 8)
    pr = mgldg.LDGPoisson(qt, 2)
    for bad in ([2], [1, 1], [0, 1], [-1]):
        with pytest.raises(mgldg.MGLDGError):
            mgldg.Multigrid(pr, "gs", pdegrees=bad)
