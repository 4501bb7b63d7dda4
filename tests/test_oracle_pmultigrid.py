"""The oracle's p- and hp-multigrid (Multigrid<N,P1,P2,...> and
InterpolationHierarchy<N,P1,P2,...>, multigrid.h:435-639,
interpolationhierarchy.h:98-154), pinned by what the mathematics fixes:
the p-interpolation reproduces the coarse degree's polynomials; the spaces
are nested on every element, so the RATRAT Galerkin operators of a p-level
equal the LDG discretisation at that degree on the same mesh; RAT equals
the dense product; the hierarchy's level sizes and the V-cycle's linearity;
MG-PCG with it matches a dense solve."""
import numpy as np
import pytest

from conftest import sine_f


def problems(oracle, n, degrees, tau0=1.5, adaptive=False):
    h = (lambda x: 1 / n) if not adaptive else (lambda x: (1 / n) if x[0] > 0.5 else 0.5 / n)
    qt = oracle.Quadtree(2, h)
    return qt, [oracle.Problem(qt, p, tau0=tau0) for p in degrees]


@pytest.mark.parametrize("pf,pc", [(3, 2), (3, 1), (2, 1), (2, 0), (1, 0)])
def test_p_interpolation_reproduces_coarse_polynomials(oracle, pf, pc):
    qt, (fine, coarse) = problems(oracle, 4, [pf, pc])
    mg = oracle.Multigrid(fine, "gs", pdegrees=[pc], hcoarsen=False)
    assert mg.nlevels == 2
    T = mg.dense(0, "interp")
    assert T.shape == (fine.ndof, coarse.ndof)
    Xf, Xc = fine.nodes(), coarse.nodes()
    for fn in (lambda X: 1.0 + 0 * X[:, 0], lambda X: X[:, 0] ** min(pc, 2) - X[:, 1] ** pc,
               lambda X: (X[:, 0] * X[:, 1]) ** pc):
        assert np.abs(T @ fn(Xc) - fn(Xf)).max() < 1e-13


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("pf,pc", [(3, 1), (2, 1), (3, 2)])
def test_ratrat_p_level_equals_rediscretisation(oracle, pf, pc, adaptive):
    qt, (fine, coarse) = problems(oracle, 8, [pf, pc], adaptive=adaptive)
    mg = oracle.Multigrid(fine, "gs", restriction="ratrat", pdegrees=[pc], hcoarsen=False)
    for w in ["M", "T", "G0", "G1", "A"]:
        Ac, Ar = mg.dense(1, w), coarse.dense(w)
        assert np.abs(Ac - Ar).max() < 1e-11 * np.abs(Ar).max(), w


def test_rat_p_level_equals_dense_product(oracle):
    qt, (fine,) = problems(oracle, 8, [3])
    mg = oracle.Multigrid(fine, "gs", restriction="rat", pdegrees=[2, 1], hcoarsen=False)
    for l in range(2):
        T = mg.dense(l, "interp")
        ref = T.T @ mg.dense(l, "A") @ T
        assert np.abs(mg.dense(l + 1, "A") - ref).max() < 1e-12 * np.abs(ref).max()


def test_hp_hierarchy_shape(oracle):
    qt, (fine,) = problems(oracle, 8, [3])
    mg = oracle.Multigrid(fine, "gs", pdegrees=[2, 1], hcoarsen=True)
    h = oracle.Multigrid(oracle.Problem(qt, 1), "gs")
    assert mg.nlevels == 2 + h.nlevels
    dims = [mg.level_dims(l) for l in range(mg.nlevels)]
    assert [d[1] for d in dims[:3]] == [16, 9, 4]
    assert dims[2:] == [h.level_dims(l) for l in range(h.nlevels)]
    with pytest.raises(ValueError):
        oracle.Multigrid(fine, "gs", pdegrees=[1, 2])   # degrees must decrease
    with pytest.raises(ValueError):
        oracle.Multigrid(fine, "gs", pdegrees=[3])      # and stay below the problem's


@pytest.mark.parametrize("hcoarsen", [True, False])
def test_hp_pcg_matches_dense_solve(oracle, hcoarsen):
    """hp: p-levels then h-levels, Dirichlet.  p only (StopCoarsening): the
    last p-level's coarse solve subtracts the mean of its right-hand side
    (multigrid.h:596-599), consistent only where the constants are the
    operator's kernel, so that case is a Neumann problem with a mean-zero
    right-hand side, compared up to a constant."""
    qt = oracle.Quadtree(2, lambda x: 1 / 8)
    bct = 0 if hcoarsen else 1
    fine = oracle.Problem(qt, 3, bctype=[bct] * 4, tau0=0.0)
    mg = oracle.Multigrid(fine, "gs", npre=2, npost=2, pdegrees=[1], hcoarsen=hcoarsen)
    b = fine.rhs(sine_f(fine.nodes()), np.zeros(fine.nqb))
    if bct:
        b = b - b.mean()
    x, it, rr, _ = oracle.pcg(fine, mg, b, tol=1e-10, maxit=100, criterion="true")
    xd = np.linalg.lstsq(fine.dense("A"), b, rcond=None)[0]
    d = x - xd
    if bct:
        d, xd = d - d.mean(), xd - xd.mean()
    assert rr <= 1e-10 and np.abs(d).max() < 1e-8 * np.abs(xd).max()
    assert it < 30
    # the V-cycle is a linear map (up to the kernel's constants on the Neumann problem)
    v1, v2 = np.random.default_rng(0).standard_normal((2, fine.ndof))
    d = mg.vcycle(2 * v1 + v2) - 2 * mg.vcycle(v1) - mg.vcycle(v2)
    if bct:
        d = d - d.mean()
    assert np.abs(d).max() < 1e-10 * np.abs(mg.vcycle(v1)).max()
