"""Known-answer and invariant checks that pin the oracle (CPU).

Each check is chosen so that a plausible slip in the restatement (a dropped
lifting term, a wrong sign or index, a transposed block) fails it."""
import numpy as np
import pytest

from conftest import sine_f


def test_matrix_tests_golden(oracle):
    """tests/cartesian/matrix_tests.cpp:31-52: y = A*1 = [22, 30, 37, 38]"""
    rowptr = np.array([0, 2, 4]); cols = np.array([0, 1, 0, 1])
    blocks = np.array([[[1, 2], [3, 4]], [[9, 10], [11, 12]], [[4, 6], [7, 8]], [[13, 14], [15, 8]]], dtype=float)
    # matrix_tests sets blocks (0,0),(0,1),(1,0),(1,1) in that order
    y = oracle.bsr_spmv(rowptr, cols, blocks, np.ones(4), 2)
    assert np.array_equal(y, [22, 30, 37, 38])


def _uniform(O, n, p, **kw):
    qt = O.Quadtree(2, lambda x: 1.0 / n)
    return qt, O.Problem(qt, p, **kw)


def test_mesh_counts(oracle):
    qt, pr = _uniform(oracle, 32, 1)
    assert qt.info() == (1365, 6)
    assert (pr.ne, pr.nf, pr.nbf, pr.b, pr.ndof) == (1024, 2112, 128, 4, 4096)
    assert len(pr.export("A")[1]) == 4992


@pytest.mark.parametrize("p", [1, 2, 3])
def test_assembly_symmetric_spd(oracle, p):
    _, pr = _uniform(oracle, 4, p)
    A = pr.dense("A")
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0
    T = pr.dense("T")
    assert np.abs(T - T.T).max() < 1e-13 * np.abs(T).max()


def test_neumann_constant_in_kernel(oracle):
    """without Dirichlet faces A*1 = 0 (the LDG Laplacian annihilates constants)"""
    _, pr = _uniform(oracle, 8, 2, bctype=[1, 1, 1, 1], tau0=3.0)
    A = pr.dense("A")
    assert np.abs(A @ np.ones(pr.ndof)).max() < 1e-10 * np.abs(A).max()


@pytest.mark.parametrize("p", [1, 2])
def test_patch_test_linear(oracle, p):
    """g = x + y, f = 0: LDG reproduces the linear solution to round-off."""
    _, pr = _uniform(oracle, 4, p, tau0=1.0)
    X = pr.nodes()
    xq, _ = pr.bc_qpoints()
    b = pr.rhs(np.zeros(pr.ndof), xq[:, 0] + xq[:, 1])
    u = np.linalg.solve(pr.dense("A"), b)
    assert np.abs(u - (X[:, 0] + X[:, 1])).max() < 1e-10


@pytest.mark.parametrize("p", [1, 2])
def test_convergence_order(oracle, p):
    """manufactured solution of tests/cartesian/dg_tests.cpp:235-237 converges at O(h^{p+1})"""
    errs = []
    for n in (4, 8, 16):
        _, pr = _uniform(oracle, n, p)
        X = pr.nodes(); xq, _ = pr.bc_qpoints()
        uf = lambda Y: np.cos(2 * np.pi * Y[:, 0]) * np.cos(2 * np.pi * Y[:, 1]) + Y[:, 0] + Y[:, 1]
        ff = 8 * np.pi ** 2 * np.cos(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1])
        u = np.linalg.solve(pr.dense("A"), pr.rhs(ff, uf(xq)))
        errs.append(np.abs(u - uf(X)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > p + 0.6, rates


def test_ratrat_equals_rediscretisation(oracle):
    """On nested uniform trees the RATRAT coarse operators equal the LDG
    discretisation of the coarse layer (multigrid.h:115-157)."""
    _, fine = _uniform(oracle, 8, 2, tau0=2.0)
    _, coarse = _uniform(oracle, 4, 2, tau0=2.0)
    mg = oracle.Multigrid(fine, "gs", restriction="ratrat")
    for w in ["M", "T", "G0", "G1", "A"]:
        Ac = mg.dense(1, w); Ar = coarse.dense(w)
        assert np.abs(Ac - Ar).max() < 1e-11 * np.abs(Ar).max(), w


def test_rat_equals_dense_product(oracle):
    _, pr = _uniform(oracle, 8, 1, tau0=1.0)
    mg = oracle.Multigrid(pr, "gs", restriction="rat")
    A = pr.dense("A")
    for l in range(mg.nlevels - 1):
        T = mg.dense(l, "interp")
        ref = T.T @ mg.dense(l, "A") @ T
        assert np.abs(mg.dense(l + 1, "A") - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(mg.dense(0, "A") - A).max() == 0


def test_interpolation_reproduces_coarse_polynomials(oracle):
    qt, pr = _uniform(oracle, 8, 2)
    mg = oracle.Multigrid(pr, "gs")
    T = mg.dense(0, "interp")
    _, pc = _uniform(oracle, 4, 2)
    Xf, Xc = pr.nodes(), pc.nodes()
    for fn in (lambda X: X[:, 0] ** 2 - X[:, 1], lambda X: X[:, 0] * X[:, 1] ** 2):
        assert np.abs(T @ fn(Xc) - fn(Xf)).max() < 1e-13


@pytest.mark.parametrize("relax", ["gs", "mcgs", "jacobi"])
def test_smoother_against_dense(oracle, relax):
    _, pr = _uniform(oracle, 8, 1)
    mg = oracle.Multigrid(pr, relax)
    A = pr.dense("A"); b = pr.rhs(sine_f(pr.nodes()))
    B = pr.b; nb = pr.ne
    rng = np.random.default_rng(0)
    x0 = rng.standard_normal(pr.ndof)
    colors, nc = mg.colors(0)
    for reverse in (False, True):
        x = x0.copy()
        if relax == "jacobi":
            r = b - A @ x
            x = x + 0.8 * np.concatenate([np.linalg.solve(A[i*B:(i+1)*B, i*B:(i+1)*B], r[i*B:(i+1)*B]) for i in range(nb)])
        else:
            order = list(range(nb))
            if relax == "mcgs":
                cs = range(nc - 1, -1, -1) if reverse else range(nc)
                order = [i for c in cs for i in range(nb) if colors[i] == c]
            elif reverse:
                order = order[::-1]
            for i in order:
                s = slice(i * B, (i + 1) * B)
                t = b[s] - A[s] @ x + A[s, s] @ x[s]
                x[s] += 0.8 * (np.linalg.solve(A[s, s], t) - x[s])
        got = mg.relax(0, b, x0, 1, reverse)
        assert np.abs(got - x).max() < 1e-12 * np.abs(x).max()


def test_multicolour_is_valid_colouring(oracle):
    qt = oracle.Quadtree(2, lambda x: 1 / 16 if x[0] < 0.5 else 1 / 4)
    pr = oracle.Problem(qt, 1)
    mg = oracle.Multigrid(pr, "mcgs")
    for l in range(mg.nlevels):
        rp, cols, _, nbr, _ = mg.export(l, "A")
        c, nc = mg.colors(l)
        for i in range(nbr):
            for k in range(rp[i], rp[i + 1]):
                assert cols[k] == i or c[cols[k]] != c[i]


def test_vcycle_is_linear_and_matches_dense(oracle):
    """dense V (multigrid.h:386-400) built column by column is a fixed linear map"""
    _, pr = _uniform(oracle, 4, 1)
    mg = oracle.Multigrid(pr, "gs", npre=2, npost=2)
    n = pr.ndof
    V = np.array([mg.vcycle(e) for e in np.eye(n)]).T
    rng = np.random.default_rng(1)
    b = rng.standard_normal(n)
    assert np.abs(mg.vcycle(b) - V @ b).max() < 1e-12 * np.abs(V @ b).max()
    # as a preconditioner the V-cycle contracts: spectrum of I - V A inside the unit disc
    A = pr.dense("A")
    assert np.abs(np.linalg.eigvals(np.eye(n) - V @ A)).max() < 0.5


def test_pcg_matches_direct_solve(oracle):
    _, pr = _uniform(oracle, 16, 2)
    mg = oracle.Multigrid(pr, "gs", npre=2, npost=2)
    b = pr.rhs(sine_f(pr.nodes()))
    x, it, rr, _ = oracle.pcg(pr, mg, b, tol=1e-12, maxit=100, criterion="true")
    xs = np.linalg.solve(pr.dense("A"), b)
    assert rr < 1e-12 and np.abs(x - xs).max() < 1e-9 * np.abs(xs).max()
    # plain CG (identity preconditioner, krylov.h:78-82) reaches the same solution
    x2, it2, rr2, _ = oracle.pcg(pr, None, b, tol=1e-12, maxit=2000, criterion="true")
    assert rr2 < 1e-12 and it2 > it


def test_config1_iterations(oracle):
    """C1: 32x32, p=1, V(2,2) block GS, MG-PCG to 1e-10 (reference criterion)"""
    _, pr = _uniform(oracle, 32, 1)
    mg = oracle.Multigrid(pr, "gs", npre=2, npost=2)
    b = pr.rhs(sine_f(pr.nodes()))
    x, it, rr, _ = oracle.pcg(pr, mg, b, tol=1e-10, maxit=200)
    assert it == 12 and rr < 1e-8
    x, it, rr, _ = oracle.pcg(pr, mg, b, tol=1e-10, maxit=200, criterion="true")
    assert it == 14 and rr <= 1e-10


def test_pcg_edge_cases(oracle):
    _, pr = _uniform(oracle, 4, 1)
    mg = oracle.Multigrid(pr, "gs")
    x, it, rr, _ = oracle.pcg(pr, mg, np.zeros(pr.ndof))
    assert it == 0 and np.all(x == 0)
    b = pr.rhs(sine_f(pr.nodes()))
    x, it, rr, _ = oracle.pcg(pr, mg, b, maxit=1)
    assert it == 1
    bn = b.copy(); bn[3] = np.nan
    x, it, rr, _ = oracle.pcg(pr, mg, bn)
    assert it == 0   # NaN makes the loop test false at once (krylov.h:42)
