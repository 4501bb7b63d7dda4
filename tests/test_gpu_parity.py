"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: block patterns bit-exact; FP64 values and iterates within 1e-12
relative (different but fixed summation orders, explicit block inverses on the
GPU vs LDLT solves in the oracle); FP32 kernels within 1e-5 relative; MG-PCG
iteration counts equal to the oracle's under the same colouring, solutions to
1e-9 relative at tol 1e-10."""
import numpy as np
import pytest

from conftest import sine_f

pytestmark = pytest.mark.gpu


def h_adaptive(dx):
    """the driver's adaptive size function (tests/cartesian/dg_tests.cpp:212-215)"""
    return lambda x: dx if x[0] > 0.5 else (0.25 * dx if x[0] > 0.25 else 0.5 * dx)


MESHES = {
    "uniform8": (2, lambda x: 1 / 8),
    "adaptive": (2, h_adaptive(0.25)),
    "uniform3d": (3, lambda x: 1 / 4),
    # refined towards (0.3, 0.7) to depth 7 (the benchmark's rule at a small size)
    "point": (2, lambda x: max(2.0 ** -7, min(float(np.hypot(x[0] - 0.3, x[1] - 0.7)) / 28.8, 1 / 16))),
}


def build_pair(gctx, oracle, mesh, p, rapid=False, tau0=0.0, bc=None, balance=False):
    import mgldg
    N, h = MESHES[mesh]
    qt = mgldg.Quadtree(gctx, N, h, coarsening="rapid" if rapid else "equal", balance=balance)
    pr = mgldg.LDGPoisson(qt, p, bc=bc or "dirichlet", tau0=tau0)
    oqt = oracle.Quadtree(N, h, rapid=rapid, balance=balance)
    opr = oracle.Problem(oqt, p, bctype=[{"dirichlet": 0, "neumann": 1}[bc or "dirichlet"]] * (2 * N), tau0=tau0)
    return qt, pr, oqt, opr


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("mesh,p", [("uniform8", 1), ("uniform8", 2), ("uniform8", 3), ("adaptive", 2), ("uniform3d", 1)])
def test_operators_match_oracle(gctx, oracle, mesh, p):
    qt, pr, oqt, opr = build_pair(gctx, oracle, mesh, p, tau0=1.5)
    assert (pr.ne, pr.nf, pr.nbf, pr.npl, pr.ndof) == (opr.ne, opr.nf, opr.nbf, opr.b, opr.ndof)
    for w in ["M", "T", "A"] + [f"G{d}" for d in range(qt.N)]:
        a, b = pr.export(w), opr.export(w)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), w   # symbolic: bit-exact
        assert rel(a[2], b[2]) < 1e-13, w
    X = pr.nodes()
    assert np.array_equal(X, opr.nodes())
    xq, tag = pr.bc_qpoints()
    oq, otag = opr.bc_qpoints()
    assert np.array_equal(xq, oq) and np.array_equal(tag, otag)
    rng = np.random.default_rng(3)
    f, g = rng.standard_normal(pr.ndof), rng.standard_normal(pr.nqb)
    assert rel(pr.rhs(f, g), opr.rhs(f, g)) < 1e-13


@pytest.mark.parametrize("restriction", ["ratrat", "rat"])
@pytest.mark.parametrize("mesh,p,rapid", [("uniform8", 2, False), ("adaptive", 1, False), ("adaptive", 2, True)])
def test_hierarchy_matches_oracle(gctx, oracle, mesh, p, rapid, restriction):
    import mgldg
    qt, pr, oqt, opr = build_pair(gctx, oracle, mesh, p, rapid=rapid)
    mg = mgldg.Multigrid(pr, "gs", restriction=restriction)
    omg = oracle.Multigrid(opr, "mcgs", restriction=restriction)
    assert mg.nlevels == omg.nlevels
    for l in range(mg.nlevels):
        a, b = mg.export(l, "A"), omg.export(l, "A")
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), l
        assert rel(a[2], b[2]) < 1e-12, l
        c1, n1 = mg.colors(l)
        c2, n2 = omg.colors(l)
        assert n1 == n2 and np.array_equal(c1, c2)
        if l + 1 < mg.nlevels:
            ti, to = mg.export(l, "interp"), omg.export(l, "interp")
            assert np.array_equal(ti[1], to[1]) and rel(ti[2], to[2]) < 1e-15
        if restriction == "ratrat":
            for w in ["M", "T", "G0", "G1"]:
                a, b = mg.export(l, w), omg.export(l, w)
                assert np.array_equal(a[1], b[1]) and rel(a[2], b[2]) < 1e-12, (l, w)


@pytest.mark.parametrize("collapse", [0, 4096])
@pytest.mark.parametrize("relax", ["gs", "jacobi"])
def test_kernels_match_oracle(gctx, oracle, relax, collapse):
    import mgldg
    import torch
    qt, pr, oqt, opr = build_pair(gctx, oracle, "adaptive", 2)
    mg = mgldg.Multigrid(pr, relax, npre=2, npost=2, precision="mixed", coarse_dense_max=collapse)
    omg = oracle.Multigrid(opr, "mcgs" if relax == "gs" else "jacobi", npre=2, npost=2)
    rng = np.random.default_rng(7)
    dev = torch.device("cuda")
    for l in range(mg.nlevels):
        n = mg.level_rows[l] * pr.npl
        x = rng.standard_normal(n)
        xt = torch.tensor(x, device=dev)
        y = torch.zeros_like(xt)
        mg.spmv(l, xt, y)
        assert rel(y.cpu().numpy(), omg.spmv(l, x)) < 1e-13
        y32 = torch.zeros(n, dtype=torch.float32, device=dev)
        mg.spmv(l, xt.float(), y32, precision="mixed")
        assert rel(y32.double().cpu().numpy(), omg.spmv(l, x)) < 1e-5
        b = rng.standard_normal(n)
        for reverse in (False, True):
            xs = torch.tensor(x, device=dev)
            mg.smooth(l, torch.tensor(b, device=dev), xs, 2, reverse)
            assert rel(xs.cpu().numpy(), omg.relax(l, b, x, 2, reverse)) < 1e-12
    b = rng.standard_normal(pr.ndof)
    z = torch.zeros(pr.ndof, dtype=torch.float64, device=dev)
    mg.vcycle(torch.tensor(b, device=dev), z)
    zo = omg.vcycle(b)
    assert rel(z.cpu().numpy(), zo) < 1e-12
    mg.vcycle(torch.tensor(b, device=dev), z, precision="mixed")
    assert rel(z.cpu().numpy(), zo) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_paired_last_colours_match_oracle(gctx, oracle, precision):
    """Levels whose last two colours run in one launch (clusters of the
    last colour's rows with their neighbours, one workgroup each): one and two
    sweeps, forward and reverse, equal the oracle's colour-by-colour sweeps"""
    import mgldg
    import torch
    qt, pr, oqt, opr = build_pair(gctx, oracle, "point", 2, rapid=True, balance=True)
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, precision="mixed", coarse_dense_max=0)
    omg = oracle.Multigrid(opr, "mcgs", npre=2, npost=2)
    paired = []
    for l in range(mg.nlevels):
        try:
            mg.profile_relax(l, "fp64", 1, part="pair")
            paired.append(l)
        except mgldg.MGLDGError:
            pass
    assert 0 in paired, ("the fine level has no paired launch", paired)
    rng = np.random.default_rng(11)
    dt = torch.float64 if precision == "fp64" else torch.float32
    tol = 1e-12 if precision == "fp64" else 1e-5
    for l in paired:
        assert mg.colors(l)[1] >= 3
        n = mg.level_rows[l] * pr.npl
        x, b = rng.standard_normal(n), rng.standard_normal(n)
        for reverse in (False, True):
            for nsw in (1, 2):
                xs = torch.tensor(x, dtype=dt, device="cuda")
                mg.smooth(l, torch.tensor(b, dtype=dt, device="cuda"), xs, nsw, reverse, precision=precision)
                assert rel(xs.double().cpu().numpy(), omg.relax(l, b, x, nsw, reverse)) < tol, (l, reverse, nsw)


@pytest.mark.parametrize("collapse", [0, 4096])
@pytest.mark.parametrize("mesh,p", [("uniform8", 1), ("adaptive", 2), ("uniform3d", 1)])
@pytest.mark.parametrize("criterion", ["reference", "true"])
def test_pcg_matches_oracle(gctx, oracle, mesh, p, criterion, collapse):
    import mgldg
    qt, pr, oqt, opr = build_pair(gctx, oracle, mesh, p)
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, precision="mixed", coarse_dense_max=collapse)
    omg = oracle.Multigrid(opr, "mcgs", npre=2, npost=2)
    rng = np.random.default_rng(11)
    f = rng.standard_normal(pr.ndof)
    b = pr.rhs(f)
    xo, ito, rro, _ = oracle.pcg(opr, omg, opr.rhs(f), tol=1e-10, maxit=200, criterion=criterion)
    x, st = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion=criterion)
    assert st["status"] == "converged"
    assert st["iterations"] == ito
    assert rel(x, xo) < 1e-9
    xm, stm = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion=criterion, precision="mixed", flexible=True)
    assert stm["status"] == "converged" and abs(stm["iterations"] - ito) <= 1
    if criterion == "true":
        assert st["relres_true"] <= 1e-10 and stm["relres_true"] <= 1e-10
    assert rel(xm, xo) < 1e-8


_C2_ORACLE = {}


@pytest.mark.parametrize("form", ["fp64", "mixed", "mixed-graph", "mixed-batched"])
def test_adaptive_point_rule_pcg_matches_oracle(gctx, oracle, form):
    """The adaptive benchmark's configuration at depth 7 below the unit square
    (distance refinement towards (0.3, 0.7), 2:1 balanced, rapid coarsening,
    p = 3: 10,192 elements, 163,072 dofs) with everything the benchmark turns
    on: pooled operators, the collapsed coarse tail (8192), V(2,2)
    multicolour GS, RATRAT.  f ~ N(0,1) per dof.  MG-PCG to 1e-10 on the true
    residual against the oracle's PCG: FP64 iterations equal, mixed (FP32
    V-cycle, flexible beta; eager, graph-replayed and as 4 batched columns on
    the autotuned kernels) within one; solutions to 1e-9 / 1e-8."""
    import mgldg
    import torch
    qt, pr, oqt, opr = build_pair(gctx, oracle, "point", 3, rapid=True, balance=True)
    assert pr.ndof == 163072
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, precision="mixed")
    assert mg.level_pool(0) != (0, 0)
    rng = np.random.default_rng(0)
    k = 4 if form == "mixed-batched" else 1
    F = rng.standard_normal((pr.ndof, 4))[:, :k]     # column 0 is the same in every form
    for c in range(k):
        if c not in _C2_ORACLE:
            omg = oracle.Multigrid(opr, "mcgs", npre=2, npost=2, restriction="ratrat")
            _C2_ORACLE[c] = oracle.pcg(opr, omg, opr.rhs(F[:, c], np.zeros(opr.nqb)), tol=1e-10, maxit=200,
                                       criterion="true")[:2]
    prec = "fp64" if form == "fp64" else "mixed"
    opts = dict(tol=1e-10, maxit=200, criterion="true", precision=prec, flexible=prec == "mixed",
                use_graph=form == "mixed-graph")
    if k == 1:
        x, st = mgldg.pcg(mg, pr.rhs(F[:, 0]), **opts)
        X, stats = x[:, None], [st]
    else:
        Bt = torch.tensor(pr.rhs(F), device="cuda")
        Xt = torch.zeros_like(Bt)
        _, stats = mgldg.pcg(mg, Bt, Xt, **opts)
        X = Xt.cpu().numpy()
    for c in range(k):
        xo, ito = _C2_ORACLE[c]
        assert stats[c]["status"] == "converged" and stats[c]["relres_true"] <= 1e-10
        if form == "fp64":
            assert stats[c]["iterations"] == ito and rel(X[:, c], xo) < 1e-9
        else:
            assert abs(stats[c]["iterations"] - ito) <= 1 and rel(X[:, c], xo) < 1e-8, (c, stats[c], ito)


@pytest.mark.parametrize("criterion", ["reference", "true"])
def test_config1_iteration_parity(gctx, oracle, criterion):
    """C1: 32x32 p=1, V(2,2) block GS, MG-PCG to 1e-10.  The lexicographic GS
    (the reference's order, run as wavefronts) gives the reference algorithm's
    iterates and iteration count; the multicolour GS equals the oracle under
    the same colouring."""
    import mgldg
    qt = mgldg.Quadtree(gctx, 2, lambda x: 1 / 32)
    pr = mgldg.LDGPoisson(qt, 1)
    oqt = oracle.Quadtree(2, lambda x: 1 / 32)
    opr = oracle.Problem(oqt, 1)
    b = pr.rhs(sine_f(pr.nodes()))
    xl, it_lex, _, _ = oracle.pcg(opr, oracle.Multigrid(opr, "gs", npre=2, npost=2), b, tol=1e-10, maxit=200,
                                  criterion=criterion)
    _, it_mc, _, _ = oracle.pcg(opr, oracle.Multigrid(opr, "mcgs", npre=2, npost=2), b, tol=1e-10, maxit=200,
                                criterion=criterion)
    for collapse in (0, 4096):
        mgl = mgldg.Multigrid(pr, "gs-lex", npre=2, npost=2, precision="mixed", coarse_dense_max=collapse)
        x, st = mgldg.pcg(mgl, b, tol=1e-10, maxit=200, criterion=criterion)
        assert st["iterations"] == it_lex and rel(x, xl) < 1e-10
        xm, stm = mgldg.pcg(mgl, b, tol=1e-10, maxit=200, criterion=criterion, precision="mixed", flexible=True)
        assert abs(stm["iterations"] - it_lex) <= 1 and rel(xm, xl) < 1e-8
    mgc = mgldg.Multigrid(pr, "gs", npre=2, npost=2)
    x, st = mgldg.pcg(mgc, b, tol=1e-10, maxit=200, criterion=criterion)
    assert st["iterations"] == it_mc


@pytest.mark.parametrize("mesh,p,rapid", [("uniform8", 2, False), ("adaptive", 2, True)])
def test_lexicographic_gs_matches_reference_order(gctx, oracle, mesh, p, rapid):
    """wavefront-scheduled GS equals the oracle's sequential lexicographic sweep
    (multigrid.h:236-262) forward and reverse, on every level"""
    import mgldg
    import torch
    qt, pr, oqt, opr = build_pair(gctx, oracle, mesh, p, rapid=rapid)
    mg = mgldg.Multigrid(pr, "gs-lex", npre=2, npost=2)
    omg = oracle.Multigrid(opr, "gs", npre=2, npost=2)
    rng = np.random.default_rng(13)
    for l in range(mg.nlevels):
        n = mg.level_rows[l] * pr.npl
        x, b = rng.standard_normal(n), rng.standard_normal(n)
        for reverse in (False, True):
            xs = torch.tensor(x, device="cuda")
            mg.smooth(l, torch.tensor(b, device="cuda"), xs, 2, reverse)
            assert rel(xs.cpu().numpy(), omg.relax(l, b, x, 2, reverse)) < 1e-12
    bb = pr.rhs(rng.standard_normal(pr.ndof))
    xo, ito, _, _ = oracle.pcg(opr, omg, bb, tol=1e-10, maxit=200, criterion="true")
    x, st = mgldg.pcg(mg, bb, tol=1e-10, maxit=200, criterion="true")
    assert st["iterations"] == ito and rel(x, xo) < 1e-9


def test_graph_replay_matches_eager(gctx, oracle):
    """the captured V-cycle graph replayed many times gives the eager result"""
    import mgldg
    qt, pr, _, _ = build_pair(gctx, oracle, "adaptive", 2)
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, precision="mixed")
    b = pr.rhs(np.random.default_rng(5).standard_normal(pr.ndof))
    for prec in ("fp64", "mixed"):
        x0, s0 = mgldg.pcg(mg, b, tol=1e-10, criterion="true", precision=prec)
        for _ in range(3):
            x1, s1 = mgldg.pcg(mg, b, tol=1e-10, criterion="true", precision=prec, use_graph=True)
            assert s1["iterations"] == s0["iterations"]
            assert np.array_equal(x1, x0)
            # eager work between replays
            mgldg.pcg(mg, 2 * b, tol=1e-6, precision=prec)


@pytest.mark.parametrize("coarse", ["cholesky", "relaxation"])
def test_collapsed_vcycle_equals_levelwise(gctx, oracle, coarse):
    """the dense collapsed tail reproduces the level-by-level V-cycle"""
    import mgldg
    import torch
    qt, pr, oqt, opr = build_pair(gctx, oracle, "adaptive", 2, rapid=True)
    b = torch.tensor(np.random.default_rng(2).standard_normal(pr.ndof), device="cuda")
    outs, outs32 = [], []
    for cdm in (0, 64, 4096, 8192):
        mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, coarse=coarse, coarse_dense_max=cdm, precision="mixed")
        z = torch.zeros_like(b)
        mg.vcycle(b, z)
        outs.append(z.cpu().numpy())
        z32 = torch.zeros_like(b)
        mg.vcycle(b, z32, "mixed")          # FP32 V-cycle (FP32 copy of the collapsed tail)
        outs32.append(z32.cpu().numpy())
    omg = oracle.Multigrid(opr, "mcgs", npre=2, npost=2, coarse=coarse)
    zo = omg.vcycle(b.cpu().numpy())
    for z in outs:
        assert rel(z, zo) < 1e-12
    for z in outs32:
        assert rel(z, zo) < 1e-5


@pytest.mark.parametrize("relaxation,orelax", [("gs", "mcgs"), ("jacobi", "jacobi")])
def test_vcycle_without_presmoothing_is_stateless(gctx, oracle, relaxation, orelax):
    """V(0,2) with relaxation on the coarsest level: every level's x starts at
    zero (multigrid.h:376) although no pre-sweep clears it, so two V-cycles
    in a row (the second after one on another right-hand side) both equal
    the oracle's"""
    import mgldg
    import torch
    qt, pr, oqt, opr = build_pair(gctx, oracle, "adaptive", 2)
    mg = mgldg.Multigrid(pr, relaxation, npre=0, npost=2, coarse="relaxation", coarse_dense_max=0, precision="mixed")
    omg = oracle.Multigrid(opr, orelax, npre=0, npost=2, coarse="relaxation")
    rng = np.random.default_rng(11)
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-5)):
        for _ in range(2):
            bh = rng.standard_normal(pr.ndof)
            b = torch.tensor(bh, device="cuda")
            z = torch.zeros_like(b)
            mg.vcycle(b, z, prec)
            assert rel(z.cpu().numpy(), omg.vcycle(bh)) < tol, prec


def test_solve_statuses(gctx, oracle):
    import mgldg
    qt, pr, _, _ = build_pair(gctx, oracle, "uniform8", 1)
    mg = mgldg.Multigrid(pr, "gs")
    x, st = mgldg.pcg(mg, np.zeros(pr.ndof))
    assert st["iterations"] == 0 and np.all(x == 0)
    b = pr.rhs(sine_f(pr.nodes()))
    x, st = mgldg.pcg(mg, b, tol=1e-14, maxit=1)
    assert st["status"] == "maxit" and st["iterations"] == 1
    bn = b.copy()
    bn[5] = np.nan
    x, st = mgldg.pcg(mg, bn)
    assert st["status"] == "nonfinite"


def test_invalid_arguments_raise(gctx, oracle):
    import mgldg
    with pytest.raises(mgldg.MGLDGError):
        mgldg.LDGPoisson(mgldg.Quadtree(gctx, 2, lambda x: 0.25), 7)
    qt = mgldg.Quadtree(gctx, 2, lambda x: 2.0)     # a single element: no hierarchy
    pr = mgldg.LDGPoisson(qt, 1)
    with pytest.raises(mgldg.MGLDGError):
        mgldg.Multigrid(pr)
