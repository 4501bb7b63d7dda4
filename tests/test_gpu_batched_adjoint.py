"""GPU: batched right-hand sides and the adjoint, against single solves and the
oracle (FP64)."""
import numpy as np
import pytest

from test_gpu_parity import build_pair, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("collapse", [0, 4096])
def test_batched_matches_single(gctx, oracle, precision, collapse):
    import mgldg
    import torch
    qt, pr, _, opr = build_pair(gctx, oracle, "adaptive", 2, rapid=True)
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, precision="mixed", coarse_dense_max=collapse)
    rng = np.random.default_rng(21)
    k = 5
    F = rng.standard_normal((pr.ndof, k))
    F[:, 3] = 0.0                                   # a zero right-hand side converges at once
    B = pr.rhs(F)
    opts = dict(tol=1e-10, maxit=200, criterion="true", precision=precision, flexible=precision == "mixed")
    Bt = torch.tensor(B, device="cuda")
    Xt = torch.zeros_like(Bt)
    _, stats = mgldg.pcg(mg, Bt, Xt, **opts)
    X = Xt.cpu().numpy()
    for c in range(k):
        x1, s1 = mgldg.pcg(mg, B[:, c].copy(), **opts)
        assert stats[c]["status"] == "converged"
        assert abs(stats[c]["iterations"] - s1["iterations"]) <= (0 if precision == "fp64" else 1)
        if c == 3:
            assert stats[c]["iterations"] == 0 and np.all(X[:, c] == 0)
            continue
        assert stats[c]["relres_true"] <= 1e-10
        assert rel(X[:, c], x1) < 1e-8


@pytest.mark.parametrize("k", [5, 40, 70])
def test_batched_collapsed_tail_equals_levelwise(gctx, oracle, k):
    """One batched PCG step (x = alpha V(b)) with the coarse tail collapsed
    into the MFMA GEMM equals the level-by-level V-cycle's: rounding apart
    in FP64, within the FP32 copy's rounding in the FP32 V-cycle.  k = 5,
    40, 70 run 1, 2 and 4 column tiles per wavefront, the last two with a
    partial last tile or tile group."""
    import mgldg
    import torch
    qt, pr, _, _ = build_pair(gctx, oracle, "adaptive", 2, rapid=True)
    B = torch.tensor(pr.rhs(np.random.default_rng(23).standard_normal((pr.ndof, k))), device="cuda")
    for prec, bound in (("fp64", 1e-12), ("mixed", 1e-5)):
        out = []
        for cdm in (0, 4096):
            mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, precision="mixed", coarse_dense_max=cdm)
            X = torch.zeros_like(B)
            mgldg.pcg(mg, B, X, tol=0.0, maxit=1, criterion="true", precision=prec, flexible=prec == "mixed")
            out.append(X.cpu().numpy())
        assert np.all(np.isfinite(out[1]))
        for c in range(k):
            assert rel(out[1][:, c], out[0][:, c]) < bound, (prec, c)


def test_batched_nonfinite_column_is_isolated(gctx, oracle):
    import mgldg
    import torch
    qt, pr, _, _ = build_pair(gctx, oracle, "uniform8", 1)
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2)
    B = pr.rhs(np.random.default_rng(2).standard_normal((pr.ndof, 3)))
    B[7, 1] = np.nan
    Bt = torch.tensor(B, device="cuda")
    Xt = torch.zeros_like(Bt)
    _, st = mgldg.pcg(mg, Bt, Xt, tol=1e-10, criterion="true")
    assert st[1]["status"] == "nonfinite"
    assert st[0]["status"] == "converged" and st[2]["status"] == "converged"


@pytest.mark.parametrize("mesh,p,k", [("uniform8", 1, 1), ("adaptive", 2, 1), ("adaptive", 2, 3)])
def test_device_rhs_and_adjoint_match_oracle(gctx, oracle, mesh, p, k):
    import mgldg
    import torch
    qt, pr, oqt, opr = build_pair(gctx, oracle, mesh, p, tau0=0.5)
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2)
    omg = oracle.Multigrid(opr, "mcgs", npre=2, npost=2)
    rng = np.random.default_rng(4)
    F = rng.standard_normal((pr.ndof, k)); G = rng.standard_normal((pr.nqb, k)); D = rng.standard_normal((pr.ndof, k))
    sq = (lambda a: a[:, 0]) if k == 1 else (lambda a: a)
    ft, gt, dt = [torch.tensor(sq(a).copy(), device="cuda") for a in (F, G, D)]
    b = mgldg.rhs_device(mg, ft, gt).cpu().numpy().reshape(pr.ndof, k)
    for c in range(k):
        assert rel(b[:, c], opr.rhs(F[:, c], G[:, c])) < 1e-13
    lam, dLdf, dLdg, st = mgldg.adjoint(mg, dt, tol=1e-12)
    lam, dLdf, dLdg = [a.cpu().numpy().reshape(-1, k) for a in (lam, dLdf, dLdg)]
    for c in range(k):
        lo, _, _, _ = oracle.pcg(opr, omg, D[:, c], tol=1e-13, maxit=300, criterion="true")
        assert rel(lam[:, c], lo) < 1e-9
        assert rel(dLdf[:, c], opr.rhs(lo, np.zeros(opr.nqb))) < 1e-9
        assert rel(dLdg[:, c], opr.rhs_g_adjoint(lo)) < 1e-9


def test_autograd_gradients_against_finite_differences(gctx, oracle):
    """L = 0.5 ||u(f, g) - u*||^2: the adjoint gradient equals central finite
    differences (eps = 1e-6, FP64) along random directions."""
    import mgldg
    import torch
    qt, pr, _, _ = build_pair(gctx, oracle, "adaptive", 2, tau0=0.5)
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2)
    rng = np.random.default_rng(9)
    dev = "cuda"
    f = torch.tensor(rng.standard_normal(pr.ndof), device=dev, requires_grad=True)
    g = torch.tensor(rng.standard_normal(pr.nqb), device=dev, requires_grad=True)
    ustar = torch.tensor(rng.standard_normal(pr.ndof), device=dev)

    def loss(ff, gg):
        u = mgldg.ldg_solve(ff, gg, mg, tol=1e-13)
        return 0.5 * torch.sum((u - ustar) ** 2)

    L = loss(f, g)
    L.backward()
    # L is quadratic in (f, g), so central differences carry no truncation
    # error; what remains is the solve error (tol 1e-13) divided by 2 eps.
    for eps, bound in ((1e-6, 1e-4), (1e-3, 1e-7)):
        for _ in range(4):
            df = torch.tensor(rng.standard_normal(pr.ndof), device=dev)
            dg = torch.tensor(rng.standard_normal(pr.nqb), device=dev)
            with torch.no_grad():
                fd = (loss(f + eps * df, g + eps * dg) - loss(f - eps * df, g - eps * dg)) / (2 * eps)
            an = torch.dot(f.grad, df) + torch.dot(g.grad, dg)
            assert abs(fd.item() - an.item()) <= bound * abs(an.item()), (eps, fd.item(), an.item())


@pytest.mark.parametrize("case", ["quadtree_p3", "quadtree_p2", "simplex_p2", "ball_p2"])
def test_mfma_batched_kernels(gctx, oracle, case):
    """The MFMA block kernels (forced 1, 2, 4 tiles of 16 columns, and the
    autotuned plan) and the vector kernels against the oracle's PCG
    (krylov.h:22-68) column by column under the same colouring and
    hierarchy: block sizes 16 (p = 3 quadtree), 9 (p = 2 quadtree), 6 / 9
    (p = 2 triangles with box coarse levels) and 10 / 27 (p = 2 tetrahedra,
    two MFMA row tiles), GS and Jacobi, FP64 and FP32 V-cycles.  FP64
    quadtree solves take the oracle's iteration count exactly; mixed solves
    and agglomerated hierarchies (whose coarse bases are chosen
    independently, so iterates agree only up to rounding) within one."""
    import os
    import mgldg
    import torch
    here = os.path.dirname(os.path.abspath(__file__))
    if case.startswith("quadtree"):
        p = 3 if case.endswith("p3") else 2
        qt, pr, _, opr = build_pair(gctx, oracle, "adaptive", p, rapid=True)
        relax = "gs"
        mg = mgldg.Multigrid(pr, relax, npre=2, npost=2, precision="mixed", coarse_dense_max=0)
        omg = oracle.Multigrid(opr, "mcgs", npre=2, npost=2, restriction="ratrat")
    elif case == "simplex_p2":
        path = os.path.join(here, "golden", "meshes", "circle2.mesh")
        wf = mgldg.Wireframe.read(gctx, path)
        pr = mgldg.SimplexLDGPoisson(wf, 2, tau0=100.0, tauD=1000.0)
        relax = "jacobi"
        parts = mgldg.agglomerate(pr)
        mg = mgldg.Multigrid(pr, relax, omega=0.6, npre=2, npost=2, restriction="rat", precision="mixed",
                             coarse_dense_max=0, partitions=parts)
        pts, tris = oracle.read_mesh(path)
        opr = oracle.SxProblem(pts, tris, 2, tau0=100.0, tauD=1000.0)
        omg = oracle.SxMultigrid(opr, parts, relaxation="jacobi", omega=0.6, npre=2, npost=2, restriction="rat")
    else:   # tetrahedra: 10-wide fine blocks, 27-wide box blocks (two MFMA row tiles)
        path = os.path.join(here, "golden", "meshes", "ball.mesh")
        wf = mgldg.Wireframe.read(gctx, path, N=3)
        pr = mgldg.SimplexLDGPoisson(wf, 2, tau0=100.0, tauD=1000.0)
        relax = "gs"
        parts = mgldg.agglomerate(pr)
        mg = mgldg.Multigrid(pr, relax, npre=2, npost=2, restriction="rat", precision="mixed", coarse_dense_max=0,
                             partitions=parts)
        pts, tets = oracle.read_mesh(path)
        opr = oracle.SxProblem(pts, tets, 2, tau0=100.0, tauD=1000.0)
        omg = oracle.SxMultigrid(opr, parts, relaxation="mcgs", npre=2, npost=2, restriction="rat")
    rng = np.random.default_rng(31)
    # a partial last tile; on the p = 3 quadtree also a k that 4 does not
    # divide (a wave's NT columns are one vector access: 4 tiles run as 2)
    k = 22 if case == "quadtree_p3" else 40
    Bh = pr.rhs(rng.standard_normal((pr.ndof, k)))
    B = torch.tensor(Bh, device="cuda")
    res = {}
    for mode in ("vector", 1, 2, 4, "auto"):
        mg.batch_kernels(mode)
        for prec in ("fp64", "mixed"):
            X = torch.zeros_like(B)
            _, st = mgldg.pcg(mg, B, X, tol=1e-10, maxit=300, criterion="true", precision=prec,
                              flexible=prec == "mixed")
            assert all(s["status"] == "converged" for s in st), (mode, prec)
            res[(mode, prec)] = (X.cpu().numpy(), [s["iterations"] for s in st])
        if mode == "auto":
            plan = mg.batch_plan()
            assert any(lv["apply"] > 0 or lv["smooth"] > 0 for lv in plan["fp64"] + plan["fp32"])
    for prec in ("fp64", "mixed"):
        X0, it0 = res[("vector", prec)]
        for mode in (1, 2, 4, "auto"):
            X1, it1 = res[(mode, prec)]
            assert max(abs(a - b) for a, b in zip(it0, it1)) <= 1, (mode, prec)
            assert rel(X1, X0) < 1e-8, (mode, prec)
    # each column against the oracle: the first and last column of the first
    # tile, the first of the second, the last of the batch (all of them on
    # the quadtrees, where the oracle is fast)
    cols = range(k) if case.startswith("quadtree") else (0, 15, 16, k - 1)
    exact = case.startswith("quadtree")
    for c in cols:
        xo, ito, _, _ = oracle.pcg(opr, omg, Bh[:, c].copy(), tol=1e-10, maxit=300, criterion="true")
        for mode in ("vector", 1, 2, 4, "auto"):
            for prec in ("fp64", "mixed"):
                X, its = res[(mode, prec)]
                slack = 0 if (exact and prec == "fp64") else 1
                assert abs(its[c] - ito) <= slack, (c, mode, prec, its[c], ito)
                assert rel(X[:, c], xo) < (1e-9 if prec == "fp64" and exact else 1e-8), (c, mode, prec)
