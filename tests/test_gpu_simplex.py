"""GPU parity of the unstructured (simplex) path: the HIP path through the C ABI
against the CPU oracle on the reference's meshes (data/*.mesh) and the
reference MGridGen's partitions (tests/golden/partitions.json).

Bars: patterns bit-exact; assembled operators within 1e-13 relative (the
library builds the nodal simplex basis from monomials, the oracle from
Koornwinder polynomials as the reference does).  Coarse levels use
M-orthonormal bases of the agglomerates' Q_p box spaces that each side picks
by its own eigen-decomposition, so coarse quantities are compared through
basis-free forms (projectors, lifted operators, kernels conjugated by the
block-orthogonal map between the bases) within 1e-9; V-cycles within 1e-9
(FP64) and 1e-5 (FP32 V-cycle).  PCG iteration counts within one of the
oracle's under the same colouring; solutions to 1e-8 relative at tol 1e-10."""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def on_circle(x):
    return math.hypot(x[0], x[1]) > 1 - 1e-3   # tests/unstructured/dg_tests.cpp:14-17


def partitions(name):
    with open(os.path.join(HERE, "golden", "partitions.json")) as f:
        return [np.array(q, dtype=np.int32) for q in json.load(f)[name]]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)




def build_pair(gctx, oracle, name, p, tau0=100.0, tauD=1000.0, bndfuns=None, bc=None):
    import mgldg
    path = os.path.join(MESHES, name + ".mesh")
    wf = mgldg.Wireframe.read(gctx, path)
    pts, tris = oracle.read_mesh(path)
    if bndfuns is None and name.startswith("circle"):
        bndfuns = [on_circle]
    nb = len(bndfuns) if bndfuns else 0
    bct = bc if bc is not None else [0] * max(nb, 1)
    pr = mgldg.SimplexLDGPoisson(wf, p, bndfuns=bndfuns, bc=bct, tau0=tau0, tauD=tauD)
    opr = oracle.SxProblem(pts, tris, p, bndfuns=bndfuns, bctype=bct, tau0=tau0, tauD=tauD)
    return wf, pr, opr


def test_wireframe_reader(gctx, oracle):
    import mgldg
    for name in ["circle", "square_8"]:
        path = os.path.join(MESHES, name + ".mesh")
        p, t = mgldg.Wireframe.read(gctx, path).arrays()
        po, to = oracle.read_mesh(path)
        assert np.array_equal(p, po) and np.array_equal(t, to)
    with pytest.raises(mgldg.MGLDGError):
        mgldg.Wireframe.read(gctx, os.path.join(MESHES, "no_such.mesh"))


@pytest.mark.parametrize("name,p", [("circle", 2), ("circle2", 0), ("circle2", 1), ("circle2", 2), ("circle2", 3),
                                    ("square_16", 2)])
def test_operators_match_oracle(gctx, oracle, name, p):
    wf, pr, opr = build_pair(gctx, oracle, name, p)
    assert (pr.ne, pr.nf, pr.nbf, pr.npl, pr.ndof, pr.nqb) == (opr.ne, opr.nf, opr.nbf, opr.b, opr.ndof, opr.nqb)
    for w in ["M", "T", "A", "G0", "G1"]:
        a, b = pr.export(w), opr.export(w)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), w
        assert rel(a[2], b[2]) < 1e-13, w
    assert rel(pr.nodes(), opr.nodes()) < 1e-15
    xq, tag = pr.bc_qpoints()
    oq, otag = opr.bc_qpoints()
    assert rel(xq, oq) < 1e-15 and np.array_equal(tag, otag)
    rng = np.random.default_rng(3)
    f, g = rng.standard_normal(pr.ndof), rng.standard_normal(pr.nqb)
    assert rel(pr.rhs(f, g), opr.rhs(f, g)) < 1e-13


def test_mixed_boundary_keys(gctx, oracle):
    fns = [lambda x: x[0] < 1e-12, lambda x: x[1] > 1 - 1e-12, lambda x: True]
    wf, pr, opr = build_pair(gctx, oracle, "square_8", 2, bndfuns=fns, bc=[0, 1, 0], tau0=10)
    for w in ["A", "T", "G0", "G1"]:
        a, b = pr.export(w), opr.export(w)
        assert np.array_equal(a[1], b[1]) and rel(a[2], b[2]) < 1e-13
    xq, tag = pr.bc_qpoints()
    assert set(tag) == {1, 2, 3}
    rng = np.random.default_rng(4)
    g = rng.standard_normal(pr.nqb)
    assert rel(pr.rhs(None, g), opr.rhs(None, g)) < 1e-13


def dense(exported):
    rp, cols, blk, nbr, nbc = exported
    br, bc = blk.shape[1], blk.shape[2]
    D = np.zeros((nbr * br, nbc * bc))
    for i in range(nbr):
        for k in range(rp[i], rp[i + 1]):
            D[i * br:(i + 1) * br, cols[k] * bc:(cols[k] + 1) * bc] += blk[k]
    return D


def composites(mg, nlevels):
    """Q_l = T_0 ... T_{l-1}: level-l coefficients -> fine nodal values"""
    Q = [None]
    for l in range(nlevels - 1):
        T = dense(mg.export(l, "interp"))
        Q.append(T if l == 0 else Q[-1] @ T)
    return Q


@pytest.mark.parametrize("name,p", [("circle2", 1), ("circle2", 2), ("square_16", 2), ("square_16", 3)])
def test_agglomerated_hierarchy_matches_oracle(gctx, oracle, name, p):
    """Coarse bases are M-orthonormal bases of the agglomerates' box spaces,
    chosen independently by the library and the oracle, so the comparison is
    through basis-free quantities: the M-orthogonal projector onto each level
    (Q Q^T M) and the coarse operator lifted to the mesh (Q A_l Q^T)."""
    import mgldg
    wf, pr, opr = build_pair(gctx, oracle, name, p)
    parts = partitions(name)
    mg = mgldg.Multigrid(pr, "gs", restriction="rat", partitions=parts, coarse_dense_max=0)
    omg = oracle.SxMultigrid(opr, parts, relaxation="mcgs", restriction="rat")
    assert mg.nlevels == omg.nlevels == len(parts) + 1
    M = opr.dense("M")
    Qg, Qo = composites(mg, mg.nlevels), composites(omg, mg.nlevels)
    Ag = [dense(mg.export(l, "A")) for l in range(mg.nlevels)]
    for l in range(mg.nlevels):
        assert (mg.level_rows[l], mg.level_B[l]) == omg.level_dims(l)
        a, b = mg.export(l, "A"), omg.export(l, "A")
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), l     # block patterns
        c1, n1 = mg.colors(l)
        c2, n2 = omg.colors(l)
        assert n1 == n2 and np.array_equal(c1, c2)
        if l == 0:
            assert rel(a[2], b[2]) < 1e-13
            continue
        Q = Qg[l]
        G = Q.T @ M @ Q
        d = np.diag(G)
        assert np.abs(G - np.diag(d)).max() < 1e-9 and np.all((np.abs(d - 1) < 1e-9) | (np.abs(d) < 1e-9)), l
        assert rel(Q @ Q.T, Qo[l] @ Qo[l].T) < 1e-9, l
        # Galerkin on the GPU: A_l = T^T A_{l-1} T, unit diagonal on dropped directions
        T = dense(mg.export(l - 1, "interp"))
        dropped = np.abs(T).max(axis=0) == 0
        assert rel(Ag[l], T.T @ Ag[l - 1] @ T + np.diag(dropped.astype(float))) < 1e-12, l
        Ao = dense(omg.export(l, "A"))
        assert rel(Q @ Ag[l] @ Q.T, Qo[l] @ Ao @ Qo[l].T) < 1e-9, l
    with pytest.raises(mgldg.MGLDGError):
        mgldg.Multigrid(pr, "gs", restriction="ratrat", partitions=parts)


@pytest.mark.parametrize("relax", ["gs", "jacobi"])
def test_agglomerated_kernels_match_oracle(gctx, oracle, relax):
    """Level kernels commute with the block-orthogonal map O = Q^T M Q_oracle
    between the two coarse bases: spmv(O x) = O spmv_oracle(x), likewise the
    smoothers (block smoothers do not depend on the basis within a block)."""
    import mgldg
    import torch
    wf, pr, opr = build_pair(gctx, oracle, "circle2", 2)
    parts = partitions("circle2")
    om = 0.6 if relax == "jacobi" else 0.8
    mg = mgldg.Multigrid(pr, relax, omega=om, npre=2, npost=2, restriction="rat",
                         precision="mixed", partitions=parts, coarse_dense_max=0)
    omg = oracle.SxMultigrid(opr, parts, relaxation="mcgs" if relax == "gs" else "jacobi",
                             omega=om, npre=2, npost=2, restriction="rat")
    M = opr.dense("M")
    Qg, Qo = composites(mg, mg.nlevels), composites(omg, mg.nlevels)
    rng = np.random.default_rng(7)
    for l in range(mg.nlevels):
        n = mg.level_rows[l] * mg.level_B[l]
        O = np.eye(n) if l == 0 else Qg[l].T @ M @ Qo[l]
        bar = 1e-12 if l == 0 else 1e-9
        x = rng.standard_normal(n)
        xt = torch.tensor(O @ x, device="cuda")
        y = torch.zeros_like(xt)
        mg.spmv(l, xt, y)
        assert rel(y.cpu().numpy(), O @ omg.spmv(l, x)) < bar, l
        b = rng.standard_normal(n)
        for reverse in (False, True):
            xs = torch.tensor(O @ x, device="cuda")
            mg.smooth(l, torch.tensor(O @ b, device="cuda"), xs, 2, reverse)
            assert rel(xs.cpu().numpy(), O @ omg.relax(l, b, x, 2, reverse)) < bar, (l, reverse)
    b = rng.standard_normal(pr.ndof)
    z = torch.zeros(pr.ndof, dtype=torch.float64, device="cuda")
    mg.vcycle(torch.tensor(b, device="cuda"), z)
    zo = omg.vcycle(b)
    assert rel(z.cpu().numpy(), zo) < 1e-9
    mg.vcycle(torch.tensor(b, device="cuda"), z, precision="mixed")
    assert rel(z.cpu().numpy(), zo) < 1e-5


@pytest.mark.parametrize("collapse", [0, 4096])
@pytest.mark.parametrize("relax", ["gs", "gs-lex", "jacobi"])
def test_agglomerated_pcg_matches_oracle(gctx, oracle, relax, collapse):
    import mgldg
    wf, pr, opr = build_pair(gctx, oracle, "circle2", 2)
    parts = partitions("circle2")
    om = 0.6 if relax == "jacobi" else 0.8
    mg = mgldg.Multigrid(pr, relax, omega=om, npre=2, npost=2, restriction="rat", precision="mixed",
                         partitions=parts, coarse_dense_max=collapse)
    orelax = {"gs": "mcgs", "gs-lex": "gs", "jacobi": "jacobi"}[relax]
    omg = oracle.SxMultigrid(opr, parts, relaxation=orelax, omega=om, npre=2, npost=2, restriction="rat")
    rng = np.random.default_rng(11)
    b = opr.rhs(rng.standard_normal(opr.ndof), np.zeros(opr.nqb))
    xo, ito, rro, _ = oracle.pcg(opr, omg, b, tol=1e-10, maxit=200, criterion="true")
    x, st = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion="true")
    assert st["status"] == "converged" and abs(st["iterations"] - ito) <= 1
    assert rel(x, xo) < 1e-8
    xm, stm = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion="true", precision="mixed", flexible=True)
    print("iterations oracle / fp64 / mixed:", ito, st["iterations"], stm["iterations"])
    assert stm["status"] == "converged" and stm["iterations"] <= ito + 3
    assert rel(xm, xo) < 1e-8


@pytest.mark.parametrize("name,p", [("circle2", 2), ("square_32", 2), ("ball", 2)])
def test_own_agglomeration(gctx, oracle, name, p):
    """The library's agglomerator (MGridGen's multilevel algorithm): valid
    nested partitions of at most 2^N agglomerates of the finer level,
    connected agglomerates, and MG-PCG within one iteration of the solver on
    MGridGen's partitions (tests/golden/partitions.json)."""
    import mgldg
    if name == "ball":
        path = os.path.join(MESHES, "ball.mesh")
        wf = mgldg.Wireframe.read(gctx, path, N=3)
        pr = mgldg.SimplexLDGPoisson(wf, p, tau0=100.0, tauD=1000.0)
        pts, tets = oracle.read_mesh(path)
        opr = oracle.SxProblem(pts, tets, p, tau0=100.0, tauD=1000.0)
    else:
        wf, pr, opr = build_pair(gctx, oracle, name, p)
    N = opr.N
    parts = mgldg.agglomerate(pr)
    assert len(parts) >= 3
    xadj, adj, w, vvol, vsurf = opr.agglom_graph()
    orig = np.arange(pr.ne)
    n = pr.ne
    for q in parts:
        assert len(q) == n
        sizes = np.bincount(q)
        assert sizes.min() >= 1 and sizes.max() <= 1 << N
        # connectivity of each agglomerate in the level graph
        lx, la, _, _, _ = opr.agglom_graph(None if n == pr.ne else orig, n if n != pr.ne else None)
        for a in range(len(sizes)):
            mem = set(np.nonzero(q == a)[0])
            start = next(iter(mem))
            seen, stack = {start}, [start]
            while stack:
                u = stack.pop()
                for k in range(lx[u], lx[u + 1]):
                    v = int(la[k])
                    if v in mem and v not in seen:
                        seen.add(v); stack.append(v)
            assert seen == mem
        orig = q[orig]
        n = len(sizes)
    with pytest.raises(mgldg.MGLDGError):
        mgldg.Multigrid(pr, "gs")      # RATRAT (the default) is rejected on agglomerated hierarchies
    mg_own = mgldg.Multigrid(pr, "gs", npre=2, npost=2, restriction="rat")
    mg_ref = mgldg.Multigrid(pr, "gs", npre=2, npost=2, restriction="rat", partitions=partitions(name))
    b = opr.rhs(np.random.default_rng(5).standard_normal(opr.ndof), np.zeros(opr.nqb))
    _, s1 = mgldg.pcg(mg_own, b, tol=1e-10, criterion="true", maxit=400)
    _, s2 = mgldg.pcg(mg_ref, b, tol=1e-10, criterion="true", maxit=400)
    print(name, "iterations: library partitions", s1["iterations"], "MGridGen's", s2["iterations"])
    assert s1["status"] == s2["status"] == "converged"
    assert abs(s1["iterations"] - s2["iterations"]) <= 1
    # the oracle on the library's partitions gives the same solver
    omg = oracle.SxMultigrid(opr, parts, relaxation="mcgs", npre=2, npost=2, restriction="rat")
    _, ito, _, _ = oracle.pcg(opr, omg, b, tol=1e-10, criterion="true", maxit=400)
    assert abs(ito - s1["iterations"]) <= 1


def test_batched_and_adjoint_on_simplices(gctx, oracle):
    import mgldg
    import torch
    fns = [lambda x: x[1] >= 0, lambda x: True]
    wf, pr, opr = build_pair(gctx, oracle, "circle2", 2, bndfuns=fns, bc=[0, 1])
    parts = partitions("circle2")
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, restriction="rat", precision="mixed", partitions=parts)
    rng = np.random.default_rng(9)
    k = 8
    B = np.stack([opr.rhs(rng.standard_normal(opr.ndof), rng.standard_normal(opr.nqb)) for _ in range(k)], axis=1)
    for prec in ("fp64", "mixed"):
        X, sts = mgldg.pcg(mg, torch.tensor(B, device="cuda"), tol=1e-10, criterion="true", precision=prec,
                           flexible=prec == "mixed")
        Xc = X.cpu().numpy()
        for c in range(k):
            xs, s1 = mgldg.pcg(mg, B[:, c], tol=1e-10, criterion="true", precision=prec, flexible=prec == "mixed")
            assert sts[c]["status"] == "converged"
            assert sts[c]["iterations"] == s1["iterations"]
            assert rel(Xc[:, c], xs) < 1e-12
    # device rhs and the adjoint maps
    f = rng.standard_normal(pr.ndof)
    g = rng.standard_normal(pr.nqb)
    bd = mgldg.rhs_device(mg, torch.tensor(f, device="cuda"), torch.tensor(g, device="cuda"))
    assert rel(bd.cpu().numpy(), opr.rhs(f, g)) < 1e-12
    dLdu = rng.standard_normal(pr.ndof)
    lam, dLdf, dLdg, st = mgldg.adjoint(mg, torch.tensor(dLdu, device="cuda"), tol=1e-12)
    A = opr.dense("A")
    lam_ref = np.linalg.solve(A, dLdu)
    assert rel(lam.cpu().numpy(), lam_ref) < 1e-9
    assert rel(dLdf.cpu().numpy(), opr.dense("M") @ lam_ref) < 1e-9
    assert rel(dLdg.cpu().numpy(), opr.rhs_g_adjoint(lam_ref)) < 1e-9


def test_invalid_simplex_inputs(gctx, oracle):
    import mgldg
    path = os.path.join(MESHES, "square_8.mesh")
    wf = mgldg.Wireframe.read(gctx, path)
    with pytest.raises(mgldg.MGLDGError):
        mgldg.SimplexLDGPoisson(wf, 1, bndfuns=[lambda x: x[0] < 1e-12])   # a face no predicate covers
    with pytest.raises(mgldg.MGLDGError):
        mgldg.SimplexLDGPoisson(wf, 6)
    pr = mgldg.SimplexLDGPoisson(wf, 1)
    with pytest.raises(mgldg.MGLDGError):
        mgldg.Multigrid(pr, "gs", partitions=[np.full(pr.ne, -1, dtype=np.int32)])
    with pytest.raises(mgldg.MGLDGError):
        mgldg.Multigrid(pr, "gs", partitions=[np.arange(pr.ne, dtype=np.int32) + pr.ne])
    pts, tris = oracle.read_mesh(path)
    with pytest.raises(mgldg.MGLDGError):
        mgldg.Wireframe(gctx, pts, np.where(tris == 0, 10 ** 6, tris))


@pytest.mark.parametrize("p", [1, 2])
def test_ball_3d_matches_oracle(gctx, oracle, p):
    """Tetrahedra (data/ball.mesh): operators, RHS and the agglomerated
    MG-PCG (MGridGen partitions, at most 8 per agglomerate, 10 x 27
    interpolation blocks at p = 2) against the oracle."""
    import mgldg
    path = os.path.join(MESHES, "ball.mesh")
    wf = mgldg.Wireframe.read(gctx, path, N=3)
    pts, tets = oracle.read_mesh(path)
    pr = mgldg.SimplexLDGPoisson(wf, p, tau0=100.0, tauD=1000.0)
    opr = oracle.SxProblem(pts, tets, p, tau0=100.0, tauD=1000.0)
    assert (pr.ne, pr.nf, pr.nbf, pr.npl, pr.nqb) == (opr.ne, opr.nf, opr.nbf, opr.b, opr.nqb)
    for w in ["M", "T", "A", "G0", "G1", "G2"]:
        a, b = pr.export(w), opr.export(w)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), w
        assert rel(a[2], b[2]) < 1e-13, w
    xq, tag = pr.bc_qpoints()
    oq, otag = opr.bc_qpoints()
    assert rel(xq, oq) < 1e-15 and np.array_equal(tag, otag)
    rng = np.random.default_rng(12)
    f, g = rng.standard_normal(pr.ndof), rng.standard_normal(pr.nqb)
    b = opr.rhs(f, g)
    assert rel(pr.rhs(f, g), b) < 1e-13
    parts = partitions("ball")
    mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, restriction="rat", precision="mixed", partitions=parts)
    assert mg.level_B == [pr.npl] + [(p + 1) ** 3] * len(parts)
    omg = oracle.SxMultigrid(opr, parts, relaxation="mcgs", npre=2, npost=2, restriction="rat")
    xo, ito, _, _ = oracle.pcg(opr, omg, b, tol=1e-10, maxit=200, criterion="true")
    x, st = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion="true")
    assert st["status"] == "converged" and abs(st["iterations"] - ito) <= 1
    assert rel(x, xo) < 1e-8
    xm, stm = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion="true", precision="mixed", flexible=True)
    assert stm["status"] == "converged" and rel(xm, xo) < 1e-8
    # the library's own agglomeration (at most 8 per agglomerate)
    own = mgldg.agglomerate(pr)
    assert max(np.bincount(q).max() for q in own) <= 8
    mg2 = mgldg.Multigrid(pr, "gs", npre=2, npost=2, restriction="rat")
    _, st2 = mgldg.pcg(mg2, b, tol=1e-10, maxit=200, criterion="true")
    assert st2["status"] == "converged"
