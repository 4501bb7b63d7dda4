"""Pins of the oracle's unstructured (simplex) restatement against closed forms,
invariants and the reference's data (data/*.mesh, MGridGen partitions).

Reference: src/unstructured/{dg/master.h, mesh/mesh.h, dg/ldgpoisson.tpp,
mesh/agglomeration.h, multigrid/interpolationhierarchy.h}."""
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def mesh(name):
    return po.read_mesh(os.path.join(MESHES, name + ".mesh"))


def partitions(name):
    with open(os.path.join(HERE, "golden", "partitions.json")) as f:
        return [np.array(q, dtype=np.int32) for q in json.load(f)[name]]


def on_circle(x):
    return math.hypot(x[0], x[1]) > 1 - 1e-3   # tests/unstructured/dg_tests.cpp:14-17


# ------------------------------------------------------------------ master element
def simplex_order(N, P):
    """SimplexRange<N,P> order (range.h:291-309): entries >= 0 with sum <= P-1,
    last coordinate fastest."""
    return [t for t in itertools.product(range(P), repeat=N) if sum(t) <= P - 1]


def monomial_moment(a):
    """exact integral of prod x_i^a_i over the unit simplex: prod a_i! / (sum a + N)!"""
    return math.prod(math.factorial(v) for v in a) / math.factorial(sum(a) + len(a))


@pytest.mark.parametrize("N,p", [(2, 0), (2, 1), (2, 2), (2, 3), (2, 4), (3, 0), (3, 1), (3, 2), (3, 3)])
def test_master_nodes_mass_diff_closed_form(N, p):
    P = p + 1
    nodes, mass, diff = po.sx_master(N, p)
    idx = simplex_order(N, P)
    assert len(idx) == math.comb(p + N, N)
    want = np.full((1, N), 1 / 3) if P == 1 else np.array(idx, dtype=float) / (P - 1)   # master.h:80-95
    assert np.array_equal(nodes, want)
    # Lagrange basis through the monomials of degree <= p (an independent construction)
    expo = simplex_order(N, P)
    V = np.array([[np.prod(x ** np.array(e)) for e in expo] for x in nodes])
    C = np.linalg.inv(V)        # phi_i = sum_k C[k, i] m_k
    Mm = np.array([[monomial_moment(np.add(a, b)) for b in expo] for a in expo])
    np.testing.assert_allclose(mass, C.T @ Mm @ C, rtol=0, atol=1e-13 * max(1, np.abs(mass).max()))
    for d in range(N):
        dV = np.zeros_like(V)
        for k, e in enumerate(expo):
            if e[d] > 0:
                e2 = list(e); e2[d] -= 1
                dV[:, k] = e[d] * np.prod(nodes ** np.array(e2), axis=1)
        np.testing.assert_allclose(diff[d], dV @ C, rtol=0, atol=1e-11)


# ------------------------------------------------------------------ mesh
@pytest.mark.parametrize("name", ["circle", "circle2", "square_8"])
def test_mesh_faces(name):
    pts, tris = mesh(name)
    pr = po.SxProblem(pts, tris, 1, tau0=100, tauD=1000)
    lr, nrm, verts = pr.faces()
    ne, nf, nbf = pr.ne, pr.nf, pr.nbf
    assert 2 * nf - nbf == 3 * ne
    # the set of faces = the set of sorted vertex pairs (mesh.h:210-226)
    edges = {tuple(sorted((t[(j) % 3], t[(j + 1) % 3]))) for t in tris for j in range(3)}
    assert {tuple(v) for v in verts} == edges and len(edges) == nf
    # first-appearance order
    seen, order = set(), []
    for t in tris:
        for j in range(3):
            e = tuple(sorted((t[j % 3], t[(j + 1) % 3])))
            if e not in seen:
                seen.add(e); order.append(e)
    assert [tuple(v) for v in verts] == order
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-14)
    for f in range(nf):
        l, r = lr[f]
        assert r < 0 or l < r          # left = smaller id (mesh.h:234-237)
        opp = [v for v in tris[l] if v not in verts[f]][0]
        assert np.dot(nrm[f], pts[opp] - pts[verts[f][0]]) < 0      # outward from left
        assert abs(np.dot(nrm[f], pts[verts[f][1]] - pts[verts[f][0]])) < 1e-12
    assert (lr[:, 1] < 0).sum() == nbf and set(lr[lr[:, 1] < 0, 1]) == {-1}


def test_boundary_tags_from_predicates():
    pts, tris = mesh("square_8")
    fns = [lambda x: x[0] < 1e-12, lambda x: x[0] > 1 - 1e-12, lambda x: True]
    pr = po.SxProblem(pts, tris, 1, bndfuns=fns, bctype=[0, 1, 0])
    lr, nrm, verts = pr.faces()
    b = lr[:, 1] < 0
    tags = -lr[b, 1]
    assert set(tags) == {1, 2, 3}
    for f in np.nonzero(b)[0]:
        x = pts[verts[f]]
        want = 1 if np.all(x[:, 0] < 1e-12) else (2 if np.all(x[:, 0] > 1 - 1e-12) else 3)
        assert -lr[f, 1] == want
    with pytest.raises(ValueError):
        po.SxProblem(pts, tris, 1, bndfuns=[lambda x: x[0] < 1e-12])   # a face no predicate covers


# ------------------------------------------------------------------ lift / assembly
def test_lift_integrals():
    pts, tris = mesh("circle")
    for p in (0, 1, 2, 3):
        pr = po.SxProblem(pts, tris, p, bndfuns=[on_circle], tau0=100, tauD=1000)
        lr, nrm, verts = pr.faces()
        for f in range(pr.nf):
            L, LL, RR, RL = pr.lift(f)
            area = np.linalg.norm(pts[verts[f][1]] - pts[verts[f][0]])
            # sum_ij int phi_i phi_j = |face| (partition of unity)
            assert LL.sum() == pytest.approx(area, rel=1e-12)
            if lr[f, 1] >= 0:
                assert RR.sum() == pytest.approx(area, rel=1e-12)
                assert RL.sum() == pytest.approx(area, rel=1e-12)
            else:
                assert L.sum() == pytest.approx(area, rel=1e-12)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_assembly_symmetric_spd_and_neumann_kernel(p):
    pts, tris = mesh("circle")
    pr = po.SxProblem(pts, tris, p, bndfuns=[on_circle], tau0=100, tauD=1000)
    A = pr.dense("A")
    assert np.abs(A - A.T).max() < 1e-10 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0
    prn = po.SxProblem(pts, tris, p, bndfuns=[on_circle], bctype=[1], tau0=100, tauD=1000)
    An = prn.dense("A")
    assert np.abs(An @ np.ones(prn.ndof)).max() < 1e-10 * np.abs(An).max()
    M = pr.dense("M")
    assert M.sum() == pytest.approx(sum(0.5 * abs(np.linalg.det(
        np.array([pts[t[1]] - pts[t[0]], pts[t[2]] - pts[t[0]]]))) for t in tris), rel=1e-12)


@pytest.mark.parametrize("name,p", [("circle", 1), ("circle", 2), ("circle", 3), ("square_8", 2)])
def test_patch_test(name, p):
    """LDG is consistent: f = 0, g = linear reproduces the nodal interpolant."""
    pts, tris = mesh(name)
    pr = po.SxProblem(pts, tris, p, tau0=100, tauD=1000)
    g = lambda X: 0.3 + X[:, 0] - 2 * X[:, 1]
    xq, _ = pr.bc_qpoints()
    u = np.linalg.solve(pr.dense("A"), pr.rhs(np.zeros(pr.ndof), g(xq)))
    assert np.abs(u - g(pr.nodes())).max() < 1e-11


@pytest.mark.parametrize("p", [1, 2])
def test_convergence_order(p):
    """-Lap u = f with u = sin(pi x) sin(pi y) on the reference's square meshes."""
    errs = []
    for m in ["square_4", "square_8", "square_16"]:
        pts, tris = mesh(m)
        pr = po.SxProblem(pts, tris, p, tau0=1, tauD=1000)
        x = pr.nodes()
        u = np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1])
        f = 2 * np.pi ** 2 * u
        xq, _ = pr.bc_qpoints()
        uh = np.linalg.solve(pr.dense("A"), pr.rhs(f, np.zeros(len(xq))))
        e = uh - u
        errs.append(math.sqrt(e @ pr.dense("M") @ e))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert rates[-1] > p + 0.7, (errs, rates)


def test_rhs_adjoint_is_transpose():
    pts, tris = mesh("circle")
    fns = [lambda x: x[1] >= 0, lambda x: True]
    pr = po.SxProblem(pts, tris, 2, bndfuns=fns, bctype=[0, 1], tau0=100, tauD=1000)
    rng = np.random.default_rng(0)
    g = rng.standard_normal(pr.nqb)
    lam = rng.standard_normal(pr.ndof)
    bg = pr.rhs(np.zeros(pr.ndof), g)
    assert lam @ bg == pytest.approx(g @ pr.rhs_g_adjoint(lam), rel=1e-12)
    bx, t = pr.bc_qpoints()
    assert set(t) == {1, 2}
    np.testing.assert_allclose(np.hypot(bx[:, 0], bx[:, 1])[t == 1] <= 1.0 + 1e-12, True)


# ------------------------------------------------------------------ agglomeration
def test_agglomeration_graph():
    pts, tris = mesh("circle2")
    pr = po.SxProblem(pts, tris, 1)
    xadj, adj, w, vvol, vsurf = pr.agglom_graph()
    lr, nrm, verts = pr.faces()
    area = np.linalg.norm(pts[verts[:, 1]] - pts[verts[:, 0]], axis=1)
    assert len(adj) == 2 * (pr.nf - pr.nbf)
    E = {}
    for i in range(pr.ne):
        for k in range(xadj[i], xadj[i + 1]):
            assert adj[k] != i
            E[(i, adj[k])] = w[k]
    for (i, j), v in E.items():
        assert E[(j, i)] == v
    assert vsurf.sum() == pytest.approx(area[lr[:, 1] < 0].sum(), rel=1e-13)
    assert vvol.sum() == pytest.approx(pr.dense("M").sum(), rel=1e-12)
    # the graph of a partition sums the fine graph
    part = partitions("circle2")[0]
    nc = int(part.max()) + 1
    cx, ca, cw, cv, cs = pr.agglom_graph(part, nc)
    Ec = {}
    for (i, j), v in E.items():
        a, b = part[i], part[j]
        if a != b:
            Ec[(a, b)] = Ec.get((a, b), 0.0) + v
    got = {(i, ca[k]): cw[k] for i in range(nc) for k in range(cx[i], cx[i + 1])}
    assert set(got) == set(Ec)
    for key in got:
        assert got[key] == pytest.approx(Ec[key], rel=1e-13)
    np.testing.assert_allclose(cv, np.bincount(part, weights=vvol), rtol=1e-13)
    np.testing.assert_allclose(cs, np.bincount(part, weights=vsurf, minlength=nc), rtol=1e-13, atol=1e-15)


@pytest.mark.skipif(not po.mgridgen_available(), reason="reference MGridGen not built (make -C oracle ref)")
def test_mgridgen_partitions_fixture():
    for m in ["circle", "square_16"]:
        pts, tris = mesh(m)
        got = po.agglomerate(po.SxProblem(pts, tris, 1))
        want = partitions(m)
        assert len(got) == len(want)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
        for q in want:
            sizes = np.bincount(q)
            assert sizes.min() >= 1 and sizes.max() <= 4      # minsize 1, maxsize 2^N


@pytest.mark.parametrize("name,p", [("circle2", 1), ("circle2", 2), ("square_16", 2)])
def test_agglomerated_spaces(name, p):
    """Each level's space holds the Q_p box functions of its agglomerates
    (interpolationhierarchy.h:99-199), here checked through global Q_p
    polynomials: the M-orthogonal projection onto the level reproduces them.
    The coarse bases are M-orthonormal: Q_l^T M Q_l is diagonal with ones for
    the kept directions and zeros for the dropped ones."""
    pts, tris = mesh(name)
    parts = partitions(name)
    pr = po.SxProblem(pts, tris, p, tau0=100, tauD=1000)
    mg = po.SxMultigrid(pr, parts, relaxation="jacobi", omega=0.6, restriction="rat")
    assert mg.nlevels == len(parts) + 1
    assert mg.level_dims(0) == (pr.ne, pr.b) and mg.level_dims(1)[1] == (p + 1) ** 2
    X = pr.nodes()
    M = pr.dense("M")
    qs = [(1 + X[:, 0]) ** p * (0.3 - X[:, 1]) ** p + 0.5 * X[:, 0] ** p * X[:, 1],
          X[:, 0] ** p * X[:, 1] ** p - 0.2 * X[:, 1]]   # in Q_p
    Q = np.eye(pr.ndof)
    for l in range(mg.nlevels - 1):
        Q = Q @ mg.dense(l, "interp")
        G = Q.T @ M @ Q
        d = np.diag(G)
        assert np.abs(G - np.diag(d)).max() < 1e-9
        assert np.all((np.abs(d - 1) < 1e-9) | (np.abs(d) < 1e-9))
        for q in qs:
            assert np.abs(Q @ (Q.T @ (M @ q)) - q).max() < 1e-7 * np.abs(q).max(), l


def test_agglomerated_galerkin():
    pts, tris = mesh("circle2")
    pr = po.SxProblem(pts, tris, 2, bndfuns=[on_circle], tau0=100, tauD=1000)
    mg = po.SxMultigrid(pr, partitions("circle2"), relaxation="jacobi", omega=0.6, restriction="rat")
    A = pr.dense("A")
    for l in range(mg.nlevels - 1):
        T = mg.dense(l, "interp")
        Ac = mg.dense(l + 1, "A")
        D = Ac - T.T @ A @ T
        # dropped directions: zero columns of T, unit diagonal of A_c
        dropped = np.abs(T).max(axis=0) == 0
        np.testing.assert_allclose(D, np.diag(dropped.astype(float)), atol=1e-11 * np.abs(Ac).max())
        assert np.linalg.eigvalsh(Ac).min() > 0
        A = Ac


def test_agglomerated_ratrat_rejected():
    """T^T M T can be singular on an agglomerated hierarchy (here: pairs of
    triangles whose four vertices do not determine a bilinear function), so
    RATRAT, which inverts it, is rejected: RAT only."""
    pts, tris = mesh("circle2")
    pr = po.SxProblem(pts, tris, 1, bndfuns=[on_circle], tau0=100, tauD=1000)
    parts = partitions("circle2")
    T0 = po.SxMultigrid(pr, parts, restriction="rat").dense(0, "interp")
    Mc = T0.T @ pr.dense("M") @ T0
    ranks = [np.linalg.matrix_rank(Mc[4 * a:4 * a + 4, 4 * a:4 * a + 4], tol=1e-12 * np.abs(Mc).max())
             for a in range(int(parts[0].max()) + 1)]
    assert min(ranks) < 4
    with pytest.raises(ValueError):
        po.SxMultigrid(pr, parts, restriction="ratrat")


@pytest.mark.parametrize("relaxation", ["jacobi", "mcgs", "gs"])
def test_agglomerated_pcg_matches_direct(relaxation):
    pts, tris = mesh("circle2")
    pr = po.SxProblem(pts, tris, 2, bndfuns=[on_circle], tau0=100, tauD=1000)
    mg = po.SxMultigrid(pr, partitions("circle2"), relaxation=relaxation,
                        omega=0.6 if relaxation == "jacobi" else 0.8, npre=2, npost=2, restriction="rat")
    rng = np.random.default_rng(1)
    b = pr.rhs(rng.standard_normal(pr.ndof), np.zeros(pr.nqb))
    u, it, rr, _ = po.pcg(pr, mg, b, tol=1e-10, maxit=200, criterion="true")
    ud = np.linalg.solve(pr.dense("A"), b)
    assert rr <= 1e-10 and it < 60
    assert np.abs(u - ud).max() < 1e-8 * np.abs(ud).max()


# ------------------------------------------------------------------ face rules and 3-D
def test_face_rules_match_reference_tables():
    """Quadrature<N-1,10> (master.h:359-442): the 10-point Gauss rule on [0,1]
    and the 19-point triangle rule equal the reference's literal tables
    (fixture tests/golden/simplex_quadrature.json) and integrate monomials up
    to their degree exactly (weights normalised to the unit simplex)."""
    with open(os.path.join(HERE, "golden", "simplex_quadrature.json")) as f:
        rules = json.load(f)["rules"]
    x1, w1 = po.sx_face_rule(2)
    np.testing.assert_allclose(x1.ravel(), rules["1"]["10"]["nodes"], rtol=0, atol=2e-16)
    np.testing.assert_allclose(w1, rules["1"]["10"]["weights"], rtol=0, atol=2e-16)
    x2, w2 = po.sx_face_rule(3)
    assert np.array_equal(x2.ravel(), np.array(rules["2"]["10"]["nodes"]))
    assert np.array_equal(w2, np.array(rules["2"]["10"]["weights"]))
    for i in range(20):
        assert (w1 * x1[:, 0] ** i).sum() == pytest.approx(1 / (i + 1), abs=1e-14)
    for i in range(10):
        for j in range(10 - i):
            exact = 2 * math.factorial(i) * math.factorial(j) / math.factorial(i + j + 2)
            assert abs((w2 * x2[:, 0] ** i * x2[:, 1] ** j).sum() - exact) < 1e-13, (i, j)


def test_ball_mesh_3d():
    pts, tets = mesh("ball")
    pr = po.SxProblem(pts, tets, 1, tau0=100, tauD=1000)
    assert (pr.ne, pr.N) == (6327, 3)
    assert 2 * pr.nf - pr.nbf == 4 * pr.ne
    lr, nrm, verts = pr.faces()
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-13)
    for f in range(0, pr.nf, 7):
        l, r = lr[f]
        assert r < 0 or l < r
        opp = [v for v in tets[l] if v not in verts[f]][0]
        assert np.dot(nrm[f], pts[opp] - pts[verts[f][0]]) < 0
    bverts = np.unique(verts[lr[:, 1] < 0])
    np.testing.assert_allclose(np.linalg.norm(pts[bverts], axis=1), 1, atol=1e-5)
    # lift sums: interior faces integrate 1 to the face area; boundary faces
    # to half of it, the reference's weight w * area * |reference triangle|
    # (mesh.h:343-344; exact for segments, half for triangles)
    for f in list(range(0, pr.nf, 97)):
        L, LL, RR, RL = pr.lift(f)
        a, b, c = pts[verts[f]]
        area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
        if lr[f, 1] >= 0:
            assert LL.sum() == pytest.approx(area, rel=1e-12) and RL.sum() == pytest.approx(area, rel=1e-12)
        else:
            assert LL.sum() == pytest.approx(area / 2, rel=1e-12)


def test_ball_assembly_3d():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    pts, tets = mesh("ball")
    pr = po.SxProblem(pts, tets, 1, tau0=100, tauD=1000)
    rp, cols, blk = pr.export("A")
    A = sp.bsr_matrix((blk, cols, rp), shape=(pr.ndof, pr.ndof)).tocsr()
    assert abs(A - A.T).max() < 1e-12 * abs(A).max()
    prn = po.SxProblem(pts, tets, 1, bctype=[1], tau0=100, tauD=1000)
    rp, cols, blk = prn.export("A")
    An = sp.bsr_matrix((blk, cols, rp), shape=(prn.ndof, prn.ndof))
    assert np.abs(An @ np.ones(prn.ndof)).max() < 1e-10 * abs(An).max()
    # the halved boundary terms keep the operator SPD and the iteration
    # converging but cost the 3-D patch test its exactness
    g = lambda X: 0.3 + X[:, 0] - 2 * X[:, 1] + 0.5 * X[:, 2]
    xq, _ = pr.bc_qpoints()
    b = pr.rhs(np.zeros(pr.ndof), g(xq))
    u = spl.spsolve(A.tocsc(), b)
    err = np.abs(u - g(pr.nodes())).max()
    assert 1e-4 < err < 1e-2
    mg = po.SxMultigrid(pr, partitions("ball"), relaxation="mcgs", npre=2, npost=2, restriction="rat")
    x, it, rr, _ = po.pcg(pr, mg, b, tol=1e-10, maxit=200, criterion="true")
    assert rr <= 1e-10 and it < 80
    assert np.abs(x - u).max() < 1e-8 * np.abs(u).max()
