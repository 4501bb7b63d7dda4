"""The library's multilevel agglomerator (host code, called through the C ABI
entry mgldg_mgridgen that stands in for MGridGen(), lib/ParMGridGen-1.0/
MGridGen/Lib/mgridgen.c:18-49) against the reference's MGridGen 1.0 compiled
from its sources (oracle/_ref), on the dual graphs of the reference's meshes
(src/unstructured/mesh/agglomeration.h:57-93, 133-185, restated by the
oracle) with the reference's parameters (minsize 1, maxsize 2^N,
agglomeration.h:96-116).

The two use different random orders, so partitions are compared by what
MGridGen optimises and by what they do for the solver: cluster counts,
largest and mean aspect ratios, and the oracle's MG-PCG iteration counts on
the two hierarchies (within one).  Over 16 random orders the library's
iteration counts spread about MGridGen's single draw (ball.mesh p = 1 block
Jacobi: 56-63, mean 58.2, MGridGen 57; circle2 p = 2 Jacobi 39-42, mean
40.1, MGridGen 40; circle2 multicolour GS 29-31 / 30; square_32 GS 28-31 /
30); the library's fixed order is one of them."""
import json
import os

import numpy as np
import pytest

import mgldg

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def fixture_partitions(name):
    with open(os.path.join(HERE, "golden", "partitions.json")) as f:
        return [np.array(q, dtype=np.int32) for q in json.load(f)[name]]


def aratio(N, s, v):
    return s * s / v if N == 2 else s ** 1.5 / v


def cluster_stats(g, part, N):
    xadj, adj, w, vv, vs = g
    nc = int(part.max()) + 1
    vol = np.bincount(part, vv, nc)
    surf = np.bincount(part, vs, nc)
    src = np.repeat(np.arange(len(vv)), np.diff(xadj))
    cut = part[adj] != part[src]
    surf += np.bincount(part[src[cut]], np.asarray(w)[cut], nc)
    ar = aratio(N, surf, vol)
    return nc, ar.max(), ar.mean(), np.bincount(part, minlength=nc)


def connected(g, part):
    xadj, adj = g[0], g[1]
    comp = -np.ones(len(part), dtype=np.int64)
    nc = 0
    for s in range(len(part)):
        if comp[s] >= 0:
            continue
        comp[s] = nc
        stack = [s]
        while stack:
            u = stack.pop()
            for k in range(xadj[u], xadj[u + 1]):
                v = int(adj[k])
                if comp[v] < 0 and part[v] == part[u]:
                    comp[v] = nc
                    stack.append(v)
        nc += 1
    return nc == int(part.max()) + 1


def hierarchy(pr, fn):
    """Agglomeration<N,P>(mesh, 10) (agglomeration.h:21-34): partitions of the
    mesh, then of each agglomerated mesh, until one agglomerate remains"""
    parts, orig, n = [], None, pr.ne
    for _ in range(10):
        if n <= 1:
            break
        g = pr.agglom_graph(orig, n if orig is not None else None)
        nc, part = fn(g)
        if nc >= n:
            break
        parts.append(np.asarray(part, dtype=np.int32))
        orig = parts[-1] if orig is None else parts[-1][orig]
        n = nc
    return parts


def lib_fn(N):
    return lambda g: mgldg.mgridgen(g[0], g[3], g[4], g[1], g[2], 1, 1 << N, N)


@pytest.mark.parametrize("name", ["circle2", "square_32", "ball"])
def test_partitions_valid_and_of_mgridgen_quality(oracle, name):
    pts, tris = oracle.read_mesh(os.path.join(MESHES, name + ".mesh"))
    pr = oracle.SxProblem(pts, tris, 1)
    N = pr.N
    mine = hierarchy(pr, lib_fn(N))
    ref = fixture_partitions(name)
    g = pr.agglom_graph()
    n1, mx1, mean1, sizes1 = cluster_stats(g, mine[0], N)
    n2, mx2, mean2, _ = cluster_stats(g, ref[0], N)
    assert sizes1.min() >= 1 and sizes1.max() <= 1 << N
    assert connected(g, mine[0])
    assert abs(n1 - n2) <= 0.03 * n2, (n1, n2)
    assert mean1 <= 1.02 * mean2 and mx1 <= 1.25 * mx2, (mean1, mean2, mx1, mx2)
    # every level: clusters of 1..2^N connected agglomerates
    orig, n = None, pr.ne
    for q in mine:
        gl = pr.agglom_graph(orig, n if orig is not None else None)
        assert len(q) == n and np.bincount(q).min() >= 1 and np.bincount(q).max() <= 1 << N
        assert connected(gl, q)
        orig = q if orig is None else q[orig]
        n = int(q.max()) + 1
    assert n == 1


@pytest.mark.parametrize("name,p,relax", [("circle2", 2, "jacobi"), ("circle2", 2, "mcgs"), ("square_32", 2, "mcgs"),
                                          ("square_32", 1, "jacobi"), ("ball", 1, "jacobi")])
def test_solver_iterations_match_mgridgen(oracle, name, p, relax):
    """the oracle's MG-PCG (RAT, V(2,2)) on the library's hierarchy takes
    MGridGen's iteration count within one"""
    pts, tris = oracle.read_mesh(os.path.join(MESHES, name + ".mesh"))
    pr = oracle.SxProblem(pts, tris, p, tau0=100.0, tauD=1000.0)
    mine = hierarchy(pr, lib_fn(pr.N))
    b = pr.rhs(np.random.default_rng(5).standard_normal(pr.ndof), np.zeros(pr.nqb))
    om = 0.6 if relax == "jacobi" else 0.8
    its = []
    for parts in (fixture_partitions(name), mine):
        mg = oracle.SxMultigrid(pr, parts, relaxation=relax, omega=om, npre=2, npost=2, restriction="rat")
        _, it, rr, _ = oracle.pcg(pr, mg, b, tol=1e-10, maxit=400, criterion="true")
        assert rr <= 1e-10
        its.append(it)
    assert abs(its[1] - its[0]) <= 1, its


def test_live_mgridgen_on_the_same_graph(oracle):
    """the compiled reference MGridGen and the library on one graph (when the
    reference build is present): same kind of output"""
    if not oracle.mgridgen_available():
        pytest.skip("oracle/_ref/libmgridgen.so not built")
    pts, tris = oracle.read_mesh(os.path.join(MESHES, "square_16.mesh"))
    pr = oracle.SxProblem(pts, tris, 1)
    xadj, adj, w, vv, vs = pr.agglom_graph()
    n_ref, p_ref = oracle.mgridgen(xadj, vv, vs, adj, w, 1, 4, [4, 6, 1, 2])
    n_lib, p_lib = mgldg.mgridgen(xadj, vv, vs, adj, w, 1, 4, 2)
    assert abs(n_lib - n_ref) <= 0.05 * n_ref
    assert p_lib.min() == 0 and p_lib.max() == n_lib - 1


def test_minsize_merges_small_clusters(oracle):
    """minsize > 1 exercises the merge / attract phases (agglomeration.h
    passes 1; MGridGen's interface takes any): no cluster below minsize"""
    pts, tris = oracle.read_mesh(os.path.join(MESHES, "circle2.mesh"))
    pr = oracle.SxProblem(pts, tris, 1)
    g = pr.agglom_graph()
    for mn, mx in ((2, 4), (3, 6), (4, 8)):
        nc, part = mgldg.mgridgen(g[0], g[3], g[4], g[1], g[2], mn, mx, 2)
        sizes = np.bincount(part)
        assert len(sizes) == nc and sizes.min() >= mn and sizes.max() <= mx, (mn, mx, sizes.min(), sizes.max())
        assert connected(g, part)


def test_deterministic_and_argument_errors(oracle):
    pts, tris = oracle.read_mesh(os.path.join(MESHES, "circle2.mesh"))
    pr = oracle.SxProblem(pts, tris, 1)
    g = pr.agglom_graph()
    a = mgldg.mgridgen(g[0], g[3], g[4], g[1], g[2], 1, 4, 2)
    b = mgldg.mgridgen(g[0], g[3], g[4], g[1], g[2], 1, 4, 2)
    assert a[0] == b[0] and np.array_equal(a[1], b[1])
    with pytest.raises(mgldg.MGLDGError):
        mgldg.mgridgen(g[0], g[3], g[4], g[1], g[2], 1, 4, 4)          # dim
    with pytest.raises(mgldg.MGLDGError):
        mgldg.mgridgen(g[0], g[3], g[4], g[1], g[2], 3, 2, 2)          # minsize > maxsize
    bad = np.array(g[1]).copy()
    bad[0] = len(g[3]) + 5
    with pytest.raises(mgldg.MGLDGError):
        mgldg.mgridgen(g[0], g[3], g[4], bad, g[2], 1, 4, 2)           # neighbour out of range
    import ctypes as C
    from mgldg._lib import lib
    opts = np.array([2, 6, 1, 2], dtype=np.int32)                       # CType 2: not built
    n = len(g[0]) - 1
    part = np.zeros(n, dtype=np.int32)
    xa = np.ascontiguousarray(g[0], dtype=np.int32)
    adj = np.ascontiguousarray(g[1], dtype=np.int32)
    w, vv, vs = [np.ascontiguousarray(x, dtype=np.float64) for x in (g[2], g[3], g[4])]
    PI = lambda x: x.ctypes.data_as(C.POINTER(C.c_int))
    PD = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    nparts = C.c_int()
    assert lib().mgldg_mgridgen(n, PI(xa), PD(vv), PD(vs), PI(adj), PD(w), 1, 4, PI(opts), None, C.byref(nparts),
                                PI(part)) == -5
