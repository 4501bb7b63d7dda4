"""ctypes binding of the C oracle (oracle/cart.c, oracle/simplex.c).  Test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")

HFN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


class MGParams(C.Structure):
    _fields_ = [("relaxation", C.c_int), ("omega", C.c_double), ("coarse", C.c_int),
                ("npre", C.c_int), ("npost", C.c_int), ("restriction", C.c_int)]


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        build()
    L = C.CDLL(_LIB)
    P, I, D, V = C.POINTER, C.c_int, C.c_double, C.c_void_p
    pd, pi = P(D), P(I)
    sig = {
        "orc_gl_nodes": (I, [I, pd]), "orc_lagrange_denom": (I, [I, pd]),
        "orc_mass1d": (I, [I, pd]), "orc_invmass1d": (I, [I, pd]), "orc_diff1d": (I, [I, pd]),
        "orc_gauss": (I, [I, pd, pd]), "orc_lagrange_eval": (I, [I, D, pd]),
        "orc_quadtree_new": (V, [I, pd, pd, HFN, V, I, I, I]),
        "orc_quadtree_free": (None, [V]), "orc_quadtree_info": (I, [V, pi, pi]),
        "orc_quadtree_layer": (I, [V, I, pi]), "orc_quadtree_node": (I, [V, I, pi, pd]),
        "orc_problem_new": (V, [V, I, pi, D, D]), "orc_problem_free": (None, [V]),
        "orc_problem_sizes": (I, [V, pi]), "orc_problem_nodes": (I, [V, pd]),
        "orc_problem_bc_qpoints": (I, [V, pd, pi]), "orc_problem_faces": (I, [V, pi, pd]),
        "orc_problem_rhs": (I, [V, pd, pd, pd]), "orc_problem_rhs_g_adjoint": (I, [V, pd, pd]),
        "orc_problem_export": (C.c_long, [V, I, pi, pi, pd]),
        "orc_mg_new": (V, [V, P(MGParams)]), "orc_mg_free": (None, [V]),
        "orc_mg_new_hp": (V, [V, P(MGParams), I, pi, I]),
        "orc_mg_nlevels": (I, [V]),
        "orc_mg_export": (C.c_long, [V, I, I, pi, pi, pd, pi, pi]),
        "orc_mg_colors": (I, [V, I, pi]), "orc_mg_vcycle": (I, [V, pd, pd]), "orc_mg_vcycles": (I, [V, pd, pd, I, pd]),
        "orc_mg_relax": (I, [V, I, I, I, pd, pd]), "orc_mg_spmv": (I, [V, I, pd, pd]),
        "orc_pcg": (I, [V, V, pd, pd, D, I, I, I, pd, pd]),
        "orc_bsr_spmv": (I, [I, I, pi, pi, pd, pd, pd]),
        "orc_mg_level_dims": (I, [V, I, pi, pi]),
        "orc_sx_master": (I, [I, I, pd, pd, pd]),
        "orc_sx_face_rule": (I, [I, pd, pd]),
        "orc_sx_problem_new": (V, [I, I, pd, I, pi, I, pi, I, pi, D, D]),
        "orc_sx_problem_free": (None, [V]), "orc_sx_sizes": (I, [V, pi]),
        "orc_sx_nodes": (I, [V, pd]), "orc_sx_faces": (I, [V, pi, pd, pi]),
        "orc_sx_bc_qpoints": (I, [V, pd, pi]), "orc_sx_lift": (I, [V, I, pd, pd, pd, pd]),
        "orc_sx_rhs": (I, [V, pd, pd, pd]), "orc_sx_rhs_g_adjoint": (I, [V, pd, pd]),
        "orc_sx_export": (C.c_long, [V, I, pi, pi, pd]),
        "orc_sx_agglom_graph": (I, [V, pi, I, pi, pi, pd, pd, pd]),
        "orc_sx_mg_new": (V, [V, I, P(pi), pi, P(MGParams)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


def basis_tables(P, Q=10):
    L = lib()
    out = {}
    for name, fn, n in [("nodes", L.orc_gl_nodes, P), ("denom", L.orc_lagrange_denom, P),
                        ("mass", L.orc_mass1d, P * P), ("invmass", L.orc_invmass1d, P * P),
                        ("diff", L.orc_diff1d, P * P)]:
        a = np.zeros(n)
        assert fn(P, _pd(a)) == 0
        out[name] = a if n == P else a.reshape(P, P)
    gx, gw = np.zeros(Q), np.zeros(Q)
    assert L.orc_gauss(Q, _pd(gx), _pd(gw)) == 0
    out["qnodes"], out["qweights"] = gx, gw
    return out


def bsr_spmv(rowptr, cols, blocks, x, b):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    blocks = np.ascontiguousarray(blocks, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    m = len(rowptr) - 1
    y = np.zeros(m * b)
    lib().orc_bsr_spmv(m, b, _pi(rowptr), _pi(cols), _pd(blocks), _pd(x), _pd(y))
    return y


class Quadtree:
    """Quadtree<N>(h, strategy, periodic) (src/cartesian/mesh/quadtree.h:211-228).

    ``h`` maps a point (numpy array of length N) to N widths (or a scalar)."""

    def __init__(self, N, h, lo=None, hi=None, rapid=False, periodic=False, balance=False):
        self.N = N

        def cb(x, out, user):
            v = h(np.array([x[i] for i in range(N)]))
            v = np.broadcast_to(np.asarray(v, dtype=np.float64), (N,))
            for i in range(N):
                out[i] = float(v[i])

        self._cb = HFN(cb)
        lo = np.zeros(N) if lo is None else np.asarray(lo, dtype=np.float64)
        hi = np.ones(N) if hi is None else np.asarray(hi, dtype=np.float64)
        self.h = lib().orc_quadtree_new(N, _pd(lo), _pd(hi), self._cb, None, int(rapid), int(periodic), int(balance))
        if not self.h:
            raise ValueError("oracle quadtree creation failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_quadtree_free(self.h)
            self.h = None

    def info(self):
        a, b = C.c_int(), C.c_int()
        lib().orc_quadtree_info(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def layer(self, l):
        n = lib().orc_quadtree_layer(self.h, l, None)
        ids = np.zeros(n, dtype=np.int32)
        lib().orc_quadtree_layer(self.h, l, _pi(ids))
        return ids

    def node(self, i):
        ilh = np.zeros(3, dtype=np.int32)
        lohi = np.zeros(2 * self.N)
        lib().orc_quadtree_node(self.h, i, _pi(ilh), _pd(lohi))
        return dict(level=int(ilh[0]), height=int(ilh[1]), parent=int(ilh[2]),
                    lo=lohi[:self.N].copy(), hi=lohi[self.N:].copy())


class Problem:
    """Mesh<N,P>(qt) + LDGPoisson<N,P>(mesh, bcs, tau0, tauD)
    (src/cartesian/mesh/mesh.h:186-265, src/cartesian/dg/ldgpoisson.h:60-77)."""

    def __init__(self, qt, p, bctype=None, tau0=0.0, tauD=1000.0):
        self.qt = qt
        self.N = qt.N
        self.p = p
        bct = np.zeros(2 * qt.N, dtype=np.int32) if bctype is None else np.asarray(bctype, dtype=np.int32)
        self.h = lib().orc_problem_new(qt.h, p, _pi(bct), tau0, tauD)
        if not self.h:
            raise ValueError("oracle problem creation failed")
        s = np.zeros(6, dtype=np.int32)
        lib().orc_problem_sizes(self.h, _pi(s))
        self.ne, self.nf, self.nbf, self.b, self.nqb, self.ndof = [int(v) for v in s]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_problem_free(self.h)
            self.h = None

    def nodes(self):
        x = np.zeros((self.ndof, self.N))
        lib().orc_problem_nodes(self.h, _pd(x))
        return x

    def bc_qpoints(self):
        x = np.zeros((self.nqb, self.N))
        t = np.zeros(self.nqb, dtype=np.int32)
        lib().orc_problem_bc_qpoints(self.h, _pd(x), _pi(t))
        return x, t

    def faces(self):
        a = np.zeros((self.nf, 4), dtype=np.int32)
        n = np.zeros(self.nf)
        lib().orc_problem_faces(self.h, _pi(a), _pd(n))
        return a, n

    def rhs(self, f=None, g=None):
        f = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        g = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        out = np.zeros(self.ndof)
        lib().orc_problem_rhs(self.h, _pd(f), _pd(g), _pd(out))
        return out

    def rhs_g_adjoint(self, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.zeros(self.nqb)
        lib().orc_problem_rhs_g_adjoint(self.h, _pd(lam), _pd(out))
        return out

    def export(self, which):
        """which in 'A','M','T','G0','G1','G2' -> (rowptr, cols, blocks[nnzb,b,b])"""
        w = {"A": 0, "M": 1, "T": 2, "G0": 3, "G1": 4, "G2": 5}[which]
        nz = lib().orc_problem_export(self.h, w, None, None, None)
        rp = np.zeros(self.ne + 1, dtype=np.int32)
        cols = np.zeros(nz, dtype=np.int32)
        blk = np.zeros((nz, self.b, self.b))
        lib().orc_problem_export(self.h, w, _pi(rp), _pi(cols), _pd(blk))
        return rp, cols, blk

    def dense(self, which):
        rp, cols, blk = self.export(which)
        return bsr_to_dense(rp, cols, blk, self.ne, self.ne)


def bsr_to_dense(rp, cols, blk, nbr, nbc):
    br, bc = blk.shape[1], blk.shape[2]
    D = np.zeros((nbr * br, nbc * bc))
    for i in range(nbr):
        for k in range(rp[i], rp[i + 1]):
            j = cols[k]
            D[i * br:(i + 1) * br, j * bc:(j + 1) * bc] += blk[k]
    return D


class Multigrid:
    """InterpolationHierarchy<N,P>(qt) + Multigrid<N,P>(ops, hierarchy, params)
    (src/cartesian/multigrid/interpolationhierarchy.h:30-82, multigrid.h:291-301).

    relaxation: 'jacobi' | 'gs' (lexicographic, the reference) | 'mcgs' (multicolour)."""

    RELAX = {"jacobi": 0, "gs": 1, "mcgs": 2}

    def __init__(self, prob, relaxation="gs", omega=0.8, coarse="cholesky", npre=3, npost=3,
                 restriction="ratrat", pdegrees=None, hcoarsen=True):
        """pdegrees: polynomial degrees of p-levels below the problem's
        (Multigrid<N,P1,P2,...>, multigrid.h:435-639); then the h-hierarchy at
        the last degree, or none if hcoarsen is False (StopCoarsening)."""
        self.prob = prob
        self.params = MGParams(self.RELAX[relaxation], omega, {"cholesky": 0, "relaxation": 1}[coarse],
                               npre, npost, {"rat": 0, "ratrat": 1}[restriction])
        self.pdegrees = None if pdegrees is None else np.ascontiguousarray(pdegrees, dtype=np.int32)
        self.hcoarsen = hcoarsen
        self.h = self._create(prob)
        if not self.h:
            raise ValueError("oracle multigrid creation failed")
        self.nlevels = lib().orc_mg_nlevels(self.h)

    def _create(self, prob):
        if self.pdegrees is not None:
            return lib().orc_mg_new_hp(prob.h, C.byref(self.params), len(self.pdegrees), _pi(self.pdegrees),
                                       int(self.hcoarsen))
        return lib().orc_mg_new(prob.h, C.byref(self.params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mg_free(self.h)
            self.h = None

    def export(self, level, which):
        w = {"A": 0, "M": 1, "T": 2, "G0": 3, "G1": 4, "G2": 5, "interp": 6}[which]
        nbr, nbc = C.c_int(), C.c_int()
        nz = lib().orc_mg_export(self.h, level, w, None, None, None, C.byref(nbr), C.byref(nbc))
        if nz < 0:
            raise ValueError("not available")
        rp = np.zeros(nbr.value + 1, dtype=np.int32)
        cols = np.zeros(nz, dtype=np.int32)
        br = self.level_dims(level)[1]
        bc = self.level_dims(level + 1)[1] if which == "interp" else br
        blk = np.zeros((nz, br, bc))
        lib().orc_mg_export(self.h, level, w, _pi(rp), _pi(cols), _pd(blk), None, None)
        return rp, cols, blk, nbr.value, nbc.value

    def level_dims(self, level):
        """(block rows, block size) of a level"""
        nb, b = C.c_int(), C.c_int()
        if lib().orc_mg_level_dims(self.h, level, C.byref(nb), C.byref(b)) != 0:
            raise ValueError("no such level")
        return nb.value, b.value

    def dense(self, level, which):
        rp, cols, blk, nbr, nbc = self.export(level, which)
        return bsr_to_dense(rp, cols, blk, nbr, nbc)

    def colors(self, level):
        nb = self.export(level, "A")[3]
        c = np.zeros(nb, dtype=np.int32)
        nc = lib().orc_mg_colors(self.h, level, _pi(c))
        return c, nc

    def vcycle(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().orc_mg_vcycle(self.h, _pd(b), _pd(x))
        return x

    def vcycles(self, b, x0, ncycle):
        """the driver's stationary V-cycle solver (tests/cartesian/dg_tests.cpp:309-332):
        ncycle V-cycles from x0; returns (x, residual norms before and after each cycle)"""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x0, dtype=np.float64)
        rn = np.zeros(ncycle + 1)
        lib().orc_mg_vcycles(self.h, _pd(b), _pd(x), int(ncycle), _pd(rn))
        return x, rn

    def relax(self, level, b, x, nsweeps=1, reverse=False):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        lib().orc_mg_relax(self.h, level, nsweeps, int(reverse), _pd(b), _pd(x))
        return x

    def spmv(self, level, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().orc_mg_spmv(self.h, level, _pd(x), _pd(y))
        return y


def pcg(prob, mg, b, x0=None, tol=1e-8, maxit=200, criterion="reference", flexible=False, history=False):
    """pcg(A, b, x, precon, tol, maxit) (src/cartesian/multigrid/krylov.h:22-68).

    Returns (x, iterations, true_relres, hist)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    stats = np.zeros(3)
    hist = np.zeros(max(maxit, 1)) if history else None
    crit = {"reference": 0, "true": 1}[criterion]
    lib().orc_pcg(prob.h, mg.h if mg is not None else None, _pd(b), _pd(x), tol, maxit, crit,
                  int(flexible), _pd(stats), _pd(hist))
    it = int(stats[0])
    return x, it, float(stats[1]), (hist[:it] if history else None)


# ---------------------------------------------------------------- simplices
def read_mesh(path):
    """Wireframe<N>(file) (src/unstructured/mesh/wireframe.h:22-68): header
    "np nt", a blank line, np coordinate lines, a blank line, nt lines of
    1-based vertex ids.  Returns (points [np,N], simplices [nt,N+1] 0-based)."""
    with open(path) as f:
        lines = f.read().split("\n")
    np_, nt = (int(v) for v in lines[0].split())
    k = 2
    pts = []
    while k < len(lines) and lines[k].strip():
        pts.append([float(v) for v in lines[k].split()])
        k += 1
    k += 1
    tris = []
    while k < len(lines) and lines[k].strip():
        tris.append([int(v) - 1 for v in lines[k].split()])
        k += 1
    p = np.array(pts, dtype=np.float64)
    t = np.array(tris, dtype=np.int32)
    assert p.shape[0] == np_ and t.shape[0] == nt and t.shape[1] == p.shape[1] + 1
    return p, t


def sx_master(N, p):
    """Master<N,p+1>: (nodes [npl,N], mass [npl,npl], diff [N,npl,npl])"""
    L = lib()
    n = L.orc_sx_master(N, p + 1, None, None, None)
    nodes, mass, diff = np.zeros((n, N)), np.zeros((n, n)), np.zeros((N, n, n))
    L.orc_sx_master(N, p + 1, _pd(nodes), _pd(mass), _pd(diff))
    return nodes, mass, diff


def sx_face_rule(N):
    """Quadrature<N-1,10> of the simplex path: (points [nq, N-1], weights [nq])"""
    n = lib().orc_sx_face_rule(N, None, None)
    x, w = np.zeros((n, N - 1)), np.zeros(n)
    lib().orc_sx_face_rule(N, _pd(x), _pd(w))
    return x, w


class SxProblem:
    """Mesh<N,P>(wireframe, bndfuns) + LDGPoisson<N,P>(mesh, bcs, tau0, tauD)
    (src/unstructured/mesh/mesh.h:185-298, src/unstructured/dg/ldgpoisson.tpp:14-99).

    bndfuns: list of predicates on a vertex (tags 1..len); bctype: 0 Dirichlet /
    1 Neumann per tag."""

    def __init__(self, pts, tris, p, bndfuns=None, bctype=None, tau0=0.0, tauD=1000.0):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        tris = np.ascontiguousarray(tris, dtype=np.int32)
        self.N = pts.shape[1]
        self.p = p
        nb = len(bndfuns) if bndfuns else 0
        vmask = np.zeros(len(pts), dtype=np.int32)
        for k, fn in enumerate(bndfuns or []):
            for v in range(len(pts)):
                if fn(pts[v]):
                    vmask[v] |= 1 << k
        bct = np.zeros(max(nb, 1), dtype=np.int32) if bctype is None else np.asarray(bctype, dtype=np.int32)
        self.vmask = vmask
        self.h = lib().orc_sx_problem_new(self.N, len(pts), _pd(pts), len(tris), _pi(tris), nb, _pi(vmask),
                                          p, _pi(bct), tau0, tauD)
        if not self.h:
            raise ValueError("oracle simplex problem creation failed")
        s = np.zeros(7, dtype=np.int32)
        lib().orc_sx_sizes(self.h, _pi(s))
        self.ne, self.nf, self.nbf, self.b, self.nqb, self.ndof, self.nq = [int(v) for v in s]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sx_problem_free(self.h)
            self.h = None

    def nodes(self):
        x = np.zeros((self.ndof, self.N))
        lib().orc_sx_nodes(self.h, _pd(x))
        return x

    def faces(self):
        lr = np.zeros((self.nf, 2), dtype=np.int32)
        n = np.zeros((self.nf, self.N))
        v = np.zeros((self.nf, self.N), dtype=np.int32)
        lib().orc_sx_faces(self.h, _pi(lr), _pd(n), _pi(v))
        return lr, n, v

    def bc_qpoints(self):
        x = np.zeros((self.nqb, self.N))
        t = np.zeros(self.nqb, dtype=np.int32)
        lib().orc_sx_bc_qpoints(self.h, _pd(x), _pi(t))
        return x, t

    def lift(self, face):
        b, nq = self.b, self.nq
        L, LL, RR, RL = np.zeros((b, nq)), np.zeros((b, b)), np.zeros((b, b)), np.zeros((b, b))
        lib().orc_sx_lift(self.h, face, _pd(L), _pd(LL), _pd(RR), _pd(RL))
        return L, LL, RR, RL

    def rhs(self, f=None, g=None):
        f = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        g = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        out = np.zeros(self.ndof)
        lib().orc_sx_rhs(self.h, _pd(f), _pd(g), _pd(out))
        return out

    def rhs_g_adjoint(self, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.zeros(self.nqb)
        lib().orc_sx_rhs_g_adjoint(self.h, _pd(lam), _pd(out))
        return out

    def export(self, which):
        w = {"A": 0, "M": 1, "T": 2, "G0": 3, "G1": 4, "G2": 5}[which]
        nz = lib().orc_sx_export(self.h, w, None, None, None)
        if nz < 0:
            raise ValueError("not available")
        rp = np.zeros(self.ne + 1, dtype=np.int32)
        cols = np.zeros(nz, dtype=np.int32)
        blk = np.zeros((nz, self.b, self.b))
        lib().orc_sx_export(self.h, w, _pi(rp), _pi(cols), _pd(blk))
        return rp, cols, blk

    def dense(self, which):
        rp, cols, blk = self.export(which)
        return bsr_to_dense(rp, cols, blk, self.ne, self.ne)

    def agglom_graph(self, part=None, nparts=None):
        """The MGridGen input graph (src/unstructured/mesh/agglomeration.h:57-93,
        133-185): (xadj, adjncy, adjwgt, vvol, vsurf)."""
        nv = self.ne if part is None else int(nparts)
        part = None if part is None else np.ascontiguousarray(part, dtype=np.int32)
        xadj = np.zeros(nv + 1, dtype=np.int32)
        adj = np.zeros(2 * self.nf, dtype=np.int32)
        w = np.zeros(2 * self.nf)
        vvol, vsurf = np.zeros(nv), np.zeros(nv)
        ne = lib().orc_sx_agglom_graph(self.h, _pi(part), nv, _pi(xadj), _pi(adj), _pd(w), _pd(vvol), _pd(vsurf))
        return xadj, adj[:ne].copy(), w[:ne].copy(), vvol, vsurf


class SxMultigrid(Multigrid):
    """Agglomerated InterpolationHierarchy<N,P>(agglomeration) + multigrid
    (src/unstructured/multigrid/interpolationhierarchy.h:88-204).  parts[l]
    maps the elements of level l to agglomerates of level l+1."""

    def __init__(self, prob, parts, **kw):
        self.parts = [np.ascontiguousarray(q, dtype=np.int32) for q in parts]
        self.nparts = np.array([int(q.max()) + 1 for q in self.parts], dtype=np.int32)
        super().__init__(prob, **kw)

    def _create(self, prob):
        arr = (C.POINTER(C.c_int) * len(self.parts))(*[_pi(q) for q in self.parts])
        return lib().orc_sx_mg_new(prob.h, len(self.parts), arr, _pi(self.nparts), C.byref(self.params))


# ---------------------------------------------------------------- MGridGen
_REF_MGG = os.path.join(_HERE, "_ref", "libmgridgen.so")
_mgg = None


def mgridgen_available():
    return os.path.exists(_REF_MGG)


def mgridgen(xadj, vvol, vsurf, adjncy, adjwgt, minsize, maxsize, options):
    """MGridGen(nvtxs, xadj, vvol, vsurf, adjncy, adjwgt, minsize, maxsize,
    options, &nmoves, &nparts, part) of the reference's MGridGen 1.0
    (lib/ParMGridGen-1.0/MGridGen/Lib/mgridgen.c:18-49), compiled from its
    sources into oracle/_ref.  Returns (nparts, part)."""
    global _mgg
    if _mgg is None:
        _mgg = C.CDLL(_REF_MGG)
        _mgg.MGridGen.restype = None
    xadj = np.array(xadj, dtype=np.int32)
    adjncy = np.array(adjncy, dtype=np.int32)
    vvol = np.array(vvol, dtype=np.float64)
    vsurf = np.array(vsurf, dtype=np.float64)
    adjwgt = np.array(adjwgt, dtype=np.float64)
    opts = np.array(options, dtype=np.int32)
    n = len(xadj) - 1
    part = np.zeros(n, dtype=np.int32)
    nmoves, nparts = C.c_int(), C.c_int()
    _mgg.MGridGen(C.c_int(n), _pi(xadj), _pd(vvol), _pd(vsurf), _pi(adjncy), _pd(adjwgt),
                  C.c_int(minsize), C.c_int(maxsize), _pi(opts), C.byref(nmoves), C.byref(nparts), _pi(part))
    return nparts.value, part


def agglomerate(prob, max_levels=10):
    """Agglomeration<N,P>(mesh, maxAgglom) (src/unstructured/mesh/agglomeration.h:21-34,
    96-116, 188-216): MGridGen with minsize 1, maxsize 2^N, options {4,6,1,N}
    on the mesh and then on each agglomerated mesh, up to max_levels
    agglomerated meshes or until one agglomerate remains.  Returns the list of
    partitions, level l elements -> level l+1 agglomerates."""
    N = prob.N
    parts = []
    orig = None
    nparts = prob.ne
    for _ in range(max_levels):
        if nparts <= 1:
            break
        xadj, adj, w, vvol, vsurf = prob.agglom_graph(orig, nparts if orig is not None else None)
        nc, part = mgridgen(xadj, vvol, vsurf, adj, w, 1, 1 << N, [4, 6, 1, N])
        parts.append(part)
        orig = part if orig is None else part[orig]
        nparts = nc
    return parts
