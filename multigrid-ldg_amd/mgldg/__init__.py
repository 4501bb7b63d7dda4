"""mgldg — MI355X-native LDG-Poisson setup + MG-preconditioned CG.

Python mirror of the reference's C++ surface (danfortunato/multigrid-ldg):

    reference (src/cartesian/...)                  here
    Quadtree<N>(h, strategy, periodic)             Quadtree(ctx, N, h, coarsening=..., periodic=...)
    Mesh<N,P>(qt) + LDGPoisson<N,P>(mesh, bcs,     LDGPoisson(qt, p, bc=..., tau0=..., tauD=...)
        tau0, tauD), computeRHS(f)                     .rhs(f, g)
    InterpolationHierarchy<N,P>(qt) +              Multigrid(prob, relaxation=..., omega=..., npre=...,
        Multigrid<N,P>(ops, hierarchy, params)         npost=..., restriction=..., precision=...)
    pcg(A, b, x, precon, tol, maxit)               pcg(mg, b, x0, tol, maxit, ...)

Everything above the ctypes calls is argument plumbing; the work happens in
libmgldg.so (host C++ + HIP kernels for gfx950).  There is no CPU fallback:
if the library or a GPU is missing, the calls raise.
"""
import ctypes as C
import os

import numpy as np

from ._lib import lib, LIB_PATH, MGParams, SolveParams, SolveStats, SIZEFN, check, MGLDGError  # noqa: F401

STATUS = {0: "converged", 1: "maxit", 2: "breakdown", 3: "nonfinite"}


class _Handle:
    """Owns one library handle; the destroy function is bound at creation so
    that release also works while the interpreter shuts down."""

    def _own(self, destroy):
        self._destroy = destroy

    def __del__(self):
        h, d = getattr(self, "h", None), getattr(self, "_destroy", None)
        if h and d is not None:
            d(h)
        self.h = None


def _pd(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


class Context(_Handle):
    """One device + one HIP stream (no reference counterpart)."""

    def __init__(self, device=0, stream=None):
        self.h = C.c_void_p()
        check(None, lib().mgldg_ctx_create(device, stream, C.byref(self.h)))
        self._own(lib().mgldg_ctx_destroy)
        self.device = device

    def sync(self):
        check(self, lib().mgldg_ctx_synchronize(self.h))


class Quadtree(_Handle):
    """Quadtree<N>(h, strategy, periodic) (src/cartesian/mesh/quadtree.h:211-228).

    h(x) -> widths (scalar or N values): a cell splits iff every width exceeds
    h(center) (quadtree.tpp:15).  balance=True enforces 2:1 leaf balance."""

    def __init__(self, ctx, N, h, lo=None, hi=None, coarsening="equal", periodic=False, balance=False):
        self.ctx, self.N = ctx, N

        def cb(x, out, user):
            v = np.broadcast_to(np.asarray(h(np.array([x[i] for i in range(N)])), dtype=np.float64), (N,))
            for i in range(N):
                out[i] = float(v[i])

        self._cb = SIZEFN(cb)
        lo_a = None if lo is None else np.ascontiguousarray(lo, dtype=np.float64)
        hi_a = None if hi is None else np.ascontiguousarray(hi, dtype=np.float64)
        self.h = C.c_void_p()
        strat = {"equal": 0, "rapid": 1}[coarsening]
        check(ctx, lib().mgldg_quadtree_create(ctx.h, N, _pd(lo_a), _pd(hi_a), self._cb, None, strat,
                                               int(periodic), int(balance), C.byref(self.h)))
        self._own(lib().mgldg_quadtree_destroy)
        info = np.zeros(3, dtype=np.int32)
        check(ctx, lib().mgldg_quadtree_info(self.h, _pi(info)))
        self.nnodes, self.num_levels, self.nleaves = [int(v) for v in info]

    def layer(self, l):
        n = C.c_int()
        check(self.ctx, lib().mgldg_quadtree_layer(self.h, l, None, C.byref(n)))
        ids = np.zeros(n.value, dtype=np.int32)
        check(self.ctx, lib().mgldg_quadtree_layer(self.h, l, _pi(ids), C.byref(n)))
        return ids


class _Problem(_Handle):
    """Operations shared by the quadtree and simplex discretisations."""

    OPS = {"A": 0, "M": 1, "T": 2, "G0": 3, "G1": 4, "G2": 5}

    def _sizes(self):
        self._own(lib().mgldg_problem_destroy)
        s = np.zeros(7, dtype=np.int32)
        check(self.ctx, lib().mgldg_problem_sizes(self.h, _pi(s)))
        self.ne, self.nf, self.nbf, self.npl, self.nqb, self.ndof, self.N = [int(v) for v in s]

    def nodes(self):
        x = np.zeros((self.ndof, self.N))
        check(self.ctx, lib().mgldg_problem_nodes(self.h, _pd(x)))
        return x

    def bc_qpoints(self):
        x = np.zeros((self.nqb, self.N))
        t = np.zeros(self.nqb, dtype=np.int32)
        check(self.ctx, lib().mgldg_problem_bc_qpoints(self.h, _pd(x), _pi(t)))
        return x, t

    def rhs(self, f=None, g=None):
        """computeRHS (ldgpoisson.tpp:167-176): b = M f + B_g g; f [ndof(,k)], g [nqb(,k)]"""
        arrs = [a for a in (f, g) if a is not None]
        k = 1 if not arrs or np.ndim(arrs[0]) == 1 else np.shape(arrs[0])[1]
        f_a = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        g_a = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        b = np.zeros((self.ndof, k)) if k > 1 else np.zeros(self.ndof)
        check(self.ctx, lib().mgldg_rhs(self.h, _pd(f_a), _pd(g_a), k, _pd(b)))
        return b

    def export(self, which):
        """(rowptr, cols, blocks[nnzb, npl, npl]) of A, M, T or G_d"""
        w = self.OPS[which]
        nz = C.c_long()
        check(self.ctx, lib().mgldg_problem_export(self.h, w, None, None, None, C.byref(nz)))
        rp = np.zeros(self.ne + 1, dtype=np.int32)
        cols = np.zeros(nz.value, dtype=np.int32)
        blk = np.zeros((nz.value, self.npl, self.npl))
        check(self.ctx, lib().mgldg_problem_export(self.h, w, _pi(rp), _pi(cols), _pd(blk), C.byref(nz)))
        return rp, cols, blk

    def setup_times(self):
        """host symbolic and device phases (s), the assembly kernel (ms) and its algorithmic bytes"""
        t = np.zeros(4)
        check(self.ctx, lib().mgldg_problem_setup_times(self.h, _pd(t)))
        return dict(host_s=t[0], device_s=t[1], assembly_kernel_ms=t[2], assembly_bytes=t[3])

    def write_mm(self, which, path):
        """SparseBlockMatrix::write (sparseblockmatrix.tpp:145-164) of A, M, T or G_d"""
        check(self.ctx, lib().mgldg_problem_write_mm(self.h, self.OPS[which], os.fsencode(path)))

    def write_fun(self, u, path):
        """Function::write (function.h:173-190): u [ndof] at the problem's nodes"""
        ua = _vec(u, self.ndof)
        check(self.ctx, lib().mgldg_problem_write_fun(self.h, _pd(ua), os.fsencode(path)))

    def dump(self, directory, rhs=None):
        """LDGPoisson::dump (ldgpoisson.h:100-109): A, M, T, G_d .mtx and rhs.fun into directory"""
        ra = None if rhs is None else _vec(rhs, self.ndof)
        check(self.ctx, lib().mgldg_problem_dump(self.h, _pd(ra), os.fsencode(directory)))


class LDGPoisson(_Problem):
    """Mesh<N,P>(qt) + BoundaryConditions + LDGPoisson<N,P>(mesh, bcs, tau0, tauD)
    (src/cartesian/mesh/mesh.h:186-265, src/cartesian/dg/ldgpoisson.h:60-77).

    bc: 'dirichlet' | 'neumann' | list of 2N types per boundary tag."""

    simplex = False

    def __init__(self, qt, p, bc="dirichlet", tau0=0.0, tauD=1000.0):
        self.qt, self.ctx, self.p = qt, qt.ctx, p
        if isinstance(bc, str):
            bc = [{"dirichlet": 0, "neumann": 1}[bc]] * (2 * qt.N)
        bct = np.ascontiguousarray(bc, dtype=np.int32)
        self.h = C.c_void_p()
        check(self.ctx, lib().mgldg_ldg_create(qt.h, p, _pi(bct), tau0, tauD, C.byref(self.h)))
        self._sizes()


class Wireframe(_Handle):
    """Wireframe<N>(p, t) / Wireframe<N>(file) (src/unstructured/mesh/wireframe.h:12-77).

    points [np, N]; simplices [nt, N+1] with 0-based vertex ids."""

    def __init__(self, ctx, points=None, simplices=None, path=None, N=2):
        self.ctx = ctx
        self.h = C.c_void_p()
        if path is not None:
            check(ctx, lib().mgldg_wireframe_read(ctx.h, N, os.fsencode(path), C.byref(self.h)))
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64)
            tri = np.ascontiguousarray(simplices, dtype=np.int32)
            check(ctx, lib().mgldg_wireframe_create(ctx.h, pts.shape[1], pts.shape[0], _pd(pts), tri.shape[0], _pi(tri),
                                                    C.byref(self.h)))
        self._own(lib().mgldg_wireframe_destroy)
        info = np.zeros(3, dtype=np.int32)
        check(ctx, lib().mgldg_wireframe_get(self.h, _pi(info), None, None))
        self.N, self.np, self.nt = [int(v) for v in info]

    @classmethod
    def read(cls, ctx, path, N=2):
        return cls(ctx, path=path, N=N)

    def arrays(self):
        p = np.zeros((self.np, self.N))
        t = np.zeros((self.nt, self.N + 1), dtype=np.int32)
        check(self.ctx, lib().mgldg_wireframe_get(self.h, None, _pd(p), _pi(t)))
        return p, t

    def write(self, path, digits=0):
        """the .mesh layout (writemesh.m:17-28); digits 0 = writemesh.m's "%f",
        1..17 significant digits (17 round-trips every double)"""
        check(self.ctx, lib().mgldg_wireframe_write(self.h, os.fsencode(path), digits))


class SimplexLDGPoisson(_Problem):
    """Mesh<N,P>(wireframe, bndfuns) + BoundaryConditions + LDGPoisson<N,P>
    (src/unstructured/mesh/mesh.h:185-298, src/unstructured/dg/ldgpoisson.tpp:14-170).

    bndfuns: predicates on a vertex (keys 1..len, first match); bc: 'dirichlet'
    | 'neumann' | one type per key.  vmask may be given instead of bndfuns."""

    simplex = True

    def __init__(self, wf, p, bndfuns=None, bc="dirichlet", tau0=0.0, tauD=1000.0, vmask=None):
        self.wf, self.ctx, self.p = wf, wf.ctx, p
        nb = len(bndfuns) if bndfuns else 0
        if vmask is None and nb:
            pts, _ = wf.arrays()
            vmask = np.zeros(len(pts), dtype=np.int32)
            for k, fn in enumerate(bndfuns):
                vmask |= np.array([bool(fn(x)) for x in pts], dtype=np.int32) << k
        elif vmask is not None:
            vmask = np.ascontiguousarray(vmask, dtype=np.int32)
            nb = int(vmask.max()).bit_length() if nb == 0 else nb
        if isinstance(bc, str):
            bc = [{"dirichlet": 0, "neumann": 1}[bc]] * max(nb, 1)
        bct = np.ascontiguousarray(bc, dtype=np.int32)
        self.h = C.c_void_p()
        check(self.ctx, lib().mgldg_ldg_create_simplex(wf.h, p, nb, _pi(vmask), _pi(bct), tau0, tauD, C.byref(self.h)))
        self._sizes()


def locality_order(points, simplices):
    """Permutation of the simplices along a Morton (Z-order) curve of their
    centroids.  Block rows follow element order, so the x_j gathers of the
    block kernels stay in cache when neighbouring elements have nearby
    indices; meshes from generators that emit elements in a scattered order
    (qhull's Delaunay triangulations, lattice splits emitted pattern by
    pattern) gain most.  Use as simplices = simplices[locality_order(points,
    simplices)] before building the Wireframe."""
    P = np.asarray(points, dtype=np.float64)
    S = np.asarray(simplices)
    c = P[S].mean(axis=1)
    d = c.shape[1]
    bits = min(21, 63 // d)
    lo, hi = c.min(axis=0), c.max(axis=0)
    q = ((c - lo) / np.maximum(hi - lo, 1e-300) * ((1 << bits) - 1)).astype(np.uint64)
    code = np.zeros(len(c), dtype=np.uint64)
    for b in range(bits):
        for a in range(d):
            code |= ((q[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(b * d + a)
    return np.argsort(code, kind="stable")


def agglomerate(prob, max_levels=10):
    """The library's Agglomeration<N,P>(mesh, max_levels) (agglomeration.h:21-34):
    list of partitions, level l elements -> level l+1 agglomerates."""
    nl = C.c_int()
    nparts = np.zeros(max_levels, dtype=np.int32)
    parts = np.zeros(max_levels * prob.ne, dtype=np.int32)
    check(prob.ctx, lib().mgldg_agglomerate(prob.h, max_levels, C.byref(nl), _pi(nparts), _pi(parts)))
    out, n = [], prob.ne
    for l in range(nl.value):
        out.append(parts[l * prob.ne:l * prob.ne + n].copy())
        n = int(nparts[l])
    return out


def mgridgen(xadj, vvol, vsurf, adjncy, adjwgt, minsize, maxsize, dim):
    """MGridGen(nvtxs, xadj, vvol, vsurf, adjncy, adjwgt, minsize, maxsize,
    {4, 6, 1, dim}, ...) (lib/ParMGridGen-1.0/MGridGen/Lib/mgridgen.c:18-49) by
    the library's multilevel agglomerator (host only).  Returns (nparts, part)."""
    xa = np.ascontiguousarray(xadj, dtype=np.int32)
    n = len(xa) - 1
    adj = np.ascontiguousarray(adjncy, dtype=np.int32)
    w = np.ascontiguousarray(adjwgt, dtype=np.float64)
    vv = np.ascontiguousarray(vvol, dtype=np.float64)
    vs = np.ascontiguousarray(vsurf, dtype=np.float64)
    if len(vv) != n or len(vs) != n or len(adj) != xa[-1] or len(w) != xa[-1]:
        raise MGLDGError(-1, "mgridgen: array lengths do not match xadj")
    opts = np.array([4, 6, 1, dim], dtype=np.int32)
    part = np.zeros(max(n, 1), dtype=np.int32)
    nparts, nmoves = C.c_int(), C.c_int()
    check(None, lib().mgldg_mgridgen(n, _pi(xa), _pd(vv), _pd(vs), _pi(adj), _pd(w), int(minsize), int(maxsize), _pi(opts),
                                     C.byref(nmoves), C.byref(nparts), _pi(part)))
    return nparts.value, part[:n]


class Multigrid(_Handle):
    """InterpolationHierarchy<N,P>(qt) + Multigrid<N,P>(ops, hierarchy, params)
    (src/cartesian/multigrid/interpolationhierarchy.h:30-82, multigrid.h:291-301).

    precision='fp64' builds the FP64 V-cycle only; 'mixed' also builds the FP32
    copies used by the mixed-precision V-cycle.

    On a simplex problem the hierarchy is agglomerated
    (src/unstructured/multigrid/interpolationhierarchy.h:88-204): by the
    library's own agglomeration, or from `partitions` (list of arrays, level l
    elements -> level l+1 agglomerates).

    pdegrees (quadtree problems): p-levels of these polynomial degrees below
    the problem's, Multigrid<N,P1,P2,...> (multigrid.h:435-639), then the
    h-levels at the last degree, or none if hcoarsen is False
    (StopCoarsening).

    pool (default True): levels whose operator repeats its blocks (quadtree
    levels) are read by the single-RHS kernels from a pool of the distinct
    blocks -- the same bits, a fraction of the HBM traffic.
    row_tiles ("auto" | "on" | "off"): on pooled levels, the GS and operator
    kernels that take rows with the same blocks 16 at a time on the matrix
    cores -- "auto" from 2048 tiles (32k block rows) in FP64 and 256 tiles
    (4k rows) in FP32.  Same operator,
    summation order differing in the last bits."""

    def __init__(self, prob, relaxation="gs", omega=0.8, coarse="cholesky", npre=3, npost=3,
                 restriction="ratrat", precision="fp64", coarse_dense_max=8192, partitions=None, pdegrees=None,
                 hcoarsen=True, pool=True, row_tiles="auto"):
        self.prob, self.ctx = prob, prob.ctx
        self.params = MGParams({"jacobi": 0, "gs": 1, "gs-lex": 2}[relaxation], omega,
                               {"cholesky": 0, "relaxation": 1}[coarse], npre, npost,
                               {"rat": 0, "ratrat": 1}[restriction], {"fp64": 0, "mixed": 1}[precision],
                               coarse_dense_max, 0 if pool else 1,
                               {"auto": 0, "on": 1, "off": 2}[row_tiles])
        self.h = C.c_void_p()
        if partitions is not None:
            parts = np.zeros(len(partitions) * prob.ne, dtype=np.int32)
            for l, q in enumerate(partitions):
                parts[l * prob.ne:l * prob.ne + len(q)] = q
            check(self.ctx, lib().mgldg_mg_create_agglomerated(prob.h, C.byref(self.params), len(partitions), _pi(parts),
                                                               C.byref(self.h)))
        elif pdegrees is not None:
            deg = np.ascontiguousarray(pdegrees, dtype=np.int32)
            check(self.ctx, lib().mgldg_mg_create_p(prob.h, C.byref(self.params), len(deg), _pi(deg), int(hcoarsen),
                                                    C.byref(self.h)))
        else:
            check(self.ctx, lib().mgldg_mg_create(prob.h, C.byref(self.params), C.byref(self.h)))
        self._own(lib().mgldg_mg_destroy)
        info = np.zeros(3, dtype=np.int64)
        check(self.ctx, lib().mgldg_mg_info(self.h, info.ctypes.data_as(C.POINTER(C.c_long)), len(info)))
        self.nlevels = int(info[0])
        self.level_rows, self.level_B, self.level_nnzb = [], [], []
        for l in range(self.nlevels):
            nb, B, nz = C.c_int(), C.c_int(), C.c_long()
            check(self.ctx, lib().mgldg_mg_level(self.h, l, C.byref(nb), C.byref(B), C.byref(nz)))
            self.level_rows.append(nb.value); self.level_B.append(B.value); self.level_nnzb.append(nz.value)

    def export(self, level, which):
        w = {"A": 0, "M": 1, "T": 2, "G0": 3, "G1": 4, "G2": 5, "interp": 6}[which]
        nz, nbr, nbc = C.c_long(), C.c_int(), C.c_int()
        check(self.ctx, lib().mgldg_mg_export(self.h, level, w, None, None, None, C.byref(nz), C.byref(nbr), C.byref(nbc)))
        rp = np.zeros(nbr.value + 1, dtype=np.int32)
        cols = np.zeros(nz.value, dtype=np.int32)
        br = self.level_B[level]
        bc = self.level_B[level + 1] if which == "interp" else br
        blk = np.zeros((nz.value, br, bc))
        check(self.ctx, lib().mgldg_mg_export(self.h, level, w, _pi(rp), _pi(cols), _pd(blk), C.byref(nz), None, None))
        return rp, cols, blk, nbr.value, nbc.value

    def write_mm(self, level, which, path):
        """MatrixMarket dump of a level operator or interpolation (which as export)"""
        w = {"A": 0, "M": 1, "T": 2, "G0": 3, "G1": 4, "G2": 5, "interp": 6}[which]
        check(self.ctx, lib().mgldg_mg_write_mm(self.h, level, w, os.fsencode(path)))

    def vcycle_matrix(self, precision="fp64"):
        """Multigrid::vcycle_matrix (multigrid.h:386-400): column i = one V-cycle
        from zero applied to e_i, [ndof, ndof]"""
        n = self.prob.ndof
        V = np.zeros((n, n))
        check(self.ctx, lib().mgldg_mg_vcycle_matrix(self.h, _prec(precision), _pd(V)))
        return V

    def level_pool(self, level):
        """(distinct A blocks, distinct D^-1 blocks) of a pooled level, (0, 0) otherwise"""
        a, d = C.c_int(), C.c_int()
        check(self.ctx, lib().mgldg_mg_level_pool(self.h, level, C.byref(a), C.byref(d)))
        return a.value, d.value

    def solve_cost(self, precision="fp64", flexible=None):
        """cost model of the single-RHS MG-PCG: kernel launches and
        algorithmic HBM bytes of one preconditioner V-cycle and of one whole
        CG iteration, and the node count of the V-cycle's captured graph (-1
        before a graph solve).  flexible defaults to the mixed form's use."""
        fl = precision == "mixed" if flexible is None else flexible
        n, b, ni, bi, g = C.c_int(), C.c_double(), C.c_int(), C.c_double(), C.c_int()
        check(self.ctx, lib().mgldg_mg_solve_cost(self.h, _prec(precision), int(fl), C.byref(n), C.byref(b), C.byref(ni),
                                                  C.byref(bi), C.byref(g)))
        return {"vcycle_launches": n.value, "vcycle_bytes": b.value, "iteration_launches": ni.value,
                "iteration_bytes": bi.value, "graph_nodes": g.value}

    def level_tiles(self, level):
        """(tiles of 16 rows, table width, used by the FP64 single-RHS
        kernels, used by the FP32 ones) of a level"""
        t, w, u = C.c_int(), C.c_int(), C.c_int()
        check(self.ctx, lib().mgldg_mg_level_tiles(self.h, level, C.byref(t), C.byref(w), C.byref(u)))
        return t.value, w.value, bool(u.value & 1), bool(u.value & 2)

    def colors(self, level):
        nc = C.c_int()
        c = np.zeros(self.level_rows[level], dtype=np.int32)
        check(self.ctx, lib().mgldg_mg_colors(self.h, level, _pi(c), C.byref(nc)))
        return c, nc.value

    # unit-parity hooks on device tensors (torch) -----------------------------
    def _level_n(self, level):
        if not 0 <= level < self.nlevels:
            raise MGLDGError(-1, f"level {level} out of range (0..{self.nlevels - 1})")
        return self.level_rows[level] * self.level_B[level]

    def spmv(self, level, x, y, precision="fp64"):
        """y = A_l x; x, y of the level's length, float64 (fp64) or float32 (mixed)"""
        n, dt = self._level_n(level), _vdtype(precision)
        check(self.ctx, lib().mgldg_spmv_device(self.h, level, _prec(precision), _ptr(x, n, dt, "x"), _ptr(y, n, dt, "y"), 1))

    def smooth(self, level, b, x, nsweeps=1, reverse=False, precision="fp64"):
        n, dt = self._level_n(level), _vdtype(precision)
        check(self.ctx, lib().mgldg_smooth_device(self.h, level, _prec(precision), nsweeps, int(reverse), _ptr(b, n, dt, "b"),
                                                  _ptr(x, n, dt, "x"), 1))

    def vcycle(self, b, x, precision="fp64"):
        """x = V(b) from zero; b and x are float64 of the fine length in either precision"""
        n = self.prob.ndof
        check(self.ctx, lib().mgldg_vcycle_device(self.h, _prec(precision), _ptr(b, n, name="b"), _ptr(x, n, name="x"), 1))

    def profile_relax(self, level=0, precision="fp64", nrep=10, part="sweep"):
        """(ms per smoother launch, algorithmic bytes per launch, launches);
        part: 'sweep' (all launches of a sweep), 'single' (the one-colour GS
        launches) or 'pair' (the launch doing the last two colours)"""
        ms, by, n = C.c_double(), C.c_double(), C.c_int()
        pk = {"sweep": 0, "single": 1, "pair": 2}[part]
        check(self.ctx, lib().mgldg_profile_relax_part(self.h, level, _prec(precision), nrep, pk, C.byref(ms), C.byref(by),
                                                      C.byref(n)))
        return ms.value, by.value, n.value

    def profile_spmv(self, level=0, precision="fp64", nrep=10):
        ms, by = C.c_double(), C.c_double()
        check(self.ctx, lib().mgldg_profile_spmv(self.h, level, _prec(precision), nrep, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def batch_kernels(self, mode="auto"):
        """kernel family of batched solves: 'auto' (autotuned per level), 'vector',
        or the number of 16-column MFMA tiles (1, 2, 4)"""
        m = {"auto": -1, "vector": 0}.get(mode, mode)
        check(self.ctx, lib().mgldg_mg_batch_kernels(self.h, int(m)))

    def batch_plan(self):
        """per precision ('fp64', 'fp32') and level: (apply, smoother) kernel
        choice (0 vector, nt MFMA tiles) and the autotuner's timings in ms"""
        nl = self.nlevels
        plan = np.zeros(4 * nl, dtype=np.int32)
        ms = np.zeros(32 * nl, dtype=np.float32)
        check(self.ctx, lib().mgldg_mg_batch_plan(self.h, _pi(plan), ms.ctypes.data_as(C.POINTER(C.c_float))))
        out = {}
        for p, name in enumerate(("fp64", "fp32")):
            out[name] = [dict(apply=int(plan[(p * nl + l) * 2]), smooth=int(plan[(p * nl + l) * 2 + 1]),
                              apply_ms=ms[((p * nl + l) * 2) * 4:((p * nl + l) * 2) * 4 + 4].tolist(),
                              smooth_ms=ms[((p * nl + l) * 2 + 1) * 4:((p * nl + l) * 2 + 1) * 4 + 4].tolist())
                         for l in range(nl)]
        return out

    def setup_times(self):
        t = np.zeros(4)
        check(self.ctx, lib().mgldg_setup_times(self.h, _pd(t)))
        return dict(assembly_host=t[0], assembly_device=t[1], hierarchy_host=t[2], hierarchy_device=t[3])


def _vec(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if a.size != n:
        raise MGLDGError(-1, f"expected {n} values, got {a.size}")
    return a


# ---------------------------------------------------------------- interop file formats
def write_mm(path, rowptr, cols, blocks, nbcols=None):
    """SparseBlockMatrix::write (sparseblockmatrix.tpp:145-164) of a block CSR
    matrix: blocks [nnzb, br, bc] row-major; nbcols defaults to the row count"""
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    cl = np.ascontiguousarray(cols, dtype=np.int32)
    bl = np.ascontiguousarray(blocks, dtype=np.float64)
    nbr = len(rp) - 1
    nbc = nbr if nbcols is None else nbcols
    if bl.ndim != 3 or bl.shape[0] != len(cl) or rp[-1] != len(cl):
        raise MGLDGError(-1, "blocks must be [nnzb, br, bc] with nnzb = rowptr[-1] = len(cols)")
    check(None, lib().mgldg_bsr_write_mm(os.fsencode(path), nbr, nbc, bl.shape[1], bl.shape[2], _pi(rp), _pi(cl), _pd(bl)))


def write_fun(path, P, xyz, u, npl):
    """Function::write (function.h:173-190): header "P N ne", per element its
    npl nodes as "x .. value" lines and an empty line; xyz [ndof, N], u [ndof]"""
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    ua = _vec(u, x.shape[0])
    if x.shape[0] % npl:
        raise MGLDGError(-1, "ndof is not a multiple of npl")
    check(None, lib().mgldg_fun_write(os.fsencode(path), P, x.shape[1], x.shape[0] // npl, npl, _pd(x), _pd(ua)))


def read_fun(path):
    """Read a .fun file as src/DGTools.m:71-83 imports it: the header
    (P, N, ne) and the per-element node lines as data [ne, npl, N + 1]
    (coordinates, then the coefficient)."""
    with open(path) as f:
        head, _, body = f.read().partition("\n")
    P, N, ne = (int(v) for v in head.split())
    elems = [e for e in body.split("\n\n") if e.strip()]
    data = np.array([[[float(v) for v in ln.split()] for ln in e.strip("\n").split("\n")] for e in elems])
    if ne == 0 and data.size == 0:
        return (P, N, ne), np.zeros((0, 0, N + 1))
    if data.ndim != 3 or data.shape[0] != ne or data.shape[2] != N + 1:
        raise MGLDGError(-1, f"{path}: {data.shape} does not match the header {(P, N, ne)}")
    return (P, N, ne), data.reshape(ne, -1, N + 1)


def write_mesh(path, points, simplices, digits=0):
    """The .mesh layout (writemesh.m:17-28, read by Wireframe.read); simplices
    0-based; digits 0 = "%f" as writemesh.m, 1..17 significant digits"""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    tri = np.ascontiguousarray(simplices, dtype=np.int32)
    check(None, lib().mgldg_mesh_write(os.fsencode(path), pts.shape[1], pts.shape[0], _pd(pts), tri.shape[0],
                                       _pi(tri), digits))


def write_dense(path, a):
    """A dense matrix as the driver writes V.dat (dg_tests.cpp:337-347)"""
    A = np.ascontiguousarray(a, dtype=np.float64)
    if A.ndim != 2:
        raise MGLDGError(-1, "expected a matrix")
    check(None, lib().mgldg_dense_write(os.fsencode(path), A.shape[0], A.shape[1], _pd(A)))


def _prec(p):
    return {"fp64": 0, "mixed": 1, "fp32": 1}[p]


def _vdtype(p):
    """element type of the level vectors of the unit hooks in a precision"""
    return "float64" if _prec(p) == 0 else "float32"


def _ptr(t, numel=None, dtype="float64", name="tensor"):
    """Device pointer of a torch tensor, checked before anything reaches the
    library: the entry points read and write numel elements of dtype
    ("float64" or "float32") at that address, so a tensor of another type or
    length would be an out-of-bounds device access.  The library works on
    its own stream and returns once that stream is idle, so torch's work
    still queued on its current stream (the kernel that filled or wrote the
    tensor) is waited for here, before the library reads or writes it."""
    if not hasattr(t, "data_ptr"):
        raise MGLDGError(-1, f"{name}: expected a torch tensor")
    if str(t.dtype) != "torch." + dtype:
        raise MGLDGError(-1, f"{name}: expected dtype {dtype}, got {t.dtype}")
    if numel is not None and t.numel() != numel:
        raise MGLDGError(-1, f"{name}: expected {numel} elements, got {t.numel()} (shape {tuple(t.shape)})")
    if not t.is_contiguous():
        raise MGLDGError(-1, f"{name}: expected a contiguous tensor")
    if not t.is_cuda:
        raise MGLDGError(-1, f"{name}: expected a device tensor")
    import torch
    torch.cuda.current_stream(t.device).synchronize()
    return C.c_void_p(t.data_ptr())


def _cols(t, rows, name):
    """number of right-hand sides of a [rows] or [rows, k] tensor"""
    if t.ndim not in (1, 2) or t.shape[0] != rows:
        raise MGLDGError(-1, f"{name}: expected shape ({rows},) or ({rows}, k), got {tuple(t.shape)}")
    return 1 if t.ndim == 1 else int(t.shape[1])


def pcg(mg, b, x0=None, tol=1e-8, maxit=200, criterion="reference", flexible=False, precision="fp64",
        use_graph=False):
    """pcg(A, b, x, precon=V-cycle, tol, maxit) (src/cartesian/multigrid/krylov.h:22-68).

    b: numpy array (host path) or torch device tensor (device path, x0 updated
    in place).  Returns (x, stats dict)."""
    sp = SolveParams(tol, maxit, {"reference": 0, "true": 1}[criterion], int(flexible), _prec(precision),
                     int(use_graph))
    is_dev = hasattr(b, "data_ptr")
    n = mg.prob.ndof
    if is_dev:
        import torch
        nrhs = _cols(b, n, "b")
        x = torch.zeros_like(b) if x0 is None else x0
        pb, px = _ptr(b, n * nrhs, name="b"), _ptr(x, n * nrhs, name="x0")
        st = (SolveStats * nrhs)()
        torch.cuda.synchronize()
        check(mg.ctx, lib().mgldg_solve_device(mg.h, pb, px, nrhs, C.byref(sp), st))
    else:
        bb = np.ascontiguousarray(b, dtype=np.float64)
        if bb.ndim not in (1, 2) or bb.shape[0] != n:
            raise MGLDGError(-1, f"b: expected shape ({n},) or ({n}, k), got {bb.shape}")
        nrhs = 1 if bb.ndim == 1 else bb.shape[1]
        x = np.zeros_like(bb) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        if x.shape != bb.shape:
            raise MGLDGError(-1, f"x0: expected shape {bb.shape}, got {x.shape}")
        st = (SolveStats * nrhs)()
        check(mg.ctx, lib().mgldg_solve(mg.h, _pd(bb), _pd(x), nrhs, C.byref(sp), st))
    stats = [dict(iterations=s.iterations, status=STATUS.get(s.status, s.status), relres_true=s.relres_true,
                  relres_prec=s.relres_prec, seconds=s.seconds) for s in st]
    return x, (stats[0] if nrhs == 1 else stats)


def vcycles(mg, b, x0=None, ncycle=10, tol=0.0, precision="fp64"):
    """The driver's stationary solver `--solver vcycles`
    (tests/cartesian/dg_tests.cpp:309-332): up to ncycle V-cycles
    (multigrid.h:361-383) from x0 (zero if None); tol > 0 stops once
    ||b - A x|| <= tol ||b||.  precision "mixed": x += V(b - A x) with an
    FP32 V-cycle, residual and update in FP64.

    b: FP64 torch device tensor (x0, if given, updated in place) or numpy
    array.  Returns (x, residual norms before and after each cycle run)."""
    import torch
    is_dev = hasattr(b, "data_ptr")
    n = mg.prob.ndof
    if is_dev:
        _ptr(b, n, name="b")
    bd = b if is_dev else torch.tensor(np.ascontiguousarray(b, dtype=np.float64), device="cuda")
    if x0 is None:
        xd = torch.zeros_like(bd)
    elif is_dev:
        xd = x0
    else:
        xd = torch.tensor(np.asarray(x0, dtype=np.float64), device="cuda")
    pb, px = _ptr(bd, n, name="b"), _ptr(xd, n, name="x0")
    rn = np.zeros(int(ncycle) + 1)
    cyc = C.c_int()
    torch.cuda.synchronize()
    check(mg.ctx, lib().mgldg_vcycles_device(mg.h, _prec(precision), pb, px, int(ncycle), float(tol), _pd(rn),
                                            C.byref(cyc)))
    return (xd if is_dev else xd.cpu().numpy()), rn[:cyc.value + 1]


def mean_zero(ctx, v):
    """Function::meanZero (src/cartesian/dg/function.h:156-170) in place on an
    FP64 device tensor [n] or [n, k] (each column); the driver applies it to
    the right-hand side of Neumann and periodic problems (dg_tests.cpp:282-286)."""
    import torch
    nrhs = _cols(v, v.shape[0] if v.ndim else -1, "v")
    pv = _ptr(v, v.shape[0] * nrhs, name="v")
    torch.cuda.synchronize()
    check(ctx, lib().mgldg_mean_zero_device(ctx.h, v.shape[0], pv, nrhs))
    return v


def rhs_device(mg, f=None, g=None):
    """b = M f + B_g g on the GPU (ldgpoisson.tpp:103-176).  f [ndof(,k)] and g
    [nqb(,k)] are FP64 device tensors (either may be None)."""
    import torch
    if f is None and g is None:
        raise MGLDGError(-1, "rhs_device: f and g are both None")
    n, nq = mg.prob.ndof, mg.prob.nqb
    k = _cols(f, n, "f") if f is not None else _cols(g, nq, "g")
    pf = None if f is None else _ptr(f, n * k, name="f")
    pg = None if g is None else _ptr(g, nq * k, name="g")
    ref = f if f is not None else g
    b = torch.empty((n,) if k == 1 else (n, k), dtype=torch.float64, device=ref.device)
    torch.cuda.synchronize()
    check(mg.ctx, lib().mgldg_rhs_device(mg.h, pf, pg, k, _ptr(b, n * k, name="b")))
    return b


def adjoint(mg, dLdu, tol=1e-12, maxit=500, criterion="true", flexible=False, precision="fp64", use_graph=False):
    """Adjoint solve through the same MG-PCG: lambda = A^{-1} dL/du, then
    dL/df = M lambda and dL/dg = B_g^T lambda.  Returns (lambda, dLdf, dLdg, stats)."""
    import torch
    n = mg.prob.ndof
    k = _cols(dLdu, n, "dLdu")
    dLdu = dLdu.contiguous()
    _ptr(dLdu, n * k, name="dLdu")
    sp = SolveParams(tol, maxit, {"reference": 0, "true": 1}[criterion], int(flexible), _prec(precision), int(use_graph))
    st = (SolveStats * k)()
    lam = torch.empty_like(dLdu)
    dLdf = torch.empty_like(dLdu)
    nq = mg.prob.nqb
    dLdg = torch.empty((nq,) if k == 1 else (nq, k), dtype=torch.float64, device=dLdu.device)
    torch.cuda.synchronize()
    check(mg.ctx, lib().mgldg_adjoint_device(mg.h, _ptr(dLdu.contiguous()), _ptr(lam), _ptr(dLdf), _ptr(dLdg), k,
                                             C.byref(sp), st))
    stats = [dict(iterations=s.iterations, status=STATUS.get(s.status, s.status), relres_true=s.relres_true) for s in st]
    return lam, dLdf, dLdg, (stats[0] if k == 1 else stats)


def adjoint_maps(mg, lam):
    """dL/df = M lam and dL/dg = B_g^T lam for adjoint states solved elsewhere
    (device tensors [ndof(,k)]); returns (dLdf, dLdg)."""
    import torch
    n = mg.prob.ndof
    k = _cols(lam, n, "lam")
    lam = lam.contiguous()
    _ptr(lam, n * k, name="lam")
    dLdf = torch.empty_like(lam)
    nq = mg.prob.nqb
    dLdg = torch.empty((nq,) if k == 1 else (nq, k), dtype=torch.float64, device=lam.device)
    torch.cuda.synchronize()
    check(mg.ctx, lib().mgldg_adjoint_maps_device(mg.h, _ptr(lam), _ptr(dLdf), _ptr(dLdg), k))
    return dLdf, dLdg


def _make_autograd():
    import torch

    class LDGSolve(torch.autograd.Function):
        """u = A^{-1} (M f + B_g g) with a hand-written adjoint backward.

        forward : device RHS, MG-PCG (FP64 or the mixed form)
        backward: dL/df = M A^{-1} dL/du, dL/dg = B_g^T A^{-1} dL/du (A symmetric)"""

        @staticmethod
        def forward(ctx, f, g, mg, opts):
            b = rhs_device(mg, f, g)
            u = torch.zeros_like(b)
            _, st = pcg(mg, b, u, **opts)
            _require_converged("forward solve", st)
            ctx.mg, ctx.opts = mg, opts
            ctx.has_f, ctx.has_g = f is not None, g is not None
            return u

        @staticmethod
        def backward(ctx, gu):
            _, dLdf, dLdg, st = adjoint(ctx.mg, gu.contiguous(), **ctx.opts)
            _require_converged("adjoint solve", st)
            return (dLdf if ctx.has_f else None), (dLdg if ctx.has_g else None), None, None

    return LDGSolve


def _require_converged(what, st):
    """a solve that stopped at maxit, broke down or met NaN/Inf must not feed
    autograd a wrong u or gradient"""
    for c, d in enumerate(st if isinstance(st, list) else [st]):
        if d["status"] != "converged":
            raise MGLDGError(-7, f"{what}: column {c} {d['status']} after {d['iterations']} iterations "
                                 f"(relative residual {d['relres_true']:.3e})")


_ldg_solve_cls = None


def ldg_solve(f, g, mg, **opts):
    """Differentiable solve u(f, g) (torch.autograd).  opts as for pcg()."""
    global _ldg_solve_cls
    if _ldg_solve_cls is None:
        _ldg_solve_cls = _make_autograd()
    o = dict(tol=1e-12, maxit=500, criterion="true")
    o.update(opts)
    return _ldg_solve_cls.apply(f, g, mg, o)
