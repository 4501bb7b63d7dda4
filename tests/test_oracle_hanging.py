"""Pins of the oracle's LDG assembly on adaptive quadtrees (CPU).

Hanging (non-canonical) faces take the reference's quadrature branch of
Mesh::lift (src/cartesian/mesh/mesh.h:312-351) with the face points of
Face::Face (mesh.h:44-77).  Three independent checks pin that branch:

* a brute-force LDG assembler written here in numpy from the mathematics
  alone: element boxes from the tree, faces found geometrically as the
  overlaps of touching element sides, Gauss-Legendre points of numpy's own
  rule on each overlap, Gauss-Lobatto nodes as numpy's Legendre roots.  It
  forms M, T (the jump penalty tau0 * int [u][v] on interior faces, tauD on
  Dirichlet faces), G_d (nodal derivative plus the one-sided lifting of
  ldgpoisson.tpp:60-99) and A = T + sum_d G_d^T M G_d, compared block for
  block with the oracle's exports;
* the patch test: g = x + y, f = 0 is reproduced to round-off;
* symmetry, positive definiteness and O(h^{p+1}) convergence of the
  manufactured solution of tests/cartesian/dg_tests.cpp:235-237 on graded
  trees.

A wrong sub-face mapping (the face points of the coarse side placed on the
wrong half, a swapped left/right evaluation, a dropped area factor) changes
T's off-diagonal blocks of every hanging pair and fails the first check; the
patch test catches a lift that is inconsistent with G_d."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from numpy.polynomial import legendre as npleg


def graded(dx):
    """the driver's adaptive size function (tests/cartesian/dg_tests.cpp:212-215):
    4:1 hanging faces at x = 0.5 when not balanced"""
    return lambda x: dx if x[0] > 0.5 else (0.25 * dx if x[0] > 0.25 else 0.5 * dx)


def point_rule(c, lo, hi):
    """distance refinement towards (0.3, 0.7) (the adaptive benchmark's rule, smaller)"""
    return lambda x: max(lo, min(float(np.hypot(x[0] - 0.3, x[1] - 0.7)) / c, hi))


TREES = {
    # name: (N, size function, balance)
    "graded4to1": (2, graded(0.25), False),
    "point2to1": (2, point_rule(4.0, 2.0 ** -6, 0.5), True),
    "graded3d": (3, lambda x: 0.5 if x[0] > 0.5 else 0.25, False),
}


# ------------------------------------------------------------------ brute force
def gl_nodes(P):
    """Gauss-Lobatto points on [0, 1]: the ends and the roots of P'_{P-1}"""
    if P == 1:
        return np.array([0.5])
    inner = npleg.Legendre.basis(P - 1).deriv().roots()
    return np.sort(np.concatenate([[-1.0], np.real(inner), [1.0]]) * 0.5 + 0.5)


def lagrange_1d(nodes, t):
    """values [len(t), P] and derivatives of the Lagrange basis at nodes"""
    t = np.atleast_1d(np.asarray(t, dtype=float))
    P = len(nodes)
    V = np.ones((len(t), P))
    dV = np.zeros((len(t), P))
    for j in range(P):
        others = [nodes[m] for m in range(P) if m != j]
        den = np.prod([nodes[j] - o for o in others]) if others else 1.0
        coef = np.poly(others) / den if others else np.array([1.0])
        V[:, j] = np.polyval(coef, t)
        dV[:, j] = np.polyval(np.polyder(coef), t) if P > 1 else 0.0
    return V, dV


class BruteLDG:
    """LDG Poisson on the leaves of a tree, all boundaries Dirichlet"""

    def __init__(self, boxes, N, p, tau0, tauD):
        self.lo, self.hi = boxes
        self.N, self.P = N, p + 1
        self.ne = len(self.lo)
        self.b = self.P ** N
        self.g = gl_nodes(self.P)
        self.idx = np.array(list(itertools.product(range(self.P), repeat=N)))   # first coordinate slowest
        xg, wg = npleg.leggauss(8)
        self.xg, self.wg = 0.5 * (xg + 1), 0.5 * wg
        b, ne = self.b, self.ne
        self.M = np.zeros((ne, b, b))
        self.T = {}
        self.G = [dict() for _ in range(N)]
        self._mass()
        for e in range(ne):
            for d in range(N):
                self._add(self.G[d], e, e, self._deriv(e, d))
        self._faces(tau0, tauD)

    def phi(self, e, X):
        """basis of element e at physical points X [m, N] -> [m, b]"""
        out = np.ones((len(X), self.b))
        for d in range(self.N):
            xi = (X[:, d] - self.lo[e, d]) / (self.hi[e, d] - self.lo[e, d])
            V, _ = lagrange_1d(self.g, xi)
            out *= V[:, self.idx[:, d]]
        return out

    def _mass(self):
        pts = np.array(list(itertools.product(self.xg, repeat=self.N)))
        wts = np.prod(np.array(list(itertools.product(self.wg, repeat=self.N))), axis=1)
        for e in range(self.ne):
            w = self.hi[e] - self.lo[e]
            X = self.lo[e] + pts * w
            F = self.phi(e, X)
            self.M[e] = (F * (wts * np.prod(w))[:, None]).T @ F

    def _deriv(self, e, d):
        """(d/dx_d phi_j)(node_i): the nodal derivative of element e"""
        _, dV = lagrange_1d(self.g, self.g)
        V = np.eye(self.P)
        D = np.ones((self.b, self.b))
        for k in range(self.N):
            Mk = dV if k == d else V
            D *= Mk[np.ix_(self.idx[:, k], self.idx[:, k])]
        return D / (self.hi[e, d] - self.lo[e, d])

    def _add(self, mat, i, j, blk):
        mat[(i, j)] = mat.get((i, j), 0) + blk

    def _face_points(self, lo, hi, d):
        """Gauss points and weights on the face box [lo, hi] (degenerate in d)"""
        dims = [k for k in range(self.N) if k != d]
        pts = np.array(list(itertools.product(self.xg, repeat=self.N - 1)))
        wts = np.prod(np.array(list(itertools.product(self.wg, repeat=self.N - 1))), axis=1)
        X = np.tile(lo, (len(pts), 1))
        area = 1.0
        for c, k in enumerate(dims):
            X[:, k] = lo[k] + pts[:, c] * (hi[k] - lo[k])
            area *= hi[k] - lo[k]
        return X, wts * area

    def _faces(self, tau0, tauD):
        lo, hi, N, tol = self.lo, self.hi, self.N, 1e-12
        self.nhanging = 0
        for d in range(N):
            for a in range(self.ne):
                # interior faces with a on the left (low side)
                for bb in np.nonzero(np.abs(lo[:, d] - hi[a, d]) < tol)[0]:
                    flo, fhi = np.maximum(lo[a], lo[bb]), np.minimum(hi[a], hi[bb])
                    if any(fhi[k] - flo[k] <= tol for k in range(N) if k != d):
                        continue
                    flo[d] = fhi[d] = hi[a, d]
                    X, w = self._face_points(flo, fhi, d)
                    Fa, Fb = self.phi(a, X), self.phi(bb, X)
                    LL, RR, RL = (Fa * w[:, None]).T @ Fa, (Fb * w[:, None]).T @ Fb, (Fb * w[:, None]).T @ Fa
                    same = np.allclose(lo[a][np.arange(N) != d], lo[bb][np.arange(N) != d]) and \
                        np.allclose(hi[a][np.arange(N) != d], hi[bb][np.arange(N) != d])
                    self.nhanging += not same
                    Minv = np.linalg.inv(self.M[bb])
                    self._add(self.G[d], bb, bb, Minv @ RR)
                    self._add(self.G[d], bb, a, -Minv @ RL)
                    self._add(self.T, bb, bb, tau0 * RR)
                    self._add(self.T, a, a, tau0 * LL)
                    self._add(self.T, bb, a, -tau0 * RL)
                    self._add(self.T, a, bb, -tau0 * RL.T)
                # Dirichlet faces of the domain [0, 1]^N
                for side, n in ((0, -1.0), (1, 1.0)):
                    at = lo[a, d] if side == 0 else hi[a, d]
                    if abs(at - side) > tol:
                        continue
                    flo, fhi = lo[a].copy(), hi[a].copy()
                    flo[d] = fhi[d] = at
                    X, w = self._face_points(flo, fhi, d)
                    Fa = self.phi(a, X)
                    LL = (Fa * w[:, None]).T @ Fa
                    self._add(self.G[d], a, a, -n * np.linalg.inv(self.M[a]) @ LL)
                    self._add(self.T, a, a, tauD * LL)

    def dense(self, which):
        b, n = self.b, self.ne * self.b
        D = np.zeros((n, n))
        if which == "M":
            for e in range(self.ne):
                D[e * b:(e + 1) * b, e * b:(e + 1) * b] = self.M[e]
            return D
        if which == "A":
            Mf = self.dense("M")
            D = self.dense("T")
            for d in range(self.N):
                Gd = self.dense(f"G{d}")
                D += Gd.T @ Mf @ Gd
            return D
        mat = self.T if which == "T" else self.G[int(which[1])]
        for (i, j), blk in mat.items():
            D[i * b:(i + 1) * b, j * b:(j + 1) * b] += blk
        return D


def boxes_of(qt):
    ids = qt.layer(0)
    nodes = [qt.node(int(i)) for i in ids]
    return np.array([n["lo"] for n in nodes]), np.array([n["hi"] for n in nodes])


def tree(O, name):
    N, h, bal = TREES[name]
    return O.Quadtree(N, h, rapid=True, balance=bal)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


# ------------------------------------------------------------------ tests
@pytest.mark.parametrize("name,p", [("graded4to1", 1), ("graded4to1", 2), ("graded4to1", 3), ("point2to1", 1),
                                    ("point2to1", 2), ("graded3d", 1)])
def test_assembly_equals_brute_force(oracle, name, p):
    qt = tree(oracle, name)
    tau0, tauD = 1.7, 1000.0
    pr = oracle.Problem(qt, p, tau0=tau0, tauD=tauD)
    bf = BruteLDG(boxes_of(qt), qt.N, p, tau0, tauD)
    assert bf.nhanging > 0
    # node positions: the element boxes at Gauss-Lobatto points, first coordinate slowest
    X = np.concatenate([bf.lo[e] + (bf.hi[e] - bf.lo[e]) * bf.g[bf.idx] for e in range(bf.ne)])
    assert np.abs(pr.nodes() - X).max() < 1e-15
    for w in ["M", "T"] + [f"G{d}" for d in range(qt.N)] + ["A"]:
        ref = bf.dense(w)
        got = pr.dense(w)
        assert rel(got, ref) < 1e-12, (w, rel(got, ref))


def test_hanging_face_blocks(oracle):
    """4:1 hanging faces at x = 0.5: T's off-diagonal block is -tau0 RL with
    RL_ij = int_f phi^right_i phi^left_j over the small element's side"""
    qt = tree(oracle, "graded4to1")
    pr = oracle.Problem(qt, 2, tau0=1.0)
    lo, hi = boxes_of(qt)
    rp, cols, blk = pr.export("T")
    bf = BruteLDG((lo, hi), 2, 2, 1.0, 1000.0)
    seen = 0
    for r in range(pr.ne):
        for k in range(rp[r], rp[r + 1]):
            l = cols[k]
            if l == r or abs(hi[l, 0] - lo[r, 0]) > 1e-12 or hi[l, 0] != 0.5:
                continue
            wl, wr = hi[l, 1] - lo[l, 1], hi[r, 1] - lo[r, 1]
            if max(wl, wr) < 3.9 * min(wl, wr):
                continue
            # the face is the small element's side
            y0, y1 = max(lo[l, 1], lo[r, 1]), min(hi[l, 1], hi[r, 1])
            assert abs((y1 - y0) - min(wl, wr)) < 1e-15
            xg, wg = npleg.leggauss(10)
            y = y0 + 0.5 * (xg + 1) * (y1 - y0)
            X = np.stack([np.full_like(y, 0.5), y], axis=1)
            w = 0.5 * wg * (y1 - y0)
            RL = (bf.phi(r, X) * w[:, None]).T @ bf.phi(l, X)
            assert np.abs(blk[k] + RL).max() < 1e-13 * np.abs(RL).max()
            seen += 1
    assert seen >= 4


@pytest.mark.parametrize("name", ["graded4to1", "point2to1", "graded3d"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_patch_test_adaptive(oracle, name, p):
    """g = x + y (+ z), f = 0: LDG reproduces the linear solution to round-off
    across hanging faces"""
    if name == "graded3d" and p == 3:
        pytest.skip("3-D p = 3 dense solve is large; p = 1, 2 cover the 3-D faces")
    qt = tree(oracle, name)
    pr = oracle.Problem(qt, p, tau0=1.0)
    X = pr.nodes()
    xq, _ = pr.bc_qpoints()
    b = pr.rhs(np.zeros(pr.ndof), xq.sum(axis=1))
    A = sp.csr_matrix(pr.dense("A"))
    u = spla.spsolve(A.tocsc(), b)
    assert np.abs(u - X.sum(axis=1)).max() < 1e-10


@pytest.mark.parametrize("name,p", [("graded4to1", 1), ("graded4to1", 3), ("point2to1", 2)])
def test_symmetric_spd_adaptive(oracle, name, p):
    qt = tree(oracle, name)
    pr = oracle.Problem(qt, p, tau0=0.0)
    A = pr.dense("A")
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0
    T = oracle.Problem(qt, p, tau0=2.0).dense("T")
    assert np.abs(T - T.T).max() < 1e-13 * np.abs(T).max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_convergence_order_adaptive(oracle, p):
    """the manufactured solution of tests/cartesian/dg_tests.cpp:235-237 on
    graded trees with 4:1 hanging faces, every cell halved per step:
    O(h^{p+1}) in the nodal max norm"""
    errs = []
    uf = lambda Y: np.cos(2 * np.pi * Y[:, 0]) * np.cos(2 * np.pi * Y[:, 1]) + Y[:, 0] + Y[:, 1]
    for dx in ((0.125, 0.0625, 0.03125) if p == 1 else (0.25, 0.125, 0.0625)):
        qt = oracle.Quadtree(2, graded(dx))
        pr = oracle.Problem(qt, p)
        X = pr.nodes(); xq, _ = pr.bc_qpoints()
        ff = 8 * np.pi ** 2 * np.cos(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1])
        rp, cols, blk = pr.export("A")
        A = sp.bsr_matrix((blk, cols, rp), shape=(pr.ndof, pr.ndof)).tocsc()
        u = spla.spsolve(A, pr.rhs(ff, uf(xq)))
        errs.append(np.abs(u - uf(X)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > p + 0.6, (errs, rates)
