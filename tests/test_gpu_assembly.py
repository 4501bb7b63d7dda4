"""GPU: the device assembly of LDGPoisson::discretize (ldgpoisson.tpp:14-176;
unstructured ldgpoisson.tpp:14-170) against the oracle's host assembly.

Edge cases of the gathered rows: hanging (non-conforming) quadtree faces,
whose face matrices the device forms by quadrature; a periodic grid two
elements wide, where two faces join the same pair of elements and their
terms accumulate into one block; 3-D quadtrees; p = 0; Dirichlet and Neumann
keys on simplices in 2-D and 3-D.  Bars: patterns bit-exact, blocks 1e-13
relative; the right-hand side with boundary data (the boundary lifts) 1e-13."""
import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MESHES = os.path.join(os.path.dirname(__file__), "golden", "meshes")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def check_ops(pr, opr, N):
    for w in ["M", "T", "A"] + [f"G{d}" for d in range(N)]:
        a, o = pr.export(w), opr.export(w)
        assert np.array_equal(a[0], o[0]) and np.array_equal(a[1], o[1]), w
        assert rel(a[2], o[2]) < 1e-13, w


def check_rhs(pr, opr):
    rng = np.random.default_rng(1)
    f = rng.standard_normal(pr.ndof)
    g = rng.standard_normal(pr.nqb)
    assert rel(pr.rhs(f, g), opr.rhs(f, g)) < 1e-13


QT_CASES = {
    "hanging-p2": (2, 2, lambda x: 1 / 8 if x[0] > 0.5 else 1 / 16, False, 0),
    "hanging-neumann-p3": (2, 3, lambda x: 1 / 4 if x[0] > 0.5 else 1 / 8, False, 1),
    "periodic-2x2": (2, 1, lambda x: 1 / 2, True, 1),
    "periodic-hanging": (2, 2, lambda x: 1 / 4 if x[0] > 0.5 else 1 / 8, True, 1),
    "3d-hanging-p1": (3, 1, lambda x: 1 / 4 if x[0] > 0.5 else 1 / 8, False, 0),
    "p0": (2, 0, lambda x: 1 / 8 if x[0] > 0.5 else 1 / 16, False, 0),
}


@pytest.mark.parametrize("case", list(QT_CASES))
def test_quadtree_assembly_matches_oracle(gctx, oracle, case):
    import mgldg
    N, p, h, periodic, bc = QT_CASES[case]
    qt = mgldg.Quadtree(gctx, N, h, periodic=periodic)
    pr = mgldg.LDGPoisson(qt, p, bc=[bc] * (2 * N), tau0=1.0 if periodic else 0.0)
    opr = oracle.Problem(oracle.Quadtree(N, h, periodic=periodic), p, bctype=[bc] * (2 * N), tau0=1.0 if periodic else 0.0)
    check_ops(pr, opr, N)
    if pr.nqb:
        check_rhs(pr, opr)


SX_CASES = {
    "circle-p2-dirichlet": ("circle.mesh", 2, 2, 0),
    "square8-p1-neumann": ("square_8.mesh", 2, 1, 1),
    "circle2-p3-dirichlet": ("circle2.mesh", 2, 3, 0),
    "ball-p2-dirichlet": ("ball.mesh", 3, 2, 0),
    "square4-p0": ("square_4.mesh", 2, 0, 0),
}


@pytest.mark.parametrize("case", list(SX_CASES))
def test_simplex_assembly_matches_oracle(gctx, oracle, case):
    import mgldg
    mesh, N, p, bc = SX_CASES[case]
    pts, tris = oracle.read_mesh(os.path.join(MESHES, mesh))
    pr = mgldg.SimplexLDGPoisson(mgldg.Wireframe(gctx, pts, tris), p, bc=[bc], tau0=100.0, tauD=1000.0)
    opr = oracle.SxProblem(pts, tris, p, bctype=[bc], tau0=100.0, tauD=1000.0)
    check_ops(pr, opr, N)
    check_rhs(pr, opr)


def test_setup_times_report_the_assembly_kernel(gctx):
    import mgldg
    qt = mgldg.Quadtree(gctx, 2, lambda x: 1 / 16)
    pr = mgldg.LDGPoisson(qt, 2)
    t = pr.setup_times()
    assert t["host_s"] > 0 and t["device_s"] > 0 and t["assembly_kernel_ms"] > 0
    blocks = sum(pr.export(w)[1].size for w in ["M", "T", "G0", "G1"])
    assert t["assembly_bytes"] >= blocks * pr.npl ** 2 * 8
