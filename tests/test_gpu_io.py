"""GPU: the interop dumps of device-resident objects through the C ABI.

LDGPoisson::dump (ldgpoisson.h:100-109) writes A, M, T, G_d as MatrixMarket
and the right-hand side as .fun; the library's operators read back from those
files equal its exports exactly and the oracle's to 1e-13.  Wireframe::write
reproduces the reference's mesh files byte for byte.  The V-cycle matrix
(Multigrid::vcycle_matrix, multigrid.h:386-400) runs its unit vectors through
the batched V-cycle; its columns equal the oracle's V-cycle of e_i to 1e-12
(FP64) and the single-vector V-cycle to 1e-13, and the FP32 form is within
1e-5 of the FP64 one."""
import glob
import os

import numpy as np
import pytest
import scipy.io

from conftest import sine_f

pytestmark = pytest.mark.gpu

GOLDEN_MESHES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "meshes", "*.mesh")))


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("kind", ["quadtree", "simplex"])
def test_problem_dump_matches_exports_and_oracle(gctx, oracle, tmp_path, kind):
    import mgldg
    if kind == "quadtree":
        h = lambda x: 1 / 8 if x[0] < 0.5 else 1 / 4
        pr = mgldg.LDGPoisson(mgldg.Quadtree(gctx, 2, h), 2)
        opr = oracle.Problem(oracle.Quadtree(2, h), 2)
    else:
        pts, tris = oracle.read_mesh(GOLDEN_MESHES[1])        # circle.mesh
        pr = mgldg.SimplexLDGPoisson(mgldg.Wireframe(gctx, pts, tris), 2)
        opr = oracle.SxProblem(pts, tris, 2)
    b = pr.rhs(sine_f(pr.nodes()))
    pr.dump(tmp_path, rhs=b)
    names = sorted(os.listdir(tmp_path))
    assert names == ["A.mtx", "G0.mtx", "G1.mtx", "M.mtx", "T.mtx", "rhs.fun"]
    for w in ["A", "M", "T", "G0", "G1"]:
        rp, cols, blk = pr.export(w)
        D = scipy.io.mmread(str(tmp_path / f"{w}.mtx")).toarray()
        assert np.array_equal(D, oracle.bsr_to_dense(rp, cols, blk, pr.ne, pr.ne)), w
        assert rel(D, opr.dense(w)) < 1e-13, w
    (P, N, ne), data = mgldg.read_fun(tmp_path / "rhs.fun")
    assert (P, N, ne) == (3, 2, pr.ne)
    assert np.array_equal(data[:, :, :2].reshape(-1, 2), pr.nodes())
    assert np.array_equal(data[:, :, 2].reshape(-1), b)
    # one operator alone, and a level operator / interpolation of the hierarchy
    pr.write_mm("T", tmp_path / "T2.mtx")
    assert (tmp_path / "T2.mtx").read_bytes() == (tmp_path / "T.mtx").read_bytes()


def test_mg_write_mm_level_operators(gctx, oracle, tmp_path):
    import mgldg
    qt = mgldg.Quadtree(gctx, 2, lambda x: 1 / 16)
    mg = mgldg.Multigrid(mgldg.LDGPoisson(qt, 1), "gs")
    for l, w in [(1, "A"), (0, "interp"), (2, "M")]:
        rp, cols, blk, nbr, nbc = mg.export(l, w)
        mg.write_mm(l, w, tmp_path / "x.mtx")
        D = scipy.io.mmread(str(tmp_path / "x.mtx")).toarray()
        assert np.array_equal(D, oracle.bsr_to_dense(rp, cols, blk, nbr, nbc)), (l, w)


@pytest.mark.parametrize("path", GOLDEN_MESHES, ids=os.path.basename)
def test_wireframe_write_reproduces_reference_files(gctx, tmp_path, path):
    import mgldg
    N = 3 if "ball" in path else 2
    wf = mgldg.Wireframe.read(gctx, path, N=N)
    wf.write(tmp_path / "m.mesh")
    assert (tmp_path / "m.mesh").read_bytes() == open(path, "rb").read()
    wf.write(tmp_path / "f.mesh", digits=17)
    p0, t0 = wf.arrays()
    p1, t1 = mgldg.Wireframe.read(gctx, tmp_path / "f.mesh", N=N).arrays()
    assert np.array_equal(p0, p1) and np.array_equal(t0, t1)


@pytest.mark.parametrize("relax,n,p", [("gs", 8, 1), ("jacobi", 8, 1), ("gs-lex", 4, 2), ("gs", 4, 3)])
def test_vcycle_matrix_matches_oracle(gctx, oracle, tmp_path, relax, n, p):
    import mgldg
    import torch
    qt = mgldg.Quadtree(gctx, 2, lambda x: 1 / n)
    pr = mgldg.LDGPoisson(qt, p)
    mg = mgldg.Multigrid(pr, relax, npre=2, npost=2, precision="mixed")
    omg = oracle.Multigrid(oracle.Problem(oracle.Quadtree(2, lambda x: 1 / n), p),
                           {"gs": "mcgs", "gs-lex": "gs", "jacobi": "jacobi"}[relax], npre=2, npost=2)
    V = mg.vcycle_matrix()
    nd = pr.ndof
    assert V.shape == (nd, nd)        # 144 dofs (p = 2) end in a partial chunk of 16 columns
    rng = np.random.default_rng(0)
    for i in sorted(set(rng.integers(0, nd, 6).tolist()) | {0, nd - 1}):
        e = np.zeros(nd)
        e[i] = 1.0
        assert rel(V[:, i], omg.vcycle(e)) < 1e-12, i
        ed = torch.tensor(e, device="cuda")
        xd = torch.zeros_like(ed)
        mg.vcycle(ed, xd)
        assert rel(V[:, i], xd.cpu().numpy()) < 1e-13, i
    V32 = mg.vcycle_matrix("mixed")
    assert rel(V32, V) < 1e-5
    # the driver's V.dat, read back exactly
    mgldg.write_dense(tmp_path / "V.dat", V)
    assert np.array_equal(np.loadtxt(tmp_path / "V.dat"), V)
