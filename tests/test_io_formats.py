"""Interop file formats written by libmgldg's host writers (no GPU):
MatrixMarket dumps (SparseBlockMatrix::write, sparseblockmatrix.tpp:145-164),
.fun (Function::write, src/cartesian/dg/function.h:173-190), .mesh
(lib/distmesh/writemesh.m:17-28, read by wireframe.h:22-68) and the dense
V-cycle matrix dump (dg_tests.cpp:337-347).

Pinned by: the reference's own mesh files (tests/golden/meshes, written by
writemesh.m) reproduced byte for byte; literal expected text for tiny known
cases; and independent readers (scipy.io.mmread, numpy.loadtxt, the
reference's .fun import rule, DGTools.m:71-83) recovering the oracle's
operators and functions exactly."""
import glob
import os

import numpy as np
import pytest
import scipy.io

import mgldg

GOLDEN_MESHES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "meshes", "*.mesh")))


def text(path):
    with open(path) as f:
        return f.read()


# ---------------------------------------------------------------- MatrixMarket
def test_mm_known_answer(tmp_path):
    # expected values are the strings a C++ stream prints under precision(17)
    # 2 x 3 block rows/cols of 1 x 2 blocks: row 0 holds blocks at columns 0
    # and 2, row 1 one block at column 1
    rp = np.array([0, 2, 3])
    cols = np.array([0, 2, 1])
    blk = np.array([[[0.1, -2.0]], [[0.0, 1e-300]], [[3.0, 2.0 / 3.0]]])
    f = tmp_path / "a.mtx"
    mgldg.write_mm(f, rp, cols, blk, nbcols=3)
    assert text(f) == ("%%MatrixMarket matrix coordinate real general\n"
                       "2 6 6\n"
                       "1 1 0.10000000000000001\n"
                       "1 2 -2\n"
                       "1 5 0\n"
                       "1 6 1e-300\n"
                       "2 3 3\n"
                       "2 4 0.66666666666666663\n")


@pytest.mark.parametrize("which", ["A", "M", "T", "G0", "G1"])
def test_mm_of_oracle_operators_reads_back_exactly(tmp_path, oracle, which):
    qt = oracle.Quadtree(2, lambda x: 1 / 8 if x[0] < 0.5 else 1 / 4)
    pr = oracle.Problem(qt, 2)
    rp, cols, blk = pr.export(which)
    f = tmp_path / f"{which}.mtx"
    mgldg.write_mm(f, rp, cols, blk)
    lines = text(f).split("\n")
    assert lines[0] == "%%MatrixMarket matrix coordinate real general"
    n = pr.ne * pr.b
    assert lines[1] == f"{n} {n} {len(cols) * pr.b * pr.b}"
    assert len(lines) - 3 == len(cols) * pr.b * pr.b      # explicit zeros are entries too
    A = scipy.io.mmread(str(f))
    assert A.shape == (n, n)
    assert np.array_equal(A.toarray(), oracle.bsr_to_dense(rp, cols, blk, pr.ne, pr.ne))


def test_mm_rectangular_interpolation_shape(tmp_path, oracle):
    qt = oracle.Quadtree(2, lambda x: 1 / 8)
    omg = oracle.Multigrid(oracle.Problem(qt, 1), "mcgs")
    rp, cols, blk, nbr, nbc = omg.export(0, "interp")
    f = tmp_path / "T0.mtx"
    mgldg.write_mm(f, rp, cols, blk, nbcols=nbc)
    A = scipy.io.mmread(str(f)).toarray()
    assert A.shape == (nbr * blk.shape[1], nbc * blk.shape[2])
    assert np.array_equal(A, oracle.bsr_to_dense(rp, cols, blk, nbr, nbc))


# ---------------------------------------------------------------- .fun
def test_fun_known_answer(tmp_path):
    xyz = np.array([[0.0, 0.5], [1.0, 0.5], [0.25, 1.0 / 3.0], [2.0, 2.0]])
    u = np.array([1.0, -0.1, 0.0, 1e20])
    f = tmp_path / "u.fun"
    mgldg.write_fun(f, 2, xyz, u, npl=2)
    assert text(f) == ("2 2 2\n"
                       "0 0.5 1\n"
                       "1 0.5 -0.10000000000000001\n"
                       "\n"
                       "0.25 0.33333333333333331 0\n"
                       "2 2 1e+20\n"
                       "\n")


@pytest.mark.parametrize("p", [0, 1, 3])
def test_fun_of_oracle_function_reads_back_exactly(tmp_path, oracle, p):
    qt = oracle.Quadtree(2, lambda x: 1 / 4 if x[1] < 0.5 else 1 / 2)
    pr = oracle.Problem(qt, p)
    X = pr.nodes()
    u = np.sin(3 * X[:, 0]) * np.exp(X[:, 1])
    f = tmp_path / "u.fun"
    mgldg.write_fun(f, p + 1, X, u, npl=pr.b)
    (P, N, ne), data = mgldg.read_fun(f)
    assert (P, N, ne) == (p + 1, 2, pr.ne)
    assert data.shape == (pr.ne, pr.b, 3)
    assert np.array_equal(data[:, :, :2].reshape(-1, 2), X)
    assert np.array_equal(data[:, :, 2].reshape(-1), u)
    # one empty line closes every element (the importer splits on them)
    assert text(f).count("\n\n") == pr.ne


def test_fun_simplex_nodes(tmp_path, oracle):
    pts, tris = oracle.read_mesh(GOLDEN_MESHES[1])          # circle.mesh
    pr = oracle.SxProblem(pts, tris, 2)
    X = pr.nodes()
    f = tmp_path / "s.fun"
    mgldg.write_fun(f, 3, X, X[:, 0] - X[:, 1], npl=pr.b)
    (P, N, ne), data = mgldg.read_fun(f)
    assert (P, N, ne) == (3, 2, len(tris)) and data.shape[1] == 6
    assert np.array_equal(data[:, :, 2].reshape(-1), X[:, 0] - X[:, 1])


# ---------------------------------------------------------------- .mesh
@pytest.mark.parametrize("path", GOLDEN_MESHES, ids=os.path.basename)
def test_mesh_writer_reproduces_reference_files(tmp_path, oracle, path):
    pts, tris = oracle.read_mesh(path)
    f = tmp_path / "m.mesh"
    mgldg.write_mesh(f, pts, tris)
    with open(path, "rb") as a, open(f, "rb") as b:
        assert a.read() == b.read()


def test_mesh_full_precision_round_trip(tmp_path, oracle):
    rng = np.random.default_rng(4)
    pts = rng.standard_normal((40, 3)) * 10.0 ** rng.integers(-8, 8, (40, 1))
    tets = rng.integers(0, 40, (25, 4))
    f = tmp_path / "r.mesh"
    mgldg.write_mesh(f, pts, tets, digits=17)
    p2, t2 = oracle.read_mesh(f)
    assert np.array_equal(p2, pts) and np.array_equal(t2, tets)


# ---------------------------------------------------------------- dense (V.dat)
def test_dense_known_answer(tmp_path):
    f = tmp_path / "V.dat"
    mgldg.write_dense(f, np.array([[1.0, -0.5], [1e-20, 2.0]]))
    # the strings a C++ stream prints under precision(17)
    w = len("9.9999999999999995e-21")
    assert text(f) == ("1".rjust(w) + " " + "-0.5".rjust(w) + "\n"
                       + "9.9999999999999995e-21" + " " + "2".rjust(w) + "\n")


def test_dense_round_trip(tmp_path):
    A = np.random.default_rng(7).standard_normal((9, 5)) ** 3
    f = tmp_path / "V.dat"
    mgldg.write_dense(f, A)
    assert np.array_equal(np.loadtxt(f), A)
    widths = {len(t) for ln in text(f).splitlines() for t in ln.split(" ") if t}
    lines = text(f).splitlines()
    assert len({len(ln) for ln in lines}) == 1 and len(lines[0]) == 5 * max(widths) + 4


# ---------------------------------------------------------------- errors
def test_writer_errors(tmp_path):
    with pytest.raises(mgldg.MGLDGError) as e:
        mgldg.write_dense(tmp_path / "missing" / "V.dat", np.eye(2))
    assert e.value.code == -6
    with pytest.raises(mgldg.MGLDGError) as e:
        mgldg.write_mesh(tmp_path / "bad.mesh", np.zeros((3, 2)), np.array([[0, 1, 3]]))
    assert e.value.code == -1
    with pytest.raises(mgldg.MGLDGError):
        mgldg.write_fun(tmp_path / "bad.fun", 2, np.zeros((5, 2)), np.zeros(5), npl=4)
