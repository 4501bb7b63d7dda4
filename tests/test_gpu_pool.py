"""GPU: pooled operators (repeated quadtree blocks read from a pool of the
distinct blocks by the single-RHS kernels) against the full block arrays.

A pool holds the same bits as the blocks it replaces.  Where a block row is
a whole number of 16-byte vectors (B = 4 and 16: p = 1, 3) the kernels do
the same arithmetic in the same order, so every result is compared for
bitwise equality: SpMV on every level, smoother sweeps (multicolour and
lexicographic Gauss-Seidel, Jacobi), the V-cycle and MG-PCG in FP64 and
mixed precision.  For B = 9 (p = 2) the pooled kernels read the block rows
padded to whole vectors and sum them in vector-wide partial sums, a
different grouping of the same products: results agree to rounding (1e-13
relative in FP64, 1e-5 in FP32) and MG-PCG takes the same iterations.
Unstructured (simplex) operators repeat nothing and are not pooled."""
import os

import numpy as np
import pytest

from conftest import sine_f

pytestmark = pytest.mark.gpu


def h_adaptive(x):
    return 1 / 16 if x[0] > 0.5 else 1 / 32


def pair(gctx, relax, p, precision="mixed"):
    import mgldg
    qt = mgldg.Quadtree(gctx, 2, h_adaptive)
    pr = mgldg.LDGPoisson(qt, p)
    kw = dict(npre=2, npost=2, precision=precision)
    return pr, mgldg.Multigrid(pr, relax, pool=True, **kw), mgldg.Multigrid(pr, relax, pool=False, **kw)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_quadtree_levels_are_pooled(gctx, p):
    pr, mg, mg0 = pair(gctx, "gs", p)
    n0, d0 = mg.level_pool(0)
    assert 0 < n0 <= mg.level_nnzb[0] // 4 and 0 < d0 <= mg.level_rows[0]
    assert mg0.level_pool(0) == (0, 0)


def same(a, b, p, prec):
    """bitwise for p = 1, 3; to rounding for p = 2 (padded 9-wide rows)"""
    import torch
    if p != 2:
        return torch.equal(a, b)
    tol = 1e-13 if prec == "fp64" else 1e-5
    return float((a - b).abs().max()) <= tol * float(b.abs().max())


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_pooled_spmv_and_smoothers_bitwise(gctx, p, prec):
    import torch
    pr, mg, mg0 = pair(gctx, "gs", p)
    dt = torch.float64 if prec == "fp64" else torch.float32
    g = torch.Generator(device="cuda").manual_seed(3)
    for l in range(mg.nlevels - 1):
        n = mg.level_rows[l] * mg.level_B[l]
        x = torch.randn(n, device="cuda", dtype=dt, generator=g)
        y1, y0 = torch.empty_like(x), torch.empty_like(x)
        mg.spmv(l, x, y1, prec)
        mg0.spmv(l, x, y0, prec)
        assert same(y1, y0, p, prec), l
        b = torch.randn(n, device="cuda", dtype=dt, generator=g)
        x1, x0 = x.clone(), x.clone()
        mg.smooth(l, b, x1, nsweeps=2, precision=prec)
        mg0.smooth(l, b, x0, nsweeps=2, precision=prec)
        assert same(x1, x0, p, prec), l
        mg.smooth(l, b, x1, nsweeps=1, reverse=True, precision=prec)
        mg0.smooth(l, b, x0, nsweeps=1, reverse=True, precision=prec)
        assert same(x1, x0, p, prec), l


@pytest.mark.parametrize("p", [2, 3])
@pytest.mark.parametrize("relax", ["gs", "gs-lex", "jacobi"])
def test_pooled_vcycle_and_pcg_bitwise(gctx, relax, p):
    import mgldg
    import torch
    pr, mg, mg0 = pair(gctx, relax, p)
    b = torch.tensor(pr.rhs(sine_f(pr.nodes())), device="cuda")
    for prec in ["fp64", "mixed"]:
        z1, z0 = torch.zeros_like(b), torch.zeros_like(b)
        mg.vcycle(b, z1, prec)
        mg0.vcycle(b, z0, prec)
        assert same(z1, z0, p, "fp64" if prec == "fp64" else "mixed"), prec
        kw = dict(tol=1e-10, maxit=200, criterion="true", precision=prec, flexible=prec == "mixed")
        x1, s1 = mgldg.pcg(mg, b, torch.zeros_like(b), **kw)
        x0, s0 = mgldg.pcg(mg0, b, torch.zeros_like(b), **kw)
        assert s1["iterations"] == s0["iterations"] and s1["status"] == "converged"
        if p == 2:
            assert float((x1 - x0).abs().max()) <= 1e-9 * float(x0.abs().max()), prec
        else:
            assert torch.equal(x1, x0), prec


def test_simplex_levels_are_not_pooled(gctx, oracle):
    import mgldg
    pts, tris = oracle.read_mesh(os.path.join(os.path.dirname(__file__), "golden", "meshes", "circle2.mesh"))
    pr = mgldg.SimplexLDGPoisson(mgldg.Wireframe(gctx, pts, tris), 2, tau0=100.0)
    mg = mgldg.Multigrid(pr, "jacobi", omega=0.6, restriction="rat")
    assert mg.level_pool(0) == (0, 0)


@pytest.mark.parametrize("relax", ["gs", "jacobi"])
def test_pooled_batched_mfma_bitwise(gctx, relax):
    """the batched MFMA kernels read the pool too (same tile family forced on both)"""
    import mgldg
    import torch
    pr, mg, mg0 = pair(gctx, relax, 3)
    rng = np.random.default_rng(4)
    B = torch.tensor(pr.rhs(rng.standard_normal((pr.ndof, 32))), device="cuda")
    for m in (mg, mg0):
        m.batch_kernels(1)                  # one 16-column MFMA tile per wavefront, every level
    for prec in ["fp64", "mixed"]:
        kw = dict(tol=1e-10, maxit=200, criterion="true", precision=prec, flexible=prec == "mixed")
        X1, s1 = mgldg.pcg(mg, B, torch.zeros_like(B), **kw)
        X0, s0 = mgldg.pcg(mg0, B, torch.zeros_like(B), **kw)
        assert [s["iterations"] for s in s1] == [s["iterations"] for s in s0]
        assert torch.equal(X1, X0), prec
