"""GPU: the row-tile kernels of pooled levels (rows with the same blocks, 16 at
a time on the matrix cores) against the CPU oracle, with the tiles forced
onto every pooled level (row_tiles="on"; by default only levels of 2048
tiles and more in FP64, 256 in FP32).

Bars as the other parity tests: SpMV, Gauss-Seidel sweeps (forward and
reverse, multicolour as the oracle's "mcgs") and the V-cycle within 1e-12
relative in FP64 and 1e-5 in FP32; MG-PCG iteration counts equal to the
oracle's in FP64 (within one in mixed precision), solutions to 1e-9.  The
uniform 32 x 32 mesh at p = 2 has vectors whose byte length is a multiple
of 4096 (1024 rows x 9 x 4 or 8 bytes), so a gather that read past its last
row would leave the allocation."""
import numpy as np
import pytest

from test_gpu_parity import build_pair, rel

pytestmark = pytest.mark.gpu

CASES = [("uniform32", 2), ("graded", 1), ("graded", 2), ("point", 3)]
EXTRA = {"uniform32": (2, lambda x: 1 / 32),
         # two sizes with a 2:1 interface at x = 1/2 (most D^-1 blocks repeat)
         "graded": (2, lambda x: 1 / 16 if x[0] > 0.5 else 1 / 32)}


def build(gctx, oracle, mesh, p, npre=2):
    import mgldg
    import test_gpu_parity as T
    for k, v in EXTRA.items():
        T.MESHES.setdefault(k, v)
    rapid = mesh == "point"
    qt, pr, oqt, opr = build_pair(gctx, oracle, mesh, p, rapid=rapid, balance=rapid)
    mg = mgldg.Multigrid(pr, "gs", npre=npre, npost=2, precision="mixed", coarse_dense_max=0, row_tiles="on",
                         restriction="ratrat")
    omg = oracle.Multigrid(opr, "mcgs", npre=npre, npost=2, restriction="ratrat")
    tiled = [l for l in range(mg.nlevels) if mg.level_tiles(l)[2]]
    assert 0 in tiled, [mg.level_tiles(l) for l in range(mg.nlevels)]
    return pr, opr, mg, omg, tiled


@pytest.mark.parametrize("mesh,p", CASES)
def test_tile_kernels_match_oracle(gctx, oracle, mesh, p):
    import torch
    pr, opr, mg, omg, tiled = build(gctx, oracle, mesh, p)
    rng = np.random.default_rng(5)
    dev = torch.device("cuda")
    for l in tiled:
        n = mg.level_rows[l] * pr.npl
        x, b = rng.standard_normal(n), rng.standard_normal(n)
        yo = omg.spmv(l, x)
        for prec, dt, tol in (("fp64", torch.float64, 1e-12), ("mixed", torch.float32, 1e-5)):
            xt = torch.tensor(x, dtype=dt, device=dev)
            y = torch.zeros_like(xt)
            mg.spmv(l, xt, y, precision=prec)
            assert rel(y.double().cpu().numpy(), yo) < tol, (l, prec)
            for reverse in (False, True):
                xs = torch.tensor(x, dtype=dt, device=dev)
                mg.smooth(l, torch.tensor(b, dtype=dt, device=dev), xs, 2, reverse, precision=prec)
                assert rel(xs.double().cpu().numpy(), omg.relax(l, b, x, 2, reverse)) < tol, (l, prec, reverse)


@pytest.mark.parametrize("npre", [2, 0])
@pytest.mark.parametrize("mesh,p", CASES)
def test_tile_vcycle_matches_oracle(gctx, oracle, mesh, p, npre):
    """from zero: the first pre-smoothing sweep's launches read only the
    colours already updated; with npre = 0 the post-smoothing starts from the
    prolonged correction alone"""
    import torch
    pr, opr, mg, omg, tiled = build(gctx, oracle, mesh, p, npre=npre)
    b = np.random.default_rng(9).standard_normal(pr.ndof)
    zo = omg.vcycle(b)
    z = torch.zeros(pr.ndof, dtype=torch.float64, device="cuda")
    mg.vcycle(torch.tensor(b, device="cuda"), z)
    assert rel(z.cpu().numpy(), zo) < 1e-12
    mg.vcycle(torch.tensor(b, device="cuda"), z, precision="mixed")
    assert rel(z.cpu().numpy(), zo) < 1e-5


@pytest.mark.parametrize("mesh,p", CASES)
def test_tile_pcg_matches_oracle(gctx, oracle, mesh, p):
    import mgldg
    pr, opr, mg, omg, tiled = build(gctx, oracle, mesh, p)
    f = np.random.default_rng(13).standard_normal(pr.ndof)
    xo, ito, _, _ = oracle.pcg(opr, omg, opr.rhs(f), tol=1e-10, maxit=200, criterion="true")
    b = pr.rhs(f)
    x, st = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion="true")
    assert st["status"] == "converged" and st["iterations"] == ito
    assert rel(x, xo) < 1e-9
    for graph in (False, True):
        xm, stm = mgldg.pcg(mg, b, tol=1e-10, maxit=200, criterion="true", precision="mixed", flexible=True,
                            use_graph=graph)
        assert stm["status"] == "converged" and abs(stm["iterations"] - ito) <= 1
        assert rel(xm, xo) < 1e-8


def test_tiles_off_auto_on(gctx):
    """row_tiles: "off" builds the tables but never uses them, "auto" uses them
    from 2048 tiles (32k rows) in FP64 and 256 tiles in FP32, "on" on every
    level that has them"""
    import mgldg
    qt = mgldg.Quadtree(gctx, 2, lambda x: 1 / 32)
    pr = mgldg.LDGPoisson(qt, 2)
    for mode, want in (("off", False), ("auto", False), ("on", True)):
        mg = mgldg.Multigrid(pr, "gs", npre=2, npost=2, row_tiles=mode, precision="mixed")
        t, w, u64, u32 = mg.level_tiles(0)
        assert 64 <= t < 256 and w == 4 and u64 == want and u32 == want
    qt = mgldg.Quadtree(gctx, 2, lambda x: 1 / 128)          # 16,384 rows: about 1,100 tiles
    mg = mgldg.Multigrid(mgldg.LDGPoisson(qt, 2), "gs", npre=2, npost=2, precision="mixed")
    t, w, u64, u32 = mg.level_tiles(0)
    assert 256 <= t < 2048 and not u64 and u32
    with pytest.raises(KeyError):
        mgldg.Multigrid(pr, "gs", row_tiles="always")
