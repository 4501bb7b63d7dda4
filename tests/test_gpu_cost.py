"""GPU: the cost model of the single-RHS MG-PCG (mgldg_mg_solve_cost), the
launch and byte counts the benchmark's whole-iteration roofline uses.

Its V-cycle walk mirrors the enqueue path, so its launch count must equal
the kernel nodes of the captured V-cycle graph, for every smoother, with and
without the collapsed coarse tail, pooled or not, with the row tiles on or
off (pooled GS levels below the finest: the restriction into the level also
does its first GS launch).  Bytes: positive, FP32 V-cycle below FP64, pooled below unpooled, the
iteration above its V-cycle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def problem(gctx, p=2):
    import mgldg
    qt = mgldg.Quadtree(gctx, 2, lambda x: 1 / 16 if x[0] > 0.5 else 1 / 32)
    return mgldg.LDGPoisson(qt, p)


@pytest.mark.parametrize("relax", ["gs", "gs-lex", "jacobi"])
@pytest.mark.parametrize("collapse", [0, 2048])
@pytest.mark.parametrize("tiles", ["off", "on"])
def test_vcycle_launches_equal_graph_nodes(gctx, relax, collapse, tiles):
    import mgldg
    import torch
    pr = problem(gctx)
    mg = mgldg.Multigrid(pr, relax, npre=2, npost=2, precision="mixed", coarse_dense_max=collapse, row_tiles=tiles)
    b = torch.tensor(pr.rhs(np.random.default_rng(1).standard_normal(pr.ndof)), device="cuda")
    for prec in ("fp64", "mixed"):
        before = mg.solve_cost(prec)
        assert before["graph_nodes"] == -1
        _, st = mgldg.pcg(mg, b, torch.zeros_like(b), tol=1e-8, maxit=100, precision=prec, flexible=prec == "mixed",
                          use_graph=True)
        assert st["status"] == "converged"
        c = mg.solve_cost(prec)
        assert c["vcycle_launches"] == c["graph_nodes"], (prec, c)
        assert c["iteration_launches"] == c["vcycle_launches"] + 4
        assert 0 < c["vcycle_bytes"] < c["iteration_bytes"]
    assert mg.solve_cost("mixed")["vcycle_bytes"] < mg.solve_cost("fp64")["vcycle_bytes"]


def test_pooled_bytes_below_unpooled(gctx):
    import mgldg
    pr = problem(gctx, 3)
    a = mgldg.Multigrid(pr, "gs", npre=2, npost=2, pool=True, coarse_dense_max=0).solve_cost()
    b = mgldg.Multigrid(pr, "gs", npre=2, npost=2, pool=False, coarse_dense_max=0).solve_cost()
    # on pooled levels the restriction also does the coarse level's first GS
    # launch: at most one launch fewer per level between the finest and the
    # coarsest
    nl = mgldg.Multigrid(pr, "gs", npre=2, npost=2, coarse_dense_max=0).nlevels
    assert b["vcycle_launches"] - (nl - 2) <= a["vcycle_launches"] < b["vcycle_launches"]
    assert a["vcycle_bytes"] < 0.5 * b["vcycle_bytes"]
