"""Benchmark: DoF-solves/s of MG-PCG to a 1e-10 relative residual on one MI355X.

Workload (default): adaptive quadtree on [0,1]^2 refined around (0.3, 0.7),
p = 3 (16-dof blocks, 1,039,408 dofs), homogeneous Dirichlet data,
f ~ N(0,1) per dof (seed 0), V(2,2) multicolour block Gauss-Seidel, RATRAT
Galerkin hierarchy with rapid (height) coarsening, MG-preconditioned CG until
the true residual ||b - A x|| / ||b|| <= 1e-10.

A second workload is reported in the same line ("unstructured"): the unit
disk, Delaunay triangulation of 200,000 uniform random interior points
(radius < 1 - h/2, seed 0) and 1,571 boundary points at spacing h ~ 0.004,
p = 2 (401,569 triangles, 2,409,414 dofs), tau0 = 100, tauD = 1000 (the
reference's unstructured driver), the library's agglomerated hierarchy (RAT),
V(2,2) block Jacobi (omega 0.6), f ~ N(0,1) per dof; single solves in FP64
and mixed precision and a batch of 32 right-hand sides solved together.

One step = one full solve from x = 0 with b resident in HBM.  `value` is the
mixed-precision form (FP32 V-cycle inside FP64 CG); the plain FP64 GPU solve
is timed beside it.  The roofline object describes the dominant kernel, the
fine-level one-colour block Gauss-Seidel launch of the FP32 V-cycle, timed
with HIP events on the library's stream, and beside it the whole CG iteration
(the library's cost model of its algorithmic bytes over the measured time per
iteration, with the launches per V-cycle).  The CPU baseline is the FP64 oracle (a C
restatement of the reference algorithm, one thread per solve as the
reference) solving the same problem on the host: one core, and as many
concurrent solves as the process has cores (all-core throughput); beside it
the plain method's 32 x 32 p = 1 configuration on one and all cores and, for
the disk, adjoint and ball workloads (too large for the time budget), the
oracle's cost per iteration on stated smaller instances.  The ball workload
is measured on a Kuhn split of the jittered lattice; the Delaunay mesh the
configuration names is run too and its status recorded (its slivers stall
MG-PCG).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "multigrid-ldg_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
CALIB_BYTES = 256 * 1024 * 1024   # plain copy used to scale the FETCH_SIZE / WRITE_SIZE counters
X0 = (0.3, 0.7)
C2_SCALE = 1.8          # refine while |x - x0| < 1.8 * 2^-l (l = depth below the 16x16 root grid)


def c2_size(x):
    d = float(np.hypot(x[0] - X0[0], x[1] - X0[1]))
    return max(2.0 ** -14, min(d / (16 * C2_SCALE), 1.0 / 16))


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


MFMA_PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}   # MI355X dense matrix-core peaks


def mfma_use(mg, plan, k):
    """Utilisation of the matrix cores by the fine-level batched operator
    application (the residual form y = b - A x on k columns, as the autotuner
    times it), from the autotuner's HIP-event timing of the chosen kernel:
    useful FLOPs nnzb * 2 B^2 k (the 16 x 16 padding of narrower blocks not
    counted) over time, against the dense MFMA peak; and the algorithmic
    bytes (A and the column indices once; x, b and y once) over the same
    time."""
    out = {}
    nnzb, B, nb = mg.level_nnzb[0], mg.level_B[0], mg.level_rows[0]
    for prec, size in (("fp64", 8), ("fp32", 4)):
        lv = plan[prec][0]
        if lv["apply"] == 0:
            out[prec] = {"kernel": "vector (MFMA slower here)"}
            continue
        ms = lv["apply_ms"][{1: 1, 2: 2, 4: 3}[lv["apply"]]]
        flops = nnzb * 2.0 * B * B * k
        byts = nnzb * B * B * size + 4.0 * nnzb + 3.0 * nb * B * k * size
        tf = flops / (ms * 1e-3) / 1e12
        out[prec] = {"kernel": f"mf_apply<{'float' if size == 4 else 'double'},{B},{lv['apply']}>", "ms": round(ms, 4),
                     "tflops": round(tf, 2), "peak_tflops": MFMA_PEAK_TFLOPS[prec],
                     "frac": round(tf / MFMA_PEAK_TFLOPS[prec], 4), "gbs": round(byts / (ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    return out


def plan_summary(plan):
    """the autotuner's per-level choices: 0 = vector kernels, n = MFMA tiles of
    16 x n columns, with the fine level's timings (ms: vector, 1, 2, 4 tiles)"""
    out = {}
    for prec in ("fp64", "fp32"):
        lv = plan[prec]
        out[prec] = {"apply": [q["apply"] for q in lv], "smooth": [q["smooth"] for q in lv],
                     "fine_apply_ms": [round(v, 4) for v in lv[0]["apply_ms"]],
                     "fine_smooth_ms": [round(v, 4) for v in lv[0]["smooth_ms"]]}
    return out


def disk_mesh(n_int=200_000, h=0.004, seed=0):
    """Unit disk: boundary points at spacing ~h on r = 1 and n_int uniform
    random points in r < 1 - h/2, Delaunay-triangulated (scipy)."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    nb = int(round(2 * np.pi / h))
    t = 2 * np.pi * np.arange(nb) / nb
    bnd = np.stack([np.cos(t), np.sin(t)], 1)
    r = (1 - h / 2) * np.sqrt(rng.random(n_int))
    a = 2 * np.pi * rng.random(n_int)
    pts = np.concatenate([bnd, np.stack([r * np.cos(a), r * np.sin(a)], 1)])
    return pts, Delaunay(pts).simplices.astype(np.int32)


def unstructured_leg(mgldg, torch, nrhs, warm):
    """The unit-disk workload: setup times, single FP64 / mixed solves, and
    nrhs right-hand sides batched in FP64 and in mixed precision."""
    t0 = time.perf_counter()
    pts, tri = disk_mesh()
    tri = tri[mgldg.locality_order(pts, tri)]
    t_mesh = time.perf_counter() - t0
    ctx = mgldg.Context(0)
    t0 = time.perf_counter()
    wf = mgldg.Wireframe(ctx, pts, tri)
    prob = mgldg.SimplexLDGPoisson(wf, 2, tau0=100.0, tauD=1000.0)
    t_prob = time.perf_counter() - t0
    t0 = time.perf_counter()
    parts = mgldg.agglomerate(prob)
    t_agg = time.perf_counter() - t0
    t0 = time.perf_counter()
    mg = mgldg.Multigrid(prob, "jacobi", omega=0.6, npre=2, npost=2, restriction="rat", precision="mixed",
                         partitions=parts)
    t_mg = time.perf_counter() - t0
    log(f"unstructured: {prob.ne} triangles, {prob.ndof} dofs, levels {mg.level_rows}; setup mesh {t_mesh:.2f}s "
        f"problem {t_prob:.2f}s agglomeration {t_agg:.2f}s hierarchy {t_mg:.2f}s")
    rng = np.random.default_rng(0)
    F = torch.tensor(rng.standard_normal((prob.ndof, nrhs)), device="cuda")
    B = mgldg.rhs_device(mg, F)
    b = B[:, 0].contiguous()
    out = {}
    xs = {}
    for form, rhs in (("single", b), ("batched", B)):
        for prec in ("fp64", "mixed"):
            kw = dict(tol=1e-10, maxit=1000, criterion="true", precision=prec, flexible=(prec == "mixed"),
                      use_graph=True)
            for _ in range(warm):
                mgldg.pcg(mg, rhs, torch.zeros_like(rhs), **kw)
            torch.cuda.synchronize()
            t = time.perf_counter()
            x = torch.zeros_like(rhs)
            _, st = mgldg.pcg(mg, rhs, x, **kw)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            sts = st if isinstance(st, list) else [st]
            for q in sts:
                if q["status"] != "converged" or q["relres_true"] > 1e-10:
                    log("UNSTRUCTURED NOT CONVERGED", form, prec, q)
                    sys.exit(1)
            k = len(sts)
            out[f"{form}_{prec}"] = {"value": round(prob.ndof * k / dt, 1), "ms_per_step": round(dt * 1e3, 3),
                                     "iterations": max(q["iterations"] for q in sts)}
            xs[(form, prec)] = x
            log(f"unstructured {form} {prec}: {dt * 1e3:.1f} ms, {prob.ndof * k / dt:.3g} DoF-solves/s, "
                f"iterations {out[f'{form}_{prec}']['iterations']}")
    plan = plan_summary(mg.batch_plan())
    log(f"unstructured batched kernel plan: {plan}")
    # the single solve's fine-level launches against the HBM roofline (HIP
    # events, 50 launches each; bytes: every block of the operator once, the
    # D^-1 blocks, indices and the vector streams)
    roof = {}
    for name, fn in (("k_jacobi<double,6>", lambda: mg.profile_relax(0, "fp64", 50, part="sweep")[:2]),
                     ("k_jacobi<float,6>", lambda: mg.profile_relax(0, "mixed", 50, part="sweep")[:2]),
                     ("k_apply<double,6>", lambda: mg.profile_spmv(0, "fp64", 50))):
        ms, by = fn()
        roof[name] = {"us_per_launch": round(ms * 1e3, 2), "bytes_per_launch": int(by),
                      "achieved": round(by / (ms * 1e-3) / 1e9, 1), "peak": 8000.0,
                      "frac": round(by / (ms * 1e-3) / 1e9 / 8000.0, 4)}
    log(f"unstructured fine-level launches: {roof}")
    x64, x32 = xs[("batched", "fp64")], xs[("batched", "mixed")]
    diff = float(torch.linalg.norm(x32 - x64) / torch.linalg.norm(x64))
    return {
        "workload": "unit disk, Delaunay of 200k random interior + 1571 boundary points, p=2, tau0=100, tauD=1000",
        "elements": prob.ne, "ndof": prob.ndof, "levels": mg.nlevels, "level_rows": mg.level_rows,
        "smoother": "V(2,2) block Jacobi, omega 0.6", "coarsening": "own agglomeration (<= 4 per level), RAT",
        "nrhs": nrhs, **out, "kernel_plan": plan, "fine_level_launches": roof,
        "mixed_batched_vs_fp64_single": round(out["batched_mixed"]["value"] / out["single_fp64"]["value"], 3),
        "mixed_vs_fp64_batched_rel_diff": diff,
        "setup_s": {"mesh": round(t_mesh, 3), "problem": round(t_prob, 3), "agglomeration": round(t_agg, 3),
                    "hierarchy": round(t_mg, 3), "assembly": assembly_info(prob)},
    }


def c1_gpu_leg(mgldg, torch, warm=3, steps=20):
    """The plain method's configuration (uniform 32 x 32, p = 1, tau0 = 0,
    tauD = 1000, f = 2 pi^2 sin(pi x) sin(pi y), g = 0, V(2,2) block GS, RATRAT,
    MG-PCG to 1e-10) as the FP64 GPU solve with the V-cycle replayed from its
    hipGraph: latency-bound (4,096 dofs, 6 levels), so reported as time per
    solve.  Gauss-Seidel in the reference's lexicographic order (wavefronts)
    and multicolour."""
    ctx = mgldg.Context(0)
    qt = mgldg.Quadtree(ctx, 2, lambda x: 1.0 / 32)
    prob = mgldg.LDGPoisson(qt, 1, tau0=0.0, tauD=1000.0)
    X = prob.nodes()
    b = torch.tensor(prob.rhs(2 * np.pi ** 2 * np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1])), device="cuda")
    out = {"ndof": prob.ndof}
    for relax in ("gs-lex", "gs"):
        mg = mgldg.Multigrid(prob, relax, npre=2, npost=2)
        kw = dict(tol=1e-10, maxit=200, criterion="true", precision="fp64", use_graph=True)
        for _ in range(warm):
            mgldg.pcg(mg, b, torch.zeros_like(b), **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            _, st = mgldg.pcg(mg, b, torch.zeros_like(b), **kw)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        out[relax] = {"ms_per_solve": round(dt * 1e3, 3), "value": round(prob.ndof / dt, 1),
                      "iterations": st["iterations"], "relres_true": st["relres_true"]}
    log(f"C1 (32x32 p=1) FP64 GPU solves: {out}")
    return out


# ---------------------------------------------------------------- CPU baseline (FP64 oracle)
def _oracle_job(job):
    """One FP64 oracle solve in a worker process (spawned, no GPU): builds the
    job's problem and hierarchy (not timed), then times PCG to 1e-10 or for
    job['maxit'] iterations.  Returns (setup_s, solve_s, iterations, relres,
    ndof)."""
    import oracle as O
    kind = job["kind"]
    t0 = time.perf_counter()
    if kind == "c1":
        qt = O.Quadtree(2, lambda x: 1.0 / 32)
        pr = O.Problem(qt, 1, tau0=0.0, tauD=1000.0)
        mg = O.Multigrid(pr, job["relax"], npre=2, npost=2)
        X = pr.nodes()
        b = pr.rhs(2 * np.pi ** 2 * np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1]), np.zeros(pr.nqb))
    elif kind in ("c2", "c4"):
        if kind == "c2":
            qt = O.Quadtree(2, c2_size, rapid=True, balance=True)
            pr = O.Problem(qt, 3)
        else:
            qt = O.Quadtree(2, lambda x: 1.0 / job["n"])
            pr = O.Problem(qt, 2, tau0=0.0, tauD=1000.0)
        mg = O.Multigrid(pr, "mcgs", npre=2, npost=2)
        b = pr.rhs(np.random.default_rng(0).standard_normal(pr.ndof), np.zeros(pr.nqb))
    else:   # c3 / c5: agglomerated hierarchies from the reference's MGridGen (oracle/_ref)
        if kind == "c3":
            pts, cells = disk_mesh(job["n_int"], job["h"])
        else:
            pts, cells = ball_mesh(job["n"])
        pr = O.SxProblem(pts, cells, 2, tau0=100.0, tauD=1000.0)
        parts = O.agglomerate(pr)
        mg = O.SxMultigrid(pr, parts, relaxation="jacobi" if kind == "c3" else "mcgs",
                           omega=0.6 if kind == "c3" else 0.8, npre=2, npost=2, restriction="rat")
        b = pr.rhs(np.random.default_rng(0).standard_normal(pr.ndof), np.zeros(pr.nqb))
    t_setup = time.perf_counter() - t0
    if job.get("barrier") is not None:
        job["barrier"].wait(timeout=600)   # a sibling job that failed breaks the barrier instead of hanging it
    t0 = time.perf_counter()
    _, it, rr, _ = O.pcg(pr, mg, b, tol=1e-10, maxit=job.get("maxit", 500), criterion="true")
    return t_setup, time.perf_counter() - t0, it, rr, pr.ndof


def _cpu_pool_run(jobs, concurrent):
    """run jobs in spawned processes; concurrent ones start their timed solves together"""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    if not concurrent or len(jobs) == 1:
        with ctx.Pool(1) as pool:
            return [pool.apply(_oracle_job, (j,)) for j in jobs]
    mgr = ctx.Manager()
    bar = mgr.Barrier(len(jobs))
    with ctx.Pool(len(jobs)) as pool:
        return pool.map(_oracle_job, [dict(j, barrier=bar) for j in jobs])


def cpu_cores():
    """host threads the CPU baseline uses: this process's CPU share (the GPU
    box grants 16 per GPU)"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_legs(gpu_iters, c2_one_core=None):
    """The reference's CPU path (the FP64 oracle, a C restatement, one thread
    per solve as the reference: Makefile THREADS=1, no OpenMP in src/) on this
    host.  C1 and C2: full solves on 1 core and `cores` concurrent solves on
    `cores` cores (all-core throughput).  C3 / C4 / C5 do not fit the time
    budget at full size (oracle setup 21 s / 56 s / minutes and 2.4 s / 9 s /
    ~20 s per PCG iteration on one core): each is timed on a stated smaller
    instance of its configuration for 3 PCG iterations, and its figure is the
    DoF-solves/s the oracle would reach on the full problem with a cost per
    iteration linear in the dofs and the full problem's iteration count
    (from the GPU solve): ndof_sub / (seconds per iteration on the subset *
    iterations)."""
    cores = cpu_cores()
    out = {"host_cpu": cpu_model(), "host_nproc": os.cpu_count(), "cores_used": cores}
    for name, job, what in (("c1", {"kind": "c1", "relax": "gs"},
                             "uniform 32x32 p=1, V(2,2) lexicographic block GS (the reference's order)"),
                            ("c2", {"kind": "c2"}, "the adaptive p=3 workload, V(2,2) multicolour block GS")):
        if name == "c2" and c2_one_core is not None:
            t_setup, t1, it, rr, ndof = c2_one_core      # measured in the parent (its solution checks the GPU's)
        else:
            t_setup, t1, it, rr, ndof = _cpu_pool_run([job], False)[0]
        res = _cpu_pool_run([job] * cores, True)
        tmax = max(r[1] for r in res)
        out[name] = {"what": what, "ndof": ndof, "iterations": it, "relres_true": rr, "oracle_setup_s": round(t_setup, 2),
                     "one_core": {"solve_s": round(t1, 4), "value": round(ndof / t1, 1)},
                     "all_cores": {"cores": cores, "solves": cores, "wall_s": round(tmax, 4),
                                   "value": round(cores * ndof / tmax, 1)}}
        log(f"cpu {name}: {out[name]}")
    subs = [("c3", {"kind": "c3", "n_int": 25_000, "h": 0.0113, "maxit": 3},
             "disk: 25k random interior points (1/8 of the workload), p=2, MGridGen hierarchy, block Jacobi"),
            ("c4", {"kind": "c4", "n": 256, "maxit": 3}, "uniform 256x256 quadtree, p=2 (1/16 of the workload)"),
            ("c5", {"kind": "c5", "n": 17, "maxit": 3},
             "ball: Kuhn split of the jittered lattice of spacing 1/17 (1/15 of the workload), p=2, MGridGen "
             "hierarchy, multicolour GS")]
    import oracle as O
    if not O.mgridgen_available():      # the agglomerated hierarchies need the reference's MGridGen (oracle/_ref)
        for name, _, what in subs:
            if name in ("c3", "c5"):
                out[name] = {"subset": what, "skipped": "oracle/_ref/libmgridgen.so not built"}
                log(f"cpu {name}: skipped (MGridGen not built)")
        subs = [q for q in subs if q[0] == "c4"]
    log(f"cpu {', '.join(q[0] for q in subs)}: {len(subs)} subset solves started")
    res = _cpu_pool_run([j for _, j, _ in subs], True)
    for (name, job, what), (t_setup, ts, it, rr, ndof) in zip(subs, res):
        per_it = ts / max(it, 1)
        iters = gpu_iters.get(name)
        o = {"subset": what, "ndof_subset": ndof, "oracle_setup_s": round(t_setup, 2), "iterations_timed": it,
             "s_per_iteration": round(per_it, 4), "dof_iterations_per_s": round(ndof / per_it, 1),
             "fits_budget_at_full_size": False}
        if iters:
            o.update({"full_iterations_from_gpu": iters, "value": round(ndof / (per_it * iters), 1),
                      "unit": "DoF-solves/s (1 core, extrapolated)"})
        out[name] = o
        log(f"cpu {name}: {o}")
    return out


def assembly_info(prob):
    """The device half of LDGPoisson::discretize: host symbolic and device
    phases (s), and the assembly kernel against the HBM roofline (its
    algorithmic bytes -- every block of M, T, G_d written once, the plan read
    once -- over its duration, HIP events)"""
    t = prob.setup_times()
    gbs = t["assembly_bytes"] / (t["assembly_kernel_ms"] * 1e-3) / 1e9
    return {"host_symbolic_s": round(t["host_s"], 3), "device_s": round(t["device_s"], 3), "kernel": "k_asm_rows",
            "kernel_ms": round(t["assembly_kernel_ms"], 3), "bytes": int(t["assembly_bytes"]),
            "achieved_gbs": round(gbs, 1), "peak_gbs": 8000.0, "frac": round(gbs / 8000.0, 4)}


def bumps(X, seed, n=16):
    """a smooth field: n Gaussian bumps with seeded centres, widths and signs"""
    rng = np.random.default_rng(seed)
    c = rng.random((n, 2))
    w = 0.03 + 0.1 * rng.random(n)
    a = rng.choice([-1.0, 1.0], n) * (0.5 + rng.random(n))
    f = np.zeros(len(X))
    for i in range(n):
        f += a[i] * np.exp(-((X[:, 0] - c[i, 0]) ** 2 + (X[:, 1] - c[i, 1]) ** 2) / (2 * w[i] ** 2))
    return f


def adjoint_leg(mgldg, torch, ndir=64, eps=1e-6, warm=1, fd=True):
    """The adjoint workload: uniform 1024 x 1024 quadtree, p = 2.  u* solves
    f* (16 Gaussian bumps, seed 0); u solves f (16 bumps, seed 1) with
    g = sin(2 pi (x + y)) on the boundary; L = 0.5 ||u - u*||^2.  dL/df and
    dL/dg come from one adjoint MG-PCG and are checked against central
    differences (eps, FP64) along ndir random directions (half in f, half in
    g), the perturbed solves batched 32 columns at a time."""
    ctx = mgldg.Context(0)
    t0 = time.perf_counter()
    qt = mgldg.Quadtree(ctx, 2, lambda x: 1.0 / 1024)
    prob = mgldg.LDGPoisson(qt, 2, tau0=0.0, tauD=1000.0)
    mg = mgldg.Multigrid(prob, "gs", npre=2, npost=2, precision="mixed")
    t_setup = time.perf_counter() - t0
    log(f"adjoint: {prob.ne} elements, {prob.ndof} dofs, {mg.nlevels} levels, setup {t_setup:.2f}s")
    X = prob.nodes()
    xq, _ = prob.bc_qpoints()
    dev = torch.device("cuda", 0)
    fs = torch.tensor(bumps(X, 0), device=dev)
    f = torch.tensor(bumps(X, 1), device=dev)
    g = torch.tensor(np.sin(2 * np.pi * (xq[:, 0] + xq[:, 1])), device=dev)
    zg = torch.zeros_like(g)
    # smooth right-hand sides put the FP64 floor of ||b - A x|| / ||b|| near
    # kappa(A) eps ~ 1e-10 at h = 1/1024: solves stop at 1e-10 or at that floor
    floor = 1e-9

    def ok(sts, what):
        for q in (sts if isinstance(sts, list) else [sts]):
            if q["relres_true"] > floor or q["status"] not in ("converged", "maxit"):
                log("ADJOINT LEG SOLVE FAILED", what, q)
                sys.exit(1)
    kw = dict(tol=1e-10, maxit=300, criterion="true")
    us, st = mgldg.pcg(mg, mgldg.rhs_device(mg, fs, zg), **kw)
    ok(st, "u*")
    u, st = mgldg.pcg(mg, mgldg.rhs_device(mg, f, g), **kw)
    ok(st, "u")
    r = u - us
    L0 = 0.5 * float(r @ r)
    out = {}
    for prec in ("fp64", "mixed"):
        akw = dict(tol=1e-10, maxit=300, criterion="true", precision=prec, flexible=prec == "mixed", use_graph=True)
        for _ in range(warm):
            mgldg.adjoint(mg, r, **akw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        lam, dLdf, dLdg, ast = mgldg.adjoint(mg, r, **akw)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        ok(ast, "adjoint " + prec)
        out[prec] = {"value": round(prob.ndof / dt, 1), "ms_per_step": round(dt * 1e3, 3), "iterations": ast["iterations"],
                     "status": ast["status"], "relres_true": ast["relres_true"]}
        if prec == "fp64":
            grads = (dLdf, dLdg)
        else:
            # the mixed form's gradients against the FP64 adjoint's, full size
            rel = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
            out["mixed_vs_fp64_gradients"] = {"dLdf_rel_diff": rel(dLdf, grads[0]), "dLdg_rel_diff": rel(dLdg, grads[1])}
        log(f"adjoint {prec}: {dt * 1e3:.1f} ms, {prob.ndof / dt:.3g} DoF-solves/s, iterations {ast['iterations']}")
    # the fine level's one-colour GS launch (row tiles where the level has
    # them), HIP-event timed, against the HBM peak: algorithmic bytes of the
    # library's model (tables, b and x of the colour's rows, x of their
    # neighbours once, the pool blocks once, x written)
    ntile, twidth, t64, t32 = mg.level_tiles(0)
    fine_gs = {}
    for prec, name, tiled in (("mixed", "fp32", t32), ("fp64", "fp64", t64)):
        ms, by, _ = mg.profile_relax(0, prec, 50, part="single")
        T = "float" if name == "fp32" else "double"
        fine_gs[name] = {"kernel": f"k_gs_tile<{T},9,0,{twidth}>" if tiled else f"k_gs<{T},9,0,true>",
                         "us_per_launch": round(ms * 1e3, 3), "bytes_per_launch": by,
                         "achieved": round(by / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "frac": round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"adjoint fine GS {name}: {ms * 1e3:.2f} us, {by / 1e6:.2f} MB -> {fine_gs[name]['achieved']:.0f} GB/s")
    if not fd:   # profiling mode: the adjoint solves alone
        return {"ndof": prob.ndof, "levels": mg.nlevels, "adjoint": out, "fine_gs_launch": fine_gs}
    dLdf, dLdg = grads
    # central differences along random smooth directions (4 Gaussian bumps in f,
    # low Fourier modes of g), 32 perturbed solves per batch.  L is quadratic in
    # (f, g), so central differences carry no truncation error: what they see
    # besides the gradient is the solves' rounding divided by 2 eps
    rng = np.random.default_rng(7)
    nf = ndir // 2
    cols = []
    for d in range(ndir):
        if d < nf:
            v = torch.tensor(bumps(X, 100 + d, n=4), device=dev)
            cols.append(("f", v, float(dLdf @ v)))
        else:
            a, m, ph = rng.standard_normal(3), rng.integers(1, 4, 3), 2 * np.pi * rng.random(3)
            gv = sum(a[i] * np.sin(2 * np.pi * m[i] * (xq[:, 0] if i % 2 else xq[:, 1]) + ph[i]) for i in range(3))
            v = torch.tensor(gv, device=dev)
            cols.append(("g", v, float(dLdg @ v)))
    fd_out = {}
    t_fd = time.perf_counter()
    # (L is quadratic in (f, g): every eps is exact but for rounding, which
    # the larger ones divide by more)
    for e in (eps, 1e-3, 1e-2):
        errs = []
        for c0 in range(0, ndir, 16):
            chunk = cols[c0:c0 + 16]
            F = torch.stack([f + sgn * e * v if kind == "f" else f for kind, v, _ in chunk for sgn in (1, -1)], 1)
            G = torch.stack([g + sgn * e * v if kind == "g" else g for kind, v, _ in chunk for sgn in (1, -1)], 1)
            U, sts = mgldg.pcg(mg, mgldg.rhs_device(mg, F.contiguous(), G.contiguous()), **kw)
            ok(sts, "finite differences")
            for j, (kind, v, ad) in enumerate(chunk):
                rp, rm = U[:, 2 * j] - us, U[:, 2 * j + 1] - us
                fd = (0.5 * float(rp @ rp) - 0.5 * float(rm @ rm)) / (2 * e)
                errs.append(abs(fd - ad) / max(abs(ad), 1e-300))
        fd_out[f"{e:g}"] = {"max_rel_err": max(errs), "median_rel_err": float(np.median(errs))}
        log(f"adjoint: L = {L0:.6e}; central differences (eps {e:g}) on {ndir} directions: max rel err "
            f"{max(errs):.3e}, median {float(np.median(errs)):.3e}")
    t_fd = time.perf_counter() - t_fd
    return {
        "workload": "uniform 1024x1024 quadtree, p=2, L = 0.5||u - u*||^2, f* / f = 16 Gaussian bumps (seeds 0 / 1), "
                    "g = sin(2 pi (x+y))",
        "ndof": prob.ndof, "levels": mg.nlevels, "setup_s": round(t_setup, 3), "adjoint": out,
        "mixed_vs_fp64": round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3),
        "fine_gs_launch": fine_gs, "loss": L0,
        "fd_check": {"directions": ndir, "kind": "smooth random (4 bumps in f, 3 Fourier modes in g)",
                     "by_eps": fd_out, "solve_tol": 1e-10, "relres_floor_accepted": floor, "seconds": round(t_fd, 2)},
    }


def periodic_leg(mgldg, torch, warm=1, n=256, p=3):
    """Periodic unit square, uniform n x n quadtree, p = 3: the operator's
    kernel is the constants, so the right-hand side is made mean-zero on the
    device (Function::meanZero, as the reference driver does for periodic and
    Neumann problems); MG-PCG in FP64 and mixed precision to 1e-10."""
    t0 = time.perf_counter()
    ctx = mgldg.Context(0)
    qt = mgldg.Quadtree(ctx, 2, lambda x: 1.0 / n, periodic=True)
    prob = mgldg.LDGPoisson(qt, p, tau0=0.0, tauD=1000.0)
    mg = mgldg.Multigrid(prob, relaxation="gs", omega=0.8, npre=2, npost=2, restriction="ratrat", precision="mixed")
    t_setup = time.perf_counter() - t0
    X = prob.nodes()
    f = np.sin(2 * np.pi * X[:, 0]) * np.cos(4 * np.pi * X[:, 1]) + np.random.default_rng(0).standard_normal(prob.ndof)
    b = torch.tensor(prob.rhs(f), device="cuda")
    mgldg.mean_zero(ctx, b)
    out = {"workload": f"periodic uniform {n}x{n} quadtree, p={p}, f = sin(2 pi x) cos(4 pi y) + N(0,1), mean-zero b",
           "ndof": prob.ndof, "levels": mg.nlevels, "setup_s": round(t_setup, 3)}
    for prec in ("fp64", "mixed"):
        kw = dict(tol=1e-10, maxit=500, criterion="true", precision=prec, flexible=prec == "mixed", use_graph=True)
        for _ in range(warm):
            mgldg.pcg(mg, b, torch.zeros_like(b), **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, st = mgldg.pcg(mg, b, torch.zeros_like(b), **kw)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        out[prec] = {"value": round(prob.ndof / dt, 1), "ms_per_step": round(dt * 1e3, 3), "iterations": st["iterations"],
                     "status": st["status"], "relres_true": st["relres_true"]}
        log(f"periodic {prec}: {dt * 1e3:.2f} ms, {prob.ndof / dt:.3g} DoF-solves/s, iterations {st['iterations']} "
            f"({st['status']}, relres {st['relres_true']:.2e})")
    return out


def ball_mesh(n=41, jitter=0.01, seed=0):
    """Unit ball: the cubic lattice of spacing 1/n, vertices jittered by
    jitter/n (seed), each lattice cube whose corners lie in r <= 1 split into
    the 6 tetrahedra of its Kuhn (Freudenthal) triangulation -- a Delaunay
    triangulation of the unjittered lattice."""
    import itertools
    rng = np.random.default_rng(seed)
    s = 1.0 / n
    m = 2 * n + 1
    g = np.arange(-n, n + 1)
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * s
    c = np.stack(np.meshgrid(np.arange(-n, n), np.arange(-n, n), np.arange(-n, n), indexing="ij"), -1).reshape(-1, 3)
    corners = np.array(list(itertools.product([0, 1], repeat=3)))
    c = c[np.all([np.linalg.norm((c + o) * s, axis=1) <= 1.0 for o in corners], axis=0)]
    vid = lambda w: ((w[:, 0] + n) * m + (w[:, 1] + n)) * m + (w[:, 2] + n)
    tets = []
    for perm in itertools.permutations(range(3)):
        cur = c.copy()
        v = [vid(cur)]
        for d in perm:
            cur = cur.copy()
            cur[:, d] += 1
            v.append(vid(cur))
        tets.append(np.stack(v, 1))
    T = np.concatenate(tets)
    used, inv = np.unique(T, return_inverse=True)
    P = P[used] + jitter * s * rng.uniform(-1, 1, (len(used), 3))
    return P, inv.reshape(T.shape).astype(np.int32)


def ball_delaunay_status(mgldg, torch, maxit=300):
    """The ball workload as the configuration names it: the Delaunay
    tetrahedralisation (scipy / qhull) of the same jittered lattice points
    inside r <= 1, not the Kuhn split the measured leg uses.  Jittered cube
    corners are nearly cospherical, so the Delaunay mesh holds slivers; the
    FP64 MG-PCG is run for at most maxit iterations and its status and final
    residual are recorded."""
    from scipy.spatial import Delaunay
    t0 = time.perf_counter()
    n, jitter = 41, 0.01
    s = 1.0 / n
    rng = np.random.default_rng(0)
    g = np.arange(-n, n + 1)
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * s
    P = P[np.linalg.norm(P, axis=1) <= 1.0]
    P = P + jitter * s * rng.uniform(-1, 1, P.shape)
    T = Delaunay(P).simplices.astype(np.int32)
    T = T[mgldg.locality_order(P, T)]
    e = [P[T[:, k]] - P[T[:, 0]] for k in (1, 2, 3)]
    vol = np.abs(np.einsum("ij,ij->i", np.cross(e[0], e[1]), e[2])) / 6
    med = float(np.median(vol))
    t_mesh = time.perf_counter() - t0
    ctx = mgldg.Context(0)
    t0 = time.perf_counter()
    prob = mgldg.SimplexLDGPoisson(mgldg.Wireframe(ctx, P, T), 2, tau0=100.0, tauD=1000.0)
    mg = mgldg.Multigrid(prob, "gs", omega=0.8, npre=2, npost=2, restriction="rat", precision="fp64")
    t_setup = time.perf_counter() - t0
    b = mgldg.rhs_device(mg, torch.tensor(np.random.default_rng(0).standard_normal(prob.ndof), device="cuda"))
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, st = mgldg.pcg(mg, b, torch.zeros_like(b), tol=1e-10, maxit=maxit, criterion="true", precision="fp64",
                      use_graph=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    out = {"mesh": "Delaunay (qhull) of the 1%-jittered lattice points (spacing 1/41) in r <= 1, p=2",
           "elements": prob.ne, "ndof": prob.ndof, "levels": mg.nlevels,
           "slivers_below_1e-3_median_volume": int((vol < 1e-3 * med).sum()),
           "min_volume_over_median": float(vol.min() / med),
           "fp64": {"status": st["status"], "iterations": st["iterations"], "relres_true": st["relres_true"],
                    "ms": round(dt * 1e3, 1)},
           "setup_s": {"mesh": round(t_mesh, 2), "problem_and_hierarchy": round(t_setup, 2)}}
    log(f"ball, Delaunay mesh (not the measured leg): {out}")
    return out


def ball_leg(mgldg, torch, nrhs=16, warm=1):
    """The 3-D workload: unit ball of tetrahedra (ball_mesh), p = 2, tau0 =
    100, tauD = 1000, the library's agglomeration (at most 8 per agglomerate),
    V(2,2) multicolour block GS, RAT.  nrhs forward right-hand sides of uneven
    support (Gaussian bumps of random centre and width 0.03..0.3) and the
    nrhs adjoint solves of the linear losses L_i = <w_i, u_i> (w_i smooth,
    so dL_i/du_i = w_i is known at once) go in one batch of 2 nrhs columns:
    mixed precision, V-cycle replayed from its hipGraph; beside it the plain
    FP64 single-column solve."""
    t0 = time.perf_counter()
    pts, tets = ball_mesh()
    tets = tets[mgldg.locality_order(pts, tets)]
    t_mesh = time.perf_counter() - t0
    ctx = mgldg.Context(0)
    t0 = time.perf_counter()
    wf = mgldg.Wireframe(ctx, pts, tets)
    prob = mgldg.SimplexLDGPoisson(wf, 2, tau0=100.0, tauD=1000.0)
    t_prob = time.perf_counter() - t0
    t0 = time.perf_counter()
    parts = mgldg.agglomerate(prob)
    t_agg = time.perf_counter() - t0
    t0 = time.perf_counter()
    mg = mgldg.Multigrid(prob, "gs", omega=0.8, npre=2, npost=2, restriction="rat", precision="mixed",
                         partitions=parts)
    t_mg = time.perf_counter() - t0
    log(f"ball: {prob.ne} tetrahedra, {prob.ndof} dofs, levels {mg.level_rows}; setup mesh {t_mesh:.2f}s "
        f"problem {t_prob:.2f}s agglomeration {t_agg:.2f}s hierarchy {t_mg:.2f}s")
    X = prob.nodes()
    rng = np.random.default_rng(3)
    dev = torch.device("cuda", 0)
    cols = []
    for i in range(nrhs):        # forward: localised sources of uneven support
        c = rng.uniform(-0.6, 0.6, 3)
        w = 0.03 + 0.27 * rng.random()
        cols.append(np.exp(-((X - c) ** 2).sum(1) / (2 * w * w)))
    F = torch.tensor(np.stack(cols, 1), device=dev)
    B = mgldg.rhs_device(mg, F)
    Wt = []
    for i in range(nrhs):        # adjoint: dL_i/du_i = w_i, smooth weights
        a = rng.standard_normal(3)
        Wt.append(np.cos(np.pi * (X @ a)))
    Badj = torch.tensor(np.stack(Wt, 1), device=dev)
    Ball = torch.cat([B, Badj], 1).contiguous()
    out = {}
    kw = dict(tol=1e-10, maxit=400, criterion="true")
    b0 = Ball[:, 0].contiguous()
    for prec in ("fp64", "mixed"):
        skw = dict(kw, precision=prec, flexible=prec == "mixed", use_graph=True)
        for _ in range(warm):
            mgldg.pcg(mg, b0, torch.zeros_like(b0), **skw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, st = mgldg.pcg(mg, b0, torch.zeros_like(b0), **skw)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        if st["status"] != "converged":
            log("BALL NOT CONVERGED", prec, st)
            sys.exit(1)
        out[f"single_{prec}"] = {"value": round(prob.ndof / dt, 1), "ms_per_step": round(dt * 1e3, 3),
                                 "iterations": st["iterations"]}
        log(f"ball single {prec}: {dt * 1e3:.1f} ms, {prob.ndof / dt:.3g} DoF-solves/s, iterations {st['iterations']}")
    bkw = dict(kw, precision="mixed", flexible=True, use_graph=True)
    mgldg.pcg(mg, Ball, torch.zeros_like(Ball), **dict(bkw, maxit=2))   # batch buffers, kernel plan, graph: not timed
    torch.cuda.synchronize()
    t = time.perf_counter()
    Xall, sts = mgldg.pcg(mg, Ball, torch.zeros_like(Ball), **bkw)
    dLdf, dLdg = mgldg.adjoint_maps(mg, Xall[:, nrhs:])           # dL_i/df = M lambda_i, dL_i/dg = B_g^T lambda_i
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    for q in sts:
        if q["status"] != "converged" or q["relres_true"] > 1e-10:
            log("BALL BATCH NOT CONVERGED", q)
            sys.exit(1)
    k = 2 * nrhs
    out["batched_mixed"] = {"value": round(prob.ndof * k / dt, 1), "ms_per_step": round(dt * 1e3, 3),
                            "iterations": max(q["iterations"] for q in sts), "columns": k}
    log(f"ball batched mixed ({nrhs} forward + {nrhs} adjoint): {dt * 1e3:.1f} ms, {prob.ndof * k / dt:.3g} DoF-solves/s, "
        f"iterations {out['batched_mixed']['iterations']}")
    x1, _ = mgldg.pcg(mg, b0, **dict(kw, precision="fp64"))
    diff = float(torch.linalg.norm(Xall[:, 0] - x1) / torch.linalg.norm(x1))
    # adjoint check on one column: dL/df . v = <w, A^{-1} M v> for a smooth v
    v = torch.tensor(np.sin(3 * X[:, 0]) * np.cos(2 * X[:, 2]), device=dev)
    uv, _ = mgldg.pcg(mg, mgldg.rhs_device(mg, v), **dict(kw, precision="fp64"))
    adj_err = abs(float(dLdf[:, 0] @ v) - float(Badj[:, 0] @ uv)) / abs(float(Badj[:, 0] @ uv))
    del dLdg
    return {
        "workload": "unit ball, Kuhn split of a 1%-jittered cubic lattice (spacing 1/41) inside r <= 1, p=2, "
                    "tau0=100, tauD=1000",
        "elements": prob.ne, "ndof": prob.ndof, "levels": mg.nlevels, "level_rows": mg.level_rows,
        "smoother": "V(2,2) multicolour block GS, omega 0.8", "coarsening": "own agglomeration (<= 8 per level), RAT",
        "nrhs": nrhs, **out,
        "mixed_batched_vs_fp64_single": round(out["batched_mixed"]["value"] / out["single_fp64"]["value"], 3),
        "mixed_batched_vs_fp64_single_rel_diff": diff,
        "adjoint_check_rel_err": adj_err,
        "kernel_plan": plan_summary(mg.batch_plan()), "mfma_fine_apply": mfma_use(mg, mg.batch_plan(), 2 * nrhs),
        "setup_s": {"mesh": round(t_mesh, 3), "problem": round(t_prob, 3), "agglomeration": round(t_agg, 3),
                    "hierarchy": round(t_mg, 3), "assembly": assembly_info(prob)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--roofline-only", action="store_true", help="setup + fine-level smoother sweeps only (profiling)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-maxit", type=int, default=60)
    ap.add_argument("--nrhs", type=int, default=32, help="right-hand sides of the batched form")
    ap.add_argument("--solve-only", action="store_true", help="setup + three mixed solves (profiling)")
    ap.add_argument("--no-unstructured", action="store_true", help="skip the unit-disk workload")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workload")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint workload")
    ap.add_argument("--adjoint-solve-only", action="store_true",
                    help="the adjoint workload's setup and adjoint solves, no finite differences (profiling)")
    ap.add_argument("--no-ball", action="store_true", help="skip the 3-D ball workload")
    ap.add_argument("--no-delaunay", action="store_true", help="skip the Delaunay-mesh status run of the ball workload")
    ap.add_argument("--ball-only", action="store_true", help="only the 3-D ball workload")
    ap.add_argument("--setup-only", action="store_true",
                    help="the device assembly: the adaptive quadtree problem and hierarchy and the ball problem (profiling)")
    ap.add_argument("--unstructured-only", action="store_true", help="only the unit-disk workload")
    ap.add_argument("--no-periodic", action="store_true", help="skip the periodic workload")
    ap.add_argument("--batched-only", action="store_true",
                    help="setup + two mixed and one FP64 batched solves of --nrhs columns (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the path shards by right-hand-side batch or by element ranges, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    import mgldg

    if not os.path.exists(mgldg.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    if args.adjoint_only or args.adjoint_solve_only:
        print(json.dumps({"adjoint": adjoint_leg(mgldg, torch, warm=args.warmup, fd=not args.adjoint_solve_only)}),
              flush=True)
        return
    if args.ball_only:
        print(json.dumps({"ball": ball_leg(mgldg, torch, warm=args.warmup)}), flush=True)
        return
    if args.unstructured_only:
        print(json.dumps({"unstructured": unstructured_leg(mgldg, torch, args.nrhs, args.warmup)}), flush=True)
        return
    if args.setup_only:
        ctx = mgldg.Context(0)
        t0 = time.perf_counter()
        qt = mgldg.Quadtree(ctx, 2, c2_size, coarsening="rapid", balance=True)
        prob = mgldg.LDGPoisson(qt, 3, bc="dirichlet", tau0=0.0, tauD=1000.0)
        mgldg.Multigrid(prob, relaxation="gs", omega=0.8, npre=2, npost=2, restriction="ratrat", precision="mixed")
        quad = {"setup_s": round(time.perf_counter() - t0, 3), "assembly": assembly_info(prob)}
        log(f"quadtree problem: {quad}")
        pts, tets = ball_mesh()
        tets = tets[mgldg.locality_order(pts, tets)]
        t0 = time.perf_counter()
        bprob = mgldg.SimplexLDGPoisson(mgldg.Wireframe(ctx, pts, tets), 2, tau0=100.0, tauD=1000.0)
        ball = {"problem_s": round(time.perf_counter() - t0, 3), "assembly": assembly_info(bprob)}
        log(f"ball problem: {ball}")
        print(json.dumps({"setup": {"quadtree": quad, "ball": ball}}), flush=True)
        return

    torch.cuda.init()
    ctx = mgldg.Context(0)
    t0 = time.perf_counter()
    qt = mgldg.Quadtree(ctx, 2, c2_size, coarsening="rapid", balance=True)
    t_qt = time.perf_counter() - t0
    t0 = time.perf_counter()
    prob = mgldg.LDGPoisson(qt, 3, bc="dirichlet", tau0=0.0, tauD=1000.0)
    t_prob = time.perf_counter() - t0
    t0 = time.perf_counter()
    mg = mgldg.Multigrid(prob, relaxation="gs", omega=0.8, npre=2, npost=2, restriction="ratrat", precision="mixed")
    t_mg = time.perf_counter() - t0
    log(f"ndof {prob.ndof} elements {prob.ne} levels {mg.nlevels} rows {mg.level_rows}")
    log(f"setup: quadtree {t_qt:.2f}s problem {t_prob:.2f}s hierarchy {t_mg:.2f}s {mg.setup_times()}")

    roofline = None
    if not (args.solve_only or args.batched_only):
        # roofline leg.  Dominant kernel: the fine level's one-colour GS
        # launches of the FP32 V-cycle (k_gs); beside it whole sweeps (those
        # launches + the one doing the last two colours, k_gs_pair)
        nrep = 50
        ms32, by32, n32 = mg.profile_relax(0, "mixed", nrep, part="single")
        ms64, by64, n64 = mg.profile_relax(0, "fp64", nrep, part="single")
        sw32 = mg.profile_relax(0, "mixed", nrep, part="sweep")
        sw64 = mg.profile_relax(0, "fp64", nrep, part="sweep")
        mss, bys = mg.profile_spmv(0, "fp64", nrep)
        gbs = lambda ms, by: by / (ms * 1e-3) / 1e9
        gs32, gs64, spmv64 = gbs(ms32, by32), gbs(ms64, by64), gbs(mss, bys)
        sweep = {}
        for name, (ms, by, nl) in (("fp32", sw32), ("fp64", sw64)):
            per = nl // nrep
            sweep[name] = {"launches": per, "us_per_sweep": round(ms * per * 1e3, 2), "MB_per_sweep": round(by * per / 1e6, 2),
                           "GBs": round(gbs(ms, by), 1)}
        log(f"GS one-colour launches fp32: {ms32 * 1e3:.1f} us/launch, {by32 / 1e6:.2f} MB -> {gs32:.0f} GB/s; "
            f"fp64: {ms64 * 1e3:.1f} us, {gs64:.0f} GB/s; SpMV fp64 {mss * 1e3:.1f} us {spmv64:.0f} GB/s; sweeps {sweep}")
        # the same launch on the unpooled operator (every block streamed from
        # HBM): the kernel's streaming efficiency, beside the pooled launch
        # the solve uses (the quadtree operator's distinct blocks from cache)
        mg_np = mgldg.Multigrid(prob, relaxation="gs", omega=0.8, npre=2, npost=2, restriction="ratrat",
                                precision="mixed", pool=False)
        msu, byu, _ = mg_np.profile_relax(0, "mixed", nrep, part="single")
        del mg_np
        gsu = gbs(msu, byu)
        log(f"unpooled GS one-colour launch fp32: {msu * 1e3:.1f} us/launch, {byu / 1e6:.2f} MB -> {gsu:.0f} GB/s")
        # the fine level's one-colour launches run on the row tiles when the
        # level has enough of them (MFMA, rows with the same blocks 16 at a time)
        ntile, twidth, _, tiled = mg.level_tiles(0)
        kname = f"k_gs_tile<float, 16, 0, {twidth}>" if tiled else "k_gs<float, 16, 0, true>"
        traffic = traffic_u = None
        tp = os.path.join(ROOT, "profiles", "k_gs_c2_fine_pmc_traffic.json")
        if os.path.exists(tp):   # measured offline: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated
            kk = json.load(open(tp))["kernels"]
            traffic = kk.get("void mgldg::dev::" + kname, {}).get("traffic_bytes")
            traffic_u = kk.get("void mgldg::dev::k_gs<float, 16, 0, false>", {}).get("traffic_bytes")
        roofline = {"bound": "hbm", "kernel": kname.replace(" ", "") + " (fine-level one-colour block GS launch, pooled "
                    "operator" + (f", row tiles: {ntile} tiles of 16 rows over the level" if tiled else "") + ")",
                    "achieved": round(gs32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gs32 / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "bytes_per_launch": by32, "us_per_launch": round(ms32 * 1e3, 3),
                    "limiter": "latency: about a thousand tiles per launch, one wavefront each (one per SIMD), each a "
                               "chain of table loads -> x gathers -> 32 MFMAs -> D^-1 MFMAs -> store; the pool removed 94% "
                               "of the launch's HBM bytes",
                    "unpooled": {"kernel": "k_gs<float,16,0,false>", "us_per_launch": round(msu * 1e3, 3),
                                 "bytes_per_launch": byu, "achieved": round(gsu, 1), "frac": round(gsu / HBM_PEAK_GBS, 4),
                                 "traffic": traffic_u},
                    "speedup_vs_unpooled": round(msu / ms32, 2),
                    "fp64_gs_GBs": round(gs64, 1), "fp64_spmv_GBs": round(spmv64, 1), "sweep": sweep}
    if args.roofline_only:
        # calibration for the PMC traffic counters: a plain device copy of known size
        src = torch.ones(CALIB_BYTES // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"roofline": roofline, "calibration_copy_bytes": CALIB_BYTES}), flush=True)
        return

    rng = np.random.default_rng(0)
    f = rng.standard_normal(prob.ndof)
    b_host = prob.rhs(f)
    dev = torch.device("cuda", 0)
    b = torch.tensor(b_host, device=dev)
    if args.solve_only:
        # profiling mode: three mixed solves, V-cycle replayed from its hipGraph
        for _ in range(3):
            _, st = mgldg.pcg(mg, b, torch.zeros_like(b), tol=1e-10, maxit=500, criterion="true", precision="mixed",
                              flexible=True, use_graph=True)
        torch.cuda.synchronize()
        print(json.dumps({"solve": st}), flush=True)
        return

    if args.batched_only:
        # profiling mode: batched solves (autotuned block kernels, V-cycle from its hipGraph)
        Bp = torch.tensor(prob.rhs(rng.standard_normal((prob.ndof, args.nrhs))), device=dev)
        for prec in ("mixed", "mixed", "fp64"):
            _, st = mgldg.pcg(mg, Bp, torch.zeros_like(Bp), tol=1e-10, maxit=500, criterion="true", precision=prec,
                              flexible=prec == "mixed", use_graph=True)
        torch.cuda.synchronize()
        print(json.dumps({"batched": st[0], "kernel_plan": plan_summary(mg.batch_plan())}), flush=True)
        return

    def run(precision, steps, warm):
        kw = dict(tol=1e-10, maxit=500, criterion="true", precision=precision, flexible=(precision == "mixed"),
                  use_graph=True)
        st = None
        for _ in range(warm):
            x = torch.zeros_like(b)
            _, st = mgldg.pcg(mg, b, x, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        xs, stats = [], []
        for _ in range(steps):
            x = torch.zeros_like(b)
            _, st = mgldg.pcg(mg, b, x, **kw)
            stats.append(st)
            xs.append(x)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        return dt, stats, xs[-1].cpu().numpy()

    dt64, st64, x64 = run("fp64", args.steps, args.warmup)
    log(f"fp64 : {dt64 / args.steps * 1e3:.2f} ms/solve  {st64[-1]}")
    dt32, st32, x32 = run("mixed", args.steps, args.warmup)
    log(f"mixed: {dt32 / args.steps * 1e3:.2f} ms/solve  {st32[-1]}")
    for st in st64 + st32:
        if st["status"] != "converged" or st["relres_true"] > 1e-10:
            log("NOT CONVERGED", st)
            sys.exit(1)
    ndof = prob.ndof
    v64 = ndof * args.steps / dt64
    v32 = ndof * args.steps / dt32
    # whole-iteration roofline: the cost model's algorithmic bytes of one CG
    # iteration (V-cycle + product + updates + reductions) over the measured
    # time per iteration; launches per V-cycle beside the captured graph's nodes
    iteration = {}
    for prec, dt, sts in (("mixed", dt32, st32), ("fp64", dt64, st64)):
        c = mg.solve_cost(prec)
        its = sts[-1]["iterations"] + 1          # the initial V-cycle and product besides the iterations
        t_it = dt / args.steps / its
        iteration[prec] = {"vcycle_launches": c["vcycle_launches"], "vcycle_graph_nodes": c["graph_nodes"],
                           "iteration_launches": c["iteration_launches"], "MB_per_iteration": round(c["iteration_bytes"] / 1e6, 2),
                           "us_per_iteration": round(t_it * 1e6, 1),
                           "achieved": round(c["iteration_bytes"] / t_it / 1e9, 1),
                           "frac": round(c["iteration_bytes"] / t_it / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"whole-iteration cost: {iteration}")
    if roofline is not None:
        roofline["iteration"] = iteration

    # batched form: k right-hand sides share every operator application
    k = args.nrhs
    Fb = rng.standard_normal((prob.ndof, k))
    Bb = torch.tensor(prob.rhs(Fb), device=dev)

    def run_batched(precision, steps, warm):
        kw = dict(tol=1e-10, maxit=500, criterion="true", precision=precision, flexible=(precision == "mixed"),
                  use_graph=True)
        for _ in range(warm):
            mgldg.pcg(mg, Bb, torch.zeros_like(Bb), **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            X = torch.zeros_like(Bb)
            _, st = mgldg.pcg(mg, Bb, X, **kw)
        torch.cuda.synchronize()
        return time.perf_counter() - t, st

    bsteps = max(1, args.steps // 2)
    dtb64, stb64 = run_batched("fp64", bsteps, 1)
    dtb32, stb32 = run_batched("mixed", bsteps, 1)
    for st in stb64 + stb32:
        if st["status"] != "converged" or st["relres_true"] > 1e-10:
            log("BATCHED NOT CONVERGED", st)
            sys.exit(1)
    vb64 = prob.ndof * k * bsteps / dtb64
    vb32 = prob.ndof * k * bsteps / dtb32
    plan = mg.batch_plan()
    log(f"batched k={k}: fp64 {dtb64 / bsteps * 1e3:.1f} ms/solve ({vb64:.3g} DoF-solves/s), "
        f"mixed {dtb32 / bsteps * 1e3:.1f} ms ({vb32:.3g}); max iterations {max(s['iterations'] for s in stb32)}; "
        f"kernel plan {plan_summary(plan)}")

    # the driver's stationary solver (--solver vcycles) on the same problem
    vcyc = {}
    for prec in ("fp64", "mixed"):
        mgldg.vcycles(mg, b, torch.zeros_like(b), ncycle=2, precision=prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        xv, rn = mgldg.vcycles(mg, b, torch.zeros_like(b), ncycle=200, tol=1e-10, precision=prec)
        torch.cuda.synchronize()
        dtv = time.perf_counter() - t
        ncyc = len(rn) - 1
        vcyc[prec] = {"cycles": ncyc, "ms_per_solve": round(dtv * 1e3, 3), "value": round(ndof / dtv, 1),
                      "mean_contraction": round(float((rn[-1] / rn[0]) ** (1.0 / max(ncyc, 1))), 4),
                      "relres": float(rn[-1] / float(torch.linalg.norm(b)))}
    log(f"vcycles (stationary solver) to 1e-10: {vcyc}")

    # hp-multigrid on the same problem: p = 3 -> 1, then the h-levels at p = 1
    t0 = time.perf_counter()
    mg_hp = mgldg.Multigrid(prob, relaxation="gs", omega=0.8, npre=2, npost=2, restriction="ratrat", precision="mixed",
                            pdegrees=[1])
    hp = {"pdegrees": [3, 1], "levels": mg_hp.nlevels, "level_B": mg_hp.level_B[:3], "setup_s": round(time.perf_counter() - t0, 3)}
    for prec in ("fp64", "mixed"):
        kw = dict(tol=1e-10, maxit=500, criterion="true", precision=prec, flexible=prec == "mixed", use_graph=True)
        mgldg.pcg(mg_hp, b, torch.zeros_like(b), **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, sth = mgldg.pcg(mg_hp, b, torch.zeros_like(b), **kw)
        torch.cuda.synchronize()
        dth = time.perf_counter() - t
        hp[prec] = {"value": round(ndof / dth, 1), "ms_per_step": round(dth * 1e3, 3), "iterations": sth["iterations"],
                    "relres_true": sth["relres_true"]}
    del mg_hp
    log(f"hp-multigrid (p 3 -> 1, then h): {hp}")
    periodic = None if args.no_periodic else periodic_leg(mgldg, torch, args.warmup)

    cpu = None
    err = {}
    if not args.no_cpu:
        import oracle as O
        log("cpu baseline: FP64 oracle (C restatement, 1 thread) on the same problem")
        t0 = time.perf_counter()
        oqt = O.Quadtree(2, c2_size, rapid=True, balance=True)
        opr = O.Problem(oqt, 3)
        omg = O.Multigrid(opr, "mcgs", npre=2, npost=2)
        t_setup = time.perf_counter() - t0
        ob = opr.rhs(f, np.zeros(opr.nqb))
        t0 = time.perf_counter()
        xo, ito, rro, _ = O.pcg(opr, omg, ob, tol=1e-10, maxit=args.cpu_maxit, criterion="true")
        t_solve = time.perf_counter() - t0
        per_it = t_solve / max(ito, 1)
        full = ito if rro <= 1e-10 else st64[-1]["iterations"]
        cpu_rate = ndof / (per_it * full)
        sample = (f"same C2 problem, oracle setup {t_setup:.1f}s (not timed), {ito} PCG iterations in {t_solve:.1f}s"
                  + ("" if rro <= 1e-10 else f", extrapolated to {full} iterations"))
        c2_one = (t_setup, t_solve, ito, rro, ndof) if rro <= 1e-10 else None
        cpu = {"value": round(cpu_rate, 1), "unit": "DoF-solves/s", "cores": 1, "kind": "port", "sample": sample,
               "host_cpu": cpu_model(), "host_nproc": os.cpu_count(), "iterations": ito, "relres_true": rro}
        nx = np.linalg.norm(xo)
        err = {"fp64_vs_oracle": float(np.linalg.norm(x64 - xo) / nx), "mixed_vs_oracle": float(np.linalg.norm(x32 - xo) / nx),
               "oracle_iterations": ito}
        log(f"cpu: {cpu_rate:.0f} DoF-solves/s, iterations {ito}, errors {err}")

    c1 = c1_gpu_leg(mgldg, torch)
    unstructured = None if args.no_unstructured else unstructured_leg(mgldg, torch, k, args.warmup)
    adjoint = None if args.no_adjoint else adjoint_leg(mgldg, torch, warm=args.warmup)
    ball = None if args.no_ball else ball_leg(mgldg, torch, warm=args.warmup)
    if ball is not None and not args.no_delaunay:
        ball["delaunay_mesh_status"] = ball_delaunay_status(mgldg, torch)
    if cpu is not None:
        gpu_iters = {"c3": unstructured and unstructured["single_fp64"]["iterations"],
                     "c4": adjoint and adjoint["adjoint"]["fp64"]["iterations"],
                     "c5": ball and ball["single_fp64"]["iterations"]}
        legs = cpu_legs(gpu_iters, c2_one)
        cpu["all_cores"] = {"value": legs["c2"]["all_cores"]["value"], "cores": legs["c2"]["all_cores"]["cores"],
                            "sample": f"{legs['c2']['all_cores']['solves']} concurrent full C2 oracle solves, one per core"}
        cpu["legs"] = legs

    line = {
        "metric": "DoF-solves/s to 1e-10 residual, mixed vs FP64 MG-PCG; smoother HBM GB/s",
        "value": round(v32, 1),
        "unit": "DoF-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt32 / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 CG",
        "data": "synthetic: f ~ N(0,1) per dof (seed 0), g = 0",
        "config": {"workload": "C2 adaptive quadtree [0,1]^2 around (0.3,0.7), depth 14, p=3, 1 RHS",
                   "ndof": ndof, "elements": prob.ne, "levels": mg.nlevels, "nrhs": 1, "tol": 1e-10,
                   "smoother": "V(2,2) multicolour block GS, omega 0.8", "coarsening": "RATRAT, rapid",
                   "parallelism": "single GPU"},
        "fp64": {"value": round(v64, 1), "ms_per_step": round(dt64 / args.steps * 1e3, 3),
                 "iterations": st64[-1]["iterations"], "relres_true": st64[-1]["relres_true"]},
        "mixed": {"value": round(v32, 1), "ms_per_step": round(dt32 / args.steps * 1e3, 3),
                  "iterations": st32[-1]["iterations"], "relres_true": st32[-1]["relres_true"]},
        "speedup_mixed_vs_fp64": round(v32 / v64, 3),
        "batched": {"nrhs": k, "fp64": {"value": round(vb64, 1), "ms_per_step": round(dtb64 / bsteps * 1e3, 3),
                                        "max_iterations": max(s["iterations"] for s in stb64)},
                    "mixed": {"value": round(vb32, 1), "ms_per_step": round(dtb32 / bsteps * 1e3, 3),
                              "max_iterations": max(s["iterations"] for s in stb32)},
                    "mixed_batched_vs_fp64_single": round(vb32 / v64, 3),
                    "mixed_vs_fp64_batched": round(vb32 / vb64, 3),
                    "kernel_plan": plan_summary(plan), "mfma_fine_apply": mfma_use(mg, plan, k)},
        "vcycles": vcyc,
        "hp_multigrid": hp,
        "periodic": periodic,
        "error_vs_fp64_oracle": err,
        "setup_s": {"quadtree": round(t_qt, 3), "problem": round(t_prob, 3), "hierarchy": round(t_mg, 3),
                    "assembly": assembly_info(prob)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "c1": c1,
        "unstructured": unstructured,
        "adjoint": adjoint,
        "ball": ball,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
