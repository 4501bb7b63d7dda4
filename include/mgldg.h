/*
 * mgldg.h — C ABI of the MI355X-native LDG-Poisson MG-PCG library (libmgldg.so).
 *
 * The reference (danfortunato/multigrid-ldg) has no C ABI: its "operator API"
 * is a C++ template surface.  Each entry point below replaces one piece of that
 * surface; the reference interface it stands in for is cited beside it
 * (paths relative to the reference repository root).
 *
 * Conventions
 *   - Every function returns an int status: 0 = OK, negative = error.  The
 *     message of the last error on a context is returned by mgldg_last_error.
 *     No C++ exception crosses this boundary.
 *   - Host buffers are caller-owned.  Functions with the _device suffix take
 *     device pointers (HBM) and enqueue on the context's stream; they return
 *     after the work is complete unless stated otherwise.
 *   - Layouts follow the reference: element-major vectors with npl = (p+1)^N
 *     nodal values per element, tensor node index with the first coordinate
 *     slowest (src/cartesian/util/range.h:187-194); element order is the
 *     quadtree depth-first order (src/cartesian/mesh/quadtree.h:176-186);
 *     exported BSR blocks are row-major (src/cartesian/util/common.h:25-32).
 *     Multiple right-hand sides are stored [ndof][nrhs] (right-hand side
 *     fastest).
 *   - A multigrid handle is not reentrant (the reference's Multigrid holds
 *     mutable x/b/r scratch, src/cartesian/multigrid/multigrid.h:268-273).
 *     Use one context per host thread.
 */
#ifndef MGLDG_H
#define MGLDG_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MGLDG_OK             0
#define MGLDG_ERR_ARG       -1   /* invalid argument (replaces std::invalid_argument / out_of_range) */
#define MGLDG_ERR_HIP       -2   /* HIP runtime failure */
#define MGLDG_ERR_NOMEM     -3   /* host or device allocation failure */
#define MGLDG_ERR_STATE     -4   /* handle in the wrong state */
#define MGLDG_ERR_UNSUPPORTED -5 /* valid request this build does not implement */
#define MGLDG_ERR_IO        -6   /* a file could not be opened or written */

/* ---- solve status (pcg returns void in the reference: krylov.h:23) ---- */
#define MGLDG_CONVERGED   0
#define MGLDG_MAXIT       1   /* maxit reached before the tolerance */
#define MGLDG_BREAKDOWN   2   /* d^T A d <= 0 or r^T z <= 0 */
#define MGLDG_NONFINITE   3   /* NaN/Inf met in the recurrences */

/* ---- enums (src/cartesian/multigrid/multigrid.h:24-63, quadtree.h:19-23) ---- */
#define MGLDG_EQUAL_COARSENING 0   /* kEqualCoarsening */
#define MGLDG_RAPID_COARSENING 1   /* kRapidCoarsening */
#define MGLDG_DIRICHLET 0          /* kDirichlet (boundaryconditions.h:10-15) */
#define MGLDG_NEUMANN   1          /* kNeumann */
#define MGLDG_JACOBI        0      /* MG::kJacobi */
#define MGLDG_GAUSS_SEIDEL  1      /* MG::kGaussSeidel, multicolour on the GPU */
#define MGLDG_GAUSS_SEIDEL_LEX 2   /* MG::kGaussSeidel in the reference's own (lexicographic)
                                      order, run as wavefronts of independent rows */
#define MGLDG_COARSE_CHOLESKY   0  /* MG::kCholesky */
#define MGLDG_COARSE_RELAXATION 1  /* MG::kRelaxation */
#define MGLDG_RAT    0             /* MG::kRAT */
#define MGLDG_RATRAT 1             /* MG::kRATRAT */
#define MGLDG_FP64   0             /* V-cycle in FP64 */
#define MGLDG_MIXED  1             /* V-cycle in FP32, CG and true residual in FP64 */
#define MGLDG_CRIT_REFERENCE 0     /* r^T z <= tol^2 r0^T z0 (krylov.h:42) */
#define MGLDG_CRIT_TRUE      1     /* ||b - A x|| <= tol ||b||, confirmed in FP64 */

typedef struct mgldg_ctx_s*      mgldg_ctx;
typedef struct mgldg_quadtree_s* mgldg_quadtree;
typedef struct mgldg_problem_s*  mgldg_problem;
typedef struct mgldg_mg_s*       mgldg_mg;
typedef struct mgldg_wireframe_s* mgldg_wireframe;

/* size function: N widths for the point x.  A cell splits iff every width of
 * the cell exceeds h(center) (src/cartesian/mesh/quadtree.tpp:15). */
typedef void (*mgldg_sizefn)(const double* x, double* h, void* user);

/* ======================= context ======================= */
/* No reference counterpart (the reference is single-threaded CPU code).
 * stream: a hipStream_t, or NULL to create a private stream. */
int mgldg_ctx_create(int device, void* stream, mgldg_ctx* out);
int mgldg_ctx_destroy(mgldg_ctx ctx);
const char* mgldg_last_error(mgldg_ctx ctx);
int mgldg_ctx_synchronize(mgldg_ctx ctx);

/* ======================= mesh ======================= */
/* Quadtree<N>(h, strategy, periodic) and Quadtree<N>(domain, h, ...)
 * (src/cartesian/mesh/quadtree.h:211-228, build quadtree.tpp:8-40).
 * lo/hi may be NULL for [0,1]^N.  balance=1 additionally enforces 2:1 face
 * balance of the leaves (not in the reference). */
int mgldg_quadtree_create(mgldg_ctx ctx, int N, const double* lo, const double* hi,
                          mgldg_sizefn h, void* user, int strategy, int periodic,
                          int balance, mgldg_quadtree* out);
/* info[0]=nodes info[1]=numLevels info[2]=leaves (Quadtree::size/numLevels) */
int mgldg_quadtree_info(mgldg_quadtree qt, int* info);
/* Quadtree::layer (quadtree.h:168-178): node ids of layer l in DFS order;
 * ids may be NULL to query the count. */
int mgldg_quadtree_layer(mgldg_quadtree qt, int l, int* ids, int* count);
int mgldg_quadtree_destroy(mgldg_quadtree qt);

/* Wireframe<N>(p, t) and Wireframe<N>(file) (src/unstructured/mesh/wireframe.h:12-77):
 * the points (np x N) and simplices (nt x (N+1)) of an unstructured mesh.
 * mgldg_wireframe_create takes 0-based vertex ids; mgldg_wireframe_read reads
 * the reference's .mesh text format (header "np nt", a blank line, np point
 * lines, a blank line, nt lines of 1-based vertex ids). */
int mgldg_wireframe_create(mgldg_ctx ctx, int N, int np, const double* points, int nt,
                           const int* simplices, mgldg_wireframe* out);
int mgldg_wireframe_read(mgldg_ctx ctx, int N, const char* path, mgldg_wireframe* out);
/* info[0]=N info[1]=np info[2]=nt; points / simplices (0-based) may be NULL */
int mgldg_wireframe_get(mgldg_wireframe wf, int* info, double* points, int* simplices);
int mgldg_wireframe_destroy(mgldg_wireframe wf);

/* ======================= assembly ======================= */
/* Mesh<N,P>(qt) + BoundaryConditions + LDGPoisson<N,P>(mesh, bcs, tau0, tauD)
 * (src/cartesian/mesh/mesh.h:186-265, src/cartesian/dg/ldgpoisson.h:60-77,
 * ldgpoisson.tpp:14-161).  bctype[2N]: type of boundary -(2d+1) (low side of
 * dimension d) at index 2d and -(2d+2) at 2d+1.  Boundary data g enters
 * later as values at the boundary quadrature points (mgldg_problem_bc_qpoints). */
int mgldg_ldg_create(mgldg_quadtree qt, int p, const int* bctype, double tau0, double tauD,
                     mgldg_problem* out);
/* Mesh<N,P>(wireframe, bndfuns) + BoundaryConditions + LDGPoisson<N,P>(mesh,
 * bcs, tau0, tauD) on simplices (src/unstructured/mesh/mesh.h:185-298,
 * src/unstructured/dg/ldgpoisson.tpp:14-170).  The boundary predicates
 * bndfuns[k] enter as vertex masks: bit k of vmask[v] says whether predicate k
 * holds at vertex v; a boundary face gets key k+1 for the first k holding at
 * all of its vertices (mesh.h:276-288).  nbnd = 0 (vmask NULL): every boundary
 * face has key 1.  bctype[max(nbnd,1)]: MGLDG_DIRICHLET / MGLDG_NEUMANN per key.
 * Faces are numbered in first-appearance order (simplex i, local face j); the
 * reference's order is that of a hash map. */
int mgldg_ldg_create_simplex(mgldg_wireframe wf, int p, int nbnd, const int* vmask,
                             const int* bctype, double tau0, double tauD, mgldg_problem* out);
/* sizes[0..6] = ne, nf, nbfaces, npl, nqb (boundary quadrature points), ndof, N */
int mgldg_problem_sizes(mgldg_problem pr, int* sizes);
/* Master::dgnodes for every element (master.h:35-43): xyz[ndof*N] */
int mgldg_problem_nodes(mgldg_problem pr, double* xyz);
/* boundary quadrature points in boundary-face order (ldgpoisson.tpp:119-129);
 * tag: the boundary key of each point (quadtree: -(2d+1) / -(2d+2); simplex:
 * 1..nbnd) */
int mgldg_problem_bc_qpoints(mgldg_problem pr, double* xyz, int* tag);
/* LDGPoisson::computeRHS (ldgpoisson.tpp:167-176) plus add_source_terms
 * (:103-161): b = M f + B_g g.  Host buffers; f[ndof*nrhs] nodal values,
 * g[nqb*nrhs] (either may be NULL = zero). */
int mgldg_rhs(mgldg_problem pr, const double* f, const double* g, int nrhs, double* b);
/* LDGOperators dump (ldgpoisson.h:100-109) as sorted CSR of row-major blocks.
 * which: 0=A 1=M 2=T 3+d=G_d.  Call with NULL arrays to get *nnzb first. */
int mgldg_problem_export(mgldg_problem pr, int which, int* rowptr, int* cols,
                         double* blocks, long* nnzb);
int mgldg_problem_destroy(mgldg_problem pr);

/* ======================= multigrid setup ======================= */
/* MG::Parameters (multigrid.h:48-62) plus the GPU forms. */
typedef struct {
    int relaxation;    /* MGLDG_JACOBI | MGLDG_GAUSS_SEIDEL | MGLDG_GAUSS_SEIDEL_LEX */
    double omega;      /* damping, reference default 0.8                   */
    int coarse;        /* MGLDG_COARSE_CHOLESKY | MGLDG_COARSE_RELAXATION  */
    int npre, npost;   /* reference default 3, 3                            */
    int restriction;   /* MGLDG_RAT | MGLDG_RATRAT                          */
    int precision;     /* MGLDG_FP64 | MGLDG_MIXED: which V-cycles to build */
    int coarse_dense_max; /* 0: off.  Otherwise the V-cycle below the first level
                             (not the finest) with at most this many dofs is
                             formed once at setup as a dense FP64 matrix (the
                             reference's Multigrid::vcycle_matrix, multigrid.h:
                             386-400, on that level) and applied as one GEMV:
                             the same linear map in one launch. <= 8192. */
    int pool_off;      /* 0: where at most a quarter of a level's blocks are
                          distinct (quadtree levels), the single-RHS kernels
                          read them from a pool of the distinct blocks (the
                          same bits, so the same results, with a fraction of
                          the HBM traffic); 1: always the full block arrays */
    int row_tiles;     /* pooled levels (no counterpart in the reference): the
                          single-RHS GS and operator kernels on rows grouped
                          in tiles of 16 with the same blocks, on the matrix
                          cores.  0: on levels with at least 2048 tiles (32k
                          block rows) in FP64, 256 tiles (4k rows) in FP32;
                          1: on every pooled level; 2: never.  The same
                          operator; summation order differs in the last bits */
} mgldg_mg_params;

/* InterpolationHierarchy<N,P>(qt) + Multigrid<N,P>(ops, hierarchy, params)
 * (interpolationhierarchy.h:30-82, multigrid.h:291-301, Level 83-171).
 * The symbolic sparsity of every Galerkin product is formed on the host; the
 * numeric block triple products, block inversions and colouring-ordered
 * operator layouts are built on the GPU. */
/* On a simplex problem: the library's own agglomeration (below) followed by
 * the agglomerated hierarchy. */
int mgldg_mg_create(mgldg_problem pr, const mgldg_mg_params* prm, mgldg_mg* out);
/* Multigrid<N,P1,P2,...>(ops, InterpolationHierarchy<N,P1,P2,...>(qt), params)
 * (src/cartesian/multigrid/multigrid.h:435-639,
 * interpolationhierarchy.h:98-154): p-levels of the polynomial degrees
 * degrees[0] > degrees[1] > ... (each below the problem's, >= 0; T on every
 * element evaluates the coarser degree's Lagrange basis at the finer degree's
 * GL nodes), then with hcoarsen = 1 the quadtree h-hierarchy at the last
 * degree, with hcoarsen = 0 none (StopCoarsening: the last p-level is the
 * coarsest).  Quadtree problems only. */
int mgldg_mg_create_p(mgldg_problem pr, const mgldg_mg_params* prm, int ndeg, const int* degrees, int hcoarsen,
                      mgldg_mg* out);
/* Agglomeration<N,P>(mesh, maxAgglom) (src/unstructured/mesh/agglomeration.h:21-34)
 * with the library's multilevel agglomerator on each level's dual graph
 * (mgldg_mgridgen below, minsize 1, maxsize 2^N, as agglomeration.h:96-116),
 * stopping at one agglomerate, max_levels partitions, or a level that no
 * longer shrinks.  Outputs: *nlevels partitions; nparts[l] agglomerates made
 * by partition l; parts[l*ne + i] = agglomerate of level-l element i
 * (capacity ne * max_levels; may be NULL). */
int mgldg_agglomerate(mgldg_problem pr, int max_levels, int* nlevels, int* nparts, int* parts);
/* MGridGen(nvtxs, xadj, vvol, vsurf, adjncy, adjwgt, minsize, maxsize,
 * options, &nmoves, &nparts, part) (lib/ParMGridGen-1.0/MGridGen/mgridgen.h:5-6,
 * Lib/mgridgen.c:18-49), the call of src/unstructured/mesh/agglomeration.h:
 * 103-116 and 196-209, by the library's own multilevel agglomerator: rounds
 * of aspect-ratio matching, then multi-objective k-way refinement with
 * component splitting, merging and attraction of clusters under minsize.
 * Same arguments and meaning: the dual graph as CSR (xadj[nvtxs+1],
 * adjncy / adjwgt = neighbour and shared face area), vvol = volume and
 * vsurf = boundary-face area per vertex; out: *nparts clusters, part[v].
 * options = {CType, RType, dbglvl, dim}; only the reference's CType 4 and
 * RType 6 are built (others: MGLDG_ERR_UNSUPPORTED).  Deterministic (own
 * random orders, not drand48's), so partitions are of MGridGen's kind and
 * quality but not its bits.  *nmoves (may be NULL) is set to 0.  Host only;
 * errors through mgldg_last_error(NULL). */
int mgldg_mgridgen(int nvtxs, const int* xadj, const double* vvol, const double* vsurf, const int* adjncy,
                   const double* adjwgt, int minsize, int maxsize, const int* options, int* nmoves, int* nparts,
                   int* part);
/* InterpolationHierarchy<N,P>(agglomeration) + multigrid for a simplex problem
 * (src/unstructured/multigrid/interpolationhierarchy.h:88-204; the reference
 * has no unstructured Multigrid and its constructor does not compile, :96,150,217).
 * parts[l*ne + i], l < nlev: level-l element i -> level-(l+1) agglomerate
 * (level 0 = the mesh, ne its element count).  Level 0 blocks are P_p simplex
 * bases; coarse levels are tensor Q_p Gauss-Lobatto bases on the agglomerates'
 * bounding boxes (of their vertices). */
int mgldg_mg_create_agglomerated(mgldg_problem pr, const mgldg_mg_params* prm, int nlev,
                                 const int* parts, mgldg_mg* out);
/* info[0]=levels, info[1]=ndof, info[2]=npl; then per level l < 16:
 * info[3+2l]=block rows, info[4+2l]=stored blocks */
int mgldg_mg_info(mgldg_mg mg, long* info, int ninfo);
/* block rows, block size and stored blocks of level l */
int mgldg_mg_level(mgldg_mg mg, int level, int* nb, int* B, long* nnzb);
/* No reference counterpart.  Kernel family of the batched (nrhs > 1) block
 * applications and smoothers: mode -1 (default) = picked per level by the
 * library's autotuner when a batch size is first used; 0 = vector kernels;
 * 1, 2, 4 = MFMA tiles of 16, 32, 64 columns on levels whose block size has
 * them (6, 8, 9, 10, 16; up to 2 tiles for 27).  A wavefront reads and writes
 * its tile columns as one vector per row, so a tile count that does not
 * divide the batch size runs as the largest of 2, 1 that does.  Takes effect
 * at the next batch-size change. */
int mgldg_mg_batch_kernels(mgldg_mg mg, int mode);
/* the choices in use for the current batch: plan[2*l] (operator
 * application) and plan[2*l+1] (smoother) per level, FP64 then FP32 (levels
 * entries each), 0 = vector kernels, nt = MFMA tiles of 16 x nt columns;
 * ms (may be NULL): the autotuner's timings, 4 candidates (vector, 1, 2, 4
 * tiles; 0 = not tried) x 2 operations per level and precision */
int mgldg_mg_batch_plan(mgldg_mg mg, int* plan, float* ms);
/* Export a level operator (copied back from the GPU): which 0=A 1=M 2=T 3+d=G_d
 * (RATRAT levels only), 6 = interpolation T_l (rows level l, cols level l+1).
 * Exported patterns are the host symbolic phase; values come from the GPU. */
int mgldg_mg_export(mgldg_mg mg, int level, int which, int* rowptr, int* cols,
                    double* blocks, long* nnzb, int* nbrows, int* nbcols);
/* no reference counterpart: the pool of level l (npool distinct A blocks,
 * ndpool distinct D^-1 blocks; 0, 0 when the level is not pooled) */
int mgldg_mg_level_pool(mgldg_mg mg, int level, int* npool, int* ndpool);
/* no reference counterpart: the row tiles of level l (mgldg_mg_params
 * row_tiles): *ntiles tiles of 16 rows over the level (0: none built),
 * *width their table width, *used bit 0 (FP64) / bit 1 (FP32) set when the
 * single-RHS GS and operator
 * kernels of the level run on them */
int mgldg_mg_level_tiles(mgldg_mg mg, int level, int* ntiles, int* width, int* used);
/* no reference counterpart: the cost model of the single-RHS MG-PCG in the
 * given precision (flexible: the mixed form's r.z_old term) -- one
 * preconditioner V-cycle's kernel launches and algorithmic HBM bytes (every
 * stream counted once, as the launch reads it), the same for one whole CG
 * iteration (V-cycle, operator product, vector updates and reductions), and
 * the kernel nodes of the V-cycle's captured graph (-1 before the first
 * graph solve) */
int mgldg_mg_solve_cost(mgldg_mg mg, int precision, int flexible, int* vc_launches, double* vc_bytes,
                        int* it_launches, double* it_bytes, int* graph_nodes);
/* colouring used by the multicolour Gauss-Seidel on level l */
int mgldg_mg_colors(mgldg_mg mg, int level, int* colors, int* ncolors);
int mgldg_mg_destroy(mgldg_mg mg);

/* ======================= solve ======================= */
typedef struct {
    double tol;        /* reference default 1e-8 (dg_tests.cpp:43)            */
    int maxit;         /* reference default 200                              */
    int criterion;     /* MGLDG_CRIT_REFERENCE | MGLDG_CRIT_TRUE              */
    int flexible;      /* 1: Polak-Ribiere beta (for the FP32 preconditioner) */
    int precision;     /* MGLDG_FP64 | MGLDG_MIXED                            */
    int use_graph;     /* 1: replay the V-cycle from a captured hipGraph       */
} mgldg_solve_params;

typedef struct {
    int iterations;    /* PCG loop bodies executed (krylov.h:41-65)  */
    int status;        /* MGLDG_CONVERGED / MAXIT / BREAKDOWN / NONFINITE */
    double relres_true;/* ||b - A x||_2 / ||b||_2 in FP64 at exit      */
    double relres_prec;/* sqrt(r^T z / r0^T z0) at exit                */
    double seconds;    /* device time of the solve                     */
} mgldg_solve_stats;

/* pcg(A, b, x, precon=V-cycle, tol, maxit) (krylov.h:22-68 with the
 * preconditioner lambda of tests/cartesian/dg_tests.cpp:300-305).
 * Device pointers, x is the initial guess on entry. nrhs right-hand sides are
 * solved together ([ndof][nrhs] layout); stats has nrhs entries. */
int mgldg_solve_device(mgldg_mg mg, const double* b, double* x, int nrhs,
                       const mgldg_solve_params* prm, mgldg_solve_stats* stats);
/* same with host buffers (copies in and out of HBM) */
int mgldg_solve(mgldg_mg mg, const double* b, double* x, int nrhs,
                const mgldg_solve_params* prm, mgldg_solve_stats* stats);

/* ======================= adjoint ======================= */
/* No reference counterpart.  For a loss L(u) with u = A^{-1} b(f, g):
 * lambda = A^{-1} dL/du by the same MG-PCG (A is symmetric), then
 * dL/df = M lambda and dL/dg = B_g^T lambda.  Device pointers; any of
 * lambda/dLdf/dLdg may be NULL.  Layout [ndof][nrhs], dLdg [nqb][nrhs]. */
int mgldg_adjoint_device(mgldg_mg mg, const double* dLdu, double* lambda, double* dLdf,
                         double* dLdg, int nrhs, const mgldg_solve_params* prm,
                         mgldg_solve_stats* stats);
/* No reference counterpart.  The maps of the adjoint alone, for adjoint
 * states lambda solved elsewhere (e.g. batched beside forward solves):
 * dL/df = M lambda, dL/dg = B_g^T lambda (either output may be NULL). */
int mgldg_adjoint_maps_device(mgldg_mg mg, const double* lambda, double* dLdf, double* dLdg, int nrhs);
/* b = M f + B_g g on the GPU (device pointers; f or g may be NULL) */
int mgldg_rhs_device(mgldg_mg mg, const double* f, const double* g, int nrhs, double* b);

/* ======================= unit-parity hooks ======================= */
/* y = A_l x on level l (multiply_mv, sparseblockmatrix.tpp:288-301) */
int mgldg_spmv_device(mgldg_mg mg, int level, int precision, const void* x, void* y, int nrhs);
/* nsweeps relaxations on level l (Level::relax, multigrid.h:202-216), x in/out */
int mgldg_smooth_device(mgldg_mg mg, int level, int precision, int nsweeps, int reverse,
                        const void* b, void* x, int nrhs);
/* x = V(b) from x = 0 (one preconditioner application, dg_tests.cpp:300-305) */
int mgldg_vcycle_device(mgldg_mg mg, int precision, const double* b, double* x, int nrhs);
/* The driver's stationary solver `--solver vcycles` (dg_tests.cpp:309-332):
 * up to ncycle V-cycles (Multigrid::vcycle, multigrid.h:361-383) from the
 * iterate in x (device, in/out).  rnorm (host, ncycle + 1 doubles, or NULL)
 * receives ||b - A x|| before and after each cycle; *cycles the number run.
 * tol > 0 stops once ||b - A x|| <= tol ||b|| (tol = 0: exactly ncycle, as
 * the driver).  precision MGLDG_MIXED: x += V(b - A x) with an FP32 V-cycle
 * and the residual and update in FP64. */
int mgldg_vcycles_device(mgldg_mg mg, int precision, const double* b, double* x, int ncycle, double tol, double* rnorm,
                         int* cycles);
/* Function::meanZero (src/cartesian/dg/function.h:156-170) on the device:
 * subtract from each of the nrhs columns of v ([n][nrhs]) its mean; the
 * driver applies it to the right-hand side of Neumann and periodic problems
 * (dg_tests.cpp:282-286). */
int mgldg_mean_zero_device(mgldg_ctx ctx, long n, double* v, int nrhs);
/* No counterpart in the reference.  mgldg_mean_zero_device for callers that
 * cannot allocate device memory (code between graph captures, serving loops,
 * harnesses that forbid hipMalloc): the same kernels and results, with the
 * reduction's partial sums in a static device array of the library (2.1 MB,
 * held by each device from the first context on it to the end of the
 * process) instead of a buffer grown on demand.  Calls no hipMalloc / hipFree.
 * Limit: nrhs <= 256 (MGLDG_ERR_ARG beyond).  Every call on one device shares
 * that array: calls that overlap on one device (other host threads or other
 * streams) race; calls in sequence are independent (nothing is carried from
 * one call to the next).  Returns after the work is complete. */
int mgldg_mean_zero_device_noalloc(mgldg_ctx ctx, long n, double* v, int nrhs);

/* ======================= interop file formats ======================= */
/* Text formats of the reference's dumps and meshes.  Entries taking no handle
 * report errors through mgldg_last_error(NULL).  Numbers are printed at 17
 * significant digits (the reference's precision(max_digits10)), so every
 * double round-trips. */

/* SparseBlockMatrix::write (src/cartesian/util/sparseblockmatrix.tpp:145-164,
 * src/unstructured/util/sparseblockmatrix.tpp:145-164): MatrixMarket
 * "coordinate real general", header "rows cols nnz" with nnz = nnzb*br*bc,
 * one 1-based line per entry of every stored block (explicit zeros
 * included).  Blocks go in row order, columns ascending (the reference walks
 * its hash map, whose order is unspecified).  blocks: nnzb row-major br x bc. */
int mgldg_bsr_write_mm(const char* path, int nbrows, int nbcols, int br, int bc, const int* rowptr,
                       const int* cols, const double* blocks);
/* the same for a problem operator (which as mgldg_problem_export) */
int mgldg_problem_write_mm(mgldg_problem pr, int which, const char* path);
/* the same for a level operator or interpolation (which as mgldg_mg_export) */
int mgldg_mg_write_mm(mgldg_mg mg, int level, int which, const char* path);
/* Function::write (src/cartesian/dg/function.h:173-190,
 * src/unstructured/dg/function.h:159-176): header "P N ne" (P = p + 1),
 * then per element one line "x_1 .. x_N value" per node in storage order
 * and an empty line.  xyz[ne*npl*N] node coordinates, u[ne*npl] values. */
int mgldg_fun_write(const char* path, int P, int N, int ne, int npl, const double* xyz, const double* u);
/* the same with the problem's nodes (mgldg_problem_nodes) and u[ndof] */
int mgldg_problem_write_fun(mgldg_problem pr, const double* u, const char* path);
/* LDGPoisson::dump (src/cartesian/dg/ldgpoisson.h:100-109, unstructured
 * ldgpoisson.h:100-109): dir/A.mtx, M.mtx, T.mtx, G0.mtx .. G{N-1}.mtx and,
 * when rhs (host, ndof) is not NULL, dir/rhs.fun.  dir must exist. */
int mgldg_problem_dump(mgldg_problem pr, const double* rhs, const char* dir);
/* The .mesh layout (lib/distmesh/writemesh.m:17-28, read back by
 * mgldg_wireframe_read and Wireframe<N>(file), wireframe.h:22-68): "np nt",
 * an empty line, np lines of N coordinates, an empty line, nt lines of N+1
 * 1-based vertex ids.  digits <= 0: coordinates as "%f" (writemesh.m);
 * digits in 1..17: that many significant digits.  simplices are 0-based. */
int mgldg_mesh_write(const char* path, int N, int np, const double* points, int nt, const int* simplices,
                     int digits);
int mgldg_wireframe_write(mgldg_wireframe wf, const char* path, int digits);
/* Multigrid::vcycle_matrix (src/cartesian/multigrid/multigrid.h:386-400,
 * :523-537): V[:, i] = one V-cycle from zero applied to the unit vector e_i,
 * for every fine dof i; V host, row-major ndof x ndof.  The columns run
 * through the batched V-cycle 32 at a time. */
int mgldg_mg_vcycle_matrix(mgldg_mg mg, int precision, double* V);
/* A dense row-major matrix as the driver writes V.dat (dg_tests.cpp:337-347,
 * the Eigen stream operator under precision 17): one line per row, every
 * coefficient right-aligned to the widest, single spaces between them. */
int mgldg_dense_write(const char* path, int rows, int cols, const double* a);

/* ======================= measurement hooks ======================= */
/* No reference counterpart (the reference only prints a Timer).  Times
 * `nrep` smoother sweeps (all colours) or SpMVs on level l with HIP events on
 * the context stream and reports the average duration of one kernel launch
 * (ms), the launches timed and the algorithmic HBM bytes one launch moves:
 *   GS launch: off-diagonal blocks and column indices of its rows, their
 *   inverse diagonal blocks and row descriptors, b, x (read) and x (write)
 *   of its rows, and x of their distinct neighbours.
 *   Jacobi sweep: all blocks + inverse diagonal blocks + indices + b, x, x_out.
 *   SpMV: all blocks + indices + x (read) + y (write).
 * On a pooled level (mgldg_mg_level_pool) the blocks count once per distinct
 * pool block a launch touches and the indices as (column, pool index) pairs. */
int mgldg_profile_relax(mgldg_mg mg, int level, int precision, int nrep, double* ms_per_launch,
                        double* bytes_per_launch, int* launches);
/* The same for one part of a multicolour GS sweep: part 0 = whole sweeps
 * (as mgldg_profile_relax), 1 = the one-colour launches only, 2 = the launch
 * that does the last two colours together (levels where it exists). */
int mgldg_profile_relax_part(mgldg_mg mg, int level, int precision, int nrep, int part, double* ms_per_launch,
                             double* bytes_per_launch, int* launches);
int mgldg_profile_spmv(mgldg_mg mg, int level, int precision, int nrep, double* ms_per_launch,
                       double* bytes_per_launch);
/* problem setup: t[0] host symbolic phase (mesh, patterns, assembly plan)
 * and t[1] device phase (upload, assembly, A = T + sum G^T M G, G_d^T) in
 * seconds; t[2] the assembly kernel's duration (ms, HIP events) and t[3] its
 * algorithmic HBM bytes (every block of M, T, G_d written once, plan data
 * read once) */
int mgldg_problem_setup_times(mgldg_problem pr, double* t);
/* setup timings: t[0] host assembly, t[1] device assembly (A), t[2] hierarchy
 * host symbolic, t[3] hierarchy device numeric (seconds) */
int mgldg_setup_times(mgldg_mg mg, double* t);

#ifdef __cplusplus
}
#endif
#endif
