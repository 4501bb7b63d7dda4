/*
 * oracle.h — CPU FP64 restatement of multigrid-ldg's Cartesian LDG-Poisson
 * setup + MG-preconditioned CG solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libmgldg.so) never links or calls it.
 *
 * Every function follows the reference file:line cited beside it
 * (paths relative to the reference repository root).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   basis tables        -> literal tables src/cartesian/dg/master.h:150-214
 *                          (golden fixture tests/golden/master_tables.json)
 *   BSR SpMV            -> tests/cartesian/matrix_tests.cpp:31-52, y=[22,30,37,38]
 *   assembly            -> symmetry, A*1 = 0 without Dirichlet faces, LDG
 *                          patch test (g=x+y reproduced), O(h^{p+1}) convergence
 *   RATRAT coarsening   -> equals LDG rediscretisation on the coarse layer
 *   RAT                 -> dense T^T A T
 *   smoothers / V-cycle -> dense brute force (block triangular solves)
 *   PCG                 -> numpy.linalg.solve
 */
#ifndef MGLDG_ORACLE_H
#define MGLDG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* size function: writes N widths for the point x (quadtree.tpp:15 semantics) */
typedef void (*orc_hfn)(const double* x, double* hout, void* user);

/* ---- basis tables (src/cartesian/dg/master.h) ---- */
int orc_gl_nodes(int P, double* out);            /* P nodes on [0,1]            */
int orc_lagrange_denom(int P, double* out);      /* P denominators              */
int orc_mass1d(int P, double* out);              /* P*P row-major               */
int orc_invmass1d(int P, double* out);           /* P*P row-major               */
int orc_diff1d(int P, double* out);              /* P*P, out[i*P+j] = l_j'(x_i) */
int orc_gauss(int Q, double* nodes, double* weights); /* Gauss-Legendre on [0,1] */
int orc_lagrange_eval(int P, double x, double* out);

/* ---- quadtree (src/cartesian/mesh/quadtree.{h,tpp}) ---- */
void* orc_quadtree_new(int N, const double* lo, const double* hi, orc_hfn h, void* user,
                       int rapid, int periodic, int balance);
void  orc_quadtree_free(void* qt);
int   orc_quadtree_info(void* qt, int* nnodes, int* nlevels);
int   orc_quadtree_layer(void* qt, int l, int* ids);          /* returns count */
int   orc_quadtree_node(void* qt, int id, int* ilevel_height_parent, double* lohi);

/* ---- LDG problem (mesh + assembly), src/cartesian/dg/ldgpoisson.tpp ---- */
/* bctype[2N]: per boundary tag -(2d+1) / -(2d+2) -> index 2d / 2d+1; 0=Dirichlet 1=Neumann */
void* orc_problem_new(void* qt, int p, const int* bctype, double tau0, double tauD);
void  orc_problem_free(void* pr);
/* sizes[0..5] = ne, nf, nbfaces, b (npl), nqb (boundary qpts total), ndof */
int   orc_problem_sizes(void* pr, int* sizes);
int   orc_problem_nodes(void* pr, double* xyz);               /* ndof*N   */
int   orc_problem_bc_qpoints(void* pr, double* xyz, int* tag); /* nqb*N, nqb */
int   orc_problem_faces(void* pr, int* lr_dim_canon, double* normal); /* nf*4, nf */
/* b = rhs_bc(g) + M f ;  f nodal (ndof), g at boundary quadrature points (nqb) */
int   orc_problem_rhs(void* pr, const double* f, const double* g, double* b);
/* transpose of the g -> b map: dg[nqb] = B_g^T lam */
int   orc_problem_rhs_g_adjoint(void* pr, const double* lam, double* dg);
/* which: 0=A 1=M 2=T 3..5=G_d.  Returns nnzb.  Outputs CSR sorted by column,
   blocks row-major b*b.  Any output pointer may be NULL. */
long  orc_problem_export(void* pr, int which, int* rowptr, int* cols, double* blocks);

/* ---- multigrid (src/cartesian/multigrid/{interpolationhierarchy,multigrid,krylov}.h) ---- */
typedef struct {
    int relaxation;   /* 0 Jacobi, 1 lexicographic GS (reference), 2 multicolour GS */
    double omega;
    int coarse;       /* 0 Cholesky (LDLT) of the coarsest level, 1 relaxation */
    int npre, npost;
    int restriction;  /* 0 RAT, 1 RATRAT */
} orc_mg_params;

void* orc_mg_new(void* pr, const orc_mg_params* prm);
void  orc_mg_free(void* mg);
int   orc_mg_nlevels(void* mg);
int   orc_mg_level_dims(void* mg, int level, int* nb, int* b);   /* block rows, block size */
/* which: 0=A 1=M 2=T 3..5=G_d of level l ; 6 = interpolation T_l (level l -> l+1 cols) */
long  orc_mg_export(void* mg, int level, int which, int* rowptr, int* cols, double* blocks,
                    int* nbrows, int* nbcols);
int   orc_mg_colors(void* mg, int level, int* colors);        /* returns ncolors */
/* one preconditioner application: x = V(b) with x0 = 0 (dg_tests.cpp:300-305) */
int   orc_mg_vcycle(void* mg, const double* b, double* x);
int orc_mg_vcycles(void* mg, const double* b, double* x, int ncycle, double* rnorm);
void* orc_mg_new_hp(void* problem, const orc_mg_params* prm, int ndeg, const int* deg, int hcoarsen);
/* nsweeps relaxations on level l (multigrid.h:202-216) with given b, x in/out */
int   orc_mg_relax(void* mg, int level, int nsweeps, int reverse, const double* b, double* x);
/* y = A_l x */
int   orc_mg_spmv(void* mg, int level, const double* x, double* y);

/* PCG (krylov.h:22-68).  criterion 0: reference r^T z test; 1: recurrence
   ||r||/||b|| <= tol then confirmed on the true residual.  flexible: Polak-Ribiere beta.
   mg may be NULL -> identity preconditioner (krylov.h:78-82).
   stats[0]=iterations, stats[1]=true relres, stats[2]=sqrt(r^T z / r0^T z0) ;
   hist (may be NULL) receives ||r|| per iteration (maxit entries). */
/* pr: any problem handle (Cartesian or simplex) */
int   orc_pcg(void* pr, void* mg, const double* b, double* x, double tol, int maxit,
              int criterion, int flexible, double* stats, double* hist);

/* ---- unstructured simplex path (src/unstructured/{mesh,dg,multigrid}) ---- */
/* Koornwinder nodal master element Master<N,P=p+1>: returns npl; nodes npl*N,
   mass npl*npl, diff N*npl*npl (any may be NULL) */
int   orc_sx_master(int N, int P, double* nodes, double* mass, double* diff);
/* face rule of an N-D simplex mesh (Quadrature<N-1,10>): returns nq; x nq*(N-1), w nq */
int   orc_sx_face_rule(int N, double* x, double* w);
/* mesh + LDG assembly (mesh.h:185-298, ldgpoisson.tpp:14-99).  tris 0-based
   (N+1 per simplex); vmask[v] bit k = boundary function k holds at vertex v
   (nbnd functions; 0 = one boundary tag); bctype[max(nbnd,1)] 0=Dirichlet 1=Neumann */
void* orc_sx_problem_new(int N, int np, const double* pts, int nt, const int* tris, int nbnd, const int* vmask,
                         int p, const int* bctype, double tau0, double tauD);
void  orc_sx_problem_free(void* pr);
/* sizes[0..6] = ne, nf, nbf, npl, nqb, ndof, nq */
int   orc_sx_sizes(void* pr, int* sizes);
int   orc_sx_nodes(void* pr, double* xyz);
int   orc_sx_faces(void* pr, int* lr, double* normal, int* verts);
int   orc_sx_bc_qpoints(void* pr, double* xyz, int* tag);
int   orc_sx_lift(void* pr, int face, double* L, double* LL, double* RR, double* RL);
int   orc_sx_rhs(void* pr, const double* f, const double* g, double* b);
int   orc_sx_rhs_g_adjoint(void* pr, const double* lam, double* dg);
long  orc_sx_export(void* pr, int which, int* rowptr, int* cols, double* blocks);
/* dual graph handed to MGridGen (agglomeration.h:57-93 for part = NULL, :133-185
   for the agglomerates of part); returns the number of directed edges */
int   orc_sx_agglom_graph(void* pr, const int* part, int nparts, int* xadj, int* adjncy,
                          double* adjwgt, double* vvol, double* vsurf);
/* agglomerated MG (interpolationhierarchy.h:88-204): parts[l] maps the
   elements of level l to nparts[l] agglomerates; usable with orc_mg_* and orc_pcg */
void* orc_sx_mg_new(void* pr, int nlev, const int* const* parts, const int* nparts, const orc_mg_params* prm);

/* generic BSR helpers used by known-answer tests (sparseblockmatrix.tpp:288-301) */
int   orc_bsr_spmv(int nbrows, int b, const int* rowptr, const int* cols,
                   const double* blocks, const double* x, double* y);

#ifdef __cplusplus
}
#endif
#endif
