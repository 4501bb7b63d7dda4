/*
 * cart.c — CPU FP64 restatement (the oracle) of multigrid-ldg's Cartesian
 * LDG-Poisson setup and MG-preconditioned CG solve.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Never linked by the product.
 *
 * Reference files restated here (paths relative to the reference root):
 *   src/cartesian/dg/master.h            basis tables (closed forms, pinned
 *                                        against the literal tables :150-214)
 *   src/cartesian/mesh/quadtree.{h,tpp}  tree, layers, neighbour search
 *   src/cartesian/mesh/mesh.h            elements, faces, lift matrices
 *   src/cartesian/dg/ldgpoisson.{h,tpp}  LDG assembly and RHS
 *   src/cartesian/util/sparseblockmatrix.tpp  block-sparse arithmetic
 *   src/cartesian/multigrid/interpolationhierarchy.h:21-96
 *   src/cartesian/multigrid/multigrid.h:24-433
 *   src/cartesian/multigrid/krylov.h:22-82
 * Additions that the reference does not have (documented where they occur):
 *   2:1 balancing of the quadtree, multicolour block Gauss-Seidel (greedy
 *   colouring), a true-residual stopping criterion and flexible (Polak-Ribiere)
 *   beta for PCG, a dense LDLT coarse solve when the coarsest level has more
 *   than one block.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define MAXP 8
#define QPTS 10 /* src/cartesian/util/common.h:15 */

static void* xmalloc(size_t n) { void* p = malloc(n ? n : 1); if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); } return p; }
static void* xcalloc(size_t n, size_t s) { void* p = calloc(n ? n : 1, s ? s : 1); if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); } return p; }
static void* xrealloc(void* p, size_t n) { p = realloc(p, n ? n : 1); if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); } return p; }

static int ipow_(int b, int e) { int r = 1; while (e-- > 0) r *= b; return r; }

/* ======================================================================
 * Dense helpers (row-major)
 * ====================================================================== */
/* C[m x n] (+)= alpha * op(A) * op(B);  op(A) is m x k */
static void gemm(int ta, int tb, int m, int n, int k, double alpha,
                 const double* A, int lda, const double* B, int ldb,
                 double beta, double* C, int ldc)
{
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < n; ++j) C[i*ldc+j] = (beta == 0.0) ? 0.0 : beta * C[i*ldc+j];
        for (int l = 0; l < k; ++l) {
            double a = ta ? A[l*lda+i] : A[i*lda+l];
            if (a == 0.0) continue;
            a *= alpha;
            if (!tb) { const double* Br = B + l*ldb; double* Cr = C + i*ldc; for (int j = 0; j < n; ++j) Cr[j] += a * Br[j]; }
            else     { for (int j = 0; j < n; ++j) C[i*ldc+j] += a * B[j*ldb+l]; }
        }
    }
}

/* LDLT with diagonal pivoting (the factorisation behind Eigen::LDLT used at
 * multigrid.h:94,169,190 and ldgpoisson.h:25).  F holds L (unit lower) and D on
 * the diagonal; perm the symmetric permutation. */
typedef struct { int n; double* F; int* perm; } ldlt_t;

static void ldlt_factor(ldlt_t* f, const double* A, int n)
{
    f->n = n;
    f->F = (double*)xmalloc(sizeof(double)*n*n);
    f->perm = (int*)xmalloc(sizeof(int)*n);
    double* W = f->F;
    memcpy(W, A, sizeof(double)*n*n);
    for (int i = 0; i < n; ++i) f->perm[i] = i;
    for (int k = 0; k < n; ++k) {
        int piv = k; double best = fabs(W[k*n+k]);
        for (int i = k+1; i < n; ++i) if (fabs(W[i*n+i]) > best) { best = fabs(W[i*n+i]); piv = i; }
        if (piv != k) {
            for (int j = 0; j < n; ++j) { double t = W[k*n+j]; W[k*n+j] = W[piv*n+j]; W[piv*n+j] = t; }
            for (int i = 0; i < n; ++i) { double t = W[i*n+k]; W[i*n+k] = W[i*n+piv]; W[i*n+piv] = t; }
            int t = f->perm[k]; f->perm[k] = f->perm[piv]; f->perm[piv] = t;
        }
        double d = W[k*n+k];
        for (int i = k+1; i < n; ++i) {
            double lik = (d != 0.0) ? W[i*n+k] / d : 0.0;
            for (int j = k+1; j <= i; ++j) W[i*n+j] -= lik * W[j*n+k];
        }
        for (int i = k+1; i < n; ++i) W[i*n+k] = (d != 0.0) ? W[i*n+k] / d : 0.0;
        for (int i = k+1; i < n; ++i) for (int j = i+1; j < n; ++j) W[i*n+j] = W[j*n+i];
    }
}

static void ldlt_free(ldlt_t* f) { free(f->F); free(f->perm); f->F = NULL; f->perm = NULL; }

/* x = A^{-1} b (b and x may alias).  Zero pivots give zero components, as in
 * Eigen's LDLT::solve. */
static void ldlt_solve(const ldlt_t* f, const double* b, double* x)
{
    int n = f->n; double y[1024]; double* t = (n <= 1024) ? y : (double*)xmalloc(sizeof(double)*n);
    const double* W = f->F;
    for (int i = 0; i < n; ++i) t[i] = b[f->perm[i]];
    for (int i = 0; i < n; ++i) { double s = t[i]; for (int j = 0; j < i; ++j) s -= W[i*n+j]*t[j]; t[i] = s; }
    for (int i = 0; i < n; ++i) { double d = W[i*n+i]; t[i] = (fabs(d) > 2.2250738585072014e-308) ? t[i]/d : 0.0; }
    for (int i = n-1; i >= 0; --i) { double s = t[i]; for (int j = i+1; j < n; ++j) s -= W[j*n+i]*t[j]; t[i] = s; }
    for (int i = 0; i < n; ++i) x[f->perm[i]] = t[i];
    if (t != y) free(t);
}

/* columnwise solve of a b x c block: X = A^{-1} B */
static void ldlt_solve_mat(const ldlt_t* f, const double* B, int c, double* X)
{
    int n = f->n; double col[1024];
    for (int j = 0; j < c; ++j) {
        for (int i = 0; i < n; ++i) col[i] = B[i*c+j];
        ldlt_solve(f, col, col);
        for (int i = 0; i < n; ++i) X[i*c+j] = col[i];
    }
}

/* ======================================================================
 * 1-D basis (src/cartesian/dg/master.h:12-214), computed from closed forms
 * ====================================================================== */
static void legendre(int n, long double x, long double* P, long double* dP)
{
    long double p0 = 1.0L, p1 = x;
    if (n == 0) { *P = 1.0L; *dP = 0.0L; return; }
    for (int k = 2; k <= n; ++k) { long double p2 = ((2*k-1)*x*p1 - (k-1)*p0)/k; p0 = p1; p1 = p2; }
    *P = p1;
    *dP = (fabsl(x) < 1.0L) ? n*(x*p1 - p0)/(x*x - 1.0L) : 0.5L*n*(n+1)*((x > 0) ? 1.0L : ((n % 2) ? 1.0L : -1.0L));
}

/* Gauss-Lobatto points on [0,1] (GaussLobatto<P>::nodes, master.h:174-179):
 * P=1 -> {1/2}; else {0, roots of P'_{P-1} mapped, 1}. */
int orc_gl_nodes(int P, double* out)
{
    if (P < 1 || P > MAXP) return -1;
    if (P == 1) { out[0] = 0.5; return 0; }
    int n = P - 1;
    long double x[MAXP];
    x[0] = -1.0L; x[n] = 1.0L;
    for (int k = 1; k < n; ++k) {
        long double z = -cosl(3.14159265358979323846264338327950288L * k / n);
        for (int it = 0; it < 100; ++it) {
            long double Pn, dPn; legendre(n, z, &Pn, &dPn);
            long double d2 = (2.0L*z*dPn - n*(n+1)*Pn)/(1.0L - z*z);
            long double dz = dPn / d2; z -= dz;
            if (fabsl(dz) < 1e-21L) break;
        }
        x[k] = z;
    }
    for (int k = 0; k < P; ++k) {
        long double a = (x[k] + 1.0L)/2.0L, b = 1.0L - (x[n-k] + 1.0L)/2.0L;
        out[k] = (double)((a + b)/2.0L);
    }
    out[0] = 0.0; out[P-1] = 1.0;
    if (P % 2 == 1) out[P/2] = 0.5;
    return 0;
}

/* Gauss-Legendre on [0,1] (Quadrature<Q>, master.h:188-214) */
int orc_gauss(int Q, double* nodes, double* weights)
{
    if (Q < 1 || Q > 64) return -1;
    long double xs[64], ws[64];
    for (int k = 0; k < Q; ++k) {
        long double z = -cosl(3.14159265358979323846264338327950288L * (k + 0.75L)/(Q + 0.5L));
        long double Pn = 0, dPn = 0;
        for (int it = 0; it < 100; ++it) {
            legendre(Q, z, &Pn, &dPn);
            long double dz = Pn/dPn; z -= dz;
            if (fabsl(dz) < 1e-21L) break;
        }
        legendre(Q, z, &Pn, &dPn);
        xs[k] = z; ws[k] = 2.0L/((1.0L - z*z)*dPn*dPn);
    }
    for (int k = 0; k < Q; ++k) {
        long double a = (xs[k] + 1.0L)/2.0L, b = 1.0L - (xs[Q-1-k] + 1.0L)/2.0L;
        nodes[k] = (double)((a + b)/2.0L);
        weights[k] = (double)((ws[k] + ws[Q-1-k])/4.0L);
    }
    if (Q % 2 == 1) nodes[Q/2] = 0.5;
    return 0;
}

/* LagrangePoly<P>::denom (master.h:181-186): 1/prod_{j!=i}(x_i - x_j) */
int orc_lagrange_denom(int P, double* out)
{
    double x[MAXP]; if (orc_gl_nodes(P, x)) return -1;
    for (int i = 0; i < P; ++i) {
        long double d = 1.0L;
        for (int j = 0; j < P; ++j) if (j != i) d *= (long double)x[i] - (long double)x[j];
        out[i] = (double)(1.0L/d);
    }
    return 0;
}

/* LagrangePoly<P>::eval(double) (master.h:78-95): left and right partial products */
int orc_lagrange_eval(int P, double x, double* out)
{
    double nd[MAXP], den[MAXP];
    if (orc_gl_nodes(P, nd) || orc_lagrange_denom(P, den)) return -1;
    double alpha = 1.0;
    out[0] = den[0];
    for (int i = 0; i < P-1; ++i) { alpha *= x - nd[i]; out[i+1] = alpha*den[i+1]; }
    double beta = 1.0;
    for (int i = P-1; i > 0; --i) { beta *= x - nd[i]; out[i-1] *= beta; }
    return 0;
}

static long double lag_ld(int P, const double* nd, int i, long double x)
{
    long double v = 1.0L;
    for (int j = 0; j < P; ++j) if (j != i) v *= (x - nd[j])/((long double)nd[i] - nd[j]);
    return v;
}

/* Master<1,P>::mass (master.h:150-155): exact integrals of products */
int orc_mass1d(int P, double* out)
{
    double nd[MAXP], gx[MAXP+2], gw[MAXP+2];
    if (orc_gl_nodes(P, nd) || orc_gauss(P+1, gx, gw)) return -1;
    for (int i = 0; i < P; ++i) for (int j = 0; j < P; ++j) {
        long double s = 0.0L;
        for (int q = 0; q < P+1; ++q) s += (long double)gw[q]*lag_ld(P, nd, i, gx[q])*lag_ld(P, nd, j, gx[q]);
        out[i*P+j] = (double)s;
    }
    return 0;
}

/* Master<1,P>::invmass (master.h:158-163) */
int orc_invmass1d(int P, double* out)
{
    double m[MAXP*MAXP];
    if (orc_mass1d(P, m)) return -1;
    long double a[MAXP][2*MAXP];
    for (int i = 0; i < P; ++i) for (int j = 0; j < 2*P; ++j) a[i][j] = (j < P) ? m[i*P+j] : (j - P == i ? 1.0L : 0.0L);
    for (int k = 0; k < P; ++k) {
        int piv = k; for (int i = k+1; i < P; ++i) if (fabsl(a[i][k]) > fabsl(a[piv][k])) piv = i;
        for (int j = 0; j < 2*P; ++j) { long double t = a[k][j]; a[k][j] = a[piv][j]; a[piv][j] = t; }
        long double d = a[k][k];
        for (int j = 0; j < 2*P; ++j) a[k][j] /= d;
        for (int i = 0; i < P; ++i) if (i != k) { long double f = a[i][k]; for (int j = 0; j < 2*P; ++j) a[i][j] -= f*a[k][j]; }
    }
    for (int i = 0; i < P; ++i) for (int j = 0; j < P; ++j) out[i*P+j] = (double)a[i][P+j];
    return 0;
}

/* GaussLobatto<P>::diff (master.h:166-171): D(i,j) = l_j'(x_i) */
int orc_diff1d(int P, double* out)
{
    double nd[MAXP];
    if (orc_gl_nodes(P, nd)) return -1;
    long double w[MAXP];
    for (int i = 0; i < P; ++i) { w[i] = 1.0L; for (int j = 0; j < P; ++j) if (j != i) w[i] *= (long double)nd[i] - nd[j]; w[i] = 1.0L/w[i]; }
    for (int i = 0; i < P; ++i) {
        long double diag = 0.0L;
        for (int j = 0; j < P; ++j) if (j != i) {
            long double v = (w[j]/w[i])/((long double)nd[i] - nd[j]);
            out[i*P+j] = (double)v; diag -= v;
        }
        out[i*P+i] = (double)diag;
    }
    if (P == 1) out[0] = 0.0;
    return 0;
}

/* ======================================================================
 * N-D tensor-product basis for one (N, P)
 * ====================================================================== */
typedef struct {
    int N, P, npl, nfl;            /* nodes per element / per face */
    double nodes[MAXP], den[MAXP];
    double m1[MAXP*MAXP], im1[MAXP*MAXP], d1[MAXP*MAXP];
    double gx[QPTS], gw[QPTS];
    double* mass;                  /* npl x npl */
    double* invmass;               /* npl x npl */
    double* diff[3];               /* npl x npl per dimension */
    double* fmass;                 /* nfl x nfl, (N-1)-D mass */
} basis_t;

/* multi-index of linear index lin, last coordinate fastest (range.h:187-194) */
static void unlin(int lin, int N, int P, int* it) { for (int k = N-1; k >= 0; --k) { it[k] = lin % P; lin /= P; } }
static int lin_(const int* it, int N, int P) { int l = it[0]; for (int k = 1; k < N; ++k) l = P*l + it[k]; return l; }

static void basis_init(basis_t* B, int N, int P)
{
    memset(B, 0, sizeof(*B));
    B->N = N; B->P = P; B->npl = ipow_(P, N); B->nfl = ipow_(P, N-1);
    orc_gl_nodes(P, B->nodes); orc_lagrange_denom(P, B->den);
    orc_mass1d(P, B->m1); orc_invmass1d(P, B->im1); orc_diff1d(P, B->d1);
    orc_gauss(QPTS, B->gx, B->gw);
    int n = B->npl, it[3], jt[3];
    B->mass = (double*)xmalloc(sizeof(double)*n*n);
    B->invmass = (double*)xmalloc(sizeof(double)*n*n);
    /* tensorProduct (master.h:130-145) */
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) {
        unlin(i, N, P, it); unlin(j, N, P, jt);
        double a = 1.0, b = 1.0;
        for (int k = 0; k < N; ++k) { a *= B->m1[it[k]*P+jt[k]]; b *= B->im1[it[k]*P+jt[k]]; }
        B->mass[i*n+j] = a; B->invmass[i*n+j] = b;
    }
    /* Master<N,P>::diff(i) (master.h:45-69) */
    for (int d = 0; d < N; ++d) {
        B->diff[d] = (double*)xcalloc((size_t)n*n, sizeof(double));
        for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) {
            unlin(i, N, P, it); unlin(j, N, P, jt);
            int nz = 1; for (int k = 0; k < N; ++k) if (k != d && it[k] != jt[k]) { nz = 0; break; }
            if (nz) B->diff[d][i*n+j] = B->d1[it[d]*P+jt[d]];
        }
    }
    int nf = B->nfl;
    B->fmass = (double*)xmalloc(sizeof(double)*nf*nf);
    for (int i = 0; i < nf; ++i) for (int j = 0; j < nf; ++j) {
        unlin(i, N-1, P, it); unlin(j, N-1, P, jt);
        double a = 1.0; for (int k = 0; k < N-1; ++k) a *= B->m1[it[k]*P+jt[k]];
        B->fmass[i*nf+j] = a;
    }
}

static void basis_free(basis_t* B) { free(B->mass); free(B->invmass); for (int d = 0; d < 3; ++d) free(B->diff[d]); free(B->fmass); }

/* LagrangePoly<P>::eval(Tuple<N>) (master.h:97-116) */
static void lagrange_eval_nd(const basis_t* B, const double* x, double* out)
{
    int N = B->N, P = B->P; double g[3][MAXP]; int it[3];
    for (int j = 0; j < N; ++j) {
        double alpha = 1.0; g[j][0] = B->den[0];
        for (int i = 0; i < P-1; ++i) { alpha *= x[j] - B->nodes[i]; g[j][i+1] = alpha*B->den[i+1]; }
        double beta = 1.0;
        for (int i = P-1; i > 0; --i) { beta *= x[j] - B->nodes[i]; g[j][i-1] *= beta; }
    }
    for (int i = 0; i < B->npl; ++i) {
        unlin(i, N, P, it);
        double v = 1.0; for (int j = 0; j < N; ++j) v *= g[j][it[j]];
        out[i] = v;
    }
}

/* ======================================================================
 * Quadtree (src/cartesian/mesh/quadtree.{h,tpp})
 * ====================================================================== */
typedef struct { int level, height, parent; double lo[3], hi[3]; int child[8]; } qnode;
typedef struct { int N; int n, cap; qnode* nd; int numLevels; int rapid, periodic; } qtree;
typedef struct { int n, cap; int* v; } ivec;

static void iv_push(ivec* a, int x) { if (a->n == a->cap) { a->cap = a->cap ? 2*a->cap : 16; a->v = (int*)xrealloc(a->v, sizeof(int)*a->cap); } a->v[a->n++] = x; }

static int qt_push(qtree* t, const double* lo, const double* hi, int parent)
{
    if (t->n == t->cap) { t->cap = t->cap ? 2*t->cap : 64; t->nd = (qnode*)xrealloc(t->nd, sizeof(qnode)*t->cap); }
    qnode* q = &t->nd[t->n];
    memset(q, 0, sizeof(*q));
    for (int i = 0; i < t->N; ++i) { q->lo[i] = lo[i]; q->hi[i] = hi[i]; }
    q->parent = parent;
    for (int c = 0; c < 8; ++c) q->child[c] = -1;
    return t->n++;
}

/* create the 2^N children of id in RangeIterator<2,N> order (quadtree.tpp:18-33) */
static void qt_split(qtree* t, int id)
{
    int N = t->N;
    double mid[3];
    for (int i = 0; i < N; ++i) mid[i] = t->nd[id].lo[i] + (t->nd[id].hi[i] - t->nd[id].lo[i])/2;
    for (int c = 0; c < (1 << N); ++c) {
        double lo[3], hi[3];
        for (int i = 0; i < N; ++i) {
            int bit = (c >> (N-1-i)) & 1;
            lo[i] = t->nd[id].lo[i]; hi[i] = t->nd[id].hi[i];
            if (bit) lo[i] = mid[i]; else hi[i] = mid[i];
        }
        int next = qt_push(t, lo, hi, id);
        t->nd[id].child[c] = next;
        t->nd[next].level = t->nd[id].level + 1;
    }
}

/* Quadtree::build (quadtree.tpp:8-40): split iff every width exceeds h(center) */
static void qt_build(qtree* t, int id, int level, orc_hfn h, void* user)
{
    int N = t->N;
    if (level + 1 > t->numLevels) t->numLevels = level + 1;
    t->nd[id].level = level; t->nd[id].height = 0;
    double mid[3], hv[3] = {0, 0, 0};
    for (int i = 0; i < N; ++i) mid[i] = t->nd[id].lo[i] + (t->nd[id].hi[i] - t->nd[id].lo[i])/2;
    h(mid, hv, user);
    int split = 1;
    for (int i = 0; i < N; ++i) if (!((t->nd[id].hi[i] - t->nd[id].lo[i]) > hv[i])) split = 0;
    if (!split) return;
    qt_split(t, id);
    int hmax = 0;
    for (int c = 0; c < (1 << N); ++c) {
        int ch = t->nd[id].child[c];
        qt_build(t, ch, level + 1, h, user);
        if (t->nd[ch].height > hmax) hmax = t->nd[ch].height;
    }
    t->nd[id].height = hmax + 1;
}

/* Quadtree::isLeaf (quadtree.h:156-165) */
static int qt_isleaf(const qtree* t, int id, int c)
{
    const qnode* q = &t->nd[id];
    if (t->rapid) return q->height == c || (q->height < c && t->nd[q->parent].height > c);
    int clevel = t->numLevels - 1 - c;
    return q->level == clevel || (q->level < clevel && q->height == 0);
}

static void qt_dfs_layer(const qtree* t, int id, int c, ivec* out)
{
    if (qt_isleaf(t, id, c)) { iv_push(out, id); return; }
    for (int k = 0; k < (1 << t->N); ++k) qt_dfs_layer(t, t->nd[id].child[k], c, out);
}

/* Quadtree::isNeighbor (quadtree.tpp:59-88) */
static int qt_isneighbor(const qtree* t, const qnode* a, const qnode* b, int dim, int dir)
{
    double pa = (dir == 0) ? a->lo[dim] : a->hi[dim];
    double pb = (dir == 0) ? b->hi[dim] : b->lo[dim];
    if (t->periodic) {
        double dl = t->nd[0].lo[dim], du = t->nd[0].hi[dim];
        if ((dir == 0 && pa == dl && pb == du) || (dir == 1 && pa == du && pb == dl)) pb = pa;
    }
    if (pa != pb) return 0;
    for (int i = 0; i < t->N; ++i) if (i != dim) if (!(b->hi[i] > a->lo[i] && b->lo[i] < a->hi[i])) return 0;
    return 1;
}

/* Quadtree::isParent (quadtree.tpp:91-98) */
static int qt_isparent(const qtree* t, const qnode* a, const qnode* b)
{
    int inside = 1, equal = 1;
    for (int i = 0; i < t->N; ++i) {
        if (!(a->lo[i] >= b->lo[i] && a->hi[i] <= b->hi[i])) inside = 0;
        if (!(a->lo[i] == b->lo[i] && a->hi[i] == b->hi[i])) equal = 0;
    }
    return inside && !equal;
}

/* Quadtree::search (quadtree.tpp:101-114) */
static void qt_search(const qtree* t, int a, int dim, int dir, int id, int c, ivec* out)
{
    const qnode* A = &t->nd[a]; const qnode* B = &t->nd[id];
    if (!(qt_isneighbor(t, A, B, dim, dir) || qt_isparent(t, A, B))) return;
    if (!qt_isleaf(t, id, c)) { for (int k = 0; k < (1 << t->N); ++k) qt_search(t, a, dim, dir, t->nd[id].child[k], c, out); }
    else iv_push(out, id);
}

static void qt_neighbors(const qtree* t, int id, int dim, int dir, int c, ivec* out) { out->n = 0; qt_search(t, id, dim, dir, 0, c, out); }

static void qt_fix_heights(qtree* t, int id)
{
    int p = t->nd[id].parent;
    while (p >= 0) {
        int hp = t->nd[p].height, hc = t->nd[id].height + 1;
        if (hc <= hp) break;
        t->nd[p].height = hc; id = p; p = t->nd[p].parent;
    }
}

/* 2:1 balance (not in the reference): split every leaf that has a face
 * neighbour leaf more than one level finer, until none is left.  New
 * children are appended; element order remains the DFS (Morton) order. */
static void qt_balance(qtree* t)
{
    ivec leaves = {0}, nb = {0};
    int changed = 1;
    while (changed) {
        changed = 0; leaves.n = 0;
        qt_dfs_layer(t, 0, 0, &leaves);
        for (int k = 0; k < leaves.n; ++k) {
            int id = leaves.v[k];
            int need = 0;
            for (int d = 0; d < t->N && !need; ++d) for (int dir = 0; dir < 2 && !need; ++dir) {
                qt_neighbors(t, id, d, dir, 0, &nb);
                for (int q = 0; q < nb.n; ++q) if (t->nd[nb.v[q]].level > t->nd[id].level + 1) { need = 1; break; }
            }
            if (need) {
                qt_split(t, id);
                t->nd[id].height = 1;
                qt_fix_heights(t, id);
                if (t->nd[id].level + 2 > t->numLevels) t->numLevels = t->nd[id].level + 2;
                changed = 1;
            }
        }
    }
    free(leaves.v); free(nb.v);
}

void* orc_quadtree_new(int N, const double* lo, const double* hi, orc_hfn h, void* user,
                       int rapid, int periodic, int balance)
{
    if (N < 2 || N > 3 || !h) return NULL;
    qtree* t = (qtree*)xcalloc(1, sizeof(qtree));
    t->N = N; t->numLevels = 1; t->rapid = rapid ? 1 : 0; t->periodic = periodic ? 1 : 0;
    double l0[3] = {0, 0, 0}, h0[3] = {1, 1, 1};
    qt_push(t, lo ? lo : l0, hi ? hi : h0, -1);
    qt_build(t, 0, 0, h, user);
    if (balance) qt_balance(t);
    return t;
}

void orc_quadtree_free(void* p) { qtree* t = (qtree*)p; if (!t) return; free(t->nd); free(t); }

int orc_quadtree_info(void* p, int* nnodes, int* nlevels) { qtree* t = (qtree*)p; *nnodes = t->n; *nlevels = t->numLevels; return 0; }

int orc_quadtree_layer(void* p, int l, int* ids)
{
    qtree* t = (qtree*)p; ivec v = {0};
    if (l < 0 || l >= t->numLevels) return -1;
    qt_dfs_layer(t, 0, l, &v);
    if (ids) memcpy(ids, v.v, sizeof(int)*v.n);
    int n = v.n; free(v.v); return n;
}

int orc_quadtree_node(void* p, int id, int* ilh, double* lohi)
{
    qtree* t = (qtree*)p; if (id < 0 || id >= t->n) return -1;
    ilh[0] = t->nd[id].level; ilh[1] = t->nd[id].height; ilh[2] = t->nd[id].parent;
    for (int i = 0; i < t->N; ++i) { lohi[i] = t->nd[id].lo[i]; lohi[t->N+i] = t->nd[id].hi[i]; }
    return 0;
}

/* ======================================================================
 * Block sparse matrix (sparseblockmatrix.{h,tpp}), rows kept sorted by column
 * ====================================================================== */
typedef struct {
    int m, n, br, bc;
    int* len; int* cap; int** col; long** blk;
    double* pool; long np, pcap;
} bsm;

static void bsm_init(bsm* A, int m, int n, int br, int bc)
{
    memset(A, 0, sizeof(*A));
    A->m = m; A->n = n; A->br = br; A->bc = bc;
    A->len = (int*)xcalloc(m, sizeof(int)); A->cap = (int*)xcalloc(m, sizeof(int));
    A->col = (int**)xcalloc(m, sizeof(int*)); A->blk = (long**)xcalloc(m, sizeof(long*));
}

static void bsm_free(bsm* A)
{
    if (!A->len) return;
    for (int i = 0; i < A->m; ++i) { free(A->col[i]); free(A->blk[i]); }
    free(A->len); free(A->cap); free(A->col); free(A->blk); free(A->pool);
    memset(A, 0, sizeof(*A));
}

static double* bsm_get(const bsm* A, int i, int j)
{
    int lo = 0, hi = A->len[i] - 1;
    while (lo <= hi) { int mid = (lo + hi)/2; int c = A->col[i][mid]; if (c == j) return A->pool + A->blk[i][mid]; if (c < j) lo = mid + 1; else hi = mid - 1; }
    return NULL;
}

/* addToBlock (sparseblockmatrix.tpp:92-100): creates the block if missing */
static double* bsm_slot(bsm* A, int i, int j)
{
    double* p = bsm_get(A, i, j);
    if (p) return p;
    long bs = (long)A->br*A->bc;
    if (A->np + bs > A->pcap) { A->pcap = (A->pcap ? 2*A->pcap : 1024*bs); if (A->pcap < A->np + bs) A->pcap = A->np + bs; A->pool = (double*)xrealloc(A->pool, sizeof(double)*A->pcap); }
    long off = A->np; A->np += bs;
    memset(A->pool + off, 0, sizeof(double)*bs);
    if (A->len[i] == A->cap[i]) { A->cap[i] = A->cap[i] ? 2*A->cap[i] : 4; A->col[i] = (int*)xrealloc(A->col[i], sizeof(int)*A->cap[i]); A->blk[i] = (long*)xrealloc(A->blk[i], sizeof(long)*A->cap[i]); }
    int k = A->len[i];
    while (k > 0 && A->col[i][k-1] > j) { A->col[i][k] = A->col[i][k-1]; A->blk[i][k] = A->blk[i][k-1]; --k; }
    A->col[i][k] = j; A->blk[i][k] = off; A->len[i]++;
    return A->pool + off;
}

static void bsm_add(bsm* A, int i, int j, double alpha, const double* X)
{
    double* p = bsm_slot(A, i, j); long bs = (long)A->br*A->bc;
    for (long k = 0; k < bs; ++k) p[k] += alpha*X[k];
}

static void bsm_copy(bsm* C, const bsm* A)
{
    bsm_init(C, A->m, A->n, A->br, A->bc);
    for (int i = 0; i < A->m; ++i) for (int k = 0; k < A->len[i]; ++k) bsm_add(C, i, A->col[i][k], 1.0, A->pool + A->blk[i][k]);
}

static void bsm_transpose(bsm* C, const bsm* A)
{
    bsm_init(C, A->n, A->m, A->bc, A->br);
    double* t = (double*)xmalloc(sizeof(double)*A->br*A->bc);
    for (int i = 0; i < A->m; ++i) for (int k = 0; k < A->len[i]; ++k) {
        const double* s = A->pool + A->blk[i][k];
        for (int r = 0; r < A->br; ++r) for (int c = 0; c < A->bc; ++c) t[c*A->br + r] = s[r*A->bc + c];
        bsm_add(C, A->col[i][k], i, 1.0, t);
    }
    free(t);
}

/* multiply_mm (sparseblockmatrix.tpp:234-253): boolean product + block GEMMs */
static void bsm_mul(bsm* C, const bsm* A, const bsm* B)
{
    bsm_init(C, A->m, B->n, A->br, B->bc);
    double* t = (double*)xmalloc(sizeof(double)*A->br*B->bc);
    for (int i = 0; i < A->m; ++i) for (int a = 0; a < A->len[i]; ++a) {
        int k = A->col[i][a]; const double* Ab = A->pool + A->blk[i][a];
        for (int b = 0; b < B->len[k]; ++b) {
            int j = B->col[k][b];
            gemm(0, 0, A->br, B->bc, A->bc, 1.0, Ab, A->bc, B->pool + B->blk[k][b], B->bc, 0.0, t, B->bc);
            bsm_add(C, i, j, 1.0, t);
        }
    }
    free(t);
}

/* multiply_mm_t (sparseblockmatrix.tpp:261-280): C = A^T B */
static void bsm_mul_t(bsm* C, const bsm* A, const bsm* B)
{
    bsm At; bsm_transpose(&At, A); bsm_mul(C, &At, B); bsm_free(&At);
}

/* multiply_add_mv (sparseblockmatrix.tpp:358-371): y = alpha A x + beta y */
static void bsm_mv(const bsm* A, double alpha, const double* x, double beta, double* y)
{
    int br = A->br, bc = A->bc;
    for (int i = 0; i < A->m; ++i) {
        double* yi = y + (long)i*br;
        for (int r = 0; r < br; ++r) yi[r] = (beta == 0.0) ? 0.0 : beta*yi[r];
        for (int k = 0; k < A->len[i]; ++k) {
            const double* Ab = A->pool + A->blk[i][k]; const double* xj = x + (long)A->col[i][k]*bc;
            for (int r = 0; r < br; ++r) { double s = 0.0; for (int c = 0; c < bc; ++c) s += Ab[r*bc+c]*xj[c]; yi[r] += alpha*s; }
        }
    }
}

/* multiply_mv_t (sparseblockmatrix.tpp:322-335): y = A^T x */
static void bsm_mv_t(const bsm* A, const double* x, double* y)
{
    int br = A->br, bc = A->bc;
    memset(y, 0, sizeof(double)*(size_t)A->n*bc);
    for (int i = 0; i < A->m; ++i) for (int k = 0; k < A->len[i]; ++k) {
        const double* Ab = A->pool + A->blk[i][k]; const double* xi = x + (long)i*br; double* yj = y + (long)A->col[i][k]*bc;
        for (int c = 0; c < bc; ++c) { double s = 0.0; for (int r = 0; r < br; ++r) s += Ab[r*bc+c]*xi[r]; yj[c] += s; }
    }
}

static long bsm_export(const bsm* A, int* rowptr, int* cols, double* blocks)
{
    long nz = 0, bs = (long)A->br*A->bc;
    for (int i = 0; i < A->m; ++i) {
        if (rowptr) rowptr[i] = (int)nz;
        for (int k = 0; k < A->len[i]; ++k) {
            if (cols) cols[nz] = A->col[i][k];
            if (blocks) memcpy(blocks + nz*bs, A->pool + A->blk[i][k], sizeof(double)*bs);
            ++nz;
        }
    }
    if (rowptr) rowptr[A->m] = (int)nz;
    return nz;
}

int orc_bsr_spmv(int m, int b, const int* rowptr, const int* cols, const double* blocks, const double* x, double* y)
{
    for (int i = 0; i < m; ++i) {
        for (int r = 0; r < b; ++r) y[i*b+r] = 0.0;
        for (int k = rowptr[i]; k < rowptr[i+1]; ++k)
            for (int r = 0; r < b; ++r) { double s = 0.0; for (int c = 0; c < b; ++c) s += blocks[(long)k*b*b + r*b + c]*x[(long)cols[k]*b + c]; y[i*b+r] += s; }
    }
    return 0;
}

/* ======================================================================
 * Mesh (src/cartesian/mesh/mesh.h:17-374)
 * ====================================================================== */
typedef struct {
    int left, right, dim, canonical;
    double nsign;
    double lo[3], hi[3];
    double* xl; double* xr;      /* Q^{N-1} x N local quadrature points (non-canonical) */
} oface;

typedef struct {
    int N, P, ne, nf, nbf;
    qtree* qt;
    basis_t B;
    int* gid;                    /* element -> tree node */
    double* elo; double* ehi;    /* element boxes, ne x N */
    oface* faces;
    int* bface;                  /* boundary face index -> face */
    int nq;                      /* Q^{N-1} */
} omesh;

static double face_area(const omesh* M, const oface* f) { double a = 1.0; for (int i = 0; i < M->N; ++i) if (i != f->dim) a *= f->hi[i] - f->lo[i]; return a; }

static void mesh_build(omesh* Msh, qtree* qt, int P)
{
    int N = qt->N;
    memset(Msh, 0, sizeof(*Msh));
    Msh->N = N; Msh->P = P; Msh->qt = qt;
    basis_init(&Msh->B, N, P);
    Msh->nq = ipow_(QPTS, N-1);
    ivec ids = {0};
    qt_dfs_layer(qt, 0, 0, &ids);
    int ne = ids.n;
    Msh->ne = ne; Msh->gid = ids.v;
    int* g2l = (int*)xmalloc(sizeof(int)*qt->n);
    for (int i = 0; i < qt->n; ++i) g2l[i] = -1;
    Msh->elo = (double*)xmalloc(sizeof(double)*ne*N); Msh->ehi = (double*)xmalloc(sizeof(double)*ne*N);
    for (int e = 0; e < ne; ++e) { g2l[ids.v[e]] = e; for (int i = 0; i < N; ++i) { Msh->elo[e*N+i] = qt->nd[ids.v[e]].lo[i]; Msh->ehi[e*N+i] = qt->nd[ids.v[e]].hi[i]; } }

    int fcap = 64, nf = 0, nbf = 0; oface* F = (oface*)xmalloc(sizeof(oface)*fcap);
    ivec nb = {0};
    for (int e = 0; e < ne; ++e) {
        for (int i = 0; i < N; ++i) for (int dir = 0; dir < 2; ++dir) {
            qt_neighbors(qt, ids.v[e], i, dir, 0, &nb);
            if (nb.n == 0) {
                if (nf == fcap) { fcap *= 2; F = (oface*)xrealloc(F, sizeof(oface)*fcap); }
                oface* f = &F[nf++]; memset(f, 0, sizeof(*f));
                f->left = e; f->right = -(2*i + 1 + dir); f->dim = i; f->canonical = 1;
                f->nsign = dir == 0 ? -1.0 : 1.0;
                for (int k = 0; k < N; ++k) { f->lo[k] = Msh->elo[e*N+k]; f->hi[k] = Msh->ehi[e*N+k]; }
                if (dir == 0) f->hi[i] = f->lo[i]; else f->lo[i] = f->hi[i];
                nbf++;
            } else {
                for (int q = 0; q < nb.n; ++q) {
                    int nl = g2l[nb.v[q]];
                    if (!(e < nl)) continue;
                    int L = (dir == 0) ? nl : e, R = (dir == 1) ? nl : e;
                    if (nf == fcap) { fcap *= 2; F = (oface*)xrealloc(F, sizeof(oface)*fcap); }
                    oface* f = &F[nf++]; memset(f, 0, sizeof(*f));
                    f->left = L; f->right = R; f->dim = i; f->nsign = 1.0;
                    for (int k = 0; k < N; ++k) {
                        f->lo[k] = fmax(Msh->elo[L*N+k], Msh->elo[R*N+k]);
                        f->hi[k] = fmin(Msh->ehi[L*N+k], Msh->ehi[R*N+k]);
                    }
                    f->lo[i] = Msh->elo[R*N+i]; f->hi[i] = Msh->ehi[L*N+i];
                    /* Face constructor: canonical iff the left element's right side
                       equals the right element's left side (mesh.h:33-43) */
                    double llo[3], lhi[3], rlo[3], rhi[3];
                    for (int k = 0; k < N; ++k) { llo[k] = Msh->elo[L*N+k]; lhi[k] = Msh->ehi[L*N+k]; rlo[k] = Msh->elo[R*N+k]; rhi[k] = Msh->ehi[R*N+k]; }
                    llo[i] = lhi[i]; rhi[i] = rlo[i];
                    int canon = 1; for (int k = 0; k < N; ++k) if (llo[k] != rlo[k] || lhi[k] != rhi[k]) canon = 0;
                    f->canonical = canon;
                    if (!canon) {
                        int nqf = Msh->nq;
                        f->xl = (double*)xmalloc(sizeof(double)*nqf*N); f->xr = (double*)xmalloc(sizeof(double)*nqf*N);
                        double al[3], bl[3], ar[3], br[3];
                        for (int k = 0; k < N; ++k) { al[k] = f->lo[k] - llo[k]; bl[k] = f->hi[k] - llo[k]; ar[k] = f->lo[k] - rlo[k]; br[k] = f->hi[k] - rlo[k]; }
                        for (int k = 0; k < N; ++k) if (k != i) { double wl = lhi[k] - llo[k], wr = rhi[k] - rlo[k]; al[k] /= wl; bl[k] /= wl; ar[k] /= wr; br[k] /= wr; }
                        int it[3];
                        for (int qq = 0; qq < nqf; ++qq) {
                            unlin(qq, N-1, QPTS, it);
                            double* tl = f->xl + qq*N; double* tr = f->xr + qq*N;
                            for (int k = 0; k < N-1; ++k) { double x = Msh->B.gx[it[k]]; int kk = k + (k >= i); tl[kk] = al[kk] + (bl[kk]-al[kk])*x; tr[kk] = ar[kk] + (br[kk]-ar[kk])*x; }
                            tl[i] = 1.0; tr[i] = 0.0;
                        }
                    }
                }
            }
        }
    }
    free(nb.v); free(g2l);
    Msh->faces = F; Msh->nf = nf; Msh->nbf = nbf;
    Msh->bface = (int*)xmalloc(sizeof(int)*(nbf ? nbf : 1));
    for (int f = 0, k = 0; f < nf; ++f) if (F[f].right < 0) Msh->bface[k++] = f;
}

static void mesh_free(omesh* M)
{
    for (int f = 0; f < M->nf; ++f) { free(M->faces[f].xl); free(M->faces[f].xr); }
    free(M->faces); free(M->gid); free(M->elo); free(M->ehi); free(M->bface);
    basis_free(&M->B);
}

/* Mesh::lift (mesh.h:282-352).  Any output may be NULL.  L, R: npl x nq. */
static void mesh_lift(const omesh* Msh, const oface* f, double* L, double* R, double* LL, double* RR, double* RL)
{
    const basis_t* B = &Msh->B;
    int N = Msh->N, P = Msh->P, n = B->npl, dim = f->dim;
    if (f->canonical && f->right >= 0) {
        int nfl = B->nfl; int sl[MAXP*MAXP*MAXP], sr[MAXP*MAXP*MAXP]; int it[3], jt[3];
        for (int i = 0; i < nfl; ++i) {
            unlin(i, N-1, P, it);
            for (int k = 0; k < dim; ++k) jt[k] = it[k];
            for (int k = dim; k < N-1; ++k) jt[k+1] = it[k];
            jt[dim] = (f->nsign > 0) ? P-1 : 0; sl[i] = lin_(jt, N, P);
            jt[dim] = (f->nsign > 0) ? 0 : P-1; sr[i] = lin_(jt, N, P);
        }
        double a = face_area(Msh, f);
        if (LL) memset(LL, 0, sizeof(double)*n*n);
        if (RR) memset(RR, 0, sizeof(double)*n*n);
        if (RL) memset(RL, 0, sizeof(double)*n*n);
        for (int i = 0; i < nfl; ++i) for (int j = 0; j < nfl; ++j) {
            double v = B->fmass[i*nfl+j]*a;
            if (LL) LL[sl[i]*n + sl[j]] = v;
            if (RR) RR[sr[i]*n + sr[j]] = v;
            if (RL) RL[sr[i]*n + sl[j]] = v;
        }
        return;
    }
    int nq = Msh->nq; int it[3];
    double* phl = (double*)xmalloc(sizeof(double)*nq*n);
    double* phr = (double*)xmalloc(sizeof(double)*nq*n);
    double* w = (double*)xmalloc(sizeof(double)*nq);
    double area = face_area(Msh, f);
    for (int q = 0; q < nq; ++q) {
        double ql[3], qr[3];
        unlin(q, N-1, QPTS, it);
        if (f->canonical) {
            for (int k = 0; k < N-1; ++k) { double x = B->gx[it[k]]; int kk = k + (k >= dim); ql[kk] = x; qr[kk] = x; }
            ql[dim] = (f->nsign > 0) ? 1.0 : 0.0; qr[dim] = (f->nsign > 0) ? 0.0 : 1.0;
        } else {
            for (int k = 0; k < N; ++k) { ql[k] = f->xl[q*N+k]; qr[k] = f->xr[q*N+k]; }
        }
        lagrange_eval_nd(B, ql, phl + q*n); lagrange_eval_nd(B, qr, phr + q*n);
        w[q] = 1.0; for (int k = 0; k < N-1; ++k) w[q] *= B->gw[it[k]];
    }
    for (int q = 0; q < nq; ++q) w[q] *= area;
    double* Lm = (double*)xmalloc(sizeof(double)*n*nq); double* Rm = (double*)xmalloc(sizeof(double)*n*nq);
    for (int i = 0; i < n; ++i) for (int q = 0; q < nq; ++q) { Lm[i*nq+q] = phl[q*n+i]*w[q]; Rm[i*nq+q] = phr[q*n+i]*w[q]; }
    if (L) memcpy(L, Lm, sizeof(double)*n*nq);
    if (R) memcpy(R, Rm, sizeof(double)*n*nq);
    if (LL) gemm(0, 0, n, n, nq, 1.0, Lm, nq, phl, n, 0.0, LL, n);
    if (RR) gemm(0, 0, n, n, nq, 1.0, Rm, nq, phr, n, 0.0, RR, n);
    if (RL) gemm(0, 0, n, n, nq, 1.0, Rm, nq, phl, n, 0.0, RL, n);
    free(phl); free(phr); free(w); free(Lm); free(Rm);
}

/* ======================================================================
 * LDG assembly (src/cartesian/dg/ldgpoisson.{h,tpp})
 * ====================================================================== */
/* the operators of one discretisation (LDGOperators, ldgpoisson.h:15-53);
 * the first member of every problem type */
typedef struct {
    int N, ne, b, ndof;
    bsm M, T, A, G[3];
} oops;

typedef struct {
    oops ops;
    omesh mesh;
    int b, ndof, nqb;
    int bctype[6];
    double tau0, tauD;
} oproblem;

/* LDGOperators::construct_laplacian (ldgpoisson.h:34-52): A = T + sum_d G_d^T M G_d */
static void construct_laplacian(bsm* A, const bsm* T, const bsm* M, const bsm* G, int N)
{
    bsm_copy(A, T);
    int b = M->br;
    double* GTM = (double*)xmalloc(sizeof(double)*b*b); double* GTMG = (double*)xmalloc(sizeof(double)*b*b);
    for (int d = 0; d < N; ++d) {
        for (int k = 0; k < M->m; ++k) {
            const double* Mk = bsm_get(M, k, k);
            for (int a = 0; a < G[d].len[k]; ++a) {
                int i = G[d].col[k][a];
                gemm(1, 0, b, b, b, 1.0, G[d].pool + G[d].blk[k][a], b, Mk, b, 0.0, GTM, b);
                for (int c = 0; c < G[d].len[k]; ++c) {
                    int j = G[d].col[k][c];
                    gemm(0, 0, b, b, b, 1.0, GTM, b, G[d].pool + G[d].blk[k][c], b, 0.0, GTMG, b);
                    bsm_add(A, i, j, 1.0, GTMG);
                }
            }
        }
    }
    free(GTM); free(GTMG);
}

void* orc_problem_new(void* qtp, int p, const int* bctype, double tau0, double tauD)
{
    qtree* qt = (qtree*)qtp;
    int P = p + 1;
    if (!qt || P < 1 || P > 6) return NULL;
    oproblem* pr = (oproblem*)xcalloc(1, sizeof(oproblem));
    mesh_build(&pr->mesh, qt, P);
    omesh* Msh = &pr->mesh; basis_t* B = &Msh->B;
    int N = qt->N, ne = Msh->ne, b = B->npl;
    pr->b = b; pr->ndof = ne*b; pr->nqb = Msh->nbf*Msh->nq; pr->tau0 = tau0; pr->tauD = tauD;
    for (int k = 0; k < 2*N; ++k) pr->bctype[k] = bctype ? bctype[k] : 0;
    for (int k = 0; k < 2*N; ++k) if (pr->bctype[k] != 0 && pr->bctype[k] != 1) { orc_problem_free(pr); return NULL; }

    bsm_init(&pr->ops.M, ne, ne, b, b); bsm_init(&pr->ops.T, ne, ne, b, b);
    for (int d = 0; d < N; ++d) bsm_init(&pr->ops.G[d], ne, ne, b, b);
    double* blk = (double*)xmalloc(sizeof(double)*b*b);
    /* construct_mass_matrix (ldgpoisson.tpp:27-39) */
    for (int e = 0; e < ne; ++e) {
        double vol = 1.0; for (int i = 0; i < N; ++i) vol *= Msh->ehi[e*N+i] - Msh->elo[e*N+i];
        bsm_add(&pr->ops.M, e, e, vol, B->mass);
    }
    /* construct_broken_gradient (ldgpoisson.tpp:42-60) */
    for (int d = 0; d < N; ++d) for (int e = 0; e < ne; ++e) {
        double jac = 1.0/(Msh->ehi[e*N+d] - Msh->elo[e*N+d]);
        bsm_add(&pr->ops.G[d], e, e, jac, B->diff[d]);
    }
    /* construct_lifting_terms (ldgpoisson.tpp:63-100) */
    double* LL = (double*)xmalloc(sizeof(double)*b*b); double* RR = (double*)xmalloc(sizeof(double)*b*b); double* RL = (double*)xmalloc(sizeof(double)*b*b);
    double* MI = (double*)xmalloc(sizeof(double)*b*b);
    for (int fi = 0; fi < Msh->nf; ++fi) {
        const oface* f = &Msh->faces[fi];
        mesh_lift(Msh, f, NULL, NULL, LL, RR, RL);
        double n = f->nsign;
        if (f->right >= 0) {
            int r = f->right, l = f->left;
            double vol = 1.0; for (int i = 0; i < N; ++i) vol *= Msh->ehi[r*N+i] - Msh->elo[r*N+i];
            for (int k = 0; k < b*b; ++k) MI[k] = B->invmass[k]*(1.0/vol);
            gemm(0, 0, b, b, b, n, MI, b, RR, b, 0.0, blk, b); bsm_add(&pr->ops.G[f->dim], r, r, 1.0, blk);
            gemm(0, 0, b, b, b, -n, MI, b, RL, b, 0.0, blk, b); bsm_add(&pr->ops.G[f->dim], r, l, 1.0, blk);
            bsm_add(&pr->ops.T, r, r, tau0, RR);
            bsm_add(&pr->ops.T, l, l, tau0, LL);
            bsm_add(&pr->ops.T, r, l, -tau0, RL);
            for (int i = 0; i < b; ++i) for (int j = 0; j < b; ++j) blk[i*b+j] = RL[j*b+i];
            bsm_add(&pr->ops.T, l, r, -tau0, blk);
        } else {
            int tag = -f->right - 1;
            if (pr->bctype[tag] == 0) {
                int l = f->left;
                double vol = 1.0; for (int i = 0; i < N; ++i) vol *= Msh->ehi[l*N+i] - Msh->elo[l*N+i];
                for (int k = 0; k < b*b; ++k) MI[k] = B->invmass[k]*(1.0/vol);
                gemm(0, 0, b, b, b, -n, MI, b, LL, b, 0.0, blk, b); bsm_add(&pr->ops.G[f->dim], l, l, 1.0, blk);
                bsm_add(&pr->ops.T, l, l, tauD, LL);
            }
        }
    }
    free(LL); free(RR); free(RL); free(MI); free(blk);
    construct_laplacian(&pr->ops.A, &pr->ops.T, &pr->ops.M, pr->ops.G, N);
    pr->ops.N = N; pr->ops.ne = ne; pr->ops.b = b; pr->ops.ndof = ne*b;
    return pr;
}

void orc_problem_free(void* p)
{
    oproblem* pr = (oproblem*)p; if (!pr) return;
    bsm_free(&pr->ops.M); bsm_free(&pr->ops.T); bsm_free(&pr->ops.A);
    for (int d = 0; d < 3; ++d) bsm_free(&pr->ops.G[d]);
    mesh_free(&pr->mesh); free(pr);
}

int orc_problem_sizes(void* p, int* s)
{
    oproblem* pr = (oproblem*)p;
    s[0] = pr->mesh.ne; s[1] = pr->mesh.nf; s[2] = pr->mesh.nbf; s[3] = pr->b; s[4] = pr->nqb; s[5] = pr->ndof;
    return 0;
}

/* Master::dgnodes (master.h:35-43) for every element, element-major */
int orc_problem_nodes(void* p, double* xyz)
{
    oproblem* pr = (oproblem*)p; omesh* M = &pr->mesh; int N = M->N, P = M->P, b = pr->b, it[3];
    for (int e = 0; e < M->ne; ++e) for (int i = 0; i < b; ++i) {
        unlin(i, N, P, it);
        for (int k = 0; k < N; ++k) { double lo = M->elo[e*N+k], w = M->ehi[e*N+k] - lo; xyz[((long)e*b + i)*N + k] = lo + w*M->B.nodes[it[k]]; }
    }
    return 0;
}

/* boundary quadrature points (ldgpoisson.tpp:119-129) in boundary-face order */
int orc_problem_bc_qpoints(void* p, double* xyz, int* tag)
{
    oproblem* pr = (oproblem*)p; omesh* M = &pr->mesh; int N = M->N, nq = M->nq, it[3];
    for (int k = 0; k < M->nbf; ++k) {
        const oface* f = &M->faces[M->bface[k]];
        for (int q = 0; q < nq; ++q) {
            unlin(q, N-1, QPTS, it);
            double* x = xyz + ((long)k*nq + q)*N;
            for (int j = 0; j < N-1; ++j) { int kk = j + (j >= f->dim); x[kk] = f->lo[kk] + (f->hi[kk] - f->lo[kk])*M->B.gx[it[j]]; }
            x[f->dim] = f->lo[f->dim];
            if (tag) tag[(long)k*nq + q] = f->right;
        }
    }
    return 0;
}

int orc_problem_faces(void* p, int* lrdc, double* normal)
{
    oproblem* pr = (oproblem*)p; omesh* M = &pr->mesh;
    for (int f = 0; f < M->nf; ++f) {
        lrdc[4*f] = M->faces[f].left; lrdc[4*f+1] = M->faces[f].right; lrdc[4*f+2] = M->faces[f].dim; lrdc[4*f+3] = M->faces[f].canonical;
        if (normal) normal[f] = M->faces[f].nsign;
    }
    return 0;
}

/* add_source_terms (ldgpoisson.tpp:103-161) and computeRHS (:167-176) */
int orc_problem_rhs(void* p, const double* fnod, const double* g, double* out)
{
    oproblem* pr = (oproblem*)p; omesh* M = &pr->mesh; int N = M->N, b = pr->b, nq = M->nq, ne = M->ne;
    double* rhs = (double*)xcalloc((size_t)pr->ndof, sizeof(double));
    double* Jg[3] = {0, 0, 0};
    for (int d = 0; d < N; ++d) Jg[d] = (double*)xcalloc((size_t)pr->ndof, sizeof(double));
    double* L = (double*)xmalloc(sizeof(double)*b*nq); double* Lg = (double*)xmalloc(sizeof(double)*b);
    for (int k = 0; k < M->nbf; ++k) {
        const oface* f = &M->faces[M->bface[k]];
        mesh_lift(M, f, L, NULL, NULL, NULL, NULL);
        const double* gv = g ? g + (long)k*nq : NULL;
        for (int i = 0; i < b; ++i) { double s = 0.0; if (gv) for (int q = 0; q < nq; ++q) s += L[i*nq+q]*gv[q]; Lg[i] = s; }
        int tag = -f->right - 1; double n = f->nsign; long off = (long)f->left*b;
        if (pr->bctype[tag] == 0) {
            for (int i = 0; i < b; ++i) { Jg[f->dim][off+i] += n*Lg[i]; rhs[off+i] += pr->tauD*Lg[i]; }
        } else {
            for (int i = 0; i < b; ++i) rhs[off+i] += Lg[i];
        }
    }
    /* rhs_e -= sum_d sum_{i in rowsInCol(e)} G_d(i,e)^T Jg_d(i) */
    double* t = (double*)xmalloc(sizeof(double)*pr->ndof);
    for (int d = 0; d < N; ++d) {
        bsm_mv_t(&pr->ops.G[d], Jg[d], t);
        for (long i = 0; i < pr->ndof; ++i) rhs[i] -= t[i];
    }
    for (int e = 0; e < ne; ++e) {
        const double* Me = bsm_get(&pr->ops.M, e, e);
        for (int i = 0; i < b; ++i) { double s = 0.0; if (fnod) for (int j = 0; j < b; ++j) s += Me[i*b+j]*fnod[(long)e*b+j]; out[(long)e*b+i] = rhs[(long)e*b+i] + s; }
    }
    free(t); free(L); free(Lg); free(rhs); for (int d = 0; d < N; ++d) free(Jg[d]);
    return 0;
}

/* transpose of the g -> rhs map above (for the adjoint): dg = B_g^T lam */
int orc_problem_rhs_g_adjoint(void* p, const double* lam, double* dg)
{
    oproblem* pr = (oproblem*)p; omesh* M = &pr->mesh; int N = M->N, b = pr->b, nq = M->nq;
    double* Glam[3] = {0, 0, 0};
    for (int d = 0; d < N; ++d) { Glam[d] = (double*)xmalloc(sizeof(double)*pr->ndof); bsm_mv(&pr->ops.G[d], 1.0, lam, 0.0, Glam[d]); }
    double* L = (double*)xmalloc(sizeof(double)*b*nq);
    for (int k = 0; k < M->nbf; ++k) {
        const oface* f = &M->faces[M->bface[k]];
        mesh_lift(M, f, L, NULL, NULL, NULL, NULL);
        int tag = -f->right - 1; double n = f->nsign; long off = (long)f->left*b;
        double v[MAXP*MAXP*MAXP];
        for (int i = 0; i < b; ++i) v[i] = (pr->bctype[tag] == 0) ? pr->tauD*lam[off+i] - n*Glam[f->dim][off+i] : lam[off+i];
        for (int q = 0; q < nq; ++q) { double s = 0.0; for (int i = 0; i < b; ++i) s += L[i*nq+q]*v[i]; dg[(long)k*nq+q] = s; }
    }
    free(L); for (int d = 0; d < N; ++d) free(Glam[d]);
    return 0;
}

long orc_problem_export(void* p, int which, int* rowptr, int* cols, double* blocks)
{
    oproblem* pr = (oproblem*)p;
    const bsm* A = which == 0 ? &pr->ops.A : which == 1 ? &pr->ops.M : which == 2 ? &pr->ops.T : (which - 3 < pr->mesh.N && which >= 3) ? &pr->ops.G[which-3] : NULL;
    if (!A) return -1;
    return bsm_export(A, rowptr, cols, blocks);
}

/* ======================================================================
 * Multigrid (src/cartesian/multigrid/{interpolationhierarchy,multigrid}.h)
 * ====================================================================== */
typedef struct {
    int nb, b;
    bsm A, M, T, G[3];
    int has_ops;           /* M, T, G present (fine level or RATRAT) */
    ldlt_t* Dinv;          /* LDLT of each diagonal block of A */
    ldlt_t* Minv;          /* LDLT of each diagonal block of M (coarse RATRAT) */
    ldlt_t dense;          /* dense LDLT of the whole coarsest A (more than one block) */
    int has_dense;
    int* color; int ncolor;
    double *x, *bv, *r;
} olevel;

typedef struct {
    oops* ops;
    int N;
    orc_mg_params prm;
    int L;                 /* number of levels */
    olevel* lev;
    bsm* Tl;               /* L-1 interpolation operators (one block per fine row) */
    int submean;           /* coarsest b -= mean(b) (the quadtree Multigrid, multigrid.h:365-368) */
} omg;

/* InterpolationHierarchy<N,P>(qt) (interpolationhierarchy.h:30-82), with
 * the degree of basis B; Tl has room for `extra` operators in front (the
 * p-levels of an hp-hierarchy) */
static bsm* build_interp_qt_basis(const oproblem* pr, const basis_t* B, int extra, int* nlev)
{
    const qtree* qt = pr->mesh.qt;
    int N = qt->N, P = B->P, b = B->npl, it[3];
    bsm* Tl = (bsm*)xcalloc(extra + qt->numLevels - 1, sizeof(bsm)) + extra;
    int* cg2l = (int*)xmalloc(sizeof(int)*qt->n);
    double* E = (double*)xmalloc(sizeof(double)*b*b);
    double* I = (double*)xcalloc((size_t)b*b, sizeof(double));
    for (int i = 0; i < b; ++i) I[i*b+i] = 1.0;
    for (int l = 0; l < qt->numLevels - 1; ++l) {
        ivec fine = {0}, coarse = {0};
        qt_dfs_layer(qt, 0, l, &fine); qt_dfs_layer(qt, 0, l+1, &coarse);
        for (int k = 0; k < coarse.n; ++k) cg2l[coarse.v[k]] = k;
        bsm_init(&Tl[l], fine.n, coarse.n, b, b);
        for (int fl = 0; fl < fine.n; ++fl) {
            int fg = fine.v[fl];
            if (qt_isleaf(qt, fg, l+1)) { bsm_add(&Tl[l], fl, cg2l[fg], 1.0, I); continue; }
            int cg = qt->nd[fg].parent;
            double rlo[3], rhi[3];
            for (int k = 0; k < N; ++k) {
                double w = qt->nd[cg].hi[k] - qt->nd[cg].lo[k];
                rlo[k] = (qt->nd[fg].lo[k] - qt->nd[cg].lo[k])/w;
                rhi[k] = (qt->nd[fg].hi[k] - qt->nd[cg].lo[k])/w;
            }
            for (int i = 0; i < b; ++i) {
                unlin(i, N, P, it); double x[3];
                for (int k = 0; k < N; ++k) x[k] = rlo[k] + (rhi[k] - rlo[k])*B->nodes[it[k]];
                lagrange_eval_nd(B, x, E + i*b);
            }
            bsm_add(&Tl[l], fl, cg2l[cg], 1.0, E);
        }
        free(fine.v); free(coarse.v);
    }
    free(cg2l); free(E); free(I);
    *nlev = qt->numLevels;
    return Tl - extra;
}

static bsm* build_interp_qt(const oproblem* pr, int* nlev) { return build_interp_qt_basis(pr, &pr->mesh.B, 0, nlev); }

/* InterpolationHierarchy<N,P1,P2,...> (interpolationhierarchy.h:98-131): on
 * every element of the fine layer, the degree-P2 Lagrange basis evaluated at
 * the degree-P1 Gauss-Lobatto nodes (a block-diagonal T of P1^N x P2^N blocks) */
static void build_pinterp(bsm* T, int ne, int N, const basis_t* B1, const basis_t* B2)
{
    int b1 = B1->npl, b2 = B2->npl, it[3];
    double* E = (double*)xmalloc(sizeof(double)*(size_t)b1*b2);
    for (int i = 0; i < b1; ++i) {
        unlin(i, N, B1->P, it); double x[3];
        for (int k = 0; k < N; ++k) x[k] = B1->nodes[it[k]];
        lagrange_eval_nd(B2, x, E + (size_t)i*b2);
    }
    bsm_init(T, ne, ne, b1, b2);
    for (int e = 0; e < ne; ++e) bsm_add(T, e, e, 1.0, E);
    free(E);
}

/* greedy colouring in block-row order: smallest colour unused by any
 * already-coloured neighbour (not in the reference; multicolour GS) */
static void color_level(olevel* lv)
{
    const bsm* A = &lv->A; int n = A->m;
    lv->color = (int*)xmalloc(sizeof(int)*(n ? n : 1));
    int maxc = 0; char used[256];
    for (int i = 0; i < n; ++i) {
        memset(used, 0, sizeof(used));
        for (int k = 0; k < A->len[i]; ++k) { int j = A->col[i][k]; if (j < i) used[lv->color[j]] = 1; }
        int c = 0; while (used[c]) ++c;
        lv->color[i] = c; if (c + 1 > maxc) maxc = c + 1;
    }
    lv->ncolor = maxc;
}

static void level_factor(olevel* lv)
{
    lv->Dinv = (ldlt_t*)xcalloc(lv->nb, sizeof(ldlt_t));
    for (int i = 0; i < lv->nb; ++i) ldlt_factor(&lv->Dinv[i], bsm_get(&lv->A, i, i), lv->b);
    long n = (long)lv->nb*lv->b;
    lv->x = (double*)xcalloc(n, sizeof(double)); lv->bv = (double*)xcalloc(n, sizeof(double)); lv->r = (double*)xcalloc(n, sizeof(double));
    color_level(lv);
}

/* Level coarse constructor (multigrid.h:105-171), fine blocks bf, coarse bc
 * (T blocks bf x bc; the reference's Level<N,P>(ops<N,P1>, T<N,P1,P>)) */
static void coarsen(olevel* c, const olevel* f, const bsm* T, int restriction, int N, const int* crank)
{
    int bf = T->br, b = T->bc;
    memset(c, 0, sizeof(*c));
    c->nb = T->n; c->b = b;
    if (restriction == 1) {
        /* M_c = T^T M_f T (multigrid.h:117-130) */
        bsm_init(&c->M, T->n, T->n, b, b);
        double* TTM = (double*)xmalloc(sizeof(double)*b*bf); double* TTMT = (double*)xmalloc(sizeof(double)*b*b);
        for (int k = 0; k < f->M.m; ++k) {
            const double* Mk = bsm_get(&f->M, k, k);
            for (int a = 0; a < T->len[k]; ++a) {
                int i = T->col[k][a];
                gemm(1, 0, b, bf, bf, 1.0, T->pool + T->blk[k][a], b, Mk, bf, 0.0, TTM, bf);
                for (int cc = 0; cc < T->len[k]; ++cc) {
                    int j = T->col[k][cc];
                    gemm(0, 0, b, b, bf, 1.0, TTM, bf, T->pool + T->blk[k][cc], b, 0.0, TTMT, b);
                    bsm_add(&c->M, i, j, 1.0, TTMT);
                }
            }
        }
        free(TTM); free(TTMT);
        c->Minv = (ldlt_t*)xcalloc(c->nb, sizeof(ldlt_t));
        for (int i = 0; i < c->nb; ++i) ldlt_factor(&c->Minv[i], bsm_get(&c->M, i, i), b);
        /* G_c = M_c^{-1} T^T M_f G_f T (multigrid.h:137-149) */
        for (int d = 0; d < N; ++d) {
            bsm t1, t2;
            bsm_mul(&t1, &f->G[d], T);
            bsm_mul(&t2, &f->M, &t1);
            bsm_mul_t(&c->G[d], T, &t2);
            bsm_free(&t1); bsm_free(&t2);
            for (int i = 0; i < c->nb; ++i) for (int k = 0; k < c->G[d].len[i]; ++k) {
                double* blk = c->G[d].pool + c->G[d].blk[i][k];
                ldlt_solve_mat(&c->Minv[i], blk, b, blk);
            }
        }
        /* T_c = T^T T_f T (multigrid.h:151-153) */
        bsm t1; bsm_mul(&t1, &f->T, T); bsm_mul_t(&c->T, T, &t1); bsm_free(&t1);
        construct_laplacian(&c->A, &c->T, &c->M, c->G, N);
        c->has_ops = 1;
    } else {
        /* A_c = T^T A T (multigrid.h:158-162) */
        bsm t1; bsm_mul(&t1, &f->A, T); bsm_mul_t(&c->A, T, &t1); bsm_free(&t1);
    }
    /* coarse directions beyond crank[i] (zero columns of T) are decoupled:
     * unit diagonal */
    if (crank) for (int i = 0; i < c->nb; ++i) {
        double* D = bsm_get(&c->A, i, i);
        for (int d = crank[i]; d < b; ++d) D[d*b + d] += 1.0;
    }
    level_factor(c);
}

/* Multigrid(ops, hierarchy, params) (multigrid.h:291-301, setup 404-411) for
 * any hierarchy of one-block-per-row interpolation operators Tl[0..L-2] */
static omg* mg_setup(oops* ops, int N, bsm* Tl, int L, const orc_mg_params* prm, int* const* crank, int submean)
{
    omg* mg = (omg*)xcalloc(1, sizeof(omg));
    mg->ops = ops; mg->N = N; mg->prm = *prm; mg->Tl = Tl; mg->L = L; mg->submean = submean;
    mg->lev = (olevel*)xcalloc(L, sizeof(olevel));
    olevel* f = &mg->lev[0];
    f->nb = ops->ne; f->b = ops->b;
    bsm_copy(&f->A, &ops->A); bsm_copy(&f->M, &ops->M); bsm_copy(&f->T, &ops->T);
    for (int d = 0; d < N; ++d) bsm_copy(&f->G[d], &ops->G[d]);
    f->has_ops = 1;
    level_factor(f);
    for (int l = 0; l + 1 < L; ++l) coarsen(&mg->lev[l+1], &mg->lev[l], &mg->Tl[l], prm->restriction, N, crank ? crank[l] : NULL);
    /* coarsest level with more than one block: dense LDLT of the assembled operator */
    olevel* c = &mg->lev[L-1];
    if (prm->coarse == 0 && c->nb > 1) {
        int n = c->nb*c->b; double* D = (double*)xcalloc((size_t)n*n, sizeof(double));
        for (int i = 0; i < c->nb; ++i) for (int k = 0; k < c->A.len[i]; ++k) {
            int j = c->A.col[i][k]; const double* B = c->A.pool + c->A.blk[i][k];
            for (int r = 0; r < c->b; ++r) for (int s = 0; s < c->b; ++s) D[(long)(i*c->b + r)*n + j*c->b + s] = B[r*c->b + s];
        }
        ldlt_factor(&c->dense, D, n); c->has_dense = 1; free(D);
    }
    return mg;
}

static int params_ok(const orc_mg_params* prm)
{
    return prm && !(prm->relaxation < 0 || prm->relaxation > 2 || prm->coarse < 0 || prm->coarse > 1 || prm->restriction < 0 || prm->restriction > 1);
}

void* orc_mg_new(void* p, const orc_mg_params* prm)
{
    oproblem* pr = (oproblem*)p;
    if (!pr || !params_ok(prm)) return NULL;
    if (pr->mesh.qt->numLevels < 2) return NULL;   /* interpolationhierarchy.h:33 */
    int L = 0;
    bsm* Tl = build_interp_qt(pr, &L);
    return mg_setup(&pr->ops, pr->mesh.N, Tl, L, prm, NULL, 1);
}

/* Multigrid<N,P1,P2,...> (multigrid.h:435-639) with InterpolationHierarchy<N,P1,P2,...>:
 * p-levels of the polynomial degrees deg[0] > deg[1] > ... (each below the
 * problem's), then, if hcoarsen, the quadtree h-hierarchy at degree
 * deg[ndeg-1]; otherwise the last p-level is the coarsest
 * (InterpolationHierarchy<N,P,StopCoarsening>) */
void* orc_mg_new_hp(void* p, const orc_mg_params* prm, int ndeg, const int* deg, int hcoarsen)
{
    oproblem* pr = (oproblem*)p;
    if (!pr || !params_ok(prm) || ndeg < 1 || !deg) return NULL;
    int N = pr->mesh.N, Pf = pr->mesh.B.P;
    for (int i = 0; i < ndeg; ++i) if (deg[i] < 0 || deg[i] + 1 >= (i ? deg[i-1] + 1 : Pf)) return NULL;
    if (hcoarsen && pr->mesh.qt->numLevels < 2) return NULL;
    basis_t* Bs = (basis_t*)xcalloc(ndeg, sizeof(basis_t));
    for (int i = 0; i < ndeg; ++i) basis_init(&Bs[i], N, deg[i] + 1);
    int nh = 1;
    bsm* Tl;
    if (hcoarsen) Tl = build_interp_qt_basis(pr, &Bs[ndeg-1], ndeg, &nh);
    else Tl = (bsm*)xcalloc(ndeg, sizeof(bsm));
    for (int i = 0; i < ndeg; ++i) build_pinterp(&Tl[i], pr->ops.ne, N, i ? &Bs[i-1] : &pr->mesh.B, &Bs[i]);
    for (int i = 0; i < ndeg; ++i) basis_free(&Bs[i]);
    free(Bs);
    return mg_setup(&pr->ops, N, Tl, ndeg + nh, prm, NULL, 1);
}

void orc_mg_free(void* p)
{
    omg* mg = (omg*)p; if (!mg) return;
    for (int l = 0; l < mg->L; ++l) {
        olevel* lv = &mg->lev[l];
        bsm_free(&lv->A); bsm_free(&lv->M); bsm_free(&lv->T); for (int d = 0; d < 3; ++d) bsm_free(&lv->G[d]);
        for (int i = 0; i < lv->nb; ++i) { if (lv->Dinv) ldlt_free(&lv->Dinv[i]); if (lv->Minv) ldlt_free(&lv->Minv[i]); }
        free(lv->Dinv); free(lv->Minv); if (lv->has_dense) ldlt_free(&lv->dense);
        free(lv->color); free(lv->x); free(lv->bv); free(lv->r);
    }
    for (int l = 0; l + 1 < mg->L; ++l) bsm_free(&mg->Tl[l]);
    free(mg->Tl); free(mg->lev); free(mg);
}

int orc_mg_nlevels(void* p) { return ((omg*)p)->L; }

int orc_mg_level_dims(void* p, int level, int* nb, int* b)
{
    omg* mg = (omg*)p; if (level < 0 || level >= mg->L) return -1;
    *nb = mg->lev[level].nb; *b = mg->lev[level].b;
    return 0;
}

long orc_mg_export(void* p, int level, int which, int* rowptr, int* cols, double* blocks, int* nbr, int* nbc)
{
    omg* mg = (omg*)p;
    if (level < 0 || level >= mg->L) return -1;
    olevel* lv = &mg->lev[level];
    const bsm* A = NULL;
    if (which == 6) { if (level >= mg->L - 1) return -1; A = &mg->Tl[level]; }
    else if (which == 0) A = &lv->A;
    else if (!lv->has_ops) return -1;
    else if (which == 1) A = &lv->M; else if (which == 2) A = &lv->T;
    else if (which >= 3 && which - 3 < mg->N) A = &lv->G[which-3];
    if (!A) return -1;
    if (nbr) *nbr = A->m;
    if (nbc) *nbc = A->n;
    return bsm_export(A, rowptr, cols, blocks);
}

int orc_mg_colors(void* p, int level, int* colors)
{
    omg* mg = (omg*)p; if (level < 0 || level >= mg->L) return -1;
    if (colors) memcpy(colors, mg->lev[level].color, sizeof(int)*mg->lev[level].nb);
    return mg->lev[level].ncolor;
}

/* Level::computeResidual (multigrid.h:174-178) */
static void lv_residual(olevel* lv) { long n = (long)lv->nb*lv->b; memcpy(lv->r, lv->bv, sizeof(double)*n); bsm_mv(&lv->A, -1.0, lv->x, 1.0, lv->r); }

/* one damped block GS row update (multigrid.h:240-249) */
static void gs_row(olevel* lv, int i, double omega, double* tmp, double* sol)
{
    int b = lv->b; const bsm* A = &lv->A;
    for (int r = 0; r < b; ++r) tmp[r] = 0.0;
    for (int k = 0; k < A->len[i]; ++k) {
        int j = A->col[i][k]; if (j == i) continue;
        const double* Ab = A->pool + A->blk[i][k]; const double* xj = lv->x + (long)j*b;
        for (int r = 0; r < b; ++r) { double s = 0.0; for (int c = 0; c < b; ++c) s += Ab[r*b+c]*xj[c]; tmp[r] += s; }
    }
    for (int r = 0; r < b; ++r) tmp[r] = lv->bv[(long)i*b + r] - tmp[r];
    ldlt_solve(&lv->Dinv[i], tmp, sol);
    for (int r = 0; r < b; ++r) lv->x[(long)i*b + r] += omega*(sol[r] - lv->x[(long)i*b + r]);
}

/* Level::relax (multigrid.h:202-262); relaxation 2 = multicolour GS: colours in
 * increasing order (forward) or decreasing order (reverse) */
static void lv_relax(olevel* lv, const orc_mg_params* prm, int n, int reverse)
{
    int b = lv->b; double tmp[1024], sol[1024];
    for (int s = 0; s < n; ++s) {
        if (prm->relaxation == 0) {
            lv_residual(lv);
            for (int i = 0; i < lv->nb; ++i) {
                ldlt_solve(&lv->Dinv[i], lv->r + (long)i*b, sol);
                for (int r = 0; r < b; ++r) lv->x[(long)i*b + r] += prm->omega*sol[r];
            }
        } else if (prm->relaxation == 1) {
            if (!reverse) for (int i = 0; i < lv->nb; ++i) gs_row(lv, i, prm->omega, tmp, sol);
            else for (int i = lv->nb - 1; i >= 0; --i) gs_row(lv, i, prm->omega, tmp, sol);
        } else {
            for (int cc = 0; cc < lv->ncolor; ++cc) {
                int c = reverse ? lv->ncolor - 1 - cc : cc;
                for (int i = 0; i < lv->nb; ++i) if (lv->color[i] == c) gs_row(lv, i, prm->omega, tmp, sol);
            }
        }
    }
}

/* Level::solve (multigrid.h:181-195) */
static void lv_solve(olevel* lv, const orc_mg_params* prm)
{
    if (prm->coarse == 1) { lv_relax(lv, prm, prm->npre, 0); lv_relax(lv, prm, prm->npost, 1); return; }
    if (lv->has_dense) ldlt_solve(&lv->dense, lv->bv, lv->x);
    else ldlt_solve(&lv->Dinv[0], lv->bv, lv->x);
}

/* Multigrid::vcycle (multigrid.h:361-383) */
static void mg_vcycle(omg* mg, int i)
{
    olevel* lv = &mg->lev[i]; long n = (long)lv->nb*lv->b;
    if (i == mg->L - 1) {
        if (mg->submean) {
            double mean = 0.0; for (long k = 0; k < n; ++k) mean += lv->bv[k]; mean /= (double)n;
            for (long k = 0; k < n; ++k) lv->bv[k] -= mean;
        }
        lv_solve(lv, &mg->prm);
        return;
    }
    lv_relax(lv, &mg->prm, mg->prm.npre, 0);
    lv_residual(lv);
    olevel* c = &mg->lev[i+1];
    bsm_mv_t(&mg->Tl[i], lv->r, c->bv);
    memset(c->x, 0, sizeof(double)*(size_t)c->nb*c->b);
    mg_vcycle(mg, i+1);
    bsm_mv(&mg->Tl[i], 1.0, c->x, 1.0, lv->x);
    lv_relax(lv, &mg->prm, mg->prm.npost, 1);
}

int orc_mg_vcycle(void* p, const double* b, double* x)
{
    omg* mg = (omg*)p; olevel* f = &mg->lev[0]; long n = (long)f->nb*f->b;
    memset(f->x, 0, sizeof(double)*n); memcpy(f->bv, b, sizeof(double)*n);
    mg_vcycle(mg, 0);
    memcpy(x, f->x, sizeof(double)*n);
    return 0;
}

/* The stationary V-cycle solver of the reference driver
 * (tests/cartesian/dg_tests.cpp:309-332): solution() = x, rhs() = b, then
 * ncycle times vcycle() from the current solution (multigrid.h:361-383);
 * rnorm[0] = ||b - A x0||, rnorm[i+1] = ||b - A x|| after cycle i. */
int orc_mg_vcycles(void* p, const double* b, double* x, int ncycle, double* rnorm)
{
    omg* mg = (omg*)p; olevel* f = &mg->lev[0]; long n = (long)f->nb*f->b;
    memcpy(f->x, x, sizeof(double)*n); memcpy(f->bv, b, sizeof(double)*n);
    lv_residual(f);
    if (rnorm) { double s = 0.0; for (long k = 0; k < n; ++k) s += f->r[k]*f->r[k]; rnorm[0] = sqrt(s); }
    for (int i = 0; i < ncycle; ++i) {
        mg_vcycle(mg, 0);
        lv_residual(f);
        if (rnorm) { double s = 0.0; for (long k = 0; k < n; ++k) s += f->r[k]*f->r[k]; rnorm[i+1] = sqrt(s); }
    }
    memcpy(x, f->x, sizeof(double)*n);
    return 0;
}

int orc_mg_relax(void* p, int level, int nsweeps, int reverse, const double* b, double* x)
{
    omg* mg = (omg*)p; if (level < 0 || level >= mg->L) return -1;
    olevel* lv = &mg->lev[level]; long n = (long)lv->nb*lv->b;
    memcpy(lv->bv, b, sizeof(double)*n); memcpy(lv->x, x, sizeof(double)*n);
    lv_relax(lv, &mg->prm, nsweeps, reverse);
    memcpy(x, lv->x, sizeof(double)*n);
    return 0;
}

int orc_mg_spmv(void* p, int level, const double* x, double* y)
{
    omg* mg = (omg*)p; if (level < 0 || level >= mg->L) return -1;
    bsm_mv(&mg->lev[level].A, 1.0, x, 0.0, y);
    return 0;
}

/* ======================================================================
 * PCG (src/cartesian/multigrid/krylov.h:22-68)
 * ====================================================================== */
static double dot(const double* a, const double* b, long n) { double s = 0.0; for (long i = 0; i < n; ++i) s += a[i]*b[i]; return s; }

static void precon(omg* mg, const double* r, double* z, long n)
{
    if (mg) orc_mg_vcycle(mg, r, z); else memcpy(z, r, sizeof(double)*n);
}

int orc_pcg(void* p, void* m, const double* b, double* x, double tol, int maxit,
            int criterion, int flexible, double* stats, double* hist)
{
    oops* ops = (oops*)p; omg* mg = (omg*)m; const bsm* A = &ops->A; long n = ops->ndof;
    double* r = (double*)xmalloc(sizeof(double)*n); double* d = (double*)xmalloc(sizeof(double)*n);
    double* q = (double*)xmalloc(sizeof(double)*n); double* s = (double*)xmalloc(sizeof(double)*n);
    double* sold = flexible ? (double*)xmalloc(sizeof(double)*n) : NULL;
    double bnorm = sqrt(dot(b, b, n));
    memcpy(r, b, sizeof(double)*n); bsm_mv(A, -1.0, x, 1.0, r);
    precon(mg, r, d, n);
    double delta_new = dot(r, d, n), delta0 = delta_new, delta_old, alpha, beta;
    if (flexible) memcpy(sold, d, sizeof(double)*n);
    int i = 0;
    for (;;) {
        int go;
        if (criterion == 0) go = delta_new > tol*tol*delta0;
        else go = sqrt(dot(r, r, n)) > tol*bnorm;
        if (i >= maxit) go = 0;
        if (!go && criterion == 1 && i < maxit && bnorm > 0) {
            /* confirm on the true residual; restart the direction if it fails */
            memcpy(r, b, sizeof(double)*n); bsm_mv(A, -1.0, x, 1.0, r);
            if (sqrt(dot(r, r, n)) > tol*bnorm && delta_new == delta_new) {
                precon(mg, r, d, n); delta_new = dot(r, d, n);
                if (flexible) memcpy(sold, d, sizeof(double)*n);
                go = 1;
            }
        }
        if (!go) break;
        bsm_mv(A, 1.0, d, 0.0, q);
        alpha = delta_new/dot(d, q, n);
        for (long k = 0; k < n; ++k) x[k] += alpha*d[k];
        if (i % 50 == 0) { memcpy(r, b, sizeof(double)*n); bsm_mv(A, -1.0, x, 1.0, r); }
        else for (long k = 0; k < n; ++k) r[k] -= alpha*q[k];
        precon(mg, r, s, n);
        delta_old = delta_new; delta_new = dot(r, s, n);
        if (flexible) { beta = (delta_new - dot(r, sold, n))/delta_old; memcpy(sold, s, sizeof(double)*n); }
        else beta = delta_new/delta_old;
        for (long k = 0; k < n; ++k) d[k] = s[k] + beta*d[k];
        if (hist) hist[i] = sqrt(dot(r, r, n));
        ++i;
    }
    memcpy(r, b, sizeof(double)*n); bsm_mv(A, -1.0, x, 1.0, r);
    stats[0] = i;
    stats[1] = bnorm > 0 ? sqrt(dot(r, r, n))/bnorm : 0.0;
    stats[2] = delta0 > 0 ? sqrt(fabs(delta_new/delta0)) : 0.0;
    free(r); free(d); free(q); free(s); free(sold);
    return 0;
}
