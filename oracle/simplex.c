/*
 * simplex.c — CPU FP64 restatement (the oracle) of multigrid-ldg's unstructured
 * simplex LDG path and of an agglomerated multigrid hierarchy on it.
 *
 * TEST INFRASTRUCTURE ONLY (compiled into liboracle.so together with cart.c).
 *
 * Reference files restated here (paths relative to the reference root):
 *   src/unstructured/dg/master.h          Koornwinder / Jacobi basis, nodal
 *                                          basis on equispaced simplex nodes
 *   src/unstructured/util/range.h:258-330 simplex multi-index order
 *   src/unstructured/mesh/wireframe.h     .mesh reader
 *   src/unstructured/mesh/geometry.h      Simplex (Jacobian, volume)
 *   src/unstructured/mesh/mesh.h          faces, normals, slices, lift
 *   src/unstructured/dg/ldgpoisson.tpp    LDG assembly and RHS
 *   src/unstructured/mesh/agglomeration.h MGridGen graph of a mesh / agglomerate mesh
 *   src/unstructured/multigrid/interpolationhierarchy.h:88-204
 *                                          agglomerated interpolation (Q_p on boxes)
 * Departures, each a definition where the reference is unspecified or does
 * not compile:
 *   - faces are numbered in first-appearance order (element i, local face j);
 *     the reference uses unordered_map iteration order;
 *   - agglomerate bounding boxes are the boxes of their vertices (the
 *     reference initialises them to [0,1]^N before taking min/max);
 *   - the dual graph lists face f of an element as +-(f+1); the reference's
 *     +-f (agglomeration.h:61,141) cannot tell the sides of face 0 apart and
 *     makes its right element a neighbour of itself;
 *   - the hierarchy's operators T0, T_l take the reference's evident intent
 *     (fine rows x coarse columns), which does not compile as written
 *     (interpolationhierarchy.h:96,150,217).
 */
#include <float.h>

#define SX_MAXNPL 64
#define SX_MAXQ 64

static long ichoose_(int n, int k) { if (k < 0 || k > n) return 0; long r = 1; for (int i = 1; i <= k; ++i) r = r*(n - k + i)/i; return r; }
static long ifac_(int n) { long r = 1; for (int i = 2; i <= n; ++i) r *= i; return r; }

/* multi-indices of SimplexRange<N,P>: entries >= 0, sum <= P-1, last
 * coordinate fastest (range.h:291-309) */
static int simplex_indices(int N, int P, int* idx)
{
    int cur[3] = {0, 0, 0}, n = 0;
    for (;;) {
        for (int k = 0; k < N; ++k) idx[n*N + k] = cur[k];
        ++n;
        int s;
        for (s = N - 1; s >= 0; --s) {
            int sum = 0; for (int k = 0; k < N; ++k) sum += cur[k];
            int other = sum - cur[s];
            if (++cur[s] != P - other) break;
            cur[s] = 0;
        }
        if (s < 0) break;
    }
    return n;
}

/* normalised Jacobi polynomial P_n^{(a,b)} at x (master.h:97-137) */
static double jacobi1(int n, int a, int b, double x)
{
    double a1 = a + 1, b1 = b + 1, ab = a + b, ab1 = a + b + 1;
    double gamma0 = pow(2.0, ab1)/ab1*tgamma(a1)*tgamma(b1)/tgamma(ab1);
    double p0 = 1.0/sqrt(gamma0);
    if (n == 0) return p0;
    double gamma1 = a1*b1/(ab + 3.0)*gamma0;
    double p1 = ((ab + 2.0)*x/2.0 + (a - b)/2.0)/sqrt(gamma1);
    if (n == 1) return p1;
    double aold = 2.0/(2.0 + ab)*sqrt(a1*b1/(ab + 3.0));
    for (int i = 1; i < n; ++i) {
        double h1 = 2.0*i + ab;
        double anew = 2.0/(h1 + 2.0)*sqrt((i + 1)*(i + ab1)*(i + a1)*(i + b1)/(h1 + 1.0)/(h1 + 3.0));
        double bnew = -(a*a - b*b)/h1/(h1 + 2.0);
        double p2 = ((x - bnew)*p1 - aold*p0)/anew;
        p0 = p1; p1 = p2; aold = anew;
    }
    return p1;
}

static double djacobi1(int n, int a, int b, double x)
{
    return n == 0 ? 0.0 : sqrt((double)n*(n + a + b + 1))*jacobi1(n - 1, a + 1, b + 1, x);
}

/* collapsed coordinates (master.h:148-160) */
static void collapse(int N, const double* xn, double* A)
{
    double X[3];
    for (int i = 0; i < N; ++i) X[i] = 2*xn[i] - 1;
    for (int i = 0; i < N - 1; ++i) {
        double denom = 3 - N + i;
        for (int j = i + 1; j < N; ++j) denom -= X[j];
        A[i] = (denom != 0) ? 2*(1.0 + X[i])/denom - 1 : -1;
    }
    A[N-1] = X[N-1];
}

/* koornwinder (master.h:146-180) */
static double koornwinder1(int N, const double* x, const int* porder)
{
    double A[3]; collapse(N, x, A);
    double v = 1.0; int poly = 0, power = 0;
    for (int i = 0; i < N; ++i) {
        v *= jacobi1(porder[i], poly, 0, A[i]);
        v *= pow(1 - A[i], power);
        poly += 2*porder[i] + 1; power += porder[i];
    }
    return v*pow(2.0, ichoose_(N + 1, 2)/2.0);
}

/* dkoornwinder (master.h:196-260) */
static double dkoornwinder1(int N, const double* x, const int* porder, int dim)
{
    double A[3]; collapse(N, x, A);
    double dv = 0.0;
    for (int d = 0; d < dim + 1; ++d) {
        double chain = (d != dim) ? 0.5*(1 + A[d]) : 1.0;
        int divide = 0; for (int i = d + 1; i < N; ++i) divide |= 1 << i;
        double da = 0.0, fac = 1.0; int poly = 0, power = 0;
        for (int i = 0; i < N; ++i) {
            double a = A[i];
            if (i != d) {
                fac *= jacobi1(porder[i], poly, 0, a);
                if (power > 0) {
                    if (divide & (1 << i)) fac = 2*fac*pow(1 - a, power - 1);
                    else fac *= pow(1 - a, power);
                }
            } else {
                da = djacobi1(porder[i], poly, 0, a);
                if (power > 0) {
                    da *= pow(1 - a, power);
                    da -= jacobi1(porder[i], poly, 0, a)*power*pow(1 - a, power - 1);
                }
            }
            poly += 2*porder[i] + 1; power += porder[i];
        }
        dv += chain*fac*da;
    }
    return dv*pow(2.0, ichoose_(N + 1, 2)/2.0 + 1);
}

/* Gauss-Jordan inverse (the reference uses Eigen's inverse(), master.h:313) */
static void dense_inv(const double* A, int n, double* X)
{
    double* W = (double*)xmalloc(sizeof(double)*n*2*n);
    for (int i = 0; i < n; ++i) for (int j = 0; j < 2*n; ++j) W[i*2*n + j] = (j < n) ? A[i*n + j] : (j - n == i ? 1.0 : 0.0);
    for (int k = 0; k < n; ++k) {
        int p = k; for (int i = k + 1; i < n; ++i) if (fabs(W[i*2*n + k]) > fabs(W[p*2*n + k])) p = i;
        if (p != k) for (int j = 0; j < 2*n; ++j) { double t = W[k*2*n + j]; W[k*2*n + j] = W[p*2*n + j]; W[p*2*n + j] = t; }
        double d = W[k*2*n + k];
        for (int j = 0; j < 2*n; ++j) W[k*2*n + j] /= d;
        for (int i = 0; i < n; ++i) if (i != k) { double f = W[i*2*n + k]; if (f != 0) for (int j = 0; j < 2*n; ++j) W[i*2*n + j] -= f*W[k*2*n + j]; }
    }
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) X[i*n + j] = W[i*2*n + n + j];
    free(W);
}

/* the simplex master element Master<N,P> (master.h:12-43, 280-356) */
typedef struct {
    int N, P, npl;
    int idx[SX_MAXNPL*3];
    double nodes[SX_MAXNPL*3];
    double V[SX_MAXNPL*SX_MAXNPL], Vinv[SX_MAXNPL*SX_MAXNPL];
    double mass[SX_MAXNPL*SX_MAXNPL], invmass[SX_MAXNPL*SX_MAXNPL];
    double diff[3][SX_MAXNPL*SX_MAXNPL];
    double volume;
} sxmaster;

static void sxmaster_init(sxmaster* m, int N, int P)
{
    memset(m, 0, sizeof(*m));
    m->N = N; m->P = P;
    m->npl = simplex_indices(N, P, m->idx);
    int n = m->npl;
    m->volume = 1.0/ifac_(N);
    for (int i = 0; i < n; ++i) for (int k = 0; k < N; ++k)
        m->nodes[i*N + k] = (P == 1) ? 1.0/3.0 : (double)m->idx[i*N + k]/(P - 1);
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) m->V[i*n + j] = koornwinder1(N, m->nodes + i*N, m->idx + j*N);
    dense_inv(m->V, n, m->Vinv);
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) { double s = 0; for (int k = 0; k < n; ++k) s += m->Vinv[k*n + i]*m->Vinv[k*n + j]; m->mass[i*n + j] = s; }
    dense_inv(m->mass, n, m->invmass);
    double dV[SX_MAXNPL*SX_MAXNPL];
    for (int d = 0; d < N; ++d) {
        for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) dV[i*n + j] = dkoornwinder1(N, m->nodes + i*N, m->idx + j*N, d);
        gemm(0, 0, n, n, n, 1.0, dV, n, m->Vinv, n, 0.0, m->diff[d], n);
    }
}

/* nodal basis at a point: Vinv^T K(x) (Phi::eval, master.h:50-59) */
static void sxmaster_eval(const sxmaster* m, const double* x, double* phi)
{
    int n = m->npl; double K[SX_MAXNPL];
    for (int j = 0; j < n; ++j) K[j] = koornwinder1(m->N, x, m->idx + j*m->N);
    for (int i = 0; i < n; ++i) { double s = 0; for (int j = 0; j < n; ++j) s += m->Vinv[j*n + i]*K[j]; phi[i] = s; }
}

/* ======================================================================
 * Mesh<N,P>(wireframe, bndfuns) (mesh.h:185-298) and Face (mesh.h:22-121)
 * ====================================================================== */
typedef struct {
    int left, right;           /* right < 0: boundary, tag = -right */
    int v[3];                  /* sorted vertex ids */
    double normal[3];
    double area;
    int sl[SX_MAXNPL], sr[SX_MAXNPL];   /* slices (interior): face node -> element node */
    double* xl;                /* boundary: quadrature points in the left element's reference coords */
} sxface;

typedef struct {
    oops ops;
    int N, P, np, ne, nf, nbf, nq, nqb;
    double* p; int* t;         /* points (np x N), simplices (ne x N+1), 0-based */
    sxmaster E, F;             /* element and face masters */
    double qx[SX_MAXQ*2], qw[SX_MAXQ];   /* face quadrature (reference face coords, N-1 per point) */
    sxface* faces;
    int* bface;
    double* vol; double* Jinv; /* per element: volume, inverse Jacobian (N x N row-major) */
    int bctype[16];
    double tau0, tauD;
} sxproblem;

static void simplex_jac(const double* p, const int* t, int N, double* J)
{   /* J(:,i) = p[t[i+1]] - p[t[0]] (geometry.h:30-38), row-major N x N */
    for (int r = 0; r < N; ++r) for (int i = 0; i < N; ++i) J[r*N + i] = p[t[i+1]*N + r] - p[t[0]*N + r];
}

static double det_n(const double* J, int N)
{
    if (N == 1) return J[0];
    if (N == 2) return J[0]*J[3] - J[1]*J[2];
    return J[0]*(J[4]*J[8] - J[5]*J[7]) - J[1]*(J[3]*J[8] - J[5]*J[6]) + J[2]*(J[3]*J[7] - J[4]*J[6]);
}

/* The reference's Quadrature<2,10> (master.h:276, 289, 302): Dunavant's
 * degree-9 rule with 19 points on the unit triangle, weights summing to 1, as
 * the centroid, four orbits (a,b), (b,a), (b,b) and one orbit (a,b), (c,a),
 * (b,c), (a,c), (c,b), (b,a), with the table's 15-digit values. */
static int triangle_rule19(double* x, double* w)
{
    static const double o3[4][3] = {   /* a, b, weight */
        {2.06349616025250E-02, 4.89682519198738E-01, 0.031334700227139},
        {1.25820817014127E-01, 4.37089591492937E-01, 0.077827541004774},
        {6.23592928761935E-01, 1.88203535619033E-01, 0.079647738927210},
        {9.10540973211095E-01, 4.47295133944530E-02, 0.025577675658698}};
    const double a6 = 3.68384120547360E-02, b6 = 2.21962989160766E-01, c6 = 7.41198598784498E-01, w6 = 0.043283539377289;
    int n = 0;
    x[0] = 3.33333333333333E-01; x[1] = 3.33333333333333E-01; w[n++] = 0.097135796282799;
    for (int k = 0; k < 4; ++k) {
        double a = o3[k][0], b = o3[k][1];
        x[2*n] = a; x[2*n+1] = b; w[n++] = o3[k][2];
        x[2*n] = b; x[2*n+1] = a; w[n++] = o3[k][2];
        x[2*n] = b; x[2*n+1] = b; w[n++] = o3[k][2];
    }
    const double p6[6][2] = {{a6, b6}, {c6, a6}, {b6, c6}, {a6, c6}, {c6, b6}, {b6, a6}};
    for (int k = 0; k < 6; ++k) { x[2*n] = p6[k][0]; x[2*n+1] = p6[k][1]; w[n++] = w6; }
    return n;
}

/* face quadrature on the reference (N-1)-simplex with weights summing to 1
 * (Quadrature<N-1,Q>, Q = 10, master.h:359-442): the 10-point Gauss rule on
 * [0,1] for triangle meshes, the 19-point triangle rule for tetrahedra */
static int face_quadrature(int N, double* x, double* w)
{
    if (N == 2) { orc_gauss(QPTS, x, w); return QPTS; }
    if (N == 3) return triangle_rule19(x, w);
    return -1;
}

int orc_sx_face_rule(int N, double* x, double* w)
{
    double xx[SX_MAXQ*2], ww[SX_MAXQ];
    int n = face_quadrature(N, xx, ww);
    if (n > 0) { if (x) memcpy(x, xx, sizeof(double)*n*(N - 1)); if (w) memcpy(w, ww, sizeof(double)*n); }
    return n;
}

/* phys = p[v0] + sum_k (p[v_{k+1}] - p[v0]) ref[k]: the face simplex map
 * (Simplex<N-1,N>::jacobian_mat, geometry.h:30-38) */
static void face_map(const sxproblem* sx, const sxface* g, const double* ref, double* phys)
{
    int N = sx->N;
    for (int r = 0; r < N; ++r) {
        double s = sx->p[g->v[0]*N + r];
        for (int k = 0; k < N - 1; ++k) s += (sx->p[g->v[k+1]*N + r] - sx->p[g->v[0]*N + r])*ref[k];
        phys[r] = s;
    }
}

static void to_local(const sxproblem* sx, int el, const double* phys, double* loc)
{
    int N = sx->N; const double* I = sx->Jinv + el*N*N; const int* te = sx->t + el*(N+1);
    for (int r = 0; r < N; ++r) { double s = 0; for (int c = 0; c < N; ++c) s += I[r*N + c]*(phys[c] - sx->p[te[0]*N + c]); loc[r] = s; }
}

/* vmask[v] bit k: boundary function k holds at vertex v; a boundary face gets
 * tag k+1 for the first k holding at all of its vertices (mesh.h:276-288);
 * nbnd = 0: every boundary face has tag 1.  bctype[tag-1]: 0 Dirichlet, 1 Neumann. */
void* orc_sx_problem_new(int N, int np, const double* pts, int nt, const int* tris, int nbnd, const int* vmask,
                         int p, const int* bctype, double tau0, double tauD)
{
    if (N < 2 || N > 3 || p < 0 || p > 5 || np < N + 1 || nt < 1 || nbnd < 0 || nbnd > 16) return NULL;
    int nbtypes = nbnd ? nbnd : 1;
    int P = p + 1;
    sxproblem* sx = (sxproblem*)xcalloc(1, sizeof(sxproblem));
    sx->N = N; sx->P = P; sx->np = np; sx->ne = nt; sx->tau0 = tau0; sx->tauD = tauD;
    for (int k = 0; k < nbtypes && k < 16; ++k) sx->bctype[k] = bctype ? bctype[k] : 0;
    sx->p = (double*)xmalloc(sizeof(double)*np*N); memcpy(sx->p, pts, sizeof(double)*np*N);
    sx->t = (int*)xmalloc(sizeof(int)*nt*(N+1)); memcpy(sx->t, tris, sizeof(int)*nt*(N+1));
    sxmaster_init(&sx->E, N, P); sxmaster_init(&sx->F, N - 1, P);
    sx->nq = face_quadrature(N, sx->qx, sx->qw);
    if (sx->nq < 0) { orc_sx_problem_free(sx); return NULL; }
    int b = sx->E.npl;
    /* elements */
    sx->vol = (double*)xmalloc(sizeof(double)*nt); sx->Jinv = (double*)xmalloc(sizeof(double)*nt*N*N);
    for (int e = 0; e < nt; ++e) {
        double J[9]; simplex_jac(sx->p, sx->t + e*(N+1), N, J);
        double d = det_n(J, N);
        sx->vol[e] = fabs(d)/ifac_(N);   /* sqrt(det(J^T J))/N! (geometry.h:65-74) */
        dense_inv(J, N, sx->Jinv + e*N*N);
    }
    /* faces: sorted vertex tuples, first-appearance order */
    int cap = nt*(N+1); sx->faces = (sxface*)xcalloc(cap, sizeof(sxface));
    long hsize = 1; while (hsize < 4L*cap) hsize <<= 1;
    int* hkey = (int*)xmalloc(sizeof(int)*hsize); for (long i = 0; i < hsize; ++i) hkey[i] = -1;
    int nf = 0;
    for (int e = 0; e < nt; ++e) for (int j = 0; j < N + 1; ++j) {
        int f[3];
        for (int k = 0; k < N; ++k) f[k] = sx->t[e*(N+1) + (k + j)%(N + 1)];
        for (int a = 0; a < N; ++a) for (int c = a + 1; c < N; ++c) if (f[c] < f[a]) { int tmp = f[a]; f[a] = f[c]; f[c] = tmp; }
        unsigned long h = 1469598103934665603UL;
        for (int k = 0; k < N; ++k) { h ^= (unsigned long)f[k]; h *= 1099511628211UL; }
        long slot = (long)(h & (hsize - 1));
        int found = -1;
        while (hkey[slot] >= 0) {
            sxface* g = &sx->faces[hkey[slot]];
            int same = 1; for (int k = 0; k < N; ++k) if (g->v[k] != f[k]) same = 0;
            if (same) { found = hkey[slot]; break; }
            slot = (slot + 1) & (hsize - 1);
        }
        if (found < 0) { hkey[slot] = nf; sxface* g = &sx->faces[nf++]; for (int k = 0; k < N; ++k) g->v[k] = f[k]; g->left = e; g->right = -1; }
        else sx->faces[found].right = e;
    }
    free(hkey);
    sx->nf = nf;
    int nbf = 0;
    for (int fi = 0; fi < nf; ++fi) {
        sxface* g = &sx->faces[fi];
        int id1 = g->left, id2 = g->right;
        int mn = id1 < id2 ? id1 : id2, mx = id1 < id2 ? id2 : id1;
        g->left = (mn >= 0) ? mn : mx; g->right = (mn >= 0) ? mx : mn;
        /* normal from the cofactors of the face Jacobian (mesh.h:249-263) */
        double fp[3][3];
        for (int k = 0; k < N; ++k) for (int r = 0; r < N; ++r) fp[k][r] = sx->p[g->v[k]*N + r];
        if (N == 2) { double dx = fp[1][0] - fp[0][0], dy = fp[1][1] - fp[0][1]; g->normal[0] = dy; g->normal[1] = -dx; }
        else {
            double a[3], c[3];
            for (int r = 0; r < 3; ++r) { a[r] = fp[1][r] - fp[0][r]; c[r] = fp[2][r] - fp[0][r]; }
            g->normal[0] = a[1]*c[2] - a[2]*c[1]; g->normal[1] = a[2]*c[0] - a[0]*c[2]; g->normal[2] = a[0]*c[1] - a[1]*c[0];
        }
        double nn = 0; for (int r = 0; r < N; ++r) nn += g->normal[r]*g->normal[r];
        nn = sqrt(nn);
        g->area = nn/ifac_(N - 1);   /* sqrt(gramian)/(N-1)!: the cofactor norm is sqrt of the Gram determinant */
        for (int r = 0; r < N; ++r) g->normal[r] /= nn;
        /* outward from the left element (mesh.h:266-273) */
        const int* tl = sx->t + g->left*(N+1); int a = tl[0];
        for (int i = 0; i < N + 1; ++i) { int in = 0; for (int k = 0; k < N; ++k) if (tl[i] == g->v[k]) in = 1; if (!in) a = tl[i]; }
        double dot = 0; for (int r = 0; r < N; ++r) dot += g->normal[r]*(sx->p[a*N + r] - fp[0][r]);
        if (dot > 0) for (int r = 0; r < N; ++r) g->normal[r] = -g->normal[r];
        if (g->right < 0) {
            int tag = 1;
            if (nbnd) {
                tag = 0;
                for (int k = 0; k < nbnd && !tag; ++k) {
                    int all = 1; for (int q = 0; q < N; ++q) if (!(vmask[g->v[q]] >> k & 1)) all = 0;
                    if (all) tag = k + 1;
                }
                if (!tag) { orc_sx_problem_free(sx); return NULL; }   /* no boundary function holds on the face */
            }
            g->right = -tag; ++nbf;
        }
    }
    sx->nbf = nbf; sx->nqb = nbf*sx->nq;
    sx->bface = (int*)xmalloc(sizeof(int)*(nbf ? nbf : 1));
    for (int fi = 0, k = 0; fi < nf; ++fi) if (sx->faces[fi].right < 0) sx->bface[k++] = fi;
    /* slices (interior) and boundary quadrature points (Face ctor, mesh.h:29-76) */
    const sxmaster* FM = &sx->F;
    for (int fi = 0; fi < nf; ++fi) {
        sxface* g = &sx->faces[fi];
        if (g->right >= 0) {
            if (P == 1) { g->sl[0] = 0; g->sr[0] = 0; continue; }
            for (int i = 0; i < FM->npl; ++i) {
                double phys[3];
                face_map(sx, g, FM->nodes + i*(N - 1), phys);
                for (int side = 0; side < 2; ++side) {
                    int el = side ? g->right : g->left;
                    double loc[3];
                    to_local(sx, el, phys, loc);
                    int m = -1;
                    for (int j = 0; j < b && m < 0; ++j) {
                        int ok = 1; for (int r = 0; r < N; ++r) if (!(fabs(loc[r] - sx->E.nodes[j*N + r]) < 1e-6)) ok = 0;
                        if (ok) m = j;
                    }
                    if (side) g->sr[i] = m; else g->sl[i] = m;
                }
            }
        } else {
            g->xl = (double*)xmalloc(sizeof(double)*sx->nq*N);
            for (int q = 0; q < sx->nq; ++q) {
                double phys[3];
                face_map(sx, g, sx->qx + q*(N - 1), phys);
                to_local(sx, g->left, phys, g->xl + q*N);
            }
        }
    }
    /* ---- assembly (ldgpoisson.tpp:26-99) ---- */
    oops* ops = &sx->ops;
    ops->N = N; ops->ne = nt; ops->b = b; ops->ndof = nt*b;
    bsm_init(&ops->M, nt, nt, b, b); bsm_init(&ops->T, nt, nt, b, b);
    for (int d = 0; d < N; ++d) bsm_init(&ops->G[d], nt, nt, b, b);
    for (int e = 0; e < nt; ++e) bsm_add(&ops->M, e, e, sx->vol[e]/sx->E.volume, sx->E.mass);
    for (int e = 0; e < nt; ++e) {
        const double* I = sx->Jinv + e*N*N;
        for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) bsm_add(&ops->G[i], e, e, I[j*N + i], sx->E.diff[j]);
    }
    double* LL = (double*)xmalloc(sizeof(double)*b*b); double* RR = (double*)xmalloc(sizeof(double)*b*b);
    double* RL = (double*)xmalloc(sizeof(double)*b*b); double* W = (double*)xmalloc(sizeof(double)*b*b);
    double* MI = (double*)xmalloc(sizeof(double)*b*b);
    for (int fi = 0; fi < nf; ++fi) {
        sxface* g = &sx->faces[fi];
        if (g->right < 0 && sx->bctype[-g->right - 1] != 0) continue;
        orc_sx_lift(sx, fi, NULL, LL, RR, RL);
        if (g->right >= 0) {
            for (int k = 0; k < b*b; ++k) MI[k] = sx->E.invmass[k]*(sx->E.volume/sx->vol[g->right]);
            for (int i = 0; i < N; ++i) {
                double n = g->normal[i];
                gemm(0, 0, b, b, b, n, MI, b, RR, b, 0.0, W, b); bsm_add(&ops->G[i], g->right, g->right, 1.0, W);
                gemm(0, 0, b, b, b, -n, MI, b, RL, b, 0.0, W, b); bsm_add(&ops->G[i], g->right, g->left, 1.0, W);
            }
            bsm_add(&ops->T, g->right, g->right, tau0, RR);
            bsm_add(&ops->T, g->left, g->left, tau0, LL);
            bsm_add(&ops->T, g->right, g->left, -tau0, RL);
            for (int i = 0; i < b; ++i) for (int j = 0; j < b; ++j) W[i*b + j] = RL[j*b + i];
            bsm_add(&ops->T, g->left, g->right, -tau0, W);
        } else {
            for (int k = 0; k < b*b; ++k) MI[k] = sx->E.invmass[k]*(sx->E.volume/sx->vol[g->left]);
            for (int i = 0; i < N; ++i) {
                gemm(0, 0, b, b, b, -g->normal[i], MI, b, LL, b, 0.0, W, b); bsm_add(&ops->G[i], g->left, g->left, 1.0, W);
            }
            bsm_add(&ops->T, g->left, g->left, tauD, LL);
        }
    }
    free(LL); free(RR); free(RL); free(W); free(MI);
    construct_laplacian(&ops->A, &ops->T, &ops->M, ops->G, N);
    return sx;
}

/* Mesh::lift (mesh.h:327-349).  Interior: face mass on the sliced nodes.
 * Boundary: quadrature with weights * area * |reference face| (as written,
 * mesh.h:343-344).  L is npl x nq. */
int orc_sx_lift(void* p, int fi, double* L, double* LL, double* RR, double* RL)
{
    sxproblem* sx = (sxproblem*)p; const sxface* g = &sx->faces[fi];
    int b = sx->E.npl, nfl = sx->F.npl;
    if (g->right >= 0) {
        double jac = g->area/sx->F.volume;
        if (LL) memset(LL, 0, sizeof(double)*b*b);
        if (RR) memset(RR, 0, sizeof(double)*b*b);
        if (RL) memset(RL, 0, sizeof(double)*b*b);
        for (int i = 0; i < nfl; ++i) for (int j = 0; j < nfl; ++j) {
            double v = sx->F.mass[i*nfl + j]*jac;
            if (LL) LL[g->sl[i]*b + g->sl[j]] += v;
            if (RR) RR[g->sr[i]*b + g->sr[j]] += v;
            if (RL) RL[g->sr[i]*b + g->sl[j]] += v;
        }
        return 0;
    }
    int nq = sx->nq; double phi[SX_MAXQ][SX_MAXNPL], w[SX_MAXQ];
    for (int q = 0; q < nq; ++q) { sxmaster_eval(&sx->E, g->xl + q*sx->N, phi[q]); w[q] = sx->qw[q]*g->area*sx->F.volume; }
    if (L) for (int i = 0; i < b; ++i) for (int q = 0; q < nq; ++q) L[i*nq + q] = phi[q][i]*w[q];
    if (LL) for (int i = 0; i < b; ++i) for (int j = 0; j < b; ++j) { double s = 0; for (int q = 0; q < nq; ++q) s += phi[q][i]*w[q]*phi[q][j]; LL[i*b + j] = s; }
    return 0;
}

void orc_sx_problem_free(void* p)
{
    sxproblem* sx = (sxproblem*)p; if (!sx) return;
    bsm_free(&sx->ops.M); bsm_free(&sx->ops.T); bsm_free(&sx->ops.A); for (int d = 0; d < 3; ++d) bsm_free(&sx->ops.G[d]);
    for (int f = 0; f < sx->nf; ++f) free(sx->faces[f].xl);
    free(sx->faces); free(sx->bface); free(sx->p); free(sx->t); free(sx->vol); free(sx->Jinv); free(sx);
}

/* sizes[0..6] = ne, nf, nbf, npl, nqb, ndof, nq */
int orc_sx_sizes(void* p, int* s)
{
    sxproblem* sx = (sxproblem*)p;
    s[0] = sx->ne; s[1] = sx->nf; s[2] = sx->nbf; s[3] = sx->E.npl; s[4] = sx->nqb; s[5] = sx->ops.ndof; s[6] = sx->nq;
    return 0;
}

/* Master::dgnodes per element (master.h:36-42) */
int orc_sx_nodes(void* p, double* xyz)
{
    sxproblem* sx = (sxproblem*)p; int N = sx->N, b = sx->E.npl;
    for (int e = 0; e < sx->ne; ++e) {
        double J[9]; simplex_jac(sx->p, sx->t + e*(N+1), N, J);
        const double* p0 = sx->p + sx->t[e*(N+1)]*N;
        for (int i = 0; i < b; ++i) for (int r = 0; r < N; ++r) {
            double s = p0[r]; for (int c = 0; c < N; ++c) s += J[r*N + c]*sx->E.nodes[i*N + c];
            xyz[((long)e*b + i)*N + r] = s;
        }
    }
    return 0;
}

/* faces: lr_tag[nf*2] = left, right ; normal[nf*N] ; verts[nf*N] */
int orc_sx_faces(void* p, int* lr, double* normal, int* verts)
{
    sxproblem* sx = (sxproblem*)p; int N = sx->N;
    for (int f = 0; f < sx->nf; ++f) {
        if (lr) { lr[2*f] = sx->faces[f].left; lr[2*f+1] = sx->faces[f].right; }
        for (int r = 0; r < N; ++r) { if (normal) normal[f*N + r] = sx->faces[f].normal[r]; if (verts) verts[f*N + r] = sx->faces[f].v[r]; }
    }
    return 0;
}

/* boundary quadrature points (ldgpoisson.tpp:119-122), boundary-face order */
int orc_sx_bc_qpoints(void* p, double* xyz, int* tag)
{
    sxproblem* sx = (sxproblem*)p; int N = sx->N, nq = sx->nq;
    for (int k = 0; k < sx->nbf; ++k) {
        const sxface* g = &sx->faces[sx->bface[k]];
        for (int q = 0; q < nq; ++q) {
            face_map(sx, g, sx->qx + q*(N - 1), xyz + ((long)k*nq + q)*N);
            if (tag) tag[(long)k*nq + q] = -g->right;
        }
    }
    return 0;
}

/* add_source_terms + computeRHS (ldgpoisson.tpp:101-170) */
int orc_sx_rhs(void* p, const double* fnod, const double* gq, double* out)
{
    sxproblem* sx = (sxproblem*)p; int N = sx->N, b = sx->E.npl, nq = sx->nq; long nd = sx->ops.ndof;
    double* rhs = (double*)xcalloc(nd, sizeof(double)); double* Jg[3] = {0, 0, 0};
    for (int d = 0; d < N; ++d) Jg[d] = (double*)xcalloc(nd, sizeof(double));
    double* L = (double*)xmalloc(sizeof(double)*b*nq);
    for (int k = 0; k < sx->nbf; ++k) {
        int fi = sx->bface[k]; const sxface* g = &sx->faces[fi];
        orc_sx_lift(sx, fi, L, NULL, NULL, NULL);
        double Lg[SX_MAXNPL];
        for (int i = 0; i < b; ++i) { double s = 0; if (gq) for (int q = 0; q < nq; ++q) s += L[i*nq + q]*gq[(long)k*nq + q]; Lg[i] = s; }
        long off = (long)g->left*b;
        if (sx->bctype[-g->right - 1] == 0) {
            for (int d = 0; d < N; ++d) for (int i = 0; i < b; ++i) Jg[d][off + i] += g->normal[d]*Lg[i];
            for (int i = 0; i < b; ++i) rhs[off + i] += sx->tauD*Lg[i];
        } else for (int i = 0; i < b; ++i) rhs[off + i] += Lg[i];
    }
    double* t = (double*)xmalloc(sizeof(double)*nd);
    for (int d = 0; d < N; ++d) { bsm_mv_t(&sx->ops.G[d], Jg[d], t); for (long i = 0; i < nd; ++i) rhs[i] -= t[i]; }
    for (int e = 0; e < sx->ne; ++e) {
        const double* Me = bsm_get(&sx->ops.M, e, e);
        for (int i = 0; i < b; ++i) { double s = 0; if (fnod) for (int j = 0; j < b; ++j) s += Me[i*b + j]*fnod[(long)e*b + j]; out[(long)e*b + i] = rhs[(long)e*b + i] + s; }
    }
    free(t); free(L); free(rhs); for (int d = 0; d < N; ++d) free(Jg[d]);
    return 0;
}

/* transpose of the g -> rhs map */
int orc_sx_rhs_g_adjoint(void* p, const double* lam, double* dg)
{
    sxproblem* sx = (sxproblem*)p; int N = sx->N, b = sx->E.npl, nq = sx->nq; long nd = sx->ops.ndof;
    double* Glam[3] = {0, 0, 0};
    for (int d = 0; d < N; ++d) { Glam[d] = (double*)xmalloc(sizeof(double)*nd); bsm_mv(&sx->ops.G[d], 1.0, lam, 0.0, Glam[d]); }
    double* L = (double*)xmalloc(sizeof(double)*b*nq);
    for (int k = 0; k < sx->nbf; ++k) {
        int fi = sx->bface[k]; const sxface* g = &sx->faces[fi];
        orc_sx_lift(sx, fi, L, NULL, NULL, NULL);
        long off = (long)g->left*b; double v[SX_MAXNPL];
        for (int i = 0; i < b; ++i) {
            if (sx->bctype[-g->right - 1] == 0) { double s = sx->tauD*lam[off + i]; for (int d = 0; d < N; ++d) s -= g->normal[d]*Glam[d][off + i]; v[i] = s; }
            else v[i] = lam[off + i];
        }
        for (int q = 0; q < nq; ++q) { double s = 0; for (int i = 0; i < b; ++i) s += L[i*nq + q]*v[i]; dg[(long)k*nq + q] = s; }
    }
    free(L); for (int d = 0; d < N; ++d) free(Glam[d]);
    return 0;
}

long orc_sx_export(void* p, int which, int* rowptr, int* cols, double* blocks)
{
    sxproblem* sx = (sxproblem*)p; oops* o = &sx->ops;
    const bsm* A = which == 0 ? &o->A : which == 1 ? &o->M : which == 2 ? &o->T : (which >= 3 && which - 3 < sx->N) ? &o->G[which-3] : NULL;
    if (!A) return -1;
    return bsm_export(A, rowptr, cols, blocks);
}

int orc_sx_master(int N, int P, double* nodes, double* mass, double* diff)
{
    sxmaster m; sxmaster_init(&m, N, P); int n = m.npl;
    if (nodes) memcpy(nodes, m.nodes, sizeof(double)*n*N);
    if (mass) memcpy(mass, m.mass, sizeof(double)*n*n);
    if (diff) for (int d = 0; d < N; ++d) memcpy(diff + (long)d*n*n, m.diff[d], sizeof(double)*n*n);
    return n;
}

/* ======================================================================
 * Agglomeration graph (agglomeration.h:57-93 and 133-185) and hierarchy
 * ====================================================================== */
/* The graph MGridGen receives for the level whose elements are the parts of
 * `part` (nparts; part = NULL: the mesh itself).  Outputs xadj (nparts+1),
 * adjncy/adjwgt (<= 2 nf), vvol, vsurf.  Returns the number of edges. */
int orc_sx_agglom_graph(void* p, const int* part, int nparts, int* xadj, int* adjncy, double* adjwgt,
                        double* vvol, double* vsurf)
{
    sxproblem* sx = (sxproblem*)p; int nv = part ? nparts : sx->ne;
    for (int i = 0; i < nv; ++i) { vvol[i] = 0; vsurf[i] = 0; }
    for (int e = 0; e < sx->ne; ++e) vvol[part ? part[e] : e] += sx->vol[e];
    /* faces listed per vertex in face order: +f from the left, -f from the right */
    int* cnt = (int*)xcalloc(nv + 1, sizeof(int));
    for (int f = 0; f < sx->nf; ++f) {
        const sxface* g = &sx->faces[f];
        int l = part ? part[g->left] : g->left;
        if (g->right >= 0) { int r = part ? part[g->right] : g->right; if (l != r) { cnt[l + 1]++; cnt[r + 1]++; } }
        else cnt[l + 1]++;
    }
    for (int i = 0; i < nv; ++i) cnt[i + 1] += cnt[i];
    int* lst = (int*)xmalloc(sizeof(int)*(cnt[nv] ? cnt[nv] : 1)); int* pos = (int*)xmalloc(sizeof(int)*(nv ? nv : 1));
    for (int i = 0; i < nv; ++i) pos[i] = cnt[i];
    for (int f = 0; f < sx->nf; ++f) {
        const sxface* g = &sx->faces[f];
        int l = part ? part[g->left] : g->left;
        if (g->right >= 0) { int r = part ? part[g->right] : g->right; if (l != r) { lst[pos[l]++] = f + 1; lst[pos[r]++] = -(f + 1); } }
        else lst[pos[l]++] = f + 1;
    }
    int ne_ = 0; xadj[0] = 0;
    int* seen = (int*)xmalloc(sizeof(int)*(nv ? nv : 1)); for (int i = 0; i < nv; ++i) seen[i] = -1;
    for (int i = 0; i < nv; ++i) {
        xadj[i + 1] = xadj[i];
        for (int t = cnt[i]; t < cnt[i + 1]; ++t) {
            int jj = lst[t], f = (jj > 0 ? jj : -jj) - 1; const sxface* g = &sx->faces[f];
            int on = jj > 0 ? g->right : g->left;
            if (on >= 0) {
                int an = part ? part[on] : on;
                if (part && seen[an] >= 0) { adjwgt[seen[an]] += g->area; continue; }
                adjncy[xadj[i + 1]] = an; adjwgt[xadj[i + 1]] = g->area;
                if (part) seen[an] = xadj[i + 1];
                xadj[i + 1]++; ne_++;
            } else vsurf[i] += g->area;
        }
        if (part) for (int t = xadj[i]; t < xadj[i + 1]; ++t) seen[adjncy[t]] = -1;
    }
    free(cnt); free(lst); free(pos); free(seen);
    return ne_;
}

/* symmetric eigenproblem by cyclic Jacobi rotations: A = V diag(w) V^T,
 * eigenvalues in descending order, V's columns the eigenvectors */
static void sym_eig(int n, const double* A, double* w, double* V)
{
    double* S = (double*)xmalloc(sizeof(double)*n*n);
    memcpy(S, A, sizeof(double)*n*n);
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) V[i*n + j] = (i == j);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) { tot += S[i*n + j]*S[i*n + j]; if (i != j) off += S[i*n + j]*S[i*n + j]; }
        if (off <= 1e-30*tot) break;
        for (int p = 0; p < n; ++p) for (int q = p + 1; q < n; ++q) {
            double apq = S[p*n + q];
            if (apq == 0.0) continue;
            double th = (S[q*n + q] - S[p*n + p])/(2.0*apq);
            double t = (th >= 0 ? 1.0 : -1.0)/(fabs(th) + sqrt(th*th + 1.0));
            double c = 1.0/sqrt(t*t + 1.0), sn = t*c;
            for (int k = 0; k < n; ++k) {   /* S <- S J */
                double skp = S[k*n + p], skq = S[k*n + q];
                S[k*n + p] = c*skp - sn*skq; S[k*n + q] = sn*skp + c*skq;
            }
            for (int k = 0; k < n; ++k) {   /* S <- J^T S */
                double spk = S[p*n + k], sqk = S[q*n + k];
                S[p*n + k] = c*spk - sn*sqk; S[q*n + k] = sn*spk + c*sqk;
            }
            for (int k = 0; k < n; ++k) {   /* V <- V J */
                double vkp = V[k*n + p], vkq = V[k*n + q];
                V[k*n + p] = c*vkp - sn*vkq; V[k*n + q] = sn*vkp + c*vkq;
            }
        }
    }
    for (int i = 0; i < n; ++i) w[i] = S[i*n + i];
    /* selection sort, descending */
    for (int i = 0; i < n; ++i) {
        int m = i; for (int j = i + 1; j < n; ++j) if (w[j] > w[m]) m = j;
        if (m != i) {
            double t = w[i]; w[i] = w[m]; w[m] = t;
            for (int k = 0; k < n; ++k) { double u = V[k*n + i]; V[k*n + i] = V[k*n + m]; V[k*n + m] = u; }
        }
    }
    free(S);
}

/* Relative eigenvalue below which a direction of an agglomerate's coarse
 * space counts as zero. */
#define SX_DROP 1e-10

/* C = [v_k / sqrt(w_k) for w_k > SX_DROP w_0, then zero columns]: an
 * orthonormal basis of the span in the inner product G; returns the rank */
static int orth_basis(int n, const double* G, double* C)
{
    double w[SX_MAXNPL], V[SX_MAXNPL*SX_MAXNPL];
    sym_eig(n, G, w, V);
    int r = 0;
    while (r < n && w[r] > SX_DROP*w[0]) ++r;
    for (int i = 0; i < n; ++i) for (int k = 0; k < n; ++k) C[i*n + k] = (k < r) ? V[i*n + k]/sqrt(w[k]) : 0.0;
    return r;
}

/* Agglomerated InterpolationHierarchy (interpolationhierarchy.h:88-204) and
 * Multigrid on it.  parts[l] partitions the elements of level l (level 0 =
 * the mesh) into nparts[l] agglomerates; nlev partitions give nlev+1 levels.
 *
 * Coarse spaces: tensor Q_p (Gauss-Lobatto) on the agglomerates' bounding
 * boxes, as in the reference, restricted to the mesh.  On a small or thin
 * agglomerate these box functions are nearly or exactly linearly dependent
 * (a lone P_p simplex has fewer nodes than Q_p has functions), so each
 * agglomerate's space is represented in an orthonormal basis of its span in
 * the mass inner product (eigenvectors of its Gram matrix over the mesh,
 * scaled by w^{-1/2}), directions with w < SX_DROP w_max dropped: their
 * columns of T are zero and their diagonal entries of A are one.  Block
 * smoothers and Galerkin products are unchanged by such block-diagonal
 * changes of basis, so where the reference's spaces are well posed the
 * V-cycle is the same.  The coarsest solve is exact without the quadtree
 * Multigrid's mean subtraction. */
void* orc_sx_mg_new(void* p, int nlev, const int* const* parts, const int* nparts, const orc_mg_params* prm)
{
    sxproblem* sx = (sxproblem*)p;
    /* RATRAT inverts the coarse mass matrices T^T M T, which Q_p box spaces
     * over a few P_p simplices can make singular (a lone simplex has fewer
     * nodes than box functions; two triangles' four vertices need not
     * determine a bilinear function), so agglomerated hierarchies coarsen by
     * RAT only */
    if (!sx || !params_ok(prm) || nlev < 1 || prm->restriction != 0) return NULL;
    int N = sx->N, P = sx->P, bf = sx->E.npl;
    basis_t Bq; basis_init(&Bq, N, P);
    int bc = Bq.npl;
    bsm* Tl = (bsm*)xcalloc(nlev, sizeof(bsm));
    int** crank = (int**)xcalloc(nlev, sizeof(int*));
    /* per agglomerate of the current coarse level: Gram matrix of its box
     * basis (G) and its orthonormal basis (C) */
    double *Gp = NULL, *Cp = NULL;
    double* lo_prev = NULL; double* hi_prev = NULL; int nprev = sx->ne;
    int* orig = (int*)xmalloc(sizeof(int)*sx->ne);   /* original element -> agglomerate of the current level */
    for (int e = 0; e < sx->ne; ++e) orig[e] = e;
    double Erow[SX_MAXNPL*SX_MAXNPL], W1[SX_MAXNPL*SX_MAXNPL], W2[SX_MAXNPL*SX_MAXNPL], Me[SX_MAXNPL*SX_MAXNPL];
    int it[3];
    for (int l = 0; l < nlev; ++l) {
        int nc = nparts[l]; const int* part = parts[l];
        double* lo = (double*)xmalloc(sizeof(double)*nc*N); double* hi = (double*)xmalloc(sizeof(double)*nc*N);
        for (int i = 0; i < nc*N; ++i) { lo[i] = DBL_MAX; hi[i] = -DBL_MAX; }
        for (int e = 0; e < sx->ne; ++e) {
            int a = part[orig[e]];
            for (int v = 0; v < N + 1; ++v) for (int r = 0; r < N; ++r) {
                double x = sx->p[sx->t[e*(N+1) + v]*N + r];
                if (x < lo[a*N + r]) lo[a*N + r] = x;
                if (x > hi[a*N + r]) hi[a*N + r] = x;
            }
        }
        int* cntc = (int*)xcalloc(nc, sizeof(int));
        for (int i = 0; i < nprev; ++i) cntc[part[i]]++;
        /* reference interpolation blocks (fine rows x box columns), one per fine row */
        int rb = (l == 0) ? bf : bc;
        double* E = (double*)xmalloc(sizeof(double)*nprev*rb*bc);
        for (int i = 0; i < nprev; ++i) {
            int a = part[i]; double* Ei = E + (size_t)i*rb*bc;
            if (l == 0) {
                /* T0: the box's Q_p Lagrange basis at the fine simplex nodes, the fine
                 * simplex first mapped into the box's [0,1]^N (:112-128) */
                double rp[4][3];
                for (int v = 0; v < N + 1; ++v) for (int r = 0; r < N; ++r)
                    rp[v][r] = (sx->p[sx->t[i*(N+1) + v]*N + r] - lo[a*N + r])/(hi[a*N + r] - lo[a*N + r]);
                for (int q = 0; q < bf; ++q) {
                    double x[3];
                    for (int r = 0; r < N; ++r) { double s = rp[0][r]; for (int c = 0; c < N; ++c) s += (rp[c+1][r] - rp[0][r])*sx->E.nodes[q*N + c]; x[r] = s; }
                    lagrange_eval_nd(&Bq, x, Ei + q*bc);
                }
            } else if (cntc[a] == 1) {
                for (int k = 0; k < bc*bc; ++k) Ei[k] = (k/bc == k%bc);   /* identity (:155-161) */
            } else {
                /* box -> box: the coarse box's basis at the finer box's nodes (:164-199) */
                for (int q = 0; q < bc; ++q) {
                    unlin(q, N, P, it); double x[3];
                    for (int r = 0; r < N; ++r) {
                        double rl = (lo_prev[i*N + r] - lo[a*N + r])/(hi[a*N + r] - lo[a*N + r]);
                        double rh = (hi_prev[i*N + r] - lo[a*N + r])/(hi[a*N + r] - lo[a*N + r]);
                        x[r] = rl + (rh - rl)*Bq.nodes[it[r]];
                    }
                    lagrange_eval_nd(&Bq, x, Ei + q*bc);
                }
            }
        }
        /* Gram matrices of the coarse box bases over the mesh: sum over the
         * finer rows of E^T G_fine E (G_fine = element mass on level 0) */
        double* G = (double*)xcalloc((size_t)nc*bc*bc, sizeof(double));
        for (int i = 0; i < nprev; ++i) {
            const double* Gi;
            if (l == 0) { for (int k = 0; k < bf*bf; ++k) Me[k] = sx->E.mass[k]*(sx->vol[i]/sx->E.volume); Gi = Me; }
            else Gi = Gp + (size_t)i*bc*bc;
            const double* Ei = E + (size_t)i*rb*bc;
            gemm(0, 0, rb, bc, rb, 1.0, Gi, rb, Ei, bc, 0.0, W1, bc);
            gemm(1, 0, bc, bc, rb, 1.0, Ei, bc, W1, bc, 1.0, G + (size_t)part[i]*bc*bc, bc);
        }
        double* C = (double*)xmalloc(sizeof(double)*(size_t)nc*bc*bc);
        crank[l] = (int*)xmalloc(sizeof(int)*nc);
        for (int a = 0; a < nc; ++a) crank[l][a] = orth_basis(bc, G + (size_t)a*bc*bc, C + (size_t)a*bc*bc);
        /* T_l rows in the orthonormal bases: (C_i^T G_i) E_i C_a on coarse
         * levels, E_i C_a from the mesh */
        bsm_init(&Tl[l], nprev, nc, rb, bc);
        for (int i = 0; i < nprev; ++i) {
            int a = part[i]; const double* Ei = E + (size_t)i*rb*bc;
            gemm(0, 0, rb, bc, bc, 1.0, Ei, bc, C + (size_t)a*bc*bc, bc, 0.0, W1, bc);
            if (l == 0) { bsm_add(&Tl[l], i, a, 1.0, W1); continue; }
            gemm(1, 0, bc, bc, bc, 1.0, Cp + (size_t)i*bc*bc, bc, Gp + (size_t)i*bc*bc, bc, 0.0, W2, bc);
            gemm(0, 0, bc, bc, bc, 1.0, W2, bc, W1, bc, 0.0, Erow, bc);
            bsm_add(&Tl[l], i, a, 1.0, Erow);
        }
        free(E); free(cntc); free(Gp); free(Cp);
        Gp = G; Cp = C;
        for (int e = 0; e < sx->ne; ++e) orig[e] = part[orig[e]];
        free(lo_prev); free(hi_prev); lo_prev = lo; hi_prev = hi; nprev = nc;
    }
    free(Gp); free(Cp);
    free(lo_prev); free(hi_prev); free(orig); basis_free(&Bq);
    omg* mg = mg_setup(&sx->ops, N, Tl, nlev + 1, prm, crank, 0);
    for (int l = 0; l < nlev; ++l) free(crank[l]);
    free(crank);
    return mg;
}
