// simplex.cpp — unstructured (simplex) host path.  See simplex.hpp.
#include "simplex.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

namespace mgldg {

// fn(lo, hi) over [0, n) on up to 16 threads (disjoint index ranges)
template <typename F>
static void parallel_for(long n, F&& fn)
{
    const long nthr = std::min<long>({16L, (long)std::max(1u, std::thread::hardware_concurrency()), std::max(1L, n / 256)});
    if (nthr <= 1) { fn(0L, n); return; }
    std::vector<std::thread> th;
    const long chunk = (n + nthr - 1) / nthr;
    for (long t = 0; t < nthr; ++t) {
        const long lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([&fn, lo, hi] { fn(lo, hi); });
    }
    for (auto& x : th) x.join();
}

static long factorial(int n) { long r = 1; for (int i = 2; i <= n; ++i) r *= i; return r; }

// multi-indices with entries >= 0 and sum <= P-1, last coordinate fastest
// (SimplexRangeIterator, src/unstructured/util/range.h:291-309)
static std::vector<int> simplex_multi_indices(int N, int P)
{
    std::vector<int> out;
    int cur[3] = {0, 0, 0};
    for (;;) {
        out.insert(out.end(), cur, cur + N);
        int s = N - 1;
        for (; s >= 0; --s) {
            int other = 0;
            for (int k = 0; k < N; ++k) if (k != s) other += cur[k];
            if (++cur[s] < P - other) break;
            cur[s] = 0;
        }
        if (s < 0) break;
    }
    return out;
}

// Gauss-Jordan inverse with partial pivoting in long double
static std::vector<long double> invert_ld(std::vector<long double> A, int n)
{
    std::vector<long double> X((size_t)n * n, 0.0L);
    for (int i = 0; i < n; ++i) X[(size_t)i * n + i] = 1.0L;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i) if (fabsl(A[(size_t)i * n + k]) > fabsl(A[(size_t)p * n + k])) p = i;
        if (A[(size_t)p * n + k] == 0.0L) throw Error(-1, "singular matrix in the simplex basis");
        if (p != k) for (int j = 0; j < n; ++j) { std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]); std::swap(X[(size_t)k * n + j], X[(size_t)p * n + j]); }
        const long double inv = 1.0L / A[(size_t)k * n + k];
        for (int j = 0; j < n; ++j) { A[(size_t)k * n + j] *= inv; X[(size_t)k * n + j] *= inv; }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            const long double f = A[(size_t)i * n + k];
            if (f == 0.0L) continue;
            for (int j = 0; j < n; ++j) { A[(size_t)i * n + j] -= f * A[(size_t)k * n + j]; X[(size_t)i * n + j] -= f * X[(size_t)k * n + j]; }
        }
    }
    return X;
}

static long double mono_ld(const long double* x, const int* e, int N)
{
    long double v = 1.0L;
    for (int d = 0; d < N; ++d) for (int k = 0; k < e[d]; ++k) v *= x[d];
    return v;
}

void SimplexBasis::init(int N_, int P_)
{
    N = N_; P = P_;
    if (N < 1 || N > 3 || P < 1 || P > 6) throw Error(-1, "simplex basis: unsupported dimension or degree");
    expo = simplex_multi_indices(N, P);
    npl = (int)expo.size() / N;
    volume = 1.0 / (double)factorial(N);
    nodes.assign((size_t)npl * N, 0.0);
    for (int i = 0; i < npl; ++i)
        for (int d = 0; d < N; ++d)
            nodes[(size_t)i * N + d] = (P == 1) ? 1.0 / 3.0 : (double)expo[(size_t)i * N + d] / (P - 1);
    const int n = npl;
    // Vandermonde of the monomials at the nodes, long double
    std::vector<long double> V((size_t)n * n), xl((size_t)n * N);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < N; ++d)
            xl[(size_t)i * N + d] = (P == 1) ? 1.0L / 3.0L : (long double)expo[(size_t)i * N + d] / (P - 1);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) V[(size_t)i * n + k] = mono_ld(&xl[(size_t)i * N], &expo[(size_t)k * N], N);
    std::vector<long double> C = invert_ld(V, n);   // phi_i = sum_k C[k][i] m_k
    coef.resize((size_t)n * n);
    for (size_t t = 0; t < coef.size(); ++t) coef[t] = (double)C[t];
    // exact moments int x^a over the unit simplex = prod a_d! / (|a| + N)!
    auto moment = [&](const int* a, const int* b) {
        long double num = 1.0L; int tot = 0;
        for (int d = 0; d < N; ++d) { int s = a[d] + b[d]; tot += s; for (int k = 2; k <= s; ++k) num *= k; }
        long double den = 1.0L;
        for (int k = 2; k <= tot + N; ++k) den *= k;
        return num / den;
    };
    std::vector<long double> Mm((size_t)n * n), mass_ld((size_t)n * n, 0.0L);
    for (int k = 0; k < n; ++k) for (int l = 0; l < n; ++l) Mm[(size_t)k * n + l] = moment(&expo[(size_t)k * N], &expo[(size_t)l * N]);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            long double s = 0.0L;
            for (int k = 0; k < n; ++k) {
                const long double cki = C[(size_t)k * n + i];
                if (cki == 0.0L) continue;
                for (int l = 0; l < n; ++l) s += cki * Mm[(size_t)k * n + l] * C[(size_t)l * n + j];
            }
            mass_ld[(size_t)i * n + j] = s;
        }
    mass.resize((size_t)n * n);
    for (size_t t = 0; t < mass.size(); ++t) mass[t] = (double)mass_ld[t];
    std::vector<long double> Mi = invert_ld(mass_ld, n);
    invmass.resize((size_t)n * n);
    for (size_t t = 0; t < invmass.size(); ++t) invmass[t] = (double)Mi[t];
    for (int d = 0; d < N; ++d) {
        diff[d].assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                long double s = 0.0L;
                for (int k = 0; k < n; ++k) {
                    const int* e = &expo[(size_t)k * N];
                    if (e[d] == 0) continue;
                    int e2[3] = {e[0], N > 1 ? e[1] : 0, N > 2 ? e[2] : 0};
                    e2[d] -= 1;
                    s += C[(size_t)k * n + j] * (long double)e[d] * mono_ld(&xl[(size_t)i * N], e2, N);
                }
                diff[d][(size_t)i * n + j] = (double)s;
            }
    }
}

void SimplexBasis::eval(const double* x, double* phi) const
{
    long double xl[3] = {x[0], N > 1 ? x[1] : 0.0, N > 2 ? x[2] : 0.0};
    long double m[64];   // npl <= C(p+3, 3) = 56 for p <= 5
    for (int k = 0; k < npl; ++k) m[k] = mono_ld(xl, &expo[(size_t)k * N], N);
    for (int i = 0; i < npl; ++i) {
        long double s = 0.0L;
        for (int k = 0; k < npl; ++k) s += (long double)coef[(size_t)k * npl + i] * m[k];
        phi[i] = (double)s;
    }
}

// ======================================================================= mesh
static double det_n(const double* J, int N)
{
    if (N == 1) return J[0];
    if (N == 2) return J[0] * J[3] - J[1] * J[2];
    return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) + J[2] * (J[3] * J[7] - J[4] * J[6]);
}

static void inv_n(const double* J, int N, double det, double* I)
{
    if (N == 1) { I[0] = 1.0 / det; return; }
    if (N == 2) { I[0] = J[3] / det; I[1] = -J[1] / det; I[2] = -J[2] / det; I[3] = J[0] / det; return; }
    I[0] = (J[4] * J[8] - J[5] * J[7]) / det; I[1] = (J[2] * J[7] - J[1] * J[8]) / det; I[2] = (J[1] * J[5] - J[2] * J[4]) / det;
    I[3] = (J[5] * J[6] - J[3] * J[8]) / det; I[4] = (J[0] * J[8] - J[2] * J[6]) / det; I[5] = (J[2] * J[3] - J[0] * J[5]) / det;
    I[6] = (J[3] * J[7] - J[4] * J[6]) / det; I[7] = (J[1] * J[6] - J[0] * J[7]) / det; I[8] = (J[0] * J[4] - J[1] * J[3]) / det;
}

void SimplexMesh::build(int N_, int np_, const double* p, int nt, const int* t, int nbnd, const int* vmask)
{
    N = N_; np = np_; ne = nt;
    if (N < 2 || N > 3) throw Error(-1, "simplex meshes are 2-D or 3-D");
    if (np < N + 1 || nt < 1 || !p || !t) throw Error(-1, "empty simplex mesh");
    if (nbnd < 0 || nbnd > 30 || (nbnd > 0 && !vmask)) throw Error(-1, "invalid boundary predicates");
    pts.assign(p, p + (size_t)np * N);
    tris.assign(t, t + (size_t)nt * (N + 1));
    for (int v : tris) if (v < 0 || v >= np) throw Error(-1, "simplex vertex index out of range");
    vol.resize(ne); J.resize((size_t)ne * N * N); Jinv.resize((size_t)ne * N * N);
    for (int e = 0; e < ne; ++e) {
        double* Je = &J[(size_t)e * N * N];
        const int* te = &tris[(size_t)e * (N + 1)];
        for (int r = 0; r < N; ++r)
            for (int i = 0; i < N; ++i) Je[r * N + i] = pts[(size_t)te[i + 1] * N + r] - pts[(size_t)te[0] * N + r];
        const double d = det_n(Je, N);
        if (!(std::fabs(d) > 0.0)) throw Error(-1, "degenerate simplex");
        vol[e] = std::fabs(d) / (double)factorial(N);   // sqrt(det(J^T J)) / N! (geometry.h:65-74)
        inv_n(Je, N, d, &Jinv[(size_t)e * N * N]);
    }
    // faces: sorted vertex N-tuples (mesh.h:210-226), numbered by first appearance
    struct Ent { int v[3]; long pos; };
    const long nent = (long)ne * (N + 1);
    std::vector<Ent> ent(nent);
    for (int e = 0; e < ne; ++e)
        for (int j = 0; j < N + 1; ++j) {
            Ent& x = ent[(long)e * (N + 1) + j];
            x.v[2] = -1;
            for (int k = 0; k < N; ++k) x.v[k] = tris[(size_t)e * (N + 1) + (k + j) % (N + 1)];
            std::sort(x.v, x.v + N);
            x.pos = (long)e * (N + 1) + j;
        }
    std::sort(ent.begin(), ent.end(), [](const Ent& a, const Ent& b) {
        for (int k = 0; k < 3; ++k) if (a.v[k] != b.v[k]) return a.v[k] < b.v[k];
        return a.pos < b.pos;
    });
    struct Grp { long first; int e1, e2; int v[3]; };
    std::vector<Grp> grp;
    grp.reserve(nent / 2 + 16);
    for (long i = 0; i < nent;) {
        long j = i + 1;
        while (j < nent && ent[j].v[0] == ent[i].v[0] && ent[j].v[1] == ent[i].v[1] && ent[j].v[2] == ent[i].v[2]) ++j;
        if (j - i > 2) throw Error(-1, "non-manifold simplex mesh (a face shared by more than two simplices)");
        Grp g;
        g.first = ent[i].pos;
        g.e1 = (int)(ent[i].pos / (N + 1));
        g.e2 = (j - i == 2) ? (int)(ent[i + 1].pos / (N + 1)) : -1;
        std::memcpy(g.v, ent[i].v, sizeof g.v);
        grp.push_back(g);
        i = j;
    }
    std::sort(grp.begin(), grp.end(), [](const Grp& a, const Grp& b) { return a.first < b.first; });
    nf = (int)grp.size();
    faces.assign(nf, SimplexFace());
    bfaces.clear();
    for (int fi = 0; fi < nf; ++fi) {
        const Grp& g = grp[fi];
        SimplexFace& f = faces[fi];
        const int mn = std::min(g.e1, g.e2), mx = std::max(g.e1, g.e2);
        f.left = (mn >= 0) ? mn : mx;          // mesh.h:234-237
        f.right = (mn >= 0) ? mx : mn;
        for (int k = 0; k < N; ++k) f.v[k] = g.v[k];
        // normal from the cofactors of the face Jacobian (mesh.h:249-263)
        const double* x0 = &pts[(size_t)f.v[0] * N];
        const double* x1 = &pts[(size_t)f.v[1] * N];
        double n[3] = {0, 0, 0};
        if (N == 2) { n[0] = x1[1] - x0[1]; n[1] = -(x1[0] - x0[0]); }
        else {
            const double* x2 = &pts[(size_t)f.v[2] * N];
            const double a[3] = {x1[0] - x0[0], x1[1] - x0[1], x1[2] - x0[2]};
            const double c[3] = {x2[0] - x0[0], x2[1] - x0[1], x2[2] - x0[2]};
            n[0] = a[1] * c[2] - a[2] * c[1]; n[1] = a[2] * c[0] - a[0] * c[2]; n[2] = a[0] * c[1] - a[1] * c[0];
        }
        double nn = 0;
        for (int d = 0; d < N; ++d) nn += n[d] * n[d];
        nn = std::sqrt(nn);
        f.area = nn / (double)factorial(N - 1);
        // outward from the left element (mesh.h:266-273)
        const int* tl = &tris[(size_t)f.left * (N + 1)];
        int opp = tl[0];
        for (int i = 0; i < N + 1; ++i) {
            bool in = false;
            for (int k = 0; k < N; ++k) in |= (tl[i] == f.v[k]);
            if (!in) opp = tl[i];
        }
        double dot = 0;
        for (int d = 0; d < N; ++d) dot += n[d] * (pts[(size_t)opp * N + d] - x0[d]);
        const double sgn = dot > 0 ? -1.0 : 1.0;
        for (int d = 0; d < N; ++d) f.normal[d] = sgn * n[d] / nn;
        if (f.right < 0) {
            int key = 1;
            if (nbnd > 0) {   // first predicate holding at all face vertices (mesh.h:276-288)
                key = 0;
                for (int k = 0; k < nbnd && !key; ++k) {
                    bool all = true;
                    for (int q = 0; q < N; ++q) all &= ((vmask[f.v[q]] >> k) & 1) != 0;
                    if (all) key = k + 1;
                }
                if (!key) throw Error(-1, "Boundary face does not satisfy any boundary function.");
            }
            f.right = -key;
            bfaces.push_back(fi);
        }
    }
    nbf = (int)bfaces.size();
}

void SimplexMesh::face_point(const SimplexFace& f, const double* ref, double* x) const
{
    const double* x0 = &pts[(size_t)f.v[0] * N];
    for (int d = 0; d < N; ++d) {
        double s = x0[d];
        for (int k = 0; k < N - 1; ++k) s += (pts[(size_t)f.v[k + 1] * N + d] - x0[d]) * ref[k];
        x[d] = s;
    }
}

void SimplexMesh::to_local(int e, const double* x, double* loc) const
{
    const double* I = &Jinv[(size_t)e * N * N];
    const double* x0 = &pts[(size_t)tris[(size_t)e * (N + 1)] * N];
    for (int r = 0; r < N; ++r) {
        double s = 0;
        for (int c = 0; c < N; ++c) s += I[r * N + c] * (x[c] - x0[c]);
        loc[r] = s;
    }
}

// ======================================================================= assembly
// Face rule on the reference (N-1)-simplex, weights summing to one
// (Quadrature<N-1,10>, src/unstructured/dg/master.h:359-442): the 10-point
// Gauss rule on [0,1] for triangle meshes; for tetrahedra Dunavant's degree-9
// rule with 19 points on the unit triangle (D. A. Dunavant, Int. J. Numer.
// Meth. Eng. 21 (1985)), in the reference table's point order: the centroid,
// four orbits (a,b), (b,a), (b,b) and one orbit (a,b), (c,a), (b,c), (a,c),
// (c,b), (b,a), with the table's 15-digit values.
static void face_rule(int N, std::vector<double>& x, std::vector<double>& w)
{
    if (N == 2) {
        x.resize(kQuad); w.resize(kQuad);
        gauss_rule(kQuad, x.data(), w.data());
        return;
    }
    struct Orbit3 { double a, b, w; };
    static const Orbit3 o3[4] = {
        {2.06349616025250E-02, 4.89682519198738E-01, 0.031334700227139},
        {1.25820817014127E-01, 4.37089591492937E-01, 0.077827541004774},
        {6.23592928761935E-01, 1.88203535619033E-01, 0.079647738927210},
        {9.10540973211095E-01, 4.47295133944530E-02, 0.025577675658698}};
    const double a6 = 3.68384120547360E-02, b6 = 2.21962989160766E-01, c6 = 7.41198598784498E-01;
    const double w6 = 0.043283539377289;
    x.clear(); w.clear();
    auto put = [&](double u, double v, double wt) { x.push_back(u); x.push_back(v); w.push_back(wt); };
    put(3.33333333333333E-01, 3.33333333333333E-01, 0.097135796282799);
    for (const Orbit3& o : o3) { put(o.a, o.b, o.w); put(o.b, o.a, o.w); put(o.b, o.b, o.w); }
    put(a6, b6, w6); put(c6, a6, w6); put(b6, c6, w6); put(a6, c6, w6); put(c6, b6, w6); put(b6, a6, w6);
}

void SimplexLdg::assemble(int N_, int np, const double* pts, int nt, const int* tris, int nbnd, const int* vmask, int p_,
                          const int* bct, double t0, double tD)
{
    N = N_; p = p_; tau0 = t0; tauD = tD;
    if (p < 0 || p > 5) throw Error(-1, "polynomial degree out of range");
    mesh.build(N, np, pts, nt, tris, nbnd, vmask);
    const int nkeys = std::max(nbnd, 1);
    bctype.assign(nkeys, 0);
    for (int k = 0; k < nkeys; ++k) {
        bctype[k] = bct ? bct[k] : 0;
        if (bctype[k] != 0 && bctype[k] != 1) throw Error(-1, "unknown boundary condition");   // ldgpoisson.tpp:145
    }
    E.init(N, p + 1);
    F.init(N - 1, p + 1);
    face_rule(N, qx, qw);
    npl = E.npl; nq = (int)qw.size();
    ne = mesh.ne; nf = mesh.nf; nbf = mesh.nbf;
    const int b = npl, nfl = F.npl;
    auto dirichlet = [&](const SimplexFace& f) { return f.right < 0 && bctype[-f.right - 1] == 0; };
    // codes: face vertex k (ascending ids) is element vertex m_k; code =
    // sum_k m_k (N+1)^k.  A face node with barycentric weights lambda over the
    // face vertices sits at sum_k lambda_k V(m_k) in the element's reference
    // coordinates (V(0) = 0, V(m) = e_{m-1}), so every code fixes a slice
    // (face node -> element node, mesh.h:32-66) and the element basis at the
    // face rule (mesh.h:67-76, 336-347).
    AsmPlan& A = plan;
    A = AsmPlan();
    int ncode = 1;
    for (int k = 0; k < N; ++k) ncode *= N + 1;
    A.N = N; A.b = b; A.nfl = nfl; A.ncode = ncode; A.nq = nq; A.gshared = 1;
    A.emass = E.mass; A.einv = E.invmass; A.fmass = F.mass;
    for (int d = 0; d < N; ++d) A.ediff.insert(A.ediff.end(), E.diff[d].begin(), E.diff[d].end());
    A.slice.assign((size_t)ncode * nfl, 0); A.islice.assign((size_t)ncode * b, -1);
    A.pool.assign((size_t)ncode * b * b, 0.0); A.npool_ref = ncode;
    A.lbref.assign((size_t)ncode * b * nq, 0.0);
    auto decode = [&](int code, int* m) {
        for (int k = 0; k < N; ++k) { m[k] = code % (N + 1); code /= N + 1; }
        for (int k = 0; k < N; ++k) for (int q = 0; q < k; ++q) if (m[k] == m[q]) return false;
        return true;
    };
    auto to_elem = [&](const int* m, const double* r, double* y) {
        double l0 = 1;
        for (int k = 0; k < N - 1; ++k) l0 -= r[k];
        for (int d = 0; d < N; ++d) y[d] = 0;
        for (int k = 0; k < N; ++k) {
            const double lam = k ? r[k - 1] : l0;
            if (m[k] > 0) y[m[k] - 1] += lam;
        }
    };
    std::vector<double> phi((size_t)nq * b);
    for (int code = 0; code < ncode; ++code) {
        int m[3];
        if (!decode(code, m)) continue;
        for (int a = 0; a < nfl; ++a) {
            int hit = 0;
            if (E.P > 1) {
                double y[3];
                to_elem(m, &F.nodes[(size_t)a * (N - 1)], y);
                hit = -1;
                for (int j = 0; j < b && hit < 0; ++j) {
                    bool ok = true;
                    for (int d = 0; d < N; ++d) ok &= std::fabs(y[d] - E.nodes[(size_t)j * N + d]) < 1e-9;
                    if (ok) hit = j;
                }
                if (hit < 0) throw Error(-1, "face nodes do not match element nodes");
            }
            A.slice[(size_t)code * nfl + a] = hit;
            A.islice[(size_t)code * b + hit] = a;
        }
        for (int q = 0; q < nq; ++q) {
            double y[3];
            to_elem(m, &qx[(size_t)q * (N - 1)], y);
            E.eval(y, &phi[(size_t)q * b]);
        }
        double* LL = &A.pool[(size_t)code * b * b];
        for (int i = 0; i < b; ++i)
            for (int j = 0; j < b; ++j) {
                double s = 0;
                for (int q = 0; q < nq; ++q) s += phi[(size_t)q * b + i] * qw[q] * phi[(size_t)q * b + j];
                LL[(size_t)i * b + j] = s;
            }
        double* Lr = &A.lbref[(size_t)code * b * nq];
        for (int i = 0; i < b; ++i) for (int q = 0; q < nq; ++q) Lr[(size_t)i * nq + q] = phi[(size_t)q * b + i] * qw[q];
    }
    auto code_of = [&](const SimplexFace& f, int e) {
        const int* t = &mesh.tris[(size_t)e * (N + 1)];
        int code = 0, pw = 1;
        for (int k = 0; k < N; ++k) {
            int mk = 0;
            while (mk <= N && t[mk] != f.v[k]) ++mk;
            if (mk > N) throw Error(-1, "face vertex not in its element");
            code += mk * pw; pw *= N + 1;
        }
        return code;
    };
    // patterns (addToBlock creates structural blocks for every normal
    // component, so the N gradient patterns are one)
    {
        PatternBuilder pm(ne, ne), pt(ne, ne), pg(ne, ne);
        for (int e = 0; e < ne; ++e) { pm.add(e, e); pg.add(e, e); }
        for (const SimplexFace& f : mesh.faces) {
            if (f.right >= 0) {
                pg.add(f.right, f.right); pg.add(f.right, f.left);
                pt.add(f.right, f.right); pt.add(f.left, f.left); pt.add(f.right, f.left); pt.add(f.left, f.right);
            } else if (dirichlet(f)) {
                pt.add(f.left, f.left);
            }
        }
        pm.finish(M, b, b, false); pt.finish(T, b, b, false);
        pg.finish(G[0], b, b, false);
        for (int d = 1; d < N; ++d) G[d] = G[0];
    }
    // elements: M_e = (|e| / |ref|) mass, M_e^-1 its inverse, G_i volume
    // term sum_j Jinv(j, i) diff_j (ldgpoisson.tpp:26-58)
    A.mscale.resize(ne); A.iscale.resize(ne);
    A.C.assign(mesh.Jinv.begin(), mesh.Jinv.end());
    for (int e = 0; e < ne; ++e) { A.mscale[e] = mesh.vol[e] / E.volume; A.iscale[e] = E.volume / mesh.vol[e]; }
    // faces: interior by slices with jacobian |f| / |ref face|, Dirichlet by
    // the reference blocks with |f| |ref face| (mesh.h:343-344 as written)
    A.fl.resize(nf); A.fr.resize(nf); A.fkind.resize(nf); A.fca.assign(nf, 0); A.fcb.assign(nf, 0);
    A.fscale.assign(nf, 0.0); A.fnormal.assign((size_t)nf * N, 0.0);
    std::vector<int> cnt(ne + 1, 0);
    for (int fi = 0; fi < nf; ++fi) {
        const SimplexFace& f = mesh.faces[fi];
        A.fl[fi] = f.left; A.fr[fi] = f.right;
        for (int d = 0; d < N; ++d) A.fnormal[(size_t)fi * N + d] = f.normal[d];
        if (f.right >= 0) {
            A.fkind[fi] = AsmPlan::kSlice; A.fca[fi] = code_of(f, f.left); A.fcb[fi] = code_of(f, f.right);
            A.fscale[fi] = f.area / F.volume;
            cnt[f.left + 1]++; cnt[f.right + 1]++;
        } else if (dirichlet(f)) {
            A.fkind[fi] = AsmPlan::kDirichlet; A.fca[fi] = code_of(f, f.left); A.fscale[fi] = f.area * F.volume;
            cnt[f.left + 1]++;
        } else {
            A.fkind[fi] = AsmPlan::kNone;
        }
    }
    for (int e = 0; e < ne; ++e) cnt[e + 1] += cnt[e];
    A.ef_off = cnt; A.ef.resize(cnt[ne]);
    for (int fi = 0; fi < nf; ++fi) {
        const int k = A.fkind[fi];
        if (k == AsmPlan::kNone) continue;
        A.ef[cnt[A.fl[fi]]++] = fi << 1;
        if (k != AsmPlan::kDirichlet) A.ef[cnt[A.fr[fi]]++] = (fi << 1) | 1;
    }
    A.finish_rows(ne, T, G);
    // boundary faces: data for the right-hand side and its adjoint
    bleft.resize(nbf); btype.resize(nbf); btag.resize(nbf); bnormal.assign((size_t)nbf * N, 0.0);
    A.bcode.resize(nbf); A.bscale.resize(nbf);
    for (int k = 0; k < nbf; ++k) {
        const SimplexFace& f = mesh.faces[mesh.bfaces[k]];
        A.bcode[k] = code_of(f, f.left); A.bscale[k] = f.area * F.volume;
        bleft[k] = f.left; btag[k] = -f.right; btype[k] = bctype[-f.right - 1];
        for (int d = 0; d < N; ++d) bnormal[(size_t)k * N + d] = f.normal[d];
    }
}

void SimplexLdg::bc_qpoints(double* xyz, int* tag) const
{
    for (int k = 0; k < nbf; ++k) {
        const SimplexFace& f = mesh.faces[mesh.bfaces[k]];
        for (int q = 0; q < nq; ++q) {
            mesh.face_point(f, &qx[(size_t)q * (N - 1)], xyz + ((size_t)k * nq + q) * N);
            if (tag) tag[(size_t)k * nq + q] = -f.right;
        }
    }
}

void SimplexLdg::nodes_xyz(double* xyz) const
{
    for (int e = 0; e < ne; ++e) {
        const double* Je = &mesh.J[(size_t)e * N * N];
        const double* x0 = &mesh.pts[(size_t)mesh.tris[(size_t)e * (N + 1)] * N];
        for (int i = 0; i < npl; ++i)
            for (int r = 0; r < N; ++r) {
                double s = x0[r];
                for (int c = 0; c < N; ++c) s += Je[r * N + c] * E.nodes[(size_t)i * N + c];
                xyz[((size_t)e * npl + i) * N + r] = s;
            }
    }
}

// ======================================================================= agglomeration
void agglomeration_graph(const SimplexMesh& m, const int* orig, int nparts, AggGraph& g)
{
    const int nv = orig ? nparts : m.ne;
    auto P = [&](int e) { return orig ? orig[e] : e; };
    g.vvol.assign(nv, 0.0); g.vsurf.assign(nv, 0.0);
    for (int e = 0; e < m.ne; ++e) g.vvol[P(e)] += m.vol[e];
    // faces of each vertex in face order, +-(f+1) for the left / right side
    std::vector<int> cnt(nv + 1, 0);
    for (const SimplexFace& f : m.faces) {
        const int l = P(f.left);
        if (f.right >= 0) { const int r = P(f.right); if (l != r) { cnt[l + 1]++; cnt[r + 1]++; } }
        else cnt[l + 1]++;
    }
    for (int i = 0; i < nv; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> lst(std::max(cnt[nv], 1)), pos(cnt.begin(), cnt.end() - 1);
    for (int fi = 0; fi < m.nf; ++fi) {
        const SimplexFace& f = m.faces[fi];
        const int l = P(f.left);
        if (f.right >= 0) { const int r = P(f.right); if (l != r) { lst[pos[l]++] = fi + 1; lst[pos[r]++] = -(fi + 1); } }
        else lst[pos[l]++] = fi + 1;
    }
    g.xadj.assign(nv + 1, 0); g.adj.clear(); g.w.clear();
    std::vector<int> seen(nv, -1);
    for (int i = 0; i < nv; ++i) {
        for (int t = cnt[i]; t < cnt[i + 1]; ++t) {
            const int jj = lst[t], fi = std::abs(jj) - 1;
            const SimplexFace& f = m.faces[fi];
            const int on = jj > 0 ? f.right : f.left;
            if (on >= 0) {
                const int an = P(on);
                if (seen[an] >= 0) { g.w[seen[an]] += f.area; continue; }
                seen[an] = (int)g.adj.size();
                g.adj.push_back(an); g.w.push_back(f.area);
            } else g.vsurf[i] += f.area;
        }
        g.xadj[i + 1] = (int)g.adj.size();
        for (int t = g.xadj[i]; t < g.xadj[i + 1]; ++t) seen[g.adj[t]] = -1;
    }
}

std::vector<std::vector<int>> agglomerate(const SimplexMesh& m, int max_levels, int maxsize)
{
    std::vector<std::vector<int>> parts;
    std::vector<int> orig(m.ne);
    std::iota(orig.begin(), orig.end(), 0);
    int n = m.ne;
    for (int l = 0; l < max_levels && n > 1; ++l) {
        AggGraph g;
        agglomeration_graph(m, l == 0 ? nullptr : orig.data(), n, g);
        std::vector<int> part;
        // MGridGen's call of agglomeration.h:96-116: minsize 1, maxsize 2^N
        const int nc = mgridgen_partition(n, g.xadj.data(), g.vvol.data(), g.vsurf.data(), g.adj.data(), g.w.data(), m.N,
                                          1, maxsize, part);
        if (nc >= n) break;
        for (int e = 0; e < m.ne; ++e) orig[e] = part[orig[e]];
        parts.push_back(std::move(part));
        n = nc;
    }
    return parts;
}

// symmetric eigenproblem by cyclic Jacobi rotations: A = V diag(w) V^T with
// the eigenvalues descending and V's columns the eigenvectors (n x n row-major)
static void sym_eig(int n, const double* A, double* w, double* V)
{
    std::vector<double> S(A, A + (size_t)n * n);
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) V[i * n + j] = (i == j);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) { const double v = S[i * n + j] * S[i * n + j]; tot += v; if (i != j) off += v; }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = S[p * n + q];
                if (apq == 0.0) continue;
                const double th = (S[q * n + q] - S[p * n + p]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {
                    const double skp = S[k * n + p], skq = S[k * n + q];
                    S[k * n + p] = c * skp - sn * skq; S[k * n + q] = sn * skp + c * skq;
                }
                for (int k = 0; k < n; ++k) {
                    const double spk = S[p * n + k], sqk = S[q * n + k];
                    S[p * n + k] = c * spk - sn * sqk; S[q * n + k] = sn * spk + c * sqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - sn * vkq; V[k * n + q] = sn * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) w[i] = S[i * n + i];
    for (int i = 0; i < n; ++i) {
        int mx = i;
        for (int j = i + 1; j < n; ++j) if (w[j] > w[mx]) mx = j;
        if (mx != i) {
            std::swap(w[i], w[mx]);
            for (int k = 0; k < n; ++k) std::swap(V[k * n + i], V[k * n + mx]);
        }
    }
}

constexpr double kDrop = 1e-10;   // relative eigenvalue of a dropped coarse direction

// An orthonormal basis of the span in the inner product G, as columns of C
// (zero columns for dropped directions, kept ones first); returns the rank.
// The Cholesky factor G = L L^T gives C = L^{-T} (C^T G C = I) whenever it
// provably drops nothing: ||G^{-1}||_2 = ||C||_2^2 <= ||C||_F^2, so
// 1/||C||_F^2 > kDrop ||G||_F >= kDrop w_max means every eigenvalue passes.
// Otherwise the eigen-decomposition: C = [v_k / sqrt(w_k) for w_k > kDrop w_0].
static int orth_basis(int n, const double* G, double* C)
{
    std::vector<double> Lc((size_t)n * n, 0.0);
    bool ok = true;
    for (int j = 0; j < n && ok; ++j) {
        double d = G[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= Lc[(size_t)j * n + k] * Lc[(size_t)j * n + k];
        if (!(d > 0.0)) { ok = false; break; }
        const double ljj = std::sqrt(d);
        Lc[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double v = G[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) v -= Lc[(size_t)i * n + k] * Lc[(size_t)j * n + k];
            Lc[(size_t)i * n + j] = v / ljj;
        }
    }
    if (ok) {
        // C = L^{-T}: solve L^T C = I column by column (C upper triangular)
        std::fill(C, C + (size_t)n * n, 0.0);
        for (int c = 0; c < n; ++c)
            for (int i = c; i >= 0; --i) {
                double v = (i == c) ? 1.0 : 0.0;
                for (int k = i + 1; k <= c; ++k) v -= Lc[(size_t)k * n + i] * C[(size_t)k * n + c];
                C[(size_t)i * n + c] = v / Lc[(size_t)i * n + i];
            }
        double gf = 0, cf = 0;
        for (size_t t = 0; t < (size_t)n * n; ++t) { gf += G[t] * G[t]; cf += C[t] * C[t]; }
        if (1.0 / cf > kDrop * std::sqrt(gf)) return n;
    }
    std::vector<double> w(n), V((size_t)n * n);
    sym_eig(n, G, w.data(), V.data());
    int r = 0;
    while (r < n && w[r] > kDrop * w[0]) ++r;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) C[(size_t)i * n + k] = (k < r) ? V[(size_t)i * n + k] / std::sqrt(w[k]) : 0.0;
    return r;
}

// C (m x n) = A B (+ C), A m x k, B k x n, row-major
static void mm_nn(int m, int n, int k, const double* A, const double* B, double* C, bool acc = false)
{
    for (int i = 0; i < m; ++i) {
        double* c = C + (size_t)i * n;
        if (!acc) std::fill(c, c + n, 0.0);
        for (int t = 0; t < k; ++t) {
            const double a = A[(size_t)i * k + t];
            const double* b = B + (size_t)t * n;
            for (int j = 0; j < n; ++j) c[j] += a * b[j];
        }
    }
}

// C (m x n) = A^T B (+ C), A k x m, B k x n, row-major
static void mm_tn(int m, int n, int k, const double* A, const double* B, double* C, bool acc = false)
{
    if (!acc) std::fill(C, C + (size_t)m * n, 0.0);
    for (int t = 0; t < k; ++t) {
        const double* a = A + (size_t)t * m;
        const double* b = B + (size_t)t * n;
        for (int i = 0; i < m; ++i) {
            const double ai = a[i];
            double* c = C + (size_t)i * n;
            for (int j = 0; j < n; ++j) c[j] += ai * b[j];
        }
    }
}


void build_interps_agglo(const SimplexMesh& m, const SimplexBasis& E, const Basis& Q,
                         const std::vector<std::vector<int>>& parts, std::vector<Interp>& Iv,
                         std::vector<std::vector<int>>& rank)
{
    const int N = m.N, bf = E.npl, bc = Q.npl, nlev = (int)parts.size();
    Iv.assign(nlev, Interp());
    rank.assign(nlev, std::vector<int>());
    std::vector<int> orig(m.ne);   // mesh element -> agglomerate of the current level
    std::iota(orig.begin(), orig.end(), 0);
    // per agglomerate of the current level: Gram matrix of its box basis and
    // its orthonormal basis (empty on the mesh level)
    std::vector<double> Gp, Cp, lop, hip;
    int nprev = m.ne;
    for (int l = 0; l < nlev; ++l) {
        const std::vector<int>& part = parts[l];
        if ((int)part.size() != nprev) throw Error(-1, "partition size does not match its level");
        int nc = 0;
        for (int v : part) { if (v < 0) throw Error(-1, "negative agglomerate id"); nc = std::max(nc, v + 1); }
        std::vector<int> cnt(nc, 0);
        for (int v : part) cnt[v]++;
        for (int c = 0; c < nc; ++c) if (!cnt[c]) throw Error(-1, "empty agglomerate in partition");
        // bounding boxes of the original vertices (AgglomeratedElement, agglomeration.h:258-276)
        std::vector<double> lo((size_t)nc * N, 1e300), hi((size_t)nc * N, -1e300);
        for (int e = 0; e < m.ne; ++e) {
            const int a = part[orig[e]];
            for (int v = 0; v < N + 1; ++v)
                for (int d = 0; d < N; ++d) {
                    const double x = m.pts[(size_t)m.tris[(size_t)e * (N + 1) + v] * N + d];
                    lo[(size_t)a * N + d] = std::min(lo[(size_t)a * N + d], x);
                    hi[(size_t)a * N + d] = std::max(hi[(size_t)a * N + d], x);
                }
        }
        // reference blocks (the box basis at the finer nodes), one per finer row
        const int rb = (l == 0) ? bf : bc;
        std::vector<double> Eb((size_t)nprev * rb * bc);
        parallel_for(nprev, [&](long i0, long i1) {
            int it[3];
            for (long i = i0; i < i1; ++i) {
                const int a = part[i];
                double* Ei = &Eb[(size_t)i * rb * bc];
                if (l == 0) {
                    // T0 (:99-138): the fine simplex mapped into the box's [0,1]^N, then its nodes
                    double rp[4][3];
                    for (int v = 0; v < N + 1; ++v)
                        for (int d = 0; d < N; ++d)
                            rp[v][d] = (m.pts[(size_t)m.tris[(size_t)i * (N + 1) + v] * N + d] - lo[(size_t)a * N + d]) /
                                       (hi[(size_t)a * N + d] - lo[(size_t)a * N + d]);
                    for (int q = 0; q < bf; ++q) {
                        double x[3];
                        for (int d = 0; d < N; ++d) {
                            double sv = rp[0][d];
                            for (int c = 0; c < N; ++c) sv += (rp[c + 1][d] - rp[0][d]) * E.nodes[(size_t)q * N + c];
                            x[d] = sv;
                        }
                        Q.lagrangeN(x, Ei + (size_t)q * bc);
                    }
                } else if (cnt[a] == 1) {
                    for (int k = 0; k < bc * bc; ++k) Ei[k] = (k / bc == k % bc) ? 1.0 : 0.0;   // identity (:155-161)
                } else {
                    // box -> box (:164-199)
                    for (int q = 0; q < bc; ++q) {
                        int lin = q;
                        for (int d = N - 1; d >= 0; --d) { it[d] = lin % Q.P; lin /= Q.P; }
                        double x[3];
                        for (int d = 0; d < N; ++d) {
                            const double wd = hi[(size_t)a * N + d] - lo[(size_t)a * N + d];
                            const double rl = (lop[(size_t)i * N + d] - lo[(size_t)a * N + d]) / wd;
                            const double rh = (hip[(size_t)i * N + d] - lo[(size_t)a * N + d]) / wd;
                            x[d] = rl + (rh - rl) * Q.nodes[it[d]];
                        }
                        Q.lagrangeN(x, Ei + (size_t)q * bc);
                    }
                }
            }
        });
        // members of each agglomerate, in row order
        std::vector<int> moff(nc + 1, 0), mem(nprev);
        for (int v : part) moff[v + 1]++;
        for (int c = 0; c < nc; ++c) moff[c + 1] += moff[c];
        {
            std::vector<int> pos(moff.begin(), moff.end() - 1);
            for (int i = 0; i < nprev; ++i) mem[pos[part[i]]++] = i;
        }
        // Gram matrices of the coarse box bases over the mesh (sum over the
        // finer rows of E^T G_fine E, G_fine = element mass on level 0) and
        // their orthonormal bases
        std::vector<double> G((size_t)nc * bc * bc), C((size_t)nc * bc * bc);
        rank[l].resize(nc);
        parallel_for(nc, [&](long a0, long a1) {
            std::vector<double> Wf((size_t)rb * bc), Me((size_t)bf * bf);
            for (long a = a0; a < a1; ++a) {
                double* Ga = &G[(size_t)a * bc * bc];
                std::fill(Ga, Ga + (size_t)bc * bc, 0.0);
                for (int t = moff[a]; t < moff[a + 1]; ++t) {
                    const int i = mem[t];
                    const double* Gi;
                    if (l == 0) {
                        const double sc = m.vol[i] / E.volume;
                        for (size_t k = 0; k < Me.size(); ++k) Me[k] = E.mass[k] * sc;
                        Gi = Me.data();
                    } else Gi = &Gp[(size_t)i * bc * bc];
                    const double* Ei = &Eb[(size_t)i * rb * bc];
                    mm_nn(rb, bc, rb, Gi, Ei, Wf.data());
                    mm_tn(bc, bc, rb, Ei, Wf.data(), Ga, true);
                }
                rank[l][a] = orth_basis(bc, Ga, &C[(size_t)a * bc * bc]);
            }
        });
        // rows in the orthonormal bases: E_i C_a from the mesh, (C_i^T G_i) E_i C_a above
        Interp& I = Iv[l];
        I.nf = nprev; I.nc = nc; I.bf = rb; I.bc = bc;
        I.parent.assign(part.begin(), part.end());
        I.type.resize(nprev);
        I.pool.resize((size_t)nprev * rb * bc);
        I.ntypes = nprev;
        parallel_for(nprev, [&](long i0, long i1) {
            std::vector<double> Wf((size_t)rb * bc), W2((size_t)bc * bc);
            for (long i = i0; i < i1; ++i) {
                const int a = part[i];
                I.type[i] = (int)i;
                double* out = &I.pool[(size_t)i * rb * bc];
                if (l == 0) { mm_nn(rb, bc, bc, &Eb[(size_t)i * rb * bc], &C[(size_t)a * bc * bc], out); continue; }
                mm_nn(rb, bc, bc, &Eb[(size_t)i * rb * bc], &C[(size_t)a * bc * bc], Wf.data());
                mm_tn(bc, bc, bc, &Cp[(size_t)i * bc * bc], &Gp[(size_t)i * bc * bc], W2.data());
                mm_nn(bc, bc, bc, W2.data(), Wf.data(), out);
            }
        });
        I.make_gather();
        Gp.swap(G); Cp.swap(C); lop.swap(lo); hip.swap(hi);
        for (int e = 0; e < m.ne; ++e) orig[e] = part[orig[e]];
        nprev = nc;
    }
}

}  // namespace mgldg
