// host.cpp — basis, quadtree, mesh, LDG face/element assembly and symbolic
// block-sparse products.  See host.hpp.
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace mgldg {

static int ipow(int b, int e) { int r = 1; while (e-- > 0) r *= b; return r; }

// multi-index of a tensor node, first coordinate slowest (src/cartesian/util/range.h:187-194)
static inline void split_index(int lin, int N, int P, int* it) { for (int k = N - 1; k >= 0; --k) { it[k] = lin % P; lin /= P; } }
static inline int join_index(const int* it, int N, int P) { int l = 0; for (int k = 0; k < N; ++k) l = l * P + it[k]; return l; }

// ======================================================================= basis
// Legendre P_n and P_n' at x in (-1,1), long double.
static void legendre_ld(int n, long double x, long double& p, long double& dp)
{
    long double a = 1, b = x;
    if (n == 0) { p = 1; dp = 0; return; }
    for (int k = 2; k <= n; ++k) { long double c = ((2 * k - 1) * x * b - (k - 1) * a) / k; a = b; b = c; }
    p = b; dp = n * (a - x * b) / (1 - x * x);
}

// Gauss-Lobatto points on [0,1] (GaussLobatto<P>::nodes, master.h:174-179):
// interior points are the zeros of P'_{P-1}; P = 1 is the midpoint.
void gl_nodes(int P, double* out)
{
    if (P == 1) { out[0] = 0.5; return; }
    const int n = P - 1;
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<long double> z(P);
    z[0] = -1; z[n] = 1;
    for (int k = 1; k < n; ++k) {
        long double t = -std::cos(pi * k / n);
        for (int it = 0; it < 200; ++it) {
            long double p, dp; legendre_ld(n, t, p, dp);
            // (1-t^2) P'' = 2 t P' - n(n+1) P
            long double ddp = (2 * t * dp - n * (n + 1) * p) / (1 - t * t);
            long double step = dp / ddp; t -= step;
            if (std::fabs(step) < 1e-22L) break;
        }
        z[k] = t;
    }
    for (int k = 0; k < P; ++k) out[k] = (double)(((z[k] + 1) / 2 + (1 - (z[n - k] + 1) / 2)) / 2);
    out[0] = 0; out[n] = 1;
    if (P % 2) out[P / 2] = 0.5;
}

// Gauss-Legendre rule on [0,1] (Quadrature<Q>, master.h:188-214)
void gauss_rule(int Q, double* x, double* w)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<long double> t(Q), wt(Q);
    for (int k = 0; k < Q; ++k) {
        long double s = -std::cos(pi * (k + 0.75L) / (Q + 0.5L)), p = 0, dp = 0;
        for (int it = 0; it < 200; ++it) {
            legendre_ld(Q, s, p, dp);
            long double step = p / dp; s -= step;
            if (std::fabs(step) < 1e-22L) break;
        }
        legendre_ld(Q, s, p, dp);
        t[k] = s; wt[k] = 2 / ((1 - s * s) * dp * dp);
    }
    for (int k = 0; k < Q; ++k) {
        x[k] = (double)(((t[k] + 1) / 2 + (1 - (t[Q - 1 - k] + 1) / 2)) / 2);
        w[k] = (double)((wt[k] + wt[Q - 1 - k]) / 4);
    }
    if (Q % 2) x[Q / 2] = 0.5;
}

void Basis::init(int N_, int P_)
{
    N = N_; P = P_; npl = ipow(P, N); nfl = ipow(P, N - 1);
    gl_nodes(P, nodes);
    gauss_rule(kQuad, qx, qw);
    // barycentric weights = Lagrange denominators (master.h:181-186)
    long double bw[kMaxP];
    for (int i = 0; i < P; ++i) {
        long double d = 1;
        for (int j = 0; j < P; ++j) if (j != i) d *= (long double)nodes[i] - nodes[j];
        bw[i] = 1 / d; denom[i] = (double)bw[i];
    }
    // 1-D mass: exact Gauss integration of l_i l_j (degree 2P-2)
    double gx[kMaxP + 1], gw[kMaxP + 1];
    gauss_rule(P + 1, gx, gw);
    std::vector<long double> lv((P + 1) * P);
    for (int q = 0; q <= P; ++q)
        for (int i = 0; i < P; ++i) {
            long double v = bw[i];
            for (int j = 0; j < P; ++j) if (j != i) v *= (long double)gx[q] - nodes[j];
            lv[q * P + i] = v;
        }
    mass1.assign(P * P, 0); invmass1.assign(P * P, 0); diff1.assign(P * P, 0);
    std::vector<long double> m(P * P);
    for (int i = 0; i < P; ++i) for (int j = 0; j < P; ++j) {
        long double s = 0; for (int q = 0; q <= P; ++q) s += (long double)gw[q] * lv[q * P + i] * lv[q * P + j];
        m[i * P + j] = s; mass1[i * P + j] = (double)s;
    }
    // inverse by Cholesky in long double
    std::vector<long double> L(P * P, 0), Li(P * P, 0);
    for (int j = 0; j < P; ++j) {
        long double s = m[j * P + j]; for (int k = 0; k < j; ++k) s -= L[j * P + k] * L[j * P + k];
        L[j * P + j] = std::sqrt(s);
        for (int i = j + 1; i < P; ++i) { long double t = m[i * P + j]; for (int k = 0; k < j; ++k) t -= L[i * P + k] * L[j * P + k]; L[i * P + j] = t / L[j * P + j]; }
    }
    for (int c = 0; c < P; ++c) for (int i = 0; i < P; ++i) {   // L^{-1}
        long double t = (i == c); for (int k = 0; k < i; ++k) t -= L[i * P + k] * Li[k * P + c];
        Li[i * P + c] = t / L[i * P + i];
    }
    for (int i = 0; i < P; ++i) for (int j = 0; j < P; ++j) {
        long double s = 0; for (int k = 0; k < P; ++k) s += Li[k * P + i] * Li[k * P + j];
        invmass1[i * P + j] = (double)s;
    }
    // differentiation D(i,j) = l_j'(x_i) (master.h:166-171)
    for (int i = 0; i < P; ++i) {
        long double dsum = 0;
        for (int j = 0; j < P; ++j) if (j != i) {
            long double v = (bw[j] / bw[i]) / ((long double)nodes[i] - nodes[j]);
            diff1[i * P + j] = (double)v; dsum += v;
        }
        diff1[i * P + i] = (double)(-dsum);
    }
    // tensor products (master.h:130-145, 45-69)
    int it[3], jt[3];
    mass.assign((size_t)npl * npl, 0); invmass.assign((size_t)npl * npl, 0);
    for (int d = 0; d < 3; ++d) diff[d].assign(d < N ? (size_t)npl * npl : 0, 0);
    for (int i = 0; i < npl; ++i) for (int j = 0; j < npl; ++j) {
        split_index(i, N, P, it); split_index(j, N, P, jt);
        double a = 1, b = 1;
        for (int k = 0; k < N; ++k) { a *= mass1[it[k] * P + jt[k]]; b *= invmass1[it[k] * P + jt[k]]; }
        mass[(size_t)i * npl + j] = a; invmass[(size_t)i * npl + j] = b;
        for (int d = 0; d < N; ++d) {
            bool same = true;
            for (int k = 0; k < N; ++k) if (k != d && it[k] != jt[k]) same = false;
            if (same) diff[d][(size_t)i * npl + j] = diff1[it[d] * P + jt[d]];
        }
    }
    fmass.assign((size_t)nfl * nfl, 0);
    for (int i = 0; i < nfl; ++i) for (int j = 0; j < nfl; ++j) {
        split_index(i, N - 1, P, it); split_index(j, N - 1, P, jt);
        double a = 1; for (int k = 0; k < N - 1; ++k) a *= mass1[it[k] * P + jt[k]];
        fmass[(size_t)i * nfl + j] = a;
    }
}

// LagrangePoly<P>::eval(double) (master.h:78-95)
void Basis::lagrange1(double x, double* out) const
{
    double a = 1; out[0] = denom[0];
    for (int i = 0; i + 1 < P; ++i) { a *= x - nodes[i]; out[i + 1] = a * denom[i + 1]; }
    double b = 1;
    for (int i = P - 1; i > 0; --i) { b *= x - nodes[i]; out[i - 1] *= b; }
}

void Basis::lagrangeN(const double* x, double* out) const
{
    double g[3][kMaxP]; int it[3];
    for (int d = 0; d < N; ++d) lagrange1(x[d], g[d]);
    for (int i = 0; i < npl; ++i) {
        split_index(i, N, P, it);
        double v = 1; for (int d = 0; d < N; ++d) v *= g[d][it[d]];
        out[i] = v;
    }
}

// ======================================================================= quadtree
void Quadtree::split(int id)
{
    const int nc = 1 << N;
    QNode parent = nodes[id];
    int first = (int)nodes.size();
    nodes[id].first_child = first;
    double mid[3];
    for (int d = 0; d < N; ++d) mid[d] = parent.lo[d] + (parent.hi[d] - parent.lo[d]) / 2;
    for (int c = 0; c < nc; ++c) {
        QNode ch{};
        for (int d = 0; d < N; ++d) {
            bool up = (c >> (N - 1 - d)) & 1;     // RangeIterator<2,N>: last coordinate fastest
            ch.lo[d] = up ? mid[d] : parent.lo[d];
            ch.hi[d] = up ? parent.hi[d] : mid[d];
        }
        ch.level = parent.level + 1; ch.height = 0; ch.parent = id; ch.first_child = -1;
        nodes.push_back(ch);
    }
}

// Quadtree::build (quadtree.tpp:8-40): a cell splits iff every width exceeds
// h(center); children are created together (ids are internal; the element
// order is the depth-first leaf order either way).
void Quadtree::build(const double* lo, const double* hi, void (*h)(const double*, double*, void*), void* user)
{
    nodes.clear();
    QNode root{};
    for (int d = 0; d < N; ++d) { root.lo[d] = lo ? lo[d] : 0.0; root.hi[d] = hi ? hi[d] : 1.0; }
    root.level = 0; root.height = 0; root.parent = -1; root.first_child = -1;
    nodes.push_back(root);
    num_levels = 1;
    // explicit stack; heights are fixed bottom-up afterwards
    std::vector<int> stack{0};
    while (!stack.empty()) {
        int id = stack.back(); stack.pop_back();
        num_levels = std::max(num_levels, nodes[id].level + 1);
        double c[3], w[3] = {0, 0, 0};
        for (int d = 0; d < N; ++d) c[d] = nodes[id].lo[d] + (nodes[id].hi[d] - nodes[id].lo[d]) / 2;
        h(c, w, user);
        bool sp = true;
        for (int d = 0; d < N; ++d) if (!((nodes[id].hi[d] - nodes[id].lo[d]) > w[d])) sp = false;
        if (!sp) continue;
        split(id);
        for (int k = (1 << N) - 1; k >= 0; --k) stack.push_back(nodes[id].first_child + k);
    }
    // heights: children always have larger ids than their parent
    for (int id = (int)nodes.size() - 1; id >= 0; --id) {
        QNode& n = nodes[id];
        n.height = 0;
        if (n.first_child >= 0)
            for (int k = 0; k < (1 << N); ++k) n.height = std::max(n.height, nodes[n.first_child + k].height + 1);
    }
}

// Quadtree::isLeaf (quadtree.h:256-265)
bool Quadtree::is_leaf(int id, int c) const
{
    const QNode& n = nodes[id];
    if (rapid) return n.height == c || (n.height < c && nodes[n.parent].height > c);
    int cl = num_levels - 1 - c;
    return n.level == cl || (n.level < cl && n.height == 0);
}

// Quadtree::layer (quadtree.h:268-278): leaves of coarsening c in DFS order
std::vector<int> Quadtree::layer(int c) const
{
    std::vector<int> out, st{0};
    while (!st.empty()) {
        int id = st.back(); st.pop_back();
        if (is_leaf(id, c)) { out.push_back(id); continue; }
        for (int k = (1 << N) - 1; k >= 0; --k) st.push_back(nodes[id].first_child + k);
    }
    return out;
}

// isNeighbor (quadtree.tpp:59-88): b lies across a's face (dim, dir)
bool Quadtree::touches(const QNode& a, const QNode& b, int dim, int dir) const
{
    double pa = dir == 0 ? a.lo[dim] : a.hi[dim];
    double pb = dir == 0 ? b.hi[dim] : b.lo[dim];
    if (periodic) {
        double dl = nodes[0].lo[dim], du = nodes[0].hi[dim];
        if ((dir == 0 && pa == dl && pb == du) || (dir == 1 && pa == du && pb == dl)) pb = pa;
    }
    if (pa != pb) return false;
    for (int i = 0; i < N; ++i)
        if (i != dim && !(b.hi[i] > a.lo[i] && b.lo[i] < a.hi[i])) return false;
    return true;
}

// isParent (quadtree.tpp:91-98): b strictly contains a
bool Quadtree::contains(const QNode& a, const QNode& b) const
{
    bool in = true, eq = true;
    for (int i = 0; i < N; ++i) {
        if (!(a.lo[i] >= b.lo[i] && a.hi[i] <= b.hi[i])) in = false;
        if (!(a.lo[i] == b.lo[i] && a.hi[i] == b.hi[i])) eq = false;
    }
    return in && !eq;
}

// neighbors/search (quadtree.tpp:45-56, 101-114): pruned descent, results in DFS order
void Quadtree::search(int a, int dim, int dir, int id, int c, std::vector<int>& out) const
{
    const QNode& A = nodes[a]; const QNode& B = nodes[id];
    if (!(touches(A, B, dim, dir) || contains(A, B))) return;
    if (is_leaf(id, c)) { out.push_back(id); return; }
    for (int k = 0; k < (1 << N); ++k) search(a, dim, dir, B.first_child + k, c, out);
}

void Quadtree::neighbors(int id, int dim, int dir, int c, std::vector<int>& out) const
{
    out.clear();
    search(id, dim, dir, 0, c, out);
}

// 2:1 balance (not in the reference): split leaves that have a face neighbour
// more than one level finer until none remains.
void Quadtree::balance()
{
    std::vector<int> nb;
    for (bool changed = true; changed;) {
        changed = false;
        std::vector<int> leaves = layer(0);
        for (int id : leaves) {
            bool need = false;
            for (int d = 0; d < N && !need; ++d)
                for (int dir = 0; dir < 2 && !need; ++dir) {
                    neighbors(id, d, dir, 0, nb);
                    for (int q : nb) if (nodes[q].level > nodes[id].level + 1) { need = true; break; }
                }
            if (!need) continue;
            split(id);
            nodes[id].height = 1;
            for (int c = id, p = nodes[id].parent; p >= 0; c = p, p = nodes[p].parent) {
                if (nodes[p].height >= nodes[c].height + 1) break;
                nodes[p].height = nodes[c].height + 1;
            }
            num_levels = std::max(num_levels, nodes[id].level + 2);
            changed = true;
        }
    }
}

// ======================================================================= BSR
int HostBSR::find(int i, int j) const
{
    auto b = cols.begin() + rowptr[i], e = cols.begin() + rowptr[i + 1];
    auto it = std::lower_bound(b, e, j);
    return (it != e && *it == j) ? (int)(it - cols.begin()) : -1;
}

void PatternBuilder::finish(HostBSR& out, int br, int bc, bool values)
{
    out.nbr = nbr; out.nbc = nbc; out.br = br; out.bc = bc;
    // counting sort of the pairs by row, then sort and deduplicate each row
    std::vector<int> off(nbr + 1, 0);
    for (int i : ri) off[i + 1]++;
    for (int i = 0; i < nbr; ++i) off[i + 1] += off[i];
    std::vector<int> tmp(ci.size());
    {
        std::vector<int> pos(off.begin(), off.end() - 1);
        for (size_t t = 0; t < ri.size(); ++t) tmp[pos[ri[t]]++] = ci[t];
    }
    ri.clear(); ri.shrink_to_fit(); ci.clear(); ci.shrink_to_fit();
    out.rowptr.assign(nbr + 1, 0);
    int w = 0;
    for (int i = 0; i < nbr; ++i) {
        auto b = tmp.begin() + off[i], e = tmp.begin() + off[i + 1];
        std::sort(b, e);
        e = std::unique(b, e);
        for (auto it = b; it != e; ++it) tmp[w++] = *it;
        out.rowptr[i + 1] = w;
    }
    tmp.resize(w);
    out.cols.swap(tmp);
    if (values) out.val.assign((size_t)out.cols.size() * br * bc, 0.0);
    else out.val.clear();
}

// ======================================================================= mesh
double Mesh::volume(int e) const { double v = 1; for (int d = 0; d < N; ++d) v *= width(e, d); return v; }

// Mesh<N,P>(qt) (mesh.h:186-265) and the Face constructor (mesh.h:21-79)
void Mesh::build(const Quadtree& qt, const Basis& B)
{
    N = qt.N;
    gid = qt.layer(0);
    ne = (int)gid.size();
    std::vector<int> g2l(qt.nodes.size(), -1);
    lo.resize((size_t)ne * N); hi.resize((size_t)ne * N);
    for (int e = 0; e < ne; ++e) {
        g2l[gid[e]] = e;
        for (int d = 0; d < N; ++d) { lo[e * N + d] = qt.nodes[gid[e]].lo[d]; hi[e * N + d] = qt.nodes[gid[e]].hi[d]; }
    }
    const int nq = (N == 2) ? kQuad : kQuad * kQuad;
    faces.clear(); bfaces.clear(); face_q.clear();
    std::vector<int> nb;
    for (int e = 0; e < ne; ++e)
        for (int d = 0; d < N; ++d)
            for (int dir = 0; dir < 2; ++dir) {
                qt.neighbors(gid[e], d, dir, 0, nb);
                if (nb.empty()) {
                    Face f{};
                    f.left = e; f.right = -(2 * d + 1 + dir); f.dim = d; f.canonical = 1; f.qoff = -1;
                    f.sign = dir == 0 ? -1.0 : 1.0;
                    for (int k = 0; k < N; ++k) { f.lo[k] = lo[e * N + k]; f.hi[k] = hi[e * N + k]; }
                    if (dir == 0) f.hi[d] = f.lo[d]; else f.lo[d] = f.hi[d];
                    bfaces.push_back((int)faces.size());
                    faces.push_back(f);
                    continue;
                }
                for (int g : nb) {
                    int o = g2l[g];
                    if (!(e < o)) continue;            // each interior face once (mesh.h:241)
                    Face f{};
                    f.left = dir == 0 ? o : e; f.right = dir == 1 ? o : e; f.dim = d; f.sign = 1.0; f.qoff = -1;
                    const int L = f.left, R = f.right;
                    for (int k = 0; k < N; ++k) {
                        f.lo[k] = std::max(lo[L * N + k], lo[R * N + k]);
                        f.hi[k] = std::min(hi[L * N + k], hi[R * N + k]);
                    }
                    f.lo[d] = lo[R * N + d]; f.hi[d] = hi[L * N + d];
                    // canonical iff the two element sides coincide
                    bool canon = true;
                    for (int k = 0; k < N; ++k) {
                        double llo = (k == d) ? hi[L * N + k] : lo[L * N + k], lhi = hi[L * N + k];
                        double rlo = lo[R * N + k], rhi = (k == d) ? lo[R * N + k] : hi[R * N + k];
                        if (llo != rlo || lhi != rhi) canon = false;
                    }
                    f.canonical = canon;
                    if (!canon) {
                        // local quadrature points w.r.t. both elements (mesh.h:44-77)
                        f.qoff = (int)face_q.size();
                        face_q.resize(face_q.size() + 2 * (size_t)nq * N);
                        double* xl = face_q.data() + f.qoff; double* xr = xl + (size_t)nq * N;
                        double al[3], bl[3], ar[3], br[3];
                        for (int k = 0; k < N; ++k) {
                            double l0 = lo[L * N + k], r0 = lo[R * N + k];
                            if (k == d) l0 = hi[L * N + k];
                            al[k] = f.lo[k] - l0; bl[k] = f.hi[k] - l0; ar[k] = f.lo[k] - r0; br[k] = f.hi[k] - r0;
                            if (k != d) { double wl = width(L, k), wr = width(R, k); al[k] /= wl; bl[k] /= wl; ar[k] /= wr; br[k] /= wr; }
                        }
                        int it[3];
                        for (int q = 0; q < nq; ++q) {
                            split_index(q, N - 1, kQuad, it);
                            for (int k = 0; k < N - 1; ++k) {
                                double x = B.qx[it[k]]; int kk = k + (k >= d);
                                xl[q * N + kk] = al[kk] + (bl[kk] - al[kk]) * x;
                                xr[q * N + kk] = ar[kk] + (br[kk] - ar[kk]) * x;
                            }
                            xl[q * N + d] = 1.0; xr[q * N + d] = 0.0;
                        }
                    }
                    faces.push_back(f);
                }
            }
    nf = (int)faces.size();
    nbf = (int)bfaces.size();
}

// ======================================================================= LDG
// Rows of the device assembly from the patterns: per element its T / G_d
// diagonal blocks and, per listed face, the block (e, other side) of T and
// (e, left) of G, with first-term flags (two faces may join the same pair
// of elements on a periodic quadtree of width two).
void AsmPlan::finish_rows(int ne, const HostBSR& T, const HostBSR* G)
{
    tdiag.assign(ne, -1); gdiag.assign((size_t)ne * N, -1);
    ef_tslot.assign(ef.size(), -1); ef_gslot.assign(ef.size(), -1); ef_first.assign(ef.size(), 0);
    std::vector<int> gdim(ef.size(), 0);      // which G_d pattern ef_gslot indexes
    for (int e = 0; e < ne; ++e) {
        tdiag[e] = T.find(e, e);
        for (int d = 0; d < N; ++d) gdiag[(size_t)e * N + d] = G[d].find(e, e);
        for (int k = ef_off[e]; k < ef_off[e + 1]; ++k) {
            const int f = ef[k] >> 1, side = ef[k] & 1;
            if (fkind[f] == kDirichlet || fkind[f] == kNone) continue;
            const int other = side ? fl[f] : fr[f];
            ef_tslot[k] = T.find(e, other);
            if (ef_tslot[k] < 0) throw Error(-4, "assembly: missing block in symbolic pattern");
            if (side) {
                int gd = 0;
                if (!gshared) while (gd < N && fnormal[(size_t)f * N + gd] == 0.0) ++gd;
                ef_gslot[k] = G[gd].find(e, other);
                gdim[k] = gd;
                if (ef_gslot[k] < 0) throw Error(-4, "assembly: missing block in symbolic pattern");
            }
            int first = 3;
            for (int q = ef_off[e]; q < k; ++q) {
                if (ef_tslot[q] == ef_tslot[k]) first &= ~1;
                if (ef_gslot[k] >= 0 && ef_gslot[q] == ef_gslot[k] && gdim[q] == gdim[k]) first &= ~2;
            }
            ef_first[k] = first;
        }
    }
}

// LDGPoisson::discretize without the Laplacian (ldgpoisson.tpp:14-100), the
// symbolic half: the mesh, the patterns of M, T and G_d (addToBlock creates
// structural blocks, sparseblockmatrix.tpp:92-100) and the plan from which
// the device forms every block (mass, broken gradient with its volume and
// lifting terms, penalty) and the boundary lifts.
void LdgHost::assemble(const Quadtree& qt, int p_, const int* bct, double t0, double tD)
{
    N = qt.N; p = p_; tau0 = t0; tauD = tD;
    for (int k = 0; k < 2 * N; ++k) {
        bctype[k] = bct ? bct[k] : 0;
        if (bctype[k] != 0 && bctype[k] != 1) throw Error(-1, "unknown boundary condition");   // ldgpoisson.tpp:145
    }
    basis.init(N, p + 1);
    mesh.build(qt, basis);
    npl = basis.npl; nq = (N == 2) ? kQuad : kQuad * kQuad;
    const int ne = mesh.ne, b = npl, P = basis.P, nfl = basis.nfl;
    this->ne = ne; nf = mesh.nf; nbf = mesh.nbf;
    auto dirichlet = [&](const Face& f) { return f.right < 0 && bctype[-f.right - 1] == 0; };

    // patterns
    {
        PatternBuilder pm(ne, ne), pt(ne, ne);
        std::vector<PatternBuilder> pg(N, PatternBuilder(ne, ne));
        for (int e = 0; e < ne; ++e) { pm.add(e, e); for (int d = 0; d < N; ++d) pg[d].add(e, e); }
        for (const Face& f : mesh.faces) {
            if (f.right >= 0) {
                pg[f.dim].add(f.right, f.right); pg[f.dim].add(f.right, f.left);
                pt.add(f.right, f.right); pt.add(f.left, f.left); pt.add(f.right, f.left); pt.add(f.left, f.right);
            } else if (dirichlet(f)) {
                pg[f.dim].add(f.left, f.left); pt.add(f.left, f.left);
            }
        }
        pm.finish(M, b, b, false); pt.finish(T, b, b, false);
        for (int d = 0; d < N; ++d) pg[d].finish(G[d], b, b, false);
    }

    // the plan.  Code 2d + s: the face normal to dimension d on the element
    // side where that index is 0 (s = 0) or P - 1 (s = 1).
    AsmPlan& A = plan;
    A = AsmPlan();
    A.N = N; A.b = b; A.nfl = nfl; A.ncode = 2 * N; A.nq = nq; A.gshared = 0;
    A.emass = basis.mass; A.einv = basis.invmass; A.fmass = basis.fmass;
    for (int d = 0; d < N; ++d) A.ediff.insert(A.ediff.end(), basis.diff[d].begin(), basis.diff[d].end());
    A.slice.assign((size_t)A.ncode * nfl, 0); A.islice.assign((size_t)A.ncode * b, -1);
    for (int code = 0; code < A.ncode; ++code) {
        const int d = code / 2;
        int it[3], jt[3];
        for (int a = 0; a < nfl; ++a) {
            split_index(a, N - 1, P, it);
            for (int k = 0, kk = 0; k < N; ++k) if (k != d) jt[k] = it[kk++];
            jt[d] = (code & 1) ? P - 1 : 0;
            const int i = join_index(jt, N, P);
            A.slice[(size_t)code * nfl + a] = i;
            A.islice[(size_t)code * b + i] = a;
        }
    }
    // reference boundary blocks per code: LL = sum_q w_q phi phi^T and the
    // lift phi_i(q) w_q at the face rule (Mesh::lift, mesh.h:282-352, on a
    // face of unit area)
    std::vector<double> wq(nq), phi((size_t)nq * b);
    A.pool.assign((size_t)A.ncode * b * b, 0.0); A.npool_ref = A.ncode;
    A.lbref.assign((size_t)A.ncode * b * nq, 0.0);
    for (int code = 0; code < A.ncode; ++code) {
        const int d = code / 2;
        int it[3];
        for (int q = 0; q < nq; ++q) {
            double x[3];
            split_index(q, N - 1, kQuad, it);
            for (int k = 0; k < N - 1; ++k) x[k + (k >= d)] = basis.qx[it[k]];
            x[d] = (code & 1) ? 1.0 : 0.0;
            basis.lagrangeN(x, &phi[(size_t)q * b]);
            double w = 1; for (int k = 0; k < N - 1; ++k) w *= basis.qw[it[k]];
            wq[q] = w;
        }
        double* LL = &A.pool[(size_t)code * b * b];
        for (int i = 0; i < b; ++i)
            for (int j = 0; j < b; ++j) {
                double s = 0;
                for (int q = 0; q < nq; ++q) s += (phi[(size_t)q * b + i] * wq[q]) * phi[(size_t)q * b + j];
                LL[(size_t)i * b + j] = s;
            }
        double* Lr = &A.lbref[(size_t)code * b * nq];
        for (int i = 0; i < b; ++i) for (int q = 0; q < nq; ++q) Lr[(size_t)i * nq + q] = phi[(size_t)q * b + i] * wq[q];
    }
    A.lnodes.assign(basis.nodes, basis.nodes + P); A.ldenom.assign(basis.denom, basis.denom + P);
    // elements: M_e = |e| mass, M_e^-1 = invmass / |e|, G_d volume term diff_d / h_d
    A.mscale.resize(ne); A.iscale.resize(ne); A.C.assign((size_t)ne * N * N, 0.0);
    for (int e = 0; e < ne; ++e) {
        A.mscale[e] = mesh.volume(e); A.iscale[e] = 1.0 / mesh.volume(e);
        for (int d = 0; d < N; ++d) A.C[((size_t)e * N + d) * N + d] = 1.0 / mesh.width(e, d);
    }
    // faces
    const int nfc = mesh.nf;
    A.fl.resize(nfc); A.fr.resize(nfc); A.fkind.resize(nfc); A.fca.assign(nfc, 0); A.fcb.assign(nfc, 0);
    A.fscale.assign(nfc, 0.0); A.fnormal.assign((size_t)nfc * N, 0.0);
    std::vector<int> cnt(ne + 1, 0);
    for (int fi = 0; fi < nfc; ++fi) {
        const Face& f = mesh.faces[fi];
        double area = 1; for (int k = 0; k < N; ++k) if (k != f.dim) area *= f.hi[k] - f.lo[k];
        A.fl[fi] = f.left; A.fr[fi] = f.right; A.fnormal[(size_t)fi * N + f.dim] = f.sign;
        if (f.right >= 0) {
            if (f.canonical) {
                A.fkind[fi] = AsmPlan::kSlice; A.fca[fi] = 2 * f.dim + 1; A.fcb[fi] = 2 * f.dim; A.fscale[fi] = area;
            } else {
                // points from the mesh (mesh.h:44-77), weights with the area
                A.fkind[fi] = AsmPlan::kDense;
                A.fca[fi] = A.npool_ref + 3 * (int)A.qface.size();
                A.qface.push_back(fi);
                const double* xq = mesh.face_q.data() + f.qoff;
                A.qpts.insert(A.qpts.end(), xq, xq + 2 * (size_t)nq * N);
                for (int q = 0; q < nq; ++q) {
                    int it[3]; split_index(q, N - 1, kQuad, it);
                    double w = 1; for (int k = 0; k < N - 1; ++k) w *= basis.qw[it[k]];
                    A.qwt.push_back(w * area);
                }
            }
            cnt[f.left + 1]++; cnt[f.right + 1]++;
        } else if (dirichlet(f)) {
            A.fkind[fi] = AsmPlan::kDirichlet; A.fca[fi] = 2 * f.dim + (f.sign > 0 ? 1 : 0); A.fscale[fi] = area;
            cnt[f.left + 1]++;
        } else {
            A.fkind[fi] = AsmPlan::kNone;
        }
    }
    for (int e = 0; e < ne; ++e) cnt[e + 1] += cnt[e];
    A.ef_off = cnt; A.ef.resize(cnt[ne]);
    for (int fi = 0; fi < nfc; ++fi) {
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
        const Face& f = mesh.faces[mesh.bfaces[k]];
        double area = 1; for (int q = 0; q < N; ++q) if (q != f.dim) area *= f.hi[q] - f.lo[q];
        A.bcode[k] = 2 * f.dim + (f.sign > 0 ? 1 : 0); A.bscale[k] = area;
        bleft[k] = f.left; btype[k] = bctype[-f.right - 1]; btag[k] = f.right;
        bnormal[(size_t)k * N + f.dim] = f.sign;
    }
}

void LdgHost::bc_qpoints(double* xyz, int* tag) const
{
    int it[3];
    for (int k = 0; k < mesh.nbf; ++k) {
        const Face& f = mesh.faces[mesh.bfaces[k]];
        for (int q = 0; q < nq; ++q) {
            split_index(q, N - 1, kQuad, it);
            double* x = xyz + ((size_t)k * nq + q) * N;
            for (int j = 0; j < N - 1; ++j) { int kk = j + (j >= f.dim); x[kk] = f.lo[kk] + (f.hi[kk] - f.lo[kk]) * basis.qx[it[j]]; }
            x[f.dim] = f.lo[f.dim];
            if (tag) tag[(size_t)k * nq + q] = f.right;
        }
    }
}

void LdgHost::nodes_xyz(double* xyz) const
{
    int it[3];
    for (int e = 0; e < mesh.ne; ++e)
        for (int i = 0; i < npl; ++i) {
            split_index(i, N, basis.P, it);
            for (int d = 0; d < N; ++d) xyz[((size_t)e * npl + i) * N + d] = mesh.lo[e * N + d] + mesh.width(e, d) * basis.nodes[it[d]];
        }
}

// ======================================================================= symbolic products
void laplacian_symbolic(const HostBSR& T, const HostBSR* G, int N, const HostBSR& M, LaplacianSymbolic& out)
{
    const int n = T.nbr;
    PatternBuilder pb(n, n);
    for (int i = 0; i < n; ++i) for (int k = T.rowptr[i]; k < T.rowptr[i + 1]; ++k) pb.add(i, T.cols[k]);
    for (int d = 0; d < N; ++d)
        for (int k = 0; k < n; ++k)
            for (int a = G[d].rowptr[k]; a < G[d].rowptr[k + 1]; ++a)
                for (int c = G[d].rowptr[k]; c < G[d].rowptr[k + 1]; ++c) pb.add(G[d].cols[a], G[d].cols[c]);
    pb.finish(out.pattern, T.br, T.bc, false);
    const HostBSR& A = out.pattern;
    out.tslot.assign(A.nnzb(), -1);
    for (int i = 0; i < n; ++i) for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) out.tslot[k] = T.find(i, A.cols[k]);
    // terms: per A block, in (d, k) order as construct_laplacian accumulates them
    std::vector<int> cnt(A.nnzb() + 1, 0);
    auto each = [&](auto&& fn) {
        long goff = 0;
        for (int d = 0; d < N; ++d) {
            for (int k = 0; k < n; ++k)
                for (int a = G[d].rowptr[k]; a < G[d].rowptr[k + 1]; ++a)
                    for (int c = G[d].rowptr[k]; c < G[d].rowptr[k + 1]; ++c)
                        fn(A.find(G[d].cols[a], G[d].cols[c]), (int)(goff + a), k, (int)(goff + c));
            goff += G[d].nnzb();
        }
    };
    each([&](int slot, int, int, int) { cnt[slot + 1]++; });
    TermList& tl = out.gmg;
    tl.off.assign(A.nnzb() + 1, 0);
    for (long s = 0; s < A.nnzb(); ++s) tl.off[s + 1] = tl.off[s] + cnt[s + 1];
    tl.l.resize(tl.off.back()); tl.x.resize(tl.off.back()); tl.r.resize(tl.off.back());
    std::vector<int> pos(tl.off.begin(), tl.off.end() - 1);
    (void)M;
    each([&](int slot, int gl, int k, int gr) { int p = pos[slot]++; tl.l[p] = gl; tl.x[p] = k; tl.r[p] = gr; });
}

void Interp::make_gather()
{
    kid_off.assign(nc + 1, 0);
    for (int i = 0; i < nf; ++i) kid_off[parent[i] + 1]++;
    for (int c = 0; c < nc; ++c) kid_off[c + 1] += kid_off[c];
    kids.resize(nf);
    std::vector<int> pos(kid_off.begin(), kid_off.end() - 1);
    for (int i = 0; i < nf; ++i) kids[pos[parent[i]]++] = i;
}

// Coarse rows are independent: row I's pattern, its blocks' slots and their
// terms come from the kids of I only, visited in ascending order (kids
// ascending, their blocks ascending), which is the order of a serial sweep
// over the fine rows -- so the rows are formed on up to 16 threads with the
// same pattern and term order.
void rap_symbolic(const Interp& T, const HostBSR& X, HostBSR& C, TermList& terms)
{
    const int nc = T.nc;
    std::vector<int> koff(nc + 1, 0), kid(X.nbr);
    for (int k = 0; k < X.nbr; ++k) koff[T.parent[k] + 1]++;
    for (int c = 0; c < nc; ++c) koff[c + 1] += koff[c];
    {
        std::vector<int> pos(koff.begin(), koff.end() - 1);
        for (int k = 0; k < X.nbr; ++k) kid[pos[T.parent[k]]++] = k;
    }
    std::vector<std::vector<int>> rows(nc);
    parallel_rows(nc, [&](long a, long b) {
        std::vector<int> tmp;
        for (long I = a; I < b; ++I) {
            tmp.clear();
            for (int q = koff[I]; q < koff[I + 1]; ++q) {
                const int k = kid[q];
                for (int s = X.rowptr[k]; s < X.rowptr[k + 1]; ++s) tmp.push_back(T.parent[X.cols[s]]);
            }
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            rows[I] = tmp;
        }
    });
    C.nbr = C.nbc = nc; C.br = C.bc = T.bc; C.val.clear();
    C.rowptr.assign(nc + 1, 0);
    for (int I = 0; I < nc; ++I) C.rowptr[I + 1] = C.rowptr[I] + (int)rows[I].size();
    C.cols.resize(C.rowptr[nc]);
    std::vector<int> cnt(C.nnzb() + 1, 0), slot(X.nnzb());
    parallel_rows(nc, [&](long a, long b) {
        for (long I = a; I < b; ++I) {
            std::copy(rows[I].begin(), rows[I].end(), C.cols.begin() + C.rowptr[I]);
            const auto rb = C.cols.begin() + C.rowptr[I], re = C.cols.begin() + C.rowptr[I + 1];
            for (int q = koff[I]; q < koff[I + 1]; ++q) {
                const int k = kid[q];
                for (int s = X.rowptr[k]; s < X.rowptr[k + 1]; ++s) {
                    slot[s] = (int)(std::lower_bound(rb, re, T.parent[X.cols[s]]) - C.cols.begin());
                    cnt[slot[s] + 1]++;          // (slots of row I: this thread's alone)
                }
            }
        }
    });
    rows.clear(); rows.shrink_to_fit();
    terms.off.assign(C.nnzb() + 1, 0);
    for (long s = 0; s < C.nnzb(); ++s) terms.off[s + 1] = terms.off[s] + cnt[s + 1];
    const long nt = terms.off.back();
    terms.l.resize(nt); terms.x.resize(nt); terms.r.resize(nt);
    std::vector<int> pos(terms.off.begin(), terms.off.end() - 1);
    parallel_rows(nc, [&](long a, long b) {
        for (long I = a; I < b; ++I)
            for (int q = koff[I]; q < koff[I + 1]; ++q) {
                const int k = kid[q];
                for (int s = X.rowptr[k]; s < X.rowptr[k + 1]; ++s) {
                    const int p = pos[slot[s]]++;
                    terms.l[p] = T.type[k]; terms.x[p] = s; terms.r[p] = T.type[X.cols[s]];
                }
            }
    });
}

int greedy_colors(const HostBSR& A, std::vector<int>& color)
{
    color.assign(A.nbr, 0);
    int nc = 0;
    std::vector<int> mark;
    for (int i = 0; i < A.nbr; ++i) {
        mark.assign(nc + 1, 0);
        for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) { int j = A.cols[k]; if (j < i) mark[color[j]] = 1; }
        int c = 0; while (mark[c]) ++c;
        color[i] = c; nc = std::max(nc, c + 1);
    }
    return nc;
}

int wavefronts(const HostBSR& A, bool reverse, std::vector<int>& level)
{
    const int n = A.nbr;
    level.assign(n, 0);
    int nl = 0;
    for (int t = 0; t < n; ++t) {
        const int i = reverse ? n - 1 - t : t;
        int lv = 0;
        for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) {
            const int j = A.cols[k];
            if (reverse ? j > i : j < i) lv = std::max(lv, level[j] + 1);
        }
        level[i] = lv;
        nl = std::max(nl, lv + 1);
    }
    return nl;
}

}  // namespace mgldg
