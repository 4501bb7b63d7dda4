// agglomerate.cpp — multilevel agglomeration of a dual graph into clusters of
// minsize..maxsize vertices with small aspect ratios.
//
// This is the algorithm of MGridGen 1.0, the library the reference calls for
// every agglomeration level (src/unstructured/mesh/agglomeration.h:96-116,
// 188-209, with minsize 1, maxsize 2^N and options {4, 6, 1, N}; the library
// lib/ParMGridGen-1.0/MGridGen/Lib/mgridgen.c:18-49), as published by
// I. Moulitsas and G. Karypis ("Multilevel algorithms for generating coarse
// grids for multigrid methods", SC 2001):
//
//   1. coarsening: rounds of matching that pair adjacent vertices in order of
//      the aspect ratio of the pair (smallest first, at most a quarter of the
//      vertices paired per round, pairs heavier than maxsize excluded), each
//      contracted into a coarser graph, until a round pairs nothing;
//   2. the coarsest graph's vertices are the initial clusters;
//   3. uncoarsening: on every level, from the coarsest to the input graph,
//      the clusters are projected to the finer graph, split into their
//      connected components, clusters under minsize are merged into a
//      neighbour, and a multi-objective k-way refinement moves single
//      vertices between adjacent clusters: first so that the largest aspect
//      ratio of any cluster falls (never rises), then so that the
//      size-weighted sum of aspect ratios falls;
//   4. a final cycle of merging undersized clusters and attracting vertices
//      into them from their neighbours, then a last refinement.
//
// Aspect ratio of a cluster of surface s (its boundary faces plus the faces
// it shares with other clusters) and volume v: s^2 / v in 2-D, s^1.5 / v in
// 3-D.  The matching key is the one MGridGen 1.0 uses (its ARATIO2 macro):
// s^3 / v^2 in 3-D and, as that macro evaluates in 2-D, s^4 (the macro
// divides by the volume and multiplies by it again), so in 2-D pairs are
// matched by the perimeter of the pair.
//
// The random orders (vertex visits of the matching and of each refinement
// pass; random transpositions like MGridGen's, applied within blocks of
// consecutive vertices, see Visit) come from this file's own generator with
// a fixed seed, so results are deterministic; they are not MGridGen's
// drand48 sequence, so partitions are of the same kind and quality but not
// the same bits.  Two additions:
// the final refinement is followed by a split into connected components, so
// every cluster is connected; and the partition is the best (smallest
// largest aspect ratio) of four runs from different random orders, which
// removes most of the run-to-run spread of a single random order.
#include "simplex.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace mgldg {

namespace {

struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

// Random vertex orders that stay cache-friendly: the positions are cut into
// blocks of kBlock consecutive vertices (graph order, which follows the
// element order of the mesh); each new order applies, inside every block,
// as many random transpositions as the block has entries to the block's
// previous order, and visits the blocks in a fresh random order.  A graph
// of at most kBlock vertices gets plain random-transposition orders; on
// large graphs the visits touch one neighbourhood at a time, which halves
// the time of a refinement pass over visits scattered across the graph.
struct Visit {
    static constexpr int kBlock = 512;
    std::vector<int> perm, blocks, order;
    explicit Visit(int n) : perm(n), blocks((n + kBlock - 1) / kBlock), order(n)
    {
        std::iota(perm.begin(), perm.end(), 0);
        std::iota(blocks.begin(), blocks.end(), 0);
    }
    const std::vector<int>& next(Rng& rng)
    {
        const int n = (int)perm.size();
        for (size_t i = blocks.size(); i > 1; --i) std::swap(blocks[i - 1], blocks[rng.next() % i]);
        int k = 0;
        for (int b : blocks) {
            const int s = b * kBlock, m = std::min(n, s + kBlock) - s;
            for (int i = 0; i < m; ++i) {
                const int a = (int)(rng.next() % m), c = (int)(rng.next() % m);
                std::swap(perm[s + a], perm[s + c]);
            }
            for (int i = 0; i < m; ++i) order[k++] = perm[s + i];
        }
        return order;
    }
};

struct MLGraph {
    int n = 0;
    std::vector<int> xadj, adj, vwgt, cmap;
    std::vector<double> w, vvol, vsurf, wsum;
};

double aratio(int N, double s, double v) { return N == 2 ? s * s / v : s * std::sqrt(s) / v; }
double match_key(int N, double s, double v) { return N == 2 ? (s * s) * (s * s) : s * s * s / (v * v); }

// one round of aspect-ratio matching; returns the contracted graph (n equal
// to g.n when nothing was paired)
MLGraph match_contract(MLGraph& g, int N, int maxsize, Rng& rng)
{
    const int n = g.n;
    Visit vis(n);
    const std::vector<int>& perm = vis.next(rng);
    struct Cand { double key; int i, k; };
    std::vector<Cand> cand;
    cand.reserve(g.adj.size() / 2);
    for (int i : perm)
        for (int t = g.xadj[i]; t < g.xadj[i + 1]; ++t) {
            const int k = g.adj[t];
            if (k > i || g.vwgt[i] + g.vwgt[k] > maxsize) continue;
            cand.push_back({match_key(N, g.vsurf[i] + g.vsurf[k] + g.wsum[i] + g.wsum[k] - 2.0 * g.w[t], g.vvol[i] + g.vvol[k]),
                            i, k});
        }
    std::stable_sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) { return a.key < b.key; });
    std::vector<int> mate(n, -1);
    g.cmap.assign(n, -1);
    std::vector<int> first;   // first member of each coarse vertex
    first.reserve(n);
    int pairs = 0;
    for (const Cand& c : cand) {
        if (pairs > n / 4) break;
        if (mate[c.i] >= 0 || mate[c.k] >= 0) continue;
        mate[c.i] = c.k; mate[c.k] = c.i;
        g.cmap[c.i] = g.cmap[c.k] = (int)first.size();
        first.push_back(c.i);
        ++pairs;
    }
    for (int i = 0; i < n; ++i)
        if (mate[i] < 0) { mate[i] = i; g.cmap[i] = (int)first.size(); first.push_back(i); }
    MLGraph c;
    c.n = (int)first.size();
    if (c.n == n) return c;
    c.xadj.assign(c.n + 1, 0);
    c.vwgt.resize(c.n); c.vvol.resize(c.n); c.vsurf.resize(c.n); c.wsum.resize(c.n);
    c.adj.reserve(g.adj.size()); c.w.reserve(g.adj.size());
    std::vector<int> slot(c.n, -1);
    for (int cv = 0; cv < c.n; ++cv) {
        const int a = first[cv], b = mate[a];
        const int m0 = (int)c.adj.size();
        c.vwgt[cv] = g.vwgt[a]; c.vvol[cv] = g.vvol[a]; c.vsurf[cv] = g.vsurf[a]; c.wsum[cv] = g.wsum[a];
        if (b != a) { c.vwgt[cv] += g.vwgt[b]; c.vvol[cv] += g.vvol[b]; c.vsurf[cv] += g.vsurf[b]; c.wsum[cv] += g.wsum[b]; }
        for (int v : {a, b}) {
            for (int t = g.xadj[v]; t < g.xadj[v + 1]; ++t) {
                const int u = g.cmap[g.adj[t]];
                if (u == cv) { c.wsum[cv] -= g.w[t]; continue; }   // the pair's own face (seen from both sides)
                if (slot[u] >= 0) { c.w[slot[u]] += g.w[t]; continue; }
                slot[u] = (int)c.adj.size();
                c.adj.push_back(u); c.w.push_back(g.w[t]);
            }
            if (b == a) break;
        }
        for (int t = m0; t < (int)c.adj.size(); ++t) slot[c.adj[t]] = -1;
        c.xadj[cv + 1] = (int)c.adj.size();
    }
    return c;
}

// clusters of one graph level
struct Parts {
    int np = 0;
    std::vector<int> where, pw;
    std::vector<double> pvol, psurf;

    void params(const MLGraph& g)
    {
        pw.assign(np, 0); pvol.assign(np, 0.0); psurf.assign(np, 0.0);
        for (int i = 0; i < g.n; ++i) {
            const int me = where[i];
            pw[me] += g.vwgt[i]; pvol[me] += g.vvol[i]; psurf[me] += g.vsurf[i];
            for (int t = g.xadj[i]; t < g.xadj[i + 1]; ++t)
                if (where[g.adj[t]] != me) psurf[me] += g.w[t];
        }
    }
    // renumber the non-empty clusters 0.. in the order of their old ids
    void compact(int n)
    {
        std::vector<int> id(np, -1);
        for (int i = 0; i < n; ++i) id[where[i]] = 0;
        int k = 0;
        for (int p = 0; p < np; ++p) if (id[p] == 0) id[p] = k++;
        for (int i = 0; i < n; ++i) where[i] = id[where[i]];
        np = k;
    }
};

// the cluster with the largest aspect ratio
void find_max(const Parts& P, int N, int& pmax, double& maxar)
{
    pmax = -1; maxar = 0.0;
    for (int p = 0; p < P.np; ++p) {
        if (P.pw[p] == 0) continue;
        const double a = aratio(N, P.psurf[p], P.pvol[p]);
        if (pmax < 0 || a > maxar) { maxar = a; pmax = p; }
    }
}

// multi-objective k-way refinement: single-vertex moves to adjacent clusters
void refine(const MLGraph& g, Parts& P, int N, int minsize, int maxsize, int npasses, Rng& rng)
{
    int pmax; double maxar;
    find_max(P, N, pmax, maxar);
    Visit vis(g.n);
    std::vector<int> slot(P.np, -1), tgt;
    std::vector<double> deg;
    for (int pass = 0; pass < npasses; ++pass) {
        int moves = 0;
        for (int i : vis.next(rng)) {
            const int from = P.where[i];
            if (P.pw[from] - g.vwgt[i] < minsize) continue;
            double id = 0.0, ed = 0.0;
            tgt.clear(); deg.clear();
            for (int t = g.xadj[i]; t < g.xadj[i + 1]; ++t) {
                const int to = P.where[g.adj[t]];
                if (to == from) { id += g.w[t]; continue; }
                ed += g.w[t];
                if (P.pw[to] + g.vwgt[i] > maxsize) continue;
                if (slot[to] < 0) { slot[to] = (int)tgt.size(); tgt.push_back(to); deg.push_back(g.w[t]); }
                else deg[slot[to]] += g.w[t];
            }
            int j1 = -1, j2 = -1;
            double best1 = 0.01, best2 = 0.1;
            const double oldFrom = aratio(N, P.psurf[from], P.pvol[from]);
            const double newFrom = aratio(N, P.psurf[from] + id - ed - g.vsurf[i], P.pvol[from] - g.vvol[i]);
            for (int j = 0; j < (int)tgt.size(); ++j) {
                const int to = tgt[j];
                const double oldTo = aratio(N, P.psurf[to], P.pvol[to]);
                const double newTo = aratio(N, P.psurf[to] + ed + id - 2.0 * deg[j] + g.vsurf[i], P.pvol[to] + g.vvol[i]);
                if (newFrom > maxar || newTo > maxar) continue;      // never raise the largest ratio
                if (to == pmax || from == pmax) { j1 = j; break; }   // a move that touches the worst cluster
                const double gain1 = std::max(oldFrom, oldTo) - std::max(newFrom, newTo);
                if (gain1 > best1) { best1 = gain1; j1 = j; }
                const double gain2 = oldFrom * P.pw[from] + oldTo * P.pw[to] -
                                     (newFrom * (P.pw[from] - g.vwgt[i]) + newTo * (P.pw[to] + g.vwgt[i]));
                if (gain2 > best2) { best2 = gain2; j2 = j; }
            }
            const int jb = j1 >= 0 ? j1 : j2;
            if (jb >= 0) {
                const int to = tgt[jb];
                P.where[i] = to;
                P.pw[to] += g.vwgt[i]; P.pw[from] -= g.vwgt[i];
                P.pvol[to] += g.vvol[i]; P.pvol[from] -= g.vvol[i];
                P.psurf[from] += id - ed - g.vsurf[i];
                P.psurf[to] += id + ed - 2.0 * deg[jb] + g.vsurf[i];
                if (from == pmax || to == pmax) find_max(P, N, pmax, maxar);
                ++moves;
            }
            for (int to : tgt) slot[to] = -1;
        }
        if (moves == 0) break;
    }
}

// split every cluster into its connected components (only when some
// cluster is disconnected); returns true if it changed the clusters
bool break_components(const MLGraph& g, Parts& P)
{
    std::vector<int> comp(g.n, -1), stack;
    int nc = 0;
    for (int s = 0; s < g.n; ++s) {
        if (comp[s] >= 0) continue;
        const int me = P.where[s];
        comp[s] = nc;
        stack.assign(1, s);
        while (!stack.empty()) {
            const int u = stack.back(); stack.pop_back();
            for (int t = g.xadj[u]; t < g.xadj[u + 1]; ++t) {
                const int v = g.adj[t];
                if (comp[v] < 0 && P.where[v] == me) { comp[v] = nc; stack.push_back(v); }
            }
        }
        ++nc;
    }
    if (nc <= P.np) return false;
    P.where.swap(comp);
    P.np = nc;
    return true;
}

// merge every cluster lighter than minsize into the adjacent cluster that
// keeps the largest aspect ratio lowest or, where several stay under the
// current largest, lowers the weighted sum most; returns the merges made
int merge_small(const MLGraph& g, Parts& P, int N, int minsize, int maxsize)
{
    P.params(g);
    std::vector<int> small;
    for (int p = 0; p < P.np; ++p) if (P.pw[p] < minsize) small.push_back(p);
    if (small.empty()) return 0;
    std::vector<std::vector<int>> mem(P.np);
    for (int i = 0; i < g.n; ++i) mem[P.where[i]].push_back(i);
    std::stable_sort(small.begin(), small.end(), [&](int a, int b) { return P.pw[a] > P.pw[b]; });   // heaviest first
    int pmax; double maxar;
    find_max(P, N, pmax, maxar);
    std::vector<int> slot(P.np, -1), tgt;
    std::vector<double> deg;
    int merges = 0;
    for (int me : small) {
        if (P.pw[me] == 0 || P.pw[me] >= minsize) continue;
        tgt.clear(); deg.clear();
        for (int i : mem[me])
            for (int t = g.xadj[i]; t < g.xadj[i + 1]; ++t) {
                const int to = P.where[g.adj[t]];
                if (to == me || P.pw[to] == 0 || P.pw[to] + P.pw[me] > maxsize) continue;
                if (slot[to] < 0) { slot[to] = (int)tgt.size(); tgt.push_back(to); deg.push_back(g.w[t]); }
                else deg[slot[to]] += g.w[t];
            }
        if (!tgt.empty()) {
            const double oldMe = aratio(N, P.psurf[me], P.pvol[me]) * P.pw[me];
            int j1 = -1, j2 = -1;
            double best1 = 0.0, best2 = 0.0;
            for (int j = 0; j < (int)tgt.size(); ++j) {
                const int to = tgt[j];
                const double nw = aratio(N, P.psurf[me] + P.psurf[to] - 2.0 * deg[j], P.pvol[me] + P.pvol[to]);
                if (j1 < 0 || nw < best1) { best1 = nw; j1 = j; }
                if (nw <= maxar) {
                    const double gain = oldMe + aratio(N, P.psurf[to], P.pvol[to]) * P.pw[to] - nw * (P.pw[me] + P.pw[to]);
                    if (j2 < 0 || gain > best2) { best2 = gain; j2 = j; }
                }
            }
            const int jb = j2 >= 0 ? j2 : j1;
            const int to = tgt[jb];
            for (int i : mem[me]) P.where[i] = to;
            mem[to].insert(mem[to].end(), mem[me].begin(), mem[me].end());
            mem[me].clear();
            P.pw[to] += P.pw[me]; P.pw[me] = 0;
            P.pvol[to] += P.pvol[me]; P.pvol[me] = 0.0;
            P.psurf[to] += P.psurf[me] - 2.0 * deg[jb]; P.psurf[me] = 0.0;
            find_max(P, N, pmax, maxar);
            ++merges;
        }
        for (int to : tgt) slot[to] = -1;
    }
    P.compact(g.n);
    return merges;
}

// clusters lighter than minsize attract single vertices from adjacent
// clusters that can spare them (by the same two objectives); returns -1 when
// no cluster is under minsize
int contribute(const MLGraph& g, Parts& P, int N, int minsize, int maxsize)
{
    P.params(g);
    std::vector<int> regular;
    bool any_small = false;
    for (int p = 0; p < P.np; ++p) {
        if (P.pw[p] > minsize) regular.push_back(p);
        else if (P.pw[p] < minsize) any_small = true;
    }
    if (!any_small) return -1;
    std::stable_sort(regular.begin(), regular.end(), [&](int a, int b) { return P.pw[a] > P.pw[b]; });
    std::vector<std::vector<int>> mem(P.np);
    for (int i = 0; i < g.n; ++i) mem[P.where[i]].push_back(i);
    int pmax; double maxar;
    find_max(P, N, pmax, maxar);
    std::vector<int> slot(P.np, -1), tgt;
    std::vector<double> deg;
    int moves = 0;
    for (int me : regular)
        for (int i : mem[me]) {
            if (P.where[i] != me || P.pw[me] - g.vwgt[i] < minsize) continue;
            double id = 0.0, ed = 0.0;
            tgt.clear(); deg.clear();
            for (int t = g.xadj[i]; t < g.xadj[i + 1]; ++t) {
                const int to = P.where[g.adj[t]];
                if (to == me) { id += g.w[t]; continue; }
                ed += g.w[t];
                if (P.pw[to] >= minsize || P.pw[to] + g.vwgt[i] > maxsize) continue;
                if (slot[to] < 0) { slot[to] = (int)tgt.size(); tgt.push_back(to); deg.push_back(g.w[t]); }
                else deg[slot[to]] += g.w[t];
            }
            if (!tgt.empty()) {
                const double oldMe = aratio(N, P.psurf[me], P.pvol[me]) * P.pw[me];
                const double newMe = aratio(N, P.psurf[me] + id - ed - g.vsurf[i], P.pvol[me] - g.vvol[i]);
                int j1 = -1, j2 = -1;
                double best1 = 0.0, best2 = 0.0;
                for (int j = 0; j < (int)tgt.size(); ++j) {
                    const int to = tgt[j];
                    const double newTo = aratio(N, P.psurf[to] + ed + id - 2.0 * deg[j] + g.vsurf[i], P.pvol[to] + g.vvol[i]);
                    const double mx = std::max(newMe, newTo);
                    if (j1 < 0 || mx < best1) { best1 = mx; j1 = j; }
                    if (mx <= maxar) {
                        const double gain = oldMe + aratio(N, P.psurf[to], P.pvol[to]) * P.pw[to] -
                                            (newMe * (P.pw[me] - g.vwgt[i]) + newTo * (P.pw[to] + g.vwgt[i]));
                        if (j2 < 0 || gain > best2) { best2 = gain; j2 = j; }
                    }
                }
                const int jb = j2 >= 0 ? j2 : j1;
                const int to = tgt[jb];
                P.where[i] = to;
                P.pw[to] += g.vwgt[i]; P.pw[me] -= g.vwgt[i];
                P.pvol[to] += g.vvol[i]; P.pvol[me] -= g.vvol[i];
                P.psurf[me] += id - ed - g.vsurf[i];
                P.psurf[to] += id + ed - 2.0 * deg[jb] + g.vsurf[i];
                find_max(P, N, pmax, maxar);
                ++moves;
            }
            for (int to : tgt) slot[to] = -1;
        }
    return moves;
}

// one multilevel run from the random orders of `seed`: the clusters of g0 in P
void partition_once(const MLGraph& g0, int N, int minsize, int maxsize, uint64_t seed, Parts& P)
{
    const int npasses = 10;
    Rng rng{seed};
    std::vector<MLGraph> G(1, g0);
    // 1. coarsening
    for (;;) {
        MLGraph c = match_contract(G.back(), N, maxsize, rng);
        if (c.n == G.back().n) break;
        G.push_back(std::move(c));
    }
    // 2. initial clusters: the coarsest vertices
    P.np = G.back().n;
    P.where.resize(P.np);
    std::iota(P.where.begin(), P.where.end(), 0);
    P.params(G.back());
    // 3. uncoarsening with refinement on every level
    for (int l = (int)G.size() - 1;; --l) {
        refine(G[l], P, N, minsize, maxsize, npasses, rng);
        if (l == 0) break;
        std::vector<int> fw(G[l - 1].n);
        for (int i = 0; i < G[l - 1].n; ++i) fw[i] = P.where[G[l - 1].cmap[i]];
        P.where.swap(fw);
        break_components(G[l - 1], P);
        merge_small(G[l - 1], P, N, minsize, maxsize);
        P.params(G[l - 1]);
    }
    // 4. final split, merge / attract cycle, refinement
    const MLGraph& g = G[0];
    break_components(g, P);
    merge_small(g, P, N, minsize, maxsize);
    P.params(g);
    refine(g, P, N, minsize, maxsize, npasses, rng);
    for (int pass = 0; pass < npasses; ++pass) {
        merge_small(g, P, N, minsize, maxsize);
        if (contribute(g, P, N, minsize, maxsize) < 0) break;
    }
    P.params(g);
    refine(g, P, N, minsize, maxsize, npasses, rng);
    break_components(g, P);
    P.compact(g.n);
}

}  // namespace

int mgridgen_partition(int nv, const int* xadj, const double* vvol, const double* vsurf, const int* adjncy,
                       const double* adjwgt, int N, int minsize, int maxsize, std::vector<int>& part, uint64_t seed)
{
    if (nv <= 0) { part.clear(); return 0; }
    if (N != 2 && N != 3) throw Error(-1, "agglomeration: dimension must be 2 or 3");
    if (minsize < 1 || maxsize < minsize) throw Error(-1, "agglomeration: need 1 <= minsize <= maxsize");
    MLGraph g0;
    g0.n = nv;
    g0.xadj.assign(xadj, xadj + nv + 1);
    if (g0.xadj[0] != 0) throw Error(-1, "agglomeration: xadj[0] must be 0");
    for (int i = 0; i < nv; ++i)
        if (g0.xadj[i + 1] < g0.xadj[i]) throw Error(-1, "agglomeration: xadj must be non-decreasing");
    const int ne = g0.xadj[nv];
    g0.adj.assign(adjncy, adjncy + ne);
    g0.w.assign(adjwgt, adjwgt + ne);
    g0.vvol.assign(vvol, vvol + nv);
    g0.vsurf.assign(vsurf, vsurf + nv);
    g0.vwgt.assign(nv, 1);
    g0.wsum.assign(nv, 0.0);
    for (int i = 0; i < nv; ++i) {
        if (!(vvol[i] > 0)) throw Error(-1, "agglomeration: vertex volumes must be positive");
        for (int t = g0.xadj[i]; t < g0.xadj[i + 1]; ++t) {
            if (g0.adj[t] < 0 || g0.adj[t] >= nv || g0.adj[t] == i)
                throw Error(-1, "agglomeration: adjacency entry out of range or a self loop");
            g0.wsum[i] += g0.w[t];
        }
    }
    Parts P;
    partition_once(g0, N, minsize, maxsize, seed, P);
    part = std::move(P.where);
    return P.np;
}

}  // namespace mgldg
