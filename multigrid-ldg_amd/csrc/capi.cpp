// capi.cpp — the extern "C" boundary of libmgldg (declared in include/mgldg.h).
// No exception crosses it: every entry point maps errors to a status code and
// records the message on the context.
#include <cstring>
#include <exception>
#include <fstream>
#include <new>
#include <sstream>

#include "../../include/mgldg.h"
#include "io.hpp"
#include "kernels.hpp"
#include "mg.hpp"

using namespace mgldg;

struct mgldg_ctx_s { Ctx c; };
struct mgldg_quadtree_s { Ctx* ctx; Quadtree qt; };
struct mgldg_problem_s { Ctx* ctx; ProblemH p; };
struct mgldg_mg_s { Ctx* ctx; MGH m; };
struct mgldg_wireframe_s { Ctx* ctx; int N; std::vector<double> p; std::vector<int> t; };

static thread_local std::string g_err;

template <typename F>
static int guard(Ctx* ctx, F&& f)
{
    try {
        if (ctx) HIPCHK(hipSetDevice(ctx->device));
        f();
        return MGLDG_OK;
    } catch (const Error& e) {
        (ctx ? ctx->err : g_err) = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        (ctx ? ctx->err : g_err) = "host allocation failed";
        return MGLDG_ERR_NOMEM;
    } catch (const std::exception& e) {
        (ctx ? ctx->err : g_err) = e.what();
        return MGLDG_ERR_HIP;
    }
}

#define ARG(cond, msg) do { if (!(cond)) throw Error(MGLDG_ERR_ARG, msg); } while (0)

extern "C" {

int mgldg_ctx_create(int device, void* stream, mgldg_ctx* out)
{
    if (!out) return MGLDG_ERR_ARG;
    *out = nullptr;
    mgldg_ctx h = new (std::nothrow) mgldg_ctx_s();
    if (!h) return MGLDG_ERR_NOMEM;
    int st = guard(nullptr, [&] {
        int n = 0;
        HIPCHK(hipGetDeviceCount(&n));
        ARG(device >= 0 && device < n, "no such device");
        h->c.init(device, stream);
    });
    if (st != MGLDG_OK) { delete h; return st; }
    *out = h;
    return MGLDG_OK;
}

int mgldg_ctx_destroy(mgldg_ctx ctx) { delete ctx; return MGLDG_OK; }

const char* mgldg_last_error(mgldg_ctx ctx) { return ctx ? ctx->c.err.c_str() : g_err.c_str(); }

int mgldg_ctx_synchronize(mgldg_ctx ctx)
{
    if (!ctx) return MGLDG_ERR_ARG;
    return guard(&ctx->c, [&] { HIPCHK(hipStreamSynchronize(ctx->c.stream)); });
}

int mgldg_quadtree_create(mgldg_ctx ctx, int N, const double* lo, const double* hi, mgldg_sizefn h, void* user,
                          int strategy, int periodic, int balance, mgldg_quadtree* out)
{
    if (!ctx || !out) return MGLDG_ERR_ARG;
    *out = nullptr;
    return guard(&ctx->c, [&] {
        ARG(N == 2 || N == 3, "dimension must be 2 or 3");
        ARG(h != nullptr, "size function required");
        ARG(strategy == MGLDG_EQUAL_COARSENING || strategy == MGLDG_RAPID_COARSENING, "unknown coarsening strategy");
        if (lo && hi) for (int d = 0; d < N; ++d) ARG(hi[d] > lo[d], "empty domain");
        auto* q = new mgldg_quadtree_s();
        q->ctx = &ctx->c;
        q->qt.N = N;
        // the reference's domain constructor ignores the strategy (quadtree.h:221-228)
        q->qt.rapid = (strategy == MGLDG_RAPID_COARSENING) && !(lo || hi);
        q->qt.periodic = periodic != 0;
        try {
            q->qt.build(lo, hi, h, user);
            if (balance) q->qt.balance();
        } catch (...) { delete q; throw; }
        *out = q;
    });
}

int mgldg_quadtree_info(mgldg_quadtree qt, int* info)
{
    if (!qt || !info) return MGLDG_ERR_ARG;
    return guard(qt->ctx, [&] {
        info[0] = (int)qt->qt.nodes.size();
        info[1] = qt->qt.num_levels;
        info[2] = (int)qt->qt.layer(0).size();
    });
}

int mgldg_quadtree_layer(mgldg_quadtree qt, int l, int* ids, int* count)
{
    if (!qt || !count) return MGLDG_ERR_ARG;
    return guard(qt->ctx, [&] {
        ARG(l >= 0 && l < qt->qt.num_levels, "layer out of range");
        std::vector<int> v = qt->qt.layer(l);
        *count = (int)v.size();
        if (ids) std::memcpy(ids, v.data(), v.size() * sizeof(int));
    });
}

int mgldg_quadtree_destroy(mgldg_quadtree qt) { delete qt; return MGLDG_OK; }

int mgldg_wireframe_create(mgldg_ctx ctx, int N, int np, const double* points, int nt, const int* simplices,
                           mgldg_wireframe* out)
{
    if (!ctx || !out) return MGLDG_ERR_ARG;
    *out = nullptr;
    return guard(&ctx->c, [&] {
        ARG(N == 2 || N == 3, "dimension must be 2 or 3");
        ARG(np >= N + 1 && nt >= 1 && points && simplices, "empty wireframe");
        auto* w = new mgldg_wireframe_s();
        w->ctx = &ctx->c; w->N = N;
        w->p.assign(points, points + (size_t)np * N);
        w->t.assign(simplices, simplices + (size_t)nt * (N + 1));
        for (int v : w->t) if (v < 0 || v >= np) { delete w; throw Error(MGLDG_ERR_ARG, "vertex id out of range"); }
        *out = w;
    });
}

// Wireframe<N>(file) (wireframe.h:22-68): header "np nt", one skipped line,
// point lines up to an empty line, simplex lines (1-based) up to an empty line
int mgldg_wireframe_read(mgldg_ctx ctx, int N, const char* path, mgldg_wireframe* out)
{
    if (!ctx || !out || !path) return MGLDG_ERR_ARG;
    *out = nullptr;
    return guard(&ctx->c, [&] {
        ARG(N == 2 || N == 3, "dimension must be 2 or 3");
        std::ifstream in(path);
        ARG(in.is_open(), "cannot open mesh file");
        std::string line;
        int np = 0, nt = 0;
        ARG((bool)std::getline(in, line), "empty mesh file");
        { std::istringstream ss(line); ARG((bool)(ss >> np >> nt), "bad mesh header"); }
        ARG(np >= N + 1 && nt >= 1, "bad mesh header");
        std::getline(in, line);
        std::vector<double> p;
        std::vector<int> t;
        p.reserve((size_t)np * N); t.reserve((size_t)nt * (N + 1));
        while (std::getline(in, line) && !line.empty()) {
            std::istringstream ss(line);
            double v; int k = 0;
            while (ss >> v) { p.push_back(v); ++k; }
            ARG(k == N, "point line with the wrong number of coordinates");
        }
        while (std::getline(in, line) && !line.empty()) {
            std::istringstream ss(line);
            int v, k = 0;
            while (ss >> v) { t.push_back(v - 1); ++k; }
            ARG(k == N + 1, "simplex line with the wrong number of vertices");
        }
        ARG((long)p.size() == (long)np * N && (long)t.size() == (long)nt * (N + 1), "mesh file counts do not match its header");
        for (int v : t) ARG(v >= 0 && v < np, "vertex id out of range");
        auto* w = new mgldg_wireframe_s();
        w->ctx = &ctx->c; w->N = N; w->p = std::move(p); w->t = std::move(t);
        *out = w;
    });
}

int mgldg_wireframe_get(mgldg_wireframe wf, int* info, double* points, int* simplices)
{
    if (!wf) return MGLDG_ERR_ARG;
    const int np = (int)(wf->p.size() / wf->N), nt = (int)(wf->t.size() / (wf->N + 1));
    if (info) { info[0] = wf->N; info[1] = np; info[2] = nt; }
    if (points) std::memcpy(points, wf->p.data(), wf->p.size() * sizeof(double));
    if (simplices) std::memcpy(simplices, wf->t.data(), wf->t.size() * sizeof(int));
    return MGLDG_OK;
}

int mgldg_wireframe_destroy(mgldg_wireframe wf) { delete wf; return MGLDG_OK; }

int mgldg_ldg_create_simplex(mgldg_wireframe wf, int p, int nbnd, const int* vmask, const int* bctype, double tau0,
                             double tauD, mgldg_problem* out)
{
    if (!wf || !out) return MGLDG_ERR_ARG;
    *out = nullptr;
    return guard(wf->ctx, [&] {
        ARG(p >= 0 && p <= 5, "Invalid polynomial order.");
        ARG(nbnd >= 0 && nbnd <= 30 && (nbnd == 0 || vmask), "invalid boundary predicates");
        const int N = wf->N, np = (int)(wf->p.size() / N), nt = (int)(wf->t.size() / (N + 1));
        auto* h = new mgldg_problem_s();
        h->ctx = wf->ctx;
        try { h->p.build_simplex(wf->ctx, N, np, wf->p.data(), nt, wf->t.data(), nbnd, vmask, p, bctype, tau0, tauD); }
        catch (...) { delete h; throw; }
        *out = h;
    });
}

int mgldg_ldg_create(mgldg_quadtree qt, int p, const int* bctype, double tau0, double tauD, mgldg_problem* out)
{
    if (!qt || !out) return MGLDG_ERR_ARG;
    *out = nullptr;
    return guard(qt->ctx, [&] {
        ARG(p >= 0 && p <= 5, "Invalid polynomial order.");   // dg_tests.cpp:384-385
        auto* h = new mgldg_problem_s();
        h->ctx = qt->ctx;
        try { h->p.build(qt->ctx, &qt->qt, p, bctype, tau0, tauD); }
        catch (...) { delete h; throw; }
        *out = h;
    });
}

int mgldg_problem_sizes(mgldg_problem pr, int* s)
{
    if (!pr || !s) return MGLDG_ERR_ARG;
    const ProblemH& p = pr->p;
    s[0] = p.ne; s[1] = p.host->nf; s[2] = p.host->nbf; s[3] = p.b; s[4] = p.nqb; s[5] = (int)p.ndof; s[6] = p.N;
    return MGLDG_OK;
}

int mgldg_problem_nodes(mgldg_problem pr, double* xyz)
{
    if (!pr || !xyz) return MGLDG_ERR_ARG;
    return guard(pr->ctx, [&] { pr->p.host->nodes_xyz(xyz); });
}

int mgldg_problem_bc_qpoints(mgldg_problem pr, double* xyz, int* tag)
{
    if (!pr || !xyz) return MGLDG_ERR_ARG;
    return guard(pr->ctx, [&] { pr->p.host->bc_qpoints(xyz, tag); });
}

int mgldg_rhs(mgldg_problem pr, const double* f, const double* g, int nrhs, double* b)
{
    if (!pr || !b || nrhs < 1) return MGLDG_ERR_ARG;
    return guard(pr->ctx, [&] {
        // the device path (ProblemH::rhs_device) on staged copies
        ProblemH& p = pr->p;
        hipStream_t st = pr->ctx->stream;
        DBuf<double> fd, gd, bd;
        if (f) fd.upload(f, (size_t)p.ndof * nrhs, st);
        if (g && p.nqb > 0) gd.upload(g, (size_t)p.nqb * nrhs, st);
        bd.alloc((size_t)p.ndof * nrhs);
        p.rhs_device(f ? fd.p : nullptr, (g && p.nqb > 0) ? gd.p : nullptr, nrhs, bd);
        bd.download(b, (size_t)p.ndof * nrhs, st);
        HIPCHK(hipStreamSynchronize(st));
    });
}

static void export_op(Ctx* ctx, const DevOp& op, int* rowptr, int* cols, double* blocks, long* nnzb)
{
    if (nnzb) *nnzb = op.nnzb();
    if (rowptr) std::memcpy(rowptr, op.pat.rowptr.data(), op.pat.rowptr.size() * sizeof(int));
    if (cols) std::memcpy(cols, op.pat.cols.data(), op.pat.cols.size() * sizeof(int));
    if (blocks) {
        op.val.download(blocks, (size_t)op.nnzb() * op.pat.br * op.pat.bc, ctx->stream);
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
}

int mgldg_problem_export(mgldg_problem pr, int which, int* rowptr, int* cols, double* blocks, long* nnzb)
{
    if (!pr) return MGLDG_ERR_ARG;
    return guard(pr->ctx, [&] {
        ProblemH& p = pr->p;
        const DevOp* op = which == 0 ? &p.A : which == 1 ? &p.M : which == 2 ? &p.T
                        : (which >= 3 && which - 3 < p.N) ? &p.G[which - 3] : nullptr;
        ARG(op, "unknown operator");
        export_op(pr->ctx, *op, rowptr, cols, blocks, nnzb);
    });
}

int mgldg_problem_destroy(mgldg_problem pr) { delete pr; return MGLDG_OK; }

static MGParams mg_params(const mgldg_mg_params* prm)
{
    MGParams m;
    m.relaxation = prm->relaxation; m.omega = prm->omega; m.coarse = prm->coarse;
    m.npre = prm->npre; m.npost = prm->npost; m.restriction = prm->restriction; m.precision = prm->precision;
    m.coarse_dense_max = prm->coarse_dense_max;
    m.pool_off = prm->pool_off != 0;
    m.row_tiles = prm->row_tiles;
    ARG(m.row_tiles >= 0 && m.row_tiles <= 2, "row_tiles must be 0, 1 or 2");
    ARG(m.coarse_dense_max >= 0 && m.coarse_dense_max <= 8192, "coarse_dense_max must be in [0, 8192]");
    return m;
}

static void mg_build(mgldg_problem pr, const MGParams& m, const std::vector<std::vector<int>>* parts, mgldg_mg* out,
                     const std::vector<int>* pdeg = nullptr, bool hcoarsen = true)
{
    auto* h = new mgldg_mg_s();
    h->ctx = pr->ctx;
    try { h->m.build(&pr->p, m, parts, pdeg, hcoarsen); }
    catch (...) { delete h; throw; }
    *out = h;
}

int mgldg_mg_create_p(mgldg_problem pr, const mgldg_mg_params* prm, int ndeg, const int* degrees, int hcoarsen,
                      mgldg_mg* out)
{
    if (!pr || !prm || !out || ndeg < 1 || !degrees) return MGLDG_ERR_ARG;
    *out = nullptr;
    return guard(pr->ctx, [&] {
        MGParams m = mg_params(prm);
        ARG(!pr->p.simplex(), "p-multigrid hierarchies apply to quadtree problems only");
        std::vector<int> deg(degrees, degrees + ndeg);
        mg_build(pr, m, nullptr, out, &deg, hcoarsen != 0);
    });
}

int mgldg_mg_create(mgldg_problem pr, const mgldg_mg_params* prm, mgldg_mg* out)
{
    if (!pr || !prm || !out) return MGLDG_ERR_ARG;
    *out = nullptr;
    return guard(pr->ctx, [&] {
        MGParams m = mg_params(prm);
        if (const SimplexLdg* sx = pr->p.simplex()) {
            std::vector<std::vector<int>> parts = agglomerate(sx->mesh, 10, 1 << sx->N);
            ARG(!parts.empty(), "the mesh cannot be agglomerated");
            mg_build(pr, m, &parts, out);
        } else {
            mg_build(pr, m, nullptr, out);
        }
    });
}

int mgldg_agglomerate(mgldg_problem pr, int max_levels, int* nlevels, int* nparts, int* parts)
{
    if (!pr || !nlevels || max_levels < 1) return MGLDG_ERR_ARG;
    return guard(pr->ctx, [&] {
        const SimplexLdg* sx = pr->p.simplex();
        ARG(sx, "agglomeration applies to simplex problems");
        std::vector<std::vector<int>> q = agglomerate(sx->mesh, max_levels, 1 << sx->N);
        *nlevels = (int)q.size();
        for (size_t l = 0; l < q.size(); ++l) {
            int nc = 0;
            for (int v : q[l]) nc = std::max(nc, v + 1);
            if (nparts) nparts[l] = nc;
            if (parts) std::memcpy(parts + l * (size_t)sx->ne, q[l].data(), q[l].size() * sizeof(int));
        }
    });
}

int mgldg_mgridgen(int nvtxs, const int* xadj, const double* vvol, const double* vsurf, const int* adjncy,
                   const double* adjwgt, int minsize, int maxsize, const int* options, int* nmoves, int* nparts, int* part)
{
    if (nvtxs < 0 || !xadj || !vvol || !vsurf || !options || !nparts || (nvtxs > 0 && !part)) return MGLDG_ERR_ARG;
    return guard(nullptr, [&] {
        if (options[0] != 4 || options[1] != 6)
            throw Error(MGLDG_ERR_UNSUPPORTED, "mgridgen: only CType 4 (aspect-ratio matching) with RType 6 "
                                               "(multi-objective refinement), the reference's options, are built");
        ARG(options[3] == 2 || options[3] == 3, "mgridgen: options[3] (dim) must be 2 or 3");
        ARG(nvtxs == 0 || xadj[nvtxs] == 0 || (adjncy && adjwgt), "mgridgen: adjacency arrays are NULL");
        std::vector<int> q;
        *nparts = mgridgen_partition(nvtxs, xadj, vvol, vsurf, adjncy, adjwgt, options[3], minsize, maxsize, q);
        if (nvtxs > 0) std::memcpy(part, q.data(), q.size() * sizeof(int));
        if (nmoves) *nmoves = 0;
    });
}

int mgldg_mg_create_agglomerated(mgldg_problem pr, const mgldg_mg_params* prm, int nlev, const int* parts, mgldg_mg* out)
{
    if (!pr || !prm || !out || !parts || nlev < 1) return MGLDG_ERR_ARG;
    *out = nullptr;
    return guard(pr->ctx, [&] {
        MGParams m = mg_params(prm);
        const SimplexLdg* sx = pr->p.simplex();
        ARG(sx, "agglomerated multigrid applies to simplex problems");
        std::vector<std::vector<int>> q(nlev);
        int n = sx->ne;
        for (int l = 0; l < nlev; ++l) {
            const int* src = parts + (size_t)l * sx->ne;
            q[l].assign(src, src + n);
            int nc = 0;
            for (int v : q[l]) { ARG(v >= 0 && v < n, "agglomerate id out of range"); nc = std::max(nc, v + 1); }
            n = nc;
        }
        mg_build(pr, m, &q, out);
    });
}

int mgldg_mg_info(mgldg_mg mg, long* info, int ninfo)
{
    if (!mg || !info || ninfo < 3) return MGLDG_ERR_ARG;
    const MGH& m = mg->m;
    info[0] = (long)m.L.size(); info[1] = m.pr->ndof; info[2] = m.pr->b;
    for (size_t l = 0; l < m.L.size() && 4 + 2 * (int)l < ninfo; ++l) { info[3 + 2 * l] = m.L[l].nb; info[4 + 2 * l] = m.L[l].A.nnzb(); }
    return MGLDG_OK;
}

int mgldg_mg_level(mgldg_mg mg, int level, int* nb, int* B, long* nnzb)
{
    if (!mg || level < 0 || level >= (int)mg->m.L.size()) return MGLDG_ERR_ARG;
    const Level& lv = mg->m.L[level];
    if (nb) *nb = lv.nb;
    if (B) *B = lv.B;
    if (nnzb) *nnzb = lv.A.nnzb();
    return MGLDG_OK;
}

int mgldg_mg_batch_kernels(mgldg_mg mg, int mode)
{
    if (!mg || !(mode == -1 || mode == 0 || mode == 1 || mode == 2 || mode == 4)) return MGLDG_ERR_ARG;
    mg->m.bforce = mode;
    mg->m.bk = 0;          // re-plan at the next batched call
    return MGLDG_OK;
}

int mgldg_mg_batch_plan(mgldg_mg mg, int* plan, float* ms)
{
    if (!mg || !plan) return MGLDG_ERR_ARG;
    const MGH& m = mg->m;
    const int nl = (int)m.L.size();
    for (int p = 0; p < 2; ++p)
        for (int l = 0; l < nl; ++l) {
            const bool has = (int)m.btune[p].size() > l;
            plan[(p * nl + l) * 2] = has ? m.btune[p][l].apply : 0;
            plan[(p * nl + l) * 2 + 1] = has ? m.btune[p][l].smooth : 0;
            if (ms)
                for (int c = 0; c < 4; ++c) {
                    ms[((p * nl + l) * 2) * 4 + c] = has ? m.btune[p][l].apply_ms[c] : 0.f;
                    ms[((p * nl + l) * 2 + 1) * 4 + c] = has ? m.btune[p][l].smooth_ms[c] : 0.f;
                }
        }
    return MGLDG_OK;
}

int mgldg_mg_export(mgldg_mg mg, int level, int which, int* rowptr, int* cols, double* blocks, long* nnzb,
                    int* nbrows, int* nbcols)
{
    if (!mg) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        MGH& m = mg->m;
        ARG(level >= 0 && level < (int)m.L.size(), "level out of range");
        Level& lv = m.L[level];
        if (which == 6) {
            ARG(level + 1 < (int)m.L.size(), "coarsest level has no interpolation");
            const Interp& I = lv.I;
            if (nnzb) *nnzb = I.nf;
            if (nbrows) *nbrows = I.nf;
            if (nbcols) *nbcols = I.nc;
            if (rowptr) for (int i = 0; i <= I.nf; ++i) rowptr[i] = i;
            if (cols) for (int i = 0; i < I.nf; ++i) cols[i] = I.parent[i];
            if (blocks) {
                const size_t bs = (size_t)I.bf * I.bc;
                for (int i = 0; i < I.nf; ++i) std::memcpy(blocks + i * bs, I.pool.data() + (size_t)I.type[i] * bs, bs * sizeof(double));
            }
            return;
        }
        const DevOp* op = which == 0 ? &lv.A : nullptr;
        if (which >= 1 && which <= 5) {
            ARG(lv.has_ops, "operator only kept on RATRAT levels");
            op = which == 1 ? &lv.M : which == 2 ? &lv.Tp : (which - 3 < m.pr->N) ? &lv.G[which - 3] : nullptr;
        }
        ARG(op, "unknown operator");
        if (nbrows) *nbrows = op->pat.nbr;
        if (nbcols) *nbcols = op->pat.nbc;
        export_op(mg->ctx, *op, rowptr, cols, blocks, nnzb);
    });
}

int mgldg_mg_level_pool(mgldg_mg mg, int level, int* npool, int* ndpool)
{
    if (!mg) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(level >= 0 && level < (int)mg->m.L.size(), "level out of range");
        if (npool) *npool = mg->m.L[level].npool;
        if (ndpool) *ndpool = mg->m.L[level].ndpool;
    });
}

int mgldg_mg_solve_cost(mgldg_mg mg, int precision, int flexible, int* vc_launches, double* vc_bytes,
                        int* it_launches, double* it_bytes, int* graph_nodes)
{
    if (!mg) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(precision == 0 || precision == 1, "unknown precision");
        if (precision == 1) ARG(mg->m.L[0].A32.p, "mixed precision not built");
        int n = 0, ni = 0;
        double b = 0, bi = 0;
        mg->m.vcycle_cost(precision, &n, &b);
        mg->m.iteration_cost(precision, flexible != 0, &ni, &bi);
        if (vc_launches) *vc_launches = n;
        if (vc_bytes) *vc_bytes = b;
        if (it_launches) *it_launches = n + ni;
        if (it_bytes) *it_bytes = b + bi;
        if (graph_nodes) *graph_nodes = mg->m.graph_nodes[precision];
    });
}

int mgldg_mg_level_tiles(mgldg_mg mg, int level, int* ntiles, int* width, int* used)
{
    if (!mg) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(level >= 0 && level < (int)mg->m.L.size(), "level out of range");
        const Level& lv = mg->m.L[level];
        if (ntiles) *ntiles = lv.a_ntiles;
        if (width) *width = lv.t_W;
        if (used) *used = (lv.use_tiles[0] != 0 ? 1 : 0) | (lv.use_tiles[1] != 0 ? 2 : 0);
    });
}

int mgldg_mg_colors(mgldg_mg mg, int level, int* colors, int* ncolors)
{
    if (!mg) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(level >= 0 && level < (int)mg->m.L.size(), "level out of range");
        const Level& lv = mg->m.L[level];
        if (ncolors) *ncolors = lv.ncolor;
        if (colors) std::memcpy(colors, lv.color.data(), lv.color.size() * sizeof(int));
    });
}

int mgldg_mg_destroy(mgldg_mg mg) { delete mg; return MGLDG_OK; }

static SolveParams to_sp(const mgldg_solve_params* p)
{
    SolveParams s;
    if (p) { s.tol = p->tol; s.maxit = p->maxit; s.criterion = p->criterion; s.flexible = p->flexible; s.precision = p->precision; s.use_graph = p->use_graph; }
    if (s.criterion != 0 && s.criterion != 1) throw Error(MGLDG_ERR_ARG, "unknown stopping criterion");
    if (s.precision != 0 && s.precision != 1) throw Error(MGLDG_ERR_ARG, "unknown precision");
    if (s.maxit < 0 || !(s.tol >= 0)) throw Error(MGLDG_ERR_ARG, "invalid tolerance or iteration limit");
    return s;
}

static void fill_stats(mgldg_solve_stats* o, const SolveStats& s)
{
    o->iterations = s.iterations; o->status = s.status; o->relres_true = s.relres_true;
    o->relres_prec = s.relres_prec; o->seconds = s.seconds;
}

int mgldg_solve_device(mgldg_mg mg, const double* b, double* x, int nrhs, const mgldg_solve_params* prm,
                       mgldg_solve_stats* stats)
{
    if (!mg || !b || !x || nrhs < 1) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        SolveParams sp = to_sp(prm);
        MGH& m = mg->m;
        const long n = m.pr->ndof;
        hipStream_t s = mg->ctx->stream;
        if (nrhs == 1) {
            SolveStats st = m.solve(b, x, sp);
            if (stats) fill_stats(stats, st);
            return;
        }
        // all columns of the [ndof][nrhs] layout in one batched solve
        ARG(nrhs <= 256, "at most 256 right-hand sides per call");
        std::vector<SolveStats> st = m.solve_batched(b, x, nrhs, sp);
        if (stats) for (int c = 0; c < nrhs; ++c) fill_stats(&stats[c], st[c]);
        (void)n; (void)s;
    });
}

int mgldg_solve(mgldg_mg mg, const double* b, double* x, int nrhs, const mgldg_solve_params* prm, mgldg_solve_stats* stats)
{
    if (!mg || !b || !x || nrhs < 1) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        const size_t n = (size_t)mg->m.pr->ndof * nrhs;
        hipStream_t s = mg->ctx->stream;
        DBuf<double> db, dx;
        db.upload(b, n, s); dx.upload(x, n, s);
        int st = mgldg_solve_device(mg, db, dx, nrhs, prm, stats);
        if (st != MGLDG_OK) throw Error(st, mg->ctx->err);
        dx.download(x, n, s);
        HIPCHK(hipStreamSynchronize(s));
    });
}

int mgldg_adjoint_device(mgldg_mg mg, const double* dLdu, double* lambda, double* dLdf, double* dLdg, int nrhs,
                         const mgldg_solve_params* prm, mgldg_solve_stats* stats)
{
    if (!mg || !dLdu || nrhs < 1) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        SolveParams sp = to_sp(prm);
        ARG(nrhs <= 256, "at most 256 right-hand sides per call");
        std::vector<SolveStats> st = mg->m.adjoint(dLdu, lambda, dLdf, dLdg, nrhs, sp);
        if (stats) for (int c = 0; c < nrhs; ++c) fill_stats(&stats[c], st[c]);
    });
}

int mgldg_adjoint_maps_device(mgldg_mg mg, const double* lambda, double* dLdf, double* dLdg, int nrhs)
{
    if (!mg || !lambda || nrhs < 1) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] { mg->m.pr->adjoint_maps(lambda, nrhs, dLdf, dLdg); });
}

int mgldg_rhs_device(mgldg_mg mg, const double* f, const double* g, int nrhs, double* b)
{
    if (!mg || !b || nrhs < 1) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] { mg->m.pr->rhs_device(f, g, nrhs, b); });
}

int mgldg_spmv_device(mgldg_mg mg, int level, int precision, const void* x, void* y, int nrhs)
{
    if (!mg || !x || !y || nrhs != 1) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(level >= 0 && level < (int)mg->m.L.size(), "level out of range");
        ARG(precision == 0 || precision == 1, "unknown precision");
        if (precision == 1) ARG(mg->m.L[level].A32.p, "mixed precision not built");
        mg->m.spmv(level, precision, x, y);
        HIPCHK(hipStreamSynchronize(mg->ctx->stream));
    });
}

int mgldg_smooth_device(mgldg_mg mg, int level, int precision, int nsweeps, int reverse, const void* b, void* x, int nrhs)
{
    if (!mg || !b || !x || nrhs != 1 || nsweeps < 0) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(level >= 0 && level < (int)mg->m.L.size(), "level out of range");
        ARG(precision == 0 || precision == 1, "unknown precision");
        if (precision == 1) ARG(mg->m.L[level].A32.p, "mixed precision not built");
        mg->m.smooth(level, precision, nsweeps, reverse != 0, b, x);
        HIPCHK(hipStreamSynchronize(mg->ctx->stream));
    });
}

int mgldg_vcycle_device(mgldg_mg mg, int precision, const double* b, double* x, int nrhs)
{
    if (!mg || !b || !x || nrhs != 1) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(precision == 0 || precision == 1, "unknown precision");
        mg->m.vcycle(precision, b, x);
        HIPCHK(hipStreamSynchronize(mg->ctx->stream));
    });
}

int mgldg_vcycles_device(mgldg_mg mg, int precision, const double* b, double* x, int ncycle, double tol, double* rnorm,
                         int* cycles)
{
    if (!mg || !b || !x || ncycle < 0) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(precision == 0 || precision == 1, "unknown precision");
        ARG(tol >= 0, "tol must be >= 0");
        mg->m.vcycles(precision, b, x, ncycle, tol, rnorm, cycles);
        HIPCHK(hipStreamSynchronize(mg->ctx->stream));
    });
}

int mgldg_mean_zero_device(mgldg_ctx ctx, long n, double* v, int nrhs)
{
    if (!ctx || !v || n <= 0 || nrhs < 1) return MGLDG_ERR_ARG;
    Ctx* c = &ctx->c;
    return guard(c, [&] {
        ARG(nrhs <= 256, "at most 256 right-hand sides per call");
        const size_t need = dev::mean_zero_part_size(n, nrhs);
        if (c->colpart.n < need) c->colpart.alloc(need);
        dev::mean_zero(c->stream, n, nrhs, v, c->colpart.p);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int mgldg_mean_zero_device_noalloc(mgldg_ctx ctx, long n, double* v, int nrhs)
{
    if (!ctx || !v || n <= 0 || nrhs < 1) return MGLDG_ERR_ARG;
    Ctx* c = &ctx->c;
    return guard(c, [&] {
        ARG(nrhs <= 256, "mean_zero_noalloc: at most 256 right-hand sides per call (the static scratch's limit)");
        ARG(dev::mean_zero_part_size(n, nrhs) <= dev::kScratchDoubles, "mean_zero_noalloc: scratch too small");
        ARG(c->scratch, "mean_zero_noalloc: scratch address not resolved");
        dev::mean_zero(c->stream, n, nrhs, v, c->scratch);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

// ---------------------------------------------------------------- interop file formats
int mgldg_bsr_write_mm(const char* path, int nbrows, int nbcols, int br, int bc, const int* rowptr, const int* cols,
                       const double* blocks)
{
    if (!path) return MGLDG_ERR_ARG;
    return guard(nullptr, [&] { io::write_mm(path, nbrows, nbcols, br, bc, rowptr, cols, blocks); });
}

static void write_op_mm(Ctx* ctx, const DevOp& op, const char* path)
{
    std::vector<double> blk((size_t)op.nnzb() * op.pat.br * op.pat.bc);
    export_op(ctx, op, nullptr, nullptr, blk.data(), nullptr);
    io::write_mm(path, op.pat.nbr, op.pat.nbc, op.pat.br, op.pat.bc, op.pat.rowptr.data(), op.pat.cols.data(),
                 blk.data());
}

static const DevOp* problem_op(ProblemH& p, int which)
{
    return which == 0 ? &p.A : which == 1 ? &p.M : which == 2 ? &p.T
         : (which >= 3 && which - 3 < p.N) ? &p.G[which - 3] : nullptr;
}

int mgldg_problem_write_mm(mgldg_problem pr, int which, const char* path)
{
    if (!pr || !path) return MGLDG_ERR_ARG;
    return guard(pr->ctx, [&] {
        const DevOp* op = problem_op(pr->p, which);
        ARG(op, "unknown operator");
        write_op_mm(pr->ctx, *op, path);
    });
}

int mgldg_mg_write_mm(mgldg_mg mg, int level, int which, const char* path)
{
    if (!mg || !path) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        MGH& m = mg->m;
        ARG(level >= 0 && level < (int)m.L.size(), "level out of range");
        Level& lv = m.L[level];
        if (which == 6) {
            ARG(level + 1 < (int)m.L.size(), "coarsest level has no interpolation");
            const Interp& I = lv.I;
            const size_t bs = (size_t)I.bf * I.bc;
            std::vector<int> rp(I.nf + 1), cols(I.parent.begin(), I.parent.begin() + I.nf);
            std::vector<double> blk((size_t)I.nf * bs);
            for (int i = 0; i <= I.nf; ++i) rp[i] = i;
            for (int i = 0; i < I.nf; ++i) std::memcpy(blk.data() + i * bs, I.pool.data() + (size_t)I.type[i] * bs, bs * sizeof(double));
            io::write_mm(path, I.nf, I.nc, I.bf, I.bc, rp.data(), cols.data(), blk.data());
            return;
        }
        const DevOp* op = which == 0 ? &lv.A : nullptr;
        if (which >= 1 && which <= 5) {
            ARG(lv.has_ops, "operator only kept on RATRAT levels");
            op = which == 1 ? &lv.M : which == 2 ? &lv.Tp : (which - 3 < m.pr->N) ? &lv.G[which - 3] : nullptr;
        }
        ARG(op, "unknown operator");
        write_op_mm(mg->ctx, *op, path);
    });
}

int mgldg_fun_write(const char* path, int P, int N, int ne, int npl, const double* xyz, const double* u)
{
    if (!path) return MGLDG_ERR_ARG;
    return guard(nullptr, [&] { io::write_fun(path, P, N, ne, npl, xyz, u); });
}

static void problem_fun(ProblemH& p, const double* u, const std::string& path)
{
    std::vector<double> xyz((size_t)p.ndof * p.N);
    p.host->nodes_xyz(xyz.data());
    io::write_fun(path, p.host->p + 1, p.N, p.ne, p.b, xyz.data(), u);
}

int mgldg_problem_write_fun(mgldg_problem pr, const double* u, const char* path)
{
    if (!pr || !u || !path) return MGLDG_ERR_ARG;
    return guard(pr->ctx, [&] { problem_fun(pr->p, u, path); });
}

int mgldg_problem_dump(mgldg_problem pr, const double* rhs, const char* dir)
{
    if (!pr || !dir) return MGLDG_ERR_ARG;
    return guard(pr->ctx, [&] {
        ProblemH& p = pr->p;
        const std::string d = std::string(dir) + "/";
        write_op_mm(pr->ctx, p.A, (d + "A.mtx").c_str());
        write_op_mm(pr->ctx, p.M, (d + "M.mtx").c_str());
        write_op_mm(pr->ctx, p.T, (d + "T.mtx").c_str());
        for (int i = 0; i < p.N; ++i) write_op_mm(pr->ctx, p.G[i], (d + "G" + std::to_string(i) + ".mtx").c_str());
        if (rhs) problem_fun(p, rhs, d + "rhs.fun");
    });
}

int mgldg_mesh_write(const char* path, int N, int np, const double* points, int nt, const int* simplices, int digits)
{
    if (!path) return MGLDG_ERR_ARG;
    return guard(nullptr, [&] { io::write_mesh(path, N, np, points, nt, simplices, digits); });
}

int mgldg_wireframe_write(mgldg_wireframe wf, const char* path, int digits)
{
    if (!wf || !path) return MGLDG_ERR_ARG;
    return guard(wf->ctx, [&] {
        const int np = (int)(wf->p.size() / wf->N), nt = (int)(wf->t.size() / (wf->N + 1));
        io::write_mesh(path, wf->N, np, wf->p.data(), nt, wf->t.data(), digits);
    });
}

int mgldg_mg_vcycle_matrix(mgldg_mg mg, int precision, double* V)
{
    if (!mg || !V) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(precision == 0 || precision == 1, "unknown precision");
        mg->m.vcycle_matrix(precision, V);
    });
}

int mgldg_dense_write(const char* path, int rows, int cols, const double* a)
{
    if (!path) return MGLDG_ERR_ARG;
    return guard(nullptr, [&] { io::write_dense(path, rows, cols, a); });
}

int mgldg_profile_relax(mgldg_mg mg, int level, int precision, int nrep, double* ms, double* bytes, int* launches)
{
    if (!mg || nrep < 1) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(level >= 0 && level < (int)mg->m.L.size(), "level out of range");
        ARG(precision == 0 || precision == 1, "unknown precision");
        if (precision == 1) ARG(mg->m.L[level].A32.p, "mixed precision not built");
        mg->m.profile_relax(level, precision, nrep, ms, bytes, launches);
    });
}

int mgldg_profile_relax_part(mgldg_mg mg, int level, int precision, int nrep, int part, double* ms, double* bytes,
                             int* launches)
{
    if (!mg || nrep < 1) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(level >= 0 && level < (int)mg->m.L.size(), "level out of range");
        ARG(precision == 0 || precision == 1, "unknown precision");
        ARG(part >= 0 && part <= 2, "part must be 0, 1 or 2");
        if (precision == 1) ARG(mg->m.L[level].A32.p, "mixed precision not built");
        ARG(part == 0 || mg->m.prm.relaxation == 1, "parts exist for multicolour Gauss-Seidel only");
        ARG(part != 2 || mg->m.L[level].pair_nwg > 0, "level has no paired launch");
        mg->m.profile_relax(level, precision, nrep, ms, bytes, launches, part);
    });
}

int mgldg_profile_spmv(mgldg_mg mg, int level, int precision, int nrep, double* ms, double* bytes)
{
    if (!mg || nrep < 1) return MGLDG_ERR_ARG;
    return guard(mg->ctx, [&] {
        ARG(level >= 0 && level < (int)mg->m.L.size(), "level out of range");
        ARG(precision == 0 || precision == 1, "unknown precision");
        if (precision == 1) ARG(mg->m.L[level].A32.p, "mixed precision not built");
        mg->m.profile_spmv(level, precision, nrep, ms, bytes);
    });
}

int mgldg_problem_setup_times(mgldg_problem pr, double* t)
{
    if (!pr || !t) return MGLDG_ERR_ARG;
    const ProblemH& p = pr->p;
    t[0] = p.setup_host_s; t[1] = p.setup_dev_s; t[2] = p.asm_ms; t[3] = p.asm_bytes;
    return MGLDG_OK;
}

int mgldg_setup_times(mgldg_mg mg, double* t)
{
    if (!mg || !t) return MGLDG_ERR_ARG;
    t[0] = mg->m.pr->setup_host_s; t[1] = mg->m.pr->setup_dev_s; t[2] = mg->m.setup_host_s; t[3] = mg->m.setup_dev_s;
    return MGLDG_OK;
}

}  // extern "C"
