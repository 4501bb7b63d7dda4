// mg.cpp — device-resident problem setup, GPU Galerkin hierarchy, V-cycle and
// MG-preconditioned CG.  Reference structure:
//   LDGPoisson / LDGOperators       src/cartesian/dg/ldgpoisson.{h,tpp}
//   InterpolationHierarchy          src/cartesian/multigrid/interpolationhierarchy.h:21-96
//   Level / Multigrid               src/cartesian/multigrid/multigrid.h:65-433
//   pcg                             src/cartesian/multigrid/krylov.h:22-68
#include "mg.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <unordered_map>

#include "kernels.hpp"

namespace mgldg {

using Clock = std::chrono::steady_clock;
static double secs(Clock::time_point a) { return std::chrono::duration<double>(Clock::now() - a).count(); }

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Error(-2, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
void DBuf<T>::alloc(size_t count)
{
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) { p = nullptr; throw Error(-3, std::string("hipMalloc failed: ") + hipGetErrorString(e)); }
    n = count;
}
template <typename T>
void DBuf<T>::upload(const T* h, size_t count, hipStream_t s)
{
    if (count > n) alloc(count);
    if (count) HIPCHK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
}
template <typename T>
void DBuf<T>::download(T* h, size_t count, hipStream_t s) const
{
    if (count) HIPCHK(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
}
template <typename T>
void DBuf<T>::release()
{
    if (p) (void)hipFree(p);
    p = nullptr; n = 0;
}
template struct DBuf<double>;
template struct DBuf<float>;
template struct DBuf<int>;
template struct DBuf<unsigned>;
template struct DBuf<long>;

void Ctx::init(int dev, void* s)
{
    device = dev;
    HIPCHK(hipSetDevice(dev));
    if (s) { stream = (hipStream_t)s; own_stream = false; }
    else { HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)); own_stream = true; }
    scratch = dev::scratch_address(dev);   // with dev current
    part.alloc(3 * 1024 + 64);
    counter.alloc(16);
    HIPCHK(hipMemsetAsync(counter.p, 0, 16 * sizeof(unsigned), stream));
    HIPCHK(hipStreamSynchronize(stream));
}

Ctx::~Ctx()
{
    if (stream) (void)hipStreamSynchronize(stream);
    part.release(); counter.release();
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}

void DevOp::upload_pattern(const HostBSR& h, hipStream_t s)
{
    pat.nbr = h.nbr; pat.nbc = h.nbc; pat.br = h.br; pat.bc = h.bc;
    pat.rowptr = h.rowptr; pat.cols = h.cols; pat.val.clear();
    rowptr.upload(h.rowptr.data(), h.rowptr.size(), s);
    cols.upload(h.cols.data(), std::max<size_t>(h.cols.size(), 0), s);
}

void DevOp::build_ell(hipStream_t s)
{
    const int n = pat.nbr;
    ellw = 1;
    for (int i = 0; i < n; ++i) ellw = std::max(ellw, pat.rowptr[i + 1] - pat.rowptr[i]);
    std::vector<int> e((size_t)n * ellw, -1);
    for (int i = 0; i < n; ++i)
        for (int k = pat.rowptr[i]; k < pat.rowptr[i + 1]; ++k) e[(size_t)i * ellw + (k - pat.rowptr[i])] = pat.cols[k];
    ell.upload(e.data(), e.size(), s);
    HIPCHK(hipStreamSynchronize(s));
}

// GS row tables of the rows `rows` (in that order): descriptors {row,
// rowptr[row], rowptr[row+1], diagonal position} and off-diagonal columns at
// stride w (-1 past the last)
// pidx / didx (pooled operator): per block slot its A pool index, per row its
// D^-1 pool index; the table then holds (column, pool index) pairs and
// desc.w the D^-1 pool index
// colour (pooled tables): the order of the sweep the tables serve (colours
// or wavefronts); a pair's pool index then carries dev::kPidLower when the
// column's colour precedes the row's, so that the first sweep from zero
// knows which neighbours are already written without reading colours
static bool pool_blocks(hipStream_t s, long nblk, int bs, const double* blocks, std::vector<int>& pidx,
                        std::vector<int>& rep);

// One-column products with a problem operator (adjoint maps): its distinct
// blocks in the row-padded pool layout and (column, pool index) pairs at the
// ELL stride, read by the pooled single-RHS operator kernel; false (nothing
// built) when the blocks do not repeat.
static bool pool_operator(hipStream_t s, DevOp& op, int b, DBuf<double>& pool, DBuf<int>& pairs)
{
    std::vector<int> pidx, rep;
    if (!pool_blocks(s, op.nnzb(), b * b, op.val, pidx, rep)) return false;
    const long np = (long)rep.size();
    DBuf<int> rd; rd.upload(rep.data(), rep.size(), s);
    DBuf<double> rm; rm.alloc((size_t)np * b * b);
    dev::gather_blocks(s, (int)np, b * b, rd, op.val, rm);
    pool.alloc((size_t)np * b * dev::pool_pad(8, b));
    dev::to_rowpad_f64(s, np, b, rm, pool);
    const int n = op.pat.nbr, W = op.ellw;
    std::vector<int> pe((size_t)n * W * 2, 0);
    for (int i = 0; i < n; ++i) {
        int m = 0;
        for (int k = op.pat.rowptr[i]; k < op.pat.rowptr[i + 1]; ++k, ++m) {
            pe[((size_t)i * W + m) * 2] = op.pat.cols[k];
            pe[((size_t)i * W + m) * 2 + 1] = pidx[k];
        }
        for (; m < W; ++m) pe[((size_t)i * W + m) * 2] = -1;
    }
    pairs.upload(pe.data(), pe.size(), s);
    HIPCHK(hipStreamSynchronize(s));
    return true;
}

static void gs_tables(const HostBSR& P, const std::vector<int>& rows, int w, hipStream_t s, DBuf<int>& desc_d,
                      DBuf<int>& gcols_d, const std::vector<int>* pidx = nullptr, const std::vector<int>* didx = nullptr,
                      const std::vector<int>* colour = nullptr)
{
    const size_t nr = rows.size();
    const int ew = pidx ? 2 : 1;
    std::vector<int> desc(4 * nr), gc(nr * w * ew, -1);
    if (pidx) for (size_t t = 1; t < gc.size(); t += 2) gc[t] = 0;
    for (size_t q = 0; q < nr; ++q) {
        const int i = rows[q];
        int dp = -1, m = 0;
        for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
            if (P.cols[k] == i) { if (dp < 0) { dp = k; continue; } }
            if (pidx) {
                const int pv = (*pidx)[k];
                if (pv < 0 || pv > dev::kPidMask) throw Error(-5, "pool index out of range");
                const bool lower = colour && (*colour)[P.cols[k]] < (*colour)[i];
                gc[(q * w + m) * 2] = P.cols[k]; gc[(q * w + m) * 2 + 1] = pv | (lower ? dev::kPidLower : 0); ++m;
            }
            else gc[q * w + m++] = P.cols[k];
        }
        desc[4 * q] = i; desc[4 * q + 1] = P.rowptr[i]; desc[4 * q + 2] = P.rowptr[i + 1];
        desc[4 * q + 3] = didx ? (*didx)[i] : dp;
    }
    desc_d.upload(desc.data(), desc.size(), s);
    gcols_d.upload(gc.data(), gc.size(), s);
    HIPCHK(hipStreamSynchronize(s));
}

// upload a term list and launch the triple-product kernel
static void run_triple(hipStream_t s, const TermList& t, int p, int q, int m, int n, int transL, const double* Lp,
                       const double* Xp, const double* Rp, const double* addpool, const std::vector<int>* addslot,
                       double* out)
{
    int nout = (int)t.off.size() - 1;
    if (nout <= 0) return;
    DBuf<int> off, l, x, r, as;
    off.upload(t.off.data(), t.off.size(), s);
    if (!t.l.empty()) { l.upload(t.l.data(), t.l.size(), s); x.upload(t.x.data(), t.x.size(), s); r.upload(t.r.data(), t.r.size(), s); }
    if (addslot) as.upload(addslot->data(), addslot->size(), s);
    dev::triple(s, nout, p, q, m, n, transL, off, l, x, r, Lp, Xp, Rp, addpool, addslot ? as.p : nullptr, out);
    HIPCHK(hipStreamSynchronize(s));   // term buffers are freed on return
}

// A = T + sum_d G_d^T M G_d on the GPU (construct_laplacian, ldgpoisson.h:34-52)
static void build_laplacian(hipStream_t s, int N, int b, DevOp& T, DevOp* G, DevOp& M, DevOp& A)
{
    HostBSR Gp[3];
    for (int d = 0; d < N; ++d) Gp[d] = G[d].pat;
    LaplacianSymbolic sym;
    laplacian_symbolic(T.pat, Gp, N, M.pat, sym);
    A.upload_pattern(sym.pattern, s);
    A.val.alloc((size_t)A.nnzb() * b * b);
    // concatenate the G pools so one term list addresses all dimensions
    long gtot = 0;
    for (int d = 0; d < N; ++d) gtot += G[d].nnzb();
    DBuf<double> gcat;
    gcat.alloc((size_t)gtot * b * b);
    long off = 0;
    for (int d = 0; d < N; ++d) {
        HIPCHK(hipMemcpyAsync(gcat.p + off * b * b, G[d].val.p, (size_t)G[d].nnzb() * b * b * sizeof(double), hipMemcpyDeviceToDevice, s));
        off += G[d].nnzb();
    }
    run_triple(s, sym.gmg, b, b, b, b, 1, gcat, M.val, gcat, T.val, &sym.tslot, A.val);
}

void ProblemH::build(Ctx* c, const Quadtree* q, int p, const int* bct, double tau0, double tauD)
{
    ctx = c; qt = q;
    auto t0 = Clock::now();
    auto h = std::make_unique<LdgHost>();
    h->assemble(*q, p, bct, tau0, tauD);
    host = std::move(h);
    setup_host_s = secs(t0);
    upload();
}

void ProblemH::build_simplex(Ctx* c, int N_, int np, const double* pts, int nt, const int* tris, int nbnd,
                             const int* vmask, int p, const int* bct, double tau0, double tauD)
{
    ctx = c; qt = nullptr;
    auto t0 = Clock::now();
    auto h = std::make_unique<SimplexLdg>();
    h->assemble(N_, np, pts, nt, tris, nbnd, vmask, p, bct, tau0, tauD);
    host = std::move(h);
    setup_host_s = secs(t0);
    upload();
}

// LDGPoisson::discretize, device half: the patterns go up, the plan's
// tables and face lists go up, and every block of M, T and G_d, the boundary
// lifts and A = T + sum_d G_d^T M G_d (construct_laplacian,
// ldgpoisson.tpp:178-190) are formed on the device.
void ProblemH::upload()
{
    LdgOps& H = *host;
    const AsmPlan& P = H.plan;
    N = H.N; b = H.npl; ne = H.ne; ndof = (long)ne * b; nqb = H.nbf * H.nq;
    hipStream_t s = ctx->stream;
    auto t1 = Clock::now();
    const size_t bb = (size_t)b * b;
    M.upload_pattern(H.M, s); T.upload_pattern(H.T, s);
    for (int d = 0; d < N; ++d) G[d].upload_pattern(H.G[d], s);
    M.val.alloc(std::max<size_t>((size_t)M.nnzb() * bb, 1));
    T.val.alloc(std::max<size_t>((size_t)T.nnzb() * bb, 1));
    for (int d = 0; d < N; ++d) G[d].val.alloc(std::max<size_t>((size_t)G[d].nnzb() * bb, 1));
    // (at least one element each, so that empty lists -- a mesh without
    // boundary faces, a periodic quadtree -- still pass valid pointers)
    auto up1 = [&](auto& D, const auto& v) {
        if (D.n < std::max<size_t>(v.size(), 1)) D.alloc(std::max<size_t>(v.size(), 1));
        D.upload(v.data(), v.size(), s);
    };
    DBuf<double> mscale, iscale, C, fscale, fnormal, emass, einv, ediff, fmass, pool, qpts, qwt, lnodes, ldenom, lbref, bscale;
    DBuf<int> ef_off, ef, ef_tslot, ef_gslot, ef_first, tdiag, gdiag, fkind, fca, fcb, slice, islice, bcode;
    up1(mscale, P.mscale); up1(iscale, P.iscale); up1(C, P.C); up1(fscale, P.fscale); up1(fnormal, P.fnormal);
    up1(emass, P.emass); up1(einv, P.einv); up1(ediff, P.ediff); up1(fmass, P.fmass);
    up1(ef_off, P.ef_off); up1(ef, P.ef); up1(ef_tslot, P.ef_tslot); up1(ef_gslot, P.ef_gslot); up1(ef_first, P.ef_first);
    up1(tdiag, P.tdiag); up1(gdiag, P.gdiag); up1(fkind, P.fkind); up1(fca, P.fca); up1(fcb, P.fcb);
    up1(slice, P.slice); up1(islice, P.islice);
    const int nqf = (int)P.qface.size();
    pool.alloc(((size_t)P.npool_ref + 3 * (size_t)nqf) * bb + 1);
    pool.upload(P.pool.data(), P.pool.size(), s);
    if (nqf) {
        up1(qpts, P.qpts); up1(qwt, P.qwt); up1(lnodes, P.lnodes); up1(ldenom, P.ldenom);
        dev::asm_quad(s, nqf, N, (int)P.lnodes.size(), P.nq, b, qpts, qwt, lnodes, ldenom, pool.p + (size_t)P.npool_ref * bb);
    }
    dev::AsmRows a;
    a.ne = ne; a.N = N; a.b = b; a.nfl = P.nfl; a.gshared = P.gshared; a.tau0 = H.tau0; a.tauD = H.tauD;
    a.mscale = mscale; a.iscale = iscale; a.C = C;
    a.ef_off = ef_off; a.ef = ef; a.ef_tslot = ef_tslot; a.ef_gslot = ef_gslot; a.ef_first = ef_first;
    a.tdiag = tdiag; a.gdiag = gdiag; a.fkind = fkind; a.fca = fca; a.fcb = fcb;
    a.fscale = fscale; a.fnormal = fnormal;
    a.emass = emass; a.einv = einv; a.ediff = ediff; a.fmass = fmass; a.pool = pool;
    a.slice = slice; a.islice = islice;
    a.M = M.val; a.T = T.val; a.G0 = G[0].val; a.G1 = N > 1 ? G[1].val.p : nullptr; a.G2 = N > 2 ? G[2].val.p : nullptr;
    hipEvent_t ea, eb;
    HIPCHK(hipEventCreate(&ea)); HIPCHK(hipEventCreate(&eb));
    HIPCHK(hipEventRecord(ea, s));
    dev::asm_rows(s, a);
    HIPCHK(hipEventRecord(eb, s));
    // algorithmic bytes: every block written once, the per-element, per-entry
    // and per-face plan data read once (reference tables are cache-resident)
    {
        long blocks = M.nnzb() + T.nnzb();
        for (int d = 0; d < N; ++d) blocks += G[d].nnzb();
        asm_bytes = (double)blocks * bb * 8
                  + (double)ne * (8.0 * (2 + N * N) + 4.0 * (2 + N))
                  + (double)P.ef.size() * 16.0
                  + (double)P.fkind.size() * (12.0 + 8.0 + 8.0 * N);
    }
    // boundary data for the device RHS / adjoint
    const int nbf = H.nbf;
    up1(lbref, P.lbref); up1(bscale, P.bscale); up1(bcode, P.bcode);
    Lb.alloc(std::max<size_t>((size_t)nbf * b * H.nq, 1));
    dev::asm_lift(s, nbf, b, H.nq, bcode, bscale, lbref, Lb);
    up1(bleft, H.bleft);
    up1(btype, H.btype);
    up1(bnormal, H.bnormal);
    std::vector<int> off(ne + 1, 0), lst(std::max(nbf, 1));
    for (int k = 0; k < nbf; ++k) off[H.bleft[k] + 1]++;
    for (int e = 0; e < ne; ++e) off[e + 1] += off[e];
    {
        std::vector<int> pos(off.begin(), off.end() - 1);
        for (int k = 0; k < nbf; ++k) lst[pos[H.bleft[k]]++] = k;
    }
    bf_off.upload(off.data(), off.size(), s); bf_list.upload(lst.data(), lst.size(), s);
    build_laplacian(s, N, b, T, G, M, A);
    // M and G_d with their ELL tables for one column of the adjoint maps on
    // the single-RHS row kernels: pooled where the blocks repeat (quadtree
    // meshes), else in the panel layout
    M.build_ell(s);
    if (!pool_operator(s, M, b, Mpool, Mpell)) {
        Mpan.alloc(M.val.n);
        dev::to_panel_f64(s, M.nnzb(), b, M.val, Mpan);
    }
    // G_d (panel) and G_d^T (pattern transpose, transposed blocks, panel) for B_g and its adjoint
    for (int d = 0; d < N; ++d) {
        Gpan[d].alloc(G[d].val.n);
        dev::to_panel_f64(s, G[d].nnzb(), b, G[d].val, Gpan[d]);
        G[d].build_ell(s);
        pool_operator(s, G[d], b, Gpool[d], Gpell[d]);
        const HostBSR& Gh = H.G[d];
        HostBSR Tr;
        Tr.nbr = ne; Tr.nbc = ne; Tr.br = b; Tr.bc = b;
        Tr.rowptr.assign(ne + 1, 0);
        for (int c : Gh.cols) Tr.rowptr[c + 1]++;
        for (int i = 0; i < ne; ++i) Tr.rowptr[i + 1] += Tr.rowptr[i];
        Tr.cols.resize(Gh.cols.size());
        std::vector<int> src(std::max<size_t>(Gh.cols.size(), 1));
        {
            std::vector<int> pos(Tr.rowptr.begin(), Tr.rowptr.end() - 1);
            for (int i = 0; i < ne; ++i)                     // rows ascending: each transposed row comes out sorted
                for (int t = Gh.rowptr[i]; t < Gh.rowptr[i + 1]; ++t) {
                    const int q = pos[Gh.cols[t]]++;
                    Tr.cols[q] = i; src[q] = t;
                }
        }
        GT[d].upload_pattern(Tr, s);
        DBuf<int> srcd; srcd.upload(src.data(), src.size(), s);
        DBuf<double> rm; rm.alloc(std::max<size_t>((size_t)Tr.nnzb() * bb, 1));
        dev::transpose_blocks(s, Tr.nnzb(), b, srcd, G[d].val, rm);
        GT[d].val.alloc(std::max<size_t>((size_t)Tr.nnzb() * bb, 1));
        dev::to_panel_f64(s, Tr.nnzb(), b, rm, GT[d].val);
        HIPCHK(hipStreamSynchronize(s));
    }
    HIPCHK(hipStreamSynchronize(s));
    {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ea, eb));
        asm_ms = ms;
        (void)hipEventDestroy(ea); (void)hipEventDestroy(eb);
    }
    H.plan = AsmPlan();                  // only the device needed it
    setup_dev_s = secs(t1);
}

void ProblemH::rhs_device(const double* f, const double* g, int k, double* bout)
{
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)ndof * k;
    if (f) dev::blockdiag(s, ne, b, M.val, f, bout, k);
    else dev::Ops<double>::fill(s, (long)n, 0.0, bout);
    if (!g) return;
    DBuf<double> r, J[3], t;
    r.alloc(n); t.alloc(n);
    for (int d = 0; d < N; ++d) J[d].alloc(n);
    dev::bc_forward(s, ne, b, host->nq, k, N, bf_off, bf_list, btype, bnormal, host->tauD, Lb, g, r, J[0], J[1], J[2]);
    dev::axpby(s, (long)n, 1.0, bout, 1.0, r, bout);
    for (int d = 0; d < N; ++d) {
        dev::MOps<double>::apply(s, b, k, ne, GT[d].rowptr, GT[d].cols, GT[d].val, J[d], nullptr, t);
        dev::axpby(s, (long)n, 1.0, bout, -1.0, t, bout);
    }
    HIPCHK(hipStreamSynchronize(s));
}

void ProblemH::adjoint_maps(const double* lam, int k, double* dLdf, double* dLdg)
{
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)ndof * k;
    if (k == 1) {
        auto apply1 = [&](const DevOp& op, const DBuf<double>& pool, const DBuf<int>& pell, const double* pan, double* y) {
            if (pool.p) dev::Ops<double>::apply(s, b, ne, op.rowptr, pell, op.ellw, pool, lam, nullptr, y, nullptr, nullptr, nullptr, 1);
            else dev::Ops<double>::apply(s, b, ne, op.rowptr, op.ell, op.ellw, pan, lam, nullptr, y, nullptr, nullptr, nullptr);
        };
        if (dLdf) apply1(M, Mpool, Mpell, Mpan, dLdf);
        if (dLdg) {
            for (int d = 0; d < N; ++d) {
                if (Gl1[d].n < n) Gl1[d].alloc(n);
                apply1(G[d], Gpool[d], Gpell[d], Gpan[d], Gl1[d]);
            }
            dev::bc_adjoint(s, host->nbf, b, host->nq, 1, N, bleft, btype, bnormal, host->tauD, Lb, lam, Gl1[0],
                            N > 1 ? Gl1[1].p : nullptr, N > 2 ? Gl1[2].p : nullptr, dLdg);
        }
        HIPCHK(hipStreamSynchronize(s));
        return;
    }
    if (dLdf) dev::blockdiag(s, ne, b, M.val, lam, dLdf, k);
    if (dLdg) {
        DBuf<double> Gl[3];
        for (int d = 0; d < N; ++d) {
            Gl[d].alloc(n);
            dev::MOps<double>::apply(s, b, k, ne, G[d].rowptr, G[d].cols, Gpan[d], lam, nullptr, Gl[d]);
        }
        dev::bc_adjoint(s, host->nbf, b, host->nq, k, N, bleft, btype, bnormal, host->tauD, Lb, lam, Gl[0],
                        N > 1 ? Gl[1].p : nullptr, N > 2 ? Gl[2].p : nullptr, dLdg);
        HIPCHK(hipStreamSynchronize(s));
    }
    HIPCHK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------- hierarchy
// InterpolationHierarchy<N,P>(qt) (interpolationhierarchy.h:30-82): one block
// per fine element, identity if the element persists, else the parent's
// Lagrange basis evaluated at the child's nodes.  Distinct blocks are pooled.
static void build_interp(const Quadtree& qt, const Basis& B, int l, Interp& I)
{
    std::vector<int> fine = qt.layer(l), coarse = qt.layer(l + 1);
    std::vector<int> cmap(qt.nodes.size(), -1);
    for (size_t k = 0; k < coarse.size(); ++k) cmap[coarse[k]] = (int)k;
    const int b = B.npl, N = qt.N;
    I.nf = (int)fine.size(); I.nc = (int)coarse.size(); I.bf = b; I.bc = b;
    I.parent.resize(I.nf); I.type.resize(I.nf);
    std::map<std::vector<double>, int> seen;
    I.pool.clear(); I.ntypes = 0;
    auto add_type = [&](const std::vector<double>& key, const std::vector<double>& blk) {
        auto it = seen.find(key);
        if (it != seen.end()) return it->second;
        I.pool.insert(I.pool.end(), blk.begin(), blk.end());
        seen[key] = I.ntypes;
        return I.ntypes++;
    };
    std::vector<double> eye((size_t)b * b, 0.0), E((size_t)b * b);
    for (int i = 0; i < b; ++i) eye[(size_t)i * b + i] = 1.0;
    for (int f = 0; f < I.nf; ++f) {
        int g = fine[f];
        if (qt.is_leaf(g, l + 1)) {
            I.parent[f] = cmap[g];
            I.type[f] = add_type(std::vector<double>{-1.0}, eye);
            continue;
        }
        int c = qt.nodes[g].parent;
        I.parent[f] = cmap[c];
        std::vector<double> key(2 * N);
        for (int d = 0; d < N; ++d) {
            double w = qt.nodes[c].hi[d] - qt.nodes[c].lo[d];
            key[d] = (qt.nodes[g].lo[d] - qt.nodes[c].lo[d]) / w;
            key[N + d] = (qt.nodes[g].hi[d] - qt.nodes[c].lo[d]) / w;
        }
        if (!seen.count(key)) {
            int it[3];
            for (int i = 0; i < b; ++i) {
                int lin = i;
                for (int d = N - 1; d >= 0; --d) { it[d] = lin % B.P; lin /= B.P; }
                double x[3];
                for (int d = 0; d < N; ++d) x[d] = key[d] + (key[N + d] - key[d]) * B.nodes[it[d]];
                B.lagrangeN(x, &E[(size_t)i * b]);
            }
        }
        I.type[f] = add_type(key, E);
    }
    I.make_gather();
}

// InterpolationHierarchy<N,P1,P2,...> (interpolationhierarchy.h:98-131): on
// every element, the degree-P2 Lagrange basis evaluated at the degree-P1
// Gauss-Lobatto nodes -- one block for all ne elements (parent = itself).
static void build_pinterp(int ne, const Basis& B1, const Basis& B2, Interp& I)
{
    const int N = B1.N, b1 = B1.npl, b2 = B2.npl;
    I.nf = ne; I.nc = ne; I.bf = b1; I.bc = b2;
    I.parent.resize(ne); I.type.assign(ne, 0);
    for (int e = 0; e < ne; ++e) I.parent[e] = e;
    I.pool.assign((size_t)b1 * b2, 0.0); I.ntypes = 1;
    int it[3];
    for (int i = 0; i < b1; ++i) {
        int lin = i;
        for (int d = N - 1; d >= 0; --d) { it[d] = lin % B1.P; lin /= B1.P; }
        double x[3];
        for (int d = 0; d < N; ++d) x[d] = B1.nodes[it[d]];
        B2.lagrangeN(x, &I.pool[(size_t)i * b2]);
    }
    I.make_gather();
}

// invert a dense SPD(-semidefinite) matrix on the host by Gauss-Jordan with
// partial pivoting.  Zero pivots give zero rows, as LDLT::solve does; a pivot
// below 1e-13 of the matrix's largest entry counts as zero.  Without that, the
// rounding-level pivot of a semidefinite matrix (the coarsest operator of a
// Neumann or periodic problem, whose kernel is the constants) puts ~1e16
// entries in the inverse, and their products with a mean-zero right-hand side
// lose every digit, not only the kernel component.  With it, the result is a
// generalised inverse whose solutions differ from the LDLT solve's only in
// the kernel.
static std::vector<double> dense_inverse(std::vector<double> A, int n)
{
    std::vector<double> X((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) X[(size_t)i * n + i] = 1.0;
    double amax = 0;
    for (double v : A) amax = std::max(amax, std::fabs(v));
    const double zero_pivot = std::max(1e-13 * amax, 2.2250738585072014e-308);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i) if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)p * n + k])) p = i;
        if (p != k) for (int j = 0; j < n; ++j) { std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]); std::swap(X[(size_t)k * n + j], X[(size_t)p * n + j]); }
        double d = A[(size_t)k * n + k];
        double inv = (std::fabs(d) > zero_pivot) ? 1.0 / d : 0.0;
        for (int j = 0; j < n; ++j) { A[(size_t)k * n + j] *= inv; X[(size_t)k * n + j] *= inv; }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            double f = A[(size_t)i * n + k];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) { A[(size_t)i * n + j] -= f * A[(size_t)k * n + j]; X[(size_t)i * n + j] -= f * X[(size_t)k * n + j]; }
        }
    }
    return X;
}

static std::vector<int> diag_slots(const HostBSR& P)
{
    std::vector<int> d(P.nbr);
    for (int i = 0; i < P.nbr; ++i) {
        d[i] = P.find(i, i);
        if (d[i] < 0) throw Error(-4, "operator has no diagonal block");
    }
    return d;
}

// Level coarse constructor (multigrid.h:105-171): symbolic on the host,
// numeric block triple products on the GPU.  Interpolation blocks are bf x bc
// (bf = bc on quadtree hierarchies).
static void coarsen(hipStream_t s, int N, Level& f, Level& c, int restriction)
{
    const Interp& I = f.I;
    const int bf = f.B, bc = c.B;
    if (restriction == 1) {
        // M_c = T^T M_f T
        HostBSR pat; TermList t;
        rap_symbolic(I, f.M.pat, pat, t);
        c.M.upload_pattern(pat, s);
        c.M.val.alloc((size_t)pat.nnzb() * bc * bc);
        run_triple(s, t, bc, bc, bf, bf, 1, f.pool64, f.M.val, f.pool64, nullptr, nullptr, c.M.val);
        // inverse coarse mass blocks (the reference's Minv LDLT, multigrid.h:133-135)
        std::vector<int> ds = diag_slots(pat);
        DBuf<int> dsl; dsl.upload(ds.data(), ds.size(), s);
        c.Minv.alloc((size_t)c.nb * bc * bc);
        dev::inverse(s, c.nb, bc, c.M.val, dsl, c.Minv);
        HIPCHK(hipStreamSynchronize(s));
        // G_c = M_c^{-1} T^T (M_f G_f) T
        std::vector<int> mslot = diag_slots(f.M.pat);
        for (int d = 0; d < N; ++d) {
            DevOp& Gf = f.G[d];
            // MG = M_f G_f (block row scaling), pattern of G_f
            TermList tm; tm.off.resize(Gf.nnzb() + 1); tm.l.assign(Gf.nnzb(), -1); tm.x.resize(Gf.nnzb()); tm.r.resize(Gf.nnzb());
            for (int i = 0; i < Gf.pat.nbr; ++i)
                for (int k = Gf.pat.rowptr[i]; k < Gf.pat.rowptr[i + 1]; ++k) { tm.off[k] = k; tm.x[k] = mslot[i]; tm.r[k] = k; }
            tm.off[Gf.nnzb()] = (int)Gf.nnzb();
            DBuf<double> MG; MG.alloc((size_t)Gf.nnzb() * bf * bf);
            run_triple(s, tm, bf, bf, bf, bf, 0, nullptr, f.M.val, Gf.val, nullptr, nullptr, MG);
            HostBSR gp; TermList ty;
            rap_symbolic(I, Gf.pat, gp, ty);
            DBuf<double> Y; Y.alloc((size_t)gp.nnzb() * bc * bc);
            run_triple(s, ty, bc, bc, bf, bf, 1, f.pool64, MG, f.pool64, nullptr, nullptr, Y);
            // G_c(i,j) = Minv_c(i) Y(i,j)
            TermList tg; tg.off.resize(gp.nnzb() + 1); tg.l.resize(gp.nnzb()); tg.x.resize(gp.nnzb()); tg.r.assign(gp.nnzb(), -1);
            for (int i = 0; i < gp.nbr; ++i)
                for (int k = gp.rowptr[i]; k < gp.rowptr[i + 1]; ++k) { tg.off[k] = k; tg.l[k] = i; tg.x[k] = k; }
            tg.off[gp.nnzb()] = (int)gp.nnzb();
            c.G[d].upload_pattern(gp, s);
            c.G[d].val.alloc((size_t)gp.nnzb() * bc * bc);
            run_triple(s, tg, bc, bc, bc, bc, 0, c.Minv, Y, nullptr, nullptr, nullptr, c.G[d].val);
        }
        // T_c = T^T T_f T
        {
            HostBSR tp; TermList tt;
            rap_symbolic(I, f.Tp.pat, tp, tt);
            c.Tp.upload_pattern(tp, s);
            c.Tp.val.alloc((size_t)tp.nnzb() * bc * bc);
            run_triple(s, tt, bc, bc, bf, bf, 1, f.pool64, f.Tp.val, f.pool64, nullptr, nullptr, c.Tp.val);
        }
        build_laplacian(s, N, bc, c.Tp, c.G, c.M, c.A);
        c.has_ops = true;
    } else {
        // A_c = T^T A_f T
        HostBSR ap; TermList ta;
        rap_symbolic(I, f.A.pat, ap, ta);
        c.A.upload_pattern(ap, s);
        c.A.val.alloc((size_t)ap.nnzb() * bc * bc);
        run_triple(s, ta, bc, bc, bf, bf, 1, f.pool64, f.A.val, f.pool64, nullptr, nullptr, c.A.val);
    }
}

// solve layouts, inverse diagonal blocks, colouring, vectors
// The last two colours a = nc-2, b = nc-1 of a multicolour level in one
// launch.  Rows of colour b are few (greedy colouring leaves the last colour
// to the rows that see all the others); each is grouped, by union-find over
// the a-b couplings, with all its colour-a neighbours into a cluster that one
// workgroup processes in two ordered ranges.  Colour-a rows with no colour-b
// neighbour need no ordering and fill plain workgroups.  Used only when every
// cluster fits one pass of a workgroup per range; otherwise the level keeps
// one launch per colour.
static void pair_schedule(Level& lv, hipStream_t s, const std::vector<int>* pidx, const std::vector<int>* didx)
{
    lv.pair_nwg = 0;
    if (lv.ncolor < 3) return;
    const int n = lv.nb, ca = lv.ncolor - 2, cb = lv.ncolor - 1;
    const int rpb = dev::rows_per_workgroup(lv.B);
    const HostBSR& P = lv.A.pat;
    std::vector<int> par(n);
    for (int i = 0; i < n; ++i) par[i] = i;
    auto find = [&](int a) { while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; } return a; };
    std::vector<char> in(n, 0);
    for (int i = 0; i < n; ++i) {
        if (lv.color[i] != cb) continue;
        in[i] = 1;
        for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
            const int j = P.cols[k];
            if (j != i && lv.color[j] == ca) { in[j] = 1; const int ri = find(i), rj = find(j); if (ri != rj) par[std::max(ri, rj)] = std::min(ri, rj); }
        }
    }
    // clusters in order of their smallest row, rows ascending in each range
    std::vector<int> cid(n, -1);
    std::vector<std::vector<int>> ra, rb;
    std::vector<int> plain;
    for (int i = 0; i < n; ++i) {
        if (lv.color[i] != ca && lv.color[i] != cb) continue;
        if (!in[i]) { plain.push_back(i); continue; }
        const int r = find(i);
        if (cid[r] < 0) { cid[r] = (int)ra.size(); ra.emplace_back(); rb.emplace_back(); }
        (lv.color[i] == ca ? ra : rb)[cid[r]].push_back(i);
    }
    for (size_t c = 0; c < ra.size(); ++c)
        if ((int)ra[c].size() > rpb || (int)rb[c].size() > rpb) return;
    std::vector<int> order, wg;
    order.reserve(plain.size() + 2 * ra.size());
    for (size_t q = 0; q < plain.size(); q += rpb) {
        const int a = (int)q, e = (int)std::min(plain.size(), q + rpb);
        wg.insert(wg.end(), {a, e, e});
    }
    order = plain;
    for (size_t c = 0; c < ra.size(); ++c) {
        const int a = (int)order.size();
        order.insert(order.end(), ra[c].begin(), ra[c].end());
        const int m = (int)order.size();
        order.insert(order.end(), rb[c].begin(), rb[c].end());
        wg.insert(wg.end(), {a, m, (int)order.size()});
    }
    gs_tables(P, order, lv.gw, s, lv.pair_desc, lv.pair_gcols, pidx, didx, &lv.color);
    lv.pair_wg.upload(wg.data(), wg.size(), s);
    HIPCHK(hipStreamSynchronize(s));
    lv.pair_nwg = (int)(wg.size() / 3);
}

// Distinct blocks of a device array of nblk row-major FP64 blocks (bs
// doubles each): pidx[k] = pool index of block k, rep[p] = first block of
// pool entry p.  Blocks are grouped by a 64-bit hash of their bits and every
// block is then compared with its entry bit for bit on the device, so the
// pool is exact.  false (no pool) when there are more than nblk / 4 distinct
// blocks -- the pool would not cut the traffic -- or a comparison fails.
static bool pool_blocks(hipStream_t s, long nblk, int bs, const double* blocks, std::vector<int>& pidx,
                        std::vector<int>& rep)
{
    pidx.clear(); rep.clear();
    if (nblk < 64) return false;
    DBuf<unsigned long long> hd;
    hd.alloc(nblk);
    dev::block_hash(s, nblk, bs, blocks, hd);
    std::vector<unsigned long long> h(nblk);
    hd.download(h.data(), h.size(), s);
    HIPCHK(hipStreamSynchronize(s));
    {
        // a strided sample first: an operator whose sample is mostly distinct
        // (unstructured meshes) is not worth the full grouping
        std::vector<unsigned long long> smp;
        const long step = std::max<long>(1, nblk / 4096);
        for (long k = 0; k < nblk; k += step) smp.push_back(h[k]);
        std::sort(smp.begin(), smp.end());
        const long nd = (long)(std::unique(smp.begin(), smp.end()) - smp.begin());
        if (nd * 2 > (long)smp.size()) return false;
    }
    std::unordered_map<unsigned long long, int> first;
    first.reserve(1024);
    pidx.resize(nblk);
    for (long k = 0; k < nblk; ++k) {
        auto it = first.find(h[k]);
        if (it == first.end()) {
            if ((long)rep.size() >= nblk / 4) { pidx.clear(); rep.clear(); return false; }
            first.emplace(h[k], (int)rep.size());
            pidx[k] = (int)rep.size();
            rep.push_back((int)k);
        } else {
            pidx[k] = it->second;
        }
    }
    DBuf<int> pd, rd, bad;
    pd.upload(pidx.data(), pidx.size(), s); rd.upload(rep.data(), rep.size(), s);
    bad.alloc(nblk);
    dev::block_verify(s, nblk, bs, blocks, pd, rd, bad);
    std::vector<int> hb(nblk);
    bad.download(hb.data(), hb.size(), s);
    HIPCHK(hipStreamSynchronize(s));
    for (int v : hb) if (v) { pidx.clear(); rep.clear(); return false; }
    return true;
}

// Rows grouped by block signature (D^-1 or diagonal pool index, blocks in
// order of pool index then column) into tiles of 16 rows, signatures in
// order of first appearance and rows in their given order within a
// signature.  blocks(i, out): the (pool index, column) pairs of row i's
// summed blocks.  Tables of width W: col[(16 t + r) W + u], pid[t W + u],
// meta[2 t] = {blocks, key}.
struct TileHost {
    std::vector<int> row, col, pid, meta;
    int n = 0;
};
template <typename F>
// colour (GS tiles): a column whose colour precedes the row's carries
// dev::kPidLower (the first sweep from zero reads those neighbours only)
static void tile_rows(const std::vector<int>& rows, int W, const std::vector<int>& dkey, F&& blocks, TileHost& th,
                      const std::vector<int>* colour = nullptr)
{
    std::map<std::vector<int>, int> sig;
    std::vector<std::vector<int>> members;
    std::vector<std::pair<int, int>> bl;
    for (int i : rows) {
        blocks(i, bl);
        std::sort(bl.begin(), bl.end());
        std::vector<int> key;
        key.reserve(bl.size() + 2);
        key.push_back(dkey.empty() ? 0 : dkey[i]);
        key.push_back((int)bl.size());
        for (const auto& e : bl) key.push_back(e.first);
        auto it = sig.find(key);
        int id;
        if (it == sig.end()) { id = (int)members.size(); sig.emplace(std::move(key), id); members.emplace_back(); }
        else id = it->second;
        members[id].push_back(i);
    }
    for (const std::vector<int>& m : members)
        for (size_t c0 = 0; c0 < m.size(); c0 += 16) {
            const int t = th.n++;
            th.row.resize((size_t)th.n * 16, -1);
            th.col.resize((size_t)th.n * W * 16, -1);
            th.pid.resize((size_t)th.n * W, -1);
            th.meta.resize((size_t)th.n * 2, 0);
            for (size_t r = 0; r < 16 && c0 + r < m.size(); ++r) {
                const int i = m[c0 + r];
                blocks(i, bl);
                std::sort(bl.begin(), bl.end());
                th.row[(size_t)t * 16 + r] = i;
                for (size_t u = 0; u < bl.size(); ++u) {
                    const int j = bl[u].second;
                    th.col[((size_t)t * 16 + r) * W + u] = j | ((colour && (*colour)[j] < (*colour)[i]) ? dev::kPidLower : 0);
                    if (r == 0) th.pid[(size_t)t * W + u] = bl[u].first;
                }
                if (r == 0) { th.meta[2 * t] = (int)bl.size(); th.meta[2 * t + 1] = dkey.empty() ? 0 : dkey[i]; }
            }
        }
}

static bool every_row_has_diagonal(const HostBSR& P)
{
    for (int i = 0; i + 1 < (int)P.rowptr.size(); ++i) {
        bool d = false;
        for (int k = P.rowptr[i]; k < P.rowptr[i + 1] && !d; ++k) d = P.cols[k] == i;
        if (!d) return false;
    }
    return true;
}

// row tiles of a pooled level: GS per colour, application over all rows
static void build_tiles(hipStream_t s, Level& lv, const std::vector<int>& pidx, const std::vector<int>& didx)
{
    const HostBSR& P = lv.A.pat;
    if (lv.nb > dev::kPidMask) return;   // (columns share their word with kPidLower)
    TileHost g, a;
    lv.t_W = dev::tile_width(lv.gw);
    lv.tile_off.assign(lv.ncolor + 1, 0);
    auto offd = [&](int i, std::vector<std::pair<int, int>>& bl) {
        bl.clear();
        bool diag = false;
        for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
            if (P.cols[k] == i && !diag) { diag = true; continue; }
            bl.push_back({pidx[k], P.cols[k]});
        }
    };
    for (int c = 0; c < lv.ncolor; ++c) {
        std::vector<int> rows;
        for (int i = 0; i < lv.nb; ++i) if (lv.color[i] == c) rows.push_back(i);
        tile_rows(rows, lv.t_W, didx, offd, g, &lv.color);
        lv.tile_off[c + 1] = g.n;
    }
    // the application's tiles: the diagonal block's pool index as the key,
    // the off-diagonal blocks in the tables
    std::vector<int> rows(lv.nb), adiag(lv.nb, -1);
    std::iota(rows.begin(), rows.end(), 0);
    for (int i = 0; i < lv.nb; ++i)
        for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k)
            if (P.cols[k] == i) { adiag[i] = pidx[k]; break; }
    tile_rows(rows, lv.t_W, adiag, offd, a);
    lv.a_ntiles = a.n;
    // what the kernels index with: rows and columns in [-1, nb), pool
    // indices in [-1, npool), keys in [0, ndpool) (GS) / [0, npool)
    auto check = [&](const TileHost& th, int nkey) {
        const int W = lv.t_W;
        bool ok = th.row.size() == (size_t)th.n * 16 && th.col.size() == (size_t)th.n * 16 * W &&
                  th.pid.size() == (size_t)th.n * W && th.meta.size() == (size_t)th.n * 2;
        for (int v : th.row) ok = ok && v >= -1 && v < lv.nb;
        for (int v : th.col) ok = ok && (v == -1 || (v >= 0 && (v & dev::kPidMask) < lv.nb));
        for (int v : th.pid) ok = ok && v >= -1 && v < lv.npool;
        for (int t = 0; t < th.n; ++t)
            ok = ok && th.meta[2 * t] >= 0 && th.meta[2 * t] <= W && th.meta[2 * t + 1] >= 0 && th.meta[2 * t + 1] < nkey;
        if (!ok) throw Error(-5, "row tiles: table out of range");
    };
    check(g, lv.ndpool);
    check(a, lv.npool);
    lv.t_row.upload(g.row.data(), g.row.size(), s); lv.t_col.upload(g.col.data(), g.col.size(), s);
    lv.t_pid.upload(g.pid.data(), g.pid.size(), s); lv.t_meta.upload(g.meta.data(), g.meta.size(), s);
    lv.a_row.upload(a.row.data(), a.row.size(), s); lv.a_col.upload(a.col.data(), a.col.size(), s);
    lv.a_pid.upload(a.pid.data(), a.pid.size(), s); lv.a_meta.upload(a.meta.data(), a.meta.size(), s);
    HIPCHK(hipStreamSynchronize(s));
}

static void finalize_level(hipStream_t s, Level& lv, int precision, bool lexicographic, bool pool_off)
{
    const int b = lv.B;
    const long nnzb = lv.A.nnzb();
    lv.A64.alloc((size_t)nnzb * b * b);
    dev::to_panel_f64(s, nnzb, b, lv.A.val, lv.A64);
    std::vector<int> ds = diag_slots(lv.A.pat);
    DBuf<int> dsl; dsl.upload(ds.data(), ds.size(), s);
    lv.Dinv_rm.alloc((size_t)lv.nb * b * b);
    dev::inverse(s, lv.nb, b, lv.A.val, dsl, lv.Dinv_rm);
    lv.D64.alloc((size_t)lv.nb * b * b);
    dev::to_panel_f64(s, lv.nb, b, lv.Dinv_rm, lv.D64);
    if (precision == 1) {
        lv.A32.alloc((size_t)nnzb * b * b);
        dev::to_panel_f32(s, nnzb, b, lv.A.val, lv.A32);
        lv.D32.alloc((size_t)lv.nb * b * b);
        dev::to_panel_f32(s, lv.nb, b, lv.Dinv_rm, lv.D32);
    }
    lv.A.build_ell(s);
    // pooled operator for the single-RHS kernels when A repeats its blocks
    std::vector<int> pidx, prep, didx, drep;
    const bool pooled = !pool_off && pool_blocks(s, nnzb, b * b, lv.A.val, pidx, prep);
    lv.npool = lv.ndpool = 0;
    if (pooled) {
        if (!pool_blocks(s, lv.nb, b * b, lv.Dinv_rm, didx, drep)) {
            didx.resize(lv.nb); drep.resize(lv.nb);
            for (int i = 0; i < lv.nb; ++i) didx[i] = drep[i] = i;
        }
        auto make_pool = [&](const std::vector<int>& rep, const double* src, DBuf<double>& p64, DBuf<float>& p32,
                             DBuf<double>& r64, DBuf<float>& r32, DBuf<double>& t64, DBuf<float>& t32, int tmode) {
            const long np = (long)rep.size();
            DBuf<int> rd; rd.upload(rep.data(), rep.size(), s);
            DBuf<double> rm; rm.alloc((size_t)np * b * b);
            dev::gather_blocks(s, (int)np, b * b, rd, src, rm);
            p64.alloc((size_t)np * b * b);
            dev::to_panel_f64(s, np, b, rm, p64);
            r64.alloc((size_t)np * b * dev::pool_pad(8, b));
            dev::to_rowpad_f64(s, np, b, rm, r64);
            t64.alloc((size_t)np * dev::tile_block_elems());
            dev::tile_pool<double>(s, np, b, rm, t64, tmode);
            if (precision == 1) {
                p32.alloc((size_t)np * b * b); dev::to_panel_f32(s, np, b, rm, p32);
                r32.alloc((size_t)np * b * dev::pool_pad(4, b)); dev::to_rowpad_f32(s, np, b, rm, r32);
                t32.alloc((size_t)np * dev::tile_block_elems());
                dev::tile_pool<float>(s, np, b, rm, t32, tmode);
            }
            HIPCHK(hipStreamSynchronize(s));
        };
        make_pool(prep, lv.A.val, lv.P64, lv.P32, lv.R64, lv.R32, lv.TA64, lv.TA32, 0);
        make_pool(drep, lv.Dinv_rm, lv.DP64, lv.DP32, lv.DR64, lv.DR32, lv.TD64, lv.TD32, 1);
        const int W = lv.A.ellw;
        std::vector<int> pe((size_t)lv.nb * W * 2, 0);
        for (int i = 0; i < lv.nb; ++i) {
            int m = 0;
            for (int k = lv.A.pat.rowptr[i]; k < lv.A.pat.rowptr[i + 1]; ++k, ++m) {
                pe[((size_t)i * W + m) * 2] = lv.A.pat.cols[k];
                pe[((size_t)i * W + m) * 2 + 1] = pidx[k];
            }
            for (; m < W; ++m) pe[((size_t)i * W + m) * 2] = -1;
        }
        lv.pell.upload(pe.data(), pe.size(), s);
        lv.pdidx.upload(didx.data(), didx.size(), s);
        HIPCHK(hipStreamSynchronize(s));
        lv.npool = (int)prep.size(); lv.ndpool = (int)drep.size();
        lv.hpidx = pidx; lv.hdidx = didx;
    }
    const std::vector<int>* pp = pooled ? &pidx : nullptr;
    const std::vector<int>* dp = pooled ? &didx : nullptr;
    lv.gw = 1;   // most off-diagonal blocks in a row
    for (int i = 0; i < lv.nb; ++i) {
        int m = 0;
        for (int k = lv.A.pat.rowptr[i]; k < lv.A.pat.rowptr[i + 1]; ++k) m += lv.A.pat.cols[k] != i;
        lv.gw = std::max(lv.gw, m);
    }
    auto group = [&](const std::vector<int>& color, int nc, std::vector<int>& off, DBuf<int>& rows_d, DBuf<int>& col_d,
                     DBuf<int>& desc_d, DBuf<int>& gcols_d) {
        off.assign(nc + 1, 0);
        for (int c : color) off[c + 1]++;
        for (int c = 0; c < nc; ++c) off[c + 1] += off[c];
        std::vector<int> rows(lv.nb), pos(off.begin(), off.end() - 1);
        for (int i = 0; i < lv.nb; ++i) rows[pos[color[i]]++] = i;
        rows_d.upload(rows.data(), rows.size(), s);
        col_d.upload(color.data(), color.size(), s);
        gs_tables(lv.A.pat, rows, lv.gw, s, desc_d, gcols_d, pp, dp, &color);
    };
    if (lexicographic) {
        // wavefronts of the lexicographic order: a row waits for its earlier
        // (forward) or later (reverse) neighbours only, so processing the
        // wavefronts in order reproduces the sequential sweep exactly
        lv.ncolor = wavefronts(lv.A.pat, false, lv.color);
        lv.rncolor = wavefronts(lv.A.pat, true, lv.rcolor);
        group(lv.rcolor, lv.rncolor, lv.rcolor_off, lv.rcolor_rows, lv.rcolor_dev, lv.rcolor_desc, lv.rcolor_gcols);
    } else {
        lv.ncolor = greedy_colors(lv.A.pat, lv.color);
    }
    group(lv.color, lv.ncolor, lv.color_off, lv.color_rows, lv.color_dev, lv.color_desc, lv.color_gcols);
    if (!lexicographic) pair_schedule(lv, s, pp, dp);
    // row tiles where the D^-1 blocks repeat too (a tile's rows share one)
    if (pooled && !lexicographic && 4 * (long)lv.ndpool <= lv.nb && dev::tile_supported(b, lv.gw) &&
        every_row_has_diagonal(lv.A.pat))
        build_tiles(s, lv, pidx, didx);
    const size_t n = (size_t)lv.nb * b;
    lv.v64.x.alloc(n); lv.v64.b.alloc(n); lv.v64.r.alloc(n); lv.v64.xalt.alloc(n);
    if (precision == 1) { lv.v32.x.alloc(n); lv.v32.b.alloc(n); lv.v32.r.alloc(n); lv.v32.xalt.alloc(n); }
    HIPCHK(hipStreamSynchronize(s));
}

static void upload_interp(hipStream_t s, Level& lv, int precision)
{
    Interp& I = lv.I;
    lv.parent.upload(I.parent.data(), I.parent.size(), s);
    lv.type.upload(I.type.data(), I.type.size(), s);
    lv.kid_off.upload(I.kid_off.data(), I.kid_off.size(), s);
    lv.kids.upload(I.kids.data(), I.kids.size(), s);
    lv.pool64.upload(I.pool.data(), I.pool.size(), s);
    if (precision == 1) {
        std::vector<float> p32(I.pool.begin(), I.pool.end());
        lv.pool32.upload(p32.data(), p32.size(), s);
        HIPCHK(hipStreamSynchronize(s));
    }
    // packed tables for the quadtree form: square blocks, few children per
    // coarse row and a small pool (a transposed copy of it is kept)
    int maxk = 0;
    for (int c = 0; c < I.nc; ++c) maxk = std::max(maxk, I.kid_off[c + 1] - I.kid_off[c]);
    lv.ilat = I.bf == I.bc && I.ntypes <= 64 && dev::interp_lat(I.bf, maxk);
    if (!lv.ilat) return;
    const int B = I.bf, KW = dev::kKidWidth;
    std::vector<int> kp((size_t)I.nc * KW * 2, -1);
    for (int c = 0; c < I.nc; ++c)
        for (int kk = I.kid_off[c], sl = 0; kk < I.kid_off[c + 1]; ++kk, ++sl) {
            const int k = I.kids[kk];
            kp[((size_t)c * KW + sl) * 2] = k;
            kp[((size_t)c * KW + sl) * 2 + 1] = I.type[k];
        }
    std::vector<double> pt(I.pool.size());
    for (int t = 0; t < I.ntypes; ++t)
        for (int c = 0; c < B; ++c)
            for (int q = 0; q < B; ++q) pt[((size_t)t * B + c) * B + q] = I.pool[((size_t)t * B + q) * B + c];
    lv.kpair.upload(kp.data(), kp.size(), s);
    DBuf<double> tr;
    tr.upload(pt.data(), pt.size(), s);
    lv.ipool64.alloc((size_t)I.ntypes * B * dev::pool_pad(8, B));
    lv.ipoolT64.alloc((size_t)I.ntypes * B * dev::pool_pad(8, B));
    dev::to_rowpad_f64(s, I.ntypes, B, lv.pool64, lv.ipool64);
    dev::to_rowpad_f64(s, I.ntypes, B, tr, lv.ipoolT64);
    if (precision == 1) {
        lv.ipool32.alloc((size_t)I.ntypes * B * dev::pool_pad(4, B));
        lv.ipoolT32.alloc((size_t)I.ntypes * B * dev::pool_pad(4, B));
        dev::to_rowpad_f32(s, I.ntypes, B, lv.pool64, lv.ipool32);
        dev::to_rowpad_f32(s, I.ntypes, B, tr, lv.ipoolT32);
    }
    HIPCHK(hipStreamSynchronize(s));
}

// directions of an agglomerate's coarse space beyond its rank have zero
// interpolation columns: give them a unit diagonal in A so they stay decoupled
static void pad_dropped(hipStream_t s, Level& c, const std::vector<int>& rank)
{
    std::vector<int> ds = diag_slots(c.A.pat);
    std::vector<long> idx;
    const int b = c.B;
    for (int a = 0; a < c.nb; ++a)
        for (int d = rank[a]; d < b; ++d) idx.push_back((long)ds[a] * b * b + (long)d * b + d);
    if (idx.empty()) return;
    DBuf<long> di;
    di.upload(idx.data(), idx.size(), s);
    dev::add_at(s, (long)idx.size(), di, 1.0, c.A.val);
    HIPCHK(hipStreamSynchronize(s));
}

void MGH::build(ProblemH* p, const MGParams& pm, const std::vector<std::vector<int>>* parts, const std::vector<int>* pdeg,
                bool hcoarsen)
{
    pr = p; ctx = p->ctx; prm = pm;
    if (prm.relaxation < 0 || prm.relaxation > 2) throw Error(-1, "Unknown relaxation method.");   // multigrid.h:213
    if (prm.coarse < 0 || prm.coarse > 1) throw Error(-1, "Unknown coarse solver.");              // multigrid.h:193
    if (prm.restriction < 0 || prm.restriction > 1) throw Error(-1, "Unknown restriction method."); // multigrid.h:164
    if (prm.precision < 0 || prm.precision > 1) throw Error(-1, "Unknown precision.");
    if (prm.npre < 0 || prm.npost < 0) throw Error(-1, "negative sweep count");
    const SimplexLdg* sx = p->simplex();
    int nl = 0;
    np = 0;
    std::vector<Interp> AI;                 // agglomerated hierarchy: all interpolations
    std::vector<std::vector<int>> crank;    // kept coarse directions per agglomerate
    submean = sx ? 0 : 1;
    if (sx) {
        if (!parts || parts->empty()) throw Error(-1, "an agglomerated hierarchy needs at least one partition");
        // RATRAT inverts the coarse mass matrices T^T M T, which Q_p box spaces
        // over a few P_p simplices can make singular (a lone simplex has fewer
        // nodes than box functions; two triangles' four vertices need not
        // determine a bilinear function)
        if (prm.restriction != 0) throw Error(-1, "agglomerated hierarchies coarsen by RAT (MGLDG_RAT) only");
        int n = sx->ne;
        for (const auto& q : *parts) {
            if ((int)q.size() != n) throw Error(-1, "partition size does not match its level");
            int nc = 0;
            for (int v : q) { if (v < 0 || v >= n) throw Error(-1, "agglomerate id out of range"); nc = std::max(nc, v + 1); }
            n = nc;
        }
        nl = (int)parts->size() + 1;
    } else {
        if (parts) throw Error(-1, "partitions apply to simplex problems only");
        const Quadtree& qt = *p->qt;
        np = pdeg ? (int)pdeg->size() : 0;
        for (int i = 0; i < np; ++i) {
            const int prev = i ? (*pdeg)[i - 1] : p->cart()->p;
            if ((*pdeg)[i] < 0 || (*pdeg)[i] >= prev)
                throw Error(-1, "p-multigrid degrees must decrease from the problem's and be >= 0");  // interpolationhierarchy.h:101
        }
        if (!pdeg) hcoarsen = true;
        if (hcoarsen && qt.num_levels < 2) throw Error(-1, "the quadtree needs at least two levels");  // interpolationhierarchy.h:33
        nl = 1 + np + (hcoarsen ? qt.num_levels - 1 : 0);
    }
    if (sx && pdeg) throw Error(-1, "p-multigrid hierarchies apply to quadtree problems only");
    hipStream_t s = ctx->stream;
    const int N = p->N;
    L.clear(); L.resize(nl);
    double th = 0;
    auto t0 = Clock::now();
    // fine level: the problem's operators (device copies share nothing with the problem)
    Level& f = L[0];
    f.nb = p->ne; f.B = p->b;
    auto clone = [&](DevOp& dst, const DevOp& src) {
        dst.upload_pattern(src.pat, s);
        dst.val.alloc(src.val.n);
        HIPCHK(hipMemcpyAsync(dst.val.p, src.val.p, src.val.n * sizeof(double), hipMemcpyDeviceToDevice, s));
    };
    clone(f.A, p->A);
    if (prm.restriction == 1) {
        clone(f.M, p->M); clone(f.Tp, p->T);
        for (int d = 0; d < N; ++d) clone(f.G[d], p->G[d]);
        f.has_ops = true;
    }
    if (sx) {
        auto th0 = Clock::now();
        Basis Q;   // coarse box basis
        Q.init(sx->N, sx->p + 1);
        build_interps_agglo(sx->mesh, sx->E, Q, *parts, AI, crank);
        th += secs(th0);
    }
    // p-levels (degrees pdeg), then h-levels at the last degree
    std::vector<Basis> pb(np);
    for (int i = 0; i < np; ++i) pb[i].init(N, (*pdeg)[i] + 1);
    for (int l = 0; l + 1 < nl; ++l) {
        auto th0 = Clock::now();
        if (sx) L[l].I = std::move(AI[l]);
        else if (l < np) build_pinterp(p->ne, l ? pb[l - 1] : p->cart()->basis, pb[l], L[l].I);
        else build_interp(*p->qt, np ? pb[np - 1] : p->cart()->basis, l - np, L[l].I);
        th += secs(th0);
        upload_interp(s, L[l], prm.precision);
        L[l + 1].nb = L[l].I.nc; L[l + 1].B = L[l].I.bc;
        coarsen(s, N, L[l], L[l + 1], prm.restriction);
        if (sx) pad_dropped(s, L[l + 1], crank[l]);
    }
    for (int l = 0; l < nl; ++l) finalize_level(s, L[l], prm.precision, prm.relaxation == 2, prm.pool_off != 0);
    // coarsest level: dense inverse of the assembled operator (FP64)
    Level& c = L[nl - 1];
    if (prm.coarse == 0) {
        const int b = c.B, n = c.nb * b;
        if (n > 6144) throw Error(-5, "coarsest level too large for the dense solve");
        std::vector<double> blk(c.A.val.n), D((size_t)n * n, 0.0);
        c.A.val.download(blk.data(), blk.size(), s);
        HIPCHK(hipStreamSynchronize(s));
        for (int i = 0; i < c.nb; ++i)
            for (int k = c.A.pat.rowptr[i]; k < c.A.pat.rowptr[i + 1]; ++k) {
                int j = c.A.pat.cols[k];
                for (int r = 0; r < b; ++r) for (int q = 0; q < b; ++q)
                    D[(size_t)(i * b + r) * n + j * b + q] = blk[(size_t)k * b * b + r * b + q];
            }
        std::vector<double> X = dense_inverse(D, n);
        c.Cinv.upload(X.data(), X.size(), s);
    }
    build_collapse();
    // PCG work space
    const size_t nd = (size_t)p->ndof;
    r.alloc(nd); zold.alloc(nd); d.alloc(nd); q.alloc(nd); xw.alloc(nd); bw.alloc(nd);
    S.alloc(16);
    part.alloc(std::max<size_t>(3 * 1024 + 64, (size_t)L[0].nb + 64));
    HIPCHK(hipMemsetAsync(S.p, 0, 16 * sizeof(double), s));
    HIPCHK(hipHostMalloc((void**)&S_host, 16 * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipStreamSynchronize(s));
    select_tiles();
    setup_dev_s = secs(t0) - th;
    setup_host_s = th;
}

MGH::~MGH()
{
    for (auto& g : graph) if (g) (void)hipGraphExecDestroy(g);
    for (auto& g : bgraph) if (g) (void)hipGraphExecDestroy(g);
    if (S_host) (void)hipHostFree(S_host);
}

template <> LevelVecs<double>& MGH::vecs<double>(int l) { return L[l].v64; }
template <> LevelVecs<float>& MGH::vecs<float>(int l) { return L[l].v32; }

template <typename T> static const T* Aop(const Level& lv);
template <> const double* Aop<double>(const Level& lv) { return lv.A64.p; }
template <> const float* Aop<float>(const Level& lv) { return lv.A32.p; }
template <typename T> static const T* Dop(const Level& lv);
template <> const double* Dop<double>(const Level& lv) { return lv.D64.p; }
template <> const float* Dop<float>(const Level& lv) { return lv.D32.p; }
// the operator as the single-RHS kernels read it: the block pool when the
// level is pooled (with the pair tables and D^-1 indices), else the full
// panel arrays
template <typename T> static const T* Sop(const Level& lv);
template <> const double* Sop<double>(const Level& lv) { return lv.npool ? lv.R64.p : lv.A64.p; }
template <> const float* Sop<float>(const Level& lv) { return lv.npool ? lv.R32.p : lv.A32.p; }
template <typename T> static const T* SDop(const Level& lv);
template <> const double* SDop<double>(const Level& lv) { return lv.npool ? lv.DR64.p : lv.D64.p; }
template <> const float* SDop<float>(const Level& lv) { return lv.npool ? lv.DR32.p : lv.D32.p; }
// the batched MFMA kernels read the pools in the panel layout
template <typename T> static const T* Mop(const Level& lv);
template <> const double* Mop<double>(const Level& lv) { return lv.npool ? lv.P64.p : lv.A64.p; }
template <> const float* Mop<float>(const Level& lv) { return lv.npool ? lv.P32.p : lv.A32.p; }
template <typename T> static const T* MDop(const Level& lv);
template <> const double* MDop<double>(const Level& lv) { return lv.npool ? lv.DP64.p : lv.D64.p; }
template <> const float* MDop<float>(const Level& lv) { return lv.npool ? lv.DP32.p : lv.D32.p; }
// the pools in the row tiles' layout
template <typename T> static const T* TAop(const Level& lv);
template <> const double* TAop<double>(const Level& lv) { return lv.TA64.p; }
template <> const float* TAop<float>(const Level& lv) { return lv.TA32.p; }
template <typename T> static const T* TDop(const Level& lv);
template <> const double* TDop<double>(const Level& lv) { return lv.TD64.p; }
template <> const float* TDop<float>(const Level& lv) { return lv.TD32.p; }
static const int* Sell(const Level& lv) { return lv.npool ? lv.pell.p : lv.A.ell.p; }
static const int* Sdidx(const Level& lv) { return lv.npool ? lv.pdidx.p : nullptr; }
static int Spool(const Level& lv) { return lv.npool ? 1 : 0; }
static size_t Sgw(const Level& lv) { return (size_t)lv.gw * (lv.npool ? 2 : 1); }   // ints per row of a GS column table
template <typename T> static const T* Pool(const Level& lv);
template <> const double* Pool<double>(const Level& lv) { return lv.pool64.p; }
template <> const float* Pool<float>(const Level& lv) { return lv.pool32.p; }
template <typename T> static const T* IPool(const Level& lv);
template <> const double* IPool<double>(const Level& lv) { return lv.ipool64.p; }
template <> const float* IPool<float>(const Level& lv) { return lv.ipool32.p; }
template <typename T> static const T* IPoolT(const Level& lv);
template <> const double* IPoolT<double>(const Level& lv) { return lv.ipoolT64.p; }
template <> const float* IPoolT<float>(const Level& lv) { return lv.ipoolT32.p; }

// The GS launches of one sweep in order.  Multicolour: colours ascending
// (forward) or descending (reverse).  Lexicographic: forward wavefronts
// ascending, or reverse-order wavefronts ascending.  skip marks the first
// sweep from x = 0 (rows after the current launch in sweep order are zero);
// fn receives the colour to pass to the kernels as `cur`, which for the first
// launch of such a sweep is dev::kFirstFwd / kFirstRev (it reads no neighbour).
template <typename F>
void MGH::sweep(const Level& lv, bool reverse, bool first_from_zero, F&& fn) const
{
    const bool lex = prm.relaxation == 2;
    const bool use_r = lex && reverse;
    const int nc = use_r ? lv.rncolor : lv.ncolor;
    const std::vector<int>& off = use_r ? lv.rcolor_off : lv.color_off;
    const int* rows = use_r ? lv.rcolor_rows.p : lv.color_rows.p;
    const int* dsc = use_r ? lv.rcolor_desc.p : lv.color_desc.p;
    const int* gcl = use_r ? lv.rcolor_gcols.p : lv.color_gcols.p;
    const int* cdev = use_r ? lv.rcolor_dev.p : lv.color_dev.p;
    const bool desc = reverse && !lex;
    for (int cc = 0; cc < nc; ++cc) {
        const int c = desc ? nc - 1 - cc : cc;
        const int skip = first_from_zero ? (desc ? 2 : 1) : 0;
        const int cur = (skip && cc == 0) ? (desc ? dev::kFirstRev : dev::kFirstFwd) : c;
        fn(rows + off[c], dsc + 4 * (size_t)off[c], gcl + Sgw(lv) * off[c], off[c + 1] - off[c], cdev, cur, skip);
    }
}

template <typename T>
void MGH::gs_colour(Level& lv, int c, int cur, int skip, const T* b, T* x)
{
    hipStream_t s = ctx->stream;
    const T omega = (T)prm.omega;
    if (lv.use_tiles[sizeof(T) == 4] & 1) {
        const int t0 = lv.tile_off[c], nt = lv.tile_off[c + 1] - t0;
        dev::tile_gs<T>(s, lv.B, nt, lv.t_row.p + 16 * (size_t)t0, lv.t_col.p + 16 * (size_t)lv.t_W * t0,
                        lv.t_pid.p + (size_t)lv.t_W * t0, lv.t_meta.p + 2 * (size_t)t0, lv.t_W, TAop<T>(lv), TDop<T>(lv), b,
                        x, omega, lv.color_dev, cur, skip);
        return;
    }
    const int o = lv.color_off[c];
    dev::Ops<T>::gs_color(s, lv.B, lv.color_off[c + 1] - o, lv.color_desc.p + 4 * (size_t)o, lv.color_gcols.p + Sgw(lv) * o,
                          lv.gw, Sop<T>(lv), SDop<T>(lv), b, x, omega, lv.color_dev, cur, skip, Spool(lv));
}

template <typename T>
int MGH::apply_level(const Level& lv, const T* x, const T* b, T* y, const T* dotv, double* dot_out, double* part)
{
    hipStream_t s = ctx->stream;
    if (lv.use_tiles[sizeof(T) == 4] & 2)
        return dev::tile_apply<T>(s, lv.B, lv.a_ntiles, lv.a_row, lv.a_col, lv.a_pid, lv.a_meta, lv.t_W, TAop<T>(lv), x, b,
                                  y, dotv, dot_out, part);
    return dev::Ops<T>::apply(s, lv.B, lv.nb, lv.A.rowptr, Sell(lv), lv.A.ellw, Sop<T>(lv), x, b, y, dotv, dot_out, part,
                              Spool(lv));
}

// Row tiles replace the vector kernels on pooled levels with at least
// 256 tiles (4k block rows) in FP32 and 2048 tiles (32k rows) in FP64, for
// both the GS sweep and the operator application.  Below that the tile
// launch has too few wavefronts to hide its chain of loads and MFMAs (FP64
// MFMAs take twice as long): timed one sweep and one application at a time
// on the adaptive p = 3 hierarchy (levels of 65k, 17k, 4.5k and 1.3k rows),
// the FP32 tiles take 0.75-0.9x the vector time from 4.5k rows up and the
// FP64 tiles 1.0-1.3x below 65k rows.  A fixed rule rather than a timed
// choice keeps the summation order, and so the iterates, the same from run
// to run.
void MGH::select_tiles()
{
    const int kTileMin[2] = {2048, 256};
    for (Level& lv : L)
        for (int p = 0; p < 2; ++p) {
            const bool on = lv.a_ntiles > 0 && prm.row_tiles != 2 && (prm.row_tiles == 1 || lv.a_ntiles >= kTileMin[p]);
            lv.use_tiles[p] = on ? 3 : 0;
        }
}

// Level::relax (multigrid.h:202-262); Gauss-Seidel is multicolour on the GPU:
// colours in increasing order forward, decreasing order in reverse.
// The first launch of a forward multicolour GS sweep from x = 0 (colour 0,
// reading no neighbour) can be done by the restriction that forms the
// level's right-hand side (Ops::restrict_lat with the colour arguments):
// quadtree interpolation into a pooled level whose colour 0 has a launch of
// its own.
bool MGH::fused_first(int l) const
{
    if (l < 1 || prm.relaxation != 1 || prm.npre < 1) return false;
    const Level& f = L[l - 1];
    const Level& c = L[l];
    return f.ilat && c.npool > 0 && c.pdidx.p && (c.pair_nwg == 0 || c.ncolor >= 3);
}

template <typename T>
void MGH::relax(int l, int n, bool reverse, bool zero_guess, bool first_done)
{
    Level& lv = L[l];
    LevelVecs<T>& v = vecs<T>(l);
    hipStream_t s = ctx->stream;
    const T omega = (T)prm.omega;
    if (Aop<T>(lv) == nullptr) throw Error(-4, "precision not built for this hierarchy");
    // zero_guess means x starts at zero whatever n is: Gauss-Seidel's first
    // sweep and Jacobi's first sweep read no x, but with no sweep at all x
    // must be cleared here (the reference zeroes x before relaxing,
    // multigrid.h:376)
    if (zero_guess && n == 0) { dev::Ops<T>::fill(s, (long)lv.nb * lv.B, (T)0, v.x); return; }
    if (prm.relaxation == 1 && lv.pair_nwg > 0) {
        // colours ascending (descending in reverse), the last two in one launch
        const int nc = lv.ncolor;
        for (int sw = 0; sw < n; ++sw) {
            const int skip = (zero_guess && sw == 0) ? (reverse ? 2 : 1) : 0;
            auto one = [&](int c) {
                const int cur = (skip && c == 0) ? dev::kFirstFwd : c;   // (forward: colour 0 goes first)
                gs_colour<T>(lv, c, cur, skip, v.b, v.x);
            };
            auto pair = [&] {
                // the pair goes first in reverse, and forward when there are only two colours
                const int c0 = reverse ? nc - 1 : nc - 2;
                const int cur0 = (skip && (reverse || nc == 2)) ? (reverse ? dev::kFirstRev : dev::kFirstFwd) : c0;
                dev::Ops<T>::gs_pair(s, lv.B, lv.pair_nwg, lv.pair_wg, reverse ? 1 : 0, lv.pair_desc, lv.pair_gcols, lv.gw,
                                     Sop<T>(lv), SDop<T>(lv), v.b, v.x, omega, lv.color_dev, cur0,
                                     reverse ? nc - 2 : nc - 1, skip, Spool(lv));
            };
            if (!reverse) { for (int c = (first_done && sw == 0) ? 1 : 0; c < nc - 2; ++c) one(c); pair(); }
            else { pair(); for (int c = nc - 3; c >= 0; --c) one(c); }
        }
    } else if (prm.relaxation == 1) {
        // multicolour, one launch per colour: ascending (descending in reverse)
        const int nc = lv.ncolor;
        for (int sw = 0; sw < n; ++sw) {
            const int skip = (zero_guess && sw == 0) ? (reverse ? 2 : 1) : 0;
            for (int cc = (first_done && sw == 0) ? 1 : 0; cc < nc; ++cc) {
                const int c = reverse ? nc - 1 - cc : cc;
                const int cur = (skip && cc == 0) ? (reverse ? dev::kFirstRev : dev::kFirstFwd) : c;
                gs_colour<T>(lv, c, cur, skip, v.b, v.x);
            }
        }
    } else if (prm.relaxation >= 1) {
        for (int sw = 0; sw < n; ++sw)
            sweep(lv, reverse, zero_guess && sw == 0,
                  [&](const int*, const int* dsc, const int* gcl, int nr, const int* cdev, int c, int skip) {
                      dev::Ops<T>::gs_color(s, lv.B, nr, dsc, gcl, lv.gw, Sop<T>(lv), SDop<T>(lv), v.b, v.x, omega, cdev, c,
                                            skip, Spool(lv));
                  });
    } else {
        // Jacobi ping-pongs between x and xalt; the result ends in x.  From
        // x = 0 the first sweep is omega D^-1 b (no A x, no clearing of x).
        T* cur = v.x.p; T* alt = v.xalt.p;
        int sw0 = 0;
        if (zero_guess) { dev::Ops<T>::jacobi0(s, lv.B, lv.nb, SDop<T>(lv), v.b, alt, omega, Sdidx(lv)); std::swap(cur, alt); sw0 = 1; }
        for (int sw = sw0; sw < n; ++sw) {
            dev::Ops<T>::jacobi(s, lv.B, lv.nb, lv.A.rowptr, Sell(lv), lv.A.ellw, Sop<T>(lv), SDop<T>(lv), v.b, cur, alt, omega,
                                Sdidx(lv));
            std::swap(cur, alt);
        }
        if (cur != v.x.p) dev::Ops<T>::copy(s, (long)lv.nb * lv.B, cur, v.x);
    }
}

// Multigrid::vcycle (multigrid.h:361-383), enqueued without host syncs
template <typename T>
void MGH::vcycle_from(int l0, bool x0_zero)
{
    hipStream_t s = ctx->stream;
    const int nl = (int)L.size();
    const int bottom = cut > l0 ? cut : nl - 1;
    bool zero = x0_zero;   // x of the starting level is zero (preconditioner use) or holds the iterate
    bool first_done = false;
    for (int l = l0; l < bottom; ++l) {
        Level& lv = L[l];
        LevelVecs<T>& v = vecs<T>(l);
        relax<T>(l, prm.npre, false, zero, first_done);
        apply_level<T>(lv, v.x, v.b, v.r, nullptr, nullptr, nullptr);
        Level& lc = L[l + 1];
        LevelVecs<T>& vc = vecs<T>(l + 1);
        // the restriction also does level l+1's first GS launch where it can
        first_done = l + 1 < bottom && fused_first(l + 1);
        if (first_done)
            dev::Ops<T>::restrict_lat(s, lv.B, lv.I.nc, lv.kpair, IPoolT<T>(lv), v.r, vc.b, lc.color_dev, lc.pdidx,
                                      SDop<T>(lc), vc.x, (T)prm.omega);
        else if (lv.ilat) dev::Ops<T>::restrict_lat(s, lv.B, lv.I.nc, lv.kpair, IPoolT<T>(lv), v.r, vc.b);
        else dev::Ops<T>::restrict_(s, lv.B, lc.B, lv.I.nc, lv.kid_off, lv.kids, lv.type, Pool<T>(lv), v.r, vc.b);
        // x_{l+1} = 0 (multigrid.h:376): relax() from a zero guess starts
        // there (its first sweep reads no x; with no pre-sweep it clears x)
        zero = true;
    }
    Level& c = L[bottom];
    LevelVecs<T>& vc = vecs<T>(bottom);
    if (bottom == cut) {
        // the V-cycle below `cut` as one precomputed linear map (FP64; the
        // FP32 V-cycle reads its FP32 copy, half the bytes)
        if (sizeof(T) == 4 && W32.p) dev::gemv_w32(s, c.nb * c.B, W32.p, (const float*)vc.b.p, (float*)vc.x.p);
        else dev::gemv_f64<T>(s, c.nb * c.B, W, vc.b, vc.x);
    } else if (prm.coarse == 0) {
        dev::Ops<T>::coarse(s, c.nb * c.B, c.Cinv, vc.b, vc.x, submean);
    } else {
        // b -= mean(b) then relaxations from x = 0 (multigrid.h:184-187, 367)
        if (submean) dev::Ops<T>::submean(s, (long)c.nb * c.B, vc.b);
        relax<T>(bottom, prm.npre, false, bottom == l0 ? x0_zero : true);
        relax<T>(bottom, prm.npost, true, false);
    }
    for (int l = bottom - 1; l >= l0; --l) {
        Level& lv = L[l];
        LevelVecs<T>& v = vecs<T>(l);
        if (lv.ilat) dev::Ops<T>::prolong_lat(s, lv.B, lv.I.nc, lv.kpair, IPool<T>(lv), vecs<T>(l + 1).x, v.x);
        else dev::Ops<T>::prolong(s, lv.B, L[l + 1].B, lv.nb, lv.parent, lv.type, Pool<T>(lv), vecs<T>(l + 1).x, v.x);
        relax<T>(l, prm.npost, true, false);
    }
}

// relaxation on k columns (FP64), used to form the collapsed coarse operator
void MGH::mrelax(int l, int k, int nsw, bool reverse, bool zero_guess, double* x, const double* b, double* alt)
{
    Level& lv = L[l];
    hipStream_t s = ctx->stream;
    const double omega = prm.omega;
    if (prm.relaxation >= 1) {
        for (int sw = 0; sw < nsw; ++sw)
            sweep(lv, reverse, zero_guess && sw == 0, [&](const int* rows, const int*, const int*, int nr, const int* cdev, int c, int skip) {
                dev::MOps<double>::gs_color(s, lv.B, k, nr, rows, lv.A.rowptr, lv.A.cols, lv.A64, lv.D64, b, x, omega, cdev, c,
                                            skip);
            });
    } else {
        const long n = (long)lv.nb * lv.B * k;
        if (zero_guess) dev::Ops<double>::fill(s, n, 0.0, x);
        double* cur = x; double* nxt = alt;
        for (int sw = 0; sw < nsw; ++sw) {
            dev::MOps<double>::jacobi(s, lv.B, k, lv.nb, lv.A.rowptr, lv.A.cols, lv.A64, lv.D64, b, cur, nxt, omega);
            std::swap(cur, nxt);
        }
        if (cur != x) dev::Ops<double>::copy(s, n, cur, x);
    }
}

// The V-cycle below a level is a fixed linear map of that level's right-hand
// side (the reference forms it column by column as Multigrid::vcycle_matrix,
// multigrid.h:386-400).  Form it once for the first level with at most
// coarse_dense_max dofs by running the sub-cycle on all identity columns at
// once, and apply it as one dense FP64 GEMV in every V-cycle.
void MGH::build_collapse()
{
    cut = -1;
    const int nl = (int)L.size();
    if (prm.coarse_dense_max <= 0) return;
    for (int l = 1; l + 2 < nl; ++l)
        if ((long)L[l].nb * L[l].B <= prm.coarse_dense_max) { cut = l; break; }
    if (cut < 0) return;
    hipStream_t s = ctx->stream;
    const int k = L[cut].nb * L[cut].B;
    std::vector<DBuf<double>> X(nl), Bv(nl), R(nl);
    for (int l = cut; l < nl; ++l) {
        const size_t n = (size_t)L[l].nb * L[l].B * k;
        X[l].alloc(n); Bv[l].alloc(n); R[l].alloc(n);
    }
    dev::eye(s, k, Bv[cut]);
    for (int l = cut; l + 1 < nl; ++l) {
        Level& lv = L[l];
        mrelax(l, k, prm.npre, false, true, X[l], Bv[l], R[l]);
        if (prm.npre == 0) dev::Ops<double>::fill(s, (long)lv.nb * lv.B * k, 0.0, X[l]);
        dev::MOps<double>::apply(s, lv.B, k, lv.nb, lv.A.rowptr, lv.A.cols, lv.A64, X[l], Bv[l], R[l]);
        dev::MOps<double>::restrict_(s, lv.B, L[l + 1].B, k, lv.I.nc, lv.kid_off, lv.kids, lv.type, lv.pool64, R[l], Bv[l + 1]);
    }
    Level& c = L[nl - 1];
    const int nc = c.nb * c.B;
    if (prm.coarse == 0) {
        dev::MOps<double>::coarse(s, nc, k, c.Cinv, Bv[nl - 1], X[nl - 1], submean);
    } else {
        if (submean) dev::MOps<double>::coarse(s, nc, k, nullptr, Bv[nl - 1], nullptr, 1);
        mrelax(nl - 1, k, prm.npre, false, true, X[nl - 1], Bv[nl - 1], R[nl - 1]);
        if (prm.npre == 0) dev::Ops<double>::fill(s, (long)nc * k, 0.0, X[nl - 1]);
        mrelax(nl - 1, k, prm.npost, true, false, X[nl - 1], Bv[nl - 1], R[nl - 1]);
    }
    for (int l = nl - 2; l >= cut; --l) {
        Level& lv = L[l];
        dev::MOps<double>::prolong(s, lv.B, L[l + 1].B, k, lv.nb, lv.parent, lv.type, lv.pool64, X[l + 1], X[l]);
        mrelax(l, k, prm.npost, true, false, X[l], Bv[l], R[l]);
    }
    HIPCHK(hipStreamSynchronize(s));
    W = std::move(X[cut]);
    if (prm.precision == 1 && k % 4 == 0) {
        W32.alloc((size_t)k * k);
        dev::f64_to_f32(s, (long)k * k, W, W32);
    }
    Wt.alloc((size_t)dev::gemm_tiles_size(k));
    dev::gemm_tiles<double>(s, k, W, Wt);
    if (prm.precision == 1) {
        Wt32.alloc((size_t)dev::gemm_tiles_size(k));
        dev::gemm_tiles<float>(s, k, W, Wt32);
    }
    HIPCHK(hipStreamSynchronize(s));
}

// z = V(r): enqueue the preconditioner on the stream
// x_0 = V(b_0) on level 0's vectors of the given precision (v32 / v64): the
// caller puts the right-hand side in L[0].v*.b and reads the result in L[0].v*.x
void MGH::precon_enqueue(int precision)
{
    if (precision == 1) {
        if (!L[0].A32.p) throw Error(-4, "mixed precision was not built (mg params precision)");
        vcycle_from<float>(0);
    } else {
        vcycle_from<double>(0);
    }
}

void MGH::precon(int precision, bool use_graph)
{
    hipStream_t s = ctx->stream;
    if (!use_graph) { precon_enqueue(precision); return; }
    hipGraphExec_t& g = graph[precision];
    if (!g) {
        hipGraph_t gr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try { precon_enqueue(precision); }
        catch (...) { hipGraph_t junk; (void)hipStreamEndCapture(s, &junk); if (junk) (void)hipGraphDestroy(junk); throw; }
        HIPCHK(hipStreamEndCapture(s, &gr));
        size_t nodes = 0;
        if (hipGraphGetNodes(gr, nullptr, &nodes) == hipSuccess) graph_nodes[precision] = (int)nodes;
        HIPCHK(hipGraphInstantiate(&g, gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
    }
    HIPCHK(hipGraphLaunch(g, s));
}

// pcg (krylov.h:22-68) with the V-cycle preconditioner; FP64 recurrences
SolveStats MGH::solve(const double* b, double* x, const SolveParams& sp)
{
    hipStream_t s = ctx->stream;
    Level& f = L[0];
    const long n = pr->ndof;
    double* part = this->part.p;
    const bool mixed = sp.precision == 1;
    if (mixed && !f.A32.p) throw Error(-4, "mixed precision was not built (mg params precision)");
    if (mixed && sp.flexible && !zold32.p) zold32.alloc(n);
    SolveStats st;
    hipEvent_t e0, e1, es;
    HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventCreateWithFlags(&es, hipEventDisableTiming));
    HIPCHK(hipEventRecord(e0, s));
    // the scalars go to pinned host memory behind an event; the host waits
    // for that copy only, so work enqueued after it keeps the GPU busy
    auto copyS = [&]() {
        HIPCHK(hipMemcpyAsync(S_host, S.p, 16 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(es, s));
    };
    auto waitS = [&]() { HIPCHK(hipEventSynchronize(es)); };
    auto fetchS = [&]() { copyS(); waitS(); };
    auto residual = [&](bool with_rr) {   // r = b - A x  (optionally r.r into S[RR])
        apply_level<double>(f, x, b, r, with_rr ? r.p : nullptr, with_rr ? S.p + dev::S_RR : nullptr, part);
    };
    // q = A d and the d.q partials: the first kernel of an iteration.  It
    // writes only q and the partials, so it is enqueued before the host has
    // tested the previous iteration's scalars and is simply not used when
    // the loop ends there.  pcg_xr / pcg_x sum the partials into S[DQ].
    int ndq = 0, nrz = 0;
    auto product = [&]() {
        ndq = apply_level<double>(f, d, nullptr, q, d, nullptr, part);
    };
    // the preconditioner reads r in its own precision from L[0].v*.b and
    // leaves z in L[0].v*.x; the CG kernels convert on the fly
    auto to_vin = [&] {
        if (mixed) dev::f64_to_f32(s, n, r, f.v32.b); else dev::Ops<double>::copy(s, n, r, f.v64.b);
    };
    bool rz_old = false;
    auto rz = [&](bool read_old) {   // r.z, r.r (+ r.zold) partials; zold = z
        rz_old = read_old && sp.flexible;
        if (mixed) nrz = dev::pcg_rz<float>(s, n, S, r, f.v32.x, sp.flexible ? zold32.p : nullptr, read_old, part);
        else nrz = dev::pcg_rz<double>(s, n, S, r, f.v64.x, sp.flexible ? zold.p : nullptr, read_old, part);
    };
    auto dupd = [&](bool init) {     // S[DELTA] = r.z, S[RR] = r.r, beta; d = z + beta d  (init: d = z)
        if (mixed) dev::pcg_d<float>(s, n, S, f.v32.x, d, init, part, nrz, rz_old);
        else dev::pcg_d<double>(s, n, S, f.v64.x, d, init, part, nrz, rz_old);
    };
    // ||b||
    dev::dot(s, n, b, b, S.p + dev::S_BB, part);
    residual(false);
    to_vin();
    precon(sp.precision, sp.use_graph != 0);
    rz(false);
    dupd(true);
    copyS();
    product();
    waitS();
    const double bnorm = std::sqrt(S_host[dev::S_BB]);
    const double delta0 = S_host[dev::S_DELTA];
    double delta = delta0, rr = S_host[dev::S_RR];
    int i = 0;
    for (;;) {
        if (!std::isfinite(delta) || !std::isfinite(rr)) { st.status = 3; break; }
        bool go = (sp.criterion == 0) ? (delta > sp.tol * sp.tol * delta0) : (std::sqrt(rr) > sp.tol * bnorm);
        if (!go && sp.criterion == 1 && bnorm > 0) {
            // confirm on the true residual; restart the direction if it fails
            residual(true);
            fetchS();
            if (std::sqrt(S_host[dev::S_RR]) > sp.tol * bnorm && i < sp.maxit) {
                to_vin();
                precon(sp.precision, sp.use_graph != 0);
                rz(false);
                dupd(true);
                copyS();
                product();                 // the product of the restarted direction
                waitS();
                delta = S_host[dev::S_DELTA]; rr = S_host[dev::S_RR];
                go = true;
            }
        }
        if (!go) { st.status = 0; break; }
        if (i >= sp.maxit) { st.status = 1; break; }
        // q = A d and S[DQ] are in flight (product() above)
        if (i % 50 == 0) {   // periodic true residual (krylov.h:47-52)
            dev::pcg_x(s, n, S, d, x, part, ndq);
            residual(true);
            to_vin();
        } else if (mixed) {
            dev::pcg_xr<float>(s, n, S, d, q, x, r, f.v32.b, part, ndq);
        } else {
            dev::pcg_xr<double>(s, n, S, d, q, x, r, f.v64.b, part, ndq);
        }
        precon(sp.precision, sp.use_graph != 0);
        rz(sp.flexible != 0);
        dupd(false);
        ++i;
        copyS();
        product();
        waitS();
        // S was copied before product() was enqueued: S_host[DQ] is d.q of
        // the direction this iteration used (as in the batched loop), not of
        // the next one, so a converged z = 0 never reads as a breakdown here
        if (!(S_host[dev::S_DQ] > 0)) { st.status = std::isfinite(S_host[dev::S_DQ]) ? 2 : 3; break; }
        delta = S_host[dev::S_DELTA]; rr = S_host[dev::S_RR];
    }
    // true residual at exit
    residual(true);
    HIPCHK(hipEventRecord(e1, s));
    fetchS();
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(es);
    st.iterations = i;
    st.relres_true = bnorm > 0 ? std::sqrt(S_host[dev::S_RR]) / bnorm : 0.0;
    st.relres_prec = delta0 > 0 ? std::sqrt(std::fabs(delta / delta0)) : 0.0;
    st.seconds = ms * 1e-3;
    if (st.status == 0 && sp.criterion == 1 && st.relres_true > sp.tol) st.status = 1;
    return st;
}

// ---------------------------------------------------------------------------- batched
template <typename T> static T* bx(MGH::BatchLevel& b);
template <> double* bx<double>(MGH::BatchLevel& b) { return b.x64.p; }
template <> float* bx<float>(MGH::BatchLevel& b) { return b.x32.p; }
template <typename T> static T* bb(MGH::BatchLevel& b);
template <> double* bb<double>(MGH::BatchLevel& b) { return b.b64.p; }
template <> float* bb<float>(MGH::BatchLevel& b) { return b.b32.p; }
template <typename T> static T* br(MGH::BatchLevel& b);
template <> double* br<double>(MGH::BatchLevel& b) { return b.r64.p; }
template <> float* br<float>(MGH::BatchLevel& b) { return b.r32.p; }
template <typename T> static T* balt(MGH::BatchLevel& b);
template <> double* balt<double>(MGH::BatchLevel& b) { return b.alt64.p; }
template <> float* balt<float>(MGH::BatchLevel& b) { return b.alt32.p; }

void MGH::ensure_batch(int k, int precision)
{
    if (k > 256) throw Error(-1, "at most 256 right-hand sides per batch");
    if (precision == 1 && !L[0].A32.p) throw Error(-4, "mixed precision was not built (mg params precision)");
    if (bk == k && bprec >= precision) return;
    for (auto& g : bgraph) if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    const int nl = (int)L.size();
    BL.clear(); BL.resize(nl);
    for (int l = 0; l < nl; ++l) {
        const size_t n = (size_t)L[l].nb * L[l].B * k;
        BL[l].x64.alloc(n); BL[l].b64.alloc(n); BL[l].r64.alloc(n); BL[l].alt64.alloc(n);
        if (precision == 1) { BL[l].x32.alloc(n); BL[l].b32.alloc(n); BL[l].r32.alloc(n); BL[l].alt32.alloc(n); }
    }
    const size_t nd = (size_t)pr->ndof * k;
    Br.alloc(nd); Bzold.alloc(nd); Bd.alloc(nd); Bq.alloc(nd);
    if (precision == 1) Bzold32.alloc(nd);
    BS.alloc((size_t)k * dev::S_COUNT);
    Bpart.alloc((size_t)1024 * 3 * k + 64);
    Bactive.alloc(k);
    if (cut >= 0) Bgemm.alloc((size_t)dev::gemm_part_size(L[cut].nb * L[cut].B, k));
    bk = k; bprec = precision;
    tune_batch(k, precision);
    // the CG product on MFMA tiles sums d.q in its own workgroups
    const int nt0 = btune[0].empty() ? 0 : btune[0][0].apply;
    if (nt0) Bdq.alloc((size_t)dev::apply_dot_part_size(nt0, k, L[0].nb));
    else Bdq.release();
}

template <typename T>
void MGH::bapply(int l, int k, const T* x, const T* b, T* y)
{
    Level& lv = L[l];
    const int nt = (int)btune[sizeof(T) == 4].size() > l ? btune[sizeof(T) == 4][l].apply : 0;
    if (nt) dev::MF<T>::apply(ctx->stream, lv.B, nt, k, lv.nb, lv.A.rowptr, Sell(lv), lv.A.ellw, Mop<T>(lv), x, b, y, lv.npool > 0);
    else dev::MOps<T>::apply(ctx->stream, lv.B, k, lv.nb, lv.A.rowptr, lv.A.cols, Aop<T>(lv), x, b, y);
}

// Autotuner of the batched block kernels: on every level and precision in
// use, time one operator application and one smoother sweep with the vector
// kernels and with MFMA tiles of 16, 32 and 64 columns (those that fit k),
// on the level's own operators and batch buffers, and keep the fastest.
// bforce >= 0 fixes the choice instead (0: vector kernels).
void MGH::tune_batch(int k, int precision)
{
    hipStream_t s = ctx->stream;
    const int nl = (int)L.size();
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
    auto timed = [&](auto&& fn) {
        fn();                                    // warm-up
        HIPCHK(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r) fn();
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        return ms / 3;
    };
    const int cand[4] = {0, 1, 2, 4};
    for (int p = 0; p <= precision; ++p) {
        btune[p].assign(nl, BTune());
        for (int l = 0; l < nl; ++l) {
            Level& lv = L[l];
            BTune& t = btune[p][l];
            if (!dev::mf_supported(lv.B)) continue;
            if (bforce >= 0) {
                int nt = std::min(bforce, dev::mf_max_tiles(lv.B));
                while (nt > 1 && k % nt) nt >>= 1;                    // (as the launchers do)
                t.apply = t.smooth = nt;
                continue;
            }
            const long m = (long)lv.nb * lv.B * k;
            auto run = [&](auto tag) {
                using T = decltype(tag);
                BatchLevel& v = BL[l];
                dev::Ops<T>::fill(s, m, (T)1, bx<T>(v));
                dev::Ops<T>::fill(s, m, (T)1, bb<T>(v));
                float best_a = 1e30f, best_s = 1e30f;
                for (int ci = 0; ci < 4; ++ci) {
                    const int nt = cand[ci];
                    if (nt > 1 && 16 * nt / 2 >= k) continue;      // more tiles than columns
                    if (nt > dev::mf_max_tiles(lv.B)) continue;
                    if (nt > 1 && k % nt) continue;                 // a wave's nt columns are one vector access
                    t.apply = nt; t.smooth = nt;
                    const float ma = timed([&] { bapply<T>(l, k, bx<T>(v), bb<T>(v), br<T>(v)); });
                    const float msm = timed([&] { brelax<T>(l, k, 1, false, false); });
                    t.apply_ms[ci] = ma; t.smooth_ms[ci] = msm;
                    if (ma < best_a) { best_a = ma; t.apply = nt; }
                    if (msm < best_s) { best_s = msm; t.smooth = nt; }
                }
                int ba = 0, bs = 0;
                for (int ci = 0; ci < 4; ++ci) {
                    if (t.apply_ms[ci] > 0 && t.apply_ms[ci] <= best_a) ba = cand[ci];
                    if (t.smooth_ms[ci] > 0 && t.smooth_ms[ci] <= best_s) bs = cand[ci];
                }
                t.apply = ba; t.smooth = bs;
            };
            if (p == 0) run(double());
            else run(float());
        }
    }
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
}

template <typename T>
void MGH::brelax(int l, int k, int n, bool reverse, bool zero_guess)
{
    Level& lv = L[l];
    BatchLevel& v = BL[l];
    hipStream_t s = ctx->stream;
    const T omega = (T)prm.omega;
    const int nt = (int)btune[sizeof(T) == 4].size() > l ? btune[sizeof(T) == 4][l].smooth : 0;
    if (prm.relaxation >= 1) {
        for (int sw = 0; sw < n; ++sw)
            sweep(lv, reverse, zero_guess && sw == 0, [&](const int* rows, const int* dsc, const int* gcl, int nr, const int* cdev, int c, int skip) {
                if (nt) dev::MF<T>::gs_color(s, lv.B, nt, k, nr, dsc, gcl, lv.gw, Mop<T>(lv), MDop<T>(lv), bb<T>(v), bx<T>(v), omega,
                                             cdev, c, skip, lv.npool > 0);
                else dev::MOps<T>::gs_color(s, lv.B, k, nr, rows, lv.A.rowptr, lv.A.cols, Aop<T>(lv), Dop<T>(lv), bb<T>(v),
                                            bx<T>(v), omega, cdev, c, skip);
            });
    } else {
        const long m = (long)lv.nb * lv.B * k;
        T* cur = bx<T>(v); T* nxt = balt<T>(v);
        int sw0 = 0;
        if (zero_guess) {   // from x = 0 the first sweep is omega D^-1 b
            if (n == 0) dev::Ops<T>::fill(s, m, (T)0, bx<T>(v));
            else { dev::MOps<T>::jacobi0(s, lv.B, k, lv.nb, Dop<T>(lv), bb<T>(v), nxt, omega); std::swap(cur, nxt); sw0 = 1; }
        }
        for (int sw = sw0; sw < n; ++sw) {
            if (nt) dev::MF<T>::jacobi(s, lv.B, nt, k, lv.nb, lv.A.rowptr, Sell(lv), lv.A.ellw, Mop<T>(lv), MDop<T>(lv), bb<T>(v),
                                       cur, nxt, omega, Sdidx(lv));
            else dev::MOps<T>::jacobi(s, lv.B, k, lv.nb, lv.A.rowptr, lv.A.cols, Aop<T>(lv), Dop<T>(lv), bb<T>(v), cur, nxt, omega);
            std::swap(cur, nxt);
        }
        if (cur != bx<T>(v)) dev::Ops<T>::copy(s, m, cur, bx<T>(v));
    }
}

// MFMA tiles for the batched restriction from / prolongation into level l:
// quadtree levels (kid tables) whose operator application runs on MFMA
// tiles take the same tile count; 0: the vector kernels
template <typename T>
int MGH::binterp(int l) const
{
    const Level& lv = L[l];
    const auto& tune = btune[sizeof(T) == 4];
    if (!lv.ilat || !dev::mf_interp_supported(lv.B) || (int)tune.size() <= l) return 0;
    return tune[l].apply;
}

// Multigrid::vcycle (multigrid.h:361-383) on k columns, x0 = 0
template <typename T>
void MGH::bvcycle(int k)
{
    hipStream_t s = ctx->stream;
    const int nl = (int)L.size();
    const int bottom = cut > 0 ? cut : nl - 1;
    for (int l = 0; l < bottom; ++l) {
        Level& lv = L[l];
        brelax<T>(l, k, prm.npre, false, true);
        if (prm.npre == 0) dev::Ops<T>::fill(s, (long)lv.nb * lv.B * k, (T)0, bx<T>(BL[l]));
        if (const int nt = binterp<T>(l)) {
            // the residual restricted in one pass (the level's application is on MFMA tiles)
            dev::MF<T>::resrestrict(s, lv.B, nt, k, lv.I.nc, lv.kpair, Pool<T>(lv), lv.A.rowptr, Sell(lv), lv.A.ellw, Mop<T>(lv),
                                    lv.npool > 0, bx<T>(BL[l]), bb<T>(BL[l]), bb<T>(BL[l + 1]));
            continue;
        }
        bapply<T>(l, k, bx<T>(BL[l]), bb<T>(BL[l]), br<T>(BL[l]));
        dev::MOps<T>::restrict_(s, lv.B, L[l + 1].B, k, lv.I.nc, lv.kid_off, lv.kids, lv.type, Pool<T>(lv), br<T>(BL[l]),
                                bb<T>(BL[l + 1]));
    }
    Level& c = L[bottom];
    const int n = c.nb * c.B;
    if (bottom == cut) {
        if (sizeof(T) == 4 && Wt32.p)
            dev::gemm_mf<float, float>(s, n, k, Wt32.p, (const float*)bb<T>(BL[bottom]), (float*)bx<T>(BL[bottom]), Bgemm);
        else
            dev::gemm_mf<T, double>(s, n, k, Wt.p, bb<T>(BL[bottom]), bx<T>(BL[bottom]), Bgemm);
    } else if (prm.coarse == 0) {
        dev::MOps<T>::coarse(s, n, k, c.Cinv, bb<T>(BL[bottom]), bx<T>(BL[bottom]), submean);
    } else {
        if (submean) dev::MOps<T>::coarse(s, n, k, nullptr, bb<T>(BL[bottom]), nullptr, 1);
        brelax<T>(bottom, k, prm.npre, false, true);
        if (prm.npre == 0) dev::Ops<T>::fill(s, (long)n * k, (T)0, bx<T>(BL[bottom]));
        brelax<T>(bottom, k, prm.npost, true, false);
    }
    for (int l = bottom - 1; l >= 0; --l) {
        Level& lv = L[l];
        if (const int nt = binterp<T>(l))
            dev::MF<T>::prolong(s, lv.B, nt, k, lv.I.nc, lv.kpair, Pool<T>(lv), bx<T>(BL[l + 1]), bx<T>(BL[l]));
        else
            dev::MOps<T>::prolong(s, lv.B, L[l + 1].B, k, lv.nb, lv.parent, lv.type, Pool<T>(lv), bx<T>(BL[l + 1]), bx<T>(BL[l]));
        brelax<T>(l, k, prm.npost, true, false);
    }
}

void MGH::bprecon(int precision, int k, bool use_graph)
{
    hipStream_t s = ctx->stream;
    // the V-cycle alone: right-hand sides in BL[0].b*, results in BL[0].x*
    auto enqueue = [&] {
        if (precision == 1) bvcycle<float>(k); else bvcycle<double>(k);
    };
    if (!use_graph) { enqueue(); return; }
    hipGraphExec_t& g = bgraph[precision];
    if (!g) {
        hipGraph_t gr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try { enqueue(); }
        catch (...) { hipGraph_t junk; (void)hipStreamEndCapture(s, &junk); if (junk) (void)hipGraphDestroy(junk); throw; }
        HIPCHK(hipStreamEndCapture(s, &gr));
        size_t nodes = 0;
        if (hipGraphGetNodes(gr, nullptr, &nodes) == hipSuccess) graph_nodes[precision] = (int)nodes;
        HIPCHK(hipGraphInstantiate(&g, gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
    }
    HIPCHK(hipGraphLaunch(g, s));
}

// k independent PCG recurrences (krylov.h:22-68 per column) sharing every
// operator application; per-column scalars and convergence masks
std::vector<SolveStats> MGH::solve_batched(const double* b, double* x, int k, const SolveParams& sp)
{
    ensure_batch(k, sp.precision);
    hipStream_t s = ctx->stream;
    const long n = pr->ndof;
    const int NS = dev::S_COUNT;
    std::vector<double> Sh((size_t)k * NS);
    std::vector<int> act(k, 1), it(k, 0);
    std::vector<SolveStats> st(k);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
    auto fetch = [&] {
        HIPCHK(hipMemcpyAsync(Sh.data(), BS.p, Sh.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    };
    auto push_active = [&] { HIPCHK(hipMemcpyAsync(Bactive.p, act.data(), k * sizeof(int), hipMemcpyHostToDevice, s)); };
    auto residual = [&] { bapply<double>(0, k, x, b, Br); };
    // the preconditioner reads r in its own precision from BL[0].b* and
    // leaves z in BL[0].x*; the CG kernels convert on the fly
    const bool mixed = sp.precision == 1;
    auto to_vin = [&] {
        if (mixed) dev::f64_to_f32(s, n * k, Br, BL[0].b32); else dev::Ops<double>::copy(s, n * k, Br, BL[0].b64);
    };
    auto rz = [&](bool read_old) {
        if (mixed) dev::bpcg_rz<float>(s, n, k, BS, Bactive, Br, BL[0].x32, sp.flexible ? Bzold32.p : nullptr, read_old, Bpart);
        else dev::bpcg_rz<double>(s, n, k, BS, Bactive, Br, BL[0].x64, sp.flexible ? Bzold.p : nullptr, read_old, Bpart);
    };
    auto dupd = [&](bool init) {
        if (mixed) dev::bpcg_d<float>(s, n, k, BS, Bactive, BL[0].x32, Bd, init);
        else dev::bpcg_d<double>(s, n, k, BS, Bactive, BL[0].x64, Bd, init);
    };
    HIPCHK(hipMemsetAsync(BS.p, 0, BS.n * sizeof(double), s));
    push_active();
    dev::coldot(s, n, k, b, b, BS.p + dev::S_BB, NS, Bpart);
    residual();
    to_vin();
    bprecon(sp.precision, k, sp.use_graph != 0);
    rz(false);
    dupd(true);
    fetch();
    std::vector<double> bnorm(k), delta0(k);
    for (int c = 0; c < k; ++c) { bnorm[c] = std::sqrt(Sh[(size_t)c * NS + dev::S_BB]); delta0[c] = Sh[(size_t)c * NS + dev::S_DELTA]; }
    auto converged = [&](int c) {
        const double* Sc = &Sh[(size_t)c * NS];
        return sp.criterion == 0 ? !(Sc[dev::S_DELTA] > sp.tol * sp.tol * delta0[c]) : !(std::sqrt(Sc[dev::S_RR]) > sp.tol * bnorm[c]);
    };
    auto update_active = [&](bool after_iter) {
        bool changed = false;
        for (int c = 0; c < k; ++c) {
            if (!act[c]) continue;
            const double* Sc = &Sh[(size_t)c * NS];
            int status = -1;
            if (!std::isfinite(Sc[dev::S_DELTA]) || !std::isfinite(Sc[dev::S_RR])) status = 3;
            else if (after_iter && !(Sc[dev::S_DQ] > 0)) status = std::isfinite(Sc[dev::S_DQ]) ? 2 : 3;
            else if (converged(c)) status = 0;
            if (status >= 0) { act[c] = 0; st[c].status = status; changed = true; }
        }
        if (changed) push_active();
    };
    update_active(false);
    int i = 0;
    for (;;) {
        bool any = false;
        for (int c = 0; c < k; ++c) any |= act[c] != 0;
        if (!any && sp.criterion == 1 && i < sp.maxit) {
            // confirm converged columns on the true residual; restart those that fail
            residual();
            dev::coldot(s, n, k, Br, Br, BS.p + dev::S_RR, NS, Bpart);
            fetch();
            std::vector<int> redo;
            for (int c = 0; c < k; ++c)
                if (st[c].status == 0 && std::sqrt(Sh[(size_t)c * NS + dev::S_RR]) > sp.tol * bnorm[c]) redo.push_back(c);
            if (!redo.empty()) {
                for (int c : redo) act[c] = 1;
                push_active();
                to_vin();
                bprecon(sp.precision, k, sp.use_graph != 0);
                rz(false);
                fetch();
                for (int c : redo) Sh[(size_t)c * NS + dev::S_BETA] = 0.0;
                HIPCHK(hipMemcpyAsync(BS.p, Sh.data(), Sh.size() * sizeof(double), hipMemcpyHostToDevice, s));
                dupd(false);
                any = true;
            }
        }
        if (!any) break;
        if (i >= sp.maxit) { for (int c = 0; c < k; ++c) if (act[c]) st[c].status = 1; break; }
        // q = A d and d.q
        const Level& l0 = L[0];
        const int nt0 = btune[0].empty() ? 0 : btune[0][0].apply;
        if (nt0 && Bdq.p) {
            dev::apply_dot(s, l0.B, nt0, k, l0.nb, l0.A.rowptr, Sell(l0), l0.A.ellw, Mop<double>(l0), Bd, Bq, l0.npool > 0, Bdq,
                           BS.p + dev::S_DQ, NS);
        } else {
            bapply<double>(0, k, Bd, nullptr, Bq);
            dev::coldot(s, n, k, Bd, Bq, BS.p + dev::S_DQ, NS, Bpart);
        }
        if (i % 50 == 0) {   // periodic true residual (krylov.h:47-52)
            dev::bpcg_x(s, n, k, BS, Bactive, Bd, x);
            residual();
            to_vin();
        } else if (mixed) {
            dev::bpcg_xr<float>(s, n, k, BS, Bactive, Bd, Bq, x, Br, BL[0].b32.p, Bpart);
        } else {
            dev::bpcg_xr<double>(s, n, k, BS, Bactive, Bd, Bq, x, Br, BL[0].b64.p, Bpart);
        }
        bprecon(sp.precision, k, sp.use_graph != 0);
        rz(sp.flexible != 0);
        dupd(false);
        ++i;
        for (int c = 0; c < k; ++c) it[c] += act[c];
        fetch();
        update_active(true);
    }
    residual();
    dev::coldot(s, n, k, Br, Br, BS.p + dev::S_RR, NS, Bpart);
    HIPCHK(hipEventRecord(e1, s));
    fetch();
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    for (int c = 0; c < k; ++c) {
        const double* Sc = &Sh[(size_t)c * NS];
        st[c].iterations = it[c];
        st[c].relres_true = bnorm[c] > 0 ? std::sqrt(Sc[dev::S_RR]) / bnorm[c] : 0.0;
        st[c].relres_prec = delta0[c] > 0 ? std::sqrt(std::fabs(Sc[dev::S_DELTA] / delta0[c])) : 0.0;
        st[c].seconds = ms * 1e-3;
        if (st[c].status == 0 && sp.criterion == 1 && st[c].relres_true > sp.tol) st[c].status = 1;
    }
    return st;
}

// adjoint solve: A is symmetric (T symmetric, G^T M G symmetric), so the
// adjoint system is solved by the same MG-PCG
std::vector<SolveStats> MGH::adjoint(const double* dLdu, double* lambda, double* dLdf, double* dLdg, int k,
                                     const SolveParams& sp)
{
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)pr->ndof * k;
    DBuf<double> own;
    double* lam = lambda;
    if (!lam) { own.alloc(n); lam = own.p; }
    dev::Ops<double>::fill(s, (long)n, 0.0, lam);
    std::vector<SolveStats> st;
    if (k == 1) st.push_back(solve(dLdu, lam, sp));
    else st = solve_batched(dLdu, lam, k, sp);
    pr->adjoint_maps(lam, k, dLdf, dLdg);
    return st;
}

void MGH::spmv(int l, int precision, const void* x, void* y)
{
    Level& lv = L[l];
    if (precision == 1) apply_level<float>(lv, (const float*)x, nullptr, (float*)y, nullptr, nullptr, nullptr);
    else apply_level<double>(lv, (const double*)x, nullptr, (double*)y, nullptr, nullptr, nullptr);
}

void MGH::smooth(int l, int precision, int n, bool reverse, const void* b, void* x)
{
    Level& lv = L[l];
    hipStream_t s = ctx->stream;
    const long m = (long)lv.nb * lv.B;
    if (precision == 1) {
        dev::Ops<float>::copy(s, m, (const float*)b, lv.v32.b); dev::Ops<float>::copy(s, m, (const float*)x, lv.v32.x);
        relax<float>(l, n, reverse, false);
        dev::Ops<float>::copy(s, m, lv.v32.x, (float*)x);
    } else {
        dev::Ops<double>::copy(s, m, (const double*)b, lv.v64.b); dev::Ops<double>::copy(s, m, (const double*)x, lv.v64.x);
        relax<double>(l, n, reverse, false);
        dev::Ops<double>::copy(s, m, lv.v64.x, (double*)x);
    }
}

// The reference driver's stationary solver (tests/cartesian/dg_tests.cpp:309-332):
// V-cycles from the current solution, ||b - A x|| before and after each.
// FP64: the reference's form, Multigrid::vcycle from x (multigrid.h:361-383).
// Mixed: the correction form x += V(b - A x) with the V-cycle in FP32 from
// zero and the residual and update in FP64 (the same iteration in exact
// arithmetic: every step of a V-cycle is affine in (x, b)).  tol > 0 stops
// once ||b - A x|| <= tol ||b||.
void MGH::vcycles(int precision, const double* b, double* x, int ncycle, double tol, double* rnorm, int* cycles)
{
    hipStream_t s = ctx->stream;
    Level& f = L[0];
    const long n = pr->ndof;
    double* part = this->part.p;
    if (precision == 1 && !f.A32.p) throw Error(-4, "mixed precision was not built (mg params precision)");
    auto fetch = [&] {
        HIPCHK(hipMemcpyAsync(S_host, S.p, 16 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    };
    auto resid = [&]() -> double {   // r = b - A x, ||r||
        apply_level<double>(f, x, b, r, r.p, S.p + dev::S_RR, part);
        fetch();
        return std::sqrt(S_host[dev::S_RR]);
    };
    dev::dot(s, n, b, b, S.p + dev::S_BB, part);
    fetch();
    const double bnorm = std::sqrt(S_host[dev::S_BB]);
    double rn = resid();
    if (rnorm) rnorm[0] = rn;
    int i = 0;
    for (; i < ncycle; ++i) {
        if (tol > 0 && rn <= tol * bnorm) break;
        if (precision == 0) {
            dev::Ops<double>::copy(s, n, b, f.v64.b);
            dev::Ops<double>::copy(s, n, x, f.v64.x);
            vcycle_from<double>(0, false);
            dev::Ops<double>::copy(s, n, f.v64.x, x);
        } else {
            dev::f64_to_f32(s, n, r, f.v32.b);
            vcycle_from<float>(0, true);
            dev::f32_to_f64(s, n, f.v32.x, q);
            dev::axpby(s, n, 1.0, q, 1.0, x, x);
        }
        rn = resid();
        if (rnorm) rnorm[i + 1] = rn;
    }
    if (cycles) *cycles = i;
}

void MGH::vcycle(int precision, const double* b, double* x)
{
    hipStream_t s = ctx->stream;
    const long n = pr->ndof;
    if (precision == 1) {
        if (!L[0].A32.p) throw Error(-4, "mixed precision was not built (mg params precision)");
        dev::f64_to_f32(s, n, b, L[0].v32.b);
        precon(precision, false);
        dev::f32_to_f64(s, n, L[0].v32.x, x);
    } else {
        dev::Ops<double>::copy(s, n, b, L[0].v64.b);
        precon(precision, false);
        dev::Ops<double>::copy(s, n, L[0].v64.x, x);
    }
}

// Multigrid::vcycle_matrix (multigrid.h:386-400): the V-cycle from a zero
// iterate applied to every unit vector.  The unit vectors go through the
// batched V-cycle k at a time (columns c0 .. c0+k-1 of the identity; a last
// partial chunk carries zero columns, whose V-cycle is zero).
void MGH::vcycle_matrix(int precision, double* V)
{
    hipStream_t s = ctx->stream;
    const long n = pr->ndof;
    const int k = (int)std::min<long>(n, 32);
    ensure_batch(k, precision);
    std::vector<double> E((size_t)n * k), X((size_t)n * k);
    BatchLevel& b0 = BL[0];
    for (long c0 = 0; c0 < n; c0 += k) {
        const int kc = (int)std::min<long>(k, n - c0);
        std::fill(E.begin(), E.end(), 0.0);
        for (int j = 0; j < kc; ++j) E[(size_t)(c0 + j) * k + j] = 1.0;
        b0.b64.upload(E.data(), E.size(), s);
        if (precision == 1) dev::f64_to_f32(s, n * k, b0.b64, b0.b32);
        bprecon(precision, k, false);
        if (precision == 1) dev::f32_to_f64(s, n * k, b0.x32, b0.x64);
        b0.x64.download(X.data(), X.size(), s);
        HIPCHK(hipStreamSynchronize(s));
        for (long i = 0; i < n; ++i)
            for (int j = 0; j < kc; ++j) V[i * n + c0 + j] = X[(size_t)i * k + j];
    }
}

// ---------------------------------------------------------------------------- measurement
// Algorithmic HBM bytes of one GS launch over the rows R of one colour (or
// of the pair launch's two colours), each stream counted once: b and x of R
// (x not read in the first sweep from zero), x written, x of the distinct
// neighbours the launch reads (skip / cur as the kernels: in the first sweep
// from zero only neighbours already updated in sweep order, none for the
// sweep's first launch; their colours are read), the operator's blocks and
// the launch's tables.  Unpooled: every off-diagonal block with its column,
// the D^-1 blocks and the 16-byte row descriptors.  Pooled: the (column,
// pool index) pair of each off-diagonal block, the distinct A and D^-1 pool
// blocks touched once each, the descriptors.  Row tiles (ntiles > 0): the
// tiles' tables (rows, columns and pool indices at width W, two ints) and
// the pool blocks in the tile layout.
static double gs_launch_bytes(const Level& lv, const int* rows, int nr, int sa, std::vector<int>& mark, int stamp,
                              int skip = 0, int cur = 0, int ntiles = 0, const std::vector<int>* colour = nullptr)
{
    const std::vector<int>& col = colour ? *colour : lv.color;
    const HostBSR& P = lv.A.pat;
    const double b2 = (double)lv.B * lv.B;
    const bool none = (skip == 1 && cur <= dev::kFirstFwd) || (skip == 2 && cur >= dev::kFirstRev);
    double offd = 0, nbrs = 0, apool = 0, dpool = 0;
    std::vector<char> seenA(lv.npool, 0), seenD(lv.ndpool, 0);
    for (int q = 0; q < nr; ++q) {
        const int i = rows[q];
        if (lv.npool && !seenD[lv.hdidx[i]]) { seenD[lv.hdidx[i]] = 1; dpool += 1; }
        for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
            const int j = P.cols[k];
            if (j == i) continue;
            offd += 1;
            if (none) continue;
            if (skip == 1 && col[j] > cur) continue;
            if (skip == 2 && col[j] < cur) continue;
            if (lv.npool && !seenA[lv.hpidx[k]]) { seenA[lv.hpidx[k]] = 1; apool += 1; }
            if (mark[j] != stamp) { mark[j] = stamp; nbrs += 1; }
        }
    }
    const double vec = nr * ((skip ? 2.0 : 3.0) * lv.B * sa) + nbrs * lv.B * sa + ((skip && !none) ? 4.0 * offd : 0.0);
    if (ntiles > 0)
        return ntiles * (64.0 + 68.0 * lv.t_W + 8.0) + (apool + dpool) * dev::tile_block_elems() * sa + vec;
    if (lv.npool) return (none ? 0.0 : offd * 8.0) + (apool + dpool) * b2 * sa + nr * 16.0 + vec;
    return (none ? 0.0 : offd * (b2 * sa + 4.0)) + nr * (b2 * sa + 16.0) + vec;
}

// Algorithmic bytes of one operator application y = A x (or b - A x) on a
// level: x once, y written (b read), the blocks as above (every block with
// its column index and the row pointers; pooled: the distinct blocks and the
// (column, pool index) pairs; row tiles: the tables and the tile-layout
// pool).
static double apply_bytes(const Level& lv, int sa, bool tiles, bool with_b)
{
    const double n = (double)lv.nb * lv.B, b2 = (double)lv.B * lv.B;
    const double vec = (with_b ? 3.0 : 2.0) * n * sa;
    if (tiles) return lv.a_ntiles * (64.0 + 68.0 * lv.t_W + 8.0) + (double)lv.npool * dev::tile_block_elems() * sa + vec;
    if (lv.npool) return (double)lv.npool * b2 * sa + 8.0 * lv.A.nnzb() + 4.0 * (lv.nb + 1) + vec;
    return (double)lv.A.nnzb() * (b2 * sa + 4.0) + 4.0 * (lv.nb + 1) + vec;
}

// The launches of one V-cycle (the preconditioner of the single-RHS solve,
// from x = 0, in the given precision) and their algorithmic bytes, walked in
// the order precon_enqueue issues them: relaxations (as relax()), residuals,
// restrictions, the coarse solve (dense inverse, collapsed map or
// relaxations) and prolongations.
void MGH::vcycle_cost(int precision, int* launches, double* bytes)
{
    const int sa = precision == 1 ? 4 : 8;
    const int nl = (int)L.size();
    const int bottom = cut > 0 ? cut : nl - 1;
    long nlaunch = 0;
    double by = 0;
    auto relax_cost = [&](int l, int n, bool reverse, bool zero, bool first_done) {
        const Level& lv = L[l];
        const double nv = (double)lv.nb * lv.B;
        if (zero && n == 0) { ++nlaunch; by += nv * sa; return; }
        if (prm.relaxation == 0) {
            // Jacobi ping-pongs between x and its partner, the first sweep
            // from zero reading only b and D^-1; an odd count ends with a copy
            const double blocks = lv.npool ? (double)(lv.npool + lv.ndpool) * lv.B * lv.B * sa + 8.0 * lv.A.nnzb() + 4.0 * lv.nb
                                           : (double)(lv.A.nnzb() + lv.nb) * lv.B * lv.B * sa + 4.0 * lv.A.nnzb();
            for (int sw = 0; sw < n; ++sw) {
                ++nlaunch;
                by += (zero && sw == 0) ? (double)(lv.npool ? lv.ndpool : lv.nb) * lv.B * lv.B * sa + 2.0 * nv * sa
                                        : blocks + 4.0 * (lv.nb + 1) + 3.0 * nv * sa;
            }
            if (n % 2 == 1) { ++nlaunch; by += 2.0 * nv * sa; }
            return;
        }
        std::vector<int> mark(lv.nb, -1);
        int stamp = 0;
        if (prm.relaxation == 2) {
            // lexicographic: the wavefronts of the sweep's direction, in
            // increasing order, the first from zero reading earlier ones
            const std::vector<int>& col = reverse ? lv.rcolor : lv.color;
            const int nw = reverse ? lv.rncolor : lv.ncolor;
            std::vector<std::vector<int>> wf(nw);
            for (int i = 0; i < lv.nb; ++i) wf[col[i]].push_back(i);
            for (int sw = 0; sw < n; ++sw) {
                const int skip = (zero && sw == 0) ? 1 : 0;
                for (int c = 0; c < nw; ++c) {
                    ++nlaunch;
                    by += gs_launch_bytes(lv, wf[c].data(), (int)wf[c].size(), sa, mark, ++stamp, skip,
                                          (skip && c == 0) ? dev::kFirstFwd : c, 0, &col);
                }
            }
            return;
        }
        const int nc = lv.ncolor;
        const bool tiles = lv.use_tiles[precision == 1] & 1;
        const bool paired = prm.relaxation == 1 && lv.pair_nwg > 0;
        std::vector<std::vector<int>> by_colour(nc);
        for (int i = 0; i < lv.nb; ++i) by_colour[lv.color[i]].push_back(i);
        for (int sw = 0; sw < n; ++sw) {
            const int skip = (zero && sw == 0) ? (reverse ? 2 : 1) : 0;
            auto single = [&](int c, int cur) {
                // (colour 0 of the first sweep done by the restriction: its
                // launch is gone, its b, D^-1 and x traffic counted there)
                if (first_done && sw == 0 && c == 0 && !reverse) return;
                ++nlaunch;
                const int nt = tiles ? lv.tile_off[c + 1] - lv.tile_off[c] : 0;
                by += gs_launch_bytes(lv, by_colour[c].data(), (int)by_colour[c].size(), sa, mark, ++stamp, skip, cur, nt);
            };
            if (paired) {
                auto pair = [&](int cur0) {
                    ++nlaunch;
                    std::vector<int> r2 = by_colour[nc - 2];
                    r2.insert(r2.end(), by_colour[nc - 1].begin(), by_colour[nc - 1].end());
                    by += gs_launch_bytes(lv, r2.data(), (int)r2.size(), sa, mark, ++stamp, skip, cur0) + 12.0 * lv.pair_nwg;
                };
                if (!reverse) {
                    for (int c = 0; c < nc - 2; ++c) single(c, (skip && c == 0) ? dev::kFirstFwd : c);
                    pair((skip && nc == 2) ? dev::kFirstFwd : nc - 2);
                } else {
                    pair(skip ? dev::kFirstRev : nc - 1);
                    for (int c = nc - 3; c >= 0; --c) single(c, c);
                }
            } else {
                for (int cc = 0; cc < nc; ++cc) {
                    const int c = reverse ? nc - 1 - cc : cc;
                    single(c, (skip && cc == 0) ? (reverse ? dev::kFirstRev : dev::kFirstFwd) : c);
                }
            }
        }
    };
    for (int l = 0; l < bottom; ++l) {
        const Level& lv = L[l];
        const Level& lc = L[l + 1];
        relax_cost(l, prm.npre, false, true, l > 0 && fused_first(l));
        ++nlaunch; by += apply_bytes(lv, sa, lv.use_tiles[precision == 1] & 2, true);
        // restriction: r of the fine level, b of the coarse, the child lists
        // and the interpolation pool; fused with the coarse level's first GS
        // launch, also the colours, D^-1 indices and pool and x of colour 0
        ++nlaunch;
        by += ((double)lv.nb * lv.B + (double)lc.nb * lc.B) * sa + 4.0 * (lc.nb + 1) + 8.0 * lv.nb +
              (double)lv.I.pool.size() * sa;
        if (l + 1 < bottom && fused_first(l + 1)) {
            const double n0 = (double)(lc.color_off[1] - lc.color_off[0]);
            by += 8.0 * lc.nb + (double)lc.ndpool * lc.B * lc.B * sa + n0 * lc.B * sa;
        }
    }
    const Level& c = L[bottom];
    const double k = (double)c.nb * c.B;
    if (bottom == cut) { ++nlaunch; by += k * k * ((precision == 1 && W32.p) ? 4.0 : 8.0) + 2.0 * k * sa; }
    else if (prm.coarse == 0) { ++nlaunch; by += k * k * 8.0 + 2.0 * k * sa; }
    else { relax_cost(bottom, prm.npre, false, true, false); relax_cost(bottom, prm.npost, true, false, false); }
    for (int l = bottom - 1; l >= 0; --l) {
        const Level& lv = L[l];
        const Level& lc = L[l + 1];
        // prolongation: x of the coarse level, x of the fine read and written,
        // parents and types, the interpolation pool
        ++nlaunch;
        by += ((double)lc.nb * lc.B + 2.0 * lv.nb * lv.B) * sa + 8.0 * lv.nb + (double)lv.I.pool.size() * sa;
        relax_cost(l, prm.npost, true, false, false);
    }
    *launches = (int)nlaunch;
    *bytes = by;
}

// The CG work of one iteration besides the V-cycle (solve()'s loop): q = A d
// with d.q (FP64, the fine level), x += alpha d and r -= alpha q with r.r and
// r written in the V-cycle's precision, r.z (r.z_old, z_old written, in the
// flexible form) with r.r, d = z + beta d; each reduction's finalize launch
// and the 128-byte copy of the scalars.
void MGH::iteration_cost(int precision, bool flexible, int* launches, double* bytes)
{
    const Level& f = L[0];
    const int sv = precision == 1 ? 4 : 8;
    const double n = (double)f.nb * f.B;
    const double part = 8.0 * 1024 * 3;
    double by = apply_bytes(f, 8, f.use_tiles[0] & 2, false);             // d read (also for d.q), q written
    by += 8.0 * n * 4 + 8.0 * n * 2 + sv * n;                             // xr: x, d, r, q read; x, r written; v
    by += 8.0 * n + sv * n + (flexible ? 2.0 * sv * n : 0.0);              // rz
    by += sv * n + 16.0 * n;                                               // d
    by += 3 * part + 128.0;
    // product, xr, rz and d; the reductions' second passes run inside xr and d
    *launches = 4;
    *bytes = by;
}

void MGH::profile_relax(int l, int precision, int nrep, double* ms, double* bytes, int* launches, int part)
{
    Level& lv = L[l];
    hipStream_t s = ctx->stream;
    const long n = (long)lv.nb * lv.B;
    const int sa = precision == 1 ? 4 : 8;
    if (precision == 1) { dev::Ops<float>::fill(s, n, 1.f, lv.v32.b); dev::Ops<float>::fill(s, n, 0.f, lv.v32.x); }
    else { dev::Ops<double>::fill(s, n, 1.0, lv.v64.b); dev::Ops<double>::fill(s, n, 0.0, lv.v64.x); }
    const bool gs = prm.relaxation >= 1, paired = prm.relaxation == 1 && lv.pair_nwg > 0;
    if (part != 0 && !paired) part = part == 1 ? 0 : -1;
    if (part < 0) throw Error(-2, "profile_relax: level has no paired launch");
    // launches of one sweep and their bytes (forward order)
    std::vector<int> mark(lv.nb, -1);
    int per = 0;
    double by = 0;
    if (!gs) {
        per = 1;
        by = lv.npool
           ? (double)(lv.npool + lv.ndpool) * lv.B * lv.B * sa + 8.0 * lv.A.nnzb() + 4.0 * lv.nb + 4.0 * (lv.nb + 1) + 3.0 * n * sa
           : (double)lv.A.nnzb() * lv.B * lv.B * sa + (double)lv.nb * lv.B * lv.B * sa + 4.0 * lv.A.nnzb() +
             4.0 * (lv.nb + 1) + 3.0 * n * sa;
    } else {
        const int nc = lv.ncolor, nsingle = paired ? nc - 2 : nc;
        std::vector<int> host_rows(lv.nb);
        HIPCHK(hipMemcpy(host_rows.data(), lv.color_rows.p, sizeof(int) * lv.nb, hipMemcpyDeviceToHost));
        const bool tiles = lv.use_tiles[precision == 1] & 1;
        if (part != 2)
            for (int c = 0; c < nsingle; ++c) {
                by += gs_launch_bytes(lv, host_rows.data() + lv.color_off[c], lv.color_off[c + 1] - lv.color_off[c], sa, mark, c,
                                      0, 0, tiles ? lv.tile_off[c + 1] - lv.tile_off[c] : 0);
                ++per;
            }
        if (paired && part != 1) {
            const int o = lv.color_off[nc - 2];
            by += gs_launch_bytes(lv, host_rows.data() + o, lv.color_off[nc] - o, sa, mark, nc) + 12.0 * lv.pair_nwg;
            ++per;
        }
    }
    auto run = [&](int reps) {
        if (part == 0) {
            if (precision == 1) relax<float>(l, reps, false, false); else relax<double>(l, reps, false, false);
            return;
        }
        const int nc = lv.ncolor;
        for (int r = 0; r < reps; ++r) {
            if (part == 1)
                for (int c = 0; c < nc - 2; ++c) {
                    if (precision == 1) gs_colour<float>(lv, c, c, 0, lv.v32.b, lv.v32.x);
                    else gs_colour<double>(lv, c, c, 0, lv.v64.b, lv.v64.x);
                }
            else if (precision == 1)
                dev::Ops<float>::gs_pair(s, lv.B, lv.pair_nwg, lv.pair_wg, 0, lv.pair_desc, lv.pair_gcols, lv.gw, Sop<float>(lv),
                                         SDop<float>(lv), lv.v32.b, lv.v32.x, (float)prm.omega, lv.color_dev, nc - 2, nc - 1, 0,
                                         Spool(lv));
            else
                dev::Ops<double>::gs_pair(s, lv.B, lv.pair_nwg, lv.pair_wg, 0, lv.pair_desc, lv.pair_gcols, lv.gw, Sop<double>(lv),
                                          SDop<double>(lv), lv.v64.b, lv.v64.x, prm.omega, lv.color_dev, nc - 2, nc - 1, 0,
                                          Spool(lv));
        }
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
    run(1);   // warm
    HIPCHK(hipEventRecord(e0, s));
    run(nrep);
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    *ms = t / (double)(nrep * per);
    *bytes = by / per;
    if (launches) *launches = nrep * per;
}

void MGH::profile_spmv(int l, int precision, int nrep, double* ms, double* bytes)
{
    Level& lv = L[l];
    hipStream_t s = ctx->stream;
    const long n = (long)lv.nb * lv.B;
    const int sa = precision == 1 ? 4 : 8;
    if (precision == 1) dev::Ops<float>::fill(s, n, 1.f, lv.v32.x); else dev::Ops<double>::fill(s, n, 1.0, lv.v64.x);
    auto once = [&] {
        if (precision == 1) apply_level<float>(lv, lv.v32.x, nullptr, lv.v32.r, nullptr, nullptr, nullptr);
        else apply_level<double>(lv, lv.v64.x, nullptr, lv.v64.r, nullptr, nullptr, nullptr);
    };
    once();
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
    for (int i = 0; i < nrep; ++i) once();
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    *ms = t / nrep;
    *bytes = apply_bytes(lv, sa, lv.use_tiles[precision == 1] & 2, false);
}

}  // namespace mgldg
