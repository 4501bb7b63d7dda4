// mg.hpp — device-resident LDG problem, multigrid hierarchy and MG-PCG driver.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "simplex.hpp"

namespace mgldg {

void hip_check(hipError_t e, const char* what);
#define HIPCHK(x) ::mgldg::hip_check((x), #x)

// owning device buffer
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept { if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; } return *this; }
    ~DBuf() { release(); }
    void alloc(size_t count);
    void upload(const T* h, size_t count, hipStream_t s);
    void download(T* h, size_t count, hipStream_t s) const;
    void release();
    operator T*() const { return p; }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    DBuf<double> part;       // reduction partials
    DBuf<unsigned> counter;  // reduction counter (kept at zero between kernels)
    DBuf<double> colpart;    // per-column reduction partials and sums (mean_zero), grown on demand
    double* scratch = nullptr;  // this device's copy of the library's static scratch array (dev::scratch_address)
    void init(int dev, void* stream);
    ~Ctx();
};

// a block-sparse operator with host pattern and device row-major FP64 values
struct DevOp {
    HostBSR pat;             // pattern (values empty)
    DBuf<int> rowptr, cols;
    DBuf<double> val;        // nnzb x br x bc row-major
    // the rows' column indices at a fixed stride ellw (-1 past a row's
    // blocks), for kernels that load them without rowptr (build_ell)
    DBuf<int> ell;
    int ellw = 0;
    void upload_pattern(const HostBSR& h, hipStream_t s);
    void build_ell(hipStream_t s);
    long nnzb() const { return pat.nnzb(); }
};

struct QuadtreeH {
    Ctx* ctx;
    Quadtree qt;
};

struct ProblemH {
    Ctx* ctx;
    const Quadtree* qt = nullptr;      // quadtree problems: kept alive by the caller (as the reference's raw pointers)
    std::unique_ptr<LdgOps> host;      // LdgHost (quadtree) or SimplexLdg
    DevOp M, T, G[3], A;
    DBuf<double> Lb;                   // boundary lifting blocks (npl x nq per boundary face)
    DBuf<int> bleft, btype;            // per boundary face: left element, bc type
    DBuf<double> bnormal;              // per boundary face: outward unit normal (N)
    DBuf<int> bf_off, bf_list;         // boundary faces of each element (CSR)
    DBuf<double> Gpan[3];              // G_d in the solve (panel) layout
    DBuf<double> Mpan;                 // M in the solve (panel) layout (when M does not pool)
    DBuf<double> Mpool, Gpool[3];      // distinct blocks of M / G_d, row-padded pool layout (when they repeat)
    DBuf<int> Mpell, Gpell[3];         // their (column, pool index) pairs at the ELL stride
    DBuf<double> Gl1[3];               // G_d lambda of one column (adjoint_maps)
    DevOp GT[3];                       // G_d^T, values in the panel layout
    int N = 2, b = 0, ne = 0, nqb = 0;
    long ndof = 0;
    double setup_host_s = 0, setup_dev_s = 0;
    double asm_ms = 0, asm_bytes = 0;   // the assembly kernel: duration and algorithmic HBM bytes
    LdgHost* cart() const { return dynamic_cast<LdgHost*>(host.get()); }
    SimplexLdg* simplex() const { return dynamic_cast<SimplexLdg*>(host.get()); }
    void build(Ctx* c, const Quadtree* q, int p, const int* bct, double tau0, double tauD);
    void build_simplex(Ctx* c, int N, int np, const double* pts, int nt, const int* tris, int nbnd, const int* vmask,
                       int p, const int* bct, double tau0, double tauD);
    void upload();                     // device copies of the host operators, A = T + sum G^T M G
    // b = M f + B_g g on the device, k columns (f or g may be null)
    void rhs_device(const double* f, const double* g, int k, double* bout);
    // dL/df = M lam, dL/dg = B_g^T lam (either output may be null)
    void adjoint_maps(const double* lam, int k, double* dLdf, double* dLdg);
};

struct MGParams {
    int relaxation = 1;
    double omega = 0.8;
    int coarse = 0;
    int npre = 3, npost = 3;
    int restriction = 1;
    int precision = 0;   // 0: FP64 only, 1: also FP32 copies for the mixed V-cycle
    int coarse_dense_max = 0;   // collapse the V-cycle below the first level with at most this many dofs (0: off)
    int pool_off = 0;           // 1: never pool repeated blocks for the single-RHS kernels
    int row_tiles = 0;          // row-tile kernels on pooled levels: 0 by size, 1 every level, 2 never
};

template <typename T>
struct LevelVecs {
    DBuf<T> x, b, r, xalt;
};

struct Level {
    int nb = 0, B = 0;
    DevOp A;                            // row-major FP64 values (setup / export)
    DevOp M, Tp, G[3];                  // RATRAT operators (when present)
    bool has_ops = false;
    DBuf<double> A64, D64, Dinv_rm;     // panel layouts and row-major inverse
    DBuf<float> A32, D32;
    DBuf<double> Minv;                  // inverse coarse mass blocks (RATRAT)
    // Gauss-Seidel schedule: rows grouped by colour (multicolour) or by
    // wavefront of the lexicographic order (forward; the reverse sweep has its
    // own wavefronts in rcolor*)
    // (*_desc: per row in that order {row, rowptr[row], rowptr[row+1],
    // position of the diagonal block}; *_gcols: per row in that order the
    // columns of its off-diagonal blocks at stride gw, -1 past the last; both
    // read by the single-RHS GS kernels)
    std::vector<int> color, color_off;
    int ncolor = 0, gw = 1;
    DBuf<int> color_rows, color_dev, color_desc, color_gcols;
    std::vector<int> rcolor, rcolor_off;
    int rncolor = 0;
    DBuf<int> rcolor_rows, rcolor_dev, rcolor_desc, rcolor_gcols;
    // multicolour GS: the last two colours in one launch (pair_nwg > 0):
    // per workgroup {s, m, e} into pair_desc / pair_gcols, see dev::Ops::gs_pair
    int pair_nwg = 0;
    DBuf<int> pair_wg, pair_desc, pair_gcols;
    // pooled operator of the single-RHS and batched kernels (npool > 0): the distinct
    // blocks of A and of D^-1 (quadtree operators repeat a few hundred
    // blocks), panel layouts; pell: A.ell as (column, pool index) pairs at
    // stride 2 * A.ellw; pdidx: each row's D^-1 pool index.  The GS tables
    // above then hold pairs and desc.w = D^-1 pool index (Ops::gs_color).
    int npool = 0, ndpool = 0;
    DBuf<double> P64, DP64;
    DBuf<float> P32, DP32;
    // the same pools with columns padded to dev::pool_pad, read by the
    // single-RHS kernels (one 16-byte vector per load whatever B is)
    DBuf<double> R64, DR64;
    DBuf<float> R32, DR32;
    DBuf<int> pell, pdidx;
    std::vector<int> hpidx, hdidx;      // host copies of the pool indices (traffic accounting)
    // row tiles of a pooled level (dev::tile_gs / tile_apply): rows with the
    // same block signature in tiles of 16, per colour for GS (tile_off[c] ..
    // tile_off[c + 1]) and over all rows for the operator application;
    // use_tiles[precision] (0 FP64, 1 FP32) bit 0: GS, bit 1: application,
    // set at setup from the level's tile count (MGH::select_tiles)
    std::vector<int> tile_off;
    DBuf<int> t_row, t_col, t_pid, t_meta;
    int a_ntiles = 0, t_W = 0;            // application tiles; table width (4, 8 or 16)
    DBuf<double> TA64, TD64;              // A and D^-1 pools in the tile layout
    DBuf<float> TA32, TD32;
    DBuf<int> a_row, a_col, a_pid, a_meta;
    int use_tiles[2] = {0, 0};
    LevelVecs<double> v64;
    LevelVecs<float> v32;
    // interpolation to the next coarser level
    Interp I;
    DBuf<int> parent, type, kid_off, kids;
    DBuf<double> pool64;
    DBuf<float> pool32;
    // quadtree interpolation (dev::interp_lat): (child, block) pairs per
    // coarse row and the pool and its transposes in the padded panel layout
    // (dev::to_rowpad_*), read by Ops::restrict_lat / prolong_lat instead of
    // the gathers above
    bool ilat = false;
    DBuf<int> kpair;
    DBuf<double> ipool64, ipoolT64;
    DBuf<float> ipool32, ipoolT32;
    DBuf<double> Cinv;                  // coarsest: dense inverse (n x n)
};

struct SolveParams {
    double tol = 1e-8;
    int maxit = 200;
    int criterion = 0;
    int flexible = 0;
    int precision = 0;
    int use_graph = 0;
};

struct SolveStats {
    int iterations = 0, status = 0;
    double relres_true = 0, relres_prec = 0, seconds = 0;
};

struct MGH {
    Ctx* ctx;
    ProblemH* pr;
    MGParams prm;
    std::vector<Level> L;
    // PCG work vectors (FP64) and scalars
    DBuf<double> r, zold, d, q, S, xw, bw;
    DBuf<float> zold32;                // flexible CG, mixed form: previous z (exact in FP32)
    DBuf<double> part;                 // reduction partials (one per workgroup of the largest launch)
    double* S_host = nullptr;          // pinned
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_nodes[2] = {-1, -1};             // nodes of the captured preconditioner graphs
    double setup_host_s = 0, setup_dev_s = 0;
    int cut = -1;                      // first collapsed level (-1: none)
    int submean = 1;                   // coarsest b -= mean(b) (the quadtree Multigrid, multigrid.h:365-368)
    DBuf<double> W;                    // collapsed V-cycle below `cut`, FP64 dense n x n
    DBuf<float> W32;                   // the same in FP32 for the FP32 V-cycle (mixed hierarchies, n % 4 == 0)
    DBuf<double> Wt;                   // the same in the MFMA tile layout (batched form)
    DBuf<float> Wt32;                  // and rounded to FP32 (the batched FP32 V-cycle, mixed hierarchies)

    // batched (k right-hand sides) state
    struct BatchLevel { DBuf<double> x64, b64, r64, alt64; DBuf<float> x32, b32, r32, alt32; };
    int bk = 0, bprec = -1;
    std::vector<BatchLevel> BL;
    DBuf<double> Br, Bzold, Bd, Bq, BS, Bpart, Bgemm;
    DBuf<double> Bdq;                   // partial sums of the fused product and d.q (dev::apply_dot)
    DBuf<float> Bzold32;
    DBuf<int> Bactive;
    hipGraphExec_t bgraph[2] = {nullptr, nullptr};

    // parts: agglomeration partitions (level l -> l + 1) of a simplex problem;
    // null: the quadtree hierarchy of a quadtree problem
    // parts: agglomerated hierarchy (simplex problems).  pdeg: p-levels of
    // these degrees below the problem's (quadtree problems), followed by the
    // h-levels at the last degree when hcoarsen.
    void build(ProblemH* p, const MGParams& prm, const std::vector<std::vector<int>>* parts = nullptr,
               const std::vector<int>* pdeg = nullptr, bool hcoarsen = true);
    int np = 0;                        // p-levels below the fine level
    void build_collapse();
    void ensure_batch(int k, int precision);
    // kernel family of each batched operator: 0 = vector kernels (MOps),
    // nt > 0 = MFMA tiles of 16 x nt columns; per precision (0 FP64, 1 FP32) and level
    struct BTune { int apply = 0, smooth = 0; float apply_ms[4] = {0, 0, 0, 0}, smooth_ms[4] = {0, 0, 0, 0}; };
    std::vector<BTune> btune[2];
    int bforce = -1;                   // -1: autotuned; 0: vector kernels; 1, 2, 4: MFMA tiles where compiled
    void tune_batch(int k, int precision);
    template <typename T> void bapply(int l, int k, const T* x, const T* b, T* y);
    template <typename T> void brelax(int l, int k, int n, bool reverse, bool zero_guess);
    template <typename T> void bvcycle(int k);
    template <typename T> int binterp(int l) const;
    void bprecon(int precision, int k, bool use_graph);     // Bz = V(Br) on k columns
    std::vector<SolveStats> solve_batched(const double* b, double* x, int k, const SolveParams& sp);
    // lambda = A^{-1} dLdu (same MG-PCG), dLdf = M lambda, dLdg = B_g^T lambda
    std::vector<SolveStats> adjoint(const double* dLdu, double* lambda, double* dLdf, double* dLdg, int k, const SolveParams& sp);
    void mrelax(int l, int k, int n, bool reverse, bool zero_guess, double* x, const double* b, double* alt);
    ~MGH();

    // first_done: the sweep's first launch (colour 0 from zero) was done by
    // the restriction into this level (fused_first)
    template <typename T> void relax(int l, int n, bool reverse, bool zero_guess, bool first_done = false);
    bool fused_first(int l) const;
    // one colour c of a single-RHS GS sweep on level l (row tiles where chosen)
    template <typename T> void gs_colour(Level& lv, int c, int cur, int skip, const T* b, T* x);
    // y = A_l x or b - A_l x (+ y . dotv into *dot_out) on a level (row tiles where chosen)
    template <typename T>
    int apply_level(const Level& lv, const T* x, const T* b, T* y, const T* dotv, double* dot_out, double* part);
    void select_tiles();               // per pooled level: row tiles or vector kernels
    // the ordered GS launches of one sweep: fn(rows, nrows, colour array, current colour, skip mode)
    template <typename F> void sweep(const Level& lv, bool reverse, bool first_from_zero, F&& fn) const;
    template <typename T> void vcycle_from(int l, bool x0_zero = true);
    template <typename T> LevelVecs<T>& vecs(int l);
    void precon(int precision, bool use_graph);           // z = V(r)
    void precon_enqueue(int precision);
    SolveStats solve(const double* b, double* x, const SolveParams& sp);   // device, one RHS
    void spmv(int l, int precision, const void* x, void* y);
    void smooth(int l, int precision, int n, bool reverse, const void* b, void* x);
    void vcycle(int precision, const double* b, double* x);
    void vcycles(int precision, const double* b, double* x, int ncycle, double tol, double* rnorm, int* cycles);
    // V[:, i] = V(e_i) from zero for every fine dof (host, row-major ndof x ndof)
    void vcycle_matrix(int precision, double* V);
    void profile_relax(int l, int precision, int nrep, double* ms, double* bytes, int* launches, int part = 0);
    // launches and algorithmic bytes of one preconditioner V-cycle (precision 0 FP64, 1 FP32),
    // and of the rest of one CG iteration (operator product, updates, reductions)
    void vcycle_cost(int precision, int* launches, double* bytes);
    void iteration_cost(int precision, bool flexible, int* launches, double* bytes);
    void profile_spmv(int l, int precision, int nrep, double* ms, double* bytes);
};

}  // namespace mgldg
