// kernels.hpp — host-callable launchers of the HIP kernels in kernels.hip.
// All pointers are device pointers; every launcher enqueues on `s` and
// returns without synchronising.
#pragma once

#include <hip/hip_runtime.h>

namespace mgldg {
namespace dev {

// `cur` of the first launch of a Gauss-Seidel sweep from x = 0 (skip 1:
// forward, skip 2: reverse): no neighbour of its rows has been written yet,
// and the colour tests of the skip modes read none of them
constexpr int kFirstFwd = -1;
// pooled GS tables: (column, pool index) pairs whose pool index carries
// kPidLower when the column's colour precedes the row's in the sweep order
constexpr int kPidMask = (1 << 30) - 1;
constexpr int kPidLower = 1 << 30;
constexpr int kFirstRev = 1 << 30;

// ---------------------------------------------------------------- setup
// out[o] = sum_{t in [off[o], off[o+1])} opL(L[l[t]]) * X[x[t]] * R[r[t]]
// (row-major FP64 blocks; l/r = -1 means identity; opL = transpose if transL).
// Block shapes: L m x p (transL) or p x m, X m x n, R n x q, out p x q.
// If addpool != nullptr, out[o] starts from addpool[addslot[o]] (addslot -1 = zero).
void triple(hipStream_t s, int nout, int p, int q, int m, int n, int transL,
            const int* off, const int* l, const int* x, const int* r,
            const double* Lp, const double* Xp, const double* Rp,
            const double* addpool, const int* addslot, double* out);
// batched inverse of nblk row-major b x b FP64 blocks (Gauss-Jordan, partial pivoting)
void inverse(hipStream_t s, int nblk, int b, const double* in, const int* slot, double* out);
// row-major FP64 blocks -> panel layout of float or double (see kernels.hip)
// pooled blocks for the single-RHS kernels: the panel layout with the columns
// padded with zeros to pool_pad(sizeof(T), b) (a multiple of 16 bytes)
int pool_pad(int tsize, int b);
void to_rowpad_f64(hipStream_t s, long nblk, int b, const double* in, double* out);
void to_rowpad_f32(hipStream_t s, long nblk, int b, const double* in, float* out);
void to_panel_f64(hipStream_t s, long nblk, int b, const double* in, double* out);
void to_panel_f32(hipStream_t s, long nblk, int b, const double* in, float* out);
// gather diagonal blocks: out[i] = in[diag[i]]
void gather_blocks(hipStream_t s, int n, int bs, const int* idx, const double* in, double* out);

// ---------------------------------------------------------------- LDG assembly (assemble.hip)
// The device half of LDGPoisson::discretize: one thread per entry of the
// block rows of one element, walking the element's faces (the host plan,
// AsmPlan in host.hpp, describes every array).  Writes M, T and G_d.
struct AsmRows {
    int ne = 0, N = 2, b = 0, nfl = 0, gshared = 1;
    double tau0 = 0, tauD = 0;
    const double *mscale = nullptr, *iscale = nullptr, *C = nullptr;
    const int *ef_off = nullptr, *ef = nullptr, *ef_tslot = nullptr, *ef_gslot = nullptr, *ef_first = nullptr;
    const int *tdiag = nullptr, *gdiag = nullptr;
    const int *fkind = nullptr, *fca = nullptr, *fcb = nullptr;
    const double *fscale = nullptr, *fnormal = nullptr;
    const double *emass = nullptr, *einv = nullptr, *ediff = nullptr, *fmass = nullptr, *pool = nullptr;
    const int *slice = nullptr, *islice = nullptr;
    double *M = nullptr, *T = nullptr, *G0 = nullptr, *G1 = nullptr, *G2 = nullptr;
};
void asm_rows(hipStream_t s, const AsmRows& a);
// LL, RR, RL (3 b x b blocks per face, into pool) of nface non-conforming
// tensor faces by quadrature: pts 2 x nq x N reference points per face (left
// element, then right), wts nq weights per face, 1-D Lagrange nodes and
// barycentric denominators of the P-point basis
void asm_quad(hipStream_t s, int nface, int N, int P, int nq, int b, const double* pts, const double* wts,
              const double* lnodes, const double* ldenom, double* pool);
// boundary lifts Lb(k) = bscale[k] lbref[bcode[k]] (b x nq per boundary face)
void asm_lift(hipStream_t s, int nbf, int b, int nq, const int* bcode, const double* bscale, const double* lbref, double* Lb);
// out block o = transpose of in block src[o] (row-major b x b)
void transpose_blocks(hipStream_t s, long nblk, int b, const int* src, const double* in, double* out);
// pooling of repeated blocks: a 64-bit hash of each row-major FP64 block's
// bit pattern, and bad[k] = 1 unless block k equals block rep[pidx[k]]
// bit for bit
void block_hash(hipStream_t s, long nblk, int bs, const double* in, unsigned long long* out);
void block_verify(hipStream_t s, long nblk, int bs, const double* in, const int* pidx, const int* rep, int* bad);

// block rows one 256-thread workgroup of the single-RHS row kernels handles per pass
int rows_per_workgroup(int B);
// restriction / prolongation with square B x B blocks and at most max_kids
// children per coarse row have the table-driven forms (Ops::restrict_lat)
bool interp_lat(int B, int max_kids);
constexpr int kKidWidth = 4;   // kpair entries per coarse row

// ---------------------------------------------------------------- solve (T = float | double)
template <typename T>
struct Ops {
    // y = A x (b == nullptr) or y = b - A x ; if dot_out: y.dotv is reduced
    // into *dot_out (FP64, deterministic; part holds one partial per workgroup)
    // (ell: the rows' column indices at stride W, -1 past a row's blocks)
    // pool != 0: A is a pool of distinct blocks and ell holds (column, pool
    // index) pairs at stride 2W (the other launchers below alike)
    // (with dotv and part, the y.dotv partials go to part, one per workgroup;
    // with dot_out they are also summed into *dot_out.  Returns the number of
    // workgroups, i.e. of partials.)
    static int apply(hipStream_t s, int B, int nrows, const int* rowptr, const int* ell, int W, const T* A,
                     const T* x, const T* b, T* y, const T* dotv, double* dot_out, double* part, int pool = 0);
    // damped block Gauss-Seidel on the rows of colour `cur` (multigrid.h:240-249).
    // desc: 4 ints per row of the colour, {row, rowptr[row], rowptr[row+1],
    // position of the diagonal block}; gcols: per row W columns of its
    // off-diagonal blocks (-1 past the last).  skip = 1 (2): first forward
    // (reverse) sweep from x = 0, where the colours after `cur` in sweep order
    // are still zero and are not read.  pool != 0: A and Dinv are pools,
    // gcols holds (column, pool index) pairs and desc.w the D^-1 pool index.
    static void gs_color(hipStream_t s, int B, int nrows, const int* desc, const int* gcols, int W, const T* A,
                         const T* Dinv, const T* b, T* x, T omega, const int* color, int cur, int skip, int pool = 0);
    // the last two colours of a sweep in one launch of nwg workgroups:
    // workgroup w does desc ranges [s, m) then [m, e) ({s, m, e} = wg[3w..]),
    // the reverse for a backward sweep, with a workgroup barrier between;
    // cur0 / cur1 = colour of the first / second range processed
    static void gs_pair(hipStream_t s, int B, int nwg, const int* wg, int reverse, const int* desc, const int* gcols,
                        int W, const T* A, const T* Dinv, const T* b, T* x, T omega, const int* color, int cur0,
                        int cur1, int skip, int pool = 0);
    // damped block Jacobi, x_out = x + omega Dinv (b - A x) (multigrid.h:223-229)
    // (didx != nullptr: A and Dinv are pools, ell pairs, didx the rows' D^-1 pool indices)
    static void jacobi(hipStream_t s, int B, int nrows, const int* rowptr, const int* ell, int W, const T* A,
                       const T* Dinv, const T* b, const T* x, T* xout, T omega, const int* didx = nullptr);
    // the first Jacobi sweep from x = 0: x_out = omega Dinv b (didx as in jacobi)
    static void jacobi0(hipStream_t s, int B, int nrows, const T* Dinv, const T* b, T* xout, T omega,
                        const int* didx = nullptr);
    // b_c = T^T r (multigrid.h:414-418) as a gather over children; T blocks
    // are BF x BC (fine x coarse basis)
    static void restrict_(hipStream_t s, int BF, int BC, int nc, const int* kid_off, const int* kids, const int* type,
                          const T* pool, const T* r, T* bc);
    // x_f += T x_c (multigrid.h:421-425)
    static void prolong(hipStream_t s, int BF, int BC, int nf, const int* parent, const int* type, const T* pool,
                        const T* xc, T* xf);
    // the same two on quadtree levels (interp_lat), both over the nc coarse
    // rows: kpair = (child, block) pairs, kKidWidth per coarse row, -1
    // padded; P / PT = the pool's blocks / their transposes in the padded
    // panel layout of T (panel_pool)
    // With color != nullptr it also does the coarse level's first GS launch
    // of a sweep from zero: x_I = omega D^-1 b_I on the rows of colour 0
    // (didx: the rows' D^-1 pool indices, Dinv: the padded D^-1 pool)
    static void restrict_lat(hipStream_t s, int B, int nc, const int* kpair, const T* PT, const T* r, T* bc,
                             const int* color = nullptr, const int* didx = nullptr, const T* Dinv = nullptr,
                             T* x = nullptr, T omega = 0);
    static void prolong_lat(hipStream_t s, int B, int nc, const int* kpair, const T* P, const T* xc, T* xf);
    // coarsest level: b -= mean(b); x = Cinv b, Cinv FP64 dense n x n (multigrid.h:365-368)
    static void coarse(hipStream_t s, int n, const double* Cinv, const T* b, T* x, int subtract_mean);
    static void submean(hipStream_t s, long n, T* b);   // b -= mean(b), one workgroup
    static void fill(hipStream_t s, long n, T v, T* x);
    static void copy(hipStream_t s, long n, const T* a, T* b);
};

// ---------------------------------------------------------------- k right-hand sides
// Vectors hold k columns, [row][B][k] (right-hand side fastest).  One thread
// owns one (block row, column) pair and keeps the B entries in registers.
template <typename T>
struct MOps {
    static void apply(hipStream_t s, int B, int k, int nrows, const int* rowptr, const int* cols, const T* A,
                      const T* x, const T* b, T* y);
    static void gs_color(hipStream_t s, int B, int k, int nrows, const int* rows, const int* rowptr, const int* cols,
                         const T* A, const T* Dinv, const T* b, T* x, T omega, const int* color, int cur, int skip);
    static void jacobi(hipStream_t s, int B, int k, int nrows, const int* rowptr, const int* cols, const T* A,
                       const T* Dinv, const T* b, const T* x, T* xout, T omega);
    // the first sweep from x = 0: xout = omega Dinv b (Dinv: one block per row)
    static void jacobi0(hipStream_t s, int B, int k, int nrows, const T* Dinv, const T* b, T* xout, T omega);
    static void restrict_(hipStream_t s, int BF, int BC, int k, int nc, const int* kid_off, const int* kids,
                          const int* type, const T* pool, const T* r, T* bc);
    static void prolong(hipStream_t s, int BF, int BC, int k, int nf, const int* parent, const int* type,
                        const T* pool, const T* xc, T* xf);
    // per column: b -= mean(b) (optional) then x = Cinv b (Cinv FP64 n x n); Cinv null: only the mean
    static void coarse(hipStream_t s, int n, int k, const double* Cinv, T* b, T* x, int subtract_mean);
};
// The same operators on the matrix cores (mfma.hip): one wavefront per block
// row and nt tiles of 16 columns, blocks padded to 16 x 16 (B <= 16) or
// 32 x 32 (B = 27) in registers.  Compiled for B in {6, 8, 9, 10, 16} with
// nt in {1, 2, 4} and B = 27 with nt in {1, 2}.
bool mf_supported(int B);
int mf_max_tiles(int B);
template <typename T>
struct MF {
    // The single-RHS kernels' row tables: ell (every block of a row at
    // stride W, -1 past the last) with rowptr for the operator and Jacobi;
    // per GS launch the colour's descriptors {row, rowptr[row],
    // rowptr[row + 1], diagonal position} and off-diagonal columns at stride
    // gw.  pool (pooled operator): the tables hold (column, pool index) pairs
    // into the pool A (GS: with kPidLower), descriptors and didx the D^-1
    // pool index into Dinv; otherwise A and Dinv hold every block.
    static void apply(hipStream_t s, int B, int nt, int k, int nrows, const int* rowptr, const int* ell, int W,
                      const T* A, const T* x, const T* b, T* y, bool pool);
    static void gs_color(hipStream_t s, int B, int nt, int k, int nrows, const int* desc, const int* gcols, int gw,
                         const T* A, const T* Dinv, const T* b, T* x, T omega, const int* color, int cur, int skip,
                         bool pool);
    static void jacobi(hipStream_t s, int B, int nt, int k, int nrows, const int* rowptr, const int* ell, int W,
                       const T* A, const T* Dinv, const T* b, const T* x, T* xout, T omega, const int* didx);
    // quadtree interpolation over the nc coarse rows (kpair: (child, block
    // type) pairs, kKidWidth per coarse row, -1 past the last; pool: the
    // row-major B x B interpolation blocks); mf_interp_supported(B):
    // prolongation xf += T xc, and the restriction of the fine residual
    // b - A x without storing it (A as for apply: rowptr / ell / W, apool:
    // pooled)
    static void prolong(hipStream_t s, int B, int nt, int k, int nc, const int* kpair, const T* pool, const T* xc, T* xf);
    static void resrestrict(hipStream_t s, int B, int nt, int k, int nc, const int* kpair, const T* pool, const int* rowptr,
                            const int* ell, int W, const T* A, bool apool, const T* x, const T* b, T* bc);
};
bool mf_interp_supported(int B);
// Y = W X on the matrix cores (FP64 16x16x4) for k columns: W in the 16 x 4
// tile layout made by gemm_tiles (gemm_tiles_size(n) doubles), the K range
// split over slices summed in order (part: gemm_part_size(n, k) doubles)
int gemm_tiles_size(int n);
template <typename TW>
void gemm_tiles(hipStream_t s, int n, const double* W, TW* Wt);
int gemm_slices(int n, int k);
long gemm_part_size(int n, int k);
// (Wt of type TW: FP64 tiles, or FP32 tiles on FP32 MFMAs for T = float)
template <typename T, typename TW>
void gemm_mf(hipStream_t s, int n, int k, const TW* Wt, const T* x, T* y, double* part);
// Y = W X for k columns ([n][k] layout), W FP64 n x n row-major
template <typename T>
void gemm_f64(hipStream_t s, int n, int k, const double* W, const T* x, T* y);
// y = W x with W FP64 n x n row-major (the collapsed coarse V-cycle); x, y of type T
template <typename T>
void gemv_f64(hipStream_t s, int n, const double* W, const T* x, T* y);
// y = W x, W FP32 row-major n x n (n a multiple of 4), FP64 sums
void gemv_w32(hipStream_t s, int n, const float* W, const float* x, float* y);
// identity columns: X[i][c] = (i == c), n x n
void eye(hipStream_t s, int n, double* X);

void f64_to_f32(hipStream_t s, long n, const double* a, float* b);
// a[idx[t]] += v for t < n (distinct indices)
void add_at(hipStream_t s, long n, const long* idx, double v, double* a);
void f32_to_f64(hipStream_t s, long n, const float* a, double* b);

// ---------------------------------------------------------------- PCG (FP64)
// Scalar slots in a device array S:
enum { S_DELTA = 0, S_DELTA_OLD, S_DQ, S_ALPHA, S_BETA, S_RR, S_RSOLD, S_DELTA0, S_BB, S_COUNT };
// a.b reduced into *out
void dot(hipStream_t s, long n, const double* a, const double* b, double* out, double* part);
// The V-cycle's level-0 vectors have the preconditioner's precision T: v
// (its right-hand side) and zs (its result, z = (double) zs exactly).
// The second pass of these reductions is done by the consumer: pcg_xr /
// pcg_x sum the product's d.q partials (dqp, ndq: Ops::apply / tile_apply
// with dotv = d and no dot_out), pcg_d sums pcg_rz's.
// S[DQ] = d.q ; x += alpha d ; r -= alpha q ; v = (T) r      (alpha = S[DELTA]/S[DQ])
template <typename T>
void pcg_xr(hipStream_t s, long n, double* S, const double* d, const double* q, double* x, double* r, T* v,
            const double* dqp, int ndq);
// S[DQ] = d.q ; x += alpha d only
void pcg_x(hipStream_t s, long n, double* S, const double* d, double* x, const double* dqp, int ndq);
// S[DELTA_OLD] = S[DELTA]; r.z, r.zold (read_old), r.r partials into part
// (returns their count); then zold = zs (zold may be null: not kept)
template <typename T>
int pcg_rz(hipStream_t s, long n, double* S, const double* r, const T* zs, T* zold, bool read_old, double* part);
// S[DELTA] = r.z, S[RSOLD] = r.zold, S[RR] = r.r, beta (flexible: with r.zold)
// from pcg_rz's nrz partials rzp; d = z + beta d   (init: d = z)
template <typename T>
void pcg_d(hipStream_t s, long n, double* S, const T* zs, double* d, bool init, const double* rzp, int nrz, bool flexible);
// ---------------------------------------------------------------- batched PCG (k columns, [n][k])
// Per-column scalars live at S[c * S_COUNT + slot]; active[c] = 0 freezes column c.
// per-column a.b into out[c * ostride] (deterministic two-pass)
void coldot(hipStream_t s, long n, int k, const double* a, const double* b, double* out, int ostride, double* part);
// out[c * ostride] = sum over b < nb of part[b * k + c] (fixed order)
void colfinalize(hipStream_t s, const double* part, int nb, int k, double* out, int ostride);
// y = A x on the matrix cores (as MF<double>::apply with nt tiles, b null)
// with the per-column x . y into out[c * ostride]: the application's
// workgroups leave their partial sums in part (apply_dot_part_size doubles),
// then colfinalize
long apply_dot_part_size(int nt, int k, int nrows);
void apply_dot(hipStream_t s, int B, int nt, int k, int nrows, const int* rowptr, const int* ell, int W, const double* A,
               const double* x, double* y, bool pool, double* part, double* out, int ostride);
// a[:, c] -= mean(a[:, c]) for each of the k columns (deterministic); part
// holds mean_zero_part_size(n, k) doubles
// *out = sum of part[0 .. n) (fixed order)
void finalize_sum(hipStream_t s, const double* part, int n, double* out);
// ---------------------------------------------------------------- row tiles of pooled levels (tiles.hip)
// Rows grouped by block signature into tiles of 16 (one wavefront each, MFMA):
// trow[16 t + n] row of tile row n (-1: none), tcol[(16 t + n) W + u] its
// u-th off-diagonal block's column, tpid[t W + u] that block's pool index,
// tmeta[2 t] = {blocks, D^-1 (GS) or diagonal (application) pool index}; W
// = tile_width(most off-diagonal blocks of a row).  At / Dt: the pools in
// the tile layout (tile_pool, tile_block_elems() elements per block).
bool tile_supported(int B, int W);
int tile_width(int W);
int tile_block_elems();
// mode 0: A pool (the A-operand layout), 1: D^-1 pool (the GS kernels' layout)
template <typename T>
void tile_pool(hipStream_t s, long np, int B, const double* src, T* dst, int mode);
template <typename T>
void tile_gs(hipStream_t s, int B, int ntiles, const int* trow, const int* tcol, const int* tpid, const int* tmeta, int W,
             const T* At, const T* Dt, const T* b, T* x, T omega, const int* color, int cur, int skip);
template <typename T>
int tile_apply(hipStream_t s, int B, int ntiles, const int* trow, const int* tcol, const int* tpid, const int* tmeta,
               int W, const T* At, const T* x, const T* b, T* y, const T* dotv, double* dot_out, double* part);
size_t mean_zero_part_size(long n, int k);
void mean_zero(hipStream_t s, long n, int k, double* a, double* part);
// Scratch of the allocation-free mean_zero entry point: a file-scope device
// array in kernels.hip of kScratchDoubles doubles (the per-column partials of
// the reduction's at most 1024 workgroups and the column sums, for up to 256
// columns; 2.1 MB), which each device holds from the library's load to the
// end of the process.  scratch_address(device) returns that device's copy,
// resolved once per device with hipGetSymbolAddress while the device is
// current (the array's name is not a device address in host code).
constexpr size_t kScratchDoubles = (size_t)1024 * 256 + 256;
double* scratch_address(int device);
template <typename T>
void bpcg_xr(hipStream_t s, long n, int k, double* S, const int* active, const double* d, const double* q, double* x,
             double* r, T* v, double* part);
void bpcg_x(hipStream_t s, long n, int k, const double* S, const int* active, const double* d, double* x);
template <typename T>
void bpcg_rz(hipStream_t s, long n, int k, double* S, const int* active, const double* r, const T* zs, T* zold,
             bool read_old, double* part);
template <typename T>
void bpcg_d(hipStream_t s, long n, int k, const double* S, const int* active, const T* zs, double* d, bool init);
// ---------------------------------------------------------------- boundary data map B_g (ldgpoisson.tpp:103-161)
// Forward gather per element e (its boundary faces bf_off[e]..bf_off[e+1]):
//   r_e  = sum_f w_f L_f g_f     (w_f = tauD for Dirichlet, 1 for Neumann)
//   J_d,e = sum_{Dirichlet f} n_{f,d} L_f g_f
// L_f is npl x nq row-major; g [nbf*nq][k]; r, J [ndof][k] (J: N arrays);
// bnormal nbf x N outward unit normals.
void bc_forward(hipStream_t s, int ne, int B, int nq, int k, int N, const int* bf_off, const int* bf_list,
                const int* btype, const double* bnormal, double tauD, const double* Lb,
                const double* g, double* r, double* J0, double* J1, double* J2);
// Adjoint per boundary face f, point q: dg = L_f^T v_f with
//   v_f = tauD lam_l - sum_d n_{f,d} (G_d lam)_l (Dirichlet), lam_l (Neumann).
// Glam: N arrays [ndof][k] holding G_d lam.
void bc_adjoint(hipStream_t s, int nbf, int B, int nq, int k, int N, const int* bleft, const int* btype,
                const double* bnormal, double tauD, const double* Lb, const double* lam, const double* G0,
                const double* G1, const double* G2, double* dg);
// y = a + alpha b (k columns flattened)
void axpby(hipStream_t s, long n, double a_scale, const double* a, double b_scale, const double* b, double* y);
// M-block-diagonal apply y = M x (block size B, row-major FP64 blocks)
void blockdiag(hipStream_t s, int nb, int B, const double* Mblk, const double* x, double* y, int nrhs);

}  // namespace dev
}  // namespace mgldg
