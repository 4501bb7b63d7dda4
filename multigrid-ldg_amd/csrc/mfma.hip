// mfma.hip — k right-hand sides on the matrix cores.
//
// For a block row i the batched operators are small dense products
//   Y_i = sum_j A_ij X_j     (A_ij: B x B, X_j: B x k, k columns contiguous)
// One wavefront owns one block row and NT tiles of 16 columns.  Each block is
// a product of RT = ceil(B / 16) row tiles of S = ceil(B / 4) K steps of
// 16 x 16 x 4 MFMAs (v_mfma_f32_16x16x4f32 / _f64_16x16x4f64), blocks of
// 6..32 rows padded with zeros in registers.  Lane l = 16 q + n (q = l / 16,
// n = l % 16) supplies, for K step s and row tile rt, A[16 rt + n][S q + s]
// and, for column tile t, X[S q + s][col0 + NT n + t]: block column S q + s
// (not 4 s + q) so that a lane's S columns of an A row are contiguous in the
// panel layout, and output column col0 + NT n + t (not col0 + 16 t + n) so
// that a lane's NT columns of X, b, x and y are contiguous; both come as
// vector loads.  The accumulator lane layout is the instruction's (f32:
// register e of lane l holds row 4 q + e; f64: row 4 e + q; tile column n).
// The smoothers' D^-1 product takes the residual tile in the accumulators as
// its B operand, K steps ordered by accumulator register (no LDS, no barrier).
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "kernels.hpp"
#include "devutil.hpp"

namespace mgldg {
namespace dev {

namespace {

__host__ __device__ constexpr int pv(int tsize, int B)
{
    return tsize == 8 ? (B % 2 == 0 ? 2 : 1) : (B % 4 == 0 ? 4 : (B % 2 == 0 ? 2 : 1));
}

// element (r, c) of a block in the solve (panel) layout
template <typename T, int B>
__device__ __forceinline__ T panel_at(const T* blk, int r, int c)
{
    constexpr int V = pv(sizeof(T), B);
    return blk[((c / V) * B + r) * V + (c % V)];
}

template <typename T> struct Mf;
template <> struct Mf<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lane, int i) { return 4 * (lane >> 4) + i; }
};
template <> struct Mf<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lane, int i) { return 4 * i + (lane >> 4); }
};

// N contiguous elements at p (N * sizeof(T)-aligned) into o[0 .. N)
template <typename T, int N>
__device__ __forceinline__ void mf_ld(T* o, const T* __restrict__ p)
{
    if constexpr (N == 1) {
        o[0] = p[0];
    } else {
        typedef T VN __attribute__((ext_vector_type(N)));
        const VN v = *reinterpret_cast<const VN*>(p);
#pragma unroll
        for (int e = 0; e < N; ++e) o[e] = v[e];
    }
}
// the same, nontemporal (a stream whose cache lines should not evict the
// operands that are read again)
template <typename T, int N>
__device__ __forceinline__ void mf_ld_nt(T* o, const T* __restrict__ p)
{
    if constexpr (N == 1) {
        o[0] = __builtin_nontemporal_load(p);
    } else {
        typedef T VN __attribute__((ext_vector_type(N)));
        const VN v = __builtin_nontemporal_load(reinterpret_cast<const VN*>(p));
#pragma unroll
        for (int e = 0; e < N; ++e) o[e] = v[e];
    }
}
template <typename T, int N>
__device__ __forceinline__ void mf_st(T* p, const T* v)
{
    if constexpr (N == 1) {
        p[0] = v[0];
    } else {
        typedef T VN __attribute__((ext_vector_type(N)));
        VN w;
#pragma unroll
        for (int e = 0; e < N; ++e) w[e] = v[e];
        *reinterpret_cast<VN*>(p) = w;
    }
}

// Operands of one block product A_blk X_j for this lane: per K step s the A
// element of each row tile and the X element of each column tile (rows padded
// to 16 RT, K to 4 S, with zeros).
template <typename T, int B, int NT>
struct BlkOps {
    static constexpr int RT = (B + 15) / 16, S = (B + 3) / 4;
    T a[S][RT];
    T xb[S][NT];
};

// Blocks whose operands are loaded before any of their products (one memory
// round trip per U blocks), sized to keep the operands within 64 VGPRs: the
// wavefronts are latency-bound, and more of them per SIMD beat more blocks in
// flight per wavefront (a 128-register budget was 5% slower in FP64, and up
// to 8 blocks in flight 2-7% slower everywhere).
template <typename T, int B, int NT>
__host__ __device__ constexpr int mf_unroll()
{
    constexpr int per = ((B + 3) / 4) * ((B + 15) / 16 + NT) * (int)(sizeof(T) / 4);
    return per * 4 <= 64 ? 4 : (per * 3 <= 64 ? 3 : (per * 2 <= 64 ? 2 : 1));
}

// The lane-dependent, block-independent parts of the operand offsets
// (elements).  X[S q + s][gc .. gc + NT) of a block column sits at xl + s k,
// gc = col0 + NT n: a group wholly at or past k (k is a multiple of NT) is
// read from columns k - NT .. k instead, in bounds, and its products only
// reach output columns that are never stored.  When V divides S, the panel
// element (16 rt + n, S q + s) of an A block sits at
// al + (s / V) B V + 16 rt V + s % V (the lane's V columns s, s+1, .. of a
// group are contiguous); otherwise it is addressed element by element.  The
// block bases themselves are wave-uniform (scalar registers).
template <typename T, int B, int NT>
struct LaneOff {
    static constexpr int S = (B + 3) / 4, V = pv(sizeof(T), B);
    unsigned xl, al;
    int q, n, gc;
    __device__ __forceinline__ LaneOff(int lane, int k, int col0)
    {
        n = lane & 15;
        q = lane >> 4;
        gc = col0 + NT * n;
        xl = (unsigned)(S * q * k + min(gc, k - NT));
        al = (unsigned)((S % V == 0 ? (S * q / V) * B * V : 0) + n * V);
    }
    // lane's K index S q + s inside the block?  (compile-time true but where q = 3 can pass B)
    __device__ __forceinline__ bool kin(int s) const { return 3 * S + s < B || S * q + s < B; }
};

template <typename T, int B, int NT>
__device__ __forceinline__ void blk_load(BlkOps<T, B, NT>& o, const T* __restrict__ blk, const T* __restrict__ xj, int k,
                                         const LaneOff<T, B, NT>& lo)
{
    constexpr int S = LaneOff<T, B, NT>::S, V = LaneOff<T, B, NT>::V, RT = BlkOps<T, B, NT>::RT;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (lo.kin(s)) mf_ld<T, NT>(o.xb[s], xj + (size_t)s * k + lo.xl);
        else {
#pragma unroll
            for (int t = 0; t < NT; ++t) o.xb[s][t] = (T)0;
        }
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const bool rin = 16 * rt + 15 < B || 16 * rt + lo.n < B;
        if constexpr (S % V == 0) {
#pragma unroll
            for (int s0 = 0; s0 < S; s0 += V) {
                T v[V];
                if (rin && lo.kin(s0)) mf_ld<T, V>(v, blk + (s0 / V) * B * V + 16 * rt * V + lo.al);   // (B % V == 0: whole group in or out)
                else {
#pragma unroll
                    for (int e = 0; e < V; ++e) v[e] = (T)0;
                }
#pragma unroll
                for (int e = 0; e < V; ++e) o.a[s0 + e][rt] = v[e];
            }
        } else {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int c = S * lo.q + s, r = 16 * rt + lo.n;
                o.a[s][rt] = (rin && lo.kin(s)) ? panel_at<T, B>(blk, r, c) : (T)0;
            }
        }
    }
}

template <typename T, int B, int NT>
__device__ __forceinline__ void blk_mac(typename Mf<T>::v4* acc, const BlkOps<T, B, NT>& o)
{
#pragma unroll
    for (int s = 0; s < BlkOps<T, B, NT>::S; ++s)
#pragma unroll
        for (int rt = 0; rt < BlkOps<T, B, NT>::RT; ++rt)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[rt * NT + t] = Mf<T>::mac(o.a[s][rt], o.xb[s][t], acc[rt * NT + t]);
}

// A row's block positions, one per lane: lane m holds position m0 + m's
// column j (-1: none) and its block's index in A (pooled operator: the pool
// index, on Gauss-Seidel tables with kPidLower when j's colour precedes the
// row's).  The tables are read in the same memory round trip as the row's own
// operands, so the neighbour operands follow one round trip later.
struct RowTab {
    int col, blk;
};

// lane m's entry, position m0 + m of nblk: Gauss-Seidel tables (d = {row,
// rowptr[row], rowptr[row + 1], diagonal position or, pooled, D^-1 pool
// index}; gr: the row's off-diagonal columns, pairs (column, pool index) when
// pooled) ...
template <bool POOL>
__device__ __forceinline__ RowTab gs_tab(const int* __restrict__ gr, int m0, int lane, int nblk, int4 d)
{
    RowTab e{-1, 0};
    const int m = m0 + lane;
    if (m < nblk) {
        if (POOL) {
            const int2 pe = reinterpret_cast<const int2*>(gr)[m];
            e.col = pe.x; e.blk = pe.y;
        } else {
            e.col = gr[m];
            e.blk = blk_at(d.y, m, d.w);
        }
    }
    return e;
}
// ... and the operator's ELL row er (every block of the row in pattern
// order; pairs when pooled), k0 = rowptr[row]
template <bool POOL>
__device__ __forceinline__ RowTab ell_tab(const int* __restrict__ er, int m0, int lane, int nblk, int k0)
{
    RowTab e{-1, 0};
    const int m = m0 + lane;
    if (m < nblk) {
        if (POOL) {
            const int2 pe = reinterpret_cast<const int2*>(er)[m];
            e.col = pe.x; e.blk = pe.y;
        } else {
            e.col = er[m];
            e.blk = k0 + m;
        }
    }
    return e;
}

// acc += sum over the nc positions held by the lanes (one row per wavefront,
// so every branch below is uniform): columns and block indices are read out
// with readlane into scalar registers; the operands of U blocks are in flight
// together.  SKIP: -1 / 0 = every position held; 1 (2) = only neighbours of
// a colour before (after) the current one (first GS sweep from x = 0), from
// the pooled tables' kPidLower bit or, unpooled, from color[j].
template <typename T, int B, int NT, int SKIP, bool POOL>
__device__ __forceinline__ void tab_mac(typename Mf<T>::v4* acc, RowTab e, int nc, const T* __restrict__ A,
                                        const T* __restrict__ x, int k, const LaneOff<T, B, NT>& lo,
                                        const int* __restrict__ color, int cur)
{
    constexpr int U = mf_unroll<T, B, NT>();
    for (int u0 = 0; u0 < nc; u0 += U) {
        BlkOps<T, B, NT> ops[U];
        bool use[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int m = min(u0 + u, 63);
            const int j = __builtin_amdgcn_readlane(e.col, m);
            int blk = __builtin_amdgcn_readlane(e.blk, m);
            bool ok = u0 + u < nc && j >= 0;
            if (POOL) {
                if (SKIP > 0) { const bool lower = (blk & kPidLower) != 0; ok = ok && (SKIP == 1 ? lower : !lower); }
                blk &= kPidMask;
            } else if (SKIP > 0 && ok) {
                const int cj = color[j];
                ok = SKIP == 1 ? cj <= cur : cj >= cur;
            }
            use[u] = ok;
            if (ok) blk_load<T, B, NT>(ops[u], A + (size_t)blk * (B * B), x + (size_t)j * B * k, k, lo);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (use[u]) blk_mac<T, B, NT>(acc, ops[u]);
    }
}

// y = A x (b null) or y = b - A x; the row's ELL entries, rowptr and b_i in
// one round trip, then the blocks.  DOT (y = A x only): also the per-column
// sums of x . y over the workgroup's rows into part[blockIdx.x][0 .. k) (x_i
// loaded in the accumulator layout beside the tables; lanes, then the four
// wavefronts, summed in a fixed order)
template <typename T, int B, int NT, bool POOL, bool DOT>
__global__ __launch_bounds__(256) void mf_apply(int k, int nrows, int ngroups, const int* __restrict__ rowptr,
                                                const int* __restrict__ ell, int W, const T* __restrict__ A,
                                                const T* __restrict__ x, const T* __restrict__ b, T* __restrict__ y,
                                                double* __restrict__ part)
{
    using V4 = typename Mf<T>::v4;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int w = __builtin_amdgcn_readfirstlane(xcd_block() * 4 + wv);   // wave-uniform
    const int i0 = w / ngroups, g = w - i0 * ngroups;
    const bool valid = i0 < nrows;
    if (!DOT && !valid) return;
    const int i = valid ? i0 : 0;
    constexpr int RT = (B + 15) / 16;
    const int col0 = g * 16 * NT;
    const int k0 = valid ? rowptr[i] : 0, nblk = valid ? rowptr[i + 1] - k0 : 0;
    const int* er = ell + (size_t)i * W * (POOL ? 2 : 1);
    RowTab e = ell_tab<POOL>(er, 0, lane, nblk, k0);
    V4 acc[RT * NT];
#pragma unroll
    for (int t = 0; t < RT * NT; ++t) acc[t] = V4{0, 0, 0, 0};
    const int gc = col0 + NT * (lane & 15);   // this lane's NT output columns (k is a multiple of NT)
    // b_i of the residual form (DOT: x_i), in flight with the tables
    const T* own = DOT ? x : b;
    T bt[RT][4][NT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = 16 * rt + Mf<T>::row(lane, q);
            if (own && valid && r < B && gc < k) mf_ld<T, NT>(bt[rt][q], own + ((size_t)i * B + r) * k + gc);
            else {
#pragma unroll
                for (int t = 0; t < NT; ++t) bt[rt][q][t] = (T)0;
            }
        }
    const LaneOff<T, B, NT> lo(lane, k, col0);
    for (int m0 = 0; m0 < nblk; m0 += 64) {
        if (m0 > 0) e = ell_tab<POOL>(er, m0, lane, nblk, k0);
        tab_mac<T, B, NT, -1, POOL>(acc, e, min(64, nblk - m0), A, x, k, lo, nullptr, 0);
    }
    if (valid && gc < k) {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = 16 * rt + Mf<T>::row(lane, q);
                if (r >= B) continue;
                const size_t o = ((size_t)i * B + r) * k + gc;
                T v[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) v[t] = (!DOT && b) ? (T)(bt[rt][q][t] - acc[rt * NT + t][q]) : acc[rt * NT + t][q];
                mf_st<T, NT>(y + o, v);
            }
    }
    if constexpr (DOT) {
        __shared__ double red[4][64];
        double sdot[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) sdot[t] = 0;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (16 * rt + Mf<T>::row(lane, q) < B) {
#pragma unroll
                    for (int t = 0; t < NT; ++t) sdot[t] += (double)bt[rt][q][t] * (double)acc[rt * NT + t][q];
                }
        // the four lanes n, n + 16, n + 32, n + 48 hold the same columns
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            sdot[t] += __shfl_xor(sdot[t], 16, 64);
            sdot[t] += __shfl_xor(sdot[t], 32, 64);
        }
        if (lane < 16) {
#pragma unroll
            for (int t = 0; t < NT; ++t) red[wv][NT * lane + t] = sdot[t];
        }
        __syncthreads();
        const int c = threadIdx.x;
        if (c < k) {
            const int cg = c / (16 * NT), cl = c - cg * 16 * NT;
            const int wb = w - wv;   // the workgroup's first wave
            double tot = 0;
            for (int v = 0; v < 4; ++v)
                if ((wb + v) % ngroups == cg) tot += red[v][cl];
            part[(size_t)blockIdx.x * k + c] = tot;
        }
    }
}

// The operands of a smoother's row that do not depend on its neighbour sum:
// b_i and x_i on this lane's rows and NT columns gc .. gc + NT (zero outside
// the block, past k, or for an invalid wave), and the row's D^-1 entries in
// the A-operand layout.  They are loaded before the neighbour sum, so that
// they are in flight with it instead of after it.  The D^-1 product takes
// the residual tile as it stands in the accumulators: in K step e of K tile
// rt2 a lane supplies accumulator register e, i.e. row 16 rt2 + row(l, e)
// of the residual, so dv[rt2][e][rt] = D^-1[16 rt + l % 16][16 rt2 + row(l, e)]
// (zero outside B x B).
template <typename T, int B, int NT>
struct RowTile {
    static constexpr int RT = (B + 15) / 16, S = (B + 3) / 4;
    T bt[RT][4][NT], xt[RT][4][NT];
    T dv[RT][4][RT];
};

template <typename T, int B, int NT, bool LOADX>
__device__ __forceinline__ void row_tile_load(RowTile<T, B, NT>& o, const T* __restrict__ b, const T* x,
                                              const T* __restrict__ D, int i, int k, int gc, bool valid, int lane)
{
    constexpr int RT = RowTile<T, B, NT>::RT, S = RowTile<T, B, NT>::S;
#pragma unroll
    for (int rr = 0; rr < RT; ++rr)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = 16 * rr + Mf<T>::row(lane, e);
            const bool in = valid && r < B && gc < k;
            const size_t off = ((size_t)i * B + r) * k + gc;
            // b nontemporal in FP64 (measured: the FP64 batch 2% faster, the
            // FP32 one slower)
            if (in) {
                if constexpr (sizeof(T) == 8) mf_ld_nt<T, NT>(o.bt[rr][e], b + off);
                else mf_ld<T, NT>(o.bt[rr][e], b + off);
            }
            if (LOADX && in) mf_ld<T, NT>(o.xt[rr][e], x + off);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if (!in) o.bt[rr][e][t] = (T)0;
                if (!LOADX || !in) o.xt[rr][e][t] = (T)0;
            }
        }
    const int r16 = lane & 15;
#pragma unroll
    for (int rt2 = 0; rt2 < RT; ++rt2)
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                const int c = 16 * rt2 + Mf<T>::row(lane, e), r = 16 * rt + r16;
                o.dv[rt2][e][rt] = (r < B && c < B) ? panel_at<T, B>(D, r, c) : (T)0;
            }
    (void)S;
}

// res = b_i - acc (acc layout; zero where row_tile_load left b zero)
template <typename T, int B, int NT>
__device__ __forceinline__ void residual_tile(typename Mf<T>::v4* res, const typename Mf<T>::v4* acc,
                                              const RowTile<T, B, NT>& o, int k, int gc, bool valid, int lane)
{
    constexpr int RT = (B + 15) / 16;
#pragma unroll
    for (int rr = 0; rr < RT; ++rr)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = 16 * rr + Mf<T>::row(lane, e);
            const bool in = valid && r < B && gc < k;
#pragma unroll
            for (int t = 0; t < NT; ++t) res[rr * NT + t][e] = in ? (T)(o.bt[rr][e][t] - acc[rr * NT + t][e]) : (T)0;
        }
}

// the D^-1 product of a smoother: D_i r with r in the accumulator layout
// (16 RT x 16 NT), which is the B-operand layout of K steps ordered by
// accumulator register (RowTile); no data leaves the lanes
template <typename T, int B, int NT>
__device__ __forceinline__ void dinv_times(typename Mf<T>::v4* out, const typename Mf<T>::v4* res,
                                           const RowTile<T, B, NT>& o)
{
    using V4 = typename Mf<T>::v4;
    constexpr int RT = (B + 15) / 16;
#pragma unroll
    for (int t = 0; t < RT * NT; ++t) out[t] = V4{0, 0, 0, 0};
#pragma unroll
    for (int rt2 = 0; rt2 < RT; ++rt2)
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    out[rt * NT + t] = Mf<T>::mac(o.dv[rt2][e][rt], res[rt2 * NT + t][e], out[rt * NT + t]);
}

// damped block GS on the rows of one colour (multigrid.h:240-249):
// x_i += omega (D_i^-1 (b_i - sum_{j != i} A_ij x_j) - x_i); skip as in the
// vector kernel (first sweep from zero: later colours are not read).  Wave
// (ri, g) reads the colour's descriptor ri and column table, and b_i, x_i in
// one round trip, D^-1 and the neighbour blocks in the next.
template <typename T, int B, int NT, int SKIP, bool POOL>
__global__ __launch_bounds__(256) void mf_gs(int k, int nrows, int ngroups, const int4* __restrict__ desc,
                                             const int* __restrict__ gcols, int gw, const T* __restrict__ A,
                                             const T* __restrict__ Dinv, const T* __restrict__ b, T* x, T omega,
                                             const int* __restrict__ color, int cur)
{
    using V4 = typename Mf<T>::v4;
    constexpr int RT = (B + 15) / 16;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int w = __builtin_amdgcn_readfirstlane(xcd_block() * 4 + wv);   // wave-uniform
    const int ri = w / ngroups, g = w - ri * ngroups;
    if (ri >= nrows) return;
    const int4 d = desc[ri];
    const int i = d.x;
    const bool reads = !reads_none<SKIP>(cur);
    // off-diagonal blocks of the row (every row has its diagonal block)
    const int nblk = reads ? d.z - d.y - (POOL || d.w >= 0 ? 1 : 0) : 0;
    const int* gr = gcols + (size_t)ri * gw * (POOL ? 2 : 1);
    RowTab e = gs_tab<POOL>(gr, 0, lane, nblk, d);
    const int col0 = g * 16 * NT, col_l = lane & 15;
    V4 acc[RT * NT];
#pragma unroll
    for (int t = 0; t < RT * NT; ++t) acc[t] = V4{0, 0, 0, 0};
    const int gc = col0 + NT * col_l;   // this lane's NT columns (k is a multiple of NT)
    RowTile<T, B, NT> rt;
    row_tile_load<T, B, NT, SKIP == 0>(rt, b, x, Dinv + (size_t)(POOL ? d.w : i) * B * B, i, k, gc, true, lane);
    const LaneOff<T, B, NT> lo(lane, k, col0);
    for (int m0 = 0; m0 < nblk; m0 += 64) {
        if (m0 > 0) e = gs_tab<POOL>(gr, m0, lane, nblk, d);
        tab_mac<T, B, NT, SKIP, POOL>(acc, e, min(64, nblk - m0), A, x, k, lo, color, cur);
    }
    V4 res[RT * NT];
    residual_tile<T, B, NT>(res, acc, rt, k, gc, true, lane);
    V4 out[RT * NT];
    dinv_times<T, B, NT>(out, res, rt);
    if (gc >= k) return;
#pragma unroll
    for (int rr = 0; rr < RT; ++rr)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = 16 * rr + Mf<T>::row(lane, q);
            if (r >= B) continue;
            const size_t o = ((size_t)i * B + r) * k + gc;
            T nv[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) nv[t] = rt.xt[rr][q][t] + omega * (out[rr * NT + t][q] - rt.xt[rr][q][t]);   // (SKIP: x_i = 0)
            mf_st<T, NT>(x + o, nv);
        }
}

// damped block Jacobi (multigrid.h:223-229): xout = x + omega D^-1 (b - A x);
// the row's ELL entries, rowptr, D^-1 pool index, b_i and x_i in one round
// trip, D^-1 and the blocks in the next
template <typename T, int B, int NT, bool POOL>
__global__ __launch_bounds__(256) void mf_jacobi(int k, int nrows, int ngroups, const int* __restrict__ rowptr,
                                                 const int* __restrict__ ell, int W, const T* __restrict__ A,
                                                 const T* __restrict__ Dinv, const T* __restrict__ b,
                                                 const T* __restrict__ x, T* __restrict__ xout, T omega,
                                                 const int* __restrict__ didx)
{
    using V4 = typename Mf<T>::v4;
    constexpr int RT = (B + 15) / 16;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int w = __builtin_amdgcn_readfirstlane(xcd_block() * 4 + wv);   // wave-uniform
    const int i = w / ngroups, g = w - i * ngroups;
    if (i >= nrows) return;
    const int k0 = rowptr[i], nblk = rowptr[i + 1] - k0;
    const int* er = ell + (size_t)i * W * (POOL ? 2 : 1);
    RowTab e = ell_tab<POOL>(er, 0, lane, nblk, k0);
    const int col0 = g * 16 * NT, col_l = lane & 15;
    V4 acc[RT * NT];
#pragma unroll
    for (int t = 0; t < RT * NT; ++t) acc[t] = V4{0, 0, 0, 0};
    const int gc = col0 + NT * col_l;   // this lane's NT columns (k is a multiple of NT)
    RowTile<T, B, NT> rt;
    row_tile_load<T, B, NT, true>(rt, b, x, Dinv + (size_t)(POOL ? didx[i] : i) * B * B, i, k, gc, true, lane);
    const LaneOff<T, B, NT> lo(lane, k, col0);
    for (int m0 = 0; m0 < nblk; m0 += 64) {
        if (m0 > 0) e = ell_tab<POOL>(er, m0, lane, nblk, k0);
        tab_mac<T, B, NT, -1, POOL>(acc, e, min(64, nblk - m0), A, x, k, lo, nullptr, 0);
    }
    V4 res[RT * NT];
    residual_tile<T, B, NT>(res, acc, rt, k, gc, true, lane);
    V4 out[RT * NT];
    dinv_times<T, B, NT>(out, res, rt);
    if (gc >= k) return;
#pragma unroll
    for (int rr = 0; rr < RT; ++rr)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = 16 * rr + Mf<T>::row(lane, q);
            if (r >= B) continue;
            const size_t o = ((size_t)i * B + r) * k + gc;
            T nv[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) nv[t] = rt.xt[rr][q][t] + omega * out[rr * NT + t][q];
            mf_st<T, NT>(xout + o, nv);
        }
}

// Quadtree interpolation on k columns (square B x B interpolation blocks
// T_t, row-major in pool, at most kKidWidth children per coarse row): one
// wavefront per coarse row c and column group, its children from the kid
// table kpair ((child, block type) pairs, -1 past the last), so that the
// tables take one round trip and every operand the next.
//   prolongation (correctSolution, multigrid.h:421-426):  xf_f += T_{t(f)} xc_c
//   restriction of the residual: mf_resrestrict below
// Operands in the layout of the block kernels above: A[n][S q + s] per lane,
// X[S q + s][gc .. gc + NT).
template <typename T, int B, int NT>
__device__ __forceinline__ void x_rows(T (&xb)[(B + 3) / 4][NT], const T* __restrict__ xj, int k, const LaneOff<T, B, NT>& lo)
{
#pragma unroll
    for (int s = 0; s < (B + 3) / 4; ++s) {
        if (lo.kin(s)) mf_ld<T, NT>(xb[s], xj + (size_t)s * k + lo.xl);
        else {
#pragma unroll
            for (int t = 0; t < NT; ++t) xb[s][t] = (T)0;
        }
    }
}

// A operand of a row-major block, zero outside B x B
template <typename T, int B>
__device__ __forceinline__ void rm_a(T (&a)[(B + 3) / 4], const T* __restrict__ blk, int lane)
{
    constexpr int S = (B + 3) / 4;
    const int n = lane & 15, q = lane >> 4;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int c = S * q + s;
        a[s] = (n < B && c < B) ? blk[n * B + c] : (T)0;
    }
}

// The residual restricted in one pass: bc_c = sum_f T_{t(f)}^T (b_f - A_f x)
// (computeResidual and restrictResidual, multigrid.h:374-375, 414-419) for
// the children f of coarse row c in kid-table order.  A child's residual
// tile is formed as by mf_apply (its ELL row, its blocks in order) and taken
// from the accumulators as the B operand of the T^T product, K steps ordered
// by accumulator register (as the smoothers' D^-1 product): the fine
// residual is never stored.
template <typename T, int B, int NT, bool POOL>
__global__ __launch_bounds__(256) void mf_resrestrict(int k, int nc, int ngroups, const int* __restrict__ kpair,
                                                      const T* __restrict__ pool, const int* __restrict__ rowptr,
                                                      const int* __restrict__ ell, int W, const T* __restrict__ A,
                                                      const T* __restrict__ x, const T* __restrict__ b,
                                                      T* __restrict__ bc)
{
    using V4 = typename Mf<T>::v4;
    constexpr int KW = kKidWidth;
    static_assert(B <= 16, "one row tile");
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(xcd_block() * 4 + (int)(threadIdx.x >> 6));   // wave-uniform
    const int c = w / ngroups, g = w - c * ngroups;
    if (c >= nc) return;
    const int col0 = g * 16 * NT, n = lane & 15;
    const int gc = col0 + NT * n;
    const int2 kp = lane < KW ? reinterpret_cast<const int2*>(kpair)[(size_t)c * KW + lane] : make_int2(-1, 0);
    const LaneOff<T, B, NT> lo(lane, k, col0);
    V4 accc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) accc[t] = V4{0, 0, 0, 0};
    for (int m = 0; m < KW; ++m) {
        const int f = __builtin_amdgcn_readlane(kp.x, m), ty = __builtin_amdgcn_readlane(kp.y, m);
        if (f < 0) continue;
        const int k0 = rowptr[f], nblk = rowptr[f + 1] - k0;
        const int* er = ell + (size_t)f * W * (POOL ? 2 : 1);
        RowTab e = ell_tab<POOL>(er, 0, lane, nblk, k0);
        T bt[4][NT];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = Mf<T>::row(lane, q);
            if (r < B && gc < k) mf_ld<T, NT>(bt[q], b + ((size_t)f * B + r) * k + gc);
            else {
#pragma unroll
                for (int t = 0; t < NT; ++t) bt[q][t] = (T)0;
            }
        }
        T ta[4];   // T^T[n][row(l, q)] = T[row(l, q)][n]
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = Mf<T>::row(lane, q);
            ta[q] = (n < B && r < B) ? pool[(size_t)ty * B * B + r * B + n] : (T)0;
        }
        V4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = V4{0, 0, 0, 0};
        for (int m0 = 0; m0 < nblk; m0 += 64) {
            if (m0 > 0) e = ell_tab<POOL>(er, m0, lane, nblk, k0);
            tab_mac<T, B, NT, -1, POOL>(acc, e, min(64, nblk - m0), A, x, k, lo, nullptr, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int t = 0; t < NT; ++t) accc[t] = Mf<T>::mac(ta[q], (T)(bt[q][t] - acc[t][q]), accc[t]);
    }
    if (gc >= k) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int rr = Mf<T>::row(lane, q);
        if (rr >= B) continue;
        T v[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) v[t] = accc[t][q];
        mf_st<T, NT>(bc + ((size_t)c * B + rr) * k + gc, v);
    }
}

template <typename T, int B, int NT>
__global__ __launch_bounds__(256) void mf_prolong(int k, int nc, int ngroups, const int* __restrict__ kpair,
                                                  const T* __restrict__ pool, const T* __restrict__ xc, T* __restrict__ xf)
{
    using V4 = typename Mf<T>::v4;
    constexpr int S = (B + 3) / 4, KW = kKidWidth;
    static_assert(B <= 16, "one row tile");
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(xcd_block() * 4 + (int)(threadIdx.x >> 6));   // wave-uniform
    const int c = w / ngroups, g = w - c * ngroups;
    if (c >= nc) return;
    const int col0 = g * 16 * NT;
    const int gc = col0 + NT * (lane & 15);
    const int2 kp = lane < KW ? reinterpret_cast<const int2*>(kpair)[(size_t)c * KW + lane] : make_int2(-1, 0);
    const LaneOff<T, B, NT> lo(lane, k, col0);
    T xb[S][NT];
    x_rows<T, B, NT>(xb, xc + (size_t)c * B * k, k, lo);
#pragma unroll
    for (int m = 0; m < KW; ++m) {
        const int f = __builtin_amdgcn_readlane(kp.x, m), ty = __builtin_amdgcn_readlane(kp.y, m);
        if (f < 0) continue;
        T a[S];
        rm_a<T, B>(a, pool + (size_t)ty * B * B, lane);
        T xo[4][NT];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rr = Mf<T>::row(lane, q);
            if (rr < B && gc < k) mf_ld<T, NT>(xo[q], xf + ((size_t)f * B + rr) * k + gc);
            else {
#pragma unroll
                for (int t = 0; t < NT; ++t) xo[q][t] = (T)0;
            }
        }
        V4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = V4{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = Mf<T>::mac(a[s], xb[s][t], acc[t]);
        if (gc < k) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = Mf<T>::row(lane, q);
                if (rr >= B) continue;
                T v[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) v[t] = xo[q][t] + acc[t][q];
                mf_st<T, NT>(xf + ((size_t)f * B + rr) * k + gc, v);
            }
        }
    }
}

// W (n x n, row-major FP64) -> 16 x 4 tiles in the A-operand lane order:
// Wt[(R * nK + K) * 64 + l] = W[16 R + l % 16][4 K + l / 16], zero-padded
// (FP64, or rounded to FP32 for the FP32 V-cycle)
template <typename TW>
__global__ void k_tile_w(int n, int nR, int nK, const double* __restrict__ W, TW* __restrict__ Wt)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)nR * nK * 64) return;
    const int l = (int)(e % 64);
    const long t = e / 64;
    const int K = (int)(t % nK), R = (int)(t / nK);
    const int r = 16 * R + (l & 15), c = 4 * K + (l >> 4);
    Wt[e] = (r < n && c < n) ? (TW)W[(size_t)r * n + c] : (TW)0;
}

// partial products of Y = W X over K slices, part[(slice * nRpad + row) * k + col].
// A wavefront owns NRW consecutive 16-row tiles of W, NCW consecutive
// 16-column tiles of Y and one K slice: each W operand is loaded once for its
// NCW column tiles and each X operand once for its NRW row tiles, and four K
// steps of operands are loaded before their products, so several loads are in
// flight per wavefront.  W is read once per call and is larger than the
// caches: its loads are nontemporal, which keeps it from evicting the rest of
// the V-cycle's data.  Products accumulate in K order within the slice, on the
// MFMAs of W's type TW (the FP32 V-cycle reads an FP32 copy of W: half the
// bytes, FP32 matrix-core rate); the slices are summed in FP64.
template <typename T, typename TW, int NCW, int NRW>
__global__ __launch_bounds__(256) void mf_gemm(int n, int k, int nR, int nK, int nslices, int kpers,
                                               const TW* __restrict__ Wt, const T* __restrict__ x,
                                               double* __restrict__ part)
{
    using V4 = typename Mf<TW>::v4;
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nC = (k + 15) / 16, ncg = (nC + NCW - 1) / NCW, nRg = (nR + NRW - 1) / NRW;
    const int tile = w / nslices, sl = w - tile * nslices;
    if (tile >= nRg * ncg) return;
    const int R0 = (tile / ncg) * NRW, G = tile - (tile / ncg) * ncg;
    const int q = lane >> 4, cl = lane & 15;
    const int K0 = sl * kpers, K1 = min(nK, K0 + kpers);
    V4 acc[NRW][NCW];
#pragma unroll
    for (int r = 0; r < NRW; ++r)
#pragma unroll
        for (int t = 0; t < NCW; ++t) acc[r][t] = V4{0, 0, 0, 0};
    // row tiles past the last one read row tile nR - 1 (their products are not stored)
    const TW* wt[NRW];
#pragma unroll
    for (int r = 0; r < NRW; ++r) wt[r] = Wt + (size_t)min(R0 + r, nR - 1) * nK * 64 + lane;
    auto xval = [&](int K, int t) {
        const int c = 4 * K + q, col = 16 * (G * NCW + t) + cl;
        return (c < n && col < k) ? (TW)x[(size_t)c * k + col] : (TW)0;
    };
    int K = K0;
    for (; K + 4 <= K1; K += 4) {
        TW a[4][NRW], b[4][NCW];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int r = 0; r < NRW; ++r) a[u][r] = __builtin_nontemporal_load(wt[r] + (size_t)(K + u) * 64);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int t = 0; t < NCW; ++t) b[u][t] = xval(K + u, t);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int r = 0; r < NRW; ++r)
#pragma unroll
                for (int t = 0; t < NCW; ++t) acc[r][t] = Mf<TW>::mac(a[u][r], b[u][t], acc[r][t]);
    }
    for (; K < K1; ++K) {
        TW b[NCW];
#pragma unroll
        for (int t = 0; t < NCW; ++t) b[t] = xval(K, t);
#pragma unroll
        for (int r = 0; r < NRW; ++r) {
            const TW a = __builtin_nontemporal_load(wt[r] + (size_t)K * 64);
#pragma unroll
            for (int t = 0; t < NCW; ++t) acc[r][t] = Mf<TW>::mac(a, b[t], acc[r][t]);
        }
    }
#pragma unroll
    for (int r = 0; r < NRW; ++r) {
        if (R0 + r >= nR) continue;
#pragma unroll
        for (int t = 0; t < NCW; ++t) {
            const int col = 16 * (G * NCW + t) + cl;
            if (col >= k) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = 16 * (R0 + r) + Mf<TW>::row(lane, e);
                part[((size_t)sl * (16 * nR) + row) * k + col] = (double)acc[r][t][e];
            }
        }
    }
}

// y = sum over slices, in slice order
template <typename T>
__global__ void k_gemm_sum(int n, int k, int nR, int nslices, const double* __restrict__ part, T* __restrict__ y)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)n * k) return;
    double s = 0;
    for (int sl = 0; sl < nslices; ++sl) s += part[(size_t)sl * (16 * nR) * k + e];
    y[e] = (T)s;
}

inline unsigned nblocks(long waves) { return (unsigned)((waves + 3) / 4); }

void check(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("kernel launch failed: ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

#define MGLDG_MF_DISPATCH(B, NT, ...)                                                          \
    switch ((B) * 8 + (NT)) {                                                                  \
        case 6 * 8 + 1: { constexpr int BB = 6, TT = 1; __VA_ARGS__; } break;                  \
        case 6 * 8 + 2: { constexpr int BB = 6, TT = 2; __VA_ARGS__; } break;                  \
        case 6 * 8 + 4: { constexpr int BB = 6, TT = 4; __VA_ARGS__; } break;                  \
        case 8 * 8 + 1: { constexpr int BB = 8, TT = 1; __VA_ARGS__; } break;                  \
        case 8 * 8 + 2: { constexpr int BB = 8, TT = 2; __VA_ARGS__; } break;                  \
        case 8 * 8 + 4: { constexpr int BB = 8, TT = 4; __VA_ARGS__; } break;                  \
        case 9 * 8 + 1: { constexpr int BB = 9, TT = 1; __VA_ARGS__; } break;                  \
        case 9 * 8 + 2: { constexpr int BB = 9, TT = 2; __VA_ARGS__; } break;                  \
        case 9 * 8 + 4: { constexpr int BB = 9, TT = 4; __VA_ARGS__; } break;                  \
        case 10 * 8 + 1: { constexpr int BB = 10, TT = 1; __VA_ARGS__; } break;                \
        case 10 * 8 + 2: { constexpr int BB = 10, TT = 2; __VA_ARGS__; } break;                \
        case 10 * 8 + 4: { constexpr int BB = 10, TT = 4; __VA_ARGS__; } break;                \
        case 16 * 8 + 1: { constexpr int BB = 16, TT = 1; __VA_ARGS__; } break;                \
        case 16 * 8 + 2: { constexpr int BB = 16, TT = 2; __VA_ARGS__; } break;                \
        case 16 * 8 + 4: { constexpr int BB = 16, TT = 4; __VA_ARGS__; } break;                \
        case 27 * 8 + 1: { constexpr int BB = 27, TT = 1; __VA_ARGS__; } break;                \
        case 27 * 8 + 2: { constexpr int BB = 27, TT = 2; __VA_ARGS__; } break;                \
        default: throw std::runtime_error("MFMA block kernels: block size " + std::to_string(B) + " / tiles " + std::to_string(NT) + " not compiled"); \
    }

bool mf_supported(int B) { return B == 6 || B == 8 || B == 9 || B == 10 || B == 16 || B == 27; }

// a wave's NT tiles read and write NT contiguous columns as one vector, so NT
// must divide k: the largest of nt, nt / 2, .., 1 that does
static int fit_tiles(int nt, int k)
{
    while (nt > 1 && k % nt) nt >>= 1;
    return nt;
}
int mf_max_tiles(int B) { return B > 16 ? 2 : 4; }

// POOL as a compile-time constant PP
#define MGLDG_MF_POOL(pool, ...)                        \
    do {                                                \
        if (pool) { constexpr bool PP = true; __VA_ARGS__; } \
        else { constexpr bool PP = false; __VA_ARGS__; }     \
    } while (0)

template <typename T>
void MF<T>::apply(hipStream_t s, int B, int nt, int k, int nrows, const int* rowptr, const int* ell, int W, const T* A,
                  const T* x, const T* b, T* y, bool pool)
{
    if (nrows <= 0 || k <= 0) return;
    nt = fit_tiles(nt, k);
    const int ng = (k + 16 * nt - 1) / (16 * nt);
    const dim3 g(nblocks((long)nrows * ng));
    MGLDG_MF_POOL(pool, MGLDG_MF_DISPATCH(B, nt, hipLaunchKernelGGL((mf_apply<T, BB, TT, PP, false>), g, dim3(256), 0, s, k,
                                                                    nrows, ng, rowptr, ell, W, A, x, b, y, nullptr)));
    check("mf_apply");
}

long apply_dot_part_size(int nt, int k, int nrows)
{
    nt = fit_tiles(nt, k);
    const long ng = (k + 16 * nt - 1) / (16 * nt);
    return (long)nblocks((long)nrows * ng) * k;
}

void apply_dot(hipStream_t s, int B, int nt, int k, int nrows, const int* rowptr, const int* ell, int W, const double* A,
               const double* x, double* y, bool pool, double* part, double* out, int ostride)
{
    if (nrows <= 0 || k <= 0) return;
    if (k > 256) throw std::runtime_error("apply_dot: more than 256 columns");
    nt = fit_tiles(nt, k);
    const int ng = (k + 16 * nt - 1) / (16 * nt);
    const unsigned nb = nblocks((long)nrows * ng);
    MGLDG_MF_POOL(pool, MGLDG_MF_DISPATCH(B, nt, hipLaunchKernelGGL((mf_apply<double, BB, TT, PP, true>), dim3(nb), dim3(256), 0, s,
                                                                    k, nrows, ng, rowptr, ell, W, A, x, nullptr, y, part)));
    check("mf_apply (dot)");
    colfinalize(s, part, (int)nb, k, out, ostride);
}

template <typename T>
void MF<T>::gs_color(hipStream_t s, int B, int nt, int k, int nrows, const int* desc, const int* gcols, int gw, const T* A,
                     const T* Dinv, const T* b, T* x, T omega, const int* color, int cur, int skip, bool pool)
{
    if (nrows <= 0 || k <= 0) return;
    nt = fit_tiles(nt, k);
    const int ng = (k + 16 * nt - 1) / (16 * nt);
    const dim3 g(nblocks((long)nrows * ng));
    const int4* d4 = reinterpret_cast<const int4*>(desc);
    if (skip == 1) {
        MGLDG_MF_POOL(pool, MGLDG_MF_DISPATCH(B, nt, hipLaunchKernelGGL((mf_gs<T, BB, TT, 1, PP>), g, dim3(256), 0, s, k, nrows, ng,
                                                                        d4, gcols, gw, A, Dinv, b, x, omega, color, cur)));
    } else if (skip == 2) {
        MGLDG_MF_POOL(pool, MGLDG_MF_DISPATCH(B, nt, hipLaunchKernelGGL((mf_gs<T, BB, TT, 2, PP>), g, dim3(256), 0, s, k, nrows, ng,
                                                                        d4, gcols, gw, A, Dinv, b, x, omega, color, cur)));
    } else {
        MGLDG_MF_POOL(pool, MGLDG_MF_DISPATCH(B, nt, hipLaunchKernelGGL((mf_gs<T, BB, TT, 0, PP>), g, dim3(256), 0, s, k, nrows, ng,
                                                                        d4, gcols, gw, A, Dinv, b, x, omega, color, cur)));
    }
    check("mf_gs");
}

template <typename T>
void MF<T>::jacobi(hipStream_t s, int B, int nt, int k, int nrows, const int* rowptr, const int* ell, int W, const T* A,
                   const T* Dinv, const T* b, const T* x, T* xout, T omega, const int* didx)
{
    if (nrows <= 0 || k <= 0) return;
    nt = fit_tiles(nt, k);
    const int ng = (k + 16 * nt - 1) / (16 * nt);
    const dim3 g(nblocks((long)nrows * ng));
    MGLDG_MF_POOL(didx != nullptr, MGLDG_MF_DISPATCH(B, nt, hipLaunchKernelGGL((mf_jacobi<T, BB, TT, PP>), g, dim3(256), 0, s, k,
                                                                               nrows, ng, rowptr, ell, W, A, Dinv, b, x, xout,
                                                                               omega, didx)));
    check("mf_jacobi");
}

bool mf_interp_supported(int B) { return B == 9 || B == 16; }

#define MGLDG_MF_INTERP_DISPATCH(B, NT, ...)                                                   \
    switch ((B) * 8 + (NT)) {                                                                   \
        case 9 * 8 + 1: { constexpr int BB = 9, TT = 1; __VA_ARGS__; } break;                   \
        case 9 * 8 + 2: { constexpr int BB = 9, TT = 2; __VA_ARGS__; } break;                   \
        case 9 * 8 + 4: { constexpr int BB = 9, TT = 4; __VA_ARGS__; } break;                   \
        case 16 * 8 + 1: { constexpr int BB = 16, TT = 1; __VA_ARGS__; } break;                 \
        case 16 * 8 + 2: { constexpr int BB = 16, TT = 2; __VA_ARGS__; } break;                 \
        case 16 * 8 + 4: { constexpr int BB = 16, TT = 4; __VA_ARGS__; } break;                 \
        default: throw std::runtime_error("MFMA interpolation kernels: block size " + std::to_string(B) + " / tiles " + std::to_string(NT) + " not compiled"); \
    }

template <typename T>
void MF<T>::resrestrict(hipStream_t s, int B, int nt, int k, int nc, const int* kpair, const T* pool, const int* rowptr,
                        const int* ell, int W, const T* A, bool apool, const T* x, const T* b, T* bc)
{
    if (nc <= 0 || k <= 0) return;
    nt = fit_tiles(nt, k);
    const int ng = (k + 16 * nt - 1) / (16 * nt);
    MGLDG_MF_POOL(apool, MGLDG_MF_INTERP_DISPATCH(B, nt, hipLaunchKernelGGL((mf_resrestrict<T, BB, TT, PP>),
                                                                            dim3(nblocks((long)nc * ng)), dim3(256), 0, s, k,
                                                                            nc, ng, kpair, pool, rowptr, ell, W, A, x, b, bc)));
    check("mf_resrestrict");
}

template <typename T>
void MF<T>::prolong(hipStream_t s, int B, int nt, int k, int nc, const int* kpair, const T* pool, const T* xc, T* xf)
{
    if (nc <= 0 || k <= 0) return;
    nt = fit_tiles(nt, k);
    const int ng = (k + 16 * nt - 1) / (16 * nt);
    MGLDG_MF_INTERP_DISPATCH(B, nt, hipLaunchKernelGGL((mf_prolong<T, BB, TT>), dim3(nblocks((long)nc * ng)), dim3(256), 0, s,
                                                       k, nc, ng, kpair, pool, xc, xf));
    check("mf_prolong");
}

int gemm_tiles_size(int n) { return ((n + 15) / 16) * ((n + 3) / 4) * 64; }

template <typename TW>
void gemm_tiles(hipStream_t s, int n, const double* W, TW* Wt)
{
    const int nR = (n + 15) / 16, nK = (n + 3) / 4;
    const long t = (long)nR * nK * 64;
    if (t <= 0) return;
    hipLaunchKernelGGL(k_tile_w<TW>, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, n, nR, nK, W, Wt);
    check("tile_w");
}
template void gemm_tiles<double>(hipStream_t, int, const double*, double*);
template void gemm_tiles<float>(hipStream_t, int, const double*, float*);

// column tiles per wavefront of mf_gemm
static int gemm_ncw(int k)
{
    const int nC = (k + 15) / 16;
    return nC >= 4 ? 4 : (nC >= 2 ? 2 : 1);
}

// row tiles per wavefront of mf_gemm
constexpr int kGemmNRW = 2;

int gemm_slices(int n, int k)
{
    const int ncw = gemm_ncw(k);
    const int tiles = ((n + 15) / 16 + kGemmNRW - 1) / kGemmNRW * (((k + 15) / 16 + ncw - 1) / ncw);
    int sl = (8192 / kGemmNRW + tiles - 1) / tiles;   // about 8192 / kGemmNRW wavefronts
    const int nK = (n + 3) / 4;
    if (sl > 64) sl = 64;
    if (sl > nK) sl = nK;
    return sl < 1 ? 1 : sl;
}

long gemm_part_size(int n, int k) { return (long)gemm_slices(n, k) * 16 * ((n + 15) / 16) * k; }

template <typename T, typename TW>
void gemm_mf(hipStream_t s, int n, int k, const TW* Wt, const T* x, T* y, double* part)
{
    if (n <= 0 || k <= 0) return;
    const int nR = (n + 15) / 16, nK = (n + 3) / 4, nC = (k + 15) / 16;
    const int sl = gemm_slices(n, k), kpers = (nK + sl - 1) / sl;
    const int ncw = gemm_ncw(k), ncg = (nC + ncw - 1) / ncw, nRg = (nR + kGemmNRW - 1) / kGemmNRW;
    const dim3 g(nblocks((long)nRg * ncg * sl));
    constexpr int R = kGemmNRW;
    if (ncw == 4) hipLaunchKernelGGL((mf_gemm<T, TW, 4, R>), g, dim3(256), 0, s, n, k, nR, nK, sl, kpers, Wt, x, part);
    else if (ncw == 2) hipLaunchKernelGGL((mf_gemm<T, TW, 2, R>), g, dim3(256), 0, s, n, k, nR, nK, sl, kpers, Wt, x, part);
    else hipLaunchKernelGGL((mf_gemm<T, TW, 1, R>), g, dim3(256), 0, s, n, k, nR, nK, sl, kpers, Wt, x, part);
    check("mf_gemm");
    const long t = (long)n * k;
    hipLaunchKernelGGL(k_gemm_sum<T>, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, n, k, nR, sl, part, y);
    check("gemm_sum");
}
template void gemm_mf<float, double>(hipStream_t, int, int, const double*, const float*, float*, double*);
template void gemm_mf<float, float>(hipStream_t, int, int, const float*, const float*, float*, double*);
template void gemm_mf<double, double>(hipStream_t, int, int, const double*, const double*, double*, double*);

template struct MF<float>;
template struct MF<double>;

}  // namespace dev
}  // namespace mgldg
