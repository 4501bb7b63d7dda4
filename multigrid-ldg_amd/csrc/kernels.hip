// kernels.hip — CDNA4 (gfx950) kernels of libmgldg.
//
// Solve kernels are HBM-bandwidth bound block-sparse streams.  Mapping: one
// block row is handled by B consecutive lanes of a 64-wide wavefront (lane r
// owns output row r of the block), so a wavefront holds floor(64/B) block
// rows and a 256-thread workgroup four wavefronts.  Blocks are stored in a
// "panel" layout: element (r, c) of a B x B block lives at
//     ((c / V) * B + r) * V + (c % V)
// with V = 16 bytes / sizeof(T) when B is a multiple of it (else 2 or 1), so
// at every step the B lanes of a row read B consecutive 16-byte vectors and
// the whole wavefront issues one contiguous stream of dwordx4 loads.
//
// Reductions (PCG dot products) are deterministic: per-workgroup partials are
// summed in index order by the last workgroup to finish.
#include "kernels.hpp"
#include "devutil.hpp"

#include <mutex>
#include <stdexcept>
#include <string>

namespace mgldg {
namespace dev {

// --------------------------------------------------------------------- helpers
__host__ __device__ constexpr int panel_v(int tsize, int B)
{
    return tsize == 8 ? (B % 2 == 0 ? 2 : 1) : (B % 4 == 0 ? 4 : (B % 2 == 0 ? 2 : 1));
}

template <typename T, int V>
struct alignas(sizeof(T) * V) VecT { T v[V]; };

// Pooled blocks (A and D^-1 of a pooled level, read from cache by the
// single-RHS kernels) are in the panel layout with the columns padded with
// zeros to pad_of<T, B>(), a multiple of the 16-byte vector VP = vp_of<T>():
// element (r, c) at ((c / VP) * B + r) * VP + c % VP, c < pad_of.  Lane r
// loads its row as pad_of / VP aligned 16-byte vectors whatever B is (9 x 9
// blocks: 5 FP64 / 3 FP32 loads per block row instead of 9), and the B lanes
// of a row read consecutive vectors (one contiguous run per load).  Where
// B * sizeof(T) is a multiple of 16 this is the panel layout itself.
template <typename T>
__host__ __device__ constexpr int vp_of() { return 16 / (int)sizeof(T); }
template <typename T, int B>
__host__ __device__ constexpr int pad_of() { return (B + vp_of<T>() - 1) / vp_of<T>() * vp_of<T>(); }

template <typename T, int V>
__device__ __forceinline__ VecT<T, V> ldv(const T* p) { return *reinterpret_cast<const VecT<T, V>*>(p); }

template <typename T, int B>
struct Map {
    static constexpr int V = panel_v(sizeof(T), B);
    static constexpr int RPW = 64 / B;        // block rows per wavefront
    static constexpr int RPB = 4 * RPW;       // block rows per 256-thread workgroup
};

static void check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("kernel launch failed: ") + what + ": " + hipGetErrorString(e));
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// sum over the workgroup (result valid in thread 0); fixed order
__device__ double block_sum(double v)
{
    __shared__ double ws[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) ws[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0) for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += ws[i];
    return t;
}

// Deterministic two-pass reductions: every workgroup writes its partial sum;
// a one-workgroup kernel adds the partials in index order.  (A fused
// "last workgroup finishes" scheme needs agent-scope fences, which write back
// the L2 of every XCD and cost more than the extra launch.)
__device__ __forceinline__ void block_partial(double v, double* part, int slot)
{
    double bs = block_sum(v);
    if (threadIdx.x == 0) part[slot] = bs;
}
__device__ __forceinline__ void block_partial(double v, double* part) { block_partial(v, part, blockIdx.x); }

// sums nval interleaved partial streams (part[i * nval + k], i < n)
__global__ __launch_bounds__(1024) void k_finalize(const double* __restrict__ part, int n, int nval, double* S, int mode,
                                                   double* out)
{
    double t[3] = {0, 0, 0};
    for (int i = threadIdx.x; i < n; i += 1024)
        for (int k = 0; k < nval; ++k) t[k] += part[(size_t)i * nval + k];
    for (int k = 0; k < nval; ++k) t[k] = block_sum(t[k]);
    if (threadIdx.x != 0) return;
    if (mode == 0) { *out = t[0]; return; }
    // mode 1: PCG r.z / r.zold / r.r update with beta (krylov.h:54-57)
    const double old = S[S_DELTA];
    S[S_DELTA_OLD] = old;
    S[S_DELTA] = t[0];
    S[S_RSOLD] = t[1];
    S[S_RR] = t[2];
    S[S_BETA] = (mode == 2) ? (t[0] - t[1]) / old : t[0] / old;
}

static void finalize(hipStream_t s, const double* part, int n, int nval, double* S, int mode, double* out)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, part, n, nval, S, mode, out);
    check_launch("finalize");
}

void finalize_sum(hipStream_t s, const double* part, int n, double* out) { finalize(s, part, n, 1, nullptr, 0, out); }

// --------------------------------------------------------------------- setup kernels
// largest block of the setup kernels: (p+1)^N for 2-D p <= 5 and 3-D p <= 2
#define TP_MAX 36
#define TP_ACC ((TP_MAX * TP_MAX + 255) / 256)   // output entries per thread of k_triple
__global__ __launch_bounds__(256) void k_triple(int p, int q, int m, int n, int transL,
                                                const int* __restrict__ off, const int* __restrict__ li,
                                                const int* __restrict__ xi, const int* __restrict__ ri,
                                                const double* __restrict__ Lp, const double* __restrict__ Xp,
                                                const double* __restrict__ Rp, const double* __restrict__ addpool,
                                                const int* __restrict__ addslot, double* __restrict__ out)
{
    __shared__ double Ls[TP_MAX * TP_MAX], Xs[TP_MAX * TP_MAX], Rs[TP_MAX * TP_MAX], Ts[TP_MAX * TP_MAX];
    const int o = blockIdx.x, tid = threadIdx.x;
    const int pq = p * q;
    double acc[TP_ACC];
    const int as = addslot ? addslot[o] : -1;
    for (int u = 0; u < TP_ACC; ++u) {
        int e = tid + u * 256;
        acc[u] = (e < pq && as >= 0) ? addpool[(size_t)as * pq + e] : 0.0;
    }
    const int lm = transL ? m * p : p * m;
    for (int t = off[o]; t < off[o + 1]; ++t) {
        const int a = li[t], xx = xi[t], c = ri[t];
        for (int e = tid; e < lm; e += 256) {
            if (a >= 0) Ls[e] = Lp[(size_t)a * lm + e];
            else { int rr = e / (transL ? p : m), cc = e % (transL ? p : m); Ls[e] = (rr == cc) ? 1.0 : 0.0; }
        }
        for (int e = tid; e < m * n; e += 256) Xs[e] = Xp[(size_t)xx * m * n + e];
        for (int e = tid; e < n * q; e += 256) {
            if (c >= 0) Rs[e] = Rp[(size_t)c * n * q + e];
            else Rs[e] = (e / q == e % q) ? 1.0 : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < m * q; e += 256) {
            int i = e / q, j = e % q; double s = 0;
            for (int k = 0; k < n; ++k) s += Xs[i * n + k] * Rs[k * q + j];
            Ts[e] = s;
        }
        __syncthreads();
        for (int u = 0; u < TP_ACC; ++u) {
            int e = tid + u * 256;
            if (e >= pq) break;
            int i = e / q, j = e % q; double s = 0;
            if (transL) for (int k = 0; k < m; ++k) s += Ls[k * p + i] * Ts[k * q + j];
            else        for (int k = 0; k < m; ++k) s += Ls[i * m + k] * Ts[k * q + j];
            acc[u] += s;
        }
        __syncthreads();
    }
    for (int u = 0; u < TP_ACC; ++u) {
        int e = tid + u * 256;
        if (e < pq) out[(size_t)o * pq + e] = acc[u];
    }
}

void triple(hipStream_t s, int nout, int p, int q, int m, int n, int transL, const int* off, const int* l,
            const int* x, const int* r, const double* Lp, const double* Xp, const double* Rp,
            const double* addpool, const int* addslot, double* out)
{
    if (nout <= 0) return;
    if (p > TP_MAX || q > TP_MAX || m > TP_MAX || n > TP_MAX) throw std::runtime_error("triple: block larger than " + std::to_string(TP_MAX));
    hipLaunchKernelGGL(k_triple, dim3(nout), dim3(256), 0, s, p, q, m, n, transL, off, l, x, r, Lp, Xp, Rp, addpool, addslot, out);
    check_launch("triple");
}

// Gauss-Jordan inverse with partial pivoting, one workgroup per block.
__global__ __launch_bounds__(256) void k_inverse(int b, const double* __restrict__ in, const int* __restrict__ slot,
                                                 double* __restrict__ out)
{
    __shared__ double W[TP_MAX * 2 * TP_MAX];
    __shared__ double fac[TP_MAX];
    __shared__ int piv;
    const int blk = blockIdx.x, tid = threadIdx.x, w = 2 * b;
    const double* A = in + (size_t)(slot ? slot[blk] : blk) * b * b;
    for (int e = tid; e < b * w; e += 256) {
        int i = e / w, j = e % w;
        W[e] = (j < b) ? A[i * b + j] : ((j - b) == i ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int k = 0; k < b; ++k) {
        if (tid == 0) {
            int pi = k; double best = fabs(W[k * w + k]);
            for (int i = k + 1; i < b; ++i) if (fabs(W[i * w + k]) > best) { best = fabs(W[i * w + k]); pi = i; }
            piv = pi;
        }
        __syncthreads();
        const int pi = piv;
        if (pi != k) for (int j = tid; j < w; j += 256) { double t = W[k * w + j]; W[k * w + j] = W[pi * w + j]; W[pi * w + j] = t; }
        __syncthreads();
        const double d = W[k * w + k];
        __syncthreads();
        for (int j = tid; j < w; j += 256) W[k * w + j] = (d != 0.0) ? W[k * w + j] / d : 0.0;
        for (int i = tid; i < b; i += 256) fac[i] = (i == k) ? 0.0 : W[i * w + k];
        __syncthreads();
        for (int e = tid; e < b * w; e += 256) {
            int i = e / w, j = e % w;
            if (i != k) W[e] -= fac[i] * W[k * w + j];
        }
        __syncthreads();
    }
    for (int e = tid; e < b * b; e += 256) out[(size_t)blk * b * b + e] = W[(e / b) * w + b + (e % b)];
}

void inverse(hipStream_t s, int nblk, int b, const double* in, const int* slot, double* out)
{
    if (nblk <= 0) return;
    if (b > TP_MAX) throw std::runtime_error("inverse: block larger than " + std::to_string(TP_MAX));
    hipLaunchKernelGGL(k_inverse, dim3(nblk), dim3(256), 0, s, b, in, slot, out);
    check_launch("inverse");
}

template <typename T>
__global__ void k_to_panel(long nblk, int B, int V, const double* __restrict__ in, T* __restrict__ out)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long bs = (long)B * B;
    if (e >= nblk * bs) return;
    long k = e / bs; int rc = (int)(e % bs), r = rc / B, c = rc % B;
    out[k * bs + ((c / V) * B + r) * V + (c % V)] = (T)in[e];
}

// row-major FP64 blocks -> the padded panel layout of T (pad_of): element
// (r, c) of block k at k * B * bp + ((c / vp) * B + r) * vp + c % vp, zero
// for c >= B
template <typename T>
__global__ void k_to_rowpad(long nblk, int B, int bp, const double* __restrict__ in, T* __restrict__ out)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nblk * B * bp) return;
    constexpr int vp = vp_of<T>();
    const long k = e / ((long)B * bp);
    const int q = (int)(e - k * B * bp);              // ((c / vp) * B + r) * vp + c % vp
    const int cm = q % vp, rest = q / vp, r = rest % B, c = (rest / B) * vp + cm;
    out[e] = c < B ? (T)in[(k * B + r) * B + c] : (T)0;
}

int pool_pad(int tsize, int b) { return tsize == 8 ? (b + 1) / 2 * 2 : (b + 3) / 4 * 4; }
static_assert(pad_of<double, 9>() == 10 && pad_of<float, 9>() == 12 && pad_of<float, 16>() == 16, "pool_pad and pad_of agree");

void to_rowpad_f64(hipStream_t s, long nblk, int b, const double* in, double* out)
{
    const int bp = pool_pad(8, b);
    const long n = nblk * b * bp;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_rowpad<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nblk, b, bp, in, out);
    check_launch("to_rowpad");
}
void to_rowpad_f32(hipStream_t s, long nblk, int b, const double* in, float* out)
{
    const int bp = pool_pad(4, b);
    const long n = nblk * b * bp;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_rowpad<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nblk, b, bp, in, out);
    check_launch("to_rowpad");
}

void to_panel_f64(hipStream_t s, long nblk, int b, const double* in, double* out)
{
    long n = nblk * b * b; if (n <= 0) return;
    hipLaunchKernelGGL(k_to_panel<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nblk, b, panel_v(8, b), in, out);
    check_launch("to_panel");
}
void to_panel_f32(hipStream_t s, long nblk, int b, const double* in, float* out)
{
    long n = nblk * b * b; if (n <= 0) return;
    hipLaunchKernelGGL(k_to_panel<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nblk, b, panel_v(4, b), in, out);
    check_launch("to_panel");
}

__global__ void k_gather_blocks(int n, int bs, const int* __restrict__ idx, const double* __restrict__ in, double* __restrict__ out)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * bs) return;
    long i = e / bs; int t = (int)(e % bs);
    out[e] = in[(long)idx[i] * bs + t];
}
void gather_blocks(hipStream_t s, int n, int bs, const int* idx, const double* in, double* out)
{
    long tot = (long)n * bs; if (tot <= 0) return;
    hipLaunchKernelGGL(k_gather_blocks, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, bs, idx, in, out);
    check_launch("gather_blocks");
}

// --------------------------------------------------------------------- solve kernels
// Blocks kept in flight per lane: the A operands of up to U blocks of a row
// are loaded before any of them is used (one HBM round trip per U blocks).
// Sized so that the A operands stay near 64 VGPRs.
template <typename T, int B>
__host__ __device__ constexpr int unroll_of()
{
    return sizeof(T) == 4 ? (B <= 16 ? 4 : (B <= 27 ? 2 : 1)) : (B <= 10 ? 4 : (B <= 16 ? 2 : 1));
}

// LDS exchange among the lanes of one wavefront: every lane that writes and
// reads a staging area here belongs to the same wavefront, so ordering within
// the wavefront is all that is needed (no workgroup barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane r of a block row returns sum_m A_k(m)[r][:] . x_col(k(m)) over the
// row's blocks k0 .. k1-1 other than dp.  Latency is what bounds these
// launches (one colour of a sweep is a few MB to a few hundred MB), so:
//  * SKIP = 0: the A operands of a round of U blocks are loaded before the
//    column indices arrive (their addresses follow from k0 and dp alone);
//  * lane r loads entry r of x_j (one coalesced B-vector per block) and
//    stages it in LDS (xs: U * B values per row), from where the B lanes of
//    the row read it back as broadcast vectors: x costs one register per
//    block in flight instead of B;
//  * SKIP = 1 (2): first sweep from x = 0; x_j of colours after (before)
//    `cur` in sweep order are still zero, and their blocks are not read.
// The column indices come from a padded table (colp: the row's W entries,
// column of the m-th summed block at colp[m], -1 past the row's blocks), so
// their loads need neither the row's descriptor nor rowptr and go out in the
// first round trip; up to B are loaded at once, one per lane, and broadcast
// with lane shuffles.  Trip counts are uniform across the wavefront, so every
// shuffle runs with all lanes active.
// POOL: the operator is a pool of distinct blocks (quadtree operators repeat
// a few hundred blocks over the whole mesh); colp holds (column, pool index)
// pairs and a block is read at A + index * B^2, from cache once the pool is
// resident, so a launch moves its vectors and index tables only.
template <typename T, int B, int SKIP, bool POOL = false>
__device__ __forceinline__ T row_sum(bool valid, int k0, int k1, int dp, const int* __restrict__ colp, int W,
                                     const T* __restrict__ A, const T* x, int r, int lane0, T* xs,
                                     const int* __restrict__ color, int cur)
{
    constexpr int V = Map<T, B>::V;
    constexpr int U = POOL ? 1 : unroll_of<T, B>();
    T acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0;
    int mycol = -1, mypid = 0;
    auto fetch = [&](int m) {
        if (POOL) {
            const int2 e = (valid && m < W) ? reinterpret_cast<const int2*>(colp)[m] : make_int2(-1, 0);
            mycol = e.x; mypid = e.y;   // (pool index | kPidLower on GS tables)
        } else {
            mycol = (valid && m < W) ? colp[m] : -1;
        }
    };
    fetch(r);
    const int n = valid ? k1 - k0 - (dp >= 0 ? 1 : 0) : 0;
    int wl = n;
    for (int o = 32; o > 0; o >>= 1) wl = max(wl, __shfl_xor(wl, o, 64));
    if constexpr (POOL) {
        T pacc[vp_of<T>()];
#pragma unroll
        for (int v = 0; v < vp_of<T>(); ++v) pacc[v] = 0;
        // Pooled blocks come from cache; the x_j gathers are the row's HBM
        // round trips.  All UX gathers of a round (and, on a first sweep,
        // the neighbours' colours) are issued together, so a row waits for
        // one round trip per UX blocks; the A operands follow one block at a
        // time from cache, which keeps the registers low (8 wavefronts per
        // SIMD on the fine levels).  A gathered x_j of a colour not yet
        // updated in this sweep is replaced by zero, never used.
        constexpr int UX = unroll_of<T, B>();
        constexpr int VP = vp_of<T>(), BP = pad_of<T, B>();
        // the pad entries of the staged x stay zero (A's pad columns are
        // zero too, and 0 * x must not meet a stale non-finite value)
        if (r < BP - B) {
#pragma unroll
            for (int u = 0; u < UX; ++u) xs[u * BP + B + r] = (T)0;
        }
        for (int m0 = 0; m0 < wl; m0 += B) {
            if (m0 > 0) fetch(m0 + r);
            const int nc = min(B, wl - m0);
            for (int u0 = 0; u0 < nc; u0 += UX) {
                int pid[UX];
                bool use[UX];
                T xv[UX];
#pragma unroll
                for (int u = 0; u < UX; ++u) {
                    const int src = lane0 + min(u0 + u, B - 1);
                    const int j = __shfl(mycol, src, 64);
                    const int pe = __shfl(mypid, src, 64);
                    pid[u] = pe & kPidMask;
                    const bool live = valid && u0 + u < nc && j >= 0;
                    xv[u] = live ? x[(size_t)j * B + r] : (T)0;
                    // first sweep: a neighbour of a preceding colour is written (forward) or not (reverse)
                    const bool lower = (pe & kPidLower) != 0;
                    use[u] = live && (SKIP == 0 || (SKIP == 1 ? lower : !lower));
                }
#pragma unroll
                for (int u = 0; u < UX; ++u) xs[u * BP + r] = use[u] ? xv[u] : (T)0;
                wave_sync();
#pragma unroll
                for (int u = 0; u < UX; ++u)
                    if (use[u]) {
                        const T* Ab = A + (size_t)pid[u] * (B * BP) + r * VP;
#pragma unroll
                        for (int cc = 0; cc < BP / VP; ++cc) {
                            const VecT<T, VP> av = ldv<T, VP>(Ab + cc * B * VP);
                            const VecT<T, VP> xq = ldv<T, VP>(xs + u * BP + cc * VP);
#pragma unroll
                            for (int v = 0; v < VP; ++v) pacc[v] += av.v[v] * xq.v[v];
                        }
                    }
                wave_sync();
            }
        }
        T s = pacc[0];
#pragma unroll
        for (int v = 1; v < VP; ++v) s += pacc[v];
        return s;
    }
    for (int m0 = 0; m0 < wl; m0 += B) {
        if (m0 > 0) fetch(m0 + r);
        const int nc = min(B, wl - m0);
        for (int u0 = 0; u0 < nc; u0 += U) {
            VecT<T, V> a[U][B / V];
            bool use[U];
            if (SKIP == 0) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int m = m0 + u0 + u;
                    use[u] = u0 + u < nc && m < n;
                    if (use[u]) {
                        const T* Ab = A + (size_t)blk_at(k0, m, dp) * (B * B) + r * V;
#pragma unroll
                        for (int cc = 0; cc < B / V; ++cc) a[u][cc] = ldv<T, V>(Ab + cc * B * V);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = __shfl(mycol, lane0 + min(u0 + u, B - 1), 64);
                bool ok = valid && u0 + u < nc && j >= 0;
                if (SKIP != 0) {
                    if (ok) { const int cj = color[j]; ok = SKIP == 1 ? cj <= cur : cj >= cur; }
                    use[u] = ok;
                    if (ok) {
                        const T* Ab = A + (size_t)blk_at(k0, m0 + u0 + u, dp) * (B * B) + r * V;
#pragma unroll
                        for (int cc = 0; cc < B / V; ++cc) a[u][cc] = ldv<T, V>(Ab + cc * B * V);
                    }
                }
                xs[u * B + r] = ok ? x[(size_t)j * B + r] : (T)0;
            }
            wave_sync();
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (use[u]) {
#pragma unroll
                    for (int cc = 0; cc < B / V; ++cc) {
                        const VecT<T, V> xv = ldv<T, V>(xs + u * B + cc * V);
#pragma unroll
                        for (int v = 0; v < V; ++v) acc[v] += a[u][cc].v[v] * xv.v[v];
                    }
                }
            wave_sync();
        }
    }
    T s = acc[0];
#pragma unroll
    for (int v = 1; v < V; ++v) s += acc[v];
    return s;
}

// Per-wavefront LDS of the row kernels: RW row slots (the last one for the
// lanes past 64/B rows, which hold no row) of U * B staged x values and B
// values for the D^-1 product.
template <typename T, int B>
struct RowLds {
    static constexpr int RW = Map<T, B>::RPW + 1;
    static constexpr int XS = unroll_of<T, B>() * pad_of<T, B>();
    static constexpr int TB = pad_of<T, B>();   // D^-1 operand staging per row
};

// s_r = sum_c Dinv[r][c] t_c with lane r's row of Dinv in registers and t in LDS
template <typename T, int B>
__device__ __forceinline__ T dinv_row(const VecT<T, Map<T, B>::V>* dv, const T* t)
{
    constexpr int V = Map<T, B>::V;
    T s[V];
#pragma unroll
    for (int v = 0; v < V; ++v) s[v] = 0;
#pragma unroll
    for (int cc = 0; cc < B / V; ++cc) {
        const VecT<T, V> tv = ldv<T, V>(t + cc * V);
#pragma unroll
        for (int v = 0; v < V; ++v) s[v] += dv[cc].v[v] * tv.v[v];
    }
    T out = s[0];
#pragma unroll
    for (int v = 1; v < V; ++v) out += s[v];
    return out;
}

template <typename T, int B>
__device__ __forceinline__ void load_dinv_row(const T* __restrict__ Dinv, int i, int r, VecT<T, Map<T, B>::V>* dv)
{
    constexpr int V = Map<T, B>::V;
    const T* Db = Dinv + (size_t)i * (B * B) + r * V;
#pragma unroll
    for (int cc = 0; cc < B / V; ++cc) dv[cc] = ldv<T, V>(Db + cc * B * V);
}

// the same for a pooled D^-1 block (padded panel layout)
template <typename T, int B>
struct PadRow { VecT<T, vp_of<T>()> v[pad_of<T, B>() / vp_of<T>()]; };

template <typename T, int B>
__device__ __forceinline__ void load_dinv_row_pad(const T* __restrict__ Dinv, int i, int r, PadRow<T, B>& dv)
{
    constexpr int VP = vp_of<T>(), BP = pad_of<T, B>();
    const T* Db = Dinv + (size_t)i * (B * BP) + r * VP;
#pragma unroll
    for (int cc = 0; cc < BP / VP; ++cc) dv.v[cc] = ldv<T, VP>(Db + cc * B * VP);
}

// s_r = sum_c Dinv[r][c] t_c, t in LDS with its pad entries zero
template <typename T, int B>
__device__ __forceinline__ T dinv_row_pad(const PadRow<T, B>& dv, const T* t)
{
    constexpr int VP = vp_of<T>(), BP = pad_of<T, B>();
    T s[VP];
#pragma unroll
    for (int v = 0; v < VP; ++v) s[v] = 0;
#pragma unroll
    for (int cc = 0; cc < BP / VP; ++cc) {
        const VecT<T, VP> tv = ldv<T, VP>(t + cc * VP);
#pragma unroll
        for (int v = 0; v < VP; ++v) s[v] += dv.v[cc].v[v] * tv.v[v];
    }
    T out = s[0];
#pragma unroll
    for (int v = 1; v < VP; ++v) out += s[v];
    return out;
}

// y = A x  or  y = b - A x ; optional deterministic reduction of y.dotv
template <typename T, int B, bool POOL = false>
__global__ __launch_bounds__(256) void k_apply(int nrows, const int* __restrict__ rowptr, const int* __restrict__ ell,
                                               int W, const T* __restrict__ A, const T* __restrict__ x,
                                               const T* __restrict__ b, T* __restrict__ y, const T* __restrict__ dotv,
                                               double* part)
{
    using M = Map<T, B>;
    using R = RowLds<T, B>;
    __shared__ __align__(16) T xs_all[4 * R::RW * R::XS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / B, r = lane - slot * B;
    const int lb = xcd_block();
    const int row = lb * M::RPB + wave * M::RPW + slot;
    const bool valid = slot < M::RPW && row < nrows;
    const int k0 = valid ? rowptr[row] : 0, k1 = valid ? rowptr[row + 1] : 0;
    const size_t o = (size_t)row * B + r;
    const T bo = (valid && b) ? b[o] : (T)0;
    // the dot's other vector: y itself (its value is `out`), or an input
    // loaded with the row's other operands
    const bool dot_self = dotv == y;
    const T dw = (valid && dotv && !dot_self) ? dotv[o] : (T)0;
    const T acc = row_sum<T, B, 0, POOL>(valid, k0, k1, -1, ell + (size_t)row * W * (POOL ? 2 : 1), W, A, x, r, slot * B,
                                         xs_all + (wave * R::RW + slot) * R::XS, nullptr, 0);
    double dv = 0;
    if (valid) {
        T out = b ? (T)(bo - acc) : acc;
        y[o] = out;
        if (dotv) dv = (double)out * (double)(dot_self ? out : dw);
    }
    if (part) block_partial(dv, part, lb);   // partials in row order, whatever the placement
}

// Damped block GS on rows [beg, end) of a descriptor list, one block row per
// B lanes.  desc[idx] = {row i, rowptr[i], rowptr[i+1], position of the
// diagonal block}: the row's operands (b_i, x_i, its row of D^-1, its A
// blocks) are all addressable from that one load; gcols[idx * W + m] =
// column of its m-th off-diagonal block (-1 past the last).  The rows must be
// independent (one colour).
// POOL: desc.w is the pool index of the row's D^-1 block and gcols holds
// (column, pool index) pairs (row_sum).
template <typename T, int B, int SKIP, bool POOL = false>
__device__ __forceinline__ void gs_rows(int beg, int end, const int4* __restrict__ desc, const int* __restrict__ gcols,
                                        int W, const T* __restrict__ A, const T* __restrict__ Dinv, const T* __restrict__ b,
                                        T* x, T omega, const int* __restrict__ color, int cur, T* xs_all, T* tb)
{
    using M = Map<T, B>;
    using R = RowLds<T, B>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / B, r = lane - slot * B;
    const int idx = beg + wave * M::RPW + slot;
    const bool valid = slot < M::RPW && idx < end;
    int4 d = make_int4(0, 0, 0, -1);
    if (valid) d = desc[idx];
    const size_t o = (size_t)d.x * B + r;
    T bi = 0, xi = 0;
    VecT<T, M::V> dv[B / M::V];
    if (valid) {
        bi = b[o];
        if (SKIP == 0) xi = x[o];
        if (!POOL) load_dinv_row<T, B>(Dinv, d.x, r, dv);
    }
    const T acc = reads_none<SKIP>(cur) ? (T)0
                : row_sum<T, B, SKIP, POOL>(valid, d.y, d.z, POOL ? 0 : d.w, gcols + (size_t)idx * W * (POOL ? 2 : 1), W, A,
                                            x, r, slot * B, xs_all + (wave * R::RW + slot) * R::XS, color, cur);
    // a pooled D^-1 block is read from cache: loaded after the sum, it holds
    // no registers across it (more wavefronts per SIMD)
    PadRow<T, B> dp;
    if (POOL && valid) load_dinv_row_pad<T, B>(Dinv, d.w, r, dp);
    T* t = tb + (wave * R::RW + slot) * R::TB;
    t[r] = bi - acc;
    if (POOL && r < R::TB - B) t[B + r] = (T)0;
    wave_sync();
    if (valid) {
        const T s = POOL ? dinv_row_pad<T, B>(dp, t) : dinv_row<T, B>(dv, t);
        x[o] = xi + omega * (s - xi);
    }
}

// One colour of a damped block GS sweep (rows desc[0 .. nrows)).
template <typename T, int B, int SKIP, bool POOL = false>
__global__ __launch_bounds__(256) void k_gs(int nrows, const int4* __restrict__ desc, const int* __restrict__ gcols,
                                            int W, const T* __restrict__ A, const T* __restrict__ Dinv,
                                            const T* __restrict__ b, T* x, T omega,
                                            const int* __restrict__ color, int cur)
{
    using R = RowLds<T, B>;
    __shared__ __align__(16) T xs_all[4 * R::RW * R::XS];
    __shared__ __align__(16) T tb[4 * R::RW * R::TB];
    const int beg = xcd_block() * Map<T, B>::RPB;
    gs_rows<T, B, SKIP, POOL>(beg, nrows, desc, gcols, W, A, Dinv, b, x, omega, color, cur, xs_all, tb);
}

// The last two colours of a sweep in one launch.  Workgroup w owns the
// descriptor ranges [s, m) (rows of the second-to-last colour) and [m, e)
// (rows of the last colour), {s, m, e} = wg[3w .. 3w+2].  Every row of the
// last colour shares a workgroup with all its neighbours of the
// second-to-last colour (clusters of the coupling graph between the two
// colours), so processing the first range, a workgroup barrier, then the
// second (the reverse for a backward sweep) gives exactly the two
// one-colour launches in sweep order.  Workgroups holding only rows with no
// neighbour in the last colour have m = e.  cur0 / cur1: colour of the
// first / second range processed.
template <typename T, int B, int SKIP, bool POOL = false>
__global__ __launch_bounds__(256) void k_gs_pair(const int* __restrict__ wg, int reverse, const int4* __restrict__ desc,
                                                 const int* __restrict__ gcols, int W, const T* __restrict__ A,
                                                 const T* __restrict__ Dinv, const T* __restrict__ b, T* x, T omega,
                                                 const int* __restrict__ color, int cur0, int cur1)
{
    using M = Map<T, B>;
    using R = RowLds<T, B>;
    __shared__ __align__(16) T xs_all[4 * R::RW * R::XS];
    __shared__ __align__(16) T tb[4 * R::RW * R::TB];
    const int lb = xcd_block();
    const int s = wg[3 * lb], m = wg[3 * lb + 1], e = wg[3 * lb + 2];
    for (int ph = 0; ph < 2; ++ph) {
        const bool lo = (ph == 0) != (reverse != 0);
        const int beg = lo ? s : m, end = lo ? m : e;
        if (ph == 1 && m != s && m != e) __syncthreads();   // (uniform: s, m, e are per workgroup)
        for (int base = beg; base < end; base += M::RPB)
            gs_rows<T, B, SKIP, POOL>(base, end, desc, gcols, W, A, Dinv, b, x, omega, color, ph ? cur1 : cur0, xs_all, tb);
    }
}

// ------------------------------------------------- pooled GS, latency form
// A pooled GS launch of a few thousand rows is bound by the chain of
// dependent loads of each row, not by bytes: descriptor and pairs -> x_j
// gathers (rounds of UX blocks) -> the pool rows of those blocks (from cache,
// after the gathers arrive) -> D^-1 (after the sum) -> store.  The latency
// form issues everything a row needs -- b_i, x_i, its D^-1 pool row, every
// x_j (table width W <= kLatW, one pair per lane) with its colour, and the
// pool rows of the first UA blocks -- at once as soon as the row's
// descriptor and pairs have arrived, so a row waits for two round trips.
// The products and the order of every sum are those of the pooled row_sum /
// gs_rows (table order, the same VP-wide partial sums, the same D^-1
// product), so the results are the same bits.  It holds more registers and
// is launched only where the launch is too small to need occupancy (and for
// the two-colour launch, whose rows are sparse per workgroup).
constexpr int kLatW = 8;

template <typename T, int B>
struct LatCfg {
    static constexpr int VP = vp_of<T>(), BP = pad_of<T, B>();
    static constexpr int NV = BP / VP;   // vector loads per pool row
    // pool rows kept in registers: at most 128 VGPRs
    static constexpr int UA0 = (128 * 4) / (BP * (int)sizeof(T));
    static constexpr int UA = UA0 < kLatW ? UA0 : kLatW;
    static constexpr int XS = kLatW * BP;   // staged x_j per row slot
};

template <typename T, int B, int SKIP>
struct LatRow {
    using C = LatCfg<T, B>;
    bool valid = false;
    int4 d = make_int4(0, 0, 0, 0);
    int2 e = make_int2(-1, 0);
    T bi = 0, xi = 0;
    PadRow<T, B> dp;
    T xv[kLatW];
    int pid[kLatW];
    bool use[kLatW];
    VecT<T, C::VP> a[C::UA][C::NV];

    // first round trip: the descriptor and the (column, pool index) pairs,
    // pair r on lane r of the row (W <= B)
    __device__ __forceinline__ void tables(int idx, bool v, const int4* __restrict__ desc, const int* __restrict__ gcols,
                                           int W, int r)
    {
        valid = v;
        if (v) d = desc[idx];
        if (v && r < W) e = reinterpret_cast<const int2*>(gcols)[(size_t)idx * W + r];
    }
    // the row's own operands (x_i only when the sweep does not start from zero)
    __device__ __forceinline__ void own(const T* __restrict__ b, const T* x, const T* __restrict__ Dinv, int r)
    {
        if (!valid) return;
        const size_t o = (size_t)d.x * B + r;
        bi = b[o];
        if (SKIP == 0) xi = x[o];
        load_dinv_row_pad<T, B>(Dinv, d.w, r, dp);
    }
    // second round trip: every x_j with its colour (first sweep) and the
    // pool rows of the first UA blocks.  Uniform trip counts: every shuffle
    // runs with all lanes active.
    __device__ __forceinline__ void nbrs(const T* __restrict__ A, const T* x, int cur, int W, int r, int lane0)
    {
        const bool none = reads_none<SKIP>(cur);
#pragma unroll
        for (int u = 0; u < kLatW; ++u) {
            const int src = lane0 + (u < B ? u : B - 1);
            const int j = __shfl(e.x, src, 64);
            const int pe = __shfl(e.y, src, 64);
            pid[u] = pe & kPidMask;
            const bool live = !none && valid && u < W && j >= 0;
            xv[u] = live ? x[(size_t)j * B + r] : (T)0;
            // first sweep: a neighbour of a preceding colour is written (forward) or not (reverse)
            const bool lower = (pe & kPidLower) != 0;
            use[u] = live && (SKIP == 0 || (SKIP == 1 ? lower : !lower));
            if (u < C::UA && live) {
                const T* Ab = A + (size_t)pid[u] * (B * C::BP) + r * C::VP;
#pragma unroll
                for (int cc = 0; cc < C::NV; ++cc) a[u < C::UA ? u : 0][cc] = ldv<T, C::VP>(Ab + cc * B * C::VP);
            }
        }
    }
    // the row sum, the D^-1 product and the damped update of x_i
    __device__ __forceinline__ void finish(const T* __restrict__ A, T* x, T omega, int r, T* xs, T* t)
    {
        constexpr int VP = C::VP, BP = C::BP;
        wave_sync();   // the slot's previous readers (an earlier row of this wavefront) are done
#pragma unroll
        for (int u = 0; u < kLatW; ++u) xs[u * BP + r] = use[u] ? xv[u] : (T)0;
        if (r < BP - B) {
#pragma unroll
            for (int u = 0; u < kLatW; ++u) xs[u * BP + B + r] = (T)0;
        }
        wave_sync();
        T pacc[VP];
#pragma unroll
        for (int v = 0; v < VP; ++v) pacc[v] = 0;
#pragma unroll
        for (int u = 0; u < kLatW; ++u)
            if (use[u]) {
                const T* Ab = A + (size_t)pid[u] * (B * BP) + r * VP;
#pragma unroll
                for (int cc = 0; cc < C::NV; ++cc) {
                    const VecT<T, VP> av = u < C::UA ? a[u < C::UA ? u : 0][cc] : ldv<T, VP>(Ab + cc * B * VP);
                    const VecT<T, VP> xq = ldv<T, VP>(xs + u * BP + cc * VP);
#pragma unroll
                    for (int v = 0; v < VP; ++v) pacc[v] += av.v[v] * xq.v[v];
                }
            }
        T acc = pacc[0];
#pragma unroll
        for (int v = 1; v < VP; ++v) acc += pacc[v];
        t[r] = bi - acc;
        if (r < RowLds<T, B>::TB - B) t[B + r] = (T)0;
        wave_sync();
        if (valid) {
            const T s = dinv_row_pad<T, B>(dp, t);
            x[(size_t)d.x * B + r] = xi + omega * (s - xi);
        }
    }
};

// k_gs<T, B, SKIP, true> in the latency form
template <typename T, int B, int SKIP>
__global__ __launch_bounds__(256) void k_gs_lat(int nrows, const int4* __restrict__ desc, const int* __restrict__ gcols,
                                                int W, const T* __restrict__ A, const T* __restrict__ Dinv,
                                                const T* __restrict__ b, T* x, T omega, int cur)
{
    using M = Map<T, B>;
    using R = RowLds<T, B>;
    using C = LatCfg<T, B>;
    __shared__ __align__(16) T xs_all[4 * R::RW * C::XS];
    __shared__ __align__(16) T tb[4 * R::RW * R::TB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / B, r = lane - slot * B;
    const int idx = xcd_block() * M::RPB + wave * M::RPW + slot;
    LatRow<T, B, SKIP> q;
    q.tables(idx, slot < M::RPW && idx < nrows, desc, gcols, W, r);
    q.own(b, x, Dinv, r);
    q.nbrs(A, x, cur, W, r, slot * B);
    q.finish(A, x, omega, r, xs_all + (wave * R::RW + slot) * C::XS, tb + (wave * R::RW + slot) * R::TB);
}

// k_gs_pair<T, B, SKIP, true> in the latency form.  Each of a workgroup's
// two ranges holds at most one workgroup of rows (pair_schedule builds them
// so), so one row per row slot in each phase: the second range's tables and
// own operands (rows of the other colour, which the first phase does not
// write) are loaded with the first's, and after the barrier only its x_j
// gathers and pool rows remain.
template <typename T, int B, int SKIP>
__global__ __launch_bounds__(256) void k_gs_pair_lat(const int* __restrict__ wg, int reverse, const int4* __restrict__ desc,
                                                     const int* __restrict__ gcols, int W, const T* __restrict__ A,
                                                     const T* __restrict__ Dinv, const T* __restrict__ b, T* x, T omega,
                                                     int cur0, int cur1)
{
    using M = Map<T, B>;
    using R = RowLds<T, B>;
    using C = LatCfg<T, B>;
    __shared__ __align__(16) T xs_all[4 * R::RW * C::XS];
    __shared__ __align__(16) T tb[4 * R::RW * R::TB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / B, r = lane - slot * B;
    const int lb = xcd_block();
    const int s = wg[3 * lb], m = wg[3 * lb + 1], e = wg[3 * lb + 2];
    const int b0 = reverse ? m : s, e0 = reverse ? e : m;   // first range processed
    const int b1 = reverse ? s : m, e1 = reverse ? m : e;   // second
    const int off = wave * M::RPW + slot;
    const bool inw = slot < M::RPW;
    LatRow<T, B, SKIP> p, q;
    p.tables(b0 + off, inw && b0 + off < e0, desc, gcols, W, r);
    q.tables(b1 + off, inw && b1 + off < e1, desc, gcols, W, r);
    p.own(b, x, Dinv, r);
    q.own(b, x, Dinv, r);
    T* xs = xs_all + (wave * R::RW + slot) * C::XS;
    T* t = tb + (wave * R::RW + slot) * R::TB;
    p.nbrs(A, x, cur0, W, r, slot * B);
    p.finish(A, x, omega, r, xs, t);
    if (m != s && m != e) __syncthreads();   // (uniform: s, m, e are per workgroup)
    q.nbrs(A, x, cur1, W, r, slot * B);
    q.finish(A, x, omega, r, xs, t);
}

// block sizes with a latency form (the 2-D quadtree blocks, p = 1, 2, 3)
constexpr bool lat_block(int B) { return B == 4 || B == 9 || B == 16; }
constexpr long kSimds = 1024;   // 256 CUs x 4 SIMDs

// POOL: ell holds (column, pool index) pairs, didx the rows' D^-1 pool indices
template <typename T, int B, bool POOL = false>
__global__ __launch_bounds__(256) void k_jacobi(int nrows, const int* __restrict__ rowptr, const int* __restrict__ ell,
                                                int W, const T* __restrict__ A, const T* __restrict__ Dinv,
                                                const int* __restrict__ didx, const T* __restrict__ b,
                                                const T* __restrict__ x, T* __restrict__ xout, T omega)
{
    using M = Map<T, B>;
    using R = RowLds<T, B>;
    __shared__ __align__(16) T xs_all[4 * R::RW * R::XS];
    __shared__ __align__(16) T tb[4 * R::RW * R::TB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / B, r = lane - slot * B;
    const int i = xcd_block() * M::RPB + wave * M::RPW + slot;
    const bool valid = slot < M::RPW && i < nrows;
    const int k0 = valid ? rowptr[i] : 0, k1 = valid ? rowptr[i + 1] : 0;
    const size_t o = (size_t)i * B + r;
    T bi = 0, xi = 0;
    VecT<T, M::V> dv[B / M::V];
    if (valid) {
        bi = b[o];
        xi = x[o];
        if (!POOL) load_dinv_row<T, B>(Dinv, i, r, dv);
    }
    const T acc = row_sum<T, B, 0, POOL>(valid, k0, k1, -1, ell + (size_t)i * W * (POOL ? 2 : 1), W, A, x, r, slot * B,
                                         xs_all + (wave * R::RW + slot) * R::XS, nullptr, 0);
    PadRow<T, B> dp;
    if (POOL && valid) load_dinv_row_pad<T, B>(Dinv, didx[i], r, dp);   // (as in gs_rows)
    T* t = tb + (wave * R::RW + slot) * R::TB;
    t[r] = bi - acc;
    if (POOL && r < R::TB - B) t[B + r] = (T)0;
    wave_sync();
    if (valid) xout[o] = xi + omega * (POOL ? dinv_row_pad<T, B>(dp, t) : dinv_row<T, B>(dv, t));
}

// The first Jacobi sweep from x = 0: xout = omega D^-1 b (A x = 0 is not
// read), the value k_jacobi gives for x = 0
// POOL: Dinv is a pool of padded row-major blocks indexed by didx
template <typename T, int B, bool POOL = false>
__global__ __launch_bounds__(256) void k_jacobi0(int nrows, const T* __restrict__ Dinv, const int* __restrict__ didx,
                                                 const T* __restrict__ b, T* __restrict__ xout, T omega)
{
    using M = Map<T, B>;
    using R = RowLds<T, B>;
    __shared__ __align__(16) T tb[4 * R::RW * R::TB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / B, r = lane - slot * B;
    const int i = blockIdx.x * M::RPB + wave * M::RPW + slot;
    const bool valid = slot < M::RPW && i < nrows;
    const size_t o = (size_t)i * B + r;
    VecT<T, M::V> dv[B / M::V];
    PadRow<T, B> dp;
    T* t = tb + (wave * R::RW + slot) * R::TB;
    t[r] = valid ? b[o] : (T)0;
    if (POOL && r < R::TB - B) t[B + r] = (T)0;
    if (valid) {
        if (POOL) load_dinv_row_pad<T, B>(Dinv, didx[i], r, dp);
        else load_dinv_row<T, B>(Dinv, i, r, dv);
    }
    wave_sync();
    if (valid) xout[o] = omega * (POOL ? dinv_row_pad<T, B>(dp, t) : dinv_row<T, B>(dv, t));
}

// Interpolation blocks are BF x BC row-major (fine simplex or box basis x
// coarse box basis); BF == BC on quadtree hierarchies.
template <typename T, int BF, int BC>
__global__ __launch_bounds__(256) void k_restrict(int nc, const int* __restrict__ kid_off, const int* __restrict__ kids,
                                                  const int* __restrict__ type, const T* __restrict__ pool,
                                                  const T* __restrict__ r, T* __restrict__ bc)
{
    using M = Map<T, BC>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / BC, c = lane - slot * BC, lane0 = slot * BC;
    const int I = blockIdx.x * M::RPB + wave * M::RPW + slot;
    if (slot >= M::RPW || I >= nc) return;
    const int t0 = kid_off[I], nk = kid_off[I + 1] - t0;
    T acc = 0;
    // the slot's lanes load up to BC children and their block types at once
    for (int m0 = 0; m0 < nk; m0 += BC) {
        const int mk = c < nk - m0 ? kids[t0 + m0 + c] : 0;
        const int mt = c < nk - m0 ? type[mk] : 0;
        const int n = min(BC, nk - m0);
        for (int m = 0; m < n; ++m) {
            const int k = __shfl(mk, lane0 + m, 64), ty = __shfl(mt, lane0 + m, 64);
            const T* Tb = pool + (size_t)ty * (BF * BC);
            const T* rk = r + (size_t)k * BF;
#pragma unroll
            for (int q = 0; q < BF; ++q) acc += Tb[q * BC + c] * rk[q];
        }
    }
    bc[(size_t)I * BC + c] = acc;
}

template <typename T, int BF, int BC>
__global__ __launch_bounds__(256) void k_prolong(int nf, const int* __restrict__ parent, const int* __restrict__ type,
                                                 const T* __restrict__ pool, const T* __restrict__ xc, T* __restrict__ xf)
{
    using M = Map<T, BF>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / BF, r = lane - slot * BF;
    const int k = blockIdx.x * M::RPB + wave * M::RPW + slot;
    if (slot >= M::RPW || k >= nf) return;
    const T* Tb = pool + (size_t)type[k] * (BF * BC) + r * BC;
    const T* xp = xc + (size_t)parent[k] * BC;
    T tr[BC];
    uld<T, BC>(tr, Tb);   // the lane's row of the block in one run
    T acc = 0;
#pragma unroll
    for (int c = 0; c < BC; ++c) acc += tr[c] * xp[c];
    xf[(size_t)k * BF + r] += acc;
}

// Quadtree interpolation in the latency form (BF == BC == B, at most kKidW
// children per coarse element, a pool of a few blocks).  The loops above
// walk a coarse row's children one dependent load after another (kid_off ->
// kids -> type -> blocks and r), and their lanes read block rows and columns
// at a stride of B elements, one cache line per lane and load.  Here a
// coarse row's table entry comes in one load and every operand in the next
// round trip: kpair[I kKidW + s] = (child, block) of coarse row I in kids[]
// order (-1: none).  The pools are in the padded panel layout of the A
// pools (row r of a block as pad_of / VP aligned 16-byte vectors, the B
// lanes of a row reading consecutive vectors): P holds the blocks, PT their
// transposes (lane c reads column c).  Vectors pass through LDS, one
// coalesced load per row.  Both walk coarse rows: the prolongation updates
// the coarse row's children, so a parent's x_c is read once.  Products and
// sums in the same order as k_restrict / k_prolong (the pad columns are
// never summed): the same bits.
constexpr int kKidW = kKidWidth;

template <typename T, int B>
struct InterpLat {
    static constexpr int VP = vp_of<T>(), BP = pad_of<T, B>(), NV = BP / VP;
};

// row r of panel block t: o[0 .. BP)
template <typename T, int B>
__device__ __forceinline__ void panel_row(T* o, const T* __restrict__ P, int t, int r)
{
    using C = InterpLat<T, B>;
    const T* p = P + (size_t)t * (B * C::BP) + r * C::VP;
#pragma unroll
    for (int cc = 0; cc < C::NV; ++cc) {
        const VecT<T, C::VP> v = ldv<T, C::VP>(p + cc * B * C::VP);
#pragma unroll
        for (int e = 0; e < C::VP; ++e) o[cc * C::VP + e] = v.v[e];
    }
}

// FUSE: the coarse level's first Gauss-Seidel launch of a sweep from x = 0
// as well.  That launch (colour 0, reading no neighbour) sets x_I = omega
// D^-1 b_I on the rows of colour 0 and nothing else, so it is done here
// from b_I as it is formed: colour[I], the row's D^-1 pool index didx[I]
// and its padded D^-1 row come with the other operands, and the update is
// the one of the latency GS form (LatRow::finish with no neighbour).
template <typename T, int B, bool FUSE>
__global__ __launch_bounds__(256) void k_restrict_lat(int nc, const int2* __restrict__ kpair, const T* __restrict__ PT,
                                                      const T* __restrict__ r, T* __restrict__ bc,
                                                      const int* __restrict__ color, const int* __restrict__ didx,
                                                      const T* __restrict__ Dinv, T* __restrict__ x, T omega)
{
    using M = Map<T, B>;
    using C = InterpLat<T, B>;
    using R = RowLds<T, B>;
    __shared__ __align__(16) T rs_all[4 * (M::RPW + 1) * kKidW * B];
    __shared__ __align__(16) T tb[FUSE ? 4 * R::RW * R::TB : 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / B, c = lane - slot * B;
    const int I = xcd_block() * M::RPB + wave * M::RPW + slot;
    const bool valid = slot < M::RPW && I < nc;
    int2 e = make_int2(-1, 0);
    if (valid && c < kKidW) e = kpair[(size_t)I * kKidW + c];
    bool gs = false;
    PadRow<T, B> dp;
    if constexpr (FUSE) {
        gs = valid && color[I] == 0;
        if (gs) load_dinv_row_pad<T, B>(Dinv, didx[I], c, dp);
    }
    T rv[kKidW], tc[kKidW][C::BP];
    bool live[kKidW];
#pragma unroll
    for (int s = 0; s < kKidW; ++s) {
        const int src = slot * B + (s < B ? s : B - 1);
        const int k = __shfl(e.x, src, 64), ty = __shfl(e.y, src, 64);
        live[s] = valid && k >= 0;
        rv[s] = live[s] ? r[(size_t)k * B + c] : (T)0;
        if (live[s]) panel_row<T, B>(tc[s], PT, ty, c);
    }
    T* rs = rs_all + (wave * (M::RPW + 1) + slot) * (kKidW * B);
#pragma unroll
    for (int s = 0; s < kKidW; ++s) rs[s * B + c] = rv[s];
    wave_sync();
    T acc = 0;
#pragma unroll
    for (int s = 0; s < kKidW; ++s)
        if (live[s]) {
#pragma unroll
            for (int q = 0; q < B; ++q) acc += tc[s][q] * rs[s * B + q];
        }
    if (valid) bc[(size_t)I * B + c] = acc;
    if constexpr (FUSE) {
        T* t = tb + (wave * R::RW + slot) * R::TB;
        t[c] = acc;
        if (c < R::TB - B) t[B + c] = (T)0;
        wave_sync();
        if (gs) {
            const T xi = 0, sv = dinv_row_pad<T, B>(dp, t);
            x[(size_t)I * B + c] = xi + omega * (sv - xi);
        }
    }
}

template <typename T, int B>
__global__ __launch_bounds__(256) void k_prolong_lat(int nc, const int2* __restrict__ kpair, const T* __restrict__ P,
                                                     const T* __restrict__ xc, T* __restrict__ xf)
{
    using M = Map<T, B>;
    using C = InterpLat<T, B>;
    __shared__ __align__(16) T xs_all[4 * (M::RPW + 1) * B];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / B, r = lane - slot * B;
    const int I = xcd_block() * M::RPB + wave * M::RPW + slot;
    const bool valid = slot < M::RPW && I < nc;
    int2 e = make_int2(-1, 0);
    if (valid && r < kKidW) e = kpair[(size_t)I * kKidW + r];
    const T xv = valid ? xc[(size_t)I * B + r] : (T)0;
    T xo[kKidW], tr[kKidW][C::BP];
    int ks[kKidW];
#pragma unroll
    for (int s = 0; s < kKidW; ++s) {
        const int src = slot * B + (s < B ? s : B - 1);
        const int k = __shfl(e.x, src, 64), ty = __shfl(e.y, src, 64);
        ks[s] = valid ? k : -1;
        xo[s] = ks[s] >= 0 ? xf[(size_t)k * B + r] : (T)0;
        if (ks[s] >= 0) panel_row<T, B>(tr[s], P, ty, r);
    }
    T* xs = xs_all + (wave * (M::RPW + 1) + slot) * B;
    xs[r] = xv;
    wave_sync();
#pragma unroll
    for (int s = 0; s < kKidW; ++s)
        if (ks[s] >= 0) {
            T acc = 0;
#pragma unroll
            for (int c = 0; c < B; ++c) acc += tr[s][c] * xs[c];
            xf[(size_t)ks[s] * B + r] = xo[s] + acc;
        }
}

template <typename T>
__global__ __launch_bounds__(256) void k_coarse(int n, const double* __restrict__ Cinv, const T* __restrict__ b,
                                                T* __restrict__ x, int subtract_mean)
{
    extern __shared__ double bs[];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 256) { double v = (double)b[i]; bs[i] = v; s += v; }
    __shared__ double mean;
    double tot = block_sum(s);
    if (threadIdx.x == 0) mean = subtract_mean ? tot / n : 0.0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) bs[i] -= mean;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) {
        double y = 0;
        for (int j = 0; j < n; ++j) y += Cinv[(size_t)i * n + j] * bs[j];
        x[i] = (T)y;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_submean(long n, T* b)
{
    double s = 0;
    for (long i = threadIdx.x; i < n; i += 256) s += (double)b[i];
    __shared__ double mean;
    double tot = block_sum(s);
    if (threadIdx.x == 0) mean = tot / (double)n;
    __syncthreads();
    for (long i = threadIdx.x; i < n; i += 256) b[i] = (T)((double)b[i] - mean);
}

template <typename T>
__global__ void k_fill(long n, T v, T* x) { long i = (long)blockIdx.x * blockDim.x + threadIdx.x; if (i < n) x[i] = v; }
template <typename T>
__global__ void k_copy(long n, const T* a, T* b) { long i = (long)blockIdx.x * blockDim.x + threadIdx.x; if (i < n) b[i] = a[i]; }
__global__ void k_d2f(long n, const double* a, float* b) { long i = (long)blockIdx.x * blockDim.x + threadIdx.x; if (i < n) b[i] = (float)a[i]; }
__global__ void k_f2d(long n, const float* a, double* b) { long i = (long)blockIdx.x * blockDim.x + threadIdx.x; if (i < n) b[i] = (double)a[i]; }

static inline unsigned nblk(long n, int per) { return (unsigned)((n + per - 1) / per); }

// dispatch over the compiled block sizes
#define MGLDG_DISPATCH(B, ...)                                                       \
    switch (B) {                                                                     \
        case 1: { constexpr int BB = 1; __VA_ARGS__; } break;                               \
        case 3: { constexpr int BB = 3; __VA_ARGS__; } break;                               \
        case 4: { constexpr int BB = 4; __VA_ARGS__; } break;                               \
        case 6: { constexpr int BB = 6; __VA_ARGS__; } break;                               \
        case 8: { constexpr int BB = 8; __VA_ARGS__; } break;                               \
        case 9: { constexpr int BB = 9; __VA_ARGS__; } break;                               \
        case 10: { constexpr int BB = 10; __VA_ARGS__; } break;                             \
        case 16: { constexpr int BB = 16; __VA_ARGS__; } break;                             \
        case 25: { constexpr int BB = 25; __VA_ARGS__; } break;                             \
        case 27: { constexpr int BB = 27; __VA_ARGS__; } break;                             \
        case 36: { constexpr int BB = 36; __VA_ARGS__; } break;                             \
        default: throw std::runtime_error("block size " + std::to_string(B) + " not compiled"); \
    }

// dispatch over the compiled (fine, coarse) interpolation block shapes: the
// square sizes, P_p simplex -> Q_p box for 2-D p = 1..3 and 3-D p = 1, 2, and
// the p-coarsenings Q_p -> Q_q (q < p) for 2-D p <= 5 and 3-D p <= 2
#define MGLDG_DISPATCH2(BF, BC, ...)                                                 \
    if ((BF) == (BC)) {                                                              \
        MGLDG_DISPATCH(BF, { constexpr int FF = BB, CC = BB; __VA_ARGS__; });        \
    } else {                                                                         \
        switch ((BF) * 1000 + (BC)) {                                                \
            case 3004: { constexpr int FF = 3, CC = 4; __VA_ARGS__; } break;         \
            case 6009: { constexpr int FF = 6, CC = 9; __VA_ARGS__; } break;         \
            case 10016: { constexpr int FF = 10, CC = 16; __VA_ARGS__; } break;      \
            case 4008: { constexpr int FF = 4, CC = 8; __VA_ARGS__; } break;         \
            case 10027: { constexpr int FF = 10, CC = 27; __VA_ARGS__; } break;      \
            case 36025: { constexpr int FF = 36, CC = 25; __VA_ARGS__; } break;      \
            case 36016: { constexpr int FF = 36, CC = 16; __VA_ARGS__; } break;      \
            case 36009: { constexpr int FF = 36, CC = 9; __VA_ARGS__; } break;      \
            case 36004: { constexpr int FF = 36, CC = 4; __VA_ARGS__; } break;      \
            case 36001: { constexpr int FF = 36, CC = 1; __VA_ARGS__; } break;      \
            case 25016: { constexpr int FF = 25, CC = 16; __VA_ARGS__; } break;      \
            case 25009: { constexpr int FF = 25, CC = 9; __VA_ARGS__; } break;      \
            case 25004: { constexpr int FF = 25, CC = 4; __VA_ARGS__; } break;      \
            case 25001: { constexpr int FF = 25, CC = 1; __VA_ARGS__; } break;      \
            case 16009: { constexpr int FF = 16, CC = 9; __VA_ARGS__; } break;      \
            case 16004: { constexpr int FF = 16, CC = 4; __VA_ARGS__; } break;      \
            case 16001: { constexpr int FF = 16, CC = 1; __VA_ARGS__; } break;      \
            case 9004: { constexpr int FF = 9, CC = 4; __VA_ARGS__; } break;      \
            case 9001: { constexpr int FF = 9, CC = 1; __VA_ARGS__; } break;      \
            case 4001: { constexpr int FF = 4, CC = 1; __VA_ARGS__; } break;      \
            case 27008: { constexpr int FF = 27, CC = 8; __VA_ARGS__; } break;      \
            case 27001: { constexpr int FF = 27, CC = 1; __VA_ARGS__; } break;      \
            case 8001: { constexpr int FF = 8, CC = 1; __VA_ARGS__; } break;      \
            default: throw std::runtime_error("interpolation block " + std::to_string(BF) + "x" + std::to_string(BC) + " not compiled"); \
        }                                                                            \
    }

template <typename T>
int Ops<T>::apply(hipStream_t s, int B, int nrows, const int* rowptr, const int* ell, int W, const T* A, const T* x,
                  const T* b, T* y, const T* dotv, double* dot_out, double* part, int pool)
{
    if (nrows <= 0) return 0;
    unsigned g = 0;
    double* pw = (dotv && part) ? part : nullptr;
    MGLDG_DISPATCH(B, {
        using M = Map<T, BB>;
        g = nblk(nrows, M::RPB);
        if (pool)
            hipLaunchKernelGGL((k_apply<T, BB, true>), dim3(g), dim3(256), 0, s, nrows, rowptr, ell, W, A, x, b, y,
                               dotv, pw);
        else
            hipLaunchKernelGGL((k_apply<T, BB>), dim3(g), dim3(256), 0, s, nrows, rowptr, ell, W, A, x, b, y,
                               dotv, pw);
    });
    check_launch("apply");
    if (dot_out) finalize(s, part, (int)g, 1, nullptr, 0, dot_out);
    return (int)g;
}

template <typename T>
void Ops<T>::gs_color(hipStream_t s, int B, int nrows, const int* desc, const int* gcols, int W, const T* A,
                      const T* Dinv, const T* b, T* x, T omega, const int* color, int cur, int skip, int pool)
{
    if (nrows <= 0) return;
    const int4* d = reinterpret_cast<const int4*>(desc);
    MGLDG_DISPATCH(B, {
        using M = Map<T, BB>;
        const dim3 g(nblk(nrows, M::RPB));
        // the latency form for launches of at most two wavefronts per SIMD
        if constexpr (lat_block(BB)) {
            if (pool && W <= std::min(BB, kLatW) && (long)g.x * 4 <= 2L * kSimds) {
                if (skip == 1)
                    hipLaunchKernelGGL((k_gs_lat<T, BB, 1>), g, dim3(256), 0, s, nrows, d, gcols, W, A, Dinv, b, x, omega, cur);
                else if (skip == 2)
                    hipLaunchKernelGGL((k_gs_lat<T, BB, 2>), g, dim3(256), 0, s, nrows, d, gcols, W, A, Dinv, b, x, omega, cur);
                else
                    hipLaunchKernelGGL((k_gs_lat<T, BB, 0>), g, dim3(256), 0, s, nrows, d, gcols, W, A, Dinv, b, x, omega, cur);
                check_launch("gs_color");
                return;
            }
        }
        if (pool) {
            if (skip == 1)
                hipLaunchKernelGGL((k_gs<T, BB, 1, true>), g, dim3(256), 0, s, nrows, d, gcols, W, A, Dinv, b, x, omega, color, cur);
            else if (skip == 2)
                hipLaunchKernelGGL((k_gs<T, BB, 2, true>), g, dim3(256), 0, s, nrows, d, gcols, W, A, Dinv, b, x, omega, color, cur);
            else
                hipLaunchKernelGGL((k_gs<T, BB, 0, true>), g, dim3(256), 0, s, nrows, d, gcols, W, A, Dinv, b, x, omega, color, cur);
        } else if (skip == 1)
            hipLaunchKernelGGL((k_gs<T, BB, 1>), g, dim3(256), 0, s, nrows, d, gcols, W, A, Dinv, b, x, omega, color, cur);
        else if (skip == 2)
            hipLaunchKernelGGL((k_gs<T, BB, 2>), g, dim3(256), 0, s, nrows, d, gcols, W, A, Dinv, b, x, omega, color, cur);
        else
            hipLaunchKernelGGL((k_gs<T, BB, 0>), g, dim3(256), 0, s, nrows, d, gcols, W, A, Dinv, b, x, omega, color, cur);
    });
    check_launch("gs_color");
}

template <typename T>
void Ops<T>::gs_pair(hipStream_t s, int B, int nwg, const int* wg, int reverse, const int* desc, const int* gcols, int W,
                     const T* A, const T* Dinv, const T* b, T* x, T omega, const int* color, int cur0, int cur1, int skip,
                     int pool)
{
    if (nwg <= 0) return;
    const int4* d = reinterpret_cast<const int4*>(desc);
    MGLDG_DISPATCH(B, {
        const dim3 g(nwg);
        if constexpr (lat_block(BB)) {
            if (pool && W <= std::min(BB, kLatW)) {
                if (skip == 1)
                    hipLaunchKernelGGL((k_gs_pair_lat<T, BB, 1>), g, dim3(256), 0, s, wg, reverse, d, gcols, W, A, Dinv, b, x,
                                       omega, cur0, cur1);
                else if (skip == 2)
                    hipLaunchKernelGGL((k_gs_pair_lat<T, BB, 2>), g, dim3(256), 0, s, wg, reverse, d, gcols, W, A, Dinv, b, x,
                                       omega, cur0, cur1);
                else
                    hipLaunchKernelGGL((k_gs_pair_lat<T, BB, 0>), g, dim3(256), 0, s, wg, reverse, d, gcols, W, A, Dinv, b, x,
                                       omega, cur0, cur1);
                check_launch("gs_pair");
                return;
            }
        }
        if (pool) {
            if (skip == 1)
                hipLaunchKernelGGL((k_gs_pair<T, BB, 1, true>), g, dim3(256), 0, s, wg, reverse, d, gcols, W, A, Dinv, b, x,
                                   omega, color, cur0, cur1);
            else if (skip == 2)
                hipLaunchKernelGGL((k_gs_pair<T, BB, 2, true>), g, dim3(256), 0, s, wg, reverse, d, gcols, W, A, Dinv, b, x,
                                   omega, color, cur0, cur1);
            else
                hipLaunchKernelGGL((k_gs_pair<T, BB, 0, true>), g, dim3(256), 0, s, wg, reverse, d, gcols, W, A, Dinv, b, x,
                                   omega, color, cur0, cur1);
        } else if (skip == 1)
            hipLaunchKernelGGL((k_gs_pair<T, BB, 1>), g, dim3(256), 0, s, wg, reverse, d, gcols, W, A, Dinv, b, x, omega, color,
                               cur0, cur1);
        else if (skip == 2)
            hipLaunchKernelGGL((k_gs_pair<T, BB, 2>), g, dim3(256), 0, s, wg, reverse, d, gcols, W, A, Dinv, b, x, omega, color,
                               cur0, cur1);
        else
            hipLaunchKernelGGL((k_gs_pair<T, BB, 0>), g, dim3(256), 0, s, wg, reverse, d, gcols, W, A, Dinv, b, x, omega, color,
                               cur0, cur1);
    });
    check_launch("gs_pair");
}

int rows_per_workgroup(int B)
{
    int r = 0;
    MGLDG_DISPATCH(B, { r = Map<float, BB>::RPB; });
    return r;
}

template <typename T>
void Ops<T>::jacobi(hipStream_t s, int B, int nrows, const int* rowptr, const int* ell, int W, const T* A,
                    const T* Dinv, const T* b, const T* x, T* xout, T omega, const int* didx)
{
    if (nrows <= 0) return;
    MGLDG_DISPATCH(B, {
        using M = Map<T, BB>;
        if (didx)
            hipLaunchKernelGGL((k_jacobi<T, BB, true>), dim3(nblk(nrows, M::RPB)), dim3(256), 0, s, nrows, rowptr, ell, W, A,
                               Dinv, didx, b, x, xout, omega);
        else
            hipLaunchKernelGGL((k_jacobi<T, BB>), dim3(nblk(nrows, M::RPB)), dim3(256), 0, s, nrows, rowptr, ell, W, A, Dinv,
                               nullptr, b, x, xout, omega);
    });
    check_launch("jacobi");
}

template <typename T>
void Ops<T>::jacobi0(hipStream_t s, int B, int nrows, const T* Dinv, const T* b, T* xout, T omega, const int* didx)
{
    if (nrows <= 0) return;
    MGLDG_DISPATCH(B, {
        using M = Map<T, BB>;
        if (didx)
            hipLaunchKernelGGL((k_jacobi0<T, BB, true>), dim3(nblk(nrows, M::RPB)), dim3(256), 0, s, nrows, Dinv, didx, b, xout,
                               omega);
        else
            hipLaunchKernelGGL((k_jacobi0<T, BB>), dim3(nblk(nrows, M::RPB)), dim3(256), 0, s, nrows, Dinv, didx, b, xout, omega);
    });
    check_launch("jacobi0");
}

template <typename T>
void Ops<T>::restrict_(hipStream_t s, int BF, int BC, int nc, const int* kid_off, const int* kids, const int* type,
                       const T* pool, const T* r, T* bc)
{
    if (nc <= 0) return;
    MGLDG_DISPATCH2(BF, BC, {
        using M = Map<T, CC>;
        hipLaunchKernelGGL((k_restrict<T, FF, CC>), dim3(nblk(nc, M::RPB)), dim3(256), 0, s, nc, kid_off, kids, type, pool, r, bc);
    });
    check_launch("restrict");
}

template <typename T>
void Ops<T>::prolong(hipStream_t s, int BF, int BC, int nf, const int* parent, const int* type, const T* pool, const T* xc,
                     T* xf)
{
    if (nf <= 0) return;
    MGLDG_DISPATCH2(BF, BC, {
        using M = Map<T, FF>;
        hipLaunchKernelGGL((k_prolong<T, FF, CC>), dim3(nblk(nf, M::RPB)), dim3(256), 0, s, nf, parent, type, pool, xc, xf);
    });
    check_launch("prolong");
}

template <typename T>
void Ops<T>::restrict_lat(hipStream_t s, int B, int nc, const int* kpair, const T* PT, const T* r, T* bc,
                          const int* color, const int* didx, const T* Dinv, T* x, T omega)
{
    if (nc <= 0) return;
    MGLDG_DISPATCH(B, {
        if constexpr (lat_block(BB)) {
            using M = Map<T, BB>;
            const int2* kp = reinterpret_cast<const int2*>(kpair);
            if (color)
                hipLaunchKernelGGL((k_restrict_lat<T, BB, true>), dim3(nblk(nc, M::RPB)), dim3(256), 0, s, nc, kp, PT, r,
                                   bc, color, didx, Dinv, x, omega);
            else
                hipLaunchKernelGGL((k_restrict_lat<T, BB, false>), dim3(nblk(nc, M::RPB)), dim3(256), 0, s, nc, kp, PT, r,
                                   bc, nullptr, nullptr, nullptr, nullptr, (T)0);
        } else {
            throw std::runtime_error("restrict_lat: block size not compiled");
        }
    });
    check_launch("restrict_lat");
}

template <typename T>
void Ops<T>::prolong_lat(hipStream_t s, int B, int nc, const int* kpair, const T* P, const T* xc, T* xf)
{
    if (nc <= 0) return;
    MGLDG_DISPATCH(B, {
        if constexpr (lat_block(BB)) {
            using M = Map<T, BB>;
            hipLaunchKernelGGL((k_prolong_lat<T, BB>), dim3(nblk(nc, M::RPB)), dim3(256), 0, s, nc,
                               reinterpret_cast<const int2*>(kpair), P, xc, xf);
        } else {
            throw std::runtime_error("prolong_lat: block size not compiled");
        }
    });
    check_launch("prolong_lat");
}

bool interp_lat(int B, int max_kids) { return lat_block(B) && max_kids <= kKidW; }

template <typename T>
void Ops<T>::coarse(hipStream_t s, int n, const double* Cinv, const T* b, T* x, int subtract_mean)
{
    if (n > 6144) throw std::runtime_error("coarse: dense coarse problem too large");
    hipLaunchKernelGGL(k_coarse<T>, dim3(1), dim3(256), n * sizeof(double), s, n, Cinv, b, x, subtract_mean);
    check_launch("coarse");
}

template <typename T>
void Ops<T>::submean(hipStream_t s, long n, T* b)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_submean<T>, dim3(1), dim3(256), 0, s, n, b);
    check_launch("submean");
}

template <typename T>
void Ops<T>::fill(hipStream_t s, long n, T v, T* x)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n, 256)), dim3(256), 0, s, n, v, x);
    check_launch("fill");
}

template <typename T>
void Ops<T>::copy(hipStream_t s, long n, const T* a, T* b)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy<T>, dim3(nblk(n, 256)), dim3(256), 0, s, n, a, b);
    check_launch("copy");
}

template struct Ops<float>;
template struct Ops<double>;

void f64_to_f32(hipStream_t s, long n, const double* a, float* b)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_d2f, dim3(nblk(n, 256)), dim3(256), 0, s, n, a, b);
    check_launch("f64_to_f32");
}
void f32_to_f64(hipStream_t s, long n, const float* a, double* b)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_f2d, dim3(nblk(n, 256)), dim3(256), 0, s, n, a, b);
    check_launch("f32_to_f64");
}

// --------------------------------------------------------------------- k right-hand sides
template <typename T, int B>
__device__ __forceinline__ T pan(const T* blk, int r, int c)
{
    constexpr int V = Map<T, B>::V;
    return blk[((c / V) * B + r) * V + (c % V)];
}

template <typename T, int B>
__global__ __launch_bounds__(256) void mk_apply(int k, int nrows, const int* __restrict__ rowptr, const int* __restrict__ cols,
                                                const T* __restrict__ A, const T* __restrict__ x, const T* __restrict__ b,
                                                T* __restrict__ y)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)nrows * k) return;
    const int i = (int)(e / k), c = (int)(e % k);
    T acc[B];
#pragma unroll
    for (int r = 0; r < B; ++r) acc[r] = 0;
    for (int t = rowptr[i]; t < rowptr[i + 1]; ++t) {
        const T* Ab = A + (size_t)t * B * B;
        const T* xj = x + (size_t)cols[t] * B * k + c;
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const T xv = xj[(size_t)q * k];
#pragma unroll
            for (int r = 0; r < B; ++r) acc[r] += pan<T, B>(Ab, r, q) * xv;
        }
    }
#pragma unroll
    for (int r = 0; r < B; ++r) {
        const size_t o = ((size_t)i * B + r) * k + c;
        y[o] = b ? (T)(b[o] - acc[r]) : acc[r];
    }
}

template <typename T, int B>
__global__ __launch_bounds__(256) void mk_gs(int k, int nrows, const int* __restrict__ rows, const int* __restrict__ rowptr,
                                             const int* __restrict__ cols, const T* __restrict__ A, const T* __restrict__ Dinv,
                                             const T* __restrict__ b, T* x, T omega, const int* __restrict__ color, int cur,
                                             int skip)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)nrows * k) return;
    const int i = rows[(int)(e / k)], c = (int)(e % k);
    T t[B];
#pragma unroll
    for (int r = 0; r < B; ++r) t[r] = b[((size_t)i * B + r) * k + c];
    for (int s = rowptr[i]; s < rowptr[i + 1]; ++s) {
        const int j = cols[s];
        if (j == i) continue;
        if (skip) { const int cj = color[j]; if (skip == 1 ? cj > cur : cj < cur) continue; }
        const T* Ab = A + (size_t)s * B * B;
        const T* xj = x + (size_t)j * B * k + c;
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const T xv = xj[(size_t)q * k];
#pragma unroll
            for (int r = 0; r < B; ++r) t[r] -= pan<T, B>(Ab, r, q) * xv;
        }
    }
    const T* D = Dinv + (size_t)i * B * B;
#pragma unroll
    for (int r = 0; r < B; ++r) {
        T sdot = 0;
#pragma unroll
        for (int q = 0; q < B; ++q) sdot += pan<T, B>(D, r, q) * t[q];
        const size_t o = ((size_t)i * B + r) * k + c;
        const T xi = skip ? (T)0 : x[o];
        x[o] = xi + omega * (sdot - xi);
    }
}

template <typename T, int B>
__global__ __launch_bounds__(256) void mk_jacobi(int k, int nrows, const int* __restrict__ rowptr, const int* __restrict__ cols,
                                                 const T* __restrict__ A, const T* __restrict__ Dinv, const T* __restrict__ b,
                                                 const T* __restrict__ x, T* __restrict__ xout, T omega)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)nrows * k) return;
    const int i = (int)(e / k), c = (int)(e % k);
    T t[B];
#pragma unroll
    for (int r = 0; r < B; ++r) t[r] = b[((size_t)i * B + r) * k + c];
    for (int s = rowptr[i]; s < rowptr[i + 1]; ++s) {
        const T* Ab = A + (size_t)s * B * B;
        const T* xj = x + (size_t)cols[s] * B * k + c;
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const T xv = xj[(size_t)q * k];
#pragma unroll
            for (int r = 0; r < B; ++r) t[r] -= pan<T, B>(Ab, r, q) * xv;
        }
    }
    const T* D = Dinv + (size_t)i * B * B;
#pragma unroll
    for (int r = 0; r < B; ++r) {
        T sdot = 0;
#pragma unroll
        for (int q = 0; q < B; ++q) sdot += pan<T, B>(D, r, q) * t[q];
        const size_t o = ((size_t)i * B + r) * k + c;
        xout[o] = x[o] + omega * sdot;
    }
}

// Batched row kernels for narrow blocks (B <= 10) and k = 16, 32 or 64
// columns: the k lanes of a block row are the k columns, 64 / k rows per
// wavefront.  Latency is what bounds the per-thread kernels above (one block
// in flight per wavefront, rowptr -> cols -> x serial per block), so here:
// the row's column indices are loaded once, one per lane, and shuffled; the A
// blocks of MK_U consecutive positions (contiguous in storage) are staged in
// LDS by the row's lanes together and read back as broadcasts; each lane
// holds only its column's x_j values, so MK_U blocks are in flight per
// wavefront.  Sums run in the same order as the per-thread kernels.
// MODE 0: y = A x or b - A x; 1: block Jacobi; 2: block GS on one colour
// (rows[]), SKIP as in mk_gs.
constexpr int MK_U = 4;

template <typename T, int B, int MODE, int SKIP>
__global__ __launch_bounds__(256) void mk_row(int k, int nrows, const int* __restrict__ rows, const int* __restrict__ rowptr,
                                              const int* __restrict__ cols, const T* __restrict__ A, const T* __restrict__ Dinv,
                                              const T* __restrict__ b, const T* x, T* y, T omega,
                                              const int* __restrict__ color, int cur)
{
    extern __shared__ __align__(16) unsigned char mk_row_lds[];
    constexpr int B2 = B * B;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rpw = 64 / k, slot = lane / k, c = lane - slot * k, lane0 = slot * k;
    T* La = reinterpret_cast<T*>(mk_row_lds) + (size_t)(wv * rpw + slot) * (MK_U * B2);
    const int ri = (xcd_block() * 4 + wv) * rpw + slot;
    const bool valid = ri < nrows;
    const int i = valid ? (MODE == 2 ? rows[ri] : ri) : 0;
    const int k0 = valid ? rowptr[i] : 0, deg = valid ? rowptr[i + 1] - k0 : 0;
    int wl = deg;
    for (int o = 32; o > 0; o >>= 1) wl = max(wl, __shfl_xor(wl, o, 64));
    T t[B];
#pragma unroll
    for (int r = 0; r < B; ++r) t[r] = (MODE != 0 && valid) ? b[((size_t)i * B + r) * k + c] : (T)0;
    if (reads_none<SKIP>(cur)) wl = 0;
    for (int m0 = 0; m0 < wl; m0 += k) {
        const int mycol = (valid && m0 + c < deg) ? cols[k0 + m0 + c] : -1;
        const int nc = min(k, wl - m0);
        for (int u0 = 0; u0 < nc; u0 += MK_U) {
            const int nb = valid ? max(0, min(MK_U, deg - (m0 + u0))) : 0;
            const T* As = A + (size_t)(k0 + m0 + u0) * B2;
            for (int e = c; e < nb * B2; e += k) La[e] = As[e];
            T xv[MK_U][B];
            bool use[MK_U];
#pragma unroll
            for (int u = 0; u < MK_U; ++u) {
                const int m = u0 + u;
                const int j = __shfl(mycol, lane0 + min(m, k - 1), 64);
                bool ok = valid && m < nc && j >= 0;
                if (MODE == 2 && ok) ok = j != i;
                if (SKIP != 0 && ok) { const int cj = color[j]; ok = SKIP == 1 ? cj <= cur : cj >= cur; }
                use[u] = ok;
#pragma unroll
                for (int q = 0; q < B; ++q) xv[u][q] = ok ? x[((size_t)j * B + q) * k + c] : (T)0;
            }
            wave_sync();
#pragma unroll
            for (int u = 0; u < MK_U; ++u)
                if (use[u]) {
                    const T* Ab = La + u * B2;
#pragma unroll
                    for (int q = 0; q < B; ++q)
#pragma unroll
                        for (int r = 0; r < B; ++r) {
                            if (MODE == 0) t[r] += pan<T, B>(Ab, r, q) * xv[u][q];
                            else t[r] -= pan<T, B>(Ab, r, q) * xv[u][q];
                        }
                }
            wave_sync();
        }
    }
    if (!valid) return;
    if (MODE == 0) {
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const size_t o = ((size_t)i * B + r) * k + c;
            y[o] = b ? (T)(b[o] - t[r]) : t[r];
        }
        return;
    }
    const T* D = Dinv + (size_t)i * B2;
#pragma unroll
    for (int r = 0; r < B; ++r) {
        T sdot = 0;
#pragma unroll
        for (int q = 0; q < B; ++q) sdot += pan<T, B>(D, r, q) * t[q];
        const size_t o = ((size_t)i * B + r) * k + c;
        if (MODE == 1) y[o] = x[o] + omega * sdot;
        else {
            const T xi = SKIP ? (T)0 : y[o];
            y[o] = xi + omega * (sdot - xi);
        }
    }
}

// launch shape of mk_row, or 0 when the per-thread kernels are to be used
template <typename T>
static size_t mk_row_shape(int B, int k)
{
    if (B > 10 || !(k == 16 || k == 32 || k == 64)) return 0;
    const size_t lds = (size_t)4 * (64 / k) * MK_U * B * B * sizeof(T);
    return lds <= 32768 ? lds : 0;
}

#define MGLDG_DISPATCH_SMALL(B, ...)                                                 \
    switch (B) {                                                                     \
        case 1: { constexpr int BB = 1; __VA_ARGS__; } break;                        \
        case 3: { constexpr int BB = 3; __VA_ARGS__; } break;                        \
        case 4: { constexpr int BB = 4; __VA_ARGS__; } break;                        \
        case 6: { constexpr int BB = 6; __VA_ARGS__; } break;                        \
        case 8: { constexpr int BB = 8; __VA_ARGS__; } break;                        \
        case 9: { constexpr int BB = 9; __VA_ARGS__; } break;                        \
        case 10: { constexpr int BB = 10; __VA_ARGS__; } break;                      \
        default: throw std::runtime_error("mk_row: block size " + std::to_string(B)); \
    }

template <typename T, int BF, int BC>
__global__ __launch_bounds__(256) void mk_restrict(int k, int nc, const int* __restrict__ kid_off, const int* __restrict__ kids,
                                                   const int* __restrict__ type, const T* __restrict__ pool,
                                                   const T* __restrict__ r, T* __restrict__ bc)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)nc * k) return;
    const int I = (int)(e / k), c = (int)(e % k);
    T acc[BC];
#pragma unroll
    for (int q = 0; q < BC; ++q) acc[q] = 0;
    for (int t = kid_off[I]; t < kid_off[I + 1]; ++t) {
        const int f = kids[t];
        const T* Tb = pool + (size_t)type[f] * BF * BC;
        const T* rf = r + (size_t)f * BF * k + c;
#pragma unroll
        for (int q = 0; q < BF; ++q) {
            const T rv = rf[(size_t)q * k];
#pragma unroll
            for (int p = 0; p < BC; ++p) acc[p] += Tb[q * BC + p] * rv;
        }
    }
#pragma unroll
    for (int q = 0; q < BC; ++q) bc[((size_t)I * BC + q) * k + c] = acc[q];
}

template <typename T, int BF, int BC>
__global__ __launch_bounds__(256) void mk_prolong(int k, int nf, const int* __restrict__ parent, const int* __restrict__ type,
                                                  const T* __restrict__ pool, const T* __restrict__ xc, T* __restrict__ xf)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)nf * k) return;
    const int f = (int)(e / k), c = (int)(e % k);
    const T* Tb = pool + (size_t)type[f] * BF * BC;
    const T* xp = xc + (size_t)parent[f] * BC * k + c;
    T xv[BC];
#pragma unroll
    for (int q = 0; q < BC; ++q) xv[q] = xp[(size_t)q * k];
#pragma unroll
    for (int r = 0; r < BF; ++r) {
        T acc = 0;
#pragma unroll
        for (int q = 0; q < BC; ++q) acc += Tb[r * BC + q] * xv[q];
        xf[((size_t)f * BF + r) * k + c] += acc;
    }
}

// The same two maps with the interpolation blocks staged in LDS: a workgroup
// takes G = 256 / k block rows (k >= 16), copies the T blocks they use into
// LDS once, and the k column threads of a row read them back as broadcasts
// (the per-thread kernels above issue one cached global load per T entry and
// column).  Sums run in the same order as above.
template <typename T, int BF, int BC>
__global__ __launch_bounds__(256) void mk_restrict_l(int k, int nc, const int* __restrict__ kid_off, const int* __restrict__ kids,
                                                     const int* __restrict__ type, const T* __restrict__ pool,
                                                     const T* __restrict__ r, T* __restrict__ bc)
{
    extern __shared__ __align__(16) unsigned char mk_lds[];
    T* Ts = reinterpret_cast<T*>(mk_lds);
    __shared__ int maxkids;
    __shared__ int kty[256];               // block type of child t of each of the G rows (-1: none)
    constexpr int TB = BF * BC;
    const int G = 256 / k;
    const int I0 = blockIdx.x * G, ng = min(G, nc - I0);
    const int g = threadIdx.x / k, c = threadIdx.x - g * k;
    const bool valid = g < ng;
    const int t0 = valid ? kid_off[I0 + g] : 0, t1 = valid ? kid_off[I0 + g + 1] : 0;
    if (threadIdx.x == 0) {
        int m = 0;
        for (int q = 0; q < ng; ++q) m = max(m, kid_off[I0 + q + 1] - kid_off[I0 + q]);
        maxkids = m;
    }
    __syncthreads();
    const int mk = maxkids;
    T acc[BC];
#pragma unroll
    for (int p = 0; p < BC; ++p) acc[p] = 0;
    for (int t = 0; t < mk; ++t) {
        if (threadIdx.x < G) {
            const int q = threadIdx.x;
            int ty = -1;
            if (q < ng) {
                const int a = kid_off[I0 + q], b = kid_off[I0 + q + 1];
                if (a + t < b) ty = type[kids[a + t]];
            }
            kty[q] = ty;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < ng * TB; e += 256) {
            const int q = e / TB, o = e - q * TB;
            const int ty = kty[q];
            Ts[e] = ty >= 0 ? pool[(size_t)ty * TB + o] : (T)0;
        }
        __syncthreads();
        if (valid && t0 + t < t1) {
            const int f = kids[t0 + t];
            const T* rf = r + (size_t)f * BF * k + c;
            const T* Tb = Ts + g * TB;
#pragma unroll
            for (int q = 0; q < BF; ++q) {
                const T rv = rf[(size_t)q * k];
#pragma unroll
                for (int p = 0; p < BC; ++p) acc[p] += Tb[q * BC + p] * rv;
            }
        }
        __syncthreads();
    }
    if (!valid) return;
#pragma unroll
    for (int p = 0; p < BC; ++p) bc[((size_t)(I0 + g) * BC + p) * k + c] = acc[p];
}

template <typename T, int BF, int BC>
__global__ __launch_bounds__(256) void mk_prolong_l(int k, int nf, const int* __restrict__ parent, const int* __restrict__ type,
                                                    const T* __restrict__ pool, const T* __restrict__ xc, T* __restrict__ xf)
{
    extern __shared__ __align__(16) unsigned char mk_lds[];
    T* Ts = reinterpret_cast<T*>(mk_lds);
    constexpr int TB = BF * BC;
    const int G = 256 / k;
    const int f0 = blockIdx.x * G, ng = min(G, nf - f0);
    for (int e = threadIdx.x; e < ng * TB; e += 256) {
        const int q = e / TB, o = e - q * TB;
        Ts[e] = pool[(size_t)type[f0 + q] * TB + o];
    }
    __syncthreads();
    const int g = threadIdx.x / k, c = threadIdx.x - g * k;
    if (g >= ng) return;
    const int f = f0 + g;
    const T* xp = xc + (size_t)parent[f] * BC * k + c;
    T xv[BC];
#pragma unroll
    for (int q = 0; q < BC; ++q) xv[q] = xp[(size_t)q * k];
    const T* Tb = Ts + g * TB;
#pragma unroll
    for (int rr = 0; rr < BF; ++rr) {
        T acc = 0;
#pragma unroll
        for (int q = 0; q < BC; ++q) acc += Tb[rr * BC + q] * xv[q];
        xf[((size_t)f * BF + rr) * k + c] += acc;
    }
}

// one workgroup per column
template <typename T>
__global__ __launch_bounds__(256) void mk_coarse(int n, int k, const double* __restrict__ Cinv, T* b, T* x, int subtract_mean)
{
    extern __shared__ double bs[];
    const int c = blockIdx.x;
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 256) { double v = (double)b[(size_t)i * k + c]; bs[i] = v; s += v; }
    __shared__ double mean;
    double tot = block_sum(s);
    if (threadIdx.x == 0) mean = subtract_mean ? tot / n : 0.0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) bs[i] -= mean;
    __syncthreads();
    if (!Cinv) { for (int i = threadIdx.x; i < n; i += 256) b[(size_t)i * k + c] = (T)bs[i]; return; }
    for (int i = threadIdx.x; i < n; i += 256) {
        double y = 0;
        for (int j = 0; j < n; ++j) y += Cinv[(size_t)i * n + j] * bs[j];
        x[(size_t)i * k + c] = (T)y;
    }
}

template <typename T>
void MOps<T>::apply(hipStream_t s, int B, int k, int nrows, const int* rowptr, const int* cols, const T* A, const T* x,
                    const T* b, T* y)
{
    long n = (long)nrows * k; if (n <= 0) return;
    if (const size_t lds = mk_row_shape<T>(B, k)) {
        const dim3 g(nblk(nblk(nrows, 64 / k), 4));
        MGLDG_DISPATCH_SMALL(B, hipLaunchKernelGGL((mk_row<T, BB, 0, 0>), g, dim3(256), lds, s, k, nrows, nullptr, rowptr, cols, A,
                                                   nullptr, b, x, y, (T)0, nullptr, 0));
        check_launch("mk_row");
        return;
    }
    MGLDG_DISPATCH(B, hipLaunchKernelGGL((mk_apply<T, BB>), dim3(nblk(n, 256)), dim3(256), 0, s, k, nrows, rowptr, cols, A, x, b, y));
    check_launch("mk_apply");
}
template <typename T>
void MOps<T>::gs_color(hipStream_t s, int B, int k, int nrows, const int* rows, const int* rowptr, const int* cols, const T* A,
                       const T* Dinv, const T* b, T* x, T omega, const int* color, int cur, int skip)
{
    long n = (long)nrows * k; if (n <= 0) return;
    if (const size_t lds = mk_row_shape<T>(B, k)) {
        const dim3 g(nblk(nblk(nrows, 64 / k), 4));
        if (skip == 1) {
            MGLDG_DISPATCH_SMALL(B, hipLaunchKernelGGL((mk_row<T, BB, 2, 1>), g, dim3(256), lds, s, k, nrows, rows, rowptr, cols,
                                                       A, Dinv, b, x, x, omega, color, cur));
        } else if (skip == 2) {
            MGLDG_DISPATCH_SMALL(B, hipLaunchKernelGGL((mk_row<T, BB, 2, 2>), g, dim3(256), lds, s, k, nrows, rows, rowptr, cols,
                                                       A, Dinv, b, x, x, omega, color, cur));
        } else {
            MGLDG_DISPATCH_SMALL(B, hipLaunchKernelGGL((mk_row<T, BB, 2, 0>), g, dim3(256), lds, s, k, nrows, rows, rowptr, cols,
                                                       A, Dinv, b, x, x, omega, color, cur));
        }
        check_launch("mk_row");
        return;
    }
    MGLDG_DISPATCH(B, hipLaunchKernelGGL((mk_gs<T, BB>), dim3(nblk(n, 256)), dim3(256), 0, s, k, nrows, rows, rowptr, cols, A,
                                         Dinv, b, x, omega, color, cur, skip));
    check_launch("mk_gs");
}
// the first batched Jacobi sweep from x = 0: xout = omega D^-1 b, one thread
// per (block row, column)
template <typename T, int B>
__global__ __launch_bounds__(256) void mk_jacobi0(int k, int nrows, const T* __restrict__ Dinv, const T* __restrict__ b,
                                                  T* __restrict__ xout, T omega)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)nrows * k) return;
    const int i = (int)(e / k), c = (int)(e % k);
    T t[B];
#pragma unroll
    for (int r = 0; r < B; ++r) t[r] = b[((size_t)i * B + r) * k + c];
    const T* D = Dinv + (size_t)i * B * B;
#pragma unroll
    for (int r = 0; r < B; ++r) {
        T sdot = 0;
#pragma unroll
        for (int q = 0; q < B; ++q) sdot += pan<T, B>(D, r, q) * t[q];
        xout[((size_t)i * B + r) * k + c] = omega * sdot;
    }
}

template <typename T>
void MOps<T>::jacobi0(hipStream_t s, int B, int k, int nrows, const T* Dinv, const T* b, T* xout, T omega)
{
    const long n = (long)nrows * k;
    if (n <= 0) return;
    MGLDG_DISPATCH(B, hipLaunchKernelGGL((mk_jacobi0<T, BB>), dim3(nblk(n, 256)), dim3(256), 0, s, k, nrows, Dinv, b, xout,
                                         omega));
    check_launch("mk_jacobi0");
}

template <typename T>
void MOps<T>::jacobi(hipStream_t s, int B, int k, int nrows, const int* rowptr, const int* cols, const T* A, const T* Dinv,
                     const T* b, const T* x, T* xout, T omega)
{
    long n = (long)nrows * k; if (n <= 0) return;
    if (const size_t lds = mk_row_shape<T>(B, k)) {
        const dim3 g(nblk(nblk(nrows, 64 / k), 4));
        MGLDG_DISPATCH_SMALL(B, hipLaunchKernelGGL((mk_row<T, BB, 1, 0>), g, dim3(256), lds, s, k, nrows, nullptr, rowptr, cols, A,
                                                   Dinv, b, x, xout, omega, nullptr, 0));
        check_launch("mk_row");
        return;
    }
    MGLDG_DISPATCH(B, hipLaunchKernelGGL((mk_jacobi<T, BB>), dim3(nblk(n, 256)), dim3(256), 0, s, k, nrows, rowptr, cols, A,
                                         Dinv, b, x, xout, omega));
    check_launch("mk_jacobi");
}
template <typename T>
void MOps<T>::restrict_(hipStream_t s, int BF, int BC, int k, int nc, const int* kid_off, const int* kids, const int* type,
                        const T* pool, const T* r, T* bc)
{
    long n = (long)nc * k; if (n <= 0) return;
    const size_t lds = (size_t)(k >= 16 && k <= 256 ? 256 / k : 0) * BF * BC * sizeof(T);
    if (lds > 0 && lds <= 65536) {
        const int G = 256 / k;
        MGLDG_DISPATCH2(BF, BC, hipLaunchKernelGGL((mk_restrict_l<T, FF, CC>), dim3((nc + G - 1) / G), dim3(256), lds, s, k, nc,
                                                   kid_off, kids, type, pool, r, bc));
    } else {
        MGLDG_DISPATCH2(BF, BC, hipLaunchKernelGGL((mk_restrict<T, FF, CC>), dim3(nblk(n, 256)), dim3(256), 0, s, k, nc, kid_off,
                                                   kids, type, pool, r, bc));
    }
    check_launch("mk_restrict");
}
template <typename T>
void MOps<T>::prolong(hipStream_t s, int BF, int BC, int k, int nf, const int* parent, const int* type, const T* pool,
                      const T* xc, T* xf)
{
    long n = (long)nf * k; if (n <= 0) return;
    const size_t lds = (size_t)(k >= 16 && k <= 256 ? 256 / k : 0) * BF * BC * sizeof(T);
    if (lds > 0 && lds <= 65536) {
        const int G = 256 / k;
        MGLDG_DISPATCH2(BF, BC, hipLaunchKernelGGL((mk_prolong_l<T, FF, CC>), dim3((nf + G - 1) / G), dim3(256), lds, s, k, nf,
                                                   parent, type, pool, xc, xf));
    } else {
        MGLDG_DISPATCH2(BF, BC, hipLaunchKernelGGL((mk_prolong<T, FF, CC>), dim3(nblk(n, 256)), dim3(256), 0, s, k, nf, parent,
                                                   type, pool, xc, xf));
    }
    check_launch("mk_prolong");
}
template <typename T>
void MOps<T>::coarse(hipStream_t s, int n, int k, const double* Cinv, T* b, T* x, int subtract_mean)
{
    if (n > 6144) throw std::runtime_error("coarse: dense coarse problem too large");
    if (k <= 0) return;
    hipLaunchKernelGGL(mk_coarse<T>, dim3(k), dim3(256), n * sizeof(double), s, n, k, Cinv, b, x, subtract_mean);
    check_launch("mk_coarse");
}
template struct MOps<float>;
template struct MOps<double>;

// y = W x: one wavefront per row, lanes stride the columns
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_f64(int n, const double* __restrict__ W, const T* __restrict__ x, T* __restrict__ y)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    const double* w = W + (size_t)row * n;
    double acc = 0;
    for (int j = lane; j < n; j += 64) acc += w[j] * (double)x[j];
    acc = wave_sum(acc);
    if (lane == 0) y[row] = (T)acc;
}
// n even: 16-byte loads of W (every row starts on a 16-byte boundary).  W is
// read once per call and is larger than the caches: nontemporal loads keep it
// from evicting the operators and vectors of the rest of the V-cycle.
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_f64x2(int n, const double* __restrict__ W, const T* __restrict__ x,
                                                    T* __restrict__ y)
{
    typedef double dv2 __attribute__((ext_vector_type(2)));
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    const dv2* w = reinterpret_cast<const dv2*>(W + (size_t)row * n);
    double acc = 0;
    for (int j = lane; j < n / 2; j += 64) {
        const dv2 a = __builtin_nontemporal_load(w + j);
        if constexpr (sizeof(T) == 8) {
            const double2 c = reinterpret_cast<const double2*>(x)[j];
            acc += a.x * c.x + a.y * c.y;
        } else {
            acc += a.x * (double)x[2 * j] + a.y * (double)x[2 * j + 1];
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) y[row] = (T)acc;
}
template <typename T>
void gemv_f64(hipStream_t s, int n, const double* W, const T* x, T* y)
{
    if (n <= 0) return;
    if (n % 2 == 0) hipLaunchKernelGGL(k_gemv_f64x2<T>, dim3(nblk(n, 4)), dim3(256), 0, s, n, W, x, y);
    else hipLaunchKernelGGL(k_gemv_f64<T>, dim3(nblk(n, 4)), dim3(256), 0, s, n, W, x, y);
    check_launch("gemv_f64");
}
template void gemv_f64<float>(hipStream_t, int, const double*, const float*, float*);
template void gemv_f64<double>(hipStream_t, int, const double*, const double*, double*);

// y = W x with W in FP32 (the collapsed tail of the FP32 V-cycle): one
// wavefront per row, 16-byte loads of W (n a multiple of 4), FP64 sums
__global__ __launch_bounds__(256) void k_gemv_w32(int n, const float* __restrict__ W, const float* __restrict__ x,
                                                  float* __restrict__ y)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    typedef float fv4 __attribute__((ext_vector_type(4)));
    const fv4* w = reinterpret_cast<const fv4*>(W + (size_t)row * n);
    const float4* xv = reinterpret_cast<const float4*>(x);
    double acc = 0;
    for (int j = lane; j < n / 4; j += 64) {
        const fv4 a = __builtin_nontemporal_load(w + j);
        const float4 c = xv[j];
        acc += (double)a.x * c.x + (double)a.y * c.y + (double)a.z * c.z + (double)a.w * c.w;
    }
    acc = wave_sum(acc);
    if (lane == 0) y[row] = (float)acc;
}
void gemv_w32(hipStream_t s, int n, const float* W, const float* x, float* y)
{
    if (n <= 0) return;
    if (n % 4) throw std::runtime_error("gemv_w32: n must be a multiple of 4");
    hipLaunchKernelGGL(k_gemv_w32, dim3(nblk(n, 4)), dim3(256), 0, s, n, W, x, y);
    check_launch("gemv_w32");
}

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_f64(int n, int k, const double* __restrict__ W, const T* __restrict__ x,
                                                  T* __restrict__ y)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)n * k) return;
    const int i = (int)(e / k), c = (int)(e % k);
    const double* w = W + (size_t)i * n;
    double acc = 0;
    for (int j = 0; j < n; ++j) acc += w[j] * (double)x[(size_t)j * k + c];
    y[e] = (T)acc;
}
template <typename T>
void gemm_f64(hipStream_t s, int n, int k, const double* W, const T* x, T* y)
{
    if (k == 1) { gemv_f64<T>(s, n, W, x, y); return; }
    long t = (long)n * k; if (t <= 0) return;
    hipLaunchKernelGGL(k_gemm_f64<T>, dim3(nblk(t, 256)), dim3(256), 0, s, n, k, W, x, y);
    check_launch("gemm_f64");
}
template void gemm_f64<float>(hipStream_t, int, int, const double*, const float*, float*);
template void gemm_f64<double>(hipStream_t, int, int, const double*, const double*, double*);

__global__ void k_eye(int n, double* X)
{
    long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)n * n) return;
    X[e] = (e / n == e % n) ? 1.0 : 0.0;
}
void eye(hipStream_t s, int n, double* X)
{
    long t = (long)n * n; if (t <= 0) return;
    hipLaunchKernelGGL(k_eye, dim3(nblk(t, 256)), dim3(256), 0, s, n, X);
    check_launch("eye");
}

// --------------------------------------------------------------------- PCG kernels
constexpr int kRedBlocks = 1024;   // grid of the reduction kernels (grid-stride)

// The grid-stride loops of the CG passes take kUnroll elements of a thread
// at a time, their loads issued together (a grid of kRedBlocks workgroups
// holds a few wavefronts per SIMD: one element per round trip would leave
// HBM idle).  Sums run in ascending i as in the plain loop.
constexpr int kUnroll = 4;
#define MGLDG_STRIDED(n, BODY_LOAD, BODY_USE)                                               \
    {                                                                                       \
        const long stride_ = (long)gridDim.x * 256;                                         \
        for (long i0_ = (long)blockIdx.x * 256 + threadIdx.x; i0_ < (n); i0_ += kUnroll * stride_) { \
            _Pragma("unroll") for (int u = 0; u < kUnroll; ++u) {                           \
                const long i = i0_ + u * stride_;                                           \
                if (i < (n)) { BODY_LOAD; }                                                 \
            }                                                                               \
            _Pragma("unroll") for (int u = 0; u < kUnroll; ++u) {                           \
                const long i = i0_ + u * stride_;                                           \
                if (i < (n)) { BODY_USE; }                                                  \
            }                                                                               \
        }                                                                                   \
    }

__global__ __launch_bounds__(256) void k_dot(long n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* part)
{
    double s = 0, av[kUnroll], bv[kUnroll];
    MGLDG_STRIDED(n, (av[u] = a[i], bv[u] = b[i]), (s += av[u] * bv[u]));
    block_partial(s, part);
}

void dot(hipStream_t s, long n, const double* a, const double* b, double* out, double* part)
{
    unsigned g = (unsigned)std::min<long>(kRedBlocks, std::max<long>(1, (n + 255) / 256));
    hipLaunchKernelGGL(k_dot, dim3(g), dim3(256), 0, s, n, a, b, part);
    check_launch("dot");
    finalize(s, part, (int)g, 1, nullptr, 0, out);
}

// The second pass of the CG's reductions is done by the kernel that needs
// the result: each of its workgroups sums the previous launch's partials
// (part[i * NV + k], i < n) in the same fixed order -- the same value in
// every workgroup -- and workgroup 0 stores it in S for the host.  No launch
// is spent on it, and the partials are read from cache.  Every thread gets
// out[0 .. NV).
template <int NV>
__device__ __forceinline__ void wg_sum_partials(const double* __restrict__ part, int n, double* out)
{
    __shared__ double res[NV];
    double t[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) t[k] = 0;
    // a thread's partials i = tid, tid + blockDim, ... are loaded four at a
    // time (one round trip) and added in that order
    for (int base = 0; base < n; base += 4 * (int)blockDim.x) {
        double v[4][NV];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = base + q * (int)blockDim.x + (int)threadIdx.x;
#pragma unroll
            for (int k = 0; k < NV; ++k) v[q][k] = i < n ? part[(size_t)i * NV + k] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int k = 0; k < NV; ++k)
                if (base + q * (int)blockDim.x + (int)threadIdx.x < n) t[k] += v[q][k];
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double v = block_sum(t[k]);
        if (threadIdx.x == 0) res[k] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = res[k];
}

// x += alpha d ; r -= alpha q ; v = (T) r, the V-cycle's level-0 right-hand
// side (one pass instead of an update and a conversion).  alpha = S[DELTA] /
// d.q, d.q summed here from the product's partials (dqp, ndq) and stored in
// S[DQ]
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_xr(long n, double* S, const double* __restrict__ d, const double* __restrict__ q,
                                                double* __restrict__ x, double* __restrict__ r, T* __restrict__ v,
                                                const double* __restrict__ dqp, int ndq)
{
    double dq;
    wg_sum_partials<1>(dqp, ndq, &dq);
    if (blockIdx.x == 0 && threadIdx.x == 0) S[S_DQ] = dq;
    const double alpha = S[S_DELTA] / dq;
    double dv[kUnroll], qv[kUnroll], xv[kUnroll], rv[kUnroll];
    MGLDG_STRIDED(n, (dv[u] = d[i], qv[u] = q[i], xv[u] = x[i], rv[u] = r[i]), {
        x[i] = xv[u] + alpha * dv[u];
        const double ri = rv[u] - alpha * qv[u];
        r[i] = ri; v[i] = (T)ri;
    });
}

static unsigned red_grid(long n) { return (unsigned)std::min<long>(kRedBlocks, std::max<long>(1, (n + 255) / 256)); }

template <typename T>
void pcg_xr(hipStream_t s, long n, double* S, const double* d, const double* q, double* x, double* r, T* v,
            const double* dqp, int ndq)
{
    hipLaunchKernelGGL(k_pcg_xr<T>, dim3(red_grid(n)), dim3(256), 0, s, n, S, d, q, x, r, v, dqp, ndq);
    check_launch("pcg_xr");
}
template void pcg_xr<float>(hipStream_t, long, double*, const double*, const double*, double*, double*, float*,
                            const double*, int);
template void pcg_xr<double>(hipStream_t, long, double*, const double*, const double*, double*, double*, double*,
                             const double*, int);

__global__ __launch_bounds__(256) void k_pcg_x(long n, double* S, const double* __restrict__ d, double* __restrict__ x,
                                               const double* __restrict__ dqp, int ndq)
{
    double dq;
    wg_sum_partials<1>(dqp, ndq, &dq);
    if (blockIdx.x == 0 && threadIdx.x == 0) S[S_DQ] = dq;
    const double alpha = S[S_DELTA] / dq;
    double dv[kUnroll], xv[kUnroll];
    MGLDG_STRIDED(n, (dv[u] = d[i], xv[u] = x[i]), (x[i] = xv[u] + alpha * dv[u]));
}

void pcg_x(hipStream_t s, long n, double* S, const double* d, double* x, const double* dqp, int ndq)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pcg_x, dim3(red_grid(n)), dim3(256), 0, s, n, S, d, x, dqp, ndq);
    check_launch("pcg_x");
}

// three reductions in one pass: r.z, r.zold, r.r (z = the V-cycle's level-0
// result zs, precision T, converted exactly); zold <- zs in the same pass.
// Partials part[3 i + {0, 1, 2}]; S[DELTA_OLD] = S[DELTA] (pcg_d reads it)
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_rz(long n, double* S, const double* __restrict__ r, const T* __restrict__ zs,
                                                T* zold, int read_old, double* part)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) S[S_DELTA_OLD] = S[S_DELTA];
    double a = 0, b = 0, c = 0;
    double rv[kUnroll];
    T zv[kUnroll], ov[kUnroll];
    MGLDG_STRIDED(n, (rv[u] = r[i], zv[u] = zs[i], ov[u] = (zold && read_old) ? zold[i] : (T)0), {
        const double ri = rv[u];
        a += ri * (double)zv[u]; c += ri * ri;
        if (zold) {
            if (read_old) b += ri * (double)ov[u];
            zold[i] = zv[u];
        }
    });
    double ba = block_sum(a), bb = block_sum(b), bc = block_sum(c);
    if (threadIdx.x == 0) { part[3 * blockIdx.x] = ba; part[3 * blockIdx.x + 1] = bb; part[3 * blockIdx.x + 2] = bc; }
}

template <typename T>
int pcg_rz(hipStream_t s, long n, double* S, const double* r, const T* zs, T* zold, bool read_old, double* part)
{
    const unsigned g = red_grid(n);
    hipLaunchKernelGGL(k_pcg_rz<T>, dim3(g), dim3(256), 0, s, n, S, r, zs, zold, (int)(read_old && zold), part);
    check_launch("pcg_rz");
    return (int)g;
}
template int pcg_rz<float>(hipStream_t, long, double*, const double*, const float*, float*, bool, double*);
template int pcg_rz<double>(hipStream_t, long, double*, const double*, const double*, double*, bool, double*);

// d = z + beta d, or d = z (init), z = (double) zs.  The r.z, r.zold, r.r
// partials of pcg_rz are summed here: S[DELTA] = r.z, S[RSOLD] = r.zold,
// S[RR] = r.r, S[BETA] = beta = r.z / S[DELTA_OLD] (flexible: (r.z - r.zold)
// / S[DELTA_OLD]; krylov.h:54-57), 0 for init
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_d(long n, double* S, const T* __restrict__ zs, double* __restrict__ d, int init,
                                               const double* __restrict__ rzp, int nrz, int flexible)
{
    double t[3];
    wg_sum_partials<3>(rzp, nrz, t);
    const double old = S[S_DELTA_OLD];
    const double beta = init ? 0.0 : (flexible ? (t[0] - t[1]) / old : t[0] / old);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S[S_DELTA] = t[0]; S[S_RSOLD] = t[1]; S[S_RR] = t[2]; S[S_BETA] = beta;
    }
    T zv[kUnroll];
    double dv[kUnroll];
    MGLDG_STRIDED(n, (zv[u] = zs[i], dv[u] = init ? 0.0 : d[i]), (d[i] = init ? (double)zv[u] : (double)zv[u] + beta * dv[u]));
}

template <typename T>
void pcg_d(hipStream_t s, long n, double* S, const T* zs, double* d, bool init, const double* rzp, int nrz, bool flexible)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pcg_d<T>, dim3(red_grid(n)), dim3(256), 0, s, n, S, zs, d, (int)init, rzp, nrz, (int)flexible);
    check_launch("pcg_d");
}
template void pcg_d<float>(hipStream_t, long, double*, const float*, double*, bool, const double*, int, bool);
template void pcg_d<double>(hipStream_t, long, double*, const double*, double*, bool, const double*, int, bool);

// ---------------------------------------------------------------------- batched PCG
// thread (group g, column c), c = tid % k, g = tid / k; k <= 256
__device__ __forceinline__ void col_partials(int k, int nval, const double* vals, double* part)
{
    __shared__ double red[3][256];
    for (int v = 0; v < nval; ++v) red[v][threadIdx.x] = vals[v];
    __syncthreads();
    const int groups = 256 / k;
    if (threadIdx.x < k)
        for (int v = 0; v < nval; ++v) {
            double t = 0;
            for (int g = 0; g < groups; ++g) t += red[v][g * k + threadIdx.x];
            part[((size_t)blockIdx.x * nval + v) * k + threadIdx.x] = t;
        }
}

// The batched passes walk rows i = g, g + G, ... (G = gridDim * groups) of
// one column per thread, kUnroll rows at a time with their loads issued
// together; sums in ascending i as in the plain loop.
#define MGLDG_COLSTRIDED(BODY_LOAD, BODY_USE)                                               \
    {                                                                                       \
        const long gs_ = (long)gridDim.x * groups;                                          \
        for (long i0_ = (long)blockIdx.x * groups + g; i0_ < n; i0_ += kUnroll * gs_) {     \
            _Pragma("unroll") for (int u = 0; u < kUnroll; ++u) {                           \
                const long o = (i0_ + u * gs_) * k + c;                                     \
                if (i0_ + u * gs_ < n) { BODY_LOAD; }                                       \
            }                                                                               \
            _Pragma("unroll") for (int u = 0; u < kUnroll; ++u) {                           \
                const long o = (i0_ + u * gs_) * k + c;                                     \
                (void)o;                                                                    \
                if (i0_ + u * gs_ < n) { BODY_USE; }                                        \
            }                                                                               \
        }                                                                                   \
    }

__global__ __launch_bounds__(256) void k_coldot(long n, int k, const double* __restrict__ a, const double* __restrict__ b,
                                                double* part)
{
    const int groups = 256 / k, g = threadIdx.x / k, c = threadIdx.x % k;
    double acc = 0;
    if (g < groups) {
        double av[kUnroll], bv[kUnroll];
        MGLDG_COLSTRIDED((av[u] = a[o], bv[u] = b[o]), (acc += av[u] * bv[u]));
    }
    col_partials(k, 1, &acc, part);
}

// sums per-column partials, one workgroup per column (fixed-order strided
// sums and a fixed tree: deterministic); mode 0: out[c * ostride] = sum ;
// mode 1/2: PCG r.z update
__global__ __launch_bounds__(256) void k_colfinalize(const double* __restrict__ part, int nb, int k, int nval, double* out,
                                                     int ostride, double* S, const int* active, int mode)
{
    __shared__ double red[3][256];
    const int c = blockIdx.x, tid = threadIdx.x;
    double t[3] = {0, 0, 0};
    // b = tid, tid + 256, ... in order; eight partials of each value loaded at once
    int b = tid;
    for (; b + 7 * 256 < nb; b += 8 * 256) {
        double u[8][3];
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int v = 0; v < 3; ++v) u[q][v] = v < nval ? part[((size_t)(b + q * 256) * nval + v) * k + c] : 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int v = 0; v < 3; ++v) t[v] += u[q][v];
    }
    for (; b < nb; b += 256)
        for (int v = 0; v < nval; ++v) t[v] += part[((size_t)b * nval + v) * k + c];
    for (int v = 0; v < nval; ++v) red[v][tid] = t[v];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int v = 0; v < nval; ++v) red[v][tid] += red[v][tid + w];
        __syncthreads();
    }
    if (tid != 0) return;
    if (mode == 0) { out[(size_t)c * ostride] = red[0][0]; return; }
    double* Sc = S + (size_t)c * S_COUNT;
    if (active && !active[c]) return;
    const double old = Sc[S_DELTA];
    Sc[S_DELTA_OLD] = old;
    Sc[S_DELTA] = red[0][0];
    Sc[S_RSOLD] = red[1][0];
    Sc[S_RR] = red[2][0];
    Sc[S_BETA] = (mode == 2) ? (red[0][0] - red[1][0]) / old : red[0][0] / old;
}

static unsigned red_grid(long n, int k)
{
    const int groups = 256 / k;
    return (unsigned)std::min<long>(kRedBlocks, std::max<long>(1, (n + groups - 1) / groups));
}

void colfinalize(hipStream_t s, const double* part, int nb, int k, double* out, int ostride)
{
    hipLaunchKernelGGL(k_colfinalize, dim3(k), dim3(256), 0, s, part, nb, k, 1, out, ostride, nullptr, nullptr, 0);
    check_launch("colfinalize");
}

void coldot(hipStream_t s, long n, int k, const double* a, const double* b, double* out, int ostride, double* part)
{
    if (k > 256) throw std::runtime_error("coldot: more than 256 columns");
    unsigned g = red_grid(n, k);
    hipLaunchKernelGGL(k_coldot, dim3(g), dim3(256), 0, s, n, k, a, b, part);
    check_launch("coldot");
    hipLaunchKernelGGL(k_colfinalize, dim3(k), dim3(256), 0, s, part, (int)g, k, 1, out, ostride, nullptr, nullptr, 0);
    check_launch("colfinalize");
}

// per-column sums of a [n][k] array, then a -= sum / n per column
// (Function::meanZero, src/cartesian/dg/function.h:156-170, for each column)
__global__ __launch_bounds__(256) void k_colsum(long n, int k, const double* __restrict__ a, double* part)
{
    const int groups = 256 / k, g = threadIdx.x / k, c = threadIdx.x % k;
    double acc = 0;
    if (g < groups)
        for (long i = (long)blockIdx.x * groups + g; i < n; i += (long)gridDim.x * groups) acc += a[i * k + c];
    col_partials(k, 1, &acc, part);
}

__global__ void k_colsub(long n, int k, const double* __restrict__ sums, double* __restrict__ a)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * k) return;
    a[e] -= sums[e % k] / (double)n;
}

size_t mean_zero_part_size(long n, int k) { return (size_t)red_grid(n, k) * k + k; }

static_assert(kScratchDoubles >= (size_t)kRedBlocks * 256 + 256, "scratch holds mean_zero's partials for 256 columns");
static __device__ __align__(256) double scratch[kScratchDoubles];

double* scratch_address(int device)
{
    static std::mutex mu;
    static double* addr[256] = {};
    if (device < 0 || device >= 256) throw std::runtime_error("scratch_address: device id out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess || cur != device) throw std::runtime_error("scratch_address: the device must be current");
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        addr[device] = (double*)p;
    }
    return addr[device];
}

void mean_zero(hipStream_t s, long n, int k, double* a, double* part)
{
    if (n <= 0 || k <= 0) return;
    if (k > 256) throw std::runtime_error("mean_zero: more than 256 columns");
    const unsigned g = red_grid(n, k);
    double* sums = part + (size_t)g * k;
    hipLaunchKernelGGL(k_colsum, dim3(g), dim3(256), 0, s, n, k, a, part);
    check_launch("colsum");
    hipLaunchKernelGGL(k_colfinalize, dim3(k), dim3(256), 0, s, part, (int)g, k, 1, sums, 1, nullptr, nullptr, 0);
    check_launch("colfinalize");
    hipLaunchKernelGGL(k_colsub, dim3(nblk(n * k, 256)), dim3(256), 0, s, n, k, sums, a);
    check_launch("colsub");
}

// per active column: x += alpha d ; r -= alpha q ; every column: v = (T) r,
// the V-cycle's level-0 right-hand sides (no r.r here: pcg_rz forms it
// before anyone reads it)
template <typename T>
__global__ __launch_bounds__(256) void k_bpcg_xr(long n, int k, const double* S, const int* active, const double* __restrict__ d,
                                                 const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r,
                                                 T* __restrict__ v)
{
    const int groups = 256 / k, g = threadIdx.x / k, c = threadIdx.x % k;
    if (g >= groups) return;
    const bool act = active[c] != 0;
    const double alpha = act ? S[(size_t)c * S_COUNT + S_DELTA] / S[(size_t)c * S_COUNT + S_DQ] : 0.0;
    double rv[kUnroll], dv[kUnroll], qv[kUnroll], xv[kUnroll];
    MGLDG_COLSTRIDED((rv[u] = r[o], dv[u] = act ? d[o] : 0.0, qv[u] = act ? q[o] : 0.0, xv[u] = act ? x[o] : 0.0), {
        double ri = rv[u];
        if (act) {
            x[o] = xv[u] + alpha * dv[u];
            ri = ri - alpha * qv[u];
            r[o] = ri;
        }
        v[o] = (T)ri;
    });
}

template <typename T>
void bpcg_xr(hipStream_t s, long n, int k, double* S, const int* active, const double* d, const double* q, double* x, double* r,
             T* v, double* part)
{
    unsigned g = red_grid(n, k);
    hipLaunchKernelGGL(k_bpcg_xr<T>, dim3(g), dim3(256), 0, s, n, k, S, active, d, q, x, r, v);
    check_launch("bpcg_xr");
}
template void bpcg_xr<float>(hipStream_t, long, int, double*, const int*, const double*, const double*, double*, double*, float*,
                             double*);
template void bpcg_xr<double>(hipStream_t, long, int, double*, const int*, const double*, const double*, double*, double*,
                              double*, double*);

__global__ void k_bpcg_x(long n, int k, const double* S, const int* active, const double* __restrict__ d, double* __restrict__ x)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * k) return;
    const int c = (int)(e % k);
    if (!active[c]) return;
    x[e] += S[(size_t)c * S_COUNT + S_DELTA] / S[(size_t)c * S_COUNT + S_DQ] * d[e];
}

void bpcg_x(hipStream_t s, long n, int k, const double* S, const int* active, const double* d, double* x)
{
    hipLaunchKernelGGL(k_bpcg_x, dim3(nblk(n * k, 256)), dim3(256), 0, s, n, k, S, active, d, x);
    check_launch("bpcg_x");
}

// per column r.z, r.zold, r.r with z = (double) zs; zold <- zs (every column)
template <typename T>
__global__ __launch_bounds__(256) void k_bpcg_rz(long n, int k, const double* __restrict__ r, const T* __restrict__ zs, T* zold,
                                                 int read_old, double* part)
{
    const int groups = 256 / k, g = threadIdx.x / k, c = threadIdx.x % k;
    double v[3] = {0, 0, 0};
    if (g < groups) {
        double rv[kUnroll];
        T zv[kUnroll], ov[kUnroll];
        MGLDG_COLSTRIDED((rv[u] = r[o], zv[u] = zs[o], ov[u] = (zold && read_old) ? zold[o] : (T)0), {
            const double ri = rv[u];
            v[0] += ri * (double)zv[u]; v[2] += ri * ri;
            if (zold) {
                if (read_old) v[1] += ri * (double)ov[u];
                zold[o] = zv[u];
            }
        });
    }
    col_partials(k, 3, v, part);
}

template <typename T>
void bpcg_rz(hipStream_t s, long n, int k, double* S, const int* active, const double* r, const T* zs, T* zold, bool read_old,
             double* part)
{
    unsigned g = red_grid(n, k);
    const int ro = read_old && zold;
    hipLaunchKernelGGL(k_bpcg_rz<T>, dim3(g), dim3(256), 0, s, n, k, r, zs, zold, ro, part);
    check_launch("bpcg_rz");
    hipLaunchKernelGGL(k_colfinalize, dim3(k), dim3(256), 0, s, part, (int)g, k, 3, nullptr, 0, S, active, ro ? 2 : 1);
    check_launch("colfinalize");
}
template void bpcg_rz<float>(hipStream_t, long, int, double*, const int*, const double*, const float*, float*, bool, double*);
template void bpcg_rz<double>(hipStream_t, long, int, double*, const int*, const double*, const double*, double*, bool, double*);

// d = z + beta d for active columns, or d = z for every column (init); z = (double) zs
template <typename T>
__global__ void k_bpcg_d(long n, int k, const double* S, const int* active, const T* __restrict__ zs, double* __restrict__ d,
                         int init)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * k) return;
    const int c = (int)(e % k);
    if (init) { d[e] = (double)zs[e]; return; }
    if (!active[c]) return;
    d[e] = (double)zs[e] + S[(size_t)c * S_COUNT + S_BETA] * d[e];
}

// the same, four elements of one row per thread (k a multiple of 4): 16-byte
// accesses; the elements of inactive columns are written back unchanged
template <typename T>
__global__ void k_bpcg_d4(long n4, int k, const double* S, const int* active, const T* __restrict__ zs,
                          double* __restrict__ d, int init)
{
    const long e4 = (long)blockIdx.x * 256 + threadIdx.x;
    if (e4 >= n4) return;
    const long e = 4 * e4;
    const int c = (int)(e % k);
    T z[4];
    double dv[4];
    if constexpr (sizeof(T) == 4) {
        const float4 v = *reinterpret_cast<const float4*>(zs + e);
        z[0] = v.x; z[1] = v.y; z[2] = v.z; z[3] = v.w;
    } else {
        const double2 v0 = *reinterpret_cast<const double2*>(zs + e), v1 = *reinterpret_cast<const double2*>(zs + e + 2);
        z[0] = v0.x; z[1] = v0.y; z[2] = v1.x; z[3] = v1.y;
    }
    if (!init) {
        const double2 v0 = *reinterpret_cast<const double2*>(d + e), v1 = *reinterpret_cast<const double2*>(d + e + 2);
        dv[0] = v0.x; dv[1] = v0.y; dv[2] = v1.x; dv[3] = v1.y;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (init) dv[u] = (double)z[u];
        else if (active[c + u]) dv[u] = (double)z[u] + S[(size_t)(c + u) * S_COUNT + S_BETA] * dv[u];
    }
    *reinterpret_cast<double2*>(d + e) = make_double2(dv[0], dv[1]);
    *reinterpret_cast<double2*>(d + e + 2) = make_double2(dv[2], dv[3]);
}

template <typename T>
void bpcg_d(hipStream_t s, long n, int k, const double* S, const int* active, const T* zs, double* d, bool init)
{
    if (k % 4 == 0) {
        const long n4 = n * k / 4;
        hipLaunchKernelGGL(k_bpcg_d4<T>, dim3(nblk(n4, 256)), dim3(256), 0, s, n4, k, S, active, zs, d, (int)init);
    } else {
        hipLaunchKernelGGL(k_bpcg_d<T>, dim3(nblk(n * k, 256)), dim3(256), 0, s, n, k, S, active, zs, d, (int)init);
    }
    check_launch("bpcg_d");
}
template void bpcg_d<float>(hipStream_t, long, int, const double*, const int*, const float*, double*, bool);
template void bpcg_d<double>(hipStream_t, long, int, const double*, const int*, const double*, double*, bool);

// ---------------------------------------------------------------------- boundary data map
__global__ void k_bc_forward(int ne, int B, int nq, int k, int N, const int* __restrict__ bf_off, const int* __restrict__ bf_list,
                             const int* __restrict__ btype, const double* __restrict__ bnormal, double tauD,
                             const double* __restrict__ Lb, const double* __restrict__ g, double* __restrict__ r,
                             double* __restrict__ J0, double* __restrict__ J1, double* __restrict__ J2)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)ne * B * k) return;
    const int c = (int)(e % k);
    const long ei = e / k;
    const int i = (int)(ei % B), el = (int)(ei / B);
    double rr = 0, jj[3] = {0, 0, 0};
    for (int t = bf_off[el]; t < bf_off[el + 1]; ++t) {
        const int f = bf_list[t];
        const double* L = Lb + (size_t)f * B * nq + (size_t)i * nq;
        const double* gf = g + (size_t)f * nq * k + c;
        double lg = 0;
        for (int q = 0; q < nq; ++q) lg += L[q] * gf[(size_t)q * k];
        if (btype[f] == 0) {
            rr += tauD * lg;
            for (int d = 0; d < N; ++d) jj[d] += bnormal[(size_t)f * N + d] * lg;
        } else rr += lg;
    }
    r[e] = rr;
    J0[e] = jj[0];
    if (N > 1) J1[e] = jj[1];
    if (N > 2) J2[e] = jj[2];
}

void bc_forward(hipStream_t s, int ne, int B, int nq, int k, int N, const int* bf_off, const int* bf_list,
                const int* btype, const double* bnormal, double tauD, const double* Lb, const double* g, double* r,
                double* J0, double* J1, double* J2)
{
    long t = (long)ne * B * k; if (t <= 0) return;
    hipLaunchKernelGGL(k_bc_forward, dim3(nblk(t, 256)), dim3(256), 0, s, ne, B, nq, k, N, bf_off, bf_list, btype, bnormal,
                       tauD, Lb, g, r, J0, J1, J2);
    check_launch("bc_forward");
}

__global__ void k_bc_adjoint(int nbf, int B, int nq, int k, int N, const int* __restrict__ bleft,
                             const int* __restrict__ btype, const double* __restrict__ bnormal, double tauD,
                             const double* __restrict__ Lb, const double* __restrict__ lam, const double* __restrict__ G0,
                             const double* __restrict__ G1, const double* __restrict__ G2, double* __restrict__ dg)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nbf * nq * k) return;
    const int c = (int)(e % k);
    const long fq = e / k;
    const int q = (int)(fq % nq), f = (int)(fq / nq);
    const int l = bleft[f];
    const bool dir = btype[f] == 0;
    double n[3] = {0, 0, 0};
    for (int d = 0; d < N; ++d) n[d] = bnormal[(size_t)f * N + d];
    const double* L = Lb + (size_t)f * B * nq + q;
    double acc = 0;
    for (int i = 0; i < B; ++i) {
        const size_t o = ((size_t)l * B + i) * k + c;
        double v = lam[o];
        if (dir) {
            v = tauD * v - n[0] * G0[o];
            if (N > 1) v -= n[1] * G1[o];
            if (N > 2) v -= n[2] * G2[o];
        }
        acc += L[(size_t)i * nq] * v;
    }
    dg[e] = acc;
}

void bc_adjoint(hipStream_t s, int nbf, int B, int nq, int k, int N, const int* bleft, const int* btype,
                const double* bnormal, double tauD, const double* Lb, const double* lam, const double* G0, const double* G1,
                const double* G2, double* dg)
{
    long t = (long)nbf * nq * k; if (t <= 0) return;
    hipLaunchKernelGGL(k_bc_adjoint, dim3(nblk(t, 256)), dim3(256), 0, s, nbf, B, nq, k, N, bleft, btype, bnormal, tauD, Lb,
                       lam, G0, G1, G2, dg);
    check_launch("bc_adjoint");
}

__global__ void k_add_at(long n, const long* __restrict__ idx, double v, double* __restrict__ a)
{
    long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t < n) a[idx[t]] += v;
}
void add_at(hipStream_t s, long n, const long* idx, double v, double* a)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_add_at, dim3(nblk(n, 256)), dim3(256), 0, s, n, idx, v, a);
    check_launch("add_at");
}

__global__ void k_axpby(long n, double as, const double* a, double bs, const double* b, double* y)
{
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = as * a[i] + bs * b[i];
}
void axpby(hipStream_t s, long n, double a_scale, const double* a, double b_scale, const double* b, double* y)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_axpby, dim3(nblk(n, 256)), dim3(256), 0, s, n, a_scale, a, b_scale, b, y);
    check_launch("axpby");
}

__global__ void k_blockdiag(int nb, int B, int nrhs, const double* __restrict__ Mb, const double* __restrict__ x,
                            double* __restrict__ y)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long tot = (long)nb * B * nrhs;
    if (e >= tot) return;
    int c = (int)(e % nrhs); long rr = e / nrhs; int r = (int)(rr % B); long i = rr / B;
    const double* blk = Mb + i * B * B + (long)r * B;
    double s = 0;
    for (int k = 0; k < B; ++k) s += blk[k] * x[(i * B + k) * nrhs + c];
    y[e] = s;
}

void blockdiag(hipStream_t s, int nb, int B, const double* Mblk, const double* x, double* y, int nrhs)
{
    long tot = (long)nb * B * nrhs; if (tot <= 0) return;
    hipLaunchKernelGGL(k_blockdiag, dim3(nblk(tot, 256)), dim3(256), 0, s, nb, B, nrhs, Mblk, x, y);
    check_launch("blockdiag");
}

}  // namespace dev
}  // namespace mgldg
