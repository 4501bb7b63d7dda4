// tiles.hip — single-right-hand-side row kernels of pooled levels on the
// matrix cores.
//
// On a pooled (quadtree) level most block rows repeat the same few block
// sequences: the same D^-1 block and the same off-diagonal blocks, only the
// neighbours' vectors differ.  The vector row kernels read a row's blocks
// from the pool in full for every row, and on large levels the vector L1
// bandwidth of those repeated reads, not HBM, bounds the launch.  Here the
// host groups the rows of a colour (or of the level, for the operator
// application) by their signature (D^-1 pool index, number of off-diagonal
// blocks and their pool indices in pool-index order) into tiles of 16 rows.
// One wavefront owns a tile and forms, for each block position u,
//     Y += A_{p_u} [x_{j(i_0,u)} .. x_{j(i_15,u)}]
// as 16 x 16 x 4 MFMAs (v_mfma_f32_16x16x4f32 / _f64_16x16x4f64): the pool
// block is the A operand (rows padded to 16, K to 4 S, S = ceil(B / 4)) and
// the gathered neighbour vectors of the 16 rows are the B operand, so a
// block is read once per 16 rows.  Lane l = 16 q + n supplies, for K step s,
// A[n][S q + s] and X[S q + s][n] (tile row n).  For Gauss-Seidel the
// residual tile s = b - Y (accumulator layout: f32 register e of lane l holds
// row 4 q + e, f64 row 4 e + q, of tile row n) is the B operand of the D^-1
// product as it stands: K index k = row(l, s') at step s'.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "kernels.hpp"
#include "devutil.hpp"

namespace mgldg {
namespace dev {

namespace {

__host__ __device__ constexpr int tpv(int tsize, int B)
{
    return tsize == 8 ? (B % 2 == 0 ? 2 : 1) : (B % 4 == 0 ? 4 : (B % 2 == 0 ? 2 : 1));
}

// element (r, c) of a block in the panel layout of the pools the batched
// kernels read (P32 / P64, DP32 / DP64)
template <typename T, int B>
__device__ __forceinline__ T tile_panel_at(const T* blk, int r, int c)
{
    constexpr int V = tpv(sizeof(T), B);
    return blk[((c / V) * B + r) * V + (c % V)];
}

template <typename T> struct Tm;
template <> struct Tm<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lane, int e) { return 4 * (lane >> 4) + e; }
};
template <> struct Tm<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lane, int e) { return 4 * e + (lane >> 4); }
};

// wave-uniform value (the same in every lane) into a scalar register
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// N contiguous elements at p (N * sizeof(T)-aligned) into o[0 .. N)
template <typename T, int N>
__device__ __forceinline__ void tld(T* o, const T* __restrict__ p)
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

// Y += sum over the tile's block positions of A_{p_u} X_u; SKIP as the
// vector GS kernels (1 / 2: first sweep from zero, neighbours of colours not
// yet updated in sweep order are zero).  Rounds of U positions run over the
// table width W (positions past the tile's nb blocks are zero), so the first
// round's column and pool-index loads do not wait for the tile's block
// count; a round's operands are loaded before any of its MFMAs.  A lane's S
// elements of an A row and of an x_j are contiguous: one vector load each
// where B makes them aligned (B = 16, and B = 4, 8 for the A rows).
template <typename T, int B, int SKIP>
__device__ __forceinline__ void tile_blocks(typename Tm<T>::v4& acc, int t, int nb, int W, int n, int q, bool rowok,
                                            const int* __restrict__ tcol, const int* __restrict__ tpid,
                                            const T* __restrict__ A, const T* x, const int* __restrict__ color, int cur)
{
    constexpr int S = (B + 3) / 4;
    constexpr int V = tpv(sizeof(T), B);
    constexpr int U = sizeof(T) == 4 ? 4 : (S > 2 ? 2 : 4);
    // vector loads: A rows when the lane's S columns are one aligned panel
    // run, x_j when B * sizeof(T) and S * sizeof(T) keep S-runs aligned
    constexpr bool AVEC = (S % V == 0) && (S == V) && (B % 4 == 0);
    constexpr bool XVEC = (B % 4 == 0) && (S * sizeof(T) % 16 == 0 || S * sizeof(T) == 8 * S);
    for (int u0 = 0; u0 < W; u0 += U) {
        T a[U][S], xb[U][S];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool slot = u0 + u < W;
            const int pid = slot ? uni(tpid[(size_t)t * W + u0 + u]) : -1;
            const int j = (slot && rowok) ? tcol[((size_t)t * W + u0 + u) * 16 + n] : -1;
            const bool live = slot && pid >= 0;
            bool use = j >= 0;
            if (SKIP != 0 && use) { const int cj = color[j]; use = SKIP == 1 ? cj <= cur : cj >= cur; }
            const T* blk = A + (size_t)(live ? pid : 0) * (B * B);
            if constexpr (AVEC) {
                if (live && n < B) tld<T, S>(a[u], blk + ((size_t)q * B + n) * V);
                else {
#pragma unroll
                    for (int e = 0; e < S; ++e) a[u][e] = (T)0;
                }
            } else {
#pragma unroll
                for (int e = 0; e < S; ++e) {
                    const int k = S * q + e;
                    a[u][e] = (live && n < B && k < B) ? tile_panel_at<T, B>(blk, n, k) : (T)0;
                }
            }
            if constexpr (XVEC) {
                if (use) tld<T, S>(xb[u], x + (size_t)j * B + S * q);
                else {
#pragma unroll
                    for (int e = 0; e < S; ++e) xb[u][e] = (T)0;
                }
            } else {
#pragma unroll
                for (int e = 0; e < S; ++e) {
                    const int k = S * q + e;
                    xb[u][e] = (use && k < B) ? x[(size_t)j * B + k] : (T)0;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int e = 0; e < S; ++e) acc = Tm<T>::mac(a[u][e], xb[u][e], acc);
    }
    (void)nb;
}

}  // namespace

// One colour of a damped block GS sweep, one tile of 16 rows per wavefront.
// trow[16 t + n]: row of tile row n (-1: none); tcol[(t W + u) 16 + n]: the
// column of its u-th off-diagonal block; tpid[t W + u]: that block's pool
// index (the same for the tile's rows); tmeta[t] = {blocks, D^-1 pool index}.
template <typename T, int B, int SKIP>
__global__ __launch_bounds__(256) void k_gs_tile(int ntiles, const int* __restrict__ trow, const int* __restrict__ tcol,
                                                 const int* __restrict__ tpid, const int2* __restrict__ tmeta, int W,
                                                 const T* __restrict__ A, const T* __restrict__ Dinv,
                                                 const T* __restrict__ b, T* x, T omega, const int* __restrict__ color,
                                                 int cur)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = xcd_block() * 4 + wave;
    if (t >= ntiles) return;                          // (whole wavefronts; no workgroup barrier follows)
    const int n = lane & 15, q = lane >> 4;
    const int i = trow[(size_t)t * 16 + n];
    const int2 meta = tmeta[t];
    const int nb = uni(meta.x), dp = uni(meta.y);
    // the row's own operands (b, x, its D^-1 block) before the neighbour sum
    T bv[4], xo[4], dd[4];
    const T* D = Dinv + (size_t)dp * (B * B);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = Tm<T>::row(lane, e);
        const bool ok = i >= 0 && r < B;
        bv[e] = ok ? b[(size_t)i * B + r] : (T)0;
        xo[e] = (ok && SKIP == 0) ? x[(size_t)i * B + r] : (T)0;
        dd[e] = (n < B && r < B) ? tile_panel_at<T, B>(D, n, r) : (T)0;
    }
    typename Tm<T>::v4 acc = {0, 0, 0, 0};
    if (!reads_none<SKIP>(cur)) tile_blocks<T, B, SKIP>(acc, t, nb, W, n, q, i >= 0, tcol, tpid, A, x, color, cur);
    // s = b - Y in the accumulator layout is the B operand of the D^-1
    // product: K index row(l, e) at step e
    typename Tm<T>::v4 y = {0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < 4; ++e) y = Tm<T>::mac(dd[e], bv[e] - acc[e], y);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = Tm<T>::row(lane, e);
        if (i >= 0 && r < B) x[(size_t)i * B + r] = xo[e] + omega * (y[e] - xo[e]);
    }
}

// y = A x or y = b - A x over tiles of rows with the same blocks (the
// diagonal block among them); dot partials (y . dotv) per workgroup in
// workgroup order
template <typename T, int B>
__global__ __launch_bounds__(256) void k_apply_tile(int ntiles, const int* __restrict__ trow, const int* __restrict__ tcol,
                                                    const int* __restrict__ tpid, const int2* __restrict__ tmeta, int W,
                                                    const T* __restrict__ A, const T* __restrict__ x,
                                                    const T* __restrict__ b, T* __restrict__ y,
                                                    const T* __restrict__ dotv, double* part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lb = xcd_block();
    const int t = lb * 4 + wave;
    const bool live = t < ntiles;
    const int n = lane & 15, q = lane >> 4;
    const int i = live ? trow[(size_t)t * 16 + n] : -1;
    const int nb = live ? uni(tmeta[t].x) : 0;
    T bv[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = Tm<T>::row(lane, e);
        bv[e] = (i >= 0 && r < B && b) ? b[(size_t)i * B + r] : (T)0;
    }
    typename Tm<T>::v4 acc = {0, 0, 0, 0};
    if (live) tile_blocks<T, B, 0>(acc, t, nb, W, n, q, i >= 0, tcol, tpid, A, x, nullptr, 0);
    double dv = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = Tm<T>::row(lane, e);
        if (i >= 0 && r < B) {
            const size_t o = (size_t)i * B + r;
            const T out = b ? (T)(bv[e] - acc[e]) : acc[e];
            y[o] = out;
            if (dotv) dv += (double)out * (double)dotv[o];   // after the store: dotv may be y itself
        }
    }
    if (part) {
        // fixed-order workgroup sum (the partials are summed in workgroup order)
        __shared__ double ws[4];
        for (int o = 32; o > 0; o >>= 1) dv += __shfl_down(dv, o, 64);
        if (lane == 0) ws[wave] = dv;
        __syncthreads();
        if (threadIdx.x == 0) part[lb] = ws[0] + ws[1] + ws[2] + ws[3];
    }
}

static void tile_check(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("kernel launch failed: ") + what + ": " + hipGetErrorString(e));
}

#define MGLDG_TILE_DISPATCH(B, ...)                                                   \
    switch (B) {                                                                     \
        case 1: { constexpr int BB = 1; __VA_ARGS__; } break;                        \
        case 4: { constexpr int BB = 4; __VA_ARGS__; } break;                        \
        case 8: { constexpr int BB = 8; __VA_ARGS__; } break;                        \
        case 9: { constexpr int BB = 9; __VA_ARGS__; } break;                        \
        case 16: { constexpr int BB = 16; __VA_ARGS__; } break;                      \
        default: throw std::runtime_error("tile kernels: block size " + std::to_string(B) + " not compiled"); \
    }

bool tile_block_supported(int B) { return B == 1 || B == 4 || B == 8 || B == 9 || B == 16; }

template <typename T>
void tile_gs(hipStream_t s, int B, int ntiles, const int* trow, const int* tcol, const int* tpid, const int* tmeta, int W,
             const T* A, const T* Dinv, const T* b, T* x, T omega, const int* color, int cur, int skip)
{
    if (ntiles <= 0) return;
    const int2* m = reinterpret_cast<const int2*>(tmeta);
    const dim3 g((unsigned)((ntiles + 3) / 4));
    MGLDG_TILE_DISPATCH(B, {
        if (skip == 1)
            hipLaunchKernelGGL((k_gs_tile<T, BB, 1>), g, dim3(256), 0, s, ntiles, trow, tcol, tpid, m, W, A, Dinv, b, x,
                               omega, color, cur);
        else if (skip == 2)
            hipLaunchKernelGGL((k_gs_tile<T, BB, 2>), g, dim3(256), 0, s, ntiles, trow, tcol, tpid, m, W, A, Dinv, b, x,
                               omega, color, cur);
        else
            hipLaunchKernelGGL((k_gs_tile<T, BB, 0>), g, dim3(256), 0, s, ntiles, trow, tcol, tpid, m, W, A, Dinv, b, x,
                               omega, color, cur);
    });
    tile_check("gs_tile");
}

template <typename T>
void tile_apply(hipStream_t s, int B, int ntiles, const int* trow, const int* tcol, const int* tpid, const int* tmeta,
                int W, const T* A, const T* x, const T* b, T* y, const T* dotv, double* dot_out, double* part)
{
    if (ntiles <= 0) return;
    const int2* m = reinterpret_cast<const int2*>(tmeta);
    const unsigned g = (unsigned)((ntiles + 3) / 4);
    MGLDG_TILE_DISPATCH(B, {
        hipLaunchKernelGGL((k_apply_tile<T, BB>), dim3(g), dim3(256), 0, s, ntiles, trow, tcol, tpid, m, W, A, x, b, y,
                           dotv, dot_out ? part : nullptr);
    });
    tile_check("apply_tile");
    if (dot_out) finalize_sum(s, part, (int)g, dot_out);
}

template void tile_gs<float>(hipStream_t, int, int, const int*, const int*, const int*, const int*, int, const float*,
                             const float*, const float*, float*, float, const int*, int, int);
template void tile_gs<double>(hipStream_t, int, int, const int*, const int*, const int*, const int*, int, const double*,
                              const double*, const double*, double*, double, const int*, int, int);
template void tile_apply<float>(hipStream_t, int, int, const int*, const int*, const int*, const int*, int, const float*,
                                const float*, const float*, float*, const float*, double*, double*);
template void tile_apply<double>(hipStream_t, int, int, const int*, const int*, const int*, const int*, int,
                                 const double*, const double*, const double*, double*, const double*, double*, double*);

}  // namespace dev
}  // namespace mgldg
