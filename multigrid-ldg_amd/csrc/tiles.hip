// tiles.hip — single-right-hand-side row kernels of pooled levels on the
// matrix cores.
//
// On a pooled (quadtree) level most block rows repeat the same few block
// sequences: the same D^-1 block and the same off-diagonal blocks, only the
// neighbours' vectors differ.  The vector row kernels read a row's blocks
// from the pool in full for every row, and on large levels the vector L1
// bandwidth of those repeated reads, not HBM, bounds the launch.  Here the
// host groups the rows of a colour (or of the level, for the operator
// application) by their signature (D^-1 or diagonal pool index, number of
// off-diagonal blocks and their pool indices in pool-index order) into
// tiles of 16 rows.  One wavefront owns a tile and forms, for each block
// position u,
//     Y += A_{p_u} [x_{j(i_0,u)} .. x_{j(i_15,u)}]
// as 16 x 16 x 4 MFMAs (v_mfma_f32_16x16x4f32 / _f64_16x16x4f64): the pool
// block is the A operand (rows padded to 16, K to 4 S, S = ceil(B / 4)) and
// the gathered neighbour vectors of the 16 rows are the B operand, so a
// block is read once per 16 rows.  Lane l = 16 q + n supplies, for K step e,
// A[n][S q + e] and X[S q + e][n] (tile row n).  For Gauss-Seidel the
// residual tile s = b - Y (accumulator layout: f32 register e of lane l holds
// row 4 q + e, f64 row 4 e + q, of tile row n) is the B operand of the D^-1
// product as it stands: K index k = row(l, e) at step e.
//
// These launches are bound by the texture addresser and the L1 behind it,
// so the layouts are made for few, wide loads over few cache lines: the
// pools are copied into a tile layout (zero-padded, 16 x 4 x 4 elements per
// block, a wavefront's load instruction reading one contiguous kilobyte), a
// tile row's columns are contiguous (one 16-byte load per four positions),
// a lane's S elements of x_j come in one unaligned load, and in FP32 a
// lane's four accumulator rows of b, x and the result are contiguous.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "kernels.hpp"
#include "devutil.hpp"

namespace mgldg {
namespace dev {

namespace {

template <typename T> struct Tm;
template <> struct Tm<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __host__ __device__ __forceinline__ int row(int lane, int e) { return 4 * (lane >> 4) + e; }
};
template <> struct Tm<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __host__ __device__ __forceinline__ int row(int lane, int e) { return 4 * e + (lane >> 4); }
};

// wave-uniform value (the same in every lane) into a scalar register
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// the tile layout of a pool block: 256 elements; lane (n, q) holds four, and
// each load instruction of a wavefront reads one contiguous kilobyte: FP32
// element e at (n 4 + q) 4 + e, FP64 element e at ((e / 2) 64 + n 4 + q) 2 +
// e % 2 (two halves of 128)
constexpr int kTileBlock = 256;
__host__ __device__ constexpr int tile_at(int tsize, int n, int q, int e)
{
    return tsize == 4 ? (n * 4 + q) * 4 + e : ((e / 2) * 64 + n * 4 + q) * 2 + e % 2;
}
template <typename T>
__device__ __forceinline__ void tile_ld4(T* o, const T* blk, int n, int q)
{
    if constexpr (sizeof(T) == 4) {
        uld<T, 4>(o, blk + tile_at(4, n, q, 0));
    } else {
        uld<T, 2>(o, blk + tile_at(8, n, q, 0));
        uld<T, 2>(o + 2, blk + tile_at(8, n, q, 2));
    }
}

// the tile's column indices (tile row n) and pool indices, slot O + u for
// table position u; slot 0 is the row's own block when O = 1.  WM is the
// table width (a multiple of 4): tcol[(16 t + n) WM + u], tpid[t WM + u]
template <int WM, int O>
struct TileSlots {
    int jj[WM + O], pp[WM + O];
    __device__ __forceinline__ void load(int t, int n, const int* __restrict__ tcol, const int* __restrict__ tpid)
    {
        const int* c = tcol + ((size_t)t * 16 + n) * WM;
#pragma unroll
        for (int u = 0; u < WM; u += 4) {
            const int4 v = *reinterpret_cast<const int4*>(c + u);
            jj[O + u] = v.x; jj[O + u + 1] = v.y; jj[O + u + 2] = v.z; jj[O + u + 3] = v.w;
        }
        const int* pt = tpid + (size_t)uni(t * WM);
#pragma unroll
        for (int u = 0; u < WM; ++u) {
            const int pv = pt[u];
            pp[O + u] = pv >= 0 ? pv : 0;
        }
    }
};

// Y += sum over the tile's block positions of A_{p_u} X_u; SKIP as the
// vector GS kernels (1 / 2: first sweep from zero, neighbours of colours not
// yet updated in sweep order are zero).  With OWN the row's own block comes
// first (x_i with the pool block pown).  Order of issue: every x_j gather,
// then the pool blocks of up to PD positions, then the MFMAs.  Loads use in-range addresses (x_0 for an absent
// neighbour, x_j's last S elements for a lane whose K range lies past B) and
// the values are zeroed afterwards; positions past the tile's blocks have
// column -1 and add zero.  At: the pool in the tile layout (zero padding).
// nblk (wave-uniform): the tile's off-diagonal blocks, positions O ..
// O + nblk - 1 of the table (a tile's rows share their block signature, so
// the table is -1 past them for every row); the positions past them are
// neither loaded nor multiplied.
template <typename T, int B, int SKIP, int WM, bool OWN, bool ALL>
__device__ __forceinline__ void tile_blocks_n(typename Tm<T>::v4& acc, TileSlots<WM, OWN ? 1 : 0>& ts, int n, int q,
                                              int iown, int pown, const T* __restrict__ At, const T* x, int nblk)
{
    constexpr int S = (B + 3) / 4;
    static_assert(B % S == 0, "a lane's K range lies wholly inside or wholly past the block");
    constexpr int O = OWN ? 1 : 0, NS = WM + O;
    int* jj = ts.jj;
    int* pp = ts.pp;
    if constexpr (OWN) { jj[0] = iown; pp[0] = pown; }
    const bool kin = S * q < B;                       // this lane's K range is inside the block
    const int koff = kin ? S * q : B - S;
    // GS tables: a column carries kPidLower when its colour precedes the
    // row's (the neighbours a first sweep from zero has already written)
    bool low[NS];
#pragma unroll
    for (int u = 0; u < NS; ++u) {
        low[u] = (jj[u] & kPidLower) != 0;
        jj[u] = jj[u] >= 0 ? (jj[u] & kPidMask) : -1;
    }
    const int nv = ALL ? NS : O + nblk;               // positions in use (uniform)
    T xb[NS][S];
#pragma unroll
    for (int u = 0; u < NS; ++u)
        if (ALL || u < nv) uld<T, S>(xb[u], x + (size_t)(jj[u] >= 0 ? jj[u] : 0) * B + koff);
    // The pool blocks of PD positions are in flight together: a position's
    // four MFMAs take far less than a load from cache, so with one block
    // ahead every position waited for its own block (a chain of NS round
    // trips).  FP32 holds them all (4 VGPRs each); FP64 eight at a time.
    constexpr int PD = sizeof(T) == 4 ? NS : (NS < 8 ? NS : 8);
    T a[PD][4];
#pragma unroll
    for (int u = 0; u < PD; ++u)
        if (ALL || u < nv) tile_ld4<T>(a[u], At + (size_t)pp[u] * kTileBlock, n, q);
    // every gather is in flight before the first MFMA waits on one (left to
    // itself the scheduler issues each position's gathers next to its MFMAs)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < NS; ++u) {
        if (ALL || u < nv) {
            bool use = jj[u] >= 0 && kin;
            if constexpr (SKIP != 0) if (u >= O) use = use && (SKIP == 1 ? low[u] : !low[u]);
            T* au = a[u % PD];
#pragma unroll
            for (int e = 0; e < S; ++e) acc = Tm<T>::mac(au[e], use ? xb[u][e] : (T)0, acc);
            if (u + PD < (ALL ? NS : nv)) tile_ld4<T>(au, At + (size_t)pp[u + PD] * kTileBlock, n, q);
        }
    }
}

// the table's full width in use (every tile of a uniform grid): no per-
// position tests at all; otherwise the positions past nblk are skipped
template <typename T, int B, int SKIP, int WM, bool OWN>
__device__ __forceinline__ void tile_blocks(typename Tm<T>::v4& acc, TileSlots<WM, OWN ? 1 : 0>& ts, int n, int q,
                                            int iown, int pown, const T* __restrict__ At, const T* x, int nblk)
{
    if (nblk >= WM) tile_blocks_n<T, B, SKIP, WM, OWN, true>(acc, ts, n, q, iown, pown, At, x, nblk);
    else tile_blocks_n<T, B, SKIP, WM, OWN, false>(acc, ts, n, q, iown, pown, At, x, nblk);
}

// a lane's four accumulator-layout elements of a row vector v (block row i,
// rows row(l, e) < B; zero elsewhere): FP32 rows 4 q .. 4 q + 3 contiguous
template <typename T, int B>
__device__ __forceinline__ void acc_load(T* o, const T* v, int i0, int q)
{
    if constexpr (sizeof(T) == 4) {
        if (4 * q + 4 <= B) { uld<T, 4>(o, v + (size_t)i0 * B + 4 * q); return; }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = Tm<T>::row(16 * q, e);
        o[e] = r < B ? v[(size_t)i0 * B + r] : (T)0;
    }
}
template <typename T, int B>
__device__ __forceinline__ void acc_store(T* v, const T* o, int i, int q)
{
    if constexpr (sizeof(T) == 4) {
        if (4 * q + 4 <= B) {
            typedef T V4 __attribute__((ext_vector_type(4)));
            const V4 w = {o[0], o[1], o[2], o[3]};
            __builtin_memcpy(v + (size_t)i * B + 4 * q, &w, sizeof(V4));
            return;
        }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = Tm<T>::row(16 * q, e);
        if (r < B) v[(size_t)i * B + r] = o[e];
    }
}

}  // namespace

// One colour of a damped block GS sweep, one tile of 16 rows per wavefront.
// trow[16 t + n]: row of tile row n (-1: none); tcol[(16 t + n) WM + u]: the
// column of its u-th off-diagonal block (-1: none); tpid[t WM + u]: that
// block's pool index (the same for the tile's rows); tmeta[t] = {blocks,
// D^-1 pool index}.  At / Dt: the A and D^-1 pools in the tile layout (Dt:
// lane (n, q) holds D^-1[n][row(l, e)], e = 0..3).
template <typename T, int B, int SKIP, int WM>
__global__ __launch_bounds__(256) void k_gs_tile(int ntiles, const int* __restrict__ trow, const int* __restrict__ tcol,
                                                 const int* __restrict__ tpid, const int2* __restrict__ tmeta,
                                                 const T* __restrict__ At, const T* __restrict__ Dt,
                                                 const T* __restrict__ b, T* x, T omega, int cur)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = uni(xcd_block() * 4 + wave);
    if (t >= ntiles) return;                          // (whole wavefronts; no workgroup barrier follows)
    const int n = lane & 15, q = lane >> 4;
    TileSlots<WM, 0> ts;
    ts.load(t, n, tcol, tpid);
    const int i = trow[(size_t)t * 16 + n];
    const int2 meta = tmeta[t];
    const int dp = uni(meta.y), nblk = uni(meta.x);
    // the row's own operands (b, x, its D^-1 block), in-range addresses
    const int i0 = i >= 0 ? i : 0;
    T bv[4], xo[4], dd[4];
    acc_load<T, B>(bv, b, i0, q);
    if constexpr (SKIP == 0) acc_load<T, B>(xo, x, i0, q);
    tile_ld4<T>(dd, Dt + (size_t)dp * kTileBlock, n, q);
    typename Tm<T>::v4 acc = {0, 0, 0, 0};
    if (!reads_none<SKIP>(cur)) tile_blocks<T, B, SKIP, WM, false>(acc, ts, n, q, -1, 0, At, x, nblk);
    // s = b - Y in the accumulator layout is the B operand of the D^-1
    // product: K index row(l, e) at step e (Dt is zero past B)
    typename Tm<T>::v4 y = {0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < 4; ++e) y = Tm<T>::mac(dd[e], i >= 0 ? (T)(bv[e] - acc[e]) : (T)0, y);
    if (i < 0) return;
    T out[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) out[e] = SKIP == 0 ? xo[e] + omega * (y[e] - xo[e]) : omega * y[e];
    acc_store<T, B>(x, out, i, q);
}

// y = A x or y = b - A x over tiles of rows with the same blocks: the
// diagonal block (pool index tmeta[t].y) and the off-diagonal blocks of the
// tables, laid out as for the GS tiles; dot partials (y . dotv) per
// workgroup in workgroup order
template <typename T, int B, int WM>
__global__ __launch_bounds__(256) void k_apply_tile(int ntiles, const int* __restrict__ trow, const int* __restrict__ tcol,
                                                    const int* __restrict__ tpid, const int2* __restrict__ tmeta,
                                                    const T* __restrict__ At, const T* __restrict__ x,
                                                    const T* __restrict__ b, T* __restrict__ y,
                                                    const T* __restrict__ dotv, double* part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lb = xcd_block();
    const int t = uni(lb * 4 + wave);
    const bool live = t < ntiles;
    const int tl = live ? t : ntiles - 1;            // (a dead wavefront computes on the last tile, stores nothing)
    const int n = lane & 15, q = lane >> 4;
    TileSlots<WM, 1> ts;
    ts.load(tl, n, tcol, tpid);
    const int iv = trow[(size_t)tl * 16 + n];
    const int i = live ? iv : -1;
    const int i0 = i >= 0 ? i : 0;
    T bv[4];
    acc_load<T, B>(bv, b ? b : x, i0, q);             // (x: an in-range source when there is no b)
    // the dot's other vector: y itself (its values are `out`), or an input
    // loaded with the row's other operands
    const bool dot_self = dotv == y;
    T dw[4] = {0, 0, 0, 0};
    if (dotv && !dot_self) acc_load<T, B>(dw, dotv, i0, q);
    typename Tm<T>::v4 acc = {0, 0, 0, 0};
    tile_blocks<T, B, 0, WM, true>(acc, ts, n, q, i, uni(tmeta[tl].y), At, x, uni(tmeta[tl].x));
    double dv = 0;
    if (i >= 0) {
        T out[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) out[e] = b ? (T)(bv[e] - acc[e]) : acc[e];
        acc_store<T, B>(y, out, i, q);
        if (dotv) {
#pragma unroll
            for (int e = 0; e < 4; ++e) dv += (double)out[e] * (double)(dot_self ? out[e] : dw[e]);
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

// pool blocks (row-major FP64, np blocks of B x B) into the tile layout:
// mode 0 the A-operand layout (lane (n, q): A[n][S q + e], e < S), mode 1 the
// D^-1 layout of the GS kernels (lane (n, q): D[n][row(l, e)], e < 4)
template <typename T>
__global__ void k_tile_pool(long np, int B, const double* __restrict__ src, T* __restrict__ dst, int mode)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= np * kTileBlock) return;
    const long p = g / kTileBlock;
    const int o = (int)(g % kTileBlock);
    int n, q, e;                                      // the inverse of tile_at
    if (sizeof(T) == 4) { n = o / 16; q = (o / 4) % 4; e = o % 4; }
    else { const int h = o / 128, l = (o % 128) / 2; n = l / 4; q = l % 4; e = 2 * h + o % 2; }
    const int S = (B + 3) / 4;
    const int c = mode == 0 ? (e < S ? S * q + e : B) : Tm<T>::row(16 * q, e);
    dst[g] = (n < B && c < B) ? (T)src[(p * B + n) * B + c] : (T)0;
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
// compile-time slot count: the table width rounded up to 4, 8 or 16
#define MGLDG_TILE_WIDTH(W, ...)                                                      \
    if ((W) == 4) { constexpr int WW = 4; __VA_ARGS__; }                              \
    else if ((W) == 8) { constexpr int WW = 8; __VA_ARGS__; }                         \
    else if ((W) == 16) { constexpr int WW = 16; __VA_ARGS__; }                       \
    else throw std::runtime_error("tile kernels: table width " + std::to_string(W) + " not compiled");

bool tile_supported(int B, int W) { return (B == 1 || B == 4 || B == 8 || B == 9 || B == 16) && W >= 1 && W <= 16; }
int tile_width(int W) { return W <= 4 ? 4 : W <= 8 ? 8 : 16; }
int tile_block_elems() { return kTileBlock; }

template <typename T>
void tile_pool(hipStream_t s, long np, int B, const double* src, T* dst, int mode)
{
    if (np <= 0) return;
    const long n = np * kTileBlock;
    hipLaunchKernelGGL(k_tile_pool<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, np, B, src, dst, mode);
    tile_check("tile_pool");
}

template <typename T>
void tile_gs(hipStream_t s, int B, int ntiles, const int* trow, const int* tcol, const int* tpid, const int* tmeta, int W,
             const T* At, const T* Dt, const T* b, T* x, T omega, const int* /*color: the tables carry kPidLower*/, int cur,
             int skip)
{
    if (ntiles <= 0) return;
    const int2* m = reinterpret_cast<const int2*>(tmeta);
    const dim3 g((unsigned)((ntiles + 3) / 4));
    MGLDG_TILE_DISPATCH(B, MGLDG_TILE_WIDTH(W, {
        if (skip == 1)
            hipLaunchKernelGGL((k_gs_tile<T, BB, 1, WW>), g, dim3(256), 0, s, ntiles, trow, tcol, tpid, m, At, Dt, b, x,
                               omega, cur);
        else if (skip == 2)
            hipLaunchKernelGGL((k_gs_tile<T, BB, 2, WW>), g, dim3(256), 0, s, ntiles, trow, tcol, tpid, m, At, Dt, b, x,
                               omega, cur);
        else
            hipLaunchKernelGGL((k_gs_tile<T, BB, 0, WW>), g, dim3(256), 0, s, ntiles, trow, tcol, tpid, m, At, Dt, b, x,
                               omega, cur);
    }));
    tile_check("gs_tile");
}

template <typename T>
int tile_apply(hipStream_t s, int B, int ntiles, const int* trow, const int* tcol, const int* tpid, const int* tmeta,
               int W, const T* At, const T* x, const T* b, T* y, const T* dotv, double* dot_out, double* part)
{
    if (ntiles <= 0) return 0;
    const int2* m = reinterpret_cast<const int2*>(tmeta);
    const unsigned g = (unsigned)((ntiles + 3) / 4);
    double* pw = (dotv && part) ? part : nullptr;
    MGLDG_TILE_DISPATCH(B, MGLDG_TILE_WIDTH(W, {
        hipLaunchKernelGGL((k_apply_tile<T, BB, WW>), dim3(g), dim3(256), 0, s, ntiles, trow, tcol, tpid, m, At, x, b, y,
                           dotv, pw);
    }));
    tile_check("apply_tile");
    if (dot_out) finalize_sum(s, part, (int)g, dot_out);
    return (int)g;
}

template void tile_pool<float>(hipStream_t, long, int, const double*, float*, int);
template void tile_pool<double>(hipStream_t, long, int, const double*, double*, int);
template void tile_gs<float>(hipStream_t, int, int, const int*, const int*, const int*, const int*, int, const float*,
                             const float*, const float*, float*, float, const int*, int, int);
template void tile_gs<double>(hipStream_t, int, int, const int*, const int*, const int*, const int*, int, const double*,
                              const double*, const double*, double*, double, const int*, int, int);
template int tile_apply<float>(hipStream_t, int, int, const int*, const int*, const int*, const int*, int, const float*,
                               const float*, const float*, float*, const float*, double*, double*);
template int tile_apply<double>(hipStream_t, int, int, const int*, const int*, const int*, const int*, int,
                                const double*, const double*, const double*, double*, const double*, double*, double*);

}  // namespace dev
}  // namespace mgldg
