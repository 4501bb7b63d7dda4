// devutil.hpp — small device helpers shared by the row kernels (device code only).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mgldg {
namespace dev {

// Workgroup renumbering for the 8 XCDs of gfx950.  The dispatcher hands
// workgroup b to XCD b % 8, and every XCD has its own L2.  A row kernel whose
// block b covers rows [b R, (b+1) R) therefore spreads neighbouring rows over
// all eight L2s, and a vector block that several neighbouring rows read is
// fetched into each of them.  xcd_block() returns a permutation of
// [0, gridDim.x) under which XCD x runs one contiguous range of logical blocks
// (its share of b % 8 == x, in order), so that neighbouring rows, and the
// vector blocks they share, stay in one L2.  It is a bijection for any grid
// size, so a kernel that indexes its work by it covers the same rows as one
// indexed by blockIdx.x; only the placement changes.
__device__ __forceinline__ int xcd_block()
{
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    // XCD x receives q + (x < r) blocks: b = x, x + 8, ...; logical blocks
    // [x q + min(x, r), ...) in the same order
    return x * q + min(x, r) + (b >> 3);
}

// exactly N contiguous elements at p (no further) into o[0 .. N): one load
// of the whole run (the hardware takes unaligned multi-dword loads).  A
// struct of N elements, not an N-vector type: a 3-vector occupies 4.
template <typename T, int N>
__device__ __forceinline__ void uld(T* o, const T* p)
{
    struct Run { T v[N]; } w;
    static_assert(sizeof(Run) == N * sizeof(T), "no padding");
    __builtin_memcpy(&w, p, sizeof(Run));
#pragma unroll
    for (int e = 0; e < N; ++e) o[e] = w.v[e];
}

// position of the m-th summed block of a row: k0 + m, stepping over the
// diagonal block at dp (dp < 0: no block is stepped over)
__device__ __forceinline__ int blk_at(int k0, int m, int dp)
{
    const int k = k0 + m;
    return (dp >= 0 && k >= dp) ? k + 1 : k;
}

// Gauss-Seidel launches of the first sweep from x = 0 read only neighbours
// already written in sweep order: SKIP 1 (forward) those of colour <= cur,
// SKIP 2 (reverse) those of colour >= cur.  The sweep's first launch reads
// none; the host marks it with cur = kFirstFwd / kFirstRev (kernels.hpp), for
// which the colour tests also fail on every neighbour, and the row kernels
// skip the neighbour pass (its index and colour loads) altogether.
template <int SKIP>
__device__ __forceinline__ bool reads_none(int cur)
{
    return SKIP == 1 ? cur <= kFirstFwd : (SKIP == 2 && cur >= kFirstRev);
}

}  // namespace dev
}  // namespace mgldg
