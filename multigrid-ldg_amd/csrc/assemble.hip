// assemble.hip — LDGPoisson::discretize (ldgpoisson.tpp:14-100, unstructured
// ldgpoisson.tpp:14-170) on the device: the mass matrix, the broken gradient
// with its volume and lifting terms and the penalty matrix, from the host's
// symbolic plan (patterns, per-element face lists, reference tables).
//
// Rows are gathered, not scattered: one thread owns entry (i, j) of every
// block in the block rows of one element and walks that element's faces in
// face order, so each block is written by exactly one thread, without
// atomics, and the result does not depend on scheduling.  The face matrices
// (LL, RR, RL, ldgpoisson.tpp:60-99) are never stored for conforming faces:
// an entry is read from the reference face mass through the slices of the
// face's codes.  Only non-conforming (hanging) faces of a quadtree get their
// three matrices formed first, by quadrature at the face's points.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "kernels.hpp"

namespace mgldg {
namespace dev {

namespace {

enum { kSlice = 0, kDense = 1, kDirichlet = 2, kNone = 3 };

void check_asm(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("kernel launch failed: ") + what + ": " + hipGetErrorString(e));
}

// LagrangePoly<P>::eval (master.h:78-116) at x: prefix and suffix products
// of (x - x_m) times the barycentric denominators
__device__ inline void lagrange1(int P, const double* nodes, const double* denom, double x, double* out)
{
    double a = 1;
    out[0] = denom[0];
    for (int i = 0; i + 1 < P; ++i) { a *= x - nodes[i]; out[i + 1] = a * denom[i + 1]; }
    double c = 1;
    for (int i = P - 1; i > 0; --i) { c *= x - nodes[i]; out[i - 1] *= c; }
}

// LL, RR, RL of one non-conforming face per workgroup (Mesh::lift,
// mesh.h:282-352): X(i, j) = sum_q (phi_a(q, i) w_q) phi_b(q, j) with the
// tensor basis of side a / b at the face's points.  LDS: the 1-D basis
// values of both sides at every point and dimension.
__global__ void k_asm_quad(int N, int P, int nq, int b, const double* __restrict__ pts, const double* __restrict__ wts,
                           const double* __restrict__ lnodes, const double* __restrict__ ldenom, double* __restrict__ pool)
{
    extern __shared__ double g[];                  // [2][nq][N][P]
    const long f = blockIdx.x;
    const double* x = pts + f * 2 * nq * N;
    const double* w = wts + f * nq;
    for (int t = threadIdx.x; t < 2 * nq * N; t += blockDim.x)
        lagrange1(P, lnodes, ldenom, x[t], g + (size_t)t * P);
    __syncthreads();
    const int bb = b * b;
    double* out = pool + f * 3 * bb;
    for (int t = threadIdx.x; t < 3 * bb; t += blockDim.x) {
        const int m = t / bb, ij = t - m * bb, i = ij / b, j = ij - i * b;
        const int sa = m == 0 ? 0 : 1, sb = m == 1 ? 1 : 0;     // LL: (L, L), RR: (R, R), RL: (R, L)
        int it[3], jt[3];
        for (int d = N - 1, a = i, c = j; d >= 0; --d) { it[d] = a % P; a /= P; jt[d] = c % P; c /= P; }
        double s = 0;
        for (int q = 0; q < nq; ++q) {
            const double* ga = g + ((size_t)(sa * nq + q) * N) * P;
            const double* gb = g + ((size_t)(sb * nq + q) * N) * P;
            double pa = 1, pb = 1;
            for (int d = 0; d < N; ++d) { pa *= ga[d * P + it[d]]; pb *= gb[d * P + jt[d]]; }
            s += (pa * w[q]) * pb;
        }
        out[ij + (size_t)m * bb] = s;
    }
}

// entry (i, j) of a face matrix.  own: LL (side 0) or RR (side 1); cross:
// RL(i, j) with row i on the right element and column j on the left.
__device__ inline double face_own(const AsmRows& a, int kind, int ca, int cb, double sc, int side, int i, int j)
{
    if (kind == kDense) return a.pool[(size_t)(ca + side) * a.b * a.b + (size_t)i * a.b + j];
    const int code = side ? cb : ca;
    const int p = a.islice[(size_t)code * a.b + i], q = a.islice[(size_t)code * a.b + j];
    return (p >= 0 && q >= 0) ? a.fmass[(size_t)p * a.nfl + q] * sc : 0.0;
}

__device__ inline double face_rl(const AsmRows& a, int kind, int ca, int cb, double sc, int i, int j)
{
    if (kind == kDense) return a.pool[(size_t)(ca + 2) * a.b * a.b + (size_t)i * a.b + j];
    const int p = a.islice[(size_t)cb * a.b + i], q = a.islice[(size_t)ca * a.b + j];
    return (p >= 0 && q >= 0) ? a.fmass[(size_t)p * a.nfl + q] * sc : 0.0;
}

__global__ void k_asm_rows(AsmRows a)
{
    const int b = a.b, bb = b * b, N = a.N;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)a.ne * bb) return;
    const int e = (int)(t / bb), ij = (int)(t - (long)e * bb), i = ij / b, j = ij - i * b;
    double* Gd[3] = {a.G0, a.G1, a.G2};
    const double isc = a.iscale[e];
    const double* mi = a.einv + (size_t)i * b;          // row i of the reference inverse mass
    // mass (ldgpoisson.tpp:20-24) and the volume term of the broken gradient
    a.M[(size_t)e * bb + ij] = a.mscale[e] * a.emass[ij];
    double td = 0, g[3] = {0, 0, 0};
    for (int d = 0; d < N; ++d) {
        double s = 0;
        for (int k = 0; k < N; ++k) s += a.C[((size_t)e * N + k) * N + d] * a.ediff[(size_t)k * bb + ij];
        g[d] = s;
    }
    for (int k = a.ef_off[e]; k < a.ef_off[e + 1]; ++k) {
        const int fe = a.ef[k], f = fe >> 1, side = fe & 1;
        const int kind = a.fkind[f], ca = a.fca[f], cb = a.fcb[f];
        const double sc = a.fscale[f];
        const double* nrm = a.fnormal + (size_t)f * N;
        if (kind == kDirichlet) {
            // T(l,l) += tauD LL; G_d(l,l) -= n_d M_l^-1 LL (ldgpoisson.tpp:84-98)
            const double* LL = a.pool + (size_t)ca * bb;
            td += a.tauD * (sc * LL[ij]);
            double w = 0;
            for (int q = 0; q < b; ++q) w += (mi[q] * isc) * (sc * LL[(size_t)q * b + j]);
            for (int d = 0; d < N; ++d) g[d] += -nrm[d] * w;
            continue;
        }
        // interior face (ldgpoisson.tpp:62-83): T(e,e) += tau0 X_own,
        // T(r,l) -= tau0 RL, T(l,r) -= tau0 RL^T
        td += a.tau0 * face_own(a, kind, ca, cb, sc, side, i, j);
        {
            const double cr = side ? face_rl(a, kind, ca, cb, sc, i, j) : face_rl(a, kind, ca, cb, sc, j, i);
            double* dst = a.T + (size_t)a.ef_tslot[k] * bb + ij;
            const double v = -a.tau0 * cr;
            *dst = (a.ef_first[k] & 1) ? v : *dst + v;
        }
        if (!side) continue;
        // e is the right element: G_d(r,r) += n_d M_r^-1 RR, G_d(r,l) -= n_d M_r^-1 RL
        double w1 = 0, w2 = 0;
        if (kind == kSlice) {
            const int* sr = a.slice + (size_t)cb * a.nfl;
            const int jr = a.islice[(size_t)cb * b + j], jl = a.islice[(size_t)ca * b + j];
            for (int p = 0; p < a.nfl; ++p) {
                const double m = mi[sr[p]] * isc;
                if (jr >= 0) w1 += m * (a.fmass[(size_t)p * a.nfl + jr] * sc);
                if (jl >= 0) w2 += m * (a.fmass[(size_t)p * a.nfl + jl] * sc);
            }
        } else {
            const double* RR = a.pool + (size_t)(ca + 1) * bb;
            const double* RL = a.pool + (size_t)(ca + 2) * bb;
            for (int q = 0; q < b; ++q) {
                const double m = mi[q] * isc;
                w1 += m * RR[(size_t)q * b + j];
                w2 += m * RL[(size_t)q * b + j];
            }
        }
        const bool first = (a.ef_first[k] & 2) != 0;
        for (int d = 0; d < N; ++d) {
            g[d] += nrm[d] * w1;
            if (!a.gshared && nrm[d] == 0.0) continue;
            double* dst = Gd[d] + (size_t)a.ef_gslot[k] * bb + ij;
            const double v = -nrm[d] * w2;
            *dst = first ? v : *dst + v;
        }
    }
    if (a.tdiag[e] >= 0) a.T[(size_t)a.tdiag[e] * bb + ij] = td;
    for (int d = 0; d < N; ++d) Gd[d][(size_t)a.gdiag[(size_t)e * N + d] * bb + ij] = g[d];
}

// Lb(face k) = bscale[k] lbref[bcode[k]] (b x nq)
__global__ void k_asm_lift(int nbf, int bnq, const int* __restrict__ bcode, const double* __restrict__ bscale,
                           const double* __restrict__ lbref, double* __restrict__ Lb)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nbf * bnq) return;
    const long k = t / bnq;
    const int r = (int)(t - k * bnq);
    Lb[t] = bscale[k] * lbref[(size_t)bcode[k] * bnq + r];
}

// out block o = (in block src[o])^T, row-major b x b
__global__ void k_transpose_blocks(long nblk, int b, const int* __restrict__ src, const double* __restrict__ in,
                                   double* __restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int bb = b * b;
    if (t >= nblk * bb) return;
    const long o = t / bb;
    const int ij = (int)(t - o * bb), i = ij / b, j = ij - i * b;
    out[t] = in[(size_t)src[o] * bb + (size_t)j * b + i];
}

// 64-bit hash of the bit pattern of each row-major FP64 block (bs words)
__global__ void k_block_hash(long nblk, int bs, const double* __restrict__ in, unsigned long long* __restrict__ out)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nblk) return;
    const unsigned long long* w = reinterpret_cast<const unsigned long long*>(in + k * bs);
    unsigned long long h = 0x243F6A8885A308D3ull ^ (unsigned long long)bs;
    for (int i = 0; i < bs; ++i) {
        h ^= w[i];
        h *= 0x9E3779B97F4A7C15ull;
        h ^= h >> 29;
    }
    out[k] = h;
}

// bad[k] = 1 unless block k equals block rep[pidx[k]] bit for bit
__global__ void k_block_verify(long nblk, int bs, const double* __restrict__ in, const int* __restrict__ pidx,
                               const int* __restrict__ rep, int* __restrict__ bad)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nblk) return;
    const unsigned long long* a = reinterpret_cast<const unsigned long long*>(in + k * bs);
    const unsigned long long* c = reinterpret_cast<const unsigned long long*>(in + (long)rep[pidx[k]] * bs);
    int d = 0;
    for (int i = 0; i < bs; ++i) d |= a[i] != c[i];
    bad[k] = d;
}

unsigned grid_of(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void block_hash(hipStream_t s, long nblk, int bs, const double* in, unsigned long long* out)
{
    if (nblk <= 0) return;
    hipLaunchKernelGGL(k_block_hash, dim3(grid_of(nblk)), dim3(256), 0, s, nblk, bs, in, out);
    check_asm("block_hash");
}

void block_verify(hipStream_t s, long nblk, int bs, const double* in, const int* pidx, const int* rep, int* bad)
{
    if (nblk <= 0) return;
    hipLaunchKernelGGL(k_block_verify, dim3(grid_of(nblk)), dim3(256), 0, s, nblk, bs, in, pidx, rep, bad);
    check_asm("block_verify");
}

void asm_quad(hipStream_t s, int nface, int N, int P, int nq, int b, const double* pts, const double* wts,
              const double* lnodes, const double* ldenom, double* pool)
{
    if (nface <= 0) return;
    const size_t lds = (size_t)2 * nq * N * P * sizeof(double);
    hipLaunchKernelGGL(k_asm_quad, dim3((unsigned)nface), dim3(256), lds, s, N, P, nq, b, pts, wts, lnodes, ldenom, pool);
    check_asm("asm_quad");
}

void asm_rows(hipStream_t s, const AsmRows& a)
{
    const long n = (long)a.ne * a.b * a.b;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_asm_rows, dim3(grid_of(n)), dim3(256), 0, s, a);
    check_asm("asm_rows");
}

void asm_lift(hipStream_t s, int nbf, int b, int nq, const int* bcode, const double* bscale, const double* lbref, double* Lb)
{
    const long n = (long)nbf * b * nq;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_asm_lift, dim3(grid_of(n)), dim3(256), 0, s, nbf, b * nq, bcode, bscale, lbref, Lb);
    check_asm("asm_lift");
}

void transpose_blocks(hipStream_t s, long nblk, int b, const int* src, const double* in, double* out)
{
    const long n = nblk * b * b;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_transpose_blocks, dim3(grid_of(n)), dim3(256), 0, s, nblk, b, src, in, out);
    check_asm("transpose_blocks");
}

}  // namespace dev
}  // namespace mgldg
