// host.hpp — host-side (CPU) part of libmgldg: basis tables, quadtree, mesh,
// LDG face/element assembly and the symbolic phase of every block-sparse
// product.  Numeric products, inversions and the solve run on the GPU
// (kernels.hip, mg.cpp).
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace mgldg {

constexpr int kQuad = 10;   // Gauss points per face dimension (src/cartesian/util/common.h:15)
constexpr int kMaxP = 8;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---------------------------------------------------------------- basis
// Nodal Gauss-Lobatto tensor basis on [0,1]^N (src/cartesian/dg/master.h:12-250).
struct Basis {
    int N = 0, P = 0, npl = 0, nfl = 0;
    double nodes[kMaxP], denom[kMaxP];
    double qx[kQuad], qw[kQuad];
    std::vector<double> mass1, invmass1, diff1;   // P x P
    std::vector<double> mass, invmass;            // npl x npl
    std::vector<double> diff[3];                  // npl x npl per dimension
    std::vector<double> fmass;                    // nfl x nfl ((N-1)-D mass)
    void init(int N, int P);
    // LagrangePoly<P>::eval (master.h:78-116)
    void lagrange1(double x, double* out) const;
    void lagrangeN(const double* x, double* out) const;
};

void gl_nodes(int P, double* out);
void gauss_rule(int Q, double* x, double* w);

// ---------------------------------------------------------------- quadtree
struct QNode {
    double lo[3], hi[3];
    int level, height, parent;
    int first_child;   // children are first_child .. first_child + 2^N - 1, or -1
};

// Quadtree<N> (src/cartesian/mesh/quadtree.{h,tpp}).
struct Quadtree {
    int N = 2;
    bool rapid = false, periodic = false;
    int num_levels = 1;
    std::vector<QNode> nodes;

    void build(const double* lo, const double* hi, void (*h)(const double*, double*, void*), void* user);
    void balance();   // 2:1 face balance (not in the reference)
    bool is_leaf(int id, int c) const;
    std::vector<int> layer(int c) const;
    void neighbors(int id, int dim, int dir, int c, std::vector<int>& out) const;

  private:
    void split(int id);
    bool touches(const QNode& a, const QNode& b, int dim, int dir) const;
    bool contains(const QNode& a, const QNode& b) const;
    void search(int a, int dim, int dir, int id, int c, std::vector<int>& out) const;
};

// ---------------------------------------------------------------- host BSR
// Compressed sparse block rows; blocks row-major br x bc.
struct HostBSR {
    int nbr = 0, nbc = 0, br = 0, bc = 0;
    std::vector<int> rowptr, cols;
    std::vector<double> val;
    long nnzb() const { return (long)cols.size(); }
    int find(int i, int j) const;                 // slot or -1
    double* block(long k) { return val.data() + k * (long)br * bc; }
    const double* block(long k) const { return val.data() + k * (long)br * bc; }
};

// fn(lo, hi) over [0, n) on up to 16 threads (disjoint index ranges)
template <typename F>
inline void parallel_rows(long n, F&& fn)
{
    const long nthr = std::min<long>({16L, (long)std::max(1u, std::thread::hardware_concurrency()), std::max(1L, n / 256)});
    if (nthr <= 1) { fn(0L, n); return; }
    std::vector<std::thread> th;
    const long chunk = (n + nthr - 1) / nthr;
    for (long t = 0; t < nthr; ++t) {
        const long lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([&fn, lo, hi] { fn(lo, hi); });
    }
    for (auto& x : th) x.join();
}

// Pattern from (row, col) pairs: sorted unique per row.
struct PatternBuilder {
    int nbr, nbc;
    std::vector<int> ri, ci;
    PatternBuilder(int m, int n) : nbr(m), nbc(n) {}
    void add(int i, int j) { ri.push_back(i); ci.push_back(j); }
    // values = false: the pattern only (val stays empty)
    void finish(HostBSR& out, int br, int bc, bool values = true);     // allocates zero values
};

// ---------------------------------------------------------------- mesh
struct Face {
    int left, right, dim;    // right < 0: boundary tag -(2d+1) / -(2d+2)
    int canonical;
    double sign;             // normal = sign * e_dim
    double lo[3], hi[3];
    int qoff;                // offset into Mesh::face_q (non-canonical interior), else -1
};

// Mesh<N,P>(qt) (src/cartesian/mesh/mesh.h:186-265)
struct Mesh {
    int N = 2, ne = 0, nf = 0, nbf = 0;
    std::vector<int> gid;                 // element -> quadtree node
    std::vector<double> lo, hi;           // ne x N
    std::vector<Face> faces;
    std::vector<int> bfaces;              // boundary faces in face order
    std::vector<double> face_q;           // local quadrature points of non-canonical faces (left then right)
    void build(const Quadtree& qt, const Basis& B);
    double volume(int e) const;
    double width(int e, int d) const { return hi[e * N + d] - lo[e * N + d]; }
};

// ---------------------------------------------------------------- device assembly plan
// The symbolic half of LDGPoisson::discretize (ldgpoisson.tpp:14-100) in
// the form the device assembly reads.  Every face carries its face
// matrices LL, RR, RL (the lifting terms, ldgpoisson.tpp:60-99) in one of
// four forms:
//   kSlice     conforming interior face: X = fscale * F_mass scattered
//              through the slices of the two sides' codes;
//   kDense     non-conforming interior face: LL, RR, RL at pool blocks
//              fca, fca+1, fca+2, formed on the device by quadrature at the
//              face's points (qface);
//   kDirichlet Dirichlet boundary face: LL = fscale * pool block fca;
//   kNone      Neumann boundary face: no operator term.
// A code names how a face's nodes sit in an element: slice[code] maps face
// node a to element node, islice[code] the element nodes back (-1 off the
// face).  Rows are gathered per element: an element lists its faces in face
// order with its side (ef = face << 1 | side, side 1 = right), the T block
// of (e, other side) and the G block of (e, left) its terms land in, and
// bit 0 / bit 1 of ef_first when it is the first term of that T / G block.
struct AsmPlan {
    enum { kSlice = 0, kDense = 1, kDirichlet = 2, kNone = 3 };
    int N = 2, b = 0, nfl = 0, ncode = 0, nq = 0;
    int gshared = 1;                          // 1: the G_d share one pattern (normals in every direction)
    std::vector<double> mscale, iscale, C;    // per element: M_e = mscale mass, M_e^-1 = iscale invmass,
                                              // G_d(e,e) volume term = sum_j C[e][j][d] diff_j (C: ne x N x N)
    std::vector<int> ef_off, ef, ef_tslot, ef_gslot, ef_first;
    std::vector<int> tdiag, gdiag;            // per element: T block (e,e) or -1; G_d block (e,e) (ne x N)
    std::vector<int> fl, fr, fkind, fca, fcb; // per face
    std::vector<double> fscale, fnormal;      // per face; normals nf x N
    std::vector<double> emass, einv, ediff, fmass;   // reference element: mass, inverse, N diff (b x b); face mass (nfl x nfl)
    std::vector<int> slice, islice;           // ncode x nfl, ncode x b
    std::vector<double> pool;                 // reference b x b blocks (npool_ref of them; the device appends 3 per qface)
    int npool_ref = 0;
    // non-conforming faces: quadrature points (nq x N in the left element,
    // then in the right), weights (area included) and the 1-D Lagrange
    // nodes and denominators of the tensor basis
    std::vector<int> qface;
    std::vector<double> qpts, qwt, lnodes, ldenom;
    // boundary lifts: Lb(boundary face k) = bscale[k] * lbref[bcode[k]] (b x nq)
    std::vector<double> lbref, bscale;
    std::vector<int> bcode;
    // slots, first-term flags, diagonals and entry list from the patterns
    void finish_rows(int ne, const HostBSR& T, const HostBSR* G);
};

// ---------------------------------------------------------------- LDG
// The assembled discretisation handed to the device, whatever the mesh:
// LDGOperators {M, G_d, T} (ldgpoisson.h:15-53; A is formed on the GPU) and
// the boundary data of add_source_terms (ldgpoisson.tpp:103-161).
struct LdgOps {
    int N = 2, p = 1, npl = 0, nq = 0, ne = 0, nf = 0, nbf = 0;
    double tau0 = 0, tauD = 0;
    HostBSR M, T, G[3];                    // patterns (values are formed on the device)
    AsmPlan plan;
    std::vector<int> bleft, btype, btag;   // per boundary face: element, 0 Dirichlet / 1 Neumann, boundary key
    std::vector<double> bnormal;           // per boundary face: outward unit normal (N)
    virtual ~LdgOps() = default;
    // boundary quadrature points (nbf*nq x N) and their boundary keys
    virtual void bc_qpoints(double* xyz, int* tag) const = 0;
    virtual void nodes_xyz(double* xyz) const = 0;
};

// Quadtree discretisation: Mesh<N,P>(qt) + LDGPoisson (src/cartesian/dg/ldgpoisson.tpp)
struct LdgHost : LdgOps {
    int bctype[6] = {0, 0, 0, 0, 0, 0};
    Basis basis;
    Mesh mesh;
    // LDGPoisson::discretize minus construct_laplacian (ldgpoisson.tpp:14-100)
    void assemble(const Quadtree& qt, int p, const int* bct, double tau0, double tauD);
    void bc_qpoints(double* xyz, int* tag) const override;
    void nodes_xyz(double* xyz) const override;
};

// Symbolic phase of C = sum over terms of opL(L) * X * R.  For each output
// block: a run of terms (iL, iX, iR); -1 = identity.
struct TermList {
    std::vector<int> off;   // nout + 1
    std::vector<int> l, x, r;
};

// A = T + sum_d G_d^T M G_d (LDGOperators::construct_laplacian, ldgpoisson.h:34-52):
// pattern of A and the term lists.  Terms reference: L,R pool = concatenated G_d
// blocks (offsets goff[d]), X pool = M blocks (one per row), and T blocks
// (separate run with iL = iR = -1 taken from the T pool).
struct LaplacianSymbolic {
    HostBSR pattern;                    // values unused
    TermList gmg;                       // G^T M G terms (pool: G concat / M)
    std::vector<int> tslot;             // for each A block: slot in T or -1
};
void laplacian_symbolic(const HostBSR& T, const HostBSR* G, int N, const HostBSR& M, LaplacianSymbolic& out);

// One-block-per-row interpolation (InterpolationHierarchy, interpolationhierarchy.h:30-82)
struct Interp {
    int nf = 0, nc = 0, bf = 0, bc = 0;
    std::vector<int> parent;     // fine row -> coarse col
    std::vector<int> type;       // fine row -> block index in pool
    std::vector<double> pool;    // ntypes x (bf x bc) row-major
    int ntypes = 0;
    // restriction gather: coarse row -> fine rows
    std::vector<int> kid_off, kids;
    void make_gather();
};

// Pattern + terms of C = T^T X T (multiply_mm + multiply_mm_t,
// sparseblockmatrix.tpp:234-280, with T one block per row).
void rap_symbolic(const Interp& T, const HostBSR& X, HostBSR& C, TermList& terms);

// greedy colouring in row order
int greedy_colors(const HostBSR& A, std::vector<int>& color);
// wavefront index of every row for the lexicographic sweep in row order
// (reverse = false: a row after all its lower-index neighbours) or in reverse
// order; returns the number of wavefronts
int wavefronts(const HostBSR& A, bool reverse, std::vector<int>& level);

}  // namespace mgldg
