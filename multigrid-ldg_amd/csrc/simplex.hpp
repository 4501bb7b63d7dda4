// simplex.hpp — host side of the unstructured (simplex) path: nodal simplex
// basis, mesh from points and simplices, LDG assembly into LdgOps, the dual
// graph of an agglomeration level, an MGridGen-style agglomerator and the
// agglomerated interpolation blocks (P_p simplex -> Q_p box, box -> box).
#pragma once

#include "host.hpp"

namespace mgldg {

// Nodal basis of P_p on the reference N-simplex {x >= 0, sum x <= 1} at the
// equispaced nodes of Master<N,P> (src/unstructured/dg/master.h:80-95; P = p+1,
// node order of SimplexRangeIterator, range.h:291-309, last coordinate
// fastest; p = 0: one node at (1/3, ..., 1/3)).  Built from the monomials of
// degree <= p with exact simplex moments (long double), not from orthogonal
// polynomials: the nodal basis, its mass and its derivative matrices are the
// same functions either way.
struct SimplexBasis {
    int N = 2, P = 1, npl = 1;
    double volume = 0.5;                  // 1/N!
    std::vector<double> nodes;            // npl x N
    std::vector<int> expo;                // npl x N monomial exponents
    std::vector<double> coef;             // phi_i = sum_k coef[k * npl + i] x^expo_k
    std::vector<double> mass, invmass;    // npl x npl
    std::vector<double> diff[3];          // diff[d](i, j) = d phi_j / d x_d at node i
    void init(int N, int P);
    void eval(const double* x, double* phi) const;
};

// Simplex mesh (Mesh<N,P>(wireframe, bndfuns), src/unstructured/mesh/mesh.h:185-298)
struct SimplexFace {
    int left = -1, right = -1;   // right < 0: boundary, key -right
    int v[3] = {0, 0, 0};        // sorted vertex ids
    double normal[3] = {0, 0, 0};
    double area = 0;
};

struct SimplexMesh {
    int N = 2, np = 0, ne = 0, nf = 0, nbf = 0;
    std::vector<double> pts;     // np x N
    std::vector<int> tris;       // ne x (N+1), 0-based
    std::vector<double> vol;     // per element
    std::vector<double> J, Jinv; // per element, N x N row-major (J(:,i) = p_{i+1} - p_0)
    std::vector<SimplexFace> faces;
    std::vector<int> bfaces;     // boundary faces in face order
    // faces in first-appearance order (element i, local face j); vmask[v] bit
    // k = boundary predicate k holds at vertex v (nbnd predicates; 0: key 1)
    void build(int N, int np, const double* p, int nt, const int* t, int nbnd, const int* vmask);
    void face_point(const SimplexFace& f, const double* ref, double* x) const;
    void to_local(int e, const double* x, double* loc) const;
};

// LDGPoisson<N,P> on a simplex mesh (src/unstructured/dg/ldgpoisson.tpp:14-170)
struct SimplexLdg : LdgOps {
    SimplexBasis E, F;              // element and face bases
    SimplexMesh mesh;
    std::vector<double> qx, qw;     // face rule: nq points (N-1 coords) on the reference face
    std::vector<int> bctype;        // per boundary key - 1
    void assemble(int N, int np, const double* pts, int nt, const int* tris, int nbnd, const int* vmask, int p,
                  const int* bct, double tau0, double tauD);
    void bc_qpoints(double* xyz, int* tag) const override;
    void nodes_xyz(double* xyz) const override;
};

// Dual graph of one agglomeration level (src/unstructured/mesh/agglomeration.h:57-93,
// 133-185): vertices are the agglomerates of `orig` (orig[e] = agglomerate of
// mesh element e; null: the elements), edges carry the shared face area,
// vvol the volume and vsurf the boundary-face area.
struct AggGraph {
    std::vector<int> xadj, adj;
    std::vector<double> w, vvol, vsurf;
};
void agglomeration_graph(const SimplexMesh& m, const int* orig, int nparts, AggGraph& g);

// Multilevel agglomeration of one graph (agglomerate.cpp): MGridGen's
// algorithm (lib/ParMGridGen-1.0/MGridGen/Lib/mgridgen.c:18-110) of
// aspect-ratio matching, then k-way multi-objective refinement with
// component splitting, merging and attraction of clusters under minsize.
// Input as MGridGen's: CSR adjacency with face-area weights, vertex volumes
// and boundary-face areas.  Returns the number of clusters; part[v] in
// [0, nparts), each cluster connected and of minsize..maxsize vertices
// where the graph allows.
// seed: the random visiting orders (the default is what every entry point uses)
int mgridgen_partition(int nv, const int* xadj, const double* vvol, const double* vsurf, const int* adjncy,
                       const double* adjwgt, int N, int minsize, int maxsize, std::vector<int>& part,
                       uint64_t seed = 0x6D67726964676E31ull);

// Agglomeration<N,P>(mesh, maxAgglom) (agglomeration.h:21-34): successive
// partitions, level l elements -> level l+1 agglomerates, until one
// agglomerate remains, max_levels are made, or a level no longer shrinks.
std::vector<std::vector<int>> agglomerate(const SimplexMesh& m, int max_levels, int maxsize);

// Agglomerated interpolation (interpolationhierarchy.h:88-204): the
// operators of every level, I[l]: level l -> l+1.  Level 0 rows are P_p
// simplex bases (bf = E.npl).  Each agglomerate's coarse space is the span of
// the tensor Q_p Gauss-Lobatto basis (Q) of its bounding box over its
// elements, represented in an orthonormal basis of that span in the mass
// inner product: eigenvectors of the box basis' Gram matrix over the mesh,
// scaled by w^{-1/2}, directions with w < 1e-10 w_max dropped (zero columns;
// rank[l][a] = kept directions of agglomerate a of level l+1, first).  Block
// smoothers and Galerkin products do not see such block-diagonal changes of
// basis, while the box bases themselves are singular on small or thin
// agglomerates.
void build_interps_agglo(const SimplexMesh& m, const SimplexBasis& E, const Basis& Q,
                         const std::vector<std::vector<int>>& parts, std::vector<Interp>& I,
                         std::vector<std::vector<int>>& rank);

}  // namespace mgldg
