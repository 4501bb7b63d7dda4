// io.hpp — the text formats the reference writes and reads: MatrixMarket
// dumps of block operators, nodal DG functions (.fun), simplex meshes
// (.mesh) and dense matrices (V.dat).  Host code only; every writer throws
// mgldg::Error on an I/O failure.
#pragma once

#include <string>

namespace mgldg {
namespace io {

// SparseBlockMatrix::write (src/cartesian/util/sparseblockmatrix.tpp:145-164):
// "%%MatrixMarket matrix coordinate real general", then "rows cols nnz" with
// nnz = nnzb * br * bc, then one 1-based "i j value" line per entry of every
// stored block (explicit zeros included), values at 17 significant digits.
// Blocks are written in row order, columns ascending within a row.
void write_mm(const std::string& path, int nbr, int nbc, int br, int bc, const int* rowptr, const int* cols,
              const double* blocks);

// Function::write (src/cartesian/dg/function.h:173-190,
// src/unstructured/dg/function.h:159-176): header "P N ne", then per element
// one line per node in storage order with its N coordinates and its
// coefficient, and an empty line after each element; 17 significant digits.
void write_fun(const std::string& path, int P, int N, int ne, int npl, const double* xyz, const double* u);

// The layout lib/distmesh/writemesh.m:17-28 writes and Wireframe<N>(file)
// (src/unstructured/mesh/wireframe.h:22-68) reads: "np nt", an empty line,
// one line of N coordinates per point, an empty line, one line of N + 1
// 1-based vertex ids per simplex.  digits <= 0 prints coordinates as "%f"
// (writemesh.m's six decimals); digits > 0 prints that many significant
// digits (17 round-trips every double).  simplices are 0-based.
void write_mesh(const std::string& path, int N, int np, const double* points, int nt, const int* simplices, int digits);

// A dense matrix as the Eigen stream operator prints it under precision 17
// (the V-cycle matrix dump, tests/cartesian/dg_tests.cpp:337-347): rows on
// lines, every coefficient right-aligned to the widest one, single spaces
// between them, a final newline.  a is row-major.
void write_dense(const std::string& path, int rows, int cols, const double* a);

}  // namespace io
}  // namespace mgldg
