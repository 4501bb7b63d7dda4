// io.cpp — text writers of the reference's dump and mesh formats (io.hpp).
// Numbers go through printf conversions: "%.17g" is what a C++ stream
// prints for a double under precision(max_digits10) in its default float
// field, and "%d" what it prints for an int.
#include "io.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/mgldg.h"
#include "host.hpp"

namespace mgldg {
namespace io {

namespace {

// a write-only FILE* with a large buffer; close() reports any lost write
struct Out {
    FILE* f = nullptr;
    std::string path;
    std::vector<char> buf;
    explicit Out(const std::string& p) : path(p), buf(1 << 20)
    {
        f = std::fopen(p.c_str(), "w");
        if (!f) throw Error(MGLDG_ERR_IO, "cannot open " + p + " for writing");
        std::setvbuf(f, buf.data(), _IOFBF, buf.size());
    }
    ~Out() { if (f) std::fclose(f); }
    void close()
    {
        const bool bad = std::ferror(f) != 0;
        const bool fail = std::fclose(f) != 0;
        f = nullptr;
        if (bad || fail) throw Error(MGLDG_ERR_IO, "write to " + path + " failed");
    }
};

}  // namespace

void write_mm(const std::string& path, int nbr, int nbc, int br, int bc, const int* rowptr, const int* cols,
              const double* blocks)
{
    if (nbr < 0 || nbc < 0 || br < 1 || bc < 1 || !rowptr || (rowptr[nbr] > 0 && (!cols || !blocks)))
        throw Error(MGLDG_ERR_ARG, "invalid block matrix");
    const long nnzb = rowptr[nbr];
    Out o(path);
    std::fprintf(o.f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(o.f, "%ld %ld %ld\n", (long)nbr * br, (long)nbc * bc, nnzb * br * bc);
    for (int I = 0; I < nbr; ++I) {
        for (long k = rowptr[I]; k < rowptr[I + 1]; ++k) {
            const long r0 = (long)I * br + 1, c0 = (long)cols[k] * bc + 1;
            const double* blk = blocks + k * br * bc;
            for (int i = 0; i < br; ++i)
                for (int j = 0; j < bc; ++j)
                    std::fprintf(o.f, "%ld %ld %.17g\n", r0 + i, c0 + j, blk[i * bc + j]);
        }
    }
    o.close();
}

void write_fun(const std::string& path, int P, int N, int ne, int npl, const double* xyz, const double* u)
{
    if (P < 1 || N < 1 || ne < 0 || npl < 1 || (ne > 0 && (!xyz || !u))) throw Error(MGLDG_ERR_ARG, "invalid function");
    Out o(path);
    std::fprintf(o.f, "%d %d %d\n", P, N, ne);
    for (long e = 0; e < ne; ++e) {
        for (long a = e * npl; a < (e + 1) * npl; ++a) {
            for (int d = 0; d < N; ++d) std::fprintf(o.f, "%.17g ", xyz[a * N + d]);
            std::fprintf(o.f, "%.17g\n", u[a]);
        }
        std::fputc('\n', o.f);
    }
    o.close();
}

void write_mesh(const std::string& path, int N, int np, const double* points, int nt, const int* simplices, int digits)
{
    if (N < 1 || np < 0 || nt < 0 || (np > 0 && !points) || (nt > 0 && !simplices) || digits > 17)
        throw Error(MGLDG_ERR_ARG, "invalid mesh");
    for (long k = 0; k < (long)nt * (N + 1); ++k)
        if (simplices[k] < 0 || simplices[k] >= np) throw Error(MGLDG_ERR_ARG, "vertex id out of range");
    char fmt[16];
    if (digits <= 0) std::strcpy(fmt, "%f");
    else std::snprintf(fmt, sizeof fmt, "%%.%dg", digits);
    Out o(path);
    std::fprintf(o.f, "%d %d\n\n", np, nt);
    for (long i = 0; i < np; ++i) {
        for (int d = 0; d < N; ++d) {
            if (d) std::fputc(' ', o.f);
            std::fprintf(o.f, fmt, points[i * N + d]);
        }
        std::fputc('\n', o.f);
    }
    std::fputc('\n', o.f);
    for (long t = 0; t < nt; ++t) {
        for (int d = 0; d <= N; ++d) std::fprintf(o.f, d ? " %d" : "%d", simplices[t * (N + 1) + d] + 1);
        std::fputc('\n', o.f);
    }
    o.close();
}

void write_dense(const std::string& path, int rows, int cols, const double* a)
{
    if (rows < 0 || cols < 0 || ((long)rows * cols > 0 && !a)) throw Error(MGLDG_ERR_ARG, "invalid dense matrix");
    const long n = (long)rows * cols;
    char s[40];
    int width = 0;
    for (long k = 0; k < n; ++k) width = std::max(width, std::snprintf(s, sizeof s, "%.17g", a[k]));
    Out o(path);
    for (long i = 0; i < rows; ++i) {
        for (long j = 0; j < cols; ++j) {
            std::snprintf(s, sizeof s, "%.17g", a[i * cols + j]);
            std::fprintf(o.f, j ? " %*s" : "%*s", width, s);
        }
        if (i + 1 < rows) std::fputc('\n', o.f);
    }
    std::fputc('\n', o.f);
    o.close();
}

}  // namespace io
}  // namespace mgldg
