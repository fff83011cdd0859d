// Core LP/QP data types of the ganit host library (same fields as the reference's include/ganit/lp.hpp:14-47,
// which its driver and reader fill and read).  The problem they hold:
//
//   minimize    c'x + 0.5 x'Qx + obj_const
//   subject to  lo <= A x <= hi      (an infinite entry leaves that side of the row open)
//               l  <=  x  <= u       (likewise for the variables)
#pragma once
#include <cstdint>
#include <limits>
#include <string>
#include <vector>

namespace ganit {

constexpr double kInf = std::numeric_limits<double>::infinity();

// Compressed sparse row matrix with 32-bit indices (reference lp.hpp:19-25).
struct Csr {
    int rows = 0, cols = 0;
    std::vector<int> ptr;     // row r's entries are [ptr[r], ptr[r + 1])
    std::vector<int> idx;     // column of each entry, ascending within a row
    std::vector<double> val;  // value of each entry
    int64_t nnz() const { return static_cast<int64_t>(val.size()); }
};

// Triplets -> CSR: rows sorted by column, duplicates summed, exact zeros dropped (reference lp.cpp:9-53).
Csr csr_from_triplets(int rows, int cols, const std::vector<int>& r, const std::vector<int>& c,
                      const std::vector<double>& v);
// Transpose; row i of the result lists the source rows in ascending order (reference lp.cpp:55-72).
Csr transpose(const Csr& a);
// An empty rows x cols matrix.
Csr empty_csr(int rows, int cols);

struct LP {
    std::string name;
    int m = 0, n = 0;
    Csr A;
    std::vector<double> c, l, u, lo, hi;
    double obj_const = 0.0;
    bool maximize = false;
    std::vector<std::string> row_names, col_names;
    std::vector<char> is_int;  // 1 for a column inside an integer MARKER block of the MPS file
    Csr Q;  // the quadratic term: both triangles stored (n x n), no entries for an LP
    bool has_q() const { return Q.nnz() > 0; }
    int num_int() const;
};

// Free-format MPS / QPS reader (reference mps.cpp:69-272), including .gz through `gzip -dc`.
LP read_mps(const std::string& path);
// Free-format MPS writer (QUADOBJ for Q's lower triangle); read_mps(write_mps(lp)) reproduces lp exactly
// (values are printed with 17 significant digits).
void write_mps(const LP& lp, const std::string& path);

}  // namespace ganit
