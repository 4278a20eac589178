// sis_internal.hpp -- host-side types shared by the C-ABI translation units.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "sis.h"

struct sis_csr {
  int64_t n = 0;               // rows (CSR::size)
  std::vector<int64_t> rp;     // n + 1 row pointers (64-bit offsets)
  std::vector<int32_t> col;    // ascending within a row, duplicates kept
  std::vector<double> val;     // FP64 values as read
  double sum = 0.0;            // sum of file values (CSR::sum)
};

struct sis_tsp {
  int64_t n = 0;
  int kind = SIS_TSP_EUC_2D;
  int euc_round = 0;
  std::vector<double> x, y;    // coordinates as read (GEO: radians)
  std::vector<double> rawx, rawy;  // coordinates as written in the file
  std::vector<double> dist;    // packed upper triangle, n(n-1)/2
  std::vector<double> totals;  // D_k
  double maxd = 0.0;
  double sumd = 0.0;
};

namespace sis {

// Thread-local last-error message; returns the status for tail calls.
int fail(int status, const std::string &msg);
void clear_error();

// Packed index of pair c1 < c2 (src/kernels.cu:420).
inline int64_t tri(int64_t c1, int64_t c2, int64_t n) {
  return c2 + c1 * n - ((c1 + 2) * (c1 + 1)) / 2;
}

// Builds CSR rows from coordinates: counting sort by row, (col, val) order
// inside a row -- the order of the reference's lexicographic tuple sort.
int build_csr(int64_t n, const std::vector<int32_t> &rows, const std::vector<int32_t> &cols,
              const std::vector<double> &vals, sis_csr *out);

// xoshiro256** seeded through splitmix64.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed);
  uint64_t next();
  double uniform();                 // [0, 1)
  uint64_t below(uint64_t bound);   // [0, bound)
};

uint64_t splitmix64(uint64_t z);

}  // namespace sis
