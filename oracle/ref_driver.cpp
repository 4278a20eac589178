// ref_driver.cpp -- fixture writer around the reference's own host-side
// readers.  TEST INFRASTRUCTURE ONLY.
//
// Linked (by oracle/Makefile, target `ref`) against the unmodified reference
// sources src/csr.cpp and src/tsp_cpu.cpp read in place, so the CSR builder
// (src/csr.cpp:15-192) and the .tsp reader (src/tsp_cpu.cpp:9-45) that run
// here are the reference's.  Output goes to stdout as one token per value with
// 17 significant digits, after a "BEGIN" marker (the reference's own readers
// print to stdout first).
//
//   ref_driver mtx <file.mtx>     -> size sum nonzeros nrowptr rowptr... col... val...
//   ref_driver dense <file.txt>   -> same, for CSR(vector<double>, dim) (src/csr.cpp:137-155);
//                                    input file: dim followed by dim*dim values
//   ref_driver tsp <file.tsp>     -> ncities dtype then x y per city
#include "csr.cuh"

#include <cstdio>

static void dump_csr(const CSR &c) {
  std::printf("BEGIN\n%d %.17g %d %zu\n", c.size, c.sum, c.nonzeros, c.csr_row_indices.size());
  for (int v : c.csr_row_indices) std::printf("%d ", v);
  std::printf("\n");
  for (int v : c.csr_columns) std::printf("%d ", v);
  std::printf("\n");
  for (double v : c.csr_data) std::printf("%.17g ", v);
  std::printf("\n");
}

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: ref_driver mtx|dense|tsp <file>\n");
    return 2;
  }
  std::string mode = argv[1], path = argv[2];
  if (mode == "mtx") {
    CSR c(path);
    dump_csr(c);
  } else if (mode == "dense") {
    std::ifstream fin(path);
    int dim = 0;
    fin >> dim;
    std::vector<double> m((size_t)dim * dim);
    for (auto &v : m) fin >> v;
    CSR c(m, dim);
    dump_csr(c);
  } else if (mode == "tsp") {
    TSP_CPU t(path);
    std::printf("BEGIN\n%d %s\n", t.ncities, t.dtype.c_str());
    for (int i = 0; i < t.ncities; i++) std::printf("%.17g %.17g\n", t.latitudes[i], t.longitudes[i]);
  } else {
    return 2;
  }
  return 0;
}
