// Host-code robustness program, built by tests/test_host_fuzz.py with g++
// -fsanitize=address,undefined: the golden reader inputs, seeded mutations of
// them (byte flips, truncations, inserted numbers, comments and huge values)
// and the generators at their edges.  Every call must return a status; the
// sanitizers abort on any invalid access, leak or undefined behaviour.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>
#include "sis.h"

static std::string slurp(const char *p) {
  std::ifstream f(p, std::ios::binary);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}
static void spit(const char *p, const std::string &s) {
  std::ofstream f(p, std::ios::binary);
  f << s;
}

// scratch files in the directory given by SIS_FUZZ_DIR (default /tmp)
static std::string wpath(const char *suffix) {
  const char *d = std::getenv("SIS_FUZZ_DIR");
  return std::string(d ? d : "/tmp") + "/sis_fuzz" + suffix;
}

static void exercise_csr(sis_csr *J) {
  int64_t n, nnz;
  double sum;
  sis_csr_shape(J, &n, &nnz, &sum);
  std::vector<int64_t> rp(n + 1);
  std::vector<int32_t> col(nnz > 0 ? nnz : 1);
  std::vector<double> val(nnz > 0 ? nnz : 1);
  sis_csr_export(J, rp.data(), col.data(), val.data());
  double xi0, cut;
  sis_maxcut_xi0(J, &xi0);
  sis_maxcut_random_cut(J, 3, &cut);
  sis_csr_write_mtx(J, wpath(".mtx").c_str());
  sis_csr *K = nullptr;
  if (sis_csr_read_mtx(wpath(".mtx").c_str(), &K) == 0) sis_csr_free(K);
}

static void exercise_tsp(sis_tsp *t) {
  int64_t n;
  double mx, sd;
  sis_tsp_info(t, &n, &mx, &sd);
  std::vector<double> x(n), y(n), d(n * (n - 1) / 2 + 1), tot(n);
  sis_tsp_export(t, x.data(), y.data(), d.data(), tot.data());
  if (n <= 40) {
    sis_csr *J = nullptr;
    std::vector<double> h(n * n);
    if (sis_tsp_qubo(t, &J, h.data()) == 0) {
      exercise_csr(J);
      sis_csr_free(J);
    }
  }
  std::vector<int32_t> tour(n);
  for (int64_t i = 0; i < n; i++) tour[i] = (int32_t)i;
  double len;
  sis_tsp_tour_length(t, tour.data(), &len);
  if (n > 1) {
    tour[0] = tour[1];
    sis_tsp_tour_length(t, tour.data(), &len);  // invalid tour: error
  }
  sis_tsp_random_length(t, 5, &len);
  sis_tsp_write(t, wpath(".tsp").c_str());
  sis_tsp *u = nullptr;
  if (sis_tsp_read(wpath(".tsp").c_str(), 0, &u) == 0) sis_tsp_free(u);
}

int main(int argc, char **argv) {
  std::mt19937_64 rng(12345);
  const int trials = std::getenv("SIS_FUZZ_TRIALS") ? std::atoi(std::getenv("SIS_FUZZ_TRIALS")) : 200;
  int ok = 0, err = 0;
  for (int a = 1; a < argc; a++) {
    const std::string path = argv[a];
    const bool is_tsp = path.size() > 4 && path.substr(path.size() - 4) == ".tsp";
    const std::string base = slurp(argv[a]);
    for (int trial = 0; trial < trials; trial++) {
      std::string s = base;
      if (trial > 0) {
        const int kind = (int)(rng() % 7);
        const size_t pos = s.empty() ? 0 : rng() % s.size();
        if (kind == 0 && !s.empty()) s[pos] = (char)(rng() % 256);                      // byte flip
        else if (kind == 1) s = s.substr(0, pos);                                      // truncate
        else if (kind == 2) s.insert(pos, std::to_string((int64_t)(rng() % 2000000) - 1000000));
        else if (kind == 3) s.insert(pos, "\n% comment\n");
        else if (kind == 4) s.insert(pos, " 99999999999999999999 ");
        else if (kind == 5) s.insert(pos, "-1 ");
        else {
          for (int k = 0; k < 5 && !s.empty(); k++) s[rng() % s.size()] = "0123456789 \n.-e"[rng() % 16];
        }
      }
      const std::string tmps = wpath(is_tsp ? "_in.tsp" : "_in.mtx");
      const char *tmp = tmps.c_str();
      spit(tmp, s);
      if (is_tsp) {
        sis_tsp *t = nullptr;
        if (sis_tsp_read(tmp, (int)(trial & 1), &t) == 0 && t) {
          int64_t n;
          double mx, sd;
          sis_tsp_info(t, &n, &mx, &sd);
          if (n <= 2000) exercise_tsp(t);
          sis_tsp_free(t);
          ok++;
        } else {
          err++;
        }
      } else {
        sis_csr *J = nullptr;
        if (sis_csr_read_mtx(tmp, &J) == 0 && J) {
          int64_t n, nnz;
          double sum;
          sis_csr_shape(J, &n, &nnz, &sum);
          if (n <= 2000000 && nnz <= 20000000) exercise_csr(J);
          sis_csr_free(J);
          ok++;
        } else {
          err++;
        }
      }
    }
  }
  // generators at their edges
  const int64_t ns[] = {0, 1, 2, 3, 17};
  for (int64_t n : ns) {
    sis_csr *J = nullptr;
    for (double p : {0.0, 0.5, 1.0}) {
      if (sis_gen_er(n, p, SIS_W_INT3, 1, &J) == 0) { exercise_csr(J); sis_csr_free(J); }
    }
    for (int d : {0, 1, 2, 3, 4, 16})
      if (sis_gen_regular(n, d, SIS_W_PM1, 2, &J) == 0) { exercise_csr(J); sis_csr_free(J); }
    if (sis_gen_torus(n, n, SIS_W_PM1, 3, &J) == 0) { exercise_csr(J); sis_csr_free(J); }
    if (sis_gen_chung_lu(n, 2.0, 2.1, SIS_W_PM1, 4, &J) == 0) { exercise_csr(J); sis_csr_free(J); }
    sis_tsp *t = nullptr;
    if (sis_gen_tsp(n, 100, 0, 5, &t) == 0) { exercise_tsp(t); sis_tsp_free(t); }
  }
  // coordinate and dense builders with odd inputs
  {
    int32_t r[] = {0, 1, 2, 1, 5}, c[] = {1, 0, 2, 1, 0};
    double v[] = {1.5, -2, 3, 0, 4};
    sis_csr *J = nullptr;
    printf("coo out-of-range: %d\n", sis_csr_from_coo(3, 5, r, c, v, 1, &J));
    if (sis_csr_from_coo(3, 4, r, c, v, 1, &J) == 0) { exercise_csr(J); sis_csr_free(J); }
    double dense[9] = {0, 1, 0, 1, 0, 2, 0, 2, 0};
    if (sis_csr_from_dense(3, dense, &J) == 0) { exercise_csr(J); sis_csr_free(J); }
    double cuts[3] = {1, 2, 2}, st[5];
    sis_cut_stats(cuts, 3, st);
    sis_cut_stats(cuts, 1, st);
    sis_cut_stats(cuts, 0, st);
  }
  printf("parsed ok %d, rejected %d\n", ok, err);
  return 0;
}
