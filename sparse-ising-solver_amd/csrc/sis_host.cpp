// sis_host.cpp -- host half of the C ABI: MatrixMarket and TSPLIB readers,
// CSR construction, the TSP one-hot QUBO, seeded generators and statistics.
// Nothing here touches the GPU.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <sstream>
#include <unordered_map>

#include "sis_internal.hpp"

namespace sis {

static thread_local std::string g_last_error;

int fail(int status, const std::string &msg) {
  g_last_error = msg;
  return status;
}

void clear_error() { g_last_error.clear(); }

uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

Rng::Rng(uint64_t seed) {
  uint64_t z = seed;
  for (int i = 0; i < 4; i++) {
    s[i] = splitmix64(z);
    z += 0x9E3779B97F4A7C15ull;
  }
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Rng::next() {
  const uint64_t r = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return r;
}

double Rng::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

uint64_t Rng::below(uint64_t bound) {
  // Lemire's multiply-shift with rejection.
  unsigned __int128 m = (unsigned __int128)next() * bound;
  uint64_t l = (uint64_t)m;
  if (l < bound) {
    uint64_t t = (0 - bound) % bound;
    while (l < t) {
      m = (unsigned __int128)next() * bound;
      l = (uint64_t)m;
    }
  }
  return (uint64_t)(m >> 64);
}

int build_csr(int64_t n, const std::vector<int32_t> &rows, const std::vector<int32_t> &cols,
              const std::vector<double> &vals, sis_csr *out) {
  const int64_t m = (int64_t)vals.size();
  out->n = n;
  out->rp.assign((size_t)n + 1, 0);
  for (int64_t e = 0; e < m; e++) {
    if (rows[e] < 0 || rows[e] >= n || cols[e] < 0 || cols[e] >= n)
      return fail(SIS_ERR_FORMAT, "entry (" + std::to_string(rows[e] + 1) + ", " +
                                      std::to_string(cols[e] + 1) + ") outside the " +
                                      std::to_string(n) + " x " + std::to_string(n) + " matrix");
    out->rp[(size_t)rows[e] + 1]++;
  }
  for (int64_t i = 0; i < n; i++) out->rp[i + 1] += out->rp[i];
  std::vector<int64_t> fill(out->rp.begin(), out->rp.end() - 1);
  std::vector<std::pair<int32_t, double>> ent((size_t)m);
  for (int64_t e = 0; e < m; e++) ent[(size_t)fill[rows[e]]++] = {cols[e], vals[e]};
  out->col.resize((size_t)m);
  out->val.resize((size_t)m);
  for (int64_t i = 0; i < n; i++) {
    auto b = ent.begin() + out->rp[i], en = ent.begin() + out->rp[i + 1];
    bool sorted = true;
    for (auto it = b; it + 1 < en && sorted; ++it) sorted = !(*(it + 1) < *it);
    if (!sorted) std::sort(b, en);
  }
  for (int64_t e = 0; e < m; e++) {
    out->col[(size_t)e] = ent[(size_t)e].first;
    out->val[(size_t)e] = ent[(size_t)e].second;
  }
  return SIS_OK;
}

// Whitespace tokenizer over an in-memory file with the line operations the
// reference's std::istream use relies on (operator>>, ignore, peek).
struct TextStream {
  std::string buf;
  size_t p = 0;
  bool load(const char *path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    buf = ss.str();
    return true;
  }
  int peek() const { return p < buf.size() ? (unsigned char)buf[p] : -1; }
  void ignore_line() {
    while (p < buf.size() && buf[p] != '\n') p++;
    if (p < buf.size()) p++;
  }
  bool token(std::string &t) {
    while (p < buf.size() && isspace((unsigned char)buf[p])) p++;
    if (p >= buf.size()) return false;
    size_t s = p;
    while (p < buf.size() && !isspace((unsigned char)buf[p])) p++;
    t.assign(buf, s, p - s);
    return true;
  }
  bool skip_ws_in_line() {
    while (p < buf.size() && buf[p] != '\n' && isspace((unsigned char)buf[p])) p++;
    return p < buf.size() && buf[p] != '\n';
  }
  bool int_tok(int64_t &v) {
    while (p < buf.size() && isspace((unsigned char)buf[p])) p++;
    if (p >= buf.size()) return false;
    char *end = nullptr;
    errno = 0;
    long long r = strtoll(buf.c_str() + p, &end, 10);
    if (end == buf.c_str() + p || errno) return false;
    p = (size_t)(end - buf.c_str());
    v = r;
    return true;
  }
  bool dbl_tok(double &v) {
    while (p < buf.size() && isspace((unsigned char)buf[p])) p++;
    if (p >= buf.size()) return false;
    char *end = nullptr;
    double r = strtod(buf.c_str() + p, &end);
    if (end == buf.c_str() + p) return false;
    p = (size_t)(end - buf.c_str());
    v = r;
    return true;
  }
};

}  // namespace sis

using namespace sis;

extern "C" {

const char *sis_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------ MaxCut

int sis_csr_read_mtx(const char *path, sis_csr **out) {
  if (!path || !out) return fail(SIS_ERR_ARG, "sis_csr_read_mtx: null argument");
  *out = nullptr;
  TextStream s;
  if (!s.load(path)) return fail(SIS_ERR_IO, std::string("cannot open ") + path);
  // Banner: words up to "symmetric" or "general" (src/csr.cpp:28-50).
  bool symmetric = true, pattern = false;
  std::string w;
  for (;;) {
    if (!s.token(w))
      return fail(SIS_ERR_FORMAT, std::string(path) +
                                      ": MatrixMarket banner names neither symmetric nor general");
    if (w == "symmetric") { symmetric = true; break; }
    if (w == "general") { symmetric = false; break; }
    if (w == "pattern") pattern = true;
    else if (w == "integer") pattern = false;
    else if (w == "complex" || w == "array" || w == "skew-symmetric" || w == "hermitian")
      return fail(SIS_ERR_FORMAT, std::string(path) + ": unsupported MatrixMarket kind '" + w + "'");
  }
  s.ignore_line();
  while (s.peek() == '%') s.ignore_line();
  int64_t M, Nc, L;
  if (!s.int_tok(M) || !s.int_tok(Nc) || !s.int_tok(L) || M < 0 || L < 0)
    return fail(SIS_ERR_FORMAT, std::string(path) + ": bad size line");
  if (M != Nc)
    return fail(SIS_ERR_FORMAT, std::string(path) + ": MaxCut needs a square matrix");
  if (M > INT32_MAX) return fail(SIS_ERR_FORMAT, std::string(path) + ": too many rows");
  std::vector<int32_t> rows, cols;
  std::vector<double> vals;
  rows.reserve((size_t)(symmetric ? 2 * L : L));
  cols.reserve(rows.capacity());
  vals.reserve(rows.capacity());
  double sum = 0.0;
  for (int64_t l = 0; l < L; l++) {
    int64_t r, c;
    double d = 1.0;
    if (!s.int_tok(r) || !s.int_tok(c) || (!pattern && !s.dbl_tok(d)))
      return fail(SIS_ERR_FORMAT, std::string(path) + ": entry " + std::to_string(l + 1) +
                                      " of " + std::to_string(L) + " is missing or malformed");
    if (r < 1 || r > M || c < 1 || c > M)
      return fail(SIS_ERR_FORMAT, std::string(path) + ": entry " + std::to_string(l + 1) +
                                      " index out of range");
    if (r != c) {  // diagonal dropped (src/csr.cpp:65,80)
      rows.push_back((int32_t)(r - 1));
      cols.push_back((int32_t)(c - 1));
      vals.push_back(d);
      sum += d;
    }
  }
  if (symmetric) {
    size_t m = vals.size();
    for (size_t e = 0; e < m; e++) {
      rows.push_back(cols[e]);
      cols.push_back(rows[e]);
      vals.push_back(vals[e]);
    }
  }
  std::unique_ptr<sis_csr> c(new (std::nothrow) sis_csr());
  if (!c) return fail(SIS_ERR_NOMEM, "out of host memory");
  int st = build_csr(M, rows, cols, vals, c.get());
  if (st) return st;
  c->sum = sum;
  *out = c.release();
  return SIS_OK;
}

int sis_csr_from_coo(int64_t n, int64_t nent, const int32_t *rows, const int32_t *cols,
                     const double *vals, int mirror, sis_csr **out) {
  if (!out || n < 0 || n > INT32_MAX || nent < 0 || (nent > 0 && (!rows || !cols || !vals)))
    return fail(SIS_ERR_ARG, "sis_csr_from_coo: bad argument");
  *out = nullptr;
  std::vector<int32_t> r(rows, rows + nent), c(cols, cols + nent);
  std::vector<double> v(vals, vals + nent);
  double sum = 0.0;
  for (int64_t e = 0; e < nent; e++) sum += vals[e];
  if (mirror) {
    for (int64_t e = 0; e < nent; e++)
      if (rows[e] != cols[e]) {
        r.push_back(cols[e]);
        c.push_back(rows[e]);
        v.push_back(vals[e]);
      }
  }
  std::unique_ptr<sis_csr> m(new (std::nothrow) sis_csr());
  if (!m) return fail(SIS_ERR_NOMEM, "out of host memory");
  int st = build_csr(n, r, c, v, m.get());
  if (st) return st;
  m->sum = sum;
  *out = m.release();
  return SIS_OK;
}

int sis_csr_from_dense(int64_t dim, const double *dense, sis_csr **out) {
  if (!out || dim < 0 || dim > INT32_MAX || (dim > 0 && !dense))
    return fail(SIS_ERR_ARG, "sis_csr_from_dense: bad argument");
  std::unique_ptr<sis_csr> m(new (std::nothrow) sis_csr());
  if (!m) return fail(SIS_ERR_NOMEM, "out of host memory");
  m->n = dim;
  m->rp.assign((size_t)dim + 1, 0);
  for (int64_t i = 0; i < dim; i++) {
    for (int64_t j = 0; j < dim; j++) {
      double v = dense[i * dim + j];
      if (v != 0) {
        m->col.push_back((int32_t)j);
        m->val.push_back(v);
      }
    }
    m->rp[(size_t)i + 1] = (int64_t)m->col.size();
  }
  m->sum = 0.0;
  *out = m.release();
  return SIS_OK;
}

int sis_csr_shape(const sis_csr *c, int64_t *n, int64_t *nnz, double *sum) {
  if (!c) return fail(SIS_ERR_ARG, "sis_csr_shape: null csr");
  if (n) *n = c->n;
  if (nnz) *nnz = (int64_t)c->col.size();
  if (sum) *sum = c->sum;
  return SIS_OK;
}

int sis_csr_export(const sis_csr *c, int64_t *row_ptr, int32_t *col, double *val) {
  if (!c) return fail(SIS_ERR_ARG, "sis_csr_export: null csr");
  if (row_ptr) std::memcpy(row_ptr, c->rp.data(), sizeof(int64_t) * c->rp.size());
  if (col && !c->col.empty()) std::memcpy(col, c->col.data(), sizeof(int32_t) * c->col.size());
  if (val && !c->val.empty()) std::memcpy(val, c->val.data(), sizeof(double) * c->val.size());
  return SIS_OK;
}

static bool csr_is_symmetric_offdiag(const sis_csr *c) {
  // Multiset of (row, col, val) equals that of (col, row, val), no diagonal.
  std::vector<std::tuple<int32_t, int32_t, double>> a, b;
  for (int64_t i = 0; i < c->n; i++)
    for (int64_t k = c->rp[i]; k < c->rp[i + 1]; k++) {
      if (c->col[k] == i) return false;
      a.emplace_back((int32_t)i, c->col[k], c->val[k]);
      b.emplace_back(c->col[k], (int32_t)i, c->val[k]);
    }
  std::sort(a.begin(), a.end());
  std::sort(b.begin(), b.end());
  return a == b;
}

int sis_csr_write_mtx(const sis_csr *c, const char *path) {
  if (!c || !path) return fail(SIS_ERR_ARG, "sis_csr_write_mtx: null argument");
  bool sym = csr_is_symmetric_offdiag(c);
  bool integral = true;
  for (double v : c->val)
    if (v != std::floor(v) || std::fabs(v) > 9.007199254740992e15) { integral = false; break; }
  FILE *f = std::fopen(path, "w");
  if (!f) return fail(SIS_ERR_IO, std::string("cannot write ") + path);
  int64_t L = 0;
  for (int64_t i = 0; i < c->n; i++)
    for (int64_t k = c->rp[i]; k < c->rp[i + 1]; k++)
      if (!sym || c->col[k] < i) L++;
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate %s %s\n", integral ? "integer" : "real",
               sym ? "symmetric" : "general");
  std::fprintf(f, "%lld %lld %lld\n", (long long)c->n, (long long)c->n, (long long)L);
  for (int64_t i = 0; i < c->n; i++)
    for (int64_t k = c->rp[i]; k < c->rp[i + 1]; k++) {
      if (sym && c->col[k] >= i) continue;
      if (integral)
        std::fprintf(f, "%lld %d %lld\n", (long long)(i + 1), c->col[k] + 1, (long long)c->val[k]);
      else
        std::fprintf(f, "%lld %d %.17g\n", (long long)(i + 1), c->col[k] + 1, c->val[k]);
    }
  bool ok = std::fclose(f) == 0;
  return ok ? SIS_OK : fail(SIS_ERR_IO, std::string("short write to ") + path);
}

void sis_csr_free(sis_csr *c) { delete c; }

int sis_maxcut_xi0(const sis_csr *c, double *xi0) {
  if (!c || !xi0) return fail(SIS_ERR_ARG, "sis_maxcut_xi0: null argument");
  const int64_t N = c->n, Z = (int64_t)c->val.size();
  if (N < 2) return fail(SIS_ERR_ARG, "sis_maxcut_xi0: need at least 2 nodes");
  double sum = 0.0, nzc = 0.0;
  for (double v : c->val) sum += v;
  const double mean = sum / (double)(N * N);
  for (double v : c->val) nzc += std::pow(v - mean, 2);
  const double zc = (double)(N * N - Z) * mean * mean;
  const double sigma = std::pow((nzc + zc) / (double)(N * N - 1), 0.5);
  *xi0 = 0.7 / (sigma * std::pow((double)N, 0.5));
  if (!std::isfinite(*xi0)) return fail(SIS_ERR_ARG, "sis_maxcut_xi0: J has zero variance");
  return SIS_OK;
}

int sis_maxcut_random_cut(const sis_csr *c, uint64_t seed, double *cut) {
  if (!c || !cut) return fail(SIS_ERR_ARG, "sis_maxcut_random_cut: null argument");
  Rng rng(seed);
  std::vector<int8_t> s((size_t)c->n);
  for (auto &v : s) v = (rng.below(10) < 5) ? 1 : -1;
  double sum = 0.0;
  for (int64_t i = 0; i < c->n; i++)
    for (int64_t k = c->rp[i]; k < c->rp[i + 1]; k++)
      if (s[(size_t)i] != s[(size_t)c->col[k]]) sum += c->val[k];
  *cut = sum * 0.5;
  return SIS_OK;
}

int sis_cut_stats(const double *cut, int64_t agents, double out[5]) {
  if (!cut || !out || agents < 1) return fail(SIS_ERR_ARG, "sis_cut_stats: bad argument");
  double mx = -1.7976931348623157e308, mn = 1.7976931348623157e308, tot = 0.0;
  for (int64_t a = 0; a < agents; a++) {
    double v = cut[a];
    if (v > mx) mx = v;
    if (v < mn) mn = v;
    tot += v;
  }
  const double avg = tot / (double)agents;
  double sd = 0.0;
  int64_t nmax = 0;
  for (int64_t a = 0; a < agents; a++) {
    sd += std::pow(cut[a] - avg, 2);
    if (cut[a] == mx) nmax++;
  }
  sd = std::pow(sd / (double)(agents - 1), 0.5);
  out[0] = mn;
  out[1] = mx;
  out[2] = avg;
  out[3] = sd;
  out[4] = (double)nmax / (double)agents;
  return SIS_OK;
}

// --------------------------------------------------------------------- TSP

static int tsp_finish(sis_tsp *t) {
  const int64_t n = t->n;
  t->dist.assign((size_t)(n * (n - 1) / 2), 0.0);
  for (int64_t c1 = 0; c1 < n; c1++)
    for (int64_t c2 = c1 + 1; c2 < n; c2++) {
      double v;
      if (t->kind == SIS_TSP_EUC_2D) {
        const double xd = t->x[c1] - t->x[c2], yd = t->y[c1] - t->y[c2];
        const double r = std::pow(xd * xd + yd * yd, 0.5);
        v = t->euc_round ? (double)(int)(r + 0.5) : (double)(int)r;
      } else {
        const double R = 6378.388;
        const double q1 = std::cos(t->y[c1] - t->y[c2]);
        const double q2 = std::cos(t->x[c1] - t->x[c2]);
        const double q3 = std::cos(t->x[c1] + t->x[c2]);
        v = (double)(int)(R * std::acos(0.5 * ((1.0 + q1) * q2 - (1.0 - q1) * q3)) + 1.0);
      }
      t->dist[(size_t)tri(c1, c2, n)] = v;
    }
  // Totals and maximum in the reference's packed order (src/tsp_gpu.cu:55-68).
  t->totals.assign((size_t)n, 0.0);
  t->maxd = 0.0;
  t->sumd = 0.0;
  for (int64_t c1 = 0; c1 < n; c1++)
    for (int64_t c2 = c1 + 1; c2 < n; c2++) {
      const double d = t->dist[(size_t)tri(c1, c2, n)];
      t->totals[(size_t)c1] += d;
      t->totals[(size_t)c2] += d;
      t->sumd += d;
      if (d > t->maxd) t->maxd = d;
    }
  return SIS_OK;
}

static void geo_to_radians(double &v) {
  const double PI = 3.14159265358979323846;
  const int deg = (int)v;
  const double mn = v - deg;
  v = PI * (deg + 5.0 * mn / 3.0) / 180.0;
}

int sis_tsp_read(const char *path, int euc_round, sis_tsp **out) {
  if (!path || !out) return fail(SIS_ERR_ARG, "sis_tsp_read: null argument");
  *out = nullptr;
  TextStream s;
  if (!s.load(path)) return fail(SIS_ERR_IO, std::string("cannot open ") + path);
  int64_t n = -1;
  std::string type;
  bool coords = false;
  std::string line;
  // Header: KEY [:] value lines until NODE_COORD_SECTION.
  while (s.p < s.buf.size() && !coords) {
    size_t e = s.buf.find('\n', s.p);
    if (e == std::string::npos) e = s.buf.size();
    line.assign(s.buf, s.p, e - s.p);
    s.p = e < s.buf.size() ? e + 1 : e;
    size_t b = line.find_first_not_of(" \t\r");
    if (b == std::string::npos) continue;
    std::string key, value;
    size_t colon = line.find(':', b);
    if (colon != std::string::npos) {
      key = line.substr(b, colon - b);
      value = line.substr(colon + 1);
    } else {
      size_t sp = line.find_first_of(" \t\r", b);
      key = line.substr(b, sp == std::string::npos ? std::string::npos : sp - b);
      value = sp == std::string::npos ? "" : line.substr(sp);
    }
    auto trim = [](std::string &q) {
      size_t x = q.find_first_not_of(" \t\r"), y = q.find_last_not_of(" \t\r");
      q = x == std::string::npos ? "" : q.substr(x, y - x + 1);
    };
    trim(key);
    trim(value);
    if (key == "DIMENSION") n = std::atoll(value.c_str());
    else if (key == "EDGE_WEIGHT_TYPE") type = value;
    else if (key == "NODE_COORD_SECTION") coords = true;
  }
  if (n < 1) return fail(SIS_ERR_FORMAT, std::string(path) + ": missing or bad DIMENSION");
  if (type != "EUC_2D" && type != "GEO")
    return fail(SIS_ERR_FORMAT, std::string(path) + ": EDGE_WEIGHT_TYPE '" + type +
                                    "' unsupported (EUC_2D or GEO)");
  if (!coords) return fail(SIS_ERR_FORMAT, std::string(path) + ": no NODE_COORD_SECTION");
  std::unique_ptr<sis_tsp> t(new (std::nothrow) sis_tsp());
  if (!t) return fail(SIS_ERR_NOMEM, "out of host memory");
  t->n = n;
  t->kind = type == "GEO" ? SIS_TSP_GEO : SIS_TSP_EUC_2D;
  t->euc_round = euc_round ? 1 : 0;
  for (int64_t i = 0; i < n; i++) {
    double id, a, b;
    if (!s.dbl_tok(id) || !s.dbl_tok(a) || !s.dbl_tok(b))
      return fail(SIS_ERR_FORMAT, std::string(path) + ": node " + std::to_string(i + 1) +
                                      " missing or malformed");
    t->rawx.push_back(a);
    t->rawy.push_back(b);
    if (t->kind == SIS_TSP_GEO) {
      geo_to_radians(a);
      geo_to_radians(b);
    }
    t->x.push_back(a);
    t->y.push_back(b);
  }
  tsp_finish(t.get());
  *out = t.release();
  return SIS_OK;
}

int sis_tsp_from_coords(int64_t n, const double *x, const double *y, int kind, int euc_round,
                        sis_tsp **out) {
  if (!out || n < 1 || !x || !y || (kind != SIS_TSP_EUC_2D && kind != SIS_TSP_GEO))
    return fail(SIS_ERR_ARG, "sis_tsp_from_coords: bad argument");
  std::unique_ptr<sis_tsp> t(new (std::nothrow) sis_tsp());
  if (!t) return fail(SIS_ERR_NOMEM, "out of host memory");
  t->n = n;
  t->kind = kind;
  t->euc_round = euc_round ? 1 : 0;
  t->x.assign(x, x + n);
  t->y.assign(y, y + n);
  t->rawx = t->x;
  t->rawy = t->y;
  tsp_finish(t.get());
  *out = t.release();
  return SIS_OK;
}

int sis_tsp_info(const sis_tsp *t, int64_t *ncities, double *max_distance, double *sum_distance) {
  if (!t) return fail(SIS_ERR_ARG, "sis_tsp_info: null tsp");
  if (ncities) *ncities = t->n;
  if (max_distance) *max_distance = t->maxd;
  if (sum_distance) *sum_distance = t->sumd;
  return SIS_OK;
}

int sis_tsp_kind_of(const sis_tsp *t, int *kind) {
  if (!t || !kind) return fail(SIS_ERR_ARG, "sis_tsp_kind_of: null argument");
  *kind = t->kind;
  return SIS_OK;
}

int sis_tsp_export(const sis_tsp *t, double *x, double *y, double *dist, double *totals) {
  if (!t) return fail(SIS_ERR_ARG, "sis_tsp_export: null tsp");
  if (x) std::memcpy(x, t->x.data(), sizeof(double) * t->x.size());
  if (y) std::memcpy(y, t->y.data(), sizeof(double) * t->y.size());
  if (dist && !t->dist.empty()) std::memcpy(dist, t->dist.data(), sizeof(double) * t->dist.size());
  if (totals) std::memcpy(totals, t->totals.data(), sizeof(double) * t->totals.size());
  return SIS_OK;
}

int sis_tsp_qubo(const sis_tsp *t, sis_csr **J, double *field) {
  if (!t || !J) return fail(SIS_ERR_ARG, "sis_tsp_qubo: null argument");
  *J = nullptr;
  const int64_t n = t->n, N = n * n;
  if (N > INT32_MAX) return fail(SIS_ERR_ARG, "sis_tsp_qubo: too many cities");
  const double A = 1, B = t->maxd, C = t->maxd;
  // Value of entry ((i,k),(j,l)) under the dense constructor's rule order
  // (src/tsp_gpu.cu:85-106).
  auto value = [&](int64_t i, int64_t k, int64_t j, int64_t l) -> double {
    const bool adj = (i == j + 1 || j == i + 1 || (i == 0 && j == n - 1) || (i == n - 1 && j == 0));
    if (i != j && k == l) return C / 4;
    if (adj && k != l) {
      const int64_t c1 = std::min(k, l), c2 = std::max(k, l);
      return A * t->dist[(size_t)tri(c1, c2, n)] / 8;
    }
    if (i == j && k != l) return B / 4;
    if (i == j && k == l) return (B + C) / 4;
    return 0.0;
  };
  std::unique_ptr<sis_csr> m(new (std::nothrow) sis_csr());
  if (!m) return fail(SIS_ERR_NOMEM, "out of host memory");
  m->n = N;
  m->rp.assign((size_t)N + 1, 0);
  m->col.reserve((size_t)(N * (4 * n - 3)));
  m->val.reserve(m->col.capacity());
  std::vector<int64_t> cand;
  for (int64_t i = 0; i < n; i++) {
    const int64_t ip = (i + 1) % n, im = (i + n - 1) % n;
    for (int64_t k = 0; k < n; k++) {
      cand.clear();
      // Candidate positions of the only possibly-nonzero entries of row
      // (i,k): same city (any position), same position (any city), and any
      // city at the two cyclically adjacent positions.
      for (int64_t j = 0; j < n; j++) cand.push_back(j * n + k);
      for (int64_t l = 0; l < n; l++) {
        cand.push_back(i * n + l);
        cand.push_back(ip * n + l);
        cand.push_back(im * n + l);
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (int64_t cidx : cand) {
        const double v = value(i, k, cidx / n, cidx % n);
        if (v != 0) {
          m->col.push_back((int32_t)cidx);
          m->val.push_back(v);
        }
      }
      m->rp[(size_t)(i * n + k) + 1] = (int64_t)m->col.size();
    }
  }
  m->sum = 0.0;
  if (field)
    for (int64_t i = 0; i < n; i++)
      for (int64_t k = 0; k < n; k++)
        field[i * n + k] = -1 * (A * t->totals[(size_t)k] + (double)(n - 2) * (B + C)) / 2;
  *J = m.release();
  return SIS_OK;
}

int sis_tsp_xi0(const sis_tsp *t, double *xi0) {
  if (!t || !xi0) return fail(SIS_ERR_ARG, "sis_tsp_xi0: null argument");
  const int64_t n = t->n;
  if (n < 2) return fail(SIS_ERR_ARG, "sis_tsp_xi0: needs at least 2 cities");
  // Sum of squares of the QUBO's stored entries (src/ising_tsp.cu:74-77) by
  // rule (src/tsp_gpu.cu:85-106): i != j, k = l: C/4 (n^2 (n-1) entries);
  // i = j, k != l: B/4 (n^2 (n-1)); the diagonal (B+C)/4 (n^2); adjacent
  // positions (a(n) of them: 2 for n >= 3, 1 for n = 2), k != l:
  // A d_kl / 8, i.e. n a(n) times both orders of every city pair.  Zero
  // entries are not stored and add nothing.  Every square is a multiple of
  // 1/64, so the sum is exact in any order while below 2^47.
  const double A = 1, B = t->maxd, C = t->maxd, nn = (double)n;
  double pairs = 0.0;
  for (int64_t c1 = 0; c1 < n; c1++)
    for (int64_t c2 = c1 + 1; c2 < n; c2++) {
      const double v = A * t->dist[(size_t)tri(c1, c2, n)] / 8;
      pairs += v * v;
    }
  const double adj = n >= 3 ? 2.0 : 1.0;
  const double sq = nn * nn * (nn - 1) * ((C / 4) * (C / 4) + (B / 4) * (B / 4)) +
                    nn * nn * ((B + C) / 4) * ((B + C) / 4) + 2.0 * nn * adj * pairs;
  // sigma = sqrt(sum / (N^2 - 1)) with N = n^2 rows (64-bit N^2; the
  // reference's int product overflows for n >= 216), xi0 = 0.5 / (sigma sqrt N)
  const double N = nn * nn;
  const double sigma = std::pow(sq / (double)(n * n * n * n - 1), 0.5);
  *xi0 = 0.5 / (sigma * std::pow(N, 0.5));
  return SIS_OK;
}

int sis_tsp_tour_length(const sis_tsp *t, const int32_t *tour, double *length) {
  if (!t || !tour || !length) return fail(SIS_ERR_ARG, "sis_tsp_tour_length: null argument");
  const int64_t n = t->n;
  double sum = 0.0;
  for (int64_t i = 0; i < n; i++) {
    const int64_t u = tour[i], v = tour[(i + 1) % n];
    if (u < 0 || u >= n || v < 0 || v >= n || (n > 1 && u == v))
      return fail(SIS_ERR_ARG, "sis_tsp_tour_length: not a tour");
    if (n > 1) sum += t->dist[(size_t)tri(std::min(u, v), std::max(u, v), n)];
  }
  *length = sum;
  return SIS_OK;
}

int sis_tsp_random_length(const sis_tsp *t, uint64_t seed, double *length) {
  if (!t || !length) return fail(SIS_ERR_ARG, "sis_tsp_random_length: null argument");
  std::vector<int32_t> tour((size_t)t->n);
  for (int64_t i = 0; i < t->n; i++) tour[(size_t)i] = (int32_t)i;
  Rng rng(seed);
  for (int64_t i = t->n - 1; i > 0; i--) std::swap(tour[(size_t)i], tour[rng.below((uint64_t)i + 1)]);
  return sis_tsp_tour_length(t, tour.data(), length);
}

int sis_tsp_write(const sis_tsp *t, const char *path) {
  if (!t || !path) return fail(SIS_ERR_ARG, "sis_tsp_write: null argument");
  FILE *f = std::fopen(path, "w");
  if (!f) return fail(SIS_ERR_IO, std::string("cannot write ") + path);
  std::fprintf(f, "NAME: synthetic%lld\nTYPE: TSP\nDIMENSION: %lld\nEDGE_WEIGHT_TYPE: %s\n"
                  "NODE_COORD_SECTION\n",
               (long long)t->n, (long long)t->n, t->kind == SIS_TSP_GEO ? "GEO" : "EUC_2D");
  for (int64_t i = 0; i < t->n; i++)
    std::fprintf(f, "%lld %.17g %.17g\n", (long long)(i + 1), t->rawx[(size_t)i],
                 t->rawy[(size_t)i]);
  std::fprintf(f, "EOF\n");
  bool ok = std::fclose(f) == 0;
  return ok ? SIS_OK : fail(SIS_ERR_IO, std::string("short write to ") + path);
}

void sis_tsp_free(sis_tsp *t) { delete t; }

// -------------------------------------------------------------- generators

static double draw_weight(Rng &rng, int weights) {
  if (weights == SIS_W_PM1) return (rng.next() >> 63) ? -1.0 : 1.0;
  if (weights == SIS_W_INT3) {
    static const double tab[6] = {-3, -2, -1, 1, 2, 3};
    return tab[rng.below(6)];
  }
  return 1.0;
}

static int edges_to_csr(int64_t n, const std::vector<int32_t> &u, const std::vector<int32_t> &v,
                        const std::vector<double> &w, sis_csr **out) {
  return sis_csr_from_coo(n, (int64_t)w.size(), u.data(), v.data(), w.data(), 1, out);
}

int sis_gen_er(int64_t n, double p, int weights, uint64_t seed, sis_csr **out) {
  if (!out || n < 1 || n > INT32_MAX || !(p >= 0.0 && p <= 1.0))
    return fail(SIS_ERR_ARG, "sis_gen_er: bad argument");
  Rng rng(seed);
  std::vector<int32_t> u, v;
  std::vector<double> w;
  if (p > 0.0) {
    if (p >= 1.0) {
      for (int64_t a = 1; a < n; a++)
        for (int64_t b = 0; b < a; b++) {
          u.push_back((int32_t)a);
          v.push_back((int32_t)b);
          w.push_back(draw_weight(rng, weights));
        }
    } else {
      // Geometric skipping over the pairs (a, b), b < a (Batagelj-Brandes).
      const double lq = std::log(1.0 - p);
      int64_t a = 1, b = -1;
      while (a < n) {
        const double r = rng.uniform();
        b += 1 + (int64_t)std::floor(std::log(1.0 - r) / lq);
        while (b >= a && a < n) {
          b -= a;
          a++;
        }
        if (a < n) {
          u.push_back((int32_t)a);
          v.push_back((int32_t)b);
          w.push_back(draw_weight(rng, weights));
        }
      }
    }
  }
  return edges_to_csr(n, u, v, w, out);
}

int sis_gen_chung_lu(int64_t n, double avg_degree, double exponent, int weights, uint64_t seed,
                     sis_csr **out) {
  if (!out || n < 2 || n > INT32_MAX || !(avg_degree > 0.0) || !(exponent > 1.0) ||
      avg_degree * (double)n > 4e9)
    return fail(SIS_ERR_ARG, "sis_gen_chung_lu: bad argument");
  Rng rng(seed);
  // Expected-degree weights (r + 1)^(-1 / (exponent - 1)) by rank r, given to
  // the nodes in a random order; endpoints are drawn from their cumulative sum.
  std::vector<int32_t> perm((size_t)n);
  for (int64_t i = 0; i < n; i++) perm[(size_t)i] = (int32_t)i;
  for (int64_t i = n - 1; i > 0; i--) std::swap(perm[(size_t)i], perm[(size_t)rng.below((uint64_t)i + 1)]);
  std::vector<double> cdf((size_t)n);
  double acc = 0.0;
  const double e = -1.0 / (exponent - 1.0);
  for (int64_t r = 0; r < n; r++) cdf[(size_t)r] = (acc += std::pow((double)(r + 1), e));
  auto draw = [&]() -> int32_t {
    const double x = rng.uniform() * acc;
    const int64_t r = std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin();
    return perm[(size_t)std::min<int64_t>(r, n - 1)];
  };
  const int64_t m = (int64_t)std::llround(avg_degree * (double)n / 2.0);
  std::vector<uint64_t> pairs;
  pairs.reserve((size_t)m);
  for (int64_t k = 0; k < m; k++) {
    const uint32_t a = (uint32_t)draw(), b = (uint32_t)draw();
    if (a == b) continue;  // no self-loops
    pairs.push_back(a > b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a);
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());  // simple graph
  std::vector<int32_t> u(pairs.size()), v(pairs.size());
  std::vector<double> w(pairs.size());
  for (size_t k = 0; k < pairs.size(); k++) {
    u[k] = (int32_t)(pairs[k] >> 32);
    v[k] = (int32_t)(pairs[k] & 0xffffffffu);
    w[k] = draw_weight(rng, weights);
  }
  return edges_to_csr(n, u, v, w, out);
}

int sis_gen_torus(int64_t rows, int64_t cols, int weights, uint64_t seed, sis_csr **out) {
  if (!out || rows < 3 || cols < 3 || rows * cols > INT32_MAX)
    return fail(SIS_ERR_ARG, "sis_gen_torus: need rows, cols >= 3");
  Rng rng(seed);
  std::vector<int32_t> u, v;
  std::vector<double> w;
  for (int64_t r = 0; r < rows; r++)
    for (int64_t c = 0; c < cols; c++) {
      const int32_t id = (int32_t)(r * cols + c);
      u.push_back(id);
      v.push_back((int32_t)(r * cols + (c + 1) % cols));
      w.push_back(draw_weight(rng, weights));
      u.push_back(id);
      v.push_back((int32_t)(((r + 1) % rows) * cols + c));
      w.push_back(draw_weight(rng, weights));
    }
  return edges_to_csr(rows * cols, u, v, w, out);
}

int sis_gen_regular(int64_t n, int d, int weights, uint64_t seed, sis_csr **out) {
  if (!out || n < 2 || d < 1 || d >= n || (n * d) % 2 || n * d > INT32_MAX)
    return fail(SIS_ERR_ARG, "sis_gen_regular: need 1 <= d < n and n*d even");
  Rng rng(seed);
  const int64_t m = n * d / 2;
  std::vector<int32_t> stubs((size_t)(n * d));
  for (int64_t i = 0; i < n * d; i++) stubs[(size_t)i] = (int32_t)(i / d);
  for (int64_t i = n * d - 1; i > 0; i--) std::swap(stubs[(size_t)i], stubs[rng.below((uint64_t)i + 1)]);
  std::vector<int32_t> eu((size_t)m), ev((size_t)m);
  auto key = [](int32_t a, int32_t b) -> uint64_t {
    if (a > b) std::swap(a, b);
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
  };
  std::unordered_map<uint64_t, int32_t> cnt;
  cnt.reserve((size_t)m * 2);
  for (int64_t e = 0; e < m; e++) {
    eu[(size_t)e] = stubs[(size_t)(2 * e)];
    ev[(size_t)e] = stubs[(size_t)(2 * e + 1)];
    cnt[key(eu[(size_t)e], ev[(size_t)e])]++;
  }
  auto bad = [&](int64_t e) {
    return eu[(size_t)e] == ev[(size_t)e] || cnt[key(eu[(size_t)e], ev[(size_t)e])] > 1;
  };
  std::vector<int64_t> todo;
  for (int64_t e = 0; e < m; e++)
    if (bad(e)) todo.push_back(e);
  int64_t guard = 0;
  while (!todo.empty()) {
    if (++guard > 1000 * (m + 10)) return fail(SIS_ERR_ARG, "sis_gen_regular: repair did not converge");
    const int64_t e = todo.back();
    if (!bad(e)) { todo.pop_back(); continue; }
    const int64_t f = (int64_t)rng.below((uint64_t)m);
    if (f == e) continue;
    int32_t a = eu[(size_t)e], b = ev[(size_t)e], c = eu[(size_t)f], dd = ev[(size_t)f];
    if (rng.next() >> 63) std::swap(c, dd);
    // Proposed rewiring (a,c), (b,dd).
    if (a == c || b == dd) continue;
    if (cnt.count(key(a, c)) && cnt[key(a, c)] > 0) continue;
    if (cnt.count(key(b, dd)) && cnt[key(b, dd)] > 0) continue;
    if (key(a, c) == key(b, dd)) continue;
    if (--cnt[key(a, b)] == 0) cnt.erase(key(a, b));
    if (--cnt[key(c, dd)] == 0) cnt.erase(key(c, dd));
    eu[(size_t)e] = a; ev[(size_t)e] = c;
    eu[(size_t)f] = b; ev[(size_t)f] = dd;
    cnt[key(a, c)]++;
    cnt[key(b, dd)]++;
    if (bad(f)) todo.push_back(f);
  }
  std::vector<double> w((size_t)m);
  for (int64_t e = 0; e < m; e++) w[(size_t)e] = draw_weight(rng, weights);
  return edges_to_csr(n, eu, ev, w, out);
}

int sis_gen_tsp(int64_t n, int64_t span, int euc_round, uint64_t seed, sis_tsp **out) {
  if (!out || n < 1 || span < 0) return fail(SIS_ERR_ARG, "sis_gen_tsp: bad argument");
  Rng rng(seed);
  std::vector<double> x((size_t)n), y((size_t)n);
  for (int64_t i = 0; i < n; i++) {
    x[(size_t)i] = (double)rng.below((uint64_t)span + 1);
    y[(size_t)i] = (double)rng.below((uint64_t)span + 1);
  }
  return sis_tsp_from_coords(n, x.data(), y.data(), SIS_TSP_EUC_2D, euc_round, out);
}

}  // extern "C"
