// sis_solver.hip -- device half of the C ABI: solver handle, state upload,
// the window-period loop (optionally replayed as a hipGraph), evaluation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "sis_hip_util.hpp"
#include "sis_internal.hpp"
#include "sis_kernels.hpp"

using namespace sis;

// Code width of a narrow-state form (12, 10 or 8 bits), 0 for the others.
static int qn_width_of(int form) {
  return form == SIS_FORM_Q12 ? 12
         : form == SIS_FORM_Q10 ? 10
         : (form == SIS_FORM_Q8 || form == SIS_FORM_Q8S) ? 8
                                                        : 0;
}

// Dwords of the narrow state per 256-agent quad and lane: 3 (12-bit), 2.5
// (10-bit: two dword planes and one of 16-bit values) or 2 (8-bit), times 2.
static int64_t qn_half_dwords(int qw) { return qw == 12 ? 6 : qw == 10 ? 5 : 4; }

struct sis_solver {
  sis_params p{};
  int dev = 0;
  hipStream_t stream = nullptr;
  int64_t N = 0, Z = 0, A = 0, lda = 0, W = 0;
  double S = 0.0;
  int vec = 1;     // agents per lane (plain / bf16) or words per wavefront (packed)
  int gather = G_DSB;
  int prob = PROB_MAXCUT;  // PROB_MAXCUT, PROB_TSP or PROB_UNFUSED
  bool f64 = false;

  DevBuf<int64_t> rp;
  DevBuf<int32_t> col;
  DevBuf<float> valf;
  DevBuf<double> vald;
  DevBuf<float> hf;
  DevBuf<double> hd;
  bool has_field = false;

  DevBuf<float> Xf[2], Yf;       // FP32 state (plain uses both X buffers)
  DevBuf<double> Xd[2], Yd;      // FP64 state
  DevBuf<uint16_t> sh[2];        // BF16 shadow
  DevBuf<uint64_t> sg[2];        // packed sign words
  DevBuf<uint32_t> Xq;           // narrow forms: X / Y codes, 3, 2.5 or 2 dword planes (12, 10, 8 bits)
  // Row storage order: node_of[r] is the node stored in row r (empty: node
  // order); a copy on the device keys the seeded initial state by node.
  std::vector<int32_t> node_of;
  DevBuf<int32_t> node_of_d;
  int64_t row_of(int64_t node) const { return node_of.empty() ? node : row_inv[(size_t)node]; }
  std::vector<int32_t> row_inv;  // row_inv[node] = its row
  int qshift = 0;                // narrow form: Y codes are round(y 2^qshift)
  int par = 0;                   // index of the current snapshot buffer

  DevBuf<uint64_t> cur, fin, newmask, part;
  DevBuf<uint64_t> tsp_lines;    // TSP window: one-hot mask per position / city line
  DevBuf<int32_t> stab, bif, valid;
  DevBuf<Ctrl> ctrl;
  Ctrl *h_ctrl = nullptr;        // pinned mirror
  int64_t scan_blocks = 0, scan_rows = 0;

  DevBuf<float> schedf;
  DevBuf<double> schedd;
  int64_t sched_T = -1, sched_len = 0;

  DevBuf<double> cutpart, outv;
  DevBuf<int64_t> best;
  int64_t cut_blocks = 0, cut_rows = 0;

  DevBuf<double> dist;
  int64_t dist_n = -1;
  std::vector<double> dist_host;

  bool pm1 = false;               // every stored weight is exactly +1 or -1
  // ELL copies of the matrix for the resident tiles, per log2 G
  bool ell_ok[4] = {false, false, false, false};
  int32_t ell_words[4] = {0, 0, 0, 0}, ell_groups[4] = {0, 0, 0, 0};
  int32_t ell_maxw[4] = {0, 0, 0, 0};  // widest group, in chunks of four entries
  DevBuf<uint16_t> ell[4];
  DevBuf<uint32_t> gofs[4];
  DevBuf<uint16_t> ell_perm[4];  // rows in degree order, when the ELL uses it

  // Row split (k_step_split): rows with more than `split` stored entries.
  std::vector<int64_t> rp_host;  // row pointers, for building the split tables
  int64_t max_deg = 0;
  int64_t split = 0;             // active threshold, 0: none
  int64_t split_nent = 0;        // SplitEnt records (0: no row exceeds the threshold)
  DevBuf<SplitEnt> split_ent;
  DevBuf<char> split_part;
  DevBuf<int32_t> split_cnt;

  // Structured TSP product (sis_solver_create_tsp with structured = 1).
  bool structured = false;
  int64_t tsp_n = 0;
  double qB = 0.0, qC = 0.0;           // B/4, C/4
  DevBuf<double> dqd;                  // A d_kl / 8, n x n
  DevBuf<float> dqf;
  DevBuf<float> dqm;                   // FP32, zero-padded for the matrix cores
  int64_t tsp_nep = 0;                 // n rounded up to 8
  int tsp_res_G = 0;                   // agents per workgroup of tile 3 (0: not available)
  DevBuf<char> tsum_r, tsum_c;         // per-step row / column sums, form precision
  DevBuf<double> esum_r, esum_c;       // the same for the energy, FP64

  TspArgs tsp_args(bool energy) {
    TspArgs ts{};
    ts.dq = energy || f64 ? (const void *)dqd.p : (const void *)dqf.p;
    ts.rsum = energy ? (void *)esum_r.p : (void *)tsum_r.p;
    ts.csum = energy ? (void *)esum_c.p : (void *)tsum_c.p;
    ts.dqm = dqm.p;
    ts.qB = qB;
    ts.qC = qC;
    ts.n = tsp_n;
    ts.nep = tsp_nep;
    return ts;
  }

  std::vector<hipEvent_t> events;
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  int64_t graph_P = -1;
  int graph_uw = -1;
  // whole-run graph (use_graph == 2): every window period of one run
  hipGraphExec_t run_graph = nullptr;
  int64_t run_P = -1, run_nwin = -1;
  int run_uw = -1, run_par = -1;
  bool state_ready = false;

  void drop_graphs() {
    for (auto &g : graph_exec)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
    if (run_graph) {
      (void)hipGraphExecDestroy(run_graph);
      run_graph = nullptr;
    }
  }

  ~sis_solver() {
    drop_graphs();
    for (auto e : events) (void)hipEventDestroy(e);
    if (h_ctrl) (void)hipHostFree(h_ctrl);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // narrow-state forms: X, Y as qw()-bit codes
  bool narrow() const { return qn_width_of(p.form) != 0; }
  int qw() const { return qn_width_of(p.form); }
  void *Xbuf(int i) { return f64 ? (void *)Xd[i].p : (void *)Xf[i].p; }
  void *Ybuf() { return f64 ? (void *)Yd.p : (void *)Yf.p; }
  // X holding the current positions.
  void *Xcur() { return p.form == SIS_FORM_F32 || p.form == SIS_FORM_F64 ? Xbuf(par) : Xbuf(0); }
};

// Narrow-state forms on the host (code width qw = 12, 10 or 8): the same
// rounding as qn_encode (device) -- X code round(x 2^(qw-2)), +-(2^(qw-2) + 1)
// beyond the wall with a 0 Y code, Y code round(y 2^qshift) clamped to
// +-(2^(qw-1) - 1), nearest even, each stored plus 2^(qw-1) in qw bits -- and
// the plane layout (value p of lane l, quad q, row i at
// p N Q 64 + (i Q + q) 64 + l in dwords; 3 dword planes for qw = 12, 2 for
// qw = 8; for qw = 10 two dword planes and a third of 16-bit values, value o
// at 16-bit offset 2 (2 N Q 64) + o).
static int qn_shift_for(int qw, double dt) {
  if (!(dt > 0)) return -1;
  const double ymax = 2.0 / dt + dt, ym = (double)((1 << (qw - 1)) - 1);
  int sh = qw - 1;
  while (sh > 0 && std::ldexp(ym, -sh) < ymax) sh--;
  return sh;
}

static uint32_t qn_unit_host(int qw, float x, float y, int qshift) {
  const int xs = qw - 2, bias = 1 << (qw - 1), ym = bias - 1;
  int xq, yq;
  if (std::fabs(x) > 1.0f) {
    xq = std::signbit(x) ? -((1 << xs) + 1) : (1 << xs) + 1;
    yq = 0;
  } else {
    xq = (int)std::rint(std::ldexp(x, xs));
    const float yr = std::rint(std::ldexp(y, qshift));
    yq = yr > (float)ym ? ym : yr < -(float)ym ? -ym : (int)yr;
  }
  return (uint32_t)(xq + bias) | ((uint32_t)(yq + bias) << qw);
}

// Value o of plane p (a 16-bit value in the 10-bit form's third plane).
static uint32_t qn_plane_get(int qw, const std::vector<uint32_t> &P, int64_t ps, int64_t p,
                             int64_t o) {
  if (qw == 10 && p == 2) {
    uint16_t h;
    std::memcpy(&h, reinterpret_cast<const char *>(P.data()) + 4 * (size_t)(2 * ps) + 2 * (size_t)o,
                2);
    return h;
  }
  return P[(size_t)(p * ps + o)];
}

static void qn_plane_set(int qw, std::vector<uint32_t> &P, int64_t ps, int64_t p, int64_t o,
                         uint32_t v) {
  if (qw == 10 && p == 2) {
    const uint16_t h = (uint16_t)v;
    std::memcpy(reinterpret_cast<char *>(P.data()) + 4 * (size_t)(2 * ps) + 2 * (size_t)o, &h, 2);
    return;
  }
  P[(size_t)(p * ps + o)] = v;
}

// Unit j of a lane occupies bits ub j .. ub j + ub - 1 (ub = 2 qw) of its
// planes taken as one bit string (plane p at bit 32 p): read and written
// through the 64-bit window of planes p, p + 1 that holds it.
static void qn_put(int qw, std::vector<uint32_t> &P, int64_t N, int64_t Q, int64_t i, int64_t a,
                   uint32_t u) {
  const int64_t ub = 2 * qw, np = qw == 8 ? 2 : 3;
  const uint64_t m = ((uint64_t)1 << ub) - 1;
  const int64_t q = a / 256, j = (a % 256) / 64, l = a % 64, ps = N * Q * 64;
  const int64_t o = (i * Q + q) * 64 + l, p = (ub * j) / 32, r = (ub * j) % 32;
  uint64_t win = qn_plane_get(qw, P, ps, p, o);
  if (p < np - 1) win |= (uint64_t)qn_plane_get(qw, P, ps, p + 1, o) << 32;
  win = (win & ~(m << r)) | (((uint64_t)u & m) << r);
  qn_plane_set(qw, P, ps, p, o, (uint32_t)win);
  if (p < np - 1) qn_plane_set(qw, P, ps, p + 1, o, (uint32_t)(win >> 32));
}

static uint32_t qn_get(int qw, const std::vector<uint32_t> &P, int64_t N, int64_t Q, int64_t i,
                       int64_t a) {
  const int64_t ub = 2 * qw, np = qw == 8 ? 2 : 3;
  const int64_t q = a / 256, j = (a % 256) / 64, l = a % 64, ps = N * Q * 64;
  const int64_t o = (i * Q + q) * 64 + l, p = (ub * j) / 32, r = (ub * j) % 32;
  uint64_t win = qn_plane_get(qw, P, ps, p, o);
  if (p < np - 1) win |= (uint64_t)qn_plane_get(qw, P, ps, p + 1, o) << 32;
  return (uint32_t)(win >> r) & (uint32_t)(((uint64_t)1 << ub) - 1);
}

// Resident tiles (k_resident): bit 20, log2 of the agents per workgroup in
// bits 21-22.  Fused MaxCut in FP32 (plain dSB / bSB, packed dSB), weights
// +-1, N x G pairs per 1024-thread workgroup within the kernel's register
// slots (12,288), the ELL matrix and two buffers of gather slots within 160 KB
// of LDS.  Bit 23 (dSB only): the ELL in registers (k_resident_reg), for
// groups one chunk wide (N G up to 12,288) or two (N G up to 4,096), with the
// two gather-slot buffers within 80 KB so that two workgroups share a CU.
constexpr int kResidentBit = 1 << 20;
constexpr int kResRegBit = 1 << 23;

static bool resident_layout(const sis_solver *s, int vec, ResArgs *r) {
  if (!(vec & kResidentBit) || (vec & ~(kResidentBit | (3 << 21) | kResRegBit)) != 0)
    return false;
  const int lg = (vec >> 21) & 3, G = 1 << lg;
  const bool reg = (vec & kResRegBit) != 0;
  if (s->prob != PROB_MAXCUT || s->structured || !s->pm1) return false;
  if (!(s->p.form == SIS_FORM_F32 || (s->p.form == SIS_FORM_PACKED && s->gather == G_DSB)))
    return false;
  if (s->N * G > resident_pairs_max() || s->lda % G != 0 || !s->ell_ok[lg])
    return false;
  // register budget (64 VGPRs at two workgroups per CU): one chunk per row
  // up to twelve register slots, two chunks up to four
  if (reg && (s->gather != G_DSB || s->ell_maxw[lg] > 2 ||
              (s->ell_maxw[lg] == 2 && s->N * G > 4 * 1024)))
    return false;
  auto al16 = [](int64_t b) { return (b + 15) / 16 * 16; };
  const int64_t off = reg ? 0 : al16(2 * (int64_t)s->ell_words[lg]);
  // gather slots: +1 and -1 weight terms of every pair, then G zeros
  const int64_t slots = 2 * s->N * G + 2 * G;
  const int64_t snap = al16(s->gather == G_DSB ? slots : 4 * slots);
  if (off + 2 * snap > (reg ? 80 : 160) * 1024) return false;
  if (r) {
    r->G = G;
    r->lgG = lg;
    r->snap_off = (int32_t)off;
    r->snap_bytes = (int32_t)snap;
    r->ell = s->ell[lg].p;
    r->gofs = s->gofs[lg].p;
    r->perm = s->ell_perm[lg].p;  // null: rows in natural order
    r->ngroups = s->ell_groups[lg];
    r->ell_words = s->ell_words[lg];
    r->ell_cw = reg ? s->ell_maxw[lg] : 0;
  }
  return true;
}

// ELL groups of R = 64 / G rows for the resident tiles (only for matrices
// whose entries fit the kernel's LDS at all), padded to whole chunks of four
// entries: entry k of row g R + l at gofs[g] + 4 ((k / 4) R + l) + k % 4.  An
// entry is the gather slot of agent 0 -- dSB (byte slots, interleaved):
// 2 column G, plus 1 for a -1 weight; bSB (float slots, split so that +1
// weights alone still spread over every LDS bank): column G, plus N G for a
// -1 weight -- and 2 N G (a zero slot) past a row's end.
static int build_ell(sis_solver *s, const sis_csr *J) {
  if (!s->pm1 || s->N > resident_pairs_max() || s->prob != PROB_MAXCUT || 2 * s->Z > 160 * 1024) return SIS_OK;
  std::vector<int64_t> byd((size_t)s->N);  // rows by descending degree (stable)
  for (int64_t i = 0; i < s->N; i++) byd[(size_t)i] = i;
  auto deg = [&](int64_t r) { return J->rp[r + 1] - J->rp[r]; };
  std::stable_sort(byd.begin(), byd.end(), [&](int64_t x, int64_t y) { return deg(x) > deg(y); });
  for (int lg = 0; lg < 4; lg++) {
    const int64_t G = int64_t(1) << lg, R = 64 >> lg, ng = (s->N + R - 1) / R;
    if (2 * s->N * G + 2 * G > 65535) continue;  // slots must fit the 16-bit entries
    // group widths for rows in natural order and in degree order; the sorted
    // layout (kernel instances up to 8 register slots) when 10% smaller
    auto layout = [&](bool sorted, std::vector<uint32_t> &go) {
      go.assign((size_t)ng + 1, 0);
      for (int64_t g = 0; g < ng; g++) {
        int64_t wmax = 0;
        for (int64_t i = g * R; i < std::min(s->N, (g + 1) * R); i++)
          wmax = std::max(wmax, deg(sorted ? byd[(size_t)i] : i));
        go[(size_t)g + 1] = go[(size_t)g] + (uint32_t)((wmax + 3) / 4 * 4 * R);
      }
      return (int64_t)go[(size_t)ng];
    };
    std::vector<uint32_t> go, gos;
    int64_t words = layout(false, go);
    const int64_t words_s = layout(true, gos);
    const bool sorted = s->N * G <= resident_sorted_pairs_max() && 10 * words_s <= 9 * words;
    if (sorted) {
      go.swap(gos);
      words = words_s;
    }
    if (2 * words > 160 * 1024) continue;
    std::vector<uint16_t> e((size_t)std::max<int64_t>(4, words), (uint16_t)(2 * s->N * G));
    for (int64_t si = 0; si < s->N; si++) {
      const int64_t row = sorted ? byd[(size_t)si] : si, g = si / R, l = si % R;
      for (int64_t k = J->rp[row]; k < J->rp[row + 1]; k++) {
        const int64_t kk = k - J->rp[row];
        e[(size_t)(go[(size_t)g] + 4 * ((kk / 4) * R + l) + kk % 4)] =
            (uint16_t)(s->gather == G_DSB
                           ? 2 * J->col[(size_t)k] * G + (J->val[(size_t)k] < 0 ? 1 : 0)
                           : J->col[(size_t)k] * G + (J->val[(size_t)k] < 0 ? s->N * G : 0));
      }
    }
    int st;
    if ((st = s->ell[lg].alloc(e.size())) || (st = s->gofs[lg].alloc(go.size()))) return st;
    HIPCHK(hipMemcpy(s->ell[lg].p, e.data(), sizeof(uint16_t) * e.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->gofs[lg].p, go.data(), sizeof(uint32_t) * go.size(), hipMemcpyHostToDevice));
    if (sorted) {
      std::vector<uint16_t> pm((size_t)s->N);
      for (int64_t si = 0; si < s->N; si++) pm[(size_t)si] = (uint16_t)byd[(size_t)si];
      if ((st = s->ell_perm[lg].alloc(pm.size()))) return st;
      HIPCHK(hipMemcpy(s->ell_perm[lg].p, pm.data(), sizeof(uint16_t) * pm.size(), hipMemcpyHostToDevice));
    }
    s->ell_words[lg] = (int32_t)words;
    s->ell_groups[lg] = (int32_t)ng;
    uint32_t maxw = 0;
    for (int64_t g = 0; g < ng; g++)
      maxw = std::max<uint32_t>(maxw, (go[(size_t)g + 1] - go[(size_t)g]) / (uint32_t)(4 * R));
    s->ell_maxw[lg] = (int32_t)maxw;
    s->ell_ok[lg] = true;
  }
  return SIS_OK;
}

// Tile encoding: plain / BF16 = agents per lane | non-temporal << 16;
// packed = words | rows << 8 | non-temporal << 16 (see sis_kernels.hpp);
// resident tiles as above.
static bool tile_ok(const sis_solver *s, int vec) {
  if (vec & kResidentBit) return resident_layout(s, vec, nullptr);
  if (!step_supported(s->p.form, s->prob, s->gather, vec)) return false;
  if (s->p.form == SIS_FORM_PACKED) return s->W % (vec & 0xff) == 0;
  if (s->narrow()) return s->lda % 256 == 0;
  return s->lda % (64 * (vec & 0xffff)) == 0;
}

static int default_tile(const sis_solver *s) {
  int v = 1;
  if (s->narrow()) {
    // staged gathers of 16 rows per wavefront; non-temporal codes once they
    // exceed the Infinity Cache (3, 2.5 or 2 bytes per node-agent)
    v = 4 | 16 << 8 | 1 << 17;
    if ((double)s->N * (double)s->lda * (double)qn_half_dwords(s->qw()) / 2.0 >
        256.0 * 1024 * 1024)
      v |= 1 << 16;
    return v;
  }
  if (s->p.form == SIS_FORM_PACKED) {
    while (v < 4 && s->W % (v * 2) == 0) v *= 2;
    v |= 1 << 8;
  } else {
    const int vmax = s->f64 ? 2 : 4;
    while (v < vmax && s->lda % (64 * v * 2) == 0) v *= 2;
  }
  // Non-temporal own-row X / Y once the state (X and Y) is larger than the
  // 256 MB Infinity Cache: streamed once per step, it then does not displace
  // the gathered neighbour data (the autotuner's pick at N = 2^20, 256 agents).
  const double state = (double)s->N * (double)s->lda * (s->f64 ? 16.0 : 8.0);
  if (state > 256.0 * 1024 * 1024 && tile_ok(s, v | (1 << 16))) v |= 1 << 16;
  return v;
}

static std::vector<int> candidate_tiles(const sis_solver *s) {
  std::vector<int> c;
  // structured TSP: 1 = VALU distance product, 2 = matrix cores (FP32)
  if (s->structured) {
    if (s->f64) return {1};
    if (s->tsp_res_G) return {1, 2, 3};
    return {1, 2};
  }
  if (s->narrow()) {
    for (int nt : {0, 1}) {
      for (int r : {1, 2, 4}) c.push_back(4 | (r << 8) | (nt << 16));
      for (int r : {4, 8, 16, 32}) c.push_back(4 | (r << 8) | (nt << 16) | (1 << 17));
    }
    return c;
  }
  if (s->p.form == SIS_FORM_PACKED) {
    for (int wt : {1, 2, 4, 8})
      for (int r : {1, 2, 4})
        for (int nt : {0, 1}) c.push_back(wt | (r << 8) | (nt << 16));
  } else {
    for (int v : {1, 2, 4})
      for (int nt : {0, 1}) c.push_back(v | (nt << 16));
  }
  for (int lg = 0; lg < 4; lg++) {
    c.push_back(kResidentBit | (lg << 21));
    c.push_back(kResidentBit | (lg << 21) | kResRegBit);
  }
  std::vector<int> ok;
  for (int v : c)
    if (tile_ok(s, v)) ok.push_back(v);
  return ok;
}

// Row-split tables for threshold tau (0: none).  A row with deg > tau gets
// ceil(deg / (kWaves tau)) groups of kWaves chunks of tau entries; rows of more
// than one group get consecutive group-sum slots and one arrival counter per
// 64-agent tile.  Waits for queued work (the old tables may be in use).
static constexpr size_t kSplitPartMax = size_t(1) << 30;  // group-sum bytes

static int build_split(sis_solver *s, int64_t tau) {
  if (tau < 0) return fail(SIS_ERR_ARG, "row split threshold must be >= 0");
  if (s->stream) HIPCHK(hipStreamSynchronize(s->stream));
  s->drop_graphs();  // captured launches hold the old tables
  s->split = 0;
  s->split_nent = 0;
  if (tau == 0 || tau >= s->max_deg) return SIS_OK;
  if (!split_supported(s->p.form, s->prob, s->gather))
    return fail(SIS_ERR_ARG, "row split: unsupported form");
  const int64_t per = (int64_t)kWaves * tau;
  std::vector<SplitEnt> ent;
  int64_t heavy = 0, slots = 0;
  for (int64_t i = 0; i < s->N; i++) {
    const int64_t k0 = s->rp_host[i], k1 = s->rp_host[i + 1];
    if (k1 - k0 <= tau) continue;
    const int64_t ng = (k1 - k0 + per - 1) / per;
    for (int64_t g = 0; g < ng; g++) {
      SplitEnt e{};
      e.row = i;
      e.kb = k0 + g * per;
      e.ke = std::min(k1, e.kb + per);
      e.pbase = ng > 1 ? slots : 0;
      e.hr = (int32_t)heavy;
      e.g = (int32_t)g;
      e.ng = (int32_t)ng;
      ent.push_back(e);
    }
    if (ng > 1) slots += ng;
    heavy++;
  }
  const int64_t nent = (int64_t)ent.size();
  const size_t esz = s->f64 ? sizeof(double) : sizeof(float);
  if (nent * s->W >= INT32_MAX || heavy * s->W >= INT32_MAX)
    return fail(SIS_ERR_ARG, "row split: too many groups for one launch");
  if ((size_t)slots * (size_t)s->lda * esz > kSplitPartMax)
    return fail(SIS_ERR_ARG, "row split: group sums exceed 1 GiB; raise the threshold");
  int st;
  if ((st = s->split_ent.alloc((size_t)nent)) ||
      (st = s->split_part.alloc(std::max<size_t>(1, (size_t)slots * (size_t)s->lda * esz))) ||
      (st = s->split_cnt.alloc((size_t)std::max<int64_t>(1, heavy * s->W))))
    return st;
  HIPCHK(hipMemcpy(s->split_ent.p, ent.data(), sizeof(SplitEnt) * nent, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(s->split_cnt.p, 0, sizeof(int32_t) * s->split_cnt.n));
  s->split = tau;
  s->split_nent = nent;
  return SIS_OK;
}

// Locality order of the rows: multilevel greedy matching.  Level 1 pairs
// each node with an unpaired neighbour (nodes visited in a fixed
// pseudo-random order, the first unpaired neighbour in CSR order); level 2
// pairs the pairs over the graph of pairs, and so on for `levels` levels, so
// clusters of up to 2^levels nodes are connected subgraphs.  The order lists
// the top-level clusters in turn, each expanded recursively.  For the
// packed-sign gathers this puts a row's neighbours, where the graph allows,
// in the row's own 128-byte sign line (4 rows of 256 agents) or among the
// rows of its wavefront, so their gathers share cache lines.  O(Z) per level.
static void locality_order(const sis_csr &J, int levels, std::vector<int32_t> &node_of) {
  const int64_t n = J.n;
  std::vector<int64_t> rp(J.rp.begin(), J.rp.begin() + n + 1);
  std::vector<int32_t> col(J.col.begin(), J.col.end());
  std::vector<std::vector<int32_t>> up;  // up[l][v]: cluster of level-l node v
  std::vector<int64_t> sizes;            // nodes per level, finest first
  int64_t cn = n;
  uint64_t z = 0x243F6A8885A308D3ull;
  auto rnd = [&]() {  // splitmix64
    z += 0x9E3779B97F4A7C15ull;
    uint64_t r = z;
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
  };
  for (int l = 0; l < levels && cn > 1; l++) {
    std::vector<int32_t> visit((size_t)cn);
    for (int64_t v = 0; v < cn; v++) visit[(size_t)v] = (int32_t)v;
    for (int64_t v = cn - 1; v > 0; v--) std::swap(visit[(size_t)v], visit[(size_t)(rnd() % (uint64_t)(v + 1))]);
    std::vector<int32_t> cid((size_t)cn, -1);
    int32_t nc = 0;
    for (int32_t u : visit) {
      if (cid[(size_t)u] >= 0) continue;
      cid[(size_t)u] = nc;
      for (int64_t k = rp[(size_t)u]; k < rp[(size_t)u + 1]; k++) {
        const int32_t v = col[(size_t)k];
        if (v != u && cid[(size_t)v] < 0) {
          cid[(size_t)v] = nc;
          break;
        }
      }
      nc++;
    }
    // the graph of clusters: distinct neighbouring clusters, no self loops
    std::vector<int64_t> crp((size_t)nc + 1, 0);
    for (int64_t v = 0; v < cn; v++) crp[(size_t)cid[(size_t)v] + 1] += rp[(size_t)v + 1] - rp[(size_t)v];
    for (int64_t c = 0; c < nc; c++) crp[(size_t)c + 1] += crp[(size_t)c];
    std::vector<int32_t> ccol((size_t)crp[(size_t)nc]);
    std::vector<int64_t> fill(crp.begin(), crp.end() - 1);
    for (int64_t v = 0; v < cn; v++)
      for (int64_t k = rp[(size_t)v]; k < rp[(size_t)v + 1]; k++)
        ccol[(size_t)fill[(size_t)cid[(size_t)v]]++] = cid[(size_t)col[(size_t)k]];
    std::vector<int64_t> nrp((size_t)nc + 1, 0);
    int64_t w = 0;
    for (int64_t c = 0; c < nc; c++) {
      auto b = ccol.begin() + crp[(size_t)c], e = ccol.begin() + crp[(size_t)c + 1];
      std::sort(b, e);
      int32_t last = -1;
      for (auto it = b; it != e; ++it)
        if (*it != c && *it != last) ccol[(size_t)w++] = last = *it;
      nrp[(size_t)c + 1] = w;
    }
    ccol.resize((size_t)w);
    up.push_back(std::move(cid));
    sizes.push_back(cn);
    rp.swap(nrp);
    col.swap(ccol);
    cn = nc;
  }
  // expand top-down: the members of each cluster in increasing id order
  std::vector<int32_t> order((size_t)cn);
  for (int64_t c = 0; c < cn; c++) order[(size_t)c] = (int32_t)c;
  for (int64_t l = (int64_t)up.size() - 1; l >= 0; l--) {
    const std::vector<int32_t> &cid = up[(size_t)l];
    const int64_t fine = sizes[(size_t)l];
    std::vector<int64_t> start((size_t)cn + 1, 0);
    for (int64_t v = 0; v < fine; v++) start[(size_t)cid[(size_t)v] + 1]++;
    for (int64_t c = 0; c < cn; c++) start[(size_t)c + 1] += start[(size_t)c];
    std::vector<int32_t> members((size_t)fine);
    std::vector<int64_t> pos(start.begin(), start.end() - 1);
    for (int64_t v = 0; v < fine; v++) members[(size_t)pos[(size_t)cid[(size_t)v]]++] = (int32_t)v;
    std::vector<int32_t> next;
    next.reserve((size_t)fine);
    for (int32_t c : order)
      for (int64_t k = start[(size_t)c]; k < start[(size_t)c + 1]; k++) next.push_back(members[(size_t)k]);
    order.swap(next);
    cn = fine;
  }
  node_of.swap(order);
}

// The rows of J in the order node_of (row r holds node node_of[r]): each row
// keeps its entries in their CSR order, with columns renamed to rows.
static void reorder_csr(const sis_csr &J, const std::vector<int32_t> &node_of,
                        const std::vector<int32_t> &row_inv, sis_csr &out) {
  const int64_t n = J.n;
  out.n = n;
  out.sum = J.sum;
  out.rp.assign((size_t)n + 1, 0);
  out.col.resize(J.col.size());
  out.val.resize(J.val.size());
  for (int64_t r = 0; r < n; r++) {
    const int64_t v = node_of[(size_t)r], b = J.rp[(size_t)v], e = J.rp[(size_t)v + 1];
    out.rp[(size_t)r + 1] = out.rp[(size_t)r] + (e - b);
    for (int64_t k = b; k < e; k++) {
      const int64_t o = out.rp[(size_t)r] + (k - b);
      out.col[(size_t)o] = row_inv[(size_t)J.col[(size_t)k]];
      out.val[(size_t)o] = J.val[(size_t)k];
    }
  }
}

static int resolve_period(const sis_solver *s, int64_t T) {
  if (s->p.window_period > 0) return s->p.window_period;
  if (s->p.problem == SIS_TSP) return (int)std::max<int64_t>(1, T / 10);
  return 100;
}

extern "C" {

int sis_params_default(int problem, sis_params *p) {
  if (!p || (problem != SIS_MAXCUT && problem != SIS_TSP))
    return fail(SIS_ERR_ARG, "sis_params_default: bad argument");
  std::memset(p, 0, sizeof(*p));
  p->problem = problem;
  p->form = SIS_FORM_F32;
  p->dt = 0.1;
  p->xi0 = 0.0;
  p->use_graph = 0;
  p->fused = 1;
  if (problem == SIS_MAXCUT) {
    p->dynamics = SIS_DSB;
    p->agents = 1024;
    p->coupling = 1.0;
    p->window_period = 100;
    p->window_threshold = 25;
  } else {
    p->dynamics = SIS_BSB;
    p->agents = 10000;
    p->coupling = 1.005;
    p->xi0 = 0.002;
    p->window_period = 0;
    p->window_threshold = 50;
  }
  return SIS_OK;
}

int sis_solver_create(const sis_csr *J, const double *field, const sis_params *pin,
                      sis_solver **out) {
  if (!J || !pin || !out) return fail(SIS_ERR_ARG, "sis_solver_create: null argument");
  *out = nullptr;
  std::unique_ptr<sis_solver> s(new (std::nothrow) sis_solver());
  if (!s) return fail(SIS_ERR_NOMEM, "out of host memory");
  sis_params p = *pin;
  if (p.problem != SIS_MAXCUT && p.problem != SIS_TSP)
    return fail(SIS_ERR_ARG, "sis_solver_create: unknown problem");
  if (p.dynamics != SIS_DSB && p.dynamics != SIS_BSB)
    return fail(SIS_ERR_ARG, "sis_solver_create: unknown dynamics");
  if (p.form == SIS_FORM_PACKED && p.dynamics != SIS_DSB)
    return fail(SIS_ERR_ARG, "sis_solver_create: the packed-sign form is dSB only");
  if (p.form == SIS_FORM_BF16 && p.dynamics != SIS_BSB)
    return fail(SIS_ERR_ARG, "sis_solver_create: the BF16 form is bSB only");
  if (p.form < SIS_FORM_F32 || p.form > SIS_FORM_Q8S)
    return fail(SIS_ERR_ARG, "sis_solver_create: unknown form");
  if (qn_width_of(p.form)) {
    if (p.problem != SIS_MAXCUT || p.dynamics != SIS_DSB || !p.fused)
      return fail(SIS_ERR_ARG, "sis_solver_create: the narrow-state forms are fused MaxCut dSB only");
    if (!(p.dt > 0))
      return fail(SIS_ERR_ARG, "sis_solver_create: the narrow-state forms need dt > 0");
    if (p.split != 0)
      return fail(SIS_ERR_ARG, "sis_solver_create: the narrow-state forms have no row split");
  }
  if (p.agents < 1 || p.agents > (1 << 18))
    return fail(SIS_ERR_ARG, "sis_solver_create: agents must be in [1, 262144]");
  if (!(p.dt == p.dt)) return fail(SIS_ERR_ARG, "sis_solver_create: dt is NaN");
  if (p.split < -1) return fail(SIS_ERR_ARG, "sis_solver_create: split must be >= -1");
  if (p.row_order < -1 || p.row_order > 1)
    return fail(SIS_ERR_ARG, "sis_solver_create: row_order must be -1, 0 or 1");
  if (p.structured)
    return fail(SIS_ERR_ARG, "sis_solver_create: structured TSP solvers come from sis_solver_create_tsp");
  if (p.window_threshold < 1) p.window_threshold = p.problem == SIS_MAXCUT ? 25 : 50;
  s->N = J->n;
  s->Z = (int64_t)J->col.size();
  s->S = J->sum;
  s->pm1 = true;
  for (double v : J->val)
    if (v != 1.0 && v != -1.0) {
      s->pm1 = false;
      break;
    }
  if (s->N < 1) return fail(SIS_ERR_ARG, "sis_solver_create: empty matrix");
  if (p.problem == SIS_TSP) {
    if (p.ncities < 1 || (int64_t)p.ncities * p.ncities != s->N)
      return fail(SIS_ERR_ARG, "sis_solver_create: TSP needs ncities^2 == N");
    if (!field) return fail(SIS_ERR_ARG, "sis_solver_create: TSP needs the field h");
  }
  if (p.xi0 <= 0.0) {
    if (p.problem == SIS_TSP) {
      p.xi0 = 0.002;
    } else {
      int st = sis_maxcut_xi0(J, &p.xi0);
      if (st) return st;
    }
  }
  s->A = p.agents;
  const int qw = qn_width_of(p.form);
  const bool narrow = qw != 0;
  s->lda = narrow ? (s->A + 255) / 256 * 256 : (s->A + 63) / 64 * 64;
  s->qshift = narrow ? qn_shift_for(qw, p.dt) : 0;
  if (narrow) {
    // The step leaves the Y clamp out, which is exact only while the Y code
    // range covers the bound 2/dt + dt of a live element's y.
    if (std::ldexp((double)((1 << (qw - 1)) - 1), -s->qshift) < 2.0 / p.dt + p.dt)
      return fail(SIS_ERR_ARG, "sis_solver_create: dt too small for the narrow-state form's "
                               "Y code range (needs (2^(w-1) - 1) >= 2/dt + dt)");
  }
  s->W = s->lda / 64;
  s->f64 = p.form == SIS_FORM_F64;
  s->gather = p.dynamics == SIS_DSB ? G_DSB : G_BSB;
  if (p.problem == SIS_TSP) p.fused = 1;
  const int prob = p.problem == SIS_TSP ? PROB_TSP : (p.fused ? PROB_MAXCUT : PROB_UNFUSED);
  if (prob == PROB_UNFUSED &&
      (s->gather != G_DSB || (p.form != SIS_FORM_F32 && p.form != SIS_FORM_F64)))
    return fail(SIS_ERR_ARG, "sis_solver_create: the unfused step is plain-form dSB only");
  s->prob = prob;
  // Row storage order (sis_params.row_order): the locality order for the
  // sign-gather forms when the weights are integers (the FP64 cut sums are
  // then exact in any row order) and, by default, when the sign words
  // outgrow an L2 slice.
  sis_csr Jr;
  {
    bool integer_w = true;
    double wsum = 0.0;
    for (double v : J->val) {
      if (!(v == std::floor(v))) integer_w = false;
      wsum += std::fabs(v);
    }
    integer_w = integer_w && wsum < 9007199254740992.0;  // 2^53
    const bool sign_gather = prob == PROB_MAXCUT && s->gather == G_DSB &&
                             (p.form == SIS_FORM_PACKED || narrow);
    const bool want = p.row_order == 1 ||
                      (p.row_order == 0 && (double)s->N * (double)s->lda / 8.0 > 4.0 * 1024 * 1024);
    if (sign_gather && integer_w && want && s->N > 1) {
      locality_order(*J, 6, s->node_of);
      s->row_inv.assign((size_t)s->N, 0);
      for (int64_t r = 0; r < s->N; r++) s->row_inv[(size_t)s->node_of[(size_t)r]] = (int32_t)r;
      reorder_csr(*J, s->node_of, s->row_inv, Jr);
      J = &Jr;
    }
    p.row_order = s->node_of.empty() ? -1 : 1;
  }
  s->p = p;
  if (int st0 = build_ell(s.get(), J)) return st0;  // resident tiles' matrix copies
  const int vec = p.tile > 0 ? p.tile : default_tile(s.get());
  if (!tile_ok(s.get(), vec))
    return fail(SIS_ERR_ARG, "sis_solver_create: unsupported tile " + std::to_string(vec) +
                                 " for " + std::to_string(s->A) + " agents");
  s->vec = vec;
  if (s->N * s->W >= INT32_MAX || s->N * (s->lda / 64) >= INT32_MAX)
    return fail(SIS_ERR_ARG, "sis_solver_create: N x agents too large for one launch");

  HIPCHK(hipGetDevice(&s->dev));
  HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  HIPCHK(hipHostMalloc((void **)&s->h_ctrl, sizeof(Ctrl), hipHostMallocDefault));

  int st;
  const int64_t N = s->N, Z = s->Z, NL = N * s->lda, NW = N * s->W;
  if ((st = s->rp.alloc((size_t)N + 1)) || (st = s->col.alloc((size_t)Z)) ||
      (st = s->vald.alloc((size_t)Z)))
    return st;
  HIPCHK(hipMemcpy(s->rp.p, J->rp.data(), sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice));
  s->rp_host.assign(J->rp.begin(), J->rp.begin() + N + 1);
  if (!s->node_of.empty()) {
    if ((st = s->node_of_d.alloc((size_t)N))) return st;
    HIPCHK(hipMemcpy(s->node_of_d.p, s->node_of.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
  }
  for (int64_t i = 0; i < N; i++) s->max_deg = std::max(s->max_deg, J->rp[i + 1] - J->rp[i]);

  if (Z) {
    HIPCHK(hipMemcpy(s->col.p, J->col.data(), sizeof(int32_t) * Z, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->vald.p, J->val.data(), sizeof(double) * Z, hipMemcpyHostToDevice));
  }
  if (!s->f64) {
    if ((st = s->valf.alloc((size_t)Z))) return st;
    std::vector<float> vf(J->val.begin(), J->val.end());
    if (Z) HIPCHK(hipMemcpy(s->valf.p, vf.data(), sizeof(float) * Z, hipMemcpyHostToDevice));
  }
  if (field) {
    s->has_field = true;
    if ((st = s->hd.alloc((size_t)N))) return st;
    HIPCHK(hipMemcpy(s->hd.p, field, sizeof(double) * N, hipMemcpyHostToDevice));
    if (!s->f64) {
      if ((st = s->hf.alloc((size_t)N))) return st;
      std::vector<float> hf(field, field + N);
      HIPCHK(hipMemcpy(s->hf.p, hf.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    }
  }
  // State.
  if (s->f64) {
    if ((st = s->Xd[0].alloc((size_t)NL)) || (st = s->Xd[1].alloc((size_t)NL)) ||
        (st = s->Yd.alloc((size_t)NL)))
      return st;
  } else if (s->narrow()) {
    // 3, 2.5 or 2 dword planes of N x (lda / 256) x 64 lanes (12, 10 or 8
    // bits): 3, 2.5 or 2 bytes per node-agent
    if ((st = s->Xq.alloc((size_t)(qn_half_dwords(s->qw()) * N * (s->lda / 8)))) ||
        (st = s->sg[0].alloc((size_t)NW)) || (st = s->sg[1].alloc((size_t)NW)))
      return st;
  } else {
    if ((st = s->Xf[0].alloc((size_t)NL)) || (st = s->Yf.alloc((size_t)NL))) return st;
    if (p.form == SIS_FORM_F32 && (st = s->Xf[1].alloc((size_t)NL))) return st;
    if (p.form == SIS_FORM_BF16 &&
        ((st = s->sh[0].alloc((size_t)NL)) || (st = s->sh[1].alloc((size_t)NL))))
      return st;
    if (p.form == SIS_FORM_PACKED &&
        ((st = s->sg[0].alloc((size_t)NW)) || (st = s->sg[1].alloc((size_t)NW))))
      return st;
  }
  // Window state.
  s->scan_rows = std::max<int64_t>(1, std::min<int64_t>(512, (N + 2047) / 2048));
  s->scan_blocks = (N + s->scan_rows - 1) / s->scan_rows;
  s->cut_rows = std::max<int64_t>(4, std::min<int64_t>(256, (N + 1023) / 1024));
  s->cut_blocks = (N + s->cut_rows - 1) / s->cut_rows;
  if ((st = s->cur.alloc((size_t)NW)) || (st = s->fin.alloc((size_t)NW)) ||
      (st = s->newmask.alloc((size_t)s->W)) ||
      (st = s->part.alloc((size_t)(s->scan_blocks * s->W))) ||
      (st = s->stab.alloc((size_t)s->lda)) || (st = s->bif.alloc((size_t)s->lda)) ||
      (st = s->valid.alloc((size_t)s->lda)) || (st = s->ctrl.alloc(1)) ||
      (st = s->cutpart.alloc((size_t)(s->cut_blocks * s->lda))) ||
      (st = s->outv.alloc((size_t)s->lda)) || (st = s->best.alloc(1)))
    return st;
  if (p.problem == SIS_TSP && (st = s->tsp_lines.alloc((size_t)(2 * p.ncities * s->W)))) return st;
  if (p.split > 0 && (st = build_split(s.get(), p.split))) return st;
  *out = s.release();
  return SIS_OK;
}

int sis_solver_create_tsp(const sis_tsp *t, const sis_params *pin, sis_solver **out) {
  if (!t || !pin || !out) return fail(SIS_ERR_ARG, "sis_solver_create_tsp: null argument");
  *out = nullptr;
  sis_params p = *pin;
  p.problem = SIS_TSP;
  p.ncities = (int32_t)t->n;
  const int64_t n = t->n, N = n * n;
  const bool structured = p.structured != 0;
  p.structured = 0;
  std::vector<double> h((size_t)N);
  if (!structured) {
    sis_csr *J = nullptr;
    int st = sis_tsp_qubo(t, &J, h.data());
    if (st) return st;
    st = sis_solver_create(J, h.data(), &p, out);
    sis_csr_free(J);
    return st;
  }
  if (n < 3) return fail(SIS_ERR_ARG, "sis_solver_create_tsp: the structured product needs n >= 3");
  if (p.form != SIS_FORM_F32 && p.form != SIS_FORM_F64)
    return fail(SIS_ERR_ARG, "sis_solver_create_tsp: the structured product runs the plain forms");
  // v (n x 64 values of the form) is staged in LDS: at most 128 KB
  if (n > (p.form == SIS_FORM_F32 ? 512 : 256))
    return fail(SIS_ERR_ARG, "sis_solver_create_tsp: structured n must be <= 512 (f32) / 256 (f64)");
  if (p.split != 0) return fail(SIS_ERR_ARG, "sis_solver_create_tsp: no row split without a matrix");
  // the QUBO's constants and field (src/tsp_gpu.cu:81-83, src/kernels.cu:591-592)
  const double A = 1, B = t->maxd, C = t->maxd;
  for (int64_t i = 0; i < n; i++)
    for (int64_t k = 0; k < n; k++)
      h[(size_t)(i * n + k)] = -1 * (A * t->totals[(size_t)k] + (double)(n - 2) * (B + C)) / 2;
  sis_csr E;  // no stored entry: the product is structured
  E.n = N;
  E.rp.assign((size_t)N + 1, 0);
  E.sum = 0.0;
  p.tile = 0;  // the structured tile (checked below) is not a row-kernel tile
  sis_solver *sp = nullptr;
  int st = sis_solver_create(&E, h.data(), &p, &sp);
  if (st) return st;
  std::unique_ptr<sis_solver> s(sp);
  s->structured = true;
  s->p.structured = 1;

  s->tsp_n = n;
  s->qB = B / 4;
  s->qC = C / 4;
  std::vector<double> dq((size_t)(n * n), 0.0);
  for (int64_t k = 0; k < n; k++)
    for (int64_t l = 0; l < n; l++)
      if (k != l)
        dq[(size_t)(k * n + l)] = A * t->dist[(size_t)tri(std::min(k, l), std::max(k, l), n)] / 8;
  const size_t esz = s->f64 ? sizeof(double) : sizeof(float);
  if ((st = s->dqd.alloc(dq.size())) || (st = s->tsum_r.alloc((size_t)(n * s->lda) * esz)) ||
      (st = s->tsum_c.alloc((size_t)(n * s->lda) * esz)) ||
      (st = s->esum_r.alloc((size_t)(n * s->lda))) || (st = s->esum_c.alloc((size_t)(n * s->lda))))
    return st;
  HIPCHK(hipMemcpy(s->dqd.p, dq.data(), sizeof(double) * dq.size(), hipMemcpyHostToDevice));
  if (!s->f64) {
    std::vector<float> dqf(dq.begin(), dq.end());
    if ((st = s->dqf.alloc(dqf.size()))) return st;
    HIPCHK(hipMemcpy(s->dqf.p, dqf.data(), sizeof(float) * dqf.size(), hipMemcpyHostToDevice));
    // (n rounded up to 32) x nep, zero outside n x n
    const int64_t nep = (n + 7) & ~(int64_t)7, rows = (n + 31) / 32 * 32;
    std::vector<float> dm((size_t)(rows * nep), 0.0f);
    for (int64_t k = 0; k < n; k++)
      for (int64_t l = 0; l < n; l++) dm[(size_t)(k * nep + l)] = dqf[(size_t)(k * n + l)];
    if ((st = s->dqm.alloc(dm.size()))) return st;
    HIPCHK(hipMemcpy(s->dqm.p, dm.data(), sizeof(float) * dm.size(), hipMemcpyHostToDevice));
    s->tsp_nep = nep;
    // resident tile: the most agents per workgroup that fit (4, 2, 1)
    for (int G : {4, 2, 1})
      if (s->lda % G == 0 && tsp_resident_ok(n, G, nep)) {
        s->tsp_res_G = G;
        break;
      }
  }
  // tile: 1 = VALU distance product, 2 = matrix cores (default for FP32),
  // 3 = resident workgroups (FP32; n <= 48 at 4 agents per workgroup, <= 96 at 1)
  if (pin->tile == 0) s->vec = s->f64 ? 1 : 2;
  else if (pin->tile == 1 || (pin->tile == 2 && !s->f64) || (pin->tile == 3 && s->tsp_res_G))
    s->vec = pin->tile;
  else
    return fail(SIS_ERR_ARG, "sis_solver_create_tsp: structured tile must be 1 (VALU), 2 (MFMA, "
                             "f32) or 3 (resident, f32, n <= 96)");
  *out = s.release();
  return SIS_OK;
}

int sis_solver_destroy(sis_solver *s) {
  if (!s) return SIS_OK;
  (void)hipStreamSynchronize(s->stream);
  delete s;
  return SIS_OK;
}

int sis_solver_params(const sis_solver *s, sis_params *p) {
  if (!s || !p) return fail(SIS_ERR_ARG, "sis_solver_params: null argument");
  *p = s->p;
  p->tile = s->vec;
  p->split = (int32_t)s->split;
  return SIS_OK;
}

}  // extern "C"

// Spins and window state back to the reference's reset values
// (src/ising.cu:127-131, plus stability = 0); step counter to zero.
static int reset_windows(sis_solver *s) {
  const int64_t NW = s->N * s->W;
  HIPCHK(hipMemsetAsync(s->cur.p, 0xff, sizeof(uint64_t) * NW, s->stream));
  HIPCHK(hipMemsetAsync(s->fin.p, 0xff, sizeof(uint64_t) * NW, s->stream));
  HIPCHK(hipMemsetAsync(s->stab.p, 0, sizeof(int32_t) * s->lda, s->stream));
  HIPCHK(hipMemsetAsync(s->bif.p, 0, sizeof(int32_t) * s->lda, s->stream));
  HIPCHK(hipMemsetAsync(s->valid.p, 0, sizeof(int32_t) * s->lda, s->stream));
  HIPCHK(hipMemsetAsync(s->newmask.p, 0, sizeof(uint64_t) * s->W, s->stream));
  Ctrl c{};
  c.merge = s->p.problem == SIS_TSP ? 1 : 0;
  HIPCHK(hipMemcpyAsync(s->ctrl.p, &c, sizeof(Ctrl), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return SIS_OK;
}

// Derived gather buffers (BF16 shadow / sign words) from the current X.
static int refresh_shadow(sis_solver *s) {
  const int64_t NL = s->N * s->lda;
  if (s->p.form == SIS_FORM_BF16) launch_bf16_shadow(s->stream, s->Xf[0].p, s->sh[s->par].p, NL);
  if (s->p.form == SIS_FORM_PACKED)
    launch_pack_signs(s->stream, SRC_F32, s->Xf[0].p, s->sg[s->par].p, s->N, s->W, s->lda);
  if (s->narrow())
    launch_pack_signs_qn(s->stream, s->qw(), s->Xq.p, s->sg[s->par].p, s->N, s->lda);
  HIPCHK(hipGetLastError());
  return SIS_OK;
}

// Pre-step (src/kernels.cu:8-22 at t = 0) on the current state.
static void enqueue_prestep(sis_solver *s) {
  if (s->narrow())
    launch_prestep_qn(s->stream, s->qw(), s->Xq.p, s->N, s->lda, s->p.dt, s->qshift);
  else
    launch_prestep(s->stream, s->f64, s->Xcur(), s->Ybuf(), s->N * s->lda, s->p.dt);
}

// Schedule table for horizon T covering steps [0, len).
static int ensure_schedule(sis_solver *s, int64_t T, int64_t len) {
  if (s->sched_T == T && s->sched_len >= len) return SIS_OK;
  len = std::max<int64_t>(len, 1);
  // The old table may still be read by queued kernels.
  HIPCHK(hipStreamSynchronize(s->stream));
  std::vector<double> c((size_t)(4 * len));
  const double dt = s->p.dt, xi0 = s->p.xi0;
  for (int64_t t = 0; t < len; t++) {
    double *q = &c[(size_t)(4 * t)];
    if (s->p.problem == SIS_MAXCUT) {
      // src/kernels.cu:225-226, 258-262; the unfused path leaves p
      // unclamped (src/kernels.cu:15)
      double pr = 1.025 * (double)t / (double)T;
      if (s->prob != PROB_UNFUSED) pr = pr < 1.0 ? pr : 1.0;
      q[0] = dt * (pr + 1);
      q[1] = dt * (pr - 1);
      q[2] = dt * xi0;
      q[3] = 0.0;
    } else {
      // src/kernels.cu:553-559, 592-595 with ncap = 2 (src/ising_tsp.cu:134)
      const int ncap = 2;
      double at = ncap * 1.04 * (double)t / (double)T;
      double bt = 0.5 + (0.5 * std::pow((double)t / (double)T, 1.9));
      at = at < ncap ? at : ncap;
      bt = bt < 1 ? bt : 1;
      q[0] = dt;
      q[1] = dt * -(1 - at);
      q[2] = xi0;
      q[3] = bt;
    }
  }
  int st;
  if (s->f64) {
    if ((st = s->schedd.alloc(c.size()))) return st;
    HIPCHK(hipMemcpy(s->schedd.p, c.data(), sizeof(double) * c.size(), hipMemcpyHostToDevice));
  } else {
    std::vector<float> cf(c.begin(), c.end());
    if ((st = s->schedf.alloc(cf.size()))) return st;
    HIPCHK(hipMemcpy(s->schedf.p, cf.data(), sizeof(float) * cf.size(), hipMemcpyHostToDevice));
  }
  s->sched_T = T;
  s->sched_len = len;
  s->drop_graphs();  // captured launches hold the old table's address
  return SIS_OK;
}

// Enqueue one fused step at offset `off` from ctrl->base, flipping par.
static void enqueue_step(sis_solver *s, int32_t off) {
  StepArgs a{};
  a.rp = s->rp.p;
  a.col = s->col.p;
  a.val = s->f64 ? (const void *)s->vald.p : (const void *)s->valf.p;
  a.field = s->has_field ? (s->f64 ? (const void *)s->hd.p : (const void *)s->hf.p) : nullptr;
  a.sched = s->f64 ? (const void *)s->schedd.p : (const void *)s->schedf.p;
  a.ctrl = s->ctrl.p;
  a.offset = off;
  a.N = s->N;
  a.lda = s->lda;
  a.W = s->W;
  a.coupling = s->p.coupling;
  a.Y = s->Ybuf();
  const int cur = s->par, nxt = s->par ^ 1;
  switch (s->p.form) {
    case SIS_FORM_F32:
    case SIS_FORM_F64:
      a.gsrc = s->Xbuf(cur);
      a.gdst = s->Xbuf(nxt);
      a.X = nullptr;
      break;
    case SIS_FORM_BF16:
      a.gsrc = s->sh[cur].p;
      a.gdst = s->sh[nxt].p;
      a.X = s->Xf[0].p;
      break;
    case SIS_FORM_PACKED:
      a.gsrc = s->sg[cur].p;
      a.gdst = s->sg[nxt].p;
      a.X = s->Xf[0].p;
      break;
    case SIS_FORM_Q12:
    case SIS_FORM_Q10:
    case SIS_FORM_Q8:
    case SIS_FORM_Q8S:
      a.gsrc = s->sg[cur].p;
      a.gdst = s->sg[nxt].p;
      a.X = s->Xq.p;
      a.Y = nullptr;
      a.qshift = s->qshift;
      a.rowid = s->node_of_d.p;  // the dither is keyed by node (nullptr: node order)
      break;
  }
  if (s->prob == PROB_UNFUSED)
    launch_xy_update(s->stream, s->f64, s->Xbuf(cur), a.Y, s->N * s->lda, a.sched, a.ctrl, off);
  if (s->structured) {
    launch_tsp_struct_step(s->stream, s->p.form, s->gather, (s->vec & 0xffff) == 2, a,
                           s->tsp_args(false));
    s->par = nxt;
    return;
  }
  a.split = s->split_nent > 0 ? s->split : 0;
  if (a.split) {
    SplitArgs sp{s->split_ent.p, s->split_part.p, s->split_cnt.p, s->split};
    launch_step_split(s->stream, s->p.form, s->prob, s->gather, s->split_nent, a, sp);
  }
  launch_step(s->stream, s->p.form, s->prob, s->gather, s->vec, a);
  s->par = nxt;
}

// Steps off0 .. off0 + n - 1 of the current window period: n fused steps, or
// with a resident tile one k_resident launch per 1000 steps (then, for the
// packed form, the sign words of the final X for the window and the next
// step).
static void enqueue_steps(sis_solver *s, int32_t off0, int64_t n) {
  ResArgs r{};
  if (s->structured && s->vec == 3) {
    StepArgs a{};
    a.field = s->hf.p;
    a.sched = s->schedf.p;
    a.ctrl = s->ctrl.p;
    a.N = s->N;
    a.lda = s->lda;
    a.W = s->W;
    a.coupling = s->p.coupling;
    a.Y = s->Yf.p;
    r.G = s->tsp_res_G;
    r.lgG = s->tsp_res_G == 4 ? 2 : s->tsp_res_G == 2 ? 1 : 0;
    for (int64_t k0 = 0; k0 < n; k0 += 1000) {
      const int32_t m = (int32_t)std::min<int64_t>(1000, n - k0);
      a.offset = (int32_t)(off0 + k0);
      r.steps = m;
      r.Xin = s->Xf[s->par].p;
      r.Xout = s->Xf[s->par ^ (m & 1)].p;
      launch_tsp_resident(s->stream, s->gather, a, s->tsp_args(false), r);
      s->par ^= (m & 1);
    }
    return;
  }
  if (!resident_layout(s, s->vec, &r)) {
    for (int64_t k = 0; k < n; k++) enqueue_step(s, (int32_t)(off0 + k));
    return;
  }
  StepArgs a{};
  a.rp = s->rp.p;
  a.col = s->col.p;
  a.val = s->valf.p;
  a.sched = s->schedf.p;
  a.ctrl = s->ctrl.p;
  a.N = s->N;
  a.lda = s->lda;
  a.W = s->W;
  a.coupling = s->p.coupling;
  a.Y = s->Yf.p;
  const bool packed = s->p.form == SIS_FORM_PACKED;
  for (int64_t k0 = 0; k0 < n; k0 += 1000) {
    const int32_t m = (int32_t)std::min<int64_t>(1000, n - k0);
    a.offset = (int32_t)(off0 + k0);
    r.steps = m;
    r.Xin = packed ? s->Xf[0].p : s->Xf[s->par].p;
    r.Xout = packed ? s->Xf[0].p : s->Xf[s->par ^ (m & 1)].p;
    if (!launch_resident(s->stream, s->gather, a, r)) {
      // no kernel instance for this size (resident_layout admits none such):
      // the remaining steps go through the row kernels
      if (packed)
        launch_pack_signs(s->stream, SRC_F32, s->Xf[0].p, s->sg[s->par].p, s->N, s->W, s->lda);
      const int keep = s->vec;
      s->vec = default_tile(s);
      for (int64_t k = k0; k < n; k++) enqueue_step(s, (int32_t)(off0 + k));
      s->vec = keep;
      return;
    }
    s->par ^= (m & 1);
  }
  if (packed)
    launch_pack_signs(s->stream, SRC_F32, s->Xf[0].p, s->sg[s->par].p, s->N, s->W, s->lda);
}

// Enqueue the window kernels that follow a period of P steps.
static void enqueue_window(sis_solver *s, int64_t P, int uw) {
  int src;
  const void *x;
  switch (s->p.form) {
    case SIS_FORM_PACKED:
    case SIS_FORM_Q12:
    case SIS_FORM_Q10:
    case SIS_FORM_Q8:
    case SIS_FORM_Q8S: src = SRC_BITS; x = s->sg[s->par].p; break;
    case SIS_FORM_F64: src = SRC_F64; x = s->Xcur(); break;
    default: src = SRC_F32; x = s->Xcur(); break;
  }
  launch_window_scan(s->stream, src, x, s->cur.p, s->part.p, s->N, s->W, s->lda, s->scan_blocks,
                     s->scan_rows, s->ctrl.p);
  if (s->p.problem == SIS_MAXCUT) {
    launch_window_finalize(s->stream, s->part.p, s->scan_blocks, s->newmask.p, s->stab.p, s->bif.p,
                           s->A, s->W, s->p.window_threshold, uw, P, s->ctrl.p);
    launch_window_merge(s->stream, s->cur.p, s->fin.p, s->newmask.p, s->N, s->W, s->ctrl.p);
  } else {
    launch_tsp_valid(s->stream, s->cur.p, s->p.ncities, s->W, s->A, s->valid.p, s->newmask.p,
                     s->tsp_lines.p, P,
                     s->ctrl.p);
    launch_window_merge(s->stream, s->cur.p, s->fin.p, s->newmask.p, s->N, s->W, s->ctrl.p);
    launch_ctrl_advance(s->stream, s->ctrl.p, P);
  }
}

static int ensure_events(sis_solver *s, size_t n) {
  while (s->events.size() < n) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    s->events.push_back(e);
  }
  return SIS_OK;
}

// Captures one window period (P steps + window) for each snapshot parity.
static int build_graphs(sis_solver *s, int64_t P, int uw) {
  if (s->graph_P == P && s->graph_uw == uw && s->graph_exec[0] && s->graph_exec[1]) return SIS_OK;
  for (auto &g : s->graph_exec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  const int saved = s->par;
  for (int ph = 0; ph < 2; ph++) {
    s->par = ph;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    enqueue_steps(s, 0, P);
    enqueue_window(s, P, uw);
    HIPCHK(hipStreamEndCapture(s->stream, &g));
    hipError_t e = hipGraphInstantiate(&s->graph_exec[ph], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(SIS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  s->par = saved;
  s->graph_P = P;
  s->graph_uw = uw;
  return SIS_OK;
}

// Captures a whole run (nwin window periods) as one graph.  Kernels after an
// update-window exit see the device flag and return at once.
static int build_run_graph(sis_solver *s, int64_t P, int64_t nwin, int uw) {
  if (s->run_graph && s->run_P == P && s->run_nwin == nwin && s->run_uw == uw &&
      s->run_par == s->par)
    return SIS_OK;
  if (s->run_graph) {
    (void)hipGraphExecDestroy(s->run_graph);
    s->run_graph = nullptr;
  }
  const int saved = s->par;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  for (int64_t w = 0; w < nwin; w++) {
    enqueue_steps(s, 0, P);
    enqueue_window(s, P, uw);
  }
  HIPCHK(hipStreamEndCapture(s->stream, &g));
  hipError_t e = hipGraphInstantiate(&s->run_graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  s->par = saved;
  if (e != hipSuccess) {
    s->run_graph = nullptr;
    return fail(SIS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  s->run_P = P;
  s->run_nwin = nwin;
  s->run_uw = uw;
  s->run_par = saved;
  return SIS_OK;
}

extern "C" {

int sis_solver_init_random(sis_solver *s, uint64_t seed) {
  if (!s) return fail(SIS_ERR_ARG, "sis_solver_init_random: null solver");
  int st = reset_windows(s);
  if (st) return st;
  s->par = 0;
  if (s->narrow())
    launch_init_qn(s->stream, s->qw(), s->Xq.p, s->N, s->A, s->lda, seed, s->qshift,
                   s->node_of_d.p);
  else
    launch_init_state(s->stream, s->f64, s->Xbuf(0), s->Ybuf(), s->N, s->A, s->lda, seed,
                      s->p.problem == SIS_TSP, s->node_of_d.p);
  HIPCHK(hipGetLastError());
  if ((st = refresh_shadow(s))) return st;
  HIPCHK(hipStreamSynchronize(s->stream));
  s->state_ready = true;
  return SIS_OK;
}

int sis_solver_set_state(sis_solver *s, const double *x, const double *y) {
  if (!s || !x || !y) return fail(SIS_ERR_ARG, "sis_solver_set_state: null argument");
  int st = reset_windows(s);
  if (st) return st;
  s->par = 0;
  const int64_t N = s->N, A = s->A, lda = s->lda;
  if (s->narrow()) {
    // FP32 values (as the other FP32 forms take them), rounded to codes;
    // padding agents 0
    const int64_t Q = lda / 256, qw = s->qw();
    std::vector<uint32_t> P((size_t)(s->Xq.n), 0u);
    for (int64_t i = 0; i < N; i++)
      for (int64_t a = 0; a < A; a++)
        qn_put(qw, P, N, Q, s->row_of(i), a,
               qn_unit_host(qw, (float)x[i * A + a], (float)y[i * A + a], s->qshift));
    HIPCHK(hipMemcpy(s->Xq.p, P.data(), sizeof(uint32_t) * P.size(), hipMemcpyHostToDevice));
  } else if (s->f64) {
    std::vector<double> bx((size_t)(N * lda), 0.0), by((size_t)(N * lda), 0.0);
    for (int64_t i = 0; i < N; i++)
      for (int64_t a = 0; a < A; a++) {
        bx[(size_t)(s->row_of(i) * lda + a)] = x[i * A + a];
        by[(size_t)(s->row_of(i) * lda + a)] = y[i * A + a];
      }
    HIPCHK(hipMemcpy(s->Xd[0].p, bx.data(), sizeof(double) * bx.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->Yd.p, by.data(), sizeof(double) * by.size(), hipMemcpyHostToDevice));
  } else {
    std::vector<float> bx((size_t)(N * lda), 0.0f), by((size_t)(N * lda), 0.0f);
    for (int64_t i = 0; i < N; i++)
      for (int64_t a = 0; a < A; a++) {
        bx[(size_t)(s->row_of(i) * lda + a)] = (float)x[i * A + a];
        by[(size_t)(s->row_of(i) * lda + a)] = (float)y[i * A + a];
      }
    HIPCHK(hipMemcpy(s->Xf[0].p, bx.data(), sizeof(float) * bx.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->Yf.p, by.data(), sizeof(float) * by.size(), hipMemcpyHostToDevice));
  }
  if ((st = refresh_shadow(s))) return st;
  HIPCHK(hipStreamSynchronize(s->stream));
  s->state_ready = true;
  return SIS_OK;
}

int sis_solver_get_state(sis_solver *s, double *x, double *y) {
  if (!s) return fail(SIS_ERR_ARG, "sis_solver_get_state: null solver");
  HIPCHK(hipStreamSynchronize(s->stream));
  const int64_t N = s->N, A = s->A, lda = s->lda;
  auto copy_out = [&](const void *src, double *dst) -> int {
    if (!dst) return SIS_OK;
    if (s->f64) {
      std::vector<double> b((size_t)(N * lda));
      HIPCHK(hipMemcpy(b.data(), src, sizeof(double) * b.size(), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < N; i++)
        for (int64_t a = 0; a < A; a++) dst[i * A + a] = b[(size_t)(s->row_of(i) * lda + a)];
    } else {
      std::vector<float> b((size_t)(N * lda));
      HIPCHK(hipMemcpy(b.data(), src, sizeof(float) * b.size(), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < N; i++)
        for (int64_t a = 0; a < A; a++) dst[i * A + a] = b[(size_t)(s->row_of(i) * lda + a)];
    }
    return SIS_OK;
  };
  if (s->narrow()) {
    const int64_t Q = lda / 256;
    const int qw = s->qw();
    const uint32_t m = (1u << qw) - 1;
    const int32_t bias = 1 << (qw - 1);
    std::vector<uint32_t> P((size_t)(s->Xq.n));
    HIPCHK(hipMemcpy(P.data(), s->Xq.p, sizeof(uint32_t) * P.size(), hipMemcpyDeviceToHost));
    const double yinv = std::ldexp(1.0, -s->qshift);
    for (int64_t i = 0; i < N; i++)
      for (int64_t a = 0; a < A; a++) {
        const uint32_t u = qn_get(qw, P, N, Q, s->row_of(i), a);
        const int32_t xq = (int32_t)(u & m) - bias, yq = (int32_t)((u >> qw) & m) - bias;
        if (x) x[i * A + a] = std::ldexp((double)xq, -(qw - 2));
        if (y) y[i * A + a] = (double)yq * yinv;
      }
    return SIS_OK;
  }
  int st = copy_out(s->Xcur(), x);
  if (st) return st;
  return copy_out(s->Ybuf(), y);
}

int sis_solver_prestep(sis_solver *s) {
  if (!s || !s->state_ready) return fail(SIS_ERR_STATE, "sis_solver_prestep: state not set");
  enqueue_prestep(s);
  HIPCHK(hipGetLastError());
  int st = refresh_shadow(s);
  if (st) return st;
  HIPCHK(hipStreamSynchronize(s->stream));
  return SIS_OK;
}

int sis_solver_advance(sis_solver *s, int64_t t0, int64_t nsteps, int64_t T) {
  if (!s || !s->state_ready) return fail(SIS_ERR_STATE, "sis_solver_advance: state not set");
  if (t0 < 0 || nsteps < 0 || T < 1 || t0 + nsteps > (int64_t)INT32_MAX)
    return fail(SIS_ERR_ARG, "sis_solver_advance: bad step range");
  int st = ensure_schedule(s, T, t0 + nsteps);
  if (st) return st;
  Ctrl c{};
  c.base = t0;
  c.merge = s->p.problem == SIS_TSP ? 1 : 0;
  HIPCHK(hipMemcpyAsync(s->ctrl.p, &c, sizeof(Ctrl), hipMemcpyHostToDevice, s->stream));
  enqueue_steps(s, 0, nsteps);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  return SIS_OK;
}

int sis_solver_autotune(sis_solver *s, int64_t trial_steps, int32_t *tile, double *ms_per_step) {
  if (!s || !s->state_ready) return fail(SIS_ERR_STATE, "sis_solver_autotune: state not set");
  if (trial_steps < 1 || trial_steps > 100000)
    return fail(SIS_ERR_ARG, "sis_solver_autotune: trial_steps must be in [1, 100000]");
  // A resident tile runs a whole window period per launch and pays its set-up
  // (matrix into LDS, state into registers) once per launch, so it is timed
  // over one period (the configured one, else 100 steps) when that is longer.
  const int64_t period = s->p.window_period > 0 ? s->p.window_period : 100;
  const int64_t res_steps = std::max<int64_t>(trial_steps, std::min<int64_t>(period, 1000));
  const int64_t warm = 2, T = warm + res_steps;
  int st = ensure_schedule(s, T, T);
  if (st) return st;
  if ((st = ensure_events(s, 2))) return st;
  // Snapshot of the state every candidate starts from (and that is restored).
  const int64_t NL = s->N * s->lda, NW = s->N * s->W;
  // the narrow form keeps X and Y codes in one buffer
  const size_t xbytes = s->narrow() ? sizeof(uint32_t) * s->Xq.n : (size_t)NL * (s->f64 ? 8 : 4);
  const size_t ybytes = s->narrow() ? 0 : xbytes;
  size_t gbytes = 0;
  void *gcur = nullptr;
  if (s->p.form == SIS_FORM_BF16) {
    gbytes = (size_t)NL * 2;
    gcur = s->sh[s->par].p;
  } else if (s->p.form == SIS_FORM_PACKED || s->narrow()) {
    gbytes = (size_t)NW * 8;
    gcur = s->sg[s->par].p;
  }
  DevBuf<char> sx, sy, sgb;
  if ((st = sx.alloc(xbytes)) || (st = sy.alloc(ybytes)) || (st = sgb.alloc(gbytes))) return st;
  const int par0 = s->par;
  void *x0 = s->narrow() ? (void *)s->Xq.p : s->Xcur();
  HIPCHK(hipMemcpyAsync(sx.p, x0, xbytes, hipMemcpyDeviceToDevice, s->stream));
  if (ybytes) HIPCHK(hipMemcpyAsync(sy.p, s->Ybuf(), ybytes, hipMemcpyDeviceToDevice, s->stream));
  if (gbytes) HIPCHK(hipMemcpyAsync(sgb.p, gcur, gbytes, hipMemcpyDeviceToDevice, s->stream));
  auto restore = [&]() -> int {
    s->par = par0;
    HIPCHK(hipMemcpyAsync(x0, sx.p, xbytes, hipMemcpyDeviceToDevice, s->stream));
    if (ybytes) HIPCHK(hipMemcpyAsync(s->Ybuf(), sy.p, ybytes, hipMemcpyDeviceToDevice, s->stream));
    if (gbytes) HIPCHK(hipMemcpyAsync(gcur, sgb.p, gbytes, hipMemcpyDeviceToDevice, s->stream));
    Ctrl c{};
    c.merge = s->p.problem == SIS_TSP ? 1 : 0;
    HIPCHK(hipMemcpyAsync(s->ctrl.p, &c, sizeof(Ctrl), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return SIS_OK;
  };
  auto trial = [&](double *per) -> int {
    if (int e = restore()) return e;
    ResArgs probe{};
    const bool res = (s->structured && s->vec == 3) || resident_layout(s, s->vec, &probe);
    const int64_t steps = res ? res_steps : trial_steps;
    enqueue_steps(s, 0, warm);
    HIPCHK(hipEventRecord(s->events[0], s->stream));
    enqueue_steps(s, (int32_t)warm, steps);
    HIPCHK(hipEventRecord(s->events[1], s->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->events[0], s->events[1]));
    *per = (double)ms / (double)steps;
    return SIS_OK;
  };
  double best_ms = 1e300;
  int best = s->vec;
  for (int v : candidate_tiles(s)) {
    s->vec = v;
    double per;
    if ((st = trial(&per))) return st;
    if (per < best_ms) {
      best_ms = per;
      best = v;
    }
  }
  s->vec = best;
  // Row-split threshold (split = -1 lets the tuner choose): powers of two
  // from 16 below the largest row, against no split, with the chosen tile.
  if (s->p.split == -1 && s->max_deg > 16) {
    int64_t best_tau = s->split;
    for (int64_t tau = 0; tau < s->max_deg; tau = tau ? tau * 2 : 16) {
      if ((st = build_split(s, tau))) return st;
      double per;
      if ((st = trial(&per))) return st;
      if (per < best_ms) {
        best_ms = per;
        best_tau = tau;
      }
    }
    if ((st = build_split(s, best_tau))) return st;
  }
  s->drop_graphs();
  if ((st = restore())) return st;
  if (tile) *tile = best;
  if (ms_per_step) *ms_per_step = best_ms;
  return SIS_OK;
}

int sis_solver_run(sis_solver *s, int64_t max_steps, int update_window, sis_run_info *info) {
  if (!s || !s->state_ready) return fail(SIS_ERR_STATE, "sis_solver_run: state not set");
  if (max_steps < 0) return fail(SIS_ERR_ARG, "sis_solver_run: negative max_steps");
  const int uw = s->p.problem == SIS_MAXCUT && update_window ? 1 : 0;
  const int64_t P = resolve_period(s, max_steps);
  const int64_t nwin = max_steps > 0 ? (max_steps + P - 1) / P : 0;
  if (nwin * P > (int64_t)INT32_MAX) return fail(SIS_ERR_ARG, "sis_solver_run: too many steps");
  int st = ensure_schedule(s, std::max<int64_t>(max_steps, 1), nwin * P);
  if (st) return st;
  if ((st = ensure_events(s, (size_t)(2 * nwin + 2)))) return st;
  // use_graph 2: the whole run as one graph when it stays below ~40k nodes;
  // otherwise (and for use_graph 1) one graph per window period.
  const bool whole = s->p.use_graph == 2 && nwin > 0 && nwin * (P + 4) <= 40000;
  if (whole && (st = build_run_graph(s, P, nwin, uw))) return st;
  if (!whole && s->p.use_graph && nwin > 0 && (st = build_graphs(s, P, uw))) return st;
  // Fresh spins, window state and counters for the run (the reference
  // resets them in reset_initialization, src/ising.cu:127-131).
  if ((st = reset_windows(s))) return st;
  const int par0 = s->par;

  auto t_start = std::chrono::high_resolution_clock::now();
  HIPCHK(hipEventRecord(s->events[0], s->stream));
  enqueue_prestep(s);
  if ((st = refresh_shadow(s))) return st;
  // Early exit is a device flag; the host checks it every `chunk` windows.
  const int64_t chunk = uw ? 4 : nwin;
  int64_t enq = 0;
  bool stopped = false;
  if (whole) {
    HIPCHK(hipGraphLaunch(s->run_graph, s->stream));
    enq = nwin;
  }
  while (enq < nwin && !stopped) {
    const int64_t stop = std::min(nwin, enq + chunk);
    for (; enq < stop; enq++) {
      if (s->p.use_graph) {
        HIPCHK(hipGraphLaunch(s->graph_exec[s->par], s->stream));
        if (P & 1) s->par ^= 1;
      } else {
        HIPCHK(hipEventRecord(s->events[2 + 2 * enq], s->stream));
        enqueue_steps(s, 0, P);
        HIPCHK(hipEventRecord(s->events[3 + 2 * enq], s->stream));
        enqueue_window(s, P, uw);
      }
    }
    HIPCHK(hipGetLastError());
    if (uw && enq < nwin) {
      HIPCHK(hipMemcpyAsync(s->h_ctrl, s->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
      HIPCHK(hipStreamSynchronize(s->stream));
      stopped = s->h_ctrl->done != 0;
    }
  }
  HIPCHK(hipEventRecord(s->events[1], s->stream));
  HIPCHK(hipMemcpyAsync(s->h_ctrl, s->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  auto t_end = std::chrono::high_resolution_clock::now();
  const int64_t executed = s->h_ctrl->base;
  // Each executed step flipped the snapshot parity once.
  s->par = (int)((par0 + executed) & 1);
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->steps_executed = executed;
    info->windows = P > 0 ? executed / P : 0;
    info->all_bifurcated = s->p.problem == SIS_MAXCUT && s->h_ctrl->nbif == s->A && executed > 0;
    info->elapsed_us =
        (double)std::chrono::duration_cast<std::chrono::microseconds>(t_end - t_start).count();
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->events[0], s->events[1]));
    info->device_ms = ms;
    double sk = 0.0;
    if (!s->p.use_graph) {
      for (int64_t w = 0; w < info->windows && w < nwin; w++) {
        float m = 0.f;
        HIPCHK(hipEventElapsedTime(&m, s->events[2 + 2 * w], s->events[3 + 2 * w]));
        sk += m;
      }
    }
    info->step_kernel_ms = sk;
  }
  return SIS_OK;
}

int sis_solver_get_spins(sis_solver *s, int which, int32_t *out) {
  if (!s || !out || (which != SIS_SPINS_CURRENT && which != SIS_SPINS_FINAL))
    return fail(SIS_ERR_ARG, "sis_solver_get_spins: bad argument");
  const int64_t N = s->N, W = s->W, A = s->A;
  std::vector<uint64_t> b((size_t)(N * W));
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(b.data(), which == SIS_SPINS_CURRENT ? s->cur.p : s->fin.p,
                   sizeof(uint64_t) * b.size(), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < N; i++)
    for (int64_t a = 0; a < A; a++)
      out[i * A + a] = ((b[(size_t)(s->row_of(i) * W + a / 64)] >> (a % 64)) & 1) ? -1 : 1;
  return SIS_OK;
}

static int eval_rows(sis_solver *s, const uint64_t *spins, int mode, double *out,
                     int64_t *best_agent, int best_mode) {
  launch_cut_rows(s->stream, s->rp.p, s->col.p, s->vald.p, mode == 1 ? s->hd.p : nullptr, spins,
                  s->N, s->W, s->lda, s->cut_blocks, s->cut_rows, s->cutpart.p);
  launch_cut_reduce(s->stream, s->cutpart.p, s->cut_blocks, s->A, s->lda, mode, s->S, s->outv.p);
  if (best_agent) launch_best(s->stream, s->outv.p, s->A, best_mode, s->best.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  if (out) HIPCHK(hipMemcpy(out, s->outv.p, sizeof(double) * s->A, hipMemcpyDeviceToHost));
  if (best_agent) HIPCHK(hipMemcpy(best_agent, s->best.p, sizeof(int64_t), hipMemcpyDeviceToHost));
  return SIS_OK;
}

int sis_solver_maxcut_cuts(sis_solver *s, double *cut, int64_t *best_agent) {
  if (!s) return fail(SIS_ERR_ARG, "sis_solver_maxcut_cuts: null solver");
  return eval_rows(s, s->cur.p, 0, cut, best_agent, 0);
}

int sis_solver_energy(sis_solver *s, int which, double *energy) {
  if (!s || !energy || (which != SIS_SPINS_CURRENT && which != SIS_SPINS_FINAL))
    return fail(SIS_ERR_ARG, "sis_solver_energy: bad argument");
  if (s->structured) {
    // per-position partials of s^T J s - h.s, reduced in position order
    if ((int64_t)s->cutpart.n < s->tsp_n * s->lda)
      return fail(SIS_ERR_STATE, "sis_solver_energy: partial buffer too small");
    launch_tsp_struct_energy(s->stream, which == SIS_SPINS_CURRENT ? s->cur.p : s->fin.p, s->hd.p,
                             s->tsp_args(true), s->W, s->lda, s->cutpart.p);
    launch_cut_reduce(s->stream, s->cutpart.p, s->tsp_n, s->A, s->lda, 1, 0.0, s->outv.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(energy, s->outv.p, sizeof(double) * s->A, hipMemcpyDeviceToHost));
    return SIS_OK;
  }
  if (!s->has_field) {
    // Without a field: s^T J s = 2 H, from the cut-mode row sums.
    std::vector<double> c((size_t)s->A);
    launch_cut_rows(s->stream, s->rp.p, s->col.p, s->vald.p, nullptr,
                    which == SIS_SPINS_CURRENT ? s->cur.p : s->fin.p, s->N, s->W, s->lda,
                    s->cut_blocks, s->cut_rows, s->cutpart.p);
    launch_cut_reduce(s->stream, s->cutpart.p, s->cut_blocks, s->A, s->lda, 1, 0.0, s->outv.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));  // the stream does not order with hipMemcpy
    HIPCHK(hipMemcpy(c.data(), s->outv.p, sizeof(double) * s->A, hipMemcpyDeviceToHost));
    for (int64_t a = 0; a < s->A; a++) energy[a] = 2.0 * c[(size_t)a];
    return SIS_OK;
  }
  return eval_rows(s, which == SIS_SPINS_CURRENT ? s->cur.p : s->fin.p, 1, energy, nullptr, 0);
}

int sis_solver_agent_state(sis_solver *s, int32_t *stability, int32_t *locked, int32_t *valid_step) {
  if (!s) return fail(SIS_ERR_ARG, "sis_solver_agent_state: null solver");
  HIPCHK(hipStreamSynchronize(s->stream));
  if (stability) HIPCHK(hipMemcpy(stability, s->stab.p, sizeof(int32_t) * s->A, hipMemcpyDeviceToHost));
  if (locked) HIPCHK(hipMemcpy(locked, s->bif.p, sizeof(int32_t) * s->A, hipMemcpyDeviceToHost));
  if (valid_step) HIPCHK(hipMemcpy(valid_step, s->valid.p, sizeof(int32_t) * s->A, hipMemcpyDeviceToHost));
  return SIS_OK;
}

int sis_solver_tsp_lengths(sis_solver *s, const sis_tsp *t, double *length, int64_t *best_agent) {
  if (!s || !t) return fail(SIS_ERR_ARG, "sis_solver_tsp_lengths: null argument");
  if (s->p.problem != SIS_TSP || t->n != s->p.ncities)
    return fail(SIS_ERR_ARG, "sis_solver_tsp_lengths: solver is not this TSP instance");
  if (s->dist_n != t->n || s->dist_host != t->dist) {
    int st = s->dist.alloc(std::max<size_t>(1, t->dist.size()));
    if (st) return st;
    if (!t->dist.empty())
      HIPCHK(hipMemcpy(s->dist.p, t->dist.data(), sizeof(double) * t->dist.size(), hipMemcpyHostToDevice));
    s->dist_n = t->n;
    s->dist_host = t->dist;
  }
  launch_tsp_len(s->stream, s->fin.p, s->valid.p, t->n, s->W, s->A, s->dist.p, s->outv.p);
  if (best_agent) launch_best(s->stream, s->outv.p, s->A, 1, s->best.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  if (length) HIPCHK(hipMemcpy(length, s->outv.p, sizeof(double) * s->A, hipMemcpyDeviceToHost));
  if (best_agent) HIPCHK(hipMemcpy(best_agent, s->best.p, sizeof(int64_t), hipMemcpyDeviceToHost));
  return SIS_OK;
}

}  // extern "C"
