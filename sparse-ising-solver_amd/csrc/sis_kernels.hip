// sis_kernels.hip -- gfx950 kernels: initial state, pre-step, the fused SB
// step in its four forms, the update window, cut / energy and tour lengths.
//
// Every `a += b * c` of the reference step is an explicit fused multiply-add
// here and nothing else is contracted (-ffp-contract=off), so the FP32 and
// FP64 forms are bit-exact with a CPU evaluation in the same order.
#include <atomic>

#include "sis_device.hpp"
#include "sis_kernels.hpp"

namespace sis {

// ------------------------------------------------------- state / prestep

template <typename T>
__global__ void k_init_state(T *X, T *Y, int64_t N, int64_t A, int64_t lda, uint64_t seed,
                             int tsp, const int32_t *rowid) {
  const int64_t n = N * lda;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / lda, a = idx - i * lda;
    T x = 0, y = 0;
    if (a < A) {
      // the element's index in the caller's node order (rows may be stored
      // in another order: rowid[stored row] = node)
      const int64_t node = rowid ? (int64_t)rowid[i] : i;
      const uint64_t e = (uint64_t)(node * A + a);
      const double ux = d_uniform(seed, 2 * e), uy = d_uniform(seed, 2 * e + 1);
      x = tsp ? (T)0 : (T)((ux * 2 - 1) / 10);
      y = (T)((uy * 2 - 1) / 10);
    }
    X[idx] = x;
    Y[idx] = y;
  }
}

// Pre-step at t = 0 (src/kernels.cu:15-20): Y += dt(0-1) X; X += dt(0+1) Y.
template <typename T>
__global__ void k_prestep(T *X, T *Y, int64_t n, T cy, T cx) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    T y = fmaT(cy, X[e], Y[e]);
    X[e] = fmaT(cx, y, X[e]);
    Y[e] = y;
  }
}

// Update_XY inside the unfused step (src/kernels.cu:15-20) with the step's
// coefficients c0 = dt(p'+1), c1 = dt(p'-1), p' = 1.025 t / T unclamped.
template <typename T>
__global__ void k_xy_update(T *X, T *Y, int64_t n, const T *sched, const Ctrl *ctrl,
                            int32_t offset) {
  if (ctrl->done) return;
  const T *sc = sched + 4 * (ctrl->base + offset);
  const T c0 = sc[0], c1 = sc[1];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    T y = fmaT(c1, X[e], Y[e]);
    X[e] = fmaT(c0, y, X[e]);
    Y[e] = y;
  }
}

__global__ void k_bf16_shadow(const float *X, uint16_t *sh, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    sh[e] = d_bf16(X[e]);
}

// One wavefront per (row, word): ballot of the sign bits of 64 agents.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pack_signs(const T *X, uint64_t *S, int64_t N,
                                                       int64_t W, int64_t lda) {
  const int lane = threadIdx.x & 63;
  const int64_t g = wave_index();
  if (g >= N * W) return;
  const int64_t i = g / W, w = g - i * W;
  const uint64_t b = __ballot(signbit(X[i * lda + 64 * w + lane]));
  if (lane == 0) S[g] = b;
}

// ------------------------------------------------------------ SB steps
//
// MaxCut (src/kernels.cu:236-262), step t with c = sched[t]:
//   |x| > 1:  x = sgn x, y = 0          else  y += c2 * F
//   x += c0 * y;  y += c1 * x           (every element)
// TSP (src/kernels.cu:564-596):
//   |x| > 1:  x = sgn x, y = 0 and nothing else
//   else:     y += c0 * (c2 * (c3 * h_i + 2 F));  x += c0 * y;  y += c1 * x
// with F = -sum_k J_k g(x_col_k) accumulated in ascending CSR order and
// g = sgn (dSB) or coupling * x (bSB).

// Plain form: X double-buffered (gather from the snapshot Xs, write Xn), Y in
// place.  One wavefront = one row x (64 V) agents; lane owns V contiguous
// agents, so own-row and neighbour-row accesses are 64 V-wide coalesced loads.
template <typename T, int PROB, int GATHER, int V, int NT>
__global__ __launch_bounds__(kBlock) void k_step_plain(StepArgs a) {
  const Ctrl *ctrl = a.ctrl;
  if (ctrl->done) return;
  const int64_t tiles = a.lda / (64 * V);
  const int64_t g = wave_index();
  const int64_t row = g / tiles;
  if (row >= a.N) return;
  const int lane = threadIdx.x & 63;
  const int64_t a0 = (g - row * tiles) * (64 * V) + lane * V;
  const int64_t t = ctrl->base + a.offset;
  const T *sc = (const T *)a.sched + 4 * t;
  const T c0 = sc[0], c1 = sc[1], c2 = sc[2], c3 = sc[3];
  const T coupling = (T)a.coupling;
  const T *__restrict__ Xs = (const T *)a.gsrc;
  const T *__restrict__ val = (const T *)a.val;
  const int32_t *__restrict__ col = a.col;
  T *Y = (T *)a.Y;
  const int64_t lda = a.lda;

  const int64_t k0 = a.rp[row], k1 = a.rp[row + 1];
  if (a.split > 0 && k1 - k0 > a.split) return;  // k_step_split owns the row
  const Vec<T, V> xo = ldvs<T, V, NT>(Xs + row * lda + a0);
  Vec<T, V> yo = ldvs<T, V, NT>(Y + row * lda + a0);
  T F[V];
#pragma unroll
  for (int u = 0; u < V; u++) F[u] = (T)0;
  int64_t k = k0;
  for (; k + 4 <= k1; k += 4) {
    const int64_t j0 = col[k], j1 = col[k + 1], j2 = col[k + 2], j3 = col[k + 3];
    const T v0 = val[k], v1 = val[k + 1], v2 = val[k + 2], v3 = val[k + 3];
    const Vec<T, V> g0 = ldv<T, V>(Xs + j0 * lda + a0);
    const Vec<T, V> g1 = ldv<T, V>(Xs + j1 * lda + a0);
    const Vec<T, V> g2 = ldv<T, V>(Xs + j2 * lda + a0);
    const Vec<T, V> g3 = ldv<T, V>(Xs + j3 * lda + a0);
#pragma unroll
    for (int u = 0; u < V; u++) {
      F[u] = sb_accum<T, GATHER>(F[u], v0, g0.v[u], coupling);
      F[u] = sb_accum<T, GATHER>(F[u], v1, g1.v[u], coupling);
      F[u] = sb_accum<T, GATHER>(F[u], v2, g2.v[u], coupling);
      F[u] = sb_accum<T, GATHER>(F[u], v3, g3.v[u], coupling);
    }
  }
  for (; k < k1; k++) {
    const int64_t j = col[k];
    const T v = val[k];
    const Vec<T, V> gj = ldv<T, V>(Xs + j * lda + a0);
#pragma unroll
    for (int u = 0; u < V; u++) F[u] = sb_accum<T, GATHER>(F[u], v, gj.v[u], coupling);
  }
  const T h = PROB == PROB_TSP ? ((const T *)a.field)[row] : (T)0;
  Vec<T, V> xn;
#pragma unroll
  for (int u = 0; u < V; u++) {
    T x = xo.v[u], y = yo.v[u];
    sb_update<T, PROB>(x, y, F[u], c0, c1, c2, c3, h);
    xn.v[u] = x;
    yo.v[u] = y;
  }
  stvs<T, V, NT>((T *)a.gdst + row * lda + a0, xn);
  stvs<T, V, NT>(Y + row * lda + a0, yo);
}

// BF16-shadow bSB form: X, Y FP32 in place; gathers read the BF16 shadow of
// the step's snapshot (gsrc) and the step writes the next shadow (gdst).
template <int PROB, int V, int NT>
__global__ __launch_bounds__(kBlock) void k_step_bf16(StepArgs a) {
  const Ctrl *ctrl = a.ctrl;
  if (ctrl->done) return;
  const int64_t tiles = a.lda / (64 * V);
  const int64_t g = wave_index();
  const int64_t row = g / tiles;
  if (row >= a.N) return;
  const int lane = threadIdx.x & 63;
  const int64_t a0 = (g - row * tiles) * (64 * V) + lane * V;
  const int64_t t = ctrl->base + a.offset;
  const float *sc = (const float *)a.sched + 4 * t;
  const float c0 = sc[0], c1 = sc[1], c2 = sc[2], c3 = sc[3];
  const float coupling = (float)a.coupling;
  const uint16_t *__restrict__ Sh = (const uint16_t *)a.gsrc;
  const float *__restrict__ val = (const float *)a.val;
  const int32_t *__restrict__ col = a.col;
  float *X = (float *)a.X, *Y = (float *)a.Y;
  const int64_t lda = a.lda;

  const int64_t k0 = a.rp[row], k1 = a.rp[row + 1];
  if (a.split > 0 && k1 - k0 > a.split) return;  // k_step_split owns the row
  Vec<float, V> xo = ldvs<float, V, NT>(X + row * lda + a0);
  Vec<float, V> yo = ldvs<float, V, NT>(Y + row * lda + a0);
  float F[V];
#pragma unroll
  for (int u = 0; u < V; u++) F[u] = 0.0f;
  int64_t k = k0;
  for (; k + 4 <= k1; k += 4) {
    const int64_t j0 = col[k], j1 = col[k + 1], j2 = col[k + 2], j3 = col[k + 3];
    const float v0 = val[k], v1 = val[k + 1], v2 = val[k + 2], v3 = val[k + 3];
    const Vec<uint16_t, V> g0 = ldv<uint16_t, V>(Sh + j0 * lda + a0);
    const Vec<uint16_t, V> g1 = ldv<uint16_t, V>(Sh + j1 * lda + a0);
    const Vec<uint16_t, V> g2 = ldv<uint16_t, V>(Sh + j2 * lda + a0);
    const Vec<uint16_t, V> g3 = ldv<uint16_t, V>(Sh + j3 * lda + a0);
#pragma unroll
    for (int u = 0; u < V; u++) {
      F[u] = sb_accum<float, G_BSB>(F[u], v0, d_from_bf16(g0.v[u]), coupling);
      F[u] = sb_accum<float, G_BSB>(F[u], v1, d_from_bf16(g1.v[u]), coupling);
      F[u] = sb_accum<float, G_BSB>(F[u], v2, d_from_bf16(g2.v[u]), coupling);
      F[u] = sb_accum<float, G_BSB>(F[u], v3, d_from_bf16(g3.v[u]), coupling);
    }
  }
  for (; k < k1; k++) {
    const int64_t j = col[k];
    const float v = val[k];
    const Vec<uint16_t, V> gj = ldv<uint16_t, V>(Sh + j * lda + a0);
#pragma unroll
    for (int u = 0; u < V; u++) F[u] = sb_accum<float, G_BSB>(F[u], v, d_from_bf16(gj.v[u]), coupling);
  }
  const float h = PROB == PROB_TSP ? ((const float *)a.field)[row] : 0.0f;
  Vec<uint16_t, V> sn;
#pragma unroll
  for (int u = 0; u < V; u++) {
    float x = xo.v[u], y = yo.v[u];
    sb_update<float, PROB>(x, y, F[u], c0, c1, c2, c3, h);
    xo.v[u] = x;
    yo.v[u] = y;
    sn.v[u] = d_bf16(x);
  }
  stvs<float, V, NT>(X + row * lda + a0, xo);
  stvs<float, V, NT>(Y + row * lda + a0, yo);
  stv<uint16_t, V>((uint16_t *)a.gdst + row * lda + a0, sn);
}

// Bit-packed dSB form: X, Y FP32 in place; the gather reads one sign bit per
// agent from the snapshot's sign words (gsrc, wave-uniform scalar loads) and
// the step writes the next sign words (gdst) with one ballot per word.  Lane
// l of word w owns agent 64 w + l.  F is the exact sum of +-J, so the
// trajectory is bit-identical to the plain FP32 dSB form.
// Wavefront tile: R consecutive rows x WT sign words (64 WT agents).  The
// X / Y loads of all R rows are issued before the gathers.
template <int PROB, int WT, int R, int NT>
__global__ __launch_bounds__(kBlock) void k_step_packed(StepArgs a) {
  const Ctrl *ctrl = a.ctrl;
  if (ctrl->done) return;
  const int64_t W = a.W;
  const int64_t tiles = W / WT;
  const int64_t g = wave_index();
  const int64_t row0 = (g / tiles) * R;
  if (row0 >= a.N) return;
  const int lane = threadIdx.x & 63;
  const int64_t w0 = (g - (g / tiles) * tiles) * WT;
  const int64_t t = ctrl->base + a.offset;
  const float *sc = (const float *)a.sched + 4 * t;
  const float c0 = sc[0], c1 = sc[1], c2 = sc[2], c3 = sc[3];
  const uint64_t *__restrict__ S = (const uint64_t *)a.gsrc;
  const float *__restrict__ val = (const float *)a.val;
  const int32_t *__restrict__ col = a.col;
  float *X = (float *)a.X, *Y = (float *)a.Y;
  const int64_t lda = a.lda;
  uint64_t *Sn = (uint64_t *)a.gdst;

  float x[R][WT], y[R][WT];
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (row0 + r < a.N) {
#pragma unroll
      for (int u = 0; u < WT; u++) {
        x[r][u] = ld_state<NT>(X + (row0 + r) * lda + 64 * (w0 + u) + lane);
        y[r][u] = ld_state<NT>(Y + (row0 + r) * lda + 64 * (w0 + u) + lane);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int64_t row = row0 + r;
    if (row >= a.N) break;
    const int64_t k0 = a.rp[row], k1 = a.rp[row + 1];
    if (a.split > 0 && k1 - k0 > a.split) continue;  // k_step_split owns the row
    float F[WT];
#pragma unroll
    for (int u = 0; u < WT; u++) F[u] = 0.0f;
    int64_t k = k0;
    for (; k + 4 <= k1; k += 4) {
      const int64_t j0 = col[k], j1 = col[k + 1], j2 = col[k + 2], j3 = col[k + 3];
      const float v0 = val[k], v1 = val[k + 1], v2 = val[k + 2], v3 = val[k + 3];
      uint64_t s0[WT], s1[WT], s2[WT], s3[WT];
#pragma unroll
      for (int u = 0; u < WT; u++) {
        s0[u] = S[j0 * W + w0 + u];
        s1[u] = S[j1 * W + w0 + u];
        s2[u] = S[j2 * W + w0 + u];
        s3[u] = S[j3 * W + w0 + u];
      }
#pragma unroll
      for (int u = 0; u < WT; u++) {
        F[u] = ((s0[u] >> lane) & 1) ? F[u] + v0 : F[u] - v0;
        F[u] = ((s1[u] >> lane) & 1) ? F[u] + v1 : F[u] - v1;
        F[u] = ((s2[u] >> lane) & 1) ? F[u] + v2 : F[u] - v2;
        F[u] = ((s3[u] >> lane) & 1) ? F[u] + v3 : F[u] - v3;
      }
    }
    for (; k < k1; k++) {
      const int64_t j = col[k];
      const float v = val[k];
#pragma unroll
      for (int u = 0; u < WT; u++) {
        const uint64_t s = S[j * W + w0 + u];
        F[u] = ((s >> lane) & 1) ? F[u] + v : F[u] - v;
      }
    }
    const float h = PROB == PROB_TSP ? ((const float *)a.field)[row] : 0.0f;
    uint64_t mine = 0;
#pragma unroll
    for (int u = 0; u < WT; u++) {
      sb_update<float, PROB>(x[r][u], y[r][u], F[u], c0, c1, c2, c3, h);
      st_state<NT>(X + row * lda + 64 * (w0 + u) + lane, x[r][u]);
      st_state<NT>(Y + row * lda + 64 * (w0 + u) + lane, y[r][u]);
      const uint64_t b = __ballot(signbit(x[r][u]));
      if (lane == u) mine = b;
    }
    // lanes 0 .. WT-1 write the row's WT contiguous sign words at once
    if (lane < WT) Sn[row * W + w0 + lane] = mine;
  }
}

// Narrow-state dSB forms (MaxCut, fused): X and Y as QW-bit codes, QW = 12,
// 10 or 8 (see sis_device.hpp), NP = 3, 3 or 2 planes of N x Q x 64 lanes
// (Q = lda / 256 quads; the 10-bit form's third plane holds 16-bit values);
// lane l of quad q owns agents 256 q + 64 j + l,
// j = 0..3, i.e. bit l of sign word 4 q + j.  The gather is the packed form's (exact sum of +-J
// from the snapshot's sign words, in CSR order); the update is the FP32 dSB
// update on the decoded values, rounded back to codes.  A wavefront owns R
// rows of one quad; their state loads are issued before the gathers.  Bytes
// per node-agent: QW / 4 read + QW / 4 written + 1/8 sign write + degree / 8
// gathers.
//
// Staged gathers (VG = 1): the R rows' entries (at most kQ12Stage) are
// fetched at once -- lane e loads column, weight and the four sign words of
// entry e with vector loads -- into the wavefront's LDS slice; each row then
// adds its entries in CSR order from LDS broadcasts.  One dependent chain of
// memory round trips per R rows instead of one per row: the scalar path is
// bound by that latency, not by HBM.  The staged words carry the weight's
// sign (complemented where v < 0), so each term is +-|v|: the same value as
// +-v, and the sum never meets a signed zero (it starts at +0 and x + (-x) is
// +0).  Where every staged weight is +-1 the sums are integers, counted from
// the bits (two bit extracts and one three-way add per two entries and
// agent) instead of three operations per term.  Waves whose rows hold more entries take
// the scalar path.
constexpr int kQ12Stage = 128;

// Plane p of this lane at row offset ro (wave-uniform, in lanes), rb = P + ro:
// dword planes at rb + p ps, a 16-bit last plane at value ro of its base.
template <int QW, int NT>
__device__ __forceinline__ uint32_t qn_ld_row(const uint32_t *rb, int64_t ro, int64_t ps, int p,
                                              int lane) {
  if (QFmt<QW>::H16 && p == QFmt<QW>::NP - 1)
    return ld_state<NT>(reinterpret_cast<const uint16_t *>(rb - ro + p * ps) + ro + lane);
  return ld_state<NT>(rb + p * ps + lane);
}
template <int QW, int NT>
__device__ __forceinline__ void qn_st_row(uint32_t *rb, int64_t ro, int64_t ps, int p, int lane,
                                          uint32_t v) {
  if (QFmt<QW>::H16 && p == QFmt<QW>::NP - 1)
    st_state<NT>(reinterpret_cast<uint16_t *>(rb - ro + p * ps) + ro + lane, (uint16_t)v);
  else
    st_state<NT>(rb + p * ps + lane, v);
}

// New codes and sign words of one row: decode, fused dSB update, encode,
// one ballot per sign word.  rb: the row's plane-0 dwords of this quad
// (wave-uniform), ro: their offset from the planes' base, ps: plane stride,
// sw: the row's four sign words; SR: Y rounded stochastically with the
// dither of the row key (wave-uniform) and the lane.
template <int QW, int NT, int SR>
__device__ __forceinline__ void qn_finish_row(const uint32_t d[QFmt<QW>::NP], const float F[4],
                                              const Q12K &k, uint32_t *rb, int64_t ro, int64_t ps,
                                              uint64_t *sw, int lane, uint32_t key) {
  constexpr int NP = QFmt<QW>::NP;
  uint32_t un[4];
  uint64_t b[4];
  uint32_t h = 0;
  if constexpr (SR) h = q8s_lane_hash(key, lane);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    float x, y;
    qn_decode_at<QW>(d, j, k, x, y);
    sb_update_maxcut_sel(x, y, F[j], k.c0, k.c1, k.c2);
    bool neg;
    if constexpr (SR) {
      un[j] = qn_encode_sr<QW>(x, y, k, neg, (h >> (8 * j)) & 0xffu);
    } else {
      un[j] = qn_encode<QW, false>(x, y, k, neg);  // no Y clamp after a fused step
    }
    b[j] = __ballot(neg);
  }
  uint32_t e[NP];
  qn_pack<QW>(un, e);
#pragma unroll
  for (int p = 0; p < NP; p++) qn_st_row<QW, NT>(rb, ro, ps, p, lane, e[p]);
  // lane 0 writes the four (wave-uniform) sign words, 32 aligned bytes
  if (lane == 0) {
    uint4 *o = reinterpret_cast<uint4 *>(sw);
    o[0] = make_uint4((uint32_t)b[0], (uint32_t)(b[0] >> 32), (uint32_t)b[1], (uint32_t)(b[1] >> 32));
    o[1] = make_uint4((uint32_t)b[2], (uint32_t)(b[2] >> 32), (uint32_t)b[3], (uint32_t)(b[3] >> 32));
  }
}

// Rows whose wavefront holds no entries at all: F = 0, the sum of no terms
// (the bit-counting path gives fma(0, 2, -0) = +0 too).
// Stochastic-rounding key of row r of the wavefront (0 unless SR): its node
// comes from lane r of nodel, loaded once for the wavefront's rows.
template <int SR>
__device__ __forceinline__ uint32_t qn_row_key(int nodel, int r, int64_t Q, int64_t q,
                                               uint32_t salt) {
  if constexpr (SR) return q8s_row_key(__builtin_amdgcn_readlane(nodel, r), Q, q, salt);
  return 0u;
}

template <int QW, int NT, int SR, int R>
__device__ __forceinline__ void qn_finish_empty(const uint32_t (&d)[R][QFmt<QW>::NP],
                                                const Q12K &k, uint32_t *P, int64_t Q, int64_t q,
                                                int64_t ps, uint64_t *Sn, int64_t W, int64_t w0,
                                                int64_t row0, int rl, int lane, int nodel,
                                                uint32_t salt) {
  const float F[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (r >= rl) continue;
    const int64_t row = row0 + r;
    qn_finish_row<QW, NT, SR>(d[r], F, k, P + (row * Q + q) * 64, (row * Q + q) * 64, ps,
                              Sn + row * W + w0, lane, qn_row_key<SR>(nodel, r, Q, q, salt));
  }
}

template <int QW, int R, int NT, int VG, int SR>
__device__ __forceinline__ void qn_step_body(const StepArgs &a) {
  constexpr int NP = QFmt<QW>::NP;
  const Ctrl *ctrl = a.ctrl;
  if (ctrl->done) return;
  const int64_t W = a.W, Q = W / 4;
  const int64_t g = wave_index();
  const int64_t row0 = (g / Q) * R;
  if (row0 >= a.N) return;
  const int lane = threadIdx.x & 63;
  const int64_t q = g - (g / Q) * Q;
  const int64_t w0 = 4 * q;
  const int64_t t = ctrl->base + a.offset;
  const float *sc = (const float *)a.sched + 4 * t;
  const Q12K kc = qn_consts<QW>(sc[0], sc[1], sc[2], a.qshift);
  const uint32_t salt = SR ? q8s_salt(t) : 0u;
  const uint64_t *__restrict__ S = (const uint64_t *)a.gsrc;
  const float *__restrict__ val = (const float *)a.val;
  const int32_t *__restrict__ col = a.col;
  const int64_t *__restrict__ rp = a.rp;
  uint32_t *P = (uint32_t *)a.X;
  const int64_t ps = a.N * Q * 64;  // plane stride (dwords)
  uint64_t *Sn = (uint64_t *)a.gdst;

  // Memory waits count loads in issue order, so a wait for one load also
  // waits for every load issued before it.  The staged path's chain of
  // dependent loads (row pointers, then columns and weights, then sign words)
  // therefore starts with the row pointers, before the state loads, and each
  // step of it issues all its loads at once, without branches.
  // Row pointers rp[row0 .. row0 + rl] in lanes 0..rl, loaded before any
  // store so they stay in registers; their offsets from the first come back
  // one row at a time by readlane.
  const int rl = (int)(a.N - row0 < R ? a.N - row0 : R);
  int64_t rpl = 0;
  if constexpr (VG) rpl = rp[row0 + (lane < rl ? lane : rl)];
  // SR: the node of row r in lane r (the dither is keyed by node), loaded
  // with the row pointers rather than once per row
  int nodel = 0;
  if constexpr (SR) {
    const int64_t rr = row0 + (lane < rl ? lane : rl - 1);
    nodel = a.rowid ? a.rowid[rr] : (int)rr;
  }
  uint32_t d[R][NP];
#pragma unroll
  for (int r = 0; r < R; r++) {
    // every row loads (one past the last reloads the last and is never
    // stored), so that no branch hides the loads from the waits after them
    const int64_t rr = row0 + r < a.N ? row0 + r : a.N - 1;
    const int64_t ro = (rr * Q + q) * 64;  // wave-uniform
#pragma unroll
    for (int p = 0; p < NP; p++) d[r][p] = qn_ld_row<QW, NT>(P + ro, ro, ps, p, lane);
  }
  if constexpr (VG) {
    const int64_t kb = (int64_t)readlane64((uint64_t)rpl, 0);
    const int64_t ne = (int64_t)readlane64((uint64_t)rpl, rl) - kb;
    if (ne <= kQ12Stage) {
      const int off = (int)(rpl - kb);
      __shared__ uint4 lsig[kWaves][kQ12Stage][2];
      __shared__ float lval[kWaves][kQ12Stage];
      const int wv = threadIdx.x >> 6, half = lane >> 5, bit = lane & 31, sh = 31 - bit;
      // Pass 1: the columns and weights of this lane's entries (lane, lane + 64).
      // A lane past the last entry reloads entry 0 of the range (its values
      // are never read back); a wavefront without entries loads nothing.
      constexpr int kSlots = kQ12Stage / 64;
      if (ne == 0) {
        qn_finish_empty<QW, NT, SR>(d, kc, P, Q, q, ps, Sn, W, w0, row0, rl, lane, nodel, salt);
        return;
      }
      int64_t ce[kSlots];
      float ve[kSlots];
#pragma unroll
      for (int u = 0; u < kSlots; u++) {
        const int64_t e = lane + 64 * u;
        const int64_t ei = kb + (e < ne ? e : 0);
        ce[u] = col[ei];
        ve[u] = val[ei];
      }
      bool unit = true;
#pragma unroll
      for (int u = 0; u < kSlots; u++)
        unit = unit && (lane + 64 * u >= ne || __builtin_fabsf(ve[u]) == 1.0f);
      const bool all_unit = __all(unit);
      // Pass 2: the four sign words of each entry, complemented where the
      // weight is negative so that a set bit means +|v|.  Lanes 0-31 use the
      // words' low dwords L0..L3, lanes 32-63 the high dwords.  With real
      // weights, entry e holds L0..L3 of each half (a lane reads the 16 bytes
      // holding its bits) and lval[e] = -|v|.  With +-1 weights, entry e
      // holds the byte transpose T0..T3 of each half's L0..L3:
      // byte j of Tb = byte b of Lj, so lane l's four bits (bit l mod 32 of
      // L0..L3) sit at bits (l mod 8) + 8 j of the one dword T(l mod 32 / 8).
      uint4 s01v[kSlots], s23v[kSlots];  // words 0,1 and 2,3 as dword pairs
#pragma unroll
      for (int u = 0; u < kSlots; u++) {
        const uint4 *sp = reinterpret_cast<const uint4 *>(S + ce[u] * W + w0);
        s01v[u] = sp[0];
        s23v[u] = sp[1];
      }
      // every lane writes its slots (e < kQ12Stage); those at or past ne are
      // never read
#pragma unroll
      for (int u = 0; u < kSlots; u++) {
        const int64_t e = lane + 64 * u;
        const uint32_t f = ve[u] < 0.0f ? 0xffffffffu : 0u;
        const uint4 s01 = s01v[u], s23 = s23v[u];
        const uint4 lo = make_uint4(s01.x ^ f, s01.z ^ f, s23.x ^ f, s23.z ^ f);
        const uint4 hi = make_uint4(s01.y ^ f, s01.w ^ f, s23.y ^ f, s23.w ^ f);
        if (all_unit) {
          auto transpose = [](const uint4 &L) {
            // 2 x 2 byte blocks, then pairs of half-words
            const uint32_t A = __builtin_amdgcn_perm(L.y, L.x, 0x06020400u);  // L0b0 L1b0 L0b2 L1b2
            const uint32_t B = __builtin_amdgcn_perm(L.y, L.x, 0x07030501u);  // L0b1 L1b1 L0b3 L1b3
            const uint32_t C = __builtin_amdgcn_perm(L.w, L.z, 0x06020400u);  // L2b0 L3b0 L2b2 L3b2
            const uint32_t D = __builtin_amdgcn_perm(L.w, L.z, 0x07030501u);  // L2b1 L3b1 L2b3 L3b3
            return make_uint4(__builtin_amdgcn_perm(C, A, 0x05040100u),   // b0 of L0..L3
                              __builtin_amdgcn_perm(D, B, 0x05040100u),   // b1
                              __builtin_amdgcn_perm(C, A, 0x07060302u),   // b2
                              __builtin_amdgcn_perm(D, B, 0x07060302u));  // b3
          };
          lsig[wv][e][0] = transpose(lo);
          lsig[wv][e][1] = transpose(hi);
        } else {
          lsig[wv][e][0] = lo;
          lsig[wv][e][1] = hi;
          lval[wv][e] = -__builtin_fabsf(ve[u]);
        }
      }
      // the wavefront's LDS writes complete before its reads below
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (all_unit) {
        // weights +-1: F = (entries with the bit set) - (the others), an
        // integer, as the FP32 sum of +-1 terms is in any order.  The four
        // counts accumulate as bytes of one dword (a row has at most
        // kQ12Stage < 256 entries here): one shift, mask and add per entry.
        const uint32_t *lt = reinterpret_cast<const uint32_t *>(&lsig[wv][0][half]) + (bit >> 3);
        const int bs = bit & 7;
#pragma unroll
        for (int r = 0; r < R; r++) {
          if (r >= rl) continue;  // (not break: a constant trip count unrolls)
          const int k0 = __builtin_amdgcn_readlane(off, r);
          const int k1 = __builtin_amdgcn_readlane(off, r + 1);
          uint32_t acc = 0u;
          int k = k0;
          for (; k + 2 <= k1; k += 2)
            acc += ((lt[8 * k] >> bs) & 0x01010101u) + ((lt[8 * k + 8] >> bs) & 0x01010101u);
          if (k < k1) acc += (lt[8 * k] >> bs) & 0x01010101u;
          const float md = -(float)(k1 - k0);
          float F[4];
#pragma unroll
          for (int j = 0; j < 4; j++) F[j] = __builtin_fmaf((float)((acc >> (8 * j)) & 0xffu), 2.0f, md);
          const int64_t row = row0 + r;
          qn_finish_row<QW, NT, SR>(d[r], F, kc, P + (row * Q + q) * 64, (row * Q + q) * 64, ps,
                                Sn + row * W + w0, lane,
                                qn_row_key<SR>(nodel, r, Q, q, salt));
        }
        return;
      }
      // F += bit ? |v| : -|v|, as -|v| with its sign flipped where the bit is set
      auto add = [&](float F[4], const uint4 &d, float nv) {
        const uint32_t b = __float_as_uint(nv);
        F[0] += __uint_as_float(b ^ ((d.x << sh) & 0x80000000u));
        F[1] += __uint_as_float(b ^ ((d.y << sh) & 0x80000000u));
        F[2] += __uint_as_float(b ^ ((d.z << sh) & 0x80000000u));
        F[3] += __uint_as_float(b ^ ((d.w << sh) & 0x80000000u));
      };
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (r >= rl) continue;  // (not break: a constant trip count unrolls)
        const int k0 = __builtin_amdgcn_readlane(off, r);
        const int k1 = __builtin_amdgcn_readlane(off, r + 1);
        float F[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        int k = k0;
        for (; k + 2 <= k1; k += 2) {
          const uint4 d0 = lsig[wv][k][half], d1 = lsig[wv][k + 1][half];
          const float v0 = lval[wv][k], v1 = lval[wv][k + 1];
          add(F, d0, v0);
          add(F, d1, v1);
        }
        if (k < k1) add(F, lsig[wv][k][half], lval[wv][k]);
        const int64_t row = row0 + r;
        qn_finish_row<QW, NT, SR>(d[r], F, kc, P + (row * Q + q) * 64, (row * Q + q) * 64, ps,
                              Sn + row * W + w0, lane,
                              qn_row_key<SR>(nodel, r, Q, q, salt));
      }
      return;
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int64_t row = row0 + r;
    if (row >= a.N) continue;
    const int64_t k0 = rp[row], k1 = rp[row + 1];
    float F[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int64_t k = k0;
    for (; k + 4 <= k1; k += 4) {
      const int64_t j0 = col[k], j1 = col[k + 1], j2 = col[k + 2], j3 = col[k + 3];
      const float v0 = val[k], v1 = val[k + 1], v2 = val[k + 2], v3 = val[k + 3];
      uint64_t s0[4], s1[4], s2[4], s3[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        s0[u] = S[j0 * W + w0 + u];
        s1[u] = S[j1 * W + w0 + u];
        s2[u] = S[j2 * W + w0 + u];
        s3[u] = S[j3 * W + w0 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        F[u] = ((s0[u] >> lane) & 1) ? F[u] + v0 : F[u] - v0;
        F[u] = ((s1[u] >> lane) & 1) ? F[u] + v1 : F[u] - v1;
        F[u] = ((s2[u] >> lane) & 1) ? F[u] + v2 : F[u] - v2;
        F[u] = ((s3[u] >> lane) & 1) ? F[u] + v3 : F[u] - v3;
      }
    }
    for (; k < k1; k++) {
      const int64_t j = col[k];
      const float v = val[k];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint64_t sw = S[j * W + w0 + u];
        F[u] = ((sw >> lane) & 1) ? F[u] + v : F[u] - v;
      }
    }
    qn_finish_row<QW, NT, SR>(d[r], F, kc, P + (row * Q + q) * 64, (row * Q + q) * 64, ps,
                          Sn + row * W + w0, lane,
                          qn_row_key<SR>(nodel, r, Q, q, salt));
  }
}

template <int QW, int R, int NT, int VG, int SR>
__global__ __launch_bounds__(kBlock) void k_step_qn(StepArgs a) {
  qn_step_body<QW, R, NT, VG, SR>(a);
}

// Initial state of the narrow forms: the counter-based stream of
// k_init_state (element e = i A + a, FP32 values), rounded to codes; padding
// agents 0.
template <int QW>
__global__ void k_init_qn(uint32_t *P, int64_t N, int64_t A, int64_t Q, uint64_t seed, int qshift,
                          const int32_t *rowid) {
  constexpr int NP = QFmt<QW>::NP;
  const int64_t n = N * Q * 64, ps = n;
  const Q12K k = qn_consts<QW>(0.0f, 0.0f, 0.0f, qshift);
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / (Q * 64), rem = idx - i * Q * 64, q = rem / 64, l = rem - q * 64;
    const int64_t node = rowid ? (int64_t)rowid[i] : i;
    uint32_t u[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int64_t ag = 256 * q + 64 * j + l;
      float x = 0.0f, y = 0.0f;
      if (ag < A) {
        const uint64_t e = (uint64_t)(node * A + ag);
        x = (float)((d_uniform(seed, 2 * e) * 2 - 1) / 10);
        y = (float)((d_uniform(seed, 2 * e + 1) * 2 - 1) / 10);
      }
      u[j] = qn_encode<QW>(x, y, k);
    }
    uint32_t d[NP];
    qn_pack<QW>(u, d);
#pragma unroll
    for (int p = 0; p < NP; p++) qn_plane_st<QW>(P, ps, p, idx, d[p]);
  }
}

// Pre-step (src/kernels.cu:15-20 at t = 0) on the decoded values, rounded.
template <int QW>
__global__ void k_prestep_qn(uint32_t *P, int64_t n, float cy, float cx, int qshift) {
  constexpr int NP = QFmt<QW>::NP;
  const Q12K k = qn_consts<QW>(0.0f, 0.0f, 0.0f, qshift);
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    uint32_t d[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) d[p] = qn_plane_ld<QW>(P, n, p, idx);
    float x[4], y[4];
    qn_unpack<QW>(d, x, y, k);
    uint32_t u[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float yy = fmaT(cy, x[j], y[j]);
      const float xx = fmaT(cx, yy, x[j]);
      u[j] = qn_encode<QW>(xx, yy, k);
    }
    qn_pack<QW>(u, d);
#pragma unroll
    for (int p = 0; p < NP; p++) qn_plane_st<QW>(P, n, p, idx, d[p]);
  }
}

// Sign words of the narrow state: one wavefront per (row, quad), one ballot
// per word (bit l of word 4 q + j = X code of agent 256 q + 64 j + l < 0).
template <int QW>
__global__ __launch_bounds__(kBlock) void k_pack_signs_qn(const uint32_t *P, uint64_t *S,
                                                          int64_t N, int64_t Q) {
  constexpr int NP = QFmt<QW>::NP;
  const int lane = threadIdx.x & 63;
  const int64_t g = wave_index();
  if (g >= N * Q) return;
  const int64_t ps = N * Q * 64, o = g * 64 + lane;
  uint32_t d[NP], u[4];
#pragma unroll
  for (int p = 0; p < NP; p++) d[p] = qn_plane_ld<QW>(P, ps, p, o);
  qn_units<QW>(d, u);
  uint64_t mine = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const uint64_t b = __ballot(qn_negative<QW>(u[j]));
    if (lane == j) mine = b;
  }
  if (lane < 4) S[g * 4 + lane] = mine;
}

// ------------------------------------------------------------ row split
//
// FORM as in launch_step: 0 plain FP32, 1 plain FP64, 2 packed signs, 3 BF16
// shadow.  Lane l of a workgroup owns agent 64 tile + l; every wavefront
// gathers with the same per-entry arithmetic as the row kernels above.
template <int FORM> struct SplitT { typedef float real; typedef float item; };
template <> struct SplitT<1> { typedef double real; typedef double item; };
template <> struct SplitT<2> { typedef float real; typedef uint64_t item; };
template <> struct SplitT<3> { typedef float real; typedef uint16_t item; };

template <int FORM>
__device__ __forceinline__ typename SplitT<FORM>::item split_load(const StepArgs &a, int64_t j,
                                                                  int64_t tile, int lane) {
  typedef typename SplitT<FORM>::item I;
  if constexpr (FORM == 2) return ((const I *)a.gsrc)[j * a.W + tile];  // wave-uniform word
  else return ((const I *)a.gsrc)[j * a.lda + tile * 64 + lane];
}

template <int FORM, int GATHER>
__device__ __forceinline__ typename SplitT<FORM>::real split_acc(typename SplitT<FORM>::real F,
                                                                 typename SplitT<FORM>::real v,
                                                                 typename SplitT<FORM>::item g,
                                                                 typename SplitT<FORM>::real c,
                                                                 int lane) {
  if constexpr (FORM == 2) return ((g >> lane) & 1) ? F + v : F - v;
  else if constexpr (FORM == 3) return sb_accum<float, G_BSB>(F, v, d_from_bf16(g), c);
  else return sb_accum<typename SplitT<FORM>::real, GATHER>(F, v, g, c);
}

// F over entries [kb, ke) of one row, in CSR order.
template <int FORM, int GATHER>
__device__ __forceinline__ typename SplitT<FORM>::real split_chunk(const StepArgs &a, int64_t kb,
                                                                   int64_t ke, int64_t tile,
                                                                   int lane) {
  typedef typename SplitT<FORM>::real T;
  typedef typename SplitT<FORM>::item I;
  const T *__restrict__ val = (const T *)a.val;
  const int32_t *__restrict__ col = a.col;
  const T c = (T)a.coupling;
  T F = (T)0;
  if constexpr (FORM == 2) {
    // 64 entries per pass on the vector memory path: lane l loads entry
    // kk + l's column, weight and sign word; the words are then taken in
    // CSR order with readlane.
    const uint64_t *__restrict__ S = (const uint64_t *)a.gsrc;
    for (int64_t kk = kb; kk < ke; kk += 64) {
      const int64_t k = kk + lane;
      const bool ok = k < ke;
      const int64_t j = ok ? (int64_t)col[k] : 0;
      const float v = ok ? val[k] : 0.0f;
      const uint64_t sw = ok ? S[j * a.W + tile] : 0;
      const int cnt = ke - kk < 64 ? (int)(ke - kk) : 64;
      for (int e = 0; e < cnt; e++) {
        const uint64_t we = readlane64(sw, e);
        const float ve = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), e));
        F = ((we >> lane) & 1) ? F + ve : F - ve;
      }
    }
    return F;
  }
  int64_t k = kb;
  for (; k + 4 <= ke; k += 4) {
    const T v0 = val[k], v1 = val[k + 1], v2 = val[k + 2], v3 = val[k + 3];
    const I g0 = split_load<FORM>(a, col[k], tile, lane);
    const I g1 = split_load<FORM>(a, col[k + 1], tile, lane);
    const I g2 = split_load<FORM>(a, col[k + 2], tile, lane);
    const I g3 = split_load<FORM>(a, col[k + 3], tile, lane);
    F = split_acc<FORM, GATHER>(F, v0, g0, c, lane);
    F = split_acc<FORM, GATHER>(F, v1, g1, c, lane);
    F = split_acc<FORM, GATHER>(F, v2, g2, c, lane);
    F = split_acc<FORM, GATHER>(F, v3, g3, c, lane);
  }
  for (; k < ke; k++) F = split_acc<FORM, GATHER>(F, val[k], split_load<FORM>(a, col[k], tile, lane), c, lane);
  return F;
}

// The element update and stores of one row for one 64-agent tile, as the
// row kernels above do them.  Needs the whole wavefront (packed: ballot).
template <int FORM, int PROB>
__device__ __forceinline__ void split_finish(const StepArgs &a, int64_t row, int64_t tile, int lane,
                                             typename SplitT<FORM>::real F) {
  typedef typename SplitT<FORM>::real T;
  const int64_t i = row * a.lda + tile * 64 + lane;
  const int64_t t = a.ctrl->base + a.offset;
  const T *sc = (const T *)a.sched + 4 * t;
  const T h = PROB == PROB_TSP ? ((const T *)a.field)[row] : (T)0;
  T *Y = (T *)a.Y;
  if constexpr (FORM == 0 || FORM == 1) {
    T x = ((const T *)a.gsrc)[i], y = Y[i];
    sb_update<T, PROB>(x, y, F, sc[0], sc[1], sc[2], sc[3], h);
    ((T *)a.gdst)[i] = x;
    Y[i] = y;
  } else {
    float *X = (float *)a.X;
    float x = X[i], y = Y[i];
    sb_update<float, PROB>(x, y, F, sc[0], sc[1], sc[2], sc[3], h);
    X[i] = x;
    Y[i] = y;
    if constexpr (FORM == 3) {
      ((uint16_t *)a.gdst)[i] = d_bf16(x);
    } else {
      const uint64_t b = __ballot(signbit(x));
      if (lane == 0) ((uint64_t *)a.gdst)[row * a.W + tile] = b;
    }
  }
}

template <int FORM, int PROB, int GATHER>
__global__ __launch_bounds__(kBlock) void k_step_split(StepArgs a, SplitArgs sp) {
  typedef typename SplitT<FORM>::real T;
  __shared__ T sF[kWaves][64];
  if (a.ctrl->done) return;
  const int64_t tiles = a.W;
  const int64_t b = blockIdx.x / tiles;
  const int64_t tile = blockIdx.x - b * tiles;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const SplitEnt e = sp.ent[b];
  const int64_t kb = e.kb + (int64_t)w * sp.chunk;
  const int64_t ke = kb + sp.chunk < e.ke ? kb + sp.chunk : e.ke;
  sF[w][lane] = kb < ke ? split_chunk<FORM, GATHER>(a, kb, ke, tile, lane) : (T)0;
  __syncthreads();
  if (w != 0) return;
  // chunk sums in chunk order (empty chunks skipped)
  T G = sF[0][lane];
  for (int q = 1; q < kWaves; q++)
    if (e.kb + (int64_t)q * sp.chunk < e.ke) G = G + sF[q][lane];
  if (e.ng > 1) {
    T *gs = (T *)sp.part;
    const int64_t a0 = tile * 64 + lane;
    gs[(e.pbase + e.g) * a.lda + a0] = G;
    __threadfence();  // the group sum is visible device-wide before the arrival
    int last = 0;
    if (lane == 0) last = atomicAdd(&sp.cnt[(int64_t)e.hr * tiles + tile], 1) == e.ng - 1;
    last = __builtin_amdgcn_readfirstlane(last);
    if (!last) return;
    __threadfence();  // every group sum of the row is visible to this workgroup
    G = gs[e.pbase * a.lda + a0];
    for (int32_t g = 1; g < e.ng; g++) G = G + gs[(e.pbase + g) * a.lda + a0];
    if (lane == 0) atomicExch(&sp.cnt[(int64_t)e.hr * tiles + tile], 0);
  }
  split_finish<FORM, PROB>(a, e.row, tile, lane, G);
}

// ------------------------------------------------------- structured TSP
//
// (J g)(i,k) = C/4 colsum_k + B/4 rowsum_i + sum_l Dq_kl (g(i-1,l) + g(i+1,l)),
// from the one-hot QUBO's entries (src/tsp_gpu.cu:85-106): same city at any
// other position C/4, same position any other city B/4, itself (B+C)/4, any
// other city at the two cyclically adjacent positions A d_kl / 8.  The
// (B+C)/4 diagonal cancels the C/4 and B/4 self terms that colsum and rowsum
// include; d_kk = 0.  Needs n >= 3 (two distinct adjacent positions).

template <typename T, int GATHER>
__device__ __forceinline__ T tsp_g(T x, T coupling) {
  if (GATHER == G_DSB) return signbit(x) ? (T)-1 : (T)1;
  return x * coupling;
}

// Row sums (position i) and column sums (city k) of g over the snapshot: one
// wavefront per (sum, 64-agent tile), terms in index order, loads issued 16
// at a time (offsets below n^2 lda <= 2^34 elements stay 64-bit).
template <typename T, int GATHER>
__global__ __launch_bounds__(kBlock) void k_tsp_sums(StepArgs a, TspArgs ts) {
  if (a.ctrl->done) return;
  const int64_t n = ts.n, lda = a.lda, tiles = lda / 64;
  const int64_t g = wave_index();
  if (g >= 2 * n * tiles) return;
  const int lane = threadIdx.x & 63;
  const int64_t which = g / (n * tiles), rem = g - which * n * tiles;
  const int64_t idx = rem / tiles, a0 = (rem - idx * tiles) * 64 + lane;
  const int64_t base = which == 0 ? idx * n : idx, stride = which == 0 ? 1 : n;
  const T *__restrict__ Xs = (const T *)a.gsrc + base * lda + a0;
  const int64_t step = stride * lda;
  const T c = (T)a.coupling;
  constexpr int B = 16;
  T acc = (T)0;
  int64_t m = 0;
  for (; m + B <= n; m += B) {
    T xv[B];
#pragma unroll
    for (int u = 0; u < B; u++) xv[u] = Xs[(m + u) * step];
#pragma unroll
    for (int u = 0; u < B; u++) acc = acc + tsp_g<T, GATHER>(xv[u], c);
  }
  for (; m < n; m++) acc = acc + tsp_g<T, GATHER>(Xs[m * step], c);
  ((T *)(which == 0 ? ts.rsum : ts.csum))[idx * lda + a0] = acc;
}

// One workgroup per (position i, 64-agent tile): v_l = g(i-1,l) + g(i+1,l) is
// staged in LDS (n x 64), then wavefront w updates cities w, w + 4, ... with
// F = -(C/4 colsum_k + B/4 rowsum_i + sum_l Dq_kl v_l) and the TSP element
// update of the row kernels.
template <typename T, int GATHER>
__global__ __launch_bounds__(kBlock) void k_tsp_step_struct(StepArgs a, TspArgs ts) {
  extern __shared__ __align__(16) unsigned char s_raw[];
  T *v = reinterpret_cast<T *>(s_raw);
  if (a.ctrl->done) return;
  const int64_t n = ts.n, lda = a.lda, tiles = lda / 64;
  const int64_t i = blockIdx.x / tiles, tile = blockIdx.x - i * tiles;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t a0 = tile * 64 + lane;
  const T *__restrict__ Xs = (const T *)a.gsrc;
  const T c = (T)a.coupling;
  const int64_t im = (i + n - 1) % n, ip = (i + 1) % n;
  // v_l = g(i-1, l) + g(i+1, l), SU rows per batch of loads
  constexpr int SU = 8;
  for (int64_t l0 = w; l0 < n; l0 += (int64_t)kWaves * SU) {
    T xm[SU], xp[SU];
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int64_t l = l0 + (int64_t)u * kWaves;
      const int64_t lc = l < n ? l : n - 1;
      xm[u] = Xs[(im * n + lc) * lda + a0];
      xp[u] = Xs[(ip * n + lc) * lda + a0];
    }
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int64_t l = l0 + (int64_t)u * kWaves;
      if (l < n) v[l * 64 + lane] = tsp_g<T, GATHER>(xm[u], c) + tsp_g<T, GATHER>(xp[u], c);
    }
  }
  __syncthreads();
  const int64_t t = a.ctrl->base + a.offset;
  const T *sc = (const T *)a.sched + 4 * t;
  const T c0 = sc[0], c1 = sc[1], c2 = sc[2], c3 = sc[3];
  const T qB = (T)ts.qB, qC = (T)ts.qC;
  const T rs = ((const T *)ts.rsum)[i * lda + a0];
  const T *__restrict__ cs = (const T *)ts.csum;
  const T *__restrict__ dq = (const T *)ts.dq;
  const T *__restrict__ h = (const T *)a.field;
  T *Y = (T *)a.Y;
  for (int64_t k = w; k < n; k += kWaves) {
    const T *dk = dq + k * n;
    // the element's loads go out before the distance product
    const int64_t e = (i * n + k) * lda + a0;
    T x = Xs[e], y = Y[e];
    const T hk = h[i * n + k];
    T acc = fmaT(qB, rs, qC * cs[k * lda + a0]);
    int64_t l = 0;
    for (; l + 4 <= n; l += 4) {
      acc = fmaT(dk[l], v[l * 64 + lane], acc);
      acc = fmaT(dk[l + 1], v[(l + 1) * 64 + lane], acc);
      acc = fmaT(dk[l + 2], v[(l + 2) * 64 + lane], acc);
      acc = fmaT(dk[l + 3], v[(l + 3) * 64 + lane], acc);
    }
    for (; l < n; l++) acc = fmaT(dk[l], v[l * 64 + lane], acc);
    sb_update<T, PROB_TSP>(x, y, -acc, c0, c1, c2, c3, hk);
    ((T *)a.gdst)[e] = x;
    Y[e] = y;
  }
}

// FP32 form on the matrix cores: per (position i, 64-agent tile) the product
// sum_l Dq_kl v_l is the (n x n) . (n x 64) GEMM Dq V.  Wavefront w takes the
// 32-city tiles kt = w, w + 4, ..., each as two 32 x 32 output tiles (agents
// 0..31 and 32..63 of the tile); each v_mfma_f32_32x32x2f32 adds two l columns: A[row][kk] = Dq[kt*32+row][l0+kk]
// in lane row + 32 kk, B[kk][col] = v[l0+kk][at*32+col] in lane col + 32 kk,
// and C[row][col] sits in lane col + 32 ((row / 4) % 2), register
// 4 (row / 8) + row % 4.  K is padded to a multiple of 8 with zero terms
// (zero distance columns, zero v rows); cities past n have zero distance rows.
//
// The kernel is bound by memory latency, not by the matrix cores: every
// phase issues its global loads in batches (indices clamped into range, the
// values of padding slots discarded afterwards) before it uses them.
typedef float sis_floatx16 __attribute__((ext_vector_type(16)));

template <int GATHER>
__global__ __launch_bounds__(kBlock) void k_tsp_step_mfma(StepArgs a, TspArgs ts) {
  extern __shared__ __align__(16) unsigned char s_raw[];
  float *v = reinterpret_cast<float *>(s_raw);  // nep x 64, rows n.. zero
  if (a.ctrl->done) return;
  const int64_t n = ts.n, lda = a.lda, tiles = lda / 64;
  const int64_t ne = ts.nep;  // K padded to a multiple of 2 U with zero terms
  const int64_t i = blockIdx.x / tiles, tile = blockIdx.x - i * tiles;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const float c = (float)a.coupling;
  const int64_t im = (i + n - 1) % n, ip = (i + 1) % n;
  // Addresses: a uniform base per position block (n x lda elements) plus a
  // 32-bit lane offset below n lda <= 2^26 (n <= 256, lda <= 2^18).
  const uint32_t ld = (uint32_t)lda;
  const float *__restrict__ xm_b = (const float *)a.gsrc + im * n * lda + tile * 64;
  const float *__restrict__ xp_b = (const float *)a.gsrc + ip * n * lda + tile * 64;
  // v_l = g(i-1, l) + g(i+1, l), SU rows per batch of loads
  constexpr int SU = 8;
  for (int l0 = w; l0 < (int)ne; l0 += kWaves * SU) {
    float xm[SU], xp[SU];
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int l = l0 + u * kWaves;
      const uint32_t o = (uint32_t)(l < (int)n ? l : (int)n - 1) * ld + (uint32_t)lane;
      xm[u] = xm_b[o];
      xp[u] = xp_b[o];
    }
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int l = l0 + u * kWaves;
      if (l < (int)ne)
        v[l * 64 + lane] =
            l < (int)n ? tsp_g<float, GATHER>(xm[u], c) + tsp_g<float, GATHER>(xp[u], c) : 0.0f;
    }
  }
  __syncthreads();
  const int64_t t = a.ctrl->base + a.offset;
  const float *sc = (const float *)a.sched + 4 * t;
  const float c0 = sc[0], c1 = sc[1], c2 = sc[2], c3 = sc[3];
  const float qB = (float)ts.qB, qC = (float)ts.qC;
  const float *__restrict__ rs_b = (const float *)ts.rsum + i * lda + tile * 64;
  const float *__restrict__ cs_b = (const float *)ts.csum + tile * 64;
  const float *__restrict__ dq = ts.dqm;
  const float *__restrict__ h_b = (const float *)a.field + i * n;
  const float *__restrict__ x_b = (const float *)a.gsrc + i * n * lda + tile * 64;
  float *__restrict__ xn_b = (float *)a.gdst + i * n * lda + tile * 64;
  float *__restrict__ y_b = (float *)a.Y + i * n * lda + tile * 64;
  const int half = lane >> 5, col = lane & 31;
  const int64_t ktiles = (n + 31) / 32;
  // element update of the output tile (cities kt*32.., agents at*32..), 4
  // elements per batch of loads (122 VGPRs: 3 wavefronts per SIMD)
  auto finish = [&](int64_t kt, int at, const sis_floatx16 &acc) {
    const uint32_t ag = (uint32_t)(at * 32 + col);
    const float rs = rs_b[ag];
#pragma unroll
    for (int q0 = 0; q0 < 16; q0 += 4) {
      float cv[4], xv[4], yv[4], hv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int q = q0 + u;
        const int k = (int)kt * 32 + 8 * (q >> 2) + 4 * half + (q & 3);
        const uint32_t kc = (uint32_t)(k < (int)n ? k : (int)n - 1);
        const uint32_t e = kc * ld + ag;
        cv[u] = cs_b[e];
        xv[u] = x_b[e];
        yv[u] = y_b[e];
        hv[u] = h_b[kc];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int q = q0 + u;
        const int k = (int)kt * 32 + 8 * (q >> 2) + 4 * half + (q & 3);
        if (k < (int)n) {
          const float S = fmaT(qB, rs, qC * cv[u]) + acc[q];
          const uint32_t e = (uint32_t)k * ld + ag;
          float x = xv[u], y = yv[u];
          sb_update<float, PROB_TSP>(x, y, -S, c0, c1, c2, c3, hv[u]);
          xn_b[e] = x;
          y_b[e] = y;
        }
      }
    }
  };
  // U K-steps of loads go out before their MFMAs
  constexpr int U = 4;
  if (ktiles >= kWaves) {
    // a wavefront takes a 32-city tile for both 32-agent halves: one A
    // operand feeds two accumulators
    for (int64_t kt = w; kt < ktiles; kt += kWaves) {
      const float *drow = dq + kt * 32 * ne;  // rows of the tile; zero rows past n
      const uint32_t dro = (uint32_t)(col * ne);
      sis_floatx16 acc0 = {0}, acc1 = {0};
      float an[U];  // A operands of the next K block, loaded one block ahead
#pragma unroll
      for (int u = 0; u < U; u++) an[u] = drow[dro + (uint32_t)(2 * u + half)];
      for (int64_t l0 = 0; l0 < ne; l0 += 2 * U) {
        float av[U], b0[U], b1[U];
#pragma unroll
        for (int u = 0; u < U; u++) av[u] = an[u];
        if (l0 + 2 * U < ne) {
#pragma unroll
          for (int u = 0; u < U; u++) an[u] = drow[dro + (uint32_t)(l0 + 2 * U + 2 * u + half)];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int64_t l = l0 + 2 * u + half;
          b0[u] = v[l * 64 + col];
          b1[u] = v[l * 64 + 32 + col];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], b0[u], acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], b1[u], acc1, 0, 0, 0);
        }
      }
      finish(kt, 0, acc0);
      finish(kt, 1, acc1);
    }
  } else {
    // few city tiles: one (city tile, agent half) per wavefront
    for (int64_t tt = w; tt < ktiles * 2; tt += kWaves) {
      const int64_t kt = tt >> 1;
      const int at = (int)(tt & 1);
      const float *drow = dq + kt * 32 * ne;  // rows of the tile; zero rows past n
      const uint32_t dro = (uint32_t)(col * ne);
      sis_floatx16 acc = {0};
      float an[U];  // A operands of the next K block, loaded one block ahead
#pragma unroll
      for (int u = 0; u < U; u++) an[u] = drow[dro + (uint32_t)(2 * u + half)];
      for (int64_t l0 = 0; l0 < ne; l0 += 2 * U) {
        float av[U], bv[U];
#pragma unroll
        for (int u = 0; u < U; u++) av[u] = an[u];
        if (l0 + 2 * U < ne) {
#pragma unroll
          for (int u = 0; u < U; u++) an[u] = drow[dro + (uint32_t)(l0 + 2 * U + 2 * u + half)];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int64_t l = l0 + 2 * u + half;
          bv[u] = v[l * 64 + at * 32 + col];
        }
#pragma unroll
        for (int u = 0; u < U; u++)
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], bv[u], acc, 0, 0, 0);
      }
      finish(kt, at, acc);
    }
  }
}

__device__ __forceinline__ double spin_of(const uint64_t *spins, int64_t e, int64_t W, int64_t a) {
  return ((spins[e * W + (a >> 6)] >> (a & 63)) & 1) ? -1.0 : 1.0;
}

// Row / column sums of the spins (FP64, exact): as k_tsp_sums.
__global__ __launch_bounds__(kBlock) void k_tsp_spin_sums(const uint64_t *spins, double *rsum,
                                                          double *csum, int64_t n, int64_t W,
                                                          int64_t lda) {
  const int64_t tiles = lda / 64;
  const int64_t g = wave_index();
  if (g >= 2 * n * tiles) return;
  const int lane = threadIdx.x & 63;
  const int64_t which = g / (n * tiles), rem = g - which * n * tiles;
  const int64_t idx = rem / tiles, a0 = (rem - idx * tiles) * 64 + lane;
  double acc = 0.0;
  for (int64_t m = 0; m < n; m++)
    acc += spin_of(spins, which == 0 ? idx * n + m : m * n + idx, W, a0);
  (which == 0 ? rsum : csum)[idx * lda + a0] = acc;
}

// Per-position energy partials: part[i][a] = sum_k s_ik ((J s)_ik - h_ik).
// Every term is a multiple of 1/8 of an integer, so the sum is exact.
__global__ __launch_bounds__(kBlock) void k_tsp_energy_struct(const uint64_t *spins,
                                                              const double *field, TspArgs ts,
                                                              int64_t W, int64_t lda,
                                                              double *part) {
  __shared__ double red[kWaves * 64];
  const int64_t n = ts.n, tiles = lda / 64;
  const int64_t i = blockIdx.x / tiles, tile = blockIdx.x - i * tiles;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t a0 = tile * 64 + lane;
  const int64_t im = (i + n - 1) % n, ip = (i + 1) % n;
  const double *rsum = (const double *)ts.rsum, *csum = (const double *)ts.csum;
  const double *dq = (const double *)ts.dq;
  const double rs = rsum[i * lda + a0];
  double e = 0.0;
  for (int64_t k = w; k < n; k += kWaves) {
    double acc = ts.qC * csum[k * lda + a0] + ts.qB * rs;
    for (int64_t l = 0; l < n; l++)
      acc += dq[k * n + l] * (spin_of(spins, im * n + l, W, a0) + spin_of(spins, ip * n + l, W, a0));
    const double s = spin_of(spins, i * n + k, W, a0);
    e += s * acc - field[i * n + k] * s;
  }
  red[w * 64 + lane] = e;
  __syncthreads();
  if (w == 0)
    part[i * lda + a0] = ((red[lane] + red[64 + lane]) + red[128 + lane]) + red[192 + lane];
}

// Dynamic LDS above 64 KB (the staged v at n > 256 in FP32, n > 128 in FP64):
// raise the kernel's limit to the 160 KB of a gfx950 CU, once per kernel and
// device (devices 0..63; the attribute belongs to the current device).
template <auto K>
static void allow_big_lds() {
  static std::atomic<uint64_t> done{0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  const uint64_t bit = uint64_t(1) << dev;
  if (!(done.load() & bit)) {
    (void)hipFuncSetAttribute((const void *)K, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    done.fetch_or(bit);
  }
}

constexpr int kResBlock = 1024;  // threads of a resident workgroup

// Position stride of the resident TSP kernel's g buffers: the smallest
// PS >= n G with PS = G (mod 32).
__host__ __device__ constexpr int tsp_res_stride(int64_t n, int G) {
  return (int)(n * G + ((G - n * G) % 32 + 32) % 32);
}

// Resident structured TSP (FP32, up to 48 output tiles of 16 x 16 and 160 KB
// of LDS: n <= 64 at G = 2, n <= 96 at G = 1): one 1024-thread workgroup owns G
// agents for a whole launch of `steps` steps.  Per step it needs, for every
// agent, the row and column sums of g and the GEMM Dq . V with
// V[l][(i, a)] = g(i-1, l, a) + g(i+1, l, a): (n x n) . (n x nG), computed
// with v_mfma_f32_16x16x4f32 (A[r][k] in lane r + 16 k, B[k][c] in lane
// c + 16 k, C[r][c] in lane c + 16 (r / 4), register r % 4, checked on the
// device).  Wavefront w owns the output tiles w, w + 16, ... (cities x
// columns), or, when there are at most 3 city tiles and 16 column tiles, the
// city tiles 0 .. 2 of column tile w, whose B operands are then the same (one
// load per K step for all three).  The X / Y / h of exactly the owned
// elements stay in registers in
// the accumulator layout, so the GEMM result needs no exchange.  LDS holds g
// of the current and next step (n positions each, element (i, k, agent) at
// i PS + k G + agent, with the position stride PS >= n G padded to PS = G
// mod 32 so that the positions a wavefront reads together fall into distinct
// banks), Dq and the sums.  Two workgroup barriers per step.
constexpr int kTspResTiles = 3;  // output tiles per wavefront (48 tiles of 16 x 16)

template <int GATHER>
__global__ __launch_bounds__(kResBlock) void k_tsp_resident(StepArgs a, TspArgs ts, ResArgs r) {
  extern __shared__ __align__(16) unsigned char s_raw[];
  if (a.ctrl->done) return;
  const int n = (int)ts.n, G = r.G, lg = r.lgG, NC = n * G, nG = n * G;
  const int nep = (int)ts.nep, CT = (n + 15) / 16, NT = (NC + 15) / 16;
  const int64_t lda = a.lda;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t g0 = (int64_t)blockIdx.x * G;
  const int PS = tsp_res_stride(n, G);
  float *snap = reinterpret_cast<float *>(s_raw);  // 2 x n PS
  // Dq rows at an odd stride (nep + 1): the 16 rows one MFMA A operand reads
  // fall into distinct LDS banks
  const int dst = nep + 1;
  float *dqs = snap + 2 * n * PS;                  // CT*16 x dst
  float *rs = dqs + CT * 16 * dst;                 // n x G row sums
  float *cs = rs + nG;                             // n x G column sums
  float *hs = cs + nG;                             // n x n field
  for (int e = t; e < CT * 16 * nep; e += kResBlock) dqs[(e / nep) * dst + e % nep] = ts.dqm[e];
  for (int e = t; e < n * n; e += kResBlock) hs[e] = ((const float *)a.field)[e];
  // owned elements: tile u of this wavefront, register q
  float x[kTspResTiles][4], y[kTspResTiles][4];
  const float *__restrict__ Xin = r.Xin;
  float *Y = (float *)a.Y;
  const float c = (float)a.coupling;
  // tile u of this wavefront: city tile ct, column tile nt (wavefront-uniform)
  const bool colown = CT <= kTspResTiles && NT <= 16;
  auto tile_of = [&](int u, int &ct, int &nt) {
    if (colown) {
      ct = u;
      nt = w;
      return u < CT && w < NT;
    }
    const int tile = w + 16 * u;
    ct = tile / NT;
    nt = tile - ct * NT;
    return tile < CT * NT;
  };
  auto elem = [&](int u, int q, int &k, int &col) {
    int ct, nt;
    const bool ok = tile_of(u, ct, nt);
    k = ct * 16 + 4 * (lane >> 4) + q;
    col = nt * 16 + (lane & 15);
    return ok && k < n && col < NC;
  };
#pragma unroll
  for (int u = 0; u < kTspResTiles; u++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int k, col;
      x[u][q] = y[u][q] = 0.0f;
      if (elem(u, q, k, col)) {
        const int i = col >> lg, ag = col & (G - 1);
        const int64_t e = (int64_t)(i * n + k) * lda + g0 + ag;
        x[u][q] = Xin[e];
        y[u][q] = Y[e];
      }
    }
  auto put = [&](float *buf) {
#pragma unroll
    for (int u = 0; u < kTspResTiles; u++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int k, col;
        if (elem(u, q, k, col)) {
          const int i = col >> lg, ag = col & (G - 1);
          buf[i * PS + k * G + ag] = tsp_g<float, GATHER>(x[u][q], c);
        }
      }
  };
  put(snap);
  __syncthreads();
  const float qB = (float)ts.qB, qC = (float)ts.qC;
  // Everything below that does not change between steps, computed once:
  // the wavefront's tiles, the B-operand rows of each tile, and per owned
  // element its field value and its row-sum, column-sum and g offsets.
  int tim[kTspResTiles], tip[kTspResTiles], rsi[kTspResTiles], pbase[kTspResTiles];
  bool tok[kTspResTiles], tcok[kTspResTiles];
  const float *drow[kTspResTiles];
  int kb[kTspResTiles], hb[kTspResTiles];  // element (u, q): cs at kb + q G, h at hb + q
  uint32_t eok = 0;  // bit 4 u + q: element (u, q) is owned
#pragma unroll
  for (int u = 0; u < kTspResTiles; u++) {
    int ct, nt;
    tok[u] = tile_of(u, ct, nt);  // wavefront-uniform
    if (!tok[u]) ct = nt = 0;
    const int col = nt * 16 + (lane & 15);
    tcok[u] = tok[u] && col < NC;
    const int i = tcok[u] ? col >> lg : 0, ag = col & (G - 1);
    tim[u] = ((i + n - 1) % n) * PS + ag;
    tip[u] = ((i + 1) % n) * PS + ag;
    rsi[u] = i * G + ag;
    pbase[u] = i * PS;
    drow[u] = dqs + (ct * 16 + (lane & 15)) * dst;
    const int k0 = ct * 16 + 4 * (lane >> 4);
    kb[u] = k0 * G + ag;
    hb[u] = i * n + k0;
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (tcok[u] && k0 + q < n) eok |= 1u << (4 * u + q);
  }
  // sums: thread t takes items t, t + 1024, ... (at most 4: 16 n G <= 4096);
  // item = 8 sum + part sums every 8th term of row or column sum `sum`
  // (sums nG .. 2 nG - 1 are the column sums: cs follows rs)
  constexpr int kSumItems = 4;
  int sb[kSumItems];
#pragma unroll
  for (int j = 0; j < kSumItems; j++) {
    const int item = t + kResBlock * j;
    const int sum = item >> 3, part = item & 7;
    const bool row = sum < nG;
    const int idx = (row ? sum : sum - nG) >> lg, ag = sum & (G - 1);
    sb[j] = row ? idx * PS + part * G + ag : part * PS + idx * G + ag;
  }
  const float *sc = (const float *)a.sched + 4 * (a.ctrl->base + a.offset);
  float n0 = sc[0], n1 = sc[1], n2 = sc[2], n3 = sc[3];
  for (int st = 0; st < r.steps; st++) {
    const float c0 = n0, c1 = n1, c2 = n2, c3 = n3;
    if (st + 1 < r.steps) {
      n0 = sc[4 * (st + 1)];
      n1 = sc[4 * (st + 1) + 1];
      n2 = sc[4 * (st + 1) + 2];
      n3 = sc[4 * (st + 1) + 3];
    }
    const float *S = snap + (st & 1) * n * PS;
    // row sums (position i, agent a) and column sums (city k, agent a): 8
    // lanes per sum, each over every 8th term, then a butterfly over the 8
#pragma unroll
    for (int j = 0; j < kSumItems; j++) {
      const int item = t + kResBlock * j;
      if (item >= 2 * nG * 8) break;  // the same for the whole wavefront
      const int part = item & 7;
      const int stride = (item >> 3) < nG ? 8 * G : 8 * PS;
      float acc = 0.0f;
      int o = sb[j];
      for (int m = part; m < n; m += 8, o += stride) acc += S[o];
      acc += __shfl_xor(acc, 1, 64);
      acc += __shfl_xor(acc, 2, 64);
      acc += __shfl_xor(acc, 4, 64);
      if (part == 0) rs[item >> 3] = acc;
    }
    __syncthreads();
    float *Sn = snap + ((st + 1) & 1) * n * PS;
    // the wavefront's tiles advance through K together: independent
    // accumulators, so consecutive MFMAs do not wait on each other
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 acc[kTspResTiles];
#pragma unroll
    for (int u = 0; u < kTspResTiles; u++) acc[u] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    if (colown) {
      // one column tile: the B operand is shared by the wavefront's tiles
      for (int l0 = 0; l0 < n; l0 += 4) {
        const int l = l0 + (lane >> 4);
        float av[kTspResTiles];
#pragma unroll
        for (int u = 0; u < kTspResTiles; u++) av[u] = drow[u][l];  // zero past n
        const float bv = (tcok[0] && l < n) ? S[tim[0] + l * G] + S[tip[0] + l * G] : 0.0f;
#pragma unroll
        for (int u = 0; u < kTspResTiles; u++)
          if (tok[u]) acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], bv, acc[u], 0, 0, 0);
      }
    } else {
      for (int l0 = 0; l0 < n; l0 += 4) {
        const int l = l0 + (lane >> 4);
        float av[kTspResTiles], bv[kTspResTiles];
#pragma unroll
        for (int u = 0; u < kTspResTiles; u++) {
          av[u] = drow[u][l];  // zero past n (padded table)
          bv[u] = (tcok[u] && l < n) ? S[tim[u] + l * G] + S[tip[u] + l * G] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kTspResTiles; u++)
          if (tok[u]) acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], bv[u], acc[u], 0, 0, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < kTspResTiles; u++) {
      if (!tok[u]) continue;
      const float rv = rs[rsi[u]];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if ((eok >> (4 * u + q)) & 1) {
          const int ci = kb[u] + q * G;
          const float Sv = fmaT(qB, rv, qC * cs[ci]) + acc[u][q];
          sb_update<float, PROB_TSP>(x[u][q], y[u][q], -Sv, c0, c1, c2, c3, hs[hb[u] + q]);
          Sn[pbase[u] + ci] = tsp_g<float, GATHER>(x[u][q], c);
        }
      }
    }
    __syncthreads();
  }
  float *Xout = r.Xout;
#pragma unroll
  for (int u = 0; u < kTspResTiles; u++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int k, col;
      if (elem(u, q, k, col)) {
        const int i = col >> lg, ag = col & (G - 1);
        const int64_t e = (int64_t)(i * n + k) * lda + g0 + ag;
        Xout[e] = x[u][q];
        Y[e] = y[u][q];
      }
    }
}

size_t tsp_resident_lds(int64_t n, int G, int64_t nep) {
  return sizeof(float) *
         (size_t)(2 * n * tsp_res_stride(n, G) + (n + 15) / 16 * 16 * (nep + 1) + 2 * n * G + n * n);
}

bool tsp_resident_ok(int64_t n, int G, int64_t nep) {
  const int64_t tiles = (n + 15) / 16 * ((n * G + 15) / 16);
  return n >= 3 && (G == 1 || G == 2 || G == 4) && tiles <= 16 * kTspResTiles &&
         16 * n * G <= 4 * kResBlock &&  // at most 4 sum items per thread
         tsp_resident_lds(n, G, nep) <= 160 * 1024;
}

void launch_tsp_resident(hipStream_t st, int gather, const StepArgs &a, const TspArgs &ts,
                         const ResArgs &r) {
  const dim3 grid((unsigned)(a.lda / r.G));
  const size_t shm = tsp_resident_lds(ts.n, r.G, ts.nep);
  if (gather == G_BSB) {
    allow_big_lds<k_tsp_resident<G_BSB>>();
    hipLaunchKernelGGL((k_tsp_resident<G_BSB>), grid, dim3(kResBlock), shm, st, a, ts, r);
  } else {
    allow_big_lds<k_tsp_resident<G_DSB>>();
    hipLaunchKernelGGL((k_tsp_resident<G_DSB>), grid, dim3(kResBlock), shm, st, a, ts, r);
  }
}

bool tsp_struct_supported(int form, int gather) {
  return (form == 0 || form == 1) && (gather == G_DSB || gather == G_BSB);
}


void launch_tsp_struct_step(hipStream_t st, int form, int gather, bool mfma, const StepArgs &a,
                            const TspArgs &ts) {
  const int64_t tiles = a.lda / 64;
  const unsigned sums_grid = grid_for(2 * ts.n * tiles, kWaves, INT32_MAX);
  const dim3 grid((unsigned)(ts.n * tiles));
  if (mfma && form == 0) {
    const size_t shm = (size_t)(ts.nep * 64 * sizeof(float));
    if (gather == G_BSB) {
      allow_big_lds<k_tsp_step_mfma<G_BSB>>();
      hipLaunchKernelGGL((k_tsp_sums<float, G_BSB>), dim3(sums_grid), dim3(kBlock), 0, st, a, ts);
      hipLaunchKernelGGL((k_tsp_step_mfma<G_BSB>), grid, dim3(kBlock), shm, st, a, ts);
    } else {
      allow_big_lds<k_tsp_step_mfma<G_DSB>>();
      hipLaunchKernelGGL((k_tsp_sums<float, G_DSB>), dim3(sums_grid), dim3(kBlock), 0, st, a, ts);
      hipLaunchKernelGGL((k_tsp_step_mfma<G_DSB>), grid, dim3(kBlock), shm, st, a, ts);
    }
    return;
  }
#define SIS_TSPS(T, G)                                                                       \
  if (form == (sizeof(T) == 8 ? 1 : 0) && gather == G) {                                    \
    allow_big_lds<k_tsp_step_struct<T, G>>();                                               \
    hipLaunchKernelGGL((k_tsp_sums<T, G>), dim3(sums_grid), dim3(kBlock), 0, st, a, ts);    \
    hipLaunchKernelGGL((k_tsp_step_struct<T, G>), grid, dim3(kBlock),                       \
                       (size_t)(ts.n * 64 * sizeof(T)), st, a, ts);                         \
    return;                                                                                 \
  }
  SIS_TSPS(float, G_BSB)
  SIS_TSPS(float, G_DSB)
  SIS_TSPS(double, G_BSB)
  SIS_TSPS(double, G_DSB)
#undef SIS_TSPS
}

void launch_tsp_struct_energy(hipStream_t st, const uint64_t *spins, const double *field,
                              const TspArgs &ts, int64_t W, int64_t lda, double *part) {
  const int64_t tiles = lda / 64;
  hipLaunchKernelGGL(k_tsp_spin_sums, dim3(grid_for(2 * ts.n * tiles, kWaves, INT32_MAX)),
                     dim3(kBlock), 0, st, spins, (double *)ts.rsum, (double *)ts.csum, ts.n, W, lda);
  hipLaunchKernelGGL(k_tsp_energy_struct, dim3((unsigned)(ts.n * tiles)), dim3(kBlock), 0, st,
                     spins, field, ts, W, lda, part);
}

// ------------------------------------------------- resident small problems
//
// Agents are independent, so one workgroup can own G agents of a small
// problem outright: their X and Y stay in registers and the gather source in
// LDS for `steps` consecutive steps, with one workgroup barrier per step and
// no HBM traffic in between.  Weights are +-1.
// Pair q = row * G + agent is owned by thread q mod 1024 (register slot
// q / 1024), so the G agents of a row sit in consecutive lanes of one
// wavefront.
// The gather source holds each term ready to add.  dSB: byte slot 2q the term
// of a +1 weight on pair q, 2q + 1 that of a -1 weight (one 2-byte write per
// pair); bSB: float slot q and q + N G (split, so that graphs with only +1
// weights still read every LDS bank).  2G zero slots at 2 N G pad the rows.  dSB: int8 -sgn(x) (+1 where x has its sign bit set),
// summed as integers; bSB: float -(x * coupling), so F + slot equals the row
// kernels' fma(-v, x * coupling, F).  The matrix is in LDS in ELL groups of the
// R = 64 / G rows of one wavefront register slot: entry k of row g R + l is
// u16 number 4 (c R + l) + (k - 4 c) past the group's start, c = k / 4, so a
// lane reads four entries with one 8-byte load and a wavefront's loads are
// contiguous.  An entry is the slot of its (column, weight) for agent 0.
// Terms are added in CSR order (bSB) or exactly (dSB integer sums):
// results are bit-identical to k_step_plain / k_step_packed.

// SORTED: register slot pairs follow the degree order of r.perm; each slot
// keeps its pair (row G + agent) in qp[m].
template <int GATHER, int M, int SORTED>
__global__ __launch_bounds__(kResBlock) void k_resident(StepArgs a, ResArgs r) {
  extern __shared__ __align__(16) unsigned char s_raw[];
  if (a.ctrl->done) return;
  const int N = (int)a.N, G = r.G, lg = r.lgG;
  const int NG = N * G;
  const int64_t lda = a.lda;
  const int t = threadIdx.x, lane = t & 63;
  const int64_t g0 = (int64_t)blockIdx.x * G;
  uint2 *sent = reinterpret_cast<uint2 *>(s_raw);  // ELL, four entries per uint2
  unsigned char *snap = s_raw + r.snap_off;         // two buffers of r.snap_bytes
  const uint2 *__restrict__ gell = reinterpret_cast<const uint2 *>(r.ell);
  for (int k = t; k < (r.ell_words >> 2); k += kResBlock) sent[k] = gell[k];
  if (t < 2 * G) {  // padding slots of both buffers
    if (GATHER == G_DSB) {
      reinterpret_cast<int8_t *>(snap)[2 * NG + t] = 0;
      reinterpret_cast<int8_t *>(snap + r.snap_bytes)[2 * NG + t] = 0;
    } else {
      reinterpret_cast<float *>(snap)[2 * NG + t] = 0.0f;
      reinterpret_cast<float *>(snap + r.snap_bytes)[2 * NG + t] = 0.0f;
    }
  }
  // register slot m of this wavefront holds ELL group wave + 16 m: its chunks
  // start at gbase[m] (in uint2), gwid[m] chunks per row
  const int R = 64 >> lg, rloc = lane >> lg;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);  // wavefront-uniform: scalar loads
  uint32_t gbase[M], gwid[M];
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int grp = wv + 16 * m;
    const bool has = grp < r.ngroups;
    gbase[m] = has ? r.gofs[grp] >> 2 : 0;
    gwid[m] = has ? ((r.gofs[grp + 1] >> 2) - gbase[m]) / (uint32_t)R : 0;
  }
  // own pairs, and the gather source of step 0
  float x[M], y[M];
  int qp[M];  // the slot's pair: row G + agent
  const float *__restrict__ Xin = r.Xin;
  float *Y = (float *)a.Y;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int q = t + kResBlock * m;
    x[m] = y[m] = 0.0f;
    qp[m] = q;
    if (q < NG) {
      if (SORTED) qp[m] = ((int)r.perm[q >> lg] << lg) + (q & (G - 1));
      const int64_t e = (int64_t)(qp[m] >> lg) * lda + g0 + (q & (G - 1));
      x[m] = Xin[e];
      y[m] = Y[e];
    }
  }
  const float coupling = (float)a.coupling;
  auto put_snapshot = [&](unsigned char *buf) {
#pragma unroll
    for (int m = 0; m < M; m++) {
      const int q = t + kResBlock * m;
      if (q < NG) {
        const int qs = SORTED ? qp[m] : q;
        if (GATHER == G_DSB) {  // bytes (sg, -sg): 0x01 0xff or 0xff 0x01
          reinterpret_cast<uint16_t *>(buf)[qs] = signbit(x[m]) ? 0xff01u : 0x01ffu;
        } else {
          const float p = x[m] * coupling;
          reinterpret_cast<float *>(buf)[qs] = -p;
          reinterpret_cast<float *>(buf)[qs + NG] = p;
        }
      }
    }
  };
  put_snapshot(snap);
  __syncthreads();
  // step coefficients, loaded one step ahead so no step waits on them
  const float *sc = (const float *)a.sched + 4 * (a.ctrl->base + a.offset);
  float n0 = sc[0], n1 = sc[1], n2 = sc[2];
  for (int st = 0; st < r.steps; st++) {
    const float c0 = n0, c1 = n1, c2 = n2;
    if (st + 1 < r.steps) {
      n0 = sc[4 * (st + 1)];
      n1 = sc[4 * (st + 1) + 1];
      n2 = sc[4 * (st + 1) + 2];
    }
    const unsigned char *cur = snap + (st & 1) * r.snap_bytes;
#pragma unroll
    for (int m = 0; m < M; m++) {
      const int q = t + kResBlock * m;
      if (q < NG) {
        const int ag = (GATHER == G_DSB ? 2 : 1) * (q & (G - 1));
        const uint2 *ent = sent + gbase[m] + rloc;
        const uint32_t w = gwid[m];
        float F;
        if (GATHER == G_DSB) {
          const int8_t *src = reinterpret_cast<const int8_t *>(cur) + ag;
          int Fi = 0;
          for (uint32_t c = 0; c < w; c++) {
            const uint2 e = ent[c * R];
            Fi += src[e.x & 0xffff];
            Fi += src[e.x >> 16];
            Fi += src[e.y & 0xffff];
            Fi += src[e.y >> 16];
          }
          F = (float)Fi;
        } else {
          const float *src = reinterpret_cast<const float *>(cur) + ag;
          F = 0.0f;
          for (uint32_t c = 0; c < w; c++) {
            const uint2 e = ent[c * R];
            const float t0 = src[e.x & 0xffff], t1 = src[e.x >> 16];
            const float t2 = src[e.y & 0xffff], t3 = src[e.y >> 16];
            F = F + t0;
            F = F + t1;
            F = F + t2;
            F = F + t3;
          }
        }
        sb_update<float, PROB_MAXCUT>(x[m], y[m], F, c0, c1, c2, 0.0f, 0.0f);
      }
    }
    put_snapshot(snap + ((st + 1) & 1) * r.snap_bytes);
    __syncthreads();
  }
  float *Xout = r.Xout;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int q = t + kResBlock * m;
    if (q < NG) {
      const int64_t e = (int64_t)((SORTED ? qp[m] : q) >> lg) * lda + g0 + (q & (G - 1));
      Xout[e] = x[m];
      Y[e] = y[m];
    }
  }
}

// The dSB resident tile with the ELL in registers (tile bit 23): for
// matrices whose ELL groups are at most CW chunks of four entries wide (the
// torus: one), each thread loads the chunks of its own register slots once,
// so LDS holds only the two int8 gather-slot buffers.  With no ELL in LDS two
// 1024-thread workgroups fit on one CU (8 waves per SIMD: 64 VGPRs), and a term
// takes one LDS round trip (the slot) instead of two (ELL entry, then slot).
// Same sums in the same order as k_resident<G_DSB>: bit-identical.
template <int M, int SORTED, int CW>
__global__ __launch_bounds__(kResBlock, 8) void k_resident_reg(StepArgs a, ResArgs r) {
  extern __shared__ __align__(16) unsigned char s_raw[];
  if (a.ctrl->done) return;
  const int N = (int)a.N, G = r.G, lg = r.lgG;
  const int NG = N * G;
  const int64_t lda = a.lda;
  const int t = threadIdx.x, lane = t & 63;
  const int64_t g0 = (int64_t)blockIdx.x * G;
  unsigned char *snap = s_raw;  // two buffers of r.snap_bytes
  if (t < 2 * G) {              // padding slots of both buffers
    reinterpret_cast<int8_t *>(snap)[2 * NG + t] = 0;
    reinterpret_cast<int8_t *>(snap + r.snap_bytes)[2 * NG + t] = 0;
  }
  const int R = 64 >> lg, rloc = lane >> lg;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const uint2 *__restrict__ gell = reinterpret_cast<const uint2 *>(r.ell);
  // chunks past a group's width, and slots past the last group, read the
  // zero slots: the gather below has no branch, so the reads of all slots
  // can be in flight together
  const uint32_t z = (uint32_t)(2 * NG), zz = z | z << 16;
  uint2 ent[M][CW];
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int grp = wv + 16 * m;
    const bool has = grp < r.ngroups;
    const uint32_t gb = has ? r.gofs[grp] >> 2 : 0;
    const uint32_t gw = has ? ((r.gofs[grp + 1] >> 2) - gb) / (uint32_t)R : 0;
#pragma unroll
    for (int c = 0; c < CW; c++)
      ent[m][c] = (uint32_t)c < gw ? gell[gb + c * R + rloc] : make_uint2(zz, zz);
  }
  float x[M], y[M];
  int qp[M];
  const float *__restrict__ Xin = r.Xin;
  float *Y = (float *)a.Y;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int q = t + kResBlock * m;
    x[m] = y[m] = 0.0f;
    qp[m] = q;
    if (q < NG) {
      if (SORTED) qp[m] = ((int)r.perm[q >> lg] << lg) + (q & (G - 1));
      const int64_t e = (int64_t)(qp[m] >> lg) * lda + g0 + (q & (G - 1));
      x[m] = Xin[e];
      y[m] = Y[e];
    }
  }
  auto put_snapshot = [&](unsigned char *buf) {
#pragma unroll
    for (int m = 0; m < M; m++) {
      const int q = t + kResBlock * m;
      if (q < NG) {
        const int qs = SORTED ? qp[m] : q;
        reinterpret_cast<uint16_t *>(buf)[qs] = signbit(x[m]) ? 0xff01u : 0x01ffu;
      }
    }
  };
  put_snapshot(snap);
  __syncthreads();
  const float *sc = (const float *)a.sched + 4 * (a.ctrl->base + a.offset);
  float n0 = sc[0], n1 = sc[1], n2 = sc[2];
  for (int st = 0; st < r.steps; st++) {
    const float c0 = n0, c1 = n1, c2 = n2;
    if (st + 1 < r.steps) {
      n0 = sc[4 * (st + 1)];
      n1 = sc[4 * (st + 1) + 1];
      n2 = sc[4 * (st + 1) + 2];
    }
    const int8_t *src = reinterpret_cast<const int8_t *>(snap + (st & 1) * r.snap_bytes) +
                        2 * (t & (G - 1));
    // opaque each step: the entry extracts are redone per step (one SDWA add
    // each) instead of being hoisted into 4 M registers of addresses
#pragma unroll
    for (int m = 0; m < M; m++)
#pragma unroll
      for (int c = 0; c < CW; c++) asm volatile("" : "+v"(ent[m][c].x), "+v"(ent[m][c].y));
    // two slots at a time: their reads in flight together, then their update
    // and the next step's gather slots (the other buffer); the scheduling
    // barrier keeps batches apart so live ranges stay within 64 VGPRs
    unsigned char *nxt = snap + ((st + 1) & 1) * r.snap_bytes;
#pragma unroll
    for (int m0 = 0; m0 < M; m0 += 2) {
      int Fi[2] = {0, 0};
#pragma unroll
      for (int b = 0; b < 2; b++) {
        if (m0 + b < M) {
#pragma unroll
          for (int c = 0; c < CW; c++) {
            const uint2 e = ent[m0 + b][c];
            Fi[b] += src[e.x & 0xffff] + src[e.x >> 16] + src[e.y & 0xffff] + src[e.y >> 16];
          }
        }
      }
      // pairs past N G update zeros and are neither stored nor gathered
#pragma unroll
      for (int b = 0; b < 2; b++)
        if (m0 + b < M) sb_update_maxcut_carried(x[m0 + b], y[m0 + b], (float)Fi[b], c0, c1, c2);
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int m = m0 + b, q = t + kResBlock * m;
        if (m < M && q < NG) {
          const int qs = SORTED ? qp[m] : q;
          reinterpret_cast<uint16_t *>(nxt)[qs] = signbit(x[m]) ? 0xff01u : 0x01ffu;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
  }
  float *Xout = r.Xout;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int q = t + kResBlock * m;
    if (q < NG) {
      const int64_t e = (int64_t)((SORTED ? qp[m] : q) >> lg) * lda + g0 + (q & (G - 1));
      Xout[e] = x[m];
      Y[e] = y[m];
    }
  }
}

int resident_pairs_max() { return kResBlock * 12; }  // the largest instance below: M = 12
int resident_sorted_pairs_max() { return kResBlock * 8; }  // sorted instances up to M = 8

bool launch_resident(hipStream_t st, int gather, const StepArgs &a, const ResArgs &r) {
  const int64_t NG = a.N * r.G;
  const dim3 grid((unsigned)(a.lda / r.G));
  const size_t shm = (size_t)r.snap_off + 2 * (size_t)r.snap_bytes;
  const int sorted = r.perm != nullptr;
  if (r.ell_cw > 0) {
    if (gather != G_DSB || r.snap_off != 0) return false;
#define SIS_RESR(MM, SO, CW)                                                                 \
  if (sorted == SO && r.ell_cw <= CW && NG <= (int64_t)kResBlock * MM) {                    \
    allow_big_lds<k_resident_reg<MM, SO, CW>>();                                            \
    hipLaunchKernelGGL((k_resident_reg<MM, SO, CW>), grid, dim3(kResBlock), shm, st, a, r); \
    return true;                                                                            \
  }
    SIS_RESR(4, 0, 1)
    SIS_RESR(8, 0, 1)
    SIS_RESR(10, 0, 1)
    SIS_RESR(12, 0, 1)
    SIS_RESR(4, 0, 2)
    SIS_RESR(4, 1, 1)
    SIS_RESR(8, 1, 1)
    SIS_RESR(4, 1, 2)
#undef SIS_RESR
    return false;
  }
#define SIS_RES(GA, MM, SO)                                                                  \
  if (gather == GA && sorted == SO && NG <= (int64_t)kResBlock * MM) {                      \
    allow_big_lds<k_resident<GA, MM, SO>>();                                                \
    hipLaunchKernelGGL((k_resident<GA, MM, SO>), grid, dim3(kResBlock), shm, st, a, r);    \
    return true;                                                                            \
  }
  SIS_RES(G_DSB, 4, 0)
  SIS_RES(G_DSB, 8, 0)
  SIS_RES(G_DSB, 10, 0)
  SIS_RES(G_DSB, 12, 0)
  SIS_RES(G_BSB, 4, 0)
  SIS_RES(G_BSB, 8, 0)
  SIS_RES(G_BSB, 10, 0)
  SIS_RES(G_BSB, 12, 0)
  SIS_RES(G_DSB, 4, 1)
  SIS_RES(G_DSB, 8, 1)
  SIS_RES(G_BSB, 4, 1)
  SIS_RES(G_BSB, 8, 1)
#undef SIS_RES
  return false;
}

// -------------------------------------------------------------- windows

// Sign scan of the window (src/kernels.cu:37-44, 84-90), race-free: each
// (row, word) is compared and overwritten by the one wavefront that owns it;
// per-block OR of the changed bits goes to part[block][word].
template <int SRC>
__global__ __launch_bounds__(kBlock) void k_window_scan(const void *xsrc, uint64_t *cur,
                                                        uint64_t *part, int64_t N, int64_t W,
                                                        int64_t lda, int64_t rpb,
                                                        const Ctrl *ctrl) {
  extern __shared__ uint64_t s_acc[];
  if (ctrl->done) return;
  for (int64_t w = threadIdx.x; w < W; w += blockDim.x) s_acc[w] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * rpb;
  const int64_t r1 = r0 + rpb < N ? r0 + rpb : N;
  if (SRC == SRC_BITS) {
    // the words are already packed: one (row, word) per thread, coalesced
    for (int64_t idx = r0 * W + threadIdx.x; idx < r1 * W; idx += blockDim.x) {
      const uint64_t b = ((const uint64_t *)xsrc)[idx];
      const uint64_t d = b ^ cur[idx];
      cur[idx] = b;
      if (d) atomicOr((unsigned long long *)&s_acc[idx % W], (unsigned long long)d);
    }
    __syncthreads();
    for (int64_t w = threadIdx.x; w < W; w += blockDim.x) part[(int64_t)blockIdx.x * W + w] = s_acc[w];
    return;
  }
  // one (row, word) per wavefront: ballot of 64 agents' signs; the X values
  // and old words of U pairs are all loaded before the first ballot.  Pair
  // indices within a block fit 32 bits (rows per block x words).
  constexpr int U = 8;
  const uint32_t npairs = (uint32_t)((r1 - r0) * W), uW = (uint32_t)W;
  for (uint32_t p0 = (uint32_t)q * U; p0 < npairs; p0 += (uint32_t)kWaves * U) {
    bool neg[U];
    uint64_t old[U];
    uint32_t wu[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t p = p0 + u < npairs ? p0 + u : npairs - 1;
      const uint32_t di = p / uW;
      wu[u] = p - di * uW;
      const int64_t i = r0 + di;
      if (SRC == SRC_F32) neg[u] = signbit(((const float *)xsrc)[i * lda + 64 * wu[u] + lane]);
      else neg[u] = signbit(((const double *)xsrc)[i * lda + 64 * wu[u] + lane]);
      old[u] = cur[i * W + wu[u]];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (p0 + u >= npairs) break;
      const int64_t idx = r0 * W + (int64_t)(p0 + u);
      const uint64_t b = __ballot(neg[u]);
      const uint64_t d = b ^ old[u];
      if (lane == 0) {
        cur[idx] = b;
        if (d) atomicOr((unsigned long long *)&s_acc[wu[u]], (unsigned long long)d);
      }
    }
  }
  __syncthreads();
  for (int64_t w = threadIdx.x; w < W; w += blockDim.x) part[(int64_t)blockIdx.x * W + w] = s_acc[w];
}

// Per-agent stability / lock update (src/kernels.cu:46-69) in one block.
__global__ __launch_bounds__(1024) void k_window_finalize(const uint64_t *part, int64_t nblk,
                                                          uint64_t *newmask, int32_t *stab,
                                                          int32_t *bif, int64_t A, int64_t W,
                                                          int K, int uw, int64_t P, Ctrl *ctrl) {
  __shared__ int s_nbif, s_any;
  if (ctrl->done) {
    if (threadIdx.x == 0) ctrl->merge = 0;
    return;
  }
  if (threadIdx.x == 0) {
    s_nbif = 0;
    s_any = 0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int64_t base = threadIdx.x & ~63; base < W * 64; base += blockDim.x) {
    const int64_t w = base >> 6, ag = base + lane;
    uint64_t m = 0;
    for (int64_t b = lane; b < nblk; b += 64) m |= part[b * W + w];
    m = wave_or64(m);
    const int changed = (int)((m >> lane) & 1);
    int now = 0, fresh = 0;
    if (ag < A) {
      const int prev = bif[ag];
      int st = stab[ag];
      if (!prev) st = changed ? 0 : st + 1;
      now = st == K - 1;
      stab[ag] = st;
      bif[ag] = now;
      fresh = now && !prev;
    }
    const uint64_t nm = __ballot(fresh), bm = __ballot(now);
    if (lane == 0) {
      newmask[w] = nm;
      if (nm) atomicOr(&s_any, 1);
      atomicAdd(&s_nbif, (int)__popcll(bm));
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ctrl->merge = s_any;
    ctrl->nbif = s_nbif;
    ctrl->base += P;
    if (uw && (int64_t)s_nbif == A) ctrl->done = 1;
  }
}

// final = cur for newly locked / valid agents (src/kernels.cu:73-79, 524-531).
__global__ void k_window_merge(const uint64_t *cur, uint64_t *fin, const uint64_t *newmask,
                               int64_t N, int64_t W, const Ctrl *ctrl) {
  if (!ctrl->merge) return;
  const int64_t n = N * W;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t nm = newmask[idx % W];
    if (nm) fin[idx] = (fin[idx] & ~nm) | (cur[idx] & nm);
  }
}

// TSP validity (src/kernels.cu:475-500): exactly one +1 per position and per
// city.  Stage 1, one wavefront per (line, word), line = position i (< n) or
// city k (>= n): over the line's n spin words y = ~word (bit set = +1),
// ones |= y after twos |= ones & y, so ones & ~twos marks the agents with
// exactly one +1 on the line.  Stage 2 ANDs the 2n line masks of each word.
__global__ __launch_bounds__(kBlock) void k_tsp_lines(const uint64_t *__restrict__ cur, int64_t n,
                                                      int64_t W, uint64_t *lines,
                                                      const Ctrl *ctrl) {
  (void)ctrl;
  const int64_t g = wave_index();
  if (g >= 2 * n * W) return;
  const int lane = threadIdx.x & 63;
  const int64_t line = g / W, w = g - line * W;
  const int64_t base = line < n ? line * n : line - n, stride = line < n ? 1 : n;
  uint64_t ones = 0, twos = 0;
  for (int64_t m = 0; m < n; m++) {
    const uint64_t y = ~cur[(base + m * stride) * W + w];
    twos |= ones & y;
    ones |= y;
  }
  if (lane == 0) lines[line * W + w] = ones & ~twos;
}

__global__ __launch_bounds__(kBlock) void k_tsp_valid(const uint64_t *__restrict__ lines, int64_t n,
                                                      int64_t W, int64_t A, int32_t *valid_list,
                                                      uint64_t *newmask, int64_t P,
                                                      const Ctrl *ctrl) {
  const int64_t w = wave_index();
  if (w >= W) return;
  const int lane = threadIdx.x & 63;
  uint64_t m = ~0ull;
  for (int64_t line = lane; line < 2 * n; line += 64) m &= lines[line * W + w];
  for (int s = 32; s >= 1; s >>= 1) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)m, s, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(m >> 32), s, 64);
    m &= ((uint64_t)hi << 32) | lo;
  }
  const int64_t ag = 64 * w + lane;
  const int valid = ((m >> lane) & 1) && ag < A;
  const uint64_t nm = __ballot(valid);
  if (lane == 0) newmask[w] = nm;
  if (valid) valid_list[ag] = (int32_t)(ctrl->base + P);
}

__global__ void k_ctrl_advance(Ctrl *ctrl, int64_t P) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctrl->base += P;
}

// ---------------------------------------------------------- evaluation

// Per-block partial sums of H = sum_i 0.5 s_i sum_k J_k s_col
// (src/kernels.cu:371-379), written as H = 0.5 sum_k J_k s_row(k) s_col(k)
// over the stored entries, or of E = sum_k J_k s_row s_col - sum_i h_i s_i.
// Every wavefront takes an equal slice of the entries (so one long row does
// not hold up the sum) and finds the row of its first entry by binary search
// on the row pointers; the field term is taken over the block's row range.
// Partials of the kWaves wavefronts are added in a fixed order:
// part[block][agent].  Integer weights give exact sums in any order.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_cut_rows(const int64_t *__restrict__ rp,
                                                     const int32_t *__restrict__ col,
                                                     const double *__restrict__ val,
                                                     const double *__restrict__ field,
                                                     const uint64_t *__restrict__ spins,
                                                     int64_t N, int64_t W, int64_t lda,
                                                     int64_t rpb, double *part) {
  __shared__ double s_red[kWaves][64];
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t Z = rp[N];
  const int64_t nchunk = (int64_t)gridDim.x * kWaves;
  const int64_t per = (Z + nchunk - 1) / nchunk;
  const int64_t c = (int64_t)blockIdx.x * kWaves + q;
  const int64_t ka = c * per < Z ? c * per : Z;
  const int64_t kb = ka + per < Z ? ka + per : Z;
  // last row i < N with rp[i] <= ka (rp[N] = Z > ka when the slice is not empty)
  int64_t i0 = 0;
  if (ka < kb) {
    int64_t lo = 0, hi = N;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (rp[mid] <= ka) lo = mid;
      else hi = mid;
    }
    i0 = lo;
  }
  const int64_t r0 = (int64_t)blockIdx.x * rpb;
  const int64_t r1 = r0 + rpb < N ? r0 + rpb : N;
  for (int64_t w = 0; w < W; w++) {
    double acc = 0.0;
    if (ka < kb) {
      // 64 entries per pass: lane l loads entry kc + l's weight and column
      // sign word, and the end and sign word of row ib + l (a block of the
      // rows the pass walks through); the terms are then added one entry at a
      // time in CSR order from readlane broadcasts.
      int64_t i = i0, ib = i0;
      auto load_rows = [&](int64_t base, int64_t &bnd, uint64_t &wi) {
        const int64_t rr = base + lane;
        bnd = rr < N ? rp[rr + 1] : INT64_MAX;
        wi = rr < N ? spins[rr * W + w] : 0;
      };
      int64_t bnd;
      uint64_t wi;
      load_rows(ib, bnd, wi);
      int64_t iend = (int64_t)readlane64((uint64_t)bnd, 0);
      double si = ((readlane64(wi, 0) >> lane) & 1) ? -1.0 : 1.0;
      for (int64_t kc = ka; kc < kb; kc += 64) {
        const int cnt = kb - kc < 64 ? (int)(kb - kc) : 64;
        const int64_t kk = lane < cnt ? kc + lane : kc;
        const double vk = val[kk];
        const uint64_t wj = spins[(int64_t)col[kk] * W + w];
        for (int e = 0; e < cnt; e++) {
          const int64_t k = kc + e;
          if (k >= iend) {
            do {
              i++;
              if (i - ib == 64) {  // past the loaded block of rows
                ib = i;
                load_rows(ib, bnd, wi);
              }
              iend = (int64_t)readlane64((uint64_t)bnd, (int)(i - ib));
            } while (k >= iend);
            si = ((readlane64(wi, (int)(i - ib)) >> lane) & 1) ? -1.0 : 1.0;
          }
          const double sj = ((readlane64(wj, e) >> lane) & 1) ? -1.0 : 1.0;
          const double v = __longlong_as_double((long long)readlane64((uint64_t)__double_as_longlong(vk), e));
          acc = __builtin_fma(v, si * sj, acc);
        }
      }
      if (MODE == 0) acc *= 0.5;
    }
    if (MODE == 1)
      for (int64_t r = r0 + q; r < r1; r += kWaves) {
        const double sr = ((spins[r * W + w] >> lane) & 1) ? -1.0 : 1.0;
        acc = __builtin_fma(-field[r], sr, acc);
      }
    s_red[q][lane] = acc;
    __syncthreads();
    if (q == 0)
      part[(int64_t)blockIdx.x * lda + 64 * w + lane] =
          ((s_red[0][lane] + s_red[1][lane]) + s_red[2][lane]) + s_red[3][lane];
    __syncthreads();
  }
}

// mode 0: cut = -0.5 (H - S) (src/kernels.cu:393-398); mode 1: energy.
// One workgroup per 64 agents: wavefront q adds the block partials
// q, q + 16, ... in order, then the 16 wavefront sums are added in order.
constexpr int kRedWaves = 16;
__global__ __launch_bounds__(64 * kRedWaves) void k_cut_reduce(const double *part, int64_t nblk,
                                                               int64_t A, int64_t lda, int mode,
                                                               double S, double *out) {
  __shared__ double s_h[kRedWaves][64];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int64_t ag = (int64_t)blockIdx.x * 64 + lane;
  double H = 0.0;
  for (int64_t b = q; b < nblk; b += kRedWaves) H += part[b * lda + ag];
  s_h[q][lane] = H;
  __syncthreads();
  if (q != 0) return;
  H = s_h[0][lane];
  for (int k = 1; k < kRedWaves; k++) H += s_h[k][lane];
  if (ag < A) out[ag] = mode == 0 ? (-0.5) * (H - S) : H;
}

__device__ __forceinline__ int64_t d_tri(int64_t c1, int64_t c2, int64_t n) {
  return c2 + c1 * n - ((c1 + 2) * (c1 + 1)) / 2;
}

// Tour of each valid agent from its final spins, in (position, city) order
// (src/ising_tsp.cu:206-214), and its cyclic length (src/ising_tsp.cu:272-287).
__global__ void k_tsp_len(const uint64_t *__restrict__ fin, const int32_t *valid_list, int64_t n,
                          int64_t W, int64_t A, const double *__restrict__ dist, double *len) {
  const int64_t ag = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = ag >> 6;
  const int lane = (int)(ag & 63);
  if (w >= W) return;
  int64_t m = 0, first = -1, prev = -1;
  double sum = 0.0;
  // spins in (position, city) order; the words (one per spin, shared by the
  // wavefront's 64 agents) are loaded 16 at a time before they are walked
  constexpr int B = 16;
  const int64_t nn = n * n;
  for (int64_t s0 = 0; s0 < nn; s0 += B) {
    uint64_t wd[B];
#pragma unroll
    for (int u = 0; u < B; u++) wd[u] = fin[(s0 + u < nn ? s0 + u : nn - 1) * W + w];
#pragma unroll
    for (int u = 0; u < B; u++) {
      if (s0 + u >= nn || ((wd[u] >> lane) & 1)) continue;
      const int64_t c = (s0 + u) % n;
      if (m == 0) first = c;
      else if (prev != c) sum += dist[d_tri(prev < c ? prev : c, prev < c ? c : prev, n)];
      prev = c;
      m++;
    }
  }
  if (ag >= A) return;
  if (valid_list[ag] <= 0 || m != n) {
    len[ag] = -1.0;
    return;
  }
  if (n > 1 && prev != first) sum += dist[d_tri(prev < first ? prev : first, prev < first ? first : prev, n)];
  len[ag] = sum;
}

// Best agent: mode 0 = largest value, mode 1 = smallest value >= 0; ties go
// to the lowest index; -1 if no candidate.
__global__ __launch_bounds__(1024) void k_best(const double *v, int64_t A, int mode, int64_t *best) {
  __shared__ double s_v[1024];
  __shared__ int64_t s_i[1024];
  double bv = 0.0;
  int64_t bi = -1;
  for (int64_t a = threadIdx.x; a < A; a += blockDim.x) {
    const double x = v[a];
    if (mode == 1 && !(x >= 0.0)) continue;
    const bool better = bi < 0 || (mode == 0 ? x > bv : x < bv);
    if (better) {
      bv = x;
      bi = a;
    }
  }
  s_v[threadIdx.x] = bv;
  s_i[threadIdx.x] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double ov = s_v[threadIdx.x + s];
      const int64_t oi = s_i[threadIdx.x + s];
      const int64_t mi = s_i[threadIdx.x];
      const double mv = s_v[threadIdx.x];
      bool take = false;
      if (oi >= 0) {
        if (mi < 0) take = true;
        else if (mode == 0 ? ov > mv : ov < mv) take = true;
        else if (ov == mv && oi < mi) take = true;
      }
      if (take) {
        s_v[threadIdx.x] = ov;
        s_i[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *best = s_i[0];
}

// ------------------------------------------------------------ launchers

void launch_init_state(hipStream_t st, bool f64, void *X, void *Y, int64_t N, int64_t A,
                       int64_t lda, uint64_t seed, bool tsp, const int32_t *rowid) {
  const unsigned gb = grid_for(N * lda, 256, 16384);
  if (f64)
    hipLaunchKernelGGL(k_init_state<double>, dim3(gb), dim3(256), 0, st, (double *)X, (double *)Y,
                       N, A, lda, seed, (int)tsp, rowid);
  else
    hipLaunchKernelGGL(k_init_state<float>, dim3(gb), dim3(256), 0, st, (float *)X, (float *)Y, N,
                       A, lda, seed, (int)tsp, rowid);
}

void launch_prestep(hipStream_t st, bool f64, void *X, void *Y, int64_t n, double dt) {
  const unsigned gb = grid_for(n, 256, 16384);
  const double cy = dt * (0.0 - 1), cx = dt * (0.0 + 1);
  if (f64)
    hipLaunchKernelGGL(k_prestep<double>, dim3(gb), dim3(256), 0, st, (double *)X, (double *)Y, n,
                       cy, cx);
  else
    hipLaunchKernelGGL(k_prestep<float>, dim3(gb), dim3(256), 0, st, (float *)X, (float *)Y, n,
                       (float)cy, (float)cx);
}

void launch_init_qn(hipStream_t st, int qw, uint32_t *P, int64_t N, int64_t A, int64_t lda,
                    uint64_t seed, int qshift, const int32_t *rowid) {
  const int64_t Q = lda / 256;
  const dim3 grid(grid_for(N * Q * 64, 256, 16384));
  if (qw == 12)
    hipLaunchKernelGGL(k_init_qn<12>, grid, dim3(256), 0, st, P, N, A, Q, seed, qshift, rowid);
  else if (qw == 10)
    hipLaunchKernelGGL(k_init_qn<10>, grid, dim3(256), 0, st, P, N, A, Q, seed, qshift, rowid);
  else
    hipLaunchKernelGGL(k_init_qn<8>, grid, dim3(256), 0, st, P, N, A, Q, seed, qshift, rowid);
}

void launch_prestep_qn(hipStream_t st, int qw, uint32_t *P, int64_t N, int64_t lda, double dt,
                       int qshift) {
  const int64_t n = N * (lda / 256) * 64;
  const double cy = dt * (0.0 - 1), cx = dt * (0.0 + 1);
  const dim3 grid(grid_for(n, 256, 16384));
  if (qw == 12)
    hipLaunchKernelGGL(k_prestep_qn<12>, grid, dim3(256), 0, st, P, n, (float)cy, (float)cx, qshift);
  else if (qw == 10)
    hipLaunchKernelGGL(k_prestep_qn<10>, grid, dim3(256), 0, st, P, n, (float)cy, (float)cx, qshift);
  else
    hipLaunchKernelGGL(k_prestep_qn<8>, grid, dim3(256), 0, st, P, n, (float)cy, (float)cx, qshift);
}

void launch_pack_signs_qn(hipStream_t st, int qw, const uint32_t *P, uint64_t *S, int64_t N,
                          int64_t lda) {
  const int64_t Q = lda / 256;
  const dim3 grid(grid_for(N * Q, kWaves, INT32_MAX));
  if (qw == 12)
    hipLaunchKernelGGL(k_pack_signs_qn<12>, grid, dim3(kBlock), 0, st, P, S, N, Q);
  else if (qw == 10)
    hipLaunchKernelGGL(k_pack_signs_qn<10>, grid, dim3(kBlock), 0, st, P, S, N, Q);
  else
    hipLaunchKernelGGL(k_pack_signs_qn<8>, grid, dim3(kBlock), 0, st, P, S, N, Q);
}

void launch_bf16_shadow(hipStream_t st, const float *X, uint16_t *sh, int64_t n) {
  hipLaunchKernelGGL(k_bf16_shadow, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, X, sh, n);
}

void launch_pack_signs(hipStream_t st, int src, const void *X, uint64_t *S, int64_t N, int64_t W,
                       int64_t lda) {
  const unsigned gb = grid_for(N * W, kWaves, INT32_MAX);
  if (src == SRC_F64)
    hipLaunchKernelGGL(k_pack_signs<double>, dim3(gb), dim3(kBlock), 0, st, (const double *)X, S,
                       N, W, lda);
  else
    hipLaunchKernelGGL(k_pack_signs<float>, dim3(gb), dim3(kBlock), 0, st, (const float *)X, S, N,
                       W, lda);
}

void launch_xy_update(hipStream_t st, bool f64, void *X, void *Y, int64_t n, const void *sched,
                      const Ctrl *ctrl, int32_t offset) {
  const unsigned gb = grid_for(n, 256, 16384);
  if (f64)
    hipLaunchKernelGGL(k_xy_update<double>, dim3(gb), dim3(256), 0, st, (double *)X, (double *)Y,
                       n, (const double *)sched, ctrl, offset);
  else
    hipLaunchKernelGGL(k_xy_update<float>, dim3(gb), dim3(256), 0, st, (float *)X, (float *)Y, n,
                       (const float *)sched, ctrl, offset);
}

bool step_supported(int form, int prob, int gather, int vec) {
  if (form == 4 || form == 5 || form == 6 || form == 7) {
    // 4 | rows << 8 | non-temporal << 16 | staged gathers << 17
    const int rows = (vec >> 8) & 0xff, vg = (vec >> 17) & 1;
    if (prob != PROB_MAXCUT || gather != G_DSB || (vec & 0xff) != 4 || (vec >> 18) != 0)
      return false;
    return vg ? (rows == 4 || rows == 8 || rows == 16 || rows == 32)
              : (rows == 1 || rows == 2 || rows == 4);
  }
  if (form == 2) {
    const int wt = vec & 0xff, rows = (vec >> 8) & 0xff;
    return (prob == PROB_MAXCUT || prob == PROB_TSP) && gather == G_DSB && (vec >> 17) == 0 &&
           (wt == 1 || wt == 2 || wt == 4 || wt == 8) && (rows == 1 || rows == 2 || rows == 4);
  }
  // plain / BF16: agents per lane | non-temporal << 16
  const int v = vec & 0xffff;
  if ((vec >> 17) != 0 || (v != 1 && v != 2 && v != 4)) return false;
  if (prob == PROB_UNFUSED) return (form == 0 || (form == 1 && v != 4)) && gather == G_DSB;
  if (prob != PROB_MAXCUT && prob != PROB_TSP) return false;
  switch (form) {
    case 0: return gather == G_DSB || gather == G_BSB;
    case 1: return (gather == G_DSB || gather == G_BSB) && v != 4;
    case 3: return gather == G_BSB;
  }
  return false;
}

template <typename K>
static void launch_waves(hipStream_t st, K kern, int64_t waves, const StepArgs &a) {
  hipLaunchKernelGGL(kern, dim3(grid_for(waves, kWaves, INT32_MAX)), dim3(kBlock), 0, st, a);
}

#define SIS_PLAIN1(T, P, G, V, NT)                                                  \
  if (prob == P && gather == G && v == V && nt == NT) {                             \
    launch_waves(st, k_step_plain<T, P, G, V, NT>, a.N * (a.lda / (64 * V)), a);    \
    return;                                                                         \
  }
#define SIS_PLAIN(T, P, G, V) SIS_PLAIN1(T, P, G, V, 0) SIS_PLAIN1(T, P, G, V, 1)

void launch_step(hipStream_t st, int form, int prob, int gather, int vec, const StepArgs &a) {
  const int v = vec & 0xffff, nt = (vec >> 16) & 1;
  if (form == 0) {
    SIS_PLAIN(float, PROB_MAXCUT, G_DSB, 1)
    SIS_PLAIN(float, PROB_MAXCUT, G_DSB, 2)
    SIS_PLAIN(float, PROB_MAXCUT, G_DSB, 4)
    SIS_PLAIN(float, PROB_MAXCUT, G_BSB, 1)
    SIS_PLAIN(float, PROB_MAXCUT, G_BSB, 2)
    SIS_PLAIN(float, PROB_MAXCUT, G_BSB, 4)
    SIS_PLAIN(float, PROB_TSP, G_DSB, 1)
    SIS_PLAIN(float, PROB_TSP, G_DSB, 2)
    SIS_PLAIN(float, PROB_TSP, G_DSB, 4)
    SIS_PLAIN(float, PROB_TSP, G_BSB, 1)
    SIS_PLAIN(float, PROB_TSP, G_BSB, 2)
    SIS_PLAIN(float, PROB_TSP, G_BSB, 4)
    SIS_PLAIN(float, PROB_UNFUSED, G_DSB, 1)
    SIS_PLAIN(float, PROB_UNFUSED, G_DSB, 2)
    SIS_PLAIN(float, PROB_UNFUSED, G_DSB, 4)
  } else if (form == 1) {
    SIS_PLAIN(double, PROB_UNFUSED, G_DSB, 1)
    SIS_PLAIN(double, PROB_UNFUSED, G_DSB, 2)
    SIS_PLAIN(double, PROB_MAXCUT, G_DSB, 1)
    SIS_PLAIN(double, PROB_MAXCUT, G_DSB, 2)
    SIS_PLAIN(double, PROB_MAXCUT, G_BSB, 1)
    SIS_PLAIN(double, PROB_MAXCUT, G_BSB, 2)
    SIS_PLAIN(double, PROB_TSP, G_DSB, 1)
    SIS_PLAIN(double, PROB_TSP, G_DSB, 2)
    SIS_PLAIN(double, PROB_TSP, G_BSB, 1)
    SIS_PLAIN(double, PROB_TSP, G_BSB, 2)
  } else if (form == 2) {
    const int wt = vec & 0xff, rows = (vec >> 8) & 0xff;
    const int64_t waves = ((a.N + rows - 1) / rows) * (a.W / wt);
#define SIS_PACKED(P, V, R, NT)                                              \
  if (prob == P && wt == V && rows == R && nt == NT) {                       \
    launch_waves(st, k_step_packed<P, V, R, NT>, waves, a);                  \
    return;                                                                  \
  }
#define SIS_PACKED_R(P, V, NT) SIS_PACKED(P, V, 1, NT) SIS_PACKED(P, V, 2, NT) SIS_PACKED(P, V, 4, NT)
#define SIS_PACKED_V(P, NT) SIS_PACKED_R(P, 1, NT) SIS_PACKED_R(P, 2, NT) SIS_PACKED_R(P, 4, NT) SIS_PACKED_R(P, 8, NT)
    SIS_PACKED_V(PROB_MAXCUT, 0)
    SIS_PACKED_V(PROB_MAXCUT, 1)
    SIS_PACKED_V(PROB_TSP, 0)
    SIS_PACKED_V(PROB_TSP, 1)
#undef SIS_PACKED_V
#undef SIS_PACKED_R
#undef SIS_PACKED
  } else if (form == 4 || form == 5 || form == 6 || form == 7) {
    const int rows = (vec >> 8) & 0xff, vg = (vec >> 17) & 1;
    const int64_t waves = ((a.N + rows - 1) / rows) * (a.W / 4);
#define SIS_QN(QW, R, NT, VG, SR)                                   \
  if (qw == QW && sr == SR && rows == R && nt == NT && vg == VG) {  \
    launch_waves(st, k_step_qn<QW, R, NT, VG, SR>, waves, a);       \
    return;                                                         \
  }
#define SIS_QN_NT(R, VG)                                                                    \
  SIS_QN(12, R, 0, VG, 0) SIS_QN(12, R, 1, VG, 0) SIS_QN(10, R, 0, VG, 0)                   \
  SIS_QN(10, R, 1, VG, 0) SIS_QN(8, R, 0, VG, 0) SIS_QN(8, R, 1, VG, 0)                     \
  SIS_QN(8, R, 0, VG, 1) SIS_QN(8, R, 1, VG, 1)
    const int qw = form == 4 ? 12 : form == 6 ? 10 : 8, sr = form == 7;
    SIS_QN_NT(1, 0)
    SIS_QN_NT(2, 0)
    SIS_QN_NT(4, 0)
    SIS_QN_NT(4, 1)
    SIS_QN_NT(8, 1)
    SIS_QN_NT(16, 1)
    SIS_QN_NT(32, 1)
#undef SIS_QN_NT
#undef SIS_QN
  } else if (form == 3) {
#define SIS_BF16(P, V, NT)                                                      \
  if (prob == P && v == V && nt == NT) {                                        \
    launch_waves(st, k_step_bf16<P, V, NT>, a.N * (a.lda / (64 * V)), a);       \
    return;                                                                     \
  }
    SIS_BF16(PROB_MAXCUT, 1, 0)
    SIS_BF16(PROB_MAXCUT, 2, 0)
    SIS_BF16(PROB_MAXCUT, 4, 0)
    SIS_BF16(PROB_TSP, 1, 0)
    SIS_BF16(PROB_TSP, 2, 0)
    SIS_BF16(PROB_TSP, 4, 0)
    SIS_BF16(PROB_MAXCUT, 1, 1)
    SIS_BF16(PROB_MAXCUT, 2, 1)
    SIS_BF16(PROB_MAXCUT, 4, 1)
    SIS_BF16(PROB_TSP, 1, 1)
    SIS_BF16(PROB_TSP, 2, 1)
    SIS_BF16(PROB_TSP, 4, 1)
#undef SIS_BF16
  }
}
#undef SIS_PLAIN
#undef SIS_PLAIN1

bool split_supported(int form, int prob, int gather) {
  switch (form) {
    case 0:
    case 1:
      return ((prob == PROB_MAXCUT || prob == PROB_TSP) && (gather == G_DSB || gather == G_BSB)) ||
             (prob == PROB_UNFUSED && gather == G_DSB);
    case 2: return (prob == PROB_MAXCUT || prob == PROB_TSP) && gather == G_DSB;
    case 3: return (prob == PROB_MAXCUT || prob == PROB_TSP) && gather == G_BSB;
  }
  return false;
}

void launch_step_split(hipStream_t st, int form, int prob, int gather, int64_t nent,
                       const StepArgs &a, const SplitArgs &sp) {
  const dim3 grid((unsigned)(nent * a.W));
#define SIS_SPLIT(F, P, G)                                                      \
  if (form == F && prob == P && gather == G) {                                  \
    hipLaunchKernelGGL((k_step_split<F, P, G>), grid, dim3(kBlock), 0, st, a, sp); \
    return;                                                                     \
  }
  SIS_SPLIT(0, PROB_MAXCUT, G_DSB)
  SIS_SPLIT(0, PROB_MAXCUT, G_BSB)
  SIS_SPLIT(0, PROB_TSP, G_DSB)
  SIS_SPLIT(0, PROB_TSP, G_BSB)
  SIS_SPLIT(0, PROB_UNFUSED, G_DSB)
  SIS_SPLIT(1, PROB_MAXCUT, G_DSB)
  SIS_SPLIT(1, PROB_MAXCUT, G_BSB)
  SIS_SPLIT(1, PROB_TSP, G_DSB)
  SIS_SPLIT(1, PROB_TSP, G_BSB)
  SIS_SPLIT(1, PROB_UNFUSED, G_DSB)
  SIS_SPLIT(2, PROB_MAXCUT, G_DSB)
  SIS_SPLIT(2, PROB_TSP, G_DSB)
  SIS_SPLIT(3, PROB_MAXCUT, G_BSB)
  SIS_SPLIT(3, PROB_TSP, G_BSB)
#undef SIS_SPLIT
}

void launch_window_scan(hipStream_t st, int src, const void *xsrc, uint64_t *cur, uint64_t *part,
                        int64_t N, int64_t W, int64_t lda, int64_t nblk, int64_t rows_per_blk,
                        const Ctrl *ctrl) {
  const size_t shm = sizeof(uint64_t) * (size_t)W;
  if (src == SRC_BITS)
    hipLaunchKernelGGL(k_window_scan<SRC_BITS>, dim3((unsigned)nblk), dim3(kBlock), shm, st, xsrc,
                       cur, part, N, W, lda, rows_per_blk, ctrl);
  else if (src == SRC_F64)
    hipLaunchKernelGGL(k_window_scan<SRC_F64>, dim3((unsigned)nblk), dim3(kBlock), shm, st, xsrc,
                       cur, part, N, W, lda, rows_per_blk, ctrl);
  else
    hipLaunchKernelGGL(k_window_scan<SRC_F32>, dim3((unsigned)nblk), dim3(kBlock), shm, st, xsrc,
                       cur, part, N, W, lda, rows_per_blk, ctrl);
}

void launch_window_finalize(hipStream_t st, const uint64_t *part, int64_t nblk, uint64_t *newmask,
                            int32_t *stab, int32_t *bif, int64_t A, int64_t W, int K, int uw,
                            int64_t P, Ctrl *ctrl) {
  hipLaunchKernelGGL(k_window_finalize, dim3(1), dim3(1024), 0, st, part, nblk, newmask, stab, bif,
                     A, W, K, uw, P, ctrl);
}

void launch_window_merge(hipStream_t st, const uint64_t *cur, uint64_t *fin,
                         const uint64_t *newmask, int64_t N, int64_t W, const Ctrl *ctrl) {
  hipLaunchKernelGGL(k_window_merge, dim3(grid_for(N * W, 256, 8192)), dim3(256), 0, st, cur, fin,
                     newmask, N, W, ctrl);
}

void launch_tsp_valid(hipStream_t st, const uint64_t *cur, int64_t n, int64_t W, int64_t A,
                      int32_t *valid_list, uint64_t *newmask, uint64_t *lines, int64_t P,
                      Ctrl *ctrl) {
  hipLaunchKernelGGL(k_tsp_lines, dim3(grid_for(2 * n * W, kWaves, INT32_MAX)), dim3(kBlock), 0, st,
                     cur, n, W, lines, ctrl);
  hipLaunchKernelGGL(k_tsp_valid, dim3(grid_for(W, kWaves)), dim3(kBlock), 0, st, lines, n, W, A,
                     valid_list, newmask, P, ctrl);
}

void launch_ctrl_advance(hipStream_t st, Ctrl *ctrl, int64_t P) {
  hipLaunchKernelGGL(k_ctrl_advance, dim3(1), dim3(64), 0, st, ctrl, P);
}

void launch_cut_rows(hipStream_t st, const int64_t *rp, const int32_t *col, const double *val,
                     const double *field, const uint64_t *spins, int64_t N, int64_t W,
                     int64_t lda, int64_t nblk, int64_t rows_per_blk, double *part) {
  if (field)
    hipLaunchKernelGGL(k_cut_rows<1>, dim3((unsigned)nblk), dim3(kBlock), 0, st, rp, col, val,
                       field, spins, N, W, lda, rows_per_blk, part);
  else
    hipLaunchKernelGGL(k_cut_rows<0>, dim3((unsigned)nblk), dim3(kBlock), 0, st, rp, col, val,
                       field, spins, N, W, lda, rows_per_blk, part);
}

void launch_cut_reduce(hipStream_t st, const double *part, int64_t nblk, int64_t A, int64_t lda,
                       int mode, double S, double *out) {
  hipLaunchKernelGGL(k_cut_reduce, dim3((unsigned)(lda / 64)), dim3(64 * kRedWaves), 0, st, part,
                     nblk, A, lda, mode, S, out);
}

void launch_tsp_len(hipStream_t st, const uint64_t *fin, const int32_t *valid_list, int64_t n,
                    int64_t W, int64_t A, const double *dist, double *len) {
  hipLaunchKernelGGL(k_tsp_len, dim3(grid_for(W * 64, 256)), dim3(256), 0, st, fin, valid_list, n,
                     W, A, dist, len);
}

void launch_best(hipStream_t st, const double *v, int64_t A, int mode, int64_t *best) {
  hipLaunchKernelGGL(k_best, dim3(1), dim3(1024), 0, st, v, A, mode, best);
}

}  // namespace sis
