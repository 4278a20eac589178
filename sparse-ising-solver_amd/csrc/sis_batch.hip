// sis_batch.hip -- continuous batching of many small MaxCut dSB solves.
//
// Layout: every resident instance owns a contiguous range of global rows and
// of nonzeros.  Column indices are stored as global rows, so one gather path
// serves all instances.  State arrays (X, Y: rows x lda FP32; sign, current
// and final spin words: rows x W) are indexed by global row.  A task is
// (slot, first row, rows) and the persistent step kernel walks the task table
// with a grid-stride loop, so the kernel and the captured graph do not change
// when instances come and go: only the task table and the slot records do.
//
// Per instance the arithmetic is exactly that of k_step_packed (MaxCut): the
// coefficients of step t are evaluated in FP64 from (t, T, dt, xi0) with the
// same operations as the single solver's host schedule, the window follows
// k_window_scan / k_window_finalize / k_window_merge, and the cut follows
// k_cut_rows / k_cut_reduce -- so an instance gives bit-identical results
// alone (sis_solver_*, SIS_FORM_PACKED) and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <unordered_map>
#include <vector>

#include "sis_device.hpp"
#include "sis_hip_util.hpp"
#include "sis_internal.hpp"

namespace sis {

enum : int32_t { SLOT_FREE = 0, SLOT_ACTIVE = 1, SLOT_DONE = 2 };

struct BInst {
  int64_t row_off;   // first global row
  int32_t n;         // rows
  int32_t agents;    // A <= 64 W
  int32_t t;         // steps executed
  int32_t T;         // max_steps (>= 1 while active)
  int32_t P, K;      // window period and threshold
  int32_t uw, status;
  double dt, xi0, S; // time step, coupling scale, sum of file weights
  int32_t nbif;
  int32_t intw;      // every weight an integer in [-3, 3] (bit-transpose path)
};

// A task: nrows consecutive rows of one instance from global row row0.  For
// tasks of at most 8 rows, kb is the first row's entry in the global CSR and
// off[r] the entry of row r relative to kb (off[nrows] = the task's entries),
// so the pipelined step finds its column range from the record alone.  64
// bytes: one scalar load.
struct BTask {
  int32_t slot, nrows;
  int64_t row0;
  int64_t kb;
  int32_t off[9];
  int32_t pad;
};

constexpr int kXcds = 8;  // MI355X: workgroups go to the 8 XCDs round-robin

struct BArgs {
  const BTask *tasks;
  const int32_t *ntasks;  // [0] tasks; [1 .. kXcds + 1] per-XCD task offsets
  int32_t xcd;            // 1: wavefronts of XCD x take tasks [off[x], off[x + 1])
  BInst *inst;
  const int64_t *rbeg, *rend;  // written only by k_batch_admit, through BCsr
  const int32_t *col;
  const float *val;
  const double *vald;
  float *X, *Y;
  float *sink;  // 16 lda floats + 16 W words that nothing reads (pipelined step)
  uint64_t *S0, *S1, *cur, *fin;
  int32_t *stab, *bif;
  double *cuts;
  int64_t lda, W;
};

// Task range of the calling wavefront: the whole table with a grid stride,
// or (XCD-affine order) the slice of tasks placed on the wavefront's XCD, so
// an instance's sign words stay in that XCD's L2.  Any placement is correct;
// only the cache reuse depends on it.
struct TaskRange {
  int first, end, stride;
};

__device__ __forceinline__ TaskRange task_range(const BArgs &b) {
  TaskRange r;
  if (b.xcd) {
    const int x = blockIdx.x % kXcds;
    const int nb = gridDim.x / kXcds;
    r.first = b.ntasks[1 + x] + (int)(blockIdx.x / kXcds) * kWaves + (int)(threadIdx.x >> 6);
    r.end = b.ntasks[2 + x];
    r.stride = nb * kWaves;
  } else {
    r.first = wave_index();
    r.end = *b.ntasks;
    r.stride = gridDim.x * kWaves;
  }
  r.first = __builtin_amdgcn_readfirstlane(r.first);
  r.end = __builtin_amdgcn_readfirstlane(r.end);
  return r;
}

// One step of every active instance.  Wavefront task = R rows x W words.
template <int R, int WT, int NT>
__global__ __launch_bounds__(kBlock) void k_batch_step(BArgs b) {
  const int lane = threadIdx.x & 63;
  const TaskRange tr = task_range(b);
  const int64_t lda = b.lda, W = WT;
  for (int task = tr.first; task < tr.end; task += tr.stride) {
    const BTask tk = b.tasks[task];
    const BInst &in = b.inst[tk.slot];
    if (in.status != SLOT_ACTIVE) continue;
    const int32_t t = in.t;
    // src/kernels.cu:225-226, 258-262, evaluated as the host schedule does
    double p = 1.025 * (double)t / (double)in.T;
    p = p < 1.0 ? p : 1.0;
    const float c0 = (float)(in.dt * (p + 1)), c1 = (float)(in.dt * (p - 1));
    const float c2 = (float)(in.dt * in.xi0);
    const uint64_t *__restrict__ S = (t & 1) ? b.S1 : b.S0;
    uint64_t *Sn = (t & 1) ? b.S0 : b.S1;
    float x[R][WT], y[R][WT];
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (r < tk.nrows) {
#pragma unroll
        for (int u = 0; u < WT; u++) {
          x[r][u] = ld_state<NT>(b.X + (tk.row0 + r) * lda + 64 * u + lane);
          y[r][u] = ld_state<NT>(b.Y + (tk.row0 + r) * lda + 64 * u + lane);
        }
      }
    }
    // The R rows' gathers are interleaved so their dependent load chains
    // (row range -> column -> sign word) overlap.  Positions past a row's end
    // use the row itself with weight 0: F never holds -0 (it starts at +0 and
    // is only changed by nonzero-sum steps), so adding +-0 leaves it unchanged
    // and each row still accumulates its own entries in ascending order.
    int64_t kb[R], deg[R];
    int64_t maxdeg = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t row = tk.row0 + (r < tk.nrows ? r : 0);
      kb[r] = b.rbeg[row];
      deg[r] = r < tk.nrows ? b.rend[row] - kb[r] : 0;
      maxdeg = deg[r] > maxdeg ? deg[r] : maxdeg;
    }
    float F[R][WT];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int u = 0; u < WT; u++) F[r][u] = 0.0f;
    for (int64_t d = 0; d < maxdeg; d += 2) {
      int64_t j[R][2];
      float v[R][2];
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
          const bool ok = d + i < deg[r];
          const int64_t k = ok ? kb[r] + d + i : 0;
          j[r][i] = ok ? (int64_t)b.col[k] : tk.row0;
          v[r][i] = ok ? b.val[k] : 0.0f;
        }
      uint64_t s[R][2][WT];
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int u = 0; u < WT; u++) s[r][i][u] = S[j[r][i] * W + u];
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int u = 0; u < WT; u++)
            F[r][u] = ((s[r][i][u] >> lane) & 1) ? F[r][u] + v[r][i] : F[r][u] - v[r][i];
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (r >= tk.nrows) break;
      const int64_t row = tk.row0 + r;
      uint64_t mine = 0;
#pragma unroll
      for (int u = 0; u < WT; u++) {
        sb_update<float, PROB_MAXCUT>(x[r][u], y[r][u], F[r][u], c0, c1, c2, 0.0f, 0.0f);
        st_state<NT>(b.X + row * lda + 64 * u + lane, x[r][u]);
        st_state<NT>(b.Y + row * lda + 64 * u + lane, y[r][u]);
        const uint64_t bits = __ballot(signbit(x[r][u]));
        if (lane == u) mine = bits;
      }
      if (lane < WT) Sn[row * W + lane] = mine;
    }
  }
}

// Same step as k_batch_step with the gathers on the vector memory path:
// lane l serves slot l of the task (row l / (64/R), neighbour position
// l % (64/R) of the pass) and loads that neighbour's column, weight and sign
// words; every lane then reads the slots back with readlane, row by row and
// in CSR order, so the arithmetic and its order are unchanged.
template <int R, int WT, int NT>
__global__ __launch_bounds__(kBlock) void k_batch_step_vg(BArgs b) {
  constexpr int SL = 64 / R;  // neighbour slots per row and pass
  const int lane = threadIdx.x & 63;
  const TaskRange tr = task_range(b);
  const int64_t lda = b.lda, W = WT;
  const int my_r = lane / SL, my_i = lane % SL;
  for (int task = tr.first; task < tr.end; task += tr.stride) {
    const BTask tk = b.tasks[task];
    const BInst &in = b.inst[tk.slot];
    if (in.status != SLOT_ACTIVE) continue;
    const int32_t t = in.t;
    double p = 1.025 * (double)t / (double)in.T;
    p = p < 1.0 ? p : 1.0;
    const float c0 = (float)(in.dt * (p + 1)), c1 = (float)(in.dt * (p - 1));
    const float c2 = (float)(in.dt * in.xi0);
    const uint64_t *S = (t & 1) ? b.S1 : b.S0;
    uint64_t *Sn = (t & 1) ? b.S0 : b.S1;
    float x[R][WT], y[R][WT];
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (r < tk.nrows) {
#pragma unroll
        for (int u = 0; u < WT; u++) {
          x[r][u] = ld_state<NT>(b.X + (tk.row0 + r) * lda + 64 * u + lane);
          y[r][u] = ld_state<NT>(b.Y + (tk.row0 + r) * lda + 64 * u + lane);
        }
      }
    }
    // this lane's row range
    const int64_t myrow = tk.row0 + (my_r < tk.nrows ? my_r : 0);
    const int64_t mykb = b.rbeg[myrow];
    const int mydeg = my_r < tk.nrows ? (int)(b.rend[myrow] - mykb) : 0;
    int maxdeg = mydeg;
    for (int m = 32; m >= 1; m >>= 1) {
      const int o = __shfl_xor(maxdeg, m, 64);
      maxdeg = o > maxdeg ? o : maxdeg;
    }
    maxdeg = __builtin_amdgcn_readfirstlane(maxdeg);
    float F[R][WT];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int u = 0; u < WT; u++) F[r][u] = 0.0f;
    for (int d0 = 0; d0 < maxdeg; d0 += SL) {
      const bool ok = d0 + my_i < mydeg;
      const int64_t k = ok ? mykb + d0 + my_i : 0;
      const int64_t j = ok ? (int64_t)b.col[k] : tk.row0;
      const float v = ok ? b.val[k] : 0.0f;
      uint64_t w[WT];
#pragma unroll
      for (int u = 0; u < WT; u++) w[u] = S[j * W + u];
#pragma unroll
      for (int r = 0; r < R; r++) {
#pragma unroll
        for (int i = 0; i < SL; i++) {
          const int l = r * SL + i;
          const float vv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
#pragma unroll
          for (int u = 0; u < WT; u++) {
            const uint64_t ww = readlane64(w[u], l);
            F[r][u] = ((ww >> lane) & 1) ? F[r][u] + vv : F[r][u] - vv;
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (r >= tk.nrows) break;
      const int64_t row = tk.row0 + r;
      uint64_t mine = 0;
#pragma unroll
      for (int u = 0; u < WT; u++) {
        sb_update<float, PROB_MAXCUT>(x[r][u], y[r][u], F[r][u], c0, c1, c2, 0.0f, 0.0f);
        st_state<NT>(b.X + row * lda + 64 * u + lane, x[r][u]);
        st_state<NT>(b.Y + row * lda + 64 * u + lane, y[r][u]);
        const uint64_t bits = __ballot(signbit(x[r][u]));
        if (lane == u) mine = bits;
      }
      if (lane < WT) Sn[row * W + lane] = mine;
    }
  }
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
  return ((uint64_t)hi << 32) | lo;
}

// 64 x 64 bit-matrix transpose across the wavefront: lane l holds word w_l
// (bit a = agent a's sign at neighbour slot l); afterwards lane a holds the
// word whose bit l is bit a of w_l.  Six butterfly stages swap the
// off-diagonal blocks of side 32, 16, ..., 1 between lanes l and l ^ s.
__device__ __forceinline__ uint64_t bit_transpose64(uint64_t w, int lane) {
  constexpr uint64_t kMask[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull,
                                 0x00FF00FF00FF00FFull, 0x0F0F0F0F0F0F0F0Full,
                                 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int st = 0; st < 6; st++) {
    const int s = 32 >> st;
    const uint64_t m = kMask[st];
    const uint64_t p = shfl_xor64(w, s);
    // both halves, then a bitwise select (no divergent branch per stage)
    const uint64_t hi = ((p >> s) & m) | (w & ~m), lo = (w & m) | ((p & m) << s);
    const uint64_t sel = 0ull - (uint64_t)((lane & s) != 0);
    w = (hi & sel) | (lo & ~sel);
  }
  return w;
}

// Same step as k_batch_step_vg.  For instances whose weights are all integers
// in [-3, 3] the neighbour sums are taken on bit masks instead of one readlane
// per entry: after the transpose lane a holds, for each of the pass's 64
// slots, bit l = 1 when agent a's neighbour at slot l is negative; with the
// slots' weights split into bit planes p and signs (wave-uniform ballot
// masks P_p, N_p),
//   sum_l v_l bit_l = sum_p 2^p (popc(m & P_p) - popc(m & N_p))
// restricted to a row's slots, and F = 2 sum_l v_l bit_l - sum_l v_l, which
// is the sequential sum of +-v exactly (every partial sum is an integer
// below 2^24), so results are bit-identical to the other step kernels.
// Instances with other weights take the readlane path of k_batch_step_vg.
template <int R, int WT, int NT, int PF>
__device__ __forceinline__ void batch_step_bt(const BArgs &b) {
  constexpr int SL = 64 / R;  // neighbour slots per row and pass
  constexpr int kPlanes = 2;  // |v| <= 3
  const int lane = threadIdx.x & 63;
  const TaskRange tr = task_range(b);
  const int64_t lda = b.lda, W = WT;
  const int my_r = lane / SL, my_i = lane % SL;
  // PF: the next task's record, state rows and row ranges are loaded while
  // this task finishes, so a task waits only on column -> sign word.
  BTask ntk{};
  bool nact = false;
  float nx[R][WT], ny[R][WT];
  int64_t nkb = 0;
  int ndeg = 0;
  auto stage = [&](int task) {
    ntk = b.tasks[task];
    nact = b.inst[ntk.slot].status == SLOT_ACTIVE;
    if (!nact) return;
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (r < ntk.nrows) {
#pragma unroll
        for (int u = 0; u < WT; u++) {
          nx[r][u] = ld_state<NT>(b.X + (ntk.row0 + r) * lda + 64 * u + lane);
          ny[r][u] = ld_state<NT>(b.Y + (ntk.row0 + r) * lda + 64 * u + lane);
        }
      }
    }
    const int64_t row = ntk.row0 + (my_r < ntk.nrows ? my_r : 0);
    nkb = b.rbeg[row];
    ndeg = my_r < ntk.nrows ? (int)(b.rend[row] - nkb) : 0;
  };
  if (PF && tr.first < tr.end) stage(tr.first);
  for (int task = tr.first; task < tr.end; task += tr.stride) {
    if (PF) {
      if (!nact) {
        if (task + tr.stride < tr.end) stage(task + tr.stride);
        continue;
      }
    } else {
      ntk = b.tasks[task];
      if (b.inst[ntk.slot].status != SLOT_ACTIVE) continue;
      stage(task);
    }
    const BTask tk = ntk;
    const BInst &in = b.inst[tk.slot];
    const int32_t t = in.t;
    double p = 1.025 * (double)t / (double)in.T;
    p = p < 1.0 ? p : 1.0;
    const float c0 = (float)(in.dt * (p + 1)), c1 = (float)(in.dt * (p - 1));
    const float c2 = (float)(in.dt * in.xi0);
    const uint64_t *S = (t & 1) ? b.S1 : b.S0;
    uint64_t *Sn = (t & 1) ? b.S0 : b.S1;
    float x[R][WT], y[R][WT];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int u = 0; u < WT; u++) {
        x[r][u] = nx[r][u];
        y[r][u] = ny[r][u];
      }
    const int64_t mykb = nkb;
    const int mydeg = ndeg;
    int maxdeg = mydeg;
    for (int m = 32; m >= 1; m >>= 1) {
      const int o = __shfl_xor(maxdeg, m, 64);
      maxdeg = o > maxdeg ? o : maxdeg;
    }
    maxdeg = __builtin_amdgcn_readfirstlane(maxdeg);
    float F[R][WT];
    if (in.intw) {
      int Sa[R][WT], Vr[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        Vr[r] = 0;
#pragma unroll
        for (int u = 0; u < WT; u++) Sa[r][u] = 0;
      }
      for (int d0 = 0; d0 < maxdeg; d0 += SL) {
        const bool ok = d0 + my_i < mydeg;
        const int64_t k = ok ? mykb + d0 + my_i : 0;
        const int64_t j = ok ? (int64_t)b.col[k] : tk.row0;
        const int vi = ok ? (int)b.val[k] : 0;
        const int av = vi < 0 ? -vi : vi;
        uint64_t w[WT];
#pragma unroll
        for (int u = 0; u < WT; u++) w[u] = S[j * W + u];
        // bit planes of the slots' weights: P_p (positive), N_p (negative)
        uint64_t Pm[kPlanes], Nm[kPlanes];
#pragma unroll
        for (int q = 0; q < kPlanes; q++) {
          Pm[q] = __ballot(vi > 0 && ((av >> q) & 1));
          Nm[q] = __ballot(vi < 0 && ((av >> q) & 1));
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
          const uint64_t rm = (SL == 64 ? ~0ull : ((1ull << SL) - 1)) << (r * SL);
#pragma unroll
          for (int q = 0; q < kPlanes; q++)
            Vr[r] += (__popcll(Pm[q] & rm) - __popcll(Nm[q] & rm)) * (1 << q);
        }
#pragma unroll
        for (int u = 0; u < WT; u++) {
          const uint64_t m = bit_transpose64(w[u], lane);
#pragma unroll
          for (int r = 0; r < R; r++) {
            const uint64_t rm = (SL == 64 ? ~0ull : ((1ull << SL) - 1)) << (r * SL);
            int acc = 0;
#pragma unroll
            for (int q = 0; q < kPlanes; q++)
              acc += (__popcll(m & Pm[q] & rm) - __popcll(m & Nm[q] & rm)) * (1 << q);
            Sa[r][u] += acc;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int u = 0; u < WT; u++) F[r][u] = (float)(2 * Sa[r][u] - Vr[r]);
    } else {
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int u = 0; u < WT; u++) F[r][u] = 0.0f;
      for (int d0 = 0; d0 < maxdeg; d0 += SL) {
        const bool ok = d0 + my_i < mydeg;
        const int64_t k = ok ? mykb + d0 + my_i : 0;
        const int64_t j = ok ? (int64_t)b.col[k] : tk.row0;
        const float v = ok ? b.val[k] : 0.0f;
        uint64_t w[WT];
#pragma unroll
        for (int u = 0; u < WT; u++) w[u] = S[j * W + u];
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
          for (int i = 0; i < SL; i++) {
            const int l = r * SL + i;
            const float vv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
#pragma unroll
            for (int u = 0; u < WT; u++) {
              const uint64_t ww = readlane64(w[u], l);
              F[r][u] = ((ww >> lane) & 1) ? F[r][u] + vv : F[r][u] - vv;
            }
          }
        }
      }
    }
    if (PF && task + tr.stride < tr.end) stage(task + tr.stride);
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (r >= tk.nrows) break;
      const int64_t row = tk.row0 + r;
      uint64_t mine = 0;
#pragma unroll
      for (int u = 0; u < WT; u++) {
        sb_update<float, PROB_MAXCUT>(x[r][u], y[r][u], F[r][u], c0, c1, c2, 0.0f, 0.0f);
        st_state<NT>(b.X + row * lda + 64 * u + lane, x[r][u]);
        st_state<NT>(b.Y + row * lda + 64 * u + lane, y[r][u]);
        const uint64_t bits = __ballot(signbit(x[r][u]));
        if (lane == u) mine = bits;
      }
      if (lane < WT) Sn[row * W + lane] = mine;
    }
  }
}

template <int R, int WT, int NT, int PF>
__global__ __launch_bounds__(kBlock) void k_batch_step_bt(BArgs b) {
  batch_step_bt<R, WT, NT, PF>(b);
}

// Scalar loads of the task and instance records: the host writes the task
// table and k_batch_post the instance records, never a step kernel, so they
// may go through the scalar cache.  Scalar loads count on lgkmcnt, apart from
// the vector memory counter, so they never wait for earlier vector stores.
template <typename T>
__device__ __forceinline__ T ld_const(const T *p) {
  return *(const __attribute__((address_space(4))) T *)(p);
}

__device__ __forceinline__ BTask ld_task(const BTask *p) {
  BTask t;
  t.slot = ld_const(&p->slot);
  t.nrows = ld_const(&p->nrows);
  t.row0 = ld_const(&p->row0);
  t.kb = ld_const(&p->kb);
#pragma unroll
  for (int i = 0; i < 9; i++) t.off[i] = ld_const(&p->off[i]);
  t.pad = 0;
  return t;
}

// Pipelined bit-transpose step (tile bit 25).  gfx9 retires vector loads and
// stores in order on one counter, so in k_batch_step_bt every load issued
// after a task's stores -- the next task's row ranges, columns and sign words
// -- waits for those stores too: each task exposes the store latency plus a
// chain of four dependent round trips.  Here, for task k, in this order:
//   (1) task k's columns and weights (issued one task earlier) have arrived;
//       issue its sign-word loads;
//   (2) issue task k+1's state rows and pass-0 columns and weights, their
//       addresses from its record alone (entry offsets in BTask);
//   (3) issue task k-1's stores, held in registers since;
//   (4) wait for task k's sign words -- older than (2) and (3), so no wait
//       covers a store -- and compute task k; its results wait for (3) of
//       task k+1.
// One round trip (the sign words) per task is exposed.  Rows longer than one
// pass (64 / R slots) take their further passes synchronously.  Same
// arithmetic, same order: bit-identical to the other step kernels.
template <int R, int WT, int NT>
__device__ __forceinline__ void batch_step_pl(const BArgs &b) {
  constexpr int SL = 64 / R;  // neighbour slots per row and pass
  constexpr int kPlanes = 2;  // |v| <= 3
  const int lane = threadIdx.x & 63;
  const TaskRange tr = task_range(b);
  const int64_t lda = b.lda, W = WT;
  const int my_r = lane / SL, my_i = lane % SL;

  // The wait counts the compiler places are exact only if every path issues
  // the same vector memory operations, so staging and storing never branch:
  // rows past a task's end load and store (again) the task's first row,
  // slots past a row's end load entry 0 and are masked when used, a stage
  // past the last task reloads row 0, and the first task's "previous"
  // stores go to a sink.
  int ntask = tr.first;
  BTask ntk{};
  int32_t nt_ = 0;
  float nx[R][WT], ny[R][WT];
  int nkb_lo = 0, ndeg = 0;  // this lane's row: entries past the task's kb, degree
  int32_t ncol = 0;
  float nval = 0.0f;
  bool nok = false;
  auto stage = [&](int task) {
    for (; task < tr.end; task += tr.stride) {
      ntk = ld_task(b.tasks + task);
      if (ld_const(&b.inst[ntk.slot].status) == SLOT_ACTIVE) break;
    }
    ntask = task;
    const bool have = task < tr.end;
    if (have) nt_ = ld_const(&b.inst[ntk.slot].t);
    const int nr = have ? ntk.nrows : 1;
    const int64_t r0 = have ? ntk.row0 : 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t row = r0 + (r < nr ? r : 0);
#pragma unroll
      for (int u = 0; u < WT; u++) {
        nx[r][u] = ld_state<NT>(b.X + row * lda + 64 * u + lane);
        ny[r][u] = ld_state<NT>(b.Y + row * lda + 64 * u + lane);
      }
    }
    int o0 = 0, o1 = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      o0 = my_r == r ? ntk.off[r] : o0;
      o1 = my_r == r ? ntk.off[r + 1] : o1;
    }
    nkb_lo = o0;
    ndeg = have && my_r < ntk.nrows ? o1 - o0 : 0;
    nok = my_i < ndeg;
    const int64_t k = nok ? ntk.kb + o0 + my_i : 0;
    ncol = b.col[k];
    nval = b.val[k];
  };
  // task k-1's results, stored after task k+1's loads are issued; before the
  // first task, R distinct sink rows (no store may fold into another)
  float *pX = b.sink, *pY = b.sink + R * lda;
  uint64_t *pS = reinterpret_cast<uint64_t *>(b.sink + 2 * R * lda);
  int64_t pSW = 2 * W;  // sign-word row stride: 2 W in the sink, so no two stores merge
  int pnrows = R;
  float px[R][WT], py[R][WT];
  uint64_t pw[R];  // this lane's sign word (word lane % WT) of each row
#pragma unroll
  for (int r = 0; r < R; r++) {
    pw[r] = 0;
#pragma unroll
    for (int u = 0; u < WT; u++) px[r][u] = py[r][u] = 0.0f;
  }
  auto flush = [&]() {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t rr = r < pnrows ? r : 0;
#pragma unroll
      for (int u = 0; u < WT; u++) {
        st_state<NT>(pX + rr * lda + 64 * u + lane, r < pnrows ? px[r][u] : px[0][u]);
        st_state<NT>(pY + rr * lda + 64 * u + lane, r < pnrows ? py[r][u] : py[0][u]);
      }
      pS[rr * pSW + lane % WT] = r < pnrows ? pw[r] : pw[0];
    }
  };

  // compiler memory barriers keep the phases in program order (the loads and
  // stores touch different types, which alias analysis would let it swap)
  stage(tr.first);
  asm volatile("" ::: "memory");
  flush();  // to the sink: the loop is entered with as many stores in flight as from its end
  while (ntask < tr.end) {
    // task k: take over the staged record, state and pass-0 slots
    const BTask tk = ntk;
    const int task = ntask;
    const int32_t t = nt_;
    float x[R][WT], y[R][WT];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int u = 0; u < WT; u++) {
        x[r][u] = nx[r][u];
        y[r][u] = ny[r][u];
      }
    const int mykb_lo = nkb_lo, mydeg = ndeg;
    const int64_t j0 = nok ? (int64_t)ncol : tk.row0;
    const float v0 = nok ? nval : 0.0f;
    const uint64_t *S = (t & 1) ? b.S1 : b.S0;
    uint64_t *Sn = (t & 1) ? b.S0 : b.S1;
    // (1) pass-0 sign words
    uint64_t w0[WT];
#pragma unroll
    for (int u = 0; u < WT; u++) w0[u] = S[j0 * W + u];
    // (2) the next task's loads, (3) the previous task's stores
    asm volatile("" ::: "memory");
    stage(task + tr.stride);
    asm volatile("" ::: "memory");
    flush();
    // (4) task k
    const int32_t T = ld_const(&b.inst[tk.slot].T);
    const double dt = ld_const(&b.inst[tk.slot].dt), xi0 = ld_const(&b.inst[tk.slot].xi0);
    const int32_t intw = ld_const(&b.inst[tk.slot].intw);
    double p = 1.025 * (double)t / (double)T;
    p = p < 1.0 ? p : 1.0;
    const float c0 = (float)(dt * (p + 1)), c1 = (float)(dt * (p - 1));
    const float c2 = (float)(dt * xi0);
    int maxdeg = 0;
#pragma unroll
    for (int r = 0; r < R; r++)
      if (r < tk.nrows) maxdeg = max(maxdeg, tk.off[r + 1] - tk.off[r]);
    float F[R][WT];
    if (intw) {
      int Sa[R][WT], Vr[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        Vr[r] = 0;
#pragma unroll
        for (int u = 0; u < WT; u++) Sa[r][u] = 0;
      }
      for (int d0 = 0; d0 < maxdeg; d0 += SL) {
        float vf = v0;
        uint64_t w[WT];
#pragma unroll
        for (int u = 0; u < WT; u++) w[u] = w0[u];
        if (d0 > 0) {  // further passes: load this pass's slots now
          const bool ok = d0 + my_i < mydeg;
          const int64_t k = ok ? tk.kb + mykb_lo + d0 + my_i : 0;
          const int64_t j = ok ? (int64_t)b.col[k] : tk.row0;
          vf = ok ? b.val[k] : 0.0f;
#pragma unroll
          for (int u = 0; u < WT; u++) w[u] = S[j * W + u];
        }
        const int vi = (int)vf;
        const int av = vi < 0 ? -vi : vi;
        uint64_t Pm[kPlanes], Nm[kPlanes];
#pragma unroll
        for (int q = 0; q < kPlanes; q++) {
          Pm[q] = __ballot(vi > 0 && ((av >> q) & 1));
          Nm[q] = __ballot(vi < 0 && ((av >> q) & 1));
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
          const uint64_t rm = (SL == 64 ? ~0ull : ((1ull << SL) - 1)) << (r * SL);
#pragma unroll
          for (int q = 0; q < kPlanes; q++)
            Vr[r] += (__popcll(Pm[q] & rm) - __popcll(Nm[q] & rm)) * (1 << q);
        }
#pragma unroll
        for (int u = 0; u < WT; u++) {
          const uint64_t m = bit_transpose64(w[u], lane);
#pragma unroll
          for (int r = 0; r < R; r++) {
            const uint64_t rm = (SL == 64 ? ~0ull : ((1ull << SL) - 1)) << (r * SL);
            int acc = 0;
#pragma unroll
            for (int q = 0; q < kPlanes; q++)
              acc += (__popcll(m & Pm[q] & rm) - __popcll(m & Nm[q] & rm)) * (1 << q);
            Sa[r][u] += acc;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int u = 0; u < WT; u++) F[r][u] = (float)(2 * Sa[r][u] - Vr[r]);
    } else {
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int u = 0; u < WT; u++) F[r][u] = 0.0f;
      for (int d0 = 0; d0 < maxdeg; d0 += SL) {
        float vf = v0;
        uint64_t w[WT];
#pragma unroll
        for (int u = 0; u < WT; u++) w[u] = w0[u];
        if (d0 > 0) {
          const bool ok = d0 + my_i < mydeg;
          const int64_t k = ok ? tk.kb + mykb_lo + d0 + my_i : 0;
          const int64_t j = ok ? (int64_t)b.col[k] : tk.row0;
          vf = ok ? b.val[k] : 0.0f;
#pragma unroll
          for (int u = 0; u < WT; u++) w[u] = S[j * W + u];
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
          for (int i = 0; i < SL; i++) {
            const int l = r * SL + i;
            const float vv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vf), l));
#pragma unroll
            for (int u = 0; u < WT; u++) {
              const uint64_t ww = readlane64(w[u], l);
              F[r][u] = ((ww >> lane) & 1) ? F[r][u] + vv : F[r][u] - vv;
            }
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      uint64_t mine = 0;
#pragma unroll
      for (int u = 0; u < WT; u++) {
        sb_update<float, PROB_MAXCUT>(x[r][u], y[r][u], F[r][u], c0, c1, c2, 0.0f, 0.0f);
        px[r][u] = x[r][u];
        py[r][u] = y[r][u];
        const uint64_t bits = __ballot(signbit(x[r][u]));
        mine = lane % WT == u ? bits : mine;
      }
      pw[r] = mine;
    }
    pX = b.X + tk.row0 * lda;
    pY = b.Y + tk.row0 * lda;
    pS = Sn + tk.row0 * W;
    pSW = W;
    pnrows = tk.nrows;
  }
  flush();
}

template <int R, int WT, int NT>
__global__ __launch_bounds__(kBlock) void k_batch_step_pl(BArgs b) {
  batch_step_pl<R, WT, NT>(b);
}

// Same step as k_batch_step_bt with each instance's sign words staged in LDS.
// The task table holds segments of up to 512 << s rows of one instance (s from
// the tile, 0..3); a
// workgroup takes a segment, copies the instance's whole sign array (n x W
// words, coalesced) into LDS, and its wavefronts run the segment's rows R at a
// time with every neighbour sign word read from LDS, so the random gathers no
// longer miss L2.  Instances larger than the staging buffer read the words
// from global memory as k_batch_step_bt does.  Same arithmetic, same order:
// bit-identical.
constexpr int kSegRows = 512;     // rows per segment at segment code 0
constexpr int kLdsWords = 4096;   // staged sign words per workgroup (32 KB)

template <int R, int WT, int NT>
__global__ __launch_bounds__(kBlock) void k_batch_step_lds(BArgs b) {
  __shared__ uint64_t s_sign[kLdsWords];
  constexpr int SL = 64 / R;
  constexpr int kPlanes = 2;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  // segment range of the workgroup: the whole table, or its XCD's slice
  int first = blockIdx.x, end = *b.ntasks, stride = gridDim.x;
  if (b.xcd) {
    first = b.ntasks[1 + blockIdx.x % kXcds] + (int)(blockIdx.x / kXcds);
    end = b.ntasks[2 + blockIdx.x % kXcds];
    stride = gridDim.x / kXcds;
  }
  const int64_t lda = b.lda, W = WT;
  const int my_r = lane / SL, my_i = lane % SL;
  for (int seg = first; seg < end; seg += stride) {
    const BTask sg = b.tasks[seg];
    const BInst &in = b.inst[sg.slot];
    if (in.status != SLOT_ACTIVE) continue;  // the same for the whole workgroup
    const int32_t t = in.t;
    const uint64_t *S = (t & 1) ? b.S1 : b.S0;
    uint64_t *Sn = (t & 1) ? b.S0 : b.S1;
    const int64_t off = in.row_off, nw = (int64_t)in.n * W;
    const bool lds = nw <= kLdsWords;
    __syncthreads();  // the previous segment is done with the staged words
    if (lds)
      for (int64_t q = threadIdx.x; q < nw; q += blockDim.x) s_sign[q] = S[off * W + q];
    __syncthreads();
    double p = 1.025 * (double)t / (double)in.T;
    p = p < 1.0 ? p : 1.0;
    const float c0 = (float)(in.dt * (p + 1)), c1 = (float)(in.dt * (p - 1));
    const float c2 = (float)(in.dt * in.xi0);
    for (int64_t r0 = (int64_t)w * R; r0 < sg.nrows; r0 += (int64_t)kWaves * R) {
      const int64_t row0 = sg.row0 + r0;
      const int nrows = (int)(sg.nrows - r0 < R ? sg.nrows - r0 : R);
      float x[R][WT], y[R][WT];
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (r < nrows) {
#pragma unroll
          for (int u = 0; u < WT; u++) {
            x[r][u] = ld_state<NT>(b.X + (row0 + r) * lda + 64 * u + lane);
            y[r][u] = ld_state<NT>(b.Y + (row0 + r) * lda + 64 * u + lane);
          }
        }
      }
      const int64_t myrow = row0 + (my_r < nrows ? my_r : 0);
      const int64_t mykb = b.rbeg[myrow];
      const int mydeg = my_r < nrows ? (int)(b.rend[myrow] - mykb) : 0;
      int maxdeg = mydeg;
      for (int m = 32; m >= 1; m >>= 1) {
        const int o = __shfl_xor(maxdeg, m, 64);
        maxdeg = o > maxdeg ? o : maxdeg;
      }
      maxdeg = __builtin_amdgcn_readfirstlane(maxdeg);
      float F[R][WT];
      if (in.intw) {
        int Sa[R][WT], Vr[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
          Vr[r] = 0;
#pragma unroll
          for (int u = 0; u < WT; u++) Sa[r][u] = 0;
        }
        for (int d0 = 0; d0 < maxdeg; d0 += SL) {
          const bool ok = d0 + my_i < mydeg;
          const int64_t k = ok ? mykb + d0 + my_i : 0;
          const int64_t j = ok ? (int64_t)b.col[k] : row0;
          const int vi = ok ? (int)b.val[k] : 0;
          const int av = vi < 0 ? -vi : vi;
          uint64_t wd[WT];
#pragma unroll
          for (int u = 0; u < WT; u++) wd[u] = lds ? s_sign[(j - off) * W + u] : S[j * W + u];
          uint64_t Pm[kPlanes], Nm[kPlanes];
#pragma unroll
          for (int q = 0; q < kPlanes; q++) {
            Pm[q] = __ballot(vi > 0 && ((av >> q) & 1));
            Nm[q] = __ballot(vi < 0 && ((av >> q) & 1));
          }
#pragma unroll
          for (int r = 0; r < R; r++) {
            const uint64_t rm = (SL == 64 ? ~0ull : ((1ull << SL) - 1)) << (r * SL);
#pragma unroll
            for (int q = 0; q < kPlanes; q++)
              Vr[r] += (__popcll(Pm[q] & rm) - __popcll(Nm[q] & rm)) * (1 << q);
          }
#pragma unroll
          for (int u = 0; u < WT; u++) {
            const uint64_t m = bit_transpose64(wd[u], lane);
#pragma unroll
            for (int r = 0; r < R; r++) {
              const uint64_t rm = (SL == 64 ? ~0ull : ((1ull << SL) - 1)) << (r * SL);
              int acc = 0;
#pragma unroll
              for (int q = 0; q < kPlanes; q++)
                acc += (__popcll(m & Pm[q] & rm) - __popcll(m & Nm[q] & rm)) * (1 << q);
              Sa[r][u] += acc;
            }
          }
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
          for (int u = 0; u < WT; u++) F[r][u] = (float)(2 * Sa[r][u] - Vr[r]);
      } else {
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
          for (int u = 0; u < WT; u++) F[r][u] = 0.0f;
        for (int d0 = 0; d0 < maxdeg; d0 += SL) {
          const bool ok = d0 + my_i < mydeg;
          const int64_t k = ok ? mykb + d0 + my_i : 0;
          const int64_t j = ok ? (int64_t)b.col[k] : row0;
          const float v = ok ? b.val[k] : 0.0f;
          uint64_t wd[WT];
#pragma unroll
          for (int u = 0; u < WT; u++) wd[u] = lds ? s_sign[(j - off) * W + u] : S[j * W + u];
#pragma unroll
          for (int r = 0; r < R; r++) {
#pragma unroll
            for (int i = 0; i < SL; i++) {
              const int l = r * SL + i;
              const float vv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
#pragma unroll
              for (int u = 0; u < WT; u++) {
                const uint64_t ww = readlane64(wd[u], l);
                F[r][u] = ((ww >> lane) & 1) ? F[r][u] + vv : F[r][u] - vv;
              }
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (r >= nrows) break;
        const int64_t row = row0 + r;
        uint64_t mine = 0;
#pragma unroll
        for (int u = 0; u < WT; u++) {
          sb_update<float, PROB_MAXCUT>(x[r][u], y[r][u], F[r][u], c0, c1, c2, 0.0f, 0.0f);
          st_state<NT>(b.X + row * lda + 64 * u + lane, x[r][u]);
          st_state<NT>(b.Y + row * lda + 64 * u + lane, y[r][u]);
          const uint64_t bits = __ballot(signbit(x[r][u]));
          if (lane == u) mine = bits;
        }
        if (lane < WT) Sn[row * W + lane] = mine;
      }
    }
  }
}

// Resident chunks (tile bit 24): one 1024-thread workgroup per (instance,
// group of G agents) runs a whole chunk of K steps with the pairs' X and Y in
// registers and the instance's matrix and gather slots in LDS, as the
// single-problem resident kernel does.  Instances need integer weights in
// [-3, 3] and n G <= 8192 pairs; the F of dSB is then an exact integer sum,
// so it equals the sequential sums of the step kernels bit for bit, whichever
// thread computes a row.
// Set-up per task: the rows are counting-sorted by degree (keys capped at
// 255), so the 64 / G rows of a wavefront register slot have nearly equal
// degrees; the matrix is laid out as ELL groups of those rows, padded to the
// group's longest row in chunks of four entries (entry k of sorted row
// g R + l is u16 number 4 (gofs[g] + (k / 4) R + l) + k % 4, gofs in chunks).
// Gather slots (int8): q = row G + agent holds -sgn x of pair q (+1 where x has
// its sign bit set), q + n G its negation.  Entry: the slot of its column for
// agent 0, plus n G for a negative weight, | |v| << 14; padding is 0 (weight
// 0).  LDS bytes: histogram 1024 (after the set-up: the coefficients of 128
// steps), group offsets 528, sorted rows and degrees 2 al16(2 n), two slot
// buffers 2 al16(2 n G), then the ELL.  The grid is persistent: workgroups
// walk the task list (largest instances first).
struct BResTask {
  int32_t slot, g0, lg, pad;
};
constexpr int kBResBlock = 1024, kBResSlots = 8, kBResPairs = kBResBlock * kBResSlots;
constexpr int kBResLds = 160 * 1024, kBResHead = 1024 + 528;

__host__ __device__ constexpr int64_t bres_al16(int64_t v) { return (v + 15) / 16 * 16; }

__device__ __forceinline__ uint32_t bres_scan64(uint32_t v, int lane) {  // inclusive
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

__global__ __launch_bounds__(kBResBlock) void k_batch_resident(BArgs b, const BResTask *rt,
                                                               const int32_t *nrt, int32_t K) {
  extern __shared__ __align__(16) unsigned char s_raw[];
  constexpr int M = kBResSlots;
  const int ntask = *nrt, t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t lda = b.lda;
  uint32_t *hist = reinterpret_cast<uint32_t *>(s_raw);
  uint32_t *gofs = hist + 256;
  for (int task = blockIdx.x; task < ntask; task += gridDim.x) {
    const BResTask tk = rt[task];
    const BInst &in = b.inst[tk.slot];
    if (in.status != SLOT_ACTIVE) continue;  // the same for the whole workgroup
    const int n = in.n, lg = tk.lg, G = 1 << lg, nG = n << lg, R = 64 >> lg;
    const int ngr = (n + R - 1) / R;  // <= 128
    const int64_t r0 = in.row_off, g0 = tk.g0;
    const int64_t kbase = b.rbeg[r0];
    uint16_t *perm = reinterpret_cast<uint16_t *>(s_raw + kBResHead);
    uint16_t *degs = reinterpret_cast<uint16_t *>(s_raw + kBResHead + bres_al16(2 * (int64_t)n));
    int8_t *slots = reinterpret_cast<int8_t *>(s_raw + kBResHead + 2 * bres_al16(2 * (int64_t)n));
    const int64_t sbytes = bres_al16(2 * (int64_t)nG);
    uint2 *ell = reinterpret_cast<uint2 *>(reinterpret_cast<unsigned char *>(slots) + 2 * sbytes);
    // counting sort of the rows by degree.  Thread t owns rows t + 1024 j
    // (j < 8, as n <= 8192): their start and degree stay in registers, and
    // their sorted positions come back from the scatter.
    constexpr int RJ = kBResPairs / kBResBlock;
    uint32_t rb[RJ];
    int dg[RJ], ps[RJ];
#pragma unroll
    for (int j = 0; j < RJ; j++) {
      const int i = t + kBResBlock * j;
      rb[j] = 0;
      dg[j] = 0;
      if (i < n) {
        const int64_t kb0 = b.rbeg[r0 + i];
        rb[j] = (uint32_t)(kb0 - kbase);
        dg[j] = (int)(b.rend[r0 + i] - kb0);
      }
    }
    if (t < 256) hist[t] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RJ; j++) {
      const int i = t + kBResBlock * j;
      if (i < n) {
        degs[i] = (uint16_t)dg[j];
        atomicAdd(&hist[dg[j] < 255 ? dg[j] : 255], 1u);
      }
    }
    __syncthreads();
    if (t < 64) {
      uint32_t a[4], sum = 0;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        a[u] = hist[4 * t + u];
        sum += a[u];
      }
      uint32_t ex = bres_scan64(sum, lane) - sum;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        hist[4 * t + u] = ex;
        ex += a[u];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RJ; j++) {
      const int i = t + kBResBlock * j;
      ps[j] = 0;
      if (i < n) {
        ps[j] = (int)atomicAdd(&hist[dg[j] < 255 ? dg[j] : 255], 1u);
        perm[ps[j]] = (uint16_t)i;
      }
    }
    __syncthreads();
    // group widths in chunks of four, and their offsets (in uint2)
    if (t < 64) {
      uint32_t c[2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int g = 2 * t + u;
        uint32_t w = 0;
        if (g < ngr)
          for (int l = 0; l < R && g * R + l < n; l++) {
            const uint32_t d = degs[perm[g * R + l]];
            w = d > w ? d : w;
          }
        c[u] = (w + 3) / 4 * (uint32_t)R;
      }
      const uint32_t inc = bres_scan64(c[0] + c[1], lane);
      gofs[2 * t] = inc - c[0] - c[1];
      gofs[2 * t + 1] = inc - c[1];
      if (t == 63) gofs[128] = inc;
    }
    __syncthreads();
    // the ELL of the thread's own rows, eight entries' loads in flight
#pragma unroll
    for (int j = 0; j < RJ; j++) {
      if (t + kBResBlock * j >= n) continue;
      const int sr = ps[j], g = sr >> (6 - lg), l = sr & (R - 1), d = dg[j];
      const int nk = (int)(gofs[g + 1] - gofs[g]) / R * 4;
      uint16_t *dst = reinterpret_cast<uint16_t *>(ell + gofs[g] + l);
      const int32_t *cp = b.col + kbase + rb[j];
      const float *vp = b.val + kbase + rb[j];
      for (int k0 = 0; k0 < nk; k0 += 8) {
        int jc[8];
        float vv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const bool in = k0 + u < d;
          jc[u] = in ? cp[k0 + u] : 0;
          vv[u] = in ? vp[k0 + u] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int k = k0 + u;
          if (k < nk) {
            uint16_t e = 0;
            if (k < d)
              e = (uint16_t)((((uint32_t)(jc[u] - (int)r0) << lg) + (vv[u] < 0.0f ? (uint32_t)nG : 0u)) |
                             ((uint32_t)fabsf(vv[u]) << 14));
            dst[4 * (k >> 2) * R + (k & 3)] = e;
          }
        }
      }
    }
    // own pairs: register slot m of this wavefront is ELL group wv + 16 m
    float x[M], y[M];
    int qsm[M];  // gather slot of the pair: its row G + agent
    uint32_t gb[M], gw[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
      const int q = t + kBResBlock * m, grp = wv + 16 * m;
      x[m] = y[m] = 0.0f;
      qsm[m] = 0;
      const bool has = grp < ngr;
      gb[m] = __builtin_amdgcn_readfirstlane(has ? gofs[grp] : 0u);
      gw[m] = __builtin_amdgcn_readfirstlane(has ? (gofs[grp + 1] - gofs[grp]) / (uint32_t)R : 0u);
      if (q < nG) {
        const int row = perm[q >> lg];
        qsm[m] = (row << lg) + (q & (G - 1));
        const int64_t e = (r0 + row) * lda + g0 + (q & (G - 1));
        x[m] = b.X[e];
        y[m] = b.Y[e];
      }
    }
    auto put_slots = [&](int8_t *buf) {
#pragma unroll
      for (int m = 0; m < M; m++) {
        const int q = t + kBResBlock * m;
        if (q < nG) {
          const int8_t sg = signbit(x[m]) ? (int8_t)1 : (int8_t)-1;
          buf[qsm[m]] = sg;
          buf[qsm[m] + nG] = (int8_t)-sg;
        }
      }
    };
    put_slots(slots);
    __syncthreads();
    const int32_t t0 = in.t, T = in.T;
    const double dt = in.dt, xi0 = in.xi0;
    const float c2 = (float)(dt * xi0);
    const int rl = lane >> lg;
    // step coefficients, 128 steps at a time, in the histogram's LDS (free
    // after the set-up)
    float *coef = reinterpret_cast<float *>(hist);
    for (int st = 0; st < K; st++) {
      if ((st & 127) == 0) {
        if (t < 128 && st + t < K) {
          // src/kernels.cu:225-226, 258-262, evaluated as the host schedule does
          double p = 1.025 * (double)(t0 + st + t) / (double)T;
          p = p < 1.0 ? p : 1.0;
          coef[2 * t] = (float)(dt * (p + 1));
          coef[2 * t + 1] = (float)(dt * (p - 1));
        }
        __syncthreads();
      }
      const float c0 = coef[2 * (st & 127)], c1 = coef[2 * (st & 127) + 1];
      const int8_t *cur = slots + (st & 1) * sbytes;
#pragma unroll
      for (int m = 0; m < M; m++) {
        const int q = t + kBResBlock * m;
        if (q < nG) {
          const int8_t *src = cur + (q & (G - 1));
          const uint2 *e4 = ell + gb[m] + rl;
          // two chunks' entries and their eight slots loaded before the adds
          auto term4 = [&](const uint2 e, const int (&g)[4]) {
            return (int)((e.x >> 14) & 3) * g[0] + (int)(e.x >> 30) * g[1] +
                   (int)((e.y >> 14) & 3) * g[2] + (int)(e.y >> 30) * g[3];
          };
          auto load4 = [&](const uint2 e, int (&g)[4]) {
            g[0] = src[e.x & 0x3fff];
            g[1] = src[(e.x >> 16) & 0x3fff];
            g[2] = src[e.y & 0x3fff];
            g[3] = src[(e.y >> 16) & 0x3fff];
          };
          int Fi = 0;
          uint32_t c = 0;
          for (; c + 2 <= gw[m]; c += 2) {
            const uint2 e = e4[c * R], f = e4[(c + 1) * R];
            int ge[4], gf[4];
            load4(e, ge);
            load4(f, gf);
            Fi += term4(e, ge) + term4(f, gf);
          }
          if (c < gw[m]) {
            const uint2 e = e4[c * R];
            int ge[4];
            load4(e, ge);
            Fi += term4(e, ge);
          }
          sb_update<float, PROB_MAXCUT>(x[m], y[m], (float)Fi, c0, c1, c2, 0.0f, 0.0f);
        }
      }
      put_slots(slots + ((st + 1) & 1) * sbytes);
      __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < M; m++) {
      const int q = t + kBResBlock * m;
      if (q < nG) {
        const int64_t e = (r0 + (qsm[m] >> lg)) * lda + g0 + (q & (G - 1));
        b.X[e] = x[m];
        b.Y[e] = y[m];
      }
    }
  }
}

// After each step (K = 1), or after a resident chunk of K steps that ends at
// or before the instance's next window boundary: one block per slot advances
// the instance's step count and, at a window boundary, runs its window
// (src/kernels.cu:24-91, race-free) and decides whether the instance is
// finished.  With `pack` it first writes the sign words of X into the buffer
// of step t + K, as the step kernels do, for the window and the next step.
__global__ __launch_bounds__(kBlock) void k_batch_post(BArgs b, int32_t nslots, int32_t K,
                                                       int32_t pack) {
  __shared__ uint64_t s_mism[4], s_new[4];
  __shared__ int s_nbif;
  const int slot = blockIdx.x;
  if (slot >= nslots) return;
  BInst &in = b.inst[slot];
  if (in.status != SLOT_ACTIVE) return;
  const int32_t t1 = in.t + K;
  const int64_t W = b.W, r0 = in.row_off, n = in.n;
  if (pack) {
    uint64_t *Sn = (t1 & 1) ? b.S1 : b.S0;
    const int lane = threadIdx.x & 63;
    const int64_t nw = blockDim.x >> 6, lda = b.lda;
    // one wavefront per (row, word), four in flight
    for (int64_t p0 = (threadIdx.x >> 6) * 4; p0 < n * W; p0 += nw * 4) {
      float xv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t p = p0 + u;
        xv[u] = p < n * W ? b.X[(r0 + p / W) * lda + 64 * (p % W) + lane] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint64_t bits = __ballot(signbit(xv[u]));
        if (lane == 0 && p0 + u < n * W) Sn[r0 * W + p0 + u] = bits;
      }
    }
  }
  __syncthreads();
  if (t1 % in.P == 0) {
    if (threadIdx.x < 4) {
      s_mism[threadIdx.x] = 0;
      s_new[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) s_nbif = 0;
    __syncthreads();
    const uint64_t *Snew = (t1 & 1) ? b.S1 : b.S0;
    for (int64_t p = threadIdx.x; p < n * W; p += blockDim.x) {
      const int64_t idx = r0 * W + p;
      const uint64_t nw = Snew[idx];
      const uint64_t d = nw ^ b.cur[idx];
      b.cur[idx] = nw;
      if (d) atomicOr((unsigned long long *)&s_mism[p % W], (unsigned long long)d);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int64_t base = threadIdx.x & ~63; base < W * 64; base += blockDim.x) {
      const int64_t w = base >> 6, ag = base + lane;
      int now = 0, fresh = 0;
      if (ag < in.agents) {
        const int64_t ai = (int64_t)slot * W * 64 + ag;
        const int prev = b.bif[ai];
        int st = b.stab[ai];
        if (!prev) st = ((s_mism[w] >> lane) & 1) ? 0 : st + 1;
        now = st == in.K - 1;
        b.stab[ai] = st;
        b.bif[ai] = now;
        fresh = now && !prev;
      }
      const uint64_t nm = __ballot(fresh), bm = __ballot(now);
      if (lane == 0) {
        s_new[w] = nm;
        atomicAdd(&s_nbif, (int)__popcll(bm));
      }
    }
    __syncthreads();
    for (int64_t p = threadIdx.x; p < n * W; p += blockDim.x) {
      const uint64_t nm = s_new[p % W];
      if (nm) {
        const int64_t idx = r0 * W + p;
        b.fin[idx] = (b.fin[idx] & ~nm) | (b.cur[idx] & nm);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      in.nbif = s_nbif;
      if (t1 >= in.T || (in.uw && s_nbif == in.agents)) in.status = SLOT_DONE;
    }
  }
  if (threadIdx.x == 0) in.t = t1;
}

// Staged description of one admitted instance.
struct BAdmit {
  BInst inst;
  int32_t slot, pad;
  uint64_t seed;
  int64_t src_rp, src_nz, nnz_off;  // offsets into the staged row pointers / nonzeros
};

// Admission of every instance staged in one round, one block per instance:
// CSR copy into its row / nonzero ranges (columns rebased to global rows),
// counter-based initial state (as k_init_state), pre-step (as k_prestep),
// the sign words of the result, reset window state, then the slot record.
struct BCsr {
  int64_t *rbeg, *rend;
  int32_t *col;
  float *val;
  double *vald;
};

__global__ __launch_bounds__(kBlock) void k_batch_admit(BArgs b, BCsr c, const BAdmit *adm,
                                                        const int64_t *srp, const int32_t *scol,
                                                        const double *sval) {
  const BAdmit d = adm[blockIdx.x];
  const BInst &in = d.inst;
  const int64_t W = b.W, r0 = in.row_off, n = in.n;
  const int64_t z = srp[d.src_rp + n] - srp[d.src_rp];
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    c.rbeg[r0 + i] = d.nnz_off + srp[d.src_rp + i];
    c.rend[r0 + i] = d.nnz_off + srp[d.src_rp + i + 1];
  }
  for (int64_t k = threadIdx.x; k < z; k += blockDim.x) {
    c.col[d.nnz_off + k] = (int32_t)(r0 + scol[d.src_nz + k]);
    c.val[d.nnz_off + k] = (float)sval[d.src_nz + k];
    c.vald[d.nnz_off + k] = sval[d.src_nz + k];
  }
  const int lane = threadIdx.x & 63;
  const float cy = (float)(in.dt * (0.0 - 1)), cx = (float)(in.dt * (0.0 + 1));
  for (int64_t g = threadIdx.x >> 6; g < n * W; g += kWaves) {
    const int64_t i = g / W, w = g - i * W;
    const int64_t row = r0 + i, a = 64 * w + lane;
    float x = 0.0f, y = 0.0f;
    if (a < in.agents) {
      const uint64_t e = (uint64_t)(i * in.agents + a);
      x = (float)((d_uniform(d.seed, 2 * e) * 2 - 1) / 10);
      y = (float)((d_uniform(d.seed, 2 * e + 1) * 2 - 1) / 10);
    }
    y = fmaT(cy, x, y);
    x = fmaT(cx, y, x);
    b.X[row * b.lda + a] = x;
    b.Y[row * b.lda + a] = y;
    const uint64_t bits = __ballot(signbit(x));
    if (lane == 0) {
      b.S0[row * W + w] = bits;
      b.cur[row * W + w] = ~0ull;
      b.fin[row * W + w] = ~0ull;
    }
  }
  for (int64_t a = threadIdx.x; a < W * 64; a += blockDim.x) {
    b.stab[(int64_t)d.slot * W * 64 + a] = 0;
    b.bif[(int64_t)d.slot * W * 64 + a] = 0;
  }
  if (threadIdx.x == 0) b.inst[d.slot] = in;
}

// Per-agent cut of a finished instance's current spins, -0.5 (H - S), one
// block per listed slot.  Each wavefront takes groups of 8 rows; lane l loads
// neighbour slot l % 8 of row l / 8 (vector gathers) and every lane then
// accumulates each row's terms in CSR order from readlane broadcasts.
// Padding slots add fma(0, +-1, tmp) = tmp (tmp starts at +0 and is never
// -0).  Exact for integer weights.
__global__ __launch_bounds__(kBlock) void k_batch_cut(BArgs b, const int32_t *slots) {
  __shared__ double s_red[kWaves][64];
  const int32_t slot = slots[blockIdx.x];
  const BInst &in = b.inst[slot];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int64_t W = b.W, r0 = in.row_off, n = in.n;
  const int my_r = lane >> 3, my_i = lane & 7;
  for (int64_t w = 0; w < W; w++) {
    double acc = 0.0;
    for (int64_t g0 = 8 * q; g0 < n; g0 += 8 * kWaves) {
      const bool rv = g0 + my_r < n;
      const int64_t myrow = r0 + (rv ? g0 + my_r : 0);
      const int64_t kb = b.rbeg[myrow];
      const int mydeg = rv ? (int)(b.rend[myrow] - kb) : 0;
      int maxdeg = mydeg;
      for (int m = 32; m >= 1; m >>= 1) {
        const int o = __shfl_xor(maxdeg, m, 64);
        maxdeg = o > maxdeg ? o : maxdeg;
      }
      maxdeg = __builtin_amdgcn_readfirstlane(maxdeg);
      double tmp[8];
#pragma unroll
      for (int rr = 0; rr < 8; rr++) tmp[rr] = 0.0;
      for (int d0 = 0; d0 < maxdeg; d0 += 8) {
        const bool ok = d0 + my_i < mydeg;
        const int64_t k = ok ? kb + d0 + my_i : 0;
        const int64_t j = ok ? (int64_t)b.col[k] : r0;
        const double v = ok ? b.vald[k] : 0.0;
        const uint64_t word = b.cur[j * W + w];
#pragma unroll
        for (int rr = 0; rr < 8; rr++)
#pragma unroll
          for (int ii = 0; ii < 8; ii++) {
            const int l = rr * 8 + ii;
            const uint64_t vb = readlane64((uint64_t)__double_as_longlong(v), l);
            const uint64_t ww = readlane64(word, l);
            const double sj = ((ww >> lane) & 1) ? -1.0 : 1.0;
            tmp[rr] = __builtin_fma(__longlong_as_double((long long)vb), sj, tmp[rr]);
          }
      }
#pragma unroll
      for (int rr = 0; rr < 8; rr++) {
        if (g0 + rr < n) {
          const double si = ((b.cur[(r0 + g0 + rr) * W + w] >> lane) & 1) ? -1.0 : 1.0;
          acc += tmp[rr] * si * 0.5;
        }
      }
    }
    s_red[q][lane] = acc;
    __syncthreads();
    if (q == 0)
      b.cuts[(int64_t)slot * W * 64 + 64 * w + lane] =
          (-0.5) * ((((s_red[0][lane] + s_red[1][lane]) + s_red[2][lane]) + s_red[3][lane]) - in.S);
    __syncthreads();
  }
}

// Gathers retired slots' cuts and current spin words into contiguous
// staging areas (one block per slot) for a single copy to the host.
__global__ __launch_bounds__(kBlock) void k_batch_collect(BArgs b, const int32_t *slots,
                                                          const int64_t *woff, uint64_t *words,
                                                          double *cuts) {
  const int32_t slot = slots[blockIdx.x];
  const BInst &in = b.inst[slot];
  const int64_t W = b.W, L = W * 64;
  for (int64_t p = threadIdx.x; p < (int64_t)in.n * W; p += blockDim.x)
    words[woff[blockIdx.x] + p] = b.cur[in.row_off * W + p];
  for (int64_t a = threadIdx.x; a < L; a += blockDim.x)
    cuts[(int64_t)blockIdx.x * L + a] = b.cuts[(int64_t)slot * L + a];
}

}  // namespace sis

using namespace sis;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit range allocator with coalescing.
struct RangeAlloc {
  std::map<int64_t, int64_t> free_;  // start -> length
  explicit RangeAlloc(int64_t cap = 0) {
    if (cap > 0) free_[0] = cap;
  }
  int64_t take(int64_t len) {
    if (len == 0) return 0;
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= len) {
        const int64_t s = it->first, l = it->second;
        free_.erase(it);
        if (l > len) free_[s + len] = l - len;
        return s;
      }
    return -1;
  }
  void give(int64_t s, int64_t len) {
    if (len == 0) return;
    auto it = free_.emplace(s, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Request {
  int64_t ticket;
  int64_t n, T;
  std::vector<int64_t> rp;
  std::vector<int32_t> col;
  std::vector<double> val;
  double S, dt, xi0;
  int32_t agents, P, K, uw;
  uint64_t seed;
  double submit_s;
};

struct Live {
  Request req;
  int32_t slot;
  int64_t row_off, nnz_off;
  double admit_s;
  int32_t res_lg;  // log2 agents per resident-chunk workgroup, -1: not eligible
};

struct Result {
  int32_t state = 0;
  int32_t agents = 0;
  int64_t n = 0, steps = 0, best = -1;
  double best_cut = 0.0, submit_s = 0.0, admit_s = 0.0, done_s = 0.0;
  std::vector<double> cuts;
  std::vector<uint64_t> words;
};

}  // namespace

struct sis_batch {
  sis_batch_params p{};
  int64_t W = 1, lda = 64;
  hipStream_t stream = nullptr;
  DevBuf<float> X, Y;
  DevBuf<uint64_t> sink;  // 8 lda + 16 W words: the pipelined step's first stores
  DevBuf<uint64_t> S0, S1, cur, fin;
  DevBuf<int64_t> rbeg, rend;
  DevBuf<int32_t> col;
  DevBuf<float> val;
  DevBuf<double> vald;
  DevBuf<BInst> inst;
  DevBuf<BTask> tasks;
  DevBuf<int32_t> ntasks, stab, bif, cutslots;
  DevBuf<double> cuts;
  // admission / retirement staging (device side; grown on demand)
  DevBuf<BAdmit> st_adm;
  DevBuf<int64_t> st_rp, st_woff;
  DevBuf<int32_t> st_col;
  DevBuf<double> st_val, st_cuts;
  DevBuf<uint64_t> st_words;
  BInst *h_inst = nullptr;  // pinned mirror of the slot table
  RangeAlloc rows, nnz;
  std::vector<int32_t> free_slots;
  std::deque<Request> queue;
  std::map<int32_t, Live> live;
  std::unordered_map<int64_t, Result> results;
  int64_t next_ticket = 1;
  int32_t tile = 1;
  int32_t cus = 1;  // compute units of the device
  int32_t ntask_cap = 0;
  bool tasks_dirty = true;
  hipGraphExec_t graph = nullptr;
  int32_t graph_tile = -1;
  // resident chunks (tile bit 24): task list, and whether every live
  // instance qualifies (otherwise chunks run the streaming tile)
  DevBuf<BResTask> rtasks;
  DevBuf<int32_t> nrtasks;
  bool res_ok = false, graph_res = false;
  int32_t res_grid = 0;

  ~sis_batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (h_inst) (void)hipHostFree(h_inst);
    if (stream) (void)hipStreamDestroy(stream);
  }

  BArgs args() {
    BArgs a{};
    a.tasks = tasks.p;
    a.ntasks = ntasks.p;
    a.xcd = (tile >> 18) & 1;
    a.inst = inst.p;
    a.rbeg = rbeg.p;
    a.rend = rend.p;
    a.col = col.p;
    a.val = val.p;
    a.vald = vald.p;
    a.X = X.p;
    a.Y = Y.p;
    a.sink = reinterpret_cast<float *>(sink.p);
    a.S0 = S0.p;
    a.S1 = S1.p;
    a.cur = cur.p;
    a.fin = fin.p;
    a.stab = stab.p;
    a.bif = bif.p;
    a.cuts = cuts.p;
    a.lda = lda;
    a.W = W;
    return a;
  }
};

namespace {

// Tile: rows per task (1, 2, 4, 8) | non-temporal << 16 | vector gathers << 17
// (vector gathers need 2, 4 or 8 rows per task) | XCD-affine task order << 18 |
// bit-transpose integer sums << 19 (2, 4 or 8 rows, no vector-gather bit) |
// next-task prefetch << 20 (bit-transpose with non-temporal state) |
// sign words staged in LDS << 21 (2, 4 or 8 rows per pass, no other bit but
// non-temporal and XCD-affine) | segment code s << 22 (LDS staging: segments
// of 512 << s rows, s = 0..3) | resident chunks << 24 (with any streaming tile,
// which runs the chunks whenever a live instance does not qualify) |
// pipelined bit-transpose step << 25 (with the bit-transpose bit, 2, 4 or 8
// rows, no vector-gather, prefetch or LDS bit).
constexpr int kBatchResBit = 1 << 24;
constexpr int kBatchPipeBit = 1 << 25;

bool batch_tile_ok(int W, int tile) {
  tile &= ~kBatchResBit;
  const int pl = (tile & kBatchPipeBit) != 0;
  tile &= ~kBatchPipeBit;
  const int r = tile & 0xffff, nt = (tile >> 16) & 1, vg = (tile >> 17) & 1;
  const int bt = (tile >> 19) & 1, pf = (tile >> 20) & 1, ls = (tile >> 21) & 1;
  if ((tile >> 24) != 0 || (nt != 0 && nt != 1)) return false;
  if (!ls && (tile >> 22) != 0) return false;  // segment code only with LDS staging
  if (!(W == 1 || W == 2 || W == 4)) return false;
  if (pl) return bt && !vg && !pf && !ls && (r == 2 || r == 4 || r == 8);
  if (ls) return !vg && !bt && !pf && (r == 2 || r == 4 || r == 8);
  if (pf && !(bt && nt)) return false;  // prefetch: bit-transpose, non-temporal only
  if (bt) return !vg && (r == 2 || r == 4 || r == 8);
  if (vg) return r == 2 || r == 4 || r == 8;
  return r == 1 || r == 2 || r == 4 || r == 8;
}

// Persistent grid of a step kernel: as many workgroups as can be resident at
// once (occupancy of that kernel), a multiple of the XCD count.  A larger
// grid would run its last workgroups in a second round with a full share of
// the task table each.
template <typename K>
unsigned persistent_grid(const sis_batch *b, K kern) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int64_t g = (int64_t)per_cu * b->cus;
  return (unsigned)std::max<int64_t>(kXcds, g / kXcds * kXcds);
}

template <int R, int WT, int NT>
void launch_bstep(const sis_batch *b, const BArgs &a, bool launch) {
  static unsigned grid = 0;  // per kernel; the engine runs on one kind of device
  if (!grid) grid = persistent_grid(b, k_batch_step<R, WT, NT>);
  if (launch) hipLaunchKernelGGL((k_batch_step<R, WT, NT>), dim3(grid), dim3(kBlock), 0, b->stream, a);
}

template <int R, int WT, int NT>
void launch_bstep_vg(const sis_batch *b, const BArgs &a, bool launch) {
  static unsigned grid = 0;  // per kernel; the engine runs on one kind of device
  if (!grid) grid = persistent_grid(b, k_batch_step_vg<R, WT, NT>);
  if (launch) hipLaunchKernelGGL((k_batch_step_vg<R, WT, NT>), dim3(grid), dim3(kBlock), 0, b->stream, a);
}

template <int R, int WT, int NT, int PF>
void launch_bstep_bt(const sis_batch *b, const BArgs &a, bool launch) {
  static unsigned grid = 0;  // per kernel; the engine runs on one kind of device
  if (!grid) grid = persistent_grid(b, k_batch_step_bt<R, WT, NT, PF>);
  if (launch)
    hipLaunchKernelGGL((k_batch_step_bt<R, WT, NT, PF>), dim3(grid), dim3(kBlock), 0, b->stream, a);
}

template <int R, int WT, int NT>
void launch_bstep_pl(const sis_batch *b, const BArgs &a, bool launch) {
  static unsigned grid = 0;  // per kernel; the engine runs on one kind of device
  if (!grid) grid = persistent_grid(b, k_batch_step_pl<R, WT, NT>);
  if (launch)
    hipLaunchKernelGGL((k_batch_step_pl<R, WT, NT>), dim3(grid), dim3(kBlock), 0, b->stream, a);
}

template <int R, int WT, int NT>
void launch_bstep_lds(const sis_batch *b, const BArgs &a, bool launch) {
  static unsigned grid = 0;  // per kernel; the engine runs on one kind of device
  if (!grid) grid = persistent_grid(b, k_batch_step_lds<R, WT, NT>);
  if (launch)
    hipLaunchKernelGGL((k_batch_step_lds<R, WT, NT>), dim3(grid), dim3(kBlock), 0, b->stream, a);
}

// launch = false only resolves the kernel's grid (done before graph capture,
// where the occupancy query should not run).
void launch_batch_step(sis_batch *b, bool launch = true) {
  const BArgs a = b->args();
  const int r = b->tile & 0xffff, nt = (b->tile >> 16) & 1, vg = (b->tile >> 17) & 1;
  const int bt = (b->tile >> 19) & 1, pf = (b->tile >> 20) & 1, ls = (b->tile >> 21) & 1;
  const int pl = (b->tile >> 25) & 1;
  const int w = (int)b->W;
#define SIS_BSTEPP(R, WT, NT) \
  if (pl && r == R && w == WT && nt == NT) return launch_bstep_pl<R, WT, NT>(b, a, launch);
#define SIS_BSTEPP_W(R, NT) SIS_BSTEPP(R, 1, NT) SIS_BSTEPP(R, 2, NT) SIS_BSTEPP(R, 4, NT)
  SIS_BSTEPP_W(2, 0)
  SIS_BSTEPP_W(4, 0)
  SIS_BSTEPP_W(8, 0)
  SIS_BSTEPP_W(2, 1)
  SIS_BSTEPP_W(4, 1)
  SIS_BSTEPP_W(8, 1)
#undef SIS_BSTEPP_W
#undef SIS_BSTEPP
#define SIS_BSTEPL(R, WT, NT) \
  if (ls && r == R && w == WT && nt == NT) return launch_bstep_lds<R, WT, NT>(b, a, launch);
#define SIS_BSTEPL_W(R, NT) SIS_BSTEPL(R, 1, NT) SIS_BSTEPL(R, 2, NT) SIS_BSTEPL(R, 4, NT)
  SIS_BSTEPL_W(2, 0)
  SIS_BSTEPL_W(4, 0)
  SIS_BSTEPL_W(8, 0)
  SIS_BSTEPL_W(2, 1)
  SIS_BSTEPL_W(4, 1)
  SIS_BSTEPL_W(8, 1)
#undef SIS_BSTEPL_W
#undef SIS_BSTEPL
#define SIS_BSTEPB(R, WT, NT, PF) \
  if (bt && r == R && w == WT && nt == NT && pf == PF) return launch_bstep_bt<R, WT, NT, PF>(b, a, launch);
#define SIS_BSTEPB_W(R, NT, PF) SIS_BSTEPB(R, 1, NT, PF) SIS_BSTEPB(R, 2, NT, PF) SIS_BSTEPB(R, 4, NT, PF)
  SIS_BSTEPB_W(2, 0, 0)
  SIS_BSTEPB_W(4, 0, 0)
  SIS_BSTEPB_W(8, 0, 0)
  SIS_BSTEPB_W(2, 1, 0)
  SIS_BSTEPB_W(4, 1, 0)
  SIS_BSTEPB_W(8, 1, 0)
  SIS_BSTEPB_W(2, 1, 1)
  SIS_BSTEPB_W(4, 1, 1)
  SIS_BSTEPB_W(8, 1, 1)
#undef SIS_BSTEPB_W
#undef SIS_BSTEPB
#define SIS_BSTEP(R, WT, NT) \
  if (!vg && r == R && w == WT && nt == NT) return launch_bstep<R, WT, NT>(b, a, launch);
#define SIS_BSTEP_W(R, NT) SIS_BSTEP(R, 1, NT) SIS_BSTEP(R, 2, NT) SIS_BSTEP(R, 4, NT)
  SIS_BSTEP_W(1, 0)
  SIS_BSTEP_W(2, 0)
  SIS_BSTEP_W(4, 0)
  SIS_BSTEP_W(8, 0)
  SIS_BSTEP_W(1, 1)
  SIS_BSTEP_W(2, 1)
  SIS_BSTEP_W(4, 1)
  SIS_BSTEP_W(8, 1)
#undef SIS_BSTEP_W
#undef SIS_BSTEP
#define SIS_BSTEPV(R, WT, NT) \
  if (vg && r == R && w == WT && nt == NT) return launch_bstep_vg<R, WT, NT>(b, a, launch);
#define SIS_BSTEPV_W(R, NT) SIS_BSTEPV(R, 1, NT) SIS_BSTEPV(R, 2, NT) SIS_BSTEPV(R, 4, NT)
  SIS_BSTEPV_W(2, 0)
  SIS_BSTEPV_W(4, 0)
  SIS_BSTEPV_W(8, 0)
  SIS_BSTEPV_W(2, 1)
  SIS_BSTEPV_W(4, 1)
  SIS_BSTEPV_W(8, 1)
#undef SIS_BSTEPV_W
#undef SIS_BSTEPV
}

void launch_batch_post(sis_batch *b, int32_t K = 1, int32_t pack = 0) {
  hipLaunchKernelGGL(k_batch_post, dim3(b->p.max_instances), dim3(kBlock), 0, b->stream, b->args(),
                     b->p.max_instances, K, pack);
}

template <typename T>
int ensure_cap(DevBuf<T> &buf, size_t n) {
  if (buf.n >= n && buf.p) return SIS_OK;
  return buf.alloc(std::max<size_t>(n, 2 * buf.n));
}

// LDS bytes of one resident task (an upper bound of what the kernel lays
// out): rows sorted by degree with keys capped at 255, every row of degree
// 255 or more counted at the largest such degree (the kernel's order among
// them is arbitrary), groups of 64 / G rows padded to chunks of four.
int64_t bres_lds(const Request &q, int lg) {
  const int64_t n = q.n, R = 64 >> lg;
  std::vector<int64_t> d((size_t)n);
  int64_t dcap = 0;
  for (int64_t i = 0; i < n; i++) {
    d[(size_t)i] = q.rp[(size_t)i + 1] - q.rp[(size_t)i];
    if (d[(size_t)i] > 65535) return INT64_MAX;  // degrees are kept as uint16
    if (d[(size_t)i] >= 255) dcap = std::max(dcap, d[(size_t)i]);
  }
  for (auto &v : d)
    if (v >= 255) v = dcap;
  std::sort(d.begin(), d.end());
  int64_t chunks = 0;
  for (int64_t g = 0; g * R < n; g++) chunks += (d[(size_t)std::min(n, (g + 1) * R) - 1] + 3) / 4 * R;
  return kBResHead + 2 * bres_al16(2 * n) + 2 * bres_al16(2 * (n << lg)) + 8 * chunks;
}

// Agents per resident workgroup: the largest power of two G <= 64 with
// n G <= kBResPairs (-1: the instance does not qualify).
int bres_lg(const Live &l, int64_t chunk) {
  const Request &q = l.req;
  if (q.n < 1 || q.n > kBResPairs || q.P < 1 || q.P % chunk != 0) return -1;
  for (double v : q.val)
    if (!(v == std::floor(v) && v >= -3.0 && v <= 3.0)) return -1;
  int lg = 6;
  while (lg > 0 && (q.n << lg) > kBResPairs) lg--;
  if (bres_lds(q, lg) > kBResLds) return -1;
  return lg;
}

// Lets k_batch_resident use 160 KB of dynamic LDS (once per device 0..63;
// call outside graph capture).
void bres_attr() {
  static std::atomic<uint64_t> done{0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  const uint64_t bit = uint64_t(1) << dev;
  if (!(done.load() & bit)) {
    (void)hipFuncSetAttribute((const void *)k_batch_resident,
                              hipFuncAttributeMaxDynamicSharedMemorySize, kBResLds);
    done.fetch_or(bit);
  }
}

void launch_batch_resident(sis_batch *b, int32_t K) {
  bres_attr();
  hipLaunchKernelGGL(k_batch_resident, dim3((unsigned)b->res_grid), dim3(kBResBlock), kBResLds,
                     b->stream, b->args(), b->rtasks.p, b->nrtasks.p, K);
}

// One chunk: with resident chunks on and every live instance qualifying, one
// resident launch of chunk_steps steps and one post; else chunk_steps
// streaming steps, each followed by its post.
void enqueue_chunk(sis_batch *b) {
  if ((b->tile & kBatchResBit) && b->res_ok) {
    launch_batch_resident(b, b->p.chunk_steps);
    launch_batch_post(b, b->p.chunk_steps, 1);
    return;
  }
  for (int32_t k = 0; k < b->p.chunk_steps; k++) {
    launch_batch_step(b);
    launch_batch_post(b);
  }
}

// Task table.  Plain order: instance after instance.  XCD-affine order:
// instances are placed whole on the least-loaded XCD (largest first, load =
// state bytes + gathers); one too large for a fair share (over a sixteenth of
// the total) has its tasks dealt round-robin to all XCDs instead.  With
// LDS-staged sign words a task is a segment of up to kSegRows << s rows.
int rebuild_tasks(sis_batch *b) {
  if (!b->tasks_dirty) return SIS_OK;
  const int R = ((b->tile >> 21) & 1) ? kSegRows << ((b->tile >> 22) & 3) : (b->tile & 0xffff);
  const bool xcd = (b->tile >> 18) & 1;
  std::vector<std::vector<BTask>> per(xcd ? kXcds : 1);
  std::vector<const Live *> order;
  for (auto &kv : b->live) order.push_back(&kv.second);
  auto load = [&](const Live *l) {
    return (double)l->req.n * (double)b->lda * 16.0 + (double)l->req.col.size() * 8.0;
  };
  double total = 0.0;
  for (const Live *l : order) total += load(l);
  if (xcd)
    std::stable_sort(order.begin(), order.end(),
                     [&](const Live *x, const Live *y) { return load(x) > load(y); });
  std::vector<double> xl(per.size(), 0.0);
  for (const Live *l : order) {
    const bool spread = xcd && load(l) > total / (2.0 * kXcds);
    size_t dst = 0;
    if (xcd && !spread)
      dst = (size_t)(std::min_element(xl.begin(), xl.end()) - xl.begin());
    int64_t k = 0;
    for (int64_t r = 0; r < l->req.n; r += R, k++) {
      const size_t x = spread ? (size_t)(k % kXcds) : dst;
      BTask tk{};
      tk.slot = l->slot;
      tk.nrows = (int32_t)std::min<int64_t>(R, l->req.n - r);
      tk.row0 = l->row_off + r;
      const std::vector<int64_t> &rp = l->req.rp;
      tk.kb = l->nnz_off + rp[(size_t)r];
      if (tk.nrows <= 8)
        for (int i = 0; i <= tk.nrows; i++)
          tk.off[i] = (int32_t)(rp[(size_t)(r + i)] - rp[(size_t)r]);
      per[x].push_back(tk);
    }
    if (spread)
      for (auto &v : xl) v += load(l) / kXcds;
    else
      xl[dst] += load(l);
  }
  std::vector<BTask> t;
  int32_t hdr[2 + kXcds] = {0};
  for (size_t x = 0; x < per.size(); x++) {
    hdr[1 + x] = (int32_t)t.size();
    t.insert(t.end(), per[x].begin(), per[x].end());
  }
  for (size_t x = per.size(); x <= (size_t)kXcds; x++) hdr[1 + x] = (int32_t)t.size();
  hdr[0] = (int32_t)t.size();
  if ((int64_t)t.size() > b->ntask_cap) return fail(SIS_ERR_STATE, "batch task table overflow");
  if (!t.empty())
    HIPCHK(hipMemcpyAsync(b->tasks.p, t.data(), sizeof(BTask) * t.size(), hipMemcpyHostToDevice,
                          b->stream));
  HIPCHK(hipMemcpyAsync(b->ntasks.p, hdr, sizeof(hdr), hipMemcpyHostToDevice, b->stream));
  // resident chunks: (slot, first agent, log2 G) per workgroup, largest
  // instances first so the persistent grid ends evenly
  b->res_ok = false;
  if (b->tile & kBatchResBit) {
    std::vector<std::pair<int64_t, const Live *>> res;
    bool ok = !b->live.empty();
    for (auto &kv : b->live) {
      const int lg = kv.second.res_lg;
      if (lg < 0) {
        ok = false;
        break;
      }
      res.push_back({kv.second.req.n << lg, &kv.second});
    }
    if (ok) {
      std::stable_sort(res.begin(), res.end(),
                       [](const std::pair<int64_t, const Live *> &x,
                          const std::pair<int64_t, const Live *> &y) { return x.first > y.first; });
      std::vector<BResTask> rt;
      for (auto &pr : res) {
        const int lg = pr.second->res_lg;
        for (int64_t g0 = 0; g0 < b->lda; g0 += (int64_t)1 << lg)
          rt.push_back(BResTask{pr.second->slot, (int32_t)g0, lg, 0});
      }
      int st;
      if (b->rtasks.n < rt.size() && b->graph) {  // the captured launch holds the old list
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
      }
      if ((st = ensure_cap(b->rtasks, rt.size())) || (st = ensure_cap(b->nrtasks, 1))) return st;
      const int32_t nrt = (int32_t)rt.size();
      HIPCHK(hipMemcpyAsync(b->rtasks.p, rt.data(), sizeof(BResTask) * rt.size(),
                            hipMemcpyHostToDevice, b->stream));
      HIPCHK(hipMemcpyAsync(b->nrtasks.p, &nrt, sizeof(nrt), hipMemcpyHostToDevice, b->stream));
      b->res_grid = b->cus;  // one 1024-thread workgroup per CU (registers, LDS)
      b->res_ok = true;
    }
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  b->tasks_dirty = false;
  return SIS_OK;
}

// Copies the results of finished slots (one cut kernel, one gather kernel and
// one copy per array for the whole round) and frees their ranges.
int retire(sis_batch *b) {
  if (b->live.empty()) return SIS_OK;
  HIPCHK(hipMemcpyAsync(b->h_inst, b->inst.p, sizeof(BInst) * b->p.max_instances,
                        hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<int32_t> done;
  for (auto &kv : b->live)
    if (b->h_inst[kv.first].status == SLOT_DONE) done.push_back(kv.first);
  if (done.empty()) return SIS_OK;
  const double t_done = now_s();
  const int64_t L = b->W * 64;
  std::vector<int64_t> woff(done.size());
  int64_t nw = 0;
  for (size_t i = 0; i < done.size(); i++) {
    woff[i] = nw;
    nw += b->live[done[i]].req.n * b->W;
  }
  int st;
  if ((st = ensure_cap(b->st_woff, done.size())) || (st = ensure_cap(b->st_words, (size_t)nw)) ||
      (st = ensure_cap(b->st_cuts, done.size() * (size_t)L)))
    return st;
  HIPCHK(hipMemcpyAsync(b->cutslots.p, done.data(), sizeof(int32_t) * done.size(),
                        hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(b->st_woff.p, woff.data(), sizeof(int64_t) * done.size(),
                        hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(k_batch_cut, dim3((unsigned)done.size()), dim3(kBlock), 0, b->stream,
                     b->args(), b->cutslots.p);
  hipLaunchKernelGGL(k_batch_collect, dim3((unsigned)done.size()), dim3(kBlock), 0, b->stream,
                     b->args(), b->cutslots.p, b->st_woff.p, b->st_words.p, b->st_cuts.p);
  HIPCHK(hipGetLastError());
  std::vector<uint64_t> words((size_t)nw);
  std::vector<double> cuts(done.size() * (size_t)L);
  HIPCHK(hipMemcpyAsync(words.data(), b->st_words.p, sizeof(uint64_t) * words.size(),
                        hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipMemcpyAsync(cuts.data(), b->st_cuts.p, sizeof(double) * cuts.size(),
                        hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < done.size(); i++) {
    const int32_t slot = done[i];
    Live &l = b->live[slot];
    Result &r = b->results[l.req.ticket];
    r.state = 2;
    r.agents = l.req.agents;
    r.n = l.req.n;
    r.steps = b->h_inst[slot].t;
    r.submit_s = l.req.submit_s;
    r.admit_s = l.admit_s;
    r.done_s = t_done;
    r.cuts.assign(cuts.begin() + (int64_t)i * L, cuts.begin() + (int64_t)i * L + l.req.agents);
    r.words.assign(words.begin() + woff[i], words.begin() + woff[i] + l.req.n * b->W);
    r.best = 0;
    r.best_cut = r.cuts[0];
    for (int32_t a = 1; a < l.req.agents; a++)
      if (r.cuts[(size_t)a] > r.best_cut) {
        r.best_cut = r.cuts[(size_t)a];
        r.best = a;
      }
    b->h_inst[slot].status = SLOT_FREE;
    b->rows.give(l.row_off, l.req.n);
    b->nnz.give(l.nnz_off, (int64_t)l.req.col.size());
    b->free_slots.push_back(slot);
    b->live.erase(slot);
  }
  // Freed slot records (status only) back to the device in one copy.
  HIPCHK(hipMemcpyAsync(b->inst.p, b->h_inst, sizeof(BInst) * b->p.max_instances,
                        hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->tasks_dirty = true;
  return SIS_OK;
}

// Admits queued requests in order while the head fits: all of a round's CSR
// data goes to the device in one staged copy per array and one kernel
// installs and initialises every admitted instance.
int admit(sis_batch *b) {
  std::vector<BAdmit> adm;
  std::vector<int64_t> srp;
  std::vector<int32_t> scol;
  std::vector<double> sval;
  const double t_admit = now_s();
  while (!b->queue.empty() && !b->free_slots.empty()) {
    Request &q = b->queue.front();
    const int64_t z = (int64_t)q.col.size();
    const int64_t r0 = b->rows.take(q.n);
    if (r0 < 0) break;
    const int64_t z0 = b->nnz.take(z);
    if (z0 < 0) {
      b->rows.give(r0, q.n);
      break;
    }
    const int32_t slot = b->free_slots.back();
    b->free_slots.pop_back();
    BAdmit d{};
    BInst &in = d.inst;
    in.row_off = r0;
    in.n = (int32_t)q.n;
    in.agents = q.agents;
    in.t = 0;
    in.T = (int32_t)std::max<int64_t>(q.T, 1);
    in.P = q.P;
    in.K = q.K;
    in.uw = q.uw;
    in.status = q.T > 0 ? SLOT_ACTIVE : SLOT_DONE;
    in.dt = q.dt;
    in.xi0 = q.xi0;
    in.S = q.S;
    in.intw = 1;
    for (double v : q.val)
      if (!(v == std::floor(v) && v >= -3.0 && v <= 3.0)) {
        in.intw = 0;
        break;
      }
    d.slot = slot;
    d.seed = q.seed;
    d.src_rp = (int64_t)srp.size();
    d.src_nz = (int64_t)scol.size();
    d.nnz_off = z0;
    srp.insert(srp.end(), q.rp.begin(), q.rp.end());
    scol.insert(scol.end(), q.col.begin(), q.col.end());
    sval.insert(sval.end(), q.val.begin(), q.val.end());
    adm.push_back(d);
    b->h_inst[slot] = in;
    Live l;
    l.req = std::move(q);
    l.slot = slot;
    l.row_off = r0;
    l.nnz_off = z0;
    l.admit_s = t_admit;
    l.res_lg = bres_lg(l, b->p.chunk_steps);
    Result &res = b->results[l.req.ticket];
    res.state = 1;
    res.admit_s = t_admit;
    b->live.emplace(slot, std::move(l));
    b->queue.pop_front();
    b->tasks_dirty = true;
  }
  if (adm.empty()) return SIS_OK;
  int st;
  if ((st = ensure_cap(b->st_adm, adm.size())) || (st = ensure_cap(b->st_rp, srp.size())) ||
      (st = ensure_cap(b->st_col, std::max<size_t>(1, scol.size()))) ||
      (st = ensure_cap(b->st_val, std::max<size_t>(1, sval.size()))))
    return st;
  HIPCHK(hipMemcpyAsync(b->st_adm.p, adm.data(), sizeof(BAdmit) * adm.size(), hipMemcpyHostToDevice,
                        b->stream));
  HIPCHK(hipMemcpyAsync(b->st_rp.p, srp.data(), sizeof(int64_t) * srp.size(), hipMemcpyHostToDevice,
                        b->stream));
  if (!scol.empty()) {
    HIPCHK(hipMemcpyAsync(b->st_col.p, scol.data(), sizeof(int32_t) * scol.size(),
                          hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->st_val.p, sval.data(), sizeof(double) * sval.size(),
                          hipMemcpyHostToDevice, b->stream));
  }
  const BCsr c{b->rbeg.p, b->rend.p, b->col.p, b->val.p, b->vald.p};
  hipLaunchKernelGGL(k_batch_admit, dim3((unsigned)adm.size()), dim3(kBlock), 0, b->stream,
                     b->args(), c, b->st_adm.p, b->st_rp.p, b->st_col.p, b->st_val.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  return SIS_OK;
}

int build_graph(sis_batch *b) {
  if (b->graph && b->graph_tile == b->tile && b->graph_res == b->res_ok) return SIS_OK;
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  launch_batch_step(b, false);  // resolve the grid outside the capture
  bres_attr();
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  enqueue_chunk(b);
  HIPCHK(hipStreamEndCapture(b->stream, &g));
  hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(SIS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  b->graph_tile = b->tile;
  b->graph_res = b->res_ok;
  return SIS_OK;
}

int run_chunks(sis_batch *b, int64_t chunks) {
  if (b->p.eager) {
    for (int64_t c = 0; c < chunks; c++) enqueue_chunk(b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    return SIS_OK;
  }
  int st = build_graph(b);
  if (st) return st;
  for (int64_t c = 0; c < chunks; c++) HIPCHK(hipGraphLaunch(b->graph, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return SIS_OK;
}

}  // namespace

extern "C" {

int sis_batch_create(const sis_batch_params *pin, sis_batch **out) {
  if (!pin || !out) return fail(SIS_ERR_ARG, "sis_batch_create: null argument");
  *out = nullptr;
  sis_batch_params p = *pin;
  if (p.agents_max != 64 && p.agents_max != 128 && p.agents_max != 256)
    return fail(SIS_ERR_ARG, "sis_batch_create: agents_max must be 64, 128 or 256");
  if (p.max_instances < 1 || p.max_instances > (1 << 20) || p.max_rows < 1 || p.max_nnz < 0 ||
      p.max_rows >= INT32_MAX || p.max_nnz >= INT32_MAX)
    return fail(SIS_ERR_ARG, "sis_batch_create: bad capacity");
  if (p.chunk_steps < 1 || p.chunk_steps > 100000)
    return fail(SIS_ERR_ARG, "sis_batch_create: chunk_steps must be in [1, 100000]");
  std::unique_ptr<sis_batch> b(new (std::nothrow) sis_batch());
  if (!b) return fail(SIS_ERR_NOMEM, "out of host memory");
  b->W = p.agents_max / 64;
  b->lda = p.agents_max;
  // default: resident chunks while every instance qualifies, else
  // bit-transpose sums (other weights take its readlane path), 8 rows per
  // task, non-temporal state -- the autotuner's pick on mixed serving sets
  if (p.tile == 0) p.tile = 8 | (1 << 16) | (1 << 19) | kBatchResBit;
  if (!batch_tile_ok((int)b->W, p.tile)) return fail(SIS_ERR_ARG, "sis_batch_create: unsupported tile");
  b->tile = p.tile;
  b->p = p;
  b->rows = RangeAlloc(p.max_rows);
  b->nnz = RangeAlloc(p.max_nnz);
  for (int32_t s = p.max_instances - 1; s >= 0; s--) b->free_slots.push_back(s);
  b->ntask_cap = (int32_t)std::min<int64_t>(INT32_MAX - 1, p.max_rows);
  HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIPCHK(hipHostMalloc((void **)&b->h_inst, sizeof(BInst) * p.max_instances, hipHostMallocDefault));
  std::memset(b->h_inst, 0, sizeof(BInst) * p.max_instances);
  const int64_t R = p.max_rows, RW = R * b->W, L = b->W * 64 * (int64_t)p.max_instances;
  int st;
  if ((st = b->X.alloc((size_t)(R * b->lda))) || (st = b->Y.alloc((size_t)(R * b->lda))) ||
      (st = b->sink.alloc((size_t)(8 * b->lda + 16 * b->W))) ||
      (st = b->S0.alloc((size_t)RW)) || (st = b->S1.alloc((size_t)RW)) ||
      (st = b->cur.alloc((size_t)RW)) || (st = b->fin.alloc((size_t)RW)) ||
      (st = b->rbeg.alloc((size_t)R)) || (st = b->rend.alloc((size_t)R)) ||
      (st = b->col.alloc((size_t)std::max<int64_t>(1, p.max_nnz))) ||
      (st = b->val.alloc((size_t)std::max<int64_t>(1, p.max_nnz))) ||
      (st = b->vald.alloc((size_t)std::max<int64_t>(1, p.max_nnz))) ||
      (st = b->inst.alloc((size_t)p.max_instances)) || (st = b->tasks.alloc((size_t)b->ntask_cap)) ||
      (st = b->ntasks.alloc(2 + kXcds)) || (st = b->stab.alloc((size_t)L)) || (st = b->bif.alloc((size_t)L)) ||
      (st = b->cutslots.alloc((size_t)p.max_instances)) || (st = b->cuts.alloc((size_t)L)))
    return st;
  HIPCHK(hipMemset(b->inst.p, 0, sizeof(BInst) * p.max_instances));
  HIPCHK(hipMemset(b->ntasks.p, 0, sizeof(int32_t) * (2 + kXcds)));
  int dev = 0, cus = 256;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      prop.multiProcessorCount > 0)
    cus = prop.multiProcessorCount;
  b->cus = cus;
  *out = b.release();
  return SIS_OK;
}

int sis_batch_destroy(sis_batch *b) {
  if (!b) return SIS_OK;
  (void)hipStreamSynchronize(b->stream);
  delete b;
  return SIS_OK;
}

int sis_batch_submit(sis_batch *b, const sis_csr *J, const sis_params *p, int64_t max_steps,
                     int update_window, uint64_t seed, int64_t *ticket) {
  if (!b || !J || !p || !ticket) return fail(SIS_ERR_ARG, "sis_batch_submit: null argument");
  if (p->problem != SIS_MAXCUT || p->dynamics != SIS_DSB || !p->fused)
    return fail(SIS_ERR_ARG, "sis_batch_submit: batched serving runs fused MaxCut dSB");
  if (p->agents < 1 || p->agents > b->p.agents_max)
    return fail(SIS_ERR_ARG, "sis_batch_submit: agents exceed the engine's agents_max");
  if (J->n < 1 || J->n > b->p.max_rows || (int64_t)J->col.size() > b->p.max_nnz)
    return fail(SIS_ERR_ARG, "sis_batch_submit: instance larger than the engine capacity");
  if (max_steps < 0 || max_steps >= INT32_MAX) return fail(SIS_ERR_ARG, "sis_batch_submit: bad max_steps");
  Request q;
  q.ticket = b->next_ticket++;
  q.n = J->n;
  q.T = max_steps;
  q.rp = J->rp;
  q.col = J->col;
  q.val = J->val;
  q.S = J->sum;
  q.dt = p->dt;
  q.xi0 = p->xi0;
  if (q.xi0 <= 0.0) {
    int st = sis_maxcut_xi0(J, &q.xi0);
    if (st) return st;
  }
  q.agents = p->agents;
  q.P = p->window_period > 0 ? p->window_period : 100;
  q.K = p->window_threshold > 0 ? p->window_threshold : 25;
  q.uw = update_window ? 1 : 0;
  q.seed = seed;
  q.submit_s = now_s();
  Result &r = b->results[q.ticket];
  r.state = 0;
  r.submit_s = q.submit_s;
  r.agents = q.agents;
  r.n = q.n;
  *ticket = q.ticket;
  b->queue.push_back(std::move(q));
  return SIS_OK;
}

int sis_batch_advance(sis_batch *b, int64_t steps, int32_t *running, int32_t *queued) {
  if (!b || steps < 0) return fail(SIS_ERR_ARG, "sis_batch_advance: bad argument");
  int st;
  if ((st = retire(b)) || (st = admit(b)) || (st = retire(b)) || (st = rebuild_tasks(b))) return st;
  if (!b->live.empty() && steps > 0) {
    const int64_t chunks = (steps + b->p.chunk_steps - 1) / b->p.chunk_steps;
    if ((st = run_chunks(b, chunks))) return st;
  }
  if (running) *running = (int32_t)b->live.size();
  if (queued) *queued = (int32_t)b->queue.size();
  return SIS_OK;
}

int sis_batch_poll(sis_batch *b, int64_t ticket, sis_batch_status *s) {
  if (!b || !s) return fail(SIS_ERR_ARG, "sis_batch_poll: null argument");
  auto it = b->results.find(ticket);
  if (it == b->results.end()) return fail(SIS_ERR_ARG, "sis_batch_poll: unknown ticket");
  const Result &r = it->second;
  std::memset(s, 0, sizeof(*s));
  s->state = r.state;
  s->agents = r.agents;
  s->n = r.n;
  s->steps_executed = r.steps;
  s->best_agent = r.best;
  s->best_cut = r.best_cut;
  s->submit_s = r.submit_s;
  s->admit_s = r.admit_s;
  s->done_s = r.done_s;
  return SIS_OK;
}

int sis_batch_result(sis_batch *b, int64_t ticket, double *cuts, int32_t *spins) {
  if (!b) return fail(SIS_ERR_ARG, "sis_batch_result: null batch");
  auto it = b->results.find(ticket);
  if (it == b->results.end() || it->second.state != 2)
    return fail(SIS_ERR_STATE, "sis_batch_result: ticket not finished");
  const Result &r = it->second;
  if (cuts) std::memcpy(cuts, r.cuts.data(), sizeof(double) * r.cuts.size());
  if (spins)
    for (int64_t i = 0; i < r.n; i++)
      for (int64_t a = 0; a < r.agents; a++)
        spins[i * r.agents + a] = ((r.words[(size_t)(i * b->W + a / 64)] >> (a % 64)) & 1) ? -1 : 1;
  return SIS_OK;
}

int sis_batch_release(sis_batch *b, int64_t ticket) {
  if (!b) return fail(SIS_ERR_ARG, "sis_batch_release: null batch");
  auto it = b->results.find(ticket);
  if (it == b->results.end() || it->second.state != 2)
    return fail(SIS_ERR_STATE, "sis_batch_release: ticket not finished");
  b->results.erase(it);
  return SIS_OK;
}

int sis_batch_autotune(sis_batch *b, int64_t trial_steps, int32_t *tile, double *ms_per_step) {
  if (!b || trial_steps < 1 || trial_steps > 100000)
    return fail(SIS_ERR_ARG, "sis_batch_autotune: bad argument");
  int st;
  if ((st = retire(b)) || (st = admit(b)) || (st = rebuild_tasks(b))) return st;
  if (b->live.empty()) return fail(SIS_ERR_STATE, "sis_batch_autotune: no resident instance");
  // Snapshot of everything a step mutates.
  const int64_t R = b->p.max_rows, RW = R * b->W, L = b->W * 64 * (int64_t)b->p.max_instances;
  DevBuf<float> sx, sy;
  DevBuf<uint64_t> s0, s1, sc, sf;
  DevBuf<int32_t> sst, sbf;
  DevBuf<BInst> si;
  if ((st = sx.alloc((size_t)(R * b->lda))) || (st = sy.alloc((size_t)(R * b->lda))) ||
      (st = s0.alloc((size_t)RW)) || (st = s1.alloc((size_t)RW)) || (st = sc.alloc((size_t)RW)) ||
      (st = sf.alloc((size_t)RW)) || (st = sst.alloc((size_t)L)) || (st = sbf.alloc((size_t)L)) ||
      (st = si.alloc((size_t)b->p.max_instances)))
    return st;
  auto copy_all = [&](bool save) -> int {
    auto cp = [&](void *dst, void *src, size_t bytes) -> int {
      HIPCHK(hipMemcpyAsync(save ? dst : src, save ? src : dst, bytes, hipMemcpyDeviceToDevice,
                            b->stream));
      return SIS_OK;
    };
    int s2;
    if ((s2 = cp(sx.p, b->X.p, sizeof(float) * R * b->lda)) ||
        (s2 = cp(sy.p, b->Y.p, sizeof(float) * R * b->lda)) ||
        (s2 = cp(s0.p, b->S0.p, sizeof(uint64_t) * RW)) || (s2 = cp(s1.p, b->S1.p, sizeof(uint64_t) * RW)) ||
        (s2 = cp(sc.p, b->cur.p, sizeof(uint64_t) * RW)) || (s2 = cp(sf.p, b->fin.p, sizeof(uint64_t) * RW)) ||
        (s2 = cp(sst.p, b->stab.p, sizeof(int32_t) * L)) || (s2 = cp(sbf.p, b->bif.p, sizeof(int32_t) * L)) ||
        (s2 = cp(si.p, b->inst.p, sizeof(BInst) * b->p.max_instances)))
      return s2;
    HIPCHK(hipStreamSynchronize(b->stream));
    return SIS_OK;
  };
  if ((st = copy_all(true))) return st;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  double best_ms = 1e300;
  int32_t best = b->tile;
  const int32_t saved = b->tile;
  // LDS-staged signs pay only when the gathers outweigh copying each
  // instance's sign array per segment: tried from a mean degree of 32 up
  int64_t rows = 0, nnz = 0;
  for (auto &kv : b->live) {
    rows += kv.second.req.n;
    nnz += (int64_t)kv.second.req.col.size();
  }
  const bool try_lds = rows > 0 && nnz >= 32 * rows;
  std::vector<int32_t> cand;
  for (int r : {1, 2, 4, 8})
    for (int nt : {0, 1})
      for (int vg : {0, 1})
        for (int xc : {0, 1})
          for (int bt : {0, 1})
            for (int pf : {0, 1})
              for (int ls = 0; ls < (try_lds ? 2 : 1); ls++)
                for (int sg = 0; sg < (ls ? 4 : 1); sg++)
                  for (int pl : {0, 1}) {
                    const int32_t tl = r | (nt << 16) | (vg << 17) | (xc << 18) | (bt << 19) |
                                       (pf << 20) | (ls << 21) | (sg << 22) | (pl << 25);
                    if (batch_tile_ok((int)b->W, tl)) cand.push_back(tl);
                  }
  for (int32_t tl : cand) {
    {
      b->tile = tl;
      b->tasks_dirty = true;
      if ((st = copy_all(false)) || (st = rebuild_tasks(b))) break;
      launch_batch_step(b);
      launch_batch_post(b);
      (void)hipEventRecord(e0, b->stream);
      for (int64_t k = 0; k < trial_steps; k++) {
        launch_batch_step(b);
        launch_batch_post(b);
      }
      (void)hipEventRecord(e1, b->stream);
      if (hipStreamSynchronize(b->stream) != hipSuccess) {
        st = fail(SIS_ERR_HIP, "sis_batch_autotune: trial failed");
        break;
      }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms / trial_steps < best_ms) {
        best_ms = ms / trial_steps;
        best = b->tile;
      }
    }
  }
  // Resident chunks on top of the best streaming tile, timed over one chunk
  // (their matrix loads are paid once per chunk), when every instance qualifies.
  if (!st) {
    b->tile = best | kBatchResBit;
    b->tasks_dirty = true;
    if (!(st = copy_all(false)) && !(st = rebuild_tasks(b)) && b->res_ok) {
      enqueue_chunk(b);
      (void)hipEventRecord(e0, b->stream);
      enqueue_chunk(b);
      (void)hipEventRecord(e1, b->stream);
      if (hipStreamSynchronize(b->stream) != hipSuccess) {
        st = fail(SIS_ERR_HIP, "sis_batch_autotune: trial failed");
      } else {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms / b->p.chunk_steps < best_ms) {
          best_ms = ms / b->p.chunk_steps;
          best = b->tile;
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->tile = st ? saved : best;
  b->tasks_dirty = true;
  int st2 = copy_all(false);
  if (!st) st = st2;
  if (!st) st = rebuild_tasks(b);
  if (st) return st;
  if (tile) *tile = b->tile;
  if (ms_per_step) *ms_per_step = best_ms;
  return SIS_OK;
}

}  // extern "C"
