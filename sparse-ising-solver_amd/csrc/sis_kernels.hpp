// sis_kernels.hpp -- gfx950 kernels of the simulated-bifurcation step, the
// update window and the cut / energy / tour evaluation.
//
// Layout in HBM (all node-major, the reference's layout, src/kernels.cu:115):
//   X, Y      N x lda     FP32 or FP64, lda = A rounded up to 64 (padding
//                         agents start at 0 and are never reported)
//   shadow    N x lda     BF16 copy of X (bSB BF16 form), double-buffered
//   signs     N x W       uint64 words, W = lda / 64; bit l of word w is
//                         signbit(X[i, 64 w + l]) -- a wavefront ballot
//   CSR       row_ptr int64 (N+1), col int32, val FP32/FP64 per form
// A wavefront owns one row and a tile of agents; CSR indices and values are
// wave-uniform (scalar loads), the agent axis is the lane axis, so every
// gather of a neighbour row is one coalesced load per wavefront.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sis {

// Device control block: step counter and loop flags.  Kernels read the
// counter instead of a per-launch argument so a captured window period can be
// replayed; only the window kernels advance it (after every step of the
// period has finished), which removes the reference's read/increment race
// (src/kernels.cu:225,232).
struct Ctrl {
  int64_t base;   // fused steps executed before the current window period
  int32_t done;   // 1: stop (all agents bifurcated with update_window)
  int32_t merge;  // 1: the last window locked new agents (copy final spins)
  int32_t nbif;   // agents bifurcated at the last window
  int32_t pad[3];
};

// PROB_UNFUSED is the second half of the reference's four-kernel MaxCut step
// (mat_mult + confine, src/kernels.cu:150-209); its first half is
// launch_xy_update (Update_XY, src/kernels.cu:8-22).
enum : int { PROB_MAXCUT = 0, PROB_TSP = 1, PROB_UNFUSED = 2 };
enum : int { G_DSB = 0, G_BSB = 1 };
enum : int { SRC_F32 = 0, SRC_F64 = 1, SRC_BITS = 2 };

constexpr int kBlock = 256;  // 4 wavefronts of 64 lanes
constexpr int kWaves = kBlock / 64;

// Launchers (defined in sis_kernels.hip).  All take the solver's stream.
// rowid: the node of each stored row (null: rows in node order); the seeded
// stream is keyed by node.
void launch_init_state(hipStream_t st, bool f64, void *X, void *Y, int64_t N, int64_t A,
                       int64_t lda, uint64_t seed, bool tsp, const int32_t *rowid);
void launch_prestep(hipStream_t st, bool f64, void *X, void *Y, int64_t n, double dt);
void launch_bf16_shadow(hipStream_t st, const float *X, uint16_t *sh, int64_t n);
void launch_pack_signs(hipStream_t st, int src, const void *X, uint64_t *S, int64_t N, int64_t W,
                       int64_t lda);

struct StepArgs {
  const int64_t *rp;
  const int32_t *col;
  const void *val;    // FP32 or FP64 per form
  const void *field;  // per-row field (TSP) or nullptr
  const void *gsrc;   // gather source: X snapshot / BF16 shadow / sign words
  void *gdst;         // next X / next shadow / next sign words
  void *X;            // own X (in place for BF16 and packed forms)
  void *Y;
  const void *sched;  // 4 coefficients per step, form precision
  const Ctrl *ctrl;
  int32_t offset;     // step index = ctrl->base + offset
  int64_t N, lda, W;
  double coupling;
  int64_t split;      // rows with more stored entries are left to k_step_split (0: none)
  int32_t qshift;     // narrow-state form: Y codes are round(y 2^qshift)
  const int32_t *rowid;  // node of each stored row (stochastic rounding), or nullptr
};

// Row split: a row with more than `chunk` entries is cut into groups of
// kWaves chunks; one workgroup per (group, 64-agent tile).  Each wavefront
// sums one chunk in CSR order, the workgroup adds its chunk sums in order
// through LDS, and a row of one group is finished by that workgroup.  For a
// row of several groups every workgroup stores its group sum and the last to
// arrive (per-row counter) adds the group sums in group order and finishes the
// row, so the result does not depend on scheduling.
struct SplitEnt {
  int64_t row, kb, ke;  // row and its entries [kb, ke) for this group
  int64_t pbase;        // first group-sum slot of the row (rows of >1 group)
  int32_t hr, g, ng;    // heavy-row index, group index, groups of the row
  int32_t pad;
};

struct SplitArgs {
  const SplitEnt *ent;
  void *part;           // group sums, (slot, lda) in the form's precision
  int32_t *cnt;         // arrivals per (heavy row, tile); the last resets it
  int64_t chunk;        // entries per wavefront
};

// Structured TSP product (n >= 3): with g the gathered value of spin
// (position i, city k), the one-hot QUBO gives
//   (J g)(i,k) = C/4 colsum_k + B/4 rowsum_i + sum_l Dq_kl (g(i-1,l) + g(i+1,l))
// with Dq_kl = A d_kl / 8 (the (B+C)/4 diagonal cancels the i = j and k = l
// terms), so no CSR is needed.  rsum / csum are n x lda per step.
struct TspArgs {
  const void *dq;  // n x n, A d_kl / 8, form precision
  // FP32 A d_kl / 8 zero-padded to (n rounded up to 32) x nep rows / columns
  // (nep = n rounded up to 8), for the matrix-core kernel
  const float *dqm;
  void *rsum, *csum;
  double qB, qC;   // B/4, C/4
  int64_t n, nep;
};

bool tsp_struct_supported(int form, int gather);
// mfma: the FP32 form's distance product on the matrix cores
// (v_mfma_f32_32x32x2f32) instead of VALU FMAs.
void launch_tsp_struct_step(hipStream_t st, int form, int gather, bool mfma, const StepArgs &a,
                            const TspArgs &ts);
// Energy s^T J s - h.s of bit-packed spins with the structured product, FP64,
// as per-position partials part[i][agent] (reduced by launch_cut_reduce).
void launch_tsp_struct_energy(hipStream_t st, const uint64_t *spins, const double *field,
                              const TspArgs &ts, int64_t W, int64_t lda, double *part);

// Resident small problems (k_resident): one 1024-thread workgroup per G
// agents runs `steps` steps with X, Y in registers and the matrix and gather
// source in LDS.  MaxCut fused step, FP32, weights +-1, N x G <= 12288.
// The gather source has 2 N G + 2 G slots: the terms of a +1 and a -1 weight
// on pair q = row G + agent at 2q and 2q + 1 (dSB, int8 -sgn x) or at q and
// q + N G (bSB, float -(x * coupling)), zeros from 2 N G.  The matrix is in ELL groups
// of R = 64 / G rows (the rows of one wavefront register slot), padded to
// whole chunks of four entries: entry k of (sorted) row g R + l is uint16
// number gofs[g] + 4 ((k / 4) R + l) + k % 4, the gather slot of its column
// and weight sign for agent 0 (2 N G past the row's end).  With perm the rows
// are in degree order (sorted row s is row perm[s]), so a group pads to
// nearly equal degrees.  LDS: the ell_words entries, then two gather-slot
// buffers of snap_bytes at snap_off; with ell_cw > 0 (dSB, groups at most
// ell_cw chunks wide) the ELL stays in registers and snap_off is 0.
struct ResArgs {
  const float *Xin;  // X at the first step (plain: the snapshot buffer)
  float *Xout;       // X after the last step
  const uint16_t *ell;
  const uint32_t *gofs;  // ngroups + 1 entry offsets
  const uint16_t *perm;  // sorted index -> row, or null (natural order)
  int32_t ngroups, ell_words;
  int32_t G, lgG;    // agents per workgroup, log2 G
  int32_t steps;
  int32_t snap_off, snap_bytes;
  int32_t ell_cw;  // > 0: dSB with the ELL in registers, groups at most ell_cw chunks wide
};
int resident_pairs_max();         // N x G limit of one workgroup
int resident_sorted_pairs_max();  // the same with rows in degree order (perm)
// false: no kernel instance for these sizes (nothing launched)
bool launch_resident(hipStream_t st, int gather, const StepArgs &a, const ResArgs &r);

// Resident structured TSP (FP32, n >= 3 up to 48 output tiles and 160 KB of
// LDS, G = 1, 2 or 4 agents per workgroup): `steps` steps per launch with the elements' X / Y in registers
// and g, the distance table and the sums in LDS.
size_t tsp_resident_lds(int64_t n, int G, int64_t nep);
bool tsp_resident_ok(int64_t n, int G, int64_t nep);
void launch_tsp_resident(hipStream_t st, int gather, const StepArgs &a, const TspArgs &ts,
                         const ResArgs &r);

bool split_supported(int form, int prob, int gather);
void launch_step_split(hipStream_t st, int form, int prob, int gather, int64_t nent,
                       const StepArgs &a, const SplitArgs &sp);

// Narrow-state dSB forms (X, Y as qw-bit codes, qw = 12, 10 or 8,
// sis_device.hpp): 3, 2.5 or 2 dword planes of N x (lda / 256) x 64 lanes
// (the 10-bit form's third plane holds 16-bit values); lda is a multiple of
// 256.
void launch_init_qn(hipStream_t st, int qw, uint32_t *P, int64_t N, int64_t A, int64_t lda,
                    uint64_t seed, int qshift, const int32_t *rowid);
void launch_prestep_qn(hipStream_t st, int qw, uint32_t *P, int64_t N, int64_t lda, double dt,
                       int qshift);
void launch_pack_signs_qn(hipStream_t st, int qw, const uint32_t *P, uint64_t *S, int64_t N,
                          int64_t lda);

// form: 0 f32, 1 f64, 2 packed, 3 bf16, 4 narrow (12-bit codes), 5 narrow
// (8-bit codes), 6 narrow (10-bit codes), 7 narrow (8-bit codes, Y rounded
// stochastically).  vec = agents per lane (f32/f64/bf16); for the packed form
// vec = words per wavefront | rows per wavefront << 8 | non-temporal state
// access << 16; for the narrow forms 4 | rows per wavefront << 8 |
// non-temporal << 16 | staged gathers << 17.
bool step_supported(int form, int prob, int gather, int vec);
void launch_step(hipStream_t st, int form, int prob, int gather, int vec, const StepArgs &a);
// Update_XY at the step index ctrl->base + offset: y += c1 x; x += c0 y.
void launch_xy_update(hipStream_t st, bool f64, void *X, void *Y, int64_t n, const void *sched,
                      const Ctrl *ctrl, int32_t offset);

void launch_window_scan(hipStream_t st, int src, const void *xsrc, uint64_t *cur, uint64_t *part,
                        int64_t N, int64_t W, int64_t lda, int64_t nblk, int64_t rows_per_blk,
                        const Ctrl *ctrl);
void launch_window_finalize(hipStream_t st, const uint64_t *part, int64_t nblk, uint64_t *newmask,
                            int32_t *stab, int32_t *bif, int64_t A, int64_t W, int K, int uw,
                            int64_t P, Ctrl *ctrl);
void launch_window_merge(hipStream_t st, const uint64_t *cur, uint64_t *fin,
                         const uint64_t *newmask, int64_t N, int64_t W, const Ctrl *ctrl);
// lines: 2 n x W words of scratch (one mask per position / city line).
void launch_tsp_valid(hipStream_t st, const uint64_t *cur, int64_t n, int64_t W, int64_t A,
                      int32_t *valid_list, uint64_t *newmask, uint64_t *lines, int64_t P,
                      Ctrl *ctrl);
void launch_ctrl_advance(hipStream_t st, Ctrl *ctrl, int64_t P);

// Cut / energy: partial row sums per block, then an ordered reduction.
void launch_cut_rows(hipStream_t st, const int64_t *rp, const int32_t *col, const double *val,
                     const double *field, const uint64_t *spins, int64_t N, int64_t W,
                     int64_t lda, int64_t nblk, int64_t rows_per_blk, double *part);
void launch_cut_reduce(hipStream_t st, const double *part, int64_t nblk, int64_t A, int64_t lda,
                       int mode, double S, double *out);
void launch_tsp_len(hipStream_t st, const uint64_t *fin, const int32_t *valid_list, int64_t n,
                    int64_t W, int64_t A, const double *dist, double *len);
void launch_best(hipStream_t st, const double *v, int64_t A, int mode, int64_t *best);

}  // namespace sis
