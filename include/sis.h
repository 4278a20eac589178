/*
 * sis.h -- C ABI of the MI355X simulated-bifurcation Ising solver.
 *
 * Drop-in boundary for the hot path of Oelnesr1/Sparse-Ising-Solver: the
 * reference's C++ classes CSR / CSR_GPU / TSP_CPU / TSP_GPU / Ising /
 * Ising_TSP (src/csr.cuh:26-227) become opaque handles plus plain-pointer
 * entry points.  Each declaration names the reference interface it replaces.
 *
 * Conventions
 *   - Every entry point returns an int status: SIS_OK (0) or an SIS_ERR_*
 *     code; nothing calls exit().  sis_last_error() gives the message of the
 *     last failure on the calling thread.
 *   - State arrays are node-major N x A (element (i, a) at i*A + a), the
 *     reference's layout (src/kernels.cu:115,130).
 *   - Handles are not shared between threads.  A solver owns one HIP stream
 *     on the device that was current when it was created.
 *   - Host-side entry points (readers, CSR, QUBO, generators, statistics)
 *     never touch the GPU.
 */
#ifndef SIS_H
#define SIS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ constants */

enum sis_status {
  SIS_OK = 0,
  SIS_ERR_ARG = 1,    /* invalid argument or size */
  SIS_ERR_IO = 2,     /* file cannot be opened or written */
  SIS_ERR_FORMAT = 3, /* malformed or unsupported input file */
  SIS_ERR_HIP = 4,    /* HIP runtime failure */
  SIS_ERR_NOMEM = 5,  /* host or device allocation failure */
  SIS_ERR_STATE = 6   /* call not valid in the handle's current state */
};

enum sis_problem { SIS_MAXCUT = 0, SIS_TSP = 1 };

/* dSB couples through sgn(x_j) (src/kernels.cu:251-254); bSB through
 * coupling * x_j (src/kernels.cu:582-585). */
enum sis_dynamics { SIS_DSB = 0, SIS_BSB = 1 };

/* State precision / gather form of the fused step. */
enum sis_form {
  SIS_FORM_F32 = 0,    /* plain: X, Y FP32, gathers read FP32 X */
  SIS_FORM_F64 = 1,    /* plain FP64, the reference's own precision */
  SIS_FORM_PACKED = 2, /* dSB only: gathers read 1 sign bit per agent */
  SIS_FORM_BF16 = 3,   /* bSB only: gathers read a BF16 shadow of X */
  SIS_FORM_Q12 = 4,    /* fused MaxCut dSB only (no reference counterpart; the
                          reference keeps X, Y in FP64, src/ising.cu:71-72):
                          X and Y stored as 12-bit fixed-point codes,
                            X: round(x 2^10) in [-1024, 1024], +-1025 for any
                               |x| > 1 (the next step confines it anyway),
                            Y: round(y 2^s), |code| <= 2047, s the largest
                               integer in [0, 11] with 2047 2^-s >= 2/dt + dt
                               (the bound of |y| while |x| <= 1; dt > 0, and
                               dt with 2047 < 2/dt + dt is refused),
                          rounded to nearest even after every step; arithmetic
                          in FP32 on the decoded values; gathers read 1 sign
                          bit per agent, so J sgn(X) is exact.  A code of 0 is
                          +0 (sign +1).  Agents are padded to a multiple of
                          256.  Not bit-identical to SIS_FORM_F32: its own
                          trajectory.  Per node-agent and step on a
                          degree-4 graph it moves 6 + 1/8 + 4/8 = 6.6 bytes
                          (state, sign write, sign gathers) against 32 for
                          F32. */
  SIS_FORM_Q8 = 5,     /* as SIS_FORM_Q12 with 8-bit codes (no reference
                          counterpart):
                            X: round(x 2^6) in [-64, 64], +-65 for any |x| > 1,
                            Y: round(y 2^s), |code| <= 127, s the largest
                               integer in [0, 7] with 127 2^-s >= 2/dt + dt
                               (dt with 127 < 2/dt + dt, i.e. dt < 0.0158, is
                               refused);
                          2 bytes per node-agent, so on a degree-4 graph
                          4 + 1/8 + 4/8 = 4.6 bytes per node-agent and step.
                          Its own trajectory; J sgn(X) exact. */
  SIS_FORM_Q10 = 6,    /* as SIS_FORM_Q12 with 10-bit codes (no reference
                          counterpart):
                            X: round(x 2^8) in [-256, 256], +-257 for any
                               |x| > 1,
                            Y: round(y 2^s), |code| <= 511, s the largest
                               integer in [0, 9] with 511 2^-s >= 2/dt + dt
                               (dt with 511 < 2/dt + dt, i.e. dt < 0.0039,
                               is refused);
                          2.5 bytes per node-agent (two planes of 32-bit and
                          one of 16-bit values per four agents), so on a
                          degree-4 graph 5 + 1/8 + 4/8 = 5.6 bytes per
                          node-agent and step.  Its own trajectory; J sgn(X)
                          exact. */
  SIS_FORM_Q8S = 7     /* as SIS_FORM_Q8 (8-bit codes, 2 bytes per node-agent)
                          with Y rounded stochastically after every step (no
                          reference counterpart): the Y code is
                          floor((w + b) / 256), w the nearest-even integer to
                          y 2^(s+8) and b a byte from a hash of (step index,
                          node, agent): y 2^s rounded down or up with the
                          probability of its fraction, so the rounding is
                          unbiased.  X, the initial state
                          and the pre-step round to nearest even as in
                          SIS_FORM_Q8.  Deterministic: the same seed and
                          steps give the same trajectory in any row order or
                          tile.  It runs within about 10% of SIS_FORM_Q8 and
                          holds plain FP32's cut quality on the graph
                          families measured, where nearest-even 8-bit
                          rounding falls short on some. */
};

enum sis_weights { SIS_W_ONE = 0, SIS_W_PM1 = 1, SIS_W_INT3 = 2 };
enum sis_tsp_kind { SIS_TSP_EUC_2D = 0, SIS_TSP_GEO = 1 };
enum sis_spins { SIS_SPINS_CURRENT = 0, SIS_SPINS_FINAL = 1 };

/* Message of the last failure on the calling thread ("" if none). */
const char *sis_last_error(void);

/* ------------------------------------------------------- host: MaxCut */

typedef struct sis_csr sis_csr;

/* Replaces CSR::CSR(string) (src/csr.cpp:3-109): MatrixMarket banner scan
 * (pattern / integer / real, symmetric / general), '%' comments, diagonal
 * dropped, sum of kept file values, symmetric entries mirrored, rows sorted
 * by (col, val).  Errors instead of the reference's endless banner loop on
 * files without symmetric/general; empty rows get correct row pointers. */
int sis_csr_read_mtx(const char *path, sis_csr **out);

/* Replaces CSR::construct_csr (src/csr.cpp:157-192) for caller-built
 * coordinate lists.  mirror != 0 appends (c, r, v) for every r != c entry
 * (src/csr.cpp:91-104); sum = sum of the given values. */
int sis_csr_from_coo(int64_t n, int64_t nent, const int32_t *rows, const int32_t *cols,
                     const double *vals, int mirror, sis_csr **out);

/* Replaces CSR::CSR(vector<double>, int) (src/csr.cpp:137-155): nonzeros of a
 * dense dim x dim row-major matrix, diagonal kept, sum = 0. */
int sis_csr_from_dense(int64_t dim, const double *dense, sis_csr **out);

/* Shape of a CSR: rows (CSR::size), stored entries (CSR::nonzeros) and the
 * sum of file values (CSR::sum, used by the cut formula). */
int sis_csr_shape(const sis_csr *c, int64_t *n, int64_t *nnz, double *sum);

/* Copies the arrays out (row_ptr has n+1 entries). */
int sis_csr_export(const sis_csr *c, int64_t *row_ptr, int32_t *col, double *val);

/* Writes a coordinate .mtx: "symmetric" with the strict upper triangle when
 * the matrix is symmetric, "general" otherwise; "integer" when every value is
 * integral.  Re-reading it reproduces the CSR and its sum. */
int sis_csr_write_mtx(const sis_csr *c, const char *path);

void sis_csr_free(sis_csr *c);

/* MaxCut coupling scale xi0 = 0.7 / (sigma sqrt(N)), sigma the sample std of
 * all N^2 entries of J (src/ising.cu:103-125), with N^2 in 64 bits. */
int sis_maxcut_xi0(const sis_csr *c, double *xi0);

/* Cut of one uniformly random +-1 assignment, 0.5 * sum of weights across it
 * (src/ising.cu:273-301, the README's `random` comparison). */
int sis_maxcut_random_cut(const sis_csr *c, uint64_t seed, double *cut);

/* Host statistics over per-agent results (src/ising.cu:238-265):
 * out = {min, max, mean, std (divisor A-1), fraction of agents at max}. */
int sis_cut_stats(const double *cut, int64_t agents, double out[5]);

/* ---------------------------------------------------------- host: TSP */

typedef struct sis_tsp sis_tsp;

/* Replaces TSP_CPU::read_file (src/tsp_cpu.cpp:9-45) and the distance part of
 * TSP_GPU::initialize (src/tsp_gpu.cu:10-74, kernels src/kernels.cu:404-442):
 * DIMENSION, EDGE_WEIGHT_TYPE (EUC_2D or GEO) and NODE_COORD_SECTION, packed
 * upper-triangle integer distances, per-city totals and the maximum.
 * euc_round = 0 truncates EUC_2D distances like the reference; 1 rounds to
 * nearest (TSPLIB nint).  Accepts "KEY : value" and indented node lines. */
int sis_tsp_read(const char *path, int euc_round, sis_tsp **out);

/* Same from coordinates (kind SIS_TSP_EUC_2D or SIS_TSP_GEO in radians). */
int sis_tsp_from_coords(int64_t n, const double *x, const double *y, int kind, int euc_round,
                        sis_tsp **out);

int sis_tsp_info(const sis_tsp *t, int64_t *ncities, double *max_distance, double *sum_distance);

/* EDGE_WEIGHT_TYPE of the instance (SIS_TSP_EUC_2D or SIS_TSP_GEO). */
int sis_tsp_kind_of(const sis_tsp *t, int *kind);

/* x, y (n each), packed distances (n(n-1)/2), per-city totals (n). */
int sis_tsp_export(const sis_tsp *t, double *x, double *y, double *dist, double *totals);

/* Replaces TSP_GPU::construct_matrix (src/tsp_gpu.cu:76-112) followed by
 * CSR(full, n^2): the n^2-spin one-hot QUBO written straight to CSR (no n^4
 * dense matrix), A = 1, B = C = max distance, zero entries skipped as the
 * dense scan does.  field (n^2 doubles, may be NULL) receives
 * h_(i,k) = -(A D_k + (n-2)(B+C)) / 2 (src/kernels.cu:591-592). */
int sis_tsp_qubo(const sis_tsp *t, sis_csr **J, double *field);

/* The reference's alternative TSP coupling scale (the optimal_xi0 = false
 * branch of Ising_TSP::reset_initialization, src/ising_tsp.cu:62-83; the
 * default branch is xi0 = 0.002): xi0 = 0.5 / (sigma sqrt N) with N = n^2 and
 * sigma = sqrt(sum of the squared stored QUBO entries / (N^2 - 1)) -- the
 * uncentred second moment (the reference computes the mean and drops it).
 * Evaluated in closed form from the distances, with N^2 in 64 bits. */
int sis_tsp_xi0(const sis_tsp *t, double *xi0);

/* Cyclic length of a tour (n city indices) from the packed table
 * (Ising_TSP::calculate_distance_of_tour, src/ising_tsp.cu:272-287). */
int sis_tsp_tour_length(const sis_tsp *t, const int32_t *tour, double *length);

/* Length of one random tour (Ising_TSP::generate_random,
 * src/ising_tsp.cu:255-270) with a seeded shuffle. */
int sis_tsp_random_length(const sis_tsp *t, uint64_t seed, double *length);

/* Writes a TSPLIB file with integer coordinates as read. */
int sis_tsp_write(const sis_tsp *t, const char *path);

void sis_tsp_free(sis_tsp *t);

/* ---------------------------------------------------- host: generators */
/* Seeded synthetic inputs (no reference counterpart; they stand in for the
 * Gset / TSPLIB files the reference's sweeps read, src/main.cu:31-51). */

/* Erdos-Renyi G(n, p), each pair i < j kept with probability p. */
int sis_gen_er(int64_t n, double p, int weights, uint64_t seed, sis_csr **out);
/* rows x cols torus, edges to the right and down neighbours (wrapping). */
int sis_gen_torus(int64_t rows, int64_t cols, int weights, uint64_t seed, sis_csr **out);
/* Random d-regular simple graph: configuration-model pairing with random
 * edge swaps that repair loops and multi-edges. */
int sis_gen_regular(int64_t n, int d, int weights, uint64_t seed, sis_csr **out);
/* Chung-Lu power-law graph: node of rank r (ranks given in a random order)
 * has expected degree proportional to (r + 1)^(-1 / (exponent - 1)), mean
 * about avg_degree; n avg_degree / 2 endpoint pairs are drawn, self-loops and
 * repeated pairs dropped.  Hub rows of such graphs exercise the row split. */
int sis_gen_chung_lu(int64_t n, double avg_degree, double exponent, int weights, uint64_t seed,
                     sis_csr **out);
/* n cities with integer coordinates uniform in [0, span]^2, EUC_2D. */
int sis_gen_tsp(int64_t n, int64_t span, int euc_round, uint64_t seed, sis_tsp **out);

/* ------------------------------------------------------- device solver */

typedef struct sis_params {
  int32_t problem;          /* SIS_MAXCUT or SIS_TSP */
  int32_t dynamics;         /* SIS_DSB or SIS_BSB */
  int32_t form;             /* SIS_FORM_* */
  int32_t agents;           /* A, independent replicas */
  double dt;                /* time step */
  double xi0;               /* coupling scale; <= 0 selects the reference's rule */
  double coupling;          /* bSB gather factor (reference TSP: 1.005) */
  int32_t window_period;    /* P; <= 0 selects 100 (MaxCut) or max(1, T/10) (TSP) */
  int32_t window_threshold; /* K: bifurcated after K-1 unchanged windows (25) */
  int32_t ncities;          /* TSP: n, with N = n^2 spins */
  int32_t use_graph;        /* 0: direct launches; 1: replay each window period as a
                               hipGraph; 2: capture the whole max_steps loop as one
                               hipGraph (up to ~40k launches, else as 1) */
  int32_t tile;             /* wavefront tile, 0 = default: agents per lane (1, 2, 4)
                               for the plain and BF16 forms; for the packed form
                               words (1, 2, 4, 8) | rows (1, 2, 4) << 8 |
                               non-temporal state (0, 1) << 16; or resident
                               1 << 20 | log2 G << 21 (one workgroup per G = 1..8
                               agents, state in registers / LDS for a window
                               period; fused MaxCut FP32, weights +-1,
                               N x G <= 12288); structured TSP: 1 VALU, 2 MFMA, 3
                               resident workgroups (FP32, n <= 96) */
  int32_t fused;            /* MaxCut: 1 fused step (src/kernels.cu:211-269), 0 the
                               four-kernel path (src/ising.cu:167-172); plain forms */
  int32_t split;            /* row split: rows with more than `split` stored entries
                               are summed by several wavefronts (chunks of `split`
                               entries, added in chunk order); 0 off, -1 chosen by
                               sis_solver_autotune.  Bit-identical to the unsplit sum
                               when every partial sum is exact (dSB with integer
                               weights); otherwise within FP rounding of it. */
  int32_t structured;       /* TSP from sis_solver_create_tsp: 1 evaluates J x from row
                               sums, column sums and the distance table instead of the
                               QUBO's CSR (n >= 3; f32 n <= 512, f64 n <= 256) */
  int32_t row_order;        /* storage order of the rows (no reference counterpart;
                               every result is the same in either order):
                               0 = chosen by the solver: a locality order for the
                               fused MaxCut dSB forms that gather packed sign words
                               (packed, 12-, 10- and 8-bit) on graphs whose sign words
                               exceed an L2 slice (N x agents / 8 > 4 MiB) and whose
                               weights are integers; 1 = the locality order whenever
                               those forms and integer weights allow it; -1 = node
                               order.  The locality order groups rows joined by
                               edges (multilevel greedy matching, pairs of pairs up
                               to 64 rows), so a wavefront's sign gathers share
                               cache lines; states, spins and results still come
                               and go in node order, each row keeps its entries in
                               their CSR order, and the seeded initial state is
                               keyed by node.  sis_solver_params reports 1 if the
                               solver stores rows in the locality order, else -1. */
} sis_params;

typedef struct sis_run_info {
  int64_t steps_executed; /* fused steps run (P * windows) */
  int64_t windows;        /* windows evaluated */
  int32_t all_bifurcated; /* MaxCut: every agent locked at the last window */
  int32_t reserved;
  double elapsed_us;      /* host clock over pre-step + loop (src/ising.cu:183-230) */
  double device_ms;       /* HIP events over the same region */
  double step_kernel_ms;  /* HIP events around the fused-step launches only */
} sis_run_info;

typedef struct sis_solver sis_solver;

/* Reference defaults for a problem: MaxCut dSB dt 0.1, 1024 agents, P 100,
 * K 25 (src/main.cu:268-274, src/csr.cuh:108-109); TSP bSB dt 0.1, 10000
 * agents, coupling 1.005, xi0 0.002, K 50 (src/csr.cuh:162-171,
 * src/ising_tsp.cu:61). */
int sis_params_default(int problem, sis_params *p);

/* Replaces CSR_GPU(CSR&) (src/csr_gpu.cu:3-15) + Ising(CSR&, CSR_GPU&, dt,
 * agents) / Ising_TSP(...) (src/ising.cu:42-83, src/ising_tsp.cu:5-48):
 * uploads J (and, for TSP, the field h), allocates the N x A state.  xi0 is
 * resolved here.  field may be NULL for MaxCut. */
int sis_solver_create(const sis_csr *J, const double *field, const sis_params *p,
                      sis_solver **out);

/* Replaces TSP_GPU(TSP_CPU&) + construct_matrix + CSR(full, n^2) + Ising_TSP
 * (src/main.cu:229-240, src/tsp_gpu.cu:76-112, src/ising_tsp.cu:5-48) in one
 * call: the solver for instance t, with p->ncities and the field set from t.
 * p->structured = 0 builds the direct-to-CSR QUBO (sis_tsp_qubo); 1 keeps no
 * matrix at all and takes J x from the QUBO's structure,
 *   (J g)(i,k) = C/4 sum_j g(j,k) + B/4 sum_l g(i,l)
 *              + sum_l (A d_kl / 8) (g(i-1,l) + g(i+1,l)),
 * which is exact algebra but adds in another order than the CSR rows (dSB:
 * identical results; bSB: within FP rounding).  Plain forms only. */
int sis_solver_create_tsp(const sis_tsp *t, const sis_params *p, sis_solver **out);

/* Replaces Ising::~Ising (src/ising.cu:303-339). */
int sis_solver_destroy(sis_solver *s);

/* Resolved parameters (xi0, period for a given T). */
int sis_solver_params(const sis_solver *s, sis_params *p);

/* Replaces the random part of reset_initialization (src/ising.cu:85-138,
 * src/ising_tsp.cu:50-109) with a counter-based stream computed on the
 * device: element e = i*A + a takes X from u(seed, 2e), Y from u(seed, 2e+1),
 * value (2u - 1) / 10 (TSP: X = 0).  Also resets spins/windows. */
int sis_solver_init_random(sis_solver *s, uint64_t seed);

/* Explicit initial state, N x A doubles each (converted to the form's
 * precision).  Also resets spins (-1), stability, locks and the step count. */
int sis_solver_set_state(sis_solver *s, const double *x, const double *y);

/* Current X and Y as N x A doubles. */
int sis_solver_get_state(sis_solver *s, double *x, double *y);

/* Replaces Ising::symplectic_update's loop (src/ising.cu:140-230,
 * src/ising_tsp.cu:111-168): pre-step, then {P fused steps, window} while
 * fewer than max_steps steps have run; with update_window (MaxCut) stop after
 * the window at which every agent has bifurcated.  info may be NULL. */
int sis_solver_run(sis_solver *s, int64_t max_steps, int update_window, sis_run_info *info);

/* Tile autotuner (no reference counterpart): times trial_steps fused steps
 * for every tile the form supports on this problem and keeps the fastest.
 * Resident tiles, which run a window period per launch, are timed over one
 * period (window_period, else 100 steps; at most 1000) when that is longer.
 * All tiles compute bit-identical results; the state is restored afterwards.
 * tile (may be NULL) receives the chosen tile, ms_per_step (may be NULL) its
 * time per step. */
int sis_solver_autotune(sis_solver *s, int64_t trial_steps, int32_t *tile, double *ms_per_step);

/* Pre-step alone (src/kernels.cu:8-22 at t = 0). */
int sis_solver_prestep(sis_solver *s);

/* nsteps fused steps at step indices t0 .. t0+nsteps-1 of horizon T, with no
 * pre-step and no windows (trajectory-parity primitive). */
int sis_solver_advance(sis_solver *s, int64_t t0, int64_t nsteps, int64_t T);

/* Spins as N x A int32 +-1: SIS_SPINS_CURRENT = sgn X at the last window
 * (current_spins), SIS_SPINS_FINAL = spins captured at lock (MaxCut) or at
 * the last valid window (TSP) (final_spins). */
int sis_solver_get_spins(sis_solver *s, int which, int32_t *out);

/* Replaces compute_max_cut + sum_max_cut (src/kernels.cu:360-400) as used at
 * src/ising.cu:232-236: per-agent cut -0.5 (H - S) of the current spins in
 * FP64 on the device, plus the best agent (lowest index among maxima).
 * cut may be NULL; best_agent may be NULL. */
int sis_solver_maxcut_cuts(sis_solver *s, double *cut, int64_t *best_agent);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the per-agent cut -0.5 (H - S) and the best agent of
 * sis_solver_maxcut_cuts, from caller-owned DEVICE arrays -- rp (n + 1 int64),
 * col / val (stored entries, int32 / double), spins (n x W uint64 words,
 * W = ceil(agents / 64), bit a % 64 of word i W + a / 64 set for spin -1: the
 * solver's layout), cut (agents doubles), best_agent (one int64, may be NULL).
 * Same kernels and summation order as sis_solver_maxcut_cuts, so the results
 * are identical.  The per-block partial sums live in a fixed device array of
 * SIS_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends; its address
 * is resolved on the first call on each device); shapes whose partial sums
 * exceed it (ceil(n / max(4, min(256, ceil(n / 1024)))) x 64 W doubles) fail
 * with SIS_ERR_ARG.  Enqueued on `stream` (a hipStream_t; NULL is the default
 * stream) without synchronizing.  Calls that overlap on one GPU share the
 * array and race.  Calls no hipMalloc / hipFree. */
#define SIS_NOALLOC_SCRATCH_BYTES (16u << 20)
int sis_maxcut_cuts_noalloc(const int64_t *rp, const int32_t *col, const double *val, int64_t n,
                            const uint64_t *spins, int64_t agents, double S, double *cut,
                            int64_t *best_agent, void *stream);

/* Per-agent Ising energy s^T J s - h.s of the chosen spins (FP64, device). */
int sis_solver_energy(sis_solver *s, int which, double *energy);

/* Per-agent bifurcation state after a run: stability counters and locks
 * (MaxCut), or valid_list (TSP: step of the last valid window, 0 if none). */
int sis_solver_agent_state(sis_solver *s, int32_t *stability, int32_t *locked, int32_t *valid_step);

/* Replaces the tour decode and length loop (src/ising_tsp.cu:172-251) on the
 * device: length of every agent's final-spin tour (-1 for agents without a
 * valid window), and the best (shortest) valid agent (-1 if none). */
int sis_solver_tsp_lengths(sis_solver *s, const sis_tsp *t, double *length, int64_t *best_agent);

/* ------------------------------------------------- batched serving (MaxCut) */
/* No reference counterpart: the reference solves one instance per process
 * (src/main.cu:214-227).  A batch engine keeps many small MaxCut dSB solves
 * resident in one concatenated CSR and one state array, runs all of them with
 * one persistent step kernel per step (replayed as a hipGraph per chunk of
 * steps), admits queued requests when rows / nonzeros / slots free up, and
 * retires each instance when its run ends (max_steps, or every agent locked
 * with update_window).  Each instance follows exactly the single-solver packed
 * dSB semantics (sis_solver_run with SIS_FORM_PACKED and sis_solver_init_random
 * with the same seed), so its results are bit-identical alone and in the mix. */

typedef struct sis_batch sis_batch;

typedef struct sis_batch_params {
  int32_t agents_max;    /* agents per instance <= agents_max; 64, 128 or 256 */
  int32_t max_instances; /* resident slots */
  int64_t max_rows;      /* capacity: sum of N over resident instances */
  int64_t max_nnz;       /* capacity: sum of stored nonzeros */
  int32_t chunk_steps;   /* steps per graph replay between admissions (>= 1) */
  int32_t tile;          /* rows per task (1, 2, 4, 8) | non-temporal << 16 |
                            vector gathers << 17 (2, 4 or 8 rows) | XCD-affine
                            task order << 18 (each instance's tasks on one XCD,
                            so its sign words stay in that XCD's L2) |
                            bit-transpose integer sums << 19 | next-task
                            prefetch << 20 | sign words staged in LDS << 21
                            (2, 4 or 8 rows) | LDS segment code s << 22
                            (segments of 512 << s rows, s = 0..3) | resident
                            chunks << 24 (one workgroup per instance and group
                            of agents runs a whole chunk with its state on
                            chip, while every resident instance has integer
                            weights in [-3, 3], N <= 8192, its matrix within
                            LDS and a window period that chunk_steps divides;
                            otherwise the chunk runs the streaming tile);
                            0 = default (resident chunks; 8 rows,
                            non-temporal, bit-transpose otherwise) */
  int32_t eager;         /* 1: launch each chunk's kernels directly instead of
                            replaying the captured hipGraph */
} sis_batch_params;

typedef struct sis_batch_status {
  int32_t state;           /* 0 queued, 1 running, 2 done */
  int32_t agents;
  int64_t n;               /* rows of the instance */
  int64_t steps_executed;
  int64_t best_agent;      /* lowest index among maximal cuts */
  double best_cut;
  double submit_s, admit_s, done_s; /* host monotonic clock, seconds */
} sis_batch_status;

int sis_batch_create(const sis_batch_params *p, sis_batch **out);
int sis_batch_destroy(sis_batch *b);

/* Queues one MaxCut solve.  p supplies agents (<= agents_max), dt, xi0 (<= 0:
 * 0.7 / (sigma sqrt N)), window period and threshold; J is copied.  The
 * instance starts from sis_solver_init_random's state for `seed`. */
int sis_batch_submit(sis_batch *b, const sis_csr *J, const sis_params *p, int64_t max_steps,
                     int update_window, uint64_t seed, int64_t *ticket);

/* Retires finished instances, admits queued ones that fit, then runs `steps`
 * steps (rounded up to whole chunks) of every resident instance.  running /
 * queued (may be NULL) receive the counts after the call. */
int sis_batch_advance(sis_batch *b, int64_t steps, int32_t *running, int32_t *queued);

/* State and timing of a ticket. */
int sis_batch_poll(sis_batch *b, int64_t ticket, sis_batch_status *st);

/* Per-agent cuts (agents doubles) and current spins (N x agents int32, may be
 * NULL) of a finished ticket. */
int sis_batch_result(sis_batch *b, int64_t ticket, double *cuts, int32_t *spins);

/* Forgets a finished ticket's stored result. */
int sis_batch_release(sis_batch *b, int64_t ticket);

/* Times each tile on the resident instances (state restored afterwards) and
 * keeps the fastest; all tiles are bit-identical.  Resident chunks, on top of
 * the fastest streaming tile, are timed over one chunk. */
int sis_batch_autotune(sis_batch *b, int64_t trial_steps, int32_t *tile, double *ms_per_step);

#ifdef __cplusplus
}
#endif

#endif /* SIS_H */
