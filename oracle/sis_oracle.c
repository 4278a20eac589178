/*
 * sis_oracle.c -- CPU restatement of the reference simulated-bifurcation solver.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for cpu_baseline, as the timed CPU port).  The product never links it.
 *
 * Semantics follow the reference (Oelnesr1/Sparse-Ising-Solver) with its data
 * races resolved deterministically:
 *   - every neighbour read inside a step uses the start-of-step snapshot of X
 *     (Jacobi; the reference reads X while other threads write it,
 *     src/kernels.cu:251 vs :261 and :582 vs :594);
 *   - the step index t is the number of fused steps already executed
 *     (src/kernels.cu:225,232 increment it while other blocks read it);
 *   - the window compares against the previous window's spins
 *     (src/kernels.cu:43 vs :89) and stability starts at zero
 *     (src/ising.cu:127-131 never clears it).
 * Every multiply-add that the reference writes as `a += b * c` is an explicit
 * fused multiply-add here (nvcc contracts those by default); nothing else is
 * fused (build with -ffp-contract=off).
 *
 * Parity pins (tests/test_oracle_*.py): SpMV part against scipy.sparse,
 * J = 0 symplectic area invariant and the p = 1 linear closed form, dt = 0
 * identity and window lock at window 24/25, xi0 against numpy's sample std of
 * the dense matrix, cuts against brute force and direct edge sums, TSP
 * distances against closed forms, the TSP QUBO against the valid-tour energy
 * identity and brute force over all tours.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#define ORC_MAXCUT 0
#define ORC_TSP 1
#define ORC_UNFUSED 2 /* MaxCut four-kernel path (src/ising.cu:167-172) */
#define ORC_DSB 0 /* gather sgn(x)  (src/kernels.cu:251-254) */
#define ORC_BSB 1 /* gather x * coupling (src/kernels.cu:582-585) */

/* Threads used by the OpenMP loops below (the CPU baseline also times one). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------ RNG */

/* splitmix64 finaliser; counter-based stream u(seed, k). */
uint64_t orc_splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double orc_uniform(uint64_t seed, uint64_t k) {
  uint64_t z = orc_splitmix(seed + k * 0x9E3779B97F4A7C15ull);
  return (double)(z >> 11) * 0x1.0p-53;
}

/* Initial state, node-major N x A.  The reference draws X then Y per element
 * as (rand()/RAND_MAX*2-1)/10 (src/ising.cu:91-95); TSP starts at X = 0 with
 * the same Y draw (src/ising_tsp.cu:55-58).  Element e = i*A + a takes X from
 * stream index 2e and Y from 2e+1. */
void orc_init_state_f64(uint64_t seed, int problem, int64_t N, int64_t A, double *X, double *Y) {
  for (int64_t e = 0; e < N * A; e++) {
    double ux = orc_uniform(seed, 2 * (uint64_t)e);
    double uy = orc_uniform(seed, 2 * (uint64_t)e + 1);
    X[e] = problem == ORC_TSP ? 0.0 : (ux * 2 - 1) / 10;
    Y[e] = (uy * 2 - 1) / 10;
  }
}

void orc_init_state_f32(uint64_t seed, int problem, int64_t N, int64_t A, float *X, float *Y) {
  for (int64_t e = 0; e < N * A; e++) {
    double ux = orc_uniform(seed, 2 * (uint64_t)e);
    double uy = orc_uniform(seed, 2 * (uint64_t)e + 1);
    X[e] = problem == ORC_TSP ? 0.0f : (float)((ux * 2 - 1) / 10);
    Y[e] = (float)((uy * 2 - 1) / 10);
  }
}

/* ------------------------------------------------------------- scales */

/* xi0 = 0.7 / (sigma * sqrt(N)) where sigma is the sample standard deviation
 * of all N^2 entries of J (src/ising.cu:103-125), with N^2 taken in 64 bits
 * (the reference's int product overflows for N >= 46341). */
double orc_xi0_maxcut(int64_t N, int64_t Z, const double *val) {
  double sum = 0.0, nzc = 0.0;
  for (int64_t k = 0; k < Z; k++) sum += val[k];
  double n2 = (double)(N * N);
  double mean = sum / n2;
  for (int64_t k = 0; k < Z; k++) nzc += pow(val[k] - mean, 2);
  double zc = (double)(N * N - Z) * mean * mean;
  double var = pow((nzc + zc) / (double)(N * N - 1), 0.5);
  return 0.7 / (var * pow((double)N, 0.5));
}

/* Per-step coefficients, evaluated in FP64.
 * MaxCut (src/kernels.cu:225-226, 258-262):
 *   p = min(1.025 t / T, 1); c0 = dt(p+1); c1 = dt(p-1); c2 = dt*xi0; c3 = 0.
 * TSP (src/kernels.cu:553-559, 592-595), ncap = 2 (src/ising_tsp.cu:134):
 *   a = min(2*1.04 t/T, 2); b = min(0.5 + 0.5 (t/T)^1.9, 1);
 *   c0 = dt; c1 = dt * -(1 - a); c2 = xi0; c3 = b. */
void orc_sched(int problem, int64_t t, int64_t T, double dt, double xi0, double *c) {
  if (problem == ORC_MAXCUT || problem == ORC_UNFUSED) {
    /* the unfused path leaves p unclamped (src/kernels.cu:15) */
    double p = 1.025 * (double)t / (double)T;
    if (problem == ORC_MAXCUT) p = p < 1.0 ? p : 1.0;
    c[0] = dt * (p + 1);
    c[1] = dt * (p - 1);
    c[2] = dt * xi0;
    c[3] = 0.0;
  } else {
    int ncap = 2;
    double a = ncap * 1.04 * (double)t / (double)T;
    double b = 0.5 + (0.5 * pow((double)t / (double)T, 1.9));
    a = a < ncap ? a : ncap;
    b = b < 1 ? b : 1;
    c[0] = dt;
    c[1] = dt * -(1 - a);
    c[2] = xi0;
    c[3] = b;
  }
}

/* Pre-step with t = 0 (src/kernels.cu:15-20, launched at src/ising.cu:188):
 * Y += dt(0-1) X; X += dt(0+1) Y. */
void orc_prestep_f64(int64_t n, double dt, double *X, double *Y) {
  double cy = dt * (0.0 - 1), cx = dt * (0.0 + 1);
  for (int64_t e = 0; e < n; e++) {
    Y[e] = fma(cy, X[e], Y[e]);
    X[e] = fma(cx, Y[e], X[e]);
  }
}

void orc_prestep_f32(int64_t n, double dt, float *X, float *Y) {
  float cy = (float)(dt * (0.0 - 1)), cx = (float)(dt * (0.0 + 1));
  for (int64_t e = 0; e < n; e++) {
    Y[e] = fmaf(cy, X[e], Y[e]);
    X[e] = fmaf(cx, Y[e], X[e]);
  }
}

/* --------------------------------------------------------------- steps */

static inline float orc_bf16_round(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) { /* inf / nan: truncate, keep nan quiet */
    u &= 0xffff0000u;
    if (x != x) u |= 0x00400000u;
  } else {
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
  }
  float r;
  memcpy(&r, &u, 4);
  return r;
}

float orc_bf16(float x) { return orc_bf16_round(x); }

/* Unfused MaxCut step (src/ising.cu:167-172): Update_XY with c0 = dt(p+1),
 * c1 = dt(p-1), p unclamped (src/kernels.cu:15-20); then mat_mult on the
 * updated X, y += (dt xi0) F with F = -sum_k J_k sgn(x_col), no confine skip
 * (src/kernels.cu:196-207); then confine (src/kernels.cu:158-165). */
static void orc_unfused_f64(const int64_t *rp, const int32_t *col, const double *val, int64_t N,
                            int64_t A, const double *Xs, double *Xn, double *Y, const double *c) {
  double *Xu = (double *)malloc(sizeof(double) * (size_t)(N * A));
  for (int64_t e = 0; e < N * A; e++) {
    double y = fma(c[1], Xs[e], Y[e]);
    Xu[e] = fma(c[0], y, Xs[e]);
    Y[e] = y;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++)
    for (int64_t a = 0; a < A; a++) {
      int64_t e = i * A + a;
      double F = 0.0;
      for (int64_t k = rp[i]; k < rp[i + 1]; k++)
        F = fma(-val[k], signbit(Xu[(int64_t)col[k] * A + a]) ? -1.0 : 1.0, F);
      double x = Xu[e], y = fma(c[2], F, Y[e]);
      if (fabs(x) > 1) {
        x = signbit(x) ? -1.0 : 1.0;
        y = 0.0;
      }
      Xn[e] = x;
      Y[e] = y;
    }
  free(Xu);
}

static void orc_unfused_f32(const int64_t *rp, const int32_t *col, const float *val, int64_t N,
                            int64_t A, const float *Xs, float *Xn, float *Y, const double *cd) {
  float c0 = (float)cd[0], c1 = (float)cd[1], c2 = (float)cd[2];
  float *Xu = (float *)malloc(sizeof(float) * (size_t)(N * A));
  for (int64_t e = 0; e < N * A; e++) {
    float y = fmaf(c1, Xs[e], Y[e]);
    Xu[e] = fmaf(c0, y, Xs[e]);
    Y[e] = y;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++)
    for (int64_t a = 0; a < A; a++) {
      int64_t e = i * A + a;
      float F = 0.0f;
      for (int64_t k = rp[i]; k < rp[i + 1]; k++)
        F = fmaf(-val[k], signbit(Xu[(int64_t)col[k] * A + a]) ? -1.0f : 1.0f, F);
      float x = Xu[e], y = fmaf(c2, F, Y[e]);
      if (fabsf(x) > 1.0f) {
        x = signbit(x) ? -1.0f : 1.0f;
        y = 0.0f;
      }
      Xn[e] = x;
      Y[e] = y;
    }
  free(Xu);
}

/* Narrow fixed-point state (the narrow-state dSB forms), code width w = 12
 * or w = 8 bits per variable.  X is stored as the integer code
 * round(x * 2^xs), xs = w - 2, in [-2^xs, 2^xs], or +-(2^xs + 1) for any
 * |x| > 1 (the value of an element beyond the wall does not matter: the next
 * step confines it to sgn x with y = 0, src/kernels.cu:236-242); Y as
 * round(y * 2^sy) clamped to [-ymax, ymax], ymax = 2^(w-1) - 1 (0 beside a
 * wall X code).  Codes are rounded to nearest even and stored as integers,
 * so a code of 0 is +0.0 (sgn +1).  sy is the largest s in [0, w - 1] with
 * ymax * 2^-s >= 2/dt + dt, the bound of |y| while |x| <= 1:
 * |x_new - x| = c0 |y'| <= 2 with c0 = dt (p + 1) >= dt gives |y'| <= 2 / dt,
 * and |c1 x| <= dt.  w = 12: X codes +-1024 (+-1025), Y codes +-2047;
 * w = 10: X codes +-256 (+-257), Y codes +-511; w = 8: X codes +-64 (+-65),
 * Y codes +-127. */
static inline int orc_qn_ymax(int w) { return (1 << (w - 1)) - 1; }

int orc_qn_shift(int w, double dt) {
  if (!(dt > 0) || (w != 12 && w != 10 && w != 8)) return -1;
  double ymax = 2.0 / dt + dt;
  int s = w - 1;
  while (s > 0 && ldexp((double)orc_qn_ymax(w), -s) < ymax) s--;
  return s;
}

int orc_q12_shift(double dt) { return orc_qn_shift(12, dt); }

static inline void orc_qn_round(int w, float *x, float *y, int sy) {
  const int xs = w - 2;
  const float ym = (float)orc_qn_ymax(w);
  if (fabsf(*x) > 1.0f) {
    const float wall = ldexpf((float)((1 << xs) + 1), -xs);
    *x = signbit(*x) ? -wall : wall;
    *y = 0.0f;
    return;
  }
  int xq = (int)rintf(ldexpf(*x, xs));
  float yr = rintf(ldexpf(*y, sy));
  int yq = yr > ym ? (int)ym : yr < -ym ? -(int)ym : (int)yr;
  *x = ldexpf((float)xq, -xs);
  *y = ldexpf((float)yq, -sy);
}

/* Rounds FP32 (x, y) pairs to the codes above in place. */
void orc_qn_round_state(int w, int64_t n, float *X, float *Y, int sy) {
  for (int64_t e = 0; e < n; e++) orc_qn_round(w, &X[e], &Y[e], sy);
}

void orc_q12_round_state(int64_t n, float *X, float *Y, int sy) {
  orc_qn_round_state(12, n, X, Y, sy);
}

/* Stochastic Y rounding of the 8-bit form (SIS_FORM_Q8S; no reference
 * counterpart).  Element (i, a) after step t takes dither byte
 * j = (a mod 256) / 64 of
 *   h = mix(key + l 0x85EBCA77),  key = (i Q + q) 0x9E3779B1 xor salt(t),
 *   salt(t) = t 0x27D4EB2F + 0x165667B1,
 *   mix: h ^= h >> 15; h *= 0x2C1B3C6D; h ^= h >> 12,
 * q = a / 256, l = a mod 64, Q = ceil(A / 256), all in wrapping uint32
 * arithmetic (i is the node, so the storage order of the rows does not
 * matter).  With w the nearest-even integer to y 2^(sy+8) and b that byte, the
 * Y code is floor((w + b) / 256): y 2^sy to 1/256, rounded down or up with
 * the probability of its fraction (b uniform on 0..255), so the rounding is
 * unbiased, and a y 2^sy on the code grid stays put.  No clamp: a live
 * element has |y| 2^sy <= YM + 1e-5 under the Y bound rule, so |w| <= 256 YM
 * and the code stays within +-YM.  X rounds to nearest even, a wall element
 * takes the wall code and Y = 0, as in the 8-bit form. */
static inline uint32_t orc_q8s_hash(int64_t t, int64_t i, int64_t a, int64_t A) {
  const uint32_t Q = (uint32_t)((A + 255) / 256), q = (uint32_t)(a / 256), l = (uint32_t)(a % 64);
  const uint32_t salt = (uint32_t)t * 0x27D4EB2Fu + 0x165667B1u;
  const uint32_t key = ((uint32_t)i * Q + q) * 0x9E3779B1u ^ salt;
  uint32_t h = key + l * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h;
}

static inline void orc_q8s_round(float *x, float *y, int sy, int64_t t, int64_t i, int64_t a,
                                 int64_t A) {
  if (fabsf(*x) > 1.0f) {
    orc_qn_round(8, x, y, sy);  // the wall code, Y = 0
    return;
  }
  const uint32_t h = orc_q8s_hash(t, i, a, A);
  const int j = (int)((a % 256) / 64);
  const int b = (int)((h >> (8 * j)) & 0xffu);
  const int w = (int)rintf(ldexpf(*y, sy + 8));  /* exact product, |w| < 2^16 */
  /* integer codes, so that a code of 0 is +0 (sign +1) as in the 8-bit
   * form; the floor of a quotient of a possibly negative numerator */
  const int yq = (w + b + 65536) / 256 - 256;
  const int xq = (int)rintf(ldexpf(*x, 6));
  *x = ldexpf((float)xq, -6);
  *y = ldexpf((float)yq, -sy);
}

/* precision codes of the FP32 core below: 2 = 12-bit state, 3 = 8-bit state,
 * 4 = 10-bit state, 5 = 8-bit state with stochastic Y rounding */
static inline int orc_qn_width(int precision) {
  return precision == 2 ? 12 : precision == 4 ? 10 : 8;
}
static inline int orc_qn_precision(int w) { return w == 12 ? 2 : w == 8 ? 3 : 4; }

/* One fused step over all (i, a), FP64.
 * MaxCut (src/kernels.cu:236-262): confine |x| > 1 to sgn x and y = 0, else
 * y += (dt xi0) F with F = -sum_k J_k g(x_col); then x += dt(p+1) y and
 * y += dt(p-1) x for every element.
 * TSP (src/kernels.cu:564-596): a confined element skips the whole update;
 * otherwise y += dt (xi0 (2 I1 + b h_i)), x += dt y, y += dt(-(1-a)) x.
 * Each row's forces are accumulated for all agents at once (agent index
 * innermost); per element the terms are still added in ascending CSR order,
 * so the result is the same as one element at a time. */
void orc_step_f64(int problem, int gather, const int64_t *rp, const int32_t *col, const double *val,
                  const double *h, int64_t N, int64_t A, const double *Xs, double *Xn, double *Y,
                  const double *c, double coupling) {
  if (problem == ORC_UNFUSED) {
    orc_unfused_f64(rp, col, val, N, A, Xs, Xn, Y, c);
    return;
  }
#pragma omp parallel
  {
    double *F = (double *)malloc(sizeof(double) * (size_t)(A > 0 ? A : 1));
#pragma omp for schedule(static)
    for (int64_t i = 0; i < N; i++) {
      for (int64_t a = 0; a < A; a++) F[a] = 0.0;
      for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
        const double *xj = Xs + (int64_t)col[k] * A;
        const double v = val[k];
        if (gather == ORC_DSB)
          for (int64_t a = 0; a < A; a++) F[a] = fma(-v, copysign(1.0, xj[a]), F[a]);
        else
          for (int64_t a = 0; a < A; a++) F[a] = fma(-v, xj[a] * coupling, F[a]);
      }
      for (int64_t a = 0; a < A; a++) {
        int64_t e = i * A + a;
        double x = Xs[e], y = Y[e];
        int conf = fabs(x) > 1;
        if (problem == ORC_MAXCUT) {
          if (conf) {
            x = signbit(x) ? -1.0 : 1.0;
            y = 0.0;
          } else {
            y = fma(c[2], F[a], y);
          }
          x = fma(c[0], y, x);
          y = fma(c[1], x, y);
        } else {
          if (conf) {
            x = signbit(x) ? -1.0 : 1.0;
            y = 0.0;
          } else {
            double q = fma(c[3], h[i], 2 * F[a]);
            y = fma(c[0], c[2] * q, y);
            x = fma(c[0], y, x);
            y = fma(c[1], x, y);
          }
        }
        Xn[e] = x;
        Y[e] = y;
      }
    }
    free(F);
  }
}

/* FP32 form of the same step.  precision 1 rounds every gathered x to
 * bfloat16 (round to nearest even) before use: the BF16-shadow bSB form.
 * precision 2 / 3 (MaxCut dSB) round every updated (x, y) to the 12-bit /
 * 8-bit codes above with Y shift sy: the narrow-state forms. */
static void orc_step_f32_core(int problem, int gather, int precision, int sy, const int64_t *rp,
                              const int32_t *col, const float *val, const float *h, int64_t N,
                              int64_t A, const float *Xs, float *Xn, float *Y, const double *cd,
                              double coupling_d, int64_t t) {
  if (problem == ORC_UNFUSED) {
    orc_unfused_f32(rp, col, val, N, A, Xs, Xn, Y, cd);
    return;
  }
  const float c0 = (float)cd[0], c1 = (float)cd[1], c2 = (float)cd[2], c3 = (float)cd[3];
  const float coupling = (float)coupling_d;
  const int bf16_gather = precision == 1;
#pragma omp parallel
  {
    float *F = (float *)malloc(sizeof(float) * (size_t)(A > 0 ? A : 1));
#pragma omp for schedule(static)
    for (int64_t i = 0; i < N; i++) {
      for (int64_t a = 0; a < A; a++) F[a] = 0.0f;
      for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
        const float *xj = Xs + (int64_t)col[k] * A;
        const float v = val[k];
        if (gather == ORC_DSB) {
          if (bf16_gather)
            for (int64_t a = 0; a < A; a++)
              F[a] = fmaf(-v, copysignf(1.0f, orc_bf16_round(xj[a])), F[a]);
          else
            for (int64_t a = 0; a < A; a++) F[a] = fmaf(-v, copysignf(1.0f, xj[a]), F[a]);
        } else {
          if (bf16_gather)
            for (int64_t a = 0; a < A; a++) F[a] = fmaf(-v, orc_bf16_round(xj[a]) * coupling, F[a]);
          else
            for (int64_t a = 0; a < A; a++) F[a] = fmaf(-v, xj[a] * coupling, F[a]);
        }
      }
      for (int64_t a = 0; a < A; a++) {
        int64_t e = i * A + a;
        float x = Xs[e], y = Y[e];
        int conf = fabsf(x) > 1.0f;
        if (problem == ORC_MAXCUT) {
          if (conf) {
            x = signbit(x) ? -1.0f : 1.0f;
            y = 0.0f;
          } else {
            y = fmaf(c2, F[a], y);
          }
          x = fmaf(c0, y, x);
          y = fmaf(c1, x, y);
        } else {
          if (conf) {
            x = signbit(x) ? -1.0f : 1.0f;
            y = 0.0f;
          } else {
            float q = fmaf(c3, h[i], 2.0f * F[a]);
            y = fmaf(c0, c2 * q, y);
            x = fmaf(c0, y, x);
            y = fmaf(c1, x, y);
          }
        }
        if (precision == 5) orc_q8s_round(&x, &y, sy, t, i, a, A);
        else if (precision >= 2) orc_qn_round(orc_qn_width(precision), &x, &y, sy);
        Xn[e] = x;
        Y[e] = y;
      }
    }
    free(F);
  }
}

void orc_step_f32(int problem, int gather, int bf16_gather, const int64_t *rp, const int32_t *col,
                  const float *val, const float *h, int64_t N, int64_t A, const float *Xs, float *Xn,
                  float *Y, const double *cd, double coupling_d) {
  orc_step_f32_core(problem, gather, bf16_gather ? 1 : 0, 0, rp, col, val, h, N, A, Xs, Xn, Y, cd,
                    coupling_d, 0);
}

/* Narrow-state dSB step (code width w): Xs / Y hold decoded codes. */
void orc_step_qn(int w, const int64_t *rp, const int32_t *col, const float *val, int64_t N,
                 int64_t A, const float *Xs, float *Xn, float *Y, const double *cd, int sy) {
  orc_step_f32_core(ORC_MAXCUT, ORC_DSB, orc_qn_precision(w), sy, rp, col, val, NULL, N, A, Xs,
                    Xn, Y, cd, 1.0, 0);
}

/* One step of the 8-bit form with stochastic Y rounding at step index t. */
void orc_step_q8s(const int64_t *rp, const int32_t *col, const float *val, int64_t N, int64_t A,
                  const float *Xs, float *Xn, float *Y, const double *cd, int sy, int64_t t) {
  orc_step_f32_core(ORC_MAXCUT, ORC_DSB, 5, sy, rp, col, val, NULL, N, A, Xs, Xn, Y, cd, 1.0, t);
}

void orc_step_q12(const int64_t *rp, const int32_t *col, const float *val, int64_t N, int64_t A,
                  const float *Xs, float *Xn, float *Y, const double *cd, int sy) {
  orc_step_qn(12, rp, col, val, N, A, Xs, Xn, Y, cd, sy);
}

/* ------------------------------------------------------------- windows */

/* MaxCut window (src/kernels.cu:24-91), per agent a:
 *   eq = all_i sgn X[i,a] == cur[i,a];  if !bif: stab = eq ? stab+1 : 0;
 *   bif = stab == K-1; on the 0->1 transition final[:,a] = sgn X[:,a];
 *   cur[:,a] = sgn X[:,a].
 * Returns the number of agents with bif set. */
static int64_t orc_window_maxcut_core(const void *Xv, int is_f32, int64_t N, int64_t A, int K,
                                      int32_t *cur, int32_t *fin, int32_t *stab, uint8_t *bif) {
  const float *Xf = (const float *)Xv;
  const double *Xd = (const double *)Xv;
#define ORC_SGN(e) ((is_f32 ? signbit(Xf[e]) : signbit(Xd[e])) ? -1 : 1)
  uint8_t *eq = (uint8_t *)malloc((size_t)(A > 0 ? A : 1));
  uint8_t *now_locked = (uint8_t *)malloc((size_t)(A > 0 ? A : 1));
  for (int64_t a = 0; a < A; a++) eq[a] = 1;
  for (int64_t i = 0; i < N; i++)
    for (int64_t a = 0; a < A; a++)
      if (ORC_SGN(i * A + a) != cur[i * A + a]) eq[a] = 0;
  int64_t nbif = 0;
  for (int64_t a = 0; a < A; a++) {
    uint8_t prev = bif[a];
    if (!prev) stab[a] = eq[a] ? stab[a] + 1 : 0;
    uint8_t now = stab[a] == K - 1;
    bif[a] = now;
    now_locked[a] = now && !prev;
    nbif += now;
  }
  for (int64_t i = 0; i < N; i++)
    for (int64_t a = 0; a < A; a++)
      if (now_locked[a]) fin[i * A + a] = ORC_SGN(i * A + a);
  for (int64_t e = 0; e < N * A; e++) cur[e] = ORC_SGN(e);
  free(eq);
  free(now_locked);
  return nbif;
}

/* TSP window (src/kernels.cu:461-541): valid iff every city and every
 * position holds exactly one +1 spin (spin index = position*n + city); a
 * valid agent records t in valid_list and copies sgn X into final (again at
 * every later valid window).  cur = sgn X for all. */
static void orc_window_tsp_core(const void *Xv, int is_f32, int64_t n, int64_t A, int64_t t,
                                int32_t *cur, int32_t *fin, int32_t *valid_list) {
  const float *Xf = (const float *)Xv;
  const double *Xd = (const double *)Xv;
  int64_t N = n * n;
  /* visits per (line, agent): lines 0..n-1 are cities, n..2n-1 positions */
  int32_t *cnt = (int32_t *)calloc((size_t)(2 * n * A), sizeof(int32_t));
  uint8_t *valid = (uint8_t *)malloc((size_t)(A > 0 ? A : 1));
  for (int64_t i = 0; i < N; i++)
    for (int64_t a = 0; a < A; a++)
      if (ORC_SGN(i * A + a) == 1) {
        cnt[(i % n) * A + a] += 1;     /* city visits */
        cnt[(n + i / n) * A + a] += 1; /* position visits */
      }
  for (int64_t a = 0; a < A; a++) {
    valid[a] = 1;
    for (int64_t k = 0; k < 2 * n; k++)
      if (cnt[k * A + a] != 1) { valid[a] = 0; break; }
    if (valid[a]) valid_list[a] = (int32_t)t;
  }
  for (int64_t i = 0; i < N; i++)
    for (int64_t a = 0; a < A; a++)
      if (valid[a]) fin[i * A + a] = ORC_SGN(i * A + a);
  for (int64_t e = 0; e < N * A; e++) cur[e] = ORC_SGN(e);
#undef ORC_SGN
  free(cnt);
  free(valid);
}

/* --------------------------------------------------------------- solve */

static int64_t orc_solve_f32_core(int problem, int gather, int precision, int sy,
                                  const int64_t *rp, const int32_t *col, const float *val,
                                  const float *h, int64_t N, int64_t A, int64_t T, int uw, int64_t P,
                                  int K, int64_t ncities, double dt, double xi0, double coupling,
                                  float *X, float *Y, int32_t *cur, int32_t *fin, int32_t *stab,
                                  uint8_t *bif, int32_t *valid_list) {
  float *Xb = (float *)malloc(sizeof(float) * (size_t)(N * A));
  orc_prestep_f32(N * A, dt, X, Y);
  if (precision >= 2) orc_qn_round_state(orc_qn_width(precision), N * A, X, Y, sy);
  int64_t steps = 0;
  double c[4];
  while (steps < T) {
    for (int64_t s = 0; s < P; s++) {
      orc_sched(problem, steps + s, T, dt, xi0, c);
      orc_step_f32_core(problem, gather, precision, sy, rp, col, val, h, N, A, X, Xb, Y, c,
                        coupling, steps + s);
      memcpy(X, Xb, sizeof(float) * (size_t)(N * A));
    }
    steps += P;
    if (problem != ORC_TSP) {
      int64_t nb = orc_window_maxcut_core(X, 1, N, A, K, cur, fin, stab, bif);
      if (uw && nb == A) break;
    } else {
      orc_window_tsp_core(X, 1, ncities, A, steps, cur, fin, valid_list);
    }
  }
  free(Xb);
  return steps;
}

/* Whole solve from a given initial state (src/ising.cu:140-226,
 * src/ising_tsp.cu:111-164): pre-step, then repeat {P steps, window} while
 * the executed step count is below T; with uw (MaxCut only) stop after a
 * window at which every agent has bifurcated.  cur/fin must arrive filled
 * with -1 (src/ising.cu:128-129), stab/bif/valid_list with 0.
 * precision: 0 = FP32, 1 = FP32 with BF16 gathers; X, Y are float.
 * Returns the number of steps executed. */
int64_t orc_solve_f32(int problem, int gather, int precision, const int64_t *rp, const int32_t *col,
                      const float *val, const float *h, int64_t N, int64_t A, int64_t T, int uw,
                      int64_t P, int K, int64_t ncities, double dt, double xi0, double coupling,
                      float *X, float *Y, int32_t *cur, int32_t *fin, int32_t *stab, uint8_t *bif,
                      int32_t *valid_list) {
  return orc_solve_f32_core(problem, gather, precision == 1 ? 1 : 0, 0, rp, col, val, h, N, A, T,
                            uw, P, K, ncities, dt, xi0, coupling, X, Y, cur, fin, stab, bif,
                            valid_list);
}

/* The same for the narrow-state dSB MaxCut forms (code width w): X, Y arrive
 * as FP32 values and are rounded to the codes before and after the
 * pre-step (the device rounds at set_state too, which leaves values already
 * on the code grid unchanged) and after every step. */
int64_t orc_solve_qn(int w, const int64_t *rp, const int32_t *col, const float *val, int64_t N,
                     int64_t A, int64_t T, int uw, int64_t P, int K, double dt, double xi0, int sy,
                     float *X, float *Y, int32_t *cur, int32_t *fin, int32_t *stab, uint8_t *bif) {
  orc_qn_round_state(w, N * A, X, Y, sy);
  return orc_solve_f32_core(ORC_MAXCUT, ORC_DSB, orc_qn_precision(w), sy, rp, col, val, NULL, N,
                            A, T, uw, P, K, 0, dt, xi0, 1.0, X, Y, cur, fin, stab, bif, NULL);
}

/* The same for the 8-bit form with stochastic Y rounding: the initial state
 * and the pre-step round to nearest even, every step after rounds Y with the
 * dither of its step index. */
int64_t orc_solve_q8s(const int64_t *rp, const int32_t *col, const float *val, int64_t N,
                      int64_t A, int64_t T, int uw, int64_t P, int K, double dt, double xi0, int sy,
                      float *X, float *Y, int32_t *cur, int32_t *fin, int32_t *stab, uint8_t *bif) {
  orc_qn_round_state(8, N * A, X, Y, sy);
  return orc_solve_f32_core(ORC_MAXCUT, ORC_DSB, 5, sy, rp, col, val, NULL, N, A, T, uw, P, K, 0,
                            dt, xi0, 1.0, X, Y, cur, fin, stab, bif, NULL);
}

int64_t orc_solve_q12(const int64_t *rp, const int32_t *col, const float *val, int64_t N,
                      int64_t A, int64_t T, int uw, int64_t P, int K, double dt, double xi0, int sy,
                      float *X, float *Y, int32_t *cur, int32_t *fin, int32_t *stab, uint8_t *bif) {
  return orc_solve_qn(12, rp, col, val, N, A, T, uw, P, K, dt, xi0, sy, X, Y, cur, fin, stab, bif);
}

int64_t orc_solve_f64(int problem, int gather, const int64_t *rp, const int32_t *col,
                      const double *val, const double *h, int64_t N, int64_t A, int64_t T, int uw,
                      int64_t P, int K, int64_t ncities, double dt, double xi0, double coupling,
                      double *X, double *Y, int32_t *cur, int32_t *fin, int32_t *stab, uint8_t *bif,
                      int32_t *valid_list) {
  double *Xb = (double *)malloc(sizeof(double) * (size_t)(N * A));
  orc_prestep_f64(N * A, dt, X, Y);
  int64_t steps = 0;
  double c[4];
  while (steps < T) {
    for (int64_t s = 0; s < P; s++) {
      orc_sched(problem, steps + s, T, dt, xi0, c);
      orc_step_f64(problem, gather, rp, col, val, h, N, A, X, Xb, Y, c, coupling);
      memcpy(X, Xb, sizeof(double) * (size_t)(N * A));
    }
    steps += P;
    if (problem != ORC_TSP) {
      int64_t nb = orc_window_maxcut_core(X, 0, N, A, K, cur, fin, stab, bif);
      if (uw && nb == A) break;
    } else {
      orc_window_tsp_core(X, 0, ncities, A, steps, cur, fin, valid_list);
    }
  }
  free(Xb);
  return steps;
}

/* Advance `nsteps` fused steps starting at step index t0 with horizon T, no
 * pre-step and no windows: the trajectory-parity primitive. */
void orc_advance_f32(int problem, int gather, int precision, const int64_t *rp, const int32_t *col,
                     const float *val, const float *h, int64_t N, int64_t A, int64_t t0,
                     int64_t nsteps, int64_t T, double dt, double xi0, double coupling, float *X,
                     float *Y) {
  float *Xb = (float *)malloc(sizeof(float) * (size_t)(N * A));
  double c[4];
  for (int64_t s = 0; s < nsteps; s++) {
    orc_sched(problem, t0 + s, T, dt, xi0, c);
    orc_step_f32(problem, gather, precision == 1, rp, col, val, h, N, A, X, Xb, Y, c, coupling);
    memcpy(X, Xb, sizeof(float) * (size_t)(N * A));
  }
  free(Xb);
}

/* Narrow-state dSB MaxCut steps t0 .. t0+nsteps-1 on decoded codes. */
void orc_advance_qn(int w, const int64_t *rp, const int32_t *col, const float *val, int64_t N,
                    int64_t A, int64_t t0, int64_t nsteps, int64_t T, double dt, double xi0, int sy,
                    float *X, float *Y) {
  float *Xb = (float *)malloc(sizeof(float) * (size_t)(N * A));
  double c[4];
  for (int64_t s = 0; s < nsteps; s++) {
    orc_sched(ORC_MAXCUT, t0 + s, T, dt, xi0, c);
    orc_step_qn(w, rp, col, val, N, A, X, Xb, Y, c, sy);
    memcpy(X, Xb, sizeof(float) * (size_t)(N * A));
  }
  free(Xb);
}

void orc_advance_q8s(const int64_t *rp, const int32_t *col, const float *val, int64_t N,
                     int64_t A, int64_t t0, int64_t nsteps, int64_t T, double dt, double xi0, int sy,
                     float *X, float *Y) {
  float *Xb = (float *)malloc(sizeof(float) * (size_t)(N * A));
  double c[4];
  for (int64_t s = 0; s < nsteps; s++) {
    orc_sched(ORC_MAXCUT, t0 + s, T, dt, xi0, c);
    orc_step_q8s(rp, col, val, N, A, X, Xb, Y, c, sy, t0 + s);
    memcpy(X, Xb, sizeof(float) * (size_t)(N * A));
  }
  free(Xb);
}

void orc_advance_q12(const int64_t *rp, const int32_t *col, const float *val, int64_t N, int64_t A,
                     int64_t t0, int64_t nsteps, int64_t T, double dt, double xi0, int sy, float *X,
                     float *Y) {
  orc_advance_qn(12, rp, col, val, N, A, t0, nsteps, T, dt, xi0, sy, X, Y);
}

/* Pre-step (src/kernels.cu:15-20 at t = 0) on decoded codes, then rounding. */
void orc_prestep_qn(int w, int64_t n, double dt, int sy, float *X, float *Y) {
  orc_prestep_f32(n, dt, X, Y);
  orc_qn_round_state(w, n, X, Y, sy);
}

void orc_prestep_q12(int64_t n, double dt, int sy, float *X, float *Y) {
  orc_prestep_qn(12, n, dt, sy, X, Y);
}

void orc_advance_f64(int problem, int gather, const int64_t *rp, const int32_t *col,
                     const double *val, const double *h, int64_t N, int64_t A, int64_t t0,
                     int64_t nsteps, int64_t T, double dt, double xi0, double coupling, double *X,
                     double *Y) {
  double *Xb = (double *)malloc(sizeof(double) * (size_t)(N * A));
  double c[4];
  for (int64_t s = 0; s < nsteps; s++) {
    orc_sched(problem, t0 + s, T, dt, xi0, c);
    orc_step_f64(problem, gather, rp, col, val, h, N, A, X, Xb, Y, c, coupling);
    memcpy(X, Xb, sizeof(double) * (size_t)(N * A));
  }
  free(Xb);
}

/* Exposed window cores for unit tests. */
int64_t orc_window_maxcut_f32(const float *X, int64_t N, int64_t A, int K, int32_t *cur,
                              int32_t *fin, int32_t *stab, uint8_t *bif) {
  return orc_window_maxcut_core(X, 1, N, A, K, cur, fin, stab, bif);
}

void orc_window_tsp_f32(const float *X, int64_t n, int64_t A, int64_t t, int32_t *cur,
                        int32_t *fin, int32_t *valid_list) {
  orc_window_tsp_core(X, 1, n, A, t, cur, fin, valid_list);
}

/* ---------------------------------------------------------------- cuts */

/* H_a = sum_i 0.5 s_i sum_k J_k s_col (src/kernels.cu:371-379, rows summed in
 * order as src/kernels.cu:394-397 does) and cut_a = -0.5 (H_a - S). */
void orc_cut(const int64_t *rp, const int32_t *col, const double *val, int64_t N, int64_t A,
             const int32_t *spins, double S, double *cut) {
  double *H = (double *)calloc((size_t)(A > 0 ? A : 1), sizeof(double));
  double *tmp = (double *)malloc(sizeof(double) * (size_t)(A > 0 ? A : 1));
  for (int64_t i = 0; i < N; i++) {
    for (int64_t a = 0; a < A; a++) tmp[a] = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
      const int32_t *sj = spins + (int64_t)col[k] * A;
      for (int64_t a = 0; a < A; a++) tmp[a] = fma(val[k], (double)sj[a], tmp[a]);
    }
    for (int64_t a = 0; a < A; a++) H[a] += tmp[a] * spins[i * A + a] * 0.5;
  }
  for (int64_t a = 0; a < A; a++) cut[a] = (-0.5) * (H[a] - S);
  free(H);
  free(tmp);
}

/* Ising energy E_a = s^T J s - h.s for the TSP encoding (FP64). */
void orc_energy_field(const int64_t *rp, const int32_t *col, const double *val, const double *h,
                      int64_t N, int64_t A, const int32_t *spins, double *E) {
  double *acc = (double *)calloc((size_t)(A > 0 ? A : 1), sizeof(double));
  double *tmp = (double *)malloc(sizeof(double) * (size_t)(A > 0 ? A : 1));
  for (int64_t i = 0; i < N; i++) {
    for (int64_t a = 0; a < A; a++) tmp[a] = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
      const int32_t *sj = spins + (int64_t)col[k] * A;
      for (int64_t a = 0; a < A; a++) tmp[a] += val[k] * (double)sj[a];
    }
    for (int64_t a = 0; a < A; a++) acc[a] += tmp[a] * spins[i * A + a] - h[i] * spins[i * A + a];
  }
  for (int64_t a = 0; a < A; a++) E[a] = acc[a];
  free(acc);
  free(tmp);
}

/* Host statistics over per-agent cuts (src/ising.cu:238-265). out = {min,
 * max, mean, std (divisor A-1), ps}. */
void orc_cut_stats(const double *cut, int64_t A, double *out) {
  double mx = -1.7976931348623157e308, mn = 1.7976931348623157e308, tot = 0.0;
  for (int64_t a = 0; a < A; a++) {
    double v = cut[a];
    if (v > mx) mx = v;
    if (v < mn) mn = v;
    tot += v;
  }
  double avg = tot / (double)A, sd = 0.0;
  int64_t nmax = 0;
  for (int64_t a = 0; a < A; a++) {
    sd += pow(cut[a] - avg, 2);
    if (cut[a] == mx) nmax++;
  }
  sd = pow(sd / (double)(A - 1), 0.5);
  out[0] = mn;
  out[1] = mx;
  out[2] = avg;
  out[3] = sd;
  out[4] = (double)nmax / (double)A;
}

/* ----------------------------------------------------------------- TSP */

static inline int64_t orc_tri(int64_t c1, int64_t c2, int64_t n) {
  return c2 + c1 * n - ((c1 + 2) * (c1 + 1)) / 2; /* src/kernels.cu:420 */
}

/* Packed upper-triangle integer distances.  kind 0 = EUC_2D, 1 = GEO;
 * euc_round 0 truncates like the reference (src/kernels.cu:437-440),
 * 1 rounds to nearest (TSPLIB nint).  GEO: src/kernels.cu:414-421. */
void orc_tsp_distances(int64_t n, const double *x, const double *y, int kind, int euc_round,
                       double *d) {
  for (int64_t c1 = 0; c1 < n; c1++)
    for (int64_t c2 = c1 + 1; c2 < n; c2++) {
      double v;
      if (kind == 0) {
        double xd = x[c1] - x[c2], yd = y[c1] - y[c2];
        double r = pow(xd * xd + yd * yd, 0.5);
        v = euc_round ? (double)(int)(r + 0.5) : (double)(int)r;
      } else {
        double R = 6378.388;
        double q1 = cos(y[c1] - y[c2]);
        double q2 = cos(x[c1] - x[c2]);
        double q3 = cos(x[c1] + x[c2]);
        double dd = (R * acos(0.5 * ((1.0 + q1) * q2 - (1.0 - q1) * q3)) + 1.0);
        v = (double)(int)dd;
      }
      d[orc_tri(c1, c2, n)] = v;
    }
}

/* Per-city totals D_k = sum_l d_kl and the maximum distance
 * (src/tsp_gpu.cu:55-68). */
void orc_tsp_totals(int64_t n, const double *d, double *D, double *maxd) {
  double m = 0.0;
  for (int64_t k = 0; k < n; k++) D[k] = 0.0;
  for (int64_t c1 = 0; c1 < n; c1++)
    for (int64_t c2 = c1 + 1; c2 < n; c2++) {
      double v = d[orc_tri(c1, c2, n)];
      D[c1] += v;
      D[c2] += v;
      if (v > m) m = v;
    }
  *maxd = m;
}

/* Dense n^2 x n^2 QUBO with A = 1, B = C = max d (src/tsp_gpu.cu:76-112);
 * row (i,k) = i*n + k, column (j,l) = j*n + l. */
void orc_tsp_qubo_dense(int64_t n, const double *d, double maxd, double *J) {
  double A = 1, B = maxd, C = maxd;
  int64_t N = n * n;
  for (int64_t i = 0; i < n; i++)
    for (int64_t k = 0; k < n; k++)
      for (int64_t j = 0; j < n; j++)
        for (int64_t l = 0; l < n; l++) {
          double v;
          int adj = (i == j + 1 || j == i + 1 || (i == 0 && j == n - 1) || (i == n - 1 && j == 0));
          if (i != j && k == l) v = C / 4;
          else if (adj && k != l) {
            int64_t c1 = k < l ? k : l, c2 = k < l ? l : k;
            v = A * d[orc_tri(c1, c2, n)] / 8;
          } else if (i == j && k != l) v = B / 4;
          else if (i == j && k == l) v = (B + C) / 4;
          else v = 0.0;
          J[(i * n + k) * N + (j * n + l)] = v;
        }
}

/* Field h_(i,k) = -(A D_k + (n-2)(B+C)) / 2 (src/kernels.cu:591-592). */
void orc_tsp_field(int64_t n, const double *D, double maxd, double *h) {
  double A = 1, B = maxd, C = maxd;
  for (int64_t i = 0; i < n; i++)
    for (int64_t k = 0; k < n; k++) h[i * n + k] = -1 * (A * D[k] + (n - 2) * (B + C)) / 2;
}

/* Tour of agent a from final spins (src/ising_tsp.cu:206-214) and its cyclic
 * length from the packed table (src/ising_tsp.cu:272-287).  Returns -1 if
 * the spins do not decode to n cities. */
double orc_tsp_tour_length(int64_t n, const double *d, const int32_t *spins, int64_t A, int64_t a,
                           int32_t *tour) {
  int64_t m = 0;
  for (int64_t o = 0; o < n; o++)
    for (int64_t c = 0; c < n; c++)
      if (spins[(o * n + c) * A + a] == 1) {
        if (m >= n) return -1.0;
        tour[m++] = (int32_t)c;
      }
  if (m != n) return -1.0;
  double sum = 0.0;
  for (int64_t i = 0; i < n; i++) {
    int64_t u = tour[i], v = tour[(i + 1) % n];
    int64_t c1 = u < v ? u : v, c2 = u < v ? v : u;
    sum += d[orc_tri(c1, c2, n)];
  }
  return sum;
}
