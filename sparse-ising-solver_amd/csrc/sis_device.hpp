// sis_device.hpp -- device helpers shared by the single-instance and the
// batched kernels: counter-based RNG, BF16 rounding, wavefront utilities,
// state loads / stores and the SB element update.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sis_kernels.hpp"

namespace sis {

// ------------------------------------------------------------ utilities

__device__ __forceinline__ uint64_t d_splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double d_uniform(uint64_t seed, uint64_t k) {
  return (double)(d_splitmix(seed + k * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}

// Round to nearest even BF16 (NaN stays quiet NaN, inf stays inf).
__device__ __forceinline__ uint16_t d_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) {
    u &= 0xffff0000u;
    if (x != x) u |= 0x00400000u;
    return (uint16_t)(u >> 16);
  }
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float d_from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ int wave_index() {
  return __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWaves + (threadIdx.x >> 6)));
}

__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  for (int m = 32; m >= 1; m >>= 1) {
    lo |= (uint32_t)__shfl_xor((int)lo, m, 64);
    hi |= (uint32_t)__shfl_xor((int)hi, m, 64);
  }
  return ((uint64_t)hi << 32) | lo;
}

// Lane l's 64-bit value, broadcast (l wave-uniform).
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

template <typename T, int V>
struct alignas(sizeof(T) * V) Vec {
  T v[V];
};

template <typename T, int V>
__device__ __forceinline__ Vec<T, V> ldv(const T *p) {
  return *reinterpret_cast<const Vec<T, V> *>(p);
}

template <typename T, int V>
__device__ __forceinline__ void stv(T *p, const Vec<T, V> &x) {
  *reinterpret_cast<Vec<T, V> *>(p) = x;
}

// Own-row state access; NT = 1 marks it non-temporal (streamed once per step).
template <typename T, int V, int NT>
__device__ __forceinline__ Vec<T, V> ldvs(const T *p) {
  if constexpr (NT == 0) {
    return ldv<T, V>(p);
  } else if constexpr (V == 1) {
    Vec<T, V> o;
    o.v[0] = __builtin_nontemporal_load(p);
    return o;
  } else {
    typedef T vt __attribute__((ext_vector_type(V)));
    const vt r = __builtin_nontemporal_load(reinterpret_cast<const vt *>(p));
    Vec<T, V> o;
#pragma unroll
    for (int u = 0; u < V; u++) o.v[u] = r[u];
    return o;
  }
}

template <typename T, int V, int NT>
__device__ __forceinline__ void stvs(T *p, const Vec<T, V> &x) {
  if constexpr (NT == 0) {
    stv<T, V>(p, x);
  } else if constexpr (V == 1) {
    __builtin_nontemporal_store(x.v[0], p);
  } else {
    typedef T vt __attribute__((ext_vector_type(V)));
    vt r;
#pragma unroll
    for (int u = 0; u < V; u++) r[u] = x.v[u];
    __builtin_nontemporal_store(r, reinterpret_cast<vt *>(p));
  }
}

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

static inline unsigned grid_for(int64_t work, int64_t per_block, int64_t cap = 1 << 20) {
  int64_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

template <typename T, int PROB>
__device__ __forceinline__ void sb_update(T &x, T &y, T F, T c0, T c1, T c2, T c3, T h) {
  const bool conf = (x < (T)0 ? -x : x) > (T)1;
  if (PROB == PROB_UNFUSED) {
    // mat_mult then confine (src/kernels.cu:198-207, 158-165); the X/Y
    // update of this step already ran in k_xy_update.
    y = fmaT(c2, F, y);
    if (conf) {
      x = signbit(x) ? (T)-1 : (T)1;
      y = (T)0;
    }
  } else if (PROB == PROB_MAXCUT) {
    if (conf) {
      x = signbit(x) ? (T)-1 : (T)1;
      y = (T)0;
    } else {
      y = fmaT(c2, F, y);
    }
    x = fmaT(c0, y, x);
    y = fmaT(c1, x, y);
  } else {
    if (conf) {
      x = signbit(x) ? (T)-1 : (T)1;
      y = (T)0;
    } else {
      const T q = fmaT(c3, h, (T)2 * F);
      y = fmaT(c0, c2 * q, y);
      x = fmaT(c0, y, x);
      y = fmaT(c1, x, y);
    }
  }
}

template <typename T, int GATHER>
__device__ __forceinline__ T sb_accum(T F, T v, T xj, T coupling) {
  if (GATHER == G_DSB) return signbit(xj) ? F + v : F - v;  // == fma(-v, sgn xj, F)
  return fmaT(-v, xj * coupling, F);
}

// Own-row state access, optionally non-temporal (streamed once per step, so
// it need not displace the sign words the gathers re-read from cache).
template <int NT>
__device__ __forceinline__ float ld_state(const float *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <int NT>
__device__ __forceinline__ void st_state(float *p, float v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
template <int NT>
__device__ __forceinline__ uint32_t ld_state(const uint32_t *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <int NT>
__device__ __forceinline__ void st_state(uint32_t *p, uint32_t v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
template <int NT>
__device__ __forceinline__ uint16_t ld_state(const uint16_t *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <int NT>
__device__ __forceinline__ void st_state(uint16_t *p, uint16_t v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// ------------------------------------------------- narrow fixed-point state
//
// Narrow-state dSB forms, code width QW = 12, 10 or 8 bits per variable.
// Element (x, y) is the 2 QW-bit unit u = (xq + B) | (yq + B) << QW of codes
// biased by B = 2^(QW-1):
//   xq = round(x 2^XS) in [-2^XS, 2^XS], XS = QW - 2, or +-(2^XS + 1) for
//        any |x| > 1 (the wall code),
//   yq = round(y 2^sy) clamped to [-YM, YM], YM = B - 1 (0 beside a wall
//        code),
// rounded to nearest even; the element is negative iff bit QW - 1 is clear.
// QW = 12: X codes +-1024 (wall +-1025), Y codes +-2047, 24-bit units;
// QW = 10: X codes +-256 (wall +-257), Y codes +-511, 20-bit units;
// QW = 8: X codes +-64 (wall +-65), Y codes +-127, 16-bit units.
// A lane keeps four elements (agents l, l + 64, l + 128, l + 192 of a
// 256-agent quad) in NP planes, the units' low 2 QW bits in order:
//   QW = 12, NP = 3: d0 = u0 | u1 << 24, d1 = u1 >> 8 | u2 << 16,
//                    d2 = u2 >> 16 | u3 << 8;
//   QW = 10, NP = 3: d0 = u0 | u1 << 20, d1 = u1 >> 12 | u2 << 8 | u3 << 28,
//                    d2 = u3 >> 4 (16 bits: the last plane holds 16-bit
//                    values, H16);
//   QW = 8,  NP = 2: d0 = u0 | u1 << 16, d1 = u2 | u3 << 16.
//
// Decoding and rounding go through magic constants.  M = 2^23 (bits
// 0x4b000000): M + k for an integer 0 <= k < 2^23 is exact with bits
// bits(M) + k, so bits QW..22 are zero for k < 2^QW; and for v s + B in
// [0, 2^23), fma(v, s, M + B) rounds v s + M + B to an integer (ulp 1, ties
// to even), whose low QW bits are the biased code round(v s) + B.
template <int QW> struct QFmt;
template <> struct QFmt<12> {
  static constexpr int NP = 3, XS = 10;
  static constexpr bool H16 = false;
  static constexpr uint32_t MASK = 0xfffu, SIGN = 0x800u;
  static constexpr uint32_t WALL_NEG = 0x800000u | 1023u, WALL_POS = 0x800000u | 3073u;
};
template <> struct QFmt<10> {
  static constexpr int NP = 3, XS = 8;
  static constexpr bool H16 = true;
  static constexpr uint32_t MASK = 0x3ffu, SIGN = 0x200u;
  static constexpr uint32_t WALL_NEG = 0x80000u | 255u, WALL_POS = 0x80000u | 769u;
};
template <> struct QFmt<8> {
  static constexpr int NP = 2, XS = 6;
  static constexpr bool H16 = false;
  static constexpr uint32_t MASK = 0xffu, SIGN = 0x80u;
  static constexpr uint32_t WALL_NEG = 0x8000u | 63u, WALL_POS = 0x8000u | 193u;
};

// Planes of the state array P (NP planes of ps lanes each, plane p at dword
// p ps; a 16-bit last plane holds ps 16-bit values from there): value o of
// plane p, read and written as its width.
template <int QW>
__device__ __forceinline__ uint32_t qn_plane_ld(const uint32_t *P, int64_t ps, int p, int64_t o) {
  if (QFmt<QW>::H16 && p == QFmt<QW>::NP - 1)
    return reinterpret_cast<const uint16_t *>(P + p * ps)[o];
  return P[p * ps + o];
}
template <int QW>
__device__ __forceinline__ void qn_plane_st(uint32_t *P, int64_t ps, int p, int64_t o, uint32_t v) {
  if (QFmt<QW>::H16 && p == QFmt<QW>::NP - 1)
    reinterpret_cast<uint16_t *>(P + p * ps)[o] = (uint16_t)v;
  else
    P[p * ps + o] = v;
}

// Per-launch constants: step coefficients; the Y code scale 2^sy, the
// largest |y| with a code (YM 2^-sy); the decoding offsets (of 2^23 + field,
// and of the field itself for byte fields); the field mask and the two wall
// units.  The mask, the wall units and the offsets are held in registers
// (opaque to the compiler), so that a field's decode is one and-or and the
// constants are not rebuilt for every row.  All are small integers times
// powers of two, so every product below is exact.
struct Q12K {
  float c0, c1, c2, yscale, ylim, yinv, xoff, yoff, xbo, ybo, mb;
  float ys8, mb8;  // 2^(sy+8) and 2^23 + 256 B: the stochastic Y encoder's
  uint32_t mask, wpos, wneg;
};

template <int QW>
__device__ __forceinline__ Q12K qn_consts(float c0, float c1, float c2, int qshift) {
  constexpr float B = (float)(1 << (QW - 1));
  Q12K k;
  k.c0 = c0;
  k.c1 = c1;
  k.c2 = c2;
  k.yscale = __builtin_ldexpf(1.0f, qshift);
  k.yinv = __builtin_ldexpf(1.0f, -qshift);
  k.ylim = (B - 1.0f) * k.yinv;
  float xoff = -(8388608.0f + B) * (1.0f / (float)(1 << QFmt<QW>::XS));  // -(2^23 + B) 2^-XS
  float xbo = -B * (1.0f / (float)(1 << QFmt<QW>::XS));                  // -B 2^-XS
  k.yoff = -(8388608.0f + B) * k.yinv;                                   // -(2^23 + B) 2^-sy
  k.ybo = -B * k.yinv;                                                   // -B 2^-sy
  uint32_t mask = QFmt<QW>::MASK, wpos = QFmt<QW>::WALL_POS, wneg = QFmt<QW>::WALL_NEG;
  float mb = 8388608.0f + B;  // 2^23 + B: the rounding addend of the encoder
  float mb8 = 8388608.0f + 256.0f * B;
  asm volatile("" : "+v"(mask), "+v"(wpos), "+v"(wneg), "+v"(xoff), "+v"(xbo), "+v"(mb));
  k.mb = mb;
  k.ys8 = k.yscale * 256.0f;
  k.mb8 = mb8;
  k.xoff = xoff;
  k.xbo = xbo;
  k.mask = mask;
  k.wpos = wpos;
  k.wneg = wneg;
  return k;
}

// The four units of a lane's NP dwords; bits above 2 QW are not cleared.
template <int QW>
__device__ __forceinline__ void qn_units(const uint32_t d[QFmt<QW>::NP], uint32_t u[4]) {
  if constexpr (QW == 12) {
    u[0] = d[0];                                       // bits 0..23
    u[1] = __builtin_amdgcn_alignbit(d[1], d[0], 24);  // bits 24..47
    u[2] = __builtin_amdgcn_alignbit(d[2], d[1], 16);  // bits 48..71
    u[3] = d[2] >> 8;                                  // bits 72..95
  } else if constexpr (QW == 10) {
    u[0] = d[0];                                       // bits 0..19
    u[1] = __builtin_amdgcn_alignbit(d[1], d[0], 20);  // bits 20..39
    u[2] = d[1] >> 8;                                  // bits 40..59
    u[3] = __builtin_amdgcn_alignbit(d[2], d[1], 28);  // bits 60..79
  } else {
    u[0] = d[0];
    u[1] = d[0] >> 16;
    u[2] = d[1];
    u[3] = d[1] >> 16;
  }
}

// x = (field - B) 2^-XS as fma(2^23 + field, 2^-XS, -(2^23 + B) 2^-XS): both
// terms and the result are exact.  Likewise y with 2^-sy.  (The Y field is
// shifted down first: at bits QW..2 QW-1 its top bit could be the float's
// exponent bit.)
template <int QW>
__device__ __forceinline__ void qn_decode(uint32_t u, const Q12K &k, float &x, float &y) {
  const float fx = __uint_as_float((u & k.mask) | 0x4b000000u);
  const float fy = __uint_as_float(((u >> QW) & k.mask) | 0x4b000000u);
  x = __builtin_fmaf(fx, 1.0f / (float)(1 << QFmt<QW>::XS), k.xoff);
  y = __builtin_fmaf(fy, k.yinv, k.yoff);
}

// Element j of a lane's NP dwords, decoded.  8-bit codes are whole bytes:
// each field converts to float in one operation (byte b of a dword:
// v_cvt_f32_ubyte<b>), then x = (field - B) 2^-XS = fma(field, 2^-XS, -B 2^-XS)
// exactly; 12- and 10-bit codes go through qn_units / qn_decode.
template <int QW>
__device__ __forceinline__ void qn_decode_at(const uint32_t d[QFmt<QW>::NP], int j,
                                             const Q12K &k, float &x, float &y) {
  if constexpr (QW == 8) {
    const uint32_t w = d[j >> 1];
    const int sh = (j & 1) * 16;
    const float fx = (float)((w >> sh) & 0xffu), fy = (float)((w >> (sh + 8)) & 0xffu);
    x = __builtin_fmaf(fx, 1.0f / (float)(1 << QFmt<QW>::XS), k.xbo);
    y = __builtin_fmaf(fy, k.yinv, k.ybo);
  } else {
    uint32_t u[4];
    qn_units<QW>(d, u);
    qn_decode<QW>(u[j], k, x, y);
  }
}

// Unit of an updated (x, y), bits above 2 QW unspecified, and the element's
// sign.  x 2^XS stays within [-2^XS, 2^XS] while |x| <= 1; y is clamped to
// +-YM codes first (clamping y or y 2^sy is the same, the scale being a power
// of two).  Both biased codes lie in [1, 2^QW - 1], so the rounded floats'
// bits QW..22 are zero and one shift-or joins them.  Elements beyond the
// wall take the wall X code and a 0 Y code.  The sign: f = fma(x, 2^XS,
// 2^23 + B) is monotonic in x and equals 2^23 + B exactly where the X code
// is 0, so the code (or, beyond the wall, x) is negative iff f < 2^23 + B,
// i.e. iff bits(f) < bits(2^23 + B) as signed integers (a negative f, for x
// far below -1, has negative bits).  Branch-free.
//
// CLAMP = false leaves the Y clamp out.  That is exact after a fused MaxCut
// step whenever YM 2^-sy >= 2/dt + dt (the rule that picks sy; the solver
// refuses a dt for which no sy >= 0 meets it): an element that stays in the
// box has |x_new - x| <= 2 with both in [-1, 1], so c0 |y'| <= 2 + 2^-24 with
// c0 = fl(dt (p + 1)) >= fl(dt), and |y''| = |y' + c1 x_new| <= |y'| + fl(dt);
// then |y''| 2^sy <= YM + 1e-5, which rounds to at most YM.  An element that
// leaves the box takes the wall unit, whatever its y.
template <int QW, bool CLAMP = true>
__device__ __forceinline__ uint32_t qn_encode(float x, float y, const Q12K &k, bool &neg) {
  const uint32_t xb = __float_as_uint(__builtin_fmaf(x, (float)(1 << QFmt<QW>::XS), k.mb));
  const float yc = CLAMP ? __builtin_amdgcn_fmed3f(y, -k.ylim, k.ylim) : y;
  const uint32_t yb = __float_as_uint(__builtin_fmaf(yc, k.yscale, k.mb));
  // bits 0 .. 2 QW - 1 valid; bits 2 QW .. 23 clear (xb's bits QW .. 23 are
  // clear, bit 23 being the low exponent bit of 2^23, and yb << QW has no bit
  // between 2 QW and 31)
  uint32_t in = (yb << QW) | xb;
  neg = (int32_t)xb < (int32_t)(0x4b000000u | (1u << (QW - 1)));
  // computed for every element, so that the wall choice below is two selects
  // rather than a branch around the encoding
  asm volatile("" : "+v"(in));
  return __builtin_fabsf(x) > 1.0f ? (neg ? k.wneg : k.wpos) : in;
}

// The same with stochastic Y rounding (SIS_FORM_Q8S): the Y code is
// floor((w + b) / 256) for w the nearest-even integer to y 2^(sy+8) and b the
// element's dither byte, i.e. y 2^sy rounded down or up with the probability
// of its fraction (to 1/256).  fma(y, 2^(sy+8), 2^23 + 256 B) is exact before
// its one rounding, so its bits are those of 2^23 plus 256 B + w; b added as
// an integer carries into bits 8..15, which then hold the biased Y code (the
// unit's bits 8..15), and the fraction bits 0..7 are masked off.  No clamp: a
// live element has |y| 2^sy <= YM + 1e-5 under the Y bound rule, so
// 256 YM + 255 bounds w + b - 256 B from above and -256 YM from below.
template <int QW>
__device__ __forceinline__ uint32_t qn_encode_sr(float x, float y, const Q12K &k, bool &neg,
                                                 uint32_t b) {
  static_assert(QW == 8, "stochastic Y rounding is defined for the 8-bit codes");
  const uint32_t xb = __float_as_uint(__builtin_fmaf(x, (float)(1 << QFmt<QW>::XS), k.mb));
  const uint32_t ys = __float_as_uint(__builtin_fmaf(y, k.ys8, k.mb8)) + b;
  uint32_t in = (ys & 0xff00u) | xb;
  neg = (int32_t)xb < (int32_t)(0x4b000000u | (1u << (QW - 1)));
  asm volatile("" : "+v"(in));
  return __builtin_fabsf(x) > 1.0f ? (neg ? k.wneg : k.wpos) : in;
}

// Dither hash of the stochastic Y rounding: row key (node i, quad q of Q,
// step salt) and lane l; element j of the lane takes byte j.
__device__ __forceinline__ uint32_t q8s_salt(int64_t t) {
  return (uint32_t)t * 0x27D4EB2Fu + 0x165667B1u;
}
__device__ __forceinline__ uint32_t q8s_row_key(int64_t node, int64_t Q, int64_t q,
                                                uint32_t salt) {
  return ((uint32_t)node * (uint32_t)Q + (uint32_t)q) * 0x9E3779B1u ^ salt;
}
__device__ __forceinline__ uint32_t q8s_lane_hash(uint32_t key, int lane) {
  uint32_t h = key + (uint32_t)lane * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h;
}

template <int QW>
__device__ __forceinline__ uint32_t qn_encode(float x, float y, const Q12K &k) {
  bool neg;
  return qn_encode<QW, true>(x, y, k, neg);
}

template <int QW>
__device__ __forceinline__ bool qn_negative(uint32_t unit) {
  return (unit & QFmt<QW>::SIGN) == 0;
}

// NP planes from four units (their low 2 QW bits; upper bits ignored, except
// that 10-bit units must have bits 20..23 clear, as qn_encode's units and the
// wall units do: bits 24 and up of u2 leave d1 by the shift).
template <int QW>
__device__ __forceinline__ void qn_pack(const uint32_t u[4], uint32_t d[QFmt<QW>::NP]) {
  if constexpr (QW == 12) {
    d[0] = __builtin_amdgcn_perm(u[1], u[0], 0x04020100u);  // u0.b0 u0.b1 u0.b2 u1.b0
    d[1] = __builtin_amdgcn_perm(u[2], u[1], 0x05040201u);  // u1.b1 u1.b2 u2.b0 u2.b1
    d[2] = __builtin_amdgcn_perm(u[3], u[2], 0x06050402u);  // u2.b2 u3.b0 u3.b1 u3.b2
  } else if constexpr (QW == 10) {
    d[0] = (u[0] & 0xfffffu) | (u[1] << 20);
    d[1] = __builtin_amdgcn_ubfe(u[1], 12, 8) | (u[2] << 8) | (u[3] << 28);
    d[2] = u[3] >> 4;  // stored as its low 16 bits
  } else {
    d[0] = __builtin_amdgcn_perm(u[1], u[0], 0x05040100u);  // u0.b0 u0.b1 u1.b0 u1.b1
    d[1] = __builtin_amdgcn_perm(u[3], u[2], 0x05040100u);  // u2.b0 u2.b1 u3.b0 u3.b1
  }
}

template <int QW>
__device__ __forceinline__ void qn_unpack(const uint32_t d[QFmt<QW>::NP], float x[4], float y[4],
                                          const Q12K &k) {
  uint32_t u[4];
  qn_units<QW>(d, u);
#pragma unroll
  for (int j = 0; j < 4; j++) qn_decode<QW>(u[j], k, x[j], y[j]);
}

// The fused MaxCut dSB update of sb_update<float, PROB_MAXCUT> without
// branches: the same operations, selected.
__device__ __forceinline__ void sb_update_maxcut_sel(float &x, float &y, float F, float c0,
                                                     float c1, float c2) {
  const bool conf = __builtin_fabsf(x) > 1.0f;
  const float yk = __builtin_fmaf(c2, F, y);
  x = __builtin_amdgcn_fmed3f(x, -1.0f, 1.0f);  // sgn(x) where |x| > 1
  y = conf ? 0.0f : yk;
  x = __builtin_fmaf(c0, y, x);
  y = __builtin_fmaf(c1, x, y);
}

// The same update for loops that carry x in a fixed register across
// steps: the x FMA is the three-operand v_fma_f32, so it writes x's register
// from the clamped copy directly (the two-operand form the compiler picks
// needs a move there).  Same operations, same rounding, same result.
__device__ __forceinline__ void sb_update_maxcut_carried(float &x, float &y, float F, float c0,
                                                         float c1, float c2) {
  const bool conf = __builtin_fabsf(x) > 1.0f;
  const float yk = __builtin_fmaf(c2, F, y);
  const float xm = __builtin_amdgcn_fmed3f(x, -1.0f, 1.0f);  // sgn(x) where |x| > 1
  y = conf ? 0.0f : yk;
  asm("v_fma_f32 %0, %1, %2, %3" : "=v"(x) : "v"(c0), "v"(y), "v"(xm));
  y = __builtin_fmaf(c1, x, y);
}

}  // namespace sis
