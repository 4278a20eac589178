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

}  // namespace sis
