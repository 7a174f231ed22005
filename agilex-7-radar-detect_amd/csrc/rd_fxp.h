// rd_fxp.h — integer quantisation primitives for the gfx950 kernels.
//
// Device twins of the reference's fixed-point rules (model/cpp/fxp/fxp.hpp):
// round-to-nearest-even on the floor remainder (fxp.hpp:198-208), round first
// then saturate, flags taken on the rounded value (fxp.hpp:221-231), and the
// packed {im, re} complex word (fxp.hpp:288-297). Written for the ranges each
// kernel actually produces, so most paths stay in 32-bit VALU ops; the 64-bit
// forms are used only where a sum can exceed 32 bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rd {

__host__ __device__ __forceinline__ int32_t cre(uint32_t p) { return (int32_t)(int16_t)(p & 0xFFFFu); }
__host__ __device__ __forceinline__ int32_t cim(uint32_t p) { return (int32_t)(int16_t)(p >> 16); }
__host__ __device__ __forceinline__ uint32_t cpack(int32_t re, int32_t im) {
  return ((uint32_t)(uint16_t)im << 16) | (uint32_t)(uint16_t)re;
}

// Saturate to int16 and record {sat_pos<<1 | sat_neg}.
__device__ __forceinline__ int32_t sat16(int32_t r, uint32_t& fl) {
  fl |= (r > 32767 ? 2u : 0u) | (r < -32768 ? 1u : 0u);
  return r > 32767 ? 32767 : (r < -32768 ? -32768 : r);
}

// Round-to-nearest-even by `sh` in {0, 1} for a sum of two int16 values:
// (v + ((v >> 1) & sh)) >> sh. For sh = 1 this adds the bit that breaks a tie
// toward the even quotient; for sh = 0 it is the identity.
__device__ __forceinline__ int32_t rne_sh01(int32_t v, uint32_t sh) {
  return (v + ((v >> 1) & (int32_t)sh)) >> sh;
}

// Round-to-nearest-even by 15 for |v| < 2^31 (Q2.30 -> Q1.15):
// q = floor(v / 2^15), rem = v mod 2^15; q + 1 iff rem + (q & 1) > 2^14.
__device__ __forceinline__ int32_t rne15_i32(int32_t v) {
  const int32_t q = v >> 15;
  const int32_t rem = v & 0x7FFF;
  return q + ((rem + (q & 1)) > 0x4000 ? 1 : 0);
}

// The same for a 64-bit exact accumulator (|v| < 2^47); the quotient of any
// value that can later pass the 16-bit saturation fits in 32 bits, and larger
// ones are clamped before the narrowing so the flags stay exact.
__device__ __forceinline__ int64_t rne15_i64(int64_t v) {
  const int64_t q = v >> 15;
  const int64_t rem = v & 0x7FFF;
  return q + ((rem + (q & 1)) > 0x4000 ? 1 : 0);
}
__device__ __forceinline__ int32_t sat16_i64(int64_t r, uint32_t& fl) {
  fl |= (r > 32767 ? 2u : 0u) | (r < -32768 ? 1u : 0u);
  return r > 32767 ? 32767 : (r < -32768 ? -32768 : (int32_t)r);
}

// ---------------------------------------------------------------------------
// General forms, any width w and shift s (fxp.hpp:87-231): wrapping 64-bit
// arithmetic, w-bit range endpoints (w >= 64: the full range; w < 2: {0}),
// floor shift, round half to even / half up on the floor remainder, round
// first then saturate with the flags of the rounded value. The stage kernels
// use the specialised forms above; these define what the specialisations must
// equal and back the element-wise fxp entry points.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ int64_t fx_add(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
__host__ __device__ __forceinline__ int64_t fx_sub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
__host__ __device__ __forceinline__ int64_t fx_neg(int64_t a) { return (int64_t)(0ull - (uint64_t)a); }
__host__ __device__ __forceinline__ int64_t fx_top(uint32_t w) {
  return w >= 64 ? INT64_MAX : (w < 2 ? 0 : (int64_t)((1ull << (w - 1)) - 1ull));
}
__host__ __device__ __forceinline__ int64_t fx_bottom(uint32_t w) {
  return w >= 64 ? INT64_MIN : (w < 2 ? 0 : fx_neg((int64_t)(1ull << (w - 1))));
}
__host__ __device__ __forceinline__ uint32_t fx_flags(int64_t v, uint32_t w) {
  return (v > fx_top(w) ? 2u : 0u) | (v < fx_bottom(w) ? 1u : 0u);
}
__host__ __device__ __forceinline__ int64_t fx_sat(int64_t v, uint32_t w) {
  return v > fx_top(w) ? fx_top(w) : (v < fx_bottom(w) ? fx_bottom(w) : v);
}
__host__ __device__ __forceinline__ int64_t fx_floor_shift(int64_t v, uint32_t s) {
  return s == 0 ? v : (s >= 64 ? (v < 0 ? -1 : 0) : (v >> s));
}
__host__ __device__ __forceinline__ int64_t fx_rne(int64_t v, uint32_t s) {
  if (s == 0) return v;
  if (s >= 64) return 0;
  const int64_t q = v >> s;
  const uint64_t rem = (uint64_t)v & ((1ull << s) - 1ull), half = 1ull << (s - 1);
  return rem > half ? fx_add(q, 1) : (rem < half ? q : fx_add(q, q & 1));
}
__host__ __device__ __forceinline__ int64_t fx_rhu(int64_t v, uint32_t s) {
  if (s == 0) return v;
  if (s >= 64) return 0;
  return fx_add(v, (int64_t)(1ull << (s - 1))) >> s;
}
__host__ __device__ __forceinline__ int64_t fx_round_sat(int64_t v, uint32_t s, uint32_t w, uint32_t& fl) {
  const int64_t r = fx_rne(v, s);
  fl = fx_flags(r, w);
  return fx_sat(r, w);
}

__host__ __device__ __forceinline__ uint32_t bitrev(uint32_t v, uint32_t w) {
#ifdef __HIP_DEVICE_COMPILE__
  return w == 0 ? 0u : (__brev(v) >> (32 - w));
#else
  uint32_t r = 0;
  for (uint32_t i = 0; i < w; ++i)
    if ((v >> i) & 1u) r |= 1u << (w - 1 - i);
  return r;
#endif
}

__host__ __device__ __forceinline__ uint32_t ilog2u(uint32_t v) {
  uint32_t r = 0;
  while ((1u << r) < v) ++r;
  return r;
}

}  // namespace rd
