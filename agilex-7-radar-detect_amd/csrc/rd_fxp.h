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
