// rd_mfma.hip — the two dense contractions of the chain on gfx950 matrix cores:
// the complex beamformer (+ power) and the Hermitian covariance.
//
// Both are exact int16 x int16 sums computed from int8 partial products with
// v_mfma_i32_16x16x32_i8 and recombined in int64.
//
// Digit form. Every int16 value a is written a = 256*h + l + 128 with
//   h = a >> 8        in [-128, 127]
//   l = (a & 255) - 128 in [-128, 127]
// so both digits are int8. For K-term sums of products a_k * b_k:
//   sum a*b = 65536*HH + 256*(HL + LH) + LL + 128*(Sa + Sb) - 16384*K
// where HH = sum h_a h_b etc. are four int8 GEMMs with int32 accumulation
// (|HH| <= K * 2^14, no overflow for K < 2^17) and Sa, Sb are the plain sums of
// the operands over k. Padding entries carry zero digits and are not counted.
//
// Negation without overflow. -x of an int16 can be +32768; instead
// -x = ~x + 1, and the digits of ~x are the bitwise complements of the digits of
// x (~x = 256*(~h) + (~l) + 128). A negated operand therefore costs one NOT and
// a constant correction.
//
// Operand layout of v_mfma_i32_16x16x32_i8 (per lane l, 8 int8 per operand):
//   A[i][k]: i = l % 16, k = 8*(l/16) + 0..7
//   B[k][j]: j = l % 16, k = 8*(l/16) + 0..7
//   C[i][j]: j = l % 16, i = 4*(l/16) + r for register r = 0..3
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rd_fxp.h"
#include "rd_internal.h"

namespace rd {

typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4i mfma8(uint64_t a, uint64_t b, v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x32_i8((long)a, (long)b, c, 0, 0, 0);
}

// v_perm selectors taking byte `byte` of two words: [p0.b, p1.b, 0, 0] and
// [0, 0, p2.b, p3.b].
__device__ __forceinline__ uint32_t sel01(uint32_t byte) { return 0x0C0C0000u | ((byte + 4) << 8) | byte; }
__device__ __forceinline__ uint32_t sel23(uint32_t byte) { return ((byte + 4) << 24) | (byte << 16) | 0x0C0Cu; }
__device__ __forceinline__ uint32_t gather4s(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3, uint32_t s01,
                                             uint32_t s23) {
  return __builtin_amdgcn_perm(p1, p0, s01) | __builtin_amdgcn_perm(p3, p2, s23);
}
// Bytes `byte` of four words, gathered into one word (word q -> byte q).
__device__ __forceinline__ uint32_t gather4(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3, uint32_t byte) {
  return gather4s(p0, p1, p2, p3, sel01(byte), sel23(byte));
}

// Digits of 8 packed complex values: H/L of the real part (bytes 1/0) and of the
// imaginary part (bytes 3/2), one int8 per value. `vmask` zeroes padded entries.
struct Dig8 {
  uint64_t hr, lr, hi, li;
};
__device__ __forceinline__ Dig8 digits8(const uint32_t* p, uint64_t vmask) {
  const uint64_t hr = (uint64_t)gather4(p[0], p[1], p[2], p[3], 1) | ((uint64_t)gather4(p[4], p[5], p[6], p[7], 1) << 32);
  const uint64_t lr = (uint64_t)gather4(p[0], p[1], p[2], p[3], 0) | ((uint64_t)gather4(p[4], p[5], p[6], p[7], 0) << 32);
  const uint64_t hi = (uint64_t)gather4(p[0], p[1], p[2], p[3], 3) | ((uint64_t)gather4(p[4], p[5], p[6], p[7], 3) << 32);
  const uint64_t li = (uint64_t)gather4(p[0], p[1], p[2], p[3], 2) | ((uint64_t)gather4(p[4], p[5], p[6], p[7], 2) << 32);
  const uint64_t flip = 0x8080808080808080ull;
  return Dig8{hr & vmask, (lr ^ flip) & vmask, hi & vmask, (li ^ flip) & vmask};
}

__device__ __forceinline__ int64_t combine(v4i hh, v4i x, v4i ll, int r) {
  return (int64_t)hh[r] * 65536 + (int64_t)x[r] * 256 + (int64_t)ll[r];
}

// round_sat(v, 15, 16) of v = 65536*hh + 256*x + ll + 128*s + c in 32-bit
// arithmetic. With 128*s = 2^15*(s >> 8) + 128*(s & 255), v = 2^15*A + R:
//   A = 2*hh + (s >> 8),  R = 256*x + ll + 128*(s & 255) + c,
// and R fits in int32 for the beamformer's ranges (|x| <= 2^22, |ll| <= 2^21,
// |c| <= 2^22). Round half to even is floor((v + 0x3FFF + p) / 2^15) with p
// the parity of floor(v / 2^15) = A + (R >> 15), that is
// A + ((R + 0x3FFF + p) >> 15); then saturate.
__device__ __forceinline__ int32_t rne15_sat_parts(int32_t hh, int32_t x, int32_t ll, int32_t s, int32_t c,
                                                   uint32_t& fl) {
  const int32_t R = (x << 8) + ll + ((s & 255) << 7) + c;
  const int32_t A = 2 * hh + (s >> 8);
  const int32_t p = ((R >> 15) ^ A) & 1;
  return sat16(A + ((R + 0x3FFF + p) >> 15), fl);
}

// The same when R = 256*x + ll + 128*s + c fits in int32 whole (at most two
// K steps, CH <= 32: |x| <= 2^21, |ll| <= 2^20, |s| <= 2^22 + 32, |c| <= 2^21,
// so |R| < 2^30.1):
// v = 2^15*(2*hh) + R, the parity of floor(v / 2^15) is bit 15 of R, and the
// rounding is 2*hh + ((R + 0x3FFF + bit15(R)) >> 15). The caller passes
// 128*s + c split as s7c_col = 128 * (the column's sum term) and
// s7c_beam = 128 * (the beam's sum term) + c.
__device__ __forceinline__ int32_t rne15_sat_small(int32_t hh, int32_t x, int32_t ll, int32_t s7c_col,
                                                   int32_t s7c_beam, uint32_t& fl) {
  const int32_t R = (x << 8) + ll + s7c_col + s7c_beam;
  const int32_t t = (R + 0x3FFF + (int32_t)__builtin_amdgcn_ubfe((uint32_t)R, 15, 1)) >> 15;
  return sat16(2 * hh + t, fl);
}

// Beamformer outputs as a compile-time set (bits: 1 y, 2 flags, 4 power,
// 8 exact accumulators); -1 reads the set from the job's pointers.
constexpr int kBfOutY = 1, kBfOutF = 2, kBfOutP = 4, kBfOutA = 8;

__host__ __device__ inline int bf_out_mask(const BfJob& J) {
  return (J.y ? kBfOutY : 0) | (J.flags ? kBfOutF : 0) | (J.power ? kBfOutP : 0) | (J.acc ? kBfOutA : 0);
}

// ============================================================================
// Beamformer: y[b][col] = round_sat(sum_a x[col][a] * W[b][a], 15, 16), no
// conjugate (beamformer_model.hpp:162-179), and p = yr^2 + yi^2
// (covar_model.hpp:99-102). As two real GEMMs over K = 2*CH:
//   Im: A = [Wi | Wr], B = [Xr ; Xi]
//   Re: A = [Wr | Wi], B = [Xr ; ~Xi], plus sum_a Wi   (-wi*xi = wi*~xi + wi)
// K index k < CHp is the real part of channel k, k >= CHp the imaginary part of
// channel k - CHp (CHp = CH rounded up to 16).
// ============================================================================

// Weight digits in LDS: per beam tile and step, the four operands
// {Are_h, Are_l, Aim_h, Aim_l}, each as 64 lane-consecutive 8-byte words (so a
// wave's fragment read is one contiguous, conflict-free 512-byte run), then
// per beam: Sa[b] = sum_a (wr + wi), Wis[b] = sum_a wi.
__host__ __device__ inline size_t bf_frag_words(uint32_t bt, uint32_t steps) { return (size_t)bt * steps * 64 * 8; }

size_t bf_prep_bytes(uint32_t, uint32_t) { return 0; }

// Every block derives the weight digits itself (B x CH words, L2-resident),
// with the block's first nthr threads (whole waves).
__device__ __forceinline__ void bf_weights_to_lds(const uint32_t* __restrict__ W, uint32_t CH, uint32_t B, uint32_t chp,
                                  uint32_t steps, uint32_t bt_n, uint32_t* frag, int32_t* tail,
                                  uint32_t nthr = blockDim.x) {
  for (uint32_t f = threadIdx.x; f < bt_n * steps * 64; f += nthr) {
    const uint32_t lane = f & 63, s = (f >> 6) % steps, bt = (f >> 6) / steps;
    const uint32_t b = bt * 16 + (lane & 15);
    const uint32_t k0 = 32 * s + 8 * (lane >> 4);
    const bool imag = k0 >= chp;
    const uint32_t a0 = imag ? k0 - chp : k0;
    uint32_t p[8];
    uint64_t vmask = 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const bool ok = b < B && a0 + t < CH;
      p[t] = ok ? W[(size_t)b * CH + a0 + t] : 0u;
      vmask |= ok ? (0xFFull << (8 * t)) : 0ull;
    }
    const Dig8 d = digits8(p, vmask);
    // Are = [Wr | Wi], Aim = [Wi | Wr]
    uint64_t* o = reinterpret_cast<uint64_t*>(frag) + (size_t)(f >> 6) * 256 + lane;
    o[0] = imag ? d.hi : d.hr;
    o[64] = imag ? d.li : d.lr;
    o[128] = imag ? d.hr : d.hi;
    o[192] = imag ? d.lr : d.li;
  }
  // per-beam sums: four threads per beam, 8 independent loads in flight each
  for (uint32_t t = threadIdx.x; t < bt_n * 64; t += nthr) {
    const uint32_t b = t >> 2, part = t & 3;
    int32_t sa = 0, wis = 0;
    if (b < B)
      for (uint32_t a0 = 8 * part; a0 < CH; a0 += 32) {
        uint32_t w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) w[q] = a0 + q < CH ? W[(size_t)b * CH + a0 + q] : 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          sa += cre(w[q]) + cim(w[q]);
          wis += cim(w[q]);
        }
      }
    sa += __shfl_xor(sa, 1);
    sa += __shfl_xor(sa, 2);
    wis += __shfl_xor(wis, 1);
    wis += __shfl_xor(wis, 2);
    if (part == 0) {
      tail[2 * b] = sa;
      tail[2 * b + 1] = wis;
    }
  }
}

// A block barrier that orders LDS only: the fences are local-address-space
// ones, so waves with global stores in flight do not wait for them (a plain
// __syncthreads would also make them visible, a full vmcnt drain).
__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt((7 << 4) | 0xC00F);  // lgkmcnt(0) only
  __builtin_amdgcn_s_barrier();
  __asm__ volatile("" ::: "memory");
}

// global_load_lds_dwordx4: lane l's 16 bytes at src go to LDS dst + 16 l (dst
// wave-uniform). Written as asm so the compiler's wait insertion, which would
// otherwise treat every later LDS read of the function as a possible reader
// of the transfer, leaves the compute waves alone; the loader counts its own
// transfers with wait_vmcnt. No register receives data, so nothing is left
// in flight behind the compiler's back.
__device__ __forceinline__ void lds_dma16(const void* src, void* dst) {
  const uint32_t m0 = __builtin_amdgcn_readfirstlane(
      (uint32_t)(uintptr_t)(__attribute__((address_space(3))) void*)dst);
  __asm__ volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(m0), "v"(src) : "memory", "m0");
}

// s_waitcnt vmcnt(n) for a wave-uniform n <= 40 (the immediate is a constant:
// vmcnt low bits 3:0, high bits 15:14; expcnt 7 and lgkmcnt 15 do not wait).
__device__ __forceinline__ void wait_vmcnt(uint32_t n) {
#define RD_VMW(k) \
  case k: __builtin_amdgcn_s_waitcnt(((k) & 15) | (((k) >> 4) << 14) | (7 << 4) | (15 << 8)); break;
  switch (n) {
    RD_VMW(1) RD_VMW(2) RD_VMW(3) RD_VMW(4) RD_VMW(5) RD_VMW(6) RD_VMW(7) RD_VMW(8)
    RD_VMW(9) RD_VMW(10) RD_VMW(11) RD_VMW(12) RD_VMW(13) RD_VMW(14) RD_VMW(15) RD_VMW(16)
    RD_VMW(17) RD_VMW(18) RD_VMW(19) RD_VMW(20) RD_VMW(21) RD_VMW(22) RD_VMW(23) RD_VMW(24)
    RD_VMW(25) RD_VMW(26) RD_VMW(27) RD_VMW(28) RD_VMW(29) RD_VMW(30) RD_VMW(31) RD_VMW(32)
    RD_VMW(33) RD_VMW(34) RD_VMW(35) RD_VMW(36) RD_VMW(37) RD_VMW(38) RD_VMW(39) RD_VMW(40)
    default: __builtin_amdgcn_s_waitcnt((7 << 4) | (15 << 8)); break;  // vmcnt(0)
  }
#undef RD_VMW
}

// LDS of a beamformer block for at most bt beam tiles and ch channels: weight
// digits, per-beam sums, and a column tile [TC][ch + 1] (odd row stride:
// conflict-free).
template <int STEPS, int TC>
struct BfLds {
  static size_t words(uint32_t bt, uint32_t ch) { return bf_frag_words(bt, STEPS) + (size_t)bt * 32 + (size_t)TC * (ch + 1); }
};

// One tile of TC columns of job J; the weights of J are in frag/tail (they
// may have been written since the caller's last barrier: the barrier below
// orders them). The caller issues a block barrier before each tile.
// CHMAJOR: X is the FFT output layout X[ch][col] (each channel row of the tile
// is TC contiguous words), and the tile is also written out as the aligned
// copy Xa[col][ch] (the alignment stage) when J.xa_out is set. Otherwise X is
// the aligned layout and the tile is one contiguous run.
// FULL: a whole tile of a CPI whose channels fill the steps and whose beams
// fill their tiles (no bounds to test).
template <int STEPS, int TC, bool CHMAJOR, int OUT, bool FULL>
__device__ __forceinline__ void bf_tile(const BfJob& J, uint64_t tile, const uint32_t* frag, const int32_t* tail,
                                        uint32_t* xt) {
  const int out = OUT >= 0 ? OUT : bf_out_mask(J);
  constexpr uint32_t chp = STEPS * 16;
  constexpr uint32_t TILES_PER_WAVE = TC / 64;
  const uint32_t CH = FULL ? STEPS * 16 : J.CH, B = J.B, bt_n = J.bt_n;
  const uint64_t cols = J.cols;
  const uint32_t* __restrict__ X = J.X;
  const uint32_t rs = CH + 1;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t grp = lane >> 4, cl = lane & 15;
  const int32_t kconst = 16384 * 2 * (int32_t)CH;
  const uint64_t c0 = tile * TC;
  const uint32_t ncol = (uint32_t)min<uint64_t>(TC, cols - c0);
  // Staging: a thread's loads (at most chp*TC/256) are all issued before the
  // first LDS write.
  constexpr uint32_t ITER = chp * TC / 256;
  uint32_t v[ITER];
  if (CHMAJOR) {
#pragma unroll
    for (uint32_t k = 0; k < ITER; ++k) {
      const uint32_t idx = threadIdx.x + 256 * k, a = idx / TC, j = idx % TC;
      v[k] = (FULL || (a < CH && j < ncol)) ? X[(size_t)a * cols + c0 + j] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < ITER; ++k) {
      const uint32_t idx = threadIdx.x + 256 * k, a = idx / TC, j = idx % TC;
      if (FULL || (a < CH && j < ncol)) xt[j * rs + a] = v[k];
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < ITER; ++k) {
      const uint32_t idx = threadIdx.x + 256 * k;
      v[k] = idx < ncol * CH ? X[c0 * CH + idx] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < ITER; ++k) {
      const uint32_t idx = threadIdx.x + 256 * k;
      if (idx < ncol * CH) xt[(idx / CH) * rs + idx % CH] = v[k];
    }
  }
  __syncthreads();
  if (CHMAJOR && J.xa_out) {
    if (CH == chp && ncol == TC && (reinterpret_cast<uintptr_t>(J.xa_out) & 15u) == 0) {
      // a full tile of a CPI whose channel count fills the steps (and a
      // 16-byte aligned output): four consecutive words of one column per
      // thread and 16-byte stores, the divisions by the compile-time count
      constexpr uint32_t PER = TC * chp / 1024;  // quads per thread (TC * chp is a multiple of 1024)
      uint4* dst = reinterpret_cast<uint4*>(J.xa_out + c0 * chp);
#pragma unroll
      for (uint32_t it = 0; it < PER; ++it) {
        const uint32_t q = threadIdx.x + 256 * it, w = 4 * q, col = w / chp, a = w % chp;
        const uint32_t* src = xt + col * rs + a;
        dst[q] = uint4{src[0], src[1], src[2], src[3]};
      }
    } else {
      for (uint32_t w = threadIdx.x; w < ncol * CH; w += 256) J.xa_out[c0 * CH + w] = xt[(w / CH) * rs + w % CH];
    }
  }

  for (uint32_t t = 0; t < TILES_PER_WAVE; ++t) {
    const uint32_t jl = (wave * TILES_PER_WAVE + t) * 16;
    if (jl >= ncol) break;
    const uint32_t j = jl + cl;
    const bool col_ok = FULL || j < ncol;
    const uint64_t col = c0 + j;
    // B fragments for every step: Bim = [Xr; Xi] digits, Bre = [Xr; ~Xi] digits.
    uint64_t bih[STEPS], bil[STEPS], brh[STEPS], brl[STEPS];
    int32_t ps_r = 0, ps_i = 0;
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const uint32_t k0 = 32 * s + 8 * grp;
      const bool imag = k0 >= chp;
      const uint32_t a0 = imag ? k0 - chp : k0;
      uint32_t p[8];
      uint64_t vmask = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const bool ok = FULL || (col_ok && a0 + q < CH);
        p[q] = ok ? xt[j * rs + a0 + q] : 0u;
        vmask |= ok ? (0xFFull << (8 * q)) : 0ull;
      }
      // only the digits this lane's K rows use: the real part's bytes 1/0 or
      // the imaginary part's 3/2 (per-lane v_perm selectors), and its sum
      const uint32_t bh = imag ? 3u : 1u, bl = bh - 1u;
      const uint32_t sh01 = sel01(bh), sh23 = sel23(bh), sl01 = sel01(bl), sl23 = sel23(bl);
      const uint64_t h = ((uint64_t)gather4s(p[0], p[1], p[2], p[3], sh01, sh23) |
                          ((uint64_t)gather4s(p[4], p[5], p[6], p[7], sh01, sh23) << 32)) &
                         vmask;
      const uint64_t l = (((uint64_t)gather4s(p[0], p[1], p[2], p[3], sl01, sl23) |
                           ((uint64_t)gather4s(p[4], p[5], p[6], p[7], sl01, sl23) << 32)) ^
                          0x8080808080808080ull) &
                         vmask;
      int32_t sp = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) sp += __builtin_amdgcn_sbfe((int32_t)p[q], imag ? 16u : 0u, 16u);
      // Bim = [Xr; Xi], Bre = [Xr; ~Xi]: the complement of masked digits is
      // an XOR with the mask
      const uint64_t inv = imag ? vmask : 0ull;
      bih[s] = h;
      bil[s] = l;
      brh[s] = h ^ inv;
      brl[s] = l ^ inv;
      if (imag)
        ps_i += sp;
      else
        ps_r += sp;
    }
    // column sums over all channels: the four lane groups of a column
    ps_r += __shfl_xor(ps_r, 16);
    ps_r += __shfl_xor(ps_r, 32);
    ps_i += __shfl_xor(ps_i, 16);
    ps_i += __shfl_xor(ps_i, 32);
    const int32_t sb_im = ps_r + ps_i;
    const int32_t sb_re = ps_r - ps_i - (int32_t)CH;

    for (uint32_t bt = 0; bt < bt_n; ++bt) {
      v4i rhh = {0, 0, 0, 0}, rx = {0, 0, 0, 0}, rll = {0, 0, 0, 0};
      v4i ihh = {0, 0, 0, 0}, ix = {0, 0, 0, 0}, ill = {0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < STEPS; ++s) {
        const uint64_t* a = reinterpret_cast<const uint64_t*>(frag) + ((size_t)bt * STEPS + s) * 256 + lane;
        const uint64_t are_h = a[0], are_l = a[64], aim_h = a[128], aim_l = a[192];
        rhh = mfma8(are_h, brh[s], rhh);
        rx = mfma8(are_h, brl[s], rx);
        rx = mfma8(are_l, brh[s], rx);
        rll = mfma8(are_l, brl[s], rll);
        ihh = mfma8(aim_h, bih[s], ihh);
        ix = mfma8(aim_h, bil[s], ix);
        ix = mfma8(aim_l, bih[s], ix);
        ill = mfma8(aim_l, bil[s], ill);
      }
      if (!col_ok) continue;
      const size_t o0 = (size_t)(bt * 16 + 4 * grp) * cols + col;  // beam 4 grp of the tile, this column
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t b = bt * 16 + 4 * grp + r;
        if (!FULL && b >= B) continue;
        const int32_t sa = tail[2 * b], wis = tail[2 * b + 1];
        uint32_t fr = 0, fi = 0;
        int32_t yr, yi;
        if constexpr (STEPS <= 2) {
          yr = rne15_sat_small(rhh[r], rx[r], rll[r], sb_re << 7, (sa << 7) + wis - kconst, fr);
          yi = rne15_sat_small(ihh[r], ix[r], ill[r], sb_im << 7, (sa << 7) - kconst, fi);
        } else {
          yr = rne15_sat_parts(rhh[r], rx[r], rll[r], sa + sb_re, wis - kconst, fr);
          yi = rne15_sat_parts(ihh[r], ix[r], ill[r], sa + sb_im, -kconst, fi);
        }
        const size_t o = o0 + (size_t)r * cols;
        if (out & kBfOutY) J.y[o] = cpack(yr, yi);
        if (out & kBfOutF) J.flags[o] = (uint8_t)(fr | (fi << 2));
        if (out & kBfOutP) J.power[o] = (uint32_t)(yr * yr) + (uint32_t)(yi * yi);
        if (out & kBfOutA) {
          J.acc[2 * o] = combine(rhh, rx, rll, r) + 128 * ((int64_t)sa + sb_re) - kconst + wis;
          J.acc[2 * o + 1] = combine(ihh, ix, ill, r) + 128 * ((int64_t)sa + sb_im) - kconst;
        }
      }
    }
  }
}

// Single CPI: a grid stride over the column tiles, the weight digits derived
// once per block.
template <int STEPS, int TC, bool CHMAJOR, int OUT>
__global__ __launch_bounds__(256, STEPS >= 3 ? 2 : 3) void bf_kernel(BfJob J) {
  extern __shared__ uint32_t lds[];
  uint32_t* frag = lds;
  int32_t* tail = reinterpret_cast<int32_t*>(lds + bf_frag_words(J.bt_n, STEPS));
  uint32_t* xt = lds + bf_frag_words(J.bt_n, STEPS) + (size_t)J.bt_n * 32;
  bf_weights_to_lds(J.W, J.CH, J.B, STEPS * 16, STEPS, J.bt_n, frag, tail);
  const uint64_t n_tiles = (J.cols + TC - 1) / TC;
  const bool full = J.CH == STEPS * 16 && J.B % 16 == 0;
  for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    __syncthreads();  // previous tile fully consumed
    if (STEPS == 1 && full && (tile + 1) * TC <= J.cols)  // (two bodies fit the registers at one step)
      bf_tile<STEPS, TC, CHMAJOR, OUT, STEPS == 1>(J, tile, frag, tail, xt);
    else
      bf_tile<STEPS, TC, CHMAJOR, OUT, false>(J, tile, frag, tail, xt);
  }
}

// Batched: a persistent grid over the column tiles of a class of CPIs (same
// step count); a block re-derives the weight digits when its CPI changes.
template <int STEPS, int TC, int OUT>
__global__ __launch_bounds__(256, STEPS >= 3 ? 2 : 3) void bf_batch_kernel(const BfJob* __restrict__ jobs,
                                                       const uint32_t* __restrict__ tile_job, uint32_t n_tiles,
                                                       uint32_t max_bt) {
  extern __shared__ uint32_t lds[];
  uint32_t* frag = lds;
  int32_t* tail = reinterpret_cast<int32_t*>(lds + bf_frag_words(max_bt, STEPS));
  uint32_t* xt = lds + bf_frag_words(max_bt, STEPS) + (size_t)max_bt * 32;
  uint32_t cur = ~0u;
  for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const uint32_t jid = tile_job[t];
    const BfJob J = jobs[jid];
    __syncthreads();  // previous tile (and weights) fully consumed
    if (jid != cur) {
      bf_weights_to_lds(J.W, J.CH, J.B, STEPS * 16, STEPS, J.bt_n, frag, tail);
      cur = jid;
    }
    const uint32_t lt = t - J.tile0;
    if (STEPS <= 2 && J.CH == STEPS * 16 && J.B % 16 == 0 && ((uint64_t)lt + 1) * TC <= J.cols)
      bf_tile<STEPS, TC, true, OUT, STEPS <= 2>(J, lt, frag, tail, xt);
    else
      bf_tile<STEPS, TC, true, OUT, false>(J, lt, frag, tail, xt);
  }
}

// ============================================================================
// Beamformer over the FFT layout for CPIs whose channels fill 3 or 4 K steps
// (48 or 64 channels), beams a multiple of 16 and columns a multiple of 128
// (at 32 channels the tiled bf_kernel measured faster):
// a persistent block of 8 compute waves and one loader wave. The loader moves
// each 128-column tile of X (one 512-byte row per channel) into an LDS ring of
// two slots with global_load_lds_dwordx4 (one instruction per channel pair)
// and counts its own transfers; the compute waves issue no vector-memory load
// in the tile loop, so their power and aligned-copy stores are never waited
// for. Per tile the compute waves copy the tile out as Xa[col][ch] (if
// requested) and build their B fragments from the slot, meet the loader at one
// LDS-only barrier (after which the slot is reloaded two tiles ahead), then run
// the MFMAs and the epilogue of their 16 columns over every beam tile.
// ============================================================================
constexpr uint32_t kBrTC = 128;        // columns per tile
constexpr uint32_t kBrBlk = 1024 + 16;  // bytes per channel-pair block (two rows of 128 words + padding)
constexpr uint32_t kBrRing = 2;

template <int STEPS>
struct BrGeo {
  static constexpr uint32_t CHP = STEPS * 16;
  static constexpr uint32_t NBLK = CHP / 2;  // loader instructions per tile
  static constexpr uint32_t SLOT = NBLK * kBrBlk;
  static_assert((kBrRing - 1) * NBLK <= 40, "wait_vmcnt covers the loader's counts");
  static size_t lds(uint32_t bt_n) {
    return (size_t)kBrRing * SLOT + bf_frag_words(bt_n, STEPS) * 4 + (size_t)bt_n * 32 * 4;
  }
};

// Byte offset of word `col` of channel row `ch` in a slot.
__device__ __forceinline__ uint32_t br_off(uint32_t ch, uint32_t col) {
  return (ch >> 1) * kBrBlk + (ch & 1) * 512 + col * 4;
}

bool bf_ring_fits(const BfJob& J, uint32_t steps) {
  return (steps == 3 || steps == 4) && J.CH == steps * 16 && J.B % 16 == 0 && J.cols % kBrTC == 0 &&
         J.cols / kBrTC > 0 && !J.acc && (reinterpret_cast<uintptr_t>(J.X) & 15u) == 0;
}

// The tiles of a block: t_k = blockIdx.x + k * gridDim.x over the launch's
// tile space. BATCH: tile t belongs to jobs[tile_job[t]] (its tiles are the
// contiguous range from its tile0); a block that moves to another CPI's tiles
// derives that CPI's weight digits between the tile's barrier and its
// products, behind one extra barrier that the loader also takes.
template <int STEPS, int OUT, bool BATCH>
__device__ __forceinline__ void bf_ring_body(const BfJob* __restrict__ jobs, const uint32_t* __restrict__ tile_job,
                                             uint64_t n_tiles, const BfJob& J0) {
  using G = BrGeo<STEPS>;
  constexpr uint32_t CHP = G::CHP;
  extern __shared__ uint32_t lds[];
  auto job_of = [&](uint64_t t) -> uint32_t {
    return BATCH ? __builtin_amdgcn_readfirstlane(tile_job[t]) : 0u;
  };
  auto job = [&](uint32_t jid) -> const BfJob& { return BATCH ? jobs[jid] : J0; };
  const uint32_t max_bt = J0.bt_n;  // the LDS layout's beam tiles (batch: the class maximum)
  unsigned char* const raw = reinterpret_cast<unsigned char*>(lds);
  uint32_t* frag = reinterpret_cast<uint32_t*>(raw + kBrRing * G::SLOT);
  int32_t* tail = reinterpret_cast<int32_t*>(frag + bf_frag_words(max_bt, STEPS));
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool loader = wave == 8;
  // this block's tiles: blockIdx.x + k * gridDim.x, k < nt
  const uint32_t nt = (uint32_t)((n_tiles - blockIdx.x + gridDim.x - 1) / gridDim.x);
  uint32_t cur = job_of(blockIdx.x);
  if (!loader) {
    const BfJob& J = job(cur);
    bf_weights_to_lds(J.W, CHP, J.B, CHP, STEPS, J.bt_n, frag, tail, 512);
  }

  if (loader) {
    // lane l: channel 2j + l / 32 of block j, columns 4 (l % 32) .. + 3
    const uint32_t half = lane >> 5, lc = (lane & 31) * 4;
    auto issue = [&](uint32_t k) {
      unsigned char* slot = raw + (k % kBrRing) * G::SLOT;
      const uint64_t t = (uint64_t)blockIdx.x + (uint64_t)k * gridDim.x;
      const BfJob& J = job(job_of(t));
      const uint64_t cols = J.cols, c0 = (t - (BATCH ? J.tile0 : 0u)) * kBrTC;
      const uint32_t* src = J.X + (size_t)half * cols + c0 + lc;
#pragma unroll
      for (uint32_t j = 0; j < G::NBLK; ++j) lds_dma16(src + (size_t)(2 * j) * cols, slot + j * kBrBlk);
    };
    uint32_t next = 0;
    for (; next < nt && next < kBrRing; ++next) issue(next);
    wait_vmcnt((next - 1) * G::NBLK);
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();  // tile 0 is in the ring
    for (uint32_t k = 0; k < nt; ++k) {
      wait_vmcnt(k + 1 < next ? (next - k - 2) * G::NBLK : 0u);  // tile k + 1 complete
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();  // the compute waves are done with slot k % kBrRing
      __asm__ volatile("" ::: "memory");
      if (next < nt) issue(next++);
      if (BATCH) {
        const uint32_t jid = job_of((uint64_t)blockIdx.x + (uint64_t)k * gridDim.x);
        if (jid != cur) {
          __builtin_amdgcn_s_barrier();  // the compute waves' weight barrier
          cur = jid;
        }
      }
    }
    return;
  }

  __syncthreads();  // weights in LDS, tile 0 in the ring
  const uint32_t grp = lane >> 4, cl = lane & 15;
  const int32_t kconst = 16384 * 2 * (int32_t)CHP;
  const uint32_t j = 16 * wave + cl;  // this lane's column of the tile
  for (uint32_t k = 0; k < nt; ++k) {
    const unsigned char* slot = raw + (k % kBrRing) * G::SLOT;
    const uint64_t t = (uint64_t)blockIdx.x + (uint64_t)k * gridDim.x;
    const uint32_t jid = job_of(t);
    const BfJob& J = job(jid);
    const uint64_t cols = J.cols;
    const uint64_t c0 = (t - (BATCH ? J.tile0 : 0u)) * kBrTC;
    if (J.xa_out) {
      // the aligned copy: 16-byte stores of four consecutive channels of one
      // column, consecutive threads on consecutive quads
      uint4* dst = reinterpret_cast<uint4*>(J.xa_out + c0 * CHP);
      constexpr uint32_t PER = kBrTC * CHP / 4 / 512;
#pragma unroll
      for (uint32_t it = 0; it < PER; ++it) {
        const uint32_t q = threadIdx.x + 512 * it, w = 4 * q, col = w / CHP, a = w % CHP;
        const unsigned char* sp = slot + br_off(a, col);
        dst[q] = uint4{*reinterpret_cast<const uint32_t*>(sp), *reinterpret_cast<const uint32_t*>(sp + 512),
                       *reinterpret_cast<const uint32_t*>(sp + kBrBlk),
                       *reinterpret_cast<const uint32_t*>(sp + kBrBlk + 512)};
      }
    }
    // B fragments for every step: Bim = [Xr; Xi] digits, Bre = [Xr; ~Xi] digits
    uint64_t bih[STEPS], bil[STEPS], brh[STEPS], brl[STEPS];
    int32_t ps_r = 0, ps_i = 0;
#pragma unroll
    for (int st = 0; st < STEPS; ++st) {
      const uint32_t k0 = 32 * st + 8 * grp;
      const bool imag = k0 >= CHP;
      const uint32_t a0 = imag ? k0 - CHP : k0;
      uint32_t p[8];
      const unsigned char* sp = slot + br_off(a0, j);
#pragma unroll
      for (int q = 0; q < 8; ++q)
        p[q] = *reinterpret_cast<const uint32_t*>(sp + (q >> 1) * kBrBlk + (q & 1) * 512);
      const uint32_t bh = imag ? 3u : 1u, bl = bh - 1u;
      const uint32_t sh01 = sel01(bh), sh23 = sel23(bh), sl01 = sel01(bl), sl23 = sel23(bl);
      const uint64_t h = (uint64_t)gather4s(p[0], p[1], p[2], p[3], sh01, sh23) |
                         ((uint64_t)gather4s(p[4], p[5], p[6], p[7], sh01, sh23) << 32);
      const uint64_t l = ((uint64_t)gather4s(p[0], p[1], p[2], p[3], sl01, sl23) |
                          ((uint64_t)gather4s(p[4], p[5], p[6], p[7], sl01, sl23) << 32)) ^
                         0x8080808080808080ull;
      int32_t sp16 = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) sp16 += __builtin_amdgcn_sbfe((int32_t)p[q], imag ? 16u : 0u, 16u);
      const uint64_t inv = imag ? ~0ull : 0ull;
      bih[st] = h;
      bil[st] = l;
      brh[st] = h ^ inv;
      brl[st] = l ^ inv;
      if (imag)
        ps_i += sp16;
      else
        ps_r += sp16;
    }
    ps_r += __shfl_xor(ps_r, 16);
    ps_r += __shfl_xor(ps_r, 32);
    ps_i += __shfl_xor(ps_i, 16);
    ps_i += __shfl_xor(ps_i, 32);
    const int32_t sb_im = ps_r + ps_i;
    const int32_t sb_re = ps_r - ps_i - (int32_t)CHP;
    lds_barrier();  // every wave is done with this slot (and with the previous tile's products)
    if (BATCH && jid != cur) {  // another CPI: its weight digits replace the previous ones
      bf_weights_to_lds(J.W, CHP, J.B, CHP, STEPS, J.bt_n, frag, tail, 512);
      __syncthreads();
      cur = jid;
    }

    const int out = OUT >= 0 ? OUT : bf_out_mask(J);
    const uint64_t col = c0 + j;
    const size_t lane_row = (size_t)(4 * grp) * cols;
    for (uint32_t bt = 0; bt < J.bt_n; ++bt) {
      v4i rhh = {0, 0, 0, 0}, rx = {0, 0, 0, 0}, rll = {0, 0, 0, 0};
      v4i ihh = {0, 0, 0, 0}, ix = {0, 0, 0, 0}, ill = {0, 0, 0, 0};
#pragma unroll
      for (int st = 0; st < STEPS; ++st) {
        const uint64_t* a = reinterpret_cast<const uint64_t*>(frag) + ((size_t)bt * STEPS + st) * 256 + lane;
        const uint64_t are_h = a[0], are_l = a[64], aim_h = a[128], aim_l = a[192];
        rhh = mfma8(are_h, brh[st], rhh);
        rx = mfma8(are_h, brl[st], rx);
        rx = mfma8(are_l, brh[st], rx);
        rll = mfma8(are_l, brl[st], rll);
        ihh = mfma8(aim_h, bih[st], ihh);
        ix = mfma8(aim_h, bil[st], ix);
        ix = mfma8(aim_l, bih[st], ix);
        ill = mfma8(aim_l, bil[st], ill);
      }
      const size_t o0 = (size_t)(bt * 16) * cols + lane_row + col;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t b = bt * 16 + 4 * grp + r;
        const int32_t sa = tail[2 * b], wis = tail[2 * b + 1];
        uint32_t fr = 0, fi = 0;
        const int32_t yr = rne15_sat_parts(rhh[r], rx[r], rll[r], sa + sb_re, wis - kconst, fr);
        const int32_t yi = rne15_sat_parts(ihh[r], ix[r], ill[r], sa + sb_im, -kconst, fi);
        const size_t o = o0 + (size_t)r * cols;
        if (out & kBfOutY) J.y[o] = cpack(yr, yi);
        if (out & kBfOutF) J.flags[o] = (uint8_t)(fr | (fi << 2));
        if (out & kBfOutP) J.power[o] = (uint32_t)(yr * yr) + (uint32_t)(yi * yi);
      }
    }
  }
}

template <int STEPS, int OUT>
__global__ __launch_bounds__(64 * 9) void bf_ring_kernel(BfJob J) {
  bf_ring_body<STEPS, OUT, false>(nullptr, nullptr, J.cols / kBrTC, J);
}

// Batched: `cls` carries the class's largest beam-tile count (the LDS layout).
template <int STEPS, int OUT>
__global__ __launch_bounds__(64 * 9) void bf_ring_batch_kernel(const BfJob* __restrict__ jobs,
                                                              const uint32_t* __restrict__ tile_job,
                                                              uint32_t n_tiles, BfJob cls) {
  bf_ring_body<STEPS, OUT, true>(jobs, tile_job, n_tiles, cls);
}

template <int STEPS>
static hipError_t launch_bf_ring(hipStream_t s, const BfJob& J) {
  const uint64_t tiles = J.cols / kBrTC;
  const dim3 grid((uint32_t)(tiles < 256 ? tiles : 256));  // one resident block per CU (LDS)
  const size_t lds = BrGeo<STEPS>::lds(J.bt_n);
  const int out = bf_out_mask(J);
  if (out == kBfOutP)
    bf_ring_kernel<STEPS, kBfOutP><<<grid, 64 * 9, lds, s>>>(J);
  else if (out == (kBfOutY | kBfOutF | kBfOutP))
    bf_ring_kernel<STEPS, kBfOutY | kBfOutF | kBfOutP><<<grid, 64 * 9, lds, s>>>(J);
  else
    bf_ring_kernel<STEPS, -1><<<grid, 64 * 9, lds, s>>>(J);
  return hipGetLastError();
}

// The class's tiles (128 columns each) on one resident block per CU.
hipError_t launch_bf_ring_batch(hipStream_t s, uint32_t steps, uint32_t max_bt, uint32_t n_cu, int out_mask,
                                const ClassLaunch& L) {
  if (L.n_tiles == 0) return hipSuccess;
  if (steps != 3 && steps != 4) return hipErrorInvalidValue;
  const dim3 grid(L.n_tiles < n_cu ? L.n_tiles : n_cu);
  BfJob cls{};
  cls.bt_n = max_bt;
  const BfJob* jobs = static_cast<const BfJob*>(L.jobs);
#define RD_BR(S)                                                                                           \
  {                                                                                                        \
    const size_t lds = BrGeo<S>::lds(max_bt);                                                              \
    if (out_mask == kBfOutP)                                                                               \
      bf_ring_batch_kernel<S, kBfOutP><<<grid, 64 * 9, lds, s>>>(jobs, L.tile_job, L.n_tiles, cls);        \
    else if (out_mask == (kBfOutY | kBfOutF | kBfOutP))                                                    \
      bf_ring_batch_kernel<S, kBfOutY | kBfOutF | kBfOutP><<<grid, 64 * 9, lds, s>>>(jobs, L.tile_job,     \
                                                                                      L.n_tiles, cls);     \
    else                                                                                                   \
      bf_ring_batch_kernel<S, -1><<<grid, 64 * 9, lds, s>>>(jobs, L.tile_job, L.n_tiles, cls);             \
  }
  if (steps == 3)
    RD_BR(3)
  else
    RD_BR(4)
#undef RD_BR
  return hipGetLastError();
}

template <int STEPS>
static hipError_t launch_bf_s(hipStream_t s, bool chmajor, const BfJob& J) {
  constexpr uint32_t TC = STEPS <= 2 ? 256 : 128;  // columns per tile
  const uint64_t tiles = (J.cols + TC - 1) / TC;
  const dim3 grid((uint32_t)(tiles < 2048 ? tiles : 2048));
  const size_t lds = BfLds<STEPS, TC>::words(J.bt_n, J.CH) * sizeof(uint32_t);
  const int out = bf_out_mask(J);
  if constexpr (STEPS == 3 || STEPS == 4) {
    if (chmajor && bf_ring_fits(J, STEPS)) return launch_bf_ring<STEPS>(s, J);
  }
  if (chmajor) {
    if (out == kBfOutP)
      bf_kernel<STEPS, TC, true, kBfOutP><<<grid, 256, lds, s>>>(J);
    else if (out == (kBfOutY | kBfOutF | kBfOutP))
      bf_kernel<STEPS, TC, true, kBfOutY | kBfOutF | kBfOutP><<<grid, 256, lds, s>>>(J);
    else
      bf_kernel<STEPS, TC, true, -1><<<grid, 256, lds, s>>>(J);
  } else {
    bf_kernel<STEPS, TC, false, -1><<<grid, 256, lds, s>>>(J);
  }
  return hipGetLastError();
}

static hipError_t launch_bf_any(hipStream_t s, bool chmajor, const BfJob& J) {
  if (J.cols == 0) return hipSuccess;
  switch ((J.CH + 15) / 16) {
    case 1: return launch_bf_s<1>(s, chmajor, J);
    case 2: return launch_bf_s<2>(s, chmajor, J);
    case 3: return launch_bf_s<3>(s, chmajor, J);
    default: return launch_bf_s<4>(s, chmajor, J);
  }
}

hipError_t launch_bf(hipStream_t s, const uint32_t* Xa, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                     uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power, int64_t* acc, void*) {
  const BfJob J{Xa, W, y, flags, power, acc, nullptr, n_cols, n_ant, n_beams, (n_beams + 15) / 16, 0};
  return launch_bf_any(s, false, J);
}

hipError_t launch_bf_chmajor(hipStream_t s, const uint32_t* X, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                             uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power, uint32_t* xa_out) {
  const BfJob J{X, W, y, flags, power, nullptr, xa_out, n_cols, n_ant, n_beams, (n_beams + 15) / 16, 0};
  return launch_bf_any(s, true, J);
}

template <int STEPS, int TC>
static hipError_t launch_bf_batch_s(hipStream_t s, uint32_t max_ch, uint32_t max_bt, int out, const ClassLaunch& L) {
  const size_t lds = BfLds<STEPS, TC>::words(max_bt, max_ch) * sizeof(uint32_t);
  const BfJob* jobs = static_cast<const BfJob*>(L.jobs);
  if (out == kBfOutP)
    bf_batch_kernel<STEPS, TC, kBfOutP><<<L.grid, 256, lds, s>>>(jobs, L.tile_job, L.n_tiles, max_bt);
  else if (out == (kBfOutY | kBfOutF | kBfOutP))
    bf_batch_kernel<STEPS, TC, kBfOutY | kBfOutF | kBfOutP><<<L.grid, 256, lds, s>>>(jobs, L.tile_job, L.n_tiles,
                                                                                    max_bt);
  else
    bf_batch_kernel<STEPS, TC, -1><<<L.grid, 256, lds, s>>>(jobs, L.tile_job, L.n_tiles, max_bt);
  return hipGetLastError();
}

hipError_t launch_bf_batch(hipStream_t s, uint32_t steps, uint32_t tc, uint32_t max_ch, uint32_t max_bt,
                           int out_mask, const ClassLaunch& L) {
  if (L.n_tiles == 0) return hipSuccess;
#define RD_BF_TC(S)                                                                            \
  return tc == 256 ? launch_bf_batch_s<S, 256>(s, max_ch, max_bt, out_mask, L)                 \
                   : (tc == 128 ? launch_bf_batch_s<S, 128>(s, max_ch, max_bt, out_mask, L)    \
                                : launch_bf_batch_s<S, 64>(s, max_ch, max_bt, out_mask, L));
  switch (steps) {
    case 1: RD_BF_TC(1)
    case 2: RD_BF_TC(2)
    case 3: RD_BF_TC(3)
    case 4: RD_BF_TC(4)
  }
#undef RD_BF_TC
  return hipErrorInvalidValue;
}

// ============================================================================
// Covariance: R[k][i<=j] = sum_m cross_power(x_i[m][k], x_j[m][k]) with
// cross_power(x, y) = x * conj(y) (covar_model.hpp:119-151), as two real GEMMs
// per bin over K = 2F (frames, real then imaginary part):
//   Re: A_i = [Xr_i | Xi_i],  B_j = [Xr_j | Xi_j]
//   Im: A_i = [Xi_i | ~Xr_i], B_j = [Xr_j | Xi_j], plus sum_m xi_j
// A block owns NB bins (NB * CHp = 64 channel rows) and their upper-triangle
// 16x16 channel tiles ("jobs"); each wave runs one job over one of FS frame
// splits. A split walks its frames 32 at a time: the block's threads load the
// 64 rows x 32 frames of a step two steps ahead (coalesced along the rows,
// two register sets in flight) and write each
// sample's int8 digits once into six LDS digit planes, [plane][row][frame]:
//   hr, lr, hi, li (digits of the real and imaginary part, low digits offset),
//   ~hr, ~lr       (the complemented real digits of the Im GEMM's A operand),
// so every MFMA operand is one 8-byte LDS read of 8 consecutive frames.
// Padded frames and channel rows hold zero bytes. Every wave turns its
// partial sums into partial int64 results (all correction terms are linear in
// the frames), and the splits of a job are added with LDS 64-bit atomics.
// ============================================================================
constexpr uint32_t kCovFS = 32;                // frames per step
constexpr uint32_t kCovRow = kCovFS + 8;       // bytes per plane row (padding spreads the banks)
constexpr uint32_t kCovPlane = 64 * kCovRow;   // bytes per plane
constexpr uint32_t kCovSplitBytes = 6 * kCovPlane + 64 * 2 * 4;  // planes + per-row frame sums

// One tile: bins [tile*nb, +nb) of job J. The caller issues a block barrier
// before each tile (the LDS of the previous tile is reused).
__device__ __forceinline__ void cov_tile(const CovJob& J, uint32_t tile, uint32_t nb, uint32_t nt, uint32_t fs_n,
                                         uint32_t* smem) {
  const uint32_t* __restrict__ Xa = J.Xa;
  const uint32_t CH = J.CH, F = J.F, N = J.N;
  const uint32_t chp = 64 / nb;
  const uint32_t ut = nt * (nt + 1) / 2;
  const uint32_t jobs = nb * ut;
  unsigned char* lds = reinterpret_cast<unsigned char*>(smem);
  unsigned long long* red = reinterpret_cast<unsigned long long*>(lds + (size_t)fs_n * kCovSplitBytes);  // [jobs][64][8]
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t grp = lane >> 4, cl = lane & 15;
  const uint32_t job = wave % jobs, fs = wave / jobs;
  const uint32_t k_base = tile * nb;
  const uint32_t bins_here = min(nb, N - k_base);
  unsigned char* planes = lds + (size_t)fs * kCovSplitBytes;
  int32_t* sums = reinterpret_cast<int32_t*>(planes + 6 * kCovPlane);  // [64 rows][2]: sum xr, sum xi

  const uint32_t bin = job / ut;
  uint32_t u = job % ut, ti = 0;
  while (u >= nt - ti) {
    u -= nt - ti;
    ++ti;
  }
  const uint32_t tj = ti + u;
  const bool live = bin < bins_here;
  const uint32_t ra = bin * chp + 16 * ti + cl, rb = bin * chp + 16 * tj + cl;  // operand rows of this lane

  if (fs_n > 1)
    for (uint32_t i = threadIdx.x; i < jobs * 64 * 8; i += blockDim.x) red[i] = 0;
  for (uint32_t i = threadIdx.x; i < fs_n * 128; i += blockDim.x)
    reinterpret_cast<int32_t*>(lds + (size_t)(i >> 7) * kCovSplitBytes + 6 * kCovPlane)[i & 127] = 0;

  const uint32_t nsteps = (F + kCovFS - 1) / kCovFS;
  const uint32_t sps = (nsteps + fs_n - 1) / fs_n;  // steps per split
  // Re: HH, HL, LL, LH; Im: HH, HL, LL, LH. Every digit product is at most
  // 2^14 in magnitude and a frame adds two of each kind (real and imaginary
  // K rows), so one accumulator grows by at most 2^15 per frame and stays
  // inside int32 for every F <= 65535 (HL and LH summed in one accumulator
  // would not past 32767 frames).
  v4i acc[8];
  for (int z = 0; z < 8; ++z) acc[z] = v4i{0, 0, 0, 0};
  int32_t nfr = 0;

  // Conversion tasks of a step: (row = lane, group g of 4 frames), g = 0..7,
  // spread over the split's waves; at most two per thread (jobs >= 4). Two
  // register sets hold the next two steps' words.
  constexpr uint32_t kTasks = 64 * (kCovFS / 4), kMaxT = 2;
  const uint32_t tsplit = (wave % jobs) * 64 + lane;  // thread index within the split
  const uint32_t stride = 64 * jobs;
  const uint32_t ch = lane % chp, rbin = lane / chp;
  const bool row_ok = ch < CH && rbin < bins_here;
  int32_t srow_r = 0, srow_i = 0;  // this thread's frame sums of its row (its tasks)
  uint32_t buf0[kMaxT][4], buf1[kMaxT][4];
  // A task's row advances by whole frames: one pointer per task at the
  // split's first frame of its group, then uniform frame offsets (no 64-bit
  // index products per load).
  const size_t fstride = (size_t)N * CH;  // words per frame
  const uint32_t* tptr[kMaxT];
  uint32_t tm0[kMaxT];
#pragma unroll
  for (uint32_t q = 0; q < kMaxT; ++q) {
    const uint32_t g = (tsplit + q * stride) >> 6;
    tm0[q] = fs * sps * kCovFS + 4 * g;
    tptr[q] = Xa + (size_t)tm0[q] * fstride + (size_t)(k_base + rbin) * CH + ch;
  }
  auto fetch = [&](uint32_t st, uint32_t (&nxt)[kMaxT][4]) {
    const size_t so = (size_t)st * kCovFS * fstride;
#pragma unroll
    for (uint32_t q = 0; q < kMaxT; ++q) {
      const uint32_t task = tsplit + q * stride;
      const bool ok = task < kTasks && row_ok && st < sps;
#pragma unroll
      for (uint32_t t = 0; t < 4; ++t) {
        const uint32_t m = tm0[q] + st * kCovFS + t;
        nxt[q][t] = (ok && m < F) ? tptr[q][so + t * fstride] : 0u;
      }
    }
  };
  fetch(0, buf0);
  fetch(1, buf1);
  // one step from the register set `cur` (even steps buf0, odd steps buf1, so
  // the sets stay in fixed registers)
  auto step = [&](uint32_t st, uint32_t (&cur)[kMaxT][4]) {
    const uint32_t m0 = (fs * sps + st) * kCovFS;
    __syncthreads();  // the previous step's planes are consumed
#pragma unroll
    for (uint32_t q = 0; q < kMaxT; ++q) {
      const uint32_t task = tsplit + q * stride, g = task >> 6;
      if (task < kTasks) {
        const uint32_t* w = cur[q];
        uint32_t bm = 0;  // byte mask of the valid frames among the four
#pragma unroll
        for (uint32_t t = 0; t < 4; ++t) bm |= (row_ok && m0 + 4 * g + t < F) ? (0xFFu << (8 * t)) : 0u;
        const uint32_t hr = gather4(w[0], w[1], w[2], w[3], 1) & bm;
        const uint32_t lr = (gather4(w[0], w[1], w[2], w[3], 0) ^ 0x80808080u) & bm;
        const uint32_t hi = gather4(w[0], w[1], w[2], w[3], 3) & bm;
        const uint32_t li = (gather4(w[0], w[1], w[2], w[3], 2) ^ 0x80808080u) & bm;
        unsigned char* o = planes + lane * kCovRow + 4 * g;
        *reinterpret_cast<uint32_t*>(o) = hr;
        *reinterpret_cast<uint32_t*>(o + kCovPlane) = lr;
        *reinterpret_cast<uint32_t*>(o + 2 * kCovPlane) = hi;
        *reinterpret_cast<uint32_t*>(o + 3 * kCovPlane) = li;
        *reinterpret_cast<uint32_t*>(o + 4 * kCovPlane) = ~hr & bm;
        *reinterpret_cast<uint32_t*>(o + 5 * kCovPlane) = ~lr & bm;
#pragma unroll
        for (uint32_t t = 0; t < 4; ++t) {  // padded words are zero
          srow_r += cre(w[t]);
          srow_i += cim(w[t]);
        }
      }
    }
    __syncthreads();
    fetch(st + 2, cur);  // the set just converted takes the step after next
    if (m0 < F) nfr += (int32_t)min(kCovFS, F - m0);
    // K-steps of 32 (16 frames, real half in lane groups 0-1, imaginary in 2-3)
#pragma unroll
    for (uint32_t ks = 0; ks < kCovFS / 16; ++ks) {
      const uint32_t fb = 16 * ks + 8 * (grp & 1);
      const bool imag = (grp >> 1) != 0;
      const unsigned char* pa = planes + ra * kCovRow + fb;
      const unsigned char* pb = planes + rb * kCovRow + fb;
      // Re: A = [Xr | Xi], B = [Xr | Xi];  Im: A = [Xi | ~Xr]
      const uint32_t pre = imag ? 2 : 0, pim = imag ? 4 : 2;
      const uint64_t reah = *reinterpret_cast<const uint64_t*>(pa + pre * kCovPlane);
      const uint64_t real_ = *reinterpret_cast<const uint64_t*>(pa + (pre + 1) * kCovPlane);
      const uint64_t imah = *reinterpret_cast<const uint64_t*>(pa + pim * kCovPlane);
      const uint64_t imal = *reinterpret_cast<const uint64_t*>(pa + (pim + 1) * kCovPlane);
      const uint64_t bh = *reinterpret_cast<const uint64_t*>(pb + pre * kCovPlane);
      const uint64_t bl = *reinterpret_cast<const uint64_t*>(pb + (pre + 1) * kCovPlane);
      acc[0] = mfma8(reah, bh, acc[0]);
      acc[1] = mfma8(reah, bl, acc[1]);
      acc[3] = mfma8(real_, bh, acc[3]);
      acc[2] = mfma8(real_, bl, acc[2]);
      acc[4] = mfma8(imah, bh, acc[4]);
      acc[5] = mfma8(imah, bl, acc[5]);
      acc[7] = mfma8(imal, bh, acc[7]);
      acc[6] = mfma8(imal, bl, acc[6]);
    }
  };
  for (uint32_t st = 0; st < sps; st += 2) {
    step(st, buf0);
    if (st + 1 < sps) step(st + 1, buf1);
  }

  // per-row frame sums of this split (the split's threads sharing a row add up)
  if (srow_r | srow_i) {
    atomicAdd(&sums[2 * lane], srow_r);
    atomicAdd(&sums[2 * lane + 1], srow_i);
  }
  __syncthreads();
  const int32_t b_r = sums[2 * rb], b_i = sums[2 * rb + 1];
  const int64_t kconst = 16384LL * 2 * nfr;
  const int64_t tjs = (int64_t)b_r + b_i;
  int64_t part[8];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t il = bin * chp + 16 * ti + 4 * grp + r;
    const int32_t ir = sums[2 * il], ii = sums[2 * il + 1];
    const int64_t tis = (int64_t)ir + ii, uis = (int64_t)ii - ir - (int64_t)nfr;
    part[2 * r] = combine(acc[0], acc[1], acc[2], r) + 256 * (int64_t)acc[3][r] + 128 * (tis + tjs) - kconst;
    part[2 * r + 1] = combine(acc[4], acc[5], acc[6], r) + 256 * (int64_t)acc[7][r] + 128 * (uis + tjs) - kconst + b_i;
  }
  if (fs_n > 1) {
    if (wave < jobs * fs_n)
#pragma unroll
      for (int q = 0; q < 8; ++q) atomicAdd(&red[(job * 64 + lane) * 8 + q], (unsigned long long)part[q]);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) part[q] = (int64_t)red[(job * 64 + lane) * 8 + q];
  }
  if (fs != 0 || !live) return;
  const uint32_t k = k_base + bin;
  const uint32_t cj = 16 * tj + cl;
  const size_t npairs = (size_t)CH * (CH + 1) / 2;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t ii = 16 * ti + 4 * grp + r;
    if (ii >= CH || cj >= CH || ii > cj) continue;
    const size_t pair = (size_t)ii * CH - (size_t)ii * (ii - 1) / 2 + (cj - ii);
    int64_t* o = J.R + ((size_t)k * npairs + pair) * 2;
    o[0] = part[2 * r];
    o[1] = part[2 * r + 1];
  }
}

// Single CPI: one block per tile.
__global__ __launch_bounds__(1024) void cov_kernel(CovJob J, uint32_t nb, uint32_t nt, uint32_t fs_n) {
  extern __shared__ uint32_t smem[];
  cov_tile(J, blockIdx.x, nb, nt, fs_n, smem);
}

// Batched: one block per bin tile of a class of CPIs (same padded channel
// count), block -> job via the tile table.
__global__ __launch_bounds__(1024) void cov_batch_kernel(const CovJob* __restrict__ jobs,
                                                         const uint32_t* __restrict__ tile_job, uint32_t n_tiles,
                                                         uint32_t nb, uint32_t nt, uint32_t fs_n) {
  extern __shared__ uint32_t smem[];
  const uint32_t t = blockIdx.x;
  if (t >= n_tiles) return;
  const CovJob& J = jobs[__builtin_amdgcn_readfirstlane(tile_job[t])];
  cov_tile(J, t - J.tile0, nb, nt, fs_n, smem);
}

struct CovGeo {
  uint32_t nb, nt, jobs;
  explicit CovGeo(uint32_t chp) : nb(64 / chp), nt(chp / 16), jobs((64 / chp) * (chp / 16) * (chp / 16 + 1) / 2) {}
  size_t lds(uint32_t fs_n) const {
    return (size_t)fs_n * kCovSplitBytes + (fs_n > 1 ? (size_t)jobs * 64 * 8 * 8 : 0);
  }
};

uint32_t cov_bins_per_tile(uint32_t chp) { return CovGeo(chp).nb; }
uint32_t cov_max_splits(uint32_t chp) { return 16 / CovGeo(chp).jobs; }

// ============================================================================
// Covariance of CPIs of 32 or 64 channels (CHP): a block owns NB = 64/CHP
// adjacent bins (64 channel rows), their upper-triangle 16x16 channel tiles
// on JOBS compute waves (6 or 10), and one loader wave. Each 32-frame step
// of the block (32 runs of 256 bytes in Xa[F][N][CHP]: the NB bins' channels
// of one frame) is moved into a ring of raw LDS slots by
// global_load_lds_dwordx4 (eight instructions a step) kCrRing steps ahead,
// with counted waits; the compute waves convert a slot into digit planes
// (double-buffered, so one barrier a step orders both the planes and the
// slot's release) and run the same MFMA products and exact int64
// recombination as cov_tile.
// ============================================================================
constexpr uint32_t kCrRing = 3;
static_assert((kCrRing - 1) * 8 <= 40, "wait_vmcnt covers the loader's counts");
constexpr uint32_t kCrSlot = kCovFS * 64 * 4;  // one step: [32 frames][64 rows] words
constexpr uint32_t kCrPlanes = 6 * kCovPlane;
constexpr size_t kCrLds = (size_t)kCrRing * kCrSlot + 2 * kCrPlanes + 128 * 4;

template <uint32_t CHP>
struct CrGeo {
  static constexpr uint32_t NB = 64 / CHP, NT = CHP / 16, UT = NT * (NT + 1) / 2, JOBS = NB * UT;
};

bool cov_ring_fits(uint32_t n_chan, uint32_t n_frames, uint32_t n_bins, const uint32_t* Xa) {
  return (n_chan == 32 || n_chan == 64) && n_frames >= 1 && n_bins >= 1 &&
         n_bins % (64 / n_chan) == 0 && (reinterpret_cast<uintptr_t>(Xa) & 15u) == 0;
}

// Bins [NB tile, +NB) of J.
template <uint32_t CHP>
__device__ __forceinline__ void cov_ring_body(const CovJob& J, uint32_t tile, uint32_t* smem) {
  using G = CrGeo<CHP>;
  unsigned char* const lds = reinterpret_cast<unsigned char*>(smem);
  unsigned char* const raw = lds;                                   // [kCrRing][kCrSlot]
  unsigned char* const planes0 = raw + kCrRing * kCrSlot;          // [2][6 planes]
  int32_t* const sums = reinterpret_cast<int32_t*>(planes0 + 2 * kCrPlanes);  // [64 rows][2]
  const uint32_t F = J.F, N = J.N;
  const uint32_t k_base = tile * G::NB;
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool loader = wave == G::JOBS;
  const uint32_t nsteps = (F + kCovFS - 1) / kCovFS;
  if (threadIdx.x < 128) sums[threadIdx.x] = 0;

  if (loader) {
    // lane l of instruction j: frame 4 j + l / 16 of the step, rows
    // 4 (l % 16) .. + 3; frames past F re-read the last one (masked later)
    const uint32_t fl = lane >> 4, c4 = (lane & 15) * 4;
    auto issue = [&](uint32_t st) {
      unsigned char* slot = raw + (st % kCrRing) * kCrSlot;
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) {
        const uint32_t m = min(st * kCovFS + 4 * j + fl, F - 1);
        lds_dma16(J.Xa + ((size_t)m * N + k_base) * CHP + c4, slot + j * 1024);
      }
    };
    uint32_t next = 0;
    for (; next < nsteps && next < kCrRing; ++next) issue(next);
    wait_vmcnt((next - 1) * 8);
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();  // step 0 is in the ring (and sums are zero)
    for (uint32_t st = 0; st < nsteps; ++st) {
      wait_vmcnt(st + 1 < next ? (next - st - 2) * 8 : 0u);  // step st + 1 complete
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();  // the compute waves are done with slot st % kCrRing
      __asm__ volatile("" ::: "memory");
      if (next < nsteps) issue(next++);
    }
    __syncthreads();  // the frame sums (below)
    return;
  }

  const uint32_t grp = lane >> 4, cl = lane & 15;
  const uint32_t bin = wave / G::UT;
  uint32_t u = wave % G::UT, ti = 0;  // upper-triangle tile (ti, tj) of this wave
  while (u >= G::NT - ti) {
    u -= G::NT - ti;
    ++ti;
  }
  const uint32_t tj = ti + u;
  const uint32_t ra = bin * CHP + 16 * ti + cl, rb = bin * CHP + 16 * tj + cl;  // operand rows of this lane
  v4i acc[8];
#pragma unroll
  for (int z = 0; z < 8; ++z) acc[z] = v4i{0, 0, 0, 0};
  // conversion tasks (row = lane, frame group g of 4 frames), g = 0..7, dealt
  // to the compute waves; the frame sums of a row accumulate per task
  int32_t srow_r = 0, srow_i = 0;
  __syncthreads();  // step 0 is in the ring (and sums are zero)
  for (uint32_t st = 0; st < nsteps; ++st) {
    unsigned char* planes = planes0 + (st & 1) * kCrPlanes;
    const uint32_t m0 = st * kCovFS;
#pragma unroll
    for (uint32_t g = wave; g < 8; g += G::JOBS) {
      const unsigned char* rs = raw + (st % kCrRing) * kCrSlot + (4 * g * 64 + lane) * 4;
      const uint32_t w[4] = {*reinterpret_cast<const uint32_t*>(rs), *reinterpret_cast<const uint32_t*>(rs + 256),
                             *reinterpret_cast<const uint32_t*>(rs + 512),
                             *reinterpret_cast<const uint32_t*>(rs + 768)};
      uint32_t bm = 0;  // byte mask of the valid frames among the four
#pragma unroll
      for (uint32_t t = 0; t < 4; ++t) bm |= m0 + 4 * g + t < F ? (0xFFu << (8 * t)) : 0u;
      const uint32_t hr = gather4(w[0], w[1], w[2], w[3], 1) & bm;
      const uint32_t lr = (gather4(w[0], w[1], w[2], w[3], 0) ^ 0x80808080u) & bm;
      const uint32_t hi = gather4(w[0], w[1], w[2], w[3], 3) & bm;
      const uint32_t li = (gather4(w[0], w[1], w[2], w[3], 2) ^ 0x80808080u) & bm;
      unsigned char* o = planes + lane * kCovRow + 4 * g;
      *reinterpret_cast<uint32_t*>(o) = hr;
      *reinterpret_cast<uint32_t*>(o + kCovPlane) = lr;
      *reinterpret_cast<uint32_t*>(o + 2 * kCovPlane) = hi;
      *reinterpret_cast<uint32_t*>(o + 3 * kCovPlane) = li;
      *reinterpret_cast<uint32_t*>(o + 4 * kCovPlane) = ~hr & bm;
      *reinterpret_cast<uint32_t*>(o + 5 * kCovPlane) = ~lr & bm;
#pragma unroll
      for (uint32_t t = 0; t < 4; ++t) {
        if ((bm >> (8 * t)) & 1u) {
          srow_r += cre(w[t]);
          srow_i += cim(w[t]);
        }
      }
    }
    lds_barrier();  // planes of step st complete; slot st % kCrRing consumed
#pragma unroll
    for (uint32_t ks = 0; ks < kCovFS / 16; ++ks) {
      const uint32_t fb = 16 * ks + 8 * (grp & 1);
      const bool imag = (grp >> 1) != 0;
      const unsigned char* pa = planes + ra * kCovRow + fb;
      const unsigned char* pb = planes + rb * kCovRow + fb;
      const uint32_t pre = imag ? 2 : 0, pim = imag ? 4 : 2;
      const uint64_t reah = *reinterpret_cast<const uint64_t*>(pa + pre * kCovPlane);
      const uint64_t real_ = *reinterpret_cast<const uint64_t*>(pa + (pre + 1) * kCovPlane);
      const uint64_t imah = *reinterpret_cast<const uint64_t*>(pa + pim * kCovPlane);
      const uint64_t imal = *reinterpret_cast<const uint64_t*>(pa + (pim + 1) * kCovPlane);
      const uint64_t bh = *reinterpret_cast<const uint64_t*>(pb + pre * kCovPlane);
      const uint64_t bl = *reinterpret_cast<const uint64_t*>(pb + (pre + 1) * kCovPlane);
      acc[0] = mfma8(reah, bh, acc[0]);
      acc[1] = mfma8(reah, bl, acc[1]);
      acc[3] = mfma8(real_, bh, acc[3]);
      acc[2] = mfma8(real_, bl, acc[2]);
      acc[4] = mfma8(imah, bh, acc[4]);
      acc[5] = mfma8(imah, bl, acc[5]);
      acc[7] = mfma8(imal, bh, acc[7]);
      acc[6] = mfma8(imal, bl, acc[6]);
    }
  }
  // per-row frame sums (the eight tasks of a row add up)
  if (wave < 8 && (srow_r | srow_i)) {
    atomicAdd(&sums[2 * lane], srow_r);
    atomicAdd(&sums[2 * lane + 1], srow_i);
  }
  __syncthreads();
  const int64_t nfr = F;
  const int32_t b_r = sums[2 * rb], b_i = sums[2 * rb + 1];
  const int64_t kconst = 16384LL * 2 * nfr;
  const int64_t tjs = (int64_t)b_r + b_i;
  const uint32_t cj = 16 * tj + cl;
  constexpr uint32_t NPAIRS = CHP * (CHP + 1) / 2;
  int64_t* const Rk = J.R + (size_t)(k_base + bin) * NPAIRS * 2;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t ii = 16 * ti + 4 * grp + r;
    const uint32_t il = bin * CHP + ii;  // its row
    const int32_t ir = sums[2 * il], iim = sums[2 * il + 1];
    const int64_t tis = (int64_t)ir + iim, uis = (int64_t)iim - ir - nfr;
    if (ii > cj) continue;
    const size_t pair = (size_t)ii * CHP - (size_t)ii * (ii - 1) / 2 + (cj - ii);
    Rk[2 * pair] = combine(acc[0], acc[1], acc[2], r) + 256 * (int64_t)acc[3][r] + 128 * (tis + tjs) - kconst;
    Rk[2 * pair + 1] =
        combine(acc[4], acc[5], acc[6], r) + 256 * (int64_t)acc[7][r] + 128 * (uis + tjs) - kconst + b_i;
  }
}

// Single CPI: one block per NB bins.
template <uint32_t CHP>
__global__ __launch_bounds__(64 * (CrGeo<CHP>::JOBS + 1)) void cov_ring_kernel(CovJob J) {
  extern __shared__ uint32_t smem[];
  cov_ring_body<CHP>(J, blockIdx.x, smem);
}

// Batched: one block per (CPI, NB bins) of a class of CHP-channel CPIs.
template <uint32_t CHP>
__global__ __launch_bounds__(64 * (CrGeo<CHP>::JOBS + 1)) void cov_ring_batch_kernel(
    const CovJob* __restrict__ jobs, const uint32_t* __restrict__ tile_job) {
  extern __shared__ uint32_t smem[];
  const CovJob& J = jobs[__builtin_amdgcn_readfirstlane(tile_job[blockIdx.x])];
  cov_ring_body<CHP>(J, blockIdx.x - J.tile0, smem);
}

hipError_t launch_cov_ring_batch(hipStream_t s, uint32_t chp, const ClassLaunch& L) {
  if (L.n_tiles == 0) return hipSuccess;
  const CovJob* jobs = static_cast<const CovJob*>(L.jobs);
  if (chp == 32)
    cov_ring_batch_kernel<32><<<L.n_tiles, 64 * (CrGeo<32>::JOBS + 1), kCrLds, s>>>(jobs, L.tile_job);
  else if (chp == 64)
    cov_ring_batch_kernel<64><<<L.n_tiles, 64 * (CrGeo<64>::JOBS + 1), kCrLds, s>>>(jobs, L.tile_job);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_cov(hipStream_t s, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                      int64_t* R) {
  if (n_bins == 0) return hipSuccess;
  if (cov_ring_fits(n_chan, n_frames, n_bins, Xa)) {
    const CovJob J{Xa, R, n_chan, n_frames, n_bins, 0};
    const uint32_t tiles = n_bins / (64 / n_chan);
    if (n_chan == 32)
      cov_ring_kernel<32><<<tiles, 64 * (CrGeo<32>::JOBS + 1), kCrLds, s>>>(J);
    else
      cov_ring_kernel<64><<<tiles, 64 * (CrGeo<64>::JOBS + 1), kCrLds, s>>>(J);
    return hipGetLastError();
  }
  const uint32_t chp = (n_chan + 15) / 16 * 16;
  const CovGeo G(chp);
  uint32_t fs_n = 16 / G.jobs;
  const uint32_t nsteps = (n_frames + kCovFS - 1) / kCovFS;
  if (fs_n > nsteps) fs_n = nsteps;
  if (fs_n < 1) fs_n = 1;
  const CovJob J{Xa, R, n_chan, n_frames, n_bins, 0};
  cov_kernel<<<(n_bins + G.nb - 1) / G.nb, 64 * G.jobs * fs_n, G.lds(fs_n), s>>>(J, G.nb, G.nt, fs_n);
  return hipGetLastError();
}

// ============================================================================
// Covariance through the Integrator frame by frame (covar_model.hpp:196-214):
// block mode adds the cross power, exponential mode moves the accumulator by
// (x - acc) >> exp_k (floor); both saturate to acc_w every frame with sticky
// flags. A recurrence, so no contraction: one thread per (bin, pair) walks the
// frames; a wave's threads share a bin, so each frame's loads hit one row.
// ============================================================================
__device__ __forceinline__ int64_t sat_w(int64_t v, uint32_t acc_w, uint32_t& fl) {
  if (acc_w >= 64) return v;
  const int64_t hi = ((int64_t)1 << (acc_w - 1)) - 1, lo = -hi - 1;
  fl |= (v > hi ? 2u : 0u) | (v < lo ? 1u : 0u);
  return v > hi ? hi : (v < lo ? lo : v);
}

__global__ __launch_bounds__(256) void cov_integrate_kernel(const uint32_t* __restrict__ Xa, uint32_t CH, uint32_t F,
                                                            uint32_t N, uint32_t mode, uint32_t exp_k,
                                                            uint32_t acc_w, int64_t* __restrict__ R,
                                                            uint8_t* __restrict__ sat_flags) {
  const uint32_t npairs = CH * (CH + 1) / 2;
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (uint64_t)N * npairs) return;
  const uint32_t k = (uint32_t)(g / npairs);
  uint32_t p = (uint32_t)(g % npairs), i = 0;
  while (p >= CH - i) {  // row-major upper triangle: row i holds CH - i pairs
    p -= CH - i;
    ++i;
  }
  const uint32_t j = i + p;
  int64_t are = 0, aim = 0;
  uint32_t fre = 0, fim = 0;
  for (uint32_t m = 0; m < F; ++m) {
    const uint32_t* row = Xa + ((size_t)m * N + k) * CH;
    const uint32_t x = row[i], y = row[j];
    // x * conj(y), exact
    const int64_t re = (int64_t)cre(x) * cre(y) + (int64_t)cim(x) * cim(y);
    const int64_t im = (int64_t)cim(x) * cre(y) - (int64_t)cre(x) * cim(y);
    const int64_t dre = mode ? (int64_t)((uint64_t)re - (uint64_t)are) >> exp_k : re;
    const int64_t dim = mode ? (int64_t)((uint64_t)im - (uint64_t)aim) >> exp_k : im;
    are = sat_w((int64_t)((uint64_t)are + (uint64_t)dre), acc_w, fre);
    aim = sat_w((int64_t)((uint64_t)aim + (uint64_t)dim), acc_w, fim);
  }
  R[2 * g] = are;
  R[2 * g + 1] = aim;
  if (sat_flags) sat_flags[g] = (uint8_t)(fre | (fim << 2));
}

hipError_t launch_cov_integrate(hipStream_t s, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames,
                                uint32_t n_bins, uint32_t mode, uint32_t exp_k, uint32_t acc_w, int64_t* R,
                                uint8_t* sat_flags) {
  const uint64_t total = (uint64_t)n_bins * n_chan * (n_chan + 1) / 2;
  if (total == 0) return hipSuccess;
  cov_integrate_kernel<<<(uint32_t)((total + 255) / 256), 256, 0, s>>>(Xa, n_chan, n_frames, n_bins, mode, exp_k,
                                                                       acc_w, R, sat_flags);
  return hipGetLastError();
}

hipError_t launch_cov_batch(hipStream_t s, uint32_t chp, uint32_t fs_n, const ClassLaunch& L) {
  if (L.n_tiles == 0) return hipSuccess;
  const CovGeo G(chp);
  if (fs_n < 1 || G.jobs * fs_n > 16) return hipErrorInvalidValue;
  cov_batch_kernel<<<L.grid, 64 * G.jobs * fs_n, G.lds(fs_n), s>>>(static_cast<const CovJob*>(L.jobs), L.tile_job,
                                                                   L.n_tiles, G.nb, G.nt, fs_n);
  return hipGetLastError();
}

// ============================================================================
// Beamformer + power and covariance in one pass over the FFT output, for CPIs
// of at most 16 channels. Both contractions read the same samples: the
// beamformer contracts over channels per column (frame, bin), the covariance
// over frames per bin. A block owns kFbNB adjacent bins over every frame and
// walks the frames FS = 4 * WAVES at a time:
//   * every thread loads the four channel words (4q .. 4q+3) of one column and
//     writes their int8 digits once, as one 32-bit word per plane, into eight
//     LDS digit planes [plane][column][16 channel bytes]:
//       hr, lr, hi, li, ~hr, ~lr, ~hi, ~li
//     (real/imaginary high and offset-low digits and their complements), so
//     a plane row is the 16 channels of one column;
//   * beamformer: wave w takes the 16 columns of frames 4w .. 4w+3; its B
//     operand (8 channels of one column) is one 8-byte LDS read per plane;
//   * covariance: wave w takes bin w % NB and the 16 frames 16 * (w / NB) of
//     the step; its operands (8 frames of one channel) are the transposes of
//     plane rows, read with ds_read_b64_tr_b8: in each 16-lane group, lane
//     2r + h addresses half h (channels 8h .. 8h+7) of row r (frame r), and
//     lane n receives channel n of the group's 8 rows;
//   * the per-column channel sums (beamformer) and per-channel frame sums
//     (covariance) of the digit identity come from extra MFMAs against an
//     all-ones operand instead of VALU additions.
// The covariance accumulators stay in registers over all frames (int32 is
// exact to 65535 frames: a frame adds at most 2^15 to each); the waves of a
// bin add theirs in LDS at the end and one wave forms the int64 results.
// The aligned copy Xa is an optional output (16-byte stores of the staged
// channel words), not an intermediate.
// Loads: one extra wave of the block (the loader) only moves the samples,
// global -> LDS without registers (global_load_lds_dwordx4), kFbRing steps
// ahead into a ring of raw step buffers, and waits for them with a counted
// s_waitcnt vmcnt before each step's barrier. The WAVES compute waves issue
// no vector-memory loads in the frame loop, so their power stores are never
// waited for and no load latency sits on their path.
// ============================================================================
constexpr uint32_t kFbNB = 4;  // bins per block
constexpr uint32_t kFbPlanes = 8;
// Raw step buffers of the loader: per channel pair one 1 KiB block (channel
// 2j: frames 0..31 of the step, 16 bytes = the block's 4 bins each; then
// channel 2j + 1), written by one global_load_lds_dwordx4; 32 bytes of
// padding after each block put the four channel quads of the conversion reads
// on distinct LDS banks.
constexpr uint32_t kFbRing = 3;                   // steps in flight
static_assert((kFbRing - 1) * 8 <= 40, "wait_vmcnt covers the loader's counts");
constexpr uint32_t kFbBlk = 1024 + 32;            // bytes per channel-pair block
constexpr uint32_t kFbSlot = 8 * kFbBlk;          // one step (16 channels)
enum : uint32_t { kPHR = 0, kPLR, kPHI, kPLI, kPNHR, kPNLR, kPNHI, kPNLI };

typedef int v2i_t __attribute__((ext_vector_type(2)));

// LDS writes of one lane become visible to the other lanes of its wave.
__device__ __forceinline__ void wave_lds_sync_mfma() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ds_read_b64_tr_b8 at LDS byte address p (8-byte aligned).
__device__ __forceinline__ uint64_t lds_tr8(const unsigned char* p) {
  const v2i_t r = __builtin_amdgcn_ds_read_tr8_b64_v2i32(
      (__attribute__((address_space(3))) v2i_t*)(__attribute__((address_space(3))) unsigned char*)p);
  return (uint64_t)(uint32_t)r.x | ((uint64_t)(uint32_t)r.y << 32);
}

template <int WAVES>
struct FbGeo {
  static constexpr uint32_t FS = 4 * WAVES;                     // frames per step
  static constexpr uint32_t COLS = FS * kFbNB;                  // columns per step
  static constexpr uint32_t PLANE = COLS * 16;                  // bytes per plane
  static constexpr uint32_t BUF = kFbPlanes * PLANE;            // one step's planes
  static constexpr uint32_t RED = (WAVES / kFbNB - 1) * kFbNB * 36 * 64 * 4;  // partials [K-step wave - 1][bin][36][64]
  static constexpr uint32_t STAGE = 2 * BUF > RED ? 2 * BUF : RED;  // double-buffered planes, aliased by RED
  static constexpr uint32_t RAW = kFbRing * kFbSlot;               // the loader's ring
  static_assert(FS == 32, "a raw block holds 32 frames of a channel pair");
  static size_t lds(uint32_t bt_n) {
    return STAGE + RAW + bf_frag_words(bt_n, 1) * 4 + (size_t)bt_n * 32 * 4 + kFbNB * 16 * 4 * 4;
  }
};


bool bfcov_fits(uint32_t n_chan, uint32_t n_bins, uint32_t n_frames, uint32_t n_beams) {
  // 32-bit byte offsets from a per-step base cover every channel row
  return n_chan >= 1 && n_chan <= 16 && n_beams >= 1 && n_beams <= 128 && n_bins % kFbNB == 0 && n_bins >= kFbNB &&
         n_frames >= 1 && n_frames <= 65535 && (uint64_t)16 * n_frames * n_bins * 4 < (1ull << 32);
}

// T1: the beams fill exactly one 16-beam tile (B = 16): no beam bounds, and a
// lane's four beam constants stay in registers over the whole frame loop.
// XCD-aware tile order: consecutive blocks are dealt to the 8 XCDs in turn,
// so the blocks one XCD runs together take adjacent tiles and share the L2
// lines of their loads and stores.
__device__ __forceinline__ uint32_t xcd_tile() {
  const uint32_t nblk = gridDim.x;
  return nblk % 8 == 0 ? (blockIdx.x % 8) * (nblk / 8) + blockIdx.x / 8 : blockIdx.x;
}

// The weight digits of a one-K-step beamformer (at most 16 channels) and the
// per-beam sums Sa = sum (wr + wi), Wis = sum wi from the same loads: the lanes
// of the real K rows (lane groups 0 and 1) hold channels 0-7 and 8-15 of one
// beam, so one shuffle adds a beam's halves (bf_weights_to_lds reads W again
// for the sums, a second dependent load round).
__device__ __forceinline__ void bf1_weights_to_lds(const uint32_t* __restrict__ W, uint32_t CH, uint32_t B,
                                                   uint32_t bt_n, uint32_t* frag, int32_t* tail, uint32_t nthr) {
  for (uint32_t f = threadIdx.x; f < bt_n * 64; f += nthr) {
    const uint32_t lane = f & 63, bt = f >> 6;
    const uint32_t b = bt * 16 + (lane & 15);
    const uint32_t k0 = 8 * (lane >> 4);
    const bool imag = k0 >= 16;
    const uint32_t a0 = imag ? k0 - 16 : k0;
    uint32_t p[8];
    uint64_t vmask = 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const bool ok = b < B && a0 + t < CH;
      p[t] = ok ? W[(size_t)b * CH + a0 + t] : 0u;
      vmask |= ok ? (0xFFull << (8 * t)) : 0ull;
    }
    const Dig8 d = digits8(p, vmask);
    uint64_t* o = reinterpret_cast<uint64_t*>(frag) + (size_t)bt * 256 + lane;
    o[0] = imag ? d.hi : d.hr;
    o[64] = imag ? d.li : d.lr;
    o[128] = imag ? d.hr : d.hi;
    o[192] = imag ? d.lr : d.li;
    int32_t sa = 0, wis = 0;  // padded entries are zero
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      sa += cre(p[t]) + cim(p[t]);
      wis += cim(p[t]);
    }
    sa += __shfl_xor(sa, 16);
    wis += __shfl_xor(wis, 16);
    if (lane < 16) {
      tail[2 * b] = sa;
      tail[2 * b + 1] = wis;
    }
  }
}

// Bins [4 tile, +4) of J; lay_bt: the beam tiles the LDS layout is sized for.
template <int WAVES, int OUT, bool T1>
__device__ __forceinline__ void bfcov_body(const BfCovJob& J, uint32_t tile, uint32_t lay_bt) {
  using G = FbGeo<WAVES>;
  constexpr uint32_t FS = G::FS;
  extern __shared__ uint32_t lds[];
  unsigned char* const sm = reinterpret_cast<unsigned char*>(lds);
  unsigned char* const raw = sm + G::STAGE;  // the loader's ring of raw steps
  uint32_t* frag = reinterpret_cast<uint32_t*>(raw + G::RAW);
  int32_t* tail = reinterpret_cast<int32_t*>(frag + bf_frag_words(lay_bt, 1));
  int32_t* sums = tail + lay_bt * 32;  // [bin][16 rows][4] digit sums of the covariance A rows
  const int out = OUT >= 0 ? OUT : bf_out_mask(J.bf);
  const uint32_t CH = J.bf.CH, B = J.bf.B, N = J.N, F = J.F;
  const uint64_t cols = (uint64_t)F * N;
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t grp = lane >> 4, cl = lane & 15;
  const bool loader = wave == WAVES;  // the last wave only loads

  const uint32_t k0 = tile * kFbNB;

  // conversion unit of a compute thread: channel quad q, bin b, frame f of the step
  const uint32_t t = threadIdx.x;
  const uint32_t q = t & 3, cb = (t >> 2) & 3, cf = (t >> 4) & (FS - 1);
  // X natural [ch][F][N]: a frame is N words on from the last; bin-quad
  // blocked [ch][N/4][F][4]: the block's 4 bins of every frame are one
  // contiguous run per channel (4 words a frame)
  const bool xb = J.x_blocked != 0;
  const uint32_t fstride = xb ? 4u : N;
  uint32_t chmask = 0;
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) chmask |= 4 * q + i < CH ? 0xFFu << (8 * i) : 0u;
  // the unit's four channel words in a raw step: channel 4q + i sits in block
  // 2q + i / 2, half i % 2
  const uint32_t roff = q * 2 * kFbBlk + cf * 16 + cb * 4;
  const uint32_t nsteps = (F + FS - 1) / FS;
  const uint32_t* const X0 = J.bf.X + (xb ? (size_t)k0 * F : (size_t)k0);  // k0 is a multiple of 4
  // the compute waves derive the weight digits while the first steps load
  if (!loader) bf1_weights_to_lds(J.bf.W, CH, B, J.bf.bt_n, frag, tail, 64 * WAVES);

  if (loader) {
    // lane l of block j moves channel 2j + l / 32, frame l % 32 of the step
    // (16 bytes: the block's 4 bins); channels past CH and frames past F
    // re-read the last ones (every block is one instruction for every lane,
    // so the counted waits below stay exact) and are masked by the conversion
    const uint32_t n_ld = (CH + 1) / 2;  // blocks per step
    const uint32_t half = lane >> 5, lf = lane & 31;
    auto issue = [&](uint32_t st) {
      unsigned char* slot = raw + (st % kFbRing) * kFbSlot;
      const uint32_t fr = min(st * FS + lf, F - 1);
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) {
        if (j >= n_ld) break;
        const uint32_t ch = min(2 * j + half, CH - 1);
        const uint32_t* src = X0 + (size_t)ch * F * N + (size_t)fr * fstride;
        lds_dma16(src, slot + j * kFbBlk);
      }
    };
    // loads issued so far cover steps [0, next); before the barrier that
    // releases step s, the loads of step s are complete
    uint32_t next = 0;
    for (; next < nsteps && next < kFbRing; ++next) issue(next);
    wait_vmcnt((next - 1) * n_ld);
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();  // releases step 0
    for (uint32_t st = 0; st < nsteps; ++st) {
      // step st + 1 complete: the later steps issued may stay in flight
      wait_vmcnt(st + 1 < next ? (next - st - 2) * n_ld : 0u);
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();  // the compute waves' barrier of step st
      __asm__ volatile("" ::: "memory");
      // slot st % kFbRing was converted by every wave before that barrier
      if (next < nsteps) issue(next++);
    }
  }

  // beamformer lane roles: column c = 16 * wave + cl  // beamformer lane roles: column c = 16 * wave + cl (frame c / 4, bin c % 4),
  // K rows: channel half grp & 1, real (grp < 2) or imaginary part
  const uint32_t bc = 16 * wave + cl, bfr = bc >> 2, bbin = bc & 3;
  const bool bim = grp >= 2;
  const uint32_t boff = bc * 16 + 8 * (grp & 1);
  const uint32_t p_bih = bim ? kPHI : kPHR, p_bil = bim ? kPLI : kPLR;
  const uint32_t p_brh = bim ? kPNHI : kPHR, p_brl = bim ? kPNLI : kPLR;
  const int32_t kconst_bf = 16384 * 2 * (int32_t)CH;
  const size_t lane_row = (size_t)(4 * grp) * cols;  // output row of the lane's first beam
  const uint64_t ONES = 0x0101010101010101ull;

  // covariance lane roles: bin cbin, frames 16 * cks + 8 * (grp & 1) + (cl >> 1)
  // (this lane's transpose-read row), channel half cl & 1
  const uint32_t cbin = wave % kFbNB, cks = wave / kFbNB;
  const uint32_t cfr = 16 * cks + 8 * (grp & 1) + (cl >> 1);
  const uint32_t coff = (cfr * kFbNB + cbin) * 16 + 8 * (cl & 1);
  const uint32_t p_reh = bim ? kPHI : kPHR, p_rel = bim ? kPLI : kPLR;
  const uint32_t p_imh = bim ? kPNHR : kPHI, p_iml = bim ? kPNLR : kPLI;
  // HH, HL, LL, LH of Re; the same of Im; acc[8] holds the frame sums of the
  // digits of the A rows: column quarter z of row i sums operand z over the
  // frames (z = Re high, Re low, Im high, Im low), since the all-ones B
  // operand of the z-th sum product is nonzero only in quarter z
  v4i acc[9];
#pragma unroll
  for (int z = 0; z < 9; ++z) acc[z] = v4i{0, 0, 0, 0};
  const uint32_t cq = cl >> 2;

  if (!loader) __syncthreads();  // the loader's step 0 is in the ring
  // T1: the rounding constants 128 * (beam sum) + c of the lane's beams
  // 4 grp + r, in the accumulator layout (row r of the lane): they start the
  // LL products' accumulation, so the epilogue does not add them
  v4i c_re = {0, 0, 0, 0}, c_im = {0, 0, 0, 0};
  if constexpr (T1) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int32_t sa = tail[2 * (4 * grp + r)], wis = tail[2 * (4 * grp + r) + 1];
      c_re[r] = (sa << 7) + wis - kconst_bf;
      c_im[r] = (sa << 7) - kconst_bf;
    }
  }
  for (uint32_t st = 0; st < nsteps && !loader; ++st) {
    unsigned char* P = sm + (st & 1) * G::BUF;
    const uint32_t m0 = st * FS;
    {  // digits of this thread's four channel words into the planes
      const unsigned char* rs = raw + (st % kFbRing) * kFbSlot + roff;
      const uint32_t w[4] = {*reinterpret_cast<const uint32_t*>(rs), *reinterpret_cast<const uint32_t*>(rs + 512),
                             *reinterpret_cast<const uint32_t*>(rs + kFbBlk),
                             *reinterpret_cast<const uint32_t*>(rs + kFbBlk + 512)};
      const uint32_t bm = m0 + cf < F ? chmask : 0u;
      const uint32_t hr = gather4(w[0], w[1], w[2], w[3], 1) & bm;
      const uint32_t lr = (gather4(w[0], w[1], w[2], w[3], 0) ^ 0x80808080u) & bm;
      const uint32_t hi = gather4(w[0], w[1], w[2], w[3], 3) & bm;
      const uint32_t li = (gather4(w[0], w[1], w[2], w[3], 2) ^ 0x80808080u) & bm;
      // plane word index of (column 4 * cf + cb, channels 4q..4q+3) is t
      uint32_t* Pw = reinterpret_cast<uint32_t*>(P) + t;
      constexpr uint32_t PW = G::PLANE / 4;
      Pw[kPHR * PW] = hr;
      Pw[kPLR * PW] = lr;
      Pw[kPHI * PW] = hi;
      Pw[kPLI * PW] = li;
      Pw[kPNHR * PW] = ~hr & bm;
      Pw[kPNLR * PW] = ~lr & bm;
      Pw[kPNHI * PW] = ~hi & bm;
      Pw[kPNLI * PW] = ~li & bm;
      if (J.bf.xa_out && m0 + cf < F) {
        uint32_t* xa = J.bf.xa_out + ((uint64_t)(m0 + cf) * N + k0 + cb) * CH + 4 * q;
        if (CH == 16 && (reinterpret_cast<uintptr_t>(J.bf.xa_out) & 15u) == 0) {
          *reinterpret_cast<uint4*>(xa) = uint4{w[0], w[1], w[2], w[3]};
        } else {
#pragma unroll
          for (uint32_t i = 0; i < 4; ++i)
            if (4 * q + i < CH) xa[i] = w[i];
        }
      }
    }
    lds_barrier();  // the planes of this step are complete (and the step before last is consumed)

    {  // beamformer: 16 columns x every beam tile
      const uint32_t m = m0 + bfr;
      const bool col_ok = m < F;
      const uint64_t col = (uint64_t)m * N + k0 + bbin;
      const uint64_t bih = *reinterpret_cast<const uint64_t*>(P + p_bih * G::PLANE + boff);
      const uint64_t bil = *reinterpret_cast<const uint64_t*>(P + p_bil * G::PLANE + boff);
      const uint64_t brh = *reinterpret_cast<const uint64_t*>(P + p_brh * G::PLANE + boff);
      const uint64_t brl = *reinterpret_cast<const uint64_t*>(P + p_brl * G::PLANE + boff);
      // column sums over the K rows: Im B = [Xr; Xi], Re B = [Xr; ~Xi]
      const v4i z4 = {0, 0, 0, 0};
      const v4i s_ih = mfma8(ONES, bih, z4), s_il = mfma8(ONES, bil, z4);
      const v4i s_rh = mfma8(ONES, brh, z4), s_rl = mfma8(ONES, brl, z4);
      const int32_t sb_im = 256 * s_ih[0] + s_il[0] + 256 * (int32_t)CH;
      const int32_t sb_re = 256 * s_rh[0] + s_rl[0] + 256 * (int32_t)CH;
      for (uint32_t bt = 0; bt < J.bf.bt_n; ++bt) {
        const uint64_t* a = reinterpret_cast<const uint64_t*>(frag) + (size_t)bt * 256 + lane;
        const uint64_t are_h = a[0], are_l = a[64], aim_h = a[128], aim_l = a[192];
        v4i rhh = mfma8(are_h, brh, z4);
        v4i rx = mfma8(are_h, brl, z4);
        rx = mfma8(are_l, brh, rx);
        v4i rll = mfma8(are_l, brl, c_re);  // + the beam constants (T1; else zero)
        v4i ihh = mfma8(aim_h, bih, z4);
        v4i ix = mfma8(aim_h, bil, z4);
        ix = mfma8(aim_l, bih, ix);
        v4i ill = mfma8(aim_l, bil, c_im);
        if (!col_ok) continue;
        const size_t o0 = (size_t)(bt * 16) * cols + lane_row + col;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t b = bt * 16 + 4 * grp + r;
          int32_t cr = 0, ci = 0;  // T1: already in rll / ill
          if constexpr (!T1) {
            if (b >= B) continue;
            const int32_t sa = tail[2 * b], wis = tail[2 * b + 1];
            cr = (sa << 7) + wis - kconst_bf;
            ci = (sa << 7) - kconst_bf;
          }
          uint32_t fr = 0, fi = 0;
          const int32_t yr = rne15_sat_small(rhh[r], rx[r], rll[r], sb_re << 7, cr, fr);
          const int32_t yi = rne15_sat_small(ihh[r], ix[r], ill[r], sb_im << 7, ci, fi);
          const size_t o = o0 + (size_t)r * cols;
          if (out & kBfOutY) J.bf.y[o] = cpack(yr, yi);
          if (out & kBfOutF) J.bf.flags[o] = (uint8_t)(fr | (fi << 2));
          if (out & kBfOutP) J.bf.power[o] = (uint32_t)(yr * yr) + (uint32_t)(yi * yi);
        }
      }
    }
    {  // covariance: one K step (16 frames) of bin cbin
      const uint64_t reh = lds_tr8(P + p_reh * G::PLANE + coff);
      const uint64_t rel = lds_tr8(P + p_rel * G::PLANE + coff);
      const uint64_t imh = lds_tr8(P + p_imh * G::PLANE + coff);
      const uint64_t iml = lds_tr8(P + p_iml * G::PLANE + coff);
      acc[0] = mfma8(reh, reh, acc[0]);
      acc[1] = mfma8(reh, rel, acc[1]);
      acc[3] = mfma8(rel, reh, acc[3]);
      acc[2] = mfma8(rel, rel, acc[2]);
      acc[4] = mfma8(imh, reh, acc[4]);
      acc[5] = mfma8(imh, rel, acc[5]);
      acc[7] = mfma8(iml, reh, acc[7]);
      acc[6] = mfma8(iml, rel, acc[6]);
      acc[8] = mfma8(reh, cq == 0 ? ONES : 0ull, acc[8]);
      acc[8] = mfma8(rel, cq == 1 ? ONES : 0ull, acc[8]);
      acc[8] = mfma8(imh, cq == 2 ? ONES : 0ull, acc[8]);
      acc[8] = mfma8(iml, cq == 3 ? ONES : 0ull, acc[8]);
    }
  }

  // the waves of a bin add their partial sums: each later K step's wave
  // stores its own copy, and the first K step's wave adds them in registers
  constexpr uint32_t KS = WAVES / kFbNB;  // K-step waves per bin
  __syncthreads();  // every wave is done with the planes (the area is reused)
  if (cks != 0 && !loader) {
    int32_t* red = reinterpret_cast<int32_t*>(sm) + ((size_t)(cks - 1) * kFbNB + cbin) * 36 * 64 + lane;
#pragma unroll
    for (int z = 0; z < 9; ++z)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[(4 * z + r) * 64] = acc[z][r];
  }
  __syncthreads();
  if (cks != 0) return;
#pragma unroll
  for (uint32_t c = 1; c < KS; ++c) {
    const int32_t* red = reinterpret_cast<const int32_t*>(sm) + ((size_t)(c - 1) * kFbNB + cbin) * 36 * 64 + lane;
#pragma unroll
    for (int z = 0; z < 9; ++z)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[z][r] += red[(4 * z + r) * 64];
  }

  // frame sums of row i = 4 grp + r: T = sum_m (xr_i + xi_i) over the Re
  // A row [Xr | Xi], U = sum_m (xi_i + ~xr_i) = sum xi - sum xr - F over the
  // Im A row [Xi | ~Xr]; each digit entry of a valid frame carries +128
  int32_t* sb = sums + (size_t)cbin * 64;
  if ((cl & 3) == 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) sb[(4 * grp + r) * 4 + cq] = acc[8][r];
  }
  wave_lds_sync_mfma();
  const int64_t Fl = F;
  auto row_t = [&](uint32_t i) { return 256 * (int64_t)sb[4 * i] + sb[4 * i + 1] + 256 * Fl; };
  auto row_u = [&](uint32_t i) { return 256 * (int64_t)sb[4 * i + 2] + sb[4 * i + 3] + 256 * Fl; };
  const int64_t Tj = row_t(cl), Uj = row_u(cl);
  const int64_t xij = (Tj + Uj + Fl) / 2;  // sum_m xi_j (exact: Tj + Uj + F = 2 sum xi_j)
  const int64_t kconst = 16384LL * 2 * Fl;
  const uint32_t k = k0 + cbin;
  const size_t npairs = (size_t)CH * (CH + 1) / 2;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t ii = 4 * grp + r, jj = cl;
    if (ii >= CH || jj >= CH || ii > jj) continue;
    const int64_t re =
        combine(acc[0], acc[1], acc[2], r) + 256 * (int64_t)acc[3][r] + 128 * (row_t(ii) + Tj) - kconst;
    const int64_t im =
        combine(acc[4], acc[5], acc[6], r) + 256 * (int64_t)acc[7][r] + 128 * (row_u(ii) + Tj) - kconst + xij;
    const size_t pair = (size_t)ii * CH - (size_t)ii * (ii - 1) / 2 + (jj - ii);
    int64_t* o = J.R + ((size_t)k * npairs + pair) * 2;
    o[0] = re;
    o[1] = im;
  }
}

template <int WAVES, int OUT, bool T1>
__global__ __launch_bounds__(64 * (WAVES + 1)) void bfcov_kernel(BfCovJob J) {
  bfcov_body<WAVES, OUT, T1>(J, xcd_tile(), J.bf.bt_n);
}

template <int WAVES, bool T1>
static hipError_t launch_bfcov_t(hipStream_t s, const BfCovJob& J) {
  const size_t lds = FbGeo<WAVES>::lds(J.bf.bt_n);
  const dim3 grid(J.N / kFbNB);
  const int out = bf_out_mask(J.bf);
  if (out == kBfOutP)
    bfcov_kernel<WAVES, kBfOutP, T1><<<grid, 64 * (WAVES + 1), lds, s>>>(J);
  else if (out == (kBfOutY | kBfOutF | kBfOutP))
    bfcov_kernel<WAVES, kBfOutY | kBfOutF | kBfOutP, T1><<<grid, 64 * (WAVES + 1), lds, s>>>(J);
  else
    bfcov_kernel<WAVES, -1, T1><<<grid, 64 * (WAVES + 1), lds, s>>>(J);
  return hipGetLastError();
}

template <int WAVES>
static hipError_t launch_bfcov_w(hipStream_t s, const BfCovJob& J) {
  return J.bf.B == 16 ? launch_bfcov_t<WAVES, true>(s, J) : launch_bfcov_t<WAVES, false>(s, J);
}

hipError_t launch_bfcov(hipStream_t s, const uint32_t* X, const uint32_t* W, uint32_t n_chan, uint32_t n_beams,
                        uint32_t n_frames, uint32_t n_bins, uint32_t* y, uint8_t* flags, uint32_t* power,
                        uint32_t* xa_out, int64_t* R, bool x_blocked) {
  // the loader's 16-byte LDS loads need X 16-byte aligned
  if (!bfcov_fits(n_chan, n_bins, n_frames, n_beams) || !R || (reinterpret_cast<uintptr_t>(X) & 15u))
    return hipErrorInvalidValue;
  BfCovJob J{};
  J.x_blocked = x_blocked ? 1u : 0u;
  J.bf = BfJob{X, W, y, flags, power, nullptr, xa_out, (uint64_t)n_frames * n_bins, n_chan, n_beams,
               (n_beams + 15) / 16, 0};
  J.R = R;
  J.N = n_bins;
  J.F = n_frames;
  // 8 waves: the block fills half of a CU's registers, so a CU can run other
  // work beside it (the chain's other stages, other streams' CPIs); 16 waves
  // finish a lone launch 8 % sooner but hold the whole CU (measured: the
  // three-stream chain runs 6 % faster with 8)
  return launch_bfcov_w<8>(s, J);
}

}  // namespace rd
