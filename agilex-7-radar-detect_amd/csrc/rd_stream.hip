// rd_stream.hip — HBM-streaming stages of the radar chain on gfx950:
// polyphase FIR bank, radix-2^2 FFT with per-sub-stage rounding, corner turn.
//
// Data are packed complex int16 words {im, re} (real low). All arithmetic is
// integer and reproduces the reference models bit for bit:
//   PFB  model/cpp/pfb/pfb_model.hpp:117-257
//   FFT  model/cpp/fft/fft_ref.hpp:250-407
//   turn model/cpp/align/align_model.hpp:172-186
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rd_fxp.h"
#include "rd_internal.h"

namespace rd {

typedef short v2s __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int32_t dot2(uint32_t a, uint32_t b, int32_t c) {
  return __builtin_amdgcn_sdot2(__builtin_bit_cast(v2s, a), __builtin_bit_cast(v2s, b), c, false);
}
// dot2(a, b, 0) in the three-operand encoding (the builtin's two-operand
// accumulate form costs a register clear per product).
__device__ __forceinline__ int32_t dot2_0(uint32_t a, uint32_t b) {
  int32_t r;
  asm("v_dot2_i32_i16 %0, %1, %2, 0" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// round-to-nearest-even by 15 of the exact value v = 256*ah + al
// (|v| < 2^38), as 64-bit adds: (v + 2^14 - 1 + bit15(v)) >> 15
// (fxp.hpp:198-208); the quotient (|q| < 2^23) fits in 32 bits.
__device__ __forceinline__ int32_t rne15_wide(int32_t ah, int32_t al) {
  const int64_t v = ((int64_t)ah << 8) + (int64_t)al;
  const int64_t t = v + 0x3FFF + ((v >> 15) & 1);
  return (int32_t)(t >> 15);
}
// The same for |al| < 2^29 - 2^14 in 32-bit arithmetic, for a caller that
// saturates the result to 16 bits: ah is clamped to +-1.5*2^22 first. An ah
// beyond that gives |256*ah| > 1.5*2^30, so |v| > 2^30 + 2^14, past the
// saturation bound 2^30 - 2^14, with the sign of ah (the clamped value keeps
// both), and the clamped sum stays below 2^31.
__device__ __forceinline__ int32_t rne15_clamp(int32_t ah, int32_t al) {
  constexpr int32_t kH = 3 << 21;
  const int32_t v = (min(max(ah, -kH), kH) << 8) + al;
  return (v + 0x3FFF + (int32_t)__builtin_amdgcn_ubfe((uint32_t)v, 15, 1)) >> 15;
}
__device__ __forceinline__ uint32_t pk16(int32_t lo, int32_t hi) {
  return ((uint32_t)hi << 16) | ((uint32_t)lo & 0xFFFFu);
}

// ============================================================================
// PFB
// ----------------------------------------------------------------------------
// Branch p of channel c owns the decimated stream x_p[m] = x[c][m*P + p] and
// computes y_p[m] = round_sat(sum_k h_p[k] x_p[m-k], 15, 16) per component with
// an exact accumulator (pfb_model.hpp:139-168).
//
// One thread per branch keeps the last T samples in registers, so each input
// word is read once (plus T-1 warm-up words per beat chunk) with coalesced
// 4-byte-per-lane loads along p. The taps stay in registers for the whole run.
//
// Exact products without 64-bit multiplies: each sample component is split
// x = 256*xh + xl (xh in [-128,127], xl in [0,255]), and
//   re = 256*sum(xrh*hr - xih*hi) + sum(xrl*hr - xil*hi)
// (and im alike). Each bracket is one packed int16 dot product
// (v_dot2_i32_i16) per tap into an int32 partial that cannot overflow for
// T <= 64 (|partial| < 2^30); the two partials are recombined exactly before
// the single rounding. With real taps (hi = 0) two taps share one dot product
// against a pair of consecutive samples (PfbRealWindow), chosen per wave.
// ============================================================================

struct Digits {
  uint32_t rh, ih, rl, il;  // (xrh,-xih) (xrh,xih) (xrl,-xil) (xrl,xil)
};

__device__ __forceinline__ Digits split(uint32_t v) {
  const int32_t xr = cre(v), xi = cim(v);
  const int32_t xrh = xr >> 8, xih = xi >> 8, xrl = xr & 255, xil = xi & 255;
  return Digits{pk16(xrh, -xih), pk16(xrh, xih), pk16(xrl, -xil), pk16(xrl, xil)};
}

// round_sat(256*ah + al, 15, 16) without 64-bit arithmetic: with
// ah = 128*qh + rh and al = 2^15*ql + rl (floor division), the exact value is
// 2^15*(qh + ql) + (256*rh + rl), and the last term lies in [0, 2^16).
__device__ __forceinline__ int32_t rne15_split(int32_t ah, int32_t al) {
  const int32_t s = ((ah & 127) << 8) + (al & 0x7FFF);
  const int32_t q = (ah >> 7) + (al >> 15) + (s >> 15);
  const int32_t rem = s & 0x7FFF;
  return q + ((rem + (q & 1)) > 0x4000 ? 1 : 0);
}

__device__ __forceinline__ uint32_t pfb_finish(int32_t arh, int32_t arl, int32_t aih, int32_t ail, uint8_t* fl) {
  uint32_t fr = 0, fi = 0;
  const int32_t yr = sat16(rne15_split(arh, arl), fr);
  const int32_t yi = sat16(rne15_split(aih, ail), fi);
  *fl = (uint8_t)(fr | (fi << 2));
  return cpack(yr, yi);
}

// x_p[m] for m in [-(T-1), n_beats): negative beats come from the imported
// history [T-1][P] (most recent first) or are zero.
__device__ __forceinline__ uint32_t pfb_sample(const uint32_t* xc, const uint32_t* hc, uint32_t P, uint32_t p,
                                               int64_t m) {
  if (m >= 0) return xc[(size_t)m * P + p];
  return hc ? hc[(size_t)(-m - 1) * P + p] : 0u;
}

// Register-window form for a compile-time T. A sample enters the window as
// two digit words, hi = (xr >> 8, xi >> 8) (one v_pk_ashrrev_i16) and
// lo = (xr & 255, xi & 255) (one AND), shared by both outputs: with the taps
// held as a = (hr, ~hi) and b = (hi, hr) (~hi = -hi - 1, so hi = -32768 needs
// no special case),
//   re = 256*sum dot2(hi, a) + sum dot2(lo, a) + sum xi
//   im = 256*sum dot2(hi, b) + sum dot2(lo, b)
// where sum xi, the sum of the imaginary parts over the window, is kept as a
// running sum (one add and one subtract per beat).
__device__ __forceinline__ uint32_t digits_hi(uint32_t v) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(v2s, v) >> (v2s){8, 8});
}
__device__ __forceinline__ uint32_t digits_lo(uint32_t v) { return v & 0x00FF00FFu; }

template <int T>
struct PfbWindow {
  uint32_t hi[T], lo[T];
  int32_t sum_xi = 0;
  // slot j takes sample v; the sample T beats older leaves it, its imaginary
  // part read back from its digits (256 * (xi >> 8) + (xi & 255))
  __device__ __forceinline__ void put(int j, uint32_t v) {
    const int32_t old = (((int32_t)hi[j] >> 16) << 8) + (int32_t)(lo[j] >> 16);
    sum_xi += cim(v) - old;
    hi[j] = digits_hi(v);
    lo[j] = digits_lo(v);
  }
  // output of the beat whose sample sits in slot j
  template <bool FLAGS>
  __device__ __forceinline__ uint32_t out(int j, const uint32_t (&ta)[T], const uint32_t (&tb)[T], uint8_t* f) const {
    int32_t arh = dot2_0(hi[j], ta[0]), arl = dot2_0(lo[j], ta[0]);
    int32_t aih = dot2_0(hi[j], tb[0]), ail = dot2_0(lo[j], tb[0]);
#pragma unroll
    for (int k = 1; k < T; ++k) {
      const int d = (j - k + T) % T;
      arh = dot2(hi[d], ta[k], arh);
      arl = dot2(lo[d], ta[k], arl);
      aih = dot2(hi[d], tb[k], aih);
      ail = dot2(lo[d], tb[k], ail);
    }
    // |arl|, |ail| < T * 2^24 and |sum_xi| <= T * 2^15: within rne15_clamp's range
    static_assert(T <= 16, "register window sized for T <= 16");
    const int32_t qr = rne15_clamp(arh, arl + sum_xi), qi = rne15_clamp(aih, ail);
    if constexpr (FLAGS) {
      uint32_t fr = 0, fi = 0;
      const int32_t yr = sat16(qr, fr), yi = sat16(qi, fi);
      *f = (uint8_t)(fr | (fi << 2));
      return cpack(yr, yi);
    } else {
      return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pk_i16(qr, qi));  // both halves clamped
    }
  }
};

// Real taps (every h_p[k] with a zero imaginary part, the windowed-sinc
// channelizer's case): re = sum hr*xr and im = sum hr*xi, so two taps share one
// packed dot product against a pair of consecutive samples. Slot j keeps the
// digit pairs formed when its sample n arrived, (xrh[n], xrh[n-1]) and
// (xih[n], xih[n-1]), and the same for the low digits; the beat in slot j
// reads slots j, j-2, ..., j-T+2 against the tap pairs (h[2i], h[2i+1]):
// T/2 dot products per digit plane instead of T, and no sum_xi term.
template <int T>
struct PfbRealWindow {
  uint32_t rh[T], ih[T], rl[T], il[T];
  uint32_t ph = 0, pl = 0;  // digit words of the previous sample
  __device__ __forceinline__ void put(int j, uint32_t v) {
    const uint32_t h = digits_hi(v), l = digits_lo(v);
    rh[j] = __builtin_amdgcn_perm(ph, h, 0x05040100u);  // (h.re, prev.re)
    ih[j] = __builtin_amdgcn_perm(ph, h, 0x07060302u);  // (h.im, prev.im)
    rl[j] = __builtin_amdgcn_perm(pl, l, 0x05040100u);
    il[j] = __builtin_amdgcn_perm(pl, l, 0x07060302u);
    ph = h;
    pl = l;
  }
  template <bool FLAGS>
  __device__ __forceinline__ uint32_t out(int j, const uint32_t (&tp)[T / 2], uint8_t* f) const {
    int32_t arh = dot2_0(rh[j], tp[0]), arl = dot2_0(rl[j], tp[0]);
    int32_t aih = dot2_0(ih[j], tp[0]), ail = dot2_0(il[j], tp[0]);
#pragma unroll
    for (int i = 1; i < T / 2; ++i) {
      const int d = (j - 2 * i + T) % T;
      arh = dot2(rh[d], tp[i], arh);
      arl = dot2(rl[d], tp[i], arl);
      aih = dot2(ih[d], tp[i], aih);
      ail = dot2(il[d], tp[i], ail);
    }
    // |arl|, |ail| < T * 2^23: within rne15_clamp's range
    static_assert(T <= 16, "register window sized for T <= 16");
    const int32_t qr = rne15_clamp(arh, arl), qi = rne15_clamp(aih, ail);
    if constexpr (FLAGS) {
      uint32_t fr = 0, fi = 0;
      const int32_t yr = sat16(qr, fr), yi = sat16(qi, fi);
      *f = (uint8_t)(fr | (fi << 2));
      return cpack(yr, yi);
    } else {
      return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pk_i16(qr, qi));  // both halves clamped
    }
  }
};

// Real taps whose magnitudes sum to at most 65535 on the branch (the
// windowed-sinc channelizer's taps: at most 57429 for the 1024- and 4096-branch
// banks): every partial sum of h * x is then within 65535 * 2^15 = 2^31 - 2^15
// for any int16 samples, so the products need no digit split and the int32
// sums are the exact accumulator. Slot j keeps the raw component pairs
// (xr[n], xr[n-1]) and (xi[n], xi[n-1]); an output is T/2 dot products per
// component (half of PfbRealWindow's) and one rounding, whose bias keeps the
// sum below 2^31.
template <int T>
struct PfbRealDirect {
  uint32_t rp[T], ip[T];
  // the previous sample's components are the low halves of the previous
  // slot's pairs, so a loaded sample is dead once it has entered the window
  __device__ __forceinline__ void put(int j, uint32_t v) {
    const int q = (j + T - 1) % T;
    rp[j] = __builtin_amdgcn_perm(rp[q], v, 0x05040100u);  // (v.re, prev.re)
    ip[j] = __builtin_amdgcn_perm(ip[q], v, 0x05040302u);  // (v.im, prev.im)
  }
  template <bool FLAGS>
  __device__ __forceinline__ uint32_t out(int j, const uint32_t (&tp)[T / 2], uint8_t* f) const {
    int32_t ar = dot2_0(rp[j], tp[0]), ai = dot2_0(ip[j], tp[0]);
#pragma unroll
    for (int i = 1; i < T / 2; ++i) {
      const int d = (j - 2 * i + T) % T;
      ar = dot2(rp[d], tp[i], ar);
      ai = dot2(ip[d], tp[i], ai);
    }
    const int32_t qr = (ar + 0x3FFF + (int32_t)__builtin_amdgcn_ubfe((uint32_t)ar, 15, 1)) >> 15;
    const int32_t qi = (ai + 0x3FFF + (int32_t)__builtin_amdgcn_ubfe((uint32_t)ai, 15, 1)) >> 15;
    if constexpr (FLAGS) {
      uint32_t fr = 0, fi = 0;
      const int32_t yr = sat16(qr, fr), yi = sat16(qi, fi);
      *f = (uint8_t)(fr | (fi << 2));
      return cpack(yr, yi);
    } else {
      return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pk_i16(qr, qi));  // both halves clamped
    }
  }
};

// The beat loop of one branch over its chunk [m0, m1) for any window:
// slot (m - m0) mod T holds x_p[m]. The warm-up feeds x_p[m0-T+1 .. m0-1]
// into slots 1..T-1 oldest first (the real windows pair each sample with the
// one before it), and slot 0 starts empty.
template <int T, bool FLAGS, class Win, class Out>
__device__ __forceinline__ void pfb_window_run(const uint32_t* xc, const uint32_t* hc, uint32_t P, uint32_t p,
                                               uint32_t m0, uint32_t m1, uint32_t* yc, uint8_t* fc, Win& w,
                                               Out out) {
#pragma unroll
  for (int k = T - 1; k >= 1; --k) w.put(T - k, pfb_sample(xc - p, hc, P, p, (int64_t)m0 - k));

  // Raw samples of the current block of T beats; as soon as sample j enters
  // the window its register is reloaded with sample j of the next block, so
  // the load has the rest of the block to land (indices clamped into the
  // chunk, so every load is in bounds; values past m1 are never used).
  uint32_t cur[T];
#pragma unroll
  for (int j = 0; j < T; ++j) cur[j] = xc[(size_t)min(m0 + j, m1 - 1) * P];
  uint32_t mb = m0;
  for (; mb + T <= m1; mb += T) {  // full blocks: no per-beat bounds
#pragma unroll
    for (int j = 0; j < T; ++j) {
      const size_t o = (size_t)(mb + j) * P;
      w.put(j, cur[j]);
      cur[j] = xc[(size_t)min(mb + T + j, m1 - 1) * P];
      uint8_t f;
      yc[o] = out(j, &f);
      if (FLAGS) fc[o] = f;
    }
  }
#pragma unroll
  for (int j = 0; j < T; ++j) {  // the last, partial block
    if (mb + j >= m1) break;
    const size_t o = (size_t)(mb + j) * P;
    w.put(j, cur[j]);
    uint8_t f;
    yc[o] = out(j, &f);
    if (FLAGS) fc[o] = f;
  }
}

// One tile: branches [bx*blockDim, +blockDim) of channel c over beat chunk z.
// FLAGS: the job has a flag buffer (otherwise the flags are not formed). A
// wave whose branches all have real taps takes a real window: the direct one
// when every branch's tap magnitudes sum to at most 65535, else the digit one.
template <int T, bool FLAGS>
__device__ __forceinline__ void pfb_reg_tile(const PfbJob& J, uint32_t bx, uint32_t c, uint32_t z) {
  const uint32_t P = J.P, n_beats = J.n_beats;
  const uint32_t p = bx * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const uint32_t m0 = z * J.chunk;
  const uint32_t m1 = min(m0 + J.chunk, n_beats);
  const size_t span = (size_t)n_beats * P;
  const uint32_t* xc = J.x + (size_t)c * span + p;
  const uint32_t* hc = J.hist ? J.hist + (size_t)c * (T - 1) * P : nullptr;
  uint32_t* yc = J.y + (size_t)c * span + p;
  uint8_t* fc = J.flags ? J.flags + (size_t)c * span + p : nullptr;

  uint32_t h[T], imag = 0, mag = 0;
#pragma unroll
  for (int k = 0; k < T; ++k) {
    h[k] = J.coeff[(size_t)p * T + k];  // (hr, hi)
    imag |= h[k] >> 16;
    mag += (uint32_t)abs(cre(h[k]));
  }
  if (__builtin_amdgcn_ballot_w64(imag != 0) == 0) {
    uint32_t tp[T / 2];
#pragma unroll
    for (int i = 0; i < T / 2; ++i) tp[i] = __builtin_amdgcn_perm(h[2 * i + 1], h[2 * i], 0x05040100u);
    if (__builtin_amdgcn_ballot_w64(mag > 65535u) == 0) {
      PfbRealDirect<T> w;
#pragma unroll
      for (int k = 0; k < T; ++k) w.rp[k] = w.ip[k] = 0u;
      pfb_window_run<T, FLAGS>(xc, hc, P, p, m0, m1, yc, fc, w,
                               [&](int j, uint8_t* f) { return w.template out<FLAGS>(j, tp, f); });
    } else {
      PfbRealWindow<T> w;
#pragma unroll
      for (int k = 0; k < T; ++k) w.rh[k] = w.ih[k] = w.rl[k] = w.il[k] = 0u;
      pfb_window_run<T, FLAGS>(xc, hc, P, p, m0, m1, yc, fc, w,
                               [&](int j, uint8_t* f) { return w.template out<FLAGS>(j, tp, f); });
    }
  } else {
    uint32_t ta[T], tb[T];
#pragma unroll
    for (int k = 0; k < T; ++k) {
      ta[k] = h[k] ^ 0xFFFF0000u;            // (hr, ~hi)
      tb[k] = (h[k] >> 16) | (h[k] << 16);  // (hi, hr)
    }
    PfbWindow<T> w;
#pragma unroll
    for (int k = 0; k < T; ++k) w.hi[k] = w.lo[k] = 0u;
    pfb_window_run<T, FLAGS>(xc, hc, P, p, m0, m1, yc, fc, w,
                             [&](int j, uint8_t* f) { return w.template out<FLAGS>(j, ta, tb, f); });
  }
}

// Any T in [1, 64]: the window is re-read from memory (L1/L2) each beat.
__device__ __forceinline__ void pfb_any_tile(const PfbJob& J, uint32_t bx, uint32_t c, uint32_t z) {
  const uint32_t P = J.P, T = J.T, n_beats = J.n_beats;
  const uint32_t p = bx * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const uint32_t m0 = z * J.chunk;
  const uint32_t m1 = min(m0 + J.chunk, n_beats);
  const size_t span = (size_t)n_beats * P;
  const uint32_t* xc = J.x + (size_t)c * span;
  const uint32_t* hc = J.hist ? J.hist + (size_t)c * (T - 1) * P : nullptr;
  for (uint32_t m = m0; m < m1; ++m) {
    int32_t arh = 0, arl = 0, aih = 0, ail = 0;
    for (uint32_t k = 0; k < T; ++k) {
      const uint32_t h = J.coeff[(size_t)p * T + k];
      const uint32_t hs = (h >> 16) | (h << 16);
      const Digits d = split(pfb_sample(xc, hc, P, p, (int64_t)m - (int64_t)k));
      arh = dot2(d.rh, h, arh);
      arl = dot2(d.rl, h, arl);
      aih = dot2(d.ih, hs, aih);
      ail = dot2(d.il, hs, ail);
    }
    uint8_t f;
    J.y[(size_t)c * span + (size_t)m * P + p] = pfb_finish(arh, arl, aih, ail, &f);
    if (J.flags) J.flags[(size_t)c * span + (size_t)m * P + p] = f;
  }
}

template <int T, bool FLAGS>
__device__ __forceinline__ void pfb_tile(const PfbJob& J, uint32_t bx, uint32_t c, uint32_t z) {
  if constexpr (T > 0)
    pfb_reg_tile<T, FLAGS>(J, bx, c, z);
  else
    pfb_any_tile(J, bx, c, z);
}

// Waves per SIMD asked of the register-window kernels: at T = 16 the real
// window's 64 pair registers need about 142 VGPRs (three waves) to run
// unspilled.
constexpr int pfb_min_waves(int T, bool flags) { return flags ? 1 : (T == 16 ? 3 : 4); }

// Single CPI: grid (nbx, n_chan, nz). T = 0 selects the generic kernel.
template <int T, bool FLAGS>
__global__ __launch_bounds__(256, pfb_min_waves(T, FLAGS)) void pfb_kernel(PfbJob J) {
  pfb_tile<T, FLAGS>(J, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Batched: one block per tile of a class of CPIs (block -> job via the tile
// table); a class holds CPIs that all keep, or all drop, the PFB flags.
template <int T, bool FLAGS>
__global__ __launch_bounds__(256, pfb_min_waves(T, FLAGS)) void pfb_batch_kernel(const PfbJob* __restrict__ jobs,
                                                                       const uint32_t* __restrict__ tile_job,
                                                                       uint32_t n_tiles) {
  const uint32_t t = blockIdx.x;
  if (t >= n_tiles) return;
  const PfbJob& J = jobs[__builtin_amdgcn_readfirstlane(tile_job[t])];
  const uint32_t local = t - J.tile0;
  const uint32_t bx = local % J.nbx, rest = local / J.nbx;
  pfb_tile<T, FLAGS>(J, bx, rest % J.n_chan, rest / J.n_chan);
}

// Final delay-line state after n_beats: hist[k-1] = x_p[n_beats - k], k = 1..T-1.
// Runs after the filter kernel (stream order); k descends so an in-place read of
// an older imported entry (index k-1-n_beats < k-1) happens before it is overwritten.
__global__ void pfb_hist_kernel(const uint32_t* __restrict__ x, uint32_t P, uint32_t T, uint32_t n_beats,
                                uint32_t* hist) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const uint32_t c = blockIdx.y;
  const uint32_t* xc = x + (size_t)c * n_beats * P;
  uint32_t* hc = hist + (size_t)c * (T - 1) * P;
  for (uint32_t k = T - 1; k >= 1; --k) {
    const int64_t m = (int64_t)n_beats - (int64_t)k;
    hc[(size_t)(k - 1) * P + p] = pfb_sample(xc, hc, P, p, m);
  }
}

hipError_t launch_pfb(hipStream_t s, const uint32_t* x, const uint32_t* coeff, uint32_t P, uint32_t T,
                      uint32_t n_chan, uint32_t n_beats, uint32_t* hist, uint32_t* y, uint8_t* flags) {
  if (n_beats == 0 || n_chan == 0) return hipSuccess;
  const uint32_t bx = (P + 255) / 256;
  // Beat chunks: about 2048 waves (8 per CU), so a chunk is usually a few
  // blocks of T beats and the reload of the next block overlaps the current
  // one (measured best between 1024 and 4096 waves on the medium workload); a
  // chunk never gets shorter than T beats, which bounds the warm-up re-reads
  // (T - 1 beats per chunk, mostly L2 hits) to under 2x.
  const uint64_t lanes = (uint64_t)P * n_chan;
  const uint64_t want = (131072 + lanes - 1) / lanes;
  uint32_t chunk = (uint32_t)((n_beats + want - 1) / want);
  if (chunk < T) chunk = T;
  if (chunk > n_beats) chunk = n_beats;
  const uint32_t nz = (n_beats + chunk - 1) / chunk;
  const PfbJob J{x, coeff, hist, y, flags, P, T, n_chan, n_beats, chunk, bx, nz, 0};
  const dim3 grid(bx, n_chan, nz);
  const dim3 block(P < 256 ? ((P + 63) / 64) * 64 : 256);
  if (flags) {
    switch (T) {
      case 4: pfb_kernel<4, true><<<grid, block, 0, s>>>(J); break;
      case 8: pfb_kernel<8, true><<<grid, block, 0, s>>>(J); break;
      case 16: pfb_kernel<16, true><<<grid, block, 0, s>>>(J); break;
      default: pfb_kernel<0, true><<<grid, block, 0, s>>>(J); break;
    }
  } else {
    switch (T) {
      case 4: pfb_kernel<4, false><<<grid, block, 0, s>>>(J); break;
      case 8: pfb_kernel<8, false><<<grid, block, 0, s>>>(J); break;
      case 16: pfb_kernel<16, false><<<grid, block, 0, s>>>(J); break;
      default: pfb_kernel<0, false><<<grid, block, 0, s>>>(J); break;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (hist && T > 1) {
    pfb_hist_kernel<<<dim3(bx, n_chan), block, 0, s>>>(x, P, T, n_beats, hist);
    e = hipGetLastError();
  }
  return e;
}

template <bool FLAGS>
static void launch_pfb_batch_f(hipStream_t s, uint32_t T, const ClassLaunch& L) {
  const PfbJob* jobs = static_cast<const PfbJob*>(L.jobs);
  switch (T) {
    case 4: pfb_batch_kernel<4, FLAGS><<<L.grid, 256, 0, s>>>(jobs, L.tile_job, L.n_tiles); break;
    case 8: pfb_batch_kernel<8, FLAGS><<<L.grid, 256, 0, s>>>(jobs, L.tile_job, L.n_tiles); break;
    case 16: pfb_batch_kernel<16, FLAGS><<<L.grid, 256, 0, s>>>(jobs, L.tile_job, L.n_tiles); break;
    default: pfb_batch_kernel<0, FLAGS><<<L.grid, 256, 0, s>>>(jobs, L.tile_job, L.n_tiles); break;
  }
}

hipError_t launch_pfb_batch(hipStream_t s, uint32_t T, bool flags, const ClassLaunch& L) {
  if (L.n_tiles == 0) return hipSuccess;
  if (flags)
    launch_pfb_batch_f<true>(s, T, L);
  else
    launch_pfb_batch_f<false>(s, T, L);
  return hipGetLastError();
}

// ============================================================================
// FFT
// ----------------------------------------------------------------------------
// One frame lives in LDS; each thread owns 4 samples of it per pass. A pass
// applies two sub-stages (BF2I at even s, BF2II at s+1 with its trivial -j by
// saturating negate) and, when s+2 < n_lane, the twiddle multiplier, all in
// registers, so a 2^k-point lane needs ceil(k/2) LDS round trips. The last odd
// sub-stage and the spc = 2 decimation-in-time merge form a final pass.
// Every butterfly output is rounded by its schedule bit then saturated; every
// twiddle product is rounded once at shift 15 (fft_ref.hpp:315-406).
//
// Lane q of an spc-lane transform is the subsequence x[spc*t + q]; the lanes sit
// back to back in LDS, which leaves every block index parity and twiddle phase
// of the second lane equal to the first's (block sizes divide the lane length).
//
// Twiddle products stay in 32 bits: every table entry has |w.re| + |w.im| <=
// 65534 (checked by the tests), so |x.re*w.re -/+ x.im*w.im| + 32768 < 2^31.
// ============================================================================

struct Cv {
  int32_t re, im;
};

__device__ __forceinline__ Cv ld_cv(uint32_t p) { return Cv{cre(p), cim(p)}; }
__device__ __forceinline__ uint32_t pk(Cv c) { return __builtin_amdgcn_perm((uint32_t)c.im, (uint32_t)c.re, 0x05040100u); }

// Per sub-stage saturation evidence: the extremes of every value that is about
// to be saturated to 16 bits; the stage's {sat_pos, sat_neg} word is then
// (max > 32767, min < -32768) — the OR over the frame (fft_ref.hpp:315-319).
struct StageFlags {
  int32_t mx = -32768, mn = 32767;
  __device__ __forceinline__ void see(int32_t a, int32_t b) {
    mx = max(mx, max(a, b));
    mn = min(mn, min(a, b));
  }
  __device__ __forceinline__ uint32_t word() const { return (mx > 32767 ? 2u : 0u) | (mn < -32768 ? 1u : 0u); }
};

__device__ __forceinline__ int32_t sat(int32_t r) { return min(max(r, -32768), 32767); }

// round-to-nearest-even by 1: (v + bit1(v)) >> 1 (the tie bit makes the quotient even)
__device__ __forceinline__ int32_t rne1(int32_t v) { return (v + (int32_t)__builtin_amdgcn_ubfe((uint32_t)v, 1, 1)) >> 1; }
// a + b as one v_add_u32 the optimiser cannot look through: with both operands
// known to be sign-extended int16 it otherwise rewrites bit 1 of the sum as a
// five-instruction 16-bit halving-add sequence.
__device__ __forceinline__ int32_t add_v(int32_t a, int32_t b) {
  int32_t r;
  asm("v_add_u32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
// round-to-nearest-even by 15 for |v| < 2^31 - 2^14: (v + 2^14 - 1 + bit15(v)) >> 15
__device__ __forceinline__ int32_t rne15(int32_t v) {
  return (v + 0x3FFF + (int32_t)__builtin_amdgcn_ubfe((uint32_t)v, 15, 1)) >> 15;
}

// lo = q(a + b), hi = q(a - b), q = round by the schedule bit then saturate.
// A scaled sub-stage needs less: the rounded half-sum of two int16 values lies
// in [-32768, 32767], and the rounded half-difference in [-32768, 32768], so
// only the difference can saturate, and only upwards.
template <bool ALL1>
__device__ __forceinline__ void bfly(Cv& a, Cv& b, bool sh, StageFlags& f) {
  int32_t sr = add_v(a.re, b.re), si = add_v(a.im, b.im), dr = a.re - b.re, di = a.im - b.im;
  if (ALL1 || sh) {
    sr = rne1(sr);
    si = rne1(si);
    dr = rne1(dr);
    di = rne1(di);
    f.mx = max(f.mx, max(dr, di));
    a = Cv{sr, si};
    b = Cv{min(dr, 32767), min(di, 32767)};
  } else {
    f.see(sr, si);
    f.see(dr, di);
    a = Cv{sat(sr), sat(si)};
    b = Cv{sat(dr), sat(di)};
  }
}

// -j * (re + j im) = im - j re with a saturating negate (fft_ref.hpp:322-331)
__device__ __forceinline__ Cv neg_j(Cv c, StageFlags& f) {
  const int32_t n = -c.re;
  f.mx = max(f.mx, n);
  return Cv{c.im, min(n, 32767)};
}

// x * w rounded once at shift 15 (fxp.hpp:314-318), as two packed int16 dot
// products against the pair the twiddle stage keeps per entry,
// wc = (w.re, ~w.im) and ws = (w.im, w.re):
//   re = x.re*w.re - x.im*w.im = dot2((re, im), wc) + x.im   (~w = -w - 1)
//   im = x.re*w.im + x.im*w.re = dot2((re, im), ws)
// so no operand is ever negated (w.im = -32768 occurs). Every table entry has
// |w.re| + |w.im| <= 65534, so every partial sum fits in int32.
__device__ __forceinline__ uint32_t tw_conj_compl(uint32_t w) { return w ^ 0xFFFF0000u; }
__device__ __forceinline__ uint32_t tw_swap(uint32_t w) { return __builtin_amdgcn_alignbit(w, w, 16); }
__device__ __forceinline__ Cv cmul_tw(Cv x, uint32_t wc, uint32_t ws, StageFlags& f) {
  const uint32_t xp = pk(x);
  const int32_t re = dot2(xp, wc, x.im);
  const int32_t im = dot2_0(xp, ws);
  const int32_t yr = rne15(re), yi = rne15(im);
  f.see(yr, yi);
  return Cv{sat(yr), sat(yi)};
}

__device__ __forceinline__ uint32_t pad(uint32_t p) { return p + (p >> 5); }

// The words at b + {0, O1, O2, O3} of a padded frame. With every O a multiple
// of 32, pad(b + O) = pad(b) + O + O / 32, so the four reads share one address.
template <uint32_t O1, uint32_t O2, uint32_t O3>
__device__ __forceinline__ uint4 lds_quad(const uint32_t* v, uint32_t b) {
  if constexpr (O1 % 32 == 0 && O2 % 32 == 0 && O3 % 32 == 0) {
    const uint32_t* q = v + pad(b);
    return make_uint4(q[0], q[O1 + O1 / 32], q[O2 + O2 / 32], q[O3 + O3 / 32]);
  } else {
    return make_uint4(v[pad(b)], v[pad(b + O1)], v[pad(b + O2)], v[pad(b + O3)]);
  }
}

// ----------------------------------------------------------------------------
// The all-ones schedule (every sub-stage scales by 1, the chain's) runs its
// butterflies on the packed words themselves, both components per instruction
// (v_pk_* on the int16 halves), with x = a ^ b and s = x >> 1 (arithmetic),
// from a + b = 2(a | b) - x and a - b = x - 2(~a & b):
//   ceil((a + b) / 2)  = (a | b) - s
//   floor((a - b) / 2) = s - (~a & b)
// x & 1 marks an odd sum (and difference). Rounded to nearest even, the
// half-sum is the ceiling with bit 0 cleared where x is odd, and the
// half-difference is the floor plus (x & floor & 1). The half-sum always fits
// in int16. The rounded half-difference reaches +32768 only for
// (32767, -32768): a saturating add clamps it, and the sign flip of the
// wrapping add (bit 15 of wrapped & ~floor) is the sat_pos evidence. These
// identities hold for every pair of int16 values (checked exhaustively against
// rne1 and the clamp when they were written).
// ----------------------------------------------------------------------------
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s16x2 S2(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ uint32_t U2(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// Saturation evidence of one packed sub-stage: bits 15 / 31 of `ev` mark an
// upward clamp in the low / high half; mx, mn bound the twiddle products
// before saturation.
// ON = false: the caller keeps no stage flags, and the evidence is not formed
// (the data path is the same).
template <bool ON>
struct PkFlagsT {
  static constexpr bool on = ON;
  uint32_t ev = 0;
  int32_t mx = -32768, mn = 32767;
  __device__ __forceinline__ uint32_t word() const {
    if constexpr (!ON) return 0u;
    return ((ev & 0x80008000u) != 0u || mx > 32767 ? 2u : 0u) | (mn < -32768 ? 1u : 0u);
  }
};

template <class F>
__device__ __forceinline__ void bfly_pk(uint32_t& a, uint32_t& b, F& f) {
  const uint32_t x = a ^ b, odd = x & 0x00010001u;
  const s16x2 sx = S2(x) >> (s16x2){1, 1};
  const uint32_t cs = U2(S2(a | b) - sx);   // ceil((a + b) / 2)
  const uint32_t fd = U2(sx - S2(~a & b));  // floor((a - b) / 2)
  const uint32_t td = odd & fd;
  if constexpr (F::on) f.ev |= U2(S2(fd) + S2(td)) & ~fd;
  a = cs & ~odd;
  b = U2(__builtin_elementwise_add_sat(S2(fd), S2(td)));
}

// -j * (re + j im) = im - j re: the halves swap and the new high half is
// negated with saturation (fft_ref.hpp:322-331); the wrapping and the
// saturating difference disagree (in bit 31) exactly when re = -32768.
template <class F>
__device__ __forceinline__ uint32_t negj_pk(uint32_t c, F& f) {
  const s16x2 lo = S2(c >> 16), hi = S2(c << 16);
  const s16x2 r = __builtin_elementwise_sub_sat(lo, hi);
  if constexpr (F::on) f.ev |= U2(lo - hi) ^ U2(r);
  return U2(r);
}

// x * W^0 = x * 0x7FFF rounded at 15 (the product is formed, never bypassed:
// 0x7FFF is not 1): per component x - 1 for x > 16384, x + 1 for x < -16384,
// x otherwise (-32768 gives -32767). Never saturates. The two masks come
// from saturating int16 arithmetic; checked against the rounded product for
// all 65536 values.
__device__ __forceinline__ uint32_t mul_w0_pk(uint32_t x) {
  const s16x2 v = S2(x), k = {16384, 16384}, sh = {15, 15};
  const s16x2 gt = __builtin_elementwise_sub_sat(k, v) >> sh;  // -1 where x > 16384
  const s16x2 lt = __builtin_elementwise_add_sat(v, k) >> sh;  // -1 where x < -16384
  return U2(v + gt - lt);
}

// x * w on a packed x (cmul_tw's arithmetic), saturated into a packed word
// by v_cvt_pk_i16_i32 (both halves clamped to int16).
template <class F>
__device__ __forceinline__ uint32_t cmul_pk(uint32_t x, uint32_t wc, uint32_t ws, F& f) {
  const int32_t re = dot2(x, wc, (int32_t)x >> 16);
  const int32_t im = dot2_0(x, ws);
  const int32_t yr = rne15(re), yi = rne15(im);
  if constexpr (F::on) {
    f.mx = max(f.mx, max(yr, yi));
    f.mn = min(f.mn, min(yr, yi));
  }
  return U2(__builtin_amdgcn_cvt_pk_i16(yr, yi));
}



// ----------------------------------------------------------------------------
// One frame per LOG2N-specialised geometry; every index below is a compile-time
// expression of the pass and the thread's position, so the passes unroll fully.
// ALL1: every sub-stage scales by 1 (the chain's schedule); otherwise the
// schedule bit of each sub-stage is read at run time.
// ----------------------------------------------------------------------------
// LDS layout of a block: [NT] wc(W_N^e) at pad(e), [NT] ws(W_N^e) at NT +
// pad(e) (NT a multiple of 64 words, so one ds_read2st64 fetches both),
// [FPB] stage-flag words, then FPB frames of NP words each.
template <int LOG2N, int PPT = 4>
struct FftGeo {
  static constexpr uint32_t N = 1u << LOG2N;
  static constexpr uint32_t TPF = N / PPT;                     // threads per frame (PPT points each)
  static constexpr uint32_t FPB = TPF >= 256 ? 1 : 256 / TPF;  // frames per block
  static constexpr uint32_t NP = N + N / 32;
  static constexpr uint32_t NT = (NP + 63) / 64 * 64;
  static constexpr size_t lds_bytes = (size_t)(2 * NT + FPB + FPB * NP) * sizeof(uint32_t);
};

// BS: the block size (a multiple of 64); entry i + k*BS pads to pad(i) +
// k*(BS + BS/32).
template <int LOG2N, uint32_t BS>
__device__ __forceinline__ void fft_stage_twiddles(const uint32_t* __restrict__ tw4096, uint32_t* tw) {
  constexpr uint32_t N = FftGeo<LOG2N>::N, NT = FftGeo<LOG2N>::NT;
  const uint32_t i = threadIdx.x, pi = pad(i);
#pragma unroll
  for (uint32_t k = 0; k < (N + BS - 1) / BS; ++k) {
    if (N % BS != 0 && i + k * BS >= N) break;
    const uint32_t w = tw4096[(i + k * BS) << (12 - LOG2N)];
    tw[pi + k * (BS + BS / 32)] = tw_conj_compl(w);
    tw[NT + pi + k * (BS + BS / 32)] = tw_swap(w);
  }
}

// One tile: frames [grp*FPB, +FPB) of job J, twiddles already in LDS. Every
// thread of the block calls it (it holds block barriers).
// Word offset of bin k of frame `frame` in the `bins` output: natural
// frame * N + k, or with J.bins_fpc set, bin-quad blocked [ch][k / 4][m][k % 4]
// for frame = ch * bins_fpc + m; `base` is the frame's part, bins_at(base, k)
// adds the bin's (four consecutive bins from a multiple of 4 stay contiguous).
struct BinsAt {
  size_t base;
  uint32_t quad;  // words per bin quad: 4 * frames per channel, or 0 (natural)
  __device__ __forceinline__ size_t at(uint32_t k) const {
    return quad ? base + (size_t)(k >> 2) * quad + (k & 3) : base + k;
  }
};
__device__ __forceinline__ BinsAt bins_at(const FftJob& J, uint32_t frame, uint32_t N) {
  if (!J.bins_fpc) return BinsAt{(size_t)frame * N, 0};
  const uint32_t ch = frame / J.bins_fpc, m = frame - ch * J.bins_fpc;
  return BinsAt{(size_t)ch * J.bins_fpc * N + 4 * m, 4 * J.bins_fpc};
}

template <int LOG2N, int SPC, bool ALL1, bool FL>
__device__ __forceinline__ void fft_tile(const FftJob& J, uint32_t grp, uint32_t* smem) {
  using G = FftGeo<LOG2N>;
  using PkFlags = PkFlagsT<FL>;
  constexpr uint32_t N = G::N;
  constexpr uint32_t M = N / SPC;
  constexpr uint32_t NL = LOG2N - (SPC == 2 ? 1 : 0);
  constexpr uint32_t TPF = G::TPF, FPB = G::FPB, NP = G::NP;
  const uint32_t* __restrict__ in = J.in;
  uint32_t* __restrict__ out = J.out;
  uint32_t* __restrict__ bins = J.bins;
  const uint32_t sched = ALL1 ? ~0u : J.sched;
  const uint32_t reorder = J.reorder;
  constexpr uint32_t NT = G::NT;
  const uint32_t* twc = smem;       // wc of W_N^e at pad(e)
  const uint32_t* tws = smem + NT;  // ws of W_N^e at pad(e)
  uint32_t* sfl = smem + 2 * NT;
  const uint32_t lf = threadIdx.x / TPF;      // local frame
  const uint32_t g = threadIdx.x % TPF;
  uint32_t* v = smem + 2 * NT + FPB + lf * NP;  // this frame's samples at pad(p)
  const uint32_t frame = grp * FPB + lf;
  const bool active = frame < J.n_frames;

  if (g == 0) sfl[lf] = 0;
  if (active) {
    const uint4 q4 = *reinterpret_cast<const uint4*>(in + (size_t)frame * N + 4 * g);
    const uint32_t w4[4] = {q4.x, q4.y, q4.z, q4.w};
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint32_t t = 4 * g + r;  // lane q of spc lanes is x[spc*t' + q]
      v[pad(SPC == 2 ? ((t & 1) * M + (t >> 1)) : t)] = w4[r];
    }
  }
  __syncthreads();

  uint32_t fword = 0;
#pragma unroll
  for (uint32_t s = 0; s + 1 < NL; s += 2) {
    const uint32_t LQ = NL - s - 2;  // log2(quarter)
    const uint32_t Q = 1u << LQ;
    const uint32_t i = g & (Q - 1);
    const uint32_t base = ((g >> LQ) << (NL - s)) + i;
    uint32_t ad[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) ad[c] = Q >= 32 ? pad(base) + c * (Q + Q / 32) : pad(base + c * Q);
    if constexpr (ALL1) {
      uint32_t p0 = v[ad[0]], p1 = v[ad[1]], p2 = v[ad[2]], p3 = v[ad[3]];
      PkFlags f0, f1;
      bfly_pk(p0, p2, f0);  // BF2I, half = l/2
      bfly_pk(p1, p3, f0);
      p3 = negj_pk(p3, f1);  // BF2II: the odd sub-block takes -j
      bfly_pk(p0, p1, f1);
      bfly_pk(p2, p3, f1);
      if (s + 2 < NL) {
        const uint32_t TSH = LOG2N - (NL - s);
        const uint32_t t1 = pad((2 * i) << TSH), t2 = pad(i << TSH), t3 = pad((3 * i) << TSH);
        p0 = mul_w0_pk(p0);  // element 0 of every group takes W^0
        p1 = cmul_pk(p1, twc[t1], tws[t1], f1);
        p2 = cmul_pk(p2, twc[t2], tws[t2], f1);
        p3 = cmul_pk(p3, twc[t3], tws[t3], f1);
      }
      v[ad[0]] = p0;
      v[ad[1]] = p1;
      v[ad[2]] = p2;
      v[ad[3]] = p3;
      fword |= (f0.word() << (2 * s)) | (f1.word() << (2 * (s + 1)));
      __syncthreads();
      continue;
    }
    Cv e0 = ld_cv(v[ad[0]]), e1 = ld_cv(v[ad[1]]), e2 = ld_cv(v[ad[2]]), e3 = ld_cv(v[ad[3]]);
    StageFlags f0, f1;
    const bool sh0 = (sched >> s) & 1u, sh1 = (sched >> (s + 1)) & 1u;
    bfly<ALL1>(e0, e2, sh0, f0);  // BF2I, half = l/2
    bfly<ALL1>(e1, e3, sh0, f0);
    e3 = neg_j(e3, f1);           // BF2II: the odd sub-block takes -j
    bfly<ALL1>(e0, e1, sh1, f1);
    bfly<ALL1>(e2, e3, sh1, f1);
    if (s + 2 < NL) {
      const uint32_t TSH = LOG2N - (NL - s);
      // element c sits at quarter c of its block: exponent i * {0, 2, 1, 3}[c]
      const uint32_t t1 = pad((2 * i) << TSH), t2 = pad(i << TSH), t3 = pad((3 * i) << TSH);
      e0 = cmul_tw(e0, twc[0], tws[0], f1);  // W^0 = 0x7FFF is multiplied, never bypassed
      e1 = cmul_tw(e1, twc[t1], tws[t1], f1);
      e2 = cmul_tw(e2, twc[t2], tws[t2], f1);
      e3 = cmul_tw(e3, twc[t3], tws[t3], f1);
    }
    // each thread writes back exactly the four positions it read
    v[ad[0]] = pk(e0);
    v[ad[1]] = pk(e1);
    v[ad[2]] = pk(e2);
    v[ad[3]] = pk(e3);
    fword |= (f0.word() << (2 * s)) | (f1.word() << (2 * (s + 1)));
    __syncthreads();
  }

  // Final pass: the last (odd-count) sub-stage and/or the DIT merge.
  constexpr uint32_t SL = NL - 1;  // index of the last lane sub-stage
  constexpr bool LAST_R2 = (NL & 1u) != 0;
  if (ALL1 && SPC == 1) {
    if (LAST_R2) {
      uint32_t a = v[pad(4 * g)], b = v[pad(4 * g + 1)], c = v[pad(4 * g + 2)], d = v[pad(4 * g + 3)];
      PkFlags f;
      bfly_pk(a, b, f);
      bfly_pk(c, d, f);
      fword |= f.word() << (2 * SL);
      v[pad(4 * g)] = a;
      v[pad(4 * g + 1)] = b;
      v[pad(4 * g + 2)] = c;
      v[pad(4 * g + 3)] = d;
    }
  } else if (ALL1) {
    uint32_t a0 = v[pad(2 * g)], a1 = v[pad(2 * g + 1)], b0 = v[pad(M + 2 * g)], b1 = v[pad(M + 2 * g + 1)];
    if (LAST_R2) {
      PkFlags f;
      bfly_pk(a0, a1, f);
      bfly_pk(b0, b1, f);
      fword |= f.word() << (2 * SL);
    }
    PkFlags fm;
    const uint32_t m0 = pad(bitrev(2 * g, NL)), m1 = pad(bitrev(2 * g + 1, NL));
    uint32_t t0 = cmul_pk(b0, twc[m0], tws[m0], fm);
    uint32_t t1 = cmul_pk(b1, twc[m1], tws[m1], fm);
    bfly_pk(a0, t0, fm);
    bfly_pk(a1, t1, fm);
    fword |= fm.word() << (2 * NL);
    __syncthreads();
    v[pad(4 * g)] = a0;
    v[pad(4 * g + 1)] = t0;
    v[pad(4 * g + 2)] = a1;
    v[pad(4 * g + 3)] = t1;
  } else if (SPC == 1) {
    if (LAST_R2) {
      Cv a = ld_cv(v[pad(4 * g)]), b = ld_cv(v[pad(4 * g + 1)]), c = ld_cv(v[pad(4 * g + 2)]),
         d = ld_cv(v[pad(4 * g + 3)]);
      StageFlags f;
      const bool sh = (sched >> SL) & 1u;
      bfly<ALL1>(a, b, sh, f);
      bfly<ALL1>(c, d, sh, f);
      fword |= f.word() << (2 * SL);
      v[pad(4 * g)] = pk(a);
      v[pad(4 * g + 1)] = pk(b);
      v[pad(4 * g + 2)] = pk(c);
      v[pad(4 * g + 3)] = pk(d);
    }
  } else {
    // lane0[2g], lane0[2g+1], lane1[2g], lane1[2g+1]
    Cv a0 = ld_cv(v[pad(2 * g)]), a1 = ld_cv(v[pad(2 * g + 1)]), b0 = ld_cv(v[pad(M + 2 * g)]),
       b1 = ld_cv(v[pad(M + 2 * g + 1)]);
    if (LAST_R2) {
      StageFlags f;
      const bool sh = (sched >> SL) & 1u;
      bfly<ALL1>(a0, a1, sh, f);
      bfly<ALL1>(b0, b1, sh, f);
      fword |= f.word() << (2 * SL);
    }
    StageFlags fm;
    const bool shm = (sched >> NL) & 1u;
    const uint32_t m0 = pad(bitrev(2 * g, NL)), m1 = pad(bitrev(2 * g + 1, NL));
    Cv t0 = cmul_tw(b0, twc[m0], tws[m0], fm);
    Cv t1 = cmul_tw(b1, twc[m1], tws[m1], fm);
    bfly<ALL1>(a0, t0, shm, fm);
    bfly<ALL1>(a1, t1, shm, fm);
    fword |= fm.word() << (2 * NL);
    __syncthreads();
    // beat order: beats[2j] = lo, beats[2j+1] = hi
    v[pad(4 * g)] = pk(a0);
    v[pad(4 * g + 1)] = pk(t0);
    v[pad(4 * g + 2)] = pk(a1);
    v[pad(4 * g + 3)] = pk(t1);
  }
  if (fword) atomicOr(&sfl[lf], fword);
  __syncthreads();
  if (!active) return;

  // v holds the beats in beat order. Natural bin k: spc = 1 -> beats[bitrev(k)];
  // spc = 2 -> k = bitrev(j) + q*m reads beats[2j + q].
  const size_t fo = (size_t)frame * N;
  const BinsAt bo = bins_at(J, frame, N);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t k = g + r * TPF;
    if (bins) {
      const uint32_t src = SPC == 1 ? bitrev(k, NL) : 2 * bitrev(k & (M - 1), NL) + (k >> NL);
      bins[bo.at(k)] = v[pad(src)];
    }
    if (out) {
      uint32_t src = k;
      if (reorder) src = SPC == 1 ? bitrev(k, NL) : 2 * bitrev(k >> 1, NL) + (k & 1);
      out[fo + k] = v[pad(src)];
    }
  }
  if (J.stage_flags && g == 0) J.stage_flags[frame] = sfl[lf];
}

// Points per thread of a transform: 16 (radix-16 passes) for the all-ones
// schedule from 64 points up, else 4.
constexpr int fft_ppt(int log2n, bool all1) { return (all1 && log2n >= 6) ? 16 : 4; }

// One tile under the all-ones schedule with 16 points per thread: each LDS
// round trip runs four sub-stages (two radix-2^2 stages with their twiddles)
// on the thread's 16 points, element (q1, q2) of a block of length L at
// q1*L/4 + q2*L/16 + r. The first stage combines over q1 (twiddle exponent
// (q2*L/16 + r) * {0, 2, 1, 3}[q1]), the second over q2 inside each quarter
// (exponent r * {0, 2, 1, 3}[q2]); the arithmetic of every sub-stage is the
// radix-4 pass's. Leftover sub-stages run as one radix-4 pass over the
// thread's four groups, then the final pass as in fft_tile.
// tw4096: the twiddle table to stage into LDS here, after the frame loads are
// issued, so the two global latencies overlap (nullptr: already staged).
template <int LOG2N, int SPC, bool FL>
__device__ __forceinline__ void fft_tile16(const FftJob& J, uint32_t grp, uint32_t* smem,
                                           const uint32_t* __restrict__ tw4096) {
  using G = FftGeo<LOG2N, 16>;
  using PkFlags = PkFlagsT<FL>;
  constexpr uint32_t N = G::N;
  constexpr uint32_t M = N / SPC;
  constexpr uint32_t NL = LOG2N - (SPC == 2 ? 1 : 0);
  constexpr uint32_t TPF = G::TPF, FPB = G::FPB, NP = G::NP, NT = G::NT;
  const uint32_t* __restrict__ in = J.in;
  uint32_t* __restrict__ out = J.out;
  uint32_t* __restrict__ bins = J.bins;
  const uint32_t reorder = J.reorder;
  const uint32_t* twc = smem;
  const uint32_t* tws = smem + NT;
  uint32_t* sfl = smem + 2 * NT;
  const uint32_t lf = threadIdx.x / TPF;
  const uint32_t g = threadIdx.x % TPF;
  uint32_t* v = smem + 2 * NT + FPB + lf * NP;
  const uint32_t frame = grp * FPB + lf;
  const bool active = frame < J.n_frames;
  // the threads of a frame share one wave when TPF <= 64: their LDS exchanges
  // then need only a wave-level sync, and the block's waves run independently
  auto frame_sync = [] {
    if constexpr (TPF <= 64) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      __syncthreads();
    }
  };

  if (g == 0) sfl[lf] = 0;
  uint4 q4v[4];
  if (active) {
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t)  // four consecutive samples 4c .. 4c+3, c = g + t*TPF
      q4v[t] = *reinterpret_cast<const uint4*>(in + (size_t)frame * N + 4 * (g + t * TPF));
  }
  if (tw4096) fft_stage_twiddles<LOG2N, TPF * FPB>(tw4096, smem);
  if (active) {
    // Sample x = 4c + r goes to lane slot (x & 1) * M + x / 2 (spc 2) or x.
    // With c = g + t*TPF: 2c and 4c are even, so the small offsets stay in
    // their 32-word row; M and, for N >= 256, 2*TPF are multiples of 32.
    constexpr bool FOLD = N >= 256;
    constexpr uint32_t CS = SPC == 2 ? 2 : 4;  // slot stride of c
    const uint32_t pb = pad(CS * g);
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
      const uint32_t c = g + t * TPF;
      const uint4 q4 = q4v[t];
      const uint32_t o = FOLD ? pb + t * (CS * TPF + CS * TPF / 32) : pad(CS * c);
      if constexpr (SPC == 2) {
        v[o] = q4.x;
        v[o + M + M / 32] = q4.y;
        v[o + 1] = q4.z;
        v[o + 1 + M + M / 32] = q4.w;
      } else {
        v[o] = q4.x;
        v[o + 1] = q4.y;
        v[o + 2] = q4.z;
        v[o + 3] = q4.w;
      }
    }
  }
  __syncthreads();

  uint32_t fword = 0;
  constexpr uint32_t SR = (NL / 4) * 4;  // sub-stages of the radix-16 passes
#pragma unroll
  for (uint32_t s = 0; s + 3 < NL; s += 4) {
    const uint32_t LL = NL - s;             // log2 of the block length
    const uint32_t L4 = 1u << (LL - 2), L16 = 1u << (LL - 4);
    const uint32_t r = g & (L16 - 1);
    const uint32_t base = ((g >> (LL - 4)) << LL) + r;
    // padded addresses: a stride that is a multiple of 32 words adds a
    // constant (so the LDS accesses take immediate offsets from one base)
    uint32_t ad[16], p[16];
    if (L16 >= 32) {
      const uint32_t pb = pad(base);
#pragma unroll
      for (uint32_t e = 0; e < 16; ++e) ad[e] = pb + (e >> 2) * (L4 + L4 / 32) + (e & 3) * (L16 + L16 / 32);
    } else if (L4 >= 32) {
#pragma unroll
      for (uint32_t e = 0; e < 16; ++e) ad[e] = pad(base + (e & 3) * L16) + (e >> 2) * (L4 + L4 / 32);
    } else if (LL <= 5) {
      // the block (length <= 32, aligned) lies inside one 32-word row
      const uint32_t pb = pad(base);
#pragma unroll
      for (uint32_t e = 0; e < 16; ++e) ad[e] = pb + (e >> 2) * L4 + (e & 3) * L16;
    } else {
#pragma unroll
      for (uint32_t e = 0; e < 16; ++e) ad[e] = pad(base + (e >> 2) * L4 + (e & 3) * L16);
    }
#pragma unroll
    for (uint32_t e = 0; e < 16; ++e) p[e] = v[ad[e]];
    // sub-stages s, s+1 over q1 (index e = 4*q1 + q2); each sub-stage's flag
    // word is folded in as soon as the sub-stage is complete
    {
      PkFlags fa;
#pragma unroll
      for (uint32_t q2 = 0; q2 < 4; ++q2) {
        bfly_pk(p[q2], p[8 + q2], fa);
        bfly_pk(p[4 + q2], p[12 + q2], fa);
      }
      fword |= fa.word() << (2 * s);
    }
    PkFlags fb;
#pragma unroll
    for (uint32_t q2 = 0; q2 < 4; ++q2) {
      p[12 + q2] = negj_pk(p[12 + q2], fb);
      bfly_pk(p[q2], p[4 + q2], fb);
      bfly_pk(p[8 + q2], p[12 + q2], fb);
    }
    {
      // exponent k*i << TSH with i = q2*L16 + r is k*q2*N/16 + (k*r << TSH);
      // for N >= 512 the first term is a multiple of 32 and pads to a constant
      const uint32_t TSH = LOG2N - LL;
      constexpr uint32_t N16 = N / 16, P16 = N16 + N16 / 32;
      const uint32_t b1 = pad((2 * r) << TSH), b2 = pad(r << TSH), b3 = pad((3 * r) << TSH);
#pragma unroll
      for (uint32_t q2 = 0; q2 < 4; ++q2) {
        const uint32_t i = q2 * L16 + r;
        const bool fold = N16 % 32 == 0;
        const uint32_t t1 = fold ? b1 + 2 * q2 * P16 : pad((2 * i) << TSH);
        const uint32_t t2 = fold ? b2 + q2 * P16 : pad(i << TSH);
        const uint32_t t3 = fold ? b3 + 3 * q2 * P16 : pad((3 * i) << TSH);
        p[q2] = mul_w0_pk(p[q2]);
        p[4 + q2] = cmul_pk(p[4 + q2], twc[t1], tws[t1], fb);
        p[8 + q2] = cmul_pk(p[8 + q2], twc[t2], tws[t2], fb);
        p[12 + q2] = cmul_pk(p[12 + q2], twc[t3], tws[t3], fb);
      }
    }
    fword |= fb.word() << (2 * (s + 1));
    // sub-stages s+2, s+3 over q2 inside each quarter
    {
      PkFlags fc;
#pragma unroll
      for (uint32_t q1 = 0; q1 < 4; ++q1) {
        bfly_pk(p[4 * q1], p[4 * q1 + 2], fc);
        bfly_pk(p[4 * q1 + 1], p[4 * q1 + 3], fc);
      }
      fword |= fc.word() << (2 * (s + 2));
    }
    PkFlags fd;
#pragma unroll
    for (uint32_t q1 = 0; q1 < 4; ++q1) {
      uint32_t* b = p + 4 * q1;
      b[3] = negj_pk(b[3], fd);
      bfly_pk(b[0], b[1], fd);
      bfly_pk(b[2], b[3], fd);
    }
    if (s + 4 < NL) {
      const uint32_t TSH = LOG2N - LL + 2;
      const uint32_t t1 = pad((2 * r) << TSH), t2 = pad(r << TSH), t3 = pad((3 * r) << TSH);
      const uint32_t c1 = twc[t1], s1 = tws[t1], c2 = twc[t2], s2 = tws[t2], c3 = twc[t3], s3 = tws[t3];
#pragma unroll
      for (uint32_t q1 = 0; q1 < 4; ++q1) {
        uint32_t* b = p + 4 * q1;
        b[0] = mul_w0_pk(b[0]);
        b[1] = cmul_pk(b[1], c1, s1, fd);
        b[2] = cmul_pk(b[2], c2, s2, fd);
        b[3] = cmul_pk(b[3], c3, s3, fd);
      }
    }
#pragma unroll
    for (uint32_t e = 0; e < 16; ++e) v[ad[e]] = p[e];
    fword |= fd.word() << (2 * (s + 3));
    frame_sync();
  }

  if constexpr (SR + 1 < NL) {  // one radix-4 pass over the thread's four groups
    constexpr uint32_t s = SR;
    constexpr uint32_t LQ = NL - s - 2, Q = 1u << LQ;
    PkFlags f0, f1;
    uint32_t ad[4][4], p[4][4];
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
      const uint32_t gi = g + t * TPF;
      const uint32_t i = gi & (Q - 1);
      const uint32_t base = ((gi >> LQ) << (NL - s)) + i;
#pragma unroll
      for (uint32_t c = 0; c < 4; ++c) ad[t][c] = pad(base + c * Q);
#pragma unroll
      for (uint32_t c = 0; c < 4; ++c) p[t][c] = v[ad[t][c]];
      bfly_pk(p[t][0], p[t][2], f0);
      bfly_pk(p[t][1], p[t][3], f0);
      p[t][3] = negj_pk(p[t][3], f1);
      bfly_pk(p[t][0], p[t][1], f1);
      bfly_pk(p[t][2], p[t][3], f1);
      if (s + 2 < NL) {
        constexpr uint32_t TSH = LOG2N - (NL - s);
        const uint32_t t1 = pad((2 * i) << TSH), t2 = pad(i << TSH), t3 = pad((3 * i) << TSH);
        p[t][0] = mul_w0_pk(p[t][0]);
        p[t][1] = cmul_pk(p[t][1], twc[t1], tws[t1], f1);
        p[t][2] = cmul_pk(p[t][2], twc[t2], tws[t2], f1);
        p[t][3] = cmul_pk(p[t][3], twc[t3], tws[t3], f1);
      }
    }
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t)
#pragma unroll
      for (uint32_t c = 0; c < 4; ++c) v[ad[t][c]] = p[t][c];
    fword |= (f0.word() << (2 * s)) | (f1.word() << (2 * (s + 1)));
    frame_sync();
  }

  // final pass: the last (odd-count) sub-stage and/or the DIT merge, four
  // groups per thread
  constexpr uint32_t SL = NL - 1;
  constexpr bool LAST_R2 = (NL & 1u) != 0;
  if (SPC == 1) {
    if (LAST_R2) {
      PkFlags f;
#pragma unroll
      for (uint32_t t = 0; t < 4; ++t) {
        const uint32_t c = 4 * (g + t * TPF);
        uint32_t a = v[pad(c)], b = v[pad(c + 1)], cc = v[pad(c + 2)], d = v[pad(c + 3)];
        bfly_pk(a, b, f);
        bfly_pk(cc, d, f);
        v[pad(c)] = a;
        v[pad(c + 1)] = b;
        v[pad(c + 2)] = cc;
        v[pad(c + 3)] = d;
      }
      fword |= f.word() << (2 * SL);
    }
  } else {
    uint32_t a0[4], a1[4], t0[4], t1[4];
    PkFlags f, fm;
    // With j = g + t*TPF (2*TPF = 2^(NL-2)): 2j is even, so pad(2j+1) =
    // pad(2j) + 1; M and, for N >= 256, 2*TPF are multiples of 32; and
    // bitrev(2j) = bitrev(2g) + bitrev(t) (two bits) with bitrev(2g) a
    // multiple of 4, bitrev(2j+1) = bitrev(2j) + M/2 (a multiple of 32 for
    // N >= 128). Those cases take constant offsets from per-thread bases.
    constexpr bool FOLD = N >= 256;
    const uint32_t pa = pad(2 * g), pm = pad(bitrev(2 * g, NL));
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
      const uint32_t j = g + t * TPF;  // lane0[2j], lane0[2j+1], lane1[2j], lane1[2j+1]
      const uint32_t a = FOLD ? pa + t * (2 * TPF + 2 * TPF / 32) : pad(2 * j);
      a0[t] = v[a];
      a1[t] = v[a + 1];
      uint32_t b0 = v[a + M + M / 32], b1 = v[a + M + M / 32 + 1];
      if (LAST_R2) {
        bfly_pk(a0[t], a1[t], f);
        bfly_pk(b0, b1, f);
      }
      const uint32_t m0 = FOLD ? pm + ((t & 1) << 1 | t >> 1) : pad(bitrev(2 * j, NL));
      const uint32_t m1 = FOLD ? m0 + M / 2 + M / 64 : pad(bitrev(2 * j + 1, NL));
      t0[t] = cmul_pk(b0, twc[m0], tws[m0], fm);
      t1[t] = cmul_pk(b1, twc[m1], tws[m1], fm);
      bfly_pk(a0[t], t0[t], fm);
      bfly_pk(a1[t], t1[t], fm);
    }
    if (LAST_R2) fword |= f.word() << (2 * SL);
    fword |= fm.word() << (2 * NL);
    frame_sync();
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {  // beat order: beats[2j] = lo, beats[2j+1] = hi
      const uint32_t j = g + t * TPF;   // 4j % 32 <= 28: the four words share a row
      const uint32_t o = pad(4 * j);
      v[o] = a0[t];
      v[o + 1] = t0[t];
      v[o + 2] = a1[t];
      v[o + 3] = t1[t];
    }
  }
  if (fword) atomicOr(&sfl[lf], fword);
  const uint32_t fpc = J.bins_fpc;
  if (bins && !out && fpc && fpc % FPB == 0) {
    // Bin-quad blocked bins whose channels hold whole blocks of frames (so
    // every frame of the block is active and in one channel): after one block
    // barrier, consecutive threads take the same bin quad of consecutive
    // frames, so each run of FPB threads stores FPB * 16 contiguous bytes.
    __syncthreads();
    constexpr uint32_t Q = SPC == 1 ? N / 4 : M / 2;
    const uint32_t f0 = grp * FPB, ch = f0 / fpc;
    uint32_t* const base = bins + (size_t)ch * fpc * N + 4 * (f0 - ch * fpc);
#pragma unroll
    for (uint32_t it = threadIdx.x; it < N / 4 * FPB; it += TPF * FPB) {
      const uint32_t j = it / FPB, fl = it % FPB, k = 4 * j;
      const uint32_t b = SPC == 1 ? bitrev(k, NL) : 2 * bitrev(k & (M - 1), NL) + (k >> NL);
      *reinterpret_cast<uint4*>(base + (size_t)j * 4 * fpc + 4 * fl) =
          lds_quad<2 * Q, Q, 3 * Q>(smem + 2 * NT + FPB + fl * NP, b);
    }
    if (J.stage_flags && g == 0) J.stage_flags[frame] = sfl[lf];
    return;
  }
  frame_sync();
  if (!active) return;

  // Each thread stores four runs of four consecutive outputs k = 4c + i, one
  // 16-byte store each. The source of output k is a base plus a constant:
  //   bins, spc 1:         bitrev(k)              = bitrev(4c) + {0,2,1,3}[i] * N/4
  //   bins, spc 2:         2 bitrev(k % M) + k / M = its value at 4c + {0,2,1,3}[i] * M/2
  //   out, reorder, spc 1: as bins
  //   out, reorder, spc 2: 2 bitrev(k / 2) + k % 2 = 2 bitrev(2c) + {0, 1, M, M+1}[i]
  //   out, natural order:  4c + i
  const size_t fo = (size_t)frame * N;
  constexpr uint32_t Q = SPC == 1 ? N / 4 : M / 2;
  if (bins) {
    const BinsAt bo = bins_at(J, frame, N);
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
      const uint32_t k = 4 * (g + t * TPF);
      const uint32_t b = SPC == 1 ? bitrev(k, NL) : 2 * bitrev(k & (M - 1), NL) + (k >> NL);
      *reinterpret_cast<uint4*>(bins + bo.at(k)) = lds_quad<2 * Q, Q, 3 * Q>(v, b);
    }
  }
  if (out) {
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
      const uint32_t k = 4 * (g + t * TPF);
      uint4 o;
      if (!reorder)
        o = lds_quad<1, 2, 3>(v, k);
      else if (SPC == 1)
        o = lds_quad<2 * Q, Q, 3 * Q>(v, bitrev(k, NL));
      else
        o = lds_quad<1, M, M + 1>(v, 2 * bitrev(k >> 1, NL));
      *reinterpret_cast<uint4*>(out + fo + k) = o;
    }
  }
  if (J.stage_flags && g == 0) J.stage_flags[frame] = sfl[lf];
}

template <int LOG2N, bool ALL1>
using FftGeoOf = FftGeo<LOG2N, fft_ppt(LOG2N, ALL1)>;

// FL: the stage flags are formed (the packed all-ones path forms them only
// when a caller keeps them; the general path always does).
// The twiddle table is staged into LDS first, except by the radix-16 tile,
// which stages it after issuing its frame loads.
template <int LOG2N, int SPC, bool ALL1, bool FL = true>
__device__ __forceinline__ void fft_tile_any(const FftJob& J, uint32_t grp, uint32_t* smem,
                                             const uint32_t* __restrict__ tw4096) {
  if constexpr (fft_ppt(LOG2N, ALL1) == 16) {
    fft_tile16<LOG2N, SPC, FL>(J, grp, smem, tw4096);
  } else {
    fft_stage_twiddles<LOG2N, FftGeoOf<LOG2N, ALL1>::TPF * FftGeoOf<LOG2N, ALL1>::FPB>(tw4096, smem);
    fft_tile<LOG2N, SPC, ALL1, FL>(J, grp, smem);
  }
}

// Single CPI: one block per tile.
template <int LOG2N, int SPC, bool ALL1, bool FL>
__global__ __launch_bounds__((FftGeoOf<LOG2N, ALL1>::TPF * FftGeoOf<LOG2N, ALL1>::FPB)) void fft_kernel(
    FftJob J, const uint32_t* __restrict__ tw4096) {
  extern __shared__ uint32_t smem[];
  // blocked bins: consecutive blocks are dealt to the 8 XCDs in turn, so
  // adjacent frames (which share the output's cache lines) go to one XCD
  const uint32_t nb = gridDim.x;
  const uint32_t grp = J.bins_fpc && nb % 8 == 0 ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;
  fft_tile_any<LOG2N, SPC, ALL1, FL>(J, grp, smem, tw4096);
}

// Batched: one block per frame group of a class of CPIs (same length and lane
// count), block -> job via the tile table.
template <int LOG2N, int SPC, bool ALL1>
__global__ __launch_bounds__((FftGeoOf<LOG2N, ALL1>::TPF * FftGeoOf<LOG2N, ALL1>::FPB)) void fft_batch_kernel(
    const FftJob* __restrict__ jobs, const uint32_t* __restrict__ tile_job, uint32_t n_tiles,
    const uint32_t* __restrict__ tw4096) {
  extern __shared__ uint32_t smem[];
  const uint32_t t = blockIdx.x;
  if (t >= n_tiles) return;
  const FftJob& J = jobs[__builtin_amdgcn_readfirstlane(tile_job[t])];
  fft_tile_any<LOG2N, SPC, ALL1>(J, t - J.tile0, smem, tw4096);
}

template <int LOG2N, int SPC, bool ALL1, bool FL>
static void launch_fft_g(hipStream_t s, const uint32_t* tw4096, const FftJob& J) {
  using G = FftGeoOf<LOG2N, ALL1>;
  const dim3 grid((J.n_frames + G::FPB - 1) / G::FPB), block(G::TPF * G::FPB);
  fft_kernel<LOG2N, SPC, ALL1, FL><<<grid, block, G::lds_bytes, s>>>(J, tw4096);
}

template <int LOG2N, int SPC>
static hipError_t launch_fft_n(hipStream_t s, const uint32_t* tw4096, const FftJob& J) {
  constexpr uint32_t used = (1u << LOG2N) - 1u;
  if ((J.sched & used) != used)
    launch_fft_g<LOG2N, SPC, false, true>(s, tw4096, J);
  else if (J.stage_flags)
    launch_fft_g<LOG2N, SPC, true, true>(s, tw4096, J);
  else
    launch_fft_g<LOG2N, SPC, true, false>(s, tw4096, J);
  return hipGetLastError();
}

template <int LOG2N, int SPC, bool ALL1>
static void launch_fft_batch_g(hipStream_t s, const uint32_t* tw4096, const ClassLaunch& L) {
  using G = FftGeoOf<LOG2N, ALL1>;
  fft_batch_kernel<LOG2N, SPC, ALL1><<<L.grid, G::TPF * G::FPB, G::lds_bytes, s>>>(
      static_cast<const FftJob*>(L.jobs), L.tile_job, L.n_tiles, tw4096);
}

template <int LOG2N, int SPC>
static hipError_t launch_fft_batch_n(hipStream_t s, const uint32_t* tw4096, bool all1, const ClassLaunch& L) {
  if (all1)
    launch_fft_batch_g<LOG2N, SPC, true>(s, tw4096, L);
  else
    launch_fft_batch_g<LOG2N, SPC, false>(s, tw4096, L);
  return hipGetLastError();
}

// Dispatch of a run-time log2 length (3..12) to the compile-time kernels.
#define RD_FFT_CASES(F, ...)                      \
  switch (log2n) {                                \
    case 3: return F<3, SPC>(__VA_ARGS__);        \
    case 4: return F<4, SPC>(__VA_ARGS__);        \
    case 5: return F<5, SPC>(__VA_ARGS__);        \
    case 6: return F<6, SPC>(__VA_ARGS__);        \
    case 7: return F<7, SPC>(__VA_ARGS__);        \
    case 8: return F<8, SPC>(__VA_ARGS__);        \
    case 9: return F<9, SPC>(__VA_ARGS__);        \
    case 10: return F<10, SPC>(__VA_ARGS__);      \
    case 11: return F<11, SPC>(__VA_ARGS__);      \
    case 12: return F<12, SPC>(__VA_ARGS__);      \
  }                                               \
  return hipErrorInvalidValue;

template <int SPC>
static hipError_t launch_fft_spc(hipStream_t s, const uint32_t* tw, uint32_t log2n, const FftJob& J) {
  RD_FFT_CASES(launch_fft_n, s, tw, J)
}

template <int SPC>
static hipError_t launch_fft_batch_spc(hipStream_t s, const uint32_t* tw, uint32_t log2n, bool all1,
                                       const ClassLaunch& L) {
  RD_FFT_CASES(launch_fft_batch_n, s, tw, all1, L)
}

hipError_t launch_fft(hipStream_t s, const uint32_t* tw4096, const uint32_t* in, uint32_t n, uint32_t spc,
                      uint32_t sched, int reorder, uint32_t n_frames, uint32_t* out, uint32_t* bins,
                      uint32_t* stage_flags, uint32_t bins_fpc) {
  if (n_frames == 0) return hipSuccess;
  const uint32_t log2n = ilog2u(n);
  if (log2n < 3 || log2n > 12) return hipErrorInvalidValue;
  if (bins_fpc && (n % 4 != 0 || n_frames % bins_fpc != 0)) return hipErrorInvalidValue;
  const FftJob J{in, out, bins, stage_flags, sched, reorder ? 1u : 0u, n_frames, 0, bins_fpc};
  return spc == 2 ? launch_fft_spc<2>(s, tw4096, log2n, J) : launch_fft_spc<1>(s, tw4096, log2n, J);
}

uint32_t fft_frames_per_tile(uint32_t log2n, bool all1) {
  const uint32_t tpf = (1u << log2n) / (uint32_t)fft_ppt((int)log2n, all1);
  return tpf >= 256 ? 1 : 256 / tpf;
}

hipError_t launch_fft_batch(hipStream_t s, const uint32_t* tw4096, uint32_t log2n, uint32_t spc, bool all1,
                            const ClassLaunch& L) {
  if (L.n_tiles == 0) return hipSuccess;
  if (log2n < 3 || log2n > 12) return hipErrorInvalidValue;
  return spc == 2 ? launch_fft_batch_spc<2>(s, tw4096, log2n, all1, L)
                  : launch_fft_batch_spc<1>(s, tw4096, log2n, all1, L);
}

// ============================================================================
// Corner turn: X[ch][col] -> Xa[col][ch]. A block moves 64 columns of every
// channel: reads are CH rows of 256 contiguous bytes, and the output tile
// [64][CH] is one contiguous run.
// ============================================================================
__global__ __launch_bounds__(256) void align_kernel(const uint32_t* __restrict__ X, uint32_t n_chan, uint64_t cols,
                                                    uint32_t* __restrict__ Xa) {
  extern __shared__ uint32_t tile[];  // [n_chan][65]
  const uint64_t c0 = (uint64_t)blockIdx.x * 64;
  const uint32_t ncol = (uint32_t)min<uint64_t>(64, cols - c0);
  for (uint32_t idx = threadIdx.x; idx < n_chan * 64; idx += blockDim.x) {
    const uint32_t ch = idx / 64, j = idx % 64;
    if (j < ncol) tile[ch * 65 + j] = X[(size_t)ch * cols + c0 + j];
  }
  __syncthreads();
  for (uint32_t idx = threadIdx.x; idx < ncol * n_chan; idx += blockDim.x) {
    const uint32_t j = idx / n_chan, ch = idx % n_chan;
    Xa[(c0 + j) * n_chan + ch] = tile[ch * 65 + j];
  }
}

hipError_t launch_align(hipStream_t s, const uint32_t* X, uint32_t n_chan, uint64_t cols, uint32_t* Xa) {
  if (cols == 0) return hipSuccess;
  const dim3 grid((uint32_t)((cols + 63) / 64));
  align_kernel<<<grid, 256, (size_t)n_chan * 65 * sizeof(uint32_t), s>>>(X, n_chan, cols, Xa);
  return hipGetLastError();
}

// ============================================================================
// Slow-time (Doppler) transform across frames. The frames of a row sit in a
// ring of `depth` slots (frame f in slot f mod depth, as hist::slot_of,
// model/cpp/history/history_model.hpp:127-129); the window is the last H
// frames, the newest in slot `newest`. Each (row, bin) column of the window is
// gathered into contiguous memory, transformed by the bit-exact FFT above
// along the frame axis, and turned back to [row][doppler][bin], optionally
// with the zero-Doppler bin moved to the centre (fftshift,
// model/python/range_doppler.py:254-265).
// ============================================================================
// in [R][depth][N] -> out [R][N][H]: out[r][k][h] = in[r][(first + h) % depth][k]
__global__ __launch_bounds__(256) void ring_to_cols_kernel(const uint32_t* __restrict__ in, uint32_t depth,
                                                           uint32_t N, uint32_t H, uint32_t first,
                                                           uint32_t* __restrict__ out) {
  __shared__ uint32_t tile[64][65];
  const uint32_t k0 = blockIdx.x * 64, h0 = blockIdx.y * 64, r = blockIdx.z;
  for (uint32_t idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const uint32_t hh = idx / 64, kk = idx % 64, h = h0 + hh, k = k0 + kk;
    if (h < H && k < N) tile[hh][kk] = in[((size_t)r * depth + (first + h) % depth) * N + k];
  }
  __syncthreads();
  for (uint32_t idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const uint32_t kk = idx / 64, hh = idx % 64, h = h0 + hh, k = k0 + kk;
    if (h < H && k < N) out[((size_t)r * N + k) * H + h] = tile[hh][kk];
  }
}

// in [R][N][H] -> out [R][H][N], doppler row h' = shift ? (h + H/2) % H : h
__global__ __launch_bounds__(256) void cols_to_rows_kernel(const uint32_t* __restrict__ in, uint32_t N, uint32_t H,
                                                           int shift, uint32_t* __restrict__ out) {
  __shared__ uint32_t tile[64][65];
  const uint32_t k0 = blockIdx.x * 64, h0 = blockIdx.y * 64, r = blockIdx.z;
  for (uint32_t idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const uint32_t kk = idx / 64, hh = idx % 64, h = h0 + hh, k = k0 + kk;
    if (h < H && k < N) tile[kk][hh] = in[((size_t)r * N + k) * H + h];
  }
  __syncthreads();
  for (uint32_t idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const uint32_t hh = idx / 64, kk = idx % 64, h = h0 + hh, k = k0 + kk;
    if (h < H && k < N) {
      const uint32_t hd = shift ? (h + H / 2) % H : h;
      out[((size_t)r * H + hd) * N + k] = tile[kk][hh];
    }
  }
}

size_t doppler_scratch_bytes(uint32_t rows, uint32_t n_bins, uint32_t H) {
  return 2 * (size_t)rows * n_bins * H * sizeof(uint32_t) + 256;
}

hipError_t launch_doppler(hipStream_t s, const uint32_t* tw4096, const uint32_t* ring, uint32_t rows, uint32_t depth,
                          uint32_t n_bins, uint32_t newest, uint32_t H, uint32_t spc, uint32_t sched, int shift,
                          uint32_t* out, uint32_t* stage_flags, void* scratch) {
  if (rows == 0 || n_bins == 0) return hipSuccess;
  uint32_t* cols = static_cast<uint32_t*>(scratch);
  uint32_t* spec = cols + (size_t)rows * n_bins * H;
  const uint32_t first = (newest + depth - (H - 1) % depth) % depth;
  const dim3 grid((n_bins + 63) / 64, (H + 63) / 64, rows);
  ring_to_cols_kernel<<<grid, 256, 0, s>>>(ring, depth, n_bins, H, first, cols);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_fft(s, tw4096, cols, H, spc, sched, 1, rows * n_bins, nullptr, spec, stage_flags);
  if (e != hipSuccess) return e;
  cols_to_rows_kernel<<<grid, 256, 0, s>>>(spec, n_bins, H, shift, out);
  return hipGetLastError();
}

}  // namespace rd
