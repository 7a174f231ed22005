// rd_exact.hip — the reference's stateful and per-element model calls as
// data-parallel kernels, the exact values the C++ mirror's scalar APIs return:
//   * covar::Integrator over a shared control stream (rd_integ_run,
//     rd_cov_integrate_windows): windows, flush, sat_clear, either mode;
//   * pfb::LaneModel::accumulate, the exact PFB accumulators (rd_pfb_acc_run);
//   * cfar::decide, the per-bin decision with its partial window sums
//     (rd_cfar_decide).
//
// Integrator. Every lane (one accumulator) of a call sees the same control
// inputs per step: valid_in, flush, sat_clear and the cfg ports. Whether a
// step accepts, closes or emits a window, the window id, sample count and the
// configuration that latches at each boundary depend only on those inputs
// (covar_model.hpp:196-262), so the host plans them once per call into one
// control word per step and the kernel applies the data path to every lane:
//   raw  = accept ? (exp ? acc + floor((x - acc) / 2^k) : acc + x) : acc   (wrapping)
//   next = sat(raw, acc_w), flags of raw when accepted
//   emit: result = next;  sticky: clear on sat_clear or flush, else merge
//   flush: acc = 0;  close: acc = exp ? next : 0;  else acc = next
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "rd_fxp.h"
#include "rd_internal.h"

namespace rd {

// Control word of one step: bit 0 accept, bit 1 exponential mode, bits 2-7
// the shift (clamped to 63: floor(v / 2^s) is the sign for every s >= 63),
// bit 8 close, bit 9 flush, bit 10 sticky clear, bit 11 emit; .y = emit slot.
constexpr uint32_t kIgAccept = 1u, kIgExp = 2u, kIgClose = 1u << 8, kIgFlush = 1u << 9, kIgClear = 1u << 10,
                   kIgEmit = 1u << 11;

uint32_t integ_plan(rd_integ_ctl& c, const rd_integ_step* steps, uint32_t n, uint2* words, rd_integ_window* windows,
                    uint32_t cap) {
  uint32_t nw = 0;
  for (uint32_t s = 0; s < n; ++s) {
    const rd_integ_step& st = steps[s];
    const uint32_t wl = c.active.window_len == 0 ? 1u : c.active.window_len;
    const bool accept = st.valid && c.active.enable;
    const uint32_t count_next = accept ? c.count + 1u : c.count;
    const bool close = accept && count_next >= wl;
    const bool flush_emit = st.flush && !close && count_next != 0u;
    const bool emit = close || flush_emit;
    const bool exp = c.active.mode == 1u;
    uint32_t w = (accept ? kIgAccept : 0u) | (exp ? kIgExp : 0u) | ((c.active.exp_k > 63u ? 63u : c.active.exp_k) << 2) |
                 (close ? kIgClose : 0u) | (st.flush ? kIgFlush : 0u) | ((st.sat_clear || st.flush) ? kIgClear : 0u) |
                 (emit ? kIgEmit : 0u);
    words[s] = make_uint2(w, nw);
    if (emit) {
      if (nw < cap && windows)
        windows[nw] = rd_integ_window{s, c.window_id, count_next, st.flush ? 1u : 0u, count_next != wl ? 1u : 0u};
      ++nw;
    }
    const bool boundary = close || st.flush || (c.count == 0u && !accept);
    if (st.flush) {
      c.count = 0;
      c.window_id = 0;
    } else if (close) {
      c.count = 0;
      c.window_id = (c.window_id + 1u) & 0xFFFFu;
    } else {
      c.count = count_next;
    }
    if (boundary) c.active = st.cfg;
  }
  return nw;
}

// One lane's data path for one step (control word w).
struct IgLane {
  int64_t acc;
  uint32_t flags, count;
};

__device__ __forceinline__ int64_t ig_sat(int64_t v, int64_t hi, int64_t lo, uint32_t& f) {
  f = (v > hi ? 2u : 0u) | (v < lo ? 1u : 0u);
  return v > hi ? hi : (v < lo ? lo : v);
}

__device__ __forceinline__ void ig_step(IgLane& L, uint2 c, int64_t x, int64_t hi, int64_t lo, int64_t* out) {
  const uint32_t w = c.x;
  int64_t raw = L.acc;
  if (w & kIgAccept) {
    if (w & kIgExp)
      raw = (int64_t)((uint64_t)L.acc + (uint64_t)((int64_t)((uint64_t)x - (uint64_t)L.acc) >> ((w >> 2) & 63u)));
    else
      raw = (int64_t)((uint64_t)L.acc + (uint64_t)x);
  }
  uint32_t f = 0;
  const int64_t next = ig_sat(raw, hi, lo, f);
  if (!(w & kIgAccept)) f = 0;
  if ((w & kIgEmit) && out) *out = next;
  if (w & kIgClear) {
    L.flags = 0;
    L.count = 0;
  } else if (w & kIgAccept) {
    L.flags |= f;
    if (f && L.count != 0xFFFFFFFFu) ++L.count;
  }
  if (w & kIgFlush)
    L.acc = 0;
  else if (w & kIgClose)
    L.acc = (w & kIgExp) ? next : 0;
  else
    L.acc = next;
}

__host__ __device__ inline void ig_bounds(uint32_t acc_w, int64_t& hi, int64_t& lo) {
  if (acc_w >= 64) {
    hi = INT64_MAX;
    lo = INT64_MIN;
  } else if (acc_w < 2) {
    hi = lo = 0;
  } else {
    hi = ((int64_t)1 << (acc_w - 1)) - 1;
    lo = -hi - 1;
  }
}

// Generic lanes: data [n_steps][n_lanes], results [cap][n_lanes].
__global__ __launch_bounds__(256) void integ_kernel(const uint2* __restrict__ ctl, uint32_t n_steps,
                                                    const int64_t* __restrict__ data, uint64_t n_lanes,
                                                    rd_integ_lane* __restrict__ lanes, int64_t* __restrict__ results,
                                                    uint32_t cap, uint32_t acc_w) {
  const uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_lanes) return;
  int64_t hi, lo;
  ig_bounds(acc_w, hi, lo);
  IgLane L{lanes[l].acc, lanes[l].flags, lanes[l].sat_count};
  for (uint32_t s = 0; s < n_steps; ++s) {
    const uint2 c = ctl[s];
    int64_t* out = (c.x & kIgEmit) && c.y < cap ? results + (size_t)c.y * n_lanes + l : nullptr;
    ig_step(L, c, data[(size_t)s * n_lanes + l], hi, lo, out);
  }
  lanes[l] = rd_integ_lane{L.acc, L.flags, L.count};
}

// Covariance lanes: thread = (bin, pair), two lanes (re, im) of
// cross_power(X_i, X_j) per frame; lanes/results [..][bin][pair][2].
__global__ __launch_bounds__(256) void cov_integ_kernel(const uint2* __restrict__ ctl, const uint32_t* __restrict__ Xa,
                                                        uint32_t CH, uint32_t F, uint32_t N,
                                                        rd_integ_lane* __restrict__ lanes,
                                                        int64_t* __restrict__ results, uint32_t cap, uint32_t acc_w) {
  const uint32_t npairs = CH * (CH + 1) / 2;
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (uint64_t)N * npairs) return;
  const uint32_t k = (uint32_t)(g / npairs);
  uint32_t p = (uint32_t)(g % npairs), i = 0;
  while (p >= CH - i) {
    p -= CH - i;
    ++i;
  }
  const uint32_t j = i + p;
  int64_t hi, lo;
  ig_bounds(acc_w, hi, lo);
  IgLane A{lanes[2 * g].acc, lanes[2 * g].flags, lanes[2 * g].sat_count};
  IgLane B{lanes[2 * g + 1].acc, lanes[2 * g + 1].flags, lanes[2 * g + 1].sat_count};
  const uint64_t lanes_all = 2ull * N * npairs;
  for (uint32_t m = 0; m < F; ++m) {
    const uint2 c = ctl[m];
    const uint32_t* row = Xa + ((size_t)m * N + k) * CH;
    const uint32_t x = row[i], y = row[j];
    const int64_t re = (int64_t)cre(x) * cre(y) + (int64_t)cim(x) * cim(y);
    const int64_t im = (int64_t)cim(x) * cre(y) - (int64_t)cre(x) * cim(y);
    int64_t* o = (c.x & kIgEmit) && c.y < cap ? results + (size_t)c.y * lanes_all + 2 * g : nullptr;
    ig_step(A, c, re, hi, lo, o);
    ig_step(B, c, im, hi, lo, o ? o + 1 : nullptr);
  }
  lanes[2 * g] = rd_integ_lane{A.acc, A.flags, A.count};
  lanes[2 * g + 1] = rd_integ_lane{B.acc, B.flags, B.count};
}

hipError_t launch_integ(hipStream_t s, const uint2* ctl, uint32_t n_steps, const int64_t* data, uint64_t n_lanes,
                        rd_integ_lane* lanes, int64_t* results, uint32_t cap, uint32_t acc_w) {
  if (n_lanes == 0) return hipSuccess;
  integ_kernel<<<(uint32_t)((n_lanes + 255) / 256), 256, 0, s>>>(ctl, n_steps, data, n_lanes, lanes, results, cap,
                                                                  acc_w);
  return hipGetLastError();
}

hipError_t launch_cov_integ(hipStream_t s, const uint2* ctl, const uint32_t* Xa, uint32_t CH, uint32_t F, uint32_t N,
                            rd_integ_lane* lanes, int64_t* results, uint32_t cap, uint32_t acc_w) {
  const uint64_t total = (uint64_t)N * CH * (CH + 1) / 2;
  if (total == 0) return hipSuccess;
  cov_integ_kernel<<<(uint32_t)((total + 255) / 256), 256, 0, s>>>(ctl, Xa, CH, F, N, lanes, results, cap, acc_w);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// pfb::LaneModel::accumulate (pfb_model.hpp:135-150): per output sample the
// exact sums sum_k cmul_q15_re_raw(x[m-k], h[k]) and the imaginary twin, in
// int64, with the history seeding x[m-k] for m-k < 0. One thread per
// (channel, beat, branch).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pfb_acc_kernel(const uint32_t* __restrict__ x, const uint32_t* __restrict__ h,
                                                      uint32_t P, uint32_t T, uint32_t n_chan, uint32_t n_beats,
                                                      const uint32_t* __restrict__ hist, int64_t* __restrict__ acc) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t per_ch = (uint64_t)n_beats * P;
  if (g >= per_ch * n_chan) return;
  const uint32_t c = (uint32_t)(g / per_ch);
  const uint32_t m = (uint32_t)((g % per_ch) / P), p = (uint32_t)(g % P);
  int64_t re = 0, im = 0;
  for (uint32_t k = 0; k < T; ++k) {
    uint32_t s = 0;
    if (k <= m)
      s = x[(size_t)c * per_ch + (size_t)(m - k) * P + p];
    else if (hist)
      s = hist[((size_t)c * (T - 1) + (k - m - 1)) * P + p];
    const uint32_t w = h[(size_t)p * T + k];
    re += (int64_t)cre(s) * cre(w) - (int64_t)cim(s) * cim(w);
    im += (int64_t)cre(s) * cim(w) + (int64_t)cim(s) * cre(w);
  }
  acc[2 * g] = re;
  acc[2 * g + 1] = im;
}

hipError_t launch_pfb_acc(hipStream_t s, const uint32_t* x, const uint32_t* coeff, uint32_t P, uint32_t T,
                          uint32_t n_chan, uint32_t n_beats, const uint32_t* hist, int64_t* acc) {
  const uint64_t total = (uint64_t)n_chan * n_beats * P;
  if (total == 0) return hipSuccess;
  pfb_acc_kernel<<<(uint32_t)((total + 255) / 256), 256, 0, s>>>(x, coeff, P, T, n_chan, n_beats, hist, acc);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// cfar::decide (cfar_model.hpp:243-285) for every bin of every row, with the
// configuration exactly as given (no clamp; run_frame clamps before it calls
// decide). The lead cells are summed upward from bin + guard_lead + 1 and stop
// at the first one past the row, the lag cells downward from bin -
// guard_lag - 1 and stop below bin 0, only when the lead window fit: the
// partial sums the reference leaves in the Decision. One thread per bin.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void cfar_decide_kernel(const uint64_t* __restrict__ cells, uint32_t n,
                                                          uint32_t n_rows, rd_cfar_cfg cfg,
                                                          rd_cfar_decision* __restrict__ out) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (uint64_t)n * n_rows) return;
  const uint64_t* row = cells + (g / n) * n;
  const int64_t idx = (int64_t)(g % n), nn = (int64_t)n;
  rd_cfar_decision d{};
  d.cut_power = row[idx];
  bool ok = true;
  for (uint32_t j = 1; j <= cfg.ref_lead; ++j) {
    const int64_t k = idx + (int64_t)(uint32_t)(cfg.guard_lead + j);
    if (k >= nn) {
      ok = false;
      break;
    }
    d.sum_lead += row[k];
  }
  if (ok)
    for (uint32_t j = 1; j <= cfg.ref_lag; ++j) {
      const int64_t k = idx - (int64_t)(uint32_t)(cfg.guard_lag + j);
      if (k < 0) {
        ok = false;
        break;
      }
      d.sum_lag += row[k];
    }
  d.window_ok = ok;
  const bool go = cfg.mode == 1;
  if (go) {
    d.take_lead = d.sum_lead * (uint64_t)cfg.ref_lag >= d.sum_lag * (uint64_t)cfg.ref_lead;
    d.sel_sum = d.take_lead ? d.sum_lead : d.sum_lag;
    d.sel_n = d.take_lead ? cfg.ref_lead : cfg.ref_lag;
  } else {
    d.take_lead = 1;
    d.sel_sum = d.sum_lead + d.sum_lag;
    d.sel_n = cfg.ref_lead + cfg.ref_lag;
  }
  const bool refs = go ? (cfg.ref_lead != 0 && cfg.ref_lag != 0) : (cfg.ref_lead + cfg.ref_lag) != 0;
  d.evaluable = ok && cfg.enable && refs;
  d.detected = d.evaluable && ((d.cut_power * (uint64_t)d.sel_n) << 8) > (uint64_t)cfg.alpha * d.sel_sum;
  d.suppressed = !d.evaluable;
  out[g] = d;
}

hipError_t launch_cfar_decide(hipStream_t s, const uint64_t* cells, uint32_t n, uint32_t n_rows,
                              const rd_cfar_cfg& cfg, rd_cfar_decision* out) {
  const uint64_t total = (uint64_t)n * n_rows;
  if (total == 0) return hipSuccess;
  cfar_decide_kernel<<<(uint32_t)((total + 255) / 256), 256, 0, s>>>(cells, n, n_rows, cfg, out);
  return hipGetLastError();
}

}  // namespace rd
