// rd_cfar.hip — CA/GO-CFAR with ordered event compaction and 176-bit packing.
//
// Semantics of cfar::run_frame (model/cpp/cfar/cfar_model.hpp:243-340):
//   * geometry clamped to the elaboration maxima (cfar_model.hpp:134-141);
//   * lead reference cells at higher bin indices, lag at lower, guards skipped;
//     a bin whose full window does not fit is suppressed;
//   * CA: S = lead + lag, N = ref_lead + ref_lag; GO: the greater mean by one
//     cross-multiply, ties to lead;
//   * detect iff (C * N) << 8 > alpha * S in u64 (strict);
//   * events in bin order, then one SUMMARY{det, sup, bin = n}; sof on the first
//     event, eof on the summary.
//
// The window sums come from an exclusive prefix sum of the row in LDS, so every
// bin costs two differences whatever the reference length; u64 wrap-around of
// the prefix differences equals the reference's u64 running sums modulo 2^64.
//
// Per-row decisions with the events staged in a fixed slot per row, a scan of
// the per-row counts into event offsets, and a gather that moves each row's
// events to its offset and recomputes only rows that overflowed their slot
// (for a single CPI with rows of several waves, two launches: a copy without
// LDS, then the recomputation of the overflowed rows). Batched launches run the same row bodies over the
// rows of many CPIs and one scan block per CPI.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "rd_internal.h"

namespace rd {

// f(integral_constant<0>), ..., f(integral_constant<N-1>): an unrolled loop
// whose index is a compile-time constant in the body.
template <typename F, uint32_t... I>
__device__ __forceinline__ void static_for_seq(F&& f, std::integer_sequence<uint32_t, I...>) {
  (f(std::integral_constant<uint32_t, I>{}), ...);
}
template <uint32_t N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_seq(f, std::make_integer_sequence<uint32_t, N>{});
}

// Events staged per row by the first pass of a single-CPI launch (summary
// included); rows with more are recomputed by the gather pass.
constexpr uint32_t kCfarSlot = 256;

CfarGeom cfar_geom(const rd_cfar_cfg& c, uint32_t max_guard, uint32_t max_ref) {
  CfarGeom g;
  g.enable = c.enable != 0;
  g.go = c.mode == 1;
  g.dense = c.out_mode == 1;
  g.gl = c.guard_lead < max_guard ? c.guard_lead : max_guard;
  g.gg = c.guard_lag < max_guard ? c.guard_lag : max_guard;
  g.rl = c.ref_lead < max_ref ? c.ref_lead : max_ref;
  g.rg = c.ref_lag < max_ref ? c.ref_lag : max_ref;
  g.alpha = c.alpha;
  g.refs_ok = g.go ? (g.rl != 0 && g.rg != 0) : (g.rl + g.rg) != 0;
  return g;
}

// ----------------------------------------------------------------------------
// Row geometry. A row of n bins is handled by W waves; a block holds R rows
// (R*W <= 4 waves, or one row of 8 waves). Wave w of a row owns the bins
// [w*64*S, (w+1)*64*S), S = 2^SLOG <= 16 a compile-time constant:
//   prefix pass:   lane l owns the S contiguous bins starting at (w*64 + l)*S:
//                  its cells go straight into registers, a sequential sum per
//                  lane and ONE wave scan per row (plus a carry across the
//                  row's waves) give the exclusive prefix P[e], e = 0..n;
//   decision pass: the wave's S chunks of 64 bins, lane l bin l of each chunk,
//                  so LDS reads and event stores are consecutive across lanes.
// P sits in LDS at pidx(e) = e + (e >> SLOG) (SLOG >= 1): one padding word per
// S entries keeps the strided segment writes of the prefix pass free of bank
// conflicts, and puts a lane's S entries at fixed offsets from one base.
// LDS per row (u64 words): P, then the per-wave statistics [W][4] = {prefix
// total, events, detections, -}, then per wave a list of up to 64 of its
// emitting bins (u32 entries; a wave with more emits chunk by chunk).
// ----------------------------------------------------------------------------
struct CfarRowGeo {
  uint32_t W, slog, words;
};

// Fewest waves per row: a lane's prefix segment holds at most 16 bins.
uint32_t cfar_min_waves(uint32_t n) {
  uint32_t W = 1;
  while (W < 8 && W * 1024u < n) W <<= 1;
  return W;
}

// W: waves per row, a power of two from cfar_min_waves(n) to 8.
__host__ __device__ inline CfarRowGeo cfar_row_geo(uint32_t n, uint32_t W) {
  CfarRowGeo G;
  G.W = W;
  const uint32_t per = (n + 64 * W - 1) / (64 * W);
  G.slog = 0;
  while ((1u << G.slog) < per) ++G.slog;
  const uint32_t pmax = n + (G.slog ? (n >> G.slog) : 0);
  G.words = pmax + 1 + W * 4 + W * 32;
  return G;
}

template <uint32_t SLOG>
__device__ __forceinline__ uint32_t pidx(uint32_t e) {
  return SLOG ? e + (e >> SLOG) : e;
}

// One step of the scan below: x += (x moved between lanes by the DPP control
// CTRL, in the rows ROWS, a 4-bit mask of the wave's rows of 16 lanes), the
// 64-bit value kept as two 32-bit halves added with a carry; lanes outside
// ROWS, and lanes whose source lies outside their row, add 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ void dpp_add(uint32_t& lo, uint32_t& hi) {
  const uint32_t dl = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, CTRL, ROWS, 0xF, true);
  const uint32_t dh = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, CTRL, ROWS, 0xF, true);
  const uint32_t s = lo + dl;
  hi = hi + dh + (s < lo ? 1u : 0u);
  lo = s;
}

// Inclusive scan over the 64 lanes of a wave, every lane active: within each
// row of 16 lanes by row shifts of 1, 2, 4 and 8, then the last lane of rows 0
// and 2 added to rows 1 and 3 (row_bcast:15) and lane 31 to rows 2 and 3
// (row_bcast:31). No LDS traffic.
__device__ __forceinline__ uint64_t wave_incscan(uint64_t x) {
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  dpp_add<0x111, 0xF>(lo, hi);
  dpp_add<0x112, 0xF>(lo, hi);
  dpp_add<0x114, 0xF>(lo, hi);
  dpp_add<0x118, 0xF>(lo, hi);
  dpp_add<0x142, 0xA>(lo, hi);
  dpp_add<0x143, 0xC>(lo, hi);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

// The value of lane 63 (uniform).
__device__ __forceinline__ uint64_t lane63(uint64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), 63);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

// LDS writes of one lane become visible to the other lanes of the wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Synchronises the waves of a row: a block barrier when a row spans several
// waves (MULTI; every thread of the block executes the same sequence), else a
// wave barrier.
template <bool MULTI>
__device__ __forceinline__ void row_sync() {
  if (MULTI)
    __syncthreads();
  else
    wave_lds_sync();
}

struct Decision {
  bool hit, eval;
  uint64_t cut, sum;
  uint32_t n;
};

// Branch-free so the six prefix reads issue back to back: window indices are
// clamped into [0, n] and the sums of windows that do not fit are masked.
template <uint32_t SLOG>
__device__ __forceinline__ Decision decide(const uint64_t* P, uint32_t n, uint32_t i, const CfarGeom& g) {
  const uint32_t a1 = min(i + g.gl + g.rl + 1, n), a0 = min(i + g.gl + 1, n);
  const uint32_t b1 = i >= g.gg ? i - g.gg : 0u, b0 = i >= g.gg + g.rg ? i - g.gg - g.rg : 0u;
  const uint64_t pc1 = P[pidx<SLOG>(i + 1)], pc0 = P[pidx<SLOG>(i)];
  const uint64_t pa1 = P[pidx<SLOG>(a1)], pa0 = P[pidx<SLOG>(a0)];
  const uint64_t pb1 = P[pidx<SLOG>(b1)], pb0 = P[pidx<SLOG>(b0)];
  const bool lead_ok = g.rl == 0 || (uint64_t)i + g.gl + g.rl < n;
  const bool lag_ok = g.rg == 0 || i >= g.gg + g.rg;
  const bool ok = lead_ok && lag_ok;
  const uint64_t slead = (ok && g.rl) ? pa1 - pa0 : 0ull;
  const uint64_t slag = (ok && g.rg) ? pb1 - pb0 : 0ull;
  Decision d;
  d.cut = pc1 - pc0;
  if (g.go) {
    const bool lead = slead * g.rg >= slag * g.rl;
    d.sum = lead ? slead : slag;
    d.n = lead ? g.rl : g.rg;
  } else {
    d.sum = slead + slag;
    d.n = g.rl + g.rg;
  }
  d.eval = ok && g.enable && g.refs_ok;
  d.hit = d.eval && ((d.cut * (uint64_t)d.n) << 8) > (uint64_t)g.alpha * d.sum;
  return d;
}

// The interior CA form reads the prefix at six fixed displacements d from the
// bin. For bin i = c + lane, c a multiple of 64 (so of S),
// pidx(i + d) = pidx(c) + pidx(lane + d) (arithmetic shift, d may be
// negative): the lane term of each displacement is computed once per wave and
// the chunk's term is uniform, a compile-time step per chunk of the wave.
template <uint32_t SLOG>
__device__ __forceinline__ uint32_t lane_run(uint32_t lane, int d) {
  const int e = (int)lane + d;
  return (uint32_t)(SLOG ? e + (e >> SLOG) : e);
}

// LDS addresses of the six runs for one lane at the wave's first chunk.
struct CaRuns {
  const uint64_t *c1, *c0, *a1, *a0, *b1, *b0;  // i+1, i, lead end/start, lag end/start
};

// CA decision of a bin whose whole window lies inside the row (the chunk test
// in the caller guarantees it): no clamping or masking, only the hit. OFF is
// the chunk's step from the wave's first chunk; n8 = (ref_lead + ref_lag) << 8.
template <typename Cell, uint32_t OFF>
__device__ __forceinline__ bool ca_hit_inner(const CaRuns& r, uint32_t alpha, uint32_t n8) {
  uint64_t lhs;
  if constexpr (sizeof(Cell) == 4) {
    // one u32 cell: the low words of its two prefix entries differ by it exactly
    const uint32_t cut = (uint32_t)r.c1[OFF] - (uint32_t)r.c0[OFF];
    lhs = (uint64_t)cut * n8;
  } else {
    lhs = (r.c1[OFF] - r.c0[OFF]) * (uint64_t)n8;
  }
  const uint64_t sum = (r.a1[OFF] - r.a0[OFF]) + (r.b1[OFF] - r.b0[OFF]);
  return lhs > (uint64_t)alpha * sum;
}

// CA decision (no dense output) of a bin in an edge chunk: the window indices
// clamped into [0, n] so every read stays in the row's prefix, and a bin whose
// window does not fit never detects (decide() gives it eval = false). For a
// bin that fits, the clamps are no-ops and the sums are the unclamped ones;
// with no lead (lag) reference cells the clamped pair reads one entry twice,
// a zero sum, as decide() has it. n8 = (ref_lead + ref_lag) << 8.
template <typename Cell, uint32_t SLOG>
__device__ __forceinline__ bool ca_hit_edge(const uint64_t* P, uint32_t n, uint32_t i, const CfarGeom& g,
                                            uint32_t n8) {
  const uint32_t a1 = min(i + g.gl + g.rl + 1, n), a0 = min(i + g.gl + 1, n);
  const uint32_t b1 = i >= g.gg ? i - g.gg : 0u, b0 = i >= g.gg + g.rg ? i - g.gg - g.rg : 0u;
  const bool fit = (g.rl == 0 || i + g.gl + g.rl < n) && (g.rg == 0 || i >= g.gg + g.rg);
  uint64_t lhs;
  if constexpr (sizeof(Cell) == 4) {
    const uint32_t cut = (uint32_t)P[pidx<SLOG>(i + 1)] - (uint32_t)P[pidx<SLOG>(i)];
    lhs = (uint64_t)cut * n8;
  } else {
    lhs = (P[pidx<SLOG>(i + 1)] - P[pidx<SLOG>(i)]) * (uint64_t)n8;
  }
  const uint64_t sum = (P[pidx<SLOG>(a1)] - P[pidx<SLOG>(a0)]) + (P[pidx<SLOG>(b1)] - P[pidx<SLOG>(b0)]);
  return fit && lhs > (uint64_t)g.alpha * sum;
}

// Bins of a row of n that are not evaluated (window does not fit, or the
// detector is off): a closed form of the reference's per-bin count.
__device__ __forceinline__ uint32_t suppressed(uint32_t n, const CfarGeom& g) {
  if (!(g.enable && g.refs_ok)) return n;
  const uint32_t lo = g.rg ? g.gg + g.rg : 0u;
  const uint32_t hi = g.rl ? (n > g.gl + g.rl ? n - g.gl - g.rl : 0u) : n;
  return n - (hi > lo ? hi - lo : 0u);
}

// 176-bit event, LSB first: kind 2 | bin 16 | frame 16 | ref 7 | alpha 16 |
// det 16 | sup 16 | cut 40 | sum 47 (cfar_model.hpp:75-84); sof bit 176, eof 177.
__device__ __forceinline__ void put_event(uint64_t* w, uint32_t kind, uint32_t bin, uint32_t frame, uint32_t ref,
                                          uint32_t alpha, uint32_t det, uint32_t sup, uint64_t cut, uint64_t sum,
                                          bool sof, bool eof) {
  const uint64_t det16 = det & 0xFFFFu;
  const uint64_t cut40 = cut & ((1ull << 40) - 1);
  const uint64_t sum47 = sum & ((1ull << 47) - 1);
  w[0] = (uint64_t)(kind & 3u) | ((uint64_t)(bin & 0xFFFFu) << 2) | ((uint64_t)(frame & 0xFFFFu) << 18) |
         ((uint64_t)(ref & 0x7Fu) << 34) | ((uint64_t)(alpha & 0xFFFFu) << 41) | (det16 << 57);
  w[1] = (det16 >> 7) | ((uint64_t)(sup & 0xFFFFu) << 9) | (cut40 << 25);
  w[2] = (cut40 >> 39) | (sum47 << 1) | ((uint64_t)sof << 48) | ((uint64_t)eof << 49);
}

__device__ __forceinline__ uint32_t frame_of(uint32_t r, const uint32_t* frame_ids, uint32_t frame_mod) {
  return frame_ids ? frame_ids[r] : (frame_mod ? r % frame_mod : r);
}

// One row: prefix, decisions, and its events in bin order with the summary
// last, written to dst[0 .. room). Every thread of the row's W waves calls it;
// `work` false makes a thread skip all memory work but keep the barriers (so
// rows of one block may differ in whether they run). wr = wave within the row.
// Returns the row's event count (summary included).
template <typename Cell, bool MULTI, uint32_t SLOG>
__device__ __forceinline__ uint64_t cfar_row_run(const CfarJob& J, uint32_t W, uint32_t r, bool work, uint32_t wr,
                                                uint32_t lane, uint64_t* lds, uint64_t* dst, uint64_t room) {
  constexpr uint32_t S = 1u << SLOG;
  const uint32_t n = J.n;
  const CfarGeom g = J.g;
  uint64_t* P = lds;
  uint64_t* stats = lds + pidx<SLOG>(n) + 1;  // [W][4]
  const uint32_t t = wr * 64 + lane;
  const uint32_t b0 = t * S;
  const bool full = (wr + 1) * 64 * S <= n;  // the wave's bins all inside the row

  // the lane's segment straight into registers, local inclusive prefix
  uint64_t loc[S];
  {
    Cell v[S];
    const Cell* row = static_cast<const Cell*>(J.cells) + (size_t)r * n;
    constexpr uint32_t NV = sizeof(Cell) * S / 16;  // 16-byte words of a lane's segment
    if (NV > 0 && work && full && (reinterpret_cast<uintptr_t>(row) & 15u) == 0) {
      // a whole segment in a 16-byte aligned row: S*sizeof(Cell) is a multiple
      // of 16 and b0 a multiple of S, so the segment is 16-byte words
      uint4 q[NV > 0 ? NV : 1];
      const uint4* src = reinterpret_cast<const uint4*>(row + b0);
#pragma unroll
      for (uint32_t k = 0; k < NV; ++k) q[k] = src[k];
      __builtin_memcpy(v, q, sizeof(v));
    } else {
#pragma unroll
      for (uint32_t j = 0; j < S; ++j) v[j] = (work && (full || b0 + j < n)) ? row[b0 + j] : Cell(0);
    }
    uint64_t run = 0;
#pragma unroll
    for (uint32_t j = 0; j < S; ++j) {
      run += (uint64_t)v[j];
      loc[j] = run;
    }
  }
  const uint64_t incl = wave_incscan(loc[S - 1]);
  uint64_t base = incl - loc[S - 1];
  if (MULTI) {
    const uint64_t wave_total = lane63(incl);
    if (lane == 0) stats[wr * 4 + 0] = wave_total;
    __syncthreads();
    for (uint32_t w = 0; w < wr; ++w) base += stats[w * 4 + 0];
  }
  if (work) {
    // P[b0 + j + 1] sits at a fixed offset from the lane's base; a wave
    // inside the row stores without per-entry tests
    uint64_t* Pl = P + pidx<SLOG>(b0 + 1) - 1;
    auto off = [](uint32_t j) { return j + 1 + ((SLOG && j + 1 == S) ? 1u : 0u); };
    if (full) {
#pragma unroll
      for (uint32_t j = 0; j < S; ++j) Pl[off(j)] = base + loc[j];
    } else {
#pragma unroll
      for (uint32_t j = 0; j < S; ++j)
        if (b0 + j + 1 <= n) Pl[off(j)] = base + loc[j];
    }
    if (t == 0) P[0] = 0;
  }
  row_sync<MULTI>();

  // decisions, first pass: which bins emit (chunk masks) and the counts. A
  // CA chunk whose 64 bins all have their whole window inside the row takes
  // the unclamped form.
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t c0 = wr * 64 * S;
  const uint32_t n_ref = g.rl + g.rg;
  const bool ca = g.enable && g.refs_ok && !g.go && n_ref < (1u << 24);
  const uint64_t* Q = P + pidx<SLOG>(c0);
  constexpr uint32_t kStep = 64 + (64 >> SLOG);  // pidx advance per chunk of 64 bins
  CaRuns runs;
  runs.c1 = Q + lane_run<SLOG>(lane, 1);
  runs.c0 = Q + lane_run<SLOG>(lane, 0);
  runs.a1 = Q + lane_run<SLOG>(lane, (int)(g.gl + g.rl + 1));
  runs.a0 = Q + lane_run<SLOG>(lane, (int)(g.gl + 1));
  runs.b1 = Q + lane_run<SLOG>(lane, -(int)g.gg);
  runs.b0 = Q + lane_run<SLOG>(lane, -(int)(g.gg + g.rg));
  // chunks [qa, qb) of the wave, 64 bins from c0 + 64 q, have every bin's
  // window inside the row: q >= qa is c0 + 64 q >= guard_lag + ref_lag, q < qb
  // is c0 + 64 q + 64 + guard_lead + ref_lead <= n
  uint32_t qa = S, qb = 0;
  if (work && ca) {
    const uint32_t lo = g.gg + g.rg, span = 64 + g.gl + g.rl;
    qa = lo > c0 ? (lo - c0 + 63) / 64 : 0u;
    qb = n >= c0 + span ? (n - c0 - span) / 64 + 1 : 0u;
  }
  uint64_t em[S];
  uint32_t cnt = 0, det = 0;
  static_for<S>([&](auto qc) {
    constexpr uint32_t q = decltype(qc)::value;
    if (q >= qa && q < qb) {  // every bin of the chunk is in the row
      const uint64_t hm = __ballot(ca_hit_inner<Cell, q * kStep>(runs, g.alpha, n_ref << 8));
      if (g.dense) {  // every bin emits
        em[q] = __ballot(true);
        det += (uint32_t)__popcll(hm);
      } else {
        em[q] = hm;
      }
    } else if (ca && !g.dense) {  // an edge chunk of a CA row: only detections emit
      const uint32_t i = c0 + q * 64 + lane;
      const bool in = work && (full || i < n);
      em[q] = __ballot(in && ca_hit_edge<Cell, SLOG>(P, n, i, g, n_ref << 8));
    } else {
      const uint32_t i = c0 + q * 64 + lane;
      const bool in = work && (full || i < n);
      Decision d{};
      if (in) d = decide<SLOG>(P, n, i, g);
      em[q] = __ballot(in && (d.hit || g.dense));
      if (g.dense) det += (uint32_t)__popcll(__ballot(in && d.hit));
    }
    cnt += (uint32_t)__popcll(em[q]);
  });
  if (!g.dense) det = cnt;
  uint64_t k = 0, K = cnt;
  uint32_t DET = det;
  if (MULTI) {
    if (lane == 0) {
      stats[wr * 4 + 1] = cnt;
      stats[wr * 4 + 2] = det;
    }
    __syncthreads();
    K = 0;
    DET = 0;
    for (uint32_t w = 0; w < W; ++w) {
      const uint64_t c = stats[w * 4 + 1];
      if (w < wr) k += c;
      K += c;
      DET += (uint32_t)stats[w * 4 + 2];
    }
  }

  // second pass: the wave's emitting bins are decided again and written at
  // their offsets. At most 64 (a sparse row's usual case): compacted in bin
  // order into the wave's LDS list, one pass with the lanes full; more: chunk
  // by chunk, the emitting lanes of each chunk at once.
  const uint32_t frame = frame_of(r, J.frame_ids, J.frame_mod);
  auto emit = [&](uint32_t i, uint64_t pos) {
    Decision d;
    if (ca && !g.dense) {
      // a CA event without the dense mode is a detection, so its window
      // fits the row: the sums need no clamping
      d.cut = P[pidx<SLOG>(i + 1)] - P[pidx<SLOG>(i)];
      const uint64_t lead = g.rl ? P[pidx<SLOG>(i + g.gl + g.rl + 1)] - P[pidx<SLOG>(i + g.gl + 1)] : 0ull;
      const uint64_t lag = g.rg ? P[pidx<SLOG>(i - g.gg)] - P[pidx<SLOG>(i - g.gg - g.rg)] : 0ull;
      d.sum = lead + lag;
      d.n = n_ref;
      d.eval = d.hit = true;
    } else {
      d = decide<SLOG>(P, n, i, g);
    }
    const uint32_t kind = d.hit ? 0u : (!d.eval ? 2u : 1u);
    if (pos < room)
      put_event(dst + 3 * pos, kind, i, frame, d.eval ? d.n : 0u, g.alpha, 0, 0, d.cut, d.eval ? d.sum : 0ull,
                pos == 0, false);
  };
  if (cnt <= 64) {
    uint32_t* lst = reinterpret_cast<uint32_t*>(stats + W * 4) + wr * 64;
    uint32_t kq = 0;
#pragma unroll
    for (uint32_t q = 0; q < S; ++q) {
      if (em[q] == 0) continue;  // uniform: most chunks of a sparse row emit nothing
      if ((em[q] >> lane) & 1ull) lst[kq + (uint32_t)__popcll(em[q] & below)] = c0 + q * 64 + lane;
      kq += (uint32_t)__popcll(em[q]);
    }
    wave_lds_sync();
    if (lane < cnt) emit(lst[lane], k + lane);
  } else {
    uint64_t kq = k;
#pragma unroll
    for (uint32_t q = 0; q < S; ++q) {
      if (em[q] == 0) continue;
      if ((em[q] >> lane) & 1ull) emit(c0 + q * 64 + lane, kq + (uint32_t)__popcll(em[q] & below));
      kq += (uint32_t)__popcll(em[q]);
    }
  }
  k += cnt;
  if (work && wr == 0 && lane == 0 && K < room)
    put_event(dst + 3 * K, 3u, n, frame, 0, g.alpha, DET, suppressed(n, g), 0, 0, K == 0, true);
  // the caller may reuse the row's LDS after this
  row_sync<MULTI>();
  return K + 1;
}

// Pass 1 for the rows of one block (R rows x W waves): events to the staging
// slots, row_count[r] = the row's event count.
template <typename Cell, bool MULTI, uint32_t SLOG>
__device__ __forceinline__ void cfar_rows_tile(const CfarJob& J, const CfarRowGeo& G, uint32_t r0, uint64_t* lds) {
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const uint32_t rl = wave / G.W, wr = wave % G.W;
  const uint32_t r = r0 + rl;
  const bool valid = r < J.n_rows;
  const uint64_t cnt = cfar_row_run<Cell, MULTI, SLOG>(J, G.W, r, valid, wr, lane, lds + (size_t)rl * G.words,
                                                       J.stage + (size_t)(valid ? r : 0) * J.slot * 3,
                                                       valid ? J.slot : 0);
  if (valid && wr == 0 && lane == 0) J.row_count[r] = (uint32_t)cnt;
}

// Pass 3: each row's staged events to its offset; rows that overflowed their
// slot are recomputed and written directly.
template <typename Cell, bool MULTI, uint32_t SLOG>
__device__ __forceinline__ void cfar_gather_tile(const CfarJob& J, const CfarRowGeo& G, uint32_t r0,
                                                 uint64_t* lds) {
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const uint32_t rl = wave / G.W, wr = wave % G.W;
  const uint32_t r = r0 + rl;
  const bool valid = r < J.n_rows;
  uint64_t base = 0, room = 0;
  uint32_t cnt = 0;
  if (valid) {
    base = J.offs[r];
    cnt = J.row_count[r];
    room = base < J.cap ? J.cap - base : 0;
  }
  if (valid && cnt <= J.slot) {  // copy, the row's waves splitting the words
    const uint64_t words = 3ull * min((uint64_t)cnt, room);
    const uint64_t* src = J.stage + (size_t)r * J.slot * 3;
    for (uint64_t w = (uint64_t)wr * 64 + lane; w < words; w += 64ull * G.W) J.events[3 * base + w] = src[w];
  }
  const bool redo = valid && cnt > J.slot;
  if (!MULTI) {
    if (redo)
      cfar_row_run<Cell, false, SLOG>(J, 1, r, true, 0, lane, lds + (size_t)rl * G.words, J.events + 3 * base,
                                      room);
  } else if (__syncthreads_or(redo)) {
    cfar_row_run<Cell, true, SLOG>(J, G.W, r, redo, wr, lane, lds + (size_t)rl * G.words, J.events + 3 * base,
                                   redo ? room : 0);
  }
}

// Pass 2, one block of 1024: exclusive scan of the row counts into 64-bit
// offsets; offs[n_rows] = total. Two forms.
// Direct, for at most kScanDirectRows rows (one pass, no LDS staging; more
// rows take more passes): each thread takes a contiguous run of at most
// kScanDirectPer rows, its loads issued at once and kept in registers, the
// waves scan their run sums with shuffles, and one barrier combines the 16
// wave totals.
constexpr uint32_t kScanDirectPer = 4;
constexpr uint32_t kScanDirectRows = 1024 * kScanDirectPer;

__device__ __forceinline__ void cfar_scan_direct(const CfarJob& J) {
  __shared__ uint64_t wsum[16];
  const uint32_t n_rows = J.n_rows;
  const uint32_t* __restrict__ row_count = J.row_count;
  const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
  uint64_t carry = 0;
  for (uint32_t c0 = 0; c0 < n_rows; c0 += kScanDirectRows) {
    const uint32_t rows = min(n_rows - c0, kScanDirectRows);
    const uint32_t per = (rows + 1023) / 1024;
    const uint32_t r0 = c0 + t * per;
    const uint32_t lim = r0 < c0 + rows ? min(per, c0 + rows - r0) : 0u;  // rows of this thread
    const uint32_t* rc = row_count + r0;
    uint64_t* of = J.offs + r0;
    uint32_t v[kScanDirectPer];
    uint64_t s = 0;
#pragma unroll
    for (uint32_t q = 0; q < kScanDirectPer; ++q) {
      v[q] = q < lim ? rc[q] : 0u;
      s += v[q];
    }
    const uint64_t x = wave_incscan(s);
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint64_t wb = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < 16; ++w) {
      const uint64_t ws = wsum[w];
      wb += w < wave ? ws : 0;
      all += ws;
    }
    uint64_t run = carry + wb + x - s;
#pragma unroll
    for (uint32_t q = 0; q < kScanDirectPer; ++q) {
      if (q < lim) of[q] = run;
      run += v[q];
    }
    carry += all;
    __syncthreads();  // wsum is rewritten by the next pass
  }
  if (t == 0) {
    J.offs[n_rows] = carry;
    if (J.total) *J.total = carry;
  }
}

// Staged, for more rows: passes of kScanRows rows, every global access
// coalesced: the counts go through LDS (thread t loads rows t, t +
// 1024, ...), each thread then scans a contiguous run of kScanPer of them
// (LDS index i + i / kScanPer: one padding word per run keeps the run reads
// free of bank conflicts), the waves scan their run sums with shuffles, one
// barrier combines the 16 wave totals, and the offsets return through LDS to
// coalesced stores. The next pass's counts are loaded before this pass's
// stores.
constexpr uint32_t kScanPer = 8;
constexpr uint32_t kScanRows = 1024 * kScanPer;
constexpr uint32_t kScanPad = kScanRows + kScanRows / kScanPer;  // padded words per array

__device__ __forceinline__ void cfar_scan_staged(const CfarJob& J) {
  __shared__ uint64_t wsum[16];
  __shared__ uint32_t cnt[kScanPad];
  __shared__ uint64_t off[kScanPad];
  const uint32_t n_rows = J.n_rows;
  const uint32_t* __restrict__ row_count = J.row_count;
  const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
  auto pad = [](uint32_t i) { return i + i / kScanPer; };
  uint32_t ld[kScanPer];
  auto fetch = [&](uint32_t c0) {
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) {
      const uint32_t r = c0 + t + 1024 * j;
      ld[j] = r < n_rows ? row_count[r] : 0u;
    }
  };
  uint64_t carry = 0;
  fetch(0);
  for (uint32_t c0 = 0; c0 < n_rows; c0 += kScanRows) {
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) cnt[pad(t + 1024 * j)] = ld[j];
    __syncthreads();
    if (c0 + kScanRows < n_rows) fetch(c0 + kScanRows);  // in flight during this pass
    uint32_t v[kScanPer];
    uint64_t sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < kScanPer; ++q) {
      v[q] = cnt[pad(kScanPer * t + q)];
      sum += v[q];
    }
    const uint64_t x = wave_incscan(sum);
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint64_t wb = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < 16; ++w) {
      const uint64_t ws = wsum[w];
      wb += w < wave ? ws : 0;
      all += ws;
    }
    uint64_t run = carry + wb + x - sum;
#pragma unroll
    for (uint32_t q = 0; q < kScanPer; ++q) {
      off[pad(kScanPer * t + q)] = run;
      run += v[q];
    }
    carry += all;
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) {
      const uint32_t r = c0 + t + 1024 * j;
      if (r < n_rows) J.offs[r] = off[pad(t + 1024 * j)];
    }
    __syncthreads();  // cnt, off and wsum are rewritten by the next pass
  }
  if (t == 0) {
    J.offs[n_rows] = carry;
    if (J.total) *J.total = carry;
  }
}

// Single CPI kernels: block b holds rows [b*R, +R).
template <typename Cell, bool MULTI, uint32_t SLOG>
__global__ __launch_bounds__(512) void cfar_rows_kernel(CfarJob J, CfarRowGeo G) {
  extern __shared__ uint64_t lds[];
  const uint32_t R = blockDim.x / (64 * G.W);
  cfar_rows_tile<Cell, MULTI, SLOG>(J, G, blockIdx.x * R, lds);
}

// STAGED selects the form; the host picks it by the row count (the staged
// form's LDS and extra barriers only slow a one-pass scan down).
template <bool STAGED>
__global__ __launch_bounds__(1024) void cfar_scan_kernel(CfarJob J) {
  if (STAGED)
    cfar_scan_staged(J);
  else
    cfar_scan_direct(J);
}

static void launch_cfar_scan(hipStream_t s, const CfarJob& J) {
  if (J.n_rows > kScanDirectRows)
    cfar_scan_kernel<true><<<1, 1024, 0, s>>>(J);
  else
    cfar_scan_kernel<false><<<1, 1024, 0, s>>>(J);
}

// Pass 3 of a single CPI for rows of one wave: one block per R rows, the
// copy or, for a row that overflowed its slot, its recomputation.
template <typename Cell, uint32_t SLOG>
__global__ __launch_bounds__(512) void cfar_gather_kernel(CfarJob J, CfarRowGeo G) {
  extern __shared__ uint64_t lds[];
  const uint32_t R = blockDim.x / 64;
  cfar_gather_tile<Cell, false, SLOG>(J, G, blockIdx.x * R, lds);
}

// Pass 3 of a single CPI for rows of several waves, as two launches, so that
// moving the staged events needs no LDS and no block barrier per row. 3a: one wave per row moves the row's staged events to its
// offset; rows that overflowed their slot are left to 3b.
__global__ __launch_bounds__(256) void cfar_copy_kernel(CfarJob J) {
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t r = blockIdx.x * 4 + wave;
  if (r >= J.n_rows) return;
  const uint32_t cnt = J.row_count[r];
  if (cnt > J.slot) return;
  const uint64_t base = J.offs[r];
  const uint64_t room = base < J.cap ? J.cap - base : 0;
  const uint64_t words = 3ull * min((uint64_t)cnt, room);
  const uint64_t* __restrict__ src = J.stage + (size_t)r * J.slot * 3;
  uint64_t* __restrict__ dst = J.events + 3 * base;
  for (uint64_t w = lane; w < words; w += 64) dst[w] = src[w];
}

// 3b: block b looks at rows [b * blockDim, +blockDim) at once (one count per
// thread) and recomputes the ones that overflowed their slot, one after the
// other with all its waves (a block is one row of W waves), writing their
// events directly at their offsets. Blocks with none return after one load.
template <typename Cell, bool MULTI, uint32_t SLOG>
__global__ __launch_bounds__(512) void cfar_redo_kernel(CfarJob J, CfarRowGeo G) {
  extern __shared__ uint64_t lds[];
  __shared__ uint64_t over_mask[8];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t r0 = blockIdx.x * blockDim.x, rt = r0 + threadIdx.x;
  const bool over = rt < J.n_rows && J.row_count[rt] > J.slot;
  const uint64_t m = __ballot(over);
  if (lane == 0) over_mask[wave] = m;
  if (!__syncthreads_or(over)) return;
  for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
    uint64_t mm = over_mask[w];  // the same for every thread: the loop is uniform
    while (mm) {
      const uint32_t r = r0 + 64 * w + (uint32_t)__builtin_ctzll(mm);
      mm &= mm - 1;
      const uint64_t base = J.offs[r];
      const uint64_t room = base < J.cap ? J.cap - base : 0;
      cfar_row_run<Cell, MULTI, SLOG>(J, G.W, r, true, wave, lane, lds, J.events + 3 * base, room);
    }
  }
}

// Batched: one block per row tile of a class of CPIs (same row length, so the
// same row geometry), block -> job via the tile table.
template <bool GATHER, bool MULTI, uint32_t SLOG>
__global__ __launch_bounds__(512) void cfar_batch_kernel(const CfarJob* __restrict__ jobs,
                                                         const uint32_t* __restrict__ tile_job, uint32_t n_tiles,
                                                         CfarRowGeo G) {
  extern __shared__ uint64_t lds[];
  const uint32_t R = blockDim.x / (64 * G.W);
  const uint32_t t = blockIdx.x;
  if (t >= n_tiles) return;
  const CfarJob& J = jobs[__builtin_amdgcn_readfirstlane(tile_job[t])];
  const uint32_t r0 = (t - J.tile0) * R;
  if (GATHER)
    cfar_gather_tile<uint32_t, MULTI, SLOG>(J, G, r0, lds);
  else
    cfar_rows_tile<uint32_t, MULTI, SLOG>(J, G, r0, lds);
}

// Batched scan: one block per CPI.
template <bool STAGED>
__global__ __launch_bounds__(1024) void cfar_scan_batch_kernel(const CfarJob* __restrict__ jobs) {
  const CfarJob J = jobs[blockIdx.x];
  if (STAGED)
    cfar_scan_staged(J);
  else
    cfar_scan_direct(J);
}

// Staging bytes a single-CPI launch asks for at most; with more rows the slot
// per row shrinks (down to none: every row is then recomputed by the gather).
constexpr size_t kCfarStageBudget = (size_t)512 << 20;

// Staging slot per row for a scratch of `bytes` after the row offsets: at
// most kCfarSlot, and never more than a row can emit (n events + the summary).
uint32_t cfar_slot_for(uint32_t n, uint32_t n_rows, size_t bytes) {
  const size_t offs = ((size_t)n_rows + 1) * sizeof(uint64_t);
  size_t slot = n + 1u < kCfarSlot ? n + 1u : kCfarSlot;
  if (n_rows == 0) return (uint32_t)slot;
  const size_t fit = bytes > offs ? (bytes - offs) / ((size_t)n_rows * 3 * sizeof(uint64_t)) : 0;
  return (uint32_t)(fit < slot ? fit : slot);
}

size_t cfar_scratch_bytes(uint32_t n, uint32_t n_rows) {
  const size_t offs = ((size_t)n_rows + 1) * sizeof(uint64_t);
  const size_t per_row = (size_t)(n_rows ? n_rows : 1) * 3 * sizeof(uint64_t);
  size_t slot = n + 1u < kCfarSlot ? n + 1u : kCfarSlot;
  if (slot * per_row > kCfarStageBudget) slot = kCfarStageBudget / per_row;
  return offs + slot * per_row;
}

// Rows per block for W waves per row: 4 waves in all (one row of 8 waves).
uint32_t cfar_rows_per_tile_max(uint32_t n, uint32_t W) {
  (void)n;
  return W >= 4 ? 1 : 4 / W;
}

static size_t cfar_lds_bytes(const CfarRowGeo& G, uint32_t R) { return (size_t)R * G.words * sizeof(uint64_t); }

// Waves per row for a single-CPI launch: the fewest that fit the row, more
// when there are too few rows to give every SIMD four waves (4096 rows of
// 1024 bins: one wave a row, 18.7 us, against 19.9 at two).
static uint32_t cfar_single_waves(uint32_t n, uint32_t n_rows) {
  uint32_t W = cfar_min_waves(n);
  while (W < 4 && (uint64_t)n_rows * W < 4096 && 128u * W <= n) W <<= 1;
  // rows of 4096 bins and more: a row's LDS (about 36-39 KB at 4096 bins)
  // allows four blocks per CU whatever W is, so eight waves a row keep the
  // most waves in flight
  if (n >= 4096) W = 8;
  return W;
}

// Run-time (W > 1, SLOG) to the compile-time kernels.
#define RD_CFAR_DISPATCH(KERNEL, ...)                             \
  switch (G.slog) {                                               \
    case 0: KERNEL(0, __VA_ARGS__); break;                        \
    case 1: KERNEL(1, __VA_ARGS__); break;                        \
    case 2: KERNEL(2, __VA_ARGS__); break;                        \
    case 3: KERNEL(3, __VA_ARGS__); break;                        \
    default: KERNEL(4, __VA_ARGS__); break;                       \
  }

// Rows pass with W waves per row; the gather (mostly a copy) with the fewest
// waves per row and as many rows per block as fit.
template <typename Cell>
static hipError_t launch_cfar_cell(hipStream_t s, const CfarJob& J, uint32_t W) {
  {
    const CfarRowGeo G = cfar_row_geo(J.n, W);
    const uint32_t R = cfar_rows_per_tile_max(J.n, W);
    const size_t lds = cfar_lds_bytes(G, R);
    const dim3 grid((J.n_rows + R - 1) / R), block(64 * G.W * R);
#define RD_ROWS(SL, M) cfar_rows_kernel<Cell, M, SL><<<grid, block, lds, s>>>(J, G)
    if (G.W > 1) {
      RD_CFAR_DISPATCH(RD_ROWS, true)
    } else {
      RD_CFAR_DISPATCH(RD_ROWS, false)
    }
#undef RD_ROWS
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  launch_cfar_scan(s, J);
  e = hipGetLastError();
  if (e != hipSuccess || J.cap == 0) return e;
  const uint32_t Wg = cfar_min_waves(J.n);
  const CfarRowGeo G = cfar_row_geo(J.n, Wg);
  if (Wg == 1) {  // one launch: a second costs more than the copy saves at this row length
    const uint32_t R = cfar_rows_per_tile_max(J.n, 1);
    const dim3 grid((J.n_rows + R - 1) / R), block(64 * R);
#define RD_GATHER(SL, M) cfar_gather_kernel<Cell, SL><<<grid, block, cfar_lds_bytes(G, R), s>>>(J, G)
    RD_CFAR_DISPATCH(RD_GATHER, false)
#undef RD_GATHER
    return hipGetLastError();
  }
  cfar_copy_kernel<<<(J.n_rows + 3) / 4, 256, 0, s>>>(J);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const dim3 grid((J.n_rows + 64 * G.W - 1) / (64 * G.W)), block(64 * G.W);
#define RD_REDO(SL, M) cfar_redo_kernel<Cell, M, SL><<<grid, block, cfar_lds_bytes(G, 1), s>>>(J, G)
  RD_CFAR_DISPATCH(RD_REDO, true)
#undef RD_REDO
  return hipGetLastError();
}

hipError_t launch_cfar(hipStream_t s, const void* cells, uint32_t cell_bytes, uint32_t n, uint32_t n_rows,
                       const rd_cfar_cfg& cfg, uint32_t max_guard, uint32_t max_ref, const uint32_t* frame_ids,
                       uint32_t frame_mod, uint64_t* events, uint64_t cap, uint32_t* row_count, void* scratch,
                       size_t scratch_bytes, uint64_t* total_dev) {
  if (scratch_bytes < ((size_t)n_rows + 1) * sizeof(uint64_t)) return hipErrorInvalidValue;
  uint64_t* offs = static_cast<uint64_t*>(scratch);
  uint64_t* stage = offs + n_rows + 1;
  const CfarJob J{cells, frame_ids, stage, row_count, offs, events, total_dev, cap,
                  cfar_geom(cfg, max_guard, max_ref), n, n_rows, frame_mod, cfar_slot_for(n, n_rows, scratch_bytes),
                  0};
  if (n_rows == 0) {  // no rows: the scan alone writes the zero total
    launch_cfar_scan(s, J);
    return hipGetLastError();
  }
  const uint32_t W = cfar_single_waves(n, n_rows);
  return cell_bytes == 4 ? launch_cfar_cell<uint32_t>(s, J, W) : launch_cfar_cell<uint64_t>(s, J, W);
}

static hipError_t launch_cfar_batch(hipStream_t s, uint32_t n, uint32_t W, uint32_t rpb, const ClassLaunch& L,
                                    bool gather) {
  if (L.n_tiles == 0) return hipSuccess;
  if (W < cfar_min_waves(n) || W > 8 || (W & (W - 1)) || rpb < 1 || rpb > cfar_rows_per_tile_max(n, W))
    return hipErrorInvalidValue;
  const CfarRowGeo G = cfar_row_geo(n, W);
  const size_t lds = cfar_lds_bytes(G, rpb);
  const CfarJob* jobs = static_cast<const CfarJob*>(L.jobs);
  const uint32_t block = 64 * W * rpb;
#define RD_BATCH(SL, GA, M) cfar_batch_kernel<GA, M, SL><<<L.grid, block, lds, s>>>(jobs, L.tile_job, L.n_tiles, G)
  if (gather) {
    if (W > 1) {
      RD_CFAR_DISPATCH(RD_BATCH, true, true)
    } else {
      RD_CFAR_DISPATCH(RD_BATCH, true, false)
    }
  } else {
    if (W > 1) {
      RD_CFAR_DISPATCH(RD_BATCH, false, true)
    } else {
      RD_CFAR_DISPATCH(RD_BATCH, false, false)
    }
  }
#undef RD_BATCH
  return hipGetLastError();
}

hipError_t launch_cfar_rows_batch(hipStream_t s, uint32_t n, uint32_t W, uint32_t rpb, const ClassLaunch& L) {
  return launch_cfar_batch(s, n, W, rpb, L, false);
}

hipError_t launch_cfar_scan_batch(hipStream_t s, const void* jobs, uint32_t n_jobs, uint32_t max_rows) {
  if (n_jobs == 0) return hipSuccess;
  if (max_rows > kScanDirectRows)
    cfar_scan_batch_kernel<true><<<n_jobs, 1024, 0, s>>>(static_cast<const CfarJob*>(jobs));
  else
    cfar_scan_batch_kernel<false><<<n_jobs, 1024, 0, s>>>(static_cast<const CfarJob*>(jobs));
  return hipGetLastError();
}

hipError_t launch_cfar_gather_batch(hipStream_t s, uint32_t n, uint32_t W, uint32_t rpb, const ClassLaunch& L) {
  return launch_cfar_batch(s, n, W, rpb, L, true);
}

}  // namespace rd
