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
// Three launches: per-row decisions with the events staged in a fixed slot per
// row, a scan of the per-row counts into event offsets, and a gather that moves
// each row's events to its offset (recomputing only rows that overflowed).
// Batched launches run the same row bodies over the rows of many CPIs and one
// scan block per CPI.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rd_internal.h"

namespace rd {

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

// One wave per row; a block holds up to 4 rows (waves), each with its own
// exclusive prefix P[0..n] in LDS. Lane l owns the interleaved bins l + 64*j, so
// global loads, LDS accesses and event stores are consecutive across lanes.
// The row is first copied into LDS with every load in flight, then scanned in
// place one 64-bin block at a time (wave scan plus carry). Nothing below needs
// a block-wide barrier: each wave only touches its own row.
__device__ __forceinline__ uint64_t wave_incscan(uint64_t v, uint32_t lane) {
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  return x;
}

// LDS writes of one lane become visible to the other lanes of the wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename Cell>
__device__ __forceinline__ void row_prefix(const Cell* row, uint32_t n, uint32_t lane, uint64_t* P) {
  for (uint32_t i = lane; i < n; i += 64) P[i + 1] = (uint64_t)row[i];
  wave_lds_sync();
  uint64_t carry = 0;
  for (uint32_t b0 = 0; b0 < n; b0 += 64) {
    const uint32_t i = b0 + lane;
    const uint64_t c = i < n ? P[i + 1] : 0ull;
    const uint64_t x = wave_incscan(c, lane);
    if (i < n) P[i + 1] = carry + x;
    carry += __shfl(x, 63);
  }
  if (lane == 0) P[0] = 0;
  wave_lds_sync();
}

struct Decision {
  bool hit, eval;
  uint64_t cut, sum;
  uint32_t n;
};

__device__ __forceinline__ Decision decide(const uint64_t* P, uint32_t n, uint32_t i, const CfarGeom& g) {
  Decision d;
  d.cut = P[i + 1] - P[i];
  const bool lead_ok = g.rl == 0 || (uint64_t)i + g.gl + g.rl < n;
  const bool lag_ok = g.rg == 0 || i >= g.gg + g.rg;
  const bool ok = lead_ok && lag_ok;
  uint64_t slead = 0, slag = 0;
  if (ok) {
    if (g.rl) slead = P[i + g.gl + g.rl + 1] - P[i + g.gl + 1];
    if (g.rg) slag = P[i - g.gg] - P[i - g.gg - g.rg];
  }
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

// The events of one row, in bin order, summary last, written to dst[0 .. room);
// returns the row's event count (summary included).
__device__ uint64_t emit_row(const uint64_t* P, uint32_t n, const CfarGeom& g, uint32_t frame, uint32_t lane,
                             uint64_t* dst, uint64_t room) {
  const uint64_t below = (1ull << lane) - 1ull;  // lanes before this one
  uint64_t k = 0;                                 // events emitted so far in the row
  uint32_t det = 0, sup = 0;
  for (uint32_t b0 = 0; b0 < n; b0 += 64) {
    const uint32_t i = b0 + lane;
    Decision d{};
    const bool in = i < n;
    if (in) d = decide(P, n, i, g);
    const bool emit = in && (d.hit || g.dense);
    const uint64_t em = __ballot(emit);
    det += (uint32_t)__popcll(__ballot(in && d.hit));
    sup += (uint32_t)__popcll(__ballot(in && !d.eval));
    if (emit) {
      const uint64_t pos = k + (uint64_t)__popcll(em & below);
      const uint32_t kind = d.hit ? 0u : (!d.eval ? 2u : 1u);
      if (pos < room)
        put_event(dst + 3 * pos, kind, i, frame, d.eval ? d.n : 0u, g.alpha, 0, 0, d.cut, d.eval ? d.sum : 0ull,
                  pos == 0, false);
    }
    k += (uint64_t)__popcll(em);
  }
  if (lane == 0 && k < room) put_event(dst + 3 * k, 3u, n, frame, 0, g.alpha, det, sup, 0, 0, k == 0, true);
  return k + 1;
}

__device__ __forceinline__ uint32_t frame_of(uint32_t r, const uint32_t* frame_ids, uint32_t frame_mod) {
  return frame_ids ? frame_ids[r] : (frame_mod ? r % frame_mod : r);
}

// Pass 1, one wave per row: the row's events go to its staging slot (as many
// as fit); row_count[r] receives the row's event count.
template <typename Cell>
__device__ __forceinline__ void cfar_row(const CfarJob& J, uint32_t r, uint32_t lane, uint64_t* P) {
  if (r >= J.n_rows) return;
  const uint32_t n = J.n;
  row_prefix(static_cast<const Cell*>(J.cells) + (size_t)r * n, n, lane, P);
  const uint64_t cnt = emit_row(P, n, J.g, frame_of(r, J.frame_ids, J.frame_mod), lane,
                                J.stage + (size_t)r * J.slot * 3, J.slot);
  if (lane == 0) J.row_count[r] = (uint32_t)cnt;
}

// Pass 3, one wave per row: copy the row's staged events to its offset; a row
// that overflowed its slot is recomputed and written directly.
template <typename Cell>
__device__ __forceinline__ void cfar_gather_row(const CfarJob& J, uint32_t r, uint32_t lane, uint64_t* P) {
  if (r >= J.n_rows) return;
  const uint64_t base = J.offs[r];
  const uint32_t cnt = J.row_count[r];
  const uint64_t room = base < J.cap ? J.cap - base : 0;
  if (cnt <= J.slot) {
    const uint64_t words = 3ull * min((uint64_t)cnt, room);
    const uint64_t* src = J.stage + (size_t)r * J.slot * 3;
    for (uint64_t w = lane; w < words; w += 64) J.events[3 * base + w] = src[w];
    return;
  }
  const uint32_t n = J.n;
  row_prefix(static_cast<const Cell*>(J.cells) + (size_t)r * n, n, lane, P);
  emit_row(P, n, J.g, frame_of(r, J.frame_ids, J.frame_mod), lane, J.events + 3 * base, room);
}

// Pass 2, one block: exclusive scan of the row counts into 64-bit offsets;
// offs[n_rows] = total.
__device__ __forceinline__ void cfar_scan_block(const CfarJob& J) {
  __shared__ uint64_t part[1024];
  const uint32_t n_rows = J.n_rows;
  const uint32_t* __restrict__ row_count = J.row_count;
  const uint32_t t = threadIdx.x;
  const uint32_t per = (n_rows + 1023) / 1024;
  const uint32_t r0 = t * per;
  uint64_t s = 0;
  for (uint32_t q = 0; q < per; ++q)
    if (r0 + q < n_rows) s += row_count[r0 + q];
  part[t] = s;
  __syncthreads();
  for (uint32_t o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
    const uint64_t v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint64_t run = part[t] - s;
  for (uint32_t q = 0; q < per; ++q) {
    if (r0 + q < n_rows) {
      J.offs[r0 + q] = run;
      run += row_count[r0 + q];
    }
  }
  if (t == 1023) {
    J.offs[n_rows] = part[1023];
    if (J.total) *J.total = part[1023];
  }
}

// Single CPI kernels: block b holds rows [b*rpb, +rpb), one per wave.
template <typename Cell>
__global__ __launch_bounds__(256) void cfar_rows_kernel(CfarJob J) {
  extern __shared__ uint64_t lds[];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  cfar_row<Cell>(J, blockIdx.x * (blockDim.x >> 6) + wave, lane, lds + (size_t)wave * (J.n + 1));
}

__global__ __launch_bounds__(1024) void cfar_scan_kernel(CfarJob J) { cfar_scan_block(J); }

template <typename Cell>
__global__ __launch_bounds__(256) void cfar_gather_kernel(CfarJob J) {
  extern __shared__ uint64_t lds[];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  cfar_gather_row<Cell>(J, blockIdx.x * (blockDim.x >> 6) + wave, lane, lds + (size_t)wave * (J.n + 1));
}

// Batched: persistent grids over the row tiles of a class of CPIs (same row
// length); every wave works on its own row, so no block barrier is needed.
template <bool GATHER>
__global__ __launch_bounds__(256) void cfar_batch_kernel(const CfarJob* __restrict__ jobs,
                                                         const uint32_t* __restrict__ tile_job, uint32_t n_tiles,
                                                         uint32_t n) {
  extern __shared__ uint64_t lds[];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, rpb = blockDim.x >> 6;
  uint64_t* P = lds + (size_t)wave * (n + 1);
  for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const CfarJob J = jobs[tile_job[t]];
    const uint32_t r = (t - J.tile0) * rpb + wave;
    if (GATHER)
      cfar_gather_row<uint32_t>(J, r, lane, P);
    else
      cfar_row<uint32_t>(J, r, lane, P);
  }
}

// Batched scan: one block per CPI.
__global__ __launch_bounds__(1024) void cfar_scan_batch_kernel(const CfarJob* __restrict__ jobs) {
  const CfarJob J = jobs[blockIdx.x];
  cfar_scan_block(J);
}

size_t cfar_scratch_bytes(uint32_t n_rows) {
  return ((size_t)n_rows + 1) * sizeof(uint64_t) + (size_t)n_rows * kCfarSlot * 3 * sizeof(uint64_t);
}

// Rows per block (waves) so the prefix rows fit 48 KiB of LDS, at most 4.
uint32_t cfar_rows_per_tile_max(uint32_t n) {
  const uint32_t rpb = (uint32_t)(49152 / (((size_t)n + 1) * sizeof(uint64_t)));
  return rpb < 1 ? 1 : (rpb > 4 ? 4 : rpb);
}

hipError_t launch_cfar(hipStream_t s, const void* cells, uint32_t cell_bytes, uint32_t n, uint32_t n_rows,
                       const rd_cfar_cfg& cfg, uint32_t max_guard, uint32_t max_ref, const uint32_t* frame_ids,
                       uint32_t frame_mod, uint64_t* events, uint64_t cap, uint32_t* row_count, void* scratch,
                       uint64_t* total_dev) {
  if (n_rows == 0) return hipSuccess;
  const uint32_t rpb = cfar_rows_per_tile_max(n);
  const size_t lds = rpb * ((size_t)n + 1) * sizeof(uint64_t);
  const dim3 grid((n_rows + rpb - 1) / rpb), block(64 * rpb);
  uint64_t* offs = static_cast<uint64_t*>(scratch);
  uint64_t* stage = offs + n_rows + 1;
  const CfarJob J{cells, frame_ids, stage, row_count, offs, events, total_dev, cap,
                  cfar_geom(cfg, max_guard, max_ref), n, n_rows, frame_mod, kCfarSlot, 0};
  if (cell_bytes == 4)
    cfar_rows_kernel<uint32_t><<<grid, block, lds, s>>>(J);
  else
    cfar_rows_kernel<uint64_t><<<grid, block, lds, s>>>(J);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  cfar_scan_kernel<<<1, 1024, 0, s>>>(J);
  e = hipGetLastError();
  if (e != hipSuccess || cap == 0) return e;
  if (cell_bytes == 4)
    cfar_gather_kernel<uint32_t><<<grid, block, lds, s>>>(J);
  else
    cfar_gather_kernel<uint64_t><<<grid, block, lds, s>>>(J);
  return hipGetLastError();
}

hipError_t launch_cfar_rows_batch(hipStream_t s, uint32_t n, uint32_t rpb, const ClassLaunch& L) {
  if (L.n_tiles == 0) return hipSuccess;
  if (rpb < 1 || rpb > cfar_rows_per_tile_max(n)) return hipErrorInvalidValue;
  cfar_batch_kernel<false><<<L.grid, 64 * rpb, rpb * ((size_t)n + 1) * sizeof(uint64_t), s>>>(
      static_cast<const CfarJob*>(L.jobs), L.tile_job, L.n_tiles, n);
  return hipGetLastError();
}

hipError_t launch_cfar_scan_batch(hipStream_t s, const void* jobs, uint32_t n_jobs) {
  if (n_jobs == 0) return hipSuccess;
  cfar_scan_batch_kernel<<<n_jobs, 1024, 0, s>>>(static_cast<const CfarJob*>(jobs));
  return hipGetLastError();
}

hipError_t launch_cfar_gather_batch(hipStream_t s, uint32_t n, uint32_t rpb, const ClassLaunch& L) {
  if (L.n_tiles == 0) return hipSuccess;
  if (rpb < 1 || rpb > cfar_rows_per_tile_max(n)) return hipErrorInvalidValue;
  cfar_batch_kernel<true><<<L.grid, 64 * rpb, rpb * ((size_t)n + 1) * sizeof(uint64_t), s>>>(
      static_cast<const CfarJob*>(L.jobs), L.tile_job, L.n_tiles, n);
  return hipGetLastError();
}

}  // namespace rd
