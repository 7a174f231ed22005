// rd_capi.hip — the C ABI (include/rd_amd.h): context, memory helpers,
// argument validation, and the per-stage and composed-chain entry points.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <mutex>
#include <new>

#include "rd_fxp.h"
#include "rd_internal.h"

namespace {

#define RD_TRY(expr)                              \
  do {                                            \
    if ((expr) != hipSuccess) return RD_E_HIP;    \
  } while (0)

// The twiddle rule of model/python/gen_fft_twiddles.py:97-159 for a master table
// of M = 4096 entries: W^e = (q15(cos), q15(-sin)) of 2*pi*e/M, evaluated on the
// first quadrant and mapped to the others by exact sign/swap symmetry; q15 is
// round-to-nearest-even of x * 2^15 then saturation to int16.
int32_t q15_real(double x) {
  const double v = nearbyint(x * 32768.0);
  return v > 32767.0 ? 32767 : (v < -32768.0 ? -32768 : (int32_t)v);
}

void make_twiddles(uint32_t* out) {
  const uint32_t M = 4096, q4 = M / 4;
  for (uint32_t e = 0; e < M; ++e) {
    const uint32_t quad = e / q4, rem = e % q4;
    double c0 = 1.0, s0 = 0.0;
    if (rem) {
      const double th = 2.0 * M_PI * (double)rem / (double)M;
      c0 = cos(th);
      s0 = sin(th);
    }
    const double c = quad == 0 ? c0 : quad == 1 ? -s0 : quad == 2 ? -c0 : s0;
    const double s = quad == 0 ? s0 : quad == 1 ? c0 : quad == 2 ? -s0 : -c0;
    out[e] = rd::cpack(q15_real(c), q15_real(-s));
  }
}

const uint32_t* host_twiddles() {
  static uint32_t table[4096];
  static std::once_flag once;
  std::call_once(once, [] { make_twiddles(table); });
  return table;
}

// A context buffer of at least `need` bytes. An outgrown buffer is freed at
// once unless a graph of the context is alive (it may have captured the
// pointer); then it is kept until the context is destroyed.
rd_status grow(rd_ctx* c, void** buf, size_t* have, size_t need) {
  if (*have >= need) return RD_OK;
  if (*buf) {
    if (c->live_graphs > 0)
      c->retired.push_back(*buf);
    else
      RD_TRY(hipFree(*buf));
  }
  *buf = nullptr;
  *have = 0;
  RD_TRY(hipMalloc(buf, need));
  *have = need;
  return RD_OK;
}

bool pow2(uint32_t v) { return v && !(v & (v - 1)); }

constexpr size_t kStaticScratchBytes = (size_t)RD_STATIC_SCRATCH_BYTES;
static_assert(kStaticScratchBytes % 256 == 0, "whole 256-byte lines");

}  // namespace

namespace rd_static {
// Scratch of the allocation-free chain entry point (rd_chain_enqueue_static):
// the CFAR pass's row offsets, then as many staged events per row as fit.
// One copy per GPU, alive until the process ends. In host code `scratch` only
// names the symbol; its device address on each GPU is taken once, at the
// first rd_create on that device, and kept in static_scratch_dev[].
__device__ __align__(256) uint64_t scratch[kStaticScratchBytes / sizeof(uint64_t)];
}  // namespace rd_static

namespace {

constexpr int kMaxDevices = 64;
std::mutex static_scratch_mu;
void* static_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on the current device `device`.
rd_status static_scratch_for(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return RD_E_NODEV;
  std::lock_guard<std::mutex> lk(static_scratch_mu);
  if (!static_scratch_dev[device]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(rd_static::scratch)) != hipSuccess || !ptr) return RD_E_HIP;
    static_scratch_dev[device] = ptr;
  }
  *out = static_scratch_dev[device];
  return RD_OK;
}

rd_status fin(hipError_t e) { return e == hipSuccess ? RD_OK : RD_E_HIP; }

rd_status enter(rd_ctx* ctx) { return rd::ctx_enter(ctx); }

}  // namespace

namespace rd {

rd_status ctx_enter(rd_ctx* ctx) {
  if (!ctx) return RD_E_ARG;
  RD_TRY(hipSetDevice(ctx->device));
  return RD_OK;
}

rd_status check_cpi(const rd_cpi_desc* d) {
  if (!d) return RD_E_ARG;
  if (d->n_chan < 1 || d->n_chan > 64 || d->n_beams < 1 || d->n_beams > 128) return RD_E_ARG;
  if (d->taps < 1 || d->taps > 64 || d->n_frames < 1 || d->n_frames > 65535) return RD_E_ARG;
  if (!pow2(d->n_fft) || d->n_fft < 8 || d->n_fft > 4096 || (d->spc != 1 && d->spc != 2)) return RD_E_GEOM;
  if (d->n_fft / d->spc < 4) return RD_E_GEOM;
  if (d->x_layout > 1) return RD_E_ARG;
  return RD_OK;
}

rd_status check_cpi_bufs(const rd_cpi_bufs* b) {
  if (!b || !b->iq || !b->taps || !b->weights || !b->u || !b->X || !b->power || !b->row_count)
    return RD_E_ARG;
  if (((uintptr_t)b->u | (uintptr_t)b->X) & 15u) return RD_E_ARG;
  if (!b->events && b->events_cap) return RD_E_ARG;
  return RD_OK;
}

}  // namespace rd

extern "C" {

int rd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* rd_status_str(rd_status s) {
  switch (s) {
    case RD_OK: return "ok";
    case RD_E_ARG: return "invalid argument";
    case RD_E_GEOM: return "geometry rejected";
    case RD_E_HIP: return "HIP runtime error";
    case RD_E_CAPACITY: return "event buffer too small";
    case RD_E_NODEV: return "no such device";
    case RD_E_SCRATCH: return "input exceeds the static scratch of the allocation-free entry point";
  }
  return "unknown";
}

rd_status rd_create(int device, rd_ctx** out) {
  if (!out) return RD_E_ARG;
  *out = nullptr;
  if (device < 0 || device >= rd_device_count()) return RD_E_NODEV;
  rd_ctx* c = new (std::nothrow) rd_ctx();
  if (!c) return RD_E_ARG;
  c->device = device;
  rd_status st = RD_OK;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->tw, 4096 * sizeof(uint32_t)) != hipSuccess ||
      hipMemcpy(c->tw, host_twiddles(), 4096 * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMalloc(&c->total_dev, sizeof(uint64_t)) != hipSuccess ||
      hipHostMalloc(&c->total_host, sizeof(uint64_t), hipHostMallocDefault) != hipSuccess)
    st = RD_E_HIP;
  if (st == RD_OK) st = static_scratch_for(device, &c->static_scratch);
  if (st != RD_OK) {
    rd_destroy(c);
    return st;
  }
  *out = c;
  return RD_OK;
}

rd_status rd_destroy(rd_ctx* c) {
  if (!c || c->live_graphs > 0) return RD_E_ARG;  // destroy the context's graphs first
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->tw) (void)hipFree(c->tw);
  if (c->prep) (void)hipFree(c->prep);
  if (c->scratch) (void)hipFree(c->scratch);
  if (c->total_dev) (void)hipFree(c->total_dev);
  if (c->total_host) (void)hipHostFree(c->total_host);
  if (c->dop) (void)hipFree(c->dop);
  if (c->integ) (void)hipFree(c->integ);
  if (c->xa) (void)hipFree(c->xa);
  for (void* p : c->retired) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return RD_OK;
}

void* rd_stream(rd_ctx* c) { return c ? (void*)c->stream : nullptr; }

rd_status rd_sync(rd_ctx* c) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  return fin(hipStreamSynchronize(c->stream));
}

rd_status rd_malloc(rd_ctx* c, size_t bytes, void** dptr) {
  rd_status st = enter(c);
  if (st != RD_OK || !dptr) return st != RD_OK ? st : RD_E_ARG;
  return fin(hipMalloc(dptr, bytes ? bytes : 4));
}

rd_status rd_free(rd_ctx* c, void* dptr) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  return fin(hipFree(dptr));
}

rd_status rd_h2d(rd_ctx* c, void* dst, const void* src, size_t bytes) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  RD_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return fin(hipStreamSynchronize(c->stream));
}

rd_status rd_d2h(rd_ctx* c, void* dst, const void* src, size_t bytes) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  RD_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return fin(hipStreamSynchronize(c->stream));
}

rd_status rd_memset(rd_ctx* c, void* dst, int value, size_t bytes) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  return fin(hipMemsetAsync(dst, value, bytes, c->stream));
}

rd_status rd_twiddle_table(uint32_t* host_out4096) {
  if (!host_out4096) return RD_E_ARG;
  memcpy(host_out4096, host_twiddles(), 4096 * sizeof(uint32_t));
  return RD_OK;
}

rd_status rd_pfb_run(rd_ctx* c, const uint32_t* x, const uint32_t* coeff, uint32_t phases, uint32_t taps,
                     uint32_t n_chan, uint32_t n_beats, uint32_t* hist, uint32_t* y, uint8_t* flags) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!x || !coeff || !y || phases < 1 || taps < 1 || taps > 64 || n_chan > 65535) return RD_E_ARG;
  return fin(rd::launch_pfb(c->stream, x, coeff, phases, taps, n_chan, n_beats, hist, y, flags));
}

rd_status rd_fft_run(rd_ctx* c, const uint32_t* in, uint32_t n_fft, uint32_t spc, uint32_t scale_sched, int reorder,
                     uint32_t n_frames, uint32_t* out, uint32_t* bins, uint32_t* stage_flags) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!in || (((uintptr_t)in | (uintptr_t)out | (uintptr_t)bins) & 15u)) return RD_E_ARG;
  if (!pow2(n_fft) || n_fft < 8 || n_fft > 4096 || (spc != 1 && spc != 2) || n_fft / spc < 4) return RD_E_GEOM;
  return fin(rd::launch_fft(c->stream, c->tw, in, n_fft, spc, scale_sched, reorder, n_frames, out, bins,
                            stage_flags));
}

rd_status rd_fft_stage(rd_ctx* c, const uint32_t* u, uint32_t n_fft, uint32_t spc, uint32_t scale_sched,
                       uint32_t n_chan, uint32_t n_frames, uint32_t x_layout, uint32_t* X, uint32_t* stage_flags) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!u || !X || ((uintptr_t)u | (uintptr_t)X) & 15u || x_layout > 1) return RD_E_ARG;
  if (!pow2(n_fft) || n_fft < 8 || n_fft > 4096 || (spc != 1 && spc != 2) || n_fft / spc < 4) return RD_E_GEOM;
  if ((uint64_t)n_chan * n_frames > 0xFFFFFFFFull) return RD_E_ARG;
  return fin(rd::launch_fft(c->stream, c->tw, u, n_fft, spc, scale_sched, 1, n_chan * n_frames, nullptr, X,
                            stage_flags, x_layout ? n_frames : 0));
}

rd_status rd_align_run(rd_ctx* c, const uint32_t* X, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                       uint32_t* Xa) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!X || !Xa || n_chan < 1 || n_chan > 64) return RD_E_ARG;
  return fin(rd::launch_align(c->stream, X, n_chan, (uint64_t)n_frames * n_bins, Xa));
}

rd_status rd_bf_run(rd_ctx* c, const uint32_t* Xa, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                    uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power, int64_t* acc) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!Xa || !W || n_ant < 1 || n_ant > 64 || n_beams < 1 || n_beams > 128) return RD_E_ARG;
  st = grow(c, &c->prep, &c->prep_bytes, rd::bf_prep_bytes(n_ant, n_beams));
  if (st != RD_OK) return st;
  return fin(rd::launch_bf(c->stream, Xa, W, n_ant, n_beams, n_cols, y, flags, power, acc, c->prep));
}

rd_status rd_bf_stage(rd_ctx* c, const uint32_t* X, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                      uint32_t n_frames, uint32_t n_bins, uint32_t* y, uint8_t* flags, uint32_t* power, uint32_t* Xa) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!X || !W || n_ant < 1 || n_ant > 64 || n_beams < 1 || n_beams > 128) return RD_E_ARG;
  return fin(rd::launch_bf_chmajor(c->stream, X, W, n_ant, n_beams, (uint64_t)n_frames * n_bins, y, flags, power, Xa));
}

rd_status rd_bf_cov_stage(rd_ctx* c, const uint32_t* X, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                          uint32_t n_frames, uint32_t n_bins, uint32_t* y, uint8_t* flags, uint32_t* power,
                          uint32_t* Xa, int64_t* R, uint32_t x_layout) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!X || !W || !R || n_ant < 1 || n_ant > 64 || n_beams < 1 || n_beams > 128 || n_frames > 65535 ||
      x_layout > 1)
    return RD_E_ARG;
  if (rd::bfcov_fits(n_ant, n_bins, n_frames, n_beams)) {
    if ((uintptr_t)X & 15u) return RD_E_ARG;  // the one-kernel form moves X in 16-byte pieces
    return fin(rd::launch_bfcov(c->stream, X, W, n_ant, n_beams, n_frames, n_bins, y, flags, power, Xa, R,
                                x_layout == 1));
  }
  if (!Xa || x_layout != 0) return RD_E_ARG;  // the two-kernel form reads natural order
  RD_TRY(rd::launch_bf_chmajor(c->stream, X, W, n_ant, n_beams, (uint64_t)n_frames * n_bins, y, flags, power, Xa));
  return fin(rd::launch_cov(c->stream, Xa, n_ant, n_frames, n_bins, R));
}

rd_status rd_doppler_run(rd_ctx* c, const uint32_t* ring, uint32_t rows, uint32_t depth, uint32_t n_bins,
                         uint32_t newest, uint32_t n_dop, uint32_t spc, uint32_t scale_sched, int shift, uint32_t* out,
                         uint32_t* stage_flags) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!ring || !out || depth < 1 || newest >= depth || n_dop > depth) return RD_E_ARG;
  if (!pow2(n_dop) || n_dop < 8 || n_dop > 4096 || (spc != 1 && spc != 2) || n_dop / spc < 4) return RD_E_GEOM;
  // rows is one grid dimension and rows * n_bins the transform's frame count (uint32)
  if (rows > 65535 || (uint64_t)rows * n_bins > 0xFFFFFFFFull) return RD_E_ARG;
  st = grow(c, &c->dop, &c->dop_bytes, rd::doppler_scratch_bytes(rows, n_bins, n_dop));
  if (st != RD_OK) return st;
  return fin(rd::launch_doppler(c->stream, c->tw, ring, rows, depth, n_bins, newest, n_dop, spc, scale_sched, shift,
                                out, stage_flags, c->dop));
}

rd_status rd_cov_run(rd_ctx* c, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                     int64_t* R) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!Xa || !R || n_chan < 1 || n_chan > 64 || n_frames < 1 || n_frames > 65535) return RD_E_ARG;
  return fin(rd::launch_cov(c->stream, Xa, n_chan, n_frames, n_bins, R));
}

rd_status rd_cov_integrate(rd_ctx* c, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                           uint32_t mode, uint32_t exp_k, uint32_t acc_w, int64_t* R, uint8_t* sat_flags) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!Xa || !R || n_chan < 1 || n_chan > 64 || n_frames < 1 || mode > 1 || exp_k > 63 || acc_w < 2 || acc_w > 64)
    return RD_E_ARG;
  if ((uint64_t)n_bins * n_chan * (n_chan + 1) / 2 > (1ull << 31)) return RD_E_ARG;  // one work-item per entry
  return fin(rd::launch_cov_integrate(c->stream, Xa, n_chan, n_frames, n_bins, mode, exp_k, acc_w, R, sat_flags));
}

rd_status rd_cfar_run(rd_ctx* c, const void* cells, uint32_t cell_bytes, uint32_t n_bins, uint32_t n_rows,
                      const rd_cfar_cfg* cfg, uint32_t max_guard, uint32_t max_ref, const uint32_t* frame_ids,
                      uint32_t frame_mod, uint64_t* events, uint64_t events_cap, uint32_t* row_count,
                      uint64_t* n_events_host) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!cells || !cfg || !row_count || (cell_bytes != 4 && cell_bytes != 8) || n_bins < 1 || n_bins > 8191)
    return RD_E_ARG;
  if (!events && events_cap) return RD_E_ARG;
  st = grow(c, &c->scratch, &c->scratch_bytes, rd::cfar_scratch_bytes(n_bins, n_rows));
  if (st != RD_OK) return st;
  RD_TRY(rd::launch_cfar(c->stream, cells, cell_bytes, n_bins, n_rows, *cfg, max_guard, max_ref, frame_ids,
                         frame_mod, events, events_cap, row_count, c->scratch, c->scratch_bytes, c->total_dev));
  RD_TRY(hipMemcpyAsync(c->total_host, c->total_dev, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  RD_TRY(hipStreamSynchronize(c->stream));
  const uint64_t tot = *c->total_host;
  if (n_events_host) *n_events_host = tot;
  return tot > events_cap ? RD_E_CAPACITY : RD_OK;
}

rd_status rd_fxp_eval(rd_ctx* c, const rd_fxp_op* ops, uint64_t n, int64_t* y, uint32_t* flags, uint32_t* forms) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if ((n && (!ops || !y || !flags)) || n > (1ull << 39)) return RD_E_ARG;
  return fin(rd::launch_fxp_eval(c->stream, ops, n, y, flags, forms));
}

rd_status rd_fxp_accum(rd_ctx* c, uint32_t acc_w, const int64_t* x, uint32_t n, int64_t* y, uint32_t* step_flags,
                       uint32_t* sticky, uint32_t* count) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (n && (!x || !y || !step_flags || !sticky || !count)) return RD_E_ARG;
  return fin(rd::launch_fxp_accum(c->stream, acc_w, x, n, y, step_flags, sticky, count));
}

rd_status rd_fxp_cmult(rd_ctx* c, const uint32_t* a, const uint32_t* b, uint64_t n, int64_t* p4, int64_t* p3,
                       uint32_t* y, uint8_t* flags) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if ((n && (!a || !b || !p4 || !p3 || !y || !flags)) || n > (1ull << 39)) return RD_E_ARG;
  return fin(rd::launch_fxp_cmult(c->stream, a, b, n, p4, p3, y, flags));
}

rd_status rd_integ_reset(const rd_integ_cfg* cfg, rd_integ_ctl* ctl) {
  if (!cfg || !ctl || cfg->mode > 1) return RD_E_ARG;
  ctl->active = *cfg;
  ctl->count = 0;
  ctl->window_id = 0;
  return RD_OK;
}

extern "C++" {
namespace {

// Plans a control stream on the host (ctl advanced only on success), uploads
// its control words and runs `launch` on them; synchronous.
template <class L>
rd_status integ_call(rd_ctx* c, rd_integ_ctl* ctl, const rd_integ_step* steps, uint32_t n_steps,
                     rd_integ_window* windows, uint32_t cap, uint32_t* n_windows, L launch) {
  if (!ctl || (!steps && n_steps) || !n_windows || (cap && !windows) || ctl->active.mode > 1) return RD_E_ARG;
  for (uint32_t s = 0; s < n_steps; ++s)
    if (steps[s].cfg.mode > 1) return RD_E_ARG;
  rd_integ_ctl next = *ctl;
  std::vector<uint2> words(n_steps ? n_steps : 1);
  const uint32_t nw = rd::integ_plan(next, steps, n_steps, words.data(), windows, cap);
  rd_status st = grow(c, &c->integ, &c->integ_bytes, words.size() * sizeof(uint2));
  if (st != RD_OK) return st;
  RD_TRY(hipMemcpyAsync(c->integ, words.data(), words.size() * sizeof(uint2), hipMemcpyHostToDevice, c->stream));
  RD_TRY(launch(static_cast<const uint2*>(c->integ)));
  RD_TRY(hipStreamSynchronize(c->stream));
  *ctl = next;
  *n_windows = nw;
  return nw > cap ? RD_E_CAPACITY : RD_OK;
}

}  // namespace
}  // extern "C++"

rd_status rd_integ_run(rd_ctx* c, uint32_t acc_w, rd_integ_ctl* ctl, const rd_integ_step* steps, uint32_t n_steps,
                       const int64_t* data, uint64_t n_lanes, rd_integ_lane* lanes, int64_t* results,
                       rd_integ_window* windows, uint32_t windows_cap, uint32_t* n_windows) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (acc_w > 64 || (n_lanes && (!lanes || (n_steps && !data))) || (windows_cap && n_lanes && !results) ||
      n_lanes > (1ull << 39))
    return RD_E_ARG;
  return integ_call(c, ctl, steps, n_steps, windows, windows_cap, n_windows, [&](const uint2* w) {
    return rd::launch_integ(c->stream, w, n_steps, data, n_lanes, lanes, results, windows_cap, acc_w);
  });
}

rd_status rd_cov_integrate_windows(rd_ctx* c, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                                   uint32_t acc_w, rd_integ_ctl* ctl, const rd_integ_step* steps,
                                   rd_integ_lane* lanes, int64_t* results, rd_integ_window* windows,
                                   uint32_t windows_cap, uint32_t* n_windows) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!Xa || !lanes || n_chan < 1 || n_chan > 64 || acc_w > 64 || (windows_cap && !results)) return RD_E_ARG;
  if ((uint64_t)n_bins * n_chan * (n_chan + 1) / 2 > (1ull << 31)) return RD_E_ARG;  // one work-item per pair
  return integ_call(c, ctl, steps, n_frames, windows, windows_cap, n_windows, [&](const uint2* w) {
    return rd::launch_cov_integ(c->stream, w, Xa, n_chan, n_frames, n_bins, lanes, results, windows_cap, acc_w);
  });
}

rd_status rd_pfb_acc_run(rd_ctx* c, const uint32_t* x, const uint32_t* coeff, uint32_t phases, uint32_t taps,
                         uint32_t n_chan, uint32_t n_beats, const uint32_t* hist, int64_t* acc) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!x || !coeff || !acc || phases < 1 || taps < 1 || taps > 64) return RD_E_ARG;
  if ((uint64_t)n_chan * n_beats * phases > (1ull << 39)) return RD_E_ARG;
  return fin(rd::launch_pfb_acc(c->stream, x, coeff, phases, taps, n_chan, n_beats, hist, acc));
}

rd_status rd_cfar_decide(rd_ctx* c, const uint64_t* cells, uint32_t n_bins, uint32_t n_rows, const rd_cfar_cfg* cfg,
                         rd_cfar_decision* out) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!cells || !cfg || !out || (uint64_t)n_bins * n_rows > (1ull << 39)) return RD_E_ARG;
  return fin(rd::launch_cfar_decide(c->stream, cells, n_bins, n_rows, *cfg, out));
}

rd_status rd_cfar_enqueue(rd_ctx* c, const void* cells, uint32_t cell_bytes, uint32_t n_bins, uint32_t n_rows,
                          const rd_cfar_cfg* cfg, uint32_t max_guard, uint32_t max_ref, const uint32_t* frame_ids,
                          uint32_t frame_mod, uint64_t* events, uint64_t events_cap, uint32_t* row_count,
                          uint64_t* n_events_dev) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!cells || !cfg || !row_count || !n_events_dev || (cell_bytes != 4 && cell_bytes != 8) || n_bins < 1 ||
      n_bins > 8191)
    return RD_E_ARG;
  if (!events && events_cap) return RD_E_ARG;
  st = grow(c, &c->scratch, &c->scratch_bytes, rd::cfar_scratch_bytes(n_bins, n_rows));
  if (st != RD_OK) return st;
  return fin(rd::launch_cfar(c->stream, cells, cell_bytes, n_bins, n_rows, *cfg, max_guard, max_ref, frame_ids,
                             frame_mod, events, events_cap, row_count, c->scratch, c->scratch_bytes, n_events_dev));
}

namespace {

// The beamformer and the covariance run as one kernel (no aligned copy
// needed between them).
bool chain_fused(const rd_cpi_desc* d, const rd_cpi_bufs* b) {
  return b->R && rd::bfcov_fits(d->n_chan, d->n_fft, d->n_frames, d->n_beams);
}

// The path needs the aligned copy as an intermediate: the plain chain
// always, the MFMA chain when it computes the covariance unfused.
bool chain_needs_xa(const rd_cpi_desc* d, const rd_cpi_bufs* b, bool plain) {
  return plain || (b->R && !chain_fused(d, b));
}

// The blocked X layout is an intermediate of the fused path only.
rd_status check_x_layout(const rd_cpi_desc* d, const rd_cpi_bufs* b, bool plain) {
  return d->x_layout == 0 || (!plain && chain_fused(d, b)) ? RD_OK : RD_E_ARG;
}

// Validates a CPI and sizes the context's scratch for it (no launches).
rd_status chain_prepare(rd_ctx* c, const rd_cpi_desc* d, const rd_cpi_bufs* b, bool plain = false) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  st = rd::check_cpi(d);
  if (st != RD_OK) return st;
  st = rd::check_cpi_bufs(b);
  if (st != RD_OK) return st;
  st = check_x_layout(d, b, plain);
  if (st != RD_OK) return st;
  st = grow(c, &c->prep, &c->prep_bytes, rd::bf_prep_bytes(d->n_chan, d->n_beams));
  if (st != RD_OK) return st;
  if (!b->Xa && chain_needs_xa(d, b, plain)) {
    st = grow(c, &c->xa, &c->xa_bytes, (size_t)4 * d->n_chan * d->n_fft * d->n_frames);
    if (st != RD_OK) return st;
  }
  return grow(c, &c->scratch, &c->scratch_bytes, rd::cfar_scratch_bytes(d->n_fft, d->n_beams * d->n_frames));
}

// The chain's launches on the context stream (prepared CPI); the CFAR pass
// keeps its row offsets and staged events in `scratch` (scratch_bytes).
rd_status chain_launch(rd_ctx* c, const rd_cpi_desc* d, const rd_cpi_bufs* b, uint64_t* n_events_dev, void* scratch,
                       size_t scratch_bytes) {
  const uint32_t CH = d->n_chan, N = d->n_fft, F = d->n_frames, B = d->n_beams;
  const uint64_t cols = (uint64_t)F * N;
  hipStream_t s = c->stream;
  RD_TRY(rd::launch_pfb(s, b->iq, b->taps, N, d->taps, CH, F, nullptr, b->u, b->pfb_flags));
  const bool xb = d->x_layout == 1;  // (validated: the fused path)
  RD_TRY(rd::launch_fft(s, c->tw, b->u, N, d->spc, d->scale_sched, 1, CH * F, nullptr, b->X, b->fft_flags,
                        xb ? F : 0));
  if (chain_fused(d, b)) {
    // beamformer + power and covariance from one read of X
    RD_TRY(rd::launch_bfcov(s, b->X, b->weights, CH, B, F, N, b->y, b->yflags, b->power, b->Xa, b->R, xb));
  } else {
    // beamformer + power on the FFT layout; the same staged tiles give the
    // aligned copy the covariance reads
    uint32_t* xa = b->Xa ? b->Xa : (b->R ? static_cast<uint32_t*>(c->xa) : nullptr);
    if (b->R && !xa) return RD_E_ARG;
    RD_TRY(rd::launch_bf_chmajor(s, b->X, b->weights, CH, B, cols, b->y, b->yflags, b->power, xa));
    if (b->R) RD_TRY(rd::launch_cov(s, xa, CH, F, N, b->R));
  }
  RD_TRY(rd::launch_cfar(s, b->power, 4, N, B * F, d->cfar, d->max_guard, d->max_ref, nullptr, F, b->events,
                         b->events ? b->events_cap : 0, b->row_count, scratch, scratch_bytes,
                         n_events_dev ? n_events_dev : c->total_dev));
  return RD_OK;
}

}  // namespace

struct rd_graph {
  rd_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

rd_status rd_chain_enqueue(rd_ctx* c, const rd_cpi_desc* d, const rd_cpi_bufs* b, uint64_t* n_events_dev) {
  rd_status st = chain_prepare(c, d, b);
  if (st != RD_OK) return st;
  return chain_launch(c, d, b, n_events_dev, c->scratch, c->scratch_bytes);
}

rd_status rd_chain_enqueue_plain(rd_ctx* c, const rd_cpi_desc* d, const rd_cpi_bufs* b, uint64_t* n_events_dev) {
  rd_status st = chain_prepare(c, d, b, true);
  if (st != RD_OK) return st;
  const uint32_t CH = d->n_chan, N = d->n_fft, F = d->n_frames, B = d->n_beams;
  const uint64_t cols = (uint64_t)F * N;
  hipStream_t s = c->stream;
  uint32_t* xa = b->Xa ? b->Xa : static_cast<uint32_t*>(c->xa);
  RD_TRY(rd::launch_pfb(s, b->iq, b->taps, N, d->taps, CH, F, nullptr, b->u, b->pfb_flags));
  RD_TRY(rd::launch_fft(s, c->tw, b->u, N, d->spc, d->scale_sched, 1, CH * F, nullptr, b->X, b->fft_flags));
  RD_TRY(rd::launch_align(s, b->X, CH, cols, xa));
  RD_TRY(rd::launch_bf_plain(s, xa, b->weights, CH, B, cols, b->y, b->yflags, b->power));
  if (b->R) RD_TRY(rd::launch_cov_plain(s, xa, CH, F, N, b->R));
  RD_TRY(rd::launch_cfar(s, b->power, 4, N, B * F, d->cfar, d->max_guard, d->max_ref, nullptr, F, b->events,
                         b->events ? b->events_cap : 0, b->row_count, c->scratch, c->scratch_bytes,
                         n_events_dev ? n_events_dev : c->total_dev));
  return RD_OK;
}

rd_status rd_bf_plain_run(rd_ctx* c, const uint32_t* Xa, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                          uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!Xa || !W || n_ant < 1 || n_ant > 64 || n_beams < 1 || n_beams > 128) return RD_E_ARG;
  return fin(rd::launch_bf_plain(c->stream, Xa, W, n_ant, n_beams, n_cols, y, flags, power));
}

rd_status rd_cov_plain_run(rd_ctx* c, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                           int64_t* R) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  if (!Xa || !R || n_chan < 1 || n_chan > 64 || n_frames > 65535) return RD_E_ARG;
  return fin(rd::launch_cov_plain(c->stream, Xa, n_chan, n_frames, n_bins, R));
}

rd_status rd_chain_enqueue_static(rd_ctx* c, const rd_cpi_desc* d, const rd_cpi_bufs* b, uint64_t* n_events_dev) {
  rd_status st = enter(c);
  if (st != RD_OK) return st;
  st = rd::check_cpi(d);
  if (st != RD_OK) return st;
  st = rd::check_cpi_bufs(b);
  if (st != RD_OK) return st;
  st = check_x_layout(d, b, false);
  if (st != RD_OK) return st;
  if (!c->static_scratch) return RD_E_HIP;
  // the row offsets must fit; staged events take what is left (rows with
  // more are recomputed by the CFAR gather)
  const uint64_t rows = (uint64_t)d->n_beams * d->n_frames;
  if ((rows + 1) * sizeof(uint64_t) > kStaticScratchBytes) return RD_E_SCRATCH;
  // no allocation here: a path that needs the aligned copy needs the caller's
  if (!b->Xa && chain_needs_xa(d, b, false)) return RD_E_ARG;
  return chain_launch(c, d, b, n_events_dev, c->static_scratch, kStaticScratchBytes);
}

rd_status rd_chain_graph(rd_ctx* c, const rd_cpi_desc* d, const rd_cpi_bufs* b, uint64_t* n_events_dev,
                         rd_graph** out) {
  if (!out) return RD_E_ARG;
  *out = nullptr;
  rd_status st = chain_prepare(c, d, b);
  if (st != RD_OK) return st;
  rd_graph* g = new (std::nothrow) rd_graph();
  if (!g) return RD_E_ARG;
  g->ctx = c;
  RD_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  st = chain_launch(c, d, b, n_events_dev ? n_events_dev : c->total_dev, c->scratch, c->scratch_bytes);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (st != RD_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    delete g;
    return st != RD_OK ? st : RD_E_HIP;
  }
  g->graph = graph;
  if (hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(graph);
    delete g;
    return RD_E_HIP;
  }
  ++c->live_graphs;
  *out = g;
  return RD_OK;
}

rd_status rd_graph_launch(rd_graph* g) {
  if (!g) return RD_E_ARG;
  rd_status st = enter(g->ctx);
  if (st != RD_OK) return st;
  return fin(hipGraphLaunch(g->exec, g->ctx->stream));
}

rd_status rd_graph_destroy(rd_graph* g) {
  if (!g) return RD_E_ARG;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  --g->ctx->live_graphs;
  delete g;
  return RD_OK;
}

rd_status rd_chain_run(rd_ctx* c, const rd_cpi_desc* d, const rd_cpi_bufs* b, uint64_t* n_events_host) {
  rd_status st = rd_chain_enqueue(c, d, b, c ? c->total_dev : nullptr);
  if (st != RD_OK) return st;
  RD_TRY(hipMemcpyAsync(c->total_host, c->total_dev, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  RD_TRY(hipStreamSynchronize(c->stream));
  if (n_events_host) *n_events_host = *c->total_host;
  return (b->events && *c->total_host > b->events_cap) ? RD_E_CAPACITY : RD_OK;
}

}  // extern "C"
