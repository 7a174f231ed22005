// rd_internal.h — launch interface between the C ABI and the kernel files.
//
// Every stage kernel is written as a per-tile body. A single-CPI launch maps
// one block to one tile; a batched launch runs a persistent grid over the
// tiles of many CPIs of one shape class, reading each tile's CPI from a device
// job table (tile -> job index, job -> pointers and sizes).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/rd_amd.h"

// The context behind rd_ctx* (rd_capi.hip).
struct rd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t* tw = nullptr;      // 4096-entry twiddle table (device)
  void* prep = nullptr;        // beamformer weight digits
  size_t prep_bytes = 0;
  void* scratch = nullptr;     // CFAR event offsets and staging
  size_t scratch_bytes = 0;
  uint64_t* total_dev = nullptr;
  uint64_t* total_host = nullptr;  // pinned
  int live_graphs = 0;             // graphs and batches hold the context's stream
  void* static_scratch = nullptr;  // this device's copy of the library's static scratch array
  void* dop = nullptr;             // slow-time transform columns
  size_t dop_bytes = 0;
  void* integ = nullptr;           // Integrator control words of the current call
  size_t integ_bytes = 0;
  void* xa = nullptr;              // aligned copy for chains whose caller passes none
  size_t xa_bytes = 0;
  std::vector<void*> retired;      // outgrown buffers a live graph may still use
};

namespace rd {

// Shared C-ABI helpers (rd_capi.hip).
rd_status ctx_enter(rd_ctx* ctx);                 // validates and makes the device current
rd_status check_cpi(const rd_cpi_desc* d);        // chain geometry limits
rd_status check_cpi_bufs(const rd_cpi_bufs* b);   // required chain buffers

// ---------------------------------------------------------------- job records
// PFB of one CPI: tiles = nbx * n_chan * nz (branch block fastest, then
// channel, then beat chunk).
struct PfbJob {
  const uint32_t* x;
  const uint32_t* coeff;
  const uint32_t* hist;
  uint32_t* y;
  uint8_t* flags;
  uint32_t P, T, n_chan, n_beats, chunk, nbx, nz, tile0;
};

// FFT of one CPI's frames: tiles = ceil(n_frames / frames per block).
struct FftJob {
  const uint32_t* in;
  uint32_t* out;
  uint32_t* bins;
  uint32_t* stage_flags;
  uint32_t sched, reorder, n_frames, tile0;
  uint32_t bins_fpc;  // 0: bins in natural order; else bin-quad blocked with this many frames per channel
};

// Beamformer of one CPI: tiles = ceil(cols / TC).
struct BfJob {
  const uint32_t* X;
  const uint32_t* W;
  uint32_t* y;
  uint8_t* flags;
  uint32_t* power;
  int64_t* acc;
  uint32_t* xa_out;
  uint64_t cols;
  uint32_t CH, B, bt_n, tile0;
};

// Covariance of one CPI: tiles = ceil(N / bins per block).
struct CovJob {
  const uint32_t* Xa;
  int64_t* R;
  uint32_t CH, F, N, tile0;
};

// Beamformer + covariance of one CPI read once from the FFT layout
// (bf.X = X[ch][frame][bin], bf.xa_out the optional aligned copy).
struct BfCovJob {
  BfJob bf;
  int64_t* R;
  uint32_t N, F;
  uint32_t x_blocked;  // X bin-quad blocked [ch][N/4][F][4] instead of [ch][F][N]
};

// CFAR geometry after clamping to the elaboration maxima.
struct CfarGeom {
  uint32_t enable, go, dense, gl, gg, rl, rg, alpha, refs_ok;
};

// CFAR of one CPI's rows: tiles = ceil(n_rows / rows per block).
struct CfarJob {
  const void* cells;
  const uint32_t* frame_ids;
  uint64_t* stage;   // [n_rows][slot][3] staged events
  uint32_t* row_count;
  uint64_t* offs;    // [n_rows + 1] event offsets
  uint64_t* events;
  uint64_t* total;   // event count out (nullable)
  uint64_t cap;
  CfarGeom g;
  uint32_t n, n_rows, frame_mod, slot, tile0;
};

CfarGeom cfar_geom(const rd_cfar_cfg& cfg, uint32_t max_guard, uint32_t max_ref);

// ---------------------------------------------------------------- single CPI
// Streaming stages (rd_stream.hip)
hipError_t launch_pfb(hipStream_t s, const uint32_t* x, const uint32_t* coeff, uint32_t P, uint32_t T,
                      uint32_t n_chan, uint32_t n_beats, uint32_t* hist, uint32_t* y, uint8_t* flags);
// bins_fpc > 0 writes `bins` bin-quad blocked: frame f = ch * bins_fpc + m,
// bin k at [ch][k / 4][m][k % 4] (the chain's layout for the fused beamformer
// + covariance); `out` stays natural.
hipError_t launch_fft(hipStream_t s, const uint32_t* tw4096, const uint32_t* in, uint32_t n, uint32_t spc,
                      uint32_t sched, int reorder, uint32_t n_frames, uint32_t* out, uint32_t* bins,
                      uint32_t* stage_flags, uint32_t bins_fpc = 0);
hipError_t launch_align(hipStream_t s, const uint32_t* X, uint32_t n_chan, uint64_t cols, uint32_t* Xa);
// Slow-time transform of the last H frames of a ring [rows][depth][n_bins]
// (newest frame in slot `newest`) -> out [rows][H][n_bins], stage_flags
// [rows][n_bins] (nullable). scratch holds doppler_scratch_bytes().
size_t doppler_scratch_bytes(uint32_t rows, uint32_t n_bins, uint32_t H);
hipError_t launch_doppler(hipStream_t s, const uint32_t* tw4096, const uint32_t* ring, uint32_t rows, uint32_t depth,
                          uint32_t n_bins, uint32_t newest, uint32_t H, uint32_t spc, uint32_t sched, int shift,
                          uint32_t* out, uint32_t* stage_flags, void* scratch);

// MFMA contractions (rd_mfma.hip)
size_t bf_prep_bytes(uint32_t n_ant, uint32_t n_beams);
hipError_t launch_bf(hipStream_t s, const uint32_t* Xa, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                     uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power, int64_t* acc, void* prep);
// The beamformer reading the FFT's channel-major X[ch][col] and writing the
// aligned copy Xa[col][ch] (nullable) from the same staged tiles.
hipError_t launch_bf_chmajor(hipStream_t s, const uint32_t* X, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                             uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power, uint32_t* xa_out);
hipError_t launch_cov(hipStream_t s, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                      int64_t* R);
// Beamformer + power and covariance in one kernel over the FFT layout X; for
// the geometries bfcov_fits accepts (at most 16 channels, bins a multiple of 4).
bool bfcov_fits(uint32_t n_chan, uint32_t n_bins, uint32_t n_frames, uint32_t n_beams);
hipError_t launch_bfcov(hipStream_t s, const uint32_t* X, const uint32_t* W, uint32_t n_chan, uint32_t n_beams,
                        uint32_t n_frames, uint32_t n_bins, uint32_t* y, uint8_t* flags, uint32_t* power,
                        uint32_t* xa_out, int64_t* R, bool x_blocked = false);
// The Integrator frame by frame: mode 0 block, 1 exponential (shift exp_k),
// saturating to acc_w (2..64) with sticky flags (nullable) per [bin][pair].
hipError_t launch_cov_integrate(hipStream_t s, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames,
                                uint32_t n_bins, uint32_t mode, uint32_t exp_k, uint32_t acc_w, int64_t* R,
                                uint8_t* sat_flags);

// The same contractions without matrix cores (rd_plain.hip): the plain-GPU
// comparison point. Xa aligned [cols][n_ant] / [n_frames][n_bins][n_chan].
hipError_t launch_bf_plain(hipStream_t s, const uint32_t* Xa, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                           uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power);
hipError_t launch_cov_plain(hipStream_t s, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                            int64_t* R);

// Exact per-element model calls (rd_exact.hip). integ_plan turns a control
// stream into one control word per step (advancing ctl) and fills the
// windows it emits (at most cap); it returns their number.
uint32_t integ_plan(rd_integ_ctl& ctl, const rd_integ_step* steps, uint32_t n_steps, uint2* words,
                    rd_integ_window* windows, uint32_t cap);
hipError_t launch_integ(hipStream_t s, const uint2* ctl, uint32_t n_steps, const int64_t* data, uint64_t n_lanes,
                        rd_integ_lane* lanes, int64_t* results, uint32_t cap, uint32_t acc_w);
hipError_t launch_cov_integ(hipStream_t s, const uint2* ctl, const uint32_t* Xa, uint32_t CH, uint32_t F, uint32_t N,
                            rd_integ_lane* lanes, int64_t* results, uint32_t cap, uint32_t acc_w);
hipError_t launch_pfb_acc(hipStream_t s, const uint32_t* x, const uint32_t* coeff, uint32_t P, uint32_t T,
                          uint32_t n_chan, uint32_t n_beats, const uint32_t* hist, int64_t* acc);
hipError_t launch_cfar_decide(hipStream_t s, const uint64_t* cells, uint32_t n, uint32_t n_rows,
                              const rd_cfar_cfg& cfg, rd_cfar_decision* out);

// Element-wise fixed-point primitives (rd_fxpgate.hip).
hipError_t launch_fxp_eval(hipStream_t s, const rd_fxp_op* ops, uint64_t n, int64_t* y, uint32_t* flags,
                           uint32_t* forms);
hipError_t launch_fxp_accum(hipStream_t s, uint32_t acc_w, const int64_t* x, uint32_t n, int64_t* y,
                            uint32_t* step_flags, uint32_t* sticky, uint32_t* count);
hipError_t launch_fxp_cmult(hipStream_t s, const uint32_t* a, const uint32_t* b, uint64_t n, int64_t* p4,
                            int64_t* p3, uint32_t* y, uint8_t* flags);

// CFAR (rd_cfar.hip). scratch holds scratch_bytes: the row offsets
// ((n_rows + 1) * 8 bytes, required) then as many staged events per row as
// fit (rows with more are recomputed by the gather pass).
// cfar_scratch_bytes(n_bins, n_rows) is the size a context allocates.
size_t cfar_scratch_bytes(uint32_t n_bins, uint32_t n_rows);
uint32_t cfar_slot_for(uint32_t n_bins, uint32_t n_rows, size_t scratch_bytes);
hipError_t launch_cfar(hipStream_t s, const void* cells, uint32_t cell_bytes, uint32_t n_bins, uint32_t n_rows,
                       const rd_cfar_cfg& cfg, uint32_t max_guard, uint32_t max_ref, const uint32_t* frame_ids,
                       uint32_t frame_mod, uint64_t* events, uint64_t events_cap, uint32_t* row_count,
                       void* scratch, size_t scratch_bytes, uint64_t* total_dev);

// Detection packets (rd_packet.hip): n_cap events, or min(n_cap, *n_dev) when
// n_dev (device) is set.
hipError_t launch_event_flits(hipStream_t s, const uint64_t* events, uint64_t n_cap, const uint64_t* n_dev,
                              uint32_t packet_w, uint32_t dest, uint32_t src, uint32_t vc, uint32_t seq0,
                              uint32_t* flits);
// Detection writeback: n_cap packets (or min(n_cap, *n_dev)) of flits into
// 64-byte ring words, checks counted into stats[RD_WB_NSTATS].
hipError_t launch_event_writeback(hipStream_t s, const uint32_t* flits, uint64_t n_cap, const uint64_t* n_dev,
                                  uint32_t packet_w, uint32_t port, uint32_t seq0, uint32_t* ring,
                                  uint64_t ring_words, uint64_t addr0, uint64_t* stats);

// ---------------------------------------------------------------- batched
// A class launch: `jobs` (device) holds n_jobs records of one shape class,
// `tile_job` (device) maps each of n_tiles tiles to its job. The beamformer's
// `grid` blocks walk the tiles with a grid stride; the other stages launch one
// block per tile (grid = n_tiles).
struct ClassLaunch {
  const void* jobs;
  const uint32_t* tile_job;
  uint32_t n_tiles, grid;
};

// PFB, taps T (4, 8, 16 use the register-window kernel; others the generic one).
hipError_t launch_pfb_batch(hipStream_t s, uint32_t T, bool flags, const ClassLaunch& L);
// FFT of log2 length `log2n`, spc lanes; all1: every job's schedule has all
// bits set for its stages. Frames per tile follow the single-CPI launch.
uint32_t fft_frames_per_tile(uint32_t log2n, bool all1);
hipError_t launch_fft_batch(hipStream_t s, const uint32_t* tw4096, uint32_t log2n, uint32_t spc, bool all1,
                            const ClassLaunch& L);
// Beamformer over the FFT layout; steps = ceil(max CH / 16) of the class,
// tc = columns per tile (64, 128 or 256), max_ch / max_bt size the LDS.
// out_mask: the outputs every job of the class writes (1 y, 2 flags, 4 power,
// 8 accumulators), or -1 when they differ between jobs.
hipError_t launch_bf_batch(hipStream_t s, uint32_t steps, uint32_t tc, uint32_t max_ch, uint32_t max_bt,
                           int out_mask, const ClassLaunch& L);
// Covariance; chp = CH rounded up to 16 (shared by the class), fs_n frame splits.
uint32_t cov_bins_per_tile(uint32_t chp);
uint32_t cov_max_splits(uint32_t chp);
hipError_t launch_cov_batch(hipStream_t s, uint32_t chp, uint32_t fs_n, const ClassLaunch& L);
// Loader-fed forms for classes whose every CPI fits them: the beamformer of
// 48/64-channel CPIs with beams a multiple of 16 and columns a multiple of 128
// (tiles of 128 columns, X 16-byte aligned), on min(tiles, n_cu) resident
// blocks; the covariance of CPIs of exactly chp = 32 or 64 channels (tiles of
// 64 / chp bins, the bin count a multiple of that, Xa 16-byte aligned).
hipError_t launch_bf_ring_batch(hipStream_t s, uint32_t steps, uint32_t max_bt, uint32_t n_cu, int out_mask,
                                const ClassLaunch& L);
hipError_t launch_cov_ring_batch(hipStream_t s, uint32_t chp, const ClassLaunch& L);
// CFAR over power rows (u32 cells) of n bins: W waves per row (a power of two
// from cfar_min_waves(n) to 8), rpb rows per tile.
uint32_t cfar_min_waves(uint32_t n);
uint32_t cfar_rows_per_tile_max(uint32_t n, uint32_t W);
hipError_t launch_cfar_rows_batch(hipStream_t s, uint32_t n, uint32_t W, uint32_t rpb, const ClassLaunch& L);
hipError_t launch_cfar_scan_batch(hipStream_t s, const void* jobs, uint32_t n_jobs, uint32_t max_rows);
hipError_t launch_cfar_gather_batch(hipStream_t s, uint32_t n, uint32_t W, uint32_t rpb, const ClassLaunch& L);

}  // namespace rd
