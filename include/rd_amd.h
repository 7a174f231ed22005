/*
 * rd_amd.h — C ABI of the MI355X (gfx950) bit-exact radar chain.
 *
 * The drop-in boundary for the reference's per-stage C++ model calls
 * (model/cpp in fpga-professional-association/agilex-7-radar-detect). The
 * reference has no FFI of its own: its callers bind the header-only classes in
 * process. Each entry point below names the reference interface it replaces
 * (file:line) and computes the same values, bit for bit, batched over
 * channels / frames / rows.
 *
 * Conventions
 *   - Complex samples are packed {im, re} in one uint32, real part in the low
 *     half (model/cpp/fxp/fxp.hpp:288-297).
 *   - Saturation flags are the reference's 2-bit word {sat_pos<<1 | sat_neg}
 *     (fxp.hpp:146-153). Per-sample flag bytes hold f_re in bits [1:0] and
 *     f_im in bits [3:2].
 *   - Every data pointer is a DEVICE pointer owned by the caller (use
 *     rd_malloc / rd_h2d / rd_d2h when the caller has no HIP headers).
 *     Nullable outputs are skipped when NULL.
 *   - Work is queued on the context's stream; nothing is synchronous unless
 *     stated. Calls on one context must not overlap from several threads.
 *   - No exceptions cross the ABI; every call returns an rd_status.
 */
#ifndef RD_AMD_H_
#define RD_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  RD_OK = 0,
  RD_E_ARG = -1,      /* null pointer or size out of range */
  RD_E_GEOM = -2,     /* geometry the reference rejects (e.g. fft::Config::valid) */
  RD_E_HIP = -3,      /* a HIP runtime call failed */
  RD_E_CAPACITY = -4, /* event buffer too small (count still reported) */
  RD_E_NODEV = -5,    /* no such device */
  RD_E_SCRATCH = -6   /* input exceeds the static scratch of rd_chain_enqueue_static */
} rd_status;

typedef struct rd_ctx rd_ctx;

/* ---- Integrator control (covar::Integrator, model/cpp/covariance/covar_model.hpp:159-262)
 * covar::Config: window_len (0 counts as 1), mode 0 block / 1 exponential,
 * exp_k the exponential shift, enable. */
typedef struct {
  uint32_t window_len, mode, exp_k, enable;
} rd_integ_cfg;
/* The control inputs of one Integrator::step edge (cfg ports, valid_in, flush,
 * sat_clear), shared by every lane of a call. */
typedef struct {
  rd_integ_cfg cfg;
  uint32_t valid, flush, sat_clear;
} rd_integ_step;
/* covar::Result without the value (covar_model.hpp:173-184): the window a step
 * emitted, the same for every lane of a call. */
typedef struct {
  uint32_t step, window_id, sample_count, flushed, truncated;
} rd_integ_window;
/* The Integrator's registers that do not depend on the data (the active
 * configuration, the sample count, the window id): host memory, carried from
 * call to call like the model object. */
typedef struct {
  rd_integ_cfg active;
  uint32_t count, window_id;
} rd_integ_ctl;
/* One lane's data registers (device memory, carried from call to call): the
 * accumulator, its sticky saturation flags {sat_pos<<1 | sat_neg} and the
 * saturating count of flagged steps (Integrator::acc, sat_flags, sat_count). */
typedef struct {
  int64_t acc;
  uint32_t flags;
  uint32_t sat_count;
} rd_integ_lane;

/* One cfar::decide result (model/cpp/cfar/cfar_model.hpp:223-234). */
typedef struct {
  uint64_t cut_power, sum_lead, sum_lag, sel_sum;
  uint32_t sel_n;
  uint8_t window_ok, evaluable, detected, suppressed, take_lead;
  uint8_t pad[7];
} rd_cfar_decision;

/* CFAR configuration as written to the register plane
 * (model/cpp/cfar/cfar_model.hpp:112-128). mode: 0 CA, 1 GO.
 * out_mode: 0 events (DETECT + SUMMARY), 1 dense (one event per bin + SUMMARY).
 * alpha: UQ8.8 threshold multiplier. */
typedef struct {
  uint32_t enable, mode, out_mode, guard_lead, guard_lag, ref_lead, ref_lag, alpha;
} rd_cfar_cfg;

/* One coherent processing interval (CPI): geometry of the composed chain. */
typedef struct {
  uint32_t n_chan;      /* antennas / channels, 1..64                    */
  uint32_t n_fft;       /* PFB branches = FFT length, 8..4096, power of 2 */
  uint32_t taps;        /* PFB taps per branch, 1..64                    */
  uint32_t n_frames;    /* PFB beats = FFT frames                        */
  uint32_t n_beams;     /* beams, 1..128                                 */
  uint32_t spc;         /* FFT samples per cycle: 1 or 2                 */
  uint32_t scale_sched; /* FFT per-sub-stage shift bits (0xFFFFFFFF = all) */
  uint32_t max_guard;   /* CFAR elaboration maxima (clamp)               */
  uint32_t max_ref;
  rd_cfar_cfg cfar;
  uint32_t x_layout;    /* layout of bufs->X in the chain: 0 natural order
                         * [n_chan][n_frames][n_fft]; 1 bin-quad blocked
                         * [n_chan][n_fft/4][n_frames][4], which the fused
                         * beamformer + covariance reads as contiguous runs
                         * (faster); allowed only where the chain fuses them
                         * (n_chan <= 16, n_fft >= 16, R set), not in
                         * rd_chain_enqueue_plain or a batch (RD_E_ARG) */
} rd_cpi_desc;

/* Device buffers for one CPI (layouts: [a][b] is row-major, a slowest).
 *   iq        [n_chan][n_frames*n_fft]      input
 *   taps      [n_fft][taps]                 phase-major, tap-minor
 *   weights   [n_beams][n_chan]             beam-major
 *   u         [n_chan][n_frames][n_fft]     PFB output (required)
 *   pfb_flags [n_chan][n_frames][n_fft]     nullable
 *   X         [n_chan][n_frames][n_fft]     FFT bins, natural order (required); bin-quad blocked
 *             [n_chan][n_fft/4][n_frames][4] when desc->x_layout is 1
 *   fft_flags [n_chan][n_frames]            stage flags, sub-stage g at bits [2g+1:2g]; nullable
 *   Xa        [n_frames][n_fft][n_chan]     aligned (bin-major, antenna-minor); nullable in the
 *             single-CPI chain (written when set; when null the context keeps its own copy
 *             where a path needs one, and rd_chain_enqueue_static then needs n_chan <= 16 and
 *             n_fft a multiple of 4 or no covariance); required by rd_batch_create
 *   y         [n_beams][n_frames][n_fft]    beam samples; nullable
 *   yflags    [n_beams][n_frames][n_fft]    nullable
 *   power     [n_beams][n_frames][n_fft]    |y|^2 as uint32 (exact, <= 2^31) (required)
 *   R         [n_fft][n_chan*(n_chan+1)/2][2] int64 covariance re/im, pairs i<=j row-major; nullable
 *   events    [events_cap][3] uint64        packed 176-bit events, rows (beam, frame) in order
 *   row_count [n_beams*n_frames]            events per row (required)
 * u and X are 16-byte aligned.
 */
typedef struct {
  const uint32_t* iq;
  const uint32_t* taps;
  const uint32_t* weights;
  uint32_t* u;
  uint8_t* pfb_flags;
  uint32_t* X;
  uint32_t* fft_flags;
  uint32_t* Xa;
  uint32_t* y;
  uint8_t* yflags;
  uint32_t* power;
  int64_t* R;
  uint64_t* events;
  uint64_t events_cap;
  uint32_t* row_count;
} rd_cpi_bufs;

/* ---- context and memory ------------------------------------------------ */
int rd_device_count(void);
rd_status rd_create(int device, rd_ctx** out);
/* Fails with RD_E_ARG (and destroys nothing) while graphs of the context live. */
rd_status rd_destroy(rd_ctx* ctx);
const char* rd_status_str(rd_status s);
rd_status rd_sync(rd_ctx* ctx);
rd_status rd_malloc(rd_ctx* ctx, size_t bytes, void** dptr);
rd_status rd_free(rd_ctx* ctx, void* dptr);
rd_status rd_h2d(rd_ctx* ctx, void* dst, const void* src, size_t bytes);
rd_status rd_d2h(rd_ctx* ctx, void* dst, const void* src, size_t bytes);
rd_status rd_memset(rd_ctx* ctx, void* dst, int value, size_t bytes);
/* The context's HIP stream (a hipStream_t), for callers that time or order work. */
void* rd_stream(rd_ctx* ctx);

/* The 4096-entry Q1.15 twiddle table the FFT uses (host copy), generated by
 * the rule of model/python/gen_fft_twiddles.py:97-159 with M = 4096; every 4th
 * entry is the reference's committed 1024-entry table
 * (model/cpp/fft/fft_twiddle_table.hpp:42). */
rd_status rd_twiddle_table(uint32_t* host_out4096);

/* ---- stages -------------------------------------------------------------- */

/* Polyphase FIR bank. Replaces pfb::PfbModel(phases, taps, coeff) + reset() +
 * step(beat) per beat (model/cpp/pfb/pfb_model.hpp:207-251), over n_chan
 * independent channels: x [n_chan][n_beats*phases] -> y [n_chan][n_beats][phases].
 * hist (nullable) [n_chan][taps-1][phases], most recent beat first: seeds the
 * delay line and receives its final state, which carries the filter across calls
 * like the model's persistent history (pfb_model.hpp:179-182, 225-233).
 * NULL means the zero history of reset(). */
rd_status rd_pfb_run(rd_ctx* ctx, const uint32_t* x, const uint32_t* coeff, uint32_t phases, uint32_t taps,
                     uint32_t n_chan, uint32_t n_beats, uint32_t* hist, uint32_t* y, uint8_t* flags);

/* Streaming R2^2 FFT with per-sub-stage rounding. Replaces
 * fft::transform(Config{n_fft, spc, scale_sched, reorder}, in)
 * (model/cpp/fft/fft_ref.hpp:409-412) for n_frames frames of n_fft samples.
 * out: Result::out (beat order), bins: Result::bins (natural order),
 * stage_flags: Result::stage_flags packed 2 bits per sub-stage (fft_vectors.hpp:27-29).
 * in, out and bins are 16-byte aligned (out, bins nullable).
 * Extends the reference's N <= 1024 to N <= 4096 with the same twiddle rule. */
rd_status rd_fft_run(rd_ctx* ctx, const uint32_t* in, uint32_t n_fft, uint32_t spc, uint32_t scale_sched,
                     int reorder, uint32_t n_frames, uint32_t* out, uint32_t* bins, uint32_t* stage_flags);

/* The chain's FFT stage (fft::transform per frame, fft_ref.hpp:250-412): u
 * [n_chan][n_frames][n_fft] -> X in natural bin order (x_layout 0; the same
 * as rd_fft_run's `bins` output with reorder 1) or bin-quad blocked
 * [n_chan][n_fft/4][n_frames][4] (x_layout 1: bins 4j..4j+3 of every frame
 * of a channel in one run, what the fused beamformer + covariance reads).
 * stage_flags [n_chan * n_frames] nullable. u and X 16-byte aligned. */
rd_status rd_fft_stage(rd_ctx* ctx, const uint32_t* u, uint32_t n_fft, uint32_t spc, uint32_t scale_sched,
                       uint32_t n_chan, uint32_t n_frames, uint32_t x_layout, uint32_t* X, uint32_t* stage_flags);

/* Frequency-bin alignment (corner turn): X [n_chan][n_frames][n_bins] ->
 * Xa [n_frames][n_bins][n_chan], the beamformer's bin-major, antenna-minor beat
 * layout (model/cpp/align/align_model.hpp:172-186). Pure permutation. */
rd_status rd_align_run(rd_ctx* ctx, const uint32_t* X, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                       uint32_t* Xa);

/* Complex beamformer + power. Replaces bf::dot(x, w, n_ant) for every beam and
 * column (model/cpp/beamformer/beamformer_model.hpp:162-179) and
 * covar::power(y) (model/cpp/covariance/covar_model.hpp:99-102).
 * Xa [n_cols][n_ant], W [n_beams][n_ant] -> y/flags/power [n_beams][n_cols],
 * acc (nullable) [n_beams][n_cols][2] = DotResult::acc_re/acc_im.
 * Computed as four int8 MFMA partial products with exact int64 recombination. */
rd_status rd_bf_run(rd_ctx* ctx, const uint32_t* Xa, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                    uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power, int64_t* acc);

/* The chain's beamformer: the same values as rd_align_run followed by rd_bf_run,
 * read straight from the FFT output X [n_ant][n_frames][n_bins] (channel-major).
 * Writes y/flags/power [n_beams][n_frames*n_bins] (each nullable) and the aligned
 * copy Xa [n_frames][n_bins][n_ant] (nullable) from the same staged tiles. */
rd_status rd_bf_stage(rd_ctx* ctx, const uint32_t* X, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                      uint32_t n_frames, uint32_t n_bins, uint32_t* y, uint8_t* flags, uint32_t* power, uint32_t* Xa);

/* The chain's beamformer and covariance together: the values of rd_bf_stage
 * and of rd_cov_run on its aligned copy (BeamformerModel::step,
 * beamformer_model.hpp:145-325; covar::power and the Integrator's block-mode
 * window sum, covar_model.hpp:99-151, 186-262), read once from X
 * [n_ant][n_frames][n_bins]. One kernel for n_ant <= 16 with n_bins a
 * multiple of 4 (X 16-byte aligned, else RD_E_ARG; Xa nullable: written when
 * set); otherwise the two stages, and then Xa is required. R [n_bins][n_ant*(n_ant+1)/2][2] is required;
 * y/flags/power are nullable. x_layout: X in natural order (0) or bin-quad
 * blocked [n_ant][n_bins/4][n_frames][4] (1, the one-kernel form only; what
 * rd_fft_stage writes with x_layout 1, the chain's fast layout). */
rd_status rd_bf_cov_stage(rd_ctx* ctx, const uint32_t* X, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                          uint32_t n_frames, uint32_t n_bins, uint32_t* y, uint8_t* flags, uint32_t* power,
                          uint32_t* Xa, int64_t* R, uint32_t x_layout);

/* Covariance: R[bin][pair i<=j] = sum over frames of covar::cross_power(X_i, X_j)
 * (covar_model.hpp:137-151), integrated by a block-mode covar::Integrator with
 * acc_w = 64 and window_len = n_frames (covar_model.hpp:186-262); at that width
 * no step can saturate, so the window result is the exact int64 sum.
 * Xa [n_frames][n_bins][n_chan] -> R [n_bins][n_chan*(n_chan+1)/2][2]. */
rd_status rd_cov_run(rd_ctx* ctx, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                     int64_t* R);

/* Covariance through covar::Integrator frame by frame, either mode
 * (model/cpp/covariance/covar_model.hpp:196-214): mode 0 (block) adds the
 * cross power, mode 1 (exponential) moves the accumulator by
 * trunc((x - acc) >> exp_k) (floor division, fxp.hpp:185: the documented dead
 * band); both saturate to acc_w bits every frame with sticky flags. For
 * acc_w = 64 block mode equals rd_cov_run. R as rd_cov_run's; sat_flags
 * (nullable) [n_bins][pairs], re in bits [1:0], im in [3:2]. The exponential
 * accumulator is not cleared at a window close (covar_model.hpp:266-268), so R
 * is its state after the n_frames frames. exp_k <= 63, acc_w 2..64,
 * n_bins * n_chan * (n_chan + 1) / 2 <= 2^31. */
rd_status rd_cov_integrate(rd_ctx* ctx, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                           uint32_t mode, uint32_t exp_k, uint32_t acc_w, int64_t* R, uint8_t* sat_flags);

/* ---- fixed-point primitives, element-wise --------------------------------
 * The operations of model/cpp/fxp/fxp.hpp (sat :175-179, trunc :185,
 * round_even / round_half_up :191-215, round_sat :221-231, add/sub/neg_sat
 * :237-254, mul_q15 :261-263, cmul_q15 :303-325, growth_bits / acc_w
 * :334-352) on the device, by the device functions the stage kernels inline.
 * One record per operation; for the multiplies a and b carry int16 operands
 * in their low 16 bits (mul) or packed complex words in their low 32 bits
 * (cmul). y and flags ({sat_pos<<1 | sat_neg}, 0 where the operation cannot
 * saturate) per record; forms (nullable) per record: bit 0 set when one of the
 * kernels' specialised forms (16-bit saturation, shift-15 rounding in 32 or
 * 64 bits, the one-bit butterfly rounding) applies to the operands, bit 1 when
 * it disagrees with the general form. ops, y, flags, forms: device. */
enum {
  RD_FXP_SAT = 0, RD_FXP_TRUNC, RD_FXP_ROUND_EVEN, RD_FXP_ROUND_HALF_UP, RD_FXP_ROUND_SAT, RD_FXP_ADD_SAT,
  RD_FXP_SUB_SAT, RD_FXP_NEG_SAT, RD_FXP_MUL_Q15, RD_FXP_MUL_Q15_RS, RD_FXP_CMUL_RE, RD_FXP_CMUL_IM, RD_FXP_ACC_W,
  RD_FXP_GROWTH_BITS
};
typedef struct {
  int64_t a, b;
  uint32_t op, sh, w, pad;
} rd_fxp_op;
rd_status rd_fxp_eval(rd_ctx* ctx, const rd_fxp_op* ops, uint64_t n, int64_t* y, uint32_t* flags, uint32_t* forms);
/* fxp::Acc (fxp.hpp:392-418) from a clear over x[0..n): per step the
 * saturated accumulator, the step's flags, the sticky flags and the saturating
 * count of flagged steps (StickyFlags, fxp.hpp:358-381). Device arrays. */
rd_status rd_fxp_accum(rd_ctx* ctx, uint32_t acc_w, const int64_t* x, uint32_t n, int64_t* y, uint32_t* step_flags,
                       uint32_t* sticky, uint32_t* count);
/* fxp::cmult::raw_mult4 / raw_mult3 / eval (model/cpp/fxp/cmult.hpp:85-168):
 * the exact products [n][2] (re, im) of both forms, the rounded Q1.15 product
 * and its flags (f_re | f_im << 2). Device arrays. */
rd_status rd_fxp_cmult(rd_ctx* ctx, const uint32_t* a, const uint32_t* b, uint64_t n, int64_t* p4, int64_t* p3,
                       uint32_t* y, uint8_t* flags);

/* ---- the Integrator over a control stream --------------------------------
 * Many covar::Integrator(acc_w) instances ("lanes") stepped together:
 * Integrator::reset(cfg) is rd_integ_reset(cfg, &ctl) plus zeroed lanes, and
 * n_steps calls of Integrator::step(steps[s].cfg, steps[s].valid, data,
 * steps[s].flush, steps[s].sat_clear, &result) (covar_model.hpp:191-262) are
 * one call, every lane with its own data and the same control inputs.
 * Each step adds (block) or moves by floor((x - acc) / 2^exp_k) (exponential),
 * saturates to acc_w bits with sticky flags, closes a window when its sample
 * count reaches window_len, emits a result on a close or a flush of a
 * non-empty window, clears the sticky flags on sat_clear or flush, and
 * latches a new configuration only at a window boundary. The emitted windows
 * (the same for every lane) are written to windows[0 .. min(n, cap)) on the
 * host and their values to results [cap][n_lanes] on the device; *n_windows
 * receives n (RD_E_CAPACITY when n > cap, values past cap not written). ctl
 * and lanes are updated in place, so consecutive calls continue the stream.
 * Synchronous. acc_w <= 64 (below 2 every value saturates to 0, as
 * fxp::max_of/min_of do), mode 0 or 1. */
rd_status rd_integ_reset(const rd_integ_cfg* cfg, rd_integ_ctl* ctl);
/* data [n_steps][n_lanes] int64 (device), lanes [n_lanes] (device). */
rd_status rd_integ_run(rd_ctx* ctx, uint32_t acc_w, rd_integ_ctl* ctl, const rd_integ_step* steps, uint32_t n_steps,
                       const int64_t* data, uint64_t n_lanes, rd_integ_lane* lanes, int64_t* results,
                       rd_integ_window* windows, uint32_t windows_cap, uint32_t* n_windows);
/* The lanes of a covariance (covar::Engine's accumulators,
 * covar_model.hpp:318-416): lane ((k * pairs + p) * 2 + c) integrates
 * component c (0 re, 1 im) of covar::cross_power(X_i, X_j) of bin k, pair
 * p = (i <= j) row-major; step m is frame m of Xa [n_frames][n_bins][n_chan].
 * lanes [n_bins][pairs][2], results [windows_cap][n_bins][pairs][2] (device).
 * With steps all valid, no flush or clear and window_len = n_frames, the
 * one window's result is rd_cov_run's R (at acc_w = 64). */
rd_status rd_cov_integrate_windows(rd_ctx* ctx, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames,
                                   uint32_t n_bins, uint32_t acc_w, rd_integ_ctl* ctl, const rd_integ_step* steps,
                                   rd_integ_lane* lanes, int64_t* results, rd_integ_window* windows,
                                   uint32_t windows_cap, uint32_t* n_windows);

/* pfb::LaneModel::accumulate (model/cpp/pfb/pfb_model.hpp:135-150) for every
 * output of rd_pfb_run: the exact int64 sums of cmul_q15_re_raw and
 * cmul_q15_im_raw over the taps, before rounding. acc [n_chan][n_beats]
 * [phases][2]; hist as rd_pfb_run's (read only, nullable). */
rd_status rd_pfb_acc_run(rd_ctx* ctx, const uint32_t* x, const uint32_t* coeff, uint32_t phases, uint32_t taps,
                         uint32_t n_chan, uint32_t n_beats, const uint32_t* hist, int64_t* acc);

/* cfar::decide(bins, i, cfg) (model/cpp/cfar/cfar_model.hpp:243-285) for every
 * bin i of every row: cells [n_rows][n_bins] uint64 magnitudes, cfg used as
 * given (no clamp: cfar::run_frame clamps before it calls decide), out
 * [n_rows][n_bins], including the partial window sums of bins whose window
 * leaves the row. */
rd_status rd_cfar_decide(rd_ctx* ctx, const uint64_t* cells, uint32_t n_bins, uint32_t n_rows,
                         const rd_cfar_cfg* cfg, rd_cfar_decision* out);

/* CA/GO-CFAR over rows of n_bins cells. Replaces cfar::run_frame(bins, cfg,
 * max_guard, max_ref, frame_id) per row (model/cpp/cfar/cfar_model.hpp:295-340),
 * with the events packed by the 176-bit layout of cfar_model.hpp:75-84 into three
 * little-endian uint64 words (sof at bit 176, eof at bit 177).
 * cells: uint32 (cell_bytes = 4) or uint64 (cell_bytes = 8) magnitudes.
 * frame_id of row r = frame_ids[r] when frame_ids != NULL, else r % frame_mod
 * (frame_mod = 0: r). row_count (device, n_rows) receives the events per row;
 * *n_events_host (synchronous) the total. Returns RD_E_CAPACITY if it exceeds
 * events_cap (no events are then written past the cap). */
rd_status rd_cfar_run(rd_ctx* ctx, const void* cells, uint32_t cell_bytes, uint32_t n_bins, uint32_t n_rows,
                      const rd_cfar_cfg* cfg, uint32_t max_guard, uint32_t max_ref, const uint32_t* frame_ids,
                      uint32_t frame_mod, uint64_t* events, uint64_t events_cap, uint32_t* row_count,
                      uint64_t* n_events_host);
/* rd_cfar_run without the final synchronisation: the event total is written
 * to device memory at *n_events_dev (uint64, required; 0 when n_rows is 0)
 * and the call returns once the work is queued. The capacity is not checked
 * on the host: a total above events_cap means the events past the cap were
 * not written. */
rd_status rd_cfar_enqueue(rd_ctx* ctx, const void* cells, uint32_t cell_bytes, uint32_t n_bins, uint32_t n_rows,
                          const rd_cfar_cfg* cfg, uint32_t max_guard, uint32_t max_ref, const uint32_t* frame_ids,
                          uint32_t frame_mod, uint64_t* events, uint64_t events_cap, uint32_t* row_count,
                          uint64_t* n_events_dev);

/* The composed chain for one CPI: PFB -> FFT -> alignment -> beamformer+power ->
 * covariance -> CFAR -> event packing, every stage as above, all on the
 * context's stream. *n_events_host (synchronous) receives the event total.
 * With at most 16 channels, n_fft a multiple of 4 and R set, the beamformer
 * and the covariance are one kernel reading X once (the aligned copy Xa is
 * then only an optional output); otherwise the beamformer writes Xa and the
 * covariance reads it. Results are the same bits either way. */
rd_status rd_chain_run(rd_ctx* ctx, const rd_cpi_desc* desc, const rd_cpi_bufs* bufs, uint64_t* n_events_host);

/* Same as rd_chain_run without the final synchronisation: the event total is
 * left in device memory at *n_events_dev (uint64). For timing loops. */
rd_status rd_chain_enqueue(rd_ctx* ctx, const rd_cpi_desc* desc, const rd_cpi_bufs* bufs, uint64_t* n_events_dev);

/* Allocation-free variant of rd_chain_enqueue, for callers that cannot
 * allocate device memory during the call: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc. No counterpart in the
 * reference (its models allocate freely with std::vector).
 * Same arguments, launches and results as rd_chain_enqueue, but the CFAR
 * pass keeps its row offsets and staged events in the library's static
 * scratch array (RD_STATIC_SCRATCH_BYTES per GPU, held until the process
 * ends) instead of the context's growable buffer, and the call allocates
 * nothing. Limit: (n_beams * n_frames + 1) * 8 <= RD_STATIC_SCRATCH_BYTES,
 * i.e. at most RD_STATIC_MAX_ROWS CFAR rows; larger CPIs get RD_E_SCRATCH.
 * Every call on one GPU shares that GPU's array (all contexts on the device):
 * calls whose work overlaps in time on one GPU race, so serialise them (one
 * stream, or events between streams). Nothing in the array needs to start at
 * zero; each call rewrites what it reads. */
#define RD_STATIC_SCRATCH_BYTES (32u << 20)
#define RD_STATIC_MAX_ROWS (RD_STATIC_SCRATCH_BYTES / 8u - 1u)
rd_status rd_chain_enqueue_static(rd_ctx* ctx, const rd_cpi_desc* desc, const rd_cpi_bufs* bufs,
                                  uint64_t* n_events_dev);

/* ---- plain (no matrix cores) contractions --------------------------------
 * The "plain GPU" comparison point of the chain's measurement. The same values,
 * bit for bit, as rd_bf_run / rd_cov_run, computed without MFMA: each
 * int16 x int16 product in a 32-bit multiply, summed in int64, rounded once,
 * tiles staged in LDS. They replace the same reference calls: bf::dot per beam
 * and column plus covar::power (model/cpp/beamformer/beamformer_model.hpp:162-179,
 * model/cpp/covariance/covar_model.hpp:99-102), and covar::cross_power summed
 * by an acc_w = 64 block-mode Integrator (covar_model.hpp:137-151, 186-262).
 * rd_bf_plain_run: Xa [n_cols][n_ant] -> y/flags/power [n_beams][n_cols] (each
 * nullable). rd_cov_plain_run: Xa [n_frames][n_bins][n_chan] -> R as rd_cov_run.
 * rd_chain_enqueue_plain: rd_chain_enqueue's chain with the stand-alone
 * alignment (rd_align_run) and these two kernels in place of the fused MFMA
 * beamformer and the MFMA covariance; same arguments and results. */
rd_status rd_bf_plain_run(rd_ctx* ctx, const uint32_t* Xa, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                          uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power);
rd_status rd_cov_plain_run(rd_ctx* ctx, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                           int64_t* R);
rd_status rd_chain_enqueue_plain(rd_ctx* ctx, const rd_cpi_desc* desc, const rd_cpi_bufs* bufs,
                                 uint64_t* n_events_dev);

/* ---- graphs --------------------------------------------------------------- */
typedef struct rd_graph rd_graph;

/* Records the launches of rd_chain_enqueue(ctx, desc, bufs, n_events_dev) into an
 * executable HIP graph (buffers and geometry are fixed at capture). Each
 * rd_graph_launch then replays the whole chain with one launch on the context's
 * stream. No graph node is a memset.
 * Lifetime: the graph holds the raw device pointers of bufs (and
 * n_events_dev): every one of those buffers must stay allocated until
 * rd_graph_destroy returns. Freeing one while the graph lives and launching
 * it again writes to freed memory. */
rd_status rd_chain_graph(rd_ctx* ctx, const rd_cpi_desc* desc, const rd_cpi_bufs* bufs, uint64_t* n_events_dev,
                         rd_graph** out);
rd_status rd_graph_launch(rd_graph* graph);
rd_status rd_graph_destroy(rd_graph* graph);

/* Slow-time (Doppler) transform: for every row (channel or beam) and range
 * bin, the bit-exact FFT of fft::transform (model/cpp/fft/fft_ref.hpp:409-412,
 * Config{n_dop, spc, scale_sched}, natural-order bins) across the last n_dop
 * frames, the range-Doppler map of model/python/range_doppler.py:254-265 in
 * the chain's fixed-point arithmetic. The frames sit in a ring
 * ring [rows][depth][n_bins] (frame f in slot f mod depth, hist::slot_of,
 * model/cpp/history/history_model.hpp:127-129), the newest in slot `newest`;
 * a plain [rows][frames][n_bins] array is the ring with depth = frames,
 * newest = frames - 1. out [rows][n_dop][n_bins]; with `shift` the Doppler
 * axis is fftshifted (bin k at row (k + n_dop/2) mod n_dop, range_doppler.py:
 * 262-265). stage_flags (nullable) [rows][n_bins], packed as rd_fft_run's.
 * n_dop: power of two, 8..4096, n_dop/spc >= 4, n_dop <= depth; rows <= 65535
 * and rows * n_bins < 2^32 (RD_E_ARG otherwise). */
rd_status rd_doppler_run(rd_ctx* ctx, const uint32_t* ring, uint32_t rows, uint32_t depth, uint32_t n_bins,
                         uint32_t newest, uint32_t n_dop, uint32_t spc, uint32_t scale_sched, int shift, uint32_t* out,
                         uint32_t* stage_flags);

/* Detection packets for the packet fabric. Replaces, per event,
 * packet_model::hdr_pack (model/cpp/packet/packet_model.hpp:150-160) and
 * make_flit (:241-260) for a detection packet of detection_flits(packet_w)
 * flits (:116-127): a header flit {dest 4, src 5, vc 2, type 0, length,
 * seq = (seq0 + e) mod 2^16} then the 176-bit event LSB first in packet_w-bit
 * slices; each flit has odd parity, vc, sof, eof in bits 0..4 and the data
 * from bit 5. events [n][3] packed as rd_cfar_run writes them (bits 176-177,
 * sof/eof of the event stream, are not payload). flits
 * [n][rd_event_flit_count(packet_w)][rd_flit_words(packet_w)] u32, the
 * reference's Flit::w words. n_events_dev (nullable, device): the count is
 * min(n_events, *n_events_dev), e.g. the chain's event total, so the call can
 * follow rd_chain_enqueue in the same stream (and graph) without a host round
 * trip. packet_w in [36, 512]; dest < 16, src < 32, vc < 4;
 * n_events * rd_event_flit_count(packet_w) <= 2^31. */
uint32_t rd_event_flit_count(uint32_t packet_w);
uint32_t rd_flit_words(uint32_t packet_w);
rd_status rd_event_flits(rd_ctx* ctx, const uint64_t* events, uint64_t n_events, const uint64_t* n_events_dev,
                         uint32_t packet_w, uint32_t dest, uint32_t src, uint32_t vc, uint32_t seq0, uint32_t* flits);

/* Detection writeback: the receive end of the detection packets, as the
 * pipeline's DMA tap does it (deserializer and memory writer,
 * rtl/top/benchmark_sim_top.sv:1994-2134; the C++ model has none). Packet k of
 * n (flits [n][rd_event_flit_count(packet_w)][rd_flit_words(packet_w)] in
 * rd_event_flits' layout; n = n_packets, or min(n_packets, *n_packets_dev)
 * when that is set) is reassembled and its 176-bit event written zero-extended
 * into one 64-byte word of a ring of ring_words words: word
 * (addr0 + k) mod ring_words, ring 16-byte aligned (with more packets than
 * ring words the ring ends as in-order writes leave it, holding the last
 * ring_words). Every packet lands, as the tap's writer checks nothing; the
 * receive checks of packet_model::Scoreboard::observe
 * (model/cpp/packet/packet_model.hpp:360-429), taken per packet on its own L
 * flits, are added to stats
 * [RD_WB_NSTATS] (device u64; added to, so zero them once):
 *   [0] packets written       [1] flits failing odd parity
 *   [2] framing violations    [3] misrouted (header dest != port)
 *   [4] sequence breaks (header seq != (seq0 + k) mod 2^16).
 * Framing violations: header flit without sof or with eof, header length !=
 * L, type != detection, header vc != flit vc; a payload flit with sof, with a
 * vc other than the header flit's, or whose eof is not set exactly on the
 * last flit (one each). packet_w in [36, 512], port < 16, ring_words in
 * [1, 2^48], n_packets <= 2^31. */
#define RD_WB_NSTATS 5
rd_status rd_event_writeback(rd_ctx* ctx, const uint32_t* flits, uint64_t n_packets, const uint64_t* n_packets_dev,
                             uint32_t packet_w, uint32_t port, uint32_t seq0, uint32_t* ring, uint64_t ring_words,
                             uint64_t addr0, uint64_t* stats);

/* ---- reference file formats (host side, no device work) -----------------
 * The readers validate as the reference's loaders do and return RD_E_ARG with
 * a diagnosis in err (nullable, err_cap bytes) on any mismatch. Called with a
 * null or too small output they fill in the sizes and return RD_E_CAPACITY, so
 * a caller can size its buffers first. */

/* .coeff, pfb::load_coeff (model/cpp/pfb/pfb_model.hpp:357-442): schema 1,
 * kind pfb_coeff, count = rows = phases*taps, rows in index order with
 * index = phase*taps + tap. coeff_out [phases*taps] packed, phase-major (the
 * layout rd_pfb_run takes). */
rd_status rd_read_coeff(const char* path, uint32_t* coeff_out, size_t cap, uint32_t* phases, uint32_t* taps,
                        char* err, size_t err_cap);
/* .vec, pfb::load_io (pfb_model.hpp:444-544): per beat the stimulus x, the
 * expected y and the flag words, [beats][phases] each (x/y packed). */
rd_status rd_read_pfb_vec(const char* path, uint32_t* x, uint32_t* y, uint8_t* f_re, uint8_t* f_im,
                          size_t cap_beats, uint32_t* phases, uint32_t* taps, uint32_t* beats, char* err,
                          size_t err_cap);
/* iq_ascii_v1 scenario files (writer model/python/gen_target_iq.py:520-617,
 * reader sim/verilator/harness/iq_loader.h): one row per (frame, antenna, beat)
 * "frame ant beat sof eof s0_re s0_im ..." with SAMPLES_PER_CYCLE samples per
 * beat; every row present once, sof/eof on the first/last beat of a frame.
 * iq [n_ant][frames*fft_size] packed, the chain's input layout. */
rd_status rd_read_iq_ascii(const char* path, uint32_t* iq, size_t cap, uint32_t* n_ant, uint32_t* fft_size,
                           uint32_t* frames, uint32_t* spc, char* err, size_t err_cap);
rd_status rd_write_iq_ascii(const char* path, const uint32_t* iq, uint32_t n_ant, uint32_t fft_size, uint32_t frames,
                            uint32_t spc, const char* scenario);

/* ---------------------------------------------------------------- batches
 * No counterpart in the reference: its models run one stream of frames at a
 * time (pfb::PfbModel::step, fft::transform per frame, cfar::run_frame per
 * row). A batch runs many CPIs of uneven geometry (channels, FFT length, taps,
 * frames, beams, CFAR settings) through the chain of rd_chain_enqueue with one
 * graph launch: the CPIs are grouped into shape classes per stage, and each
 * class runs as one persistent kernel over the tiles of all its CPIs. Every
 * CPI's outputs are identical to rd_chain_enqueue on the same desc and bufs.
 *
 * rd_batch_create validates every CPI (as rd_chain_enqueue), builds the device
 * job tables and the graph. With RD_BATCH_AUTOTUNE it first times candidate
 * tile and grid shapes per class on the given buffers (running the chain's
 * stages on them, so inputs must be loaded) and keeps the fastest; otherwise
 * it uses fixed heuristics. The buffers of different CPIs must not overlap.
 * Lifetime: the batch's graph and job tables hold the raw device pointers of
 * every CPI's buffers: they must stay allocated until rd_batch_destroy
 * returns (refill them in place between launches instead). */
typedef struct rd_batch rd_batch;

#define RD_BATCH_AUTOTUNE 1u

rd_status rd_batch_create(rd_ctx* ctx, uint32_t n_cpi, const rd_cpi_desc* descs, const rd_cpi_bufs* bufs,
                          uint32_t flags, rd_batch** out);
/* Enqueues the whole batch (one graph launch) on the context's stream. */
rd_status rd_batch_launch(rd_batch* batch);
/* Waits for the context's stream and copies the per-CPI event totals
 * (n_cpi values); RD_E_CAPACITY if any CPI's total exceeds its events_cap. */
rd_status rd_batch_events(rd_batch* batch, uint64_t* n_events_host);
/* The shape classes and the tile/grid shape chosen for each, as text. */
rd_status rd_batch_plan(const rd_batch* batch, char* text, size_t cap);
rd_status rd_batch_destroy(rd_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* RD_AMD_H_ */
