// rd_amd.hpp — C++ host mirror of the reference's per-stage model API over the
// gfx950 C ABI (rd_amd.h).
//
// Same namespaces, class and function names, argument meaning and result types
// as the reference's header-only models, so a reference-style harness compiles
// against either implementation by switching the include and the namespace
// prefix (rd_amd::pfb::PfbModel for pfb::PfbModel, ...). Every call computes on
// the GPU through the C ABI; the values are bit-identical to the reference's.
//
//   pfb::PfbModel           model/cpp/pfb/pfb_model.hpp:207-257
//   fft::Config/Result/transform   model/cpp/fft/fft_ref.hpp:199-240, 409-412
//   bf::dot, BeamformerModel       model/cpp/beamformer/beamformer_model.hpp:145-325
//   covar::power/cross_power       model/cpp/covariance/covar_model.hpp:99-151
//   cfar::Config/Event/run_frame   model/cpp/cfar/cfar_model.hpp:112-340
//
// Errors: the reference asserts nothing and returns values; here a failing
// runtime call throws rd_amd::Error (host side only — nothing throws across the
// C ABI). Per-call device allocation keeps the mirror simple; batched callers
// use the C ABI directly.
#ifndef RD_AMD_HPP_
#define RD_AMD_HPP_

#include <algorithm>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "rd_amd.h"

namespace rd_amd {

struct Error : std::runtime_error {
  rd_status code;
  Error(const std::string& what, rd_status c) : std::runtime_error(what + ": " + rd_status_str(c)), code(c) {}
};

inline void check(rd_status s, const char* what) {
  if (s != RD_OK) throw Error(what, s);
}

// Packed {im, re}, real low — fxp::Complex (model/cpp/fxp/fxp.hpp:279-298).
struct Complex {
  int16_t re = 0;
  int16_t im = 0;
  bool operator==(const Complex& o) const { return re == o.re && im == o.im; }
  bool operator!=(const Complex& o) const { return !(*this == o); }
  uint32_t packed() const { return ((uint32_t)(uint16_t)im << 16) | (uint32_t)(uint16_t)re; }
  static Complex from_packed(uint32_t p) { return Complex{(int16_t)(uint16_t)(p & 0xFFFFu), (int16_t)(uint16_t)(p >> 16)}; }
};

// fxp::Flags (fxp.hpp:136-154).
struct Flags {
  bool sat_pos = false;
  bool sat_neg = false;
  bool any() const { return sat_pos || sat_neg; }
  bool operator==(const Flags& o) const { return sat_pos == o.sat_pos && sat_neg == o.sat_neg; }
  unsigned packed() const { return (sat_pos ? 2u : 0u) | (sat_neg ? 1u : 0u); }
  static Flags from_packed(unsigned p) { return Flags{(p & 2u) != 0, (p & 1u) != 0}; }
};

// One shared context per process, created on first use: device 0 unless
// rd_amd::Device::use() picked another before that. Asking for a different
// device once the context exists throws (the mirror never moves devices).
class Device {
 public:
  static rd_ctx* ctx() {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.ctx) check(rd_create(s.device, &s.ctx), "rd_create");
    return s.ctx;
  }
  static void use(int device) {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.ctx && device != s.device)
      throw Error("rd_amd::Device::use(" + std::to_string(device) + ") after the context was created on device " +
                      std::to_string(s.device),
                  RD_E_ARG);
    s.device = device;
  }

 private:
  struct State {
    std::mutex mu;
    int device = 0;
    rd_ctx* ctx = nullptr;
    ~State() {
      if (ctx) rd_destroy(ctx);
    }
  };
  static State& state() {
    static State s;
    return s;
  }
};

// Device buffer with host copies.
template <typename T>
class DevVec {
 public:
  explicit DevVec(size_t n) : n_(n) {
    void* p = nullptr;
    check(rd_malloc(Device::ctx(), n * sizeof(T) + 16, &p), "rd_malloc");
    p_ = p;
  }
  DevVec(const T* src, size_t n) : DevVec(n) { put(src, n); }
  ~DevVec() { rd_free(Device::ctx(), p_); }
  DevVec(const DevVec&) = delete;
  DevVec& operator=(const DevVec&) = delete;
  void put(const T* src, size_t n) { check(rd_h2d(Device::ctx(), p_, src, n * sizeof(T)), "rd_h2d"); }
  std::vector<T> get(size_t n) const {
    std::vector<T> v(n);
    if (n) check(rd_d2h(Device::ctx(), v.data(), p_, n * sizeof(T)), "rd_d2h");
    return v;
  }
  T* ptr() const { return static_cast<T*>(p_); }
  size_t size() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

inline std::vector<uint32_t> pack_all(const std::vector<Complex>& v) {
  std::vector<uint32_t> o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = v[i].packed();
  return o;
}
inline std::vector<Complex> unpack_all(const std::vector<uint32_t>& v) {
  std::vector<Complex> o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = Complex::from_packed(v[i]);
  return o;
}

// ---------------------------------------------------------------------------
namespace pfb {

inline size_t coeff_index(unsigned taps, unsigned phase, unsigned tap) { return (size_t)phase * taps + tap; }

struct BeatResult {
  std::vector<Complex> y;
  std::vector<Flags> flags_re;
  std::vector<Flags> flags_im;
  Flags merged() const {
    Flags m;
    for (size_t i = 0; i < flags_re.size(); ++i) {
      m.sat_pos = m.sat_pos || flags_re[i].sat_pos || flags_im[i].sat_pos;
      m.sat_neg = m.sat_neg || flags_re[i].sat_neg || flags_im[i].sat_neg;
    }
    return m;
  }
};

// pfb::PfbModel: the delay line lives on the device and persists across step()
// calls and coefficient swaps; reset() zeroes it.
class PfbModel {
 public:
  PfbModel(unsigned phases, unsigned taps, const std::vector<Complex>& coeff)
      : phases_(phases), taps_(taps), hist_(std::max<size_t>(1, (size_t)(taps > 1 ? taps - 1 : 0) * phases)) {
    set_coeff(coeff);
    reset();
  }
  void set_coeff(const std::vector<Complex>& coeff) {
    const std::vector<uint32_t> c = pack_all(coeff);
    coeff_ = std::make_unique<DevVec<uint32_t>>(c.data(), (size_t)phases_ * taps_);
  }
  void swap_coeff(const std::vector<Complex>& coeff) { set_coeff(coeff); }
  void reset() {
    const std::vector<uint32_t> z(hist_.size(), 0u);
    hist_.put(z.data(), z.size());
  }
  unsigned phases() const { return phases_; }
  unsigned taps() const { return taps_; }

  // Consumes `n_beats` beats (n_beats * phases samples) in one launch.
  std::vector<BeatResult> run(const std::vector<Complex>& beats, unsigned n_beats) {
    const size_t S = (size_t)n_beats * phases_;
    const std::vector<uint32_t> x = pack_all(beats);
    DevVec<uint32_t> dx(x.data(), S), dy(S);
    DevVec<uint8_t> df(S);
    check(rd_pfb_run(Device::ctx(), dx.ptr(), coeff_->ptr(), phases_, taps_, 1, n_beats,
                     taps_ > 1 ? hist_.ptr() : nullptr, dy.ptr(), df.ptr()),
          "rd_pfb_run");
    const std::vector<uint32_t> y = dy.get(S);
    const std::vector<uint8_t> f = df.get(S);
    std::vector<BeatResult> out(n_beats);
    for (unsigned b = 0; b < n_beats; ++b) {
      BeatResult& r = out[b];
      r.y.resize(phases_);
      r.flags_re.resize(phases_);
      r.flags_im.resize(phases_);
      for (unsigned p = 0; p < phases_; ++p) {
        const size_t i = (size_t)b * phases_ + p;
        r.y[p] = Complex::from_packed(y[i]);
        r.flags_re[p] = Flags::from_packed(f[i] & 3u);
        r.flags_im[p] = Flags::from_packed((f[i] >> 2) & 3u);
      }
    }
    return out;
  }
  BeatResult step(const std::vector<Complex>& beat) { return run(beat, 1)[0]; }

 private:
  unsigned phases_;
  unsigned taps_;
  std::unique_ptr<DevVec<uint32_t>> coeff_;
  DevVec<uint32_t> hist_;
};

// pfb::LaneModel (pfb_model.hpp:117-183): one branch. The history stays on the
// host (most recent sample first, taps - 1 entries) and travels with each call.
class LaneModel {
 public:
  LaneModel() = default;
  LaneModel(unsigned taps, const Complex* coeff)
      : taps_(taps), coeff_(coeff, coeff + taps), history_(taps ? taps - 1 : 0, Complex{}) {}
  void reset() { history_.assign(history_.size(), Complex{}); }
  void set_coeff(const Complex* coeff) { coeff_.assign(coeff, coeff + taps_); }
  unsigned taps() const { return taps_; }
  const std::vector<Complex>& coeff() const { return coeff_; }

  // The exact accumulators of one sample, before rounding (rd_pfb_acc_run).
  void accumulate(Complex x, int64_t* acc_re, int64_t* acc_im) const {
    const uint32_t xp = x.packed();
    const std::vector<uint32_t> c = pack_all(coeff_), h = pack_all(history_);
    DevVec<uint32_t> dx(&xp, 1), dc(c.data(), c.size()), dh(h.data(), h.size());
    DevVec<int64_t> da(2);
    check(rd_pfb_acc_run(Device::ctx(), dx.ptr(), dc.ptr(), 1, taps_, 1, 1, h.empty() ? nullptr : dh.ptr(), da.ptr()),
          "rd_pfb_acc_run");
    const std::vector<int64_t> a = da.get(2);
    *acc_re = a[0];
    *acc_im = a[1];
  }

  // One sample through the branch (rd_pfb_run with the history imported and
  // exported); the history advances after the evaluation.
  Complex step(Complex x, Flags* f_re = nullptr, Flags* f_im = nullptr) {
    const uint32_t xp = x.packed();
    const std::vector<uint32_t> c = pack_all(coeff_), h = pack_all(history_);
    DevVec<uint32_t> dx(&xp, 1), dc(c.data(), c.size()), dh(h.data(), h.size()), dy(1);
    DevVec<uint8_t> df(1);
    check(rd_pfb_run(Device::ctx(), dx.ptr(), dc.ptr(), 1, taps_, 1, 1, h.empty() ? nullptr : dh.ptr(), dy.ptr(),
                     df.ptr()),
          "rd_pfb_run");
    if (!h.empty()) history_ = unpack_all(dh.get(h.size()));
    const uint8_t f = df.get(1)[0];
    if (f_re) *f_re = Flags::from_packed(f & 3u);
    if (f_im) *f_im = Flags::from_packed((f >> 2) & 3u);
    return Complex::from_packed(dy.get(1)[0]);
  }

 private:
  unsigned taps_ = 0;
  std::vector<Complex> coeff_;
  std::vector<Complex> history_;  // history_[k-1] == x[n-k]
};

}  // namespace pfb

// ---------------------------------------------------------------------------
namespace fft {

inline constexpr unsigned kSchedAll = 0xFFFFFFFFu;

struct Config {
  unsigned n_fft = 64;
  unsigned spc = 2;
  unsigned scale_sched = kSchedAll;
  bool reorder = true;
  unsigned lane() const { return n_fft / spc; }
  unsigned stages() const {
    unsigned r = 0;
    while ((1u << r) < n_fft) ++r;
    return r;
  }
  // The reference's rule (fft_ref.hpp:216-219) with the table extended to 4096.
  bool valid() const {
    const bool p2 = n_fft && !(n_fft & (n_fft - 1));
    return p2 && (spc == 1 || spc == 2) && n_fft >= 8 && n_fft <= 4096 && n_fft / spc >= 4;
  }
};

struct Result {
  std::vector<Complex> out;
  std::vector<Flags> stage_flags;
  std::vector<Complex> bins;
};

// Transforms `frames` consecutive frames of in (frames * n_fft samples).
inline std::vector<Result> transform_batch(const Config& cfg, const std::vector<Complex>& in, unsigned frames) {
  const size_t S = (size_t)frames * cfg.n_fft;
  const std::vector<uint32_t> x = pack_all(in);
  DevVec<uint32_t> dx(x.data(), S), dout(S), dbins(S), dfl(frames ? frames : 1);
  check(rd_fft_run(Device::ctx(), dx.ptr(), cfg.n_fft, cfg.spc, cfg.scale_sched, cfg.reorder ? 1 : 0, frames,
                   dout.ptr(), dbins.ptr(), dfl.ptr()),
        "rd_fft_run");
  const std::vector<uint32_t> o = dout.get(S), b = dbins.get(S), f = dfl.get(frames);
  std::vector<Result> res(frames);
  for (unsigned fr = 0; fr < frames; ++fr) {
    Result& r = res[fr];
    r.out.resize(cfg.n_fft);
    r.bins.resize(cfg.n_fft);
    for (unsigned i = 0; i < cfg.n_fft; ++i) {
      r.out[i] = Complex::from_packed(o[(size_t)fr * cfg.n_fft + i]);
      r.bins[i] = Complex::from_packed(b[(size_t)fr * cfg.n_fft + i]);
    }
    for (unsigned g = 0; g < cfg.stages(); ++g) r.stage_flags.push_back(Flags::from_packed((f[fr] >> (2 * g)) & 3u));
  }
  return res;
}

inline Result transform(const Config& cfg, const std::vector<Complex>& in) { return transform_batch(cfg, in, 1)[0]; }

}  // namespace fft

// ---------------------------------------------------------------------------
namespace bf {

inline size_t weight_index(unsigned n_ant, unsigned beam, unsigned ant) { return (size_t)beam * n_ant + ant; }

struct DotResult {
  Complex y;
  Flags f_re;
  Flags f_im;
  int64_t acc_re = 0;
  int64_t acc_im = 0;
  Flags merged() const { return Flags{f_re.sat_pos || f_im.sat_pos, f_re.sat_neg || f_im.sat_neg}; }
  bool saturating() const { return merged().any(); }
};

// Many dot products in one launch: x is [cols][n_ant], w is [n_beams][n_ant];
// result [n_beams][cols].
inline std::vector<DotResult> dot_matrix(const Complex* x, size_t cols, const Complex* w, unsigned n_ant,
                                         unsigned n_beams) {
  std::vector<uint32_t> xp(cols * n_ant), wp((size_t)n_beams * n_ant);
  for (size_t i = 0; i < xp.size(); ++i) xp[i] = x[i].packed();
  for (size_t i = 0; i < wp.size(); ++i) wp[i] = w[i].packed();
  const size_t n = (size_t)n_beams * cols;
  DevVec<uint32_t> dx(xp.data(), xp.size()), dw(wp.data(), wp.size()), dy(n), dp(n);
  DevVec<uint8_t> df(n);
  DevVec<int64_t> da(2 * n);
  check(rd_bf_run(Device::ctx(), dx.ptr(), dw.ptr(), n_ant, n_beams, cols, dy.ptr(), df.ptr(), dp.ptr(), da.ptr()),
        "rd_bf_run");
  const std::vector<uint32_t> y = dy.get(n);
  const std::vector<uint8_t> f = df.get(n);
  const std::vector<int64_t> a = da.get(2 * n);
  std::vector<DotResult> r(n);
  for (size_t i = 0; i < n; ++i) {
    r[i].y = Complex::from_packed(y[i]);
    r[i].f_re = Flags::from_packed(f[i] & 3u);
    r[i].f_im = Flags::from_packed((f[i] >> 2) & 3u);
    r[i].acc_re = a[2 * i];
    r[i].acc_im = a[2 * i + 1];
  }
  return r;
}

// bf::dot (beamformer_model.hpp:162-179)
inline DotResult dot(const Complex* x, const Complex* w, unsigned n_ant) { return dot_matrix(x, 1, w, n_ant, 1)[0]; }

struct BeamBeat {
  std::vector<std::vector<Complex>> y;
  std::vector<std::vector<Flags>> f_re, f_im;
  std::vector<std::vector<int64_t>> acc_re, acc_im;
  bool saturating = false;
};

class BeamformerModel {
 public:
  BeamformerModel(unsigned n_ant, unsigned n_beams, unsigned bin_par, const std::vector<Complex>& w)
      : n_ant_(n_ant), n_beams_(n_beams), bin_par_(bin_par), w_(w) {
    w_.resize((size_t)n_beams * n_ant);
  }
  void swap_weights(const std::vector<Complex>& w) {
    w_ = w;
    w_.resize((size_t)n_beams_ * n_ant_);
  }
  // beat is x[bin][antenna], bin_par x n_ant
  BeamBeat step(const std::vector<std::vector<Complex>>& beat) const {
    std::vector<Complex> x((size_t)bin_par_ * n_ant_);
    for (unsigned j = 0; j < bin_par_; ++j)
      for (unsigned a = 0; a < n_ant_; ++a) x[(size_t)j * n_ant_ + a] = beat[j][a];
    const std::vector<DotResult> d = dot_matrix(x.data(), bin_par_, w_.data(), n_ant_, n_beams_);
    BeamBeat r;
    r.y.assign(n_beams_, std::vector<Complex>(bin_par_));
    r.f_re.assign(n_beams_, std::vector<Flags>(bin_par_));
    r.f_im.assign(n_beams_, std::vector<Flags>(bin_par_));
    r.acc_re.assign(n_beams_, std::vector<int64_t>(bin_par_));
    r.acc_im.assign(n_beams_, std::vector<int64_t>(bin_par_));
    for (unsigned b = 0; b < n_beams_; ++b)
      for (unsigned j = 0; j < bin_par_; ++j) {
        const DotResult& e = d[(size_t)b * bin_par_ + j];
        r.y[b][j] = e.y;
        r.f_re[b][j] = e.f_re;
        r.f_im[b][j] = e.f_im;
        r.acc_re[b][j] = e.acc_re;
        r.acc_im[b][j] = e.acc_im;
        r.saturating = r.saturating || e.saturating();
      }
    return r;
  }

 private:
  unsigned n_ant_, n_beams_, bin_par_;
  std::vector<Complex> w_;
};

}  // namespace bf

// ---------------------------------------------------------------------------
namespace covar {

struct Cross {
  int64_t re = 0;
  int64_t im = 0;
  bool operator==(const Cross& o) const { return re == o.re && im == o.im; }
};

// Exact power and cross power (covar_model.hpp:99-151): plain integer
// identities, evaluated on the host like the reference's inline functions.
inline int64_t power(Complex s) { return (int64_t)s.re * s.re + (int64_t)s.im * s.im; }
inline Cross cross_power(Complex x, Complex y) {
  return Cross{(int64_t)x.re * y.re + (int64_t)x.im * y.im, (int64_t)x.im * y.re - (int64_t)x.re * y.im};
}

// The window result of a block-mode Integrator(acc_w = 64, window = frames) per
// channel pair and bin: xa is [frames][bins][n_chan]; result [bins][pairs].
inline std::vector<Cross> covariance(const std::vector<Complex>& xa, unsigned n_chan, unsigned frames, unsigned bins) {
  const std::vector<uint32_t> x = pack_all(xa);
  const size_t npairs = (size_t)n_chan * (n_chan + 1) / 2;
  DevVec<uint32_t> dx(x.data(), x.size());
  DevVec<int64_t> dr(2 * npairs * bins);
  check(rd_cov_run(Device::ctx(), dx.ptr(), n_chan, frames, bins, dr.ptr()), "rd_cov_run");
  const std::vector<int64_t> r = dr.get(2 * npairs * bins);
  std::vector<Cross> out(npairs * bins);
  for (size_t i = 0; i < out.size(); ++i) out[i] = Cross{r[2 * i], r[2 * i + 1]};
  return out;
}

// covar::Integrator (covar_model.hpp:157-288) over rd_integ_run: the control
// registers stay on the host, the accumulator, sticky flags and event count
// on the device. step() is one clock edge (one launch); run() takes a whole
// control stream with its data in one launch.
enum class Mode : unsigned { kBlock = 0, kExp = 1 };

struct Config {
  unsigned window_len = 1;
  Mode mode = Mode::kBlock;
  unsigned exp_k = 0;
  bool enable = true;
  bool operator==(const Config& o) const {
    return window_len == o.window_len && mode == o.mode && exp_k == o.exp_k && enable == o.enable;
  }
};

struct Result {
  int64_t acc = 0;
  unsigned window_id = 0;
  unsigned sample_count = 0;
  bool flushed = false;
  bool truncated = false;
  bool operator==(const Result& o) const {
    return acc == o.acc && window_id == o.window_id && sample_count == o.sample_count && flushed == o.flushed &&
           truncated == o.truncated;
  }
  bool operator!=(const Result& o) const { return !(*this == o); }
};

struct Step {
  Config cfg;
  bool valid = true;
  int64_t data = 0;
  bool flush = false;
  bool sat_clear = false;
};

class Integrator {
 public:
  Integrator(unsigned acc_w, unsigned window_w) : acc_w_(acc_w), window_w_(window_w), lane_(1) { reset(Config{}); }

  void reset(const Config& cfg) {
    const rd_integ_cfg c = to_c(cfg);
    check(rd_integ_reset(&c, &ctl_), "rd_integ_reset");
    const rd_integ_lane z{0, 0, 0};
    lane_.put(&z, 1);
  }

  bool step(const Config& cfg, bool valid_in, int64_t data, bool flush, bool sat_clear, Result* out) {
    std::vector<std::pair<size_t, Result>> r = run({Step{cfg, valid_in, data, flush, sat_clear}});
    if (!r.empty() && out) *out = r[0].second;
    return !r.empty();
  }

  // Every step of `steps` in one launch; returns (step index, result) of the
  // windows they emit, in order.
  std::vector<std::pair<size_t, Result>> run(const std::vector<Step>& steps) {
    const uint32_t n = (uint32_t)steps.size();
    std::vector<rd_integ_step> cs(n ? n : 1);
    std::vector<int64_t> d(n ? n : 1);
    for (uint32_t s = 0; s < n; ++s) {
      cs[s] = rd_integ_step{to_c(steps[s].cfg), steps[s].valid, steps[s].flush, steps[s].sat_clear};
      d[s] = steps[s].data;
    }
    const uint32_t cap = n ? n : 1;
    DevVec<int64_t> dd(d.data(), d.size()), dr(cap);
    std::vector<rd_integ_window> w(cap);
    uint32_t nw = 0;
    check(rd_integ_run(Device::ctx(), acc_w_, &ctl_, cs.data(), n, dd.ptr(), 1, lane_.ptr(), dr.ptr(), w.data(),
                       cap, &nw),
          "rd_integ_run");
    const std::vector<int64_t> v = dr.get(nw);
    std::vector<std::pair<size_t, Result>> out(nw);
    for (uint32_t k = 0; k < nw; ++k)
      out[k] = {w[k].step, Result{v[k], w[k].window_id, w[k].sample_count, w[k].flushed != 0, w[k].truncated != 0}};
    return out;
  }

  int64_t acc() const { return lane().acc; }
  unsigned count() const { return ctl_.count; }
  unsigned window_id() const { return ctl_.window_id; }
  Config active() const {
    return Config{ctl_.active.window_len, ctl_.active.mode ? Mode::kExp : Mode::kBlock, ctl_.active.exp_k,
                  ctl_.active.enable != 0};
  }
  unsigned effective_len() const { return ctl_.active.window_len == 0 ? 1u : ctl_.active.window_len; }
  Flags sat_flags() const { return Flags::from_packed(lane().flags); }
  bool sat_sticky() const { return sat_flags().any(); }
  uint32_t sat_count() const { return lane().sat_count; }
  unsigned acc_w() const { return acc_w_; }
  unsigned window_w() const { return window_w_; }

 private:
  static rd_integ_cfg to_c(const Config& c) {
    return rd_integ_cfg{c.window_len, (unsigned)c.mode, c.exp_k, c.enable ? 1u : 0u};
  }
  rd_integ_lane lane() const { return lane_.get(1)[0]; }
  unsigned acc_w_, window_w_;
  rd_integ_ctl ctl_{};
  DevVec<rd_integ_lane> lane_;
};

}  // namespace covar

// ---------------------------------------------------------------------------
namespace cfar {

enum class Mode : unsigned { kCA = 0, kGO = 1 };
enum class OutMode : unsigned { kEvents = 0, kDense = 1 };
enum Kind : unsigned { kEvDetect = 0, kEvBin = 1, kEvSuppressed = 2, kEvSummary = 3 };

struct Config {
  bool enable = true;
  Mode mode = Mode::kCA;
  OutMode out_mode = OutMode::kEvents;
  unsigned guard_lead = 2, guard_lag = 2, ref_lead = 8, ref_lag = 8;
  unsigned alpha = 256;
};

struct Event {
  unsigned kind = kEvSummary, bin = 0, frame_id = 0, ref_count = 0, alpha = 0, det_count = 0, sup_count = 0;
  uint64_t cut_power = 0, noise_sum = 0;
  bool sof = false, eof = false;
  bool operator==(const Event& o) const {
    return kind == o.kind && bin == o.bin && frame_id == o.frame_id && ref_count == o.ref_count &&
           alpha == o.alpha && det_count == o.det_count && sup_count == o.sup_count && cut_power == o.cut_power &&
           noise_sum == o.noise_sum && sof == o.sof && eof == o.eof;
  }
};

// Unpacks one 176-bit event word (cfar_model.hpp:75-84) plus the framing bits.
inline Event unpack_event(const uint64_t* w) {
  auto field = [w](unsigned lsb, unsigned n) {
    uint64_t v = 0;
    for (unsigned i = 0; i < n; ++i) v |= ((w[(lsb + i) / 64] >> ((lsb + i) % 64)) & 1ull) << i;
    return v;
  };
  Event e;
  e.kind = (unsigned)field(0, 2);
  e.bin = (unsigned)field(2, 16);
  e.frame_id = (unsigned)field(18, 16);
  e.ref_count = (unsigned)field(34, 7);
  e.alpha = (unsigned)field(41, 16);
  e.det_count = (unsigned)field(57, 16);
  e.sup_count = (unsigned)field(73, 16);
  e.cut_power = field(89, 40);
  e.noise_sum = field(129, 47);
  e.sof = field(176, 1) != 0;
  e.eof = field(177, 1) != 0;
  return e;
}

// cfar::Decision / cfar::decide (cfar_model.hpp:223-285): cfg as given, no clamp.
struct Decision {
  bool window_ok = false, evaluable = false, detected = false, suppressed = false;
  uint64_t cut_power = 0, sum_lead = 0, sum_lag = 0, sel_sum = 0;
  unsigned sel_n = 0;
  bool take_lead = false;
};

inline rd_cfar_cfg to_c(const Config& cfg) {
  return rd_cfar_cfg{cfg.enable ? 1u : 0u, (unsigned)cfg.mode, (unsigned)cfg.out_mode, cfg.guard_lead,
                     cfg.guard_lag, cfg.ref_lead, cfg.ref_lag, cfg.alpha};
}

// Every bin's decision in one launch.
inline std::vector<Decision> decide_all(const std::vector<uint64_t>& bins, const Config& cfg) {
  const unsigned n = (unsigned)bins.size();
  const rd_cfar_cfg c = to_c(cfg);
  DevVec<uint64_t> dc(bins.data(), n ? n : 1);
  DevVec<rd_cfar_decision> dd(n ? n : 1);
  check(rd_cfar_decide(Device::ctx(), dc.ptr(), n, 1, &c, dd.ptr()), "rd_cfar_decide");
  const std::vector<rd_cfar_decision> r = dd.get(n);
  std::vector<Decision> out(n);
  for (unsigned i = 0; i < n; ++i)
    out[i] = Decision{r[i].window_ok != 0, r[i].evaluable != 0, r[i].detected != 0, r[i].suppressed != 0,
                      r[i].cut_power,      r[i].sum_lead,       r[i].sum_lag,       r[i].sel_sum,
                      r[i].sel_n,          r[i].take_lead != 0};
  return out;
}

inline Decision decide(const std::vector<uint64_t>& bins, size_t i, const Config& cfg) { return decide_all(bins, cfg)[i]; }

// cfar::run_frame (cfar_model.hpp:295-340)
inline std::vector<Event> run_frame(const std::vector<uint64_t>& bins, const Config& cfg, unsigned max_guard,
                                    unsigned max_ref, unsigned frame_id) {
  const unsigned n = (unsigned)bins.size();
  rd_cfar_cfg c{cfg.enable ? 1u : 0u, (unsigned)cfg.mode, (unsigned)cfg.out_mode, cfg.guard_lead, cfg.guard_lag,
                cfg.ref_lead, cfg.ref_lag, cfg.alpha};
  DevVec<uint64_t> dc(bins.data(), n), dev(3 * ((size_t)n + 1));
  DevVec<uint32_t> dfid(&frame_id, 1), drc(1);
  uint64_t total = 0;
  check(rd_cfar_run(Device::ctx(), dc.ptr(), 8, n, 1, &c, max_guard, max_ref, dfid.ptr(), 0, dev.ptr(), n + 1,
                    drc.ptr(), &total),
        "rd_cfar_run");
  const std::vector<uint64_t> w = dev.get(3 * total);
  std::vector<Event> out(total);
  for (size_t i = 0; i < total; ++i) out[i] = unpack_event(&w[3 * i]);
  return out;
}

}  // namespace cfar

}  // namespace rd_amd

#endif  // RD_AMD_HPP_
