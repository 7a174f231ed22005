// -----------------------------------------------------------------------------
// ref_shim.cpp — C entry points over the REFERENCE's own header-only models.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile against the reference
// sources where they lie (-I <reference>/model/cpp) into oracle/_ref/, which is
// never committed. It is used to check the CPU restatement (rd_oracle.cpp)
// against the real models, and optionally as the timed CPU baseline. Nothing
// here restates arithmetic: every value comes from a call into the reference.
//
// The reference caps the FFT at N <= 1024 (fft_ref.hpp:216-219), so this shim
// covers the chain only up to that size.
// -----------------------------------------------------------------------------
#include <cctype>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "beamformer/beamformer_model.hpp"
#include "cfar/cfar_model.hpp"
#include "covariance/covar_model.hpp"
#include "fft/fft_ref.hpp"
#include "fxp/cmult.hpp"
#include "fxp/fxp.hpp"
#include "fxp/fxp_ops.hpp"
#include "packet/packet_model.hpp"
#include "pfb/pfb_model.hpp"

namespace {

template <class F>
void par(size_t n, unsigned threads, F f) {
  if (threads <= 1 || n <= 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  const unsigned nt = (unsigned)(threads < n ? threads : n);
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      for (size_t i = t; i < n; i += nt) f(i);
    });
  for (auto& th : pool) th.join();
}

std::vector<fxp::Complex> to_cx(const uint32_t* p, size_t n) {
  std::vector<fxp::Complex> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = fxp::Complex::from_packed(p[i]);
  return v;
}

// Pack one reference event by the reference's own layout constants
// (cfar_model.hpp:75-84); framing bits sof/eof at 176/177.
void put(uint64_t* w, unsigned lsb, unsigned n, uint64_t v) {
  if (n < 64) v &= (1ULL << n) - 1ULL;
  for (unsigned i = 0; i < n; ++i)
    if ((v >> i) & 1ULL) w[(lsb + i) / 64] |= 1ULL << ((lsb + i) % 64);
}
void pack_ref(const cfar::Event& e, uint64_t* w) {
  w[0] = w[1] = w[2] = 0;
  put(w, cfar::kEvKindLsb, cfar::kEvKindW, e.kind);
  put(w, cfar::kEvBinLsb, cfar::kBinW, e.bin);
  put(w, cfar::kEvFrameIdLsb, cfar::kFrameIdW, e.frame_id);
  put(w, cfar::kEvRefCntLsb, cfar::kTotRefW, e.ref_count);
  put(w, cfar::kEvAlphaLsb, cfar::kAlphaW, e.alpha);
  put(w, cfar::kEvDetLsb, cfar::kCntW, e.det_count);
  put(w, cfar::kEvSupLsb, cfar::kCntW, e.sup_count);
  put(w, cfar::kEvCellLsb, cfar::kPowerW, e.cut_power);
  put(w, cfar::kEvSumLsb, cfar::kSumW, e.noise_sum);
  put(w, cfar::kEventW, 1, e.sof ? 1 : 0);
  put(w, cfar::kEventW + 1, 1, e.eof ? 1 : 0);
}

cfar::Config cfg_of(const uint32_t* c) {
  cfar::Config o;
  o.enable = c[0] != 0;
  o.mode = c[1] ? cfar::Mode::kGO : cfar::Mode::kCA;
  o.out_mode = c[2] ? cfar::OutMode::kDense : cfar::OutMode::kEvents;
  o.guard_lead = c[3];
  o.guard_lag = c[4];
  o.ref_lead = c[5];
  o.ref_lag = c[6];
  o.alpha = c[7];
  return o;
}

}  // namespace

extern "C" {

int rdr_fxp_op(const char* op, int64_t a, int64_t b, unsigned sh, unsigned w, int64_t* y, unsigned* flags) {
  const unsigned code = fxp::ops::code_of(op);
  if (code >= fxp::ops::kOpCount) return 0;
  const fxp::ops::Result r = fxp::ops::apply(code, a, b, sh, w);
  *y = r.y;
  *flags = r.flags.packed();
  return 1;
}

void rdr_cmult(uint32_t a, uint32_t b, int64_t* p_re, int64_t* p_im, int64_t* p3_re, int64_t* p3_im, uint32_t* y,
               unsigned* f_re, unsigned* f_im) {
  const fxp::Complex ca = fxp::Complex::from_packed(a), cb = fxp::Complex::from_packed(b);
  const auto r4 = fxp::cmult::eval(fxp::cmult::Variant::kMult4, ca, cb);
  const auto r3 = fxp::cmult::raw_mult3(ca, cb);
  *p_re = r4.p_re;
  *p_im = r4.p_im;
  *p3_re = r3.re;
  *p3_im = r3.im;
  *y = fxp::Complex{r4.y_re, r4.y_im}.packed();
  *f_re = r4.f_re.packed();
  *f_im = r4.f_im.packed();
}

void rdr_twiddle_table(uint32_t* out) {
  for (unsigned e = 0; e < fft::kTwN; ++e) out[e] = fft::kTw[e];
}

void rdr_pfb_run(const uint32_t* x, const uint32_t* coeff, unsigned P, unsigned T, unsigned n_chan,
                 unsigned n_beats, uint32_t* y, uint8_t* flags, unsigned threads) {
  const std::vector<fxp::Complex> h = to_cx(coeff, (size_t)P * T);
  const size_t span = (size_t)n_beats * P;
  par(n_chan, threads, [&](size_t c) {
    pfb::PfbModel m(P, T, h);
    m.reset();
    for (unsigned b = 0; b < n_beats; ++b) {
      const pfb::BeatResult r = m.step(to_cx(x + c * span + (size_t)b * P, P));
      for (unsigned p = 0; p < P; ++p) {
        const size_t o = c * span + (size_t)b * P + p;
        y[o] = r.y[p].packed();
        if (flags) flags[o] = (uint8_t)(r.flags_re[p].packed() | (r.flags_im[p].packed() << 2));
      }
    }
  });
}

int rdr_fft_run(const uint32_t* in, unsigned n, unsigned spc, uint32_t sched, int reorder, unsigned n_frames,
                uint32_t* out, uint32_t* bins, uint32_t* stage_flags, unsigned threads) {
  fft::Config cfg;
  cfg.n_fft = n;
  cfg.spc = spc;
  cfg.scale_sched = sched;
  cfg.reorder = reorder != 0;
  if (!cfg.valid()) return -1;
  par(n_frames, threads, [&](size_t f) {
    const fft::Result r = fft::transform(cfg, to_cx(in + f * n, n));
    for (unsigned i = 0; i < n; ++i) {
      if (out) out[f * n + i] = r.out[i].packed();
      if (bins) bins[f * n + i] = r.bins[i].packed();
    }
    if (stage_flags) {
      uint32_t w = 0;
      for (size_t g = 0; g < r.stage_flags.size(); ++g) w |= r.stage_flags[g].packed() << (2 * g);
      stage_flags[f] = w;
    }
  });
  return 0;
}

void rdr_bf_run(const uint32_t* Xa, const uint32_t* W, unsigned CH, unsigned B, size_t cols, uint32_t* y,
                uint8_t* flags, uint32_t* power, int64_t* acc, unsigned threads) {
  const std::vector<fxp::Complex> w = to_cx(W, (size_t)B * CH);
  par(B, threads, [&](size_t b) {
    std::vector<fxp::Complex> x(CH);
    for (size_t col = 0; col < cols; ++col) {
      for (unsigned a = 0; a < CH; ++a) x[a] = fxp::Complex::from_packed(Xa[col * CH + a]);
      const bf::DotResult d = bf::dot(x.data(), &w[bf::weight_index(CH, (unsigned)b, 0)], CH);
      const size_t o = b * cols + col;
      if (y) y[o] = d.y.packed();
      if (flags) flags[o] = (uint8_t)(d.f_re.packed() | (d.f_im.packed() << 2));
      if (power) power[o] = (uint32_t)cfar::clamp_cell(covar::power(d.y));
      if (acc) {
        acc[2 * o] = d.acc_re;
        acc[2 * o + 1] = d.acc_im;
      }
    }
  });
}

// Covariance through the reference Integrator: block mode, window = F frames,
// one accumulator per component; the result is what the window close emits.
void rdr_cov_integrate(const uint32_t* Xa, unsigned CH, unsigned F, unsigned N, unsigned mode, unsigned exp_k,
                       unsigned acc_w, int64_t* R, uint8_t* sat_flags, unsigned threads) {
  const size_t npairs = (size_t)CH * (CH + 1) / 2;
  covar::Config cfg;
  cfg.window_len = F;
  cfg.mode = mode == 1 ? covar::Mode::kExp : covar::Mode::kBlock;
  cfg.exp_k = exp_k;
  par(N, threads, [&](size_t k) {
    size_t pi = 0;
    for (unsigned i = 0; i < CH; ++i) {
      for (unsigned j = i; j < CH; ++j, ++pi) {
        covar::Integrator ire(acc_w, 16), iim(acc_w, 16);
        ire.reset(cfg);
        iim.reset(cfg);
        covar::Result rre{}, rim{};
        for (unsigned m = 0; m < F; ++m) {
          const size_t base = ((size_t)m * N + k) * CH;
          const covar::Cross c = covar::cross_power(fxp::Complex::from_packed(Xa[base + i]),
                                                    fxp::Complex::from_packed(Xa[base + j]));
          ire.step(cfg, true, c.re, false, false, &rre);
          iim.step(cfg, true, c.im, false, false, &rim);
          if (m + 1 == F) {
            // sticky flags of the closing window, read before the next edge
            if (sat_flags) sat_flags[k * npairs + pi] = (uint8_t)(ire.sat_flags().packed() | (iim.sat_flags().packed() << 2));
          }
        }
        R[(k * npairs + pi) * 2] = rre.acc;
        R[(k * npairs + pi) * 2 + 1] = rim.acc;
      }
    }
  });
}

void rdr_cov_run(const uint32_t* Xa, unsigned CH, unsigned F, unsigned N, unsigned acc_w, int64_t* R,
                 uint8_t* sat_flags, unsigned threads) {
  rdr_cov_integrate(Xa, CH, F, N, 0, 0, acc_w, R, sat_flags, threads);
}

void rdr_cross_power(uint32_t x, uint32_t y, int64_t* sre, int64_t* sim, int64_t* dre, int64_t* dim) {
  const covar::Cross s = covar::cross_power_swap(fxp::Complex::from_packed(x), fxp::Complex::from_packed(y));
  const covar::Cross d = covar::cross_power_direct(fxp::Complex::from_packed(x), fxp::Complex::from_packed(y));
  *sre = s.re;
  *sim = s.im;
  *dre = d.re;
  *dim = d.im;
}

int64_t rdr_power(uint32_t s) { return covar::power(fxp::Complex::from_packed(s)); }

int64_t rdr_cfar_run(const uint64_t* bins, unsigned n, unsigned n_rows, const uint32_t* cfg8, unsigned max_guard,
                     unsigned max_ref, const uint32_t* frame_ids, uint64_t* events, int64_t events_cap,
                     uint32_t* row_count) {
  const cfar::Config c = cfg_of(cfg8);
  int64_t total = 0;
  for (unsigned r = 0; r < n_rows; ++r) {
    std::vector<uint64_t> row(bins + (size_t)r * n, bins + (size_t)(r + 1) * n);
    const std::vector<cfar::Event> ev = cfar::run_frame(row, c, max_guard, max_ref, frame_ids ? frame_ids[r] : r);
    if (row_count) row_count[r] = (uint32_t)ev.size();
    if (total + (int64_t)ev.size() > events_cap) return -1;
    for (size_t e = 0; e < ev.size(); ++e) pack_ref(ev[e], events + 3 * (size_t)(total + (int64_t)e));
    total += (int64_t)ev.size();
  }
  return total;
}

// The composed chain for one CPI with the reference models (N <= 1024), same
// glue and output layout as rdo_chain_run.
int64_t rdr_chain_run(const uint32_t* iq, const uint32_t* taps, const uint32_t* W, unsigned CH, unsigned N, unsigned T,
                      unsigned F, unsigned B, unsigned spc, uint32_t sched, const uint32_t* cfar_cfg8,
                      unsigned max_guard, unsigned max_ref, uint32_t* u, uint8_t* pfb_flags, uint32_t* X,
                      uint32_t* fft_flags, uint32_t* Xa, uint32_t* y, uint8_t* yflags, uint32_t* power, int64_t* R,
                      uint64_t* events, int64_t events_cap, uint32_t* row_count, unsigned threads) {
  const size_t S = (size_t)CH * F * N, cols = (size_t)F * N;
  std::vector<uint32_t> u_l, X_l, Xa_l, pw_l;
  if (!u) { u_l.resize(S); u = u_l.data(); }
  if (!X) { X_l.resize(S); X = X_l.data(); }
  if (!Xa) { Xa_l.resize(S); Xa = Xa_l.data(); }
  if (!power) { pw_l.resize((size_t)B * cols); power = pw_l.data(); }
  rdr_pfb_run(iq, taps, N, T, CH, F, u, pfb_flags, threads);
  if (rdr_fft_run(u, N, spc, sched, 1, CH * F, nullptr, X, fft_flags, threads) != 0) return -2;
  for (unsigned c = 0; c < CH; ++c)  // alignment: bin-major, antenna-minor (align_model.hpp:172-186)
    for (size_t col = 0; col < cols; ++col) Xa[col * CH + c] = X[(size_t)c * cols + col];
  rdr_bf_run(Xa, W, CH, B, cols, y, yflags, power, nullptr, threads);
  if (R) rdr_cov_run(Xa, CH, F, N, 64, R, nullptr, threads);
  const size_t rows = (size_t)B * F;
  const cfar::Config c = cfg_of(cfar_cfg8);
  std::vector<std::vector<cfar::Event>> per_row(rows);
  par(rows, threads, [&](size_t r) {
    std::vector<uint64_t> cells(N);
    for (unsigned i = 0; i < N; ++i) cells[i] = cfar::clamp_cell((int64_t)power[r * N + i]);
    per_row[r] = cfar::run_frame(cells, c, max_guard, max_ref, (unsigned)(r % F));
  });
  int64_t total = 0;
  for (size_t r = 0; r < rows; ++r) {
    if (row_count) row_count[r] = (uint32_t)per_row[r].size();
    if (events) {
      if (total + (int64_t)per_row[r].size() > events_cap) return -1;
      for (size_t e = 0; e < per_row[r].size(); ++e) pack_ref(per_row[r][e], events + 3 * (size_t)(total + (int64_t)e));
    }
    total += (int64_t)per_row[r].size();
  }
  return total;
}

// Detection packets through packet_model::hdr_pack and make_flit: the payload
// flit f carries event bits [(f-1)*packet_w, f*packet_w) in 64-bit slices.
int rdr_event_flits(const uint64_t* events, uint64_t n_events, unsigned packet_w, unsigned dest, unsigned src,
                    unsigned vc, unsigned seq0, uint32_t* out) {
  namespace pm = packet_model;
  if (packet_w < pm::kHdrW || packet_w > 512) return -1;
  const unsigned L = pm::detection_flits(packet_w);
  const unsigned fw = (packet_w + pm::kFlitCtrlW + 31) / 32;
  auto ev_bits = [&](const uint64_t* ev, unsigned lsb, unsigned n) {
    uint64_t v = 0;
    for (unsigned i = 0; i < n; ++i) {
      const unsigned b = lsb + i;
      if (b < pm::kDetectionPayloadW && ((ev[b / 64] >> (b % 64)) & 1u)) v |= 1ull << i;
    }
    return v;
  };
  for (uint64_t e = 0; e < n_events; ++e) {
    const uint64_t* ev = events + 3 * e;
    pm::Header h;
    h.dest = dest;
    h.src = src;
    h.vc = vc;
    h.ptype = pm::kTypeDetection;
    h.length = L;
    h.seq = (unsigned)((seq0 + e) & 0xFFFFu);
    for (unsigned f = 0; f < L; ++f) {
      std::vector<uint64_t> data;
      for (unsigned k = 0; k * 64 < packet_w; ++k) {
        const unsigned n = (packet_w - k * 64 >= 64) ? 64 : packet_w - k * 64;
        if (f == 0)
          data.push_back(k == 0 ? pm::hdr_pack(h) : 0ull);
        else
          data.push_back(ev_bits(ev, (f - 1) * packet_w + k * 64, n));
      }
      const pm::Flit fl = pm::make_flit(packet_w, vc, f == 0, f == L - 1, data);
      if (fl.w.size() != fw || !fl.parity_ok()) return -2;
      for (unsigned k = 0; k < fw; ++k) out[((size_t)e * L + f) * fw + k] = fl.w[k];
    }
  }
  return (int)L;
}

// The reference's receive side: packets built from `events` as rdr_event_flits
// builds them are injected into a packet_model::Scoreboard (a 32-port,
// one-stage fabric, so every src < 32 to dest < 16 has a route), then every
// flit of `flits` [n_packets][L][fw] is observed at egress `port` and the run
// finished. counts: delivered, lost, duplicated, misrouted, corrupted,
// reordered, framing, parity, unexpected, injected.
int rdr_scoreboard(const uint32_t* flits, uint64_t n_packets, unsigned packet_w, unsigned port,
                   const uint64_t* events, uint64_t n_events, unsigned dest, unsigned src, unsigned vc, unsigned seq0,
                   uint64_t* counts) {
  namespace pm = packet_model;
  if (packet_w < pm::kHdrW || packet_w > 512) return -1;
  const unsigned L = pm::detection_flits(packet_w);
  const unsigned fw = (packet_w + pm::kFlitCtrlW + 31) / 32;
  pm::Scoreboard sb(packet_w, 32, 1);
  for (uint64_t e = 0; e < n_events; ++e) {
    const uint64_t* ev = events + 3 * e;
    pm::Packet p;
    p.hdr.dest = dest;
    p.hdr.src = src;
    p.hdr.vc = vc;
    p.hdr.ptype = pm::kTypeDetection;
    p.hdr.length = L;
    p.hdr.seq = (unsigned)((seq0 + e) & 0xFFFFu);
    for (unsigned f = 1; f < L; ++f) {
      std::vector<uint64_t> words;
      for (unsigned k = 0; k * 64 < packet_w; ++k) {
        const unsigned n = (packet_w - k * 64 >= 64) ? 64 : packet_w - k * 64;
        uint64_t v = 0;
        for (unsigned i = 0; i < n; ++i) {
          const unsigned b = (f - 1) * packet_w + k * 64 + i;
          if (b < pm::kDetectionPayloadW && ((ev[b / 64] >> (b % 64)) & 1u)) v |= 1ull << i;
        }
        words.push_back(v);
      }
      p.payload.push_back(words);
    }
    sb.inject(p);
  }
  for (uint64_t k = 0; k < n_packets * L; ++k) {
    pm::Flit fl;
    fl.resize_for(packet_w);
    if (fl.w.size() != fw) return -2;
    for (unsigned j = 0; j < fw; ++j) fl.w[j] = flits[k * fw + j];
    sb.observe(port, fl);
  }
  sb.finish();
  const pm::Counts& c = sb.counts();
  counts[0] = c.delivered;
  counts[1] = c.lost;
  counts[2] = c.duplicated;
  counts[3] = c.misrouted;
  counts[4] = c.corrupted;
  counts[5] = c.reordered;
  counts[6] = c.framing;
  counts[7] = c.parity;
  counts[8] = c.unexpected;
  counts[9] = sb.injected();
  return (int)L;
}

}  // extern "C"

namespace {
covar::Config icfg_of(const uint32_t* c) {
  covar::Config o;
  o.window_len = c[0];
  o.mode = c[1] == 1 ? covar::Mode::kExp : covar::Mode::kBlock;
  o.exp_k = c[2];
  o.enable = c[3] != 0;
  return o;
}

// One reference Integrator per lane, stepped with the control stream; see
// rdo_integ_run for the layouts.
template <class G>
uint32_t ref_lane(covar::Integrator& I, const uint32_t* steps7, uint32_t n_steps, G get, int64_t* results,
                  size_t n_lanes, size_t lane, uint32_t* win5, uint32_t cap) {
  uint32_t nw = 0;
  for (uint32_t s = 0; s < n_steps; ++s) {
    const uint32_t* t = steps7 + 7 * (size_t)s;
    covar::Result r{};
    if (I.step(icfg_of(t), t[4] != 0, get(s), t[5] != 0, t[6] != 0, &r)) {
      if (nw < cap) {
        if (results) results[(size_t)nw * n_lanes + lane] = r.acc;
        if (win5 && lane == 0) {
          const uint32_t w[5] = {s, r.window_id, r.sample_count, r.flushed ? 1u : 0u, r.truncated ? 1u : 0u};
          std::memcpy(win5 + 5 * (size_t)nw, w, sizeof(w));
        }
      }
      ++nw;
    }
  }
  return nw;
}

void ctl_of(const covar::Integrator& I, uint32_t* ctl6) {
  const covar::Config& a = I.active();
  const uint32_t c[6] = {a.window_len, a.mode == covar::Mode::kExp ? 1u : 0u, a.exp_k, a.enable ? 1u : 0u,
                         I.count(), I.window_id()};
  std::memcpy(ctl6, c, sizeof(c));
}
}  // namespace

extern "C" {

uint32_t rdr_integ_run(unsigned acc_w, const uint32_t* cfg4, const uint32_t* steps7, uint32_t n_steps,
                       const int64_t* data, size_t n_lanes, int64_t* lane_acc, uint32_t* lane_flags,
                       uint32_t* lane_count, int64_t* results, uint32_t* win5, uint32_t cap, uint32_t* ctl6) {
  uint32_t nw = 0;
  for (size_t l = 0; l < n_lanes; ++l) {
    covar::Integrator I(acc_w, 16);
    I.reset(icfg_of(cfg4));
    nw = ref_lane(I, steps7, n_steps, [&](uint32_t s) { return (fxp::wide_t)data[(size_t)s * n_lanes + l]; },
                  results, n_lanes, l, win5, cap);
    lane_acc[l] = I.acc();
    lane_flags[l] = I.sat_flags().packed();
    lane_count[l] = I.sat_count();
    if (l == 0 && ctl6) ctl_of(I, ctl6);
  }
  return nw;
}

uint32_t rdr_cov_integrate_windows(const uint32_t* Xa, unsigned CH, unsigned F, unsigned N, unsigned acc_w,
                                   const uint32_t* cfg4, const uint32_t* steps7, int64_t* lane_acc,
                                   uint32_t* lane_flags, uint32_t* lane_count, int64_t* results, uint32_t* win5,
                                   uint32_t cap, uint32_t* ctl6, unsigned threads) {
  const size_t npairs = (size_t)CH * (CH + 1) / 2, lanes = 2 * (size_t)N * npairs;
  std::vector<uint32_t> nws(N, 0);
  par(N, threads, [&](size_t k) {
    size_t pi = 0;
    for (unsigned i = 0; i < CH; ++i)
      for (unsigned j = i; j < CH; ++j, ++pi)
        for (unsigned c = 0; c < 2; ++c) {
          const size_t l = (k * npairs + pi) * 2 + c;
          covar::Integrator I(acc_w, 16);
          I.reset(icfg_of(cfg4));
          auto get = [&](uint32_t m) {
            const size_t base = ((size_t)m * N + k) * CH;
            const covar::Cross x = covar::cross_power(fxp::Complex::from_packed(Xa[base + i]),
                                                      fxp::Complex::from_packed(Xa[base + j]));
            return c ? x.im : x.re;
          };
          nws[k] = ref_lane(I, steps7, F, get, results, lanes, l, l == 0 ? win5 : nullptr, cap);
          lane_acc[l] = I.acc();
          lane_flags[l] = I.sat_flags().packed();
          lane_count[l] = I.sat_count();
          if (l == 0 && ctl6) ctl_of(I, ctl6);
        }
  });
  return N ? nws[0] : 0;
}

// pfb::LaneModel::accumulate per branch; the history is seeded by stepping the
// model through the hist beats, oldest first.
void rdr_pfb_acc(const uint32_t* x, const uint32_t* coeff, unsigned P, unsigned T, unsigned n_chan,
                 unsigned n_beats, const uint32_t* hist, int64_t* acc) {
  for (unsigned c = 0; c < n_chan; ++c)
    for (unsigned p = 0; p < P; ++p) {
      std::vector<fxp::Complex> h(T);
      for (unsigned k = 0; k < T; ++k) h[k] = fxp::Complex::from_packed(coeff[(size_t)p * T + k]);
      pfb::LaneModel L(T, h.data());
      if (hist)
        for (unsigned j = T - 1; j-- > 0;) L.step(fxp::Complex::from_packed(hist[((size_t)c * (T - 1) + j) * P + p]));
      for (unsigned m = 0; m < n_beats; ++m) {
        const fxp::Complex s = fxp::Complex::from_packed(x[((size_t)c * n_beats + m) * P + p]);
        fxp::wide_t re = 0, im = 0;
        L.accumulate(s, &re, &im);
        const size_t o = (((size_t)c * n_beats + m) * P + p) * 2;
        acc[o] = re;
        acc[o + 1] = im;
        L.step(s);
      }
    }
}

void rdr_cfar_decide(const uint64_t* bins, unsigned n, unsigned n_rows, const uint32_t* cfg8, uint64_t* dec) {
  const cfar::Config c = cfg_of(cfg8);
  for (size_t r = 0; r < n_rows; ++r) {
    std::vector<uint64_t> row(bins + r * n, bins + (r + 1) * n);
    for (size_t i = 0; i < n; ++i) {
      const cfar::Decision d = cfar::decide(row, i, c);
      uint64_t* o = dec + (r * n + i) * 8;
      o[0] = d.cut_power;
      o[1] = d.sum_lead;
      o[2] = d.sum_lag;
      o[3] = d.sel_sum;
      o[4] = d.sel_n;
      o[5] = d.window_ok;
      o[6] = (d.evaluable ? 1u : 0u) | (d.detected ? 2u : 0u) | (d.suppressed ? 4u : 0u);
      o[7] = d.take_lead;
    }
  }
}

}  // extern "C"
