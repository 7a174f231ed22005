// test_mirror.cpp — a reference-style harness over the C++ mirror (rd_amd.hpp).
//
// Shaped like the reference's own gates (model/cpp/test/test_fft_ref.cpp:60-90,
// the PFB vector check of sim/tests/test_pfb_bank.cpp, test_beamformer.cpp's
// unit-weight case, test_cfar.cpp's flat-spectrum and decide cases,
// test_covariance.cpp's window, saturation and exponential passes), but every
// call goes through rd_amd::pfb / fft / bf / covar / cfar on the GPU. It reads the reference's
// committed vectors from tests/golden/ref_vectors.
//
//   test_mirror <vector dir>      prints RESULT: PASS / FAIL, exit 0 / 1
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rd_amd.hpp"

using rd_amd::Complex;

namespace {

size_t g_checks = 0, g_fail = 0;

void fail(const std::string& m) {
  ++g_fail;
  if (g_fail < 20) std::fprintf(stderr, "FAIL: %s\n", m.c_str());
}

// Data lines of a vector file, '#' header lines as key/value pairs.
struct VecFile {
  std::vector<std::pair<std::string, std::string>> header;
  std::vector<std::vector<std::string>> rows;
  std::string get(const std::string& k) const {
    for (auto& kv : header)
      if (kv.first == k) return kv.second;
    return "";
  }
};

VecFile read_vec(const std::string& path) {
  VecFile v;
  std::ifstream f(path);
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path.c_str());
    std::exit(2);
  }
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty()) continue;
    if (line[0] == '#') {
      const size_t c = line.find(':');
      if (c == std::string::npos) continue;
      auto trim = [](std::string s) {
        const size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t");
        return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
      };
      v.header.emplace_back(trim(line.substr(1, c - 1)), trim(line.substr(c + 1)));
      continue;
    }
    std::istringstream is(line);
    std::vector<std::string> t;
    std::string w;
    while (is >> w) t.push_back(w);
    v.rows.push_back(t);
  }
  return v;
}

Complex cx(const std::string& re, const std::string& im) {
  return Complex{(int16_t)std::stol(re), (int16_t)std::stol(im)};
}

void check_fft(const std::string& dir) {
  const VecFile v = read_vec(dir + "/fft64.vec");
  if (v.get("twiddle_digest") != "c756048fbc00d64f") fail("fft64.vec digest");
  const unsigned n = (unsigned)std::stoul(v.get("fft_size")), spc = (unsigned)std::stoul(v.get("spc"));
  size_t i = 0, records = 0;
  while (i < v.rows.size()) {
    const auto& h = v.rows[i];
    rd_amd::fft::Config cfg;
    cfg.n_fft = n;
    cfg.spc = spc;
    cfg.scale_sched = (unsigned)std::stoul(h[2], nullptr, 16);
    cfg.reorder = h[3] == "1";
    const unsigned flags = (unsigned)std::stoul(h[4], nullptr, 16);
    std::vector<Complex> in(n), exp(n);
    for (unsigned s = 0; s < n; ++s) {
      const auto& r = v.rows[i + 1 + s];
      in[s] = cx(r[2], r[3]);
      exp[s] = cx(r[4], r[5]);
    }
    i += 1 + n;
    ++records;
    const rd_amd::fft::Result res = rd_amd::fft::transform(cfg, in);
    ++g_checks;
    if (res.out != exp) fail("fft record " + h[1] + " samples");
    unsigned got = 0;
    for (size_t g = 0; g < res.stage_flags.size(); ++g) got |= res.stage_flags[g].packed() << (2 * g);
    if (got != flags) fail("fft record " + h[1] + " stage flags");
  }
  if (records != 36) fail("fft64.vec record count");
}

void check_pfb(const std::string& dir, const std::string& set, const std::string& geom) {
  const VecFile c = read_vec(dir + "/pfb_" + set + "_" + geom + ".coeff");
  const VecFile v = read_vec(dir + "/pfb_" + set + "_" + geom + ".vec");
  const unsigned P = (unsigned)std::stoul(c.get("phases")), T = (unsigned)std::stoul(c.get("taps"));
  std::vector<Complex> coeff;
  for (auto& r : c.rows) coeff.push_back(cx(r[3], r[4]));
  rd_amd::pfb::PfbModel m(P, T, coeff);
  m.reset();
  for (auto& r : v.rows) {  // one step() per beat: exercises the device-resident delay line
    std::vector<Complex> beat(P);
    for (unsigned p = 0; p < P; ++p) beat[p] = cx(r[1 + 2 * p], r[2 + 2 * p]);
    const rd_amd::pfb::BeatResult br = m.step(beat);
    ++g_checks;
    for (unsigned p = 0; p < P; ++p) {
      const Complex y = cx(r[1 + 2 * P + 2 * p], r[2 + 2 * P + 2 * p]);
      const unsigned fr = (unsigned)std::stoul(r[1 + 4 * P + 2 * p]), fi = (unsigned)std::stoul(r[2 + 4 * P + 2 * p]);
      if (br.y[p] != y || br.flags_re[p].packed() != fr || br.flags_im[p].packed() != fi)
        fail("pfb " + set + "_" + geom + " beat " + r[0]);
    }
  }
}

void check_bf_unit() {
  const unsigned n_ant = 8, n_beams = 8;
  std::vector<Complex> w((size_t)n_ant * n_beams);
  for (unsigned b = 0; b < n_beams; ++b) w[rd_amd::bf::weight_index(n_ant, b, (b * 3) % n_ant)] = Complex{32767, 0};
  rd_amd::bf::BeamformerModel m(n_ant, n_beams, 4, w);
  std::vector<std::vector<Complex>> beat(4, std::vector<Complex>(n_ant));
  for (unsigned j = 0; j < 4; ++j)
    for (unsigned a = 0; a < n_ant; ++a) beat[j][a] = Complex{(int16_t)(1000 * a - 3001 * j), (int16_t)(-77 * a + j)};
  const rd_amd::bf::BeamBeat r = m.step(beat);
  for (unsigned b = 0; b < n_beams; ++b)
    for (unsigned j = 0; j < 4; ++j) {
      const Complex x = beat[j][(b * 3) % n_ant];
      // round_sat(x * 0x7FFF, 15, 16): x*32767/32768 rounded to nearest even
      auto q = [](int v) {
        const long long p = (long long)v * 32767;
        long long qq = p >> 15;
        const long long rem = p & 0x7FFF;
        if (rem > 0x4000 || (rem == 0x4000 && (qq & 1))) ++qq;
        return (int16_t)qq;
      };
      ++g_checks;
      if (r.y[b][j] != Complex{q(x.re), q(x.im)}) fail("bf unit weight passthrough");
    }
}

void check_cfar_flat() {
  rd_amd::cfar::Config cfg;
  cfg.alpha = 0x100;
  const std::vector<uint64_t> flat(64, 1000);
  std::vector<rd_amd::cfar::Event> ev = rd_amd::cfar::run_frame(flat, cfg, 31, 63, 9);
  ++g_checks;
  if (ev.size() != 1 || ev[0].kind != rd_amd::cfar::kEvSummary || !ev[0].sof || !ev[0].eof || ev[0].frame_id != 9 ||
      ev[0].sup_count != 20)
    fail("cfar flat alpha=1.0");
  cfg.alpha = 0xFF;
  ev = rd_amd::cfar::run_frame(flat, cfg, 31, 63, 9);
  ++g_checks;
  if (ev.size() != 64 - 20 + 1 || ev.back().det_count != 44) fail("cfar flat alpha=255/256");
}

// ---- covar::Integrator: windows, reconfiguration, saturation, exponential,
// the shapes of sim/tests/test_covariance.cpp passes 4-6, with the expected
// values stated from the semantics (sums, clamps, settle bounds).
using rd_amd::covar::Config;
using rd_amd::covar::Integrator;
using rd_amd::covar::Mode;
using rd_amd::covar::Result;

uint64_t g_rng = 0x9E3779B97F4A7C15ull;
int64_t rnd(int64_t lo, int64_t hi) {  // splitmix64, [lo, hi]
  uint64_t z = (g_rng += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return lo + (int64_t)(z % (uint64_t)(hi - lo + 1));
}

void check_integrator_windows() {
  Integrator it(40, 16);
  Config cfg{1, Mode::kBlock, 0, true};
  it.reset(cfg);
  unsigned expect_id = 0;
  for (unsigned len : {1u, 2u, 3u, 5u, 8u, 1u, 16u, 4u}) {
    cfg.window_len = len;
    std::vector<int64_t> win;
    for (unsigned i = 0; i < 3 * len + 2; ++i) {
      const int64_t x = rnd(-500000, 500000);
      Result r;
      const bool emit = it.step(cfg, true, x, false, false, &r);
      win.push_back(x);
      if (emit) {  // a new length latches at a boundary, so window sizes are not asserted here

        int64_t sum = 0;
        for (int64_t v : win) sum += v;
        ++g_checks;
        if (r.acc != sum || r.sample_count != win.size() || r.flushed || r.truncated || r.window_id != expect_id)
          fail("integrator window len " + std::to_string(len) + " sample " + std::to_string(i));
        expect_id = (expect_id + 1) & 0xFFFFu;
        win.clear();
      }
    }
    Result r;
    const bool emit = it.step(cfg, false, 0, true, false, &r);  // flush the partial window
    int64_t sum = 0;
    for (int64_t v : win) sum += v;
    ++g_checks;
    if (emit != !win.empty() || (emit && (r.acc != sum || !r.flushed || !r.truncated || r.sample_count != win.size())))
      fail("integrator flush after len " + std::to_string(len));
    if (it.count() != 0 || it.window_id() != 0) fail("integrator state after flush");
    expect_id = 0;
  }
  // a reconfiguration written mid-window does not apply until the window closes
  cfg.window_len = 4;
  it.reset(cfg);
  Result r;
  for (int i = 0; i < 3; ++i)
    if (it.step(cfg, true, 100 + i, false, false, &r)) fail("integrator closed early");
  Config one = cfg;
  one.window_len = 1;
  ++g_checks;
  if (!it.step(one, true, 103, false, false, &r) || r.sample_count != 4 || r.acc != 406)
    fail("integrator mid-window reconfiguration applied early");
  ++g_checks;
  if (!it.step(one, true, 7, false, false, &r) || r.sample_count != 1 || r.acc != 7)
    fail("integrator reconfiguration not applied at the boundary");
}

void check_integrator_saturation() {
  const int64_t kMax = (int64_t)1 << 31;
  Integrator n34(34, 16);
  n34.reset(Config{8, Mode::kBlock, 0, true});
  for (int sign = 0; sign < 2; ++sign) {
    Result r;
    bool emit = false;
    for (int i = 0; i < 8; ++i) emit = n34.step(Config{8, Mode::kBlock, 0, true}, true, sign ? -kMax : kMax, false,
                                                false, &r);
    const int64_t clamp = sign ? -((int64_t)1 << 33) : ((int64_t)1 << 33) - 1;
    ++g_checks;
    if (!emit || r.acc != clamp) fail("integrator 34-bit clamp, sign " + std::to_string(sign));
  }
  ++g_checks;
  if (!n34.sat_sticky() || n34.sat_count() == 0 || !n34.sat_flags().sat_pos || !n34.sat_flags().sat_neg)
    fail("integrator sticky flags");
  Result r;
  n34.step(Config{8, Mode::kBlock, 0, true}, false, 0, false, true, &r);  // sat_clear
  ++g_checks;
  if (n34.sat_sticky() || n34.sat_count() != 0) fail("integrator sat_clear");

  Integrator w40(40, 16);
  for (unsigned n : {256u, 255u}) {
    const Config c{n, Mode::kBlock, 0, true};
    w40.reset(c);
    std::vector<rd_amd::covar::Step> st(n, rd_amd::covar::Step{c, true, kMax, false, false});
    const auto out = w40.run(st);
    const int64_t expect = n == 256 ? ((int64_t)1 << 39) - 1 : 255 * kMax;
    ++g_checks;
    if (out.size() != 1 || out[0].first != n - 1 || out[0].second.acc != expect)
      fail("integrator 40-bit bound at " + std::to_string(n) + " extreme terms");
  }
}

void check_integrator_exponential() {
  const int64_t hi = 1 << 20, lo = -(1 << 18);
  for (unsigned k = 0; k <= 6; ++k) {
    const Config c{4, Mode::kExp, k, true};
    Integrator it(40, 16);
    it.reset(c);
    const unsigned n = (30u << k) + 64u;
    std::vector<rd_amd::covar::Step> up(n, rd_amd::covar::Step{c, true, hi, false, false});
    std::vector<rd_amd::covar::Step> down(2 * n, rd_amd::covar::Step{c, true, lo, false, false});
    const auto a = it.run(up), b = it.run(down);
    const int64_t q = (int64_t)1 << k;
    ++g_checks;
    if (a.empty() || b.empty() || a.back().second.acc > hi || a.back().second.acc <= hi - q ||
        b.back().second.acc > lo || b.back().second.acc <= lo - q || (k == 0 && b.back().second.acc != lo))
      fail("integrator exponential settle, k = " + std::to_string(k));
  }
}

void check_integrator_step_matches_run() {
  std::vector<rd_amd::covar::Step> st;
  Config c{3, Mode::kBlock, 2, true};
  for (int i = 0; i < 60; ++i) {
    if (rnd(0, 9) == 0) c = Config{(unsigned)rnd(0, 5), rnd(0, 1) ? Mode::kExp : Mode::kBlock, (unsigned)rnd(0, 8), rnd(0, 9) != 0};
    st.push_back(rd_amd::covar::Step{c, rnd(0, 4) != 0, rnd(-(1LL << 34), 1LL << 34), rnd(0, 19) == 0, rnd(0, 19) == 0});
  }
  Integrator a(36, 16), b(36, 16);
  a.reset(st[0].cfg);
  b.reset(st[0].cfg);
  const auto batch = b.run(st);
  size_t k = 0;
  for (size_t i = 0; i < st.size(); ++i) {
    Result r;
    if (a.step(st[i].cfg, st[i].valid, st[i].data, st[i].flush, st[i].sat_clear, &r)) {
      ++g_checks;
      if (k >= batch.size() || batch[k].first != i || batch[k].second != r) fail("integrator step vs run");
      ++k;
    }
  }
  ++g_checks;
  if (k != batch.size() || a.acc() != b.acc() || a.sat_count() != b.sat_count() || !(a.active() == b.active()))
    fail("integrator step vs run final state");
}

// ---- cfar::decide: brute-force window sums (the test_cfar.cpp oracle
// pattern), partial sums at the row ends, and agreement with run_frame.
void check_cfar_decide() {
  std::vector<uint64_t> bins(200);
  for (auto& b : bins) b = (uint64_t)rnd(0, 1 << 16);
  bins[60] = bins[130] = 1 << 22;
  rd_amd::cfar::Config cfg;
  cfg.guard_lead = 2;
  cfg.guard_lag = 3;
  cfg.ref_lead = 8;
  cfg.ref_lag = 5;
  cfg.alpha = 0x300;
  const auto d = rd_amd::cfar::decide_all(bins, cfg);
  unsigned det = 0;
  for (size_t i = 0; i < bins.size(); ++i) {
    uint64_t lead = 0, lag = 0;
    bool ok = true;
    for (unsigned j = 1; j <= cfg.ref_lead && ok; ++j)
      if (i + cfg.guard_lead + j >= bins.size()) ok = false; else lead += bins[i + cfg.guard_lead + j];
    for (unsigned j = 1; j <= cfg.ref_lag && ok; ++j)
      if (i < cfg.guard_lag + j) ok = false; else lag += bins[i - cfg.guard_lag - j];
    ++g_checks;
    if (d[i].window_ok != ok || d[i].sum_lead != lead || d[i].sum_lag != lag || d[i].cut_power != bins[i])
      fail("cfar decide bin " + std::to_string(i));
    const bool hit = ok && ((bins[i] * 13) << 8) > 0x300 * (lead + lag);
    if (d[i].detected != hit || d[i].suppressed != !ok) fail("cfar decide decision bin " + std::to_string(i));
    det += hit;
  }
  const auto ev = rd_amd::cfar::run_frame(bins, cfg, 31, 63, 0);
  ++g_checks;
  if (ev.empty() || ev.back().det_count != det) fail("cfar decide vs run_frame detections");
  ++g_checks;
  if (!(rd_amd::cfar::decide(bins, 60, cfg).detected)) fail("cfar decide single bin");
}

// ---- pfb::LaneModel: one branch stepped sample by sample equals the bank.
void check_lane_model() {
  const unsigned P = 4, T = 8, beats = 12;
  std::vector<Complex> coeff((size_t)P * T), x((size_t)P * beats);
  for (auto& c : coeff) c = Complex{(int16_t)rnd(-32768, 32767), (int16_t)rnd(-32768, 32767)};
  for (auto& v : x) v = Complex{(int16_t)rnd(-32768, 32767), (int16_t)rnd(-32768, 32767)};
  rd_amd::pfb::PfbModel bank(P, T, coeff);
  bank.reset();
  std::vector<rd_amd::pfb::LaneModel> lanes;
  for (unsigned p = 0; p < P; ++p) lanes.emplace_back(T, &coeff[(size_t)p * T]);
  for (unsigned b = 0; b < beats; ++b) {
    const std::vector<Complex> beat(x.begin() + (size_t)b * P, x.begin() + (size_t)(b + 1) * P);
    const rd_amd::pfb::BeatResult br = bank.step(beat);
    for (unsigned p = 0; p < P; ++p) {
      int64_t re = 0, im = 0;
      lanes[p].accumulate(beat[p], &re, &im);
      rd_amd::Flags fr, fi;
      const Complex y = lanes[p].step(beat[p], &fr, &fi);
      auto rne = [](int64_t v) {
        int64_t q = v >> 15;
        const int64_t rem = v & 0x7FFF;
        if (rem > 0x4000 || (rem == 0x4000 && (q & 1))) ++q;
        return q > 32767 ? (int64_t)32767 : (q < -32768 ? (int64_t)-32768 : q);
      };
      ++g_checks;
      if (y != br.y[p] || fr.packed() != br.flags_re[p].packed() || fi.packed() != br.flags_im[p].packed() ||
          rne(re) != y.re || rne(im) != y.im)
        fail("lane model beat " + std::to_string(b) + " branch " + std::to_string(p));
    }
  }
}

void check_device_use() {
  rd_amd::Device::ctx();
  bool threw = false;
  try {
    rd_amd::Device::use(1);
  } catch (const rd_amd::Error&) {
    threw = true;
  }
  ++g_checks;
  if (!threw) fail("Device::use on another device after the context exists must throw");
  rd_amd::Device::use(0);  // the same device is fine
}

}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "tests/golden/ref_vectors";
  try {
    check_fft(dir);
    for (const char* set : {"ident", "max", "mixed", "proto", "random"})
      for (const char* geom : {"p4t8", "p8t16"}) check_pfb(dir, set, geom);
    check_bf_unit();
    check_cfar_flat();
    check_integrator_windows();
    check_integrator_saturation();
    check_integrator_exponential();
    check_integrator_step_matches_run();
    check_cfar_decide();
    check_lane_model();
    check_device_use();
  } catch (const std::exception& e) {
    std::printf("RESULT: FAIL exception %s\n", e.what());
    return 1;
  }
  std::printf("checks: %zu, failures: %zu\n", g_checks, g_fail);
  std::printf("RESULT: %s\n", g_fail ? "FAIL" : "PASS");
  return g_fail ? 1 : 0;
}
