// test_mirror.cpp — a reference-style harness over the C++ mirror (rd_amd.hpp).
//
// Shaped like the reference's own gates (model/cpp/test/test_fft_ref.cpp:60-90,
// the PFB vector check of sim/tests/test_pfb_bank.cpp, test_beamformer.cpp's
// unit-weight case, test_cfar.cpp's flat-spectrum cases), but every call goes
// through rd_amd::pfb / fft / bf / cfar on the GPU. It reads the reference's
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

}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "tests/golden/ref_vectors";
  try {
    check_fft(dir);
    for (const char* set : {"ident", "max", "mixed", "proto", "random"})
      for (const char* geom : {"p4t8", "p8t16"}) check_pfb(dir, set, geom);
    check_bf_unit();
    check_cfar_flat();
  } catch (const std::exception& e) {
    std::printf("RESULT: FAIL exception %s\n", e.what());
    return 1;
  }
  std::printf("checks: %zu, failures: %zu\n", g_checks, g_fail);
  std::printf("RESULT: %s\n", g_fail ? "FAIL" : "PASS");
  return g_fail ? 1 : 0;
}
