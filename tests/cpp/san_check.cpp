// san_check.cpp — drives the CPU side under -fsanitize=address,undefined
// (tests/cpp/Makefile.san): the oracle's restatement (oracle/rd_oracle.cpp)
// on random and extreme inputs of every stage, the committed fxp vectors
// through it, and the host-side file readers of the library (rd_io.hip:
// .coeff, .vec, iq_ascii_v1) on the reference's fixtures, on truncated and
// corrupted copies of them and on a write -> read round trip. The reference
// keeps its models free of undefined behaviour by construction
// (model/cpp/fxp/fxp.hpp:24-33, 80-110); this is the same claim for the CPU
// code here. Exit 0 when every check holds; a sanitizer report aborts.
//
//   san_check <vector dir> <scratch dir>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rd_amd.h"

extern "C" {
int rdo_fxp_op(const char* op, int64_t a, int64_t b, unsigned sh, unsigned w, int64_t* y, unsigned* flags);
void rdo_acc_run(unsigned acc_w, const int64_t* x, unsigned n, int64_t* y, unsigned* step_flags, unsigned* sticky,
                 uint32_t* count);
void rdo_cmult(uint32_t a, uint32_t b, int64_t* p_re, int64_t* p_im, int64_t* p3_re, int64_t* p3_im, uint32_t* y,
               unsigned* f_re, unsigned* f_im);
void rdo_pfb_run(const uint32_t* x, const uint32_t* coeff, unsigned P, unsigned T, unsigned n_chan, unsigned n_beats,
                 uint32_t* hist, uint32_t* y, uint8_t* flags, unsigned threads);
int rdo_fft_run(const uint32_t* in, unsigned n, unsigned spc, uint32_t sched, int reorder, unsigned n_frames,
                uint32_t* out, uint32_t* bins, uint32_t* stage_flags, unsigned threads);
void rdo_bf_run(const uint32_t* Xa, const uint32_t* W, unsigned CH, unsigned B, size_t cols, uint32_t* y,
                uint8_t* flags, uint32_t* power, int64_t* acc, unsigned threads);
void rdo_cov_integrate(const uint32_t* Xa, unsigned CH, unsigned F, unsigned N, unsigned mode, unsigned exp_k,
                       unsigned acc_w, int64_t* R, uint8_t* sat_flags, unsigned threads);
int64_t rdo_cfar_run(const uint64_t* bins, unsigned n, unsigned n_rows, const uint32_t* cfg8, unsigned max_guard,
                     unsigned max_ref, const uint32_t* frame_ids, uint64_t* events, int64_t events_cap,
                     uint32_t* row_count);
void rdo_cfar_decide(const uint64_t* bins, unsigned n, unsigned n_rows, const uint32_t* cfg8, uint64_t* dec);
int rdo_event_flits(const uint64_t* events, uint64_t n_events, unsigned packet_w, unsigned dest, unsigned src,
                    unsigned vc, unsigned seq0, uint32_t* flits);
int rdo_event_writeback(const uint32_t* flits, uint64_t n, unsigned packet_w, unsigned port, unsigned seq0,
                        uint32_t* ring, uint64_t ring_words, uint64_t addr0, uint64_t* stats);
int64_t rdo_chain_run(const uint32_t* iq, const uint32_t* taps, const uint32_t* W, unsigned CH, unsigned N, unsigned T,
                      unsigned F, unsigned B, unsigned spc, uint32_t sched, const uint32_t* cfar_cfg8,
                      unsigned max_guard, unsigned max_ref, uint32_t* u, uint8_t* pfb_flags, uint32_t* X,
                      uint32_t* fft_flags, uint32_t* Xa, uint32_t* y, uint8_t* yflags, uint32_t* power, int64_t* R,
                      uint64_t* events, int64_t events_cap, uint32_t* row_count, unsigned threads);
uint32_t rdo_integ_run(unsigned acc_w, const uint32_t* cfg4, const uint32_t* steps7, uint32_t n_steps,
                       const int64_t* data, size_t n_lanes, int64_t* lane_acc, uint32_t* lane_flags,
                       uint32_t* lane_count, int64_t* results, uint32_t* win5, uint32_t cap, uint32_t* ctl6);
void rdo_pfb_acc(const uint32_t* x, const uint32_t* coeff, unsigned P, unsigned T, unsigned n_chan, unsigned n_beats,
                 const uint32_t* hist, int64_t* acc);
}

namespace {

size_t g_checks = 0, g_fail = 0;
void check(bool ok, const std::string& what) {
  ++g_checks;
  if (!ok && ++g_fail < 20) std::fprintf(stderr, "FAIL: %s\n", what.c_str());
}

uint64_t g_s = 0x2545F4914F6CDD1Dull;
uint64_t next() {
  g_s ^= g_s << 13;
  g_s ^= g_s >> 7;
  g_s ^= g_s << 17;
  return g_s;
}
uint32_t cx() {  // random packed complex, a quarter of them corner values
  static const int16_t c[] = {-32768, -32767, -1, 0, 1, 32766, 32767};
  if (next() % 4 == 0) return (uint32_t)(uint16_t)c[next() % 7] | ((uint32_t)(uint16_t)c[next() % 7] << 16);
  return (uint32_t)next();
}

std::vector<std::vector<std::string>> rows_of(const std::string& path) {
  std::vector<std::vector<std::string>> out;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    std::vector<std::string> t;
    std::string w;
    while (is >> w) t.push_back(w);
    out.push_back(t);
  }
  return out;
}

void fxp_vectors(const std::string& dir) {
  for (auto& r : rows_of(dir + "/fxp_ops.vec")) {
    int64_t y = 0;
    unsigned fl = 0;
    const bool ok = rdo_fxp_op(r[1].c_str(), std::stoll(r[2]), std::stoll(r[3]), (unsigned)std::stoul(r[4]),
                               (unsigned)std::stoul(r[5]), &y, &fl) != 0;
    check(ok && y == std::stoll(r[6]) && fl == std::stoul(r[7]), "fxp op " + r[0]);
  }
  std::vector<int64_t> xs;
  std::string cur;
  auto run = [&](const std::vector<std::vector<std::string>>& seq) {
    std::vector<int64_t> x, y(seq.size());
    std::vector<unsigned> sf(seq.size()), st(seq.size());
    std::vector<uint32_t> cnt(seq.size());
    for (auto& s : seq) x.push_back(std::stoll(s[3]));
    rdo_acc_run((unsigned)std::stoul(seq[0][2]), x.data(), (unsigned)x.size(), y.data(), sf.data(), st.data(),
                cnt.data());
    for (size_t i = 0; i < seq.size(); ++i)
      check(y[i] == std::stoll(seq[i][4]) && sf[i] == std::stoul(seq[i][5]) && st[i] == std::stoul(seq[i][6]) &&
                cnt[i] == std::stoul(seq[i][7]),
            "accum " + seq[i][0]);
  };
  std::vector<std::vector<std::string>> seq;
  for (auto& r : rows_of(dir + "/fxp_accum.vec")) {
    if (!seq.empty() && r[0] != seq[0][0]) {
      run(seq);
      seq.clear();
    }
    seq.push_back(r);
  }
  if (!seq.empty()) run(seq);
  for (auto& r : rows_of(dir + "/cmult.vec")) {
    auto pk = [](const std::string& re, const std::string& im) {
      return (uint32_t)(uint16_t)std::stoi(re) | ((uint32_t)(uint16_t)std::stoi(im) << 16);
    };
    int64_t pr, pi, p3r, p3i;
    uint32_t y;
    unsigned fr, fi;
    rdo_cmult(pk(r[1], r[2]), pk(r[3], r[4]), &pr, &pi, &p3r, &p3i, &y, &fr, &fi);
    check(pr == std::stoll(r[5]) && pi == std::stoll(r[6]) && p3r == pr && p3i == pi && y == pk(r[7], r[8]) &&
              fr == std::stoul(r[9]) && fi == std::stoul(r[10]),
          "cmult " + r[0]);
  }
}

void oracle_stages() {
  // PFB with history, FFT at every length and both lane counts, beamformer
  for (unsigned T : {1u, 5u, 16u, 64u}) {
    const unsigned P = 32, nb = 9, ch = 2;
    std::vector<uint32_t> x(ch * nb * P), h(P * T), hist(ch * (T > 1 ? T - 1 : 0) * P + 1), y(x.size());
    std::vector<uint8_t> fl(x.size());
    for (auto& v : x) v = cx();
    for (auto& v : h) v = cx();
    for (auto& v : hist) v = cx();
    rdo_pfb_run(x.data(), h.data(), P, T, ch, nb, T > 1 ? hist.data() : nullptr, y.data(), fl.data(), 2);
    std::vector<int64_t> acc(2 * x.size());
    rdo_pfb_acc(x.data(), h.data(), P, T, ch, nb, nullptr, acc.data());
  }
  for (unsigned n = 8; n <= 4096; n *= 2)
    for (unsigned spc : {1u, 2u}) {
      const unsigned fr = 3;
      std::vector<uint32_t> in(n * fr), out(in.size()), bins(in.size()), sf(fr);
      for (auto& v : in) v = cx();
      const int rc = rdo_fft_run(in.data(), n, spc, (uint32_t)next(), (int)(next() & 1), fr, out.data(), bins.data(),
                                 sf.data(), 2);
      check(rc == (n / spc >= 4 ? 0 : -1), "fft rc n=" + std::to_string(n));
    }
  for (unsigned CH : {1u, 7u, 64u}) {
    const unsigned B = 9;
    const size_t cols = 33;
    std::vector<uint32_t> xa(cols * CH), w(B * CH), y(B * cols), pw(B * cols);
    std::vector<uint8_t> fl(B * cols);
    std::vector<int64_t> acc(2 * B * cols);
    for (auto& v : xa) v = cx();
    for (auto& v : w) v = cx();
    rdo_bf_run(xa.data(), w.data(), CH, B, cols, y.data(), fl.data(), pw.data(), acc.data(), 2);
    const unsigned F = 11, N = 3;
    std::vector<uint32_t> xc(F * N * CH);
    for (auto& v : xc) v = cx();
    std::vector<int64_t> R(N * CH * (CH + 1));
    std::vector<uint8_t> sfl(N * CH * (CH + 1) / 2);
    for (unsigned mode : {0u, 1u})
      for (unsigned aw : {2u, 33u, 64u}) rdo_cov_integrate(xc.data(), CH, F, N, mode, 63, aw, R.data(), sfl.data(), 2);
  }
  // CFAR: CA/GO, dense, disabled, clamped and huge geometries, u64 extremes
  const uint32_t cfgs[][8] = {{1, 0, 0, 2, 2, 8, 8, 0x800},     {1, 1, 1, 4, 1, 2, 6, 0x300},
                              {0, 0, 0, 2, 2, 8, 8, 0x800},     {1, 0, 1, 0, 0, 0, 0, 0x100},
                              {1, 1, 0, 99, 0xFFFFFFFF, 70, 70, 0xFFFF}};
  for (auto& c : cfgs)
    for (unsigned n : {1u, 7u, 100u, 1024u}) {
      const unsigned rows = 3;
      std::vector<uint64_t> bins(n * rows), ev(3 * (size_t)rows * (n + 1)), dec(8 * (size_t)n * rows);
      std::vector<uint32_t> rc(rows);
      for (auto& b : bins) b = next() % 3 == 0 ? next() : next() % 100000;
      const int64_t t = rdo_cfar_run(bins.data(), n, rows, c, 31, 63, nullptr, ev.data(), (int64_t)rows * (n + 1),
                                     rc.data());
      check(t >= rows, "cfar total");
      rdo_cfar_decide(bins.data(), n, rows, c, dec.data());
      if (t > 0) {
        for (unsigned pw : {36u, 64u, 512u}) {
          const unsigned L = 1 + (176 + pw - 1) / pw, fw = (pw + 5 + 31) / 32;
          std::vector<uint32_t> flits((size_t)t * L * fw), ring(16 * (size_t)t);
          rdo_event_flits(ev.data(), (uint64_t)t, pw, 3, 7, 1, 100, flits.data());
          for (size_t k = 0; k < flits.size(); k += 7) flits[k] ^= (uint32_t)next();  // corrupt some flits
          uint64_t stats[5] = {0, 0, 0, 0, 0};
          rdo_event_writeback(flits.data(), (uint64_t)t, pw, 3, 100, ring.data(), (uint64_t)t, 5, stats);
        }
      }
    }
  // the whole chain on a tiny CPI
  {
    const unsigned CH = 4, N = 64, T = 8, F = 16, B = 4;
    std::vector<uint32_t> iq(CH * N * F), taps(N * T), W(B * CH), u(iq.size()), X(iq.size()), Xa(iq.size()),
        y(B * N * F), pw(B * N * F), ff(CH * F), rc(B * F);
    std::vector<uint8_t> pf(iq.size()), yf(B * N * F);
    std::vector<int64_t> R(N * CH * (CH + 1));
    std::vector<uint64_t> ev(3 * (size_t)B * F * (N + 1));
    for (auto& v : iq) v = cx();
    for (auto& v : taps) v = cx();
    for (auto& v : W) v = cx();
    const int64_t t = rdo_chain_run(iq.data(), taps.data(), W.data(), CH, N, T, F, B, 2, 0xFFFFFFFF, cfgs[0], 31, 63,
                                    u.data(), pf.data(), X.data(), ff.data(), Xa.data(), y.data(), yf.data(),
                                    pw.data(), R.data(), ev.data(), (int64_t)B * F * (N + 1), rc.data(), 2);
    check(t >= B * F, "chain total");
  }
  // Integrator lanes over a random control stream
  {
    const uint32_t n = 300, lanes = 5, cap = 400;
    std::vector<uint32_t> steps(7 * n), lf(lanes), lc(lanes), win(5 * cap), ctl(6);
    std::vector<int64_t> data(n * lanes), la(lanes), res((size_t)cap * lanes);
    for (uint32_t s = 0; s < n; ++s) {
      uint32_t* t = &steps[7 * s];
      t[0] = next() % 7;
      t[1] = next() % 2;
      t[2] = next() % 80;
      t[3] = next() % 10 != 0;
      t[4] = next() % 5 != 0;
      t[5] = next() % 17 == 0;
      t[6] = next() % 13 == 0;
    }
    for (auto& d : data) d = (int64_t)next();
    const uint32_t cfg4[4] = {3, 1, 70, 1};
    for (unsigned aw : {0u, 1u, 2u, 40u, 63u, 64u})
      rdo_integ_run(aw, cfg4, steps.data(), n, data.data(), lanes, la.data(), lf.data(), lc.data(), res.data(),
                    win.data(), cap, ctl.data());
  }
}

std::string slurp(const std::string& p) {
  std::ifstream f(p, std::ios::binary);
  return std::string((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}
void spit(const std::string& p, const std::string& s) {
  std::ofstream f(p, std::ios::binary);
  f << s;
}

// The readers must accept the fixtures and reject (never crash on) damaged copies.
void readers(const std::string& dir, const std::string& tmp) {
  char err[256];
  for (const char* set : {"ident", "max", "mixed", "proto", "random"})
    for (const char* geom : {"p4t8", "p8t16"}) {
      const std::string c = dir + "/pfb_" + set + "_" + geom + ".coeff", v = dir + "/pfb_" + set + "_" + geom + ".vec";
      uint32_t P = 0, T = 0, nb = 0;
      check(rd_read_coeff(c.c_str(), nullptr, 0, &P, &T, err, sizeof(err)) == RD_E_CAPACITY, "coeff sizes");
      std::vector<uint32_t> co((size_t)P * T);
      check(rd_read_coeff(c.c_str(), co.data(), co.size(), &P, &T, err, sizeof(err)) == RD_OK, "coeff read");
      check(rd_read_pfb_vec(v.c_str(), nullptr, nullptr, nullptr, nullptr, 0, &P, &T, &nb, err, sizeof(err)) ==
                RD_E_CAPACITY,
            "vec sizes");
      std::vector<uint32_t> x((size_t)nb * P), y(x.size());
      std::vector<uint8_t> fr(x.size()), fi(x.size());
      check(rd_read_pfb_vec(v.c_str(), x.data(), y.data(), fr.data(), fi.data(), nb, &P, &T, &nb, err, sizeof(err)) ==
                RD_OK,
            "vec read");
      for (const std::string& src : {c, v}) {
        const std::string body = slurp(src);
        for (int k = 0; k < 12; ++k) {
          std::string bad = body;
          switch (k % 4) {
            case 0: bad.resize(next() % (bad.size() + 1)); break;                  // truncated
            case 1: bad[next() % bad.size()] = (char)(next() & 0xFF); break;       // one byte damaged
            case 2: bad.insert(next() % bad.size(), "\n# count: 99999999999\n"); break;
            default: bad.insert(next() % bad.size(), " -99999999999999999999 "); break;
          }
          const std::string p = tmp + "/bad_input";
          spit(p, bad);
          std::vector<uint32_t> b1(4096), b2(4096);
          std::vector<uint8_t> b3(4096), b4(4096);
          uint32_t a1 = 0, a2 = 0, a3 = 0;
          (void)rd_read_coeff(p.c_str(), b1.data(), b1.size(), &a1, &a2, err, sizeof(err));
          (void)rd_read_pfb_vec(p.c_str(), b1.data(), b2.data(), b3.data(), b4.data(), 16, &a1, &a2, &a3, err,
                                sizeof(err));
          (void)rd_read_iq_ascii(p.c_str(), b1.data(), b1.size(), &a1, &a2, &a3, &a3, err, sizeof(err));
        }
      }
    }
  // iq_ascii_v1: write -> read round trip, then damaged copies
  const uint32_t A = 3, N = 16, H = 4;
  for (uint32_t spc : {1u, 2u}) {
    std::vector<uint32_t> iq(A * N * H), back(iq.size());
    for (auto& v : iq) v = cx();
    const std::string p = tmp + "/scene.iq";
    check(rd_write_iq_ascii(p.c_str(), iq.data(), A, N, H, spc, "sanitizer") == RD_OK, "iq write");
    uint32_t a = 0, n = 0, h = 0, s = 0;
    check(rd_read_iq_ascii(p.c_str(), back.data(), back.size(), &a, &n, &h, &s, err, sizeof(err)) == RD_OK &&
              back == iq && a == A && n == N && h == H && s == spc,
          "iq round trip");
    const std::string body = slurp(p);
    for (int k = 0; k < 16; ++k) {
      std::string bad = body;
      if (k % 2)
        bad.resize(next() % (bad.size() + 1));
      else
        bad[next() % bad.size()] = "0123456789 -x\n"[next() % 14];
      spit(tmp + "/bad_scene.iq", bad);
      (void)rd_read_iq_ascii((tmp + "/bad_scene.iq").c_str(), back.data(), back.size(), &a, &n, &h, &s, err,
                             sizeof(err));
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "tests/golden/ref_vectors";
  const std::string tmp = argc > 2 ? argv[2] : "/tmp";
  fxp_vectors(dir);
  oracle_stages();
  readers(dir, tmp);
  std::printf("checks: %zu, failures: %zu\nRESULT: %s\n", g_checks, g_fail, g_fail ? "FAIL" : "PASS");
  return g_fail ? 1 : 0;
}
