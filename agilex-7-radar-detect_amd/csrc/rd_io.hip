// rd_io.hip — the reference's file formats on the host side of the library, so
// reference-format scenes and coefficient sets feed the GPU path directly
// (rd_read_coeff / rd_read_pfb_vec / rd_read_iq_ascii / rd_write_iq_ascii in
// include/rd_amd.h). Host code only.
//
//   .coeff   pfb::load_coeff   model/cpp/pfb/pfb_model.hpp:357-442
//   .vec     pfb::load_io      model/cpp/pfb/pfb_model.hpp:444-544
//   iq_ascii_v1  writer model/python/gen_target_iq.py:520-617,
//                reader sim/verilator/harness/iq_loader.h
//
// Every file is line oriented: "# key: value" header lines, other "#" lines
// are comments, data rows are whitespace-separated signed decimals. Complex
// values are packed {im, re} (real low); a value outside int16 keeps its low
// 16 bits, as the reference's static_cast<i16> does.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cctype>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "rd_internal.h"

namespace {

uint32_t pack16(long long re, long long im) {
  return ((uint32_t)((uint64_t)im & 0xFFFFu) << 16) | (uint32_t)((uint64_t)re & 0xFFFFu);
}

rd_status fail(char* err, size_t cap, const std::string& msg) {
  if (err && cap) {
    const size_t k = msg.size() < cap - 1 ? msg.size() : cap - 1;
    memcpy(err, msg.data(), k);
    err[k] = 0;
  }
  return RD_E_ARG;
}

// A parsed file: header keys and the data rows (as integers).
struct TextFile {
  std::map<std::string, std::string> hdr;
  std::vector<std::vector<long long>> rows;
};

bool trim_kv(const std::string& line, std::string* k, std::string* v) {
  if (line.empty() || line[0] != '#') return false;
  const size_t colon = line.find(':');
  if (colon == std::string::npos) return false;
  size_t ks = 1, ke = colon, vs = colon + 1, ve = line.size();
  while (ks < ke && isspace((unsigned char)line[ks])) ++ks;
  while (ke > ks && isspace((unsigned char)line[ke - 1])) --ke;
  while (vs < ve && isspace((unsigned char)line[vs])) ++vs;
  while (ve > vs && isspace((unsigned char)line[ve - 1])) --ve;
  *k = line.substr(ks, ke - ks);
  *v = line.substr(vs, ve - vs);
  return true;
}

bool read_text(const char* path, TextFile* out, std::string* err) {
  std::ifstream f(path);
  if (!f) {
    *err = std::string("cannot open ") + path;
    return false;
  }
  std::string line;
  size_t lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::string k, v;
    if (trim_kv(line, &k, &v)) {
      if (!out->hdr.count(k)) out->hdr[k] = v;  // the first occurrence of a key wins
      continue;
    }
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    std::string tok;
    std::vector<long long> row;
    while (is >> tok) {
      char* end = nullptr;
      const long long x = strtoll(tok.c_str(), &end, 10);
      if (end == tok.c_str() || *end != '\0') {
        *err = std::string(path) + ":" + std::to_string(lineno) + ": cannot parse \"" + tok + "\"";
        return false;
      }
      row.push_back(x);
    }
    out->rows.push_back(std::move(row));
  }
  return true;
}

unsigned hdr_u(const TextFile& t, const char* k) {
  auto it = t.hdr.find(k);
  return it == t.hdr.end() ? 0u : (unsigned)strtoul(it->second.c_str(), nullptr, 10);
}
std::string hdr_s(const TextFile& t, const char* k) {
  auto it = t.hdr.find(k);
  return it == t.hdr.end() ? std::string() : it->second;
}

}  // namespace

extern "C" {

rd_status rd_read_coeff(const char* path, uint32_t* coeff_out, size_t cap, uint32_t* phases, uint32_t* taps,
                        char* err, size_t err_cap) {
  if (!path || !phases || !taps) return fail(err, err_cap, "null argument");
  TextFile t;
  std::string e;
  if (!read_text(path, &t, &e)) return fail(err, err_cap, e);
  const unsigned P = hdr_u(t, "phases"), T = hdr_u(t, "taps");
  if (hdr_u(t, "schema") != 1) return fail(err, err_cap, std::string(path) + ": schema is not 1");
  if (hdr_s(t, "kind") != "pfb_coeff") return fail(err, err_cap, std::string(path) + ": kind is not pfb_coeff");
  if (hdr_u(t, "count") != t.rows.size())
    return fail(err, err_cap, std::string(path) + ": declared count differs from the rows held");
  if ((size_t)P * T != t.rows.size()) return fail(err, err_cap, std::string(path) + ": rows != phases x taps");
  for (size_t i = 0; i < t.rows.size(); ++i) {
    const std::vector<long long>& r = t.rows[i];
    if (r.size() != 5) return fail(err, err_cap, std::string(path) + ": a row without 5 fields");
    if (r[0] != (long long)i) return fail(err, err_cap, std::string(path) + ": rows out of index order");
    if (r[0] != r[1] * (long long)T + r[2])
      return fail(err, err_cap, std::string(path) + ": index is not phase*taps + tap");
  }
  *phases = P;
  *taps = T;
  if (!coeff_out || cap < t.rows.size()) return RD_E_CAPACITY;
  for (size_t i = 0; i < t.rows.size(); ++i) coeff_out[i] = pack16(t.rows[i][3], t.rows[i][4]);
  return RD_OK;
}

rd_status rd_read_pfb_vec(const char* path, uint32_t* x, uint32_t* y, uint8_t* f_re, uint8_t* f_im,
                          size_t cap_beats, uint32_t* phases, uint32_t* taps, uint32_t* beats, char* err,
                          size_t err_cap) {
  if (!path || !phases || !taps || !beats) return fail(err, err_cap, "null argument");
  TextFile t;
  std::string e;
  if (!read_text(path, &t, &e)) return fail(err, err_cap, e);
  const unsigned P = hdr_u(t, "phases");
  if (hdr_u(t, "schema") != 1) return fail(err, err_cap, std::string(path) + ": schema is not 1");
  if (hdr_s(t, "kind") != "pfb_io") return fail(err, err_cap, std::string(path) + ": kind is not pfb_io");
  if (P == 0) return fail(err, err_cap, std::string(path) + ": no phases");
  if (hdr_u(t, "count") != t.rows.size())
    return fail(err, err_cap, std::string(path) + ": declared count differs from the rows held");
  for (size_t b = 0; b < t.rows.size(); ++b) {
    if (t.rows[b].size() != 1 + 6 * (size_t)P)
      return fail(err, err_cap, std::string(path) + ": a row without 1 + 6*phases fields");
    if (t.rows[b][0] != (long long)b) return fail(err, err_cap, std::string(path) + ": beats out of order");
  }
  *phases = P;
  *taps = hdr_u(t, "taps");
  *beats = (uint32_t)t.rows.size();
  if (!x || !y || !f_re || !f_im || cap_beats < t.rows.size()) return RD_E_CAPACITY;
  for (size_t b = 0; b < t.rows.size(); ++b) {
    const std::vector<long long>& r = t.rows[b];
    for (unsigned p = 0; p < P; ++p) {
      x[b * P + p] = pack16(r[1 + 2 * p], r[2 + 2 * p]);
      y[b * P + p] = pack16(r[1 + 2 * P + 2 * p], r[2 + 2 * P + 2 * p]);
      f_re[b * P + p] = (uint8_t)(r[1 + 4 * P + 2 * p] & 3);
      f_im[b * P + p] = (uint8_t)(r[2 + 4 * P + 2 * p] & 3);
    }
  }
  return RD_OK;
}

rd_status rd_read_iq_ascii(const char* path, uint32_t* iq, size_t cap, uint32_t* n_ant, uint32_t* fft_size,
                           uint32_t* frames, uint32_t* spc, char* err, size_t err_cap) {
  if (!path || !n_ant || !fft_size || !frames || !spc) return fail(err, err_cap, "null argument");
  TextFile t;
  std::string e;
  if (!read_text(path, &t, &e)) return fail(err, err_cap, e);
  if (hdr_s(t, "format") != "iq_ascii_v1") return fail(err, err_cap, std::string(path) + ": format is not iq_ascii_v1");
  const unsigned A = hdr_u(t, "N_ANTENNAS"), P = hdr_u(t, "SAMPLES_PER_CYCLE"), N = hdr_u(t, "FFT_SIZE");
  const unsigned H = t.hdr.count("total_frames") ? hdr_u(t, "total_frames") : hdr_u(t, "HISTORY_FRAMES");
  if (!A || !P || !N || !H || N % P) return fail(err, err_cap, std::string(path) + ": bad geometry in the header");
  const unsigned bpf = N / P;
  if ((size_t)H * A * bpf != t.rows.size())
    return fail(err, err_cap, std::string(path) + ": rows != frames x antennas x beats per frame");
  *n_ant = A;
  *fft_size = N;
  *frames = H;
  *spc = P;
  const size_t need = (size_t)A * H * N;
  std::vector<uint8_t> seen((size_t)H * A * bpf, 0);
  for (const std::vector<long long>& r : t.rows) {
    if (r.size() != 5 + 2 * (size_t)P) return fail(err, err_cap, std::string(path) + ": a row with the wrong width");
    const long long f = r[0], a = r[1], b = r[2];
    if (f < 0 || f >= H || a < 0 || a >= A || b < 0 || b >= bpf)
      return fail(err, err_cap, std::string(path) + ": frame, antenna or beat out of range");
    if (r[3] != (b == 0 ? 1 : 0) || r[4] != (b == bpf - 1 ? 1 : 0))
      return fail(err, err_cap, std::string(path) + ": sof/eof disagree with the beat index");
    uint8_t& s = seen[((size_t)f * A + a) * bpf + b];
    if (s) return fail(err, err_cap, std::string(path) + ": a (frame, antenna, beat) appears twice");
    s = 1;
  }
  if (!iq || cap < need) return RD_E_CAPACITY;
  for (const std::vector<long long>& r : t.rows) {
    const size_t f = (size_t)r[0], a = (size_t)r[1], b = (size_t)r[2];
    for (unsigned k = 0; k < P; ++k) iq[a * (size_t)H * N + f * N + b * P + k] = pack16(r[5 + 2 * k], r[6 + 2 * k]);
  }
  return RD_OK;
}

rd_status rd_write_iq_ascii(const char* path, const uint32_t* iq, uint32_t n_ant, uint32_t fft_size, uint32_t frames,
                            uint32_t spc, const char* scenario) {
  if (!path || !iq || !n_ant || !fft_size || !frames || !spc || fft_size % spc) return RD_E_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return RD_E_ARG;
  const uint32_t bpf = fft_size / spc;
  fprintf(f, "# target-iq scenario file\n# schema: 1\n# format: iq_ascii_v1\n");
  fprintf(f, "# generator: librd_amd rd_write_iq_ascii\n# scenario: %s\n", scenario ? scenario : "unnamed");
  fprintf(f, "# N_ANTENNAS: %u\n# SAMPLES_PER_CYCLE: %u\n# FFT_SIZE: %u\n# HISTORY_FRAMES: %u\n# SAMPLE_W: 16\n",
          n_ant, spc, fft_size, frames);
  fprintf(f, "# beats_per_frame: %u\n# samples_per_beat: %u\n# total_frames: %u\n#\n# columns: frame antenna beat sof eof",
          bpf, spc, frames);
  for (uint32_t k = 0; k < spc; ++k) fprintf(f, " s%u_re s%u_im", k, k);
  fprintf(f, "\n#\n");
  for (uint32_t fr = 0; fr < frames; ++fr)
    for (uint32_t a = 0; a < n_ant; ++a)
      for (uint32_t b = 0; b < bpf; ++b) {
        fprintf(f, "%u %u %u %u %u", fr, a, b, b == 0 ? 1u : 0u, b == bpf - 1 ? 1u : 0u);
        for (uint32_t k = 0; k < spc; ++k) {
          const uint32_t w = iq[(size_t)a * frames * fft_size + (size_t)fr * fft_size + b * spc + k];
          fprintf(f, " %d %d", (int)(int16_t)(w & 0xFFFFu), (int)(int16_t)(w >> 16));
        }
        fputc('\n', f);
      }
  return fclose(f) == 0 ? RD_OK : RD_E_ARG;
}

}  // extern "C"
