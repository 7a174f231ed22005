// rd_packet.hip — detection packets for the packet fabric (rd_event_flits in
// include/rd_amd.h).
//
// Layout of model/cpp/packet/packet_model.hpp (widths :74-100, detection_flits
// :116-127, hdr_pack :150-160, make_flit :241-260): one packet per 176-bit
// event, a header flit then ceil(176 / packet_w) payload flits carrying the
// event LSB first. Flit bits: 0 odd parity over the flit, 1-2 vc, 3 sof,
// 4 eof, data from bit 5. The header is the low 36 data bits of the header
// flit: dest 4 | src 5 | vc 2 | type 3 (0 = detection) | length 6 (total
// flits) | seq 16. A flit is stored as ceil((packet_w + 5) / 32) little-endian
// 32-bit words, the reference's Flit::w.
//
// The writeback is the receive end, as the pipeline's DMA tap does it
// (rtl/top/benchmark_sim_top.sv:1994-2134): each packet's payload is
// reassembled into the 176-bit event and written zero-extended as one 64-byte
// word at consecutive ring addresses, with the receive checks of
// packet_model::Scoreboard::observe (packet_model.hpp:360-429) counted per
// packet.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rd_internal.h"

namespace rd {

namespace {

// 32 bits of the 176-bit payload starting at bit b (bits at or above `end`,
// and above 176, read as zero).
__device__ __forceinline__ uint32_t payload_bits(uint64_t w0, uint64_t w1, uint64_t w2, uint32_t b, uint32_t end) {
  const uint32_t lim = end < 176u ? end : 176u;
  if (b >= lim) return 0u;
  const uint32_t k = b >> 6, sh = b & 63u;
  const uint64_t lo = k == 0 ? w0 : (k == 1 ? w1 : w2);
  const uint64_t hi = k == 0 ? w1 : (k == 1 ? w2 : 0ull);
  uint64_t v = sh ? ((lo >> sh) | (hi << (64u - sh))) : lo;
  const uint32_t n = lim - b;  // valid bits from b
  if (n < 32) v &= (1ull << n) - 1ull;
  return (uint32_t)v;
}

// One thread per flit.
__global__ __launch_bounds__(256) void event_flits_kernel(const uint64_t* __restrict__ events, uint64_t n_cap,
                                                          const uint64_t* __restrict__ n_dev, uint32_t packet_w,
                                                          uint32_t L, uint32_t fw, uint64_t hdr_base, uint32_t vc,
                                                          uint32_t seq0, uint32_t* __restrict__ out) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t n = n_cap;
  if (n_dev) n = min(n, *n_dev);
  if (g >= n * L) return;
  const uint64_t e = g / L;
  const uint32_t f = (uint32_t)(g % L);
  uint64_t w0 = 0, w1 = 0, w2 = 0;
  if (f == 0) {
    w0 = hdr_base | ((uint64_t)((seq0 + e) & 0xFFFFu) << 20);  // the header as the data field
  } else {
    w0 = events[3 * e];
    w1 = events[3 * e + 1];
    w2 = events[3 * e + 2];
  }
  const uint32_t base = f == 0 ? 0u : (f - 1) * packet_w;  // payload bit of data bit 0
  const uint32_t end = f == 0 ? 36u : base + packet_w;     // first payload bit past this flit
  uint32_t* o = out + g * fw;
  uint32_t ones = 0, word0 = 0;
  for (uint32_t k = 0; k < fw; ++k) {
    // flit bits [32k, 32k + 32) = data bits [32k - 5, 32k + 27)
    uint32_t v;
    if (k == 0) {
      v = (f == 0 ? (uint32_t)w0 & ((1u << 27) - 1u) : payload_bits(w0, w1, w2, base, end) & ((1u << 27) - 1u)) << 5;
      v |= ((vc & 3u) << 1) | ((f == 0 ? 1u : 0u) << 3) | ((f == L - 1 ? 1u : 0u) << 4);
    } else {
      const uint32_t d = 32 * k - 5;
      if (f == 0)
        v = d < 36 ? (uint32_t)(w0 >> d) : 0u;
      else
        v = payload_bits(w0, w1, w2, base + d, end);
    }
    ones += __popc(v);
    if (k == 0)
      word0 = v;
    else
      o[k] = v;
  }
  o[0] = word0 | ((ones & 1u) ? 0u : 1u);  // odd parity
}

// Bits [b, b + 32) of the 176-bit payload (bits at or above 176 read as zero)
// from a packet's payload flits: payload bit q sits in flit 1 + q / packet_w
// at flit bit 5 + q % packet_w. packet_w >= 36, so at most two flits.
__device__ __forceinline__ uint32_t payload_from_flits(const uint32_t* pkt, uint32_t fw, uint32_t packet_w,
                                                       uint32_t b) {
  uint32_t out = 0, got = 0;
  while (got < 32 && b + got < 176) {
    const uint32_t q = b + got, f = 1 + q / packet_w, d = q % packet_w;
    uint32_t take = 32 - got;
    take = min(take, packet_w - d);
    take = min(take, 176u - q);
    const uint32_t* w = pkt + f * fw;
    const uint32_t fb = 5 + d, i = fb >> 5, sh = fb & 31u;
    uint32_t v = w[i] >> sh;
    if (sh && i + 1 < fw) v |= w[i + 1] << (32 - sh);
    if (take < 32) v &= (1u << take) - 1u;
    out |= v << got;
    got += take;
  }
  return out;
}

// One thread per packet: checks, reassembly, one 64-byte ring word.
__global__ __launch_bounds__(256) void event_writeback_kernel(const uint32_t* __restrict__ flits, uint64_t n_cap,
                                                              const uint64_t* __restrict__ n_dev, uint32_t packet_w,
                                                              uint32_t L, uint32_t fw, uint32_t port, uint32_t seq0,
                                                              uint4* __restrict__ ring, uint64_t ring_words,
                                                              uint64_t addr0, unsigned long long* __restrict__ stats) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t n = n_cap;
  if (n_dev) n = min(n, *n_dev);
  const bool live = k < n;
  const uint64_t written = __popcll(__ballot(live));
  if ((threadIdx.x & 63u) == 0 && written) atomicAdd(stats + 0, (unsigned long long)written);
  if (!live) return;
  const uint32_t* pkt = flits + k * L * fw;
  const uint32_t flit_w = packet_w + 5, last_bits = flit_w - 32 * (fw - 1);
  const uint32_t last_mask = last_bits == 32 ? ~0u : (1u << last_bits) - 1u;

  uint32_t parity = 0, framing = 0, vc0 = 0;
  for (uint32_t f = 0; f < L; ++f) {
    const uint32_t* w = pkt + f * fw;
    uint32_t ones = 0;
    for (uint32_t i = 0; i + 1 < fw; ++i) ones += __popc(w[i]);
    ones += __popc(w[fw - 1] & last_mask);
    parity += (ones & 1u) ? 0u : 1u;
    const uint32_t c = w[0];
    const uint32_t vc = (c >> 1) & 3u, sof = (c >> 3) & 1u, eof = (c >> 4) & 1u;
    if (f == 0) {
      vc0 = vc;
      framing += (sof ^ 1u) + eof;
    } else {
      framing += sof + (vc != vc0 ? 1u : 0u) + (eof ^ (f == L - 1 ? 1u : 0u));
    }
  }
  // the 36-bit header: data bits 0..35 of the header flit = flit bits 5..40
  const uint64_t hdr = ((uint64_t)pkt[0] >> 5 | (uint64_t)pkt[1] << 27) & ((1ull << 36) - 1);
  const uint32_t h_dest = (uint32_t)(hdr & 15u), h_vc = (uint32_t)(hdr >> 9) & 3u;
  const uint32_t h_type = (uint32_t)(hdr >> 11) & 7u, h_len = (uint32_t)(hdr >> 14) & 63u;
  const uint32_t h_seq = (uint32_t)(hdr >> 20) & 0xFFFFu;
  framing += (h_len != L ? 1u : 0u) + (h_type != 0 ? 1u : 0u) + (h_vc != vc0 ? 1u : 0u);
  const uint32_t misrouted = h_dest != port ? 1u : 0u;
  const uint32_t seq_break = h_seq != ((seq0 + (uint32_t)k) & 0xFFFFu) ? 1u : 0u;
  if (parity) atomicAdd(stats + 1, (unsigned long long)parity);
  if (framing) atomicAdd(stats + 2, (unsigned long long)framing);
  if (misrouted) atomicAdd(stats + 3, 1ull);
  if (seq_break) atomicAdd(stats + 4, 1ull);

  // a ring shorter than the stream keeps the last ring_words packets, as
  // writing them in order would
  if (k + ring_words < n) return;
  uint32_t e[6];
#pragma unroll
  for (uint32_t j = 0; j < 6; ++j) e[j] = payload_from_flits(pkt, fw, packet_w, 32 * j);
  const uint64_t a = addr0 + k;
  uint4* o = ring + (a % ring_words) * 4;
  o[0] = make_uint4(e[0], e[1], e[2], e[3]);
  o[1] = make_uint4(e[4], e[5], 0u, 0u);
  o[2] = make_uint4(0u, 0u, 0u, 0u);
  o[3] = make_uint4(0u, 0u, 0u, 0u);
}

}  // namespace

hipError_t launch_event_writeback(hipStream_t s, const uint32_t* flits, uint64_t n_cap, const uint64_t* n_dev,
                                  uint32_t packet_w, uint32_t port, uint32_t seq0, uint32_t* ring,
                                  uint64_t ring_words, uint64_t addr0, uint64_t* stats) {
  if (n_cap == 0) return hipSuccess;
  const uint32_t L = 1 + (176 + packet_w - 1) / packet_w, fw = (packet_w + 5 + 31) / 32;
  event_writeback_kernel<<<(uint32_t)((n_cap + 255) / 256), 256, 0, s>>>(
      flits, n_cap, n_dev, packet_w, L, fw, port, seq0, reinterpret_cast<uint4*>(ring), ring_words, addr0 % ring_words,
      reinterpret_cast<unsigned long long*>(stats));
  return hipGetLastError();
}

hipError_t launch_event_flits(hipStream_t s, const uint64_t* events, uint64_t n_cap, const uint64_t* n_dev,
                              uint32_t packet_w, uint32_t dest, uint32_t src, uint32_t vc, uint32_t seq0,
                              uint32_t* flits) {
  const uint32_t L = 1 + (176 + packet_w - 1) / packet_w, fw = (packet_w + 5 + 31) / 32;
  const uint64_t total = n_cap * L;
  if (total == 0) return hipSuccess;
  const uint64_t hdr = (uint64_t)(dest & 15u) | ((uint64_t)(src & 31u) << 4) | ((uint64_t)(vc & 3u) << 9) |
                       ((uint64_t)(L & 63u) << 14);
  event_flits_kernel<<<(uint32_t)((total + 255) / 256), 256, 0, s>>>(events, n_cap, n_dev, packet_w, L, fw, hdr, vc,
                                                                    seq0, flits);
  return hipGetLastError();
}

}  // namespace rd

extern "C" {

uint32_t rd_event_flit_count(uint32_t packet_w) { return packet_w ? 1 + (176 + packet_w - 1) / packet_w : 0; }

uint32_t rd_flit_words(uint32_t packet_w) { return (packet_w + 5 + 31) / 32; }

rd_status rd_event_flits(rd_ctx* c, const uint64_t* events, uint64_t n_events, const uint64_t* n_events_dev,
                         uint32_t packet_w, uint32_t dest, uint32_t src, uint32_t vc, uint32_t seq0, uint32_t* flits) {
  rd_status st = rd::ctx_enter(c);
  if (st != RD_OK) return st;
  if (packet_w < 36 || packet_w > 512 || dest > 15 || src > 31 || vc > 3) return RD_E_ARG;
  if (n_events && (!events || !flits)) return RD_E_ARG;
  // One work-item per flit; a dispatch holds fewer than 2^32 work-items.
  if (n_events > (1ull << 31) || n_events * rd_event_flit_count(packet_w) > (1ull << 31)) return RD_E_ARG;
  return rd::launch_event_flits(c->stream, events, n_events, n_events_dev, packet_w, dest, src, vc, seq0, flits) ==
                 hipSuccess
             ? RD_OK
             : RD_E_HIP;
}

rd_status rd_event_writeback(rd_ctx* c, const uint32_t* flits, uint64_t n_packets, const uint64_t* n_packets_dev,
                             uint32_t packet_w, uint32_t port, uint32_t seq0, uint32_t* ring, uint64_t ring_words,
                             uint64_t addr0, uint64_t* stats) {
  rd_status st = rd::ctx_enter(c);
  if (st != RD_OK) return st;
  if (packet_w < 36 || packet_w > 512 || port > 15 || ring_words == 0 || ring_words > (1ull << 48) || !stats)
    return RD_E_ARG;
  if (n_packets && (!flits || !ring || (reinterpret_cast<uintptr_t>(ring) & 15u))) return RD_E_ARG;
  if (n_packets > (1ull << 31)) return RD_E_ARG;  // one work-item per packet
  return rd::launch_event_writeback(c->stream, flits, n_packets, n_packets_dev, packet_w, port, seq0, ring,
                                    ring_words, addr0, stats) == hipSuccess
             ? RD_OK
             : RD_E_HIP;
}

}  // extern "C"
