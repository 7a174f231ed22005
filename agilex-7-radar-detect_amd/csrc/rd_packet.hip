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

}  // namespace

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
  return rd::launch_event_flits(c->stream, events, n_events, n_events_dev, packet_w, dest, src, vc, seq0, flits) ==
                 hipSuccess
             ? RD_OK
             : RD_E_HIP;
}

}  // extern "C"
