// rd_plain.hip — the beamformer and covariance without matrix cores: the
// "plain GPU" comparison point of the chain's measurement (rd_bf_plain_run,
// rd_cov_plain_run, rd_chain_enqueue_plain in include/rd_amd.h).
//
// Same values as the split-int8 MFMA kernels, computed the direct way: every
// int16 x int16 product in a 32-bit multiply (|product| <= 2^30, exact),
// summed in int64, rounded once. Tiles are staged in LDS so each operand is
// read from HBM once per tile; nothing else is specialised.
//   beamformer: y = round_sat(sum_a x[a] * W[b][a], 15, 16), no conjugate
//               (model/cpp/beamformer/beamformer_model.hpp:162-179), and
//               power = yr^2 + yi^2 (model/cpp/covariance/covar_model.hpp:99-102)
//   covariance: R[k][i<=j] = sum_m x_i * conj(x_j)
//               (covar_model.hpp:137-151), the exact window sum of a
//               block-mode Integrator at acc_w = 64 (covar_model.hpp:186-262)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rd_fxp.h"
#include "rd_internal.h"

namespace rd {

constexpr uint32_t kPlainCols = 64;  // beamformer columns per block

// Block: 64 columns; thread t owns column t % 64 and beams t / 64, +4, ...
__global__ __launch_bounds__(256) void plain_bf_kernel(const uint32_t* __restrict__ Xa,
                                                       const uint32_t* __restrict__ W, uint32_t CH, uint32_t B,
                                                       uint64_t cols, uint32_t* __restrict__ y,
                                                       uint8_t* __restrict__ flags, uint32_t* __restrict__ power) {
  extern __shared__ uint32_t lds[];
  uint32_t* w = lds;                 // [B][CH]
  uint32_t* xt = lds + B * CH;       // [64][CH + 1]
  const uint32_t rs = CH + 1;
  const uint64_t c0 = (uint64_t)blockIdx.x * kPlainCols;
  const uint32_t ncol = (uint32_t)min<uint64_t>(kPlainCols, cols - c0);
  for (uint32_t i = threadIdx.x; i < B * CH; i += blockDim.x) w[i] = W[i];
  for (uint32_t i = threadIdx.x; i < ncol * CH; i += blockDim.x) xt[(i / CH) * rs + i % CH] = Xa[c0 * CH + i];
  __syncthreads();
  const uint32_t c = threadIdx.x % kPlainCols;
  if (c >= ncol) return;
  const uint64_t col = c0 + c;
  const uint32_t* xc = xt + c * rs;
  for (uint32_t b = threadIdx.x / kPlainCols; b < B; b += blockDim.x / kPlainCols) {
    const uint32_t* wb = w + b * CH;
    int64_t re = 0, im = 0;
    for (uint32_t a = 0; a < CH; ++a) {
      const uint32_t x = xc[a], v = wb[a];
      const int32_t xr = cre(x), xi = cim(x), wr = cre(v), wi = cim(v);
      re += (int64_t)(xr * wr);
      re -= (int64_t)(xi * wi);
      im += (int64_t)(xr * wi);
      im += (int64_t)(xi * wr);
    }
    uint32_t fr = 0, fi = 0;
    const int32_t yr = sat16_i64(rne15_i64(re), fr), yi = sat16_i64(rne15_i64(im), fi);
    const size_t o = (size_t)b * cols + col;
    if (y) y[o] = cpack(yr, yi);
    if (flags) flags[o] = (uint8_t)(fr | (fi << 2));
    if (power) power[o] = (uint32_t)(yr * yr) + (uint32_t)(yi * yi);
  }
}

constexpr uint32_t kPlainFrames = 32;  // covariance frames staged per step

// Block: one bin; thread t owns the pairs t, t + 256, ... (upper triangle,
// row-major) and walks the frames 32 at a time through LDS.
template <uint32_t MAXP>
__global__ __launch_bounds__(256) void plain_cov_kernel(const uint32_t* __restrict__ Xa, uint32_t CH, uint32_t F,
                                                        uint32_t N, int64_t* __restrict__ R) {
  __shared__ uint32_t xs[kPlainFrames * 64];  // [frame][channel]
  const uint32_t k = blockIdx.x;
  const uint32_t npairs = CH * (CH + 1) / 2;
  uint32_t pi[MAXP], pj[MAXP];
  int64_t re[MAXP], im[MAXP];
#pragma unroll
  for (uint32_t q = 0; q < MAXP; ++q) {
    uint32_t p = threadIdx.x + q * 256, i = 0;
    if (p < npairs)
      while (p >= CH - i) {
        p -= CH - i;
        ++i;
      }
    pi[q] = i;
    pj[q] = i + p;
    re[q] = 0;
    im[q] = 0;
  }
  for (uint32_t m0 = 0; m0 < F; m0 += kPlainFrames) {
    const uint32_t nf = min(kPlainFrames, F - m0);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nf * CH; i += blockDim.x) {
      const uint32_t f = i / CH, a = i % CH;
      xs[f * CH + a] = Xa[((size_t)(m0 + f) * N + k) * CH + a];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < MAXP; ++q) {
      if (threadIdx.x + q * 256 >= npairs) continue;
      for (uint32_t f = 0; f < nf; ++f) {
        const uint32_t x = xs[f * CH + pi[q]], v = xs[f * CH + pj[q]];
        const int32_t xr = cre(x), xi = cim(x), vr = cre(v), vi = cim(v);
        // x * conj(v)
        re[q] += (int64_t)(xr * vr);
        re[q] += (int64_t)(xi * vi);
        im[q] += (int64_t)(xi * vr);
        im[q] -= (int64_t)(xr * vi);
      }
    }
  }
#pragma unroll
  for (uint32_t q = 0; q < MAXP; ++q) {
    const uint32_t p = threadIdx.x + q * 256;
    if (p >= npairs) continue;
    int64_t* o = R + ((size_t)k * npairs + p) * 2;
    o[0] = re[q];
    o[1] = im[q];
  }
}

hipError_t launch_bf_plain(hipStream_t s, const uint32_t* Xa, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                           uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power) {
  if (n_cols == 0) return hipSuccess;
  const uint64_t blocks = (n_cols + kPlainCols - 1) / kPlainCols;
  if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
  const size_t lds = ((size_t)n_beams * n_ant + (size_t)kPlainCols * (n_ant + 1)) * sizeof(uint32_t);
  plain_bf_kernel<<<(uint32_t)blocks, 256, lds, s>>>(Xa, W, n_ant, n_beams, n_cols, y, flags, power);
  return hipGetLastError();
}

hipError_t launch_cov_plain(hipStream_t s, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                            int64_t* R) {
  if (n_bins == 0) return hipSuccess;
  const uint32_t npairs = n_chan * (n_chan + 1) / 2;  // <= 2080
  const uint32_t per = (npairs + 255) / 256;
  if (per <= 1)
    plain_cov_kernel<1><<<n_bins, 256, 0, s>>>(Xa, n_chan, n_frames, n_bins, R);
  else if (per <= 3)
    plain_cov_kernel<3><<<n_bins, 256, 0, s>>>(Xa, n_chan, n_frames, n_bins, R);
  else
    plain_cov_kernel<9><<<n_bins, 256, 0, s>>>(Xa, n_chan, n_frames, n_bins, R);
  return hipGetLastError();
}

}  // namespace rd
