// rd_shadow.hip — FP32 float shadow of PFB -> FFT -> beamformer -> power, with
// hand-written backward kernels (include/rd_shadow.h).
//
// The shadow is the float model of the bit-exact chain that gradients flow
// through: the same linear maps without rounding or saturation. Complex values
// are interleaved float pairs (re, im). Gradients follow the convention
// grad(z) = dL/dRe(z) + i dL/dIm(z) for a real loss L, under which a linear map
// y = A z back-propagates as grad(z) = A^H grad(y).
//
//   PFB   u[r][m][p] = sum_k h[p][k] x[r][(m-k)N + p]       (pfb_model.hpp:139-168, no rounding)
//   FFT   X[k] = scale * sum_n u[n] exp(-+2 pi i nk/N)      (the fixed-point transform with every
//         stage scaled is the DFT / N: fft_ref.hpp:250-407)
//   BF    y[b][col] = sum_c W[b][c] X[c][col]               (beamformer_model.hpp:162-179, no conj)
//   power p = |y|^2                                         (covar_model.hpp:99-102)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rd_shadow.h"

namespace {

struct cf {
  float re, im;
};
__device__ __forceinline__ cf ld(const float* p, size_t i) {
  const float2 v = reinterpret_cast<const float2*>(p)[i];
  return cf{v.x, v.y};
}
__device__ __forceinline__ void st(float* p, size_t i, cf v) { reinterpret_cast<float2*>(p)[i] = make_float2(v.re, v.im); }
__device__ __forceinline__ void cfma(cf& acc, cf a, cf b) {
  acc.re = fmaf(a.re, b.re, fmaf(-a.im, b.im, acc.re));
  acc.im = fmaf(a.re, b.im, fmaf(a.im, b.re, acc.im));
}
// acc += conj(a) * b
__device__ __forceinline__ void cjfma(cf& acc, cf a, cf b) {
  acc.re = fmaf(a.re, b.re, fmaf(a.im, b.im, acc.re));
  acc.im = fmaf(a.re, b.im, fmaf(-a.im, b.re, acc.im));
}

rd_status fin(hipError_t e) { return e == hipSuccess ? RD_OK : RD_E_HIP; }

// ---------------------------------------------------------------- PFB
// One thread per output sample (row r, beat m, branch p).
__global__ __launch_bounds__(256) void sh_pfb_kernel(const float* __restrict__ x, const float* __restrict__ h,
                                                     uint32_t N, uint32_t T, uint32_t F, uint64_t total,
                                                     float* __restrict__ u) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint32_t p = (uint32_t)(i % N);
  const uint64_t rm = i / N;
  const uint32_t m = (uint32_t)(rm % F);
  const uint64_t r = rm / F;
  const size_t xr = (size_t)r * F * N;
  cf acc{0.f, 0.f};
  for (uint32_t k = 0; k < T && k <= m; ++k) cfma(acc, ld(h, (size_t)p * T + k), ld(x, xr + (size_t)(m - k) * N + p));
  st(u, i, acc);
}

// grad x[r][j N + p] = sum_{k < T, j + k < F} conj(h[p][k]) grad u[r][j + k][p]
__global__ __launch_bounds__(256) void sh_pfb_bwd_kernel(const float* __restrict__ gu, const float* __restrict__ h,
                                                         uint32_t N, uint32_t T, uint32_t F, uint64_t total,
                                                         float* __restrict__ gx) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint32_t p = (uint32_t)(i % N);
  const uint64_t rj = i / N;
  const uint32_t j = (uint32_t)(rj % F);
  const uint64_t r = rj / F;
  const size_t ur = (size_t)r * F * N;
  cf acc{0.f, 0.f};
  for (uint32_t k = 0; k < T && j + k < F; ++k) cjfma(acc, ld(h, (size_t)p * T + k), ld(gu, ur + (size_t)(j + k) * N + p));
  st(gx, i, acc);
}

// ---------------------------------------------------------------- FFT
// Radix-2 decimation in time in LDS: bit-reversed load, log2(N) in-place
// stages, natural-order output. A block holds FPB = max(1, 512 / N) frames.
// tw[j] = exp(-2 pi i j / N) for j < N/2; `inverse` conjugates it.
__global__ __launch_bounds__(256) void sh_fft_kernel(const float* __restrict__ in, const float* __restrict__ tw,
                                                     uint32_t log2n, uint32_t n_frames, int inverse, float scale,
                                                     float* __restrict__ out) {
  extern __shared__ float2 sm[];
  const uint32_t N = 1u << log2n, half = N / 2;
  const uint32_t fpb = N >= 512 ? 1 : 512 / N;
  const uint32_t f0 = blockIdx.x * fpb;
  const uint32_t nf = min(fpb, n_frames - f0);
  float2* tws = sm;           // [N/2]
  float2* v = sm + half;      // [fpb][N]
  for (uint32_t j = threadIdx.x; j < half; j += blockDim.x) {
    float2 w = reinterpret_cast<const float2*>(tw)[j];
    if (inverse) w.y = -w.y;
    tws[j] = w;
  }
  for (uint32_t idx = threadIdx.x; idx < nf * N; idx += blockDim.x) {
    const uint32_t lf = idx / N, n = idx % N;
    const uint32_t rn = __brev(n) >> (32 - log2n);
    v[lf * N + rn] = reinterpret_cast<const float2*>(in)[(size_t)(f0 + lf) * N + n];
  }
  __syncthreads();
  for (uint32_t s = 0; s < log2n; ++s) {
    const uint32_t hs = 1u << s;  // half size of the butterfly groups
    for (uint32_t b = threadIdx.x; b < nf * half; b += blockDim.x) {
      const uint32_t lf = b / half, j = b % half;
      const uint32_t pos = j & (hs - 1), grp = j >> s;
      const uint32_t i0 = lf * N + grp * 2 * hs + pos, i1 = i0 + hs;
      const float2 w = tws[pos << (log2n - 1 - s)];
      const float2 a = v[i0], c = v[i1];
      const float2 t = make_float2(w.x * c.x - w.y * c.y, w.x * c.y + w.y * c.x);
      v[i0] = make_float2(a.x + t.x, a.y + t.y);
      v[i1] = make_float2(a.x - t.x, a.y - t.y);
    }
    __syncthreads();
  }
  for (uint32_t idx = threadIdx.x; idx < nf * N; idx += blockDim.x) {
    const float2 r = v[idx];
    reinterpret_cast<float2*>(out)[(size_t)f0 * N + idx] = make_float2(r.x * scale, r.y * scale);
  }
}

// ---------------------------------------------------------------- beamformer + power
// Forward: one thread per column; the column's channels stay in registers and
// the weights are broadcast from LDS.
template <int MAXC>
__global__ __launch_bounds__(256) void sh_bf_kernel(const float* __restrict__ X, const float* __restrict__ W,
                                                    uint32_t C, uint32_t B, uint64_t cols, uint32_t n_cpi,
                                                    float* __restrict__ y, float* __restrict__ p) {
  extern __shared__ float2 ws[];  // [B][C]
  for (uint32_t i = threadIdx.x; i < B * C; i += blockDim.x) ws[i] = reinterpret_cast<const float2*>(W)[i];
  __syncthreads();
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= cols * n_cpi) return;
  const uint64_t q = g / cols, col = g % cols;
  const size_t xb = (size_t)q * C * cols, yb = (size_t)q * B * cols;
  cf xr[MAXC];
#pragma unroll
  for (int c = 0; c < MAXC; ++c) xr[c] = c < (int)C ? ld(X, xb + (size_t)c * cols + col) : cf{0.f, 0.f};
  for (uint32_t b = 0; b < B; ++b) {
    cf acc{0.f, 0.f};
#pragma unroll
    for (int c = 0; c < MAXC; ++c)
      if (c < (int)C) {
        const float2 w = ws[b * C + c];
        cfma(acc, cf{w.x, w.y}, xr[c]);
      }
    if (y) st(y, yb + (size_t)b * cols + col, acc);
    p[yb + (size_t)b * cols + col] = acc.re * acc.re + acc.im * acc.im;
  }
}

// Backward with y recomputed from X and W (nothing of the forward pass is
// kept but X): g_y = 2 g_p y; g_X = W^H g_y; g_W = sum over columns of
// g_y X^H. A block walks column chunks of 64 with a grid stride; its partial
// g_W (one (b, c) pair set per thread) goes to work[block][B][C], and a second
// kernel sums the partials in block order (deterministic).
constexpr uint32_t kShChunk = 64;

__global__ __launch_bounds__(256) void sh_bf_bwd_kernel(const float* __restrict__ X, const float* __restrict__ W,
                                                        const float* __restrict__ gp, uint32_t C, uint32_t B,
                                                        uint64_t cols, uint32_t n_cpi, float* __restrict__ gX,
                                                        float* __restrict__ work) {
  extern __shared__ float2 sm[];
  float2* ws = sm;                      // [B][C]
  float2* xs = ws + B * C;              // [C][64]
  float2* gys = xs + C * kShChunk;      // [B][64]
  for (uint32_t i = threadIdx.x; i < B * C; i += blockDim.x) ws[i] = reinterpret_cast<const float2*>(W)[i];
  const uint32_t pairs = B * C, ppt = (pairs + 255) / 256;  // pairs per thread, <= 16
  cf gw[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) gw[k] = cf{0.f, 0.f};
  const uint64_t chunks_per_cpi = (cols + kShChunk - 1) / kShChunk;
  const uint64_t n_chunks = chunks_per_cpi * n_cpi;
  const uint32_t lc = threadIdx.x % kShChunk, lr = threadIdx.x / kShChunk;  // 4 row lanes
  for (uint64_t ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {
    const uint64_t q = ch / chunks_per_cpi, c0 = (ch % chunks_per_cpi) * kShChunk;
    const uint32_t ncol = (uint32_t)min<uint64_t>(kShChunk, cols - c0);
    const size_t xb = (size_t)q * C * cols, yb = (size_t)q * B * cols;
    __syncthreads();  // previous chunk consumed (and the weights staged)
    for (uint32_t c = lr; c < C; c += 4)
      xs[c * kShChunk + lc] = lc < ncol ? reinterpret_cast<const float2*>(X)[xb + (size_t)c * cols + c0 + lc]
                                        : make_float2(0.f, 0.f);
    __syncthreads();
    for (uint32_t b = lr; b < B; b += 4) {
      cf acc{0.f, 0.f};
      for (uint32_t c = 0; c < C; ++c) {
        const float2 w = ws[b * C + c], xv = xs[c * kShChunk + lc];
        cfma(acc, cf{w.x, w.y}, cf{xv.x, xv.y});
      }
      const float g = lc < ncol ? 2.f * gp[yb + (size_t)b * cols + c0 + lc] : 0.f;
      gys[b * kShChunk + lc] = make_float2(g * acc.re, g * acc.im);
    }
    __syncthreads();
    if (gX) {
      for (uint32_t c = lr; c < C; c += 4) {
        cf acc{0.f, 0.f};
        for (uint32_t b = 0; b < B; ++b) {
          const float2 w = ws[b * C + c], gy = gys[b * kShChunk + lc];
          cjfma(acc, cf{w.x, w.y}, cf{gy.x, gy.y});
        }
        if (lc < ncol) st(gX, xb + (size_t)c * cols + c0 + lc, acc);
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t pr = threadIdx.x * ppt + k;
      if (k < (int)ppt && pr < pairs) {
        const uint32_t b = pr / C, c = pr % C;
        cf acc = gw[k];
        for (uint32_t j = 0; j < ncol; ++j) {
          const float2 gy = gys[b * kShChunk + j], xv = xs[c * kShChunk + j];
          cjfma(acc, cf{xv.x, xv.y}, cf{gy.x, gy.y});  // g_y * conj(x)
        }
        gw[k] = acc;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint32_t pr = threadIdx.x * ppt + k;
    if (k < (int)ppt && pr < pairs) st(work, (size_t)blockIdx.x * pairs + pr, gw[k]);
  }
}

__global__ __launch_bounds__(256) void sh_sum_partials_kernel(const float* __restrict__ work, uint32_t n_part,
                                                              uint32_t pairs, float* __restrict__ gW) {
  const uint32_t pr = blockIdx.x * blockDim.x + threadIdx.x;
  if (pr >= pairs) return;
  cf acc{0.f, 0.f};
  for (uint32_t i = 0; i < n_part; ++i) {
    const cf v = ld(work, (size_t)i * pairs + pr);
    acc.re += v.re;
    acc.im += v.im;
  }
  st(gW, pr, acc);
}

constexpr uint32_t kShBwdBlocks = 1024;

}  // namespace

extern "C" {

rd_status rd_sh_pfb(void* stream, const float* x, const float* h, uint32_t n_rows, uint32_t n_fft, uint32_t taps,
                    uint32_t n_frames, float* u) {
  if (!x || !h || !u || n_fft < 1 || taps < 1 || n_frames < 1) return RD_E_ARG;
  const uint64_t total = (uint64_t)n_rows * n_frames * n_fft;
  if (total == 0) return RD_OK;
  if (total > (1ull << 31)) return RD_E_ARG;  // one work-item per output
  sh_pfb_kernel<<<(uint32_t)((total + 255) / 256), 256, 0, (hipStream_t)stream>>>(x, h, n_fft, taps, n_frames, total, u);
  return fin(hipGetLastError());
}

rd_status rd_sh_pfb_bwd(void* stream, const float* g_u, const float* h, uint32_t n_rows, uint32_t n_fft,
                        uint32_t taps, uint32_t n_frames, float* g_x) {
  if (!g_u || !h || !g_x || n_fft < 1 || taps < 1 || n_frames < 1) return RD_E_ARG;
  const uint64_t total = (uint64_t)n_rows * n_frames * n_fft;
  if (total == 0) return RD_OK;
  if (total > (1ull << 31)) return RD_E_ARG;  // one work-item per output
  sh_pfb_bwd_kernel<<<(uint32_t)((total + 255) / 256), 256, 0, (hipStream_t)stream>>>(g_u, h, n_fft, taps, n_frames,
                                                                                       total, g_x);
  return fin(hipGetLastError());
}

rd_status rd_sh_fft(void* stream, const float* in, const float* tw, uint32_t n_fft, uint32_t n_frames, int inverse,
                    float scale, float* out) {
  if (!in || !tw || !out) return RD_E_ARG;
  if (n_fft < 2 || n_fft > 4096 || (n_fft & (n_fft - 1))) return RD_E_GEOM;
  if (n_frames == 0) return RD_OK;
  uint32_t log2n = 0;
  while ((1u << log2n) < n_fft) ++log2n;
  const uint32_t fpb = n_fft >= 512 ? 1 : 512 / n_fft;
  const size_t lds = ((size_t)n_fft / 2 + (size_t)fpb * n_fft) * sizeof(float2);
  sh_fft_kernel<<<(n_frames + fpb - 1) / fpb, 256, lds, (hipStream_t)stream>>>(in, tw, log2n, n_frames, inverse, scale,
                                                                             out);
  return fin(hipGetLastError());
}

rd_status rd_sh_bf_power(void* stream, const float* X, const float* W, uint32_t n_chan, uint32_t n_beams,
                         uint64_t n_cols, uint32_t n_cpi, float* y, float* p) {
  if (!X || !W || !p || n_chan < 1 || n_chan > 64 || n_beams < 1 || n_beams > 64) return RD_E_ARG;
  if (n_cols > (1ull << 31)) return RD_E_ARG;
  const uint64_t total = n_cols * n_cpi;
  if (total == 0) return RD_OK;
  if (total > (1ull << 31)) return RD_E_ARG;  // one work-item per column
  const dim3 grid((uint32_t)((total + 255) / 256));
  const size_t lds = (size_t)n_beams * n_chan * sizeof(float2);
  hipStream_t s = (hipStream_t)stream;
  if (n_chan <= 16)
    sh_bf_kernel<16><<<grid, 256, lds, s>>>(X, W, n_chan, n_beams, n_cols, n_cpi, y, p);
  else if (n_chan <= 32)
    sh_bf_kernel<32><<<grid, 256, lds, s>>>(X, W, n_chan, n_beams, n_cols, n_cpi, y, p);
  else
    sh_bf_kernel<64><<<grid, 256, lds, s>>>(X, W, n_chan, n_beams, n_cols, n_cpi, y, p);
  return fin(hipGetLastError());
}

size_t rd_sh_bf_bwd_work_bytes(uint32_t n_chan, uint32_t n_beams) {
  return (size_t)kShBwdBlocks * n_chan * n_beams * sizeof(float2);
}

rd_status rd_sh_bf_power_bwd(void* stream, const float* X, const float* W, const float* g_p, uint32_t n_chan,
                             uint32_t n_beams, uint64_t n_cols, uint32_t n_cpi, float* g_X, float* g_W, float* work,
                             size_t work_bytes) {
  if (!X || !W || !g_p || !g_W || !work || n_chan < 1 || n_chan > 64 || n_beams < 1 || n_beams > 64) return RD_E_ARG;
  if (work_bytes < rd_sh_bf_bwd_work_bytes(n_chan, n_beams)) return RD_E_CAPACITY;
  const uint64_t chunks = (n_cols + kShChunk - 1) / kShChunk * n_cpi;
  const uint32_t blocks = (uint32_t)(chunks < kShBwdBlocks ? (chunks ? chunks : 1) : kShBwdBlocks);
  const size_t lds = ((size_t)n_beams * n_chan + (size_t)(n_chan + n_beams) * kShChunk) * sizeof(float2);
  hipStream_t s = (hipStream_t)stream;
  sh_bf_bwd_kernel<<<blocks, 256, lds, s>>>(X, W, g_p, n_chan, n_beams, n_cols, n_cpi, g_X, work);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return RD_E_HIP;
  const uint32_t pairs = n_chan * n_beams;
  sh_sum_partials_kernel<<<(pairs + 255) / 256, 256, 0, s>>>(work, blocks, pairs, g_W);
  return fin(hipGetLastError());
}

}  // extern "C"
