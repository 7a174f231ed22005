/*
 * rd_shadow.h — C ABI of the FP32 float shadow of the chain (PFB -> FFT ->
 * beamformer -> power) and its hand-written backward kernels, on gfx950.
 *
 * No counterpart in the reference: its only float evaluations are
 * bf::dot_float / step_float, a double-precision "is this a dot product"
 * cross-check (model/cpp/beamformer/beamformer_model.hpp:202-223,307-318), and
 * the numpy float FFT checks (model/python/gen_fft_vectors.py:291-310). The
 * shadow computes the same linear maps as the bit-exact stages without
 * rounding or saturation, so gradients can flow through them:
 *   PFB   pfb_model.hpp:139-168        u[r][m][p] = sum_k h[p][k] x[r][(m-k)N+p]
 *   FFT   fft_ref.hpp:250-407          DFT; with every stage scaled it is DFT/N
 *   BF    beamformer_model.hpp:162-179 y[b][col] = sum_c W[b][c] X[c][col]
 *   power covar_model.hpp:99-102       p = |y|^2
 *
 * Conventions: complex values are interleaved float pairs (re, im); device
 * pointers; every call is queued on the HIP stream given as `stream` (a
 * hipStream_t, e.g. torch.cuda.current_stream().cuda_stream). For a real loss
 * L, gradients are grad(z) = dL/dRe(z) + i dL/dIm(z); a linear map y = A z
 * back-propagates as grad(z) = A^H grad(y).
 */
#ifndef RD_SHADOW_H_
#define RD_SHADOW_H_

#include "rd_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* x [n_rows][n_frames*n_fft], h [n_fft][taps] -> u [n_rows][n_frames][n_fft].
 * n_rows * n_frames * n_fft <= 2^31 (here and in the adjoint). */
rd_status rd_sh_pfb(void* stream, const float* x, const float* h, uint32_t n_rows, uint32_t n_fft, uint32_t taps,
                    uint32_t n_frames, float* u);
/* grad u [n_rows][n_frames][n_fft] -> grad x [n_rows][n_frames*n_fft] (the adjoint of rd_sh_pfb). */
rd_status rd_sh_pfb_bwd(void* stream, const float* g_u, const float* h, uint32_t n_rows, uint32_t n_fft,
                        uint32_t taps, uint32_t n_frames, float* g_x);

/* n_frames transforms of n_fft points (power of two, 2..4096), natural order in
 * and out: out[k] = scale * sum_n in[n] tw^(nk), tw = exp(-2 pi i / N), or its
 * conjugate when `inverse`. tw holds exp(-2 pi i j / N) for j < N/2. The
 * adjoint of the forward transform with scale s is the inverse with scale s. */
rd_status rd_sh_fft(void* stream, const float* in, const float* tw, uint32_t n_fft, uint32_t n_frames, int inverse,
                    float scale, float* out);

/* n_cpi independent CPIs sharing the weights: X [n_cpi][n_chan][n_cols],
 * W [n_beams][n_chan] -> y (nullable) / p [n_cpi][n_beams][n_cols].
 * n_chan, n_beams <= 64; n_cols * n_cpi <= 2^31. */
rd_status rd_sh_bf_power(void* stream, const float* X, const float* W, uint32_t n_chan, uint32_t n_beams,
                         uint64_t n_cols, uint32_t n_cpi, float* y, float* p);
/* Backward of rd_sh_bf_power through p (y is recomputed from X and W):
 * grad p [n_cpi][n_beams][n_cols] -> grad X (nullable, as X) and grad W
 * [n_beams][n_chan] (summed over every CPI and column, in a fixed order).
 * work: rd_sh_bf_bwd_work_bytes(n_chan, n_beams) bytes of device scratch. */
size_t rd_sh_bf_bwd_work_bytes(uint32_t n_chan, uint32_t n_beams);
rd_status rd_sh_bf_power_bwd(void* stream, const float* X, const float* W, const float* g_p, uint32_t n_chan,
                             uint32_t n_beams, uint64_t n_cols, uint32_t n_cpi, float* g_X, float* g_W, float* work,
                             size_t work_bytes);

#ifdef __cplusplus
}
#endif

#endif /* RD_SHADOW_H_ */
