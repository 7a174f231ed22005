// rd_fxpgate.hip — the fixed-point primitives of model/cpp/fxp (fxp.hpp,
// cmult.hpp) evaluated element-wise on the device by the same device
// functions the stage kernels inline (rd_fxp.h): rd_fxp_eval, rd_fxp_accum,
// rd_fxp_cmult. Besides the value and flags of each operation, rd_fxp_eval
// reports, per element, whether a specialised form the kernels use
// (sat16, rne15_i32, rne15_i64 + sat16_i64, rne_sh01) applies to that
// element's operands and whether it disagrees with the general form.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rd_fxp.h"
#include "rd_internal.h"

namespace rd {

__device__ __forceinline__ int64_t growth_bits(uint32_t n) {
  if (n <= 1) return 0;
  return 32 - __clz(n - 1);
}

__global__ __launch_bounds__(256) void fxp_eval_kernel(const rd_fxp_op* __restrict__ ops, uint64_t n,
                                                       int64_t* __restrict__ y, uint32_t* __restrict__ flags,
                                                       uint32_t* __restrict__ forms) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const rd_fxp_op o = ops[i];
  const int64_t a = o.a, b = o.b;
  const uint32_t s = o.sh, w = o.w;
  int64_t v = 0;
  uint32_t fl = 0, used = 0, bad = 0;
  // a specialised form applies (bit 0) and disagrees (bit 1)
  auto form = [&](bool applies, int64_t got, uint32_t got_fl, int64_t want, uint32_t want_fl) {
    if (!applies) return;
    used = 1;
    if (got != want || got_fl != want_fl) bad = 1;
  };
  const bool fits32 = a >= INT32_MIN && a <= INT32_MAX;
  switch (o.op) {
    case RD_FXP_SAT: {
      v = fx_sat(a, w);
      fl = fx_flags(a, w);
      uint32_t f16 = 0;
      const int32_t r16 = fits32 ? sat16((int32_t)a, f16) : 0;
      form(w == 16 && fits32, r16, f16, v, fl);
      break;
    }
    case RD_FXP_TRUNC: v = fx_floor_shift(a, s); break;
    case RD_FXP_ROUND_EVEN: {
      v = fx_rne(a, s);
      form(s == 15 && fits32, fits32 ? rne15_i32((int32_t)a) : 0, 0, v, 0);
      form(s == 15 && a > -(1ll << 47) && a < (1ll << 47), rne15_i64(a), 0, v, 0);
      form(s <= 1 && a >= -65536 && a <= 65536, rne_sh01((int32_t)a, s), 0, v, 0);
      break;
    }
    case RD_FXP_ROUND_HALF_UP: v = fx_rhu(a, s); break;
    case RD_FXP_ROUND_SAT: {
      v = fx_round_sat(a, s, w, fl);
      uint32_t f1 = 0, f2 = 0;
      const bool ok32 = s == 15 && w == 16 && fits32;
      const bool ok64 = s == 15 && w == 16 && a > -(1ll << 47) && a < (1ll << 47);
      form(ok32, ok32 ? sat16(rne15_i32((int32_t)a), f1) : 0, f1, v, fl);
      form(ok64, ok64 ? sat16_i64(rne15_i64(a), f2) : 0, f2, v, fl);
      break;
    }
    case RD_FXP_ADD_SAT: v = fx_sat(fx_add(a, b), w); fl = fx_flags(fx_add(a, b), w); break;
    case RD_FXP_SUB_SAT: v = fx_sat(fx_sub(a, b), w); fl = fx_flags(fx_sub(a, b), w); break;
    case RD_FXP_NEG_SAT: v = fx_sat(fx_neg(a), w); fl = fx_flags(fx_neg(a), w); break;
    case RD_FXP_MUL_Q15: v = (int64_t)cre((uint32_t)a) * cre((uint32_t)b); break;
    case RD_FXP_MUL_Q15_RS: {
      const int32_t p = cre((uint32_t)a) * cre((uint32_t)b);  // |p| <= 2^30: exact in int32
      v = fx_round_sat(p, 15, 16, fl);
      uint32_t f1 = 0;
      form(true, sat16(rne15_i32(p), f1), f1, v, fl);
      break;
    }
    case RD_FXP_CMUL_RE:
    case RD_FXP_CMUL_IM: {
      const uint32_t x = (uint32_t)a, h = (uint32_t)b;
      const int64_t p = o.op == RD_FXP_CMUL_RE ? (int64_t)cre(x) * cre(h) - (int64_t)cim(x) * cim(h)
                                               : (int64_t)cre(x) * cim(h) + (int64_t)cim(x) * cre(h);
      v = fx_round_sat(p, 15, 16, fl);
      uint32_t f1 = 0;
      form(true, sat16_i64(rne15_i64(p), f1), f1, v, fl);
      break;
    }
    case RD_FXP_ACC_W: v = (int64_t)w + growth_bits((uint32_t)a); break;
    case RD_FXP_GROWTH_BITS: v = growth_bits((uint32_t)a); break;
    default: v = 0; fl = 0xFFFFFFFFu; break;
  }
  y[i] = v;
  flags[i] = fl;
  if (forms) forms[i] = used | (bad << 1);
}

// The saturating accumulator from a clear (fxp.hpp:392-418 with the sticky
// flags of :358-381): per step the flags of acc + x at acc_w, the clamped
// sum, the merged sticky flags and the saturating count of flagged steps.
__global__ void fxp_accum_kernel(uint32_t acc_w, const int64_t* __restrict__ x, uint32_t n, int64_t* __restrict__ y,
                                 uint32_t* __restrict__ step_flags, uint32_t* __restrict__ sticky,
                                 uint32_t* __restrict__ count) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t acc = 0;
  uint32_t st = 0, cnt = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const int64_t raw = fx_add(acc, x[i]);
    const uint32_t f = fx_flags(raw, acc_w);
    acc = fx_sat(raw, acc_w);
    st |= f;
    if (f && cnt != 0xFFFFFFFFu) ++cnt;
    y[i] = acc;
    step_flags[i] = f;
    sticky[i] = st;
    count[i] = cnt;
  }
}

// Both complex-multiply forms of cmult.hpp:85-168 (four multiplies and the
// three-multiply Karatsuba form, equal in Z) and the rounded Q1.15 product.
__global__ __launch_bounds__(256) void fxp_cmult_kernel(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                                        uint64_t n, int64_t* __restrict__ p4, int64_t* __restrict__ p3,
                                                        uint32_t* __restrict__ y, uint8_t* __restrict__ flags) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t ar = cre(a[i]), ai = cim(a[i]), br = cre(b[i]), bi = cim(b[i]);
  const int64_t re = ar * br - ai * bi, im = ar * bi + ai * br;
  const int64_t k1 = ar * (br + bi), k2 = bi * (ar + ai), k3 = br * (ai - ar);
  p4[2 * i] = re;
  p4[2 * i + 1] = im;
  p3[2 * i] = k1 - k2;
  p3[2 * i + 1] = k1 + k3;
  uint32_t fr = 0, fi = 0;
  const int64_t yr = fx_round_sat(re, 15, 16, fr), yi = fx_round_sat(im, 15, 16, fi);
  y[i] = cpack((int32_t)yr, (int32_t)yi);
  flags[i] = (uint8_t)(fr | (fi << 2));
}

hipError_t launch_fxp_eval(hipStream_t s, const rd_fxp_op* ops, uint64_t n, int64_t* y, uint32_t* flags,
                           uint32_t* forms) {
  if (n == 0) return hipSuccess;
  fxp_eval_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(ops, n, y, flags, forms);
  return hipGetLastError();
}

hipError_t launch_fxp_accum(hipStream_t s, uint32_t acc_w, const int64_t* x, uint32_t n, int64_t* y,
                            uint32_t* step_flags, uint32_t* sticky, uint32_t* count) {
  if (n == 0) return hipSuccess;
  fxp_accum_kernel<<<1, 64, 0, s>>>(acc_w, x, n, y, step_flags, sticky, count);
  return hipGetLastError();
}

hipError_t launch_fxp_cmult(hipStream_t s, const uint32_t* a, const uint32_t* b, uint64_t n, int64_t* p4,
                            int64_t* p3, uint32_t* y, uint8_t* flags) {
  if (n == 0) return hipSuccess;
  fxp_cmult_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(a, b, n, p4, p3, y, flags);
  return hipGetLastError();
}

}  // namespace rd
