// Single-precision path math on the gfx950 transcendental unit.
//
// The FP32 stepping kernels draw the same integers as the FP64 kernels (same
// PCG outputs, same Sobol point bits) and map them to normals in FP32:
//   * Box-Muller with v_log_f32 / v_sqrt_f32 and v_sin_f32 / v_cos_f32,
//     whose inputs are in turns, so the reference's turn-based angle
//     (monte_carlo.rs:1852-1894) maps onto them directly;
//   * an inverse normal of the form x * P(w), w = -ln(4 p (1 - p)), after
//     M. Giles (GPU Computing Gems Jade Edition, 2011), with coefficients
//     fitted by kiyosi_amd/csrc/fit_inverse_normal.py (degree 7 central,
//     degree 8 tail; max abs error 1.2e-6 over all 32-bit inputs, including
//     both tails).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace kg {
namespace f32 {

constexpr float LN2 = 0.693147180559945309f;
constexpr float LOG2E = 1.44269504088896341f;

__device__ inline float log(float x) { return __builtin_amdgcn_logf(x) * LN2; }
__device__ inline float exp(float x) { return __builtin_amdgcn_exp2f(x * LOG2E); }
__device__ inline float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// exp(x) - 1 without cancellation for |x| < 1/2 (Taylor to x^9).
__device__ inline float expm1(float x) {
  if (fabsf(x) < 0.5f) {
    float p = 1.0f / 362880.0f;
    p = fmaf(p, x, 1.0f / 40320.0f);
    p = fmaf(p, x, 1.0f / 5040.0f);
    p = fmaf(p, x, 1.0f / 720.0f);
    p = fmaf(p, x, 1.0f / 120.0f);
    p = fmaf(p, x, 1.0f / 24.0f);
    p = fmaf(p, x, 1.0f / 6.0f);
    p = fmaf(p, x, 0.5f);
    return fmaf(p * x, x, x);
  }
  return exp(x) - 1.0f;
}

// Inverse normal CDF of the midpoint of a 32-bit uniform:
// p = (bits + 1/2) / 2^32. Read as a signed integer s, the bits give
// q = -(s + 1/2) / 2^32 with |q| = min(p, 1 - p) and the sign of the normal
// (s = -(k + 1) for p > 1/2, with k = ~bits the distance to the top end), so
// one conversion and one fma carry both. With l = log2(4 |q| (1 - |q|)) <= 0,
// the central polynomial runs in l itself (its constant term is sqrt(pi/2),
// the slope of the inverse normal at p = 1/2) and the tail one
// (w = -ln2 l >= 6) in t = sqrt(ln2) sqrt(-l) - 3. The tail is a
// wave-uniform branch taken only when some lane of the wave needs it (one
// draw in 800, so about one wave step in 13), and the common path carries no
// exec-mask bookkeeping.
// `sign` (+1 or -1) multiplies the result through q, which costs nothing
// once it is a constant.
__device__ inline float normal_from_bits(uint32_t bits, float sign = 1.0f) {
  constexpr float CENTRAL[8] = {-3.670217907e-08f, -1.419892442e-06f, -1.972815153e-05f,
                                -6.814792869e-05f, 1.058840542e-03f,  7.816414349e-03f,
                                -2.274419963e-01f, 1.253313780e+00f};
  constexpr float TAIL[9] = {5.152326776e-05f,  -4.903596127e-04f, 2.017864492e-03f,
                             -4.844579846e-03f, 7.992940024e-03f,  -1.083856169e-02f,
                             1.336772740e-02f,  1.416583896e+00f,  4.006433964e+00f};
  constexpr float SQRT_LN2 = 0.832554611157697756f;
  constexpr float TAIL_L = -8.65617024533378f;  // log2 value where w = 6
  const float q = fmaf((float)(int32_t)bits, -0x1p-32f * sign, -0x1p-33f * sign);
  const float ps = fabsf(q);                       // min(p, 1 - p) in (0, 1/2]
  const float ax = fmaf(-2.0f, ps, 1.0f);          // |2p - 1|
  const float l = __builtin_amdgcn_logf(ps * fmaf(-4.0f, ps, 4.0f));
  float r = CENTRAL[0];
#pragma unroll
  for (int i = 1; i < 8; ++i) r = fmaf(r, l, CENTRAL[i]);
  const bool tail = l <= TAIL_L;
  if (wave_any(tail)) {
    const float t = fmaf(SQRT_LN2, sqrt(-l), -3.0f);
    float v = TAIL[0];
#pragma unroll
    for (int i = 1; i < 9; ++i) v = fmaf(v, t, TAIL[i]);
    r = tail ? v : r;
  }
  return copysignf(ax * r, q);
}

// Uniform midpoint in (0, 1) from the high PCG word of a 52-bit uniform.
__device__ inline float uniform_hi(uint32_t hi) { return fmaf((float)hi, 0x1p-32f, 0x1p-33f); }

// One Box-Muller pair from four PCG outputs, consumed exactly as the FP64
// form consumes them (two 52-bit uniforms), so both forms stay on one
// stream. Only the high word of each uniform is used, so the low words'
// outputs are skipped by two-step LCG jumps.
__device__ inline void box_muller_pair(Pcg32& g, float* even, float* odd) {
  const uint32_t h1 = g.next_skip_one();
  const uint32_t h2 = g.next_skip_one();
  // -2 ln u = -2 ln 2 * log2 u: one multiply after the hardware log2.
  const float radius = sqrt(__builtin_amdgcn_logf(uniform_hi(h1)) * -1.38629436111989061f);
  const float turns = uniform_hi(h2);
  *even = radius * __builtin_amdgcn_cosf(turns);
  *odd = radius * __builtin_amdgcn_sinf(turns);
}

}  // namespace f32
}  // namespace kg
