// MLP continuation regressor of the American engine (KG_BASIS_MLP).
//
// The network is 3-64-64-1 with tanh-form GELU on the features
//   u0 = 4 (S/K - 1),  u1 = (2 t - T) / T,  u2 = (sigma_loc(t, S) - c) / h
// (c, h: midpoint and half-range of the surface's volatilities; 0 under
// BSM). The continuation value is parametrized as the intrinsic value plus
// the network: C(x) / K = intrinsic / K + y_scale * net(x), fitted to the
// paths' discounted cash flows, so the network carries only the timing
// value and the exercise rule is net(x) < 0.
//
// Each exercise date, from the last to the first, fits the network by Adam
// on that date's in-the-money training samples, warm-started from the
// later date's weights and Adam moments (the first fitted date starts cold
// with four times the epochs). The date's network is the average of the
// last epoch's iterates (Polyak-Ruppert averaging), and its output replaces
// the polynomial continuation value in the reference's backward induction
// (engines/monte_carlo/american.rs:1380-1436). Out-of-sample pricing
// streams fresh paths against the per-date networks as
// price_american_out_of_sample does (:574-637).
//
// Precision: FP32 master weights and Adam moments; the compute copy holds
// bf16 W2 (and its transpose) and bf16-rounded W1, W3; hidden activations
// and layer-2 gradients enter MFMA as bf16 and accumulate in FP32.
//
// MFMA mapping (v_mfma_f32_16x16x32_bf16; lane l = 16 g + c). A wave works
// on tiles of 16 samples, sample c of a tile living in lanes c, 16+c, 32+c
// and 48+c. Lane (g, c) owns the 16 hidden units u(q, i) = 16 q + 4 g + i
// (q, i in 0..3) of its sample:
//   * layer 1 runs on the VALU straight into the B operand of
//     Z2^T = W2 . H1^T (k = hidden unit, n = sample; k-slot e of step kt is
//     unit u(2 kt + e / 4, e % 4));
//   * the 16x16 output tiles of Z2^T hold exactly the units u(mt, i), so
//     they are the B operand of dH1^T = W2^T . dZ2^T, whose output tiles
//     again hold u(mt, i) for the GELU derivative of layer 1;
//   * dW2 = dZ2^T . H1 contracts over samples: two waves stage their
//     tiles' transposes of dZ2 and H1 as one 32-sample slot in LDS, and
//     each wave forms the dW2 rows of its own 16 output units over all the
//     workgroup's slots.
// Gradient partials are summed over the waves of a workgroup in wave order
// and over workgroups in a fixed slice order by the update kernel, so
// training is deterministic.
#include "lsmc_device.hpp"

namespace kg {
namespace mlp {

constexpr int HID = 64;
constexpr int P_W1 = 0, P_B1 = 192, P_W2 = 256, P_B2 = 4352, P_W3 = 4416, P_B3 = 4480, P_COUNT = 4481;
constexpr int P_LOSS = P_COUNT;  // loss slot of a partial
constexpr int P_STRIDE = 4608;   // floats per partial
constexpr int W2_ROW = 68;       // padded bf16 row of W2 / W2^T in LDS
// Packed compute copy: bf16 W2 [64][68], bf16 W2^T [64][68], then floats
// W1 [192], b1 [64], b2 [64], W3 [64], b3.
constexpr int PK_W2T = HID * W2_ROW * 2;        // byte offsets
constexpr int PK_FLOATS = 2 * HID * W2_ROW * 2;
constexpr int PK_W1 = 0, PK_B1 = 192, PK_B2 = 256, PK_W3 = 320, PK_B3 = 384;  // float offsets
constexpr int PACK_BYTES = ((PK_FLOATS + 385 * 4) + 15) & ~15;
constexpr int PACK_WORDS = PACK_BYTES / 16;
constexpr int T_ROW = 40;  // padded row of the 32-sample transposes (bf16)
constexpr int GRAD_WAVES = 2048;  // waves of a gradient launch (two per SIMD) before tiles per wave grow

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even FP32 -> bf16 bits (finite inputs).
__host__ __device__ inline uint32_t bf16_bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}
__host__ __device__ inline float bf16_value(uint32_t bits) { return __builtin_bit_cast(float, bits << 16); }

// Two FP32 values rounded to nearest-even bf16 and packed (lo in the low
// half): one v_cvt_pk_bf16_f32.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ inline uint32_t bf16_pack(float lo, float hi) {
  const f32x2 v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// GELU(x) = 0.5 x (1 + tanh(a)), a = k0 (x + k1 x^3), and its derivative, in
// the logistic form: 0.5 (1 + tanh(a)) = s = 1 / (1 + e^{-2a}), so
// GELU(x) = x s and GELU'(x) = s + x s (1 - s) 2 k0 (1 + 3 k1 x^2).
__device__ inline float gelu(float x, float* derivative) {
  constexpr float k0 = 0.7978845608028654f, k1 = 0.044715f;
  constexpr float c0 = -2.0f * 1.4426950408889634f * k0;  // -2 a log2(e) = x (c0 + c0 k1 x^2)
  const float x2 = x * x;
  const float s = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(x * fmaf(c0 * k1, x2, c0)));
  const float v = x * s;
  *derivative = fmaf(v * (1.0f - s), fmaf(6.0f * k0 * k1, x2, 2.0f * k0), s);
  return v;
}

__device__ inline f32x4 mfma(u32x4 a, u32x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                 0, 0, 0);
}

struct Pack {
  const uint16_t* w2;
  const uint16_t* w2t;
  const float* f;
  __device__ explicit Pack(const uint4* lds)
      : w2((const uint16_t*)lds),
        w2t((const uint16_t*)((const unsigned char*)lds + PK_W2T)),
        f((const float*)((const unsigned char*)lds + PK_FLOATS)) {}
};

__device__ inline void stage(uint4* lds, const uint4* __restrict__ src) {
  for (int k = threadIdx.x; k < PACK_WORDS; k += blockDim.x) lds[k] = src[k];
}

// Two 8-byte pieces of a padded bf16 row: columns base..base+3 and base+16..base+19.
__device__ inline u32x4 row_pair(const uint16_t* row, int base) {
  const uint2 lo = *(const uint2*)(row + base), hi = *(const uint2*)(row + base + 16);
  return u32x4{lo.x, lo.y, hi.x, hi.y};
}

struct Fwd {
  float dg1[16];   // GELU'(z1) of u(q, i) at 4 q + i
  uint32_t h1[8];  // bf16 H1 of u(q, i): pair (q, 2 p), (q, 2 p + 1) at 2 q + p
  uint32_t h2[8];  // bf16 GELU(z2) of u(mt, i): pair (mt, 2 p), (mt, 2 p + 1) at 2 mt + p
  float dg2[16];
  float out;
  bool finite;
  __device__ float h2_at(int r) const {  // unit u(r / 4, r % 4)
    const uint32_t w = h2[r >> 1];
    return bf16_value((r & 1) ? (w >> 16) : (w & 0xffffu));
  }
};

// Forward pass of the tile sample held by column c; every lane of the wave
// must call it (the MFMAs span the wave). With `hs`, H1 of the sample is
// also staged into the 32-sample transpose at column 16 tp + c (training),
// so it need not stay in registers until the backward pass.
__device__ inline void forward(const Pack& P, float u0, float u1, float u2, Fwd& f, uint16_t* hs = nullptr,
                               int tp = 0) {
  const int lane = threadIdx.x & (WAVE - 1), g = lane >> 4, c = lane & 15;
  const float* w1 = P.f + PK_W1;
  const float* b1 = P.f + PK_B1;
  for (int q = 0; q < 4; ++q) {
    float h[4];
    for (int i = 0; i < 4; ++i) {
      const int u = 16 * q + 4 * g + i;
      const float z = fmaf(w1[3 * u + 2], u2, fmaf(w1[3 * u + 1], u1, fmaf(w1[3 * u], u0, b1[u])));
      h[i] = gelu(z, &f.dg1[4 * q + i]);
    }
    f.h1[2 * q] = bf16_pack(h[0], h[1]);
    f.h1[2 * q + 1] = bf16_pack(h[2], h[3]);
    if (hs)
      for (int i = 0; i < 4; ++i) {
        const uint32_t w = f.h1[2 * q + (i >> 1)];
        hs[(16 * q + 4 * g + i) * T_ROW + 16 * tp + c] = (uint16_t)((i & 1) ? w >> 16 : w);
      }
  }
  const float* b2 = P.f + PK_B2;
  const float* w3 = P.f + PK_W3;
  float part = 0.0f;
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* row = P.w2 + (16 * mt + c) * W2_ROW;
    for (int kt = 0; kt < 2; ++kt) {
      const u32x4 b = {f.h1[4 * kt], f.h1[4 * kt + 1], f.h1[4 * kt + 2], f.h1[4 * kt + 3]};
      acc = mfma(row_pair(row, 32 * kt + 4 * g), b, acc);
    }
    float h[4];
    for (int i = 0; i < 4; ++i) h[i] = gelu(acc[i] + b2[16 * mt + 4 * g + i], &f.dg2[4 * mt + i]);
    f.h2[2 * mt] = bf16_pack(h[0], h[1]);
    f.h2[2 * mt + 1] = bf16_pack(h[2], h[3]);
    for (int i = 0; i < 4; ++i) part = fmaf(w3[16 * mt + 4 * g + i], f.h2_at(4 * mt + i), part);
  }
  part += __shfl_xor(part, 16, WAVE);
  part += __shfl_xor(part, 32, WAVE);
  f.out = part + P.f[PK_B3];
  f.finite = isfinite(f.out);
}

// Per-lane gradient accumulators of a wave. dW2 is held only for the
// wave's own 16 output units, summed over the samples of all the waves of
// the workgroup.
struct Acc {
  f32x4 w2[4];  // dW2 [16 wave + 4 g + i][16 nt + c]
  float w1[16][3], b1[16], b2[16], w3[16];
  float b3, loss;
};

// Backward pass of one tile: accumulates the per-lane gradients and stages
// dZ2 of the tile into the wave's 32-sample transpose at column 16 tp + c
// (the forward pass staged H1).
__device__ inline void backward(const Pack& P, const Fwd& f, float gs, float u0, float u1, float u2, Acc& a,
                                uint16_t* dz, int tp) {
  const int lane = threadIdx.x & (WAVE - 1), g = lane >> 4, c = lane & 15;
  const float* w3 = P.f + PK_W3;
  uint32_t d2[8];
  for (int mt = 0; mt < 4; ++mt) {
    float d[4];
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * mt + i, u = 16 * mt + 4 * g + i;
      d[i] = gs * w3[u] * f.dg2[r];
      a.b2[r] += d[i];
      a.w3[r] = fmaf(gs, f.h2_at(r), a.w3[r]);
    }
    d2[2 * mt] = bf16_pack(d[0], d[1]);
    d2[2 * mt + 1] = bf16_pack(d[2], d[3]);
    for (int i = 0; i < 4; ++i) {
      const uint32_t w = d2[2 * mt + (i >> 1)];
      dz[(16 * mt + 4 * g + i) * T_ROW + 16 * tp + c] = (uint16_t)((i & 1) ? w >> 16 : w);
    }
  }
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* row = P.w2t + (16 * mt + c) * W2_ROW;
    for (int kt = 0; kt < 2; ++kt) {
      const u32x4 b = {d2[4 * kt], d2[4 * kt + 1], d2[4 * kt + 2], d2[4 * kt + 3]};
      acc = mfma(row_pair(row, 32 * kt + 4 * g), b, acc);
    }
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * mt + i;
      const float d1 = acc[i] * f.dg1[r];
      a.b1[r] += d1;
      a.w1[r][0] = fmaf(d1, u0, a.w1[r][0]);
      a.w1[r][1] = fmaf(d1, u1, a.w1[r][1]);
      a.w1[r][2] = fmaf(d1, u2, a.w1[r][2]);
    }
  }
}

// Lane permutation within rows of 16 lanes by a DPP control word.
template <int CTRL>
__device__ inline float dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
// Sum over the 16 columns c of a lane group, left in every lane of the row:
// quad swaps (1 0 3 2), (2 3 0 1), then the half-row and row mirrors.
__device__ inline float sum_columns(float v) {
  v += dpp<0xB1>(v);
  v += dpp<0x4E>(v);
  v += dpp<0x141>(v);
  v += dpp<0x140>(v);
  return v;
}

// Small gradients of a partial: W1, b1, b2, W3 (64 x 6), then b3 and the loss.
constexpr int SMALL = 6 * HID + 2;

constexpr int GRAD_WG_WAVES = 4;  // waves of a gradient workgroup

// Gradient partials of the batch samples[begin, begin + n): one partial of
// P_STRIDE floats per workgroup (gradient of sum((out - y)^2) * inv_n, loss
// in slot P_LOSS). Each wave runs `tpw` consecutive 16-sample tiles, one per
// round. In a round the waves 2 s and 2 s + 1 stage their tiles' dZ2 and H1
// transposes as the two halves of the 32-sample slot s, and wave w then
// forms the dW2 rows of its own 16 output units over both slots, so dW2
// needs no cross-wave reduction. The small gradients are summed over the
// waves in wave order. Every wave runs every round so the block barriers
// are uniform.
__global__ void __launch_bounds__(GRAD_WG_WAVES * WAVE, 2)
    mlp_grad(const uint4* __restrict__ packed, const float4* __restrict__ samples, int64_t begin, int n,
             int tpw, float inv_n, float* __restrict__ parts, uint32_t* __restrict__ flags) {
  __shared__ uint4 s_pack[PACK_WORDS];
  __shared__ __attribute__((aligned(16))) uint16_t s_dz[GRAD_WG_WAVES / 2][HID * T_ROW];
  __shared__ __attribute__((aligned(16))) uint16_t s_hs[GRAD_WG_WAVES / 2][HID * T_ROW];
  __shared__ float s_small[GRAD_WG_WAVES][SMALL];
  stage(s_pack, packed);
  __syncthreads();
  const Pack P(s_pack);
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1), g = lane >> 4, c = lane & 15;
  Acc a;
  for (int nt = 0; nt < 4; ++nt) a.w2[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int r = 0; r < 16; ++r) {
    a.w1[r][0] = a.w1[r][1] = a.w1[r][2] = 0.0f;
    a.b1[r] = a.b2[r] = a.w3[r] = 0.0f;
  }
  a.b3 = a.loss = 0.0f;
  bool finite = true;
  for (int j = 0; j < tpw; ++j) {
    const int64_t tile = ((int64_t)blockIdx.x * GRAD_WG_WAVES + wave) * tpw + j;
    const int64_t local = tile * 16 + c;
    const bool valid = local < n;
    const float4 x = valid ? samples[begin + local] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    Fwd f;
    forward(P, x.x, x.y, x.z, f, s_hs[wave >> 1], wave & 1);
    finite = finite && f.finite;
    const float r = valid ? f.out - x.w : 0.0f;
    const float gs = 2.0f * r * inv_n;
    a.loss = fmaf(r * inv_n, r, a.loss);
    a.b3 += gs;
    backward(P, f, gs, x.x, x.y, x.z, a, s_dz[wave >> 1], wave & 1);
    __syncthreads();
    for (int slot = 0; slot < GRAD_WG_WAVES / 2; ++slot) {
      const uint4 av = *(const uint4*)(s_dz[slot] + (16 * wave + c) * T_ROW + 8 * g);
      for (int nt = 0; nt < 4; ++nt) {
        const uint4 bv = *(const uint4*)(s_hs[slot] + (16 * nt + c) * T_ROW + 8 * g);
        a.w2[nt] = mfma(u32x4{av.x, av.y, av.z, av.w}, u32x4{bv.x, bv.y, bv.z, bv.w}, a.w2[nt]);
      }
    }
    __syncthreads();
  }
  if (!finite) raise_flag(flags, FLAG_NONFINITE);
  float* out = parts + (int64_t)blockIdx.x * P_STRIDE;
  for (int nt = 0; nt < 4; ++nt)
    for (int i = 0; i < 4; ++i) out[P_W2 + (16 * wave + 4 * g + i) * HID + 16 * nt + c] = a.w2[nt][i];
  for (int r = 0; r < 16; ++r) {
    for (int k = 0; k < 3; ++k) a.w1[r][k] = sum_columns(a.w1[r][k]);
    a.b1[r] = sum_columns(a.b1[r]);
    a.b2[r] = sum_columns(a.b2[r]);
    a.w3[r] = sum_columns(a.w3[r]);
  }
  a.b3 = sum_columns(a.b3);
  a.loss = sum_columns(a.loss);
  float* mine = s_small[wave];
  if (c == 0)
    for (int r = 0; r < 16; ++r) {
      const int u = 16 * (r >> 2) + 4 * g + (r & 3);
      for (int k = 0; k < 3; ++k) mine[3 * u + k] = a.w1[r][k];
      mine[3 * HID + u] = a.b1[r];
      mine[4 * HID + u] = a.b2[r];
      mine[5 * HID + u] = a.w3[r];
    }
  if (lane == 0) {
    mine[6 * HID] = a.b3;
    mine[6 * HID + 1] = a.loss;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < SMALL; k += blockDim.x) {
    float v = s_small[0][k];
    for (int w = 1; w < GRAD_WG_WAVES; ++w) v += s_small[w][k];
    int slot;
    if (k < 3 * HID)
      slot = P_W1 + k;
    else if (k < 4 * HID)
      slot = P_B1 + k - 3 * HID;
    else if (k < 5 * HID)
      slot = P_B2 + k - 4 * HID;
    else if (k < 6 * HID)
      slot = P_W3 + k - 5 * HID;
    else
      slot = k == 6 * HID ? P_B3 : P_LOSS;
    out[slot] = v;
  }
}

enum : int { UPDATE_ADAM = 0, UPDATE_GRADIENT = 1, UPDATE_PACK = 2, UPDATE_ADOPT = 3 };

// Sums the partials of each parameter in workgroup order, then (ADAM) takes one
// Adam step and refreshes the packed compute copy, (GRADIENT) writes the
// summed gradient, (PACK) only packs `params`, or (ADOPT) replaces the
// parameters by their running average and packs them. With avg_k > 0 an
// Adam step also folds the new weights into the running average of the
// last avg_k iterates. Slot P_LOSS adds the batch loss to *loss.
constexpr int UPDATE_PARAMS = 16;                  // parameters per update workgroup
constexpr int UPDATE_SLICES = 256 / UPDATE_PARAMS;  // slices of the partials summed in parallel

__global__ void __launch_bounds__(256)
    mlp_update(int mode, const float* __restrict__ parts, int n_parts, float* __restrict__ params,
               float* __restrict__ m1, float* __restrict__ m2, float* __restrict__ avg, int avg_k,
               uint4* __restrict__ packed, float lr, float c1, float c2, float* __restrict__ gradient,
               float* __restrict__ loss) {
  __shared__ double s_sum[UPDATE_SLICES][UPDATE_PARAMS];
  const int slice = threadIdx.x / UPDATE_PARAMS, j = threadIdx.x % UPDATE_PARAMS;
  const int p = blockIdx.x * UPDATE_PARAMS + j;
  // Slice s sums partials s, s + UPDATE_SLICES, ... in order; the slices
  // are then added in slice order, so the sum is the same on every run.
  double part = 0.0;
  if ((mode == UPDATE_ADAM || mode == UPDATE_GRADIENT) && p <= P_COUNT)
#pragma unroll 4
    for (int w = slice; w < n_parts; w += UPDATE_SLICES) part += (double)parts[(int64_t)w * P_STRIDE + p];
  s_sum[slice][j] = part;
  __syncthreads();
  if (slice != 0 || p > P_COUNT) return;
  double sum = 0.0;
  for (int k = 0; k < UPDATE_SLICES; ++k) sum += s_sum[k][j];
  if (p == P_LOSS) {
    if (loss) *loss += (float)sum;
    return;
  }
  if (mode == UPDATE_GRADIENT) {
    gradient[p] = (float)sum;
    return;
  }
  float w = params[p];
  if (mode == UPDATE_ADAM) {
    const float gr = (float)sum;
    const float a = 0.9f * m1[p] + 0.1f * gr;
    const float b = 0.999f * m2[p] + 0.001f * gr * gr;
    m1[p] = a;
    m2[p] = b;
    w = w - lr * (a * c1) / (sqrtf(b * c2) + 1e-8f);
    params[p] = w;
    if (avg_k == 1) avg[p] = w;
    if (avg_k > 1) avg[p] += (w - avg[p]) / (float)avg_k;
  } else if (mode == UPDATE_ADOPT) {
    w = avg[p];
    params[p] = w;
  }
  uint16_t* w2 = (uint16_t*)packed;
  uint16_t* w2t = (uint16_t*)((unsigned char*)packed + PK_W2T);
  float* f = (float*)((unsigned char*)packed + PK_FLOATS);
  if (p < P_B1) {
    f[PK_W1 + p] = bf16_value(bf16_bits(w));
  } else if (p < P_W2) {
    f[PK_B1 + p - P_B1] = w;
  } else if (p < P_B2) {
    const int o = (p - P_W2) / HID, i = (p - P_W2) % HID;
    const uint16_t bits = (uint16_t)bf16_bits(w);
    w2[o * W2_ROW + i] = bits;
    w2t[i * W2_ROW + o] = bits;
  } else if (p < P_W3) {
    f[PK_B2 + p - P_B2] = w;
  } else if (p < P_B3) {
    f[PK_W3 + p - P_W3] = bf16_value(bf16_bits(w));
  } else {
    f[PK_B3] = w;
  }
}

__global__ void fill_floats(float* p, int64_t n, float v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Forward pass over samples[0, n), one 16-sample tile per wave. With `out`
// the outputs are written; with `cash` the exercise decision of the date is
// applied: a slot whose continuation value falls below its intrinsic value
// (a negative network output) takes the intrinsic value as its cash.
__global__ void __launch_bounds__(256)
    mlp_eval(const uint4* __restrict__ packed, const float4* __restrict__ samples, int n, float* __restrict__ out,
             LsmcModel m, const int32_t* __restrict__ idx, const double* __restrict__ spot,
             double* __restrict__ cash, uint32_t* __restrict__ flags) {
  __shared__ uint4 s_pack[PACK_WORDS];
  stage(s_pack, packed);
  __syncthreads();
  const Pack P(s_pack);
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1), g = lane >> 4, c = lane & 15;
  const int64_t local = ((int64_t)blockIdx.x * 4 + wave) * 16 + c;
  const bool valid = local < n;
  const float4 x = valid ? samples[local] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  Fwd f;
  forward(P, x.x, x.y, x.z, f);
  if (!valid || g != 0) return;
  if (!f.finite) raise_flag(flags, FLAG_NONFINITE);
  if (out) out[local] = f.out;
  if (cash) {
    const int32_t slot = idx[local];
    if (f.out < 0.0f) cash[slot] = payoff(m, spot[slot]);
  }
}

// Feature scaling: u0 = 4 (S/K - 1), u1 = (t - t_mid) t_scale, u2 = (sigma - s_mid) s_scale.
struct Scaling {
  double t_mid, t_scale, s_mid, s_scale, y_scale;
};

__device__ inline double date_spot(const LsmcModel& m, int step, bool fp32, bool lv, const void* coords,
                                   int64_t samples, int64_t slots, int64_t n, int64_t slot, int leg) {
  if (lv) {
    const double x = fp32 ? (double)((const float*)coords)[(size_t)step * slots + slot]
                          : ((const double*)coords)[(size_t)step * slots + slot];
    return m.spot * (fp32 ? (double)f32::exp((float)x) : exp(x));
  }
  const double w = fp32 ? (double)((const float*)coords)[(size_t)step * samples + n]
                        : ((const double*)coords)[(size_t)step * samples + n];
  return fp32 ? bsm_spot32(m, step, (float)(leg ? -w : w)) : bsm_spot(m, step, leg ? -w : w);
}

// Block sizes of the compaction: each thread owns sample n and its legs, so
// a block's slots are the reference block's slots in order.
__device__ inline int block_exclusive_scan(int v, int* lds, int* total) {
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  int x = v;
  for (int o = 1; o < WAVE; o <<= 1) {
    const int y = __shfl_up(x, o, WAVE);
    if (lane >= o) x += y;
  }
  if (lane == WAVE - 1) lds[wave] = x;
  __syncthreads();
  int base = 0;
  const int waves = blockDim.x / WAVE;
  for (int w = 0; w < waves; ++w) base += w < wave ? lds[w] : 0;
  int t = 0;
  for (int w = 0; w < waves; ++w) t += lds[w];
  *total = t;
  __syncthreads();
  return base + x - v;
}

// Date `step`: discount the cash one interval, rebuild the spot, count the
// in-the-money slots of the block.
template <bool FP32, bool LV>
__global__ void __launch_bounds__(256)
    mlp_mark(LsmcModel m, int step, int64_t samples, int64_t paths, int antithetic, const void* __restrict__ coords,
             double* __restrict__ spot, double* __restrict__ cash, int32_t* __restrict__ counts,
             uint32_t* __restrict__ flags) {
  __shared__ int lds[4];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const int legs = antithetic ? 2 : 1;
  const int64_t slots = samples * legs;
  int itm = 0;
  bool bad = false;
  for (int leg = 0; leg < legs; ++leg) {
    const bool on = n < samples && (leg == 0 || n < paths / 2);
    if (!on) continue;
    const int64_t slot = n * legs + leg;
    const double s = date_spot(m, step, FP32, LV, coords, samples, slots, n, slot, leg);
    const double c = cash[slot] * m.discount;
    bad = bad || !isfinite(c) || !isfinite(s);
    spot[slot] = s;
    cash[slot] = c;
    itm += payoff(m, s) > 0.0;
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
  int total;
  block_exclusive_scan(itm, lds, &total);
  if (j == 0) counts[block] = total;
}

// Exclusive scan of the block counts into offsets[0..blocks]; one workgroup.
__global__ void __launch_bounds__(1024) mlp_offsets(const int32_t* __restrict__ counts, int64_t blocks,
                                                    int32_t* __restrict__ offsets) {
  __shared__ int lds[16];
  int carry = 0;
  for (int64_t b0 = 0; b0 < blocks; b0 += blockDim.x) {
    const int64_t b = b0 + threadIdx.x;
    const int v = b < blocks ? counts[b] : 0;
    int total;
    const int ex = block_exclusive_scan(v, lds, &total);
    if (b < blocks) offsets[b] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) offsets[blocks] = carry;
}

// Writes the features and targets of the in-the-money slots of date
// `step` in slot order, with their slot indices.
template <bool LV>
__global__ void __launch_bounds__(256)
    mlp_scatter(LsmcModel m, int step, Scaling sc, int64_t samples, int64_t paths, int antithetic,
                const double* __restrict__ spot, const double* __restrict__ cash,
                const int32_t* __restrict__ offsets, float4* __restrict__ out, int32_t* __restrict__ idx) {
  __shared__ int lds[4];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const int legs = antithetic ? 2 : 1;
  bool itm[2] = {false, false};
  for (int leg = 0; leg < legs; ++leg) {
    const bool on = n < samples && (leg == 0 || n < paths / 2);
    itm[leg] = on && payoff(m, spot[n * legs + leg]) > 0.0;
  }
  int total;
  int at = offsets[block] + block_exclusive_scan((int)itm[0] + (int)itm[1], lds, &total);
  const double t = m.dt * (step + 1);
  const float u1 = (float)((t - sc.t_mid) * sc.t_scale);
  for (int leg = 0; leg < legs; ++leg) {
    if (!itm[leg]) continue;
    const int64_t slot = n * legs + leg;
    const double s = spot[slot];
    const double sigma = LV ? lv_sigma(m, m.year0 + t, s) : m.sigma;
    const double timing = (cash[slot] - payoff(m, s)) / m.strike;
    out[at] = make_float4((float)(4.0 * (s / m.strike - 1.0)), u1, (float)((sigma - sc.s_mid) * sc.s_scale),
                          (float)(timing / sc.y_scale));
    idx[at] = (int32_t)slot;
    ++at;
  }
}

// Out-of-sample valuation with the per-date networks: every lane streams
// one sample (both legs of an antithetic pair); at each date the wave
// evaluates the network for its 64 paths as four 16-sample tiles when any
// of its live paths is in the money.
template <bool FP32, bool LV>
__global__ void __launch_bounds__(256)
    mlp_price(LsmcModel m, const uint4* __restrict__ nets, const int32_t* __restrict__ fitted, Scaling sc,
              uint64_t seed, int64_t samples, int64_t paths, int antithetic, Moments* __restrict__ parts,
              uint32_t* __restrict__ flags) {
  __shared__ uint4 s_pack[PACK_WORDS];
  __shared__ Moments smem[2 * 4];
  __shared__ double lv_lds[LV ? 3 * LV_LDS_NODES : 1];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int lane = j & (WAVE - 1), g = lane >> 4, c = lane & 15;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const bool live = n < samples;
  const bool paired = antithetic && live && n < paths / 2;
  Normals<FP32> z;
  z.start(seed, (uint64_t)block, (uint64_t)j * (uint64_t)m.steps);
  double current[2] = {m.spot, m.spot}, x[2] = {0.0, 0.0}, value[2] = {0.0, 0.0};
  int bracket[2] = {-1, -1};
  bool alive[2] = {live, paired};
  bool bad = false;
  LvRows rows{};
  rows.row = -1;
  for (int step = 0; step < m.steps; ++step) {
    const bool fit = step < m.steps - 1 && fitted[step] != 0;
    __syncthreads();
    if (fit) stage(s_pack, nets + (size_t)step * PACK_WORDS);
    __syncthreads();
    if (LV) lv_stage<FP32>(m, step, lv_lds, rows);  // block-uniform
    const double e = live ? z.next() : 0.0;
    for (int leg = 0; leg < 2; ++leg) {
      if (!alive[leg]) continue;
      const double d = leg ? -e : e;
      if (LV) {
        x[leg] = lv_step<FP32>(m, rows, x[leg], d, bracket[leg]);
        current[leg] = m.spot * (FP32 ? (double)f32::exp((float)x[leg]) : exp(x[leg]));
      } else if (FP32) {
        current[leg] = (double)((float)current[leg] * f32::exp(fmaf((float)m.diffusion, (float)d, (float)m.drift)));
      } else {
        current[leg] *= exp(fma(m.diffusion, d, m.drift));
      }
      bad = bad || !isfinite(current[leg]) || current[leg] <= 0.0;
    }
    if (step == m.steps - 1) {
      for (int leg = 0; leg < 2; ++leg)
        if (alive[leg]) value[leg] = discounted_product(1.0, payoff(m, current[leg]), -m.rate * m.time);
      break;
    }
    if (!fit) continue;
    const Pack P(s_pack);
    const double t = m.dt * (step + 1);
    if (LV) lv_stage<FP32>(m, step + 1, lv_lds, rows);  // the date's rows; also the next step's start
    for (int leg = 0; leg < 2; ++leg) {
      const double intrinsic = payoff(m, current[leg]);
      const bool ask = alive[leg] && intrinsic > 0.0;
      if (__ballot(ask) == 0) continue;
      const double sigma = !LV ? m.sigma
                           : rows.f32 ? (double)lv_sigma_at32(m, rows, (float)current[leg], (float)x[leg], bracket[leg])
                                      : lv_sigma_at(m, rows, current[leg], x[leg], bracket[leg]);
      const float u0 = (float)(4.0 * (current[leg] / m.strike - 1.0));
      const float u1 = (float)((t - sc.t_mid) * sc.t_scale);
      const float u2 = (float)((sigma - sc.s_mid) * sc.s_scale);
      float cont = 0.0f;
      for (int nt = 0; nt < 4; ++nt) {
        const int src = 16 * nt + c;
        Fwd f;
        forward(P, __shfl(u0, src, WAVE), __shfl(u1, src, WAVE), __shfl(u2, src, WAVE), f);
        if (nt == g) cont = f.out;
      }
      bad = bad || (ask && !isfinite(cont));
      if (ask && cont < 0.0f) {
        value[leg] = discounted_product(1.0, intrinsic, -m.rate * t);
        alive[leg] = false;
      }
    }
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
  Moments mm[2] = {moments_zero(), moments_zero()};
  if (live) {
    moments_add(mm[0], value[0], true);
    if (paired) {
      moments_add(mm[0], value[1], true);
      moments_add(mm[1], midpoint(value[0], value[1]), true);
    }
  }
  block_merge<2>(mm, smem);
  if (threadIdx.x == 0) {
    parts[block * 2] = mm[0];
    parts[block * 2 + 1] = mm[1];
  }
}

// ------------------------------------------------------------------ host
// Glorot-uniform weights and zero biases from a splitmix64 stream of `seed`.
static std::vector<float> initial_params(uint64_t seed) {
  std::vector<float> p(P_COUNT, 0.0f);
  uint64_t s = seed ^ 0x6a09e667f3bcc909ULL;
  auto uniform = [&]() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1p-53 * 2.0 - 1.0;
  };
  const double a1 = std::sqrt(6.0 / (3 + HID)), a2 = std::sqrt(6.0 / (HID + HID)), a3 = std::sqrt(6.0 / (HID + 1));
  for (int k = P_W1; k < P_B1; ++k) p[k] = (float)(a1 * uniform());
  for (int k = P_W2; k < P_B2; ++k) p[k] = (float)(a2 * uniform());
  for (int k = P_W3; k < P_B3; ++k) p[k] = (float)(a3 * uniform());
  return p;
}

// Device state of one network: master weights, moments, packed copy and
// the partials of the largest batch.
struct Net {
  float* params;
  float* m1;
  float* m2;
  float* avg;
  uint4* packed;
  float* parts;
  float* scalars;  // [0] loss accumulator
  int max_parts;
};

// 16-sample tiles per wave of a gradient launch over n samples: one until
// the batch gives every SIMD a wave, then as many as keep it so.
static int tiles_per_wave(int n) {
  const int tiles = (n + 15) / 16;
  return std::max(1, (tiles + GRAD_WAVES - 1) / GRAD_WAVES);
}
// Partials (workgroups of GRAD_WG_WAVES waves) of a gradient launch over n samples.
static int parts_for(int n) {
  const int tiles = (n + 15) / 16, tpw = tiles_per_wave(n);
  const int waves = (tiles + tpw - 1) / tpw;
  return (waves + GRAD_WG_WAVES - 1) / GRAD_WG_WAVES;
}

static Net net_buffers(kg_context* ctx, DeviceBuffer& buf, int max_batch, size_t extra_bytes, void** extra) {
  Net net{};
  net.max_parts = parts_for(max_batch);
  const size_t p_bytes = P_STRIDE * sizeof(float);
  const size_t head = 4 * p_bytes + PACK_BYTES + 256;
  const size_t total = head + (size_t)net.max_parts * p_bytes + extra_bytes;
  unsigned char* base = (unsigned char*)buf.reserve(total);
  net.params = (float*)base;
  net.m1 = (float*)(base + p_bytes);
  net.m2 = (float*)(base + 2 * p_bytes);
  net.avg = (float*)(base + 3 * p_bytes);
  net.packed = (uint4*)(base + 4 * p_bytes);
  net.scalars = (float*)(base + 4 * p_bytes + PACK_BYTES);
  net.parts = (float*)(base + head);
  if (extra) *extra = base + head + (size_t)net.max_parts * p_bytes;
  (void)ctx;
  return net;
}

static void launch_update(kg_context* ctx, const Net& net, int mode, int n_parts, float lr, int step,
                          float* gradient, float* loss, int avg_k = 0) {
  const float c1 = step > 0 ? (float)(1.0 / (1.0 - std::pow(0.9, step))) : 1.0f;
  const float c2 = step > 0 ? (float)(1.0 / (1.0 - std::pow(0.999, step))) : 1.0f;
  hipLaunchKernelGGL(mlp_update, dim3((P_COUNT + UPDATE_PARAMS) / UPDATE_PARAMS), dim3(256), 0, ctx->stream, mode,
                     net.parts,
                     n_parts, net.params, net.m1, net.m2, net.avg, avg_k, net.packed, lr, c1, c2, gradient, loss);
  KG_HIP(hipGetLastError());
}

static void clear(kg_context* ctx, float* p, int64_t n) {
  hipLaunchKernelGGL(fill_floats, dim3(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1024))), dim3(256),
                     0, ctx->stream, p, n, 0.0f);
  KG_HIP(hipGetLastError());
}

// Gradient partials of samples[begin, begin + n) into net.parts.
static int launch_grad(kg_context* ctx, const Net& net, const float4* samples, int64_t begin, int n,
                       uint32_t* flags) {
  const int parts = parts_for(n);
  if (parts > net.max_parts) fail(KG_INVALID_INPUT, "batch larger than the reserved partials");
  hipLaunchKernelGGL(mlp_grad, dim3(parts), dim3(GRAD_WG_WAVES * WAVE), 0, ctx->stream, net.packed, samples, begin, n,
                     tiles_per_wave(n), 1.0f / (float)n, net.parts, flags);
  KG_HIP(hipGetLastError());
  return parts;
}

// Adam over `n` samples: `epochs` passes of consecutive batches. `step`
// counts Adam steps since the moments were last cleared; with `average`
// the weights become the mean of the last epoch's iterates.
static int train(kg_context* ctx, const Net& net, const float4* samples, int n, int batch, int epochs, float lr,
                 bool reset_moments, bool average, int& step, uint32_t* flags, float* loss_last_epoch,
                 double* gradient_samples) {
  if (reset_moments) {
    clear(ctx, net.m1, 2 * (int64_t)P_STRIDE);  // m1 and m2 are contiguous
    step = 0;
  }
  int last = 0;
  for (int e = 0; e < epochs; ++e) {
    const bool final_epoch = e == epochs - 1;
    for (int b = 0; b < n; b += batch) {
      const int nb = std::min(batch, n - b);
      const int parts = launch_grad(ctx, net, samples, b, nb, flags);
      *gradient_samples += nb;
      ++step;
      if (final_epoch) ++last;
      launch_update(ctx, net, UPDATE_ADAM, parts, lr, step, nullptr, final_epoch ? loss_last_epoch : nullptr,
                    average && final_epoch ? last : 0);
    }
  }
  if (average) launch_update(ctx, net, UPDATE_ADOPT, 0, 0.0f, 0, nullptr, nullptr);
  return last;
}

template <bool FP32, bool LV>
static void run(kg_context* ctx, LsmcHost& h, const kg_lsmc_options& opt, kg_lsmc_report* out) {
  const LsmcModel& m = h.model;
  const Sampler& tr = h.oos ? h.training : h.pricing;
  const int epochs = opt.epochs > 0 ? opt.epochs : 4;
  // Large batches keep every SIMD busy on each Adam step and need a quarter
  // of the sequential steps of 16384; the doubled rate offsets the fewer
  // steps (on the 2^20-path local-volatility put the out-of-sample price
  // moves within its paired error).
  const int batch = opt.batch > 0 ? opt.batch : 65536;
  const float lr = opt.learning_rate > 0.0 ? (float)opt.learning_rate : 2e-3f;
  if (epochs > 1000 || batch < 32 || batch > (1 << 20) || !(lr < 1.0f))
    fail(KG_INVALID_INPUT, "MLP options: epochs <= 1000, batch in 32..=1048576, learning rate below 1");
  const int S = m.steps;
  Scaling sc{};
  sc.t_mid = 0.5 * m.time;
  sc.t_scale = 2.0 / m.time;
  double s_lo = m.sigma, s_hi = m.sigma;
  if (LV) {
    std::vector<double> vols((size_t)m.nt * m.ns);
    KG_HIP(hipMemcpy(vols.data(), m.vols, vols.size() * sizeof(double), hipMemcpyDeviceToHost));
    s_lo = *std::min_element(vols.begin(), vols.end());
    s_hi = *std::max_element(vols.begin(), vols.end());
  }
  sc.s_mid = 0.5 * (s_lo + s_hi);
  sc.s_scale = s_hi > s_lo ? 2.0 / (s_hi - s_lo) : 0.0;
  // Timing values are a fraction of an at-the-money option: about 0.1 sigma sqrt(T) of the strike.
  sc.y_scale = std::max(0.1 * sc.s_mid * std::sqrt(m.time), 1e-3);

  const int64_t samples = tr.samples(), blocks = tr.blocks();
  const int legs = tr.antithetic ? 2 : 1;
  const int64_t slots = samples * legs;
  if (slots > (int64_t)INT32_MAX) fail(KG_INVALID_INPUT, "too many training slots for the MLP regressor");
  const size_t word = FP32 ? 4 : 8;
  const size_t coord_bytes = ((size_t)S * (size_t)(LV ? slots : samples) * word + 255) & ~size_t(255);
  const size_t slot_bytes = ((size_t)slots * 8 + 255) & ~size_t(255);
  const size_t sample_bytes = ((size_t)slots * 16 + 255) & ~size_t(255);
  const size_t idx_bytes = ((size_t)slots * 4 + 255) & ~size_t(255);
  const size_t count_bytes = (((size_t)blocks + 1) * 4 + 255) & ~size_t(255);
  unsigned char* base = (unsigned char*)ctx->scratch.reserve(coord_bytes + 2 * slot_bytes + sample_bytes +
                                                             idx_bytes + 2 * count_bytes);
  void* coords = base;
  double* spot = (double*)(base + coord_bytes);
  double* cash = (double*)(base + coord_bytes + slot_bytes);
  float4* data = (float4*)(base + coord_bytes + 2 * slot_bytes);
  int32_t* idx = (int32_t*)(base + coord_bytes + 2 * slot_bytes + sample_bytes);
  int32_t* counts = (int32_t*)(base + coord_bytes + 2 * slot_bytes + sample_bytes + idx_bytes);
  int32_t* offsets = (int32_t*)(base + coord_bytes + 2 * slot_bytes + sample_bytes + idx_bytes + count_bytes);

  const size_t flag_bytes = ((size_t)S * 4 + 255) / 256 * 256;
  const size_t nets_bytes = (size_t)S * PACK_BYTES + 256 + 2 * flag_bytes;
  void* extra = nullptr;
  const Net net = net_buffers(ctx, ctx->mlp_state, batch, nets_bytes, &extra);
  uint4* nets = (uint4*)extra;
  int32_t* d_fitted = (int32_t*)((unsigned char*)extra + (size_t)S * PACK_BYTES + 256);
  // Per-date sum of the last epoch's batch losses, read once after training.
  float* d_losses = (float*)((unsigned char*)d_fitted + flag_bytes);

  const std::vector<float> init = initial_params(tr.seed);
  KG_HIP(hipMemcpyAsync(net.params, init.data(), P_COUNT * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  launch_update(ctx, net, UPDATE_PACK, 0, 0.0f, 0, nullptr, nullptr);
  uint32_t* flags = reset_flags(ctx);

  timing_begin(ctx);
  hipLaunchKernelGGL((lsmc_generate<FP32, LV ? PATHS_LOCAL_VOL : PATHS_BROWNIAN>), dim3(blocks), dim3(256), 0,
                     ctx->stream, m, tr.seed, samples, tr.path_count, (int)tr.antithetic, SobolRef{nullptr, nullptr, 0},
                     coords, spot, cash, flags);
  KG_HIP(hipGetLastError());
  clear(ctx, d_losses, S);
  std::vector<int32_t> fitted(S, 0), steps_last(S, 0);
  bool first = true;
  int adam_step = 0;
  for (int step = S - 2; step >= 0; --step) {
    check_cancel(ctx);  // the loop synchronises once per date
    hipLaunchKernelGGL((mlp_mark<FP32, LV>), dim3(blocks), dim3(256), 0, ctx->stream, m, step, samples,
                       tr.path_count, (int)tr.antithetic, coords, spot, cash, counts, flags);
    KG_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_offsets, dim3(1), dim3(1024), 0, ctx->stream, counts, blocks, offsets);
    KG_HIP(hipGetLastError());
    hipLaunchKernelGGL((mlp_scatter<LV>), dim3(blocks), dim3(256), 0, ctx->stream, m, step, sc, samples,
                       tr.path_count, (int)tr.antithetic, spot, cash, offsets, data, idx);
    KG_HIP(hipGetLastError());
    int32_t n = 0;
    KG_HIP(hipMemcpyAsync(&n, offsets + blocks, sizeof(n), hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
    if (n == 0) continue;
    // The first fitted date starts cold with four times the epochs; later
    // dates continue from the previous network and its Adam moments.
    steps_last[step] = train(ctx, net, data, n, batch, first ? 4 * epochs : epochs, lr, first, true, adam_step,
                             flags, d_losses + step, &out->gradient_samples);
    first = false;
    hipLaunchKernelGGL(mlp_eval, dim3((n + 63) / 64), dim3(256), 0, ctx->stream, net.packed, data, n, nullptr, m,
                       idx, spot, cash, flags);
    KG_HIP(hipGetLastError());
    KG_HIP(hipMemcpyAsync(nets + (size_t)step * PACK_WORDS, net.packed, PACK_BYTES, hipMemcpyDeviceToDevice,
                          ctx->stream));
    fitted[step] = 1;
  }
  std::vector<float> losses(S, 0.0f);
  KG_HIP(hipMemcpyAsync(losses.data(), d_losses, S * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * 2 * blocks);
  hipLaunchKernelGGL((lsmc_final<KG_BASIS_QUADRATIC, false>), dim3(blocks), dim3(256), 0, ctx->stream, m, nullptr,
                     1, nullptr, 0, 0, samples, tr.path_count, (int)tr.antithetic, spot, cash, parts);
  KG_HIP(hipGetLastError());
  check_flags(ctx);
  Moments r[2];
  merge_and_fetch(ctx, parts, blocks, 2, r);
  out->training_ms = timing_end(ctx);
  int32_t dates = 0;
  for (int k = 0; k < S; ++k) dates += fitted[k];
  out->dates_fitted = dates;
  double loss_sum = 0.0;
  for (int k = 0; k < S; ++k)
    if (fitted[k]) loss_sum += losses[k] / steps_last[k];
  out->mean_loss = dates ? loss_sum / dates : 0.0;
  out->in_sample_price = std::max(scaled_value(r[0].scale, r[0].mean, m.log_discount), h.immediate);
  out->price = out->in_sample_price;
  out->path_steps = (double)tr.path_count * S;
  if (!h.oos) return;
  KG_HIP(hipMemcpyAsync(d_fitted, fitted.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
  const Sampler& pr = h.pricing;
  const int64_t pblocks = pr.blocks();
  Moments* pparts = (Moments*)ctx->partials.reserve(sizeof(Moments) * 2 * pblocks);
  flags = reset_flags(ctx);
  timing_begin(ctx);
  hipLaunchKernelGGL((mlp_price<FP32, LV>), dim3(pblocks), dim3(256), 0, ctx->stream, m, nets, d_fitted, sc,
                     pr.seed, pr.samples(), pr.path_count, (int)pr.antithetic, pparts, flags);
  KG_HIP(hipGetLastError());
  check_flags(ctx);
  Moments q[2];
  merge_and_fetch(ctx, pparts, pblocks, 2, q);
  out->pricing_ms = timing_end(ctx);
  kg_estimate e{};
  fill_estimate(pr, q[0], q[1], 0.0, false, true, &e);
  out->price = std::max(e.price, h.immediate);
  out->standard_error = e.has_standard_error ? e.standard_error : e.path_standard_error;
  out->path_steps += (double)pr.path_count * S;
}

// Diagnostics on caller-given parameters and samples (u0, u1, u2, y rows).
static const float4* upload_samples(kg_context* ctx, const float* samples, int64_t n) {
  if (n < 1 || n > (1 << 24) || !samples) fail(KG_INVALID_INPUT, "1..=16777216 samples are required");
  float4* d = (float4*)ctx->scratch.reserve((size_t)n * 16 + 256);
  KG_HIP(hipMemcpyAsync(d, samples, (size_t)n * 16, hipMemcpyHostToDevice, ctx->stream));
  return d;
}

static Net diag_net(kg_context* ctx, const float* params, int max_batch) {
  if (!params) fail(KG_INVALID_INPUT, "parameters must not be null");
  const Net net = net_buffers(ctx, ctx->mlp_state, max_batch, 0, nullptr);
  KG_HIP(hipMemcpyAsync(net.params, params, P_COUNT * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  launch_update(ctx, net, UPDATE_PACK, 0, 0.0f, 0, nullptr, nullptr);
  return net;
}

}  // namespace mlp

void american_mlp(kg_context* ctx, LsmcHost& h, const kg_lsmc_options& opt, kg_lsmc_report* out) {
  const bool fp32 = h.pricing.fp32, lv = h.model.lv != 0;
  if (fp32)
    lv ? mlp::run<true, true>(ctx, h, opt, out) : mlp::run<true, false>(ctx, h, opt, out);
  else
    lv ? mlp::run<false, true>(ctx, h, opt, out) : mlp::run<false, false>(ctx, h, opt, out);
}

}  // namespace kg

using namespace kg;

extern "C" {

kg_status kg_mlp_gradient(kg_context* ctx, const float* params, const float* samples, int64_t n, float* gradient,
                          double* loss) {
  return guarded(ctx, [&] {
    if (!gradient) fail(KG_INVALID_INPUT, "gradient must not be null");
    if (n > (1 << 20)) fail(KG_INVALID_INPUT, "at most 1,048,576 samples per gradient");
    const float4* d = mlp::upload_samples(ctx, samples, n);
    const mlp::Net net = mlp::diag_net(ctx, params, (int)n);
    uint32_t* flags = reset_flags(ctx);
    mlp::clear(ctx, net.scalars, 1);
    const int parts = mlp::launch_grad(ctx, net, d, 0, (int)n, flags);
    float* d_grad = net.m1;  // free in this diagnostic
    mlp::launch_update(ctx, net, mlp::UPDATE_GRADIENT, parts, 0.0f, 0, d_grad, net.scalars);
    check_flags(ctx);
    KG_HIP(hipMemcpyAsync(gradient, d_grad, mlp::P_COUNT * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    float l = 0.0f;
    KG_HIP(hipMemcpyAsync(&l, net.scalars, sizeof(l), hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l;
  });
}

kg_status kg_mlp_train(kg_context* ctx, float* params, const float* samples, int64_t n, int32_t batch,
                       int32_t steps, double learning_rate) {
  return guarded(ctx, [&] {
    if (batch < 1 || batch > (1 << 20) || steps < 0 || steps > 100000 || !(learning_rate > 0.0))
      fail(KG_INVALID_INPUT, "batch in 1..=1048576, steps in 0..=100000 and a positive learning rate");
    const float4* d = mlp::upload_samples(ctx, samples, n);
    const mlp::Net net = mlp::diag_net(ctx, params, batch);
    uint32_t* flags = reset_flags(ctx);
    mlp::clear(ctx, net.m1, 2 * (int64_t)mlp::P_STRIDE);
    int64_t begin = 0;
    for (int t = 1; t <= steps; ++t) {
      const int nb = (int)std::min<int64_t>(batch, n - begin);
      const int parts = mlp::launch_grad(ctx, net, d, begin, nb, flags);
      mlp::launch_update(ctx, net, mlp::UPDATE_ADAM, parts, (float)learning_rate, t, nullptr, nullptr);
      begin = begin + batch < n ? begin + batch : 0;
    }
    check_flags(ctx);
    KG_HIP(hipMemcpyAsync(params, net.params, mlp::P_COUNT * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

kg_status kg_mlp_forward(kg_context* ctx, const float* params, const float* samples, int64_t n, float* out) {
  return guarded(ctx, [&] {
    if (!out) fail(KG_INVALID_INPUT, "out must not be null");
    const float4* d = mlp::upload_samples(ctx, samples, n);
    const mlp::Net net = mlp::diag_net(ctx, params, 32);
    float* d_out = (float*)ctx->results.reserve((size_t)n * sizeof(float));
    uint32_t* flags = reset_flags(ctx);
    hipLaunchKernelGGL(mlp::mlp_eval, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, ctx->stream, net.packed, d,
                       (int)n, d_out, LsmcModel{}, nullptr, nullptr, nullptr, flags);
    KG_HIP(hipGetLastError());
    check_flags(ctx);
    KG_HIP(hipMemcpyAsync(out, d_out, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}
}
