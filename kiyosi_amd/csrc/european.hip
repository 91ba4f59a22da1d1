// European vanilla engine: MonteCarloVanillaEngine (monte_carlo.rs:832-1329).
//
// Plan building restates EuropeanPayoff::new (:1165-1223) on the host; the
// kernel restates TerminalPayoff::at_normal (:1238-1252) over the block
// sampler of sample_paths (:1761-1811): one workgroup per 256-sample block,
// PCG stream = block index, normal j of the block for sample j (one
// dimension), antithetic samples paired while the sample index is below
// paths/2. Each thread owns two consecutive samples, i.e. one Box-Muller pair.
#include "host_common.hpp"
#include "fp32_math.hpp"

namespace kg {

struct EuroParams {
  double sign, diffusion, half_variance, moneyness, log_scale;
  double peak_shift;  // log scale of the second peak-scaling pass
  int mode;           // 0: payoff, 1: peak pass (:1118-1125), 2: rescaled pass (:1127-1129)
};

__device__ inline double euro_value(const EuroParams& p, double z, uint32_t& bad) {
  double distance = fma(p.diffusion, z, p.moneyness);
  if (p.mode == 0) {
    double v;
    if (p.sign * distance <= 0.0)
      v = 0.0;
    else if (p.sign < 0.0)
      v = -expm1(distance);
    else
      v = -expm1(-distance) * exp(fma(p.diffusion, z, -p.half_variance));
    if (!isfinite(v)) bad = 1;
    return v;
  }
  double lv = distance > 0.0
                  ? fma(p.diffusion, z, p.log_scale - p.half_variance) + log(-expm1(-distance))
                  : -INFINITY;
  if (p.mode == 1) {
    if (lv > 1.7976931348623157e308 || isnan(lv)) bad = 1;
    return fmax(lv + 746.0, 0.0);
  }
  return exp(lv - p.peak_shift);
}

// FP32 stepping: the payoff of one normal in single precision; the value
// leaves in FP64 for accumulation.
__device__ inline double euro_value_fp32(const EuroParams& p, float z, uint32_t& bad) {
  float diffusion = (float)p.diffusion;
  float distance = fmaf(diffusion, z, (float)p.moneyness);
  float v;
  if (p.sign * (double)distance <= 0.0)
    v = 0.f;
  else if (p.sign < 0.0)
    v = -f32::expm1(distance);
  else
    v = -f32::expm1(-distance) * f32::exp(fmaf(diffusion, z, -(float)p.half_variance));
  if (!isfinite(v)) bad = 1;
  return (double)v;
}

__device__ inline uint64_t sobol_dim0(const uint64_t* __restrict__ dirs, uint64_t shift,
                                      uint64_t n) {
  uint64_t gray = n ^ (n >> 1);
  uint64_t v = shift;
  while (gray) {
    v ^= dirs[__builtin_ctzll(gray)];
    gray &= gray - 1;
  }
  return v;
}

// The normals of a thread's two samples n0, n0 + 1: Sobol dimension 0, or
// one Box-Muller pair of the block's PCG stream (pair j/2 starts at output
// 4 * (j/2)).
template <bool FP32>
__device__ inline void euro_normals(uint64_t seed, int64_t block, int j, int64_t n0, int sobol,
                                    const uint64_t* __restrict__ dirs, const uint64_t* __restrict__ shift,
                                    double z[2], float zf[2]) {
  if (sobol) {
    for (int s = 0; s < 2; ++s) {
      uint64_t bits = sobol_dim0(dirs, *shift, (uint64_t)(n0 + s));
      if (FP32)
        zf[s] = f32::normal_from_bits((uint32_t)(bits >> 20));
      else
        z[s] = normal_from_bits(bits);
    }
  } else {
    Pcg32 g;
    g.init(seed, (uint64_t)block);
    g.advance((uint64_t)j * 2);
    if (FP32)
      f32::box_muller_pair(g, &zf[0], &zf[1]);
    else
      box_muller_pair(g, &z[0], &z[1]);
  }
}

template <bool FP32>
__global__ void __launch_bounds__(128)
    european_kernel(EuroParams p, uint64_t seed, int64_t samples, int64_t paths, int antithetic,
                    int sobol, const uint64_t* __restrict__ dirs, const uint64_t* __restrict__ shift,
                    Moments* __restrict__ partials, uint32_t* __restrict__ flags) {
  __shared__ Moments smem[2 * 2];
  const int64_t block = blockIdx.x;
  const int j = 2 * threadIdx.x;
  const int64_t n0 = block * BLOCK_SAMPLES + j;
  // Path values (both legs) and pair means of the thread's two samples.
  double vals[2][4];
  int count[2] = {0, 0};
  uint32_t bad = 0;
  if (n0 < samples) {
    double z[2];
    float zf[2];
    euro_normals<FP32>(seed, block, j, n0, sobol, dirs, shift, z, zf);
    for (int s = 0; s < 2; ++s) {
      const int64_t n = n0 + s;
      if (n >= samples) break;
      double pos = FP32 ? euro_value_fp32(p, zf[s], bad) : euro_value(p, z[s], bad);
      vals[0][count[0]++] = pos;
      if (antithetic && n < paths / 2) {
        double neg = FP32 ? euro_value_fp32(p, -zf[s], bad) : euro_value(p, -z[s], bad);
        vals[0][count[0]++] = neg;
        vals[1][count[1]++] = midpoint(pos, neg);
      }
    }
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
  Moments m[2];
  block_moments<2, 4>(vals, count, m, smem);
  if (threadIdx.x == 0) {
    partials[block * 2 + 0] = m[0];
    partials[block * 2 + 1] = m[1];
  }
}

// Shared-draw scenarios (the batched Greek stencil of
// MonteCarloVanillaEngine::european_greek_prices, monte_carlo.rs:1143-1162):
// each thread draws its two normals once and evaluates every scenario's
// payoff on them. Scenario s's block moments, formed exactly as the
// single-scenario kernel forms its totals, go to partials[block * n + s].
template <bool FP32>
__global__ void __launch_bounds__(128)
    european_scenarios_kernel(const EuroParams* __restrict__ params, int n_scen, uint64_t seed, int64_t samples,
                              int64_t paths, int antithetic, int sobol, const uint64_t* __restrict__ dirs,
                              const uint64_t* __restrict__ shift, Moments* __restrict__ partials,
                              uint32_t* __restrict__ flags) {
  __shared__ Moments smem[1 * 2];
  const int64_t block = blockIdx.x;
  const int j = 2 * threadIdx.x;
  const int64_t n0 = block * BLOCK_SAMPLES + j;
  double z[2] = {0.0, 0.0};
  float zf[2] = {0.f, 0.f};
  if (n0 < samples) euro_normals<FP32>(seed, block, j, n0, sobol, dirs, shift, z, zf);
  uint32_t bad = 0;
  for (int sc = 0; sc < n_scen; ++sc) {
    const EuroParams p = params[sc];
    double vals[1][4];
    int count[1] = {0};
    if (n0 < samples) {
      for (int s = 0; s < 2; ++s) {
        const int64_t n = n0 + s;
        if (n >= samples) break;
        vals[0][count[0]++] = FP32 ? euro_value_fp32(p, zf[s], bad) : euro_value(p, z[s], bad);
        if (antithetic && n < paths / 2)
          vals[0][count[0]++] = FP32 ? euro_value_fp32(p, -zf[s], bad) : euro_value(p, -z[s], bad);
      }
    }
    Moments m[1];
    block_moments<1, 4>(vals, count, m, smem);
    if (threadIdx.x == 0) partials[block * n_scen + sc] = m[0];
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
}

// One pass of the sampler: returns total and pair moments.
static void european_pass(kg_context* ctx, const Sampler& s, const EuroParams& p,
                          Moments out[2], double* device_ms) {
  const int64_t blocks = s.blocks();
  Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * 2 * blocks);
  uint32_t* flags = reset_flags(ctx);
  const uint64_t* dirs = nullptr;
  const uint64_t* shift = nullptr;
  if (s.sobol) {
    const SobolSetup& so = sobol_setup(ctx, 1, s.seed);
    dirs = (const uint64_t*)so.d_directions.ptr;
    shift = (const uint64_t*)so.d_shifts.ptr;
  }
  timing_begin(ctx);
  if (s.fp32)
    hipLaunchKernelGGL(european_kernel<true>, dim3(blocks), dim3(128), 0, ctx->stream, p, s.seed,
                       s.samples(), s.path_count, (int)s.antithetic, (int)s.sobol, dirs, shift,
                       parts, flags);
  else
    hipLaunchKernelGGL(european_kernel<false>, dim3(blocks), dim3(128), 0, ctx->stream, p, s.seed,
                       s.samples(), s.path_count, (int)s.antithetic, (int)s.sobol, dirs, shift,
                       parts, flags);
  KG_HIP(hipGetLastError());
  double ms = 0.0;
  finish_sampling(ctx, parts, blocks, 2, out, &ms);
  *device_ms += ms;
}

// EuropeanPayoff::new (monte_carlo.rs:1171-1252) for a flat BSM market:
// either a known value or the sampled payoff's parameters.
static bool plan_european(const kg_european_terms& o, const kg_bsm_market& m, EuroParams* p, double* known) {
  if (o.option_type != 1 && o.option_type != -1)
    fail(KG_INVALID_INPUT, "option_type must be 1 (call) or -1 (put)");
  if (!(m.spot > 0.0) || !std::isfinite(m.spot)) fail(KG_INVALID_INPUT, "spot must be positive");
  if (!(m.volatility >= 0.0)) fail(KG_INVALID_INPUT, "volatility must be nonnegative");
  if (!(o.strike > 0.0)) fail(KG_INVALID_INPUT, "strike must be positive");
  // validate_life (instruments/monitoring.rs:470-481): a European option has
  // no valuation start (instruments/vanilla.rs:198-200), so only valuation
  // after expiry is rejected.
  if (m.valuation_time > o.expiry_time) fail(KG_INVALID_INPUT, "valuation outside contract life");
  const double sign = o.option_type;
  const double time = (double)(o.expiry_time - m.valuation_time) / (365.0 * 86400e6);
  if (time == 0.0) {  // monte_carlo.rs:1188-1192
    *known = std::max(sign * (m.spot - o.strike), 0.0);
    return true;
  }
  if (m.volatility == 0.0) {  // deterministic::vanilla_to_expiry for flat curves
    double fwd = m.spot * std::exp(-m.dividend * time) - o.strike * std::exp(-m.rate * time);
    *known = std::max(sign * fwd, 0.0);
    return true;
  }
  *p = EuroParams{};
  p->sign = sign;
  p->diffusion = m.volatility * std::sqrt(time);
  const double carry = (m.rate - m.dividend) * time;
  p->half_variance = (0.5 * p->diffusion) * p->diffusion;
  const double drift = carry - p->half_variance;
  if (!std::isfinite(drift) || !std::isfinite(p->diffusion)) unstable();
  p->moneyness = log_ratio(m.spot, o.strike) + carry - p->half_variance;
  p->log_scale = sign > 0.0 ? std::log(m.spot) - m.dividend * time : std::log(o.strike) - m.rate * time;
  return false;
}

static bool needs_peak_scaling(const EuroParams& p) { return p.sign > 0.0 && p.diffusion > 25.0; }

// ------------------------------------------------------------------ no-allocation entry
// Device memory of kg_vanilla_price_noalloc, one copy per GPU, held until the
// process ends: the flag word, the merged Moments, Sobol dimension 0 (52
// direction words and the shift) and the per-block partials.
constexpr int64_t NOALLOC_MAX_BLOCKS = 32768;  // 8,388,608 samples
constexpr size_t NOALLOC_FLAGS = 0, NOALLOC_RESULTS = 256, NOALLOC_SOBOL = 512, NOALLOC_PARTIALS = 1024;
constexpr size_t NOALLOC_BYTES = NOALLOC_PARTIALS + (size_t)NOALLOC_MAX_BLOCKS * 2 * sizeof(Moments);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static void* g_scratch[NOALLOC_MAX_DEVICES] = {};

// Called by kg_context_create with `device` current: the device address of
// this GPU's copy of the array (the array's name is a host-side stand-in).
void noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES || g_scratch[device]) return;
  void* ptr = nullptr;
  KG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch[device] = ptr;
}

struct SobolWords {
  uint64_t dirs[52];
  uint64_t shift;
};

// Stores Sobol dimension 0, passed by value, into device memory.
__global__ void store_sobol_words(SobolWords w, uint64_t* __restrict__ out) {
  const int t = threadIdx.x;
  if (t < 52) out[t] = w.dirs[t];
  if (t == 52) out[52] = w.shift;
}

}  // namespace kg

using namespace kg;

extern "C" kg_status kg_vanilla_price(kg_context* ctx, const kg_settings* settings,
                                      const kg_european_terms* option,
                                      const kg_bsm_market* market, kg_estimate* out) {
  return guarded(ctx, [&] {
    if (!option || !market || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
    Sampler s = validate_settings(settings, 1);
    const kg_european_terms& o = *option;
    const kg_bsm_market& m = *market;
    EuroParams p;
    double known = 0.0;
    *out = kg_estimate{};
    if (plan_european(o, m, &p, &known)) {
      fill_known(known, settings->sampling == KG_SOBOL, out);
      return;
    }
    double log_scale = p.log_scale;
    Moments r[2];
    double ms = 0.0;
    if (needs_peak_scaling(p)) {
      // Peak scaling (monte_carlo.rs:1104-1131): a first pass finds the
      // largest log payoff, a second pass accumulates in units of it.
      p.mode = 1;
      european_pass(ctx, s, p, r, &ms);
      log_scale = r[0].scale - 746.0;
      p.mode = 2;
      p.peak_shift = log_scale;
    }
    european_pass(ctx, s, p, r, &ms);
    out->device_ms = ms;
    out->path_steps = (double)s.path_count;
    fill_estimate(s, r[0], r[1], log_scale, m.valuation_time == o.expiry_time, true, out);
  });
}

extern "C" kg_status kg_vanilla_scenarios(kg_context* ctx, const kg_settings* settings,
                                          const kg_european_terms* option, const kg_bsm_market* markets,
                                          int32_t n, kg_estimate* out) {
  return guarded(ctx, [&] {
    if (!option || !markets || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
    if (n < 1 || n > KG_MAX_SCENARIOS) fail(KG_INVALID_INPUT, "1..=64 scenarios per call");
    Sampler s = validate_settings(settings, 1);
    std::vector<EuroParams> params;
    std::vector<int> index(n, -1);
    std::vector<double> known(n, 0.0);
    for (int i = 0; i < n; ++i) {
      EuroParams p;
      if (plan_european(*option, markets[i], &p, &known[i])) continue;
      if (needs_peak_scaling(p))
        fail(KG_UNSUPPORTED_OPERATION, "peak-scaled payoffs are priced one scenario at a time");
      index[i] = (int)params.size();
      params.push_back(p);
    }
    std::vector<Moments> r(params.size());
    double ms = 0.0;
    if (!params.empty()) {
      const int64_t blocks = s.blocks();
      const int sc = (int)params.size();
      EuroParams* d_params = (EuroParams*)ctx->scratch.reserve(params.size() * sizeof(EuroParams));
      upload_async(ctx, d_params, params.data(), params.size() * sizeof(EuroParams));
      Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * (size_t)sc * blocks);
      const uint64_t* dirs = nullptr;
      const uint64_t* shift = nullptr;
      if (s.sobol) {
        const SobolSetup& so = sobol_setup(ctx, 1, s.seed);
        dirs = (const uint64_t*)so.d_directions.ptr;
        shift = (const uint64_t*)so.d_shifts.ptr;
      }
      uint32_t* flags = reset_flags(ctx);
      timing_begin(ctx);
      if (s.fp32)
        hipLaunchKernelGGL(european_scenarios_kernel<true>, dim3(blocks), dim3(128), 0, ctx->stream, d_params, sc,
                           s.seed, s.samples(), s.path_count, (int)s.antithetic, (int)s.sobol, dirs, shift, parts,
                           flags);
      else
        hipLaunchKernelGGL(european_scenarios_kernel<false>, dim3(blocks), dim3(128), 0, ctx->stream, d_params, sc,
                           s.seed, s.samples(), s.path_count, (int)s.antithetic, (int)s.sobol, dirs, shift, parts,
                           flags);
      KG_HIP(hipGetLastError());
      finish_sampling(ctx, parts, blocks, sc, r.data(), &ms);
    }
    for (int i = 0; i < n; ++i) {
      out[i] = kg_estimate{};
      if (index[i] < 0) {
        fill_known(known[i], settings->sampling == KG_SOBOL, &out[i]);
        continue;
      }
      const Moments none{};
      fill_estimate(s, r[index[i]], none, params[index[i]].log_scale,
                    markets[i].valuation_time == option->expiry_time, true, &out[i]);
      // Scenario estimates carry the price and the path standard error only.
      out[i].has_standard_error = 0;
      out[i].standard_error = out[i].ci_lower = out[i].ci_upper = 0.0;
      out[i].device_ms = ms;
      out[i].path_steps = (double)s.path_count;
    }
  });
}

extern "C" kg_status kg_vanilla_price_noalloc(kg_context* ctx, const kg_settings* settings,
                                              const kg_european_terms* option, const kg_bsm_market* market,
                                              kg_estimate* out) {
  return guarded(ctx, [&] {
    if (!option || !market || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
    Sampler s = validate_settings(settings, 1);
    if (s.blocks() > NOALLOC_MAX_BLOCKS)
      fail(KG_INVALID_INPUT, "kg_vanilla_price_noalloc takes at most 8388608 samples (32768 blocks of 256)");
    unsigned char* base = ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES
                              ? (unsigned char*)g_scratch[ctx->device]
                              : nullptr;
    if (!base) fail(KG_DEVICE_ERROR, "the no-allocation workspace of this device is not set up");
    uint32_t* flags = (uint32_t*)(base + NOALLOC_FLAGS);
    Moments* results = (Moments*)(base + NOALLOC_RESULTS);
    uint64_t* sobol = (uint64_t*)(base + NOALLOC_SOBOL);
    Moments* parts = (Moments*)(base + NOALLOC_PARTIALS);
    const kg_european_terms& o = *option;
    const kg_bsm_market& m = *market;
    EuroParams p;
    double known = 0.0;
    *out = kg_estimate{};
    if (plan_european(o, m, &p, &known)) {
      fill_known(known, settings->sampling == KG_SOBOL, out);
      return;
    }
    const uint64_t* shift = sobol + 52;
    if (s.sobol) {
      SobolWords w;
      sobol_dimension_zero(s.seed, w.dirs, &w.shift);
      hipLaunchKernelGGL(store_sobol_words, dim3(1), dim3(64), 0, ctx->stream, w, sobol);
      KG_HIP(hipGetLastError());
    }
    double log_scale = p.log_scale;
    double ms = 0.0;
    Moments r[2];
    auto pass = [&](const EuroParams& q) {
      clear_flag_word(ctx, flags);
      timing_begin(ctx);
      if (s.fp32)
        hipLaunchKernelGGL(european_kernel<true>, dim3(s.blocks()), dim3(128), 0, ctx->stream, q, s.seed,
                           s.samples(), s.path_count, (int)s.antithetic, (int)s.sobol, sobol, shift, parts, flags);
      else
        hipLaunchKernelGGL(european_kernel<false>, dim3(s.blocks()), dim3(128), 0, ctx->stream, q, s.seed,
                           s.samples(), s.path_count, (int)s.antithetic, (int)s.sobol, sobol, shift, parts, flags);
      KG_HIP(hipGetLastError());
      double t = 0.0;
      finish_sampling_into(ctx, parts, s.blocks(), 2, results, flags, r, &t);
      ms += t;
    };
    if (needs_peak_scaling(p)) {
      p.mode = 1;
      pass(p);
      log_scale = r[0].scale - 746.0;
      p.mode = 2;
      p.peak_shift = log_scale;
    }
    pass(p);
    out->device_ms = ms;
    out->path_steps = (double)s.path_count;
    fill_estimate(s, r[0], r[1], log_scale, m.valuation_time == o.expiry_time, true, out);
  });
}
