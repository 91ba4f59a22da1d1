// Autocallable engine: snowball, binary/ternary snowball and Phoenix notes
// (MonteCarloAutocallableEngine, crates/kiyosi/src/engines/monte_carlo/
// autocallable.rs).
//
// Host: plan building restates prepare_sample_on_grid (:293-430),
// pricing::initial (instruments/autocallable/pricing.rs:12-46), CashEvent
// (:524-562) and BridgePlan::new (:463-515).
//
// Device: one thread per sample, one 256-thread workgroup per sample block.
// Sample i of block b uses normal dimensions 0..dims-1 of its own draw:
//   Sobol - point 256*b + j, dimension d, built from two LDS tables
//           (low and high four Gray-code bits of j, the block's base point
//           folded into the high table), so no per-path cursor exists;
//   PCG   - the block's stream, skip-ahead to normal j*dims; skipped
//           dimensions advance the stream (monte_carlo.rs:1689-1703).
// The path loop restates bridge_payoffs::<1, true> (:628-736) when the plan
// has a bridge and payoff (:823-862) otherwise. Control flow is uniform per
// wave: every lane walks the same steps with per-lane predicates, so plan
// tables are read once per wave.
#include "autocall_common.hpp"

namespace kg {

template <bool SOBOL, bool FP32>
__global__ void __launch_bounds__(256)
    autocall_kernel(AcPlanDev p, uint64_t seed, int64_t samples, int64_t paths, int antithetic,
                    const uint64_t* __restrict__ dirs, const uint64_t* __restrict__ shifts,
                    int sobol_lds, Moments* __restrict__ partials, uint32_t* __restrict__ flags) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ Moments smem[2 * 4];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const bool live = n < samples;
  const int dims = p.n_steps;
  Draws<SOBOL, FP32> dr;
  dr.dims = dims;
  dr.direct = !sobol_lds;
  dr.dirs = dirs;
  dr.shifts = shifts;
  dr.point = (uint64_t)n;
  if (SOBOL) {
    if (sobol_lds) {
      const uint32_t g = (uint32_t)(j ^ (j >> 1));
      dr.glo = g & 15u;
      dr.ghi = g >> 4;
      if (FP32)
        sobol_build_lds32(dirs, shifts, dims, (uint64_t)block * BLOCK_SAMPLES, lds, &dr.lo, &dr.hi);
      else
        sobol_build_lds64(dirs, shifts, dims, (uint64_t)block * BLOCK_SAMPLES, lds, &dr.lo, &dr.hi);
    }
  } else {
    const uint64_t first = (uint64_t)j * (uint64_t)dims;
    if (FP32) {
      dr.g32.init(seed, (uint64_t)block);
      dr.g32.advance((first >> 1) * 4);
      dr.has32 = false;
      if (first & 1) {
        float e;
        f32::box_muller_pair(dr.g32, &e, &dr.spare32);
        dr.has32 = true;
      }
    } else {
      dr.s.start(seed, (uint64_t)block, first);
    }
  }
  uint32_t bad = 0;
  const bool paired = antithetic && live && n < paths / 2;
  const Draws<SOBOL, FP32> origin = dr;
  double pos = p.bridge ? bridge_path<SOBOL, FP32>(p, dr, live, 1.0f, bad)
                        : step_path<SOBOL, FP32>(p, dr, live, 1.0f, bad);
  double neg = 0.0;
  if (antithetic && __ballot(paired) != 0) {
    dr = origin;
    neg = p.bridge ? bridge_path<SOBOL, FP32>(p, dr, paired, -1.0f, bad)
                   : step_path<SOBOL, FP32>(p, dr, paired, -1.0f, bad);
  }
  Moments m[2] = {moments_zero(), moments_zero()};
  if (live) {
    moments_add(m[0], pos, true);
    if (paired) {
      moments_add(m[0], neg, true);
      moments_add(m[1], midpoint(pos, neg), true);
    }
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
  block_merge<2>(m, smem);
  if (threadIdx.x == 0) {
    partials[block * 2 + 0] = m[0];
    partials[block * 2 + 1] = m[1];
  }
}

template <bool SOBOL, bool FP32>
static void launch(kg_context* ctx, const AcPlanDev& p, const Sampler& s, const uint64_t* dirs,
                   const uint64_t* shifts, Moments* parts, uint32_t* flags) {
  const int dims = p.n_steps;
  size_t lds = 0;
  int sobol_lds = 0;
  if (SOBOL) {
    lds = sobol_lds_bytes(dims, FP32);
    sobol_lds = lds <= SOBOL_LDS_LIMIT;
    if (!sobol_lds) lds = 0;
  }
  auto kernel = autocall_kernel<SOBOL, FP32>;
  if (lds > 64 * 1024)
    KG_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
  hipLaunchKernelGGL(kernel, dim3(s.blocks()), dim3(256), lds, ctx->stream, p, s.seed, s.samples(),
                     s.path_count, (int)s.antithetic, dirs, shifts, sobol_lds, parts, flags);
  KG_HIP(hipGetLastError());
}

}  // namespace kg

using namespace kg;

extern "C" kg_status kg_autocallable_price(kg_context* ctx, const kg_settings* settings,
                                           const kg_autocall_terms* note,
                                           const kg_bsm_market* market, kg_estimate* out) {
  return guarded(ctx, [&] {
    if (!note || !market || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
    AcHostPlan h = build_plan(*note, *market);
    *out = kg_estimate{};
    if (h.fixed) {
      fill_known(h.fixed_value, settings && settings->sampling == KG_SOBOL, out);
      return;
    }
    Sampler s = validate_settings(settings, (uint64_t)h.dev.n_steps);
    AcPlanDev p = upload(ctx, h);
    const uint64_t* dirs = nullptr;
    const uint64_t* shifts = nullptr;
    if (s.sobol) {
      const SobolSetup& so = sobol_setup(ctx, (uint32_t)p.n_steps, s.seed);
      dirs = (const uint64_t*)so.d_directions.ptr;
      shifts = (const uint64_t*)so.d_shifts.ptr;
    }
    const int64_t blocks = s.blocks();
    Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * 2 * blocks);
    uint32_t* flags = reset_flags(ctx);
    timing_begin(ctx);
    if (s.sobol && s.fp32)
      launch<true, true>(ctx, p, s, dirs, shifts, parts, flags);
    else if (s.sobol)
      launch<true, false>(ctx, p, s, dirs, shifts, parts, flags);
    else if (s.fp32)
      launch<false, true>(ctx, p, s, dirs, shifts, parts, flags);
    else
      launch<false, false>(ctx, p, s, dirs, shifts, parts, flags);
    Moments r[2];
    finish_sampling(ctx, parts, blocks, 2, r, &out->device_ms);
    out->path_steps = (double)s.path_count * (double)p.n_steps;
    fill_estimate(s, r[0], r[1], 0.0, false, true, out);
  });
}
