// Host-side runtime shared by the engine entry points: the context (device,
// stream, workspace, Sobol setup cache), error propagation to the C ABI and
// the estimate logic of MonteCarloEstimate (monte_carlo.rs:859-1001,
// 1412-1452).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kiyosi_amd.h"
#include "device_common.hpp"

namespace kg {

struct Error : std::runtime_error {
  kg_status status;
  Error(kg_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};
[[noreturn]] inline void fail(kg_status s, const std::string& m) { throw Error(s, m); }
[[noreturn]] inline void unstable() {
  fail(KG_COMPUTATION_FAILURE, "Monte Carlo produced a non-finite parameter, path or result");
}

#define KG_HIP(call)                                                                     \
  do {                                                                                   \
    hipError_t kg_err_ = (call);                                                         \
    if (kg_err_ != hipSuccess)                                                           \
      ::kg::fail(KG_DEVICE_ERROR, std::string(#call) + ": " + hipGetErrorString(kg_err_)); \
  } while (0)

// A grow-only device buffer owned by a context.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  void* reserve(size_t n) {
    if (n > bytes) {
      if (ptr) KG_HIP(hipFree(ptr));
      ptr = nullptr;
      bytes = 0;
      KG_HIP(hipMalloc(&ptr, n));
      bytes = n;
    }
    return ptr;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

// Host-built scrambled Sobol directions for (dims, seed), cached like the
// reference's request cache (engines/monte_carlo/sobol.rs:12-69).
struct BridgeBaseKey {
  int dims = 0;
  uint64_t seed = 0;
  int64_t blocks = 0;
  bool fp32 = false;
  bool valid = false;
};

// Scrambled directions [column * dims + dim] (52 columns) and shifts [dim]
// of the last (dims, seed), built on the device on the context's stream.
struct SobolSetup {
  uint32_t dims = 0;
  uint64_t seed = 0;
  bool valid = false;
  DeviceBuffer d_directions, d_shifts;
};

// The book's persistent kernels: one per path class (barrier, digital,
// geometric Asian, arithmetic Asian, stepped note, bridge note, accumulator).
constexpr int BOOK_CLASSES = 7;
// Launch-graph key: precision, sizes, buffer addresses (10 words), then each
// class's tile run and shape (4 words per class).
constexpr int BOOK_GRAPH_KEY = 10 + 4 * BOOK_CLASSES;

}  // namespace kg

// A cancellation token: one int32 in pinned, device-visible host memory.
struct kg_cancel_token {
  volatile int32_t* host = nullptr;
  int32_t* device = nullptr;
};

struct kg_context {
  int device = 0;
  kg_cancel_token* cancel = nullptr;  // bound token, or none
  hipStream_t stream = nullptr;
  hipEvent_t start = nullptr, stop = nullptr;
  std::mutex mu;
  std::string error;
  kg::DeviceBuffer partials, scratch, flags, results;
  // Bridge notes in two passes: per-leg values, the fill queue, the plan's
  // interval rows, and the sample blocks' Sobol base points (kept for the
  // key's dimensions, seed, block count and precision).
  kg::DeviceBuffer bridge_values, bridge_queue, bridge_base;
  kg::BridgeBaseKey bridge_base_key;
  kg::SobolSetup sobol;
  std::vector<uint32_t> joe_kuo;  // direction-number records, loaded once
  kg::DeviceBuffer sobol_raw;     // unscrambled directions [dim][52], seed-independent
  uint32_t sobol_raw_dims = 0;
  // Book engine: device buffers and the autotuned tile shape per precision
  // and path class (0 = not tuned yet).
  kg::DeviceBuffer book_blob, book_items, book_tiles, book_parts, book_out, book_counter;
  // American engine: exercise policy rows, model tables and MLP state.
  kg::DeviceBuffer policy, lsmc_tables, mlp_state;
  int tuned_spt[2][kg::BOOK_CLASSES] = {};
  int tuned_wpc[2][kg::BOOK_CLASSES] = {};
  // The last autocallable note's host plan and staged tables, reused by a
  // call on the same terms, market and precision (any seed); its type lives
  // in autocall.hip.
  std::shared_ptr<void> note_plan;
  // The book's instantiated launch graph, kept across calls while its launch
  // shape and buffers (the key) stay the same; the contracts and tiles it
  // reads are uploaded before each launch.
  hipGraphExec_t book_graph = nullptr;
  uint64_t book_graph_key[kg::BOOK_GRAPH_KEY] = {};
  // Pinned host memory (hipHostMalloc): results and flags of a call in the
  // first PINNED_RESULT_BYTES, then a staging area for asynchronous uploads
  // whose bump offset restarts with every call.
  void* pinned = nullptr;
  size_t pinned_bytes = 0, staging_used = 0;
};

namespace kg {

// Runs `body` with the context's device current and its lock held,
// translating exceptions into status codes and messages.
// The bound token's state; the device word the persistent kernels poll.
inline bool cancel_requested(const kg_context* ctx) { return ctx->cancel && *ctx->cancel->host != 0; }
inline const int32_t* cancel_word(const kg_context* ctx) { return ctx->cancel ? ctx->cancel->device : nullptr; }
// Raises KG_INTERRUPTED when the bound token is raised: called when a call
// starts and between the launches of multi-launch calls.
inline void check_cancel(const kg_context* ctx) {
  if (cancel_requested(ctx)) fail(KG_INTERRUPTED, "operation interrupted");
}

template <class F>
kg_status guarded(kg_context* ctx, F&& body) {
  if (!ctx) return KG_INVALID_INPUT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    check_cancel(ctx);
    KG_HIP(hipSetDevice(ctx->device));
    ctx->staging_used = 0;  // the previous call ended with a synchronisation
    body();
    ctx->error.clear();
    return KG_OK;
  } catch (const Error& e) {
    ctx->error = e.what();
    (void)hipStreamSynchronize(ctx->stream);  // nothing queued may outlive the call
    return e.status;
  } catch (const std::bad_alloc&) {
    ctx->error = "host allocation failed";
    (void)hipStreamSynchronize(ctx->stream);
    return KG_COMPUTATION_FAILURE;
  } catch (const std::exception& e) {
    ctx->error = e.what();
    (void)hipStreamSynchronize(ctx->stream);
    return KG_DEVICE_ERROR;
  }
}

// ------------------------------------------------------------------ settings
struct Sampler {
  int64_t path_count;
  uint64_t seed;
  bool sobol;
  bool antithetic;
  bool fp32;
  int64_t samples() const { return antithetic ? (path_count + 1) / 2 : path_count; }
  int64_t blocks() const { return (samples() + BLOCK_SAMPLES - 1) / BLOCK_SAMPLES; }
  // statistics flag of sample_paths (monte_carlo.rs:1771-1774)
  bool statistics() const { return !sobol && (!antithetic || path_count % 2 == 0); }
};

constexpr int64_t MAX_GPU_PATHS = int64_t(1) << 26;

// MonteCarloSettings::validate (monte_carlo.rs:198-217) and the path-step cap
// (:219-227); allow_large lifts both to the GPU limits.
inline Sampler validate_settings(const kg_settings* s, uint64_t dims) {
  if (!s) fail(KG_INVALID_INPUT, "settings must not be null");
  bool large = s->allow_large != 0;
  int64_t cap = large ? MAX_GPU_PATHS : 10000000;
  if (s->path_count < 1 || s->path_count > cap)
    fail(KG_INVALID_INPUT, large ? "path_count must be between 1 and 67108864"
                                 : "path_count must be between 1 and 10000000");
  if (s->max_active_workers < 0) fail(KG_INVALID_INPUT, "max_active_workers must be nonnegative");
  if (s->sampling != KG_PSEUDO_RANDOM && s->sampling != KG_SOBOL)
    fail(KG_INVALID_INPUT, "unknown sampling");
  if (s->precision != KG_FP64 && s->precision != KG_FP32_STEP)
    fail(KG_INVALID_INPUT, "unknown precision");
  if (s->sampling == KG_SOBOL && (s->path_count & (s->path_count - 1)))
    fail(KG_INVALID_INPUT, "path_count must be a power of two for Sobol sampling");
  if (!s->has_seed)
    fail(KG_INVALID_INPUT,
         "the GPU engine needs a resolved seed; call with_resolved_seed() before pricing");
  if (!large && (double)s->path_count * (double)dims > 1e9)
    fail(KG_INVALID_INPUT,
         "Monte Carlo simulation exceeds one billion path steps; reduce paths or sampling "
         "dimensions");
  if (s->sampling == KG_SOBOL && (dims < 1 || dims > 21201))
    fail(KG_INVALID_INPUT, "Sobol dimension must be in 1..=21,201");
  return Sampler{s->path_count, s->seed, s->sampling == KG_SOBOL, s->antithetic != 0,
                 s->precision == KG_FP32_STEP};
}

// ------------------------------------------------------------------ numerics
inline double h_discounted_product(double scale, double weight, double log_discount) {
  double value = scale * weight;
  double discount = std::exp(log_discount);
  if (scale == 0.0 || weight == 0.0) return 0.0;
  if (std::isnormal(discount) && std::isnormal(value)) return value * discount;
  return std::copysign(std::exp(std::log(scale) + std::log(std::fabs(weight)) + log_discount),
                       weight);
}
inline double scaled_value(double scale, double weight, double log_discount) {
  double r = h_discounted_product(scale, weight, log_discount);
  if (!std::isfinite(r)) unstable();
  return r;
}
// scaled_cash (numerics.rs:175-186): amount * exp(exponent) without
// spurious overflow or underflow.
inline double h_scaled_cash(double amount, double exponent) {
  if (amount == 0.0) return 0.0;
  const double discount = std::exp(exponent), value = amount * discount;
  if (std::isnormal(discount) && std::isnormal(value)) return value;
  return std::exp(std::log(amount) + exponent);
}
inline double log_ratio(double num, double den) {  // numerics.rs:80-92
  double rel = (num - den) / den;
  if (std::fabs(rel) < 0.5) return std::log1p(rel);
  double ratio = num / den;
  if (std::isnormal(ratio)) return std::log(ratio);
  return std::log(num) - std::log(den);
}

// ------------------------------------------------------------------ estimate
// estimate + sampling_standard_error + MonteCarloEstimate::from_sampling
// (monte_carlo.rs:1412-1452, 938-960, 897-936).
inline void fill_estimate(const Sampler& s, const Moments& sample, const Moments& pairs,
                          double log_discount, bool known, bool has_seed, kg_estimate* out) {
  kg_estimate e{};
  e.price = scaled_value(sample.scale, sample.mean, log_discount);
  e.has_seed = has_seed;
  e.seed = has_seed ? s.seed : 0;
  bool has_se = false;
  double se = 0.0;
  if (s.sobol) {
  } else if (known) {
    has_se = true;
  } else {
    const Moments& u = s.antithetic ? pairs : sample;
    if (u.count >= 2.0 && !(s.antithetic && s.path_count % 2 != 0)) {
      double dev = std::sqrt(std::max(u.m2, 0.0) / (u.count - 1.0) / u.count);
      se = scaled_value(u.scale, dev, log_discount);
      has_se = true;
    }
  }
  if (has_seed && has_se && se == 0.0) {
    e.uncertainty_status = KG_ZERO_OBSERVED_VARIANCE;
  } else if (has_se) {
    e.uncertainty_status = KG_AVAILABLE;
    e.has_standard_error = 1;
    e.standard_error = se;
    double hw = 1.959963984540054 * se;
    e.ci_lower = e.price - hw;
    e.ci_upper = e.price + hw;
    if (!std::isfinite(e.ci_lower) || !std::isfinite(e.ci_upper)) unstable();
  } else if (s.sobol) {
    e.uncertainty_status = KG_UNSUPPORTED_SAMPLING;
  } else if (s.antithetic && s.path_count % 2 != 0) {
    e.uncertainty_status = KG_INCOMPLETE_ANTITHETIC_PAIR;
  } else {
    e.uncertainty_status = KG_INSUFFICIENT_OBSERVATIONS;
  }
  // Diagnostic: standard error of the per-path values as if independent.
  if (sample.count >= 2.0) {
    double dev = std::sqrt(std::max(sample.m2, 0.0) / (sample.count - 1.0) / sample.count);
    e.path_standard_error = h_discounted_product(sample.scale, dev, log_discount);
  }
  double keep_steps = out->path_steps, keep_ms = out->device_ms;
  *out = e;
  out->path_steps = keep_steps;
  out->device_ms = keep_ms;
}

// A known value: one Moments add, no sampling (e.g. monte_carlo.rs:1099-1103);
// estimate() reports zero uncertainty for it except under Sobol sampling.
inline void fill_known(double value, bool sobol, kg_estimate* out) {
  if (!std::isfinite(value)) unstable();
  kg_estimate e{};
  e.price = value;
  if (sobol) {
    e.uncertainty_status = KG_UNSUPPORTED_SAMPLING;
  } else {
    e.has_standard_error = 1;
    e.standard_error = 0.0;
    e.ci_lower = e.ci_upper = value;
    e.uncertainty_status = KG_AVAILABLE;
  }
  *out = e;
}

// ------------------------------------------------------------------ runtime
// Merges per-block partials [block][slot] into `slots` results on the device
// and copies them back (see merge_partials in context.hip).
void merge_and_fetch(kg_context* ctx, const Moments* partials, int64_t blocks, int slots,
                     Moments* host_out);
// Clears the flag word, returns a device pointer to it.
uint32_t* reset_flags(kg_context* ctx);
// Reads the flag word back (synchronising the stream) and throws on a raised flag.
void check_flags(kg_context* ctx);
// Builds (or reuses) the scrambled Sobol directions on the device.
const SobolSetup& sobol_setup(kg_context* ctx, uint32_t dims, uint64_t seed);
// Builds scrambled Sobol directions ([column * dims + dim], 52 columns) and
// shifts on the host without touching the cache.
void sobol_tables(kg_context* ctx, uint32_t dims, uint64_t seed, std::vector<uint64_t>& dirs,
                  std::vector<uint64_t>& shifts);
// Scrambled directions (52 columns) and shift of Sobol dimension 0 for a
// seed, on the host without allocating.
void sobol_dimension_zero(uint64_t seed, uint64_t (&dirs)[52], uint64_t* shift);
// finish_sampling with caller-given device memory: merges the partials into
// d_out (slots Moments, then the flag word copied from d_flags), reads them
// back and raises on a flag; allocates nothing. At most 4 slots.
void finish_sampling_into(kg_context* ctx, const Moments* partials, int64_t blocks, int slots, Moments* d_out,
                          const uint32_t* d_flags, Moments* host_out, double* device_ms);
// Per-device set-up of kg_vanilla_price_noalloc's workspace (european.hip),
// with the device current; once per device.
void noalloc_setup(int device);
// Clears a caller-given device flag word with a kernel on the context stream.
void clear_flag_word(kg_context* ctx, uint32_t* word);
// Event timing on the context stream.
void timing_begin(kg_context* ctx);
double timing_end(kg_context* ctx);
// Copies n host bytes to device memory through the pinned staging area:
// asynchronous on the context's stream, and `src` may be released at once.
void upload_async(kg_context* ctx, void* dst, const void* src, size_t n);
// Ends a sampling launch with one synchronisation: records the stop event
// (call right after the sampling kernel), merges the partials, fetches the
// merged Moments and the device flags, raises on a flag and returns the
// sampling kernel's time since timing_begin.
void finish_sampling(kg_context* ctx, const Moments* partials, int64_t blocks, int slots, Moments* host_out,
                     double* device_ms);

}  // namespace kg
