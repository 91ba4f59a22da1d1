// Barrier (single and double), digital, Asian and accumulator engines, the
// heterogeneous book, and the spot-Greek kernel of the barrier and digital
// engines.
//
// Every contract becomes a BookItem: a plan built on the host (restating the
// reference engine's own plan building, cited per builder below) plus its
// sampler settings. The book's tiles (contract, 256-sample block) are grouped
// by path class; each class has its own persistent kernel, which pulls the
// class's tiles from an atomic queue, longest contracts first, and writes one
// partial per tile; a last kernel merges each contract's partials in block
// order. Tile identity fixes the stream of every sample (PCG stream = block
// index of the contract, Sobol point = sample index), so results do not
// depend on scheduling. The single-contract entry points run a book of one.
//
// Within a tile the 256 samples are spread over 256/spt threads (spt =
// samples per thread); per-sample values go through LDS and are folded by
// the first wave in sample order, so the tile shape chosen by the autotuner
// does not change the arithmetic. The reset, class kernels and merge are
// captured into one hipGraph and replayed.
#include <algorithm>
#include <numeric>

#include "book_device.hpp"
#include "hit_value.hpp"

namespace kg {

enum : int32_t { IK_BARRIER = 0, IK_GEO_ASIAN = 1, IK_ARITH_ASIAN = 2, IK_AUTOCALL = 3, IK_ACCUM = 4, IK_DIGITAL = 5 };

struct BookItem {
  int32_t kind, dims;
  int32_t antithetic, sobol;
  int64_t samples, paths;
  int64_t first_block;  // offset of this contract's partials
  uint64_t seed;
  const uint64_t* dirs;    // Sobol directions [column * dims + dim]
  const uint64_t* shifts;  // Sobol shifts [dim]
  union {
    BarrierDev b;
    GeoAsianDev g;
    ArithAsianDev a;
    AcPlanDev ac;
    AccumDev acc;
    DigitalDev dg;
  } u;
};

// Path classes of the persistent kernel. Each class has a kernel of its own,
// so a kernel's register count is that of one path loop, not of the union of
// all of them.
enum : int {
  BC_BARRIER = 0,
  BC_DIGITAL,
  BC_GEO_ASIAN,
  BC_ARITH_ASIAN,
  BC_AUTOCALL_STEP,
  BC_AUTOCALL_BRIDGE,
  BC_ACCUM,
  BC_COUNT
};
static_assert(BC_COUNT == BOOK_CLASSES, "one tuned shape per path class");

__host__ __device__ inline int book_class(const BookItem& it) {
  switch (it.kind) {
    case IK_BARRIER:
      return BC_BARRIER;
    case IK_DIGITAL:
      return BC_DIGITAL;
    case IK_GEO_ASIAN:
      return BC_GEO_ASIAN;
    case IK_ARITH_ASIAN:
      return BC_ARITH_ASIAN;
    case IK_AUTOCALL:
      return it.u.ac.bridge ? BC_AUTOCALL_BRIDGE : BC_AUTOCALL_STEP;
    default:
      return BC_ACCUM;
  }
}

// ------------------------------------------------------------------ device
template <bool FP32, bool SOBOL, int CLS>
__device__ __attribute__((always_inline)) void sample_body(const BookItem& it, int64_t block, int j, double* pos,
                                                           double* neg, bool* live_out, bool* paired_out,
                                                           uint32_t& bad) {
  const int64_t n = block * BLOCK_SAMPLES + j;
  const bool live = n < it.samples;
  const bool paired = it.antithetic && live && n < it.paths / 2;
  Draws<SOBOL, FP32, 1> dr;
  dr.dims = it.dims;
  dr.direct = true;
  dr.dirs = it.dirs;
  dr.shifts = it.shifts;
  dr.point = (uint64_t)n;
  if (!SOBOL) {
    const uint64_t first = (uint64_t)j * (uint64_t)it.dims;
    if (FP32) {
      dr.start32(it.seed, (uint64_t)block, first);
    } else {
      dr.s.start(it.seed, (uint64_t)block, first);
    }
  }
  const auto origin = dr;
  auto run = [&](bool on, float sign) -> double {
    if constexpr (CLS == BC_BARRIER)
      return barrier_path<FP32, false, false>(it.u.b, dr, on, sign, bad);
    else if constexpr (CLS == BC_DIGITAL)
      return digital_path<FP32, false>(it.u.dg, dr, on, sign, bad);
    else if constexpr (CLS == BC_GEO_ASIAN)
      return geo_asian<FP32>(it.u.g, dr, on, sign, bad);
    else if constexpr (CLS == BC_ARITH_ASIAN)
      return arith_asian<FP32>(it.u.a, dr, on, sign, bad);
    else if constexpr (CLS == BC_AUTOCALL_BRIDGE)
      return bridge_path<SOBOL, FP32>(it.u.ac, dr, on, sign, bad);
    else if constexpr (CLS == BC_AUTOCALL_STEP)
      return step_path<SOBOL, FP32>(it.u.ac, dr, on, sign, bad);
    else
      return accum_path<FP32>(it.u.acc, dr, on, sign, bad);
  };
  *pos = run(live, 1.0f);
  *neg = 0.0;
  if (it.antithetic && __ballot(paired) != 0) {
    dr = origin;
    *neg = run(paired, -1.0f);
  }
  *live_out = live;
  *paired_out = paired;
}

// Moments of one block's 256 sample values staged in LDS (pos, neg and the
// live/paired bits of each sample), formed by the first wave: lane l folds
// samples 4l .. 4l+3 in sample order, then the wave merges. Every kernel that
// prices these contracts uses this one order, so their results agree bit
// for bit. `pairs` receives the antithetic pair means.
__device__ inline void fold_block(const double* pos, const double* neg, const unsigned char* state, Moments* out,
                                  Moments* pairs) {
  Moments m = moments_zero(), p = moments_zero();
  for (int k = 0; k < 4; ++k) {
    const int j = threadIdx.x * 4 + k;
    const unsigned char st = state[j];
    if (st & 1) moments_add(m, pos[j], true);
    if (st & 2) {
      moments_add(m, neg[j], true);
      moments_add(p, midpoint(pos[j], neg[j]), true);
    }
  }
  *out = wave_merge(m);
  *pairs = wave_merge(p);
}

// One persistent kernel per path class. The uniform contract fields and step
// tables are read with scalar loads (the tile index is wave-uniform). At
// least two waves per SIMD: at most 256 VGPRs in either precision.
template <bool FP32, int SPT, int CLS>
__global__ void __launch_bounds__(256, 2)
    book_kernel(const BookItem* __restrict__ items, const int2* __restrict__ tiles, int n_tiles,
                int* __restrict__ counter, Moments* __restrict__ partials,
                uint32_t* __restrict__ flags, const volatile int32_t* cancel) {
  __shared__ int s_tile;
  __shared__ double s_pos[BLOCK_SAMPLES], s_neg[BLOCK_SAMPLES];
  __shared__ unsigned char s_state[BLOCK_SAMPLES];
  uint32_t bad = 0;
  for (int taken = 0;; ++taken) {
    if (threadIdx.x == 0) {
      // Every eighth tile a workgroup takes, thread 0 reads the cancellation
      // word (host memory) first; a raised word ends the workgroup's loop
      // like an exhausted queue.
      if (cancel && (taken & 7) == 0 && *cancel != 0) {
        s_tile = n_tiles;
        raise_flag(flags, FLAG_CANCELLED);
      } else {
        s_tile = atomicAdd(counter, 1);
      }
    }
    __syncthreads();
    // The tile index is the same in every lane; saying so lets the contract's
    // fields and step tables be read with scalar loads.
    const int t = __builtin_amdgcn_readfirstlane(s_tile);
    if (t >= n_tiles) break;  // every wave reads the same index: a uniform exit
    const int2 tb = tiles[t];
    const BookItem& it = items[tb.x];
    for (int s = 0; s < SPT; ++s) {
      const int j = threadIdx.x + s * (BLOCK_SAMPLES / SPT);
      double pos, neg;
      bool live, paired;
      if (it.sobol)
        sample_body<FP32, true, CLS>(it, tb.y, j, &pos, &neg, &live, &paired, bad);
      else
        sample_body<FP32, false, CLS>(it, tb.y, j, &pos, &neg, &live, &paired, bad);
      s_pos[j] = pos;
      s_neg[j] = neg;
      s_state[j] = (unsigned char)(live | (paired << 1));
    }
    __syncthreads();
    if (threadIdx.x < WAVE) {
      Moments m, p;
      fold_block(s_pos, s_neg, s_state, &m, &p);
      if (threadIdx.x == 0) {
        partials[(it.first_block + tb.y) * 2 + 0] = m;
        partials[(it.first_block + tb.y) * 2 + 1] = p;
      }
    }
    __syncthreads();
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
}

// One workgroup per contract: its block partials merged as merge_partials
// merges them (merge_run), so a book result equals the single-contract one.
__global__ void __launch_bounds__(MERGE_THREADS)
    merge_book(const BookItem* __restrict__ items, const Moments* __restrict__ partials,
               Moments* __restrict__ out) {
  __shared__ double smem[4 * MERGE_THREADS / WAVE];
  const BookItem& it = items[blockIdx.x];
  const int64_t blocks = (it.samples + BLOCK_SAMPLES - 1) / BLOCK_SAMPLES;
  for (int s = 0; s < 2; ++s) {
    const Moments m = merge_run(partials + it.first_block * 2 + s, 2, blocks, smem);
    if (threadIdx.x == 0) out[blockIdx.x * 2 + s] = m;
  }
}

__global__ void book_reset(int* counters, uint32_t* flags) {
  if (threadIdx.x < BC_COUNT) counters[threadIdx.x] = 0;
  if (threadIdx.x == 0) *flags = 0u;
}

// Value and likelihood-ratio delta/gamma of one barrier or digital contract
// (barrier.rs:440-472, digital.rs:116-136): one sample per thread, one
// 256-sample block per workgroup, four partials per block in the order of
// sample_paths::<3> (monte_carlo.rs:1777-1796): the three path means, then
// the antithetic pair means of the value.
template <bool FP32, bool SOBOL>
__global__ void __launch_bounds__(256)
    spot_greek_kernel(const BookItem* __restrict__ item, Moments* __restrict__ partials,
                      uint32_t* __restrict__ flags) {
  const BookItem& it = *item;
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const bool live = n < it.samples;
  const bool paired = it.antithetic && live && n < it.paths / 2;
  Draws<SOBOL, FP32, 1> dr;
  dr.dims = it.dims;
  dr.direct = true;
  dr.dirs = it.dirs;
  dr.shifts = it.shifts;
  dr.point = (uint64_t)n;
  if (!SOBOL) {
    const uint64_t first = (uint64_t)j * (uint64_t)it.dims;
    if (FP32) {
      dr.start32(it.seed, (uint64_t)block, first);
    } else {
      dr.s.start(it.seed, (uint64_t)block, first);
    }
  }
  const auto origin = dr;
  uint32_t bad = 0;
  auto run = [&](bool on, float sign, double* out) {
    switch (it.kind) {
      case IK_DIGITAL:
        out[0] = digital_path<FP32, true>(it.u.dg, dr, on, sign, bad, out + 1);
        break;
      case IK_GEO_ASIAN:
        out[0] = geo_asian<FP32, true>(it.u.g, dr, on, sign, bad, out + 1);
        break;
      case IK_ARITH_ASIAN: {
        const ArithPath a = arith_path<FP32, false>(it.u.a, dr, on, sign);
        out[0] = out[1] = out[2] = 0.0;
        if (on) out[0] = arith_payoff_greeks(it.u.a, a, bad, out + 1);
        break;
      }
      default:
        out[0] = barrier_path<FP32, true, true>(it.u.b, dr, on, sign, bad, out + 1);
    }
  };
  double pos[3], neg[3] = {0.0, 0.0, 0.0};
  run(live, 1.0f, pos);
  if (it.antithetic && __ballot(paired) != 0) {
    dr = origin;
    run(paired, -1.0f, neg);
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
  // Each output is folded in the book kernel's order (fold_block), so the
  // price equals kg_barrier_price / kg_digital_price bit for bit.
  __shared__ double s_pos[BLOCK_SAMPLES], s_neg[BLOCK_SAMPLES];
  __shared__ unsigned char s_state[BLOCK_SAMPLES];
  s_state[j] = (unsigned char)(live | (paired << 1));
  for (int k = 0; k < 3; ++k) {
    s_pos[j] = pos[k];
    s_neg[j] = neg[k];
    __syncthreads();
    if (threadIdx.x < WAVE) {
      Moments m, p;
      fold_block(s_pos, s_neg, s_state, &m, &p);
      if (threadIdx.x == 0) {
        partials[block * 4 + k] = m;
        if (k == 0) partials[block * 4 + 3] = p;
      }
    }
    __syncthreads();
  }
}

// Arithmetic Asian strike lanes (MonteCarloAsianEngine::price_strikes,
// asian.rs:197-345): one path per sample, shared by up to STRIKE_LANES strike
// payoffs; lane l's total and pair moments of the block go to
// partials[block * 2n + 2l] and [... + 1], each folded in the book's sample
// order, so a lane equals pricing its strike alone bit for bit.
constexpr int STRIKE_LANES = 32;  // asian.rs:12 (STRIKE_TILE)

template <bool FP32, bool SOBOL>
__global__ void __launch_bounds__(256)
    asian_strikes_kernel(const BookItem* __restrict__ item, const ArithAsianDev* __restrict__ lanes, int n_lanes,
                         Moments* __restrict__ partials, uint32_t* __restrict__ flags) {
  __shared__ double s_pos[BLOCK_SAMPLES], s_neg[BLOCK_SAMPLES];
  __shared__ unsigned char s_state[BLOCK_SAMPLES];
  const BookItem& it = *item;
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const bool live = n < it.samples;
  const bool paired = it.antithetic && live && n < it.paths / 2;
  Draws<SOBOL, FP32, 1> dr;
  dr.dims = it.dims;
  dr.direct = true;
  dr.dirs = it.dirs;
  dr.shifts = it.shifts;
  dr.point = (uint64_t)n;
  if (!SOBOL) {
    const uint64_t first = (uint64_t)j * (uint64_t)it.dims;
    if (FP32) {
      dr.start32(it.seed, (uint64_t)block, first);
    } else {
      dr.s.start(it.seed, (uint64_t)block, first);
    }
  }
  const auto origin = dr;
  // The lanes share the path grid; the control's geometric sum is formed for
  // all of them (it changes no other part of the path state).
  ArithAsianDev grid = lanes[0];
  grid.control = 1;
  const ArithPath pos = arith_path<FP32>(grid, dr, live, 1.0f);
  ArithPath neg{0.0, -INFINITY, 0.0};
  if (it.antithetic && __ballot(paired) != 0) {
    dr = origin;
    neg = arith_path<FP32>(grid, dr, paired, -1.0f);
  }
  uint32_t bad = 0;
  s_state[j] = (unsigned char)(live | (paired << 1));
  for (int l = 0; l < n_lanes; ++l) {
    s_pos[j] = live ? arith_payoff(lanes[l], pos, bad) : 0.0;
    s_neg[j] = paired ? arith_payoff(lanes[l], neg, bad) : 0.0;
    __syncthreads();
    if (threadIdx.x < WAVE) {
      Moments m, pm;
      fold_block(s_pos, s_neg, s_state, &m, &pm);
      if (threadIdx.x == 0) {
        partials[block * 2 * n_lanes + 2 * l] = m;
        partials[block * 2 * n_lanes + 2 * l + 1] = pm;
      }
    }
    __syncthreads();
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
}

// ------------------------------------------------------------------ host
// Result of planning one contract: either a known value or a sampled item.
enum class Estimator { PathMeans, Estimate };
struct Planned {
  bool known = false;
  double value = 0.0;
  Estimator kind = Estimator::Estimate;
  // Known addend of the path mean, with its delta and gamma: the exact
  // mean of the arithmetic Asian's geometric control (asian.rs:612-616), or
  // the analytic at-hit rebate value of a continuous knock-out
  // (barrier.rs:486-531).
  double offset = 0.0;
  double offset_delta = 0.0, offset_gamma = 0.0;
  // Spot Greeks requested and kept after the engine's filters, and the
  // values of a contract resolved before sampling.
  int32_t greeks = 0;
  double known_delta = 0.0, known_gamma = 0.0;
  // Priced by the spot-Greek kernel instead of the persistent kernel
  // (continuously monitored double barriers).
  bool separate = false;
  BookItem item{};
  Sampler sampler{};
  // Tables staged into the shared blob: (blob offset, byte offset of the
  // pointer field inside `item`). Offsets, not addresses, because the plan
  // is moved after it is built.
  std::vector<std::pair<size_t, size_t>> fixups;
  template <class T>
  void stage(size_t blob_offset, T* const& field) {
    fixups.push_back({blob_offset, (size_t)((const unsigned char*)&field - (const unsigned char*)&item)});
  }
};

struct Blob {
  std::vector<unsigned char> bytes;
  template <class T>
  size_t add(const std::vector<T>& v) {
    const size_t off = (bytes.size() + 255) & ~size_t(255);
    // Whole words, so that uload_u8 may read the word holding any entry.
    bytes.resize(off + ((v.size() * sizeof(T) + 4) & ~size_t(3)));
    if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

static inline int64_t floor_div64(int64_t a, int64_t b) {
  int64_t q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

static void check_market(const kg_bsm_market& m) {
  if (!(m.spot > 0.0) || !std::isfinite(m.spot)) fail(KG_INVALID_INPUT, "spot must be positive");
  if (!(m.volatility >= 0.0) || !std::isfinite(m.volatility))
    fail(KG_INVALID_INPUT, "volatility must be nonnegative");
  if (!std::isfinite(m.rate) || !std::isfinite(m.dividend)) fail(KG_INVALID_INPUT, "rates must be finite");
  if (m.volatility == 0.0)
    fail(KG_UNSUPPORTED_OPERATION,
         "zero volatility selects the deterministic engine, which is not part of the GPU engine");
}

// Plain BSM price through lognormal_price's main branch (barrier resolution).
static double vanilla_bsm(double S, double K, double r, double q, double sigma, double T, double sign) {
  if (T == 0.0) return std::max(sign * (S - K), 0.0);
  const double w = sigma * std::sqrt(T);
  const double lm = log_ratio(S, K) + (r - q) * T;
  const double d1 = lm / w + 0.5 * w, d2 = d1 - w;
  auto cdf = [](double x) { return 0.5 * std::erfc(-x * 0.7071067811865476); };
  return sign * (S * std::exp(-q * T) * cdf(sign * d1) - K * std::exp(-r * T) * cdf(sign * d2));
}

// A barrier contract as MonteCarloBarrierEngine::native reads it
// (barrier.rs:170-178): a single barrier, or a double barrier as its lower
// (down-type) barrier and the upper level.
struct BarrierContract {
  int32_t option_type, barrier_type, observation, rebate_timing;
  double strike, barrier, rebate;
  bool has_upper;
  double upper;
  int64_t effective_time, expiry_time;
  int32_t n_observations, touch_state;
  const int64_t* observation_times;
};
static BarrierContract contract_of(const kg_barrier_terms& o) {
  if (o.barrier_type < KG_DOWN_AND_OUT || o.barrier_type > KG_UP_AND_IN)
    fail(KG_INVALID_INPUT, "unknown barrier type");
  return {o.option_type,    o.barrier_type, o.observation, o.rebate_timing,  o.strike,
          o.barrier,        o.rebate,       false,         0.0,              o.effective_time,
          o.expiry_time,    o.n_observations, o.touch_state, o.observation_times};
}
// DoubleBarrierOption::new (instruments/double_barrier.rs:119-144).
static BarrierContract contract_of(const kg_double_barrier_terms& o) {
  if (o.knock_type != KG_DOUBLE_KNOCK_OUT && o.knock_type != KG_DOUBLE_KNOCK_IN)
    fail(KG_INVALID_INPUT, "unknown double-barrier type");
  if (!(o.lower_barrier > 0.0) || !std::isfinite(o.lower_barrier) || !std::isfinite(o.upper_barrier) ||
      !(o.upper_barrier > o.lower_barrier))
    fail(KG_INVALID_INPUT, "double barriers need finite levels with 0 < lower < upper");
  if (o.knock_type == KG_DOUBLE_KNOCK_IN && o.rebate_timing == KG_REBATE_AT_HIT)
    fail(KG_INVALID_INPUT, "knock-in rebates are paid at expiry");
  return {o.option_type,
          o.knock_type == KG_DOUBLE_KNOCK_IN ? KG_DOWN_AND_IN : KG_DOWN_AND_OUT,
          o.observation,
          o.rebate_timing,
          o.strike,
          o.lower_barrier,
          o.rebate,
          true,
          o.upper_barrier,
          o.effective_time,
          o.expiry_time,
          o.n_observations,
          o.touch_state,
          o.observation_times};
}

// Delta and gamma of the Black-Scholes vanilla, the Greeks of a barrier
// resolved to its vanilla payoff (barrier.rs:206-220).
static void vanilla_spot_greeks(double S, double K, double r, double q, double sigma, double T, double sign,
                                double* delta, double* gamma) {
  const double w = sigma * std::sqrt(T);
  const double d1 = (log_ratio(S, K) + (r - q) * T) / w + 0.5 * w;
  const double pdf = std::exp(-0.5 * d1 * d1) * 0.3989422804014327;
  *delta = sign * std::exp(-q * T) * 0.5 * std::erfc(-sign * d1 * 0.7071067811865476);
  *gamma = std::exp(-q * T) * pdf / (S * w);
}

// MonteCarloBarrierEngine::native plan (engines/monte_carlo/barrier.rs:160-290)
// for flat rates. `request` (KG_GREEK_* bits) and `radius` drive the
// spot-Greek filter of :186-205; a continuous knock-out with an at-hit
// rebate adds the analytic hit value and its central spot differences
// over the same radius (:486-531, risk/numerical.rs:359-430).
static Planned plan_barrier(const BarrierContract& o, const kg_bsm_market& m, Blob& blob, int32_t request = 0,
                            double radius = 0.0) {
  Planned pl;
  pl.kind = Estimator::PathMeans;
  if (o.option_type != 1 && o.option_type != -1) fail(KG_INVALID_INPUT, "option_type must be 1 or -1");
  if (!(o.strike > 0.0) || !(o.barrier > 0.0) || !(o.rebate >= 0.0) || !std::isfinite(o.rebate))
    fail(KG_INVALID_INPUT, "strike and barrier must be positive and the rebate nonnegative");
  if (o.observation == KG_OBSERVE_SCHEDULED && (o.n_observations < 1 || !o.observation_times))
    fail(KG_INVALID_INPUT, "scheduled monitoring needs observation times");
  for (int i = 1; i < o.n_observations; ++i)
    if (o.observation_times[i] <= o.observation_times[i - 1])
      fail(KG_INVALID_INPUT, "observation times must be strictly increasing");
  if (!(radius >= 0.0) || !std::isfinite(radius)) fail(KG_INVALID_INPUT, "spot_radius must be nonnegative");
  if (o.touch_state < KG_TOUCH_UNKNOWN || o.touch_state > KG_TOUCH_TOUCHED)
    fail(KG_INVALID_INPUT, "unknown touch state");
  const int64_t val = m.valuation_time, expiry = o.expiry_time;
  if (val < o.effective_time || val > expiry) fail(KG_INVALID_INPUT, "valuation outside contract life");
  check_market(m);
  const bool up = o.barrier_type >= KG_UP_AND_OUT;
  const bool knock_in = o.barrier_type == KG_DOWN_AND_IN || o.barrier_type == KG_UP_AND_IN;
  const bool continuous = o.observation == KG_OBSERVE_CONTINUOUS;
  const bool at_hit = o.rebate_timing == KG_REBATE_AT_HIT;
  const double sign = o.option_type, r = m.rate, q = m.dividend, sigma = m.volatility, S = m.spot;
  const double T = (double)(expiry - val) / (365.0 * 86400e6);
  // Monitoring state (instruments/monitoring.rs:272-313): the prior touch
  // from the supplied history, then a touch at the valuation itself.
  const bool had = continuous ? val > o.effective_time : o.observation_times[0] < val;
  if (!had && o.touch_state == KG_TOUCH_TOUCHED) fail(KG_INVALID_INPUT, "touch predates monitoring");
  if (had && o.touch_state == KG_TOUCH_UNKNOWN) fail(KG_MISSING_HISTORY, "prior barrier touch state is required");
  const bool prior = o.touch_state == KG_TOUCH_TOUCHED;
  bool observed_now = continuous;
  for (int i = 0; i < o.n_observations && !continuous; ++i) observed_now |= o.observation_times[i] == val;
  const bool touched =
      prior || (observed_now && ((up ? S >= o.barrier : S <= o.barrier) || (o.has_upper && S >= o.upper)));
  const bool finished =
      val == expiry || (!continuous && o.observation_times[o.n_observations - 1] <= val);
  const bool last_is_expiry =
      continuous || (o.n_observations > 0 && o.observation_times[o.n_observations - 1] == expiry);
  // payoff_requires_touch (monitoring.rs:314-321; double_barrier.rs:216-224)
  const bool requires_touch =
      o.has_upper ? last_is_expiry && (sign > 0 ? o.strike >= o.upper : o.strike <= o.barrier)
                  : last_is_expiry && (up ? (sign > 0 && o.strike >= o.barrier) : (sign < 0 && o.strike <= o.barrier));
  // Spot Greeks: never at expiry, nor when the stencil crosses a monitored
  // level (spot_discontinuity, monitoring.rs:430-445).
  auto crosses = [&](double level) { return S - radius <= level && level <= S + radius; };
  const bool discontinuity = !prior && observed_now && (crosses(o.barrier) || (o.has_upper && crosses(o.upper)));
  pl.greeks = (val != expiry && !discontinuity) ? request & (KG_GREEK_DELTA | KG_GREEK_GAMMA) : 0;
  const double cash = o.rebate == 0.0 ? 0.0 : o.rebate * std::exp(-r * T);
  const bool fixed = o.rebate == 0.0 && requires_touch;
  // resolved_price_with_state (engines/analytic/barrier.rs:32-80).
  if ((touched && !knock_in) || (!touched && finished && knock_in) || (fixed && !knock_in)) {
    pl.known = true;
    // A prior touch has already paid an at-hit rebate.
    pl.value = (touched && !knock_in) ? (at_hit ? (prior ? 0.0 : o.rebate) : cash)
               : (!touched && finished && knock_in) ? cash
                                                    : 0.0;
    return pl;
  }
  if (touched || finished || fixed) {
    pl.known = true;
    pl.value = vanilla_bsm(S, o.strike, r, q, sigma, T, sign);
    if (pl.greeks) {
      double d, g;
      vanilla_spot_greeks(S, o.strike, r, q, sigma, T, sign, &d, &g);
      pl.known_delta = (pl.greeks & KG_GREEK_DELTA) ? d : 0.0;
      pl.known_gamma = (pl.greeks & KG_GREEK_GAMMA) ? g : 0.0;
    }
    return pl;
  }
  const double carry = r - q, hvr = 0.5 * sigma * sigma, drift = carry - hvr;
  if (continuous && !knock_in && at_hit && o.rebate != 0.0) {
    // The paths carry no rebate; its value at the first hit is analytic.
    auto hit_at = [&](double spot) {
      if (o.has_upper) return hitv::Band(o.barrier, o.upper, spot, r, q, sigma, T).hit(o.rebate, r);
      return hitv::single(o.rebate, std::fabs(log_ratio(o.barrier, spot)), up, drift, sigma, T, r);
    };
    pl.offset = hit_at(S);
    if (pl.greeks) {
      // Central differences over the stencil radius; without a usable
      // radius the spot Greeks are unavailable.
      const double h = radius;
      if (h > 0.0 && S > h) {
        const double above = hit_at(S + h), below = hit_at(S - h), center = pl.offset;
        if (pl.greeks & KG_GREEK_DELTA) pl.offset_delta = (above - below) / h / 2.0;
        if (pl.greeks & KG_GREEK_GAMMA) pl.offset_gamma = ((above - center) - (center - below)) / h / h;
      } else {
        pl.greeks = 0;
      }
    }
  }
  auto years = [](int64_t a, int64_t b) { return (double)(b - a) / (365.0 * 86400e6); };
  auto mean_distance = [&](int64_t date, double level) {
    const double e = years(val, date);
    return log_ratio(S, level) + carry * e - hvr * e;
  };
  std::vector<std::pair<int64_t, bool>> dates;
  if (!continuous)
    for (int i = 0; i < o.n_observations; ++i)
      if (o.observation_times[i] > val) dates.push_back({o.observation_times[i], true});
  if (dates.empty() || dates.back().first != expiry) dates.push_back({expiry, false});
  std::vector<BarrierStep> steps;
  std::vector<float> steps32;
  int64_t prev = val;
  double width_sum = 0.0;
  for (auto [date, observe] : dates) {
    BarrierStep s{};
    const double dt = years(prev, date);
    s.width = sigma * std::sqrt(dt);
    const double elapsed = years(val, date);
    s.mean = drift * elapsed;
    if (s.width <= 0.0 || !std::isfinite(s.width) || !std::isfinite(s.mean)) unstable();
    s.discount = -r * elapsed;
    s.observe = observe;
    s.lower = observe ? mean_distance(date, o.barrier) : 0.0;
    s.upper = observe && o.has_upper ? mean_distance(date, o.upper) : 0.0;
    s.remaining = years(date, expiry);
    steps.push_back(s);
    // FP32 rows (mean, width, lower, upper): the levels of a step without a
    // fixing are NaN, so no comparison with them can report a crossing.
    steps32.push_back((float)s.mean);
    steps32.push_back((float)s.width);
    steps32.push_back(observe ? (float)s.lower : NAN);
    steps32.push_back(observe ? (float)s.upper : NAN);
    width_sum += s.width;
    prev = date;
  }
  BarrierDev b{};
  b.n_steps = (int32_t)steps.size();
  b.up = up;
  b.knock_in = knock_in;
  b.continuous = continuous;
  b.at_hit = at_hit;
  b.has_upper = o.has_upper;
  b.request = pl.greeks;
  b.bounded = std::fabs(std::log(S)) + 10.0 * width_sum + std::fabs(drift * T) < 600.0;
  b.condition_on_hit = knock_in && !continuous && steps.size() > 8 && b.bounded;
  if (b.condition_on_hit)
    for (const BarrierStep& s : steps)
      if (s.observe && sigma * std::sqrt(s.remaining) < 1e-5 && s.remaining > 0)
        fail(KG_UNSUPPORTED_OPERATION, "conditional prices below width 1e-5 are not supported");
  b.sign = sign;
  b.strike = o.strike;
  b.rebate = o.rebate;
  b.spot = S;
  b.q = q;
  b.sigma = sigma;
  b.hvr = hvr;
  b.moneyness = mean_distance(expiry, o.strike);
  b.barrier_abs = std::fabs(log_ratio(o.barrier, S));
  b.initial = -log_ratio(o.barrier, S);
  b.band = o.has_upper ? log_ratio(o.upper, o.barrier) : 0.0;
  b.terminal_barrier = mean_distance(expiry, o.barrier);
  b.log_discount = -r * T;
  pl.item.kind = IK_BARRIER;
  pl.item.dims = b.n_steps;
  pl.item.u.b = b;
  const size_t o1 = blob.add(steps), o2 = blob.add(steps32);
  pl.stage(o1, pl.item.u.b.steps);
  pl.stage(o2, pl.item.u.b.steps32);
  return pl;
}

// MonteCarloDigitalEngine::native plan (engines/monte_carlo/digital.rs:55-108).
static Planned plan_digital(const kg_digital_terms& o, const kg_bsm_market& m, Blob&, int32_t request = 0) {
  Planned pl;
  pl.kind = Estimator::PathMeans;
  if (o.option_type != 1 && o.option_type != -1) fail(KG_INVALID_INPUT, "option_type must be 1 or -1");
  if (o.payout_kind != KG_CASH_OR_NOTHING && o.payout_kind != KG_ASSET_OR_NOTHING)
    fail(KG_INVALID_INPUT, "unknown digital payout kind");
  if (!(o.strike > 0.0) || !std::isfinite(o.strike)) fail(KG_INVALID_INPUT, "strike must be positive and finite");
  if (o.expiry_time < o.effective_time) fail(KG_INVALID_INPUT, "expiry precedes the effective time");
  const bool asset = o.payout_kind == KG_ASSET_OR_NOTHING;
  if (!asset && (!(o.cash_payout >= 0.0) || !std::isfinite(o.cash_payout)))
    fail(KG_INVALID_INPUT, "payout must be finite and nonnegative");
  // No valuation start (instruments/vanilla.rs:375-377): only expiry bounds it.
  if (m.valuation_time > o.expiry_time) fail(KG_INVALID_INPUT, "valuation outside contract life");
  check_market(m);
  const bool call = o.option_type > 0;
  const double S = m.spot, sigma = m.volatility, r = m.rate, q = m.dividend;
  const double time = (double)(o.expiry_time - m.valuation_time) / (365.0 * 86400e6);
  if (time == 0.0) {  // Payoff::digital (engines.rs:212-227); the expiry stencil gives the Greeks
    pl.known = true;
    pl.value = (call ? S - o.strike : o.strike - S) > 0.0 ? (asset ? S : o.cash_payout) : 0.0;
    return pl;
  }
  pl.greeks = request & (KG_GREEK_DELTA | KG_GREEK_GAMMA);
  const double root = std::sqrt(time), width = sigma * root, carry = (r - q) * time;
  const double lm = log_ratio(S, o.strike) + carry;
  const double standardized = width < std::numeric_limits<double>::min() ? (lm / sigma) / root : lm / width;
  DigitalDev d{};
  d.call = call;
  d.asset = asset;
  d.request = pl.greeks;
  d.d2 = standardized - 0.5 * width;
  d.amount = asset ? S : o.cash_payout;
  d.drift = asset ? (-q - 0.5 * sigma * sigma) * time : -r * time;
  if (!std::isfinite(width) || !std::isfinite(carry) || !std::isfinite(d.drift)) unstable();
  d.width = width;
  d.log_spot = std::log(S);
  d.log_width = std::log(sigma) + std::log(root);
  d.log_amount = std::log(d.amount);
  pl.item.kind = IK_DIGITAL;
  pl.item.dims = 1;
  pl.item.u.dg = d;
  return pl;
}

// MonteCarloAsianEngine::prepare (engines/monte_carlo/asian.rs:374-551) and
// geometric_control (:1016-1061) for flat rates: seasoned averages (a
// realized average over the elapsed part of the window), forward starts and
// single fixings included.
static Planned plan_asian(const kg_asian_terms& o, const kg_bsm_market& m, Blob&, int32_t request = 0) {
  Planned pl;
  const int32_t spot_greeks = request & (KG_GREEK_DELTA | KG_GREEK_GAMMA);
  pl.kind = Estimator::PathMeans;
  if (o.option_type != 1 && o.option_type != -1) fail(KG_INVALID_INPUT, "option_type must be 1 or -1");
  if (o.step_count < 2 || o.step_count > 10000)
    fail(KG_INVALID_INPUT, "step_count is outside the supported Monte Carlo range");
  if (!(o.strike > 0.0)) fail(KG_INVALID_INPUT, "strike must be positive");
  const int64_t val = m.valuation_time;
  if (val < o.effective_time || val > o.expiry_time)
    fail(KG_INVALID_INPUT, "valuation must be within contract life");
  if (o.averaging_start_time < o.effective_time || o.averaging_start_time > o.expiry_time)
    fail(KG_INVALID_INPUT, "averaging must start within the contract life");
  // AverageTerms::validate (instruments/asian.rs:105-124)
  if (val > o.averaging_start_time && !o.has_realized)
    fail(KG_MISSING_HISTORY, "prior realized average is required");
  if (val <= o.averaging_start_time && o.has_realized)
    fail(KG_INVALID_INPUT, "realized average must match the elapsed averaging period");
  if (o.has_realized && (!(o.realized_average > 0.0) || !std::isfinite(o.realized_average)))
    fail(KG_INVALID_INPUT, "realized average must be finite and positive");
  const double sign = o.option_type;
  auto years = [](int64_t a, int64_t b) { return (double)(b - a) / (365.0 * 86400e6); };
  const double tau = years(val, o.expiry_time);
  if (tau == 0.0) {  // AverageTerms::settlement (:74-79)
    pl.known = true;
    pl.value = std::max(sign * ((o.has_realized ? o.realized_average : m.spot) - o.strike), 0.0);
    return pl;
  }
  check_market(m);
  const double period = years(o.averaging_start_time, o.expiry_time);
  const bool single = period == 0.0;
  const bool geometric = o.averaging == KG_AVERAGE_GEOMETRIC;
  const double sigma = m.volatility, r = m.rate, q = m.dividend;
  const double carry = r - q, drift_rate = carry - 0.5 * sigma * sigma;
  const double log_discount = -r * tau;
  const double elapsed = years(o.averaging_start_time, val);
  const double realized = o.has_realized ? o.realized_average : 0.0;
  if (!geometric && !single) {
    // adjusted_strike (instruments/asian.rs:221-233); at or below zero the
    // call is exercised on every path and priced in closed form
    // (TurnbullWakemanArithmeticAveragePriceEngine, engines/analytic/asian.rs
    // :279-292), the put is worthless; delta = exp(forward_log) for calls.
    const double adjusted = elapsed > 0.0 ? o.strike + elapsed / tau * (o.strike - realized) : o.strike;
    if (adjusted <= 0.0) {
      const double x = carry * tau;  // exprel (engines/analytic/asian.rs:21-23)
      const double m1 = x == 0.0 ? 1.0 : std::expm1(x) / x;
      if (!std::isfinite(m1) || m1 <= 0.0) fail(KG_COMPUTATION_FAILURE, "invalid arithmetic moment");
      const double scale = tau / period;
      const double forward_log = std::log(m1) + std::log(scale) - r * tau;
      pl.known = true;
      if (sign < 0.0) {
        pl.value = 0.0;
      } else {
        const double accrued = std::max(realized * (elapsed / period) - o.strike, 0.0);
        pl.value = h_scaled_cash(accrued, -r * tau) + h_scaled_cash(m.spot, forward_log);
      }
      if (!std::isfinite(pl.value)) unstable();
      pl.greeks = spot_greeks;
      pl.known_delta = sign > 0.0 ? std::exp(forward_log) : 0.0;
      pl.known_gamma = 0.0;
      return pl;
    }
  }
  const double lead = single ? 0.0 : std::max(years(val, o.averaging_start_time), 0.0);
  const double future = single ? tau : years(std::max(val, o.averaging_start_time), o.expiry_time);
  const double weight = single ? 1.0 : future / period;
  const double prior_weight = val > o.averaging_start_time ? elapsed / period : 0.0;
  if (geometric || single) {
    // GeometricParameters (asian.rs:452-484): a single fixing is the
    // terminal spot (variance time tau, forward moneyness, drift -sigma^2 tau / 2).
    const double vt = single ? tau : lead + future / 3.0;
    GeoAsianDev g{};
    g.im = (single ? log_ratio(m.spot, o.strike) + (-q * tau + r * tau) : weight * log_ratio(m.spot, o.strike)) +
           (val > o.averaging_start_time ? prior_weight * log_ratio(realized, o.strike) : 0.0);
    g.drift = single ? -0.5 * sigma * sigma * tau : weight * drift_rate * (lead + future / 2.0);
    g.width = sigma * (weight * std::sqrt(vt));
    g.log_width = std::log(sigma) + std::log(weight) + 0.5 * std::log(vt);
    g.log_strike = std::log(o.strike);
    g.sign = sign;
    g.log_discount = log_discount;
    g.log_spot = std::log(m.spot);
    g.log_weight = std::log(weight);
    g.log_base_width = std::log(sigma) + 0.5 * std::log(vt);
    g.request = spot_greeks;
    pl.greeks = spot_greeks;
    if (!std::isfinite(g.drift) || !std::isfinite(g.width)) unstable();
    pl.item.kind = IK_GEO_ASIAN;
    pl.item.dims = 1;
    pl.item.u.g = g;
    return pl;
  }
  ArithAsianDev a{};
  a.n = o.step_count;
  a.spot = m.spot;
  a.strike = o.strike;
  a.weight = weight;
  a.log_discount = log_discount;
  const double dt = future / o.step_count;
  a.interval_forward = carry * dt;
  a.drift = drift_rate * dt;
  a.width = sigma * std::sqrt(dt);
  a.log_width = std::log(sigma) + 0.5 * std::log(dt);
  a.request = spot_greeks;
  a.lead = lead > 0.0;
  a.lead_drift = drift_rate * lead;
  a.lead_width = sigma * std::sqrt(lead);
  a.log_history = realized > 0.0 && prior_weight > 0.0 ? std::log(realized) + std::log(prior_weight) : -INFINITY;
  a.prior_baseline = prior_weight * ((realized - o.strike) / m.spot);
  a.adjusted_strike = std::fma(-prior_weight, realized, o.strike);
  pl.greeks = spot_greeks;
  a.sign = sign;
  if (!std::isfinite(a.drift) || !std::isfinite(a.width) || !std::isfinite(a.lead_drift) ||
      !std::isfinite(a.lead_width) || !std::isfinite(a.interval_forward))
    unstable();
  if (a.width < 1e-5) fail(KG_UNSUPPORTED_OPERATION, "arithmetic widths below 1e-5 are not supported");
  const double n = o.step_count, fw = 1.0 - 0.5 / n;
  const double las = std::log(a.spot) + std::log(a.weight), ls = std::log(a.adjusted_strike);
  if (!(a.width > 0.8 || std::fabs(a.log_discount) > 50.0 || std::fabs(std::log(a.spot)) > 200.0 ||
        std::fabs(las) > 200.0 || !std::isfinite(ls))) {
    double mean = a.lead_drift + fw * a.drift;
    for (int s = 1; s < o.step_count; ++s) mean += (fw - (double)s / n) * a.drift;
    const double variance = a.lead_width * a.lead_width + (a.width * a.width) * (n / 3.0 - 1.0 / (12.0 * n));
    const double w = std::sqrt(variance);
    const double lf = las + mean + 0.5 * variance;
    if (w >= 0.02 && w <= 0.8 && std::fabs(lf - ls) <= 2.0 * w) {
      a.control = 1;
      a.c_log_strike = ls;
      a.c_las = las;
      a.c_fw = fw;
      const double d1 = (lf - ls) / w + 0.5 * w, d2 = d1 - w;
      auto lcdf = [](double x) {
        if (x >= -10.0) return std::log(0.5 * std::erfc(-x * 0.7071067811865476));
        const double inv_sq = (1.0 / x) / x;
        double term = 1.0, sum = 1.0;
        for (int i = 1; i <= 20; ++i) {
          term *= -(double)(2 * i - 1) * inv_sq;
          sum += term;
        }
        return -0.5 * x * x - std::log(-x) + std::log(sum / std::sqrt(2.0 * M_PI));
      };
      auto pdiff = [](double x, double y) {
        if (x == INFINITY) return x;
        if (y == -INFINITY) return std::exp(x);
        const double d = -std::expm1(y - x);
        return d > 0.0 ? std::exp(x + std::log(d)) : 0.0;
      };
      const double la = lf + log_discount + lcdf(sign * d1);
      const double lc = ls + log_discount + lcdf(sign * d2);
      pl.offset = sign > 0 ? pdiff(la, lc) : pdiff(lc, la);
      // geometric_values' delta and gamma (asian.rs:1084-1096)
      pl.offset_delta = sign * std::exp(la - std::log(m.spot));
      pl.offset_gamma = std::exp(lf + log_discount - 2.0 * std::log(m.spot) - std::log(w) - 0.5 * d1 * d1 -
                                  0.5 * std::log(2.0 * M_PI));
    }
  }
  pl.item.kind = IK_ARITH_ASIAN;
  pl.item.dims = o.step_count - 1 + a.lead;
  pl.item.u.a = a;
  return pl;
}

// The accumulator's schedules and state (instruments/accumulator.rs).
struct AccumTerms {
  const kg_accumulator_terms& o;
  static constexpr int64_t DAY = 86400LL * 1000000LL;
  bool trading(int64_t t) const {  // TradingCalendar::is_trading_day of t's date
    const int64_t d = floor_div64(t, DAY);
    return o.calendar == KG_CALENDAR_ALL_DAYS || ((d % 7) + 7 + 3) % 7 < 5;
  }
  int64_t at_clock(int64_t day) const { return day * DAY + o.observation_clock_micros; }
  bool on_clock(int64_t t) const { return t - floor_div64(t, DAY) * DAY == o.observation_clock_micros; }
  bool explicit_ko() const { return o.has_knock_out_schedule != 0; }
  // ObservationSchedule::daily_times (time/schedule.rs:126-151)
  std::vector<int64_t> daily_times(int64_t start, int64_t end, bool include_start) const {
    std::vector<int64_t> out;
    for (int64_t d = floor_div64(start, DAY);; ++d) {
      const int64_t t = at_clock(d);
      if ((t > start || (include_start && t == start)) && t <= end && trading(t)) out.push_back(t);
      if (d == floor_div64(end, DAY)) break;
    }
    return out;
  }
  // ObservationSchedule::any_daily (:153-172) over [start, end].
  template <class P>
  bool any_daily(int64_t start, int64_t end, P pred) const {
    for (int64_t d = floor_div64(start, DAY);; ++d) {
      const int64_t t = at_clock(d);
      if (t > end) break;
      if (t >= start && trading(t) && pred(t)) return true;
      if (d == floor_div64(end, DAY)) break;
    }
    return false;
  }
  template <class P>
  bool any_accumulation(int64_t start, int64_t end, P pred) const {  // :494-507
    return any_daily(std::max(start, o.effective_time), std::min(end, o.expiry_time), pred);
  }
  bool observes_knock_out(int64_t t) const {  // :512-517
    if (!explicit_ko()) return on_clock(t);
    return std::binary_search(o.knock_out_observation_times, o.knock_out_observation_times + o.n_knock_out_observations,
                              t);
  }
  bool accumulates_at(int64_t t) const {  // :521-526
    return o.effective_time <= t && t <= o.expiry_time && on_clock(t) && trading(t);
  }
  bool is_knock_out(double spot) const { return o.option_type > 0 ? spot >= o.knock_out_level : spot <= o.knock_out_level; }
  double unit_payoff(double spot) const { return o.option_type > 0 ? spot - o.strike : o.strike - spot; }
  double purchase(double spot) const {  // :508-511
    const double distance = spot - o.strike;
    const bool accelerated = o.option_type > 0 ? distance < 0.0 : distance > 0.0;
    return accelerated ? o.daily_quantity * o.acceleration_factor : o.daily_quantity;
  }
  bool accrues(int64_t val) const {  // :462-476
    if (o.knock_out_settled == KG_SETTLED_YES || o.daily_quantity == 0.0 || val >= o.expiry_time) return false;
    return any_accumulation(val, o.expiry_time, [&](int64_t t) {
      return t > val && (o.acceleration_factor != 0.0 || !is_knock_out(o.strike) || !observes_knock_out(t));
    });
  }
};

// MonteCarloAccumulatorEngine::sample plan (engines/monte_carlo/accumulator.rs:91-180):
// history validation (validate_pricing, instruments/accumulator.rs:319-358),
// the valuation-date state (initial, :528-568), the fixing dates (future
// observation times :360-390, or the fixed purchase events of an explicit
// knock-out schedule, :401-454) and the per-step transition tables.
static Planned plan_accumulator(const kg_accumulator_terms& o, const kg_bsm_market& m, Blob& blob) {
  Planned pl;
  if (o.option_type != 1 && o.option_type != -1) fail(KG_INVALID_INPUT, "option_type must be 1 or -1");
  if (!(o.strike > 0.0) || !std::isfinite(o.strike)) fail(KG_INVALID_INPUT, "strike must be positive and finite");
  if (!(o.knock_out_level > 0.0) || !std::isfinite(o.knock_out_level))
    fail(KG_INVALID_INPUT, "knock-out level must be positive and finite");
  if (!(o.daily_quantity >= 0.0) || !std::isfinite(o.daily_quantity) || !(o.acceleration_factor >= 0.0) ||
      !std::isfinite(o.acceleration_factor))
    fail(KG_INVALID_INPUT, "quantities must be finite and nonnegative");
  if (o.has_accumulated_quantity && (!(o.accumulated_quantity >= 0.0) || !std::isfinite(o.accumulated_quantity)))
    fail(KG_INVALID_INPUT, "accumulated_quantity must be finite and nonnegative");
  if (o.expiry_time < o.effective_time) fail(KG_INVALID_INPUT, "expiry precedes the effective time");
  if (o.knock_out_settled < KG_SETTLED_UNKNOWN || o.knock_out_settled > KG_SETTLED_YES)
    fail(KG_INVALID_INPUT, "unknown knock-out settlement state");
  if (o.has_knock_out_schedule != 0 && o.has_knock_out_schedule != 1)
    fail(KG_INVALID_INPUT, "has_knock_out_schedule must be 0 or 1");
  if (o.n_knock_out_observations < 0 || (o.n_knock_out_observations > 0 && !o.knock_out_observation_times) ||
      (!o.has_knock_out_schedule && o.n_knock_out_observations != 0))
    fail(KG_INVALID_INPUT, "knock-out observation times are missing or unexpected");
  const AccumTerms a{o};
  for (int i = 0; i < o.n_knock_out_observations; ++i) {
    const int64_t t = o.knock_out_observation_times[i];
    if (i && t <= o.knock_out_observation_times[i - 1])
      fail(KG_INVALID_INPUT, "knock-out observation times must be strictly increasing");
    if (t < o.effective_time || t > o.expiry_time)
      fail(KG_INVALID_INPUT, "knock-out observation times must lie within the contract life");
    if (!a.trading(t)) fail(KG_INVALID_INPUT, "knock-out observation times must be trading days");
  }
  check_market(m);
  const int64_t val = m.valuation_time;
  if (val < o.effective_time || val > o.expiry_time)  // validate_life (monitoring.rs:470-481)
    fail(KG_INVALID_INPUT, "valuation outside contract life");
  // validate_pricing (:319-358)
  const bool had_accumulation = a.any_accumulation(o.effective_time, val, [&](int64_t t) { return t < val; });
  bool had_knock_out = had_accumulation;
  if (a.explicit_ko()) {
    had_knock_out = false;
    for (int i = 0; i < o.n_knock_out_observations; ++i) had_knock_out |= o.knock_out_observation_times[i] < val;
  }
  if (had_knock_out && o.knock_out_settled == KG_SETTLED_UNKNOWN)
    fail(KG_MISSING_HISTORY, "prior knock-out settlement state is required");
  if ((had_accumulation || had_knock_out) && o.knock_out_settled != KG_SETTLED_YES && !o.has_accumulated_quantity)
    fail(KG_MISSING_HISTORY, "prior accumulated quantity is required");
  if (o.knock_out_settled == KG_SETTLED_YES) {
    if (!had_knock_out) fail(KG_INVALID_INPUT, "knock-out settlement predates knock-out observations");
  } else if (o.has_accumulated_quantity && o.accumulated_quantity > 0.0 && !had_accumulation) {
    fail(KG_INVALID_INPUT, "accumulated_quantity predates accumulation observations");
  }
  const double S = m.spot;
  // initial (:528-568)
  auto known = [&](double value) {
    if (!std::isfinite(value)) unstable();
    pl.known = true;
    pl.value = value;
    return pl;
  };
  if (o.knock_out_settled == KG_SETTLED_YES) return known(0.0);
  double quantity = o.has_accumulated_quantity ? o.accumulated_quantity : 0.0;
  if (a.observes_knock_out(val) && a.is_knock_out(S) && a.trading(val)) return known(quantity * a.unit_payoff(S));
  if (a.accumulates_at(val)) quantity += a.purchase(S);
  const bool accrues = a.accrues(val);
  if (quantity == 0.0 && !accrues) return known(0.0);
  if (val == o.expiry_time || (!accrues && m.rate == 0.0 && m.dividend == 0.0))
    return known(quantity * a.unit_payoff(S));
  if (!std::isfinite(quantity)) unstable();
  // fixed_purchase_events (:401-454)
  struct Event {
    int64_t time;
    double before, at;
  };
  std::vector<Event> events;
  bool fixed = false;
  if (a.explicit_ko() && !(o.acceleration_factor != 1.0 && o.daily_quantity != 0.0)) {
    // Valuation is at or after the effective time, so fixings start after it.
    std::vector<int64_t> fixings;
    if (o.daily_quantity != 0.0) fixings = a.daily_times(val, o.expiry_time, false);
    if ((double)fixings.size() * o.daily_quantity <= std::numeric_limits<double>::max()) {
      const int64_t* ko = o.knock_out_observation_times;
      const int nk = o.n_knock_out_observations;
      const int64_t* first = std::upper_bound(ko, ko + nk, val);
      std::vector<int64_t> times(first, ko + nk);
      if (times.empty() || times.back() != o.expiry_time) times.push_back(o.expiry_time);
      size_t previous = 0, grouped = 0;
      for (int64_t t : times) {
        const size_t before =
            previous + (size_t)(std::lower_bound(fixings.begin() + previous, fixings.end(), t) - (fixings.begin() + previous));
        const bool at = before < fixings.size() && fixings[before] == t;
        grouped += before - previous;
        events.push_back(Event{t, (double)(before - previous) * o.daily_quantity, at ? o.daily_quantity : 0.0});
        previous = before + (at ? 1 : 0);
      }
      fixed = grouped >= 2 || o.daily_quantity == 0.0;
    }
  }
  std::vector<int64_t> dates;
  if (fixed) {
    for (const Event& e : events) dates.push_back(e.time);
  } else {
    // future_observation_times (:360-390)
    dates = a.daily_times(val, o.expiry_time, false);
    for (int i = 0; i < o.n_knock_out_observations; ++i)
      if (o.knock_out_observation_times[i] > val) dates.push_back(o.knock_out_observation_times[i]);
    dates.push_back(o.expiry_time);
    std::sort(dates.begin(), dates.end());
    dates.erase(std::unique(dates.begin(), dates.end()), dates.end());
  }
  if (dates.empty()) unstable();
  std::vector<double> steps;
  std::vector<float> steps32;
  std::vector<uint8_t> flags;
  int64_t prev = val;
  for (size_t k = 0; k < dates.size(); ++k) {
    const int64_t at = dates[k];
    double drift, diff, disc;
    transition(m.rate, m.dividend, m.volatility, years(prev, at), years(val, at), &drift, &diff, &disc);
    steps.push_back(drift);
    steps.push_back(diff);
    steps.push_back(disc);
    steps.push_back(fixed ? events[k].before : 0.0);
    steps.push_back(fixed ? events[k].at : 0.0);
    steps32.push_back((float)drift);
    steps32.push_back((float)diff);
    flags.push_back((uint8_t)((a.observes_knock_out(at) ? ACC_OBSERVE : 0) |
                              (!fixed && a.accumulates_at(at) ? ACC_ACCUMULATE : 0)));
    prev = at;
  }
  AccumDev d{};
  d.n_steps = (int32_t)dates.size();
  d.fixed = fixed;
  d.sign = o.option_type;
  d.strike_log = log_ratio(o.strike, S);
  d.ko_log = log_ratio(o.knock_out_level, S);
  d.strike = o.strike;
  d.spot = S;
  d.daily_qty = o.daily_quantity;
  d.accel = o.acceleration_factor;
  d.acc_qty = quantity;
  const double lo = std::log(std::numeric_limits<double>::min()) + 1.0;
  const double hi = std::log(std::numeric_limits<double>::max()) - 1.0;
  d.safe_lo = std::max(lo, lo - std::log(S));
  d.safe_hi = std::min(hi, hi - std::log(S));
  d.terminal_discount = steps[5 * (dates.size() - 1) + 2];
  pl.item.kind = IK_ACCUM;
  pl.item.dims = d.n_steps;
  pl.item.u.acc = d;
  const size_t o1 = blob.add(steps), o2 = blob.add(steps32), o3 = blob.add(flags);
  pl.stage(o1, pl.item.u.acc.steps);
  pl.stage(o2, pl.item.u.acc.steps32);
  pl.stage(o3, pl.item.u.acc.flags);
  return pl;
}

static Planned plan_autocall(const kg_autocall_terms& t, const kg_bsm_market& m, Blob& blob) {
  Planned pl;
  AcHostPlan h = build_plan(t, m);
  if (h.fixed) {
    pl.known = true;
    pl.value = h.fixed_value;
    return pl;
  }
  pl.item.kind = IK_AUTOCALL;
  pl.item.dims = h.dev.n_steps;
  pl.item.u.ac = h.dev;
  AcPlanDev& p = pl.item.u.ac;
  pl.stage(blob.add(h.steps), p.steps);
  pl.stage(blob.add(h.step_event), p.step_event);
  pl.stage(blob.add(h.step_observe), p.step_observe);
  pl.stage(blob.add(h.events), p.events);
  pl.stage(blob.add(h.intervals), p.intervals);
  pl.stage(blob.add(h.weight), p.weight);
  pl.stage(blob.add(h.fill), p.fill);
  pl.stage(blob.add(h.steps32), p.steps32);
  pl.stage(blob.add(h.wf32), p.wf32);
  pl.stage(blob.add(h.events32), p.events32);
  pl.stage(blob.add(h.intervals32), p.intervals32);
  return pl;
}

static int cu_count(kg_context* ctx) {
  int cus = 0;
  KG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  return std::max(cus, 1);
}

// The tiles of one path class: a contiguous run of the tile list, and the
// tile shape its kernel runs with.
struct ClassRun {
  int begin = 0, count = 0, spt = 0, wpc = 0;
};

template <bool FP32, int CLS>
static void launch_class(kg_context* ctx, int spt, int grid, const BookItem* items, const int2* tiles,
                         int n_tiles, int* counter, Moments* parts, uint32_t* flags) {
  const int32_t* cancel = cancel_word(ctx);
  const dim3 threads(BLOCK_SAMPLES / spt);
  switch (spt) {
    case 1:
      hipLaunchKernelGGL((book_kernel<FP32, 1, CLS>), dim3(grid), threads, 0, ctx->stream, items, tiles, n_tiles,
                         counter, parts, flags, cancel);
      break;
    case 2:
      hipLaunchKernelGGL((book_kernel<FP32, 2, CLS>), dim3(grid), threads, 0, ctx->stream, items, tiles, n_tiles,
                         counter, parts, flags, cancel);
      break;
    default:
      hipLaunchKernelGGL((book_kernel<FP32, 4, CLS>), dim3(grid), threads, 0, ctx->stream, items, tiles, n_tiles,
                         counter, parts, flags, cancel);
      break;
  }
  KG_HIP(hipGetLastError());
}

// Launches the persistent kernel of class `cls` over `n_tiles` tiles,
// pulling them through counters[cls].
template <bool FP32>
static void launch_book(kg_context* ctx, int cls, int spt, int wpc, const BookItem* items, const int2* tiles,
                        int n_tiles, int* counters, Moments* parts, uint32_t* flags) {
  const int grid = std::max(1, std::min(n_tiles, cu_count(ctx) * wpc));
  int* counter = counters + cls;
  switch (cls) {
    case BC_BARRIER:
      launch_class<FP32, BC_BARRIER>(ctx, spt, grid, items, tiles, n_tiles, counter, parts, flags);
      break;
    case BC_DIGITAL:
      launch_class<FP32, BC_DIGITAL>(ctx, spt, grid, items, tiles, n_tiles, counter, parts, flags);
      break;
    case BC_GEO_ASIAN:
      launch_class<FP32, BC_GEO_ASIAN>(ctx, spt, grid, items, tiles, n_tiles, counter, parts, flags);
      break;
    case BC_ARITH_ASIAN:
      launch_class<FP32, BC_ARITH_ASIAN>(ctx, spt, grid, items, tiles, n_tiles, counter, parts, flags);
      break;
    case BC_AUTOCALL_STEP:
      launch_class<FP32, BC_AUTOCALL_STEP>(ctx, spt, grid, items, tiles, n_tiles, counter, parts, flags);
      break;
    case BC_AUTOCALL_BRIDGE:
      launch_class<FP32, BC_AUTOCALL_BRIDGE>(ctx, spt, grid, items, tiles, n_tiles, counter, parts, flags);
      break;
    default:
      launch_class<FP32, BC_ACCUM>(ctx, spt, grid, items, tiles, n_tiles, counter, parts, flags);
      break;
  }
}

// One book execution: the counter reset, then each class's persistent
// kernel over its run of tiles, in class order.
static void run_book_once(kg_context* ctx, bool fp32, const ClassRun* runs, const BookItem* items,
                          const int2* tiles, int* counters, Moments* parts, uint32_t* flags) {
  hipLaunchKernelGGL(book_reset, dim3(1), dim3(64), 0, ctx->stream, counters, flags);
  KG_HIP(hipGetLastError());
  for (int c = 0; c < BC_COUNT; ++c) {
    const ClassRun& r = runs[c];
    if (r.count == 0) continue;
    if (fp32)
      launch_book<true>(ctx, c, r.spt, r.wpc, items, tiles + r.begin, r.count, counters, parts, flags);
    else
      launch_book<false>(ctx, c, r.spt, r.wpc, items, tiles + r.begin, r.count, counters, parts, flags);
  }
}

// Tile shape of a class when it has too few tiles to time: one wave per
// workgroup, eight workgroups per CU.
constexpr int DEFAULT_SPT = 4, DEFAULT_WPC = 8;
constexpr size_t TUNE_MIN_TILES = 4096;

// Times each tile shape (samples per thread x workgroups per CU) on a strided
// sample of one class's tiles and keeps the fastest for that class and
// precision. Classes with fewer than TUNE_MIN_TILES tiles run the default
// shape and leave the choice open for a larger book.
static void autotune(kg_context* ctx, bool fp32, int cls, const BookItem* items, const int2* tiles, size_t n_tiles,
                     int* counters, Moments* parts, uint32_t* flags, int* spt_out, int* wpc_out) {
  const int slot = fp32 ? 1 : 0;
  if (ctx->tuned_spt[slot][cls]) {
    *spt_out = ctx->tuned_spt[slot][cls];
    *wpc_out = ctx->tuned_wpc[slot][cls];
    return;
  }
  *spt_out = DEFAULT_SPT;
  *wpc_out = DEFAULT_WPC;
  if (n_tiles < TUNE_MIN_TILES) return;
  std::vector<int2> sample;
  const size_t stride = n_tiles >= 8 * TUNE_MIN_TILES ? 8 : (n_tiles >= 2 * TUNE_MIN_TILES ? 2 : 1);
  for (size_t i = 0; i < n_tiles; i += stride) sample.push_back(tiles[i]);
  int2* d_sample = (int2*)ctx->scratch.reserve(sample.size() * sizeof(int2));
  KG_HIP(hipMemcpyAsync(d_sample, sample.data(), sample.size() * sizeof(int2), hipMemcpyHostToDevice,
                        ctx->stream));
  ClassRun runs[BC_COUNT];
  runs[cls].count = (int)sample.size();
  double best = 1e300;
  for (int spt : {1, 2, 4})
    for (int wpc : {2, 4, 8, 16}) {
      check_cancel(ctx);
      runs[cls].spt = spt;
      runs[cls].wpc = wpc;
      run_book_once(ctx, fp32, runs, items, d_sample, counters, parts, flags);
      timing_begin(ctx);
      run_book_once(ctx, fp32, runs, items, d_sample, counters, parts, flags);
      const double ms = timing_end(ctx);
      if (ms < best) {
        best = ms;
        *spt_out = spt;
        *wpc_out = wpc;
      }
    }
  check_flags(ctx);
  ctx->tuned_spt[slot][cls] = *spt_out;
  ctx->tuned_wpc[slot][cls] = *wpc_out;
}

// Every table a kernel dereferences must point into the uploaded blob.
static void check_tables(const BookItem& it) {
  bool ok = !it.sobol || (it.dirs && it.shifts);
  switch (it.kind) {
    case IK_BARRIER:
      ok = ok && it.u.b.steps && it.u.b.steps32;
      break;
    case IK_ACCUM:
      ok = ok && it.u.acc.steps && it.u.acc.steps32 && it.u.acc.flags;
      break;
    case IK_AUTOCALL: {
      const AcPlanDev& p = it.u.ac;
      ok = ok && p.steps && p.step_event && p.step_observe && p.events && p.intervals && p.weight &&
           p.fill && p.steps32 && p.wf32 && p.events32 && p.intervals32;
      break;
    }
    default:
      break;
  }
  if (!ok) fail(KG_COMPUTATION_FAILURE, "a contract table was not staged on the device");
}

struct BookRun {
  std::vector<Planned> plans;
  std::vector<int> sampled;  // indices into plans with sampling work
};

static void price_spot_greeks(kg_context* ctx, const kg_settings* settings, Planned p, Blob& blob,
                              kg_spot_greeks* greeks, kg_estimate* out);

static void price_book(kg_context* ctx, const kg_book_contract* contracts, size_t n,
                       int32_t precision, kg_book_tuning* tuning, kg_estimate* out) {
  if (n == 0) return;
  if (!contracts || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
  if (precision != KG_FP64 && precision != KG_FP32_STEP) fail(KG_INVALID_INPUT, "unknown precision");
  if (n > (size_t(1) << 20)) fail(KG_INVALID_INPUT, "a book holds at most 1,048,576 contracts");
  const bool fp32 = precision == KG_FP32_STEP;
  Blob blob;
  std::vector<Planned> plans(n);
  std::vector<std::vector<uint64_t>> sob_dirs(n), sob_shifts(n);
  std::vector<size_t> sob_off(n, 0), sob_shift_off(n, 0);
  double path_steps = 0.0;
  int64_t total_blocks = 0;
  for (size_t i = 0; i < n; ++i) {
    const kg_book_contract& c = contracts[i];
    if (!c.terms) fail(KG_INVALID_INPUT, "contract terms must not be null");
    validate_settings(&c.settings, 1);
    switch (c.kind) {
      case KG_BOOK_BARRIER:
        plans[i] = plan_barrier(contract_of(*(const kg_barrier_terms*)c.terms), c.market, blob);
        break;
      case KG_BOOK_DOUBLE_BARRIER:
        plans[i] = plan_barrier(contract_of(*(const kg_double_barrier_terms*)c.terms), c.market, blob);
        break;
      case KG_BOOK_DIGITAL:
        plans[i] = plan_digital(*(const kg_digital_terms*)c.terms, c.market, blob);
        break;
      case KG_BOOK_ASIAN:
        plans[i] = plan_asian(*(const kg_asian_terms*)c.terms, c.market, blob);
        break;
      case KG_BOOK_AUTOCALL:
        plans[i] = plan_autocall(*(const kg_autocall_terms*)c.terms, c.market, blob);
        break;
      case KG_BOOK_ACCUMULATOR:
        plans[i] = plan_accumulator(*(const kg_accumulator_terms*)c.terms, c.market, blob);
        break;
      default:
        fail(KG_INVALID_INPUT, "unknown contract kind");
    }
    Planned& p = plans[i];
    if (p.known) continue;
    // Continuously monitored double barriers need the survival series, which
    // the persistent kernel leaves out; they run on the spot-Greek kernel
    // after the book, as kg_double_barrier_price runs them.
    if (p.item.kind == IK_BARRIER && p.item.u.b.has_upper && p.item.u.b.continuous) {
      p.separate = true;
      continue;
    }
    p.sampler = validate_settings(&c.settings, (uint64_t)p.item.dims);
    const Sampler& s = p.sampler;
    p.item.samples = s.samples();
    p.item.paths = s.path_count;
    p.item.seed = s.seed;
    p.item.antithetic = s.antithetic;
    p.item.sobol = s.sobol;
    p.item.first_block = total_blocks;
    total_blocks += s.blocks();
    path_steps += (double)s.path_count * (double)p.item.dims;
    if (s.sobol) {
      sobol_tables(ctx, (uint32_t)p.item.dims, s.seed, sob_dirs[i], sob_shifts[i]);
      sob_off[i] = blob.add(sob_dirs[i]);
      sob_shift_off[i] = blob.add(sob_shifts[i]);
    }
  }
  std::vector<int> sampled;
  for (size_t i = 0; i < n; ++i)
    if (!plans[i].known && !plans[i].separate) sampled.push_back((int)i);
  std::vector<Moments> results(2 * sampled.size());
  double device_ms = 0.0;
  if (!sampled.empty()) {
    if (total_blocks > (int64_t(1) << 31) - 1) fail(KG_INVALID_INPUT, "the book has too many sample blocks");
    unsigned char* d_blob = (unsigned char*)ctx->book_blob.reserve(blob.bytes.size());
    std::vector<BookItem> items;
    for (int i : sampled) {
      Planned& p = plans[i];
      for (auto& f : p.fixups) {
        const void* at = d_blob + f.first;
        std::memcpy((unsigned char*)&p.item + f.second, &at, sizeof(at));
      }
      if (p.item.sobol) {
        p.item.dirs = (const uint64_t*)(d_blob + sob_off[i]);
        p.item.shifts = (const uint64_t*)(d_blob + sob_shift_off[i]);
      }
      check_tables(p.item);
      items.push_back(p.item);
    }
    // Tiles grouped by path class; within a class, most expensive contracts
    // first (LPT order for the class's queue).
    std::vector<int2> tiles;
    tiles.reserve((size_t)total_blocks);
    for (size_t k = 0; k < items.size(); ++k) {
      const int64_t blocks = (items[k].samples + BLOCK_SAMPLES - 1) / BLOCK_SAMPLES;
      for (int64_t b = 0; b < blocks; ++b) tiles.push_back(make_int2((int)k, (int)b));
    }
    std::stable_sort(tiles.begin(), tiles.end(), [&](const int2& x, const int2& y) {
      const int cx = book_class(items[x.x]), cy = book_class(items[y.x]);
      if (cx != cy) return cx < cy;
      return items[x.x].dims > items[y.x].dims;
    });
    ClassRun runs[BC_COUNT];
    for (size_t t = 0; t < tiles.size(); ++t) {
      ClassRun& r = runs[book_class(items[tiles[t].x])];
      if (r.count++ == 0) r.begin = (int)t;
    }
    BookItem* d_items = (BookItem*)ctx->book_items.reserve(items.size() * sizeof(BookItem));
    int2* d_tiles = (int2*)ctx->book_tiles.reserve(tiles.size() * sizeof(int2));
    Moments* d_parts = (Moments*)ctx->book_parts.reserve((size_t)total_blocks * 2 * sizeof(Moments));
    Moments* d_out = (Moments*)ctx->book_out.reserve(items.size() * 2 * sizeof(Moments));
    int* d_counters = (int*)ctx->book_counter.reserve(256);
    uint32_t* flags = (uint32_t*)ctx->flags.reserve(256);
    if (!blob.bytes.empty())
      KG_HIP(hipMemcpyAsync(d_blob, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice, ctx->stream));
    KG_HIP(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(BookItem), hipMemcpyHostToDevice,
                          ctx->stream));
    KG_HIP(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice,
                          ctx->stream));
    const int spt = tuning ? tuning->samples_per_thread : 0;
    const int wpc = tuning ? tuning->workgroups_per_cu : 0;
    if (spt != 0 && spt != 1 && spt != 2 && spt != 4) fail(KG_INVALID_INPUT, "samples_per_thread must be 1, 2 or 4");
    if (wpc < 0 || wpc > 16) fail(KG_INVALID_INPUT, "workgroups_per_cu must be in 0..16");
    // The shape reported back is that of the class with the most path-steps.
    int major = -1;
    double major_steps = -1.0;
    for (int c = 0; c < BC_COUNT; ++c) {
      ClassRun& r = runs[c];
      if (r.count == 0) continue;
      int tuned_spt = spt, tuned_wpc = wpc;
      if (!spt || !wpc)
        autotune(ctx, fp32, c, d_items, tiles.data() + r.begin, (size_t)r.count, d_counters, d_parts, flags,
                 &tuned_spt, &tuned_wpc);
      r.spt = spt ? spt : tuned_spt;
      r.wpc = wpc ? wpc : tuned_wpc;
      double steps = 0.0;
      for (int t = r.begin; t < r.begin + r.count; ++t) steps += (double)items[tiles[t].x].dims;
      if (steps > major_steps) {
        major_steps = steps;
        major = c;
      }
    }
    const bool use_graph = tuning && tuning->use_graph;
    const int replays = tuning ? std::max(0, tuning->replays) : 0;
    if (use_graph) {
      uint64_t key[BOOK_GRAPH_KEY];
      size_t nk = 0;
      for (uint64_t v : {(uint64_t)fp32, (uint64_t)items.size(), (uint64_t)tiles.size(),
                         (uint64_t)(uintptr_t)d_items, (uint64_t)(uintptr_t)d_tiles, (uint64_t)(uintptr_t)d_counters,
                         (uint64_t)(uintptr_t)d_parts, (uint64_t)(uintptr_t)d_out, (uint64_t)(uintptr_t)flags,
                         (uint64_t)(uintptr_t)cancel_word(ctx)})
        key[nk++] = v;
      for (int c = 0; c < BC_COUNT; ++c)
        for (int v : {runs[c].begin, runs[c].count, runs[c].spt, runs[c].wpc}) key[nk++] = (uint64_t)v;
      static_assert(10 + 4 * BC_COUNT == BOOK_GRAPH_KEY, "graph key size");
      if (!ctx->book_graph || std::memcmp(key, ctx->book_graph_key, sizeof(key)) != 0) {
        if (ctx->book_graph) (void)hipGraphExecDestroy(ctx->book_graph);
        ctx->book_graph = nullptr;
        // Destroys the captured graph on every exit; only an instantiated
        // graph is kept.
        struct GraphGuard {
          hipGraph_t graph = nullptr;
          ~GraphGuard() {
            if (graph) (void)hipGraphDestroy(graph);
          }
        } guard;
        KG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        run_book_once(ctx, fp32, runs, d_items, d_tiles, d_counters, d_parts, flags);
        hipLaunchKernelGGL(merge_book, dim3(items.size()), dim3(MERGE_THREADS), 0, ctx->stream, d_items, d_parts,
                           d_out);
        KG_HIP(hipStreamEndCapture(ctx->stream, &guard.graph));
        hipGraphExec_t made = nullptr;
        KG_HIP(hipGraphInstantiate(&made, guard.graph, nullptr, nullptr, 0));
        ctx->book_graph = made;
        std::memcpy(ctx->book_graph_key, key, sizeof(key));
      }
      hipGraphExec_t exec = ctx->book_graph;
      timing_begin(ctx);
      KG_HIP(hipGraphLaunch(exec, ctx->stream));
      device_ms = timing_end(ctx);
      check_flags(ctx);  // a cancelled first run ends the call before its replays
      if (replays) {
        timing_begin(ctx);
        for (int k = 0; k < replays; ++k) KG_HIP(hipGraphLaunch(exec, ctx->stream));
        device_ms = timing_end(ctx) / replays;
      }
    } else {
      for (int k = 0; k <= replays; ++k) {
        if (k) check_flags(ctx);
        timing_begin(ctx);
        run_book_once(ctx, fp32, runs, d_items, d_tiles, d_counters, d_parts, flags);
        hipLaunchKernelGGL(merge_book, dim3(items.size()), dim3(MERGE_THREADS), 0, ctx->stream, d_items, d_parts, d_out);
        KG_HIP(hipGetLastError());
        device_ms = timing_end(ctx);
      }
    }
    check_flags(ctx);
    KG_HIP(hipMemcpyAsync(results.data(), d_out, results.size() * sizeof(Moments), hipMemcpyDeviceToHost,
                          ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
    if (tuning) {
      tuning->samples_per_thread = runs[major].spt;
      tuning->workgroups_per_cu = runs[major].wpc;
    }
    // Arithmetic Asians whose controlled price is negative are resampled
    // without the control on the same seed (asian.rs:744-752).
    std::vector<kg_book_contract> redo;
    std::vector<int> redo_index;
    for (size_t k = 0; k < sampled.size(); ++k) {
      Planned& p = plans[sampled[k]];
      if (p.item.kind == IK_ARITH_ASIAN && p.item.u.a.control) {
        const double mean = scaled_value(results[2 * k].scale, results[2 * k].mean, 0.0);
        if (mean + p.offset < 0.0) redo_index.push_back((int)k);
      }
    }
    for (int k : redo_index) {
      Planned& p = plans[sampled[k]];
      BookItem it = items[k];
      it.u.a.control = 0;
      it.first_block = 0;
      const int64_t blocks = (it.samples + BLOCK_SAMPLES - 1) / BLOCK_SAMPLES;
      std::vector<int2> t2;
      for (int64_t b = 0; b < blocks; ++b) t2.push_back(make_int2(0, (int)b));
      KG_HIP(hipMemcpyAsync(d_items, &it, sizeof(BookItem), hipMemcpyHostToDevice, ctx->stream));
      KG_HIP(hipMemcpyAsync(d_tiles, t2.data(), t2.size() * sizeof(int2), hipMemcpyHostToDevice, ctx->stream));
      ClassRun one[BC_COUNT];
      one[BC_ARITH_ASIAN] = runs[BC_ARITH_ASIAN];
      one[BC_ARITH_ASIAN].begin = 0;
      one[BC_ARITH_ASIAN].count = (int)t2.size();
      run_book_once(ctx, fp32, one, d_items, d_tiles, d_counters, d_parts, flags);
      hipLaunchKernelGGL(merge_book, dim3(1), dim3(MERGE_THREADS), 0, ctx->stream, d_items, d_parts, d_out);
      check_flags(ctx);
      KG_HIP(hipMemcpyAsync(&results[2 * k], d_out, 2 * sizeof(Moments), hipMemcpyDeviceToHost, ctx->stream));
      KG_HIP(hipStreamSynchronize(ctx->stream));
      p.offset = 0.0;
    }
  }
  // Estimates in input order.
  size_t k = 0;
  for (size_t i = 0; i < n; ++i) {
    Planned& p = plans[i];
    const kg_settings& cs = contracts[i].settings;
    out[i] = kg_estimate{};
    if (p.known) {
      fill_known(p.value, cs.sampling == KG_SOBOL, &out[i]);
      continue;
    }
    if (p.separate) {
      kg_settings own = cs;
      own.precision = precision;
      Blob own_blob;
      kg_spot_greeks none{};
      price_spot_greeks(ctx, &own,
                        plan_barrier(contract_of(*(const kg_double_barrier_terms*)contracts[i].terms),
                                     contracts[i].market, own_blob),
                        own_blob, &none, &out[i]);
      continue;
    }
    const Moments& m = results[2 * k];
    const Moments& pm = results[2 * k + 1];
    ++k;
    fill_estimate(p.sampler, m, pm, 0.0, false, true, &out[i]);
    if (p.kind == Estimator::PathMeans && p.offset != 0.0) {
      out[i].price += p.offset;
      if (out[i].has_standard_error) {
        const double hw = 1.959963984540054 * out[i].standard_error;
        out[i].ci_lower = out[i].price - hw;
        out[i].ci_upper = out[i].price + hw;
      }
    }
    out[i].path_steps = (double)p.sampler.path_count * (double)p.item.dims;
    out[i].device_ms = device_ms;
  }
  if (tuning) {
    tuning->device_ms = device_ms;
    tuning->path_steps = path_steps;
  }
}

template <class Terms>
static void price_single(kg_context* ctx, int kind, const kg_settings* settings, const Terms* option,
                         const kg_bsm_market* market, kg_estimate* out) {
  if (!settings || !option || !market || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
  kg_book_contract c{};
  c.kind = kind;
  c.settings = *settings;
  c.market = *market;
  c.terms = option;
  kg_book_tuning t{};
  t.samples_per_thread = 1;
  t.workgroups_per_cu = 4;
  price_book(ctx, &c, 1, settings->precision, &t, out);
}

// Prices one barrier or digital contract with its likelihood-ratio delta and
// gamma: plan as in the book, then one launch of spot_greek_kernel and a
// block-ordered merge of its four partial sets.
static void price_spot_greeks(kg_context* ctx, const kg_settings* settings, Planned p, Blob& blob,
                              kg_spot_greeks* greeks, kg_estimate* out) {
  *out = kg_estimate{};
  greeks->available = 0;
  greeks->delta = greeks->gamma = 0.0;
  if (p.known) {
    fill_known(p.value, settings->sampling == KG_SOBOL, out);
    greeks->available = p.greeks;
    greeks->delta = p.known_delta;
    greeks->gamma = p.known_gamma;
    return;
  }
  const Sampler s = validate_settings(settings, (uint64_t)p.item.dims);
  p.item.samples = s.samples();
  p.item.paths = s.path_count;
  p.item.seed = s.seed;
  p.item.antithetic = s.antithetic;
  p.item.sobol = s.sobol;
  p.item.first_block = 0;
  size_t dirs_off = 0, shifts_off = 0;
  if (s.sobol) {
    std::vector<uint64_t> dirs, shifts;
    sobol_tables(ctx, (uint32_t)p.item.dims, s.seed, dirs, shifts);
    dirs_off = blob.add(dirs);
    shifts_off = blob.add(shifts);
  }
  const size_t item_off = (blob.bytes.size() + 255) & ~size_t(255);
  blob.bytes.resize(item_off + sizeof(BookItem));
  unsigned char* d = (unsigned char*)ctx->scratch.reserve(blob.bytes.size());
  for (auto& f : p.fixups) {
    const void* at = d + f.first;
    std::memcpy((unsigned char*)&p.item + f.second, &at, sizeof(at));
  }
  if (s.sobol) {
    p.item.dirs = (const uint64_t*)(d + dirs_off);
    p.item.shifts = (const uint64_t*)(d + shifts_off);
  }
  check_tables(p.item);
  std::memcpy(blob.bytes.data() + item_off, &p.item, sizeof(BookItem));
  upload_async(ctx, d, blob.bytes.data(), blob.bytes.size());
  const BookItem* d_item = (const BookItem*)(d + item_off);
  const int64_t blocks = s.blocks();
  if (blocks > (int64_t(1) << 31) - 1) fail(KG_INVALID_INPUT, "too many sample blocks");
  Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * 4 * blocks);
  uint32_t* flags = reset_flags(ctx);
  timing_begin(ctx);
  auto launch = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(BLOCK_SAMPLES), 0, ctx->stream, d_item, parts, flags);
  };
  if (s.sobol && s.fp32)
    launch(spot_greek_kernel<true, true>);
  else if (s.sobol)
    launch(spot_greek_kernel<false, true>);
  else if (s.fp32)
    launch(spot_greek_kernel<true, false>);
  else
    launch(spot_greek_kernel<false, false>);
  KG_HIP(hipGetLastError());
  Moments r[4];
  finish_sampling(ctx, parts, blocks, 4, r, &out->device_ms);
  const bool control = p.item.kind == IK_ARITH_ASIAN && p.item.u.a.control;
  if (control && scaled_value(r[0].scale, r[0].mean, 0.0) + p.offset < 0.0) {
    // A controlled price below zero is resampled without the control on the
    // same seed (asian.rs:618-630).
    p.item.u.a.control = 0;
    p.offset = p.offset_delta = p.offset_gamma = 0.0;
    kg_settings same = *settings;
    same.seed = s.seed;
    same.has_seed = 1;
    Blob fresh = blob;
    fresh.bytes.resize(item_off);
    price_spot_greeks(ctx, &same, std::move(p), fresh, greeks, out);
    return;
  }
  fill_estimate(s, r[0], r[3], 0.0, false, true, out);
  out->path_steps = (double)s.path_count * (double)p.item.dims;
  if (p.offset != 0.0) {  // the control's exact mean, or the analytic hit value
    out->price += p.offset;
    if (out->has_standard_error) {
      const double hw = 1.959963984540054 * out->standard_error;
      out->ci_lower = out->price - hw;
      out->ci_upper = out->price + hw;
    }
  }
  greeks->available = p.greeks;
  if (p.greeks & KG_GREEK_DELTA) greeks->delta = scaled_value(r[1].scale, r[1].mean, 0.0) + p.offset_delta;
  if (p.greeks & KG_GREEK_GAMMA) greeks->gamma = scaled_value(r[2].scale, r[2].mean, 0.0) + p.offset_gamma;
  if (!std::isfinite(greeks->delta) || !std::isfinite(greeks->gamma)) unstable();
}

// MonteCarloAsianEngine::price_strikes (asian.rs:197-345) for an arithmetic
// average: the distinct strikes (quote_distinct_strikes, engines.rs:38-75)
// are planned one by one; those that sample share each path as lanes of
// asian_strikes_kernel, STRIKE_LANES per launch. A lane whose controlled
// price is negative is priced alone, as the reference does, which resamples
// it without the control (asian.rs:744-752).
static void asian_price_strikes(kg_context* ctx, const kg_settings* settings, const kg_asian_terms* option,
                                const kg_bsm_market* market, const double* strikes, size_t n, double* prices) {
  if (!settings || !option || !market) fail(KG_INVALID_INPUT, "arguments must not be null");
  if (n == 0) return;
  if (!strikes || !prices) fail(KG_INVALID_INPUT, "arguments must not be null");
  if (n > (size_t(1) << 30) / 24) fail(KG_INVALID_INPUT, "Monte Carlo strike batch workspace exceeds 1 GiB");
  if (option->averaging != KG_AVERAGE_ARITHMETIC)
    fail(KG_UNSUPPORTED_OPERATION, "strike batches price arithmetic averages");
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(strikes[i]) || !(strikes[i] > 0.0)) fail(KG_INVALID_INPUT, "strikes must be positive and finite");
  std::vector<size_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return strikes[a] < strikes[b]; });
  std::vector<double> unique;
  std::vector<size_t> slot(n);
  for (size_t k = 0; k < n; ++k) {
    const double v = strikes[order[k]];
    if (unique.empty() || unique.back() != v) unique.push_back(v);
    slot[order[k]] = unique.size() - 1;
  }
  std::vector<double> value(unique.size());
  std::vector<Planned> plans(unique.size());
  std::vector<size_t> sampled;
  Blob blob;  // every plan's tables (arithmetic plans stage none)
  for (size_t u = 0; u < unique.size(); ++u) {
    kg_asian_terms t = *option;
    t.strike = unique[u];
    plans[u] = plan_asian(t, *market, blob);
    if (plans[u].known)
      value[u] = plans[u].value;
    else
      sampled.push_back(u);
  }
  if (!sampled.empty()) {
    const Planned& p0 = plans[sampled[0]];
    const ArithAsianDev& a0 = p0.item.u.a;
    for (size_t u : sampled) {
      const ArithAsianDev& a = plans[u].item.u.a;
      if (plans[u].item.kind != IK_ARITH_ASIAN || a.n != a0.n || a.spot != a0.spot || a.weight != a0.weight ||
          a.log_discount != a0.log_discount || a.drift != a0.drift || a.width != a0.width || a.sign != a0.sign ||
          a.interval_forward != a0.interval_forward || a.lead != a0.lead || a.lead_drift != a0.lead_drift ||
          a.lead_width != a0.lead_width)
        fail(KG_COMPUTATION_FAILURE, "strike lanes do not share their path grid");
    }
    const Sampler s = validate_settings(settings, (uint64_t)p0.item.dims);
    BookItem item = p0.item;
    item.samples = s.samples();
    item.paths = s.path_count;
    item.seed = s.seed;
    item.antithetic = s.antithetic;
    item.sobol = s.sobol;
    item.first_block = 0;
    size_t dirs_off = 0, shifts_off = 0;
    if (s.sobol) {
      std::vector<uint64_t> dirs, shifts;
      sobol_tables(ctx, (uint32_t)item.dims, s.seed, dirs, shifts);
      dirs_off = blob.add(dirs);
      shifts_off = blob.add(shifts);
    }
    std::vector<ArithAsianDev> lanes;
    for (size_t u : sampled) lanes.push_back(plans[u].item.u.a);
    const size_t lanes_off = blob.add(lanes);
    const size_t item_off = (blob.bytes.size() + 255) & ~size_t(255);
    blob.bytes.resize(item_off + sizeof(BookItem));
    unsigned char* d = (unsigned char*)ctx->scratch.reserve(blob.bytes.size());
    if (s.sobol) {
      item.dirs = (const uint64_t*)(d + dirs_off);
      item.shifts = (const uint64_t*)(d + shifts_off);
    }
    check_tables(item);
    std::memcpy(blob.bytes.data() + item_off, &item, sizeof(BookItem));
    upload_async(ctx, d, blob.bytes.data(), blob.bytes.size());
    const BookItem* d_item = (const BookItem*)(d + item_off);
    const ArithAsianDev* d_lanes = (const ArithAsianDev*)(d + lanes_off);
    const int64_t blocks = s.blocks();
    if (blocks > (int64_t(1) << 31) - 1) fail(KG_INVALID_INPUT, "too many sample blocks");
    Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * 2 * STRIKE_LANES * blocks);
    std::vector<size_t> redo;
    for (size_t g = 0; g < sampled.size(); g += STRIKE_LANES) {
      const int m = (int)std::min<size_t>(STRIKE_LANES, sampled.size() - g);
      uint32_t* flags = reset_flags(ctx);
      timing_begin(ctx);
      auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(BLOCK_SAMPLES), 0, ctx->stream, d_item, d_lanes + g,
                           m, parts, flags);
      };
      if (s.sobol && s.fp32)
        launch(asian_strikes_kernel<true, true>);
      else if (s.sobol)
        launch(asian_strikes_kernel<false, true>);
      else if (s.fp32)
        launch(asian_strikes_kernel<true, false>);
      else
        launch(asian_strikes_kernel<false, false>);
      KG_HIP(hipGetLastError());
      Moments r[2 * STRIKE_LANES];
      double ms = 0.0;
      finish_sampling(ctx, parts, blocks, 2 * m, r, &ms);
      for (int l = 0; l < m; ++l) {
        const size_t u = sampled[g + l];
        value[u] = scaled_value(r[2 * l].scale, r[2 * l].mean, 0.0) + plans[u].offset;
        if (plans[u].item.u.a.control && value[u] < 0.0) redo.push_back(u);
      }
    }
    for (size_t u : redo) {
      kg_asian_terms t = *option;
      t.strike = unique[u];
      kg_estimate e{};
      price_single(ctx, KG_BOOK_ASIAN, settings, &t, market, &e);
      value[u] = e.price;
    }
  }
  for (size_t i = 0; i < n; ++i) {
    if (!std::isfinite(value[slot[i]])) unstable();
    prices[i] = value[slot[i]];
  }
}

template <class Terms>
static void barrier_with_greeks(kg_context* ctx, const kg_settings* settings, const Terms* option,
                                const kg_bsm_market* market, kg_spot_greeks* greeks, kg_estimate* out) {
  if (!settings || !option || !market || !greeks || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
  Blob blob;
  Planned p = plan_barrier(contract_of(*option), *market, blob, greeks->request, greeks->spot_radius);
  price_spot_greeks(ctx, settings, std::move(p), blob, greeks, out);
}

}  // namespace kg

using namespace kg;

extern "C" {

kg_status kg_barrier_price(kg_context* ctx, const kg_settings* settings, const kg_barrier_terms* option,
                           const kg_bsm_market* market, kg_estimate* out) {
  return guarded(ctx, [&] { price_single(ctx, KG_BOOK_BARRIER, settings, option, market, out); });
}

kg_status kg_barrier_price_with_greeks(kg_context* ctx, const kg_settings* settings,
                                       const kg_barrier_terms* option, const kg_bsm_market* market,
                                       kg_spot_greeks* greeks, kg_estimate* out) {
  return guarded(ctx, [&] { barrier_with_greeks(ctx, settings, option, market, greeks, out); });
}

kg_status kg_double_barrier_price(kg_context* ctx, const kg_settings* settings,
                                  const kg_double_barrier_terms* option, const kg_bsm_market* market,
                                  kg_estimate* out) {
  return guarded(ctx, [&] { price_single(ctx, KG_BOOK_DOUBLE_BARRIER, settings, option, market, out); });
}

kg_status kg_double_barrier_price_with_greeks(kg_context* ctx, const kg_settings* settings,
                                              const kg_double_barrier_terms* option,
                                              const kg_bsm_market* market, kg_spot_greeks* greeks,
                                              kg_estimate* out) {
  return guarded(ctx, [&] { barrier_with_greeks(ctx, settings, option, market, greeks, out); });
}

kg_status kg_digital_price(kg_context* ctx, const kg_settings* settings, const kg_digital_terms* option,
                           const kg_bsm_market* market, kg_estimate* out) {
  return guarded(ctx, [&] { price_single(ctx, KG_BOOK_DIGITAL, settings, option, market, out); });
}

kg_status kg_digital_price_with_greeks(kg_context* ctx, const kg_settings* settings,
                                       const kg_digital_terms* option, const kg_bsm_market* market,
                                       kg_spot_greeks* greeks, kg_estimate* out) {
  return guarded(ctx, [&] {
    if (!settings || !option || !market || !greeks || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
    Blob blob;
    Planned p = plan_digital(*option, *market, blob, greeks->request);
    price_spot_greeks(ctx, settings, std::move(p), blob, greeks, out);
  });
}

kg_status kg_asian_price_strikes(kg_context* ctx, const kg_settings* settings, const kg_asian_terms* option,
                                 const kg_bsm_market* market, const double* strikes, size_t n, double* prices) {
  return guarded(ctx, [&] { asian_price_strikes(ctx, settings, option, market, strikes, n, prices); });
}

kg_status kg_asian_price(kg_context* ctx, const kg_settings* settings, const kg_asian_terms* option,
                         const kg_bsm_market* market, kg_estimate* out) {
  return guarded(ctx, [&] { price_single(ctx, KG_BOOK_ASIAN, settings, option, market, out); });
}

kg_status kg_asian_price_with_greeks(kg_context* ctx, const kg_settings* settings, const kg_asian_terms* option,
                                     const kg_bsm_market* market, kg_spot_greeks* greeks, kg_estimate* out) {
  return guarded(ctx, [&] {
    if (!settings || !option || !market || !greeks || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
    Blob blob;
    Planned p = plan_asian(*option, *market, blob, greeks->request);
    price_spot_greeks(ctx, settings, std::move(p), blob, greeks, out);
  });
}

kg_status kg_accumulator_price(kg_context* ctx, const kg_settings* settings,
                               const kg_accumulator_terms* option, const kg_bsm_market* market,
                               kg_estimate* out) {
  return guarded(ctx, [&] { price_single(ctx, KG_BOOK_ACCUMULATOR, settings, option, market, out); });
}

kg_status kg_price_book(kg_context* ctx, const kg_book_contract* contracts, size_t n, int32_t precision,
                        kg_book_tuning* tuning, kg_estimate* out) {
  return guarded(ctx, [&] { price_book(ctx, contracts, n, precision, tuning, out); });
}
}
