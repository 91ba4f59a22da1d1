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
#include <unordered_map>

#include "autocall_common.hpp"

namespace kg {

// The draw source of sample j of `block`, from a prototype that holds the
// Sobol tables (if any): the lane's Gray-code digits of the block's LDS
// tables, its point index for direct Sobol points, or the block's PCG stream
// advanced to normal j * dims.
template <bool SOBOL, bool FP32, class D>
__device__ inline D sample_draws(D dr, uint64_t seed, int64_t block, int j) {
  dr.point = (uint64_t)(block * BLOCK_SAMPLES + j);
  if (SOBOL) {
    if (!dr.is_direct()) {
      const uint32_t g = (uint32_t)(j ^ (j >> 1));
      dr.glo = g & 15u;
      dr.ghi = g >> 4;
    }
  } else {
    const uint64_t first = (uint64_t)j * (uint64_t)dr.dims;
    if (FP32) {
      dr.start32(seed, (uint64_t)block, first);
    } else {
      dr.s.start(seed, (uint64_t)block, first);
    }
  }
  return dr;
}

// Builds the block's Sobol LDS tables (all threads of the block take part)
// and returns the draw source of this thread's sample.
template <bool SOBOL, bool FP32, bool LDS>
__device__ inline Draws<SOBOL, FP32, LDS ? 0 : 1> setup_draws(int dims, uint64_t seed, int64_t block, int j,
                                                              const uint64_t* __restrict__ dirs,
                                                              const uint64_t* __restrict__ shifts,
                                                              unsigned char* lds) {
  Draws<SOBOL, FP32, LDS ? 0 : 1> dr;
  dr.dims = dims;
  dr.direct = !LDS;
  dr.dirs = dirs;
  dr.shifts = shifts;
  if (SOBOL && LDS) {
    const void* lo = nullptr;
    const void* hi = nullptr;
    if (FP32)
      sobol_build_lds32(dirs, shifts, dims, (uint64_t)block * BLOCK_SAMPLES, lds, &lo, &hi);
    else
      sobol_build_lds64(dirs, shifts, dims, (uint64_t)block * BLOCK_SAMPLES, lds, &lo, &hi);
    dr.lo = lo;
    dr.hi = hi;
  }
  return sample_draws<SOBOL, FP32>(dr, seed, block, j);
}

// One thread per sample, whole paths in one pass: plans without a bridge
// (step_path) and bridge plans too large for the two-pass kernels.
template <bool SOBOL, bool FP32, bool LDS>
__global__ void __launch_bounds__(256)
    autocall_kernel(AcPlanDev p, uint64_t seed, int64_t samples, int64_t paths, int antithetic,
                    const uint64_t* __restrict__ dirs, const uint64_t* __restrict__ shifts,
                    Moments* __restrict__ partials, uint32_t* __restrict__ flags) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ Moments smem[2 * 4];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const bool live = n < samples;
  auto dr = setup_draws<SOBOL, FP32, LDS>(p.n_steps, seed, block, j, dirs, shifts, lds);
  uint32_t bad = 0;
  const bool paired = antithetic && live && n < paths / 2;
  const auto origin = dr;
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

// Bridge plans in two passes (bridge_events, bridge_knock_in): the daily
// fill runs only for the legs that reach expiry still monitoring the
// knock-in, gathered per sample block. The host builds the interval rows
// (interval_row's fields) and the Sobol table rows of the intervals' first
// dimensions without a base point. The launches:
//   sobol_base_rows_kernel - (Sobol) every sample block's base point, one
//                            word per dimension; kept across calls with the
//                            same dimensions, seed and block count;
//   bridge_events_kernel   - pass one per sample and leg; writes each leg's
//                            value (the final value, or the cash before the
//                            expiry payoff of a leg that waits on its fill)
//                            and the waiting legs (sample << 1 | leg) of each
//                            wave into the wave's own queue region of 128
//                            items, with their count;
//   bridge_fill_kernel     - pass two over the queued legs of a run of
//                            sample blocks (adding the expiry payoff), then
//                            each block's moments, merged exactly as
//                            autocall_kernel merges them.
template <bool FP32>
using SobolWord = typename std::conditional<FP32, uint32_t, uint64_t>::type;
constexpr int QUEUE_WAVES = BLOCK_SAMPLES / 64;  // pass-one waves (queue regions) per sample block
constexpr int REGION = 128;  // queue items per pass-one wave: 64 samples, two legs
// Sample blocks (one wave each) per fill workgroup: as many waves as a CU
// holds at the kernel's register budget, FP32 at most 128 VGPRs (16 waves),
// FP64 at most 256 (8 waves).
template <bool FP32>
constexpr int fill_waves() {
  return FP32 ? 16 : 8;
}
// s_getreg operand of HW_ID's SIMD_ID field (bits 5:4): register 4, offset
// 4, width 2.
constexpr int HW_ID_SIMD = 4 | 4 << 6 | (2 - 1) << 11;

template <bool FP32>
__global__ void __launch_bounds__(256)
    sobol_base_rows_kernel(const uint64_t* __restrict__ dirs, const uint64_t* __restrict__ shifts, int dims,
                           int64_t blocks, SobolWord<FP32>* __restrict__ rows) {
  constexpr int SHIFT = FP32 ? 20 : 0;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= blocks * dims) return;
  const int64_t b = i / dims;
  const int d = (int)(i - b * dims);
  rows[i] = (SobolWord<FP32>)(sobol_block_base(dirs, shifts, dims, (uint64_t)b * BLOCK_SAMPLES, d) >> SHIFT);
}

template <bool SOBOL, bool FP32, bool LDS>
__global__ void __launch_bounds__(256)
    bridge_events_kernel(AcPlanDev p, uint64_t seed, int64_t samples, int64_t paths, int antithetic,
                         const uint64_t* __restrict__ dirs, const uint64_t* __restrict__ shifts,
                         double* __restrict__ values, const void* __restrict__ rows_in,
                         const void* __restrict__ table_in, const void* __restrict__ base_rows,
                         uint32_t* __restrict__ queue, uint32_t* __restrict__ counts, uint32_t* __restrict__ flags) {
  using R = typename std::conditional<FP32, float, double>::type;
  using W = SobolWord<FP32>;
  extern __shared__ __align__(16) unsigned char lds[];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const bool live = n < samples;
  const int dims = p.n_steps;
  // LDS: the interval rows, then (Sobol) the table rows of the intervals'
  // first dimensions (row k holds the first dimension of interval k) with
  // this block's base point folded into `hi`.
  IntervalRow<R>* rows = (IntervalRow<R>*)lds;
  const IntervalRow<R>* rows_g = (const IntervalRow<R>*)rows_in;
  for (int k = j; k < p.n_intervals; k += BLOCK_SAMPLES) rows[k] = rows_g[k];
  Draws<SOBOL, FP32, LDS ? 0 : 1> dr;
  dr.dims = dims;
  dr.direct = !LDS;
  dr.dirs = dirs;
  dr.shifts = shifts;
  if (SOBOL && LDS) {
    W* lo = (W*)(rows + p.n_intervals);
    W* hi = lo + (size_t)p.n_intervals * 16;
    const W* base = (const W*)base_rows + block * dims;
    const W* table = (const W*)table_in;
    for (int e = j; e < p.n_intervals * 32; e += BLOCK_SAMPLES) {
      const int k = e >> 4;  // lo rows, then hi rows
      W v = table[e];
      if (k >= p.n_intervals) v ^= base[rows_g[k - p.n_intervals].first];
      lo[e] = v;
    }
    dr.lo = lo;
    dr.hi = hi;
  }
  __syncthreads();
  dr = sample_draws<SOBOL, FP32>(dr, seed, block, j);
  const auto origin = dr;
  uint32_t bad = 0;
  const bool paired = antithetic && live && n < paths / 2;
  const BridgeEvents<R> up = bridge_events<SOBOL, FP32, SOBOL && LDS>(p, rows, dr, live, 1.0f);
  BridgeEvents<R> down{};
  const bool two_legs = antithetic && __ballot(paired) != 0;
  if (two_legs) {
    dr = origin;
    down = bridge_events<SOBOL, FP32, SOBOL && LDS>(p, rows, dr, paired, -1.0f);
  }
  const bool want_up = live && up.need_fill;
  const bool want_down = two_legs && paired && down.need_fill;
  if (live) values[n] = want_up ? up.cash : bridge_finish(p, up, false, true, bad);
  if (two_legs && paired) values[samples + n] = want_down ? down.cash : bridge_finish(p, down, false, true, bad);
  // This wave's region: its leg-0 items in sample order, then its leg-1
  // items; the count covers both.
  const uint64_t m_up = __ballot(want_up), m_down = __ballot(want_down);
  const int lane = j & 63;
  const uint64_t lower = (1ull << lane) - 1ull;
  uint32_t* region = queue + (size_t)(block * QUEUE_WAVES + (j >> 6)) * REGION;
  if (want_up) region[__popcll(m_up & lower)] = (uint32_t)(n << 1);
  if (want_down) region[__popcll(m_up) + __popcll(m_down & lower)] = (uint32_t)(n << 1 | 1);
  if (lane == 0) counts[block * QUEUE_WAVES + (j >> 6)] = (uint32_t)(__popcll(m_up) + __popcll(m_down));
  if (bad) raise_flag(flags, FLAG_NONFINITE);
}

// One butterfly level of wave_merge on four sets at once (lane l holds the
// values of lanes l of four waves), so the four merge chains interleave.
__device__ inline void wave_merge4(Moments (&m)[4], bool full) {
  const int lane = threadIdx.x & (WAVE - 1);
  for (int mask = 1; mask < WAVE; mask <<= 1) {
    Moments o[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) o[w] = moments_shfl_xor(m[w], mask);
    if (full) {
#pragma unroll
      for (int w = 0; w < 4; ++w) m[w] = (lane & mask) ? moments_merge_full(o[w], m[w]) : moments_merge_full(m[w], o[w]);
    } else {
#pragma unroll
      for (int w = 0; w < 4; ++w) m[w] = (lane & mask) ? moments_merge(o[w], m[w]) : moments_merge(m[w], o[w]);
    }
  }
}

// Pass two over the queue. A workgroup takes WAVES consecutive sample
// blocks, one wave each. Each wave copies its block's base row into LDS (so
// every lane's point is two table reads and one read of its block's row,
// XORed); the queued legs of all the blocks (their regions, concatenated)
// are cut into chunks of 64, and the chunks go to the waves in turn, spread
// over the CU's four SIMDs, so the SIMDs share the fill evenly whatever the
// blocks' counts (a leg's value does not depend on the lane that fills it).
// After a workgroup barrier each wave forms its block's moments from the
// 256 leg values as block_merge over 256 threads forms them: lane l holds
// samples l, 64 + l, 128 + l and 192 + l, merges each set across the wave
// (wave_merge), then the four results ((0, 1), (2, 3)).
template <bool SOBOL, bool FP32, bool LDS>
__global__ void __launch_bounds__(64 * fill_waves<FP32>())
    bridge_fill_kernel(AcPlanDev p, uint64_t seed, int64_t samples, int64_t paths, int antithetic, int64_t blocks,
                       const uint64_t* __restrict__ dirs, const uint64_t* __restrict__ shifts,
                       const void* __restrict__ base_rows, const uint32_t* __restrict__ queue,
                       const uint32_t* __restrict__ counts, double* __restrict__ values,
                       Moments* __restrict__ partials, uint32_t* __restrict__ flags) {
  using R = typename std::conditional<FP32, float, double>::type;
  using W = SobolWord<FP32>;
  constexpr int SHIFT = FP32 ? 20 : 0;
  constexpr int WAVES = fill_waves<FP32>();
  constexpr int THREADS = 64 * WAVES;
  constexpr int REGIONS = WAVES * QUEUE_WAVES;
  static_assert(REGIONS <= 64, "the region scan runs in one wave");
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ uint32_t edge[REGIONS + 1];
  __shared__ int per_simd[4];
  const int dims = p.n_steps;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t first = (int64_t)blockIdx.x * WAVES;
  const int64_t block = first + wave;
  if (threadIdx.x < 4) per_simd[threadIdx.x] = 0;
  // LDS tables without a base point (block-independent), then each wave's
  // base row of dims + 1 words (the last one for the fill loop's final
  // look-ahead, whose value goes unused).
  Draws<SOBOL, FP32, LDS ? 2 : 1> proto;
  proto.dims = dims;
  proto.direct = !LDS;
  proto.dirs = dirs;
  proto.shifts = shifts;
  proto.base = nullptr;
  W* rows = nullptr;
  if (SOBOL && LDS) {
    W* lo = (W*)lds;
    W* hi = lo + (size_t)dims * 16;
    for (int d = threadIdx.x; d < dims; d += THREADS)
      sobol_table_rows<W, SHIFT>(dirs, dims, d, 0ull, lo + d * 16, hi + d * 16);
    rows = hi + (size_t)dims * 16 + 16;
    if (block < blocks)
      for (int d = lane; d <= dims; d += 64) rows[wave * (dims + 1) + d] = ((const W*)base_rows)[block * dims + d];
    proto.lo = lo;
    proto.hi = hi;
    if constexpr (FP32) {
      // The fill coefficients too: read from LDS, the fill loop waits on
      // LDS reads alone, in order, instead of on every scalar load as well.
      float* wf = (float*)(rows + (size_t)WAVES * (dims + 1));
      for (int i = threadIdx.x; i < 2 * dims; i += THREADS) wf[i] = p.wf32[i];
      p.wf32 = wf;
    }
  }
  // The workgroup's regions (QUEUE_WAVES per block, empty past the last
  // block) and their running sums: region k holds items edge[k]..edge[k+1].
  if (wave == 0) {
    uint32_t c = lane < REGIONS && first * QUEUE_WAVES + lane < blocks * QUEUE_WAVES
                     ? counts[first * QUEUE_WAVES + lane] : 0u;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(c, o, 64);
      if (lane >= o) c += t;
    }
    if (lane < REGIONS) edge[lane + 1] = c;
    if (lane == 0) edge[0] = 0u;
  }
  __syncthreads();
  // The wave's place in the order (rank 0 on SIMDs 0..3, then rank 1, ...):
  // chunk i goes to place i mod WAVES, so consecutive chunks land on
  // different SIMDs. Any SIMD numbering gives a permutation of the places.
  const int simd = __builtin_amdgcn_s_getreg(HW_ID_SIMD) & 3;
  int rank = 0;
  if (lane == 0) rank = atomicAdd(&per_simd[simd], 1);
  rank = __shfl(rank, 0, 64);
  __syncthreads();
  int place = 0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int n = per_simd[s];
    place += (n < rank ? n : rank) + (s < simd && n > rank ? 1 : 0);
  }
  const uint32_t total = edge[REGIONS];
  uint32_t bad = 0;
  for (uint32_t c = (uint32_t)place * 64 + lane; c < total; c += THREADS) {
    // The region of item c: the last k with edge[k] <= c.
    int k = 0;
#pragma unroll
    for (int step = REGIONS / 2; step > 0; step >>= 1) k += edge[k + step] <= c ? step : 0;
    const uint32_t item = queue[(size_t)(first * QUEUE_WAVES + k) * REGION + (c - edge[k])];
    const int64_t n = item >> 1;
    const int leg = item & 1;
    const int64_t b = first + k / QUEUE_WAVES;
    auto dr = sample_draws<SOBOL, FP32>(proto, seed, b, (int)(n - b * BLOCK_SAMPLES));
    if constexpr (SOBOL && LDS) dr.base = rows + (size_t)(k / QUEUE_WAVES) * (dims + 1);
    R lr;
    const bool ki = bridge_knock_in<SOBOL, FP32>(p, dr, leg ? -1.0f : 1.0f, lr);
    double& value = values[leg * samples + n];
    const double cash = value + terminal_value(p, (double)lr, ki, bad);
    if (!isfinite(cash)) bad = 1;
    value = cash;
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
  // Every wave's stores above are read by the block's wave below.
  __syncthreads();
  if (block >= blocks) return;
  // Set 0 (every leg value), then set 1 (the pair midpoints): one set's
  // registers at a time.
  Moments m[4];
  bool full = true;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int64_t n = block * BLOCK_SAMPLES + w * 64 + lane;
    const bool live = n < samples;
    full = full && live;
    m[w] = moments_zero();
    if (live) {
      moments_add(m[w], values[n], true);
      if (antithetic && n < paths / 2) moments_add(m[w], values[samples + n], true);
    }
  }
  // Without an empty set anywhere in the block the merges need no empty
  // checks (moments_merge_full).
  wave_merge4(m, wave_all(full));
  if (lane == 0) partials[block * 2] = moments_merge(moments_merge(m[0], m[1]), moments_merge(m[2], m[3]));
  // Merging empty sets gives empty sets: without antithetic pairs set 1 is
  // empty.
  full = true;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int64_t n = block * BLOCK_SAMPLES + w * 64 + lane;
    const bool paired = antithetic && n < samples && n < paths / 2;
    full = full && paired;
    m[w] = moments_zero();
    if (paired) moments_add(m[w], midpoint(values[n], values[samples + n]), true);
  }
  if (antithetic) wave_merge4(m, wave_all(full));
  if (lane == 0) partials[block * 2 + 1] = moments_merge(moments_merge(m[0], m[1]), moments_merge(m[2], m[3]));
}

// The cached base rows of `ctx` for (dims, seed, blocks, precision), built
// on a miss.
template <bool FP32>
static const SobolWord<FP32>* bridge_base_rows(kg_context* ctx, int dims, uint64_t seed, int64_t blocks,
                                               const uint64_t* dirs, const uint64_t* shifts) {
  BridgeBaseKey& key = ctx->bridge_base_key;
  if (key.valid && key.dims == dims && key.seed == seed && key.blocks == blocks && key.fp32 == FP32)
    return (const SobolWord<FP32>*)ctx->bridge_base.ptr;
  key.valid = false;
  // One word per dimension per sample block, plus one: the fill kernel
  // copies dims + 1 words of a row.
  SobolWord<FP32>* base =
      (SobolWord<FP32>*)ctx->bridge_base.reserve(sizeof(SobolWord<FP32>) * ((size_t)blocks * dims + 1));
  const int64_t cells = blocks * dims;
  hipLaunchKernelGGL(sobol_base_rows_kernel<FP32>, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, ctx->stream,
                     dirs, shifts, dims, blocks, base);
  KG_HIP(hipGetLastError());
  key = BridgeBaseKey{dims, seed, blocks, FP32, true};
  return base;
}

// Interval rows (and, for LDS Sobol, the first-dimension table rows) of a
// bridge plan, built on the host.
template <bool SOBOL, bool FP32, bool LDS>
static void bridge_rows(kg_context* ctx, const AcHostPlan& h, int dims, uint64_t seed,
                        std::vector<unsigned char>& blob, size_t* table_offset) {
  using R = typename std::conditional<FP32, float, double>::type;
  using W = SobolWord<FP32>;
  const size_t n = h.intervals.size();
  std::vector<IntervalRow<R>> rows(n);
  for (size_t k = 0; k < n; ++k) {
    IntervalRow<R>& r = rows[k];
    r = IntervalRow<R>{};
    r.first = k == 0 ? 0 : h.intervals[k - 1].end + 1;
    r.end = h.intervals[k].end;
    r.drift = FP32 ? (R)h.intervals32[2 * k] : (R)h.intervals[k].drift;
    r.diffusion = FP32 ? (R)h.intervals32[2 * k + 1] : (R)h.intervals[k].diffusion;
    r.observe = h.step_observe[r.end];
    r.event = h.step_event[r.end];
    if (r.event >= 0) {
      const AcEventD& ev = h.events[r.event];
      r.pay_coupon = ev.pay_coupon;
      r.pay_out = ev.pay_out;
      r.pay_both = ev.pay_both;
      r.knock_out = FP32 ? (R)h.events32[2 * r.event] : (R)ev.knock_out;
      r.coupon_barrier = FP32 ? (R)h.events32[2 * r.event + 1] : (R)ev.coupon_barrier;
      r.has_knock_out = ev.has_knock_out;
    }
  }
  blob.assign((const unsigned char*)rows.data(), (const unsigned char*)(rows.data() + n));
  *table_offset = (blob.size() + 255) & ~(size_t)255;
  // Room for the first-dimension Sobol tables, which bridge_table_kernel
  // fills on the device from the scrambled directions.
  if (SOBOL && LDS) blob.resize(*table_offset + n * 32 * sizeof(W), 0);
}

// The Sobol table rows of each interval's first dimension d = rows[k].first
// (lo[k][h] = directions 0..3 selected by h, hi[k][h] = directions 4..7),
// stored >> SHIFT: one thread per (interval, h).
template <bool FP32>
__global__ void __launch_bounds__(256)
    bridge_table_kernel(const IntervalRow<typename std::conditional<FP32, float, double>::type>* __restrict__ rows,
                        int n, const uint64_t* __restrict__ dirs, int dims, SobolWord<FP32>* __restrict__ table) {
  constexpr int SHIFT = FP32 ? 20 : 0;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * 16) return;
  const int k = t >> 4, hix = t & 15;
  const int d = rows[k].first;
  uint64_t a = 0, b = 0;
  for (int bit = 0; bit < 4; ++bit)
    if (hix & (1 << bit)) {
      a ^= dirs[(size_t)bit * dims + d];
      b ^= dirs[(size_t)(4 + bit) * dims + d];
    }
  table[k * 16 + hix] = (SobolWord<FP32>)(a >> SHIFT);
  table[(n + k) * 16 + hix] = (SobolWord<FP32>)(b >> SHIFT);
}

// What a Sobol bridge note needs before its two passes besides the scrambled
// directions, in one launch: the call's flag word is cleared; workgroup 0
// writes the table rows of the intervals' first dimensions as
// bridge_table_kernel writes them; and with `base`, workgroup g writes the
// base points of sample blocks [g, g + 1) * BASE_ROWS_PER_GROUP, thread d
// taking dimension d of each block in turn, so that a block's row is
// written by consecutive threads. The base point of block m is the Sobol
// point m * 256: its Gray code is gray(m) shifted up 8 bits with bit 7 = m
// & 1, and gray(m) differs from gray(m - 1) in bit ctz(m), so one XOR moves a
// thread from block m - 1 to block m (the same point sobol_block_base forms).
constexpr int BASE_ROWS_PER_GROUP = 32;
constexpr int BASE_ROW_COLS = 21;  // directions 7 .. 27: up to 2^20 sample blocks

template <bool FP32>
__global__ void __launch_bounds__(256)
    bridge_setup_kernel(int dims, const uint64_t* __restrict__ dirs, const uint64_t* __restrict__ shifts,
                        int64_t blocks, SobolWord<FP32>* __restrict__ base,
                        const IntervalRow<typename std::conditional<FP32, float, double>::type>* __restrict__ rows,
                        int n_intervals, SobolWord<FP32>* __restrict__ table, uint32_t* __restrict__ flags) {
  constexpr int SHIFT = FP32 ? 20 : 0;
  static_assert(BLOCK_SAMPLES == 256, "the base points step by 2^8");
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) *flags = 0u;
    for (int t = threadIdx.x; t < n_intervals * 16; t += blockDim.x) {
      const int k = t >> 4, hix = t & 15;
      const int d = rows[k].first;
      uint64_t v[8];  // all eight loads in flight at once
#pragma unroll
      for (int c = 0; c < 8; ++c) v[c] = dirs[(size_t)c * dims + d];
      uint64_t a = 0, b = 0;
#pragma unroll
      for (int bit = 0; bit < 4; ++bit) {
        const uint64_t on = (hix >> bit) & 1 ? ~0ull : 0ull;
        a ^= v[bit] & on;
        b ^= v[4 + bit] & on;
      }
      table[k * 16 + hix] = (SobolWord<FP32>)(a >> SHIFT);
      table[(n_intervals + k) * 16 + hix] = (SobolWord<FP32>)(b >> SHIFT);
    }
  }
  if (!base) return;
  // Thread t keeps the directions 7 .. 7 + cols - 1 of its dimension in LDS
  // column t (only it reads them): gray(m) for m < blocks has bits below
  // cols - 1.
  __shared__ uint64_t col[BASE_ROW_COLS][256];
  const int cols = 1 + (64 - __builtin_clzll((uint64_t)blocks));
  const int64_t m0 = (int64_t)blockIdx.x * BASE_ROWS_PER_GROUP;
  const int64_t m1 = m0 + BASE_ROWS_PER_GROUP < blocks ? m0 + BASE_ROWS_PER_GROUP : blocks;
  for (int d = threadIdx.x; d < dims; d += blockDim.x) {
    uint64_t* mine = &col[0][threadIdx.x];
    for (int c = 0; c < cols; ++c) mine[c * 256] = dirs[(size_t)(7 + c) * dims + d];
    const uint64_t shift = shifts[d], odd = mine[0];
    uint64_t g = 0;  // XOR of the directions 8 + i over the bits i of gray(m)
    for (uint64_t bits = (uint64_t)m0 ^ ((uint64_t)m0 >> 1); bits; bits &= bits - 1)
      g ^= mine[(__builtin_ctzll(bits) + 1) * 256];
    for (int64_t m = m0; m < m1; ++m) {
      if (m > m0) g ^= mine[(__builtin_ctzll((uint64_t)m) + 1) * 256];
      base[m * dims + d] = (SobolWord<FP32>)((shift ^ g ^ ((m & 1) ? odd : 0)) >> SHIFT);
    }
  }
}

// Scenario lanes on shared draws: lane l's total and pair moments of the
// block go to partials[block * slots + slot_base + 2l] and [... + 1],
// formed exactly as the single-plan kernel forms its own.
template <int L, bool SOBOL, bool FP32, bool LDS>
__global__ void __launch_bounds__(256)
    autocall_lanes_kernel(const AcLanes* __restrict__ lanes, uint64_t seed, int64_t samples, int64_t paths, int antithetic,
                          const uint64_t* __restrict__ dirs, const uint64_t* __restrict__ shifts,
                          Moments* __restrict__ partials, uint32_t* __restrict__ flags) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ Moments smem[2 * 4];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const bool live = n < samples;
  const AcLanes& P = *lanes;
  auto dr = setup_draws<SOBOL, FP32, LDS>(P.n_steps, seed, block, j, dirs, shifts, lds);
  uint32_t bad = 0;
  const bool paired = antithetic && live && n < paths / 2;
  const auto origin = dr;
  double pos[L], neg[L];
  if (P.bridge)
    bridge_lanes<L, SOBOL, FP32>(P, dr, live, 1.0f, bad, pos);
  else
    step_lanes<L, SOBOL, FP32>(P, dr, live, 1.0f, bad, pos);
#pragma unroll
  for (int l = 0; l < L; ++l) neg[l] = 0.0;
  if (antithetic && __ballot(paired) != 0) {
    dr = origin;
    if (P.bridge)
      bridge_lanes<L, SOBOL, FP32>(P, dr, paired, -1.0f, bad, neg);
    else
      step_lanes<L, SOBOL, FP32>(P, dr, paired, -1.0f, bad, neg);
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
#pragma unroll
  for (int l = 0; l < L; ++l) {
    if (l >= P.n) break;
    Moments m[2] = {moments_zero(), moments_zero()};
    if (live) {
      moments_add(m[0], pos[l], true);
      if (paired) {
        moments_add(m[0], neg[l], true);
        moments_add(m[1], midpoint(pos[l], neg[l]), true);
      }
    }
    block_merge<2>(m, smem);
    if (threadIdx.x == 0) {
      partials[block * P.slots + P.slot_base + 2 * l] = m[0];
      partials[block * P.slots + P.slot_base + 2 * l + 1] = m[1];
    }
  }
}

static int cu_count(kg_context* ctx) {
  int cus = 0;
  KG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  return std::max(cus, 1);
}

// Dynamic LDS of bridge_events_kernel: the plan's interval rows, and with
// Sobol tables in LDS 32 table words per interval.
static size_t bridge_events_lds(int n_intervals, int dims, bool sobol, bool fp32) {
  size_t bytes = (size_t)n_intervals * (fp32 ? sizeof(IntervalRow<float>) : sizeof(IntervalRow<double>));
  if (sobol && sobol_lds_bytes(dims, fp32) <= SOBOL_LDS_LIMIT)
    bytes += (size_t)n_intervals * 32 * (fp32 ? sizeof(uint32_t) : sizeof(uint64_t));
  return bytes;
}

// Largest bridge_events_kernel LDS taken without an opt-in; plans with more
// interval rows (many knock-out events after the lockout) price in one pass.
constexpr size_t BRIDGE_EVENTS_LDS_LIMIT = 64 * 1024;

// The two bridge passes (events, then the daily fill of the queued legs)
// over the interval rows at d_rows; `base` holds the sample blocks' Sobol
// base points and the table rows follow d_rows at table_offset (LDS Sobol).
template <bool SOBOL, bool FP32, bool LDS>
static void launch_bridge_passes(kg_context* ctx, const AcPlanDev& p, const Sampler& s, const uint64_t* dirs,
                                 const uint64_t* shifts, Moments* parts, uint32_t* flags,
                                 const unsigned char* d_rows, size_t table_offset, const SobolWord<FP32>* base) {
  using R = typename std::conditional<FP32, float, double>::type;
  using W = SobolWord<FP32>;
  const int dims = p.n_steps;
  const int64_t samples = s.samples(), blocks = s.blocks();
  const int legs = s.antithetic ? 2 : 1;
  const size_t regions = (size_t)blocks * QUEUE_WAVES;
  double* values = (double*)ctx->bridge_values.reserve(sizeof(double) * (size_t)samples * legs);
  // Per pass-one wave: a region of REGION items and its count.
  uint32_t* queue = (uint32_t*)ctx->bridge_queue.reserve(sizeof(uint32_t) * regions * (REGION + 1));
  uint32_t* counts = queue + regions * REGION;
  size_t events_lds = (size_t)p.n_intervals * sizeof(IntervalRow<R>), fill_lds = 0;
  if (SOBOL && LDS) {
    events_lds += (size_t)p.n_intervals * 32 * sizeof(W);
    fill_lds = sobol_lds_bytes(dims, FP32) + (size_t)fill_waves<FP32>() * (dims + 1) * sizeof(W) +
               (FP32 ? (size_t)2 * dims * sizeof(float) : 0);
  }
  if (events_lds > BRIDGE_EVENTS_LDS_LIMIT) fail(KG_DEVICE_ERROR, "bridge plan exceeds its pass-one LDS");
  // The call's device time covers the two sampling kernels only.
  timing_begin(ctx);
  hipLaunchKernelGGL((bridge_events_kernel<SOBOL, FP32, LDS>), dim3(blocks), dim3(256), events_lds, ctx->stream, p,
                     s.seed, samples, s.path_count, (int)s.antithetic, dirs, shifts, values, d_rows,
                     d_rows + table_offset, base, queue, counts, flags);
  KG_HIP(hipGetLastError());
  if (fill_lds > 64 * 1024)
    KG_HIP(hipFuncSetAttribute((const void*)bridge_fill_kernel<SOBOL, FP32, LDS>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)fill_lds));
  constexpr int WAVES = fill_waves<FP32>();
  hipLaunchKernelGGL((bridge_fill_kernel<SOBOL, FP32, LDS>), dim3((unsigned)((blocks + WAVES - 1) / WAVES)),
                     dim3(64 * WAVES), fill_lds, ctx->stream, p, s.seed, samples, s.path_count, (int)s.antithetic,
                     blocks, dirs, shifts, base, queue, counts, values, parts, flags);
  KG_HIP(hipGetLastError());
}

// The two-pass launches of a bridge plan (see bridge_events_kernel): for LDS
// Sobol, the base rows (kept per dimensions, seed and block count) and the
// interval table rows first.
template <bool SOBOL, bool FP32, bool LDS>
static void launch_bridge(kg_context* ctx, const AcPlanDev& p, const Sampler& s, const uint64_t* dirs,
                          const uint64_t* shifts, Moments* parts, uint32_t* flags, const unsigned char* d_rows,
                          size_t table_offset) {
  using R = typename std::conditional<FP32, float, double>::type;
  using W = SobolWord<FP32>;
  const int dims = p.n_steps;
  const W* base = SOBOL && LDS ? bridge_base_rows<FP32>(ctx, dims, s.seed, s.blocks(), dirs, shifts) : nullptr;
  if (SOBOL && LDS)
    hipLaunchKernelGGL(bridge_table_kernel<FP32>, dim3((unsigned)((p.n_intervals * 16 + 255) / 256)), dim3(256), 0,
                       ctx->stream, (const IntervalRow<R>*)d_rows, p.n_intervals, dirs, dims,
                       (W*)const_cast<unsigned char*>(d_rows + table_offset));
  launch_bridge_passes<SOBOL, FP32, LDS>(ctx, p, s, dirs, shifts, parts, flags, d_rows, table_offset, base);
}

// A Sobol bridge note with LDS tables: the scrambled directions (kept per
// dimensions and seed), then bridge_setup_kernel (flag word, base rows unless
// kept for dimensions, seed and block count, interval table rows), then the
// two passes.
template <bool FP32>
static void launch_bridge_fused(kg_context* ctx, const AcPlanDev& p, const Sampler& s, Moments* parts,
                                uint32_t* flags, const unsigned char* d_rows, size_t table_offset) {
  using R = typename std::conditional<FP32, float, double>::type;
  using W = SobolWord<FP32>;
  const int dims = p.n_steps;
  const int64_t blocks = s.blocks();
  const SobolSetup& so = sobol_setup(ctx, (uint32_t)dims, s.seed);
  const uint64_t* dirs = (const uint64_t*)so.d_directions.ptr;
  const uint64_t* shifts = (const uint64_t*)so.d_shifts.ptr;
  BridgeBaseKey& key = ctx->bridge_base_key;
  const bool have_base =
      key.valid && key.dims == dims && key.seed == s.seed && key.blocks == blocks && key.fp32 == FP32;
  key.valid = false;
  // One word per dimension per sample block, plus one: the fill kernel
  // copies dims + 1 words of a row.
  W* base = (W*)(have_base ? ctx->bridge_base.ptr
                           : ctx->bridge_base.reserve(sizeof(W) * ((size_t)blocks * dims + 1)));
  if (1 + (64 - __builtin_clzll((uint64_t)blocks)) > BASE_ROW_COLS)
    fail(KG_INVALID_INPUT, "too many sample blocks for the bridge set-up");
  const unsigned groups = have_base ? 1u : (unsigned)((blocks + BASE_ROWS_PER_GROUP - 1) / BASE_ROWS_PER_GROUP);
  hipLaunchKernelGGL(bridge_setup_kernel<FP32>, dim3(groups), dim3(256), 0, ctx->stream, dims, dirs, shifts, blocks,
                     have_base ? (W*)nullptr : base,
                     (const IntervalRow<R>*)d_rows, p.n_intervals,
                     (W*)const_cast<unsigned char*>(d_rows + table_offset), flags);
  KG_HIP(hipGetLastError());
  key = BridgeBaseKey{dims, s.seed, blocks, FP32, true};
  launch_bridge_passes<true, FP32, true>(ctx, p, s, dirs, shifts, parts, flags, d_rows, table_offset, base);
}

template <bool SOBOL, bool FP32>
static void launch_bridge_any(kg_context* ctx, const AcPlanDev& p, const Sampler& s, const uint64_t* dirs,
                              const uint64_t* shifts, Moments* parts, uint32_t* flags, const unsigned char* d_rows,
                              size_t table_offset) {
  if (SOBOL && sobol_lds_bytes(p.n_steps, FP32) <= SOBOL_LDS_LIMIT)
    launch_bridge<SOBOL, FP32, true>(ctx, p, s, dirs, shifts, parts, flags, d_rows, table_offset);
  else
    launch_bridge<SOBOL, FP32, false>(ctx, p, s, dirs, shifts, parts, flags, d_rows, table_offset);
}

// The interval rows of launch_bridge_any's kernels (same table choice).
template <bool SOBOL, bool FP32>
static void bridge_rows_any(kg_context* ctx, const AcHostPlan& h, uint64_t seed, std::vector<unsigned char>& blob,
                            size_t* table_offset) {
  const int dims = h.dev.n_steps;
  if (SOBOL && sobol_lds_bytes(dims, FP32) <= SOBOL_LDS_LIMIT)
    bridge_rows<SOBOL, FP32, true>(ctx, h, dims, seed, blob, table_offset);
  else
    bridge_rows<SOBOL, FP32, false>(ctx, h, dims, seed, blob, table_offset);
}

template <bool SOBOL, bool FP32, int L, class Plan>
static void launch(kg_context* ctx, const Plan& p, int dims, const Sampler& s, const uint64_t* dirs,
                   const uint64_t* shifts, Moments* parts, uint32_t* flags) {
  size_t lds = 0;
  bool sobol_lds = false;
  if (SOBOL) {
    lds = sobol_lds_bytes(dims, FP32);
    sobol_lds = lds <= SOBOL_LDS_LIMIT;
    if (!sobol_lds) lds = 0;
  }
  using Kernel = void (*)(Plan, uint64_t, int64_t, int64_t, int, const uint64_t*, const uint64_t*, Moments*,
                          uint32_t*);
  Kernel kernel;
  if constexpr (std::is_same<Plan, const AcLanes*>::value)
    kernel = sobol_lds ? autocall_lanes_kernel<L, SOBOL, FP32, true> : autocall_lanes_kernel<L, SOBOL, FP32, false>;
  else
    kernel = sobol_lds ? autocall_kernel<SOBOL, FP32, true> : autocall_kernel<SOBOL, FP32, false>;
  if (lds > 64 * 1024)
    KG_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
  hipLaunchKernelGGL(kernel, dim3(s.blocks()), dim3(256), lds, ctx->stream, p, s.seed, s.samples(),
                     s.path_count, (int)s.antithetic, dirs, shifts, parts, flags);
  KG_HIP(hipGetLastError());
}

template <int L = 1, class Plan>
static void launch_any(kg_context* ctx, const Plan& p, int dims, const Sampler& s, Moments* parts,
                       uint32_t* flags) {
  const uint64_t* dirs = nullptr;
  const uint64_t* shifts = nullptr;
  if (s.sobol) {
    const SobolSetup& so = sobol_setup(ctx, (uint32_t)dims, s.seed);
    dirs = (const uint64_t*)so.d_directions.ptr;
    shifts = (const uint64_t*)so.d_shifts.ptr;
  }
  if (s.sobol && s.fp32)
    launch<true, true, L>(ctx, p, dims, s, dirs, shifts, parts, flags);
  else if (s.sobol)
    launch<true, false, L>(ctx, p, dims, s, dirs, shifts, parts, flags);
  else if (s.fp32)
    launch<false, true, L>(ctx, p, dims, s, dirs, shifts, parts, flags);
  else
    launch<false, false, L>(ctx, p, dims, s, dirs, shifts, parts, flags);
}

// The lanes' tables interleaved row by row ([row][lane]) in one staging
// blob; table pointers are offsets into the blob until bind_lanes.
template <class T>
static std::vector<T> interleave(const std::vector<const AcHostPlan*>& L, std::vector<T> AcHostPlan::*table,
                                 size_t width) {
  const size_t n = L.size(), rows = (L[0]->*table).size() / width;
  std::vector<T> out(rows * n * width);
  for (size_t l = 0; l < n; ++l)
    for (size_t r = 0; r < rows; ++r)
      for (size_t c = 0; c < width; ++c) out[(r * n + l) * width + c] = (L[l]->*table)[r * width + c];
  return out;
}

struct LaneOffsets {
  size_t o[13];
};

static AcLanes stage_lanes(const std::vector<const AcHostPlan*>& L, std::vector<unsigned char>& blob,
                           LaneOffsets& off) {
  const AcHostPlan& a = *L[0];
  AcLanes P{};
  P.n = (int32_t)L.size();
  P.n_steps = a.dev.n_steps;
  P.n_intervals = a.dev.n_intervals;
  P.bridge = a.dev.bridge;
  for (size_t l = 0; l < L.size(); ++l) {
    const AcPlanDev& d = L[l]->dev;
    P.s[l] = AcLaneScalars{d.knock_in,  d.coupon,     d.initial_spot,  d.safe_lo, d.safe_hi,
                           d.final_log_discount,   d.term_base,     d.term_no_ki, d.term_ki_fixed,
                           d.lower,     d.upper,      d.inv_initial,   d.knocked_in, d.has_ki,
                           d.expiry_knock_in,      d.downside};
  }
  std::vector<int32_t> ends;
  for (const AcIntervalD& b : a.intervals) ends.push_back(b.end);
  std::vector<AcStepD> intervals(a.intervals.size() * L.size());
  for (size_t l = 0; l < L.size(); ++l)
    for (size_t k = 0; k < a.intervals.size(); ++k)
      intervals[k * L.size() + l] = AcStepD{L[l]->intervals[k].drift, L[l]->intervals[k].diffusion};
  size_t* o = off.o;
  stage(blob, a.step_event, &o[0]);
  stage(blob, a.step_observe, &o[1]);
  stage(blob, ends, &o[2]);
  stage(blob, interleave(L, &AcHostPlan::steps, 1), &o[3]);
  stage(blob, interleave(L, &AcHostPlan::weight, 1), &o[4]);
  stage(blob, interleave(L, &AcHostPlan::fill, 1), &o[5]);
  stage(blob, interleave(L, &AcHostPlan::events, 1), &o[6]);
  stage(blob, intervals, &o[7]);
  stage(blob, interleave(L, &AcHostPlan::steps32, 2), &o[8]);
  stage(blob, interleave(L, &AcHostPlan::keep32, 1), &o[9]);
  stage(blob, interleave(L, &AcHostPlan::fill32, 1), &o[10]);
  stage(blob, interleave(L, &AcHostPlan::events32, 2), &o[11]);
  stage(blob, interleave(L, &AcHostPlan::intervals32, 2), &o[12]);
  return P;
}

static void bind_lanes(AcLanes& P, unsigned char* d, const LaneOffsets& off) {
  const size_t* o = off.o;
  P.step_event = (const int32_t*)(d + o[0]);
  P.step_observe = (const uint8_t*)(d + o[1]);
  P.interval_end = (const int32_t*)(d + o[2]);
  P.steps = (const AcStepD*)(d + o[3]);
  P.weight = (const double*)(d + o[4]);
  P.fill = (const double*)(d + o[5]);
  P.events = (const AcEventD*)(d + o[6]);
  P.intervals = (const AcStepD*)(d + o[7]);
  P.steps32 = (const float*)(d + o[8]);
  P.keep32 = (const float*)(d + o[9]);
  P.fill32 = (const float*)(d + o[10]);
  P.events32 = (const float*)(d + o[11]);
  P.intervals32 = (const float*)(d + o[12]);
}

}  // namespace kg

using namespace kg;

namespace kg {

// One sampled note on its own: the single-plan kernel and the block merge.
// The tables of one sampled note as uploaded: the plan's tables and, for
// two-pass bridge plans, the interval rows (and room for the Sobol table
// rows that bridge_table_kernel fills).
struct StagedNote {
  std::vector<unsigned char> blob;
  AcOffsets off{};
  size_t rows_at = 0, table_offset = 0;
  bool two_pass = false;
};

static void stage_note(kg_context* ctx, const Sampler& s, const AcHostPlan& h, StagedNote& st) {
  // Bridge plans run in two passes (queue items carry the sample index in
  // 31 bits) while pass one's interval rows fit its LDS; other plans, and
  // larger ones, in one. The two-pass interval rows follow the plan's tables
  // in the same upload.
  st.two_pass = h.dev.bridge && s.samples() < ((int64_t)1 << 31) &&
                bridge_events_lds(h.dev.n_intervals, h.dev.n_steps, s.sobol, s.fp32) <= BRIDGE_EVENTS_LDS_LIMIT;
  st.blob.clear();
  st.off = stage_plan(st.blob, h);
  st.rows_at = st.table_offset = 0;
  if (st.two_pass) {
    std::vector<unsigned char> rows;
    if (s.sobol && s.fp32)
      bridge_rows_any<true, true>(ctx, h, s.seed, rows, &st.table_offset);
    else if (s.sobol)
      bridge_rows_any<true, false>(ctx, h, s.seed, rows, &st.table_offset);
    else if (s.fp32)
      bridge_rows_any<false, true>(ctx, h, s.seed, rows, &st.table_offset);
    else
      bridge_rows_any<false, false>(ctx, h, s.seed, rows, &st.table_offset);
    st.rows_at = (st.blob.size() + 255) & ~(size_t)255;
    st.blob.resize(st.rows_at);
    st.blob.insert(st.blob.end(), rows.begin(), rows.end());
  }
}

static void run_note(kg_context* ctx, const Sampler& s, const AcHostPlan& h, const StagedNote& st,
                     kg_estimate* out) {
  unsigned char* d = (unsigned char*)ctx->scratch.reserve(std::max<size_t>(st.blob.size(), 1));
  upload_async(ctx, d, st.blob.data(), st.blob.size());
  const AcPlanDev p = bind_plan(h, d, st.off);
  const unsigned char* d_rows = d + st.rows_at;
  const size_t table_offset = st.table_offset;
  const int64_t blocks = s.blocks();
  Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * 2 * blocks);
  if (st.two_pass && s.sobol && sobol_lds_bytes(p.n_steps, s.fp32) <= SOBOL_LDS_LIMIT) {
    // One set-up launch, then the two passes (the flag word is cleared by
    // the set-up kernel).
    uint32_t* flags = (uint32_t*)ctx->flags.reserve(256);
    if (s.fp32)
      launch_bridge_fused<true>(ctx, p, s, parts, flags, d_rows, table_offset);
    else
      launch_bridge_fused<false>(ctx, p, s, parts, flags, d_rows, table_offset);
  } else {
    uint32_t* flags = reset_flags(ctx);
    const uint64_t* dirs = nullptr;
    const uint64_t* shifts = nullptr;
    if (s.sobol) {
      const SobolSetup& so = sobol_setup(ctx, (uint32_t)p.n_steps, s.seed);
      dirs = (const uint64_t*)so.d_directions.ptr;
      shifts = (const uint64_t*)so.d_shifts.ptr;
    }
    if (st.two_pass) {
      if (s.sobol && s.fp32)
        launch_bridge_any<true, true>(ctx, p, s, dirs, shifts, parts, flags, d_rows, table_offset);
      else if (s.sobol)
        launch_bridge_any<true, false>(ctx, p, s, dirs, shifts, parts, flags, d_rows, table_offset);
      else if (s.fp32)
        launch_bridge_any<false, true>(ctx, p, s, dirs, shifts, parts, flags, d_rows, table_offset);
      else
        launch_bridge_any<false, false>(ctx, p, s, dirs, shifts, parts, flags, d_rows, table_offset);
    } else {
      timing_begin(ctx);
      launch_any(ctx, p, p.n_steps, s, parts, flags);
    }
  }
  Moments r[2];
  *out = kg_estimate{};
  finish_sampling(ctx, parts, blocks, 2, r, &out->device_ms);
  out->path_steps = (double)s.path_count * (double)p.n_steps;
  fill_estimate(s, r[0], r[1], 0.0, false, true, out);
}

static void price_note(kg_context* ctx, const Sampler& s, AcHostPlan& h, kg_estimate* out) {
  StagedNote st;
  stage_note(ctx, s, h, st);
  run_note(ctx, s, h, st, out);
}

// The context's last note plan: the terms (with the events they point to),
// the market, and the settings bits the staged tables depend on, as bytes.
// A call whose key matches reuses the plan and its staged tables; neither
// depends on the seed.
struct NotePlan {
  std::vector<unsigned char> key;
  AcHostPlan plan;
  StagedNote staged;
  uint32_t staged_bits = 0;
  bool staged_ok = false;
};

static std::vector<unsigned char> note_key(const kg_autocall_terms& t, const kg_bsm_market& m) {
  kg_autocall_terms copy = t;
  copy.events = nullptr;
  const size_t events = t.n_events > 0 && t.events ? (size_t)t.n_events * sizeof(kg_autocall_event) : 0;
  std::vector<unsigned char> key(sizeof copy + events + sizeof m);
  std::memcpy(key.data(), &copy, sizeof copy);
  if (events) std::memcpy(key.data() + sizeof copy, t.events, events);
  std::memcpy(key.data() + sizeof copy + events, &m, sizeof m);
  return key;
}

// Up to AC_MAX_SCENARIOS sampled plans on one step grid (same_grid) as lanes
// on the same draws, in launches of AC_MAX_LANES lanes; the lane sets follow
// their tables in one upload. Lane l's estimate equals price_note of plan l
// bit for bit.
static void price_lane_set(kg_context* ctx, const Sampler& s, const std::vector<AcHostPlan*>& plans,
                           const std::vector<kg_estimate*>& outs) {
  const int m = (int)plans.size();
  const int dims = plans[0]->dev.n_steps;
  const int slots = 2 * m;
  std::vector<unsigned char> blob;
  std::vector<AcLanes> groups;
  std::vector<LaneOffsets> offs;
  for (int g = 0; g < m; g += AC_MAX_LANES) {
    std::vector<const AcHostPlan*> L(plans.begin() + g, plans.begin() + std::min(m, g + AC_MAX_LANES));
    offs.emplace_back();
    groups.push_back(stage_lanes(L, blob, offs.back()));
    groups.back().slot_base = 2 * g;
    groups.back().slots = slots;
  }
  const size_t at = (blob.size() + 255) & ~size_t(255);
  blob.resize(at + sizeof(AcLanes) * groups.size());
  unsigned char* d = (unsigned char*)ctx->scratch.reserve(blob.size());
  for (size_t g = 0; g < groups.size(); ++g) {
    bind_lanes(groups[g], d, offs[g]);
    std::memcpy(blob.data() + at + g * sizeof(AcLanes), &groups[g], sizeof(AcLanes));
  }
  upload_async(ctx, d, blob.data(), blob.size());
  const int64_t blocks = s.blocks();
  Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * slots * blocks);
  uint32_t* flags = reset_flags(ctx);
  if (s.sobol) sobol_setup(ctx, (uint32_t)dims, s.seed);
  timing_begin(ctx);
  for (size_t g = 0; g < groups.size(); ++g) {
    const AcLanes* d_lanes = (const AcLanes*)(d + at + g * sizeof(AcLanes));
    if (groups[g].n <= 2)
      launch_any<2>(ctx, d_lanes, dims, s, parts, flags);
    else
      launch_any<4>(ctx, d_lanes, dims, s, parts, flags);
  }
  Moments r[2 * AC_MAX_SCENARIOS];
  double ms = 0.0;
  finish_sampling(ctx, parts, blocks, slots, r, &ms);
  for (int l = 0; l < m; ++l) {
    kg_estimate* o = outs[l];
    *o = kg_estimate{};
    o->device_ms = ms;
    o->path_steps = (double)s.path_count * (double)dims;
    fill_estimate(s, r[2 * l], r[2 * l + 1], 0.0, false, true, o);
  }
}

// Hash of the step grid that same_grid compares.
static uint64_t grid_hash(const AcHostPlan& h) {
  uint64_t x = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    x ^= v;
    x *= 1099511628211ull;
  };
  mix((uint64_t)h.dev.n_steps);
  mix((uint64_t)h.dev.bridge);
  for (int32_t e : h.step_event) mix((uint64_t)(uint32_t)e);
  for (uint8_t o : h.step_observe) mix(o);
  for (const AcIntervalD& b : h.intervals) mix((uint64_t)(uint32_t)b.end);
  return x;
}

constexpr size_t BATCH_NOTES = size_t(1) << 20;  // binding_support::validate_contract_batch_size
constexpr size_t PLAN_CHUNK = 1024;              // notes planned and held at once

}  // namespace kg

extern "C" kg_status kg_autocallable_price(kg_context* ctx, const kg_settings* settings,
                                           const kg_autocall_terms* note,
                                           const kg_bsm_market* market, kg_estimate* out) {
  return guarded(ctx, [&] {
    if (!note || !market || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
    std::vector<unsigned char> key = note_key(*note, *market);
    auto cached = std::static_pointer_cast<NotePlan>(ctx->note_plan);
    if (!cached || cached->key != key) {
      ctx->note_plan.reset();
      auto made = std::make_shared<NotePlan>();
      made->plan = build_plan(*note, *market);
      made->key = std::move(key);
      ctx->note_plan = made;
      cached = made;
    }
    const AcHostPlan& h = cached->plan;
    *out = kg_estimate{};
    if (h.fixed) {
      fill_known(h.fixed_value, settings && settings->sampling == KG_SOBOL, out);
      return;
    }
    Sampler s = validate_settings(settings, (uint64_t)h.dev.n_steps);
    // The staged tables depend on the sampling, precision and whether the
    // samples fit the two-pass queue items; a call differing in those
    // restages.
    StagedNote& st = cached->staged;
    const uint32_t bits = (uint32_t)s.sobol | ((uint32_t)s.fp32 << 1) |
                          ((uint32_t)(s.samples() < ((int64_t)1 << 31)) << 2);
    if (!cached->staged_ok || cached->staged_bits != bits) {
      cached->staged_ok = false;
      stage_note(ctx, s, h, st);
      cached->staged_bits = bits;
      cached->staged_ok = true;
    }
    run_note(ctx, s, h, st, out);
  });
}

extern "C" kg_status kg_autocallable_scenarios(kg_context* ctx, const kg_settings* settings,
                                               const kg_autocall_terms* note, const kg_bsm_market* markets,
                                               int32_t n, kg_estimate* out) {
  return guarded(ctx, [&] {
    if (!note || !markets || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
    if (n < 1 || n > AC_MAX_SCENARIOS) fail(KG_INVALID_INPUT, "scenario count must be between 1 and 16");
    std::vector<AcHostPlan> h(n);
    for (int i = 0; i < n; ++i) {
      h[i] = build_plan(*note, markets[i]);
      out[i] = kg_estimate{};
    }
    // Scenarios with a known value take it; the others become lanes.
    std::vector<AcHostPlan*> lanes;
    std::vector<kg_estimate*> outs;
    const bool sobol = settings && settings->sampling == KG_SOBOL;
    for (int i = 0; i < n; ++i) {
      if (h[i].fixed) {
        fill_known(h[i].fixed_value, sobol, &out[i]);
        continue;
      }
      if (!lanes.empty() && !same_grid(*lanes[0], h[i]))
        fail(KG_UNSUPPORTED_OPERATION, "the scenarios do not share the path grid; price them one at a time");
      lanes.push_back(&h[i]);
      outs.push_back(&out[i]);
    }
    if (lanes.empty()) return;
    Sampler s = validate_settings(settings, (uint64_t)lanes[0]->dev.n_steps);
    price_lane_set(ctx, s, lanes, outs);
  });
}

// price_many and price_many_on_shared_grid: notes grouped by step grid, each
// group's notes as lanes of one path set. On shared grids a note with a
// daily grid is planned without the bridge (build_plan).
static void price_many(kg_context* ctx, const kg_settings* settings, const kg_autocall_terms* notes, size_t n,
                       const kg_bsm_market* market, kg_estimate* out, bool shared_grid) {
  {
    if (n == 0) return;
    if (!notes || !market || !out || !settings) fail(KG_INVALID_INPUT, "arguments must not be null");
    if (n > BATCH_NOTES) fail(KG_INVALID_INPUT, "a batch holds at most 1,048,576 notes");
    const bool sobol = settings->sampling == KG_SOBOL;
    for (size_t c0 = 0; c0 < n; c0 += PLAN_CHUNK) {
      const size_t c1 = std::min(n, c0 + PLAN_CHUNK);
      std::vector<AcHostPlan> h(c1 - c0);
      // Groups of sampled notes with one step grid, in order of first note.
      std::vector<std::vector<size_t>> groups;
      std::unordered_map<uint64_t, std::vector<size_t>> by_hash;
      for (size_t i = c0; i < c1; ++i) {
        AcHostPlan& p = h[i - c0];
        p = build_plan(notes[i], *market, shared_grid);
        out[i] = kg_estimate{};
        if (p.fixed) {
          fill_known(p.fixed_value, sobol, &out[i]);
          continue;
        }
        std::vector<size_t>& candidates = by_hash[grid_hash(p)];
        size_t g = groups.size();
        for (size_t k : candidates)
          if (same_grid(h[groups[k][0] - c0], p)) {
            g = k;
            break;
          }
        if (g == groups.size()) {
          candidates.push_back(g);
          groups.emplace_back();
        }
        groups[g].push_back(i);
      }
      for (const std::vector<size_t>& g : groups) {
        const Sampler s = validate_settings(settings, (uint64_t)h[g[0] - c0].dev.n_steps);
        if (g.size() == 1) {
          price_note(ctx, s, h[g[0] - c0], &out[g[0]]);
          continue;
        }
        for (size_t k = 0; k < g.size(); k += AC_MAX_SCENARIOS) {
          std::vector<AcHostPlan*> lanes;
          std::vector<kg_estimate*> outs;
          for (size_t l = k; l < std::min(g.size(), k + AC_MAX_SCENARIOS); ++l) {
            lanes.push_back(&h[g[l] - c0]);
            outs.push_back(&out[g[l]]);
          }
          price_lane_set(ctx, s, lanes, outs);
        }
      }
    }
  }
}

extern "C" kg_status kg_autocallable_price_many(kg_context* ctx, const kg_settings* settings,
                                                const kg_autocall_terms* notes, size_t n,
                                                const kg_bsm_market* market, kg_estimate* out) {
  return guarded(ctx, [&] { price_many(ctx, settings, notes, n, market, out, false); });
}

extern "C" kg_status kg_autocallable_price_many_on_shared_grid(kg_context* ctx, const kg_settings* settings,
                                                               const kg_autocall_terms* notes, size_t n,
                                                               const kg_bsm_market* market, kg_estimate* out) {
  return guarded(ctx, [&] { price_many(ctx, settings, notes, n, market, out, true); });
}
