// Device building blocks shared by the Monte Carlo kernels (gfx950, wave64).
//
// Generators restate kiyosi 0.23.0 bit for bit:
//   Pcg32            crates/kiyosi/src/engines/monte_carlo.rs:2224-2277
//   Box-Muller       monte_carlo.rs:1830-1894 (sin_cos_turns polynomial)
//   Cephes ndtri     engines/monte_carlo/sobol.rs:249-346
//   Moments          monte_carlo.rs:2155-2222 (merge formula, applied as a tree)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kg {

constexpr int BLOCK_SAMPLES = 256;  // monte_carlo.rs:20
constexpr int WAVE = 64;
// Threads of every kernel that merges per-block partials (merge_run); one
// layout everywhere keeps the summation order, and so the result bits, the
// same.
constexpr int MERGE_THREADS = 256;

// True when `p` holds in some active lane of the wave. The builtin takes the
// predicate as a lane mask, so no VGPR copy of it is made (HIP's __ballot
// takes an int and materialises one).
__device__ inline bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }
// True when p holds on every active lane of the wave.
__device__ inline bool wave_all(bool p) { return __builtin_amdgcn_ballot_w64(!p) == 0; }

// ------------------------------------------------------------------ PCG32
struct Pcg32 {
  uint64_t state, inc;
  uint64_t inc2;  // increment of two LCG steps: inc * (MUL + 1)
  static constexpr uint64_t MUL = 6364136223846793005ULL;
  static constexpr uint64_t MUL2 = MUL * MUL;  // multiplier of two steps (mod 2^64)
  __host__ __device__ inline void init(uint64_t seed, uint64_t stream) {
    state = 0;
    inc = (stream << 1) | 1;
    inc2 = inc * (MUL + 1);
    step();
    state += seed;
    step();
  }
  __host__ __device__ inline void step() { state = state * MUL + inc; }
  __host__ __device__ static inline uint32_t output(uint64_t old) {
    uint32_t x = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t r = (uint32_t)(old >> 59);
    return (x >> r) | (x << ((32u - r) & 31u));
  }
  __host__ __device__ inline uint32_t next() {
    uint64_t old = state;
    state = old * MUL + inc;
    return output(old);
  }
  // The next output, then one more output skipped: two LCG steps in one
  // affine jump (same stream position as two next() calls).
  __host__ __device__ inline uint32_t next_skip_one() {
    uint64_t old = state;
    state = old * MUL2 + inc2;
    return output(old);
  }
  // O(log n) affine jump of the LCG (monte_carlo.rs:2233-2245).
  __host__ __device__ inline void advance(uint64_t distance) {
    uint64_t m = MUL, c = inc;
    while (distance) {
      if (distance & 1) state = state * m + c;
      c = c * (m + 1);
      m = m * m;
      distance >>= 1;
    }
  }
  // 52-bit uniform strictly inside (0, 1) (monte_carlo.rs:2272-2276).
  __device__ inline double uniform() {
    uint64_t hi = next();
    uint64_t lo = next();
    uint64_t bits = (hi << 20) | (lo >> 12);
    return ((double)bits + 0.5) * 0x1p-52;
  }
};

// ------------------------------------------------------------------ Box-Muller
// Sine and cosine of 2*pi*turns by quadrant reduction in turns and Taylor
// polynomials of degree 15/16 on [-pi/4, pi/4] (monte_carlo.rs:1852-1894).
__device__ inline void sin_cos_turns(double turns, double* s, double* c) {
  double q = turns * 4.0 + 0.5;
  uint32_t quadrant = (uint32_t)q;  // turns lies in (0, 1): q in (0.5, 4.5)
  double angle = 6.283185307179586 * (turns - 0.25 * (double)quadrant);
  double sq = angle * angle;
  double ps = -1.0 / 1307674368000.0;
  ps = ps * sq + 1.0 / 6227020800.0;
  ps = ps * sq + -1.0 / 39916800.0;
  ps = ps * sq + 1.0 / 362880.0;
  ps = ps * sq + -1.0 / 5040.0;
  ps = ps * sq + 1.0 / 120.0;
  ps = ps * sq + -1.0 / 6.0;
  double pc = 1.0 / 20922789888000.0;
  pc = pc * sq + -1.0 / 87178291200.0;
  pc = pc * sq + 1.0 / 479001600.0;
  pc = pc * sq + -1.0 / 3628800.0;
  pc = pc * sq + 1.0 / 40320.0;
  pc = pc * sq + -1.0 / 720.0;
  pc = pc * sq + 1.0 / 24.0;
  pc = pc * sq + -0.5;
  double sn = angle + angle * sq * ps;
  double cs = 1.0 + sq * pc;
  if (quadrant & 1) {
    double t = sn;
    sn = cs;
    cs = -t;
  }
  if (quadrant & 2) {
    sn = -sn;
    cs = -cs;
  }
  *s = sn;
  *c = cs;
}

// One Box-Muller pair: normal 2k is r*cos, normal 2k+1 is r*sin.
__device__ inline void box_muller_pair(Pcg32& g, double* even, double* odd) {
  double radius = sqrt(-2.0 * log(g.uniform()));
  double s, c;
  sin_cos_turns(g.uniform(), &s, &c);
  *even = radius * c;
  *odd = radius * s;
}

// Stream of normals for one sample: normal index k of a block stream comes
// from Box-Muller pair k/2, i.e. PCG outputs 4*(k/2) .. 4*(k/2)+3
// (monte_carlo.rs:1689-1703, 1830-1845). Positioning is by skip-ahead, so a
// thread can start at any sample of the block.
// FP64 normals of a PCG stream, Box-Muller pairs in stream order. They
// queue up to three deep so that whenever some lane of the wave needs a new
// pair every lane draws its next pair on the same call: lanes whose stream
// starts on an odd normal stay in step with the others instead of diverging
// on every call. The values and their order are those of the stream.
struct NormalStream {
  Pcg32 g;
  double q[3];
  int avail;
  __device__ inline void start(uint64_t seed, uint64_t block, uint64_t first_normal) {
    g.init(seed, block);
    g.advance((first_normal >> 1) * 4);
    avail = 0;
    if (first_normal & 1) {
      double e;
      box_muller_pair(g, &e, &q[0]);
      avail = 1;
    }
  }
  __device__ inline double next() {
    // Lanes consume in step, so at a refill every lane holds zero or one.
    if (__ballot(avail == 0) != 0 && avail <= 1) {
      double e, o;
      box_muller_pair(g, &e, &o);
      if (avail == 0) {
        q[0] = e;
        q[1] = o;
      } else {
        q[1] = e;
        q[2] = o;
      }
      avail += 2;
    }
    const double v = q[0];
    q[0] = q[1];
    q[1] = q[2];
    --avail;
    return v;
  }
  // Moves the stream forward by `count` normals without drawing them.
  __device__ inline void skip(uint64_t count) {
    for (; count > 0 && avail > 0; --count) {
      q[0] = q[1];
      q[1] = q[2];
      --avail;
    }
    g.advance((count >> 1) * 4);
    if (count & 1) {  // draw the pair and keep its second normal
      double e;
      box_muller_pair(g, &e, &q[0]);
      avail = 1;
    }
  }
};

// A load through the constant address space: tables that no kernel writes
// while it runs (per-step plans staged before the launch). With a uniform
// address the compiler then issues a scalar load instead of a vector load
// that waits on memory every step; a divergent address still loads per lane.
template <class T>
__device__ __forceinline__ T uload(const T* p) {
  return *(const __attribute__((address_space(4))) T*)p;
}
// Entry i of a byte table through the word that holds it (scalar memory
// loads are whole words); the table starts word-aligned and is staged in
// whole words.
__device__ __forceinline__ uint8_t uload_u8(const uint8_t* p, int i) {
  const uint32_t w = uload((const uint32_t*)(p + (i & ~3)));
  return (uint8_t)(w >> (8 * (i & 3)));
}

// ------------------------------------------------------------------ ndtri
// Cephes ndtri in FP64 (engines/monte_carlo/sobol.rs:262-346).
__device__ inline double inverse_normal_cdf(double p) {
  const double TAIL = 0.1353352832366127;
  bool upper = p > 1.0 - TAIL;
  double tail = upper ? 1.0 - p : p;
  if (tail > TAIL) {
    double c = tail - 0.5;
    double s = c * c;
    double num = -59.96335010141079;
    num = num * s + 98.00107541859997;
    num = num * s + -56.67628574690703;
    num = num * s + 13.931260938727968;
    num = num * s + -1.2391658386738125;
    double den = 1.0;
    den = den * s + 1.9544885833814176;
    den = den * s + 4.676279128988815;
    den = den * s + 86.36024213908905;
    den = den * s + -225.46268785411937;
    den = den * s + 200.26021238006066;
    den = den * s + -82.03722561683334;
    den = den * s + 15.90562251262117;
    den = den * s + -1.1833162112133;
    return (c + c * s * num / den) * 2.5066282746310002;
  }
  double root = sqrt(-2.0 * log(tail));
  double inv = 1.0 / root;
  double num, den;
  if (root < 8.0) {
    num = 4.0554489230596245;
    num = num * inv + 31.525109459989388;
    num = num * inv + 57.16281922464213;
    num = num * inv + 44.08050738932008;
    num = num * inv + 14.684956192885803;
    num = num * inv + 2.1866330685079025;
    num = num * inv + -0.1402560791713545;
    num = num * inv + -0.03504246268278482;
    num = num * inv + -0.0008574567851546854;
    den = 1.0;
    den = den * inv + 15.779988325646675;
    den = den * inv + 45.39076351288792;
    den = den * inv + 41.3172038254672;
    den = den * inv + 15.04253856929075;
    den = den * inv + 2.504649462083094;
    den = den * inv + -0.14218292285478779;
    den = den * inv + -0.03808064076915783;
    den = den * inv + -0.0009332594808954574;
  } else {
    num = 3.2377489177694603;
    num = num * inv + 6.915228890689842;
    num = num * inv + 3.9388102529247444;
    num = num * inv + 1.3330346081580755;
    num = num * inv + 0.20148538954917908;
    num = num * inv + 0.012371663481782003;
    num = num * inv + 0.00030158155350823543;
    num = num * inv + 2.6580697468673755e-06;
    num = num * inv + 6.239745391849833e-09;
    den = 1.0;
    den = den * inv + 6.02427039364742;
    den = den * inv + 3.6798356385616087;
    den = den * inv + 1.3770209948908132;
    den = den * inv + 0.21623699359449663;
    den = den * inv + 0.013420400608854318;
    den = den * inv + 0.00032801446468212774;
    den = den * inv + 2.8924786474538068e-06;
    den = den * inv + 6.790194080099813e-09;
  }
  double q = root - log(root) / root - inv * (num / den);
  return upper ? q : -q;
}
__device__ inline double normal_from_bits(uint64_t bits) {
  return inverse_normal_cdf(((double)bits + 0.5) * 0x1p-52);
}

// ------------------------------------------------------------------ Moments
// Scale-normalised count/mean/M2 with the reference's add and merge rules.
struct Moments {
  double count;  // exact below 2^53
  double scale, mean, m2;
};

__device__ inline Moments moments_zero() { return Moments{0.0, 0.0, 0.0, 0.0}; }

// The moment updates round every product and sum on its own, as the
// reference's Rust does (no contraction into FMAs): the result is then the
// same in every kernel that inlines them.
__device__ inline void moments_add_general(Moments& m, double value, double a, bool variance) {
#pragma clang fp contract(off)
  if (a > m.scale) {
    double ratio = m.scale / a;
    m.mean *= ratio;
    if (variance) m.m2 *= ratio * ratio;
    m.scale = a;
  }
  m.count += 1.0;
  double v = m.scale == 0.0 ? 0.0 : value / m.scale;
  double delta = v - m.mean;
  m.mean += delta / m.count;
  if (variance) m.m2 += delta * (v - m.mean);
}

__device__ inline void moments_add(Moments& m, double value, bool variance) {
#pragma clang fp contract(off)
  double a = fabs(value);
  // The first finite nonzero value into empty moments: the general update
  // gives exactly {1, |value|, +-1, 0} (ratio 0, v = value / |value| = +-1
  // exactly, mean v, M2 v * (v - v) = 0), here without its divisions.
  const bool first = m.count == 0.0 && m.scale == 0.0 && a > 0.0 && a <= 1.7976931348623157e308;
  if (first) {
    m = Moments{1.0, a, value > 0.0 ? 1.0 : -1.0, 0.0};
  } else {
    moments_add_general(m, value, a, variance);
  }
}

__device__ inline Moments moments_merge(const Moments& a, const Moments& b) {
#pragma clang fp contract(off)
  if (a.count == 0.0) return b;
  if (b.count == 0.0) return a;
  double sc = fmax(a.scale, b.scale);
  // left = a.scale / sc and right = b.scale / sc. For a finite nonzero sc
  // the operand that equals sc gives exactly 1, so one quotient serves; and
  // equal counts give wb = b.count / n = 1/2 exactly. The results are the
  // plain quotients, bit for bit.
  double left, right;
  if (sc == 0.0) {
    left = right = 0.0;
  } else if (sc <= 1.7976931348623157e308) {
    const bool a_top = a.scale == sc;
    const double q = (a_top ? b.scale : a.scale) / sc;
    left = a_top ? 1.0 : q;
    right = a_top ? q : 1.0;
  } else {
    left = a.scale / sc;
    right = b.scale / sc;
  }
  double delta = b.mean * right - a.mean * left;
  double n = a.count + b.count;
  double wb;
  if (wave_all(a.count == b.count))
    wb = 0.5;
  else
    wb = b.count / n;
  Moments r;
  r.mean = a.mean * left + delta * wb;
  r.m2 = a.m2 * left * left + b.m2 * right * right + delta * delta * (a.count * wb);
  r.count = n;
  r.scale = sc;
  return r;
}

// moments_merge for operands that both hold values (count > 0): the same
// arithmetic, bit for bit, without the empty-operand branches, so that
// independent merges can interleave.
__device__ inline Moments moments_merge_full(const Moments& a, const Moments& b) {
#pragma clang fp contract(off)
  const double sc = fmax(a.scale, b.scale);
  // As in moments_merge: one quotient, of the operand below sc; x / x is 1
  // for a finite nonzero sc and NaN for an infinite or NaN one.
  const bool a_top = a.scale == sc;
  const double q = (a_top ? b.scale : a.scale) / sc;
  const double top = sc <= 1.7976931348623157e308 ? 1.0 : __builtin_nan("");
  const double left = sc == 0.0 ? 0.0 : (a_top ? top : q);
  const double right = sc == 0.0 ? 0.0 : (a_top ? q : top);
  const double delta = b.mean * right - a.mean * left;
  const double n = a.count + b.count;
  const double wb = b.count / n;
  Moments r;
  r.mean = a.mean * left + delta * wb;
  r.m2 = a.m2 * left * left + b.m2 * right * right + delta * delta * (a.count * wb);
  r.count = n;
  r.scale = sc;
  return r;
}

__device__ inline Moments moments_shfl_xor(const Moments& m, int mask) {
  Moments r;
  r.count = __shfl_xor(m.count, mask, WAVE);
  r.scale = __shfl_xor(m.scale, mask, WAVE);
  r.mean = __shfl_xor(m.mean, mask, WAVE);
  r.m2 = __shfl_xor(m.m2, mask, WAVE);
  return r;
}

// Merges the Moments of lanes 0..lanes-1 of a wave (lanes a power of two);
// the lower lane of each pair is the left operand, so the order is fixed by
// lane index.
__device__ inline Moments wave_merge(Moments m, int lanes = WAVE) {
  const int lane = threadIdx.x & (WAVE - 1);
  for (int mask = 1; mask < lanes; mask <<= 1) {
    Moments o = moments_shfl_xor(m, mask);
    m = (lane & mask) ? moments_merge(o, m) : moments_merge(m, o);
  }
  return m;
}

// Block-wide merge of N independent Moments sets; returns the result in
// thread 0. Each wave merges its lanes, then wave 0 merges the waves' results
// the same way. `smem` needs N * (blockDim.x / 64) entries.
template <int N>
__device__ inline void block_merge(Moments (&m)[N], Moments* smem) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = threadIdx.x / WAVE;
  const int waves = blockDim.x / WAVE;
#pragma unroll
  for (int i = 0; i < N; ++i) m[i] = wave_merge(m[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) smem[i * waves + wave] = m[i];
  __syncthreads();
  if (wave == 0 && waves > 1) {
    int lanes = 1;
    while (lanes < waves) lanes <<= 1;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      Moments acc = lane < waves ? smem[i * waves + lane] : moments_zero();
      acc = wave_merge(acc, lanes);
      if (lane == 0) m[i] = acc;
    }
  }
  __syncthreads();
}

// Lane l's copy of v from the lane that the DPP pattern CTRL names.
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const uint32_t hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (uint64_t)hi << 32 | lo);
}

__device__ inline double read_lane(double v, int lane) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, (uint64_t)hi << 32 | lo);
}

// Sum (with MAX the largest) of v over the wave, the same bits in every
// lane: four pairwise levels inside each 16-lane row by DPP (lanes l ^ 1 and
// l ^ 2, then the half-row and row mirrors; a + b and b + a are the same
// bits, so both lanes of a pair hold the same value), then the four rows'
// results in row order.
template <bool MAX>
__device__ inline double wave_reduce(double v) {
#pragma clang fp contract(off)
  auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
  v = op(v, dpp_move<0xB1>(v));   // quad_perm [1, 0, 3, 2]
  v = op(v, dpp_move<0x4E>(v));   // quad_perm [2, 3, 0, 1]
  v = op(v, dpp_move<0x141>(v));  // row_half_mirror
  v = op(v, dpp_move<0x140>(v));  // row_mirror
  return op(op(read_lane(v, 0), read_lane(v, 16)), op(read_lane(v, 32), read_lane(v, 48)));
}

// The waves' results of one workgroup-wide reduction, folded in wave order
// in every thread. `slot` holds THREADS / 64 doubles used by no other
// reduction in flight.
template <bool MAX, int THREADS>
__device__ inline double across_waves(double w, double* slot) {
#pragma clang fp contract(off)
  if ((threadIdx.x & (WAVE - 1)) == 0) slot[threadIdx.x / WAVE] = w;
  __syncthreads();
  double r = slot[0];
#pragma unroll
  for (int i = 1; i < THREADS / WAVE; ++i) r = MAX ? fmax(r, slot[i]) : r + slot[i];
  return r;
}

// The partials p[b * stride], b in [0, blocks), merged into one set: the
// reference's merge (Moments::merge, monte_carlo.rs:2192-2222) in its
// parallel form. With S the largest scale and r_b = scale_b / S, the set has
// count N = sum n_b, scale S, mean sum(n_b mean_b r_b) / N and
// M2 sum(m2_b r_b^2 + n_b (mean_b r_b - mean)^2): Chan's update with each
// block's deviation taken from the final mean, so M2 still comes from
// deviations, while the chain of dependent pairwise merges (each with its
// divisions) becomes three workgroup sums. Thread t takes blocks
// t, t + THREADS, ... in that order (held in registers when there are at most
// CACHED per thread), so the result depends only on the partials. Called by
// all THREADS threads; every thread returns the result. `smem`: 4 * THREADS
// / 64 doubles.
template <int THREADS = MERGE_THREADS, int CACHED = 16>
__device__ inline Moments merge_run(const Moments* __restrict__ p, int64_t stride, int64_t blocks,
                                    double* smem) {
#pragma clang fp contract(off)
  constexpr int W = THREADS / WAVE;
  const bool cached = blocks <= (int64_t)CACHED * THREADS;
  Moments v[CACHED];
  if (cached) {
#pragma unroll
    for (int k = 0; k < CACHED; ++k) {
      const int64_t b = threadIdx.x + (int64_t)k * THREADS;
      v[k] = b < blocks ? p[b * stride] : moments_zero();
    }
  }
  // f over this thread's blocks in order. Held in registers, the entries
  // past the last block are empty sets, whose terms below are exact zeros.
  auto each = [&](auto&& f) {
    if (cached) {
#pragma unroll
      for (int k = 0; k < CACHED; ++k) f(v[k]);
    } else {
      for (int64_t b = threadIdx.x; b < blocks; b += THREADS) f(p[b * stride]);
    }
  };
  double n = 0.0, top = 0.0;
  each([&](const Moments& m) {
    n += m.count;
    top = fmax(top, m.scale);
  });
  const double count = across_waves<false, THREADS>(wave_reduce<false>(n), smem);
  const double scale = across_waves<true, THREADS>(wave_reduce<true>(top), smem + W);
  if (count == 0.0) return moments_zero();
  // r_b is 1 for the block that holds S, and 0 for an empty block or when
  // every value is zero (S = 0).
  double t = 0.0;
  each([&](const Moments& m) {
    const double r = scale == 0.0 ? 0.0 : m.scale / scale;
    t += m.count * (m.mean * r);
  });
  const double mean = across_waves<false, THREADS>(wave_reduce<false>(t), smem + 2 * W) / count;
  double q = 0.0;
  each([&](const Moments& m) {
    const double r = scale == 0.0 ? 0.0 : m.scale / scale;
    const double d = m.mean * r - mean;
    q += m.m2 * r * r + m.count * (d * d);
  });
  Moments out;
  out.count = count;
  out.scale = scale;
  out.mean = mean;
  out.m2 = across_waves<false, THREADS>(wave_reduce<false>(q), smem + 3 * W);
  return out;
}

// Moments of the values a block's threads hold (at most K per thread and
// output; n[i] of them valid): each thread adds its values in order with the
// reference's scale-normalised Welford update (Moments::add,
// monte_carlo.rs:2163-2190), and the block's threads are then merged with
// the Chan tree of block_merge (Moments::merge, :2192-2222) in thread order.
// M2 is thus formed from deviations, never from raw sums of squares, so it
// keeps its relative precision for payoffs whose spread is tiny next to
// their level. The result is in thread 0; `smem` needs N * (blockDim.x / 64)
// entries.
template <int N, int K>
__device__ inline void block_moments(const double (&v)[N][K], const int (&n)[N], Moments (&out)[N],
                                     Moments* smem) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    Moments m = moments_zero();
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (k < n[i]) moments_add(m, v[i][k], true);
    out[i] = m;
  }
  block_merge<N>(out, smem);
}

// f64::midpoint for finite inputs well inside the exponent range.
__device__ inline double midpoint(double a, double b) {
  const double HI = 8.98846567431158e307;  // f64::MAX / 2
  if (fabs(a) <= HI && fabs(b) <= HI) return (a + b) / 2.0;
  return a / 2.0 + b / 2.0;
}

// discounted_product (crates/kiyosi/src/numerics.rs:226-238).
__device__ inline double discounted_product(double scale, double weight, double log_discount) {
  double value = scale * weight;
  double discount = exp(log_discount);
  if (scale == 0.0 || weight == 0.0) return 0.0;
  if (isnormal(discount) && isnormal(value)) return value * discount;
  return copysign(exp(log(scale) + log(fabs(weight)) + log_discount), weight);
}

// Flags raised by kernels and read back by the host.
enum : uint32_t { FLAG_NONFINITE = 1u, FLAG_CANCELLED = 2u };

__device__ inline void raise_flag(uint32_t* flags, uint32_t bit) { atomicOr(flags, bit); }

}  // namespace kg
