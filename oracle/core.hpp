// Test infrastructure, not product code.
//
// CPU FP64 restatement of the kiyosi 0.23.0 Monte Carlo sampler, used only as
// the checker of the HIP engine: tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it; the library never does. Compiled with
// -ffp-contract=off so that every a*b+c rounds twice, as Rust does, except
// where the reference calls mul_add (written here as std::fma).
//
// Citations are to crates/kiyosi/src/engines/monte_carlo.rs unless a file is
// named.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace oracle {

constexpr size_t BLOCK = 256;  // monte_carlo.rs:20

struct Failure : std::runtime_error {
  int code;
  Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
constexpr int INVALID_INPUT = 1, UNSUPPORTED = 2, COMPUTATION_FAILURE = 3, MISSING_HISTORY = 6,
              NON_CONVERGENCE = 7;
// scaled_cash (numerics.rs:175-186)
inline double scaled_cash(double amount, double exponent) {
  if (amount == 0.0) return 0.0;
  const double discount = std::exp(exponent), value = amount * discount;
  if (std::isnormal(discount) && std::isnormal(value)) return value;
  return std::exp(std::log(amount) + exponent);
}
[[noreturn]] inline void unstable() {  // monte_carlo.rs:2119-2121
  throw Failure(COMPUTATION_FAILURE,
                "Monte Carlo produced a non-finite parameter, path or result");
}

// ---------------------------------------------------------------- PCG32
// Pcg32 (monte_carlo.rs:2224-2277): PCG-XSH-RR 64/32.
struct Pcg32 {
  uint64_t state = 0, inc = 1;
  static constexpr uint64_t MUL = 6364136223846793005ULL;
  Pcg32() = default;
  Pcg32(uint64_t seed, uint64_t stream) {  // :2247-2256
    state = 0;
    inc = (stream << 1) | 1;
    next();
    state += seed;
    next();
  }
  uint32_t next() {  // :2261-2267
    uint64_t old = state;
    state = old * MUL + inc;
    uint32_t x = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t r = (uint32_t)(old >> 59);
    return (x >> r) | (x << ((32 - r) & 31));
  }
  void advance(uint64_t distance) {  // :2233-2245
    uint64_t m = MUL, c = inc;
    while (distance) {
      if (distance & 1) state = state * m + c;
      c = c * (m + 1);
      m = m * m;
      distance >>= 1;
    }
  }
  double uniform() {  // :2272-2276
    uint64_t hi = next();
    uint64_t lo = next();
    uint64_t bits = (hi << 20) | (lo >> 12);
    return ((double)bits + 0.5) / 4503599627370496.0;
  }
};

// ---------------------------------------------------------------- Box-Muller
// sin_cos_turns (monte_carlo.rs:1852-1894): quadrant reduction in turns and
// Taylor polynomials evaluated by Horner without fused multiply-add.
inline void sin_cos_turns(double turns, double* s, double* c) {
  static const double SINE[7] = {-1.0 / 1307674368000.0, 1.0 / 6227020800.0,
                                 -1.0 / 39916800.0,      1.0 / 362880.0,
                                 -1.0 / 5040.0,          1.0 / 120.0,
                                 -1.0 / 6.0};
  static const double COSINE[8] = {1.0 / 20922789888000.0, -1.0 / 87178291200.0,
                                   1.0 / 479001600.0,      -1.0 / 3628800.0,
                                   1.0 / 40320.0,          -1.0 / 720.0,
                                   1.0 / 24.0,             -0.5};
  const double TAU = 6.283185307179586;
  double q = turns * 4.0 + 0.5;
  uint32_t quadrant = q <= 0.0 ? 0u : (q >= 4294967295.0 ? 0xffffffffu : (uint32_t)q);
  double angle = TAU * (turns - 0.25 * (double)quadrant);
  double sq = angle * angle;
  double ps = 0.0, pc = 0.0;
  for (double k : SINE) ps = ps * sq + k;
  for (double k : COSINE) pc = pc * sq + k;
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

// ---------------------------------------------------------------- ndtri
// inverse_normal_cdf (monte_carlo/sobol.rs:262-346): Cephes ndtri.
inline double poly(double v, const double* k, int n, double lead) {
  double r = lead;
  for (int i = 0; i < n; ++i) r = r * v + k[i];
  return r;
}
inline double inverse_normal_cdf(double p) {
  static const double P0[5] = {-59.96335010141079, 98.00107541859997, -56.67628574690703,
                               13.931260938727968, -1.2391658386738125};
  static const double Q0[8] = {1.9544885833814176, 4.676279128988815,  86.36024213908905,
                               -225.46268785411937, 200.26021238006066, -82.03722561683334,
                               15.90562251262117,  -1.1833162112133};
  static const double P1[9] = {4.0554489230596245, 31.525109459989388, 57.16281922464213,
                               44.08050738932008,  14.684956192885803, 2.1866330685079025,
                               -0.1402560791713545, -0.03504246268278482,
                               -0.0008574567851546854};
  static const double Q1[8] = {15.779988325646675, 45.39076351288792,  41.3172038254672,
                               15.04253856929075,  2.504649462083094,  -0.14218292285478779,
                               -0.03808064076915783, -0.0009332594808954574};
  static const double P2[9] = {3.2377489177694603, 6.915228890689842, 3.9388102529247444,
                               1.3330346081580755, 0.20148538954917908, 0.012371663481782003,
                               0.00030158155350823543, 2.6580697468673755e-06,
                               6.239745391849833e-09};
  static const double Q2[8] = {6.02427039364742,      3.6798356385616087,   1.3770209948908132,
                               0.21623699359449663,   0.013420400608854318, 0.00032801446468212774,
                               2.8924786474538068e-06, 6.790194080099813e-09};
  const double TAIL = 0.1353352832366127;
  bool upper = p > 1.0 - TAIL;
  double tail = upper ? 1.0 - p : p;
  if (tail > TAIL) {
    double c = tail - 0.5;
    double s = c * c;
    return (c + c * s * poly(s, P0, 5, 0.0) / poly(s, Q0, 8, 1.0)) * std::sqrt(2.0 * M_PI);
  }
  double root = std::sqrt(-2.0 * std::log(tail));
  double inv = 1.0 / root;
  double corr = inv * (root < 8.0 ? poly(inv, P1, 9, 0.0) / poly(inv, Q1, 8, 1.0)
                                  : poly(inv, P2, 9, 0.0) / poly(inv, Q2, 8, 1.0));
  double q = root - std::log(root) / root - corr;
  return upper ? q : -q;
}
inline double normal_from_bits(uint64_t bits) {  // sobol.rs:249-252
  return inverse_normal_cdf(((double)bits + 0.5) / 4503599627370496.0);
}

// ---------------------------------------------------------------- Sobol
// SobolSequence::new (sobol.rs:85-134): Joe-Kuo directions with 52 bits,
// linear matrix scrambling and a digital shift seeded by Pcg32(seed, dim).
constexpr int SOBOL_BITS = 52;
constexpr uint64_t SOBOL_MASK = (1ULL << SOBOL_BITS) - 1;

struct SobolSequence {
  size_t dims = 0;
  std::vector<uint64_t> directions;  // [column * dims + dim]
  std::vector<uint64_t> shifts;
};

inline SobolSequence sobol_new(size_t dims, uint64_t seed, const uint32_t* table,
                               size_t table_len) {
  if (dims < 1 || dims > 21201)
    throw Failure(INVALID_INPUT, "Sobol dimension must be in 1..=21,201");
  SobolSequence s;
  s.dims = dims;
  s.directions.assign(dims * SOBOL_BITS, 0);
  s.shifts.resize(dims);
  size_t pos = 0;
  for (size_t d = 0; d < dims; ++d) {
    uint64_t raw[SOBOL_BITS] = {};
    if (d == 0) {
      for (int c = 0; c < SOBOL_BITS; ++c) raw[c] = 1ULL << (SOBOL_BITS - c - 1);
    } else {
      if (pos + 2 > table_len) throw Failure(INVALID_INPUT, "Sobol table truncated");
      int degree = (int)table[pos];
      uint32_t coeff = table[pos + 1];
      for (int c = 0; c < degree && c < SOBOL_BITS; ++c)
        raw[c] = (uint64_t)table[pos + 2 + c] << (SOBOL_BITS - c - 1);
      pos += degree + 2;
      for (int c = degree; c < SOBOL_BITS; ++c) {
        uint64_t v = raw[c - degree] ^ (raw[c - degree] >> degree);
        for (int o = 1; o < degree; ++o)
          if ((coeff >> (degree - o - 1)) & 1) v ^= raw[c - o];
        raw[c] = v;
      }
    }
    Pcg32 g(seed, d);
    auto word = [&]() {
      uint64_t hi = g.next();
      uint64_t lo = g.next();
      return (hi << 32) | lo;
    };
    s.shifts[d] = word() & SOBOL_MASK;
    uint64_t matrix[SOBOL_BITS];
    for (int row = 0; row < SOBOL_BITS; ++row) {
      int bit = SOBOL_BITS - row - 1;
      matrix[row] = (word() & (SOBOL_MASK << (bit + 1))) | (1ULL << bit);
    }
    for (int c = 0; c < SOBOL_BITS; ++c) {
      uint64_t sc = 0;
      for (int row = 0; row < SOBOL_BITS; ++row)
        sc |= (uint64_t)(__builtin_popcountll(raw[c] & matrix[row]) & 1) << (SOBOL_BITS - row - 1);
      s.directions[(size_t)c * dims + d] = sc;
    }
  }
  return s;
}

// SobolCursor (sobol.rs:136-243), eager form: the deferred-group form yields
// the same bits (sobol.rs:366-385).
struct SobolCursor {
  const SobolSequence* seq = nullptr;
  std::vector<uint64_t> values;
  uint64_t index = 0;
  size_t dimension = 0;
  SobolCursor() = default;
  SobolCursor(const SobolSequence* s, uint64_t idx) : seq(s), values(s->shifts), index(idx) {
    uint64_t gray = idx ^ (idx >> 1);
    while (gray) {
      int c = __builtin_ctzll(gray);
      for (size_t d = 0; d < seq->dims; ++d) values[d] ^= seq->directions[(size_t)c * seq->dims + d];
      gray &= gray - 1;
    }
  }
  double normal() { return normal_from_bits(values[dimension++]); }
  void next_path() {
    index += 1;
    dimension = 0;
    int c = __builtin_ctzll(index);
    for (size_t d = 0; d < seq->dims; ++d) values[d] ^= seq->directions[(size_t)c * seq->dims + d];
  }
  void seek(size_t d) { dimension = d; }
};

// ---------------------------------------------------------------- draws
// NormalDraws + sample_normal (monte_carlo.rs:1678-1703, 1830-1845).
struct NormalDraws {
  Pcg32 gen;
  bool has_spare = false;
  double spare = 0.0;
  SobolCursor* cursor = nullptr;
  double normal() {
    if (cursor) return cursor->normal();
    if (has_spare) {
      has_spare = false;
      return spare;
    }
    double radius = std::sqrt(-2.0 * std::log(gen.uniform()));
    double s, c;
    sin_cos_turns(gen.uniform(), &s, &c);
    spare = radius * s;
    has_spare = true;
    return radius * c;
  }
  void skip_normals(size_t remaining) {  // :1689-1703
    if (remaining == 0 || cursor) return;
    if (has_spare) {
      has_spare = false;
      remaining -= 1;
    }
    gen.advance((remaining / 2) * 4);
    if (remaining % 2) normal();
  }
};

// ---------------------------------------------------------------- Moments
// Moments (monte_carlo.rs:2155-2222): scale-normalised Welford mean and M2.
struct Moments {
  uint64_t count = 0;
  double scale = 0.0, mean = 0.0, m2 = 0.0;
  void add(double value, bool variance) {
    if (std::fabs(value) > scale) {
      double ratio = scale / std::fabs(value);
      mean *= ratio;
      if (variance) m2 *= ratio * ratio;
      scale = std::fabs(value);
    }
    count += 1;
    double v = scale == 0.0 ? 0.0 : value / scale;
    double delta = v - mean;
    mean += delta / (double)count;
    if (variance) m2 += delta * (v - mean);
  }
  void merge(const Moments& o, bool variance) {
    if (count == 0) {
      *this = o;
      return;
    }
    double sc = std::max(scale, o.scale);
    double left = sc == 0.0 ? 0.0 : scale / sc;
    double right = sc == 0.0 ? 0.0 : o.scale / sc;
    double delta = o.mean * right - mean * left;
    uint64_t n = count + o.count;
    mean = mean * left + delta * ((double)o.count / (double)n);
    if (variance)
      m2 = m2 * left * left + o.m2 * right * right +
           delta * delta * ((double)count * ((double)o.count / (double)n));
    count = n;
    scale = sc;
  }
};

// f64::midpoint as used for antithetic pair means (monte_carlo.rs:1793).
inline double midpoint(double a, double b) {
  const double LO = std::numeric_limits<double>::min() * 2.0;
  const double HI = std::numeric_limits<double>::max() / 2.0;
  double aa = std::fabs(a), ab = std::fabs(b);
  if (aa <= HI && ab <= HI) return (a + b) / 2.0;
  if (aa < LO) return a + b / 2.0;
  if (ab < LO) return a / 2.0 + b;
  return a / 2.0 + b / 2.0;
}

// numerics.rs:226-238
inline double discounted_product(double scale, double weight, double log_discount) {
  double value = scale * weight;
  double discount = std::exp(log_discount);
  if (scale == 0.0 || weight == 0.0) return 0.0;
  if (std::isnormal(discount) && std::isnormal(value)) return value * discount;
  return std::copysign(std::exp(std::log(scale) + std::log(std::fabs(weight)) + log_discount), weight);
}
inline double scaled_value(double scale, double weight, double log_discount) {  // :2147-2153
  double r = discounted_product(scale, weight, log_discount);
  if (!std::isfinite(r)) unstable();
  return r;
}
inline double log_ratio(double num, double den) {  // numerics.rs:80-92
  double rel = (num - den) / den;
  if (std::fabs(rel) < 0.5) return std::log1p(rel);
  double ratio = num / den;
  if (std::isnormal(ratio)) return std::log(ratio);
  return std::log(num) - std::log(den);
}

// ---------------------------------------------------------------- settings
enum Sampling { PSEUDO_RANDOM = 0, SOBOL = 1 };
struct Settings {
  int64_t path_count = 131072;
  uint64_t seed = 0;
  int sampling = PSEUDO_RANDOM;
  bool antithetic = false;
  int workers = 1;
  size_t sample_count() const {
    return antithetic ? (size_t)((path_count + 1) / 2) : (size_t)path_count;
  }
  size_t blocks() const { return (sample_count() + BLOCK - 1) / BLOCK; }
};

struct Context {  // Sobol setup shared by every block of one request
  const uint32_t* table = nullptr;
  size_t table_len = 0;
};

// Contiguous block chunks per worker (BlockRunner, monte_carlo.rs:1925-1997).
template <class F>
void run_blocks(size_t blocks, int workers, F&& simulate) {
  size_t w = std::max(1, workers);
  w = std::min(w, std::max<size_t>(blocks, 1));
  if (w <= 1) {
    for (size_t b = 0; b < blocks; ++b) simulate(b);
    return;
  }
  size_t chunk = (blocks + w - 1) / w;
  std::vector<std::thread> threads;
  std::vector<std::exception_ptr> errors(w);
  for (size_t t = 0; t < w; ++t) {
    threads.emplace_back([&, t]() {
      try {
        for (size_t b = t * chunk; b < std::min(blocks, (t + 1) * chunk); ++b) simulate(b);
      } catch (...) {
        errors[t] = std::current_exception();
      }
    });
  }
  for (auto& th : threads) th.join();
  for (auto& e : errors)
    if (e) std::rethrow_exception(e);
}

// sample_paths (monte_carlo.rs:1761-1811). `sample(draws, normals, paired,
// pos, neg)` returns whether a negative leg was produced.
template <int N, class F>
void sample_paths(const Settings& st, size_t dims, bool statistics, size_t buffer_len,
                  const Context& cx, F&& sample, Moments (&total)[N], Moments (&pairs)[N]) {
  // The reference caps one request at 1e9 path-steps (:219-227). The oracle
  // does not, so that it can check the 2^22 x 252 Heston workload.
  SobolSequence seq;
  if (st.sampling == SOBOL) seq = sobol_new(dims, st.seed, cx.table, cx.table_len);
  bool stats = statistics && st.sampling == PSEUDO_RANDOM &&
               (!st.antithetic || st.path_count % 2 == 0);
  bool variance = stats && !st.antithetic;
  size_t nblocks = st.blocks();
  std::vector<Moments> part_m((size_t)N * nblocks), part_p((size_t)N * nblocks);
  size_t samples = st.sample_count();
  size_t paths = (size_t)st.path_count;
  run_blocks(nblocks, st.workers, [&](size_t b) {
    Moments* m = &part_m[b * N];
    Moments* p = &part_p[b * N];
    std::vector<double> normals(buffer_len);
    size_t first = b * BLOCK;
    SobolCursor cursor;
    NormalDraws draws;
    draws.gen = Pcg32(st.seed, b);
    if (st.sampling == SOBOL) {
      cursor = SobolCursor(&seq, first);
      draws.cursor = &cursor;
    }
    double pos[N], neg[N];
    for (size_t path = first; path < std::min(first + BLOCK, samples); ++path) {
      bool paired = st.antithetic && path < paths / 2;
      bool has_neg = sample(draws, normals.data(), paired, pos, neg);
      for (int i = 0; i < N; ++i) m[i].add(pos[i], variance);
      if (has_neg)
        for (int i = 0; i < N; ++i) {
          m[i].add(neg[i], variance);
          if (stats) p[i].add(midpoint(pos[i], neg[i]), true);
        }
      if (draws.cursor) cursor.next_path();
    }
  });
  for (int i = 0; i < N; ++i) total[i] = pairs[i] = Moments();
  for (size_t b = 0; b < nblocks; ++b)
    for (int i = 0; i < N; ++i) {
      total[i].merge(part_m[b * N + i], variance);
      if (stats && st.antithetic) pairs[i].merge(part_p[b * N + i], true);
    }
}

// path_samples (monte_carlo.rs:1498-1530): eager draws, negated replay.
template <int N, class P>
void path_samples(const Settings& st, size_t dims, bool statistics, const Context& cx, P&& payoff,
                  Moments (&total)[N], Moments (&pairs)[N]) {
  sample_paths<N>(
      st, dims, statistics, dims, cx,
      [&](NormalDraws& d, double* normals, bool paired, double* pos, double* neg) {
        for (size_t k = 0; k < dims; ++k) normals[k] = d.normal();
        payoff(normals, pos);
        if (!paired) return false;
        for (size_t k = 0; k < dims; ++k) normals[k] = -normals[k];
        payoff(normals, neg);
        return true;
      },
      total, pairs);
}

// LazyPath (monte_carlo.rs:1636-1676).
struct LazyPath {
  NormalDraws* draws;
  double* cache;  // null when unpaired
  size_t consumed = 0, generated = 0;
  bool opposite = false;
  double normal() {
    if (!cache) {
      generated += 1;
      return draws->normal();
    }
    double v;
    if (consumed < generated) {
      v = cache[consumed];
    } else {
      v = draws->normal();
      cache[generated] = v;
      generated += 1;
    }
    consumed += 1;
    return opposite ? -v : v;
  }
  void finish(size_t dims) {
    if (draws->cursor) return;  // deferred Sobol groups change no bits
    draws->skip_normals(dims - generated);
  }
};

// path_samples_lazy (monte_carlo.rs:1532-1565).
template <int N, class P>
void path_samples_lazy(const Settings& st, size_t dims, bool statistics, const Context& cx,
                       P&& payoff, Moments (&total)[N], Moments (&pairs)[N]) {
  sample_paths<N>(
      st, dims, statistics, st.antithetic ? dims : 0, cx,
      [&](NormalDraws& d, double* normals, bool paired, double* pos, double* neg) {
        LazyPath path{&d, paired ? normals : nullptr};
        payoff(path, pos);
        bool has_neg = false;
        if (paired) {
          path.consumed = 0;
          path.opposite = true;
          payoff(path, neg);
          has_neg = true;
        }
        path.finish(dims);
        return has_neg;
      },
      total, pairs);
}

// SkippingPath (monte_carlo.rs:1618-1634).
struct SkippingPath {
  NormalDraws* draws;
  size_t used = 0;
  double sign = 1.0;
  double normal() {
    used += 1;
    return sign * draws->normal();
  }
  void skip(size_t count) {
    used += count;
    if (draws->cursor)
      draws->cursor->seek(used);
    else
      draws->skip_normals(count);
  }
};

// path_samples_skipping + skipped_sample_pair (monte_carlo.rs:1567-1616).
template <int N, class P>
void path_samples_skipping(const Settings& st, size_t dims, bool statistics, const Context& cx,
                           P&& payoff, Moments (&total)[N], Moments (&pairs)[N]) {
  sample_paths<N>(
      st, dims, statistics, 0, cx,
      [&](NormalDraws& d, double*, bool paired, double* pos, double* neg) {
        Pcg32 origin_gen = d.gen;
        bool origin_has = d.has_spare;
        double origin_spare = d.spare;
        SkippingPath path{&d};
        payoff(path, pos);
        if (paired) {
          d.gen = origin_gen;
          d.has_spare = origin_has;
          d.spare = origin_spare;
          if (d.cursor) d.cursor->seek(0);
          path.used = 0;
          path.sign = -1.0;
          payoff(path, neg);
        }
        path.skip(dims - path.used);
        return paired;
      },
      total, pairs);
}

// ---------------------------------------------------------------- estimate
enum UncertaintyStatus {
  AVAILABLE = 0,
  UNSUPPORTED_SAMPLING = 1,
  INSUFFICIENT_OBSERVATIONS = 2,
  INCOMPLETE_ANTITHETIC_PAIR = 3,
  ZERO_OBSERVED_VARIANCE = 4,
};

struct Estimate {
  double price = 0.0;
  double standard_error = 0.0;
  int has_standard_error = 0;
  int status = INSUFFICIENT_OBSERVATIONS;
  uint64_t seed = 0;
  int has_seed = 0;
  // Diagnostic beside the reference's fields: the standard error of the
  // per-path values as if they were independent (both antithetic legs).
  double path_standard_error = 0.0;
};

// sampling_standard_error (monte_carlo.rs:1426-1452).
inline bool sampling_standard_error(const Settings& st, const Moments& sample,
                                    const Moments& pairs, double log_discount, double* out) {
  const Moments& u = st.antithetic ? pairs : sample;
  if (st.sampling == SOBOL || u.count < 2 || (st.antithetic && st.path_count % 2 != 0))
    return false;
  double dev = std::sqrt(std::max(u.m2, 0.0) / (double)(u.count - 1) / (double)u.count);
  *out = scaled_value(u.scale, dev, log_discount);
  return true;
}

// estimate + MonteCarloEstimate::from_sampling (monte_carlo.rs:1412-1424, 938-960).
inline Estimate estimate(const Settings& st, const Moments& sample, const Moments& pairs,
                         double log_discount, bool known, bool has_seed, uint64_t seed) {
  Estimate e;
  e.price = scaled_value(sample.scale, sample.mean, log_discount);
  e.has_seed = has_seed;
  e.seed = seed;
  if (sample.count >= 2) {
    double dev = std::sqrt(std::max(sample.m2, 0.0) / (double)(sample.count - 1) / (double)sample.count);
    e.path_standard_error = discounted_product(sample.scale, dev, log_discount);
  }
  bool has_se = false;
  double se = 0.0;
  if (st.sampling == SOBOL) {
  } else if (known) {
    has_se = true;
    se = 0.0;
  } else {
    has_se = sampling_standard_error(st, sample, pairs, log_discount, &se);
  }
  if (has_seed && has_se && se == 0.0) {
    e.status = ZERO_OBSERVED_VARIANCE;
    return e;
  }
  if (has_se) {
    e.status = AVAILABLE;
    e.has_standard_error = 1;
    e.standard_error = se;
  } else if (st.sampling == SOBOL) {
    e.status = UNSUPPORTED_SAMPLING;
  } else if (st.antithetic && st.path_count % 2 != 0) {
    e.status = INCOMPLETE_ANTITHETIC_PAIR;
  } else {
    e.status = INSUFFICIENT_OBSERVATIONS;
  }
  if (!std::isfinite(e.price)) unstable();
  return e;
}

}  // namespace oracle
