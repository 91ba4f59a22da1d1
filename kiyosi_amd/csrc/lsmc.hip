// Least-squares Monte Carlo for American options
// (engines/monte_carlo/american.rs: american_scenarios :955-1455,
// price_american_out_of_sample :574-637, solve_quadratic :1693-1749).
//
// Data in HBM for one pricing (P = samples, S = exercise dates):
//   coordinates  BSM: Brownian coordinate W_k per sample, the antithetic
//                leg is -W_k, [S][P] in forward mode and [P] (the current
//                date only) in reverse mode; local vol: log-spot [S][slots];
//                anchored BSM: spot [S][slots] in FP64
//   spot, cash   [slots] FP64, slots = P or 2P with antithetic pairs
//   partials     [blocks][entries] per-block normal equations
//   policy       [S-1][8] FP64: valid flag and continuation coefficients
// One thread per sample and one 256-thread workgroup per reference block.
// The backward loop launches one regression kernel and one solve kernel per
// exercise date. For the reference's quadratic basis each normal-equation
// entry is summed in the reference's order (paths in block order, then
// blocks in order), so the FP64 engine reproduces its coefficients.
#include "lsmc_device.hpp"

namespace kg {

// ------------------------------------------------------------------ device
__global__ void clear_doubles(double* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// Entries of the normal equations: quadratic keeps the reference's nine
// sums (x^0..x^4 moments and three right-hand sides); poly4 holds the 15
// distinct Gram entries of (1, x, .., x^4) and five right-hand sides.
template <int BASIS>
struct Entries;
template <>
struct Entries<KG_BASIS_QUADRATIC> {
  static constexpr int N = 9;
};
template <>
struct Entries<KG_BASIS_POLY4> {
  static constexpr int N = 20;
};

template <int BASIS>
__device__ inline void contributions(double x, double t, double* e) {
  if (BASIS == KG_BASIS_QUADRATIC) {
    const double sq = x * x;
    e[0] = 1.0;      // [0][0]
    e[1] = x;        // [0][1]
    e[2] = sq;       // [0][2]
    e[3] = sq;       // [1][1]
    e[4] = x * sq;   // [1][2]
    e[5] = sq * sq;  // [2][2]
    e[6] = t;        // [0][3]
    e[7] = x * t;    // [1][3]
    e[8] = sq * t;   // [2][3]
  } else {
    double p[9];
    p[0] = 1.0;
    for (int i = 1; i < 9; ++i) p[i] = p[i - 1] * x;
    int k = 0;
    for (int r = 0; r < 5; ++r)
      for (int c = r; c < 5; ++c) e[k++] = p[r + c];
    for (int r = 0; r < 5; ++r) e[15 + r] = p[r] * t;
  }
}

template <bool FP32, int KIND, int BASIS>
__global__ void __launch_bounds__(256)
    lsmc_backward(LsmcModel m, int step, const double* __restrict__ pol, int have, int64_t samples,
                  int64_t paths, int antithetic, uint64_t seed, SobolRef so, void* __restrict__ coords,
                  double* __restrict__ spot, double* __restrict__ cash, double* __restrict__ partials,
                  uint32_t* __restrict__ flags) {
  constexpr int E = Entries<BASIS>::N;
  extern __shared__ double lds[];  // quadratic: [E][512] in slot order
  __shared__ double red[E * 4];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const int legs = antithetic ? 2 : 1;
  const int64_t slots = samples * legs;
  double acc[E];
  for (int e = 0; e < E; ++e) acc[e] = 0.0;
  uint32_t bad = 0;
  // Reverse mode: this date's coordinate from the next date's, in place.
  double w_rev = 0.0;
  if (KIND == PATHS_BROWNIAN && m.reverse && n < samples) {
    const int count = (int)(samples - block * BLOCK_SAMPLES < BLOCK_SAMPLES ? samples - block * BLOCK_SAMPLES
                                                                            : BLOCK_SAMPLES);
    const int row = m.steps - 1 - step;
    const double prev = FP32 ? (double)((const float*)coords)[n] : ((const double*)coords)[n];
    w_rev = bridge_back<FP32>(prev, reverse_normal<FP32>(seed, block, j, count, row, n, so), step, m.steps);
    if (FP32)
      ((float*)coords)[n] = (float)w_rev;
    else
      ((double*)coords)[n] = w_rev;
  }
  for (int leg = 0; leg < 2; ++leg) {
    double e[E];
    for (int k = 0; k < E; ++k) e[k] = 0.0;
    const bool on = n < samples && (leg == 0 || (antithetic && n < paths / 2));
    if (on) {
      const int64_t slot = n * legs + leg;
      double c = cash[slot];
      if (have && pol[0] != 0.0) {
        const double sp = spot[slot];
        const double intrinsic = payoff(m, sp);
        if (intrinsic > 0.0) {
          const double cont = continuation(pol, BASIS, sp / m.strike);
          if (!isfinite(cont)) bad = 1;
          if (intrinsic / m.strike > cont) c = intrinsic;
        }
      }
      double s;
      if (KIND == PATHS_LOCAL_VOL) {
        const double x = FP32 ? (double)((const float*)coords)[(size_t)step * slots + slot]
                              : ((const double*)coords)[(size_t)step * slots + slot];
        s = m.spot * (FP32 ? (double)f32::exp((float)x) : exp(x));
      } else if (KIND == PATHS_ANCHORED) {
        s = ((const double*)coords)[(size_t)step * slots + slot];
      } else {
        const double w = m.reverse ? w_rev
                         : FP32    ? (double)((const float*)coords)[(size_t)step * samples + n]
                                   : ((const double*)coords)[(size_t)step * samples + n];
        s = FP32 ? bsm_spot32(m, step, (float)(leg ? -w : w)) : bsm_spot(m, step, leg ? -w : w);
      }
      c *= m.discount;
      if (!isfinite(c)) bad = 1;
      spot[slot] = s;
      cash[slot] = c;
      if (payoff(m, s) > 0.0) contributions<BASIS>(s / m.strike, c / m.strike, e);
    }
    if (BASIS == KG_BASIS_QUADRATIC) {
      for (int k = 0; k < E; ++k) lds[k * 512 + 2 * j + leg] = e[k];
    } else {
      for (int k = 0; k < E; ++k) acc[k] += e[k];
    }
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
  if (BASIS == KG_BASIS_QUADRATIC) {
    __syncthreads();
    if (j < E) {  // reference order: samples in block order, a pair's legs together
      double sum = 0.0;
      for (int s = 0; s < 512; ++s) sum += lds[j * 512 + s];
      partials[block * E + j] = sum;
    }
  } else {
    const int lane = j & (WAVE - 1), wave = j / WAVE;
    for (int k = 0; k < E; ++k) {
      double v = acc[k];
      for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
      if (lane == 0) red[k * 4 + wave] = v;
    }
    __syncthreads();
    if (j < E) partials[block * E + j] = ((red[j * 4] + red[j * 4 + 1]) + red[j * 4 + 2]) + red[j * 4 + 3];
  }
}

// solve_quadratic (american.rs:1693-1749).
__device__ inline bool solve_quadratic_d(const double (&s)[3][4], double* out) {
  for (int degree = 2; degree >= 0; --degree) {
    double m[3][4];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 4; ++c) m[r][c] = s[r][c];
    double scale = 0.0;
    for (int r = 0; r <= degree; ++r)
      for (int c = 0; c <= degree; ++c) scale = fmax(scale, fabs(m[r][c]));
    bool singular = false;
    for (int col = 0; col <= degree; ++col) {
      int pivot = col;
      for (int r = col + 1; r <= degree; ++r)
        if (fabs(m[r][col]) >= fabs(m[pivot][col])) pivot = r;
      if (!isfinite(m[pivot][col])) return false;
      if (fabs(m[pivot][col]) <= 1e-14 * scale) {
        singular = true;
        break;
      }
      for (int c = 0; c < 4; ++c) {
        const double t = m[col][c];
        m[col][c] = m[pivot][c];
        m[pivot][c] = t;
      }
      double prow[4];
      for (int c = 0; c < 4; ++c) prow[c] = m[col][c];
      for (int r = col + 1; r <= degree; ++r) {
        const double f = m[r][col] / m[col][col];
        for (int c = col; c <= degree; ++c) m[r][c] -= f * prow[c];
        m[r][3] -= f * m[col][3];
      }
    }
    if (singular) continue;
    double coef[3] = {0.0, 0.0, 0.0};
    for (int r = degree; r >= 0; --r) {
      double v = m[r][3];
      for (int c = r + 1; c <= degree; ++c) v -= m[r][c] * coef[c];
      coef[r] = v / m[r][r];
    }
    for (int k = 0; k < 3; ++k)
      if (!isfinite(coef[k])) return false;
    for (int k = 0; k < 3; ++k) out[k] = coef[k];
    return true;
  }
  return false;
}

// Cholesky solve of the degree-4 normal equations, dropping the top degree
// while a pivot falls below 1e-13 of the largest diagonal entry.
__device__ inline bool solve_poly4_d(const double* e, double* out) {
  double g[5][5], b[5];
  int k = 0;
  for (int r = 0; r < 5; ++r)
    for (int c = r; c < 5; ++c) g[r][c] = g[c][r] = e[k++];
  for (int r = 0; r < 5; ++r) b[r] = e[15 + r];
  for (int degree = 4; degree >= 0; --degree) {
    double l[5][5] = {};
    double scale = 0.0;
    for (int r = 0; r <= degree; ++r) scale = fmax(scale, g[r][r]);
    bool ok = true;
    for (int c = 0; c <= degree && ok; ++c) {
      double d = g[c][c];
      for (int p = 0; p < c; ++p) d -= l[c][p] * l[c][p];
      if (!(d > 1e-13 * scale)) {
        ok = false;
        break;
      }
      l[c][c] = sqrt(d);
      for (int r = c + 1; r <= degree; ++r) {
        double v = g[r][c];
        for (int p = 0; p < c; ++p) v -= l[r][p] * l[c][p];
        l[r][c] = v / l[c][c];
      }
    }
    if (!ok) continue;
    double y[5], x[5] = {0, 0, 0, 0, 0};
    for (int r = 0; r <= degree; ++r) {
      double v = b[r];
      for (int p = 0; p < r; ++p) v -= l[r][p] * y[p];
      y[r] = v / l[r][r];
    }
    for (int r = degree; r >= 0; --r) {
      double v = y[r];
      for (int p = r + 1; p <= degree; ++p) v -= l[p][r] * x[p];
      x[r] = v / l[r][r];
    }
    for (int r = 0; r < 5; ++r) {
      if (!isfinite(x[r])) return false;
      out[r] = x[r];
    }
    return true;
  }
  return false;
}

// Sums each entry over the blocks, then solves; writes policy row `step`:
// [valid, c0 .. c4]. The quadratic basis adds the block sums in block order,
// as the reference does, streaming them through LDS so the sequential adds
// wait on LDS rather than on global loads; the poly-4 sums use a fixed
// tree (thread t folds blocks t, t + 256, ..., then waves and threads
// combine in a fixed order).
constexpr int SOLVE_CHUNK = 128;

template <int BASIS>
__global__ void __launch_bounds__(256)
    lsmc_solve(const double* __restrict__ partials, int64_t blocks, double* __restrict__ pol_row,
               uint32_t* __restrict__ flags) {
  constexpr int E = Entries<BASIS>::N;
  __shared__ double sums[E];
  const int t = threadIdx.x;
  if (BASIS == KG_BASIS_QUADRATIC) {
    __shared__ double buf[SOLVE_CHUNK * E];
    double s = 0.0;
    for (int64_t b0 = 0; b0 < blocks; b0 += SOLVE_CHUNK) {
      const int nb = (int)(blocks - b0 < SOLVE_CHUNK ? blocks - b0 : SOLVE_CHUNK);
      for (int i = t; i < nb * E; i += blockDim.x) buf[i] = partials[b0 * E + i];
      __syncthreads();
      if (t < E)
        for (int k = 0; k < nb; ++k) s += buf[k * E + t];
      __syncthreads();
    }
    if (t < E) sums[t] = s;
  } else {
    __shared__ double red[E * 4];
    double acc[E];
    for (int e = 0; e < E; ++e) acc[e] = 0.0;
    for (int64_t b = t; b < blocks; b += blockDim.x)
      for (int e = 0; e < E; ++e) acc[e] += partials[b * E + e];
    const int lane = t & (WAVE - 1), wave = t / WAVE;
    for (int e = 0; e < E; ++e) {
      double v = acc[e];
      for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
      if (lane == 0) red[e * 4 + wave] = v;
    }
    __syncthreads();
    if (t < E) sums[t] = ((red[t * 4] + red[t * 4 + 1]) + red[t * 4 + 2]) + red[t * 4 + 3];
  }
  __syncthreads();
  if (t != 0) return;
  for (int k = 0; k < 8; ++k) pol_row[k] = 0.0;
  if (sums[0] == 0.0) return;  // no path in the money
  for (int k = 0; k < E; ++k)
    if (!isfinite(sums[k])) {
      raise_flag(flags, FLAG_NONFINITE);
      return;
    }
  double coef[5] = {0, 0, 0, 0, 0};
  bool ok;
  if (BASIS == KG_BASIS_QUADRATIC) {
    double mat[3][4] = {{sums[0], sums[1], sums[2], sums[6]},
                        {sums[1], sums[3], sums[4], sums[7]},
                        {sums[2], sums[4], sums[5], sums[8]}};
    ok = solve_quadratic_d(mat, coef);
  } else {
    ok = solve_poly4_d(sums, coef);
  }
  if (!ok) {
    raise_flag(flags, FLAG_NONFINITE);
    return;
  }
  pol_row[0] = 1.0;
  for (int k = 0; k < 5; ++k) pol_row[1 + k] = coef[k];
}

// Out-of-sample valuation (price_american_out_of_sample, american.rs:574-637):
// each sample streams until the fitted rule exercises; the two legs of an
// antithetic pair step together on the same normals with opposite signs.
// The step loop is block-uniform so the local-volatility rows can be
// staged in LDS.
template <bool FP32, bool LV, int BASIS>
__global__ void __launch_bounds__(256)
    lsmc_price(LsmcModel m, const double* __restrict__ policy, uint64_t seed, int64_t samples,
               int64_t paths, int antithetic, SobolRef so, Moments* __restrict__ parts,
               uint32_t* __restrict__ flags) {
  __shared__ Moments smem[2 * 4];
  __shared__ double lv_lds[LV ? 3 * LV_LDS_NODES : 1];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const bool live = n < samples;
  const bool paired = antithetic && live && n < paths / 2;
  uint32_t bad = 0;
  PathDraws<FP32> z;
  if (live) z.start(seed, (uint64_t)block, j, m.steps, n, so);
  LvRows rows{};
  rows.row = -1;
  double current[2] = {m.spot, m.spot}, x[2] = {0.0, 0.0}, values[2] = {0.0, 0.0};
  int bracket[2] = {-1, -1};
  bool alive[2] = {live, paired};
  for (int step = 0; step < m.steps; ++step) {
    if (LV) lv_stage<FP32>(m, step, lv_lds, rows);  // block-uniform
    if (!alive[0] && !alive[1]) continue;
    const double e = z.next();
    for (int leg = 0; leg < 2; ++leg) {
      if (!alive[leg]) continue;
      const double d = leg ? -e : e;
      if (LV) {
        x[leg] = lv_step<FP32>(m, rows, x[leg], d, bracket[leg]);
        current[leg] = m.spot * (FP32 ? (double)f32::exp((float)x[leg]) : exp(x[leg]));
      } else if (FP32) {
        current[leg] =
            (double)((float)current[leg] * f32::exp(fmaf((float)m.diffusion, (float)d, (float)m.drift)));
      } else {
        current[leg] *= exp(fma(m.diffusion, d, m.drift));
      }
      if (!isfinite(current[leg]) || current[leg] <= 0.0) bad = 1;
      if (step == m.steps - 1) continue;
      const double intrinsic = payoff(m, current[leg]);
      const double* pol = policy + (size_t)step * 8;
      if (intrinsic > 0.0 && pol[0] != 0.0) {
        const double cont = continuation(pol, BASIS, current[leg] / m.strike);
        if (!isfinite(cont)) bad = 1;
        if (intrinsic / m.strike > cont) {
          values[leg] = discounted_product(1.0, intrinsic, -m.rate * (m.dt * (double)(step + 1)));
          alive[leg] = false;
        }
      }
    }
  }
  for (int leg = 0; leg < 2; ++leg)
    if (alive[leg]) values[leg] = discounted_product(1.0, payoff(m, current[leg]), -m.rate * m.time);
  Moments mm[2] = {moments_zero(), moments_zero()};
  if (live) {
    moments_add(mm[0], values[0], true);
    if (paired) {
      moments_add(mm[0], values[1], true);
      moments_add(mm[1], midpoint(values[0], values[1]), true);
    }
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
  block_merge<2>(mm, smem);
  if (threadIdx.x == 0) {
    parts[block * 2] = mm[0];
    parts[block * 2 + 1] = mm[1];
  }
}

// ------------------------------------------------------------------ host
static SobolRef sobol_ref(kg_context* ctx, const Sampler& s, int steps) {
  if (!s.sobol) return SobolRef{nullptr, nullptr, 0};
  const SobolSetup& so = sobol_setup(ctx, (uint32_t)steps, s.seed);
  return SobolRef{(const uint64_t*)so.d_directions.ptr, (const uint64_t*)so.d_shifts.ptr, steps};
}

template <bool FP32, int KIND, int BASIS>
static double run_lsmc(kg_context* ctx, const LsmcModel& m, const Sampler& s, double* d_policy,
                       int32_t* dates_fitted, double* immediate_out) {
  const int64_t samples = s.samples(), blocks = s.blocks();
  const int legs = s.antithetic ? 2 : 1;
  const size_t word = FP32 ? 4 : 8;
  const bool reverse = KIND == PATHS_BROWNIAN && m.reverse;
  // Stored values: one coordinate per sample in reverse mode, else one per
  // date and sample (Brownian) or per date and slot (local vol, anchored).
  const size_t values = reverse ? (size_t)samples
                        : (size_t)m.steps * (size_t)(KIND == PATHS_BROWNIAN ? samples : samples * legs);
  // Rounded up so the FP64 arrays after the coordinates stay aligned.
  const size_t coord_bytes = (values * word + 255) & ~size_t(255);
  constexpr int E = Entries<BASIS>::N;
  unsigned char* base = (unsigned char*)ctx->scratch.reserve(coord_bytes + 2 * (size_t)samples * legs * 8 +
                                                             (size_t)blocks * E * 8 + 1024);
  void* coords = base;
  double* spot = (double*)(base + coord_bytes);
  double* cash = spot + samples * legs;
  double* partials = cash + samples * legs;
  const SobolRef so = sobol_ref(ctx, s, m.steps);
  uint32_t* flags = reset_flags(ctx);
  if (reverse)
    hipLaunchKernelGGL((lsmc_reverse_start<FP32>), dim3(blocks), dim3(256), 0, ctx->stream, m, s.seed, samples,
                       s.path_count, (int)s.antithetic, so, coords, spot, cash);
  else
    hipLaunchKernelGGL((lsmc_generate<FP32, KIND>), dim3(blocks), dim3(256), 0, ctx->stream, m, s.seed, samples,
                       s.path_count, (int)s.antithetic, so, coords, spot, cash, flags);
  KG_HIP(hipGetLastError());
  const size_t lds = BASIS == KG_BASIS_QUADRATIC ? (size_t)E * 512 * 8 : 0;
  int have = 0;
  int fitted = 0;
  for (int step = m.steps - 2; step >= 0; --step) {
    hipLaunchKernelGGL((lsmc_backward<FP32, KIND, BASIS>), dim3(blocks), dim3(256), lds, ctx->stream, m, step,
                       d_policy + (size_t)(step + 1) * 8, have, samples, s.path_count, (int)s.antithetic, s.seed,
                       so, coords, spot, cash, partials, flags);
    KG_HIP(hipGetLastError());
    hipLaunchKernelGGL((lsmc_solve<BASIS>), dim3(1), dim3(256), 0, ctx->stream, partials, blocks,
                       d_policy + (size_t)step * 8, flags);
    KG_HIP(hipGetLastError());
    have = 1;  // the kernels read the valid flag of the row they use
    ++fitted;
  }
  Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * 2 * blocks);
  hipLaunchKernelGGL((lsmc_final<BASIS>), dim3(blocks), dim3(256), 0, ctx->stream, m, d_policy, 1, samples,
                     s.path_count, (int)s.antithetic, spot, cash, parts);
  KG_HIP(hipGetLastError());
  check_flags(ctx);
  Moments r[2];
  merge_and_fetch(ctx, parts, blocks, 2, r);
  *dates_fitted = fitted;
  return std::max(scaled_value(r[0].scale, r[0].mean, m.log_discount), *immediate_out);
}

template <bool FP32, bool LV, int BASIS>
static void run_price(kg_context* ctx, const LsmcModel& m, const Sampler& s, const double* d_policy,
                      Moments out[2]) {
  const int64_t blocks = s.blocks();
  Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * 2 * blocks);
  const SobolRef so = sobol_ref(ctx, s, m.steps);
  uint32_t* flags = reset_flags(ctx);
  hipLaunchKernelGGL((lsmc_price<FP32, LV, BASIS>), dim3(blocks), dim3(256), 0, ctx->stream, m, d_policy, s.seed,
                     s.samples(), s.path_count, (int)s.antithetic, so, parts, flags);
  KG_HIP(hipGetLastError());
  check_flags(ctx);
  merge_and_fetch(ctx, parts, blocks, 2, out);
}

template <bool FP32, int KIND, int BASIS>
static void american(kg_context* ctx, const LsmcModel& m, const Sampler& pricing, const Sampler& training,
                     bool oos, double immediate, kg_lsmc_report* out) {
  constexpr bool LV = KIND == PATHS_LOCAL_VOL;
  double* policy = (double*)ctx->policy.reserve((size_t)m.steps * 8 * sizeof(double));
  hipLaunchKernelGGL(clear_doubles, dim3(1), dim3(256), 0, ctx->stream, policy, m.steps * 8);
  KG_HIP(hipGetLastError());
  int32_t fitted = 0;
  timing_begin(ctx);
  const double in_sample = run_lsmc<FP32, KIND, BASIS>(ctx, m, oos ? training : pricing, policy, &fitted, &immediate);
  out->training_ms = timing_end(ctx);
  out->in_sample_price = in_sample;
  out->dates_fitted = fitted;
  out->path_steps = (double)(oos ? training : pricing).path_count * m.steps;
  out->price = in_sample;
  if (oos) {
    Moments r[2];
    timing_begin(ctx);
    run_price<FP32, LV, BASIS>(ctx, m, pricing, policy, r);
    out->pricing_ms = timing_end(ctx);
    kg_estimate e{};
    fill_estimate(pricing, r[0], r[1], 0.0, false, true, &e);
    out->price = std::max(e.price, immediate);
    out->standard_error = e.has_standard_error ? e.standard_error : e.path_standard_error;
    out->path_steps += (double)pricing.path_count * m.steps;
  }
}

template <int BASIS>
static void american_basis(kg_context* ctx, const LsmcHost& h, kg_lsmc_report* out) {
  const LsmcModel& m = h.model;
  // Anchored spots can leave the FP32 range, so those paths step in FP64.
  const bool fp32 = h.pricing.fp32 && !m.anchored;
#define KG_AMERICAN(F, K) american<F, K, BASIS>(ctx, m, h.pricing, h.training, h.oos, h.immediate, out)
  if (m.lv) {
    fp32 ? KG_AMERICAN(true, PATHS_LOCAL_VOL) : KG_AMERICAN(false, PATHS_LOCAL_VOL);
  } else if (m.anchored) {
    KG_AMERICAN(false, PATHS_ANCHORED);
  } else {
    fp32 ? KG_AMERICAN(true, PATHS_BROWNIAN) : KG_AMERICAN(false, PATHS_BROWNIAN);
  }
#undef KG_AMERICAN
}

}  // namespace kg

using namespace kg;

extern "C" kg_status kg_american_price(kg_context* ctx, const kg_settings* settings,
                                       const kg_american_terms* option, const kg_american_model* model,
                                       const kg_lsmc_options* options, kg_lsmc_report* out) {
  return guarded(ctx, [&] {
    if (!settings || !option || !model || !options || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
    *out = kg_lsmc_report{};
    LsmcHost h = lsmc_plan(ctx, *settings, *option, *model, *options);
    if (h.known) {
      out->price = out->in_sample_price = h.known_value;
      return;
    }
    if (options->basis == KG_BASIS_MLP) {
      american_mlp(ctx, h, *options, out);
      return;
    }
    if (options->basis == KG_BASIS_QUADRATIC)
      american_basis<KG_BASIS_QUADRATIC>(ctx, h, out);
    else
      american_basis<KG_BASIS_POLY4>(ctx, h, out);
  });
}
