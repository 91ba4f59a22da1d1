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
__global__ void clear_doubles(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}
static void clear_policy(kg_context* ctx, double* p, int64_t n) {
  const int64_t grid = std::min<int64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(clear_doubles, dim3((unsigned)grid), dim3(256), 0, ctx->stream, p, n);
  KG_HIP(hipGetLastError());
}

// Entries of the normal equations: quadratic keeps the reference's nine
// sums (x^0..x^4 moments and three right-hand sides); poly4 holds the nine
// power sums x^0..x^8 (Gram entry [r][c] is power r + c) and the five
// right-hand sides x^r t.
template <int BASIS>
struct Entries;
template <>
struct Entries<KG_BASIS_QUADRATIC> {
  static constexpr int N = 9;
};
template <>
struct Entries<KG_BASIS_POLY4> {
  static constexpr int N = 14;
};

// Quadratic block partials: the slot threads write entry k of slot s to
// lds[k * QUAD_ROW + s]; thread k then sums row k in slot order (the
// reference's order). The row pitch of 514 doubles puts the nine rows in
// different LDS banks.
constexpr int QUAD_ROW = 514;
// Poly-4 block partials: 14 rows of the 256 threads' sums; the pitch of
// 264 doubles staggers the rows' banks.
constexpr int POLY4_ROW = 264;

template <int CTRL>
__device__ inline double dpp_double(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// Sum over the 16 lanes of a row, left in every lane of the row: quad swaps
// (1 0 3 2), (2 3 0 1), then the half-row and row mirrors (each pair adds
// the same two values, so every lane forms the same bits).
__device__ inline double sum_row16(double v) {
  v += dpp_double<0xB1>(v);
  v += dpp_double<0x4E>(v);
  v += dpp_double<0x141>(v);
  v += dpp_double<0x140>(v);
  return v;
}

// s + p[0] + p[stride] + ... + p[(n - 1) stride], added in that order.
// Each batch of 16 values is read before its adds, so the reads overlap;
// the adds themselves are one dependent chain.
__device__ inline double ordered_sum(double s, const double* p, int stride, int n) {
  int k = 0;
  for (; k + 16 <= n; k += 16) {
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = p[(k + i) * stride];
#pragma unroll
    for (int i = 0; i < 16; ++i) s += v[i];
  }
  for (; k < n; ++k) s += p[k * stride];
  return s;
}

__device__ inline double quadratic_row_sum(const double* row) { return ordered_sum(0.0, row, 1, 512); }

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
    e[0] = 1.0;
    for (int i = 1; i < 9; ++i) e[i] = e[i - 1] * x;
    for (int r = 0; r < 5; ++r) e[9 + r] = e[r] * t;
  }
}

// One exercise date of one lane for the slots of this thread's sample:
// the exercise decision at the later date (its spot and policy row `pol`),
// the spot at this date from the slot's coordinate `coord[leg]`, the
// discounted cash, and the normal-equation contributions of in-the-money
// slots, summed in the reference's order into this lane's block partials.
template <bool FP32, int KIND, int BASIS>
__device__ inline void backward_lane(const LsmcModel& m, int step, const double* pol, int have, int64_t n,
                                     int legs, const bool (&on)[2], const double (&coord)[2],
                                     double* __restrict__ spot, double* __restrict__ cash,
                                     double* __restrict__ partials, double* lds, uint32_t& bad) {
  constexpr int E = Entries<BASIS>::N;
  const int j = threadIdx.x;
  double acc[E];
  for (int e = 0; e < E; ++e) acc[e] = 0.0;
  for (int leg = 0; leg < 2; ++leg) {
    double e[E];
    for (int k = 0; k < E; ++k) e[k] = 0.0;
    if (on[leg]) {
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
        s = m.spot * (FP32 ? (double)f32::exp((float)coord[leg]) : exp(coord[leg]));
      } else if (KIND == PATHS_ANCHORED) {
        s = coord[leg];
      } else {
        s = FP32 ? bsm_spot32(m, step, (float)coord[leg]) : bsm_spot(m, step, coord[leg]);
      }
      c *= m.discount;
      if (!isfinite(c)) bad = 1;
      spot[slot] = s;
      cash[slot] = c;
      if (payoff(m, s) > 0.0) contributions<BASIS>(s / m.strike, c / m.strike, e);
    }
    if (BASIS == KG_BASIS_QUADRATIC) {
      for (int k = 0; k < E; ++k) lds[k * QUAD_ROW + 2 * j + leg] = e[k];
    } else {
      for (int k = 0; k < E; ++k) acc[k] += e[k];
    }
  }
  const int64_t block = blockIdx.x;
  if (BASIS == KG_BASIS_QUADRATIC) {
    __syncthreads();
    // Reference order: samples in block order, a pair's legs together.
    if (j < E) partials[block * E + j] = quadratic_row_sum(lds + j * QUAD_ROW);
  } else {
    // Row k of LDS holds entry k of the 256 threads; the 16 lanes of row
    // group k sum its elements p, p + 16, .. and the group's DPP butterfly
    // leaves the block's sum in every lane of the group.
    for (int k = 0; k < E; ++k) lds[k * POLY4_ROW + j] = acc[k];
    __syncthreads();
    if (j < E * 16) {
      const int k = j >> 4, p = j & 15;
      double v[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = lds[k * POLY4_ROW + p + 16 * i];
      double sum = 0.0;
#pragma unroll
      for (int i = 0; i < 16; ++i) sum += v[i];
      sum = sum_row16(sum);
      if (p == 0) partials[block * E + k] = sum;
    }
  }
}

// One exercise date of the backward loop. With LANES, `nl` lanes (strikes
// or market scenarios, american_scenarios::<N, S>) share the coordinates:
// lane l reads its model from lanes[l] and keeps its spot and cash at
// l * slots, its partials at l * blocks * E and its policy at l * pol_stride
// doubles; each lane computes exactly what the lane priced alone computes.
template <bool FP32, int KIND, int BASIS, bool LANES>
__global__ void __launch_bounds__(256)
    lsmc_backward(LsmcModel m, const LsmcModel* __restrict__ lanes, int nl, int step, const double* __restrict__ pol,
                  int64_t pol_stride, int have, int64_t samples, int64_t paths, int antithetic, uint64_t seed,
                  SobolRef so, void* __restrict__ coords, double* __restrict__ spot, double* __restrict__ cash,
                  double* __restrict__ partials, uint32_t* __restrict__ flags) {
  constexpr int E = Entries<BASIS>::N;
  __shared__ __align__(16) double lds[BASIS == KG_BASIS_QUADRATIC ? E * QUAD_ROW : E * POLY4_ROW];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const int legs = antithetic ? 2 : 1;
  const int64_t slots = samples * legs;
  uint32_t bad = 0;
  bool on[2];
  double coord[2] = {0.0, 0.0};
  for (int leg = 0; leg < 2; ++leg) on[leg] = n < samples && (leg == 0 || (antithetic && n < paths / 2));
  if (KIND == PATHS_BROWNIAN) {
    double w = 0.0;
    if (m.reverse && n < samples) {
      // Reverse mode: this date's coordinate from the next date's, in place.
      const int count = (int)(samples - block * BLOCK_SAMPLES < BLOCK_SAMPLES ? samples - block * BLOCK_SAMPLES
                                                                              : BLOCK_SAMPLES);
      const int row = m.steps - 1 - step;
      const double prev = FP32 ? (double)((const float*)coords)[n] : ((const double*)coords)[n];
      w = bridge_back<FP32>(prev, reverse_normal<FP32>(seed, block, j, count, row, n, so), step, m.steps);
      if (FP32)
        ((float*)coords)[n] = (float)w;
      else
        ((double*)coords)[n] = w;
    } else if (n < samples) {
      w = FP32 ? (double)((const float*)coords)[(size_t)step * samples + n]
               : ((const double*)coords)[(size_t)step * samples + n];
    }
    coord[0] = w;
    coord[1] = -w;
  } else {
    for (int leg = 0; leg < 2; ++leg)
      if (on[leg]) {
        const size_t at = (size_t)step * slots + n * legs + leg;
        coord[leg] = KIND == PATHS_LOCAL_VOL && FP32 ? (double)((const float*)coords)[at] : ((const double*)coords)[at];
      }
  }
  if (LANES) {
    const int64_t blocks = gridDim.x;
    for (int l = 0; l < nl; ++l) {
      const LsmcModel ml = lanes[l];
      backward_lane<FP32, KIND, BASIS>(ml, step, pol + l * pol_stride, have, n, legs, on, coord, spot + l * slots,
                                       cash + l * slots, partials + l * blocks * E, lds, bad);
      __syncthreads();  // the next lane rewrites the LDS rows
    }
  } else {
    backward_lane<FP32, KIND, BASIS>(m, step, pol, have, n, legs, on, coord, spot, cash, partials, lds, bad);
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
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
  for (int r = 0; r < 5; ++r)
    for (int c = 0; c < 5; ++c) g[r][c] = e[r + c];
  for (int r = 0; r < 5; ++r) b[r] = e[9 + r];
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
// combine in a fixed order). Workgroup l solves lane l: its partials at
// l * blocks * E, its row at l * pol_stride.
constexpr int SOLVE_CHUNK = 512;

template <int BASIS>
__global__ void __launch_bounds__(256)
    lsmc_solve(const double* __restrict__ partials, int64_t blocks, double* __restrict__ pol_row,
               int64_t pol_stride, uint32_t* __restrict__ flags) {
  constexpr int E = Entries<BASIS>::N;
  __shared__ double sums[E];
  const int t = threadIdx.x;
  partials += (int64_t)blockIdx.x * blocks * E;
  pol_row += (int64_t)blockIdx.x * pol_stride;
  if (BASIS == KG_BASIS_QUADRATIC) {
    // Chunks of 512 blocks pass through LDS; every thread loads the next
    // chunk into registers while threads 0..8 add the current one, so the
    // global loads overlap the sequential adds.
    __shared__ double buf[SOLVE_CHUNK * E];
    constexpr int PER = SOLVE_CHUNK * E / 256;
    static_assert(PER * 256 == SOLVE_CHUNK * E, "a chunk is a whole number of loads per thread");
    double next[PER];
    auto fetch = [&](int64_t b0) {
      const int64_t count = (blocks - b0 < SOLVE_CHUNK ? blocks - b0 : SOLVE_CHUNK) * E;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int idx = t + i * 256;
        next[i] = idx < count ? partials[b0 * E + idx] : 0.0;
      }
    };
    fetch(0);
    double s = 0.0;
    for (int64_t b0 = 0; b0 < blocks; b0 += SOLVE_CHUNK) {
      const int nb = (int)(blocks - b0 < SOLVE_CHUNK ? blocks - b0 : SOLVE_CHUNK);
#pragma unroll
      for (int i = 0; i < PER; ++i) buf[t + i * 256] = next[i];
      __syncthreads();
      if (b0 + SOLVE_CHUNK < blocks) fetch(b0 + SOLVE_CHUNK);
      if (t < E) s = ordered_sum(s, buf + t, E, nb);
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

// In-sample LSMC of `nl` lanes on one set of paths: ml[l] is lane l's
// model (host copy) and d_lanes their device copy, or nullptr for a single
// lane, which runs the single-lane kernels. values[l] receives lane l's
// price: max(discounted mean, immediate[l]).
template <bool FP32, int KIND, int BASIS>
static void run_lsmc(kg_context* ctx, const LsmcModel* ml, const LsmcModel* d_lanes, int nl, const Sampler& s,
                     double* d_policy, int32_t* dates_fitted, const double* immediate, double* values) {
  const LsmcModel& m = ml[0];
  const bool lanes = d_lanes != nullptr;
  const int64_t samples = s.samples(), blocks = s.blocks();
  const int legs = s.antithetic ? 2 : 1;
  const size_t word = FP32 ? 4 : 8;
  const bool reverse = KIND == PATHS_BROWNIAN && m.reverse;
  // Stored values: one coordinate per sample in reverse mode, else one per
  // date and sample (Brownian) or per date and slot (local vol, anchored).
  const size_t values_stored = reverse ? (size_t)samples
                               : (size_t)m.steps * (size_t)(KIND == PATHS_BROWNIAN ? samples : samples * legs);
  // Rounded up so the FP64 arrays after the coordinates stay aligned.
  const size_t coord_bytes = (values_stored * word + 255) & ~size_t(255);
  constexpr int E = Entries<BASIS>::N;
  const int64_t slots = samples * legs;
  unsigned char* base = (unsigned char*)ctx->scratch.reserve(coord_bytes + (size_t)nl * 2 * (size_t)slots * 8 +
                                                             (size_t)nl * (size_t)blocks * E * 8 + 1024);
  void* coords = base;
  double* spot = (double*)(base + coord_bytes);
  double* cash = spot + (size_t)nl * slots;
  double* partials = cash + (size_t)nl * slots;
  const int64_t pol_stride = (int64_t)m.steps * 8;
  const SobolRef so = sobol_ref(ctx, s, m.steps);
  uint32_t* flags = reset_flags(ctx);
  if (reverse)
    hipLaunchKernelGGL((lsmc_reverse_start<FP32>), dim3(blocks), dim3(256), 0, ctx->stream, m, s.seed, samples,
                       s.path_count, (int)s.antithetic, so, coords, spot, cash);
  else
    hipLaunchKernelGGL((lsmc_generate<FP32, KIND>), dim3(blocks), dim3(256), 0, ctx->stream, m, s.seed, samples,
                       s.path_count, (int)s.antithetic, so, coords, spot, cash, flags);
  KG_HIP(hipGetLastError());
  if (lanes) {
    hipLaunchKernelGGL((lsmc_lane_terminal<FP32, KIND>), dim3(blocks), dim3(256), 0, ctx->stream, d_lanes, nl,
                       samples, s.path_count, (int)s.antithetic, (int)reverse, (const void*)coords, spot, cash);
    KG_HIP(hipGetLastError());
  }
  const size_t lds = 0;
  int have = 0;
  int fitted = 0;
  for (int step = m.steps - 2; step >= 0; --step) {
    if (lanes)
      hipLaunchKernelGGL((lsmc_backward<FP32, KIND, BASIS, true>), dim3(blocks), dim3(256), lds, ctx->stream, m,
                         d_lanes, nl, step, d_policy + (size_t)(step + 1) * 8, pol_stride, have, samples,
                         s.path_count, (int)s.antithetic, s.seed, so, coords, spot, cash, partials, flags);
    else
      hipLaunchKernelGGL((lsmc_backward<FP32, KIND, BASIS, false>), dim3(blocks), dim3(256), lds, ctx->stream, m,
                         (const LsmcModel*)nullptr, 1, step, d_policy + (size_t)(step + 1) * 8, pol_stride, have,
                         samples, s.path_count, (int)s.antithetic, s.seed, so, coords, spot, cash, partials, flags);
    KG_HIP(hipGetLastError());
    hipLaunchKernelGGL((lsmc_solve<BASIS>), dim3(nl), dim3(256), 0, ctx->stream, partials, blocks,
                       d_policy + (size_t)step * 8, pol_stride, flags);
    KG_HIP(hipGetLastError());
    have = 1;  // the kernels read the valid flag of the row they use
    ++fitted;
  }
  Moments* parts = (Moments*)ctx->partials.reserve(sizeof(Moments) * 2 * (size_t)nl * blocks);
  if (lanes)
    hipLaunchKernelGGL((lsmc_final<BASIS, true>), dim3(blocks), dim3(256), 0, ctx->stream, m, d_lanes, nl,
                       (const double*)d_policy, pol_stride, 1, samples, s.path_count, (int)s.antithetic,
                       (const double*)spot, (const double*)cash, parts);
  else
    hipLaunchKernelGGL((lsmc_final<BASIS, false>), dim3(blocks), dim3(256), 0, ctx->stream, m,
                       (const LsmcModel*)nullptr, 1, (const double*)d_policy, pol_stride, 1, samples, s.path_count,
                       (int)s.antithetic, (const double*)spot, (const double*)cash, parts);
  KG_HIP(hipGetLastError());
  check_flags(ctx);
  std::vector<Moments> r(2 * (size_t)nl);
  merge_and_fetch(ctx, parts, blocks, 2 * nl, r.data());
  *dates_fitted = fitted;
  for (int l = 0; l < nl; ++l)
    values[l] = std::max(scaled_value(r[2 * l].scale, r[2 * l].mean, ml[l].log_discount), immediate[l]);
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
  clear_policy(ctx, policy, (int64_t)m.steps * 8);
  int32_t fitted = 0;
  timing_begin(ctx);
  double in_sample = 0.0;
  run_lsmc<FP32, KIND, BASIS>(ctx, &m, nullptr, 1, oos ? training : pricing, policy, &fitted, &immediate,
                              &in_sample);
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

// Lanes that can share one set of stored paths: in-sample polynomial fits
// at one valuation time, normalised (not anchored) paths, and either BSM
// Brownian coordinates (each lane its own spot, rates, volatility and
// strike) or the log-spots of one local-volatility market (strikes differ).
static bool lanes_share_paths(const std::vector<LsmcHost>& hosts, const kg_american_model* models,
                              const kg_lsmc_options& opt) {
  if (hosts.size() < 2 || opt.basis == KG_BASIS_MLP || opt.out_of_sample) return false;
  const LsmcHost& h0 = hosts[0];
  if (h0.known) return false;
  for (size_t l = 0; l < hosts.size(); ++l) {
    const LsmcHost& h = hosts[l];
    if (h.known || h.model.anchored || models[l].valuation_time != models[0].valuation_time) return false;
    if (h.model.lv != h0.model.lv || h.model.reverse != h0.model.reverse) return false;
    if (h.model.lv) {
      auto same = [](double a, double b) { return memcmp(&a, &b, sizeof(a)) == 0; };
      if (!same(h.model.spot, h0.model.spot) || !same(h.model.rate, h0.model.rate) ||
          !same(h.model.dividend, h0.model.dividend) || h.model.nt != h0.model.nt || h.model.ns != h0.model.ns)
        return false;
      const size_t n = h.tables.size() - h.lv_offset;
      if (n != h0.tables.size() - h0.lv_offset ||
          memcmp(h.tables.data() + h.lv_offset, h0.tables.data() + h0.lv_offset, n * sizeof(double)) != 0)
        return false;
    }
  }
  return true;
}

template <bool FP32, int KIND, int BASIS>
static void shared_lanes(kg_context* ctx, std::vector<LsmcHost>& hosts, kg_lsmc_report* out) {
  const int nl = (int)hosts.size();
  const int S = hosts[0].model.steps;
  // One upload: every lane's tables (256-byte aligned), then the lane models.
  std::vector<size_t> offset(nl);
  size_t doubles = 0;
  for (int l = 0; l < nl; ++l) {
    offset[l] = doubles;
    doubles += (hosts[l].tables.size() + 31) & ~size_t(31);
  }
  const size_t bytes = doubles * sizeof(double) + (size_t)nl * sizeof(LsmcModel);
  unsigned char* d = (unsigned char*)ctx->lsmc_tables.reserve(bytes);
  std::vector<unsigned char> blob(bytes, 0);
  std::vector<LsmcModel> ml(nl);
  std::vector<double> immediate(nl), values(nl);
  for (int l = 0; l < nl; ++l) {
    memcpy(blob.data() + offset[l] * sizeof(double), hosts[l].tables.data(), hosts[l].tables.size() * sizeof(double));
    ml[l] = hosts[l].model;
    bind_tables(ml[l], (const double*)d + offset[l], hosts[l].lv_offset);
    immediate[l] = hosts[l].immediate;
  }
  memcpy(blob.data() + doubles * sizeof(double), ml.data(), (size_t)nl * sizeof(LsmcModel));
  upload_async(ctx, d, blob.data(), bytes);
  const LsmcModel* d_lanes = (const LsmcModel*)(d + doubles * sizeof(double));
  double* policy = (double*)ctx->policy.reserve((size_t)nl * S * 8 * sizeof(double));
  clear_policy(ctx, policy, (int64_t)nl * S * 8);
  int32_t fitted = 0;
  timing_begin(ctx);
  run_lsmc<FP32, KIND, BASIS>(ctx, ml.data(), d_lanes, nl, hosts[0].pricing, policy, &fitted, immediate.data(),
                              values.data());
  const double ms = timing_end(ctx);
  for (int l = 0; l < nl; ++l) {
    out[l].price = out[l].in_sample_price = values[l];
    out[l].training_ms = ms;
    out[l].dates_fitted = fitted;
    out[l].path_steps = (double)hosts[0].pricing.path_count * S;
  }
}

template <int BASIS>
static void shared_lanes_basis(kg_context* ctx, std::vector<LsmcHost>& hosts, kg_lsmc_report* out) {
  const bool fp32 = hosts[0].pricing.fp32;
  if (hosts[0].model.lv)
    fp32 ? shared_lanes<true, PATHS_LOCAL_VOL, BASIS>(ctx, hosts, out)
         : shared_lanes<false, PATHS_LOCAL_VOL, BASIS>(ctx, hosts, out);
  else
    fp32 ? shared_lanes<true, PATHS_BROWNIAN, BASIS>(ctx, hosts, out)
         : shared_lanes<false, PATHS_BROWNIAN, BASIS>(ctx, hosts, out);
}

static void american_one(kg_context* ctx, LsmcHost& h, const kg_lsmc_options& options, kg_lsmc_report* out) {
  if (h.known) {
    out->price = out->in_sample_price = h.known_value;
    return;
  }
  if (options.basis == KG_BASIS_MLP)
    american_mlp(ctx, h, options, out);
  else if (options.basis == KG_BASIS_QUADRATIC)
    american_basis<KG_BASIS_QUADRATIC>(ctx, h, out);
  else
    american_basis<KG_BASIS_POLY4>(ctx, h, out);
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
    american_one(ctx, h, *options, out);
  });
}

extern "C" kg_status kg_american_price_lanes(kg_context* ctx, const kg_settings* settings,
                                             const kg_american_terms* option, const kg_american_model* models,
                                             const double* strikes, int32_t n_lanes, const kg_lsmc_options* options,
                                             kg_lsmc_report* out) {
  return guarded(ctx, [&] {
    if (!settings || !option || !models || !options || !out) fail(KG_INVALID_INPUT, "arguments must not be null");
    if (n_lanes < 1 || n_lanes > KG_MAX_AMERICAN_LANES)
      fail(KG_INVALID_INPUT, "n_lanes must be between 1 and KG_MAX_AMERICAN_LANES");
    for (int l = 0; l < n_lanes; ++l) out[l] = kg_lsmc_report{};
    // Every lane is validated before any is priced: a failed request
    // returns no partial prices.
    std::vector<LsmcHost> hosts;
    hosts.reserve(n_lanes);
    for (int l = 0; l < n_lanes; ++l) {
      kg_american_terms o = *option;
      if (strikes) o.strike = strikes[l];
      hosts.push_back(lsmc_plan_host(*settings, o, models[l], *options));
    }
    if (lanes_share_paths(hosts, models, *options)) {
      if (options->basis == KG_BASIS_QUADRATIC)
        shared_lanes_basis<KG_BASIS_QUADRATIC>(ctx, hosts, out);
      else
        shared_lanes_basis<KG_BASIS_POLY4>(ctx, hosts, out);
      return;
    }
    std::vector<kg_lsmc_report> priced(n_lanes);
    for (int l = 0; l < n_lanes; ++l) {
      check_cancel(ctx);
      kg_american_terms o = *option;
      if (strikes) o.strike = strikes[l];
      LsmcHost h = lsmc_plan(ctx, *settings, o, models[l], *options);
      american_one(ctx, h, *options, &priced[l]);
    }
    for (int l = 0; l < n_lanes; ++l) out[l] = priced[l];
  });
}
