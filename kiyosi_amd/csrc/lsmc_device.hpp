// Device code shared by the polynomial (lsmc.hip) and MLP (mlp.hip)
// regressors of the American engine: the local-volatility table, the
// reference's spot reconstruction, path generation and the final mean.
#pragma once

#include "fp32_math.hpp"
#include "lsmc.hpp"
#include "sobol_device.hpp"

namespace kg {

// LocalVolatilitySurface::at_year (market/local_volatility.rs:322-359).
__device__ inline int lv_bracket(const double* nodes, int n, double v, double* w) {
  int lo = 0, hi = n;  // partition_point(node <= v)
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (nodes[mid] <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  int left = lo > 0 ? lo - 1 : 0;
  if (left > n - 2) left = n - 2;
  double weight = (v - nodes[left]) / (nodes[left + 1] - nodes[left]);
  *w = fmin(fmax(weight, 0.0), 1.0);
  return left;
}
__device__ inline double lv_interp(double low, double high, double w) {
  return w <= 0.5 ? low + w * (high - low) : high + (1.0 - w) * (low - high);
}
__device__ inline double lv_sigma(const LsmcModel& m, double year, double spot) {
  double wt, ws;
  const int r = lv_bracket(m.years, m.nt, year, &wt);
  const int c = lv_bracket(m.spots, m.ns, spot, &ws);
  const double* v0 = m.vols + (size_t)r * m.ns;
  const double* v1 = v0 + m.ns;
  return lv_interp(lv_interp(v0[c], v0[c + 1], ws), lv_interp(v1[c], v1[c + 1], ws), wt);
}

// The two rows of a local-volatility table that bracket one time, with the
// spot nodes: in LDS when the table has at most LV_LDS_NODES spot nodes,
// otherwise in global memory. `row` is the staged row (-1 before the first).
constexpr int LV_LDS_NODES = 1024;

struct LvRows {
  const double* spots;
  const double* row0;
  const double* row1;
  double wt;
  int row;
  // FP32 copies for the FP32-step kernels (tables staged in LDS only): the
  // spot nodes, the reciprocal node spacings and the two rows.
  const float* spots32;
  const float* inv32;
  const float* row0_32;
  const float* row1_32;
  float wt32;
  bool f32;
};

// Stages the rows of time index i (year0 + dt i). Every thread of the block
// calls it with the same i; it synchronises the block only when the row
// changes. FP32 stages single-precision tables when the spot nodes fit in
// LDS, so that the FP32 steps interpolate without FP64 arithmetic.
template <bool FP32>
__device__ inline void lv_stage(const LsmcModel& m, int i, double* lds, LvRows& t) {
  double wt;
  const int r = lv_bracket(m.years, m.nt, m.year0 + m.dt * i, &wt);
  t.wt = wt;
  t.wt32 = (float)wt;
  t.f32 = false;
  if (m.ns > LV_LDS_NODES) {
    t.spots = m.spots;
    t.row0 = m.vols + (size_t)r * m.ns;
    t.row1 = t.row0 + m.ns;
    t.row = r;
    return;
  }
  if (FP32) {
    float* f = (float*)lds;  // spots, reciprocal spacings, row r, row r + 1
    t.f32 = true;
    t.spots32 = f;
    t.inv32 = f + m.ns;
    t.row0_32 = f + 2 * m.ns;
    t.row1_32 = f + 3 * m.ns;
    if (t.row == r) return;
    __syncthreads();
    if (t.row < 0)
      for (int k = threadIdx.x; k < m.ns; k += blockDim.x) {
        f[k] = (float)m.spots[k];
        f[m.ns + k] = k + 1 < m.ns ? (float)(1.0 / (m.spots[k + 1] - m.spots[k])) : 0.0f;
      }
    for (int k = threadIdx.x; k < 2 * m.ns; k += blockDim.x) f[2 * m.ns + k] = (float)m.vols[(size_t)r * m.ns + k];
    __syncthreads();
    t.row = r;
    return;
  }
  if (t.row == r) return;
  __syncthreads();
  if (t.row < 0)
    for (int k = threadIdx.x; k < m.ns; k += blockDim.x) lds[k] = m.spots[k];
  for (int k = threadIdx.x; k < 2 * m.ns; k += blockDim.x) lds[m.ns + k] = m.vols[(size_t)r * m.ns + k];
  __syncthreads();
  t.spots = lds;
  t.row0 = lds + m.ns;
  t.row1 = lds + 2 * m.ns;
  t.row = r;
}

// lv_sigma at the staged time for `spot` = S0 exp(log_rel). `c` carries
// the spot bracket between calls (-1 before the first: a binary search); on
// a log-uniform grid it starts from the grid index of log_rel. The walk
// leaves c = clamp(partition_point - 1, 0, ns - 2), the bracket lv_bracket
// returns.
__device__ inline double lv_sigma_at(const LsmcModel& m, const LvRows& t, double spot, double log_rel, int& c) {
  if (m.log_uniform) {
    const double g = floor((log_rel + m.log_spot - m.log_s0) * m.inv_h);
    c = (int)fmin(fmax(g, 0.0), (double)(m.ns - 2));
  } else if (c < 0) {
    double w;
    c = lv_bracket(t.spots, m.ns, spot, &w);
  }
  while (c > 0 && spot < t.spots[c]) --c;
  while (c < m.ns - 2 && spot >= t.spots[c + 1]) ++c;
  const double ws = fmin(fmax((spot - t.spots[c]) / (t.spots[c + 1] - t.spots[c]), 0.0), 1.0);
  return lv_interp(lv_interp(t.row0[c], t.row0[c + 1], ws), lv_interp(t.row1[c], t.row1[c + 1], ws), t.wt);
}

// The same in FP32 on the single-precision tables: the spot weight takes the
// staged reciprocal spacing instead of a divide.
__device__ inline float lv_interp32(float low, float high, float w) {
  return w <= 0.5f ? fmaf(w, high - low, low) : fmaf(1.0f - w, low - high, high);
}
__device__ inline float lv_sigma_at32(const LsmcModel& m, const LvRows& t, float spot, float log_rel, int& c) {
  const int ns = m.ns;
  if (m.log_uniform) {
    const float g = floorf((log_rel + (float)(m.log_spot - m.log_s0)) * (float)m.inv_h);
    c = (int)fminf(fmaxf(g, 0.0f), (float)(ns - 2));
  } else if (c < 0) {
    int lo = 0, hi = ns;  // first node above spot
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (t.spots32[mid] <= spot)
        lo = mid + 1;
      else
        hi = mid;
    }
    c = lo > 0 ? lo - 1 : 0;
    if (c > ns - 2) c = ns - 2;
  }
  while (c > 0 && spot < t.spots32[c]) --c;
  while (c < ns - 2 && spot >= t.spots32[c + 1]) ++c;
  const float ws = fminf(fmaxf((spot - t.spots32[c]) * t.inv32[c], 0.0f), 1.0f);
  return lv_interp32(lv_interp32(t.row0_32[c], t.row0_32[c + 1], ws), lv_interp32(t.row1_32[c], t.row1_32[c + 1], ws),
                     t.wt32);
}

// spots_at (american.rs:1166-1215): spot after step+1 intervals.
__device__ inline double bsm_spot(const LsmcModel& m, int step, double w) {
  const double lg = m.diffusion * w;
  if (m.screen[step] && m.sign * lg < m.group_cutoff[step]) return m.strike;
  return m.scale[step] * exp(lg);
}
__device__ inline double bsm_spot32(const LsmcModel& m, int step, float w) {
  const float lg = (float)m.diffusion * w;
  if (m.screen[step] && m.sign * (double)lg < m.group_cutoff[step]) return m.strike;
  return (double)((float)m.scale[step] * f32::exp(lg));
}

__device__ inline double payoff(const LsmcModel& m, double s) { return fmax(m.sign * (s - m.strike), 0.0); }

__device__ inline double continuation(const double* pol, int basis, double x) {
  if (basis == KG_BASIS_QUADRATIC) return pol[1] + x * (pol[2] + x * pol[3]);
  return pol[1] + x * (pol[2] + x * (pol[3] + x * (pol[4] + x * pol[5])));
}

// One log-Euler step of x = ln(S / S0) under local volatility from the
// start-of-step sigma; `t` holds the rows staged for the step's start.
template <bool FP32>
__device__ inline double lv_step(const LsmcModel& m, const LvRows& t, double x, double z, int& c) {
  if (FP32) {
    const float xf = (float)x;
    const float sig = t.f32 ? lv_sigma_at32(m, t, (float)m.spot * f32::exp(xf), xf, c)
                            : (float)lv_sigma_at(m, t, m.spot * (double)f32::exp(xf), x, c);
    const float dt = (float)m.dt;
    return (double)((float)x + fmaf(sig * f32::sqrt(dt), (float)z,
                                    ((float)(m.rate - m.dividend) - 0.5f * sig * sig) * dt));
  }
  const double sig = lv_sigma_at(m, t, m.spot * exp(x), x, c);
  return x + fma(sig * sqrt(m.dt), z, (m.rate - m.dividend - 0.5 * sig * sig) * m.dt);
}

template <bool FP32>
struct Normals {
  NormalStream s;
  Pcg32 g;
  float spare;
  bool has;
  __device__ inline void start(uint64_t seed, uint64_t block, uint64_t first) {
    if (FP32) {
      g.init(seed, block);
      g.advance((first >> 1) * 4);
      has = false;
      if (first & 1) {
        float e;
        f32::box_muller_pair(g, &e, &spare);
        has = true;
      }
    } else {
      s.start(seed, block, first);
    }
  }
  __device__ inline double next() {
    if (!FP32) return s.next();
    if (has) {
      has = false;
      return (double)spare;
    }
    float e;
    f32::box_muller_pair(g, &e, &spare);
    has = true;
    return (double)e;
  }
};

// The normals of one forward path: the sample's run of its block's PCG
// stream (normal j * steps onwards for sample j), or dimensions 0, 1, ..
// of Sobol point n (the reference's cursor, engines/monte_carlo/sobol.rs).
template <bool FP32>
struct PathDraws {
  Normals<FP32> pcg;
  SobolRef so;
  uint64_t point;
  int dim;
  __device__ inline void start(uint64_t seed, uint64_t block, int j, int steps, int64_t n, const SobolRef& s) {
    so = s;
    point = (uint64_t)n;
    dim = 0;
    if (!so.dirs) pcg.start(seed, block, (uint64_t)j * (uint64_t)steps);
  }
  __device__ inline double next() {
    if (!so.dirs) return pcg.next();
    const uint64_t bits = sobol_point_direct(so.dirs, so.shifts, so.dims, point, dim++);
    return FP32 ? (double)f32::normal_from_bits((uint32_t)(bits >> 20)) : normal_from_bits(bits);
  }
};

// Normal `row` of a block's reverse-mode draws for lane j
// (ExercisePaths::reverse_row, american.rs:278-316): pseudorandom draws take
// the block's stream step-major, normal row * count + j of Pcg32(seed,
// block) with `count` samples in the block; Sobol draws take dimension
// `row` of point n. Each lane forms its own Box-Muller pair and keeps the
// half its index selects.
template <bool FP32>
__device__ inline double reverse_normal(uint64_t seed, int64_t block, int j, int count, int row, int64_t n,
                                        const SobolRef& so) {
  if (so.dirs) {
    const uint64_t bits = sobol_point_direct(so.dirs, so.shifts, so.dims, (uint64_t)n, row);
    return FP32 ? (double)f32::normal_from_bits((uint32_t)(bits >> 20)) : normal_from_bits(bits);
  }
  const uint64_t k = (uint64_t)row * (uint64_t)count + (uint64_t)j;
  Pcg32 g;
  g.init(seed, (uint64_t)block);
  g.advance((k >> 1) * 4);
  if (FP32) {
    float c, s;
    f32::box_muller_pair(g, &c, &s);
    return (double)((k & 1) ? s : c);
  }
  double c, s;
  box_muller_pair(g, &c, &s);
  return (k & 1) ? s : c;
}

// One reverse-mode row: the coordinate after step + 1 intervals from the
// one after step + 2 (W_m | W_(m+1) has mean m/(m+1) W_(m+1) and variance
// m/(m+1) in units of one interval), or the terminal coordinate
// sqrt(steps) z at the last step.
template <bool FP32>
__device__ inline double bridge_back(double w, double z, int step, int steps) {
  const int elapsed = step + 1;
  double weight, deviation;
  if (elapsed == steps) {
    weight = 0.0;
    deviation = sqrt((double)elapsed);
  } else {
    const double fraction = (double)elapsed / (double)(elapsed + 1);
    weight = fraction;
    deviation = sqrt(fraction);
  }
  if (FP32) return (double)fmaf((float)weight, (float)w, (float)deviation * (float)z);
  return fma(weight, w, deviation * z);
}

// Path generation in forward mode: Brownian coordinates [step][sample],
// local-vol log-spots [step][slot], or anchored spots [step][slot] (FP64),
// then each slot's terminal spot and payoff.
template <bool FP32, int KIND>
__global__ void __launch_bounds__(256)
    lsmc_generate(LsmcModel m, uint64_t seed, int64_t samples, int64_t paths, int antithetic, SobolRef so,
                  void* __restrict__ coords, double* __restrict__ spot, double* __restrict__ cash,
                  uint32_t* __restrict__ flags) {
  constexpr bool LV = KIND == PATHS_LOCAL_VOL;
  __shared__ double lv_lds[LV ? 3 * LV_LDS_NODES : 1];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  const bool live = n < samples;
  const bool paired = antithetic && live && n < paths / 2;
  const int legs = antithetic ? 2 : 1;
  const int64_t slots = samples * legs;
  PathDraws<FP32> z;
  if (live) z.start(seed, (uint64_t)block, j, m.steps, n, so);
  LvRows rows{};
  rows.row = -1;
  int cp = -1, cn = -1;  // spot brackets of the two legs
  double w = 0.0, xp = 0.0, xn = 0.0;
  double ap = m.spot, an = m.spot;  // anchored spots of the two legs
  bool bad = false;
  for (int k = 0; k < m.steps; ++k) {
    if (LV) lv_stage<FP32>(m, k, lv_lds, rows);  // block-uniform
    if (!live) continue;
    const double e = z.next();
    if (KIND == PATHS_ANCHORED) {
      // advance + evolve_spot (american.rs:355-362, 546-552)
      ap *= exp(fma(m.diffusion, e, m.drift));
      bad = bad || !isfinite(ap) || ap <= 0.0;
      ((double*)coords)[(size_t)k * slots + n * legs] = ap;
      if (paired) {
        an *= exp(fma(m.diffusion, -e, m.drift));
        bad = bad || !isfinite(an) || an <= 0.0;
        ((double*)coords)[(size_t)k * slots + n * legs + 1] = an;
      }
      continue;
    }
    if (LV) {
      xp = lv_step<FP32>(m, rows, xp, e, cp);
      if (paired) xn = lv_step<FP32>(m, rows, xn, -e, cn);
      if (FP32) {
        ((float*)coords)[(size_t)k * slots + n * legs] = (float)xp;
        if (antithetic) ((float*)coords)[(size_t)k * slots + n * legs + 1] = (float)xn;
      } else {
        ((double*)coords)[(size_t)k * slots + n * legs] = xp;
        if (antithetic) ((double*)coords)[(size_t)k * slots + n * legs + 1] = xn;
      }
    } else {
      w += e;
      if (FP32)
        ((float*)coords)[(size_t)k * samples + n] = (float)w;
      else
        ((double*)coords)[(size_t)k * samples + n] = w;
    }
  }
  if (bad) raise_flag(flags, FLAG_NONFINITE);
  if (!live) return;
  for (int leg = 0; leg < (paired ? 2 : 1); ++leg) {
    double s;
    if (KIND == PATHS_ANCHORED)
      s = leg ? an : ap;
    else if (LV)
      s = m.spot * exp(leg ? xn : xp);
    else
      s = FP32 ? bsm_spot32(m, m.steps - 1, (float)(leg ? -w : w)) : bsm_spot(m, m.steps - 1, leg ? -w : w);
    spot[n * legs + leg] = s;
    cash[n * legs + leg] = payoff(m, s);
  }
}

// Reverse mode, first row: each sample's terminal Brownian coordinate
// (one value per sample, FP32 or FP64) and its slots' terminal spot and
// payoff.
template <bool FP32>
__global__ void __launch_bounds__(256)
    lsmc_reverse_start(LsmcModel m, uint64_t seed, int64_t samples, int64_t paths, int antithetic, SobolRef so,
                       void* __restrict__ coords, double* __restrict__ spot, double* __restrict__ cash) {
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  if (n >= samples) return;
  const int count = (int)(samples - block * BLOCK_SAMPLES < BLOCK_SAMPLES ? samples - block * BLOCK_SAMPLES
                                                                          : BLOCK_SAMPLES);
  const int last = m.steps - 1;
  const double w = bridge_back<FP32>(0.0, reverse_normal<FP32>(seed, block, j, count, 0, n, so), last, m.steps);
  if (FP32)
    ((float*)coords)[n] = (float)w;
  else
    ((double*)coords)[n] = w;
  const bool paired = antithetic && n < paths / 2;
  const int legs = antithetic ? 2 : 1;
  for (int leg = 0; leg < (paired ? 2 : 1); ++leg) {
    const double s = FP32 ? bsm_spot32(m, last, (float)(leg ? -w : w)) : bsm_spot(m, last, leg ? -w : w);
    spot[n * legs + leg] = s;
    cash[n * legs + leg] = payoff(m, s);
  }
}

// Final exercise at the first date and the mean of the path cash in slot
// order (american.rs:1437-1454), for one lane: its moments go to
// out[0] (block values) and out[1] (unused, zero).
template <int BASIS>
__device__ inline void final_lane(const LsmcModel& m, const double* pol, int have, int64_t n, int64_t samples,
                                  int64_t paths, int antithetic, const double* __restrict__ spot,
                                  const double* __restrict__ cash, double* vals, Moments* smem,
                                  Moments* __restrict__ out) {
  const int j = threadIdx.x;
  const int legs = antithetic ? 2 : 1;
  for (int leg = 0; leg < 2; ++leg) {
    double v = NAN;
    const bool on = n < samples && (leg == 0 || (antithetic && n < paths / 2));
    if (on) {
      const int64_t slot = n * legs + leg;
      double c = cash[slot];
      if (have && pol[0] != 0.0) {
        const double sp = spot[slot];
        const double intrinsic = payoff(m, sp);
        if (intrinsic > 0.0 && intrinsic / m.strike > continuation(pol, BASIS, sp / m.strike)) c = intrinsic;
      }
      v = c;
    }
    vals[2 * j + leg] = v;
  }
  __syncthreads();
  if (j < WAVE) {
    Moments acc = moments_zero();
    for (int s = 8 * j; s < 8 * j + 8; ++s)
      if (!isnan(vals[s])) moments_add(acc, vals[s], true);
    acc = wave_merge(acc);
    if (j == 0) {
      out[0] = acc;
      out[1] = moments_zero();
    }
  }
  __syncthreads();  // vals is rewritten by the next lane
}

// One workgroup per sample block. With LANES, lane l reads its model from
// lanes[l], its policy row 0 at l * pol_stride and its spot and cash at
// l * slots; block b writes lane l's moments to parts[(b * nl + l) * 2].
template <int BASIS, bool LANES>
__global__ void __launch_bounds__(256)
    lsmc_final(LsmcModel m, const LsmcModel* __restrict__ lanes, int nl, const double* __restrict__ pol,
               int64_t pol_stride, int have, int64_t samples, int64_t paths, int antithetic,
               const double* __restrict__ spot, const double* __restrict__ cash, Moments* __restrict__ parts) {
  __shared__ double vals[512];
  __shared__ Moments smem[4];
  const int64_t block = blockIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + threadIdx.x;
  if (LANES) {
    const int64_t slots = samples * (antithetic ? 2 : 1);
    for (int l = 0; l < nl; ++l) {
      const LsmcModel ml = lanes[l];
      final_lane<BASIS>(ml, pol + l * pol_stride, have, n, samples, paths, antithetic, spot + l * slots,
                        cash + l * slots, vals, smem, parts + (block * nl + l) * 2);
    }
  } else {
    final_lane<BASIS>(m, pol, have, n, samples, paths, antithetic, spot, cash, vals, smem, parts + block * 2);
  }
}

// Lanes 1 .. nl - 1 of a shared-path request: each slot's terminal spot and
// payoff from the stored coordinates, as lsmc_generate and
// lsmc_reverse_start form them for lane 0.
template <bool FP32, int KIND>
__global__ void __launch_bounds__(256)
    lsmc_lane_terminal(const LsmcModel* __restrict__ lanes, int nl, int64_t samples, int64_t paths, int antithetic,
                       int reverse, const void* __restrict__ coords, double* __restrict__ spot,
                       double* __restrict__ cash) {
  const int64_t n = (int64_t)blockIdx.x * BLOCK_SAMPLES + threadIdx.x;
  if (n >= samples) return;
  const bool paired = antithetic && n < paths / 2;
  const int legs = antithetic ? 2 : 1;
  const int64_t slots = samples * legs;
  for (int l = 1; l < nl; ++l) {
    const LsmcModel m = lanes[l];
    const int last = m.steps - 1;
    for (int leg = 0; leg < (paired ? 2 : 1); ++leg) {
      const int64_t slot = n * legs + leg;
      double s;
      if (KIND == PATHS_LOCAL_VOL) {
        const size_t at = (size_t)last * slots + slot;
        const double x = FP32 ? (double)((const float*)coords)[at] : ((const double*)coords)[at];
        s = m.spot * exp(x);
      } else {
        const size_t at = reverse ? (size_t)n : (size_t)last * samples + n;
        const double w = FP32 ? (double)((const float*)coords)[at] : ((const double*)coords)[at];
        s = FP32 ? bsm_spot32(m, last, (float)(leg ? -w : w)) : bsm_spot(m, last, leg ? -w : w);
      }
      spot[l * slots + slot] = s;
      cash[l * slots + slot] = payoff(m, s);
    }
  }
}

}  // namespace kg
