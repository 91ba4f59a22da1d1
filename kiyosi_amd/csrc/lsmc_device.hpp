// Device code shared by the polynomial (lsmc.hip) and MLP (mlp.hip)
// regressors of the American engine: the local-volatility table, the
// reference's spot reconstruction, path generation and the final mean.
#pragma once

#include "fp32_math.hpp"
#include "lsmc.hpp"

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

// One log-Euler step under local volatility from the start-of-step sigma.
template <bool FP32>
__device__ inline double lv_step(const LsmcModel& m, int k, double x, double z) {
  const double year = m.year0 + m.dt * k;
  if (FP32) {
    const float sig = (float)lv_sigma(m, year, m.spot * (double)f32::exp((float)x));
    const float dt = (float)m.dt;
    return (double)((float)x + fmaf(sig * f32::sqrt(dt), (float)z,
                                    ((float)(m.rate - m.dividend) - 0.5f * sig * sig) * dt));
  }
  const double sig = lv_sigma(m, year, m.spot * exp(x));
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

template <bool FP32, bool LV>
__global__ void __launch_bounds__(256)
    lsmc_generate(LsmcModel m, uint64_t seed, int64_t samples, int64_t paths, int antithetic,
                  void* __restrict__ coords, double* __restrict__ spot, double* __restrict__ cash) {
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
  if (n >= samples) return;
  const bool paired = antithetic && n < paths / 2;
  const int legs = antithetic ? 2 : 1;
  const int64_t slots = samples * legs;
  Normals<FP32> z;
  z.start(seed, (uint64_t)block, (uint64_t)j * (uint64_t)m.steps);
  double w = 0.0, xp = 0.0, xn = 0.0;
  for (int k = 0; k < m.steps; ++k) {
    const double e = z.next();
    if (LV) {
      xp = lv_step<FP32>(m, k, xp, e);
      if (paired) xn = lv_step<FP32>(m, k, xn, -e);
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
  for (int leg = 0; leg < (paired ? 2 : 1); ++leg) {
    double s;
    if (LV)
      s = m.spot * exp(leg ? xn : xp);
    else
      s = FP32 ? bsm_spot32(m, m.steps - 1, (float)(leg ? -w : w)) : bsm_spot(m, m.steps - 1, leg ? -w : w);
    spot[n * legs + leg] = s;
    cash[n * legs + leg] = payoff(m, s);
  }
}

// Final exercise at the first date and the mean of the path cash in slot
// order (american.rs:1437-1454).
template <int BASIS>
__global__ void __launch_bounds__(256)
    lsmc_final(LsmcModel m, const double* __restrict__ pol, int have, int64_t samples, int64_t paths,
               int antithetic, const double* __restrict__ spot, const double* __restrict__ cash,
               Moments* __restrict__ parts) {
  __shared__ double vals[512];
  __shared__ Moments smem[4];
  const int64_t block = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t n = block * BLOCK_SAMPLES + j;
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
      parts[block * 2] = acc;
      parts[block * 2 + 1] = moments_zero();
    }
  }
}

}  // namespace kg
