// Test infrastructure, not product code: FP64 CPU Heston path stepping, the
// checker of the HIP Heston kernels.
//
// kiyosi has no Heston Monte Carlo; this restates the scheme the GPU engine
// documents (full-truncation Euler in log-spot, Lord, Koekkoek and van Dijk
// 2010) on kiyosi's sampler: path_samples (monte_carlo.rs:1498-1530) with
// 2 normals per step, normal 2k driving variance and 2k+1 the independent
// part of the spot shock.
#pragma once

#include "../include/kiyosi_amd.h"
#include "core.hpp"
#include "plans.hpp"

namespace oracle {

struct HestonPlan {
  int payoff, steps;
  double sign, strike, spot;
  double a, hdt, kdt, kth, xi, rho, rho_c, sdt, v0, log_discount;
};

inline HestonPlan heston_plan(const kg_heston_terms& o, const kg_heston_market& m) {
  if (o.option_type != 1 && o.option_type != -1)
    throw Failure(INVALID_INPUT, "option_type must be 1 (call) or -1 (put)");
  if (o.step_count < 1 || o.step_count > 10000)
    throw Failure(INVALID_INPUT, "step_count is outside the supported Monte Carlo range");
  if (m.valuation_time < o.effective_time || m.valuation_time >= o.expiry_time)
    throw Failure(INVALID_INPUT, "valuation outside contract life");
  double T = years_until(m.valuation_time, o.expiry_time);
  double dt = T / o.step_count;
  HestonPlan p;
  p.payoff = o.payoff;
  p.steps = o.step_count;
  p.sign = o.option_type;
  p.strike = o.strike;
  p.spot = m.spot;
  p.a = (m.rate - m.dividend) * dt;
  p.hdt = 0.5 * dt;
  p.kdt = m.kappa * dt;
  p.kth = m.kappa * m.theta * dt;
  p.xi = m.xi;
  p.rho = m.rho;
  p.rho_c = std::sqrt(1.0 - m.rho * m.rho);
  p.sdt = std::sqrt(dt);
  p.v0 = m.v0;
  p.log_discount = -m.rate * T;
  return p;
}

// One full-truncation Euler step of log-spot x and variance v on the normal
// pair (z[0] drives the variance, z[1] the independent part of the spot).
inline void heston_step(const HestonPlan& p, const double* z, double& x, double& v) {
  double vp = std::max(v, 0.0);
  double sq = std::sqrt(vp) * p.sdt;
  double zs = std::fma(p.rho, z[0], p.rho_c * z[1]);
  x += std::fma(sq, zs, std::fma(-p.hdt, vp, p.a));
  v += std::fma(p.xi * sq, z[0], std::fma(-p.kdt, vp, p.kth));
}

// One path over 2*steps normals; returns the discounted payoff.
inline double heston_path(const HestonPlan& p, const double* z) {
  double x = 0.0, v = p.v0, sum = 0.0;
  for (int k = 0; k < p.steps; ++k) {
    heston_step(p, z + 2 * k, x, v);
    if (p.payoff == KG_HESTON_ASIAN_ARITHMETIC)
      sum += std::exp(x);
    else if (p.payoff == KG_HESTON_ASIAN_GEOMETRIC)
      sum += x;
  }
  double underlying;
  if (p.payoff == KG_HESTON_EUROPEAN)
    underlying = p.spot * std::exp(x);
  else if (p.payoff == KG_HESTON_ASIAN_ARITHMETIC)
    underlying = p.spot * (sum / p.steps);
  else
    underlying = p.spot * std::exp(sum / p.steps);
  double pay = std::max(p.sign * (underlying - p.strike), 0.0);
  double r = pay * std::exp(p.log_discount);
  if (!std::isfinite(r)) unstable();
  return r;
}

inline Estimate heston_price(const Settings& st, const kg_heston_terms& o, const kg_heston_market& m,
                             const Context& cx) {
  HestonPlan p = heston_plan(o, m);
  Moments total[1], pairs[1];
  path_samples<1>(
      st, 2 * (size_t)p.steps, true, cx,
      [&](const double* z, double* out) { out[0] = heston_path(p, z); }, total, pairs);
  return estimate(st, total[0], pairs[0], 0.0, false, true, st.seed);
}

// Diagnostic of the scheme: estimates of E[S_T] and E[v_T] over the same
// paths heston_price draws (the terminal spot and variance of each path).
inline void heston_terminal(const Settings& st, const kg_heston_terms& o, const kg_heston_market& m,
                            const Context& cx, Estimate out[2]) {
  HestonPlan p = heston_plan(o, m);
  Moments total[2], pairs[2];
  path_samples<2>(
      st, 2 * (size_t)p.steps, true, cx,
      [&](const double* z, double* res) {
        double x = 0.0, v = p.v0;
        for (int k = 0; k < p.steps; ++k) heston_step(p, z + 2 * k, x, v);
        res[0] = p.spot * std::exp(x);
        res[1] = v;
      },
      total, pairs);
  for (int i = 0; i < 2; ++i) out[i] = estimate(st, total[i], pairs[i], 0.0, false, true, st.seed);
}

}  // namespace oracle
