// Device path code of the barrier, Asian and accumulator engines, written
// against the Draws normal source of autocall_common.hpp so that one
// persistent kernel (book.hip) runs every contract type.
//
//   barrier_path   MonteCarloBarrierEngine sample closure
//                  (engines/monte_carlo/barrier.rs:291-477), single and
//                  double barriers, optionally with likelihood-ratio Greeks
//   digital_path   MonteCarloDigitalEngine (engines/monte_carlo/digital.rs)
//   geo_asian      GeometricParameters::geometric (asian.rs:741-784)
//   arith_asian    ArithmeticGrid::path + ArithmeticParameters::arithmetic
//                  with the geometric control residual (asian.rs:640-712,
//                  1100-1256)
//   accum_path     MonteCarloAccumulatorEngine sample closure
//                  (engines/monte_carlo/accumulator.rs:180-216)
#pragma once

#include "autocall_common.hpp"

namespace kg {

// ------------------------------------------------------------------ numerics
// numerics.rs:192-215, 320-334; asian.rs:971-984; numerics.rs:160-167, 217-224
__device__ inline double d_normal_cdf(double x) { return 0.5 * erfc(-x * 0.7071067811865476); }
__device__ inline double d_log_normal_cdf(double value) {
  if (value >= -10.0) return log(d_normal_cdf(value));
  const double inv_sq = (1.0 / value) / value;
  double term = 1.0, sum = 1.0;
  for (int i = 1; i <= 20; ++i) {
    term *= -(double)(2 * i - 1) * inv_sq;
    sum += term;
  }
  return -0.5 * value * value - log(-value) + log(sum / 2.5066282746310002);
}
__device__ inline double d_log_add(double a, double b) {
  const double high = fmax(a, b);
  if (high == -INFINITY) return high;
  return high + log1p(exp(-fabs(a - b)));
}
__device__ inline double d_log_difference(double high, double low) {
  if (high == -INFINITY) return -INFINITY;
  return high + log(-expm1(low - high));
}
__device__ inline double d_positive_difference(double a, double b) {
  if (a == INFINITY) return a;
  if (b == -INFINITY) return exp(a);
  const double d = -expm1(b - a);
  return d > 0.0 ? exp(a + log(d)) : 0.0;
}
__device__ inline double d_weighted_cdf(double amount, double exponent, double argument) {
  const double discount = exp(exponent);
  const double discounted = amount * discount;
  if (argument < -10.0 || !isnormal(discount) || !isfinite(discounted))
    return exp(log(amount) + exponent + d_log_normal_cdf(argument));
  return discounted * d_normal_cdf(argument);
}
// lognormal_price (engines/analytic/european.rs:317-400) for widths >= 1e-5;
// narrower widths are rejected on the host.
__device__ inline double d_lognormal_price(double spot, double strike, double log_moneyness,
                                           double asset_log, double risk_log, double vol,
                                           double root_time, double sign) {
  const double width = vol * root_time;
  const double st = log_moneyness / width;
  const double d1 = st + 0.5 * width, d2 = st - 0.5 * width;
  const double asset = d_weighted_cdf(spot, asset_log, sign * d1);
  const double cash = d_weighted_cdf(strike, risk_log, sign * d2);
  const double direct = sign * (asset - cash);
  if (isfinite(direct)) return direct;
  const double a = log_moneyness + d_log_normal_cdf(sign * d1);
  const double c = d_log_normal_cdf(sign * d2);
  return exp(log(strike) + risk_log + (sign > 0 ? d_log_difference(a, c) : d_log_difference(c, a)));
}

// ------------------------------------------------------------------ barrier
// double_barrier::survival (engines/monte_carlo/double_barrier.rs:6-112): the
// probability that the Brownian bridge from `initial` to `terminal` (log
// distances above the lower level) stays inside (0, band), and its first two
// derivatives in `initial` with the terminal value held fixed. Image series
// for small variance, spectral series otherwise; a series that does not
// converge raises `bad`, the reference's unstable().
__device__ inline void d_double_survival(double initial, double terminal, double band, double deviation,
                                         bool derivatives, double out[3], uint32_t& bad) {
  out[0] = out[1] = out[2] = 0.0;
  if (terminal <= 0.0 || terminal >= band) return;
  double x = initial / band, y = terminal / band;
  const double ratio = deviation / band;
  const double variance = ratio * ratio;
  if (variance == 0.0) {
    out[0] = 1.0;
    return;
  }
  if (!isfinite(variance) || variance <= 0.0) {
    bad = 1;
    return;
  }
  const bool reflected = x + y > 1.0;
  if (reflected) {
    x = 1.0 - x;
    y = 1.0 - y;
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  bool converged = false;
  if (variance < 0.25) {
    const double inverse_variance = isnormal(variance) ? 1.0 / variance : INFINITY;
    s0 = -expm1(-2.0 * (x / ratio) * (y / ratio));
    // One image term: k shifts the band, `direct` picks the unreflected image.
    auto add = [&](double k, bool direct) {
      const double exponent = direct ? -2.0 * (k / ratio) * ((y - x + k) / ratio)
                                     : -2.0 * ((x + k) / ratio) * ((y + k) / ratio);
      const double sign = direct ? 1.0 : -1.0;
      const double weight = exp(exponent);
      if (k != 0.0) s0 += sign * weight;
      if (derivatives) {
        const double numerator = direct ? 2.0 * k : -2.0 * (y + k);
        if (numerator != 0.0) {
          const double slope = numerator * inverse_variance;
          const double first = weight * slope;
          const double second = first * slope;
          if (isnormal(weight) && isnormal(first) && isnormal(second)) {
            s1 += sign * first;
            s2 += sign * second;
          } else {
            const double log_slope = log(fabs(numerator)) - 2.0 * log(ratio);
            s1 += sign * copysign(exp(exponent + log_slope), numerator);
            s2 += sign * exp(exponent + 2.0 * log_slope);
          }
        }
      }
      return weight;
    };
    add(0.0, false);
    for (int n = 1; n <= 64; ++n) {
      // The four images of n in the reference's order, (k, direct), (k,
      // reflected), (-k, direct), (-k, reflected), summed left to right; one
      // at a time, so that their exponentials do not overlap in registers
      // (the book kernel inlines this beside every contract type).
      double tail = 0.0;
#pragma unroll 1
      for (int t = 0; t < 4; ++t) tail += add(t < 2 ? (double)n : -(double)n, (t & 1) == 0);
      if (n >= 2 && tail < 1e-18) {
        converged = true;
        break;
      }
    }
  } else {
    const double distance = y - x;
    const double log_scale = log(2.0 * 2.5066282746310002 * ratio) + 0.5 * distance * distance / variance;
    const double score = -distance / variance;
    for (int n = 1; n <= 256; ++n) {
      const double frequency = (double)n * M_PI;
      const double exponent = log_scale - 0.5 * frequency * frequency * variance;
      const double weight = exp(exponent);
      const double sy = sin(frequency * y);
      const double product = sin(frequency * x) * sy;
      s0 += weight * product;
      if (derivatives) {
        const double first = frequency * cos(frequency * x) * sy;
        s1 += weight * (first + score * product);
        s2 += weight * (2.0 * score * first + (score * score + 1.0 / variance - frequency * frequency) * product);
      }
      if (n >= 4 && exponent + 4.0 * log(frequency) < -40.0) {
        converged = true;
        break;
      }
    }
  }
  if (!converged || !(s0 >= -1e-13 && s0 <= 1.0 + 1e-13)) {
    bad = 1;
    return;
  }
  out[0] = fmin(fmax(s0, 0.0), 1.0);
  if (derivatives) {
    out[1] = (reflected ? -s1 : s1) / band;
    out[2] = (s2 / band) / band;
  }
}

struct BarrierStep {
  double mean, width, discount, lower, upper, remaining;
  int32_t observe;
  int32_t pad;
};
// A single barrier, or a double barrier as its lower (down-type) level plus
// `upper` (has_upper): instruments/double_barrier.rs:119-144.
struct BarrierDev {
  int32_t n_steps, up, knock_in, continuous, at_hit, condition_on_hit, bounded, has_upper;
  int32_t request, pad;  // KG_GREEK_* bits of the spot-Greek kernel; 0 in the book
  double sign, strike, rebate, spot, q, sigma, hvr, moneyness, barrier_abs, terminal_barrier;
  double log_discount, initial, band;
  const BarrierStep* steps;
  const float* steps32;  // mean, width, lower, upper per step
};

// MonteCarloBarrierEngine's sample closure (barrier.rs:291-477). With LR the
// likelihood-ratio delta and gamma of the path go to lr[0], lr[1]
// (:440-472); otherwise only the value is formed. SURVIVAL compiles the
// double-barrier survival series for continuously monitored double barriers;
// the persistent book kernel leaves it out (its FP64 temporaries would raise
// the register count of every contract type there) and never receives such
// contracts.
template <bool FP32, bool LR, bool SURVIVAL, class D>
__device__ double barrier_path(const BarrierDev& b, D& dr, bool live, float sign, uint32_t& bad,
                               double* lr = nullptr) {
  using R = typename std::conditional<FP32, float, double>::type;
  R noise = 0, terminal = 0;
  bool touched = false, has_cond = false, active = live;
  double knocked_out = 0.0, conditioned = 0.0;
  // The step and state where the path knocked out or was conditioned on the
  // hit; the FP64 values are formed after the loop, once for all lanes,
  // instead of once per lane inside it.
  int hit = -1;
  R hit_noise = 0, hit_terminal = 0, z0 = 0;
  // The loop body is branch-free per lane: every lane draws its normals and
  // a finished lane keeps its state through selects, so the only branches
  // are on contract fields (wave-uniform) and the wave's early exit. The
  // lane's flags live in one integer (kept in a vector register) and the
  // steps go two per iteration, one Box-Muller pair per iteration.
  enum : uint32_t { ACTIVE = 1, TOUCHED = 2, COND = 4 };
  uint32_t st = active ? ACTIVE : 0u;
  // The reference tests every step's spot for overflow; the spot is
  // monotone in the log coordinate, so the lowest and highest visited
  // coordinates decide that test for all of them (a NaN reaches the last
  // coordinate, which the payoff tests).
  R low = terminal, high = terminal;
  const int n_steps = b.n_steps;
  const bool up = b.up, has_upper = b.has_upper, knock_in = b.knock_in;
  const bool condition_on_hit = b.condition_on_hit, bounded = b.bounded;
  auto step = [&](int i, R z) {
    // FP32: the step's row (16 bytes); FP64: the step record.
    R width, mean, lower = 0, upper = 0;
    bool observe;
    if constexpr (FP32) {
      const float* row = b.steps32 + 4 * i;  // one 16-byte scalar load
      mean = uload(row);
      width = uload(row + 1);
      lower = uload(row + 2);
      upper = uload(row + 3);
      observe = true;  // rows of steps without a fixing hold NaN levels
    } else {
      const BarrierStep& s = b.steps[i];
      width = uload(&s.width);
      mean = uload(&s.mean);
      observe = uload(&s.observe);
      if (observe) {
        lower = uload(&s.lower);
        upper = uload(&s.upper);
      }
    }
    if (LR && i == 0) z0 = z;
    const bool on = st & ACTIVE;
    const R next = FP32 ? (R)fmaf((float)width, (float)z, (float)noise) : (R)fma((double)width, (double)z, (double)noise);
    noise = on ? next : noise;
    terminal = on ? mean + next : terminal;
    low = fmin(low, terminal);
    high = fmax(high, terminal);
    if (observe) {
      bool crossed;
      if (has_upper)
        crossed = noise <= -lower || noise >= -upper;
      else
        crossed = up ? noise >= -lower : noise <= -lower;
      crossed = crossed && on;
      // A knock-in conditioned on the hit stops the path at the hit.
      const bool cond = crossed && condition_on_hit && n_steps - i > 8;
      hit = cond ? i : hit;
      hit_noise = cond ? noise : hit_noise;
      hit_terminal = cond ? terminal : hit_terminal;
      st = (st | (crossed ? TOUCHED : 0u) | (cond ? COND : 0u)) & (cond ? ~ACTIVE : ~0u);
    }
    const bool out = !knock_in && (st & (ACTIVE | TOUCHED)) == (ACTIVE | TOUCHED);
    hit = out ? i : hit;
    st &= out ? ~ACTIVE : ~0u;
  };
  if constexpr (LR) {
    // The spot-Greek kernel steps one day at a time (its register budget).
    for (int i = 0; i < n_steps; ++i) {
      if (__ballot(st & ACTIVE) == 0) break;
      step(i, (R)sign * dr.template normal<R>(i));
    }
  } else {
    int i = 0;
    bool running = true;
    for (; i + 1 < n_steps; i += 2) {
      if (__ballot(st & ACTIVE) == 0) {
        running = false;
        break;
      }
      R za, zb;
      dr.template normal_pair<R>(i, za, zb);
      step(i, (R)sign * za);
      step(i + 1, (R)sign * zb);
    }
    if (running && i < n_steps && __ballot(st & ACTIVE) != 0) step(i, (R)sign * dr.template normal<R>(i));
  }
  active = st & ACTIVE;
  touched = st & TOUCHED;
  has_cond = st & COND;
  if (!bounded && live) {
    for (const R t : {low, high}) {
      const double sp = discounted_product(b.spot, 1.0, (double)t);
      if (!isfinite(sp) || sp <= 0.0) bad = 1;
    }
  }
  if (has_cond) {
    const BarrierStep& s = b.steps[hit];
    const double asset = discounted_product(b.spot, 1.0, (double)hit_terminal);
    conditioned = d_lognormal_price(asset, b.strike, b.moneyness + (double)hit_noise + b.hvr * s.remaining,
                                    s.discount - b.q * s.remaining, b.log_discount, b.sigma, sqrt(s.remaining),
                                    b.sign);
  } else if (hit >= 0) {
    knocked_out = discounted_product(b.rebate, 1.0, b.at_hit ? b.steps[hit].discount : b.log_discount);
  }
  if (LR) lr[0] = lr[1] = 0.0;
  if (!live) return 0.0;
  const bool want_delta = LR && (b.request & 1), want_gamma = LR && (b.request & 2);
  const double asset = discounted_product(b.spot, 1.0, (double)terminal);
  const double distance = b.moneyness + (double)noise;
  const double payoff = fabs(distance) < 0.5 ? b.strike * expm1(distance) : asset - b.strike;
  const double intrinsic = fmax(b.sign * payoff, 0.0);
  auto weighted = [&](double amount, double lw) {
    return discounted_product(fabs(amount), amount < 0.0 ? -1.0 : 1.0, b.log_discount + lw);
  };
  double value, first = 0.0, second = 0.0;
  if (has_cond) {
    value = conditioned;
  } else if (SURVIVAL && b.continuous && b.has_upper) {
    double sv[3];
    d_double_survival(b.initial, b.terminal_barrier + (double)noise, b.band, b.steps[0].width,
                      want_delta || want_gamma, sv, bad);
    const double rebate = (!b.knock_in && b.at_hit) ? 0.0 : b.rebate;
    value = weighted(intrinsic, b.knock_in ? log1p(-sv[0]) : log(sv[0])) +
            weighted(rebate, b.knock_in ? log(sv[0]) : log1p(-sv[0]));
    if (want_delta || want_gamma) {
      const double difference = (b.knock_in ? -1.0 : 1.0) * (intrinsic - rebate);
      auto derivative = [&](double w) {
        return discounted_product(fabs(difference), (difference < 0.0 ? -1.0 : 1.0) * (w < 0.0 ? -1.0 : 1.0),
                                  b.log_discount + log(fabs(w)));
      };
      first = derivative(sv[1]);
      if (want_gamma) second = derivative(sv[2]);
    }
  } else if (b.continuous) {
    const double w0 = b.steps[0].width;
    const double dtb = b.terminal_barrier + (double)noise;
    const bool alive = b.up ? dtb < 0.0 : dtb > 0.0;
    const double slope = alive ? 2.0 * (fabs(dtb) / w0) / w0 : 0.0;
    const double exponent = -b.barrier_abs * slope;
    const double survival = alive ? log(-expm1(exponent)) : -INFINITY;
    const double rebate = (!b.knock_in && b.at_hit) ? 0.0 : b.rebate;
    value = weighted(intrinsic, b.knock_in ? exponent : survival) +
            weighted(rebate, b.knock_in ? survival : exponent);
    if ((want_delta || want_gamma) && slope > 0.0 && exponent > -INFINITY) {
      const double difference = (b.knock_in ? -1.0 : 1.0) * (intrinsic - rebate);
      first = weighted(b.up ? -difference : difference, exponent + log(slope));
      if (want_gamma) second = weighted(-difference, exponent + 2.0 * log(slope));
    }
  } else if (touched && !b.knock_in) {
    value = knocked_out;
  } else {
    value = weighted(touched == (bool)b.knock_in ? intrinsic : b.rebate, 0.0);
  }
  if (!isfinite(value)) bad = 1;
  if (LR) {
    const double width = b.steps[0].width, z = (double)z0, score = z / width;
    if (want_delta) lr[0] = (value * score + first) / b.spot;
    if (want_gamma) {
      const double second_score = ((z * z - 1.0) / width - z) / width;
      lr[1] = ((value * second_score + first * (2.0 * score - 1.0) + second) / b.spot) / b.spot;
    }
    if (!isfinite(lr[0]) || !isfinite(lr[1]) || !isfinite(first) || !isfinite(second)) bad = 1;
  }
  return value;
}

// ------------------------------------------------------------------ digital
// MonteCarloDigitalEngine's path function (engines/monte_carlo/digital.rs
// :110-140): strict in-the-money test in standardised coordinates, payoff
// amount * exp(exponent), likelihood-ratio delta and gamma in log space.
struct DigitalDev {
  int32_t call, asset, request, pad;
  double d2, drift, width, amount, log_spot, log_width, log_amount;
};

template <bool FP32, bool LR, class D>
__device__ double digital_path(const DigitalDev& g, D& dr, bool live, float sign, uint32_t& bad,
                               double* lr = nullptr) {
  if (LR) lr[0] = lr[1] = 0.0;
  if (!live) return 0.0;
  const double normal = (double)sign * (FP32 ? (double)dr.template normal<float>(0) : dr.template normal<double>(0));
  const double distance = g.d2 + normal;
  if (!(g.call ? distance > 0.0 : distance < 0.0)) return 0.0;
  const double exponent = g.drift + (g.asset ? g.width * normal : 0.0);
  const double payoff = discounted_product(g.amount, 1.0, exponent);
  if (!isfinite(payoff)) bad = 1;
  if (LR) {
    const double log_payoff = g.log_amount + exponent;
    if ((g.request & 1) && normal != 0.0)
      lr[0] = copysign(exp(log_payoff + log(fabs(normal)) - g.log_spot - g.log_width), normal);
    const double score = fma(-g.width, normal, normal * normal - 1.0);
    if ((g.request & 2) && score != 0.0)
      lr[1] = copysign(exp(log_payoff + log(fabs(score)) - 2.0 * g.log_spot - 2.0 * g.log_width), score);
    if (!isfinite(lr[0]) || !isfinite(lr[1])) bad = 1;
  }
  return payoff;
}

// ------------------------------------------------------------------ Asians
struct GeoAsianDev {
  double im, drift, width, log_width, log_strike, sign, log_discount;
  double log_spot, log_weight, log_base_width;  // for the likelihood-ratio Greeks
  int32_t request, pad;                         // KG_GREEK_* bits
};

// GeometricParameters::geometric (asian.rs:741-784). LR: also the per-path
// delta and gamma, lr[0] and lr[1].
template <bool FP32, bool LR = false, class D>
__device__ double geo_asian(const GeoAsianDev& g, D& dr, bool live, float sign, uint32_t& bad,
                            double* lr = nullptr) {
  if (LR) lr[0] = lr[1] = 0.0;
  if (!live) return 0.0;
  const double z = (double)sign * (double)(FP32 ? (double)dr.template normal<float>(0)
                                                 : dr.template normal<double>(0));
  const double distance = g.im + g.drift + g.width * z;
  const double center = (g.im + g.drift) / g.width;
  const double standardized = center + z;
  if (g.sign * standardized <= 0.0) return 0.0;
  double lp;
  if (fabs(distance) < 1e-5 && isfinite(standardized))
    lp = g.log_strike + g.log_width + log(fabs(standardized)) +
         (distance == 0.0 ? 0.0 : log(fabs(expm1(distance)) / fabs(distance)));
  else if (g.sign > 0)
    lp = g.log_strike + distance + log(-expm1(-distance));
  else
    lp = g.log_strike + log(-expm1(distance));
  const double v = exp(lp + g.log_discount);
  if (!isfinite(v)) bad = 1;
  if (LR) {
    const double log_asset = g.log_strike + distance + g.log_discount + g.log_weight;
    if (g.request & 1) lr[0] = g.sign * exp(log_asset - g.log_spot);
    const double numerator = z - exp(g.log_base_width);
    if ((g.request & 2) && numerator != 0.0)
      lr[1] = g.sign * copysign(exp(log_asset - 2.0 * g.log_spot + log(fabs(numerator)) - g.log_base_width),
                                numerator);
    if (!isfinite(lr[0]) || !isfinite(lr[1])) bad = 1;
  }
  return v;
}

struct ArithAsianDev {
  int32_t n, control;
  double spot, strike, weight, log_discount, drift, width, sign, interval_forward;
  double c_log_strike, c_las, c_fw;
  double log_width;  // ln(sigma) + ln(dt) / 2, for the gamma
  int32_t request;
  int32_t lead;  // a forward start: dimension 0 steps from valuation to the averaging start
  double lead_drift, lead_width;
  // Seasoning: ln(prior_weight * realized) (-inf when none), the realized
  // part of the baseline, prior_weight * (realized - strike) / spot, and the
  // adjusted strike strike - prior_weight * realized.
  double log_history, prior_baseline, adjusted_strike;
};

__device__ inline double geometric_values_d(const ArithAsianDev& p, double lf, double ls,
                                            double width) {
  const double d1 = (lf - ls) / width + 0.5 * width;
  const double d2 = d1 - width;
  const double la = lf + p.log_discount + d_log_normal_cdf(p.sign * d1);
  const double lc = ls + p.log_discount + d_log_normal_cdf(p.sign * d2);
  return p.sign > 0 ? d_positive_difference(la, lc) : d_positive_difference(lc, la);
}

// State of one arithmetic-average path after its loop: the relative sum of
// expm1 growths, the log-sum once growth passed e^600, and the geometric
// sum of the control variate. Nothing in it depends on the strike.
struct ArithPath {
  double rel, log_sum, gg, lead_growth;
};

// ArithmeticGrid::path (asian.rs:640-712): the strike-independent loop; a
// forward start takes its lead step from dimension 0. PAIRS steps two days
// per iteration (one Box-Muller pair); the spot-Greek kernel steps one at a
// time to stay within its register budget.
template <bool FP32, bool PAIRS = true, class D>
__device__ ArithPath arith_path(const ArithAsianDev& p, D& dr, bool live, float sign) {
  using R = typename std::conditional<FP32, float, double>::type;
  const int dims = p.n - 1;
  R growth = 0;
  double rel = 0.0, log_sum = -INFINITY, gg = 0.0, lead_growth = 0.0;
  const R drift = (R)p.drift, width = (R)p.width;
  const double inv_n = 1.0 / (double)p.n;
  if (p.lead && live) {
    const R z = (R)sign * dr.template normal<R>(0);
    lead_growth = FP32 ? (double)((float)p.lead_drift + (float)p.lead_width * (float)z)
                       : p.lead_drift + p.lead_width * (double)z;
  }
  // Lanes past the contract's last sample step along; their state is
  // discarded by the caller. Two steps per iteration, one Box-Muller pair.
  if (__ballot(live) == 0) return ArithPath{rel, log_sum, gg, lead_growth};
  const bool control = p.control;
  auto step = [&](int i, R z) {
    growth += drift + width * z;
    const double fraction = (i + 2 == p.n ? 0.5 : 1.0) * inv_n;
    if (control) gg += fraction * (double)growth;
    if ((double)growth > 600.0 || log_sum != -INFINITY) {
      if (log_sum == -INFINITY) log_sum = log(1.0 - 0.5 * inv_n + rel);
      log_sum = d_log_add(log_sum, log(fraction) + (double)growth);
    } else {
      rel += fraction * (FP32 ? (double)f32::expm1((float)growth) : expm1((double)growth));
    }
  };
  int i = 0;
  if constexpr (PAIRS) {
    for (; i + 1 < dims; i += 2) {
      R za, zb;
      dr.template normal_pair<R>(i + p.lead, za, zb);
      step(i, (R)sign * za);
      step(i + 1, (R)sign * zb);
    }
  }
  for (; i < dims; ++i) step(i, (R)sign * dr.template normal<R>(i + p.lead));
  return ArithPath{rel, log_sum, gg, lead_growth};
}

// ArithmeticParameters::arithmetic with the geometric control residual
// (asian.rs:1100-1256, 1016-1061): the discounted payoff of one path state
// for the strike terms of `p`.
__device__ __attribute__((always_inline)) inline double arith_payoff(const ArithAsianDev& p, const ArithPath& a,
                                                                   uint32_t& bad) {
  const double inv_n = 1.0 / (double)p.n;
  const double rel = a.rel, log_sum = a.log_sum, gg = a.gg, lg = a.lead_growth;
  const double half = 0.5 * inv_n, rest = 1.0 - half;
  const double log_spot = log(p.spot), log_weight = log(p.weight);
  const double log_rest = log_sum == -INFINITY ? log(rest) + log1p(rel / rest) : log_sum;
  const double log_q = log_spot + log_weight + lg + log_rest + p.interval_forward;
  const double log_c = log_spot + log_weight + lg + log(half);
  const double log_accrued = d_log_add(p.log_history, log_c);
  const bool has_extra = log_sum == -INFINITY;
  const double extra =
      has_extra ? expm1(lg) + exp(lg) * (rest * expm1(p.interval_forward) + rel * exp(p.interval_forward)) : 0.0;
  double residual = 0.0;
  if (p.control) {
    const double w = p.c_fw * p.width;
    const double lf = p.c_las + lg + gg + p.c_fw * p.drift + 0.5 * w * w;
    residual = geometric_values_d(p, lf, p.c_log_strike, w);
  }
  const double log_strike = log(p.strike);
  double price;
  if (log_accrued >= log_strike) {
    price = p.sign < 0 ? 0.0
                       : exp(log_q + p.log_discount) +
                             d_positive_difference(log_accrued + p.log_discount, log_strike + p.log_discount);
  } else {
    const double log_b = log_strike + log(-expm1(log_accrued - log_strike));
    double forward = log_q - log_b;
    const double baseline = p.prior_baseline + p.weight * ((p.spot - p.strike) / p.spot);
    if (has_extra) {
      const double excess = baseline + p.weight * extra;
      const double bb = exp(log_b - log_spot);
      if (bb > 0.0 && fabs(excess / bb) < 0.5) forward = log1p(excess / bb);
    }
    const double center = forward / p.width;
    const double d1 = center + 0.5 * p.width, d2 = center - 0.5 * p.width;
    const double la = log_q + p.log_discount + d_log_normal_cdf(p.sign * d1);
    const double lc = log_b + p.log_discount + d_log_normal_cdf(p.sign * d2);
    price = p.sign > 0 ? d_positive_difference(la, lc) : d_positive_difference(lc, la);
  }
  const double v = price - residual;
  if (!isfinite(v)) bad = 1;
  return v;
}

// The same payoff with its conditional delta and gamma (asian.rs:1233-1255,
// and geometric_values' for the control residual, :1063-1098) in lr[0] and
// lr[1]: the last interval is integrated in closed form, so both are
// smooth functions of the path.
// A call, not inlined: inlined, it spills the FP32 spot-Greek kernel.
__device__ __attribute__((noinline)) double arith_payoff_greeks(const ArithAsianDev& p, const ArithPath& a, uint32_t& bad,
                                             double* lr) {
  constexpr double HALF_LN_2PI = 0.9189385332046727;
  const double inv_n = 1.0 / (double)p.n;
  const double rel = a.rel, log_sum = a.log_sum, gg = a.gg, lg = a.lead_growth;
  const double half = 0.5 * inv_n, rest = 1.0 - half;
  const double log_spot = log(p.spot), log_weight = log(p.weight);
  const double log_rest = log_sum == -INFINITY ? log(rest) + log1p(rel / rest) : log_sum;
  const double log_q = log_spot + log_weight + lg + log_rest + p.interval_forward;
  const double log_c = log_spot + log_weight + lg + log(half);
  const double log_accrued = d_log_add(p.log_history, log_c);
  const bool has_extra = log_sum == -INFINITY;
  const double extra =
      has_extra ? expm1(lg) + exp(lg) * (rest * expm1(p.interval_forward) + rel * exp(p.interval_forward)) : 0.0;
  const bool want_delta = p.request & 1, want_gamma = p.request & 2;
  double res[3] = {0.0, 0.0, 0.0};
  if (p.control) {
    const double w = p.c_fw * p.width;
    const double lf = p.c_las + lg + gg + p.c_fw * p.drift + 0.5 * w * w;
    const double d1 = (lf - p.c_log_strike) / w + 0.5 * w;
    const double d2 = d1 - w;
    const double la = lf + p.log_discount + d_log_normal_cdf(p.sign * d1);
    const double lc = p.c_log_strike + p.log_discount + d_log_normal_cdf(p.sign * d2);
    res[0] = p.sign > 0 ? d_positive_difference(la, lc) : d_positive_difference(lc, la);
    if (want_delta) res[1] = p.sign * exp(la - log_spot);
    if (want_gamma) res[2] = exp(lf + p.log_discount - 2.0 * log_spot - log(w) - 0.5 * d1 * d1 - HALF_LN_2PI);
  }
  const double log_strike = log(p.strike);
  double price, delta = 0.0, gamma = 0.0;
  if (log_accrued >= log_strike) {
    if (p.sign < 0) {
      price = 0.0;
    } else {
      price = exp(log_q + p.log_discount) +
              d_positive_difference(log_accrued + p.log_discount, log_strike + p.log_discount);
      if (want_delta) delta = exp(log_q + p.log_discount - log_spot) + exp(log_c + p.log_discount - log_spot);
    }
  } else {
    const double log_b = log_strike + log(-expm1(log_accrued - log_strike));
    double forward = log_q - log_b;
    const double baseline = p.prior_baseline + p.weight * ((p.spot - p.strike) / p.spot);
    if (has_extra) {
      const double excess = baseline + p.weight * extra;
      const double bb = exp(log_b - log_spot);
      if (bb > 0.0 && fabs(excess / bb) < 0.5) forward = log1p(excess / bb);
    }
    const double center = forward / p.width;
    const double d1 = center + 0.5 * p.width, d2 = center - 0.5 * p.width;
    const double n1 = d_log_normal_cdf(p.sign * d1), n2 = d_log_normal_cdf(p.sign * d2);
    const double la = log_q + p.log_discount + n1;
    const double lc = log_b + p.log_discount + n2;
    price = p.sign > 0 ? d_positive_difference(la, lc) : d_positive_difference(lc, la);
    if (want_delta)
      delta = p.sign * (exp(log_c + p.log_discount - log_spot + n2) + exp(log_q + p.log_discount - log_spot + n1));
    if (want_gamma)
      gamma = exp(p.log_discount + 2.0 * log(p.adjusted_strike) - 2.0 * log_spot - log_b - p.log_width -
                  0.5 * d2 * d2 - HALF_LN_2PI);
  }
  const double v = price - res[0];
  lr[0] = delta - res[1];
  lr[1] = gamma - res[2];
  if (!isfinite(v) || !isfinite(lr[0]) || !isfinite(lr[1])) bad = 1;
  return v;
}

// The book keeps the strike terms out of line: inlined into the sample body
// they would set the kernel's register count (FP32: 171 VGPRs inlined, 123
// out of line; FP64: past 256).
__device__ __attribute__((noinline)) double arith_payoff_call(const ArithAsianDev& p, const ArithPath& a,
                                                              uint32_t& bad) {
  return arith_payoff(p, a, bad);
}
template <bool FP32, class D>
__device__ double arith_asian(const ArithAsianDev& p, D& dr, bool live, float sign, uint32_t& bad) {
  const ArithPath a = arith_path<FP32>(p, dr, live, sign);
  if (!live) return 0.0;
  return arith_payoff_call(p, a, bad);
}

// ------------------------------------------------------------------ accumulator
// Step flags of the accumulator path.
enum : uint8_t { ACC_OBSERVE = 1, ACC_ACCUMULATE = 2 };

struct AccumDev {
  int32_t n_steps;
  int32_t fixed;  // purchases grouped onto knock-out dates (before / at each step)
  double sign, strike_log, ko_log, strike, spot, daily_qty, accel, acc_qty, safe_lo, safe_hi;
  double terminal_discount;
  const double* steps;   // drift, diffusion, discount, before, at per step
  const float* steps32;  // drift, diffusion per step
  const uint8_t* flags;  // ACC_* per step
};

template <bool FP32, class D>
__device__ double accum_path(const AccumDev& a, D& dr, bool live, float sign, uint32_t& bad) {
  using R = typename std::conditional<FP32, float, double>::type;
  R lr = 0;
  double quantity = a.acc_qty;
  bool active = live;
  double value = 0.0;
  // Knock-out step and state; the settlement is formed after the loop, once
  // for all lanes.
  int ko_step = -1;
  R ko_lr = 0;
  double ko_quantity = 0.0;
  const R ko = (R)a.ko_log, sk = (R)a.strike_log, sg = (R)a.sign;
  auto unit = [&](double l) {
    const double spot = discounted_product(a.spot, 1.0, l);
    if (!isfinite(spot) || spot <= 0.0) bad = 1;
    const double distance = l - a.strike_log;
    if (fabs(distance) < 0.5) return a.sign * a.strike * expm1(distance);
    return a.sign > 0 ? spot - a.strike : a.strike - spot;
  };
  // Branch-free per lane, two steps per iteration, as in barrier_path: a
  // finished lane draws and discards, and keeps its state through selects.
  const int n_steps = a.n_steps;
  const bool fixed = a.fixed;
  uint32_t on = active ? 1u : 0u;
  // Overflow test of every visited spot through the lowest and highest log
  // returns (the spot is monotone in them), as in barrier_path.
  R low = lr, high = lr;
  // The daily purchase below and above the strike (the products the
  // reference forms each day). A non-finite quantity stays non-finite, so
  // its test is made once, after the loop.
  const double buy_accelerated = a.daily_qty * a.accel, buy = a.daily_qty * 1.0;
  auto step = [&](int i, R z) {
    const bool act = on;
    R next;
    if constexpr (FP32) {
      const float* row = a.steps32 + 2 * i;  // drift, diffusion
      next = lr + fmaf(uload(row + 1), (float)z, uload(row));
    } else {
      next = lr + fma(uload(a.steps + 5 * i + 1), (double)z, uload(a.steps + 5 * i));
    }
    lr = act ? next : lr;
    low = fmin(low, lr);
    high = fmax(high, lr);
    const uint8_t f = uload_u8(a.flags, i);
    if (fixed) quantity = act ? quantity + uload(a.steps + 5 * i + 3) : quantity;
    bool still = act;
    if (f & ACC_OBSERVE) {
      const bool out = act && sg * lr >= sg * ko;
      ko_step = out ? i : ko_step;
      ko_lr = out ? lr : ko_lr;
      ko_quantity = out ? quantity : ko_quantity;
      still = act && !out;
    }
    if (fixed) {
      quantity = still ? quantity + uload(a.steps + 5 * i + 4) : quantity;
    } else if (f & ACC_ACCUMULATE) {
      const double purchase = sg * lr < sg * sk ? buy_accelerated : buy;
      quantity = still ? quantity + purchase : quantity;
    }
    on = still ? 1u : 0u;
  };
  int i = 0;
  bool running = true;
  for (; i + 1 < n_steps; i += 2) {
    if (__ballot(on) == 0) {
      running = false;
      break;
    }
    R za, zb;
    dr.template normal_pair<R>(i, za, zb);
    step(i, (R)sign * za);
    step(i + 1, (R)sign * zb);
  }
  if (running && i < n_steps && __ballot(on) != 0) step(i, (R)sign * dr.template normal<R>(i));
  active = on;
  if (live && !isfinite(quantity)) bad = 1;
  if (live && !((double)low >= a.safe_lo && (double)high <= a.safe_hi)) {
    for (const R t : {low, high}) {
      const double sp = discounted_product(a.spot, 1.0, (double)t);
      if (!isfinite(sp) || sp <= 0.0) bad = 1;
    }
  }
  if (ko_step >= 0)
    value = discounted_product(ko_quantity, unit((double)ko_lr), a.steps[5 * ko_step + 2]);
  else if (active)
    value = discounted_product(quantity, unit((double)lr), a.terminal_discount);
  if (live && !isfinite(value)) bad = 1;
  return live ? value : 0.0;
}

}  // namespace kg
