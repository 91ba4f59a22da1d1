// Test infrastructure, not product code: FP64 CPU restatements of kiyosi's
// barrier (single and double, with likelihood-ratio delta/gamma), digital,
// Asian and accumulator Monte Carlo estimators, the checker of the HIP book
// and spot-Greek kernels.
#pragma once

#include "../include/kiyosi_amd.h"
#include "core.hpp"
#include "hit.hpp"
#include "plans.hpp"

namespace oracle {

// ---------------------------------------------------------------- numerics
// numerics.rs:192-215
inline double normal_cdf(double x) { return 0.5 * std::erfc(-x * 0.7071067811865476); }
inline double log_normal_tail(double value, double exponent) {
  double inv_sq = (1.0 / value) / value;
  double term = 1.0, sum = 1.0;
  for (int i = 1; i <= 20; ++i) {
    term *= -(double)(2 * i - 1) * inv_sq;
    sum += term;
  }
  return exponent - std::log(-value) + std::log(sum / std::sqrt(2.0 * M_PI));
}
inline double log_normal_cdf(double value) {
  if (value >= -10.0) return std::log(normal_cdf(value));
  return log_normal_tail(value, -0.5 * value * value);
}
inline double log_add(double a, double b) {  // numerics.rs:320-327
  double high = std::max(a, b);
  if (high == -INFINITY) return high;
  return high + std::log1p(std::exp(-std::fabs(a - b)));
}
inline double log_difference(double high, double low) {  // numerics.rs:329-334
  if (high == -INFINITY) return -INFINITY;
  return high + std::log(-std::expm1(low - high));
}
inline double positive_difference(double a, double b) {  // asian.rs:971-984
  if (a == INFINITY) return a;
  if (b == -INFINITY) return std::exp(a);
  double d = -std::expm1(b - a);
  return d > 0.0 ? std::exp(a + std::log(d)) : 0.0;
}
inline double weighted_cdf(double amount, double exponent, double argument) {  // numerics.rs:217-224
  double discount = std::exp(exponent);
  double discounted = amount * discount;
  if (argument < -10.0 || !std::isnormal(discount) || !std::isfinite(discounted))
    return std::exp(std::log(amount) + exponent + log_normal_cdf(argument));
  return discounted * normal_cdf(argument);
}
inline double standardize(double value, double vol, double root) {  // numerics.rs:160-167
  double width = vol * root;
  if (width < std::numeric_limits<double>::min()) return (value / vol) / root;
  return value / width;
}
// lognormal_price (engines/analytic/european.rs:317-400), the branch for
// widths of at least 1e-5.
inline double lognormal_price(double spot, double strike, double log_moneyness, double asset_log,
                              double risk_log, double vol, double root_time, double sign) {
  double width = vol * root_time;
  if (!std::isfinite(width)) unstable();
  if (width < 1e-5) throw Failure(UNSUPPORTED, "lognormal price below width 1e-5");
  double st = standardize(log_moneyness, vol, root_time);
  double d1 = st + 0.5 * width, d2 = st - 0.5 * width;
  double asset = weighted_cdf(spot, asset_log, sign * d1);
  double cash = weighted_cdf(strike, risk_log, sign * d2);
  double direct = sign * (asset - cash);
  double value;
  if (std::isfinite(direct)) {
    value = direct;
  } else {
    double a = log_moneyness + log_normal_cdf(sign * d1);
    double c = log_normal_cdf(sign * d2);
    value = std::exp(std::log(strike) + risk_log + (sign > 0 ? log_difference(a, c) : log_difference(c, a)));
  }
  if (!std::isfinite(value)) unstable();
  return value;
}

// Means and pair statistics of N path functionals as path_means returns them
// (monte_carlo.rs:1475-1496).
struct Means {
  double mean;
  bool has_se;
  double se;
};
inline Means path_mean(const Settings& st, const Moments& m, const Moments& p) {
  Means r;
  r.mean = scaled_value(m.scale, m.mean, 0.0);
  r.has_se = sampling_standard_error(st, m, p, 0.0, &r.se);
  return r;
}
inline Estimate from_sampling(const Settings& st, double price, bool has_se, double se) {
  Estimate e;
  e.price = price;
  e.has_seed = 1;
  e.seed = st.seed;
  if (has_se && se == 0.0) {
    e.status = ZERO_OBSERVED_VARIANCE;
  } else if (has_se) {
    e.status = AVAILABLE;
    e.has_standard_error = 1;
    e.standard_error = se;
  } else if (st.sampling == SOBOL) {
    e.status = UNSUPPORTED_SAMPLING;
  } else if (st.antithetic && st.path_count % 2) {
    e.status = INCOMPLETE_ANTITHETIC_PAIR;
  } else {
    e.status = INSUFFICIENT_OBSERVATIONS;
  }
  return e;
}
inline Estimate known_estimate(const Settings& st, double value) {
  if (!std::isfinite(value)) unstable();
  Estimate e;
  e.price = value;
  e.has_seed = 0;
  if (st.sampling == PSEUDO_RANDOM) {
    e.status = AVAILABLE;
    e.has_standard_error = 1;
    e.standard_error = 0.0;
  } else {
    e.status = UNSUPPORTED_SAMPLING;
  }
  return e;
}

// ---------------------------------------------------------------- barrier
// double_barrier::survival (engines/monte_carlo/double_barrier.rs:6-112):
// probability that a Brownian bridge from `initial` to `terminal` (log
// distances above the lower level) stays inside (0, band), with its first
// and second derivatives in `initial` at a fixed terminal value. Image
// series for small variance, spectral series otherwise. `branch` forces the
// image (0) or spectral (1) series for tests; -1 chooses as the reference.
inline void double_survival(double initial, double terminal, double band, double deviation,
                            bool derivatives, double out[3], int branch = -1) {
  out[0] = out[1] = out[2] = 0.0;
  if (terminal <= 0.0 || terminal >= band) return;
  double x = initial / band, y = terminal / band;
  const double ratio = deviation / band;
  const double variance = ratio * ratio;
  if (variance == 0.0) {
    out[0] = 1.0;
    return;
  }
  if (!std::isfinite(variance) || variance <= 0.0) unstable();
  const bool reflected = x + y > 1.0;
  if (reflected) {
    x = 1.0 - x;
    y = 1.0 - y;
  }
  double sum[3] = {0.0, 0.0, 0.0};
  bool converged = false;
  const bool image = branch < 0 ? variance < 0.25 : branch == 0;
  if (image) {
    const double inverse_variance = std::isnormal(variance) ? 1.0 / variance : INFINITY;
    sum[0] = -std::expm1(-2.0 * (x / ratio) * (y / ratio));
    auto add = [&](double k, bool direct) {
      const double exponent = direct ? -2.0 * (k / ratio) * ((y - x + k) / ratio)
                                     : -2.0 * ((x + k) / ratio) * ((y + k) / ratio);
      const double sign = direct ? 1.0 : -1.0;
      const double weight = std::exp(exponent);
      if (k != 0.0) sum[0] += sign * weight;
      if (derivatives) {
        const double numerator = direct ? 2.0 * k : -2.0 * (y + k);
        if (numerator != 0.0) {
          const double slope = numerator * inverse_variance;
          const double first = weight * slope;
          const double second = first * slope;
          if (std::isnormal(weight) && std::isnormal(first) && std::isnormal(second)) {
            sum[1] += sign * first;
            sum[2] += sign * second;
          } else {
            const double log_slope = std::log(std::fabs(numerator)) - 2.0 * std::log(ratio);
            sum[1] += sign * std::copysign(std::exp(exponent + log_slope), numerator);
            sum[2] += sign * std::exp(exponent + 2.0 * log_slope);
          }
        }
      }
      return weight;
    };
    add(0.0, false);
    for (int n = 1; n <= 64; ++n) {
      const double k = n;
      const double tail = add(k, true) + add(k, false) + add(-k, true) + add(-k, false);
      if (n >= 2 && tail < 1e-18) {
        converged = true;
        break;
      }
    }
  } else {
    const double distance = y - x;
    const double log_scale =
        std::log(2.0 * std::sqrt(2.0 * M_PI) * ratio) + 0.5 * distance * distance / variance;
    for (int n = 1; n <= 256; ++n) {
      const double frequency = (double)n * M_PI;
      const double exponent = log_scale - 0.5 * frequency * frequency * variance;
      const double weight = std::exp(exponent);
      const double sy = std::sin(frequency * y);
      const double product = std::sin(frequency * x) * sy;
      sum[0] += weight * product;
      if (derivatives) {
        const double first = frequency * std::cos(frequency * x) * sy;
        const double score = -distance / variance;
        sum[1] += weight * (first + score * product);
        sum[2] += weight * (2.0 * score * first +
                            (score * score + 1.0 / variance - frequency * frequency) * product);
      }
      if (n >= 4 && exponent + 4.0 * std::log(frequency) < -40.0) {
        converged = true;
        break;
      }
    }
  }
  if (!converged || !(sum[0] >= -1e-13 && sum[0] <= 1.0 + 1e-13)) unstable();
  out[0] = std::min(std::max(sum[0], 0.0), 1.0);
  if (derivatives) {
    out[1] = (reflected ? -sum[1] : sum[1]) / band;
    out[2] = (sum[2] / band) / band;
  }
}

// One barrier contract as MonteCarloBarrierEngine::native sees it
// (barrier.rs:160-178): a single barrier, or a double barrier as its lower
// (down-type) barrier plus the upper level.
struct BarrierSpec {
  int option_type, barrier_type, observation, rebate_timing;
  double strike, barrier, rebate;
  bool has_upper;
  double upper;
  int64_t effective_time, expiry_time;
  int n_observations, touch_state;
  const int64_t* observation_times;
};
inline BarrierSpec barrier_spec(const kg_barrier_terms& o) {
  return {o.option_type, o.barrier_type, o.observation, o.rebate_timing, o.strike, o.barrier, o.rebate, false,
          0.0, o.effective_time, o.expiry_time, o.n_observations, o.touch_state, o.observation_times};
}
inline BarrierSpec barrier_spec(const kg_double_barrier_terms& o) {
  return {o.option_type, o.knock_type == KG_DOUBLE_KNOCK_IN ? KG_DOWN_AND_IN : KG_DOWN_AND_OUT, o.observation,
          o.rebate_timing, o.strike, o.lower_barrier, o.rebate, true, o.upper_barrier, o.effective_time,
          o.expiry_time, o.n_observations, o.touch_state, o.observation_times};
}

// Delta and gamma of the Black-Scholes vanilla (AnalyticEuropeanEngine's
// closed form), for barriers resolved to their vanilla payoff.
inline void vanilla_spot_greeks(double S, double K, double r, double q, double sigma, double T, double sign,
                                double* delta, double* gamma) {
  const double w = sigma * std::sqrt(T);
  const double d1 = (log_ratio(S, K) + (r - q) * T) / w + 0.5 * w;
  const double pdf = std::exp(-0.5 * d1 * d1) / std::sqrt(2.0 * M_PI);
  *delta = sign * std::exp(-q * T) * normal_cdf(sign * d1);
  *gamma = std::exp(-q * T) * pdf / (S * w);
}

// Result of the barrier and digital engines' native pricing: the estimate
// and the likelihood-ratio delta/gamma (`available` bits say which).
struct SpotGreeks {
  int available = 0;
  double delta = 0.0, gamma = 0.0;
};

// MonteCarloBarrierEngine::native (barrier.rs:160-541) for flat BSM rates.
// `request` holds KG_GREEK_* bits; `radius` is the symmetric spot shift
// whose interval, when it crosses a monitored level, makes the spot Greeks
// unavailable (at_spot_discontinuity, instruments/monitoring.rs:430-445),
// and over which the at-hit rebate value of a continuous knock-out is
// differenced (complete_greeks, risk/numerical.rs:359-430).
inline Estimate barrier_native(const Settings& st, const BarrierSpec& o, const kg_bsm_market& m, const Context& cx,
                               int request, double radius, SpotGreeks* greeks) {
  const int64_t val = m.valuation_time, expiry = o.expiry_time;
  if (val < o.effective_time || val > expiry)
    throw Failure(INVALID_INPUT, "valuation outside contract life");
  if (m.volatility == 0.0) throw Failure(UNSUPPORTED, "zero volatility is deterministic");
  const bool up = o.barrier_type >= KG_UP_AND_OUT;
  const bool knock_in = o.barrier_type == KG_DOWN_AND_IN || o.barrier_type == KG_UP_AND_IN;
  const bool continuous = o.observation == KG_OBSERVE_CONTINUOUS;
  const bool at_hit = o.rebate_timing == KG_REBATE_AT_HIT;
  const double sign = o.option_type;
  const double r = m.rate, q = m.dividend, sigma = m.volatility, S = m.spot;
  const double T = years_until(val, expiry);
  // Monitoring state (instruments/monitoring.rs:272-321): history before
  // valuation, then a touch at the valuation.
  const bool had = continuous ? val > o.effective_time : (o.n_observations > 0 && o.observation_times[0] < val);
  if (!had && o.touch_state == KG_TOUCH_TOUCHED) throw Failure(INVALID_INPUT, "touch predates monitoring");
  if (had && o.touch_state == KG_TOUCH_UNKNOWN) throw Failure(MISSING_HISTORY, "prior touch state is required");
  const bool prior = o.touch_state == KG_TOUCH_TOUCHED;
  bool observed_now = continuous;
  for (int i = 0; i < o.n_observations && !continuous; ++i) observed_now |= o.observation_times[i] == val;
  const bool touched =
      prior || (observed_now && ((up ? S >= o.barrier : S <= o.barrier) || (o.has_upper && S >= o.upper)));
  const bool finished = val == expiry || (!continuous && o.n_observations > 0 &&
                                          o.observation_times[o.n_observations - 1] <= val);
  const bool last_is_expiry =
      continuous || (o.n_observations > 0 && o.observation_times[o.n_observations - 1] == expiry);
  bool requires_touch;
  if (o.has_upper)  // DoubleBarrierOption::payoff_requires_touch (instruments/double_barrier.rs:216-224)
    requires_touch = last_is_expiry && (sign > 0 ? o.strike >= o.upper : o.strike <= o.barrier);
  else
    requires_touch = last_is_expiry && (up ? (sign > 0 && o.strike >= o.barrier) : (sign < 0 && o.strike <= o.barrier));
  // The request keeps delta and gamma off expiry and off a spot stencil
  // that crosses a monitored level (barrier.rs:194-205).
  auto crosses = [&](double level) { return S - radius <= level && level <= S + radius; };
  const bool discontinuity = !prior && observed_now && (crosses(o.barrier) || (o.has_upper && crosses(o.upper)));
  int req = (val != expiry && !discontinuity) ? request & (KG_GREEK_DELTA | KG_GREEK_GAMMA) : 0;
  if (greeks) *greeks = SpotGreeks{};
  auto cash = [&] { return o.rebate == 0.0 ? 0.0 : o.rebate * std::exp(-r * T); };
  auto vanilla = [&] {
    if (T == 0.0) return std::max(sign * (S - o.strike), 0.0);
    return lognormal_price(S, o.strike, log_ratio(S, o.strike) + (r - q) * T, -q * T, -r * T, sigma,
                           std::sqrt(T), sign);
  };
  // resolved_price_with_state (engines/analytic/barrier.rs:32-80) and the
  // Greeks of a resolved contract (barrier.rs:206-229).
  auto resolved = [&](double value, bool vanilla_payoff) {
    if (greeks && req) {
      greeks->available = req;
      if (vanilla_payoff) {  // volatility_affects_price: the vanilla's analytic delta/gamma
        double d, g;
        vanilla_spot_greeks(S, o.strike, r, q, sigma, T, sign, &d, &g);
        greeks->delta = (req & KG_GREEK_DELTA) ? d : 0.0;
        greeks->gamma = (req & KG_GREEK_GAMMA) ? g : 0.0;
      }
    }
    return known_estimate(st, value);
  };
  if (touched && !knock_in) return resolved(at_hit ? (prior ? 0.0 : o.rebate) : cash(), false);
  if (!touched && finished && knock_in) return resolved(cash(), false);
  const bool fixed = o.rebate == 0.0 && requires_touch;
  if (fixed && !knock_in) return resolved(0.0, false);
  if (touched || finished || fixed) return resolved(vanilla(), true);
  // Plan (barrier.rs:228-290).
  const double carry = r - q, hvr = 0.5 * sigma * sigma, drift = carry - hvr;
  // A continuous knock-out paid at hit: its paths carry no rebate and the
  // hit value is added, with central spot differences for the Greeks.
  double hit = 0.0, hit_delta = 0.0, hit_gamma = 0.0;
  if (continuous && !knock_in && at_hit && o.rebate != 0.0) {
    auto hit_at = [&](double spot) {
      if (o.has_upper) return double_hit_value(o.rebate, o.barrier, o.upper, spot, r, q, sigma, T);
      return single_hit_value(o.rebate, std::fabs(std::log(o.barrier / spot)), up, drift, sigma, T, r);
    };
    hit = hit_at(S);
    if (req) {
      if (radius > 0.0 && S > radius) {
        const double hi = hit_at(S + radius), lo = hit_at(S - radius);
        hit_delta = (hi - lo) / radius / 2.0;
        hit_gamma = ((hi - hit) - (hit - lo)) / radius / radius;
      } else {
        req = 0;
      }
    }
  }
  const bool delta = req & KG_GREEK_DELTA, gamma = req & KG_GREEK_GAMMA;
  auto mean_distance = [&](int64_t date, double level) {
    double e = years_until(val, date);
    return log_ratio(S, level) + carry * e - hvr * e;
  };
  struct Step {
    double mean, width, discount, lower, upper;
    bool observe;
    int64_t date;
  };
  std::vector<Step> steps;
  int64_t prev = val;
  double width_sum = 0.0;
  std::vector<int64_t> dates;
  if (!continuous)
    for (int i = 0; i < o.n_observations; ++i)
      if (o.observation_times[i] > val) dates.push_back(o.observation_times[i]);
  std::vector<std::pair<int64_t, bool>> all;
  for (int64_t d : dates) all.push_back({d, true});
  if (dates.empty() || dates.back() != expiry) all.push_back({expiry, false});
  for (auto [date, observe] : all) {
    double dt = years_until(prev, date);
    double width = sigma * std::sqrt(dt);
    double elapsed = years_until(val, date);
    double mean = drift * elapsed;
    if (width <= 0.0 || !std::isfinite(width) || !std::isfinite(mean)) unstable();
    steps.push_back({mean, width, -r * elapsed, observe ? mean_distance(date, o.barrier) : 0.0,
                     observe && o.has_upper ? mean_distance(date, o.upper) : 0.0, observe, date});
    width_sum += width;
    prev = date;
  }
  const double moneyness = mean_distance(expiry, o.strike);
  const double barrier = log_ratio(o.barrier, S);
  const double terminal_barrier = mean_distance(expiry, o.barrier);
  const double band = o.has_upper ? log_ratio(o.upper, o.barrier) : 0.0;
  const double log_discount = -r * T;
  const bool bounded = std::fabs(std::log(S)) + 10.0 * width_sum + std::fabs(drift * T) < 600.0;
  const bool condition_on_hit = knock_in && !continuous && steps.size() > 8 && bounded;
  auto weighted = [&](double amount, double lw) {
    return scaled_value(std::fabs(amount), amount < 0 ? -1.0 : 1.0, log_discount + lw);
  };
  Moments total[3], pairs[3];
  path_samples_lazy<3>(
      st, steps.size(), true, cx,
      [&](LazyPath& normals, double* out) {
        double terminal = 0.0, noise = 0.0, knocked_out = 0.0;
        bool touched_p = false, has_cond = false;
        double conditioned = 0.0;
        const double first_normal = normals.normal();
        for (size_t index = 0; index < steps.size(); ++index) {
          const Step& s = steps[index];
          double z = index == 0 ? first_normal : normals.normal();
          noise = std::fma(s.width, z, noise);
          terminal = s.mean + noise;
          if (!bounded && scaled_value(S, 1.0, terminal) <= 0.0) unstable();
          const bool hit = o.has_upper ? (noise <= -s.lower || noise >= -s.upper)
                                       : (up ? noise >= -s.lower : noise <= -s.lower);
          if (s.observe && hit) {
            touched_p = true;
            if (condition_on_hit && steps.size() - index > 8) {
              double asset = scaled_value(S, 1.0, terminal);
              double remaining = years_until(s.date, expiry);
              conditioned = lognormal_price(asset, o.strike, moneyness + noise + hvr * remaining,
                                            s.discount - q * remaining, log_discount, sigma,
                                            std::sqrt(remaining), sign);
              has_cond = true;
              break;
            }
          }
          if (touched_p && !knock_in) {
            knocked_out = scaled_value(o.rebate, 1.0, at_hit ? s.discount : log_discount);
            break;
          }
        }
        double asset = scaled_value(S, 1.0, terminal);
        double distance = moneyness + noise;
        double payoff = std::fabs(distance) < 0.5 ? o.strike * std::expm1(distance) : asset - o.strike;
        double intrinsic = std::max(sign * payoff, 0.0);
        double value, first = 0.0, second = 0.0;
        if (has_cond) {
          value = conditioned;
        } else if (continuous && o.has_upper) {
          double sv[3];
          double_survival(-barrier, terminal_barrier + noise, band, steps[0].width, delta || gamma, sv);
          double rebate = (!knock_in && at_hit) ? 0.0 : o.rebate;
          value = weighted(intrinsic, knock_in ? std::log1p(-sv[0]) : std::log(sv[0])) +
                  weighted(rebate, knock_in ? std::log(sv[0]) : std::log1p(-sv[0]));
          const double difference = (knock_in ? -1.0 : 1.0) * (intrinsic - rebate);
          auto derivative = [&](double w) {
            return scaled_value(std::fabs(difference), (difference < 0 ? -1.0 : 1.0) * (w < 0 ? -1.0 : 1.0),
                                log_discount + std::log(std::fabs(w)));
          };
          if (delta || gamma) first = derivative(sv[1]);
          if (gamma) second = derivative(sv[2]);
        } else if (continuous) {
          double w0 = steps[0].width;
          double dtb = terminal_barrier + noise;
          bool live = up ? dtb < 0.0 : dtb > 0.0;
          double slope = live ? 2.0 * (std::fabs(dtb) / w0) / w0 : 0.0;
          double exponent = -std::fabs(barrier) * slope;
          double survival = live ? std::log(-std::expm1(exponent)) : -INFINITY;
          double rebate = (!knock_in && at_hit) ? 0.0 : o.rebate;
          value = weighted(intrinsic, knock_in ? exponent : survival) +
                  weighted(rebate, knock_in ? survival : exponent);
          if ((delta || gamma) && slope > 0.0 && exponent > -INFINITY) {
            const double difference = (knock_in ? -1.0 : 1.0) * (intrinsic - rebate);
            first = weighted(up ? -difference : difference, exponent + std::log(slope));
            if (gamma) second = weighted(-difference, exponent + 2.0 * std::log(slope));
          }
        } else if (touched_p && !knock_in) {
          value = knocked_out;
        } else {
          value = weighted(touched_p == knock_in ? intrinsic : o.rebate, 0.0);
        }
        // Likelihood-ratio delta and gamma (barrier.rs:459-472).
        const double width = steps[0].width, z = first_normal, score = z / width;
        const double d = delta ? (value * score + first) / S : 0.0;
        double g = 0.0;
        if (gamma) {
          const double second_score = ((z * z - 1.0) / width - z) / width;
          g = ((value * second_score + first * (2.0 * score - 1.0) + second) / S) / S;
        }
        if (!std::isfinite(value) || !std::isfinite(d) || !std::isfinite(g)) unstable();
        out[0] = value;
        out[1] = d;
        out[2] = g;
      },
      total, pairs);
  Means mm = path_mean(st, total[0], pairs[0]);
  if (greeks) {
    greeks->available = req;
    greeks->delta = delta ? scaled_value(total[1].scale, total[1].mean, 0.0) + hit_delta : 0.0;
    greeks->gamma = gamma ? scaled_value(total[2].scale, total[2].mean, 0.0) + hit_gamma : 0.0;
  }
  return from_sampling(st, mm.mean + hit, mm.has_se, mm.se);
}

inline Estimate barrier_price(const Settings& st, const kg_barrier_terms& o, const kg_bsm_market& m,
                              const Context& cx) {
  return barrier_native(st, barrier_spec(o), m, cx, 0, 0.0, nullptr);
}

// ---------------------------------------------------------------- digital
// MonteCarloDigitalEngine::native (engines/monte_carlo/digital.rs:55-150):
// one terminal normal, strict in-the-money test in standardised
// coordinates, likelihood-ratio delta and gamma.
inline Estimate digital_native(const Settings& st, const kg_digital_terms& o, const kg_bsm_market& m,
                               const Context& cx, int request, SpotGreeks* greeks) {
  if (m.valuation_time > o.expiry_time) throw Failure(INVALID_INPUT, "valuation outside contract life");
  if (m.volatility == 0.0) throw Failure(UNSUPPORTED, "zero volatility is deterministic");
  const bool call = o.option_type > 0, asset = o.payout_kind == KG_ASSET_OR_NOTHING;
  const double S = m.spot, sigma = m.volatility, r = m.rate, q = m.dividend;
  if (greeks) *greeks = SpotGreeks{};
  const double time = years_until(m.valuation_time, o.expiry_time);
  const int req = request & (KG_GREEK_DELTA | KG_GREEK_GAMMA);
  if (time == 0.0) {  // Payoff::digital value at expiry (engines.rs:212-227)
    // No native Greeks at expiry: the expiry stencil supplies them.
    const double value = (call ? S - o.strike : o.strike - S) > 0.0 ? (asset ? S : o.cash_payout) : 0.0;
    return known_estimate(st, value);
  }
  const double root = std::sqrt(time), width = sigma * root, carry = (r - q) * time;
  const double d2 = standardize(log_ratio(S, o.strike) + carry, sigma, root) - 0.5 * width;
  const double amount = asset ? S : o.cash_payout;
  const double drift = asset ? (-q - 0.5 * sigma * sigma) * time : -r * time;
  if (!std::isfinite(width) || !std::isfinite(carry) || !std::isfinite(drift)) unstable();
  const double log_spot = std::log(S), log_width = std::log(sigma) + std::log(root);
  const bool want_delta = req & KG_GREEK_DELTA, want_gamma = req & KG_GREEK_GAMMA;
  Moments total[3], pairs[3];
  path_samples<3>(
      st, 1, true, cx,
      [&](const double* normals, double* out) {
        const double normal = normals[0];
        const double distance = d2 + normal;
        out[0] = out[1] = out[2] = 0.0;
        if (!(call ? distance > 0.0 : distance < 0.0)) return;
        const double exponent = drift + (asset ? width * normal : 0.0);
        const double payoff = scaled_value(amount, 1.0, exponent);
        const double log_payoff = std::log(amount) + exponent;
        const double d = want_delta && normal != 0.0
                             ? std::copysign(std::exp(log_payoff + std::log(std::fabs(normal)) - log_spot - log_width),
                                             normal)
                             : 0.0;
        const double score = std::fma(-width, normal, normal * normal - 1.0);
        const double g = want_gamma && score != 0.0
                             ? std::copysign(std::exp(log_payoff + std::log(std::fabs(score)) - 2.0 * log_spot -
                                                      2.0 * log_width),
                                             score)
                             : 0.0;
        if (!std::isfinite(d) || !std::isfinite(g)) unstable();
        out[0] = payoff;
        out[1] = d;
        out[2] = g;
      },
      total, pairs);
  Means mm = path_mean(st, total[0], pairs[0]);
  if (greeks) {
    greeks->available = req;
    greeks->delta = want_delta ? scaled_value(total[1].scale, total[1].mean, 0.0) : 0.0;
    greeks->gamma = want_gamma ? scaled_value(total[2].scale, total[2].mean, 0.0) : 0.0;
  }
  return from_sampling(st, mm.mean, mm.has_se, mm.se);
}

// ---------------------------------------------------------------- Asian
struct AsianArith {
  double spot, strike, weight, log_discount, drift, width, log_width, sign;
  int n;
  double interval_forward;
};


// geometric_values (asian.rs:1063-1098): price, delta and gamma of the
// geometric control for log forward lf, log strike ls and width.
inline void geometric_values(const AsianArith& p, double lf, double ls, double width, double* out) {
  double d1 = (lf - ls) / width + 0.5 * width;
  double d2 = d1 - width;
  double la = lf + p.log_discount + log_normal_cdf(p.sign * d1);
  double lc = ls + p.log_discount + log_normal_cdf(p.sign * d2);
  out[0] = p.sign > 0 ? positive_difference(la, lc) : positive_difference(lc, la);
  const double log_spot = std::log(p.spot);
  out[1] = p.sign * std::exp(la - log_spot);
  out[2] = std::exp(lf + p.log_discount - 2.0 * log_spot - std::log(width) - 0.5 * d1 * d1 -
                    0.5 * std::log(2.0 * M_PI));
}

// MonteCarloAsianEngine::native (asian.rs:347-372, prepare :374-560) for
// flat BSM rates: seasoned averages, forward starts and single fixings. `request`
// holds KG_GREEK_* bits; delta and gamma are the path means of the engine's
// conditional estimators (GeometricParameters::geometric :741-784,
// ArithmeticParameters::arithmetic :1100-1256), plus the geometric
// control's exact values when the control is used.
inline Estimate asian_native(const Settings& st, const kg_asian_terms& o, const kg_bsm_market& m,
                             const Context& cx, int request, SpotGreeks* greeks) {
  const int64_t val = m.valuation_time;
  if (val < o.effective_time || val > o.expiry_time)
    throw Failure(INVALID_INPUT, "valuation must be within contract life");
  if (o.step_count < 2 || o.step_count > 10000) throw Failure(INVALID_INPUT, "step_count out of range");
  // AverageTerms::validate (instruments/asian.rs:105-124)
  if (val > o.averaging_start_time && !o.has_realized)
    throw Failure(MISSING_HISTORY, "prior realized average is required");
  if (val <= o.averaging_start_time && o.has_realized)
    throw Failure(INVALID_INPUT, "realized average must match the elapsed averaging period");
  const double realized = o.has_realized ? o.realized_average : 0.0;
  const double tau = years_until(val, o.expiry_time);
  const double sign = o.option_type;
  const int req = request & (KG_GREEK_DELTA | KG_GREEK_GAMMA);
  const bool want_delta = req & KG_GREEK_DELTA, want_gamma = req & KG_GREEK_GAMMA;
  if (greeks) *greeks = SpotGreeks{};
  // At expiry the price is the settlement (:74-79) and the native estimator
  // has no Greeks.
  if (tau == 0.0) return known_estimate(st, std::max(sign * ((o.has_realized ? realized : m.spot) - o.strike), 0.0));
  if (m.volatility == 0.0) throw Failure(UNSUPPORTED, "zero volatility is deterministic");
  const double period = years_until(o.averaging_start_time, o.expiry_time);
  const bool single = period == 0.0;
  const bool geometric_kind = o.averaging == KG_AVERAGE_GEOMETRIC;
  const double sigma = m.volatility, r = m.rate, q = m.dividend;
  const double carry = r - q;
  const double drift_rate = carry - 0.5 * sigma * sigma;
  const double log_discount = -r * tau;
  const double elapsed = years_until(o.averaging_start_time, val);
  if (!geometric_kind && !single) {
    // asian.rs:399-414: a non-positive adjusted strike (instruments/asian.rs
    // :221-233) prices the certain exercise in closed form, Turnbull-Wakeman
    // under flat rates (engines/analytic/asian.rs:279-292).
    const double adjusted = elapsed > 0.0 ? o.strike + elapsed / tau * (o.strike - realized) : o.strike;
    if (adjusted <= 0.0) {
      const double x = carry * tau;
      const double m1 = x == 0.0 ? 1.0 : std::expm1(x) / x;
      const double log_m1 = std::log(m1) + std::log(tau / period) - r * tau;
      double value = 0.0;
      if (sign > 0.0)
        value = scaled_cash(std::max(realized * (elapsed / period) - o.strike, 0.0), -r * tau) +
                scaled_cash(m.spot, log_m1);
      if (greeks) {
        greeks->available = req;
        greeks->delta = want_delta && sign > 0.0 ? std::exp(log_m1) : 0.0;
        greeks->gamma = 0.0;
      }
      return known_estimate(st, value);
    }
  }
  const double lead = single ? 0.0 : std::max(years_until(val, o.averaging_start_time), 0.0);
  const double future = single ? tau : years_until(std::max(val, o.averaging_start_time), o.expiry_time);
  const double weight = single ? 1.0 : future / period;
  const double prior_weight = val > o.averaging_start_time ? elapsed / period : 0.0;
  Moments total[3], pairs[3];
  double expected[3] = {0.0, 0.0, 0.0};
  if (geometric_kind || single) {
    // GeometricParameters (asian.rs:452-484); a single fixing is the
    // terminal spot.
    const double vt = single ? tau : lead + future / 3.0;
    const double im = (single ? log_ratio(m.spot, o.strike) + (-q * tau + r * tau)
                              : weight * log_ratio(m.spot, o.strike)) +
                      (val > o.averaging_start_time ? prior_weight * log_ratio(realized, o.strike) : 0.0);
    const double drift = single ? -0.5 * sigma * sigma * tau : weight * drift_rate * (lead + future / 2.0);
    const double width = sigma * (weight * std::sqrt(vt));
    const double log_width = std::log(sigma) + std::log(weight) + 0.5 * std::log(vt);
    const double log_base_width = std::log(sigma) + 0.5 * std::log(vt);
    const double log_strike = std::log(o.strike), log_spot = std::log(m.spot), log_weight = std::log(weight);
    path_samples<3>(
        st, 1, true, cx,
        [&](const double* n, double* out) {
          // GeometricParameters::geometric (asian.rs:741-784)
          out[0] = out[1] = out[2] = 0.0;
          double distance = im + drift + width * n[0];
          double center = (im + drift) / width;
          double standardized = center + n[0];
          if (sign * standardized <= 0.0) return;
          double lp;
          if (std::fabs(distance) < 1e-5 && std::isfinite(standardized))
            lp = log_strike + log_width + std::log(std::fabs(standardized)) +
                 (distance == 0.0 ? 0.0 : std::log(std::fabs(std::expm1(distance)) / std::fabs(distance)));
          else if (sign > 0)
            lp = log_strike + distance + std::log(-std::expm1(-distance));
          else
            lp = log_strike + std::log(-std::expm1(distance));
          out[0] = std::exp(lp + log_discount);
          const double log_asset = log_strike + distance + log_discount + log_weight;
          if (want_delta) out[1] = sign * std::exp(log_asset - log_spot);
          const double numerator = n[0] - std::exp(log_base_width);
          if (want_gamma && numerator != 0.0)
            out[2] = sign * std::copysign(std::exp(log_asset - 2.0 * log_spot + std::log(std::fabs(numerator)) -
                                                   log_base_width),
                                          numerator);
          if (!std::isfinite(out[0]) || !std::isfinite(out[1]) || !std::isfinite(out[2])) unstable();
        },
        total, pairs);
  } else {
    AsianArith p;
    p.spot = m.spot;
    p.strike = o.strike;
    p.weight = weight;
    p.log_discount = log_discount;
    p.n = o.step_count;
    const double dt = future / o.step_count;
    p.interval_forward = carry * dt;
    p.drift = drift_rate * dt;
    p.width = sigma * std::sqrt(dt);
    p.log_width = std::log(sigma) + 0.5 * std::log(dt);
    p.sign = sign;
    const bool has_lead = lead > 0.0;
    const double lead_drift = drift_rate * lead, lead_width = sigma * std::sqrt(lead);
    const double log_history =
        realized > 0.0 && prior_weight > 0.0 ? std::log(realized) + std::log(prior_weight) : -INFINITY;
    const double adjusted_strike = std::fma(-prior_weight, realized, p.strike);
    const size_t dims = (size_t)o.step_count - 1 + (has_lead ? 1 : 0);
    // geometric_control (asian.rs:1016-1061)
    bool control = false;
    double c_log_strike = 0, c_las = 0, c_fw = 0;
    {
      const double n = o.step_count;
      const double fw = 1.0 - 0.5 / n;
      const double las = std::log(p.spot) + std::log(p.weight);
      const double ls = std::log(adjusted_strike);
      if (!(p.width < 1e-5 || p.width > 0.8 || std::fabs(p.log_discount) > 50.0 ||
            std::fabs(std::log(p.spot)) > 200.0 || std::fabs(las) > 200.0 || !std::isfinite(ls))) {
        double mean = lead_drift + fw * p.drift;
        for (int s = 1; s < o.step_count; ++s) mean += (fw - (double)s / n) * p.drift;
        double variance = lead_width * lead_width + (p.width * p.width) * (n / 3.0 - 1.0 / (12.0 * n));
        double w = std::sqrt(variance);
        double lf = las + mean + 0.5 * variance;
        if (w >= 0.02 && w <= 0.8 && std::fabs(lf - ls) <= 2.0 * w) {
          control = true;
          c_log_strike = ls;
          c_las = las;
          c_fw = fw;
          geometric_values(p, lf, ls, w, expected);
        }
      }
    }
    auto run = [&](bool use_control) {
      path_samples<3>(
          st, dims, true, cx,
          [&](const double* normals, double* out) {
            // ArithmeticGrid::path (asian.rs:640-712), flat rates; a forward
            // start's lead step takes normal 0.
            const size_t first = has_lead ? 1 : 0;
            const double lg = lead_drift + lead_width * (has_lead ? normals[0] : 0.0);
            double growth = 0.0, rel = 0.0, log_sum = -INFINITY, gg = 0.0;
            for (size_t i = 0; i + first < dims; ++i) {
              growth += p.drift + p.width * normals[i + first];
              double fraction = ((int)i + 2 == p.n ? 0.5 : 1.0) / (double)p.n;
              if (use_control) gg += fraction * growth;
              if (growth > 600.0 || log_sum != -INFINITY) {
                if (log_sum == -INFINITY) log_sum = std::log(1.0 - 0.5 / (double)p.n + rel);
                log_sum = log_add(log_sum, std::log(fraction) + growth);
              } else {
                rel += fraction * std::expm1(growth);
              }
            }
            const double half = 0.5 / (double)p.n, rest = 1.0 - half;
            const double log_spot = std::log(p.spot);
            const double log_rest = log_sum == -INFINITY ? std::log(rest) + std::log1p(rel / rest) : log_sum;
            const double log_q = log_spot + std::log(p.weight) + lg + log_rest + p.interval_forward;
            const double log_c = log_spot + std::log(p.weight) + lg + std::log(half);
            const double log_accrued = log_add(log_history, log_c);
            const bool has_extra = log_sum == -INFINITY;
            const double extra = has_extra ? std::expm1(lg) + std::exp(lg) * (rest * std::expm1(p.interval_forward) +
                                                                              rel * std::exp(p.interval_forward))
                                           : 0.0;
            // ArithmeticParameters::arithmetic (asian.rs:1100-1256)
            double residual[3] = {0.0, 0.0, 0.0};
            if (use_control) {
              double width = c_fw * p.width;
              double lf = c_las + lg + gg + c_fw * p.drift + 0.5 * width * width;
              geometric_values(p, lf, c_log_strike, width, residual);
              if (!want_delta) residual[1] = 0.0;
              if (!want_gamma) residual[2] = 0.0;
            }
            const double log_strike = std::log(p.strike);
            double price, delta = 0.0, gamma = 0.0;
            if (log_accrued >= log_strike) {
              if (p.sign < 0) {
                price = 0.0;
              } else {
                double c = positive_difference(log_accrued + p.log_discount, log_strike + p.log_discount);
                price = std::exp(log_q + p.log_discount) + c;
                if (want_delta)
                  delta = std::exp(log_q + p.log_discount - log_spot) + std::exp(log_c + p.log_discount - log_spot);
              }
            } else {
              const double log_b = log_strike + std::log(-std::expm1(log_accrued - log_strike));
              double forward = log_q - log_b;
              const double baseline = prior_weight * ((realized - p.strike) / p.spot) +
                                      p.weight * ((p.spot - p.strike) / p.spot);
              if (has_extra) {
                double excess = baseline + p.weight * extra;
                double b = std::exp(log_b - log_spot);
                if (b > 0.0 && std::fabs(excess / b) < 0.5) forward = std::log1p(excess / b);
              }
              if (!(p.width >= 1e-5)) throw Failure(UNSUPPORTED, "arithmetic widths below 1e-5");
              double center = forward / p.width;
              double d1 = center + 0.5 * p.width, d2 = center - 0.5 * p.width;
              double la = log_q + p.log_discount + log_normal_cdf(p.sign * d1);
              double lc = log_b + p.log_discount + log_normal_cdf(p.sign * d2);
              price = p.sign > 0 ? positive_difference(la, lc) : positive_difference(lc, la);
              if (want_delta)
                delta = p.sign * (std::exp(log_c + p.log_discount - log_spot + log_normal_cdf(p.sign * d2)) +
                                  std::exp(log_q + p.log_discount - log_spot + log_normal_cdf(p.sign * d1)));
              if (want_gamma) {
                const double k = adjusted_strike;
                gamma = std::exp(p.log_discount + 2.0 * std::log(k) - 2.0 * log_spot - log_b - p.log_width -
                                 0.5 * d2 * d2 - 0.5 * std::log(2.0 * M_PI));
              }
            }
            out[0] = price - residual[0];
            out[1] = delta - residual[1];
            out[2] = gamma - residual[2];
            if (!std::isfinite(out[0]) || !std::isfinite(out[1]) || !std::isfinite(out[2])) unstable();
          },
          total, pairs);
    };
    run(control);
    if (!control) expected[0] = expected[1] = expected[2] = 0.0;
    Means mm = path_mean(st, total[0], pairs[0]);
    if (mm.mean + expected[0] < 0.0) {  // asian.rs:618-630: resample without the control
      run(false);
      expected[0] = expected[1] = expected[2] = 0.0;
    }
  }
  Means mm = path_mean(st, total[0], pairs[0]);
  if (greeks) {
    greeks->available = req;
    greeks->delta = want_delta ? scaled_value(total[1].scale, total[1].mean, 0.0) + expected[1] : 0.0;
    greeks->gamma = want_gamma ? scaled_value(total[2].scale, total[2].mean, 0.0) + expected[2] : 0.0;
  }
  return from_sampling(st, mm.mean + expected[0], mm.has_se, mm.se);
}

inline Estimate asian_price(const Settings& st, const kg_asian_terms& o, const kg_bsm_market& m,
                            const Context& cx) {
  return asian_native(st, o, m, cx, 0, nullptr);
}

// ---------------------------------------------------------------- accumulator
// Accumulator schedule and history (instruments/accumulator.rs:319-568),
// restated with plain scans over the calendar days of the contract life.
struct AccumulatorSchedule {
  const kg_accumulator_terms& o;
  int64_t clock_at(int64_t day) const { return day * MICROS_PER_DAY + o.observation_clock_micros; }
  bool at_clock(int64_t t) const { return t - day_of(t) * MICROS_PER_DAY == o.observation_clock_micros; }
  bool trading(int64_t t) const { return is_trading_day(o.calendar, day_of(t)); }
  // Every accumulation fixing of the life (accumulation_schedule, :309-317).
  std::vector<int64_t> fixings() const {
    std::vector<int64_t> out;
    for (int64_t d = day_of(o.effective_time); d <= day_of(o.expiry_time); ++d) {
      const int64_t t = clock_at(d);
      if (t >= o.effective_time && t <= o.expiry_time && is_trading_day(o.calendar, d)) out.push_back(t);
    }
    return out;
  }
  bool explicit_schedule() const { return o.has_knock_out_schedule != 0; }
  bool observes(int64_t t) const {  // observes_knock_out (:512-517)
    if (!explicit_schedule()) return at_clock(t);
    for (int i = 0; i < o.n_knock_out_observations; ++i)
      if (o.knock_out_observation_times[i] == t) return true;
    return false;
  }
  bool accumulates(int64_t t) const {  // accumulates_at (:521-526)
    return o.effective_time <= t && t <= o.expiry_time && at_clock(t) && trading(t);
  }
  bool knocked(double spot) const { return o.option_type > 0 ? spot >= o.knock_out_level : spot <= o.knock_out_level; }
  double unit(double spot) const { return o.option_type > 0 ? spot - o.strike : o.strike - spot; }
  double purchase(double spot) const {  // purchase (:508-511)
    const bool accelerated = o.option_type > 0 ? spot - o.strike < 0.0 : spot - o.strike > 0.0;
    return accelerated ? o.daily_quantity * o.acceleration_factor : o.daily_quantity;
  }
};

inline Estimate accumulator_price(const Settings& st, const kg_accumulator_terms& o,
                                  const kg_bsm_market& m, const Context& cx) {
  if (o.option_type != 1 && o.option_type != -1) throw Failure(INVALID_INPUT, "option type");
  for (double v : {o.strike, o.knock_out_level})
    if (!(v > 0.0) || !std::isfinite(v)) throw Failure(INVALID_INPUT, "strike and level must be positive");
  for (double v : {o.daily_quantity, o.acceleration_factor, o.has_accumulated_quantity ? o.accumulated_quantity : 0.0})
    if (!(v >= 0.0) || !std::isfinite(v)) throw Failure(INVALID_INPUT, "quantities must be finite and nonnegative");
  if (o.expiry_time < o.effective_time) throw Failure(INVALID_INPUT, "expiry precedes effective date");
  const AccumulatorSchedule a{o};
  for (int i = 0; i < o.n_knock_out_observations; ++i) {
    const int64_t t = o.knock_out_observation_times[i];
    if ((i > 0 && t <= o.knock_out_observation_times[i - 1]) || t < o.effective_time || t > o.expiry_time ||
        !a.trading(t))
      throw Failure(INVALID_INPUT, "knock-out times");
  }
  if (m.volatility == 0.0) throw Failure(UNSUPPORTED, "zero volatility is deterministic");
  const int64_t val = m.valuation_time;
  if (val < o.effective_time || val > o.expiry_time) throw Failure(INVALID_INPUT, "valuation outside contract life");
  const std::vector<int64_t> all = a.fixings();
  // validate_pricing (:319-358)
  bool had_accumulation = false, had_knock_out = false;
  for (int64_t t : all) had_accumulation |= t < val;
  if (a.explicit_schedule()) {
    for (int i = 0; i < o.n_knock_out_observations; ++i) had_knock_out |= o.knock_out_observation_times[i] < val;
  } else {
    had_knock_out = had_accumulation;
  }
  if (had_knock_out && o.knock_out_settled == KG_SETTLED_UNKNOWN)
    throw Failure(MISSING_HISTORY, "prior knock-out settlement state is required");
  if ((had_accumulation || had_knock_out) && o.knock_out_settled != KG_SETTLED_YES && !o.has_accumulated_quantity)
    throw Failure(MISSING_HISTORY, "prior accumulated quantity is required");
  if (o.knock_out_settled == KG_SETTLED_YES && !had_knock_out)
    throw Failure(INVALID_INPUT, "settlement predates knock-out observations");
  if (o.knock_out_settled != KG_SETTLED_YES && o.has_accumulated_quantity && o.accumulated_quantity > 0.0 &&
      !had_accumulation)
    throw Failure(INVALID_INPUT, "accumulated quantity predates accumulation");
  // initial (:528-568)
  if (o.knock_out_settled == KG_SETTLED_YES) return known_estimate(st, 0.0);
  const double S = m.spot;
  double quantity = o.has_accumulated_quantity ? o.accumulated_quantity : 0.0;
  if (a.observes(val) && a.knocked(S) && a.trading(val)) return known_estimate(st, quantity * a.unit(S));
  if (a.accumulates(val)) quantity += a.purchase(S);
  bool accrues = false;  // accrues (:462-476)
  if (o.daily_quantity != 0.0 && val < o.expiry_time)
    for (int64_t t : all)
      accrues |= t > val && (o.acceleration_factor != 0.0 || !a.knocked(o.strike) || !a.observes(t));
  if (quantity == 0.0 && !accrues) return known_estimate(st, 0.0);
  if (val == o.expiry_time || (!accrues && m.rate == 0.0 && m.dividend == 0.0))
    return known_estimate(st, quantity * a.unit(S));
  if (!std::isfinite(quantity)) unstable();
  // Simulated dates and their purchases (future_observation_times :360-390,
  // fixed_purchase_events :401-454).
  std::vector<int64_t> future;
  for (int64_t t : all)
    if (t > val) future.push_back(t);
  std::vector<int64_t> knock_out_dates;
  for (int i = 0; i < o.n_knock_out_observations; ++i)
    if (o.knock_out_observation_times[i] > val) knock_out_dates.push_back(o.knock_out_observation_times[i]);
  struct Date {
    int64_t time;
    double before = 0.0, at = 0.0;
    bool observe = false, accumulate = false;
  };
  std::vector<Date> dates;
  bool fixed = false;
  if (a.explicit_schedule() && (o.acceleration_factor == 1.0 || o.daily_quantity == 0.0)) {
    std::vector<int64_t> events = knock_out_dates;
    if (events.empty() || events.back() != o.expiry_time) events.push_back(o.expiry_time);
    const std::vector<int64_t> purchases = o.daily_quantity == 0.0 ? std::vector<int64_t>{} : future;
    int64_t previous = val;
    size_t grouped = 0;
    for (int64_t e : events) {
      Date d;
      d.time = e;
      size_t strictly_between = 0;
      for (int64_t t : purchases) {
        if (t > previous && t < e) ++strictly_between;
        if (t == e) d.at = o.daily_quantity;
      }
      grouped += strictly_between;
      d.before = (double)strictly_between * o.daily_quantity;
      dates.push_back(d);
      previous = e;
    }
    fixed = (grouped >= 2 || o.daily_quantity == 0.0) &&
            std::isfinite((double)purchases.size() * o.daily_quantity);
  }
  if (!fixed) {
    dates.clear();
    std::vector<int64_t> times = future;
    times.insert(times.end(), knock_out_dates.begin(), knock_out_dates.end());
    times.push_back(o.expiry_time);
    std::sort(times.begin(), times.end());
    times.erase(std::unique(times.begin(), times.end()), times.end());
    for (int64_t t : times) {
      Date d;
      d.time = t;
      d.accumulate = a.accumulates(t);
      dates.push_back(d);
    }
  }
  struct Step {
    double drift, diffusion, discount;
  };
  std::vector<Step> steps;
  int64_t prev = val;
  for (Date& d : dates) {
    d.observe = a.observes(d.time);
    Step s;
    transition(m.rate, m.dividend, m.volatility, years_until(prev, d.time), years_until(val, d.time), &s.drift,
               &s.diffusion, &s.discount);
    steps.push_back(s);
    prev = d.time;
  }
  const double sign = o.option_type;
  const double log_spot = std::log(S);
  const double lo = std::log(std::numeric_limits<double>::min()) + 1.0;
  const double hi = std::log(std::numeric_limits<double>::max()) - 1.0;
  const double safe_lo = std::max(lo, lo - log_spot), safe_hi = std::min(hi, hi - log_spot);
  const double strike = log_ratio(o.strike, S), knock_out = log_ratio(o.knock_out_level, S);
  const double terminal_discount = steps.back().discount;
  auto unit_payoff = [&](double spot, double lr) {
    double distance = lr - strike;
    if (std::fabs(distance) < 0.5) return sign * o.strike * std::expm1(distance);
    return sign > 0 ? spot - o.strike : o.strike - spot;
  };
  auto spot_at = [&](double lr) {
    double s = scaled_value(S, 1.0, lr);
    if (s <= 0.0) unstable();
    return s;
  };
  // MonteCarloAccumulatorEngine::sample's path (engines/monte_carlo/accumulator.rs:181-214).
  Moments total[1], pairs[1];
  path_samples_lazy<1>(
      st, steps.size(), true, cx,
      [&](LazyPath& normals, double* out) {
        double lr = 0.0, held = quantity;
        for (size_t i = 0; i < steps.size(); ++i) {
          const Step& s = steps[i];
          const Date& d = dates[i];
          lr += std::fma(s.diffusion, normals.normal(), s.drift);
          if (!(lr >= safe_lo && lr <= safe_hi)) spot_at(lr);
          held += d.before;
          if (d.observe && sign * lr >= sign * knock_out) {
            out[0] = scaled_value(held, unit_payoff(spot_at(lr), lr), s.discount);
            return;
          }
          if (fixed)
            held += d.at;
          else if (d.accumulate)
            held += o.daily_quantity * (sign * lr < sign * strike ? o.acceleration_factor : 1.0);
          if (!std::isfinite(held)) unstable();
        }
        out[0] = scaled_value(held, unit_payoff(spot_at(lr), lr), terminal_discount);
      },
      total, pairs);
  return estimate(st, total[0], pairs[0], 0.0, false, true, st.seed);
}

}  // namespace oracle
