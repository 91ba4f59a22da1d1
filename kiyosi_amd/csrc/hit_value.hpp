// Present value of a rebate paid at the first hit of a continuously
// monitored barrier, E[amount e^{-r tau} 1{tau <= T}], for the log spot
// X_t = x0 + mu t + sigma W_t. MonteCarloBarrierEngine adds it to the path
// mean of a continuous knock-out whose rebate is paid at hit, which its
// paths then carry as zero (engines/monte_carlo/barrier.rs:486-531).
//   single barrier: the closed form in the drift-tilted speed
//     lambda = sqrt(mu^2 + 2 r sigma^2) (engines/analytic/barrier.rs:361-417),
//     or, where lambda is not real, the first-passage density integrated in
//     1/sqrt(t) (:420-491);
//   double barrier: the first-exit flux of the absorbed density integrated
//     over time panels (engines/analytic/double_barrier.rs:423-520), its
//     density an image series for short times and a sine series for long
//     ones (:329-420); at zero rate the exit probability, one minus the
//     surviving mass (:321-328, :133-312).
// Host code: it runs once per pricing (three times with spot Greeks).
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "host_common.hpp"

namespace kg {
namespace hitv {

constexpr double kPi = 3.14159265358979323846;
constexpr double kLn2 = 0.69314718055994530942;
constexpr double kHalfLog2Pi = 0.91893853320467274178;  // ln(2 pi) / 2

// Neumaier-compensated running sum (numerics.rs:251-273).
struct Compensated {
  double value = 0.0, correction = 0.0;
  void add(double term) {
    const double next = value + term;
    correction += std::fabs(value) >= std::fabs(term) ? (value - next) + term : (term - next) + value;
    value = next;
  }
  double total() const { return value + correction; }
};

inline double normal_cdf(double x) { return 0.5 * std::erfc(-x * 0.70710678118654752440); }
// ln Phi(x) with the asymptotic erfc series below -10 (numerics.rs:196-214),
// where `exponent` carries -x^2/2 or a combined Gaussian exponent.
inline double log_tail(double x, double exponent) {
  const double inv_sq = (1.0 / x) / x;
  double term = 1.0, sum = 1.0;
  for (int i = 1; i <= 20; ++i) {
    term *= -(double)(2 * i - 1) * inv_sq;
    sum += term;
  }
  return exponent - std::log(-x) + std::log(sum / std::sqrt(2.0 * kPi));
}
inline double log_normal_cdf(double x) { return x >= -10.0 ? std::log(normal_cdf(x)) : log_tail(x, -0.5 * x * x); }
inline double log_difference(double high, double low) {  // ln(e^high - e^low)
  if (high == -INFINITY) return -INFINITY;
  return high + std::log(-std::expm1(low - high));
}
// ln of the standard normal mass on [lo, hi] (numerics.rs:275-318): a Taylor
// series about the midpoint for narrow intervals, else a difference of the
// tails on the far side of zero.
inline double log_interval_mass(double lo, double hi) {
  if (lo >= hi) return -INFINITY;
  if (lo == -INFINITY) return log_normal_cdf(hi);
  if (hi == INFINITY) return log_normal_cdf(-lo);
  const double width = hi - lo, half = 0.5 * width, center = 0.5 * (lo + hi);
  if (std::fabs(center) * half + 0.5 * half * half < 0.25) {
    // integral over [-1, 1] of e^{-center h u - h^2 u^2 / 2} du by the
    // recurrence of its Taylor coefficients in u.
    double integral = 0.0, previous = 0.0, term = 1.0;
    for (int order = 0; order < 24; ++order) {
      if (order % 2 == 0) integral += 2.0 * term / (double)(order + 1);
      const double next = (-center * half * term - half * half * previous) / (double)(order + 1);
      previous = term;
      term = next;
    }
    return -0.5 * center * center - kHalfLog2Pi + std::log(width) + std::log(0.5 * integral);
  }
  if (lo >= 0.0) return log_difference(log_normal_cdf(-lo), log_normal_cdf(-hi));
  return log_difference(log_normal_cdf(hi), log_normal_cdf(lo));
}
inline double scaled_cdf(double amount, double exponent, double argument) {  // amount e^exponent Phi(argument)
  return amount == 0.0 ? 0.0 : std::exp(std::log(amount) + exponent + log_normal_cdf(argument));
}
inline double scaled_signed_cash(double amount, double exponent) {
  return std::copysign(h_scaled_cash(std::fabs(amount), exponent), amount);
}
inline double standardize(double value, double volatility, double root) {
  const double width = volatility * root;
  return width < std::numeric_limits<double>::min() ? (value / volatility) / root : value / width;
}
inline double nonnegative(double value) {  // checked (numerics.rs:64-78)
  if (!std::isfinite(value) || value < 0.0) fail(KG_COMPUTATION_FAILURE, "negative or non-finite price");
  return value;
}

// Adaptive Simpson on [a, b] given f(a), f(mid), f(b) and the panel's
// Simpson estimate, with Richardson's correction (numerics.rs:12-50).
template <class F>
double simpson(const F& f, double a, double b, const double v[3], double estimate, double tolerance, int depth) {
  const double c = a + (b - a) / 2.0;
  const double fl = f(a + (c - a) / 2.0), fr = f(c + (b - c) / 2.0);
  const double left = (c - a) * (v[0] + 4.0 * fl + v[1]) / 6.0;
  const double right = (b - c) * (v[1] + 4.0 * fr + v[2]) / 6.0;
  const double refined = left + right;
  if (!std::isfinite(refined)) fail(KG_COMPUTATION_FAILURE, "barrier first-exit integral is not representable");
  if (std::fabs(refined - estimate) <= 15.0 * tolerance || c == a || c == b)
    return refined + (refined - estimate) / 15.0;
  if (depth == 0) fail(KG_NON_CONVERGENCE, "barrier first-exit integral did not converge");
  const double lv[3] = {v[0], fl, v[1]}, rv[3] = {v[1], fr, v[2]};
  return simpson(f, a, c, lv, left, tolerance / 2.0, depth - 1) + simpson(f, c, b, rv, right, tolerance / 2.0, depth - 1);
}

// Mills-ratio form of amount e^{e2 d} Phi(-(lambda T + d)/w) once e2
// overflows (reflection_tail, engines/analytic/barrier.rs:331-359, with zero
// strike offset and drift shift).
inline double reflected_tail(double amount, double log_weight, double distance, double speed, double sigma,
                             double time) {
  if (amount == 0.0 || std::signbit(speed) != std::signbit(distance)) return 0.0;
  const double root = std::sqrt(time);
  const double base = speed * time - distance;
  const double z = standardize(base, sigma, root);
  return std::exp(std::log(amount) + log_weight - 0.5 * z * z + std::log(sigma) + std::log(root) -
                  std::log(std::fabs(distance + speed * time)) - kHalfLog2Pi);
}

// E[amount e^{-r tau} 1{tau <= t}] by the first-passage density written in
// s = d / (sigma sqrt(u)), which removes its 1/u^{3/2} singularity
// (engines/analytic/barrier.rs:420-491).
inline double integrated_single(double amount, double distance, double mu, double sigma, double time, double rate) {
  if (amount == 0.0 || distance == 0.0) return amount;
  const double root = std::sqrt(time);
  const double lower = distance / sigma / root;  // s at u = time
  const double drift_width = mu / sigma * root;
  if (lower == 0.0) return amount;
  const double endpoint = -0.5 * (lower + drift_width) * (lower + drift_width) - rate * time;
  const double coefficient = 0.5 * drift_width * drift_width + rate * time;
  const double stationary = coefficient > 0.0 ? std::sqrt(lower * std::sqrt(2.0 * coefficient)) : 0.0;
  const double peak =
      stationary > lower ? -lower * drift_width - lower * std::sqrt(2.0 * coefficient) : endpoint;
  if (!std::isfinite(lower) || !std::isfinite(peak))
    fail(KG_COMPUTATION_FAILURE, "barrier hit integral is not representable");
  const double upper = lower + (lower < 1.0 ? 12.0 : 40.0 / lower);
  auto density = [&](double s) {
    const double ratio = lower / s;
    const double z = s + drift_width * ratio;
    return std::exp(-0.5 * z * z - rate * time * ratio * ratio - peak);
  };
  double integral = 0.0;
  for (double left = lower; left < upper;) {
    const double right = std::min(upper, 2.0 * left);
    const double v[3] = {density(left), density(left + (right - left) / 2.0), density(right)};
    const double estimate = (right - left) * (v[0] + 4.0 * v[1] + v[2]) / 6.0;
    integral += simpson(density, left, right, v, estimate, 1e-13 * (right - left) / std::max(upper - lower, 1.0), 20);
    left = right;
  }
  if (integral == 0.0) return 0.0;
  return nonnegative(h_scaled_cash(amount, peak + 0.5 * std::log(2.0 / kPi) + std::log(integral)));
}

// Single barrier at log distance `distance` >= 0 from the spot
// (hit_value, engines/analytic/barrier.rs:361-417). `drift` is
// r - q - sigma^2 / 2; the passage is toward the barrier, so an up barrier
// sees the drift reversed.
inline double single(double amount, double distance, bool up, double drift, double sigma, double time, double rate) {
  const double variance = sigma * sigma;
  const double mu = up ? -drift : drift;
  if (variance == 0.0)
    return mu >= 0.0 ? 0.0
                     : nonnegative(scaled_cdf(amount, rate * (distance / mu),
                                              standardize(-mu * time - distance, sigma, std::sqrt(time))));
  const double disc = mu * mu + 2.0 * rate * variance;
  const double scale = std::max({1.0, mu * mu, std::fabs(2.0 * rate * variance)});
  if (disc < 0.0 || (rate < 0.0 && disc <= 16.0 * std::numeric_limits<double>::epsilon() * scale))
    return integrated_single(amount, distance, mu, sigma, time, rate);
  const double lambda = std::sqrt(disc);
  const double width = sigma * std::sqrt(time);
  // (-mu -/+ lambda) / sigma^2, each in the form that does not cancel.
  const double e1 = mu < 0.0 ? -2.0 * rate / (lambda - mu) : (-mu - lambda) / variance;
  const double e2 = mu > 0.0 ? 2.0 * rate / (lambda + mu) : (-mu + lambda) / variance;
  const double first = scaled_cdf(amount, e1 * distance, (lambda * time - distance) / width);
  const double second = std::isfinite(e2) ? scaled_cdf(amount, e2 * distance, (-lambda * time - distance) / width)
                                          : reflected_tail(amount, e1 * distance, distance, lambda, sigma, time);
  return nonnegative(first + second);
}

// Log spot absorbed at both barriers (AbsorbingDistribution,
// engines/analytic/double_barrier.rs:101-131): distances from the lower
// barrier (initial) and to the upper one, the band width, ln(lower), and
// the drift, volatility and time to expiry.
struct Band {
  double initial, upper_distance, band, log_lower, sigma, drift, time;

  Band(double lower, double upper, double spot, double r, double q, double vol, double t)
      : initial(log_ratio(spot, lower)),
        upper_distance(log_ratio(upper, spot)),
        band(log_ratio(upper, lower)),
        log_lower(std::log(lower)),
        sigma(vol),
        drift(r - q - 0.5 * vol * vol),
        time(t) {}

  // Discounted moment e^{discount} E[S_T^order 1{ln(S_T/lower) in
  // [lo, hi]}] over paths that survive both barriers, or over those that
  // touched one (complement) (moment, :133-320).
  double moment(int order_i, double lo, double hi, double discount, bool complement) const {
    if (lo >= hi) return 0.0;
    const double order = (double)order_i;
    const double dev = sigma * std::sqrt(time), variance = dev * dev, ratio = dev / band;
    if (dev <= 0.0 || !std::isfinite(dev) || !std::isfinite(drift * time))
      fail(KG_COMPUTATION_FAILURE, "double-barrier transition is not representable");
    const double a = std::max(lo, 0.0), b = std::min(hi, band);
    const double base = order * (log_lower + initial + drift * time) + 0.5 * order * order * variance + discount;
    const double mean = initial + drift * time + order * variance;
    auto free_mass = [&](double x, double y) {
      return std::exp(base + log_interval_mass((x - mean) / dev, (y - mean) / dev));
    };
    if (a >= b) return complement ? nonnegative(free_mass(lo, hi)) : 0.0;
    Compensated sum;
    double scale = 0.0;
    bool converged = false;
    if (ratio < 0.5) {
      // Method of images: the free density reflected k bands away, and
      // through the lower barrier.
      auto image = [&](double k, bool reflected) {
        const double shift = k * band;
        const double distance = reflected ? initial + shift : shift;
        const double center = mean - 2.0 * distance;
        const double exponent = distance == 0.0 ? base : base - 2.0 * distance * ((drift / sigma) / sigma + order);
        const double zl = (a - center) / dev, zh = (b - center) / dev;
        if (zl >= zh) return 0.0;
        if (zl <= 10.0 && zh >= -10.0) return std::exp(exponent + log_interval_mass(zl, zh));
        // Far tails: fold the reflection into the Gaussian exponent before
        // forming the tail, so a small volatility does not cancel it.
        auto tail = [&](double boundary, double z) {
          const double free_z = (boundary - mean) / dev;
          const double rb = reflected ? boundary + shift : boundary - initial + shift;
          const double reflection = (distance == 0.0 || rb == 0.0) ? 0.0 : 2.0 * (distance / dev) * (rb / dev);
          return log_tail(z, base - 0.5 * free_z * free_z - reflection);
        };
        const double large = zh < 0.0 ? tail(b, zh) : tail(a, -zl);
        const double small = zh < 0.0 ? tail(a, zl) : tail(b, -zh);
        return std::exp(log_difference(large, small));
      };
      if (complement) {
        // Outside the band every path touched; inside, the free density
        // cancels the direct image.
        sum.add(free_mass(lo, a));
        sum.add(free_mass(b, hi));
      } else {
        sum.add(image(0.0, false));
      }
      const double orientation = complement ? -1.0 : 1.0;
      sum.add(-orientation * image(0.0, true));
      scale = std::max(1.0, std::fabs(sum.total()));
      for (int i = 1; i <= 64; ++i) {
        const double k = (double)i;
        const double t[4] = {image(k, false), image(k, true), image(-k, false), image(-k, true)};
        sum.add(orientation * t[0]);
        sum.add(-orientation * t[1]);
        sum.add(orientation * t[2]);
        sum.add(-orientation * t[3]);
        if (i >= 2 && t[0] + t[1] + t[2] + t[3] <= 1e-16 * scale) {
          converged = true;
          break;
        }
      }
    } else {
      // Eigenfunction (sine) series of the absorbed density.
      const double drift_ratio = (drift / sigma) / sigma;
      const double slope = (drift_ratio + order) * band;
      auto weight = [&](double y) {
        return discount + order * (log_lower + y) + drift_ratio * (y - initial) -
               0.5 * (drift / sigma) * (drift / sigma) * time;
      };
      const double wa = weight(a), wb = weight(b);
      for (int i = 1; i <= 256; ++i) {
        const double freq = (double)i * kPi;
        const double decay = 0.5 * (freq * ratio) * (freq * ratio);
        const double denom = std::hypot(slope, freq);
        auto endpoint = [&](double y, double log_weight) {
          const double sine = (y == 0.0 || y == band) ? 0.0 : std::sin(freq * (y / band));
          const double cosine = y == 0.0 ? 1.0 : y == band ? (i % 2 == 0 ? 1.0 : -1.0) : std::cos(freq * (y / band));
          const double coefficient = ((slope / denom) * sine - (freq / denom) * cosine) / denom;
          return scaled_signed_cash(coefficient, log_weight - decay);
        };
        const double term = 2.0 * std::sin(freq * (initial / band)) * (endpoint(b, wb) - endpoint(a, wa));
        sum.add(term);
        scale = std::max({1.0, scale, std::fabs(term)});
        const double next = (double)(i + 1) * kPi;
        if (i >= 4 && std::max(wa, wb) - 0.5 * (next * ratio) * (next * ratio) < std::log(scale) - 40.0) {
          converged = true;
          break;
        }
      }
      if (complement) {
        Compensated difference;
        difference.add(free_mass(lo, hi));
        difference.add(-sum.total());
        sum = difference;
      }
    }
    const double value = sum.total();
    if (!std::isfinite(value) || value < -1e-12 * scale)
      fail(KG_COMPUTATION_FAILURE, "double-barrier density series produced an invalid moment");
    if (!converged) fail(KG_NON_CONVERGENCE, "double-barrier density series did not converge");
    return std::max(value, 0.0);
  }

  // ln of the first-exit density at time origin + offset, discounted at
  // `rate`, summed over both barriers (log_exit_density, :329-420).
  double log_exit_density(double origin, double offset, double rate) const {
    const double t = origin + offset;
    if (t == 0.0) return -INFINITY;
    const double dev = sigma * std::sqrt(t), ratio = dev / band;
    auto flux = [&](double distance, double opposite, double mu) {
      Compensated sum;
      bool converged = false;
      double base;
      if (ratio < 0.5) {
        // Inverse-Gaussian density times its image corrections. A narrow
        // drift-driven peak is centred on the time offset first.
        const double z = mu > 0.0 ? -mu * ((origin - distance / mu) + offset) / dev : (distance - mu * t) / dev;
        base = std::log(distance) - std::log(sigma) - 1.5 * std::log(t) - kHalfLog2Pi - 0.5 * z * z - rate * t;
        sum.add(1.0);
        for (int i = 1; i <= 64; ++i) {
          const double shift = (double)i * band;
          const double reflected = (double)(i - 1) * band + opposite;
          const double plus = scaled_signed_cash(1.0 + 2.0 * shift / distance,
                                                 -2.0 * (shift / dev) * ((shift + distance) / dev));
          const double minus = scaled_signed_cash(-1.0 - 2.0 * reflected / distance,
                                                  -2.0 * (shift / dev) * (reflected / dev));
          sum.add(plus);
          sum.add(minus);
          if (i >= 2 && std::fabs(plus) + std::fabs(minus) <= 1e-16) {
            converged = true;
            break;
          }
        }
      } else {
        const double first_decay = 0.5 * (kPi * ratio) * (kPi * ratio);
        base = 2.0 * std::log(sigma / band) - first_decay + (mu / sigma) * ((distance - 0.5 * mu * t) / sigma) -
               rate * t;
        const double near = std::min(distance, opposite) / band;
        for (int i = 1; i <= 256; ++i) {
          const double freq = (double)i * kPi;
          double sine = std::sin(freq * near);
          if (distance > opposite && i % 2 == 0) sine = -sine;  // sin(k pi (1 - x)) = (-1)^{k+1} sin(k pi x)
          sum.add(freq * sine * std::exp(first_decay - 0.5 * (freq * ratio) * (freq * ratio)));
          const double next = (double)(i + 1) * kPi;
          if (i >= 4 && std::log(next) + first_decay - 0.5 * (next * ratio) * (next * ratio) < -40.0) {
            converged = true;
            break;
          }
        }
      }
      const double value = sum.total();
      if (!std::isfinite(value) || value < -1e-12 || std::isnan(base) || base == INFINITY)
        fail(KG_COMPUTATION_FAILURE, "double-barrier exit density is not representable");
      if (!converged) fail(KG_NON_CONVERGENCE, "double-barrier exit density did not converge");
      return base + std::log(std::max(value, 0.0));
    };
    const double lower = flux(initial, upper_distance, -drift);
    const double upper = flux(upper_distance, initial, drift);
    const double top = std::max(lower, upper);
    if (top == -INFINITY) return top;
    return top + std::log(std::exp(lower - top) + std::exp(upper - top));
  }

  // hit_value (:423-520).
  double hit(double amount, double rate) const {
    if (rate == 0.0) return moment(0, -INFINITY, INFINITY, std::log(amount), true);
    const double width = sigma * std::sqrt(time);
    // Upper bounds on ln P(hit) from each side, from the reflection principle
    // and, against the drift, from the probability of ever hitting.
    auto log_bound_from = [&](double distance, double toward) {
      const double z = std::max(distance - std::max(toward, 0.0) * time, 0.0) / width;
      const double reflected = kLn2 - 0.5 * z * z;
      return toward < 0.0 ? std::min(reflected, 2.0 * ((toward / sigma) * (distance / sigma))) : reflected;
    };
    const double largest = std::max(log_bound_from(initial, -drift), log_bound_from(upper_distance, drift));
    const double log_bound = std::log(amount) + std::max(-rate * time, 0.0) + kLn2 + largest;
    if (log_bound < std::log(std::numeric_limits<double>::denorm_min()) - kLn2) return 0.0;
    // Panel breaks where the flux changes shape: diffusive times of both
    // distances, drift and tilted-drift arrival times and their spreads,
    // and the discount's time scale before expiry.
    std::vector<double> breaks{0.0, time};
    auto add = [&](double t) {
      if (t > 0.0 && t < time) breaks.push_back(t);
    };
    const double tilted = std::sqrt(std::max(drift * drift + 2.0 * rate * sigma * sigma, 0.0));
    for (double distance : {initial, upper_distance}) {
      for (double f : {0.0625, 1.0, 16.0}) add(f * (distance / sigma) * (distance / sigma));
      for (double speed : {std::fabs(drift), tilted}) {
        if (!(speed > 0.0)) continue;
        const double arrival = distance / speed;
        const double spread = sigma * std::sqrt(arrival) / speed;
        add(arrival);
        for (double f : {-12.0, -4.0, -1.0, 1.0, 4.0, 12.0}) add(arrival + f * spread);
      }
    }
    for (double f : {1.0, 4.0, 16.0, 64.0}) add(time - f / std::fabs(rate));
    std::sort(breaks.begin(), breaks.end());
    breaks.erase(std::unique(breaks.begin(), breaks.end()), breaks.end());
    struct Panel {
      double left, width, logs[3], scale;
    };
    std::vector<Panel> panels;
    double integral_scale = -INFINITY;
    for (size_t k = 0; k + 1 < breaks.size(); ++k) {
      const double end = breaks[k + 1];
      for (double left = breaks[k]; left < end;) {
        // Geometric panels so the early exit density does not set the
        // error scale of a long tail.
        const double right = left > 0.0 ? std::min(end, 4.0 * left) : end;
        Panel p{left, right - left, {}, -INFINITY};
        p.logs[0] = log_exit_density(left, 0.0, rate);
        p.logs[1] = log_exit_density(left, p.width / 2.0, rate);
        p.logs[2] = log_exit_density(left, p.width, rate);
        p.scale = std::max({p.logs[0], p.logs[1], p.logs[2]});
        if (p.scale != -INFINITY) {
          integral_scale = std::max(integral_scale, std::log(p.width) + p.scale);
          panels.push_back(p);
        }
        left = right;
      }
    }
    Compensated result;
    const double count = (double)panels.size();
    for (const Panel& p : panels) {
      auto f = [&](double u) { return std::exp(log_exit_density(p.left, p.width * u, rate) - p.scale); };
      const double v[3] = {std::exp(p.logs[0] - p.scale), std::exp(p.logs[1] - p.scale),
                           std::exp(p.logs[2] - p.scale)};
      const double integral =
          simpson(f, 0.0, 1.0, v, (v[0] + 4.0 * v[1] + v[2]) / 6.0,
                  1e-13 * std::exp(integral_scale - std::log(p.width) - p.scale) / count, 32);
      result.add(nonnegative(h_scaled_cash(integral, std::log(amount) + std::log(p.width) + p.scale)));
    }
    const double value = nonnegative(result.total());
    if (rate > 0.0 && value > amount) {
      if (value - amount > 1e-12 * amount)
        fail(KG_COMPUTATION_FAILURE, "discounted first-exit value exceeds its payout");
      return amount;
    }
    return value;
  }
};

}  // namespace hitv
}  // namespace kg
