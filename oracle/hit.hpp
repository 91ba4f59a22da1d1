// Test infrastructure: the checker's value of a rebate paid at the first hit
// of a continuously monitored barrier, E[A e^{-r tau} 1{tau <= T}], for the
// log spot x0 + mu t + sigma W_t. MonteCarloBarrierEngine adds this value to
// the path mean of a continuous knock-out paid at hit
// (engines/monte_carlo/barrier.rs:486-531). The derivations here are the
// checker's own, independent of the reference's numerics
// (engines/analytic/barrier.rs:361-491, double_barrier.rs:329-520):
//   single barrier: the first-passage Laplace transform on [0, T] in closed
//     form when mu^2 + 2 r sigma^2 > 0, else adaptive quadrature of the
//     inverse-Gaussian density in u = sqrt(t);
//   double barrier: integration by parts against the survival probability
//     S(t) of the killed motion, A [1 - e^{-rT} S(T) - r int_0^T e^{-rt} S(t) dt],
//     with S from its sine (eigenfunction) series.
#pragma once

#include <algorithm>
#include <cmath>

#include "core.hpp"

namespace oracle {

inline double hit_log_normal_cdf(double x) {
  if (x >= -10.0) return std::log(0.5 * std::erfc(-x / std::sqrt(2.0)));
  // Asymptotic series of the Mills ratio.
  double term = 1.0, sum = 1.0;
  for (int i = 1; i <= 20; ++i) {
    term *= -(2.0 * i - 1.0) / (x * x);
    sum += term;
  }
  return -0.5 * x * x - std::log(-x) - 0.5 * std::log(2.0 * M_PI) + std::log(sum);
}

template <class F>
double hit_simpson(const F& f, double a, double fa, double m, double fm, double b, double fb, double whole,
                   double tol, int depth) {
  const double lm = 0.5 * (a + m), rm = 0.5 * (m + b);
  const double flm = f(lm), frm = f(rm);
  const double left = (m - a) / 6.0 * (fa + 4.0 * flm + fm), right = (b - m) / 6.0 * (fm + 4.0 * frm + fb);
  if (depth <= 0 || std::fabs(left + right - whole) <= 15.0 * tol) return left + right + (left + right - whole) / 15.0;
  return hit_simpson(f, a, fa, lm, flm, m, fm, left, tol / 2, depth - 1) +
         hit_simpson(f, m, fm, rm, frm, b, fb, right, tol / 2, depth - 1);
}
template <class F>
double hit_integrate(const F& f, double a, double b, double tol) {
  // Forty equal panels, each refined adaptively.
  double total = 0.0;
  const int panels = 40;
  for (int k = 0; k < panels; ++k) {
    const double x0 = a + (b - a) * k / panels, x1 = a + (b - a) * (k + 1) / panels, xm = 0.5 * (x0 + x1);
    const double f0 = f(x0), fm = f(xm), f1 = f(x1);
    total += hit_simpson(f, x0, f0, xm, fm, x1, f1, (x1 - x0) / 6.0 * (f0 + 4.0 * fm + f1), tol / panels, 40);
  }
  return total;
}

// Single barrier at log distance d >= 0; `drift` = r - q - sigma^2/2. An up
// barrier is a down barrier for the mirrored motion.
inline double single_hit_value(double amount, double d, bool up, double drift, double sigma, double T, double r) {
  if (amount == 0.0) return 0.0;
  if (d == 0.0) return amount;
  const double mu = up ? -drift : drift, s2 = sigma * sigma;
  const double disc = mu * mu + 2.0 * r * s2;
  const double scale = std::max({1.0, mu * mu, std::fabs(2.0 * r * s2)});
  if (disc > 16.0 * 2.220446049250313e-16 * scale || (r >= 0.0 && disc >= 0.0)) {
    const double lam = std::sqrt(disc), w = sigma * std::sqrt(T);
    // The first exponent (-mu - lam)/s2 in the form without cancellation.
    const double e1 = mu < 0.0 ? -2.0 * r / (lam - mu) : (-mu - lam) / s2;
    const double e2 = (-mu + lam) / s2;
    const double a = std::log(amount) + e1 * d + hit_log_normal_cdf((lam * T - d) / w);
    const double b = std::log(amount) + e2 * d + hit_log_normal_cdf((-lam * T - d) / w);
    return std::exp(a) + std::exp(b);
  }
  // Density of tau: d / (sigma sqrt(2 pi) t^{3/2}) exp(-(d + mu t)^2 / (2 sigma^2 t)), t = u^2.
  auto f = [&](double u) {
    if (u == 0.0) return 0.0;
    const double t = u * u;
    const double e = -(d + mu * t) * (d + mu * t) / (2.0 * s2 * t) - r * t;
    return 2.0 * d / (sigma * std::sqrt(2.0 * M_PI) * t) * std::exp(e);
  };
  return amount * hit_integrate(f, 0.0, std::sqrt(T), 1e-15);
}

// Probability that the log spot stays inside (ln lower, ln upper) up to t:
// the Girsanov-weighted Dirichlet sine series of the killed transition
// density, integrated over the band in closed form.
struct BandSurvival {
  double L, x0, mu, s2, k;
  BandSurvival(double lower, double upper, double spot, double r, double q, double sigma)
      : L(std::log(upper / lower)),
        x0(std::log(spot / lower)),
        mu(r - q - 0.5 * sigma * sigma),
        s2(sigma * sigma),
        k(mu / s2) {}
  double operator()(double t) const {
    const double decay_unit = M_PI * M_PI * s2 * t / (2.0 * L * L);
    const int n_max = 8 + (int)std::ceil(std::sqrt(80.0 / decay_unit));
    double sum = 0.0;
    for (int n = 1; n <= n_max; ++n) {
      const double w = n * M_PI / L;
      // int_0^L e^{k y} sin(w y) dy = w (1 - (-1)^n e^{kL}) / (k^2 + w^2).
      const double alt = (n % 2 == 0) ? 1.0 : -1.0;
      const double base = -k * x0 - 0.5 * mu * mu * t / s2 - decay_unit * n * n;
      const double c = w / (k * k + w * w);
      sum += (2.0 / L) * std::sin(w * x0) * c * (std::exp(base) - alt * std::exp(base + k * L));
    }
    return sum;
  }
};

// Double barrier: spot between `lower` and `upper`.
inline double double_hit_value(double amount, double lower, double upper, double spot, double r, double q,
                               double sigma, double T) {
  if (amount == 0.0) return 0.0;
  const BandSurvival survival(lower, upper, spot, r, q, sigma);
  // Before t0 the motion is more than ten deviations from either barrier.
  const double m = std::min(survival.x0, survival.L - survival.x0);
  double t0 = std::pow(m / (20.0 * sigma), 2);
  if (std::fabs(survival.mu) > 0.0) t0 = std::min(t0, m / (2.0 * std::fabs(survival.mu)));
  t0 = std::min(t0, T);
  const double sT = survival(T);
  if (r == 0.0) return amount * (1.0 - sT);
  auto g = [&](double t) { return r * std::exp(-r * t) * survival(t); };
  const double head = -std::expm1(-r * t0);  // int_0^t0 r e^{-rt} dt with S = 1
  const double tail = t0 < T ? hit_integrate(g, t0, T, 1e-16) : 0.0;
  return amount * (1.0 - std::exp(-r * T) * sT - head - tail);
}

}  // namespace oracle
