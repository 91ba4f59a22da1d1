// Test infrastructure, not product code: local volatility on the CPU in
// FP64. It holds the reference's surface interpolation
// (market/local_volatility.rs:322-359), a Crank-Nicolson finite-difference
// American pricer that judges the Monte Carlo prices, and a restatement of
// the GPU engine's own local-volatility LSMC scheme (log-Euler steps with
// start-of-step sigma, normal equations of the chosen basis) for parity.
#pragma once

#include "../include/kiyosi_amd.h"
#include "american.hpp"
#include "core.hpp"
#include "plans.hpp"

namespace oracle {

struct Surface {
  std::vector<double> years, spots, values;  // values[time * ns + spot]
  int nt = 0, ns = 0;
};

inline Surface make_surface(const kg_local_vol_surface& s) {
  Surface out;
  out.nt = s.n_times;
  out.ns = s.n_spots;
  for (int i = 0; i < s.n_times; ++i) out.years.push_back(years_until(s.times[0], s.times[i]));
  out.spots.assign(s.spots, s.spots + s.n_spots);
  out.values.assign(s.values, s.values + (size_t)s.n_times * s.n_spots);
  return out;
}

// bracket + interpolate (local_volatility.rs:348-359)
inline std::pair<int, double> bracket(const std::vector<double>& nodes, double v) {
  int lo = (int)(std::upper_bound(nodes.begin(), nodes.end(), v) - nodes.begin());
  int left = std::min(std::max(lo - 1, 0), (int)nodes.size() - 2);
  double w = (v - nodes[left]) / (nodes[left + 1] - nodes[left]);
  return {left, std::min(std::max(w, 0.0), 1.0)};
}
inline double interp(double low, double high, double w) {
  return w <= 0.5 ? low + w * (high - low) : high + (1.0 - w) * (low - high);
}
inline double at_year(const Surface& s, double year, double spot) {
  auto [r, wt] = bracket(s.years, year);
  auto [c, ws] = bracket(s.spots, spot);
  const double* v0 = &s.values[(size_t)r * s.ns];
  const double* v1 = v0 + s.ns;
  return interp(interp(v0[c], v0[c + 1], ws), interp(v1[c], v1[c + 1], ws), wt);
}

// Crank-Nicolson in log-spot with four implicit half steps at the start
// (Rannacher) and projected SOR for the exercise constraint; sigma is read
// from the surface at the node spot and the time level being solved.
inline double fd_american(const Surface& s, double year0, double spot, double strike, double rate,
                          double dividend, double T, double sign, int nx, int nt) {
  const double x0 = std::log(spot);
  const double half = 3.0;
  const double dx = 2.0 * half / nx;
  std::vector<double> x(nx + 1), S(nx + 1), V(nx + 1), pay(nx + 1);
  for (int i = 0; i <= nx; ++i) {
    x[i] = x0 - half + i * dx;
    S[i] = std::exp(x[i]);
    pay[i] = std::max(sign * (S[i] - strike), 0.0);
    V[i] = pay[i];
  }
  auto coeffs = [&](double t, std::vector<double>& a, std::vector<double>& b, std::vector<double>& c) {
    for (int i = 1; i < nx; ++i) {
      const double sig = at_year(s, year0 + t, S[i]);
      const double v = sig * sig;
      const double mu = rate - dividend - 0.5 * v;
      a[i] = 0.5 * v / (dx * dx) - 0.5 * mu / dx;
      c[i] = 0.5 * v / (dx * dx) + 0.5 * mu / dx;
      b[i] = -v / (dx * dx) - rate;
    }
  };
  std::vector<double> a0(nx + 1), b0(nx + 1), c0(nx + 1), a1(nx + 1), b1(nx + 1), c1(nx + 1), rhs(nx + 1);
  std::vector<double> steps;
  const int ramp = 4;
  const double dtau = T / nt;
  for (int k = 0; k < ramp; ++k) steps.push_back(0.5 * dtau);
  for (int k = 2; k < nt; ++k) steps.push_back(dtau);
  double tau = 0.0;
  for (size_t n = 0; n < steps.size(); ++n) {
    const double h = steps[n];
    const double theta = n < (size_t)ramp ? 1.0 : 0.5;
    coeffs(T - tau, a0, b0, c0);
    coeffs(T - tau - h, a1, b1, c1);
    for (int i = 1; i < nx; ++i)
      rhs[i] = V[i] + (1.0 - theta) * h * (a0[i] * V[i - 1] + b0[i] * V[i] + c0[i] * V[i + 1]);
    // boundaries: deep in the money exercises, far out of the money is worthless
    V[0] = pay[0];
    V[nx] = pay[nx];
    const double omega = 1.4;
    for (int it = 0; it < 10000; ++it) {
      double err = 0.0;
      for (int i = 1; i < nx; ++i) {
        const double lower = -theta * h * a1[i], diag = 1.0 - theta * h * b1[i], upper = -theta * h * c1[i];
        const double gs = (rhs[i] - lower * V[i - 1] - upper * V[i + 1]) / diag;
        const double nv = std::max(pay[i], V[i] + omega * (gs - V[i]));
        err = std::max(err, std::fabs(nv - V[i]));
        V[i] = nv;
      }
      if (err < 1e-13) break;
    }
    tau += h;
  }
  return V[nx / 2];
}

struct LvSetup {
  Surface surface;
  double year0, spot, strike, sign, rate, dividend, time, dt, discount, log_discount;
  int steps;
};

inline double lv_step(const LvSetup& L, int k, double x, double z) {
  const double sig = at_year(L.surface, L.year0 + L.dt * k, L.spot * std::exp(x));
  return x + std::fma(sig * std::sqrt(L.dt), z, (L.rate - L.dividend - 0.5 * sig * sig) * L.dt);
}

inline bool solve_poly4(const double* e, double* out) {
  double g[5][5], b[5];
  int k = 0;
  for (int r = 0; r < 5; ++r)
    for (int c = r; c < 5; ++c) g[r][c] = g[c][r] = e[k++];
  for (int r = 0; r < 5; ++r) b[r] = e[15 + r];
  for (int degree = 4; degree >= 0; --degree) {
    double l[5][5] = {};
    double scale = 0.0;
    for (int r = 0; r <= degree; ++r) scale = std::max(scale, g[r][r]);
    bool ok = true;
    for (int c = 0; c <= degree && ok; ++c) {
      double d = g[c][c];
      for (int p = 0; p < c; ++p) d -= l[c][p] * l[c][p];
      if (!(d > 1e-13 * scale)) {
        ok = false;
        break;
      }
      l[c][c] = std::sqrt(d);
      for (int r = c + 1; r <= degree; ++r) {
        double v = g[r][c];
        for (int p = 0; p < c; ++p) v -= l[r][p] * l[c][p];
        l[r][c] = v / l[c][c];
      }
    }
    if (!ok) continue;
    double y[5], xs[5] = {0, 0, 0, 0, 0};
    for (int r = 0; r <= degree; ++r) {
      double v = b[r];
      for (int p = 0; p < r; ++p) v -= l[r][p] * y[p];
      y[r] = v / l[r][r];
    }
    for (int r = degree; r >= 0; --r) {
      double v = y[r];
      for (int p = r + 1; p <= degree; ++p) v -= l[p][r] * xs[p];
      xs[r] = v / l[r][r];
    }
    for (int r = 0; r < 5; ++r) {
      if (!std::isfinite(xs[r])) return false;
      out[r] = xs[r];
    }
    return true;
  }
  return false;
}

inline double continuation(const double* pol, int basis, double x) {
  if (basis == KG_BASIS_QUADRATIC) return pol[1] + x * (pol[2] + x * pol[3]);
  return pol[1] + x * (pol[2] + x * (pol[3] + x * (pol[4] + x * pol[5])));
}

// In-sample LSMC under local volatility with the GPU engine's stream layout
// (sample j of block b starts at normal j * steps of Pcg32(seed, b)); fills
// policy rows [valid, c0..c4, -, -].
inline double lv_lsmc(const Settings& st, const LvSetup& L, int basis, std::vector<double>* policy) {
  const size_t samples = st.sample_count(), blocks = st.blocks();
  const size_t paths = st.path_count;
  const int S = L.steps, legs = st.antithetic ? 2 : 1;
  std::vector<double> X((size_t)S * samples * legs, 0.0), spot(samples * legs), cash(samples * legs);
  std::vector<char> live(samples * legs, 0);
  auto payoff = [&](double s) { return std::max(L.sign * (s - L.strike), 0.0); };
  run_blocks(blocks, st.workers, [&](size_t b) {
    for (size_t j = 0; j < BLOCK && b * BLOCK + j < samples; ++j) {
      const size_t n = b * BLOCK + j;
      NormalDraws d;
      d.gen = Pcg32(st.seed, b);
      d.skip_normals(j * S);
      const bool paired = st.antithetic && n < paths / 2;
      double xp = 0.0, xn = 0.0;
      for (int k = 0; k < S; ++k) {
        const double z = d.normal();
        xp = lv_step(L, k, xp, z);
        if (paired) xn = lv_step(L, k, xn, -z);
        X[(size_t)k * samples * legs + n * legs] = xp;
        if (st.antithetic) X[(size_t)k * samples * legs + n * legs + 1] = xn;
      }
      for (int leg = 0; leg < (paired ? 2 : 1); ++leg) {
        const size_t slot = n * legs + leg;
        live[slot] = 1;
        spot[slot] = L.spot * std::exp(leg ? xn : xp);
        cash[slot] = payoff(spot[slot]);
      }
    }
  });
  const int E = basis == KG_BASIS_QUADRATIC ? 9 : 20;
  std::vector<double> pol((size_t)S * 8, 0.0);
  bool have = false;
  const double* prev = nullptr;
  for (int step = S - 2; step >= 0; --step) {
    std::vector<double> sums(E, 0.0);
    for (size_t slot = 0; slot < samples * legs; ++slot) {
      if (!live[slot]) continue;
      double c = cash[slot];
      if (have && prev[0] != 0.0) {
        const double intrinsic = payoff(spot[slot]);
        if (intrinsic > 0.0 && intrinsic / L.strike > continuation(prev, basis, spot[slot] / L.strike))
          c = intrinsic;
      }
      const double s = L.spot * std::exp(X[(size_t)step * samples * legs + slot]);
      c *= L.discount;
      spot[slot] = s;
      cash[slot] = c;
      if (payoff(s) > 0.0) {
        const double x = s / L.strike, t = c / L.strike;
        if (basis == KG_BASIS_QUADRATIC) {
          const double sq = x * x;
          const double e[9] = {1.0, x, sq, sq, x * sq, sq * sq, t, x * t, sq * t};
          for (int k = 0; k < 9; ++k) sums[k] += e[k];
        } else {
          double p[9];
          p[0] = 1.0;
          for (int i = 1; i < 9; ++i) p[i] = p[i - 1] * x;
          int k = 0;
          for (int r = 0; r < 5; ++r)
            for (int cc = r; cc < 5; ++cc) sums[k++] += p[r + cc];
          for (int r = 0; r < 5; ++r) sums[15 + r] += p[r] * t;
        }
      }
    }
    double* row = &pol[(size_t)step * 8];
    if (sums[0] != 0.0) {
      double coef[5] = {0, 0, 0, 0, 0};
      bool ok;
      if (basis == KG_BASIS_QUADRATIC) {
        double mat[3][4] = {{sums[0], sums[1], sums[2], sums[6]},
                            {sums[1], sums[3], sums[4], sums[7]},
                            {sums[2], sums[4], sums[5], sums[8]}};
        double c3[3];
        ok = solve_quadratic(mat, c3);
        for (int k = 0; k < 3; ++k) coef[k] = c3[k];
      } else {
        ok = solve_poly4(sums.data(), coef);
      }
      if (!ok) unstable();
      row[0] = 1.0;
      for (int k = 0; k < 5; ++k) row[1 + k] = coef[k];
    }
    have = true;
    prev = row;
  }
  Moments mean;
  for (size_t slot = 0; slot < samples * legs; ++slot) {
    if (!live[slot]) continue;
    double c = cash[slot];
    if (prev && prev[0] != 0.0) {
      const double intrinsic = payoff(spot[slot]);
      if (intrinsic > 0.0 && intrinsic / L.strike > continuation(prev, basis, spot[slot] / L.strike)) c = intrinsic;
    }
    mean.add(c, false);
  }
  if (policy) *policy = pol;
  return std::max(scaled_value(mean.scale, mean.mean, L.log_discount), payoff(L.spot));
}

// Out-of-sample valuation of a fitted policy under local volatility.
inline double lv_price(const Settings& st, const LvSetup& L, int basis, const std::vector<double>& pol) {
  auto payoff = [&](double s) { return std::max(L.sign * (s - L.strike), 0.0); };
  Moments total[1], pairs[1];
  path_samples_skipping<1>(
      st, L.steps, true, Context{},
      [&](SkippingPath& draws, double* out) {
        double x = 0.0, current = L.spot;
        for (int step = 0; step < L.steps; ++step) {
          x = lv_step(L, step, x, draws.normal());
          current = L.spot * std::exp(x);
          if (step == L.steps - 1) break;
          const double intrinsic = payoff(current);
          const double* row = &pol[(size_t)step * 8];
          if (intrinsic > 0.0 && row[0] != 0.0 &&
              intrinsic / L.strike > continuation(row, basis, current / L.strike)) {
            out[0] = scaled_value(1.0, intrinsic, -L.rate * (L.dt * (double)(step + 1)));
            return;
          }
        }
        out[0] = scaled_value(1.0, payoff(current), -L.rate * L.time);
      },
      total, pairs);
  return std::max(scaled_value(total[0].scale, total[0].mean, 0.0), payoff(L.spot));
}

// Diagnostic of the log-Euler local-volatility scheme (lv_step, the stepping
// of lv_lsmc and of the GPU engine) on a European payoff: kind 0 pays
// max(sign (S_T - K), 0), kind 1 pays `cash` when sign (S_T - K) > 0.
inline Estimate lv_european(const Settings& st, const LvSetup& L, int kind, double cash) {
  Moments total[1], pairs[1];
  path_samples_skipping<1>(
      st, L.steps, true, Context{},
      [&](SkippingPath& draws, double* out) {
        double x = 0.0;
        for (int step = 0; step < L.steps; ++step) x = lv_step(L, step, x, draws.normal());
        const double d = L.sign * (L.spot * std::exp(x) - L.strike);
        const double pay = kind == 0 ? std::max(d, 0.0) : (d > 0.0 ? cash : 0.0);
        out[0] = scaled_value(1.0, pay, -L.rate * L.time);
      },
      total, pairs);
  return estimate(st, total[0], pairs[0], 0.0, false, true, st.seed);
}

}  // namespace oracle
