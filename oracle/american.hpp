// Test infrastructure, not product code: FP64 CPU restatement of kiyosi's
// least-squares Monte Carlo for American options (one strike, BSM; forward,
// reverse and anchored paths; pseudorandom or Sobol draws), the checker of
// the HIP LSMC kernels.
// Citations are to crates/kiyosi/src/engines/monte_carlo/american.rs.
#pragma once

#include "../include/kiyosi_amd.h"
#include "core.hpp"
#include "plans.hpp"

namespace oracle {

// solve_quadratic (:1693-1749): Gaussian elimination with partial pivoting on
// the 3x4 normal equations, dropping to degree 1 and 0 when singular.
inline bool solve_quadratic(const double (&samples)[3][4], double (&out)[3]) {
  for (int degree = 2; degree >= 0; --degree) {
    double m[3][4];
    std::memcpy(m, samples, sizeof(m));
    double scale = 0.0;
    for (int r = 0; r <= degree; ++r)
      for (int c = 0; c <= degree; ++c) scale = std::max(scale, std::fabs(m[r][c]));
    bool singular = false;
    for (int col = 0; col <= degree; ++col) {
      int pivot = col;
      for (int r = col + 1; r <= degree; ++r)
        if (std::fabs(m[r][col]) >= std::fabs(m[pivot][col])) pivot = r;  // max_by keeps the last maximum
      if (!std::isfinite(m[pivot][col])) return false;
      if (std::fabs(m[pivot][col]) <= 1e-14 * scale) {
        singular = true;
        break;
      }
      for (int c = 0; c < 4; ++c) std::swap(m[col][c], m[pivot][c]);
      double prow[4];
      std::memcpy(prow, m[col], sizeof(prow));
      for (int r = col + 1; r <= degree; ++r) {
        double factor = m[r][col] / m[col][col];
        for (int c = col; c <= degree; ++c) m[r][c] -= factor * prow[c];
        m[r][3] -= factor * m[col][3];
      }
    }
    if (singular) continue;
    double coef[3] = {0.0, 0.0, 0.0};
    for (int r = degree; r >= 0; --r) {
      double v = m[r][3];
      for (int c = r + 1; c <= degree; ++c) v -= m[r][c] * coef[c];
      coef[r] = v / m[r][r];
    }
    for (double c : coef)
      if (!std::isfinite(c)) return false;
    std::memcpy(out, coef, sizeof(coef));
    return true;
  }
  return false;
}

struct AmericanSetup {
  double sign, strike, spot, drift, diffusion, discount, log_discount, time;
  int intervals;
  double cutoff;
};

inline AmericanSetup american_setup(const kg_european_terms& o, const kg_bsm_market& m, int steps) {
  AmericanSetup a;
  a.sign = o.option_type;
  a.strike = o.strike;
  a.spot = m.spot;
  a.time = years_until(m.valuation_time, o.expiry_time);
  a.intervals = steps;
  const double dt = a.time / steps;
  a.drift = (m.rate - m.dividend - 0.5 * m.volatility * m.volatility) * dt;  // :1229-1235
  a.diffusion = m.volatility * std::sqrt(dt);
  a.log_discount = -m.rate * dt;
  a.discount = std::exp(a.log_discount);
  double lg = log_ratio(o.strike, m.spot);
  a.cutoff = a.sign * lg - 32.0 * std::numeric_limits<double>::epsilon() * (1.0 + std::fabs(lg));
  return a;
}

// spots_at (:1166-1215) for one scenario: the spot after step+1 intervals
// from the normalised Brownian coordinate, or the strike for paths the
// screen proves out of the money.
struct SpotsAt {
  double scale, group_cutoff, diffusion, sign, strike;
  bool screen;
  double operator()(double coordinate) const {
    double lg = diffusion * coordinate;
    if (screen && sign * lg < group_cutoff) return strike;
    return scale * std::exp(lg);
  }
};
inline SpotsAt spots_at(const AmericanSetup& a, int step) {
  SpotsAt s;
  double elapsed = step + 1;
  double drifts = a.drift * elapsed;
  s.group_cutoff = a.cutoff - a.sign * drifts;
  s.screen = s.group_cutoff > 0.5 * a.diffusion * std::sqrt(elapsed);
  s.scale = a.spot * std::exp(drifts);
  s.diffusion = a.diffusion;
  s.sign = a.sign;
  s.strike = a.strike;
  return s;
}

// Paths of american_scenarios (:1184-1219): the normalised case stores one
// Brownian coordinate per sample (the antithetic leg is its negative), in
// forward mode as the running sum of the path's normals (ExercisePaths::fill,
// :365-447) or, when samples x intervals exceeds 2^25 path values
// (PREFERRED_PATH_WORKSPACE / 8, :22, :1218-1219), in reverse mode: the
// terminal coordinate first and each earlier one from the Brownian bridge
// (reverse_row, :278-316), step-major within a block. The anchored case
// (spot or drift too large for exp of a coordinate, :1149-1157) stores each
// leg's spot, stepped forward by exp(diffusion z + drift) (advance and
// evolve_spot, :355-362, :546-552).
constexpr size_t REVERSE_PATH_VALUES = (size_t(256) << 20) / sizeof(double);

inline bool american_normalized(double spot, double drift, double diffusion, int steps) {
  return std::fabs(std::log(spot)) + (std::fabs(drift) + 10.0 * std::fabs(diffusion)) * steps < 600.0;
}

inline bool american_reverse(const Settings& st, double spot, double drift, double diffusion, int steps) {
  return american_normalized(spot, drift, diffusion, steps) && st.sample_count() * (size_t)steps > REVERSE_PATH_VALUES;
}

// reverse_row (:278-316) for one block: moves `coords` (the block's samples)
// from the coordinate after step + 2 intervals to the one after step + 1,
// or draws the terminal coordinate when step is the last.
inline void reverse_row(const Settings& st, const SobolSequence* seq, size_t block, int intervals, int step,
                        std::vector<double>& coords, NormalDraws& draws) {
  const int elapsed = step + 1;
  const int dimension = intervals - elapsed;
  double weight, deviation;
  if (dimension == 0) {
    weight = 0.0;
    deviation = std::sqrt((double)elapsed);
  } else {
    const double fraction = (double)elapsed / (double)(elapsed + 1);
    weight = fraction;
    deviation = std::sqrt(fraction);
  }
  const size_t begin = block * BLOCK;
  if (st.sampling == SOBOL) {
    // dimension_normals (sobol.rs:155-174): the points begin.. of one
    // dimension in Gray-code order.
    uint64_t bits = seq->shifts[dimension];
    uint64_t gray = begin ^ (begin >> 1);
    while (gray) {
      bits ^= seq->directions[(size_t)__builtin_ctzll(gray) * seq->dims + dimension];
      gray &= gray - 1;
    }
    for (size_t i = 0; i < coords.size(); ++i) {
      const double normal = normal_from_bits(bits);
      bits ^= seq->directions[(size_t)__builtin_ctzll(begin + i + 1) * seq->dims + dimension];
      coords[i] = std::fma(weight, coords[i], deviation * normal);
    }
  } else {
    for (double& c : coords) c = std::fma(weight, c, deviation * draws.normal());
  }
}

// american_scenarios::<1, 1> (:955-1455) for one strike under BSM, in the
// mode the reference picks: forward with the full path matrix (windows and
// replay change no value), reverse, or anchored. Returns the in-sample price;
// `policy` receives the per-date coefficients (valid flag + 3 values per
// date).
// `mode` -1 picks the reference's mode; 0 and 1 force forward and reverse
// paths (a diagnostic for the tests: the reference decides by size alone).
inline double american_lsmc(const Settings& st, const kg_european_terms& o, const kg_bsm_market& m,
                            int steps, const Context& cx, std::vector<double>* policy, int mode = -1) {
  const double sign = o.option_type;
  const double immediate = std::max(sign * (m.spot - o.strike), 0.0);
  AmericanSetup a = american_setup(o, m, steps);
  if (a.time == 0.0) return immediate;
  if (steps < 2) throw Failure(INVALID_INPUT, "American Monte Carlo requires at least two exercise intervals");
  if (m.volatility == 0.0) throw Failure(UNSUPPORTED, "zero volatility is deterministic");
  if (o.option_type > 0 && m.dividend <= 0.0 && m.rate >= 0.0)
    throw Failure(UNSUPPORTED, "calls without dividends are priced as Europeans");
  const bool normalized = american_normalized(m.spot, a.drift, a.diffusion, steps);
  const bool reverse =
      mode < 0 ? american_reverse(st, m.spot, a.drift, a.diffusion, steps) : (normalized && mode == 1);
  const size_t paths = st.path_count;
  const size_t samples = st.sample_count();
  const size_t blocks = st.blocks();
  SobolSequence seq;
  if (st.sampling == SOBOL) seq = sobol_new(steps, st.seed, cx.table, cx.table_len);
  auto paired_in = [&](size_t begin, size_t end) -> size_t {
    if (!st.antithetic) return 0;
    const size_t top = std::min(end, paths / 2);
    return top > begin ? top - begin : 0;
  };
  // Coordinates per block: forward [step][lane]; reverse [lane], one row;
  // anchored [step][slot] with a paired sample's two legs side by side.
  std::vector<std::vector<double>> coords(blocks);
  std::vector<NormalDraws> reverse_draws(blocks);
  std::vector<std::vector<double>> spots(blocks), cash(blocks);
  auto payoff = [&](double s) { return std::max(sign * (s - a.strike), 0.0); };
  auto advance = [&](double& value, double normal) {
    value *= std::exp(std::fma(a.diffusion, normal, a.drift));
    if (!std::isfinite(value) || value <= 0.0) unstable();
    return value;
  };
  SpotsAt terminal = spots_at(a, steps - 1);
  run_blocks(blocks, st.workers, [&](size_t b) {
    size_t begin = b * BLOCK, end = std::min(begin + BLOCK, samples);
    size_t count = end - begin;
    const size_t paired = paired_in(begin, end);
    if (reverse) {
      coords[b].assign(count, 0.0);
      reverse_draws[b].gen = Pcg32(st.seed, b);
      reverse_row(st, &seq, b, steps, steps - 1, coords[b], reverse_draws[b]);
    } else {
      const size_t width = normalized ? count : count + paired;
      coords[b].assign(width * steps, 0.0);
      SobolCursor cursor;
      NormalDraws d;
      d.gen = Pcg32(st.seed, b);
      if (st.sampling == SOBOL) {
        cursor = SobolCursor(&seq, begin);
        d.cursor = &cursor;
      }
      size_t slot = 0;
      for (size_t lane = 0; lane < count; ++lane) {
        if (normalized) {
          double w = 0.0;
          for (int k = 0; k < steps; ++k) {
            w += d.normal();
            coords[b][k * width + lane] = w;
          }
        } else {
          const bool two = lane < paired;
          double pos = m.spot, neg = m.spot;
          for (int k = 0; k < steps; ++k) {
            const double normal = d.normal();
            coords[b][k * width + slot] = advance(pos, normal);
            if (two) coords[b][k * width + slot + 1] = advance(neg, -normal);
          }
          slot += two ? 2 : 1;
        }
        if (d.cursor) cursor.next_path();
      }
    }
    // visit_row: paired samples visit +W then -W; anchored rows hold both legs.
    const size_t row = reverse ? 0 : (size_t)(steps - 1);
    const size_t width = reverse || normalized ? count : count + paired;
    for (size_t lane = 0; lane < width; ++lane) {
      double w = coords[b][row * width + lane];
      double s = normalized ? terminal(w) : w;
      spots[b].push_back(s);
      cash[b].push_back(payoff(s));
      if (normalized && lane < paired) {
        double s2 = terminal(-w);
        spots[b].push_back(s2);
        cash[b].push_back(payoff(s2));
      }
    }
  });
  bool have = false;
  double coef[3] = {0, 0, 0};
  auto apply = [&](double s, double& c) {
    if (!have) return;
    double intrinsic = payoff(s);
    if (intrinsic <= 0.0) return;
    double x = s / a.strike;
    double cont = coef[0] + x * (coef[1] + x * coef[2]);
    if (!std::isfinite(cont)) unstable();
    if (intrinsic / a.strike > cont) c = intrinsic;
  };
  if (policy) policy->assign((size_t)(steps - 1) * 4, 0.0);
  std::vector<std::array<double, 12>> parts(blocks);
  for (int step = steps - 2; step >= 0; --step) {
    SpotsAt at = spots_at(a, step);
    run_blocks(blocks, st.workers, [&](size_t b) {
      size_t begin = b * BLOCK, end = std::min(begin + BLOCK, samples);
      size_t count = end - begin;
      const size_t paired = paired_in(begin, end);
      if (reverse) reverse_row(st, &seq, b, steps, step, coords[b], reverse_draws[b]);
      double mat[3][4] = {};
      size_t state = 0;
      auto visit = [&](double v) {
        double& s = spots[b][state];
        double& c = cash[b][state];
        ++state;
        apply(s, c);
        s = normalized ? at(v) : v;
        c *= a.discount;
        if (!std::isfinite(c)) unstable();
        if (payoff(s) > 0.0) {
          double x = s / a.strike, sq = x * x, t = c / a.strike;
          mat[0][0] += 1.0;
          mat[0][1] += x;
          mat[0][2] += sq;
          mat[1][1] += sq;
          mat[1][2] += x * sq;
          mat[2][2] += sq * sq;
          mat[0][3] += t;
          mat[1][3] += x * t;
          mat[2][3] += sq * t;
        }
      };
      if (normalized) {
        const size_t row = reverse ? 0 : (size_t)step;
        for (size_t lane = 0; lane < count; ++lane) {
          double w = coords[b][row * count + lane];
          visit(w);
          if (lane < paired) visit(-w);
        }
      } else {
        const size_t width = count + paired;
        for (size_t slot = 0; slot < width; ++slot) visit(coords[b][(size_t)step * width + slot]);
      }
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) parts[b][r * 4 + c] = mat[r][c];
    });
    double mat[3][4] = {};
    for (size_t b = 0; b < blocks; ++b)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) mat[r][c] += parts[b][r * 4 + c];
    have = false;
    if (mat[0][0] != 0.0) {
      for (auto& row : mat)
        for (double v : row)
          if (!std::isfinite(v)) unstable();
      mat[1][0] = mat[0][1];
      mat[2][0] = mat[0][2];
      mat[2][1] = mat[1][2];
      if (!solve_quadratic(mat, coef)) unstable();
      have = true;
    }
    if (policy) {
      (*policy)[step * 4] = have;
      for (int k = 0; k < 3; ++k) (*policy)[step * 4 + 1 + k] = have ? coef[k] : 0.0;
    }
  }
  Moments mean;
  for (size_t b = 0; b < blocks; ++b)
    for (size_t i = 0; i < cash[b].size(); ++i) {
      double c = cash[b][i];
      apply(spots[b][i], c);
      mean.add(c, false);
    }
  return std::max(scaled_value(mean.scale, mean.mean, a.log_discount), immediate);
}

// Diagnostic of the path generation: the normalised Brownian coordinates
// after 1..steps intervals for every sample, out[step * samples + sample],
// drawn in forward or reverse mode exactly as american_lsmc draws them.
inline void american_coordinates(const Settings& st, int steps, bool reverse, const Context& cx, double* out) {
  const size_t samples = st.sample_count();
  const size_t blocks = st.blocks();
  SobolSequence seq;
  if (st.sampling == SOBOL) seq = sobol_new(steps, st.seed, cx.table, cx.table_len);
  run_blocks(blocks, st.workers, [&](size_t b) {
    const size_t begin = b * BLOCK, count = std::min(begin + BLOCK, samples) - begin;
    NormalDraws d;
    d.gen = Pcg32(st.seed, b);
    if (reverse) {
      std::vector<double> coords(count, 0.0);
      for (int step = steps - 1; step >= 0; --step) {
        reverse_row(st, &seq, b, steps, step, coords, d);
        for (size_t i = 0; i < count; ++i) out[(size_t)step * samples + begin + i] = coords[i];
      }
      return;
    }
    SobolCursor cursor;
    if (st.sampling == SOBOL) {
      cursor = SobolCursor(&seq, begin);
      d.cursor = &cursor;
    }
    for (size_t i = 0; i < count; ++i) {
      double w = 0.0;
      for (int k = 0; k < steps; ++k) {
        w += d.normal();
        out[(size_t)k * samples + begin + i] = w;
      }
      if (d.cursor) cursor.next_path();
    }
  });
}

// price_american_out_of_sample (:574-637): fit on training paths with seed
// ^ 0x9e3779b97f4a7c15, then stream independent valuation paths through the
// fitted exercise rule.
inline double american_out_of_sample(const Settings& st, int64_t training_paths,
                                     const kg_european_terms& o, const kg_bsm_market& m, int steps,
                                     const Context& cx) {
  Settings train = st;
  train.path_count = training_paths;
  train.seed = st.seed ^ 0x9e3779b97f4a7c15ULL;
  std::vector<double> policy;
  american_lsmc(train, o, m, steps, cx, &policy);
  AmericanSetup a = american_setup(o, m, steps);
  const double sign = o.option_type;
  auto payoff = [&](double s) { return std::max(sign * (s - a.strike), 0.0); };
  const double rate = m.rate;
  const double dt = a.time / steps;
  Moments total[1], pairs[1];
  path_samples_skipping<1>(
      st, steps, false, cx,
      [&](SkippingPath& draws, double* out) {
        double current = a.spot;
        for (int step = 0; step < steps - 1; ++step) {
          current *= std::exp(std::fma(a.diffusion, draws.normal(), a.drift));
          if (!std::isfinite(current) || current <= 0.0) unstable();
          double intrinsic = payoff(current);
          if (intrinsic > 0.0 && policy[step * 4] != 0.0) {
            double x = current / a.strike;
            double cont = policy[step * 4 + 1] + x * (policy[step * 4 + 2] + x * policy[step * 4 + 3]);
            if (!std::isfinite(cont)) unstable();
            if (intrinsic / a.strike > cont) {
              out[0] = scaled_value(1.0, intrinsic, -rate * (dt * (double)(step + 1)));
              return;
            }
          }
        }
        current *= std::exp(std::fma(a.diffusion, draws.normal(), a.drift));
        if (!std::isfinite(current) || current <= 0.0) unstable();
        out[0] = scaled_value(1.0, payoff(current), -rate * a.time);
      },
      total, pairs);
  return std::max(scaled_value(total[0].scale, total[0].mean, 0.0), payoff(a.spot));
}

}  // namespace oracle
