// Test infrastructure, not product code: the C entry points through which
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg drive the
// FP64 CPU restatement of kiyosi's Monte Carlo engines (core.hpp, plans.hpp).
#include <cstdio>
#include <mutex>

#include "american.hpp"
#include "book.hpp"
#include "core.hpp"
#include "heston.hpp"
#include "lv.hpp"
#include "plans.hpp"

using namespace oracle;

namespace {
std::vector<uint32_t> g_table;
thread_local std::string g_error;

Settings to_settings(const kg_settings* s, int workers) {
  if (!s->has_seed) throw Failure(INVALID_INPUT, "the oracle needs an explicit seed");
  if (s->path_count < 1) throw Failure(INVALID_INPUT, "path_count must be positive");
  Settings st;
  st.path_count = s->path_count;
  st.seed = s->seed;
  st.sampling = s->sampling;
  st.antithetic = s->antithetic != 0;
  st.workers = workers;
  if (st.sampling == SOBOL && (st.path_count & (st.path_count - 1)))
    throw Failure(INVALID_INPUT, "path_count must be a power of two for Sobol sampling");
  return st;
}
Context context() { return Context{g_table.data(), g_table.size()}; }

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_error.clear();
    return 0;
  } catch (const Failure& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 3;
  }
}

void write(const Estimate& e, kg_estimate* out) {
  *out = kg_estimate{};
  out->price = e.price;
  out->has_standard_error = e.has_standard_error;
  out->standard_error = e.standard_error;
  if (e.has_standard_error) {
    double hw = 1.959963984540054 * e.standard_error;
    out->ci_lower = e.price - hw;
    out->ci_upper = e.price + hw;
  }
  out->uncertainty_status = e.status;
  out->has_seed = e.has_seed;
  out->seed = e.seed;
  out->path_standard_error = e.path_standard_error;
}
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_error.c_str(); }

void oracle_set_sobol_table(const uint32_t* words, size_t n) { g_table.assign(words, words + n); }

void oracle_pcg32_outputs(uint64_t seed, uint64_t stream, uint64_t first, uint64_t count,
                          uint32_t* out) {
  Pcg32 g(seed, stream);
  g.advance(first);
  for (uint64_t i = 0; i < count; ++i) out[i] = g.next();
}

void oracle_pcg32_uniforms(uint64_t seed, uint64_t stream, uint64_t count, double* out) {
  Pcg32 g(seed, stream);
  for (uint64_t i = 0; i < count; ++i) out[i] = g.uniform();
}

// Normals first..first+count of a block stream, drawn one by one as
// NormalDraws::normal does (monte_carlo.rs:1678-1687).
void oracle_pcg_normals(uint64_t seed, uint64_t stream, uint64_t first, uint64_t count,
                        double* out) {
  NormalDraws d;
  d.gen = Pcg32(seed, stream);
  d.skip_normals(first);
  for (uint64_t i = 0; i < count; ++i) out[i] = d.normal();
}

void oracle_sin_cos_turns(const double* turns, size_t n, double* s, double* c) {
  for (size_t i = 0; i < n; ++i) sin_cos_turns(turns[i], &s[i], &c[i]);
}

void oracle_inverse_normal_cdf(const double* p, size_t n, double* out) {
  for (size_t i = 0; i < n; ++i) out[i] = inverse_normal_cdf(p[i]);
}

int oracle_sobol_points(uint32_t dims, uint64_t seed, uint64_t first, uint64_t count,
                        uint64_t* out) {
  return guarded([&] {
    SobolSequence s = sobol_new(dims, seed, g_table.data(), g_table.size());
    SobolCursor c(&s, first);
    for (uint64_t p = 0; p < count; ++p) {
      for (uint32_t d = 0; d < dims; ++d) out[p * dims + d] = c.values[d];
      c.next_path();
    }
  });
}

// Unscrambled, unshifted direction integers (52 bits) for dimension checks.
int oracle_sobol_raw_directions(uint32_t dims, uint64_t* out) {
  return guarded([&] {
    // Build with an identity scramble by reusing sobol_new's recurrence on a
    // copy of the table: the scramble is the identity when every matrix row
    // is a single bit, which sobol_new cannot express, so redo the recurrence.
    size_t pos = 0;
    for (uint32_t d = 0; d < dims; ++d) {
      uint64_t raw[SOBOL_BITS] = {};
      if (d == 0) {
        for (int c = 0; c < SOBOL_BITS; ++c) raw[c] = 1ULL << (SOBOL_BITS - c - 1);
      } else {
        int degree = (int)g_table.at(pos);
        uint32_t coeff = g_table.at(pos + 1);
        for (int c = 0; c < degree && c < SOBOL_BITS; ++c)
          raw[c] = (uint64_t)g_table.at(pos + 2 + c) << (SOBOL_BITS - c - 1);
        pos += degree + 2;
        for (int c = degree; c < SOBOL_BITS; ++c) {
          uint64_t v = raw[c - degree] ^ (raw[c - degree] >> degree);
          for (int o = 1; o < degree; ++o)
            if ((coeff >> (degree - o - 1)) & 1) v ^= raw[c - o];
          raw[c] = v;
        }
      }
      for (int c = 0; c < SOBOL_BITS; ++c) out[(size_t)d * SOBOL_BITS + c] = raw[c];
    }
  });
}

int oracle_vanilla_price(const kg_settings* s, const kg_european_terms* o,
                         const kg_bsm_market* m, int workers, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    write(european_price(st, *o, *m, context()), out);
  });
}

int oracle_autocallable_price(const kg_settings* s, const kg_autocall_terms* t,
                              const kg_bsm_market* m, int workers, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    write(autocall_price(st, *t, *m, context()), out);
  });
}

// The price of one note of a price_many_on_shared_grid batch: the note
// planned on its shared daily grid when it has one (the batch's prices are
// each note's price on the shared draws; its coupon-family and threshold-
// strip estimators reproduce them to rounding).
int oracle_autocallable_price_on_shared_grid(const kg_settings* s, const kg_autocall_terms* t,
                                             const kg_bsm_market* m, int workers, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    write(autocall_price(st, *t, *m, context(), true), out);
  });
}

int oracle_heston_price(const kg_settings* s, const kg_heston_terms* o, const kg_heston_market* m,
                        int workers, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    write(heston_price(st, *o, *m, context()), out);
  });
}

int oracle_barrier_price(const kg_settings* s, const kg_barrier_terms* o, const kg_bsm_market* m,
                         int workers, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    write(barrier_price(st, *o, *m, context()), out);
  });
}

void write_greeks(const SpotGreeks& g, kg_spot_greeks* out) {
  out->available = g.available;
  out->delta = g.delta;
  out->gamma = g.gamma;
}

int oracle_barrier_price_with_greeks(const kg_settings* s, const kg_barrier_terms* o, const kg_bsm_market* m,
                                     int workers, kg_spot_greeks* greeks, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    SpotGreeks g;
    write(barrier_native(st, barrier_spec(*o), *m, context(), greeks->request, greeks->spot_radius, &g), out);
    write_greeks(g, greeks);
  });
}

int oracle_double_barrier_price(const kg_settings* s, const kg_double_barrier_terms* o, const kg_bsm_market* m,
                                int workers, kg_spot_greeks* greeks, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    if (!(o->lower_barrier > 0.0) || !(o->upper_barrier > o->lower_barrier))
      throw Failure(INVALID_INPUT, "double barriers need 0 < lower < upper");
    SpotGreeks g;
    write(barrier_native(st, barrier_spec(*o), *m, context(), greeks ? greeks->request : 0,
                         greeks ? greeks->spot_radius : 0.0, &g),
          out);
    if (greeks) write_greeks(g, greeks);
  });
}

int oracle_digital_price(const kg_settings* s, const kg_digital_terms* o, const kg_bsm_market* m, int workers,
                         kg_spot_greeks* greeks, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    SpotGreeks g;
    write(digital_native(st, *o, *m, context(), greeks ? greeks->request : 0, &g), out);
    if (greeks) write_greeks(g, greeks);
  });
}

int oracle_double_survival(double initial, double terminal, double band, double deviation, int derivatives,
                           int branch, double* out) {
  return guarded([&] { double_survival(initial, terminal, band, deviation, derivatives != 0, out, branch); });
}

int oracle_single_hit_value(double amount, double distance, int up, double drift, double sigma, double time,
                            double rate, double* out) {
  return guarded([&] { *out = single_hit_value(amount, distance, up != 0, drift, sigma, time, rate); });
}

int oracle_double_hit_value(double amount, double lower, double upper, double spot, double rate, double dividend,
                            double sigma, double time, double* out) {
  return guarded([&] { *out = double_hit_value(amount, lower, upper, spot, rate, dividend, sigma, time); });
}

int oracle_double_exit_probability(double lower, double upper, double spot, double rate, double dividend,
                                   double sigma, double time, double* out) {
  return guarded([&] { *out = 1.0 - BandSurvival(lower, upper, spot, rate, dividend, sigma)(time); });
}

int oracle_asian_price(const kg_settings* s, const kg_asian_terms* o, const kg_bsm_market* m,
                       int workers, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    write(asian_price(st, *o, *m, context()), out);
  });
}

int oracle_asian_price_with_greeks(const kg_settings* s, const kg_asian_terms* o, const kg_bsm_market* m,
                                   int workers, kg_spot_greeks* greeks, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    SpotGreeks g;
    write(asian_native(st, *o, *m, context(), greeks->request, &g), out);
    write_greeks(g, greeks);
  });
}

int oracle_accumulator_price(const kg_settings* s, const kg_accumulator_terms* o,
                             const kg_bsm_market* m, int workers, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    write(accumulator_price(st, *o, *m, context()), out);
  });
}

// In-sample LSMC price (american_scenarios) and the fitted policy
// (steps-1 rows of: valid, c0, c1, c2).
int oracle_american_lsmc(const kg_settings* s, const kg_european_terms* o, const kg_bsm_market* m,
                         int steps, int workers, int mode, double* price, double* policy) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    std::vector<double> pol;
    *price = american_lsmc(st, *o, *m, steps, context(), &pol, mode);
    if (policy) std::copy(pol.begin(), pol.end(), policy);
  });
}

// Brownian coordinates of the LSMC paths, [step][sample] (diagnostic).
int oracle_american_coordinates(const kg_settings* s, int steps, int reverse, int workers, double* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    american_coordinates(st, steps, reverse != 0, context(), out);
  });
}

int oracle_american_out_of_sample(const kg_settings* s, int64_t training_paths,
                                  const kg_european_terms* o, const kg_bsm_market* m, int steps,
                                  int workers, double* price) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    *price = american_out_of_sample(st, training_paths, *o, *m, steps, context());
  });
}

// Finite-difference American price under a local-volatility surface.
int oracle_fd_american(const kg_local_vol_surface* surface, int64_t valuation, double spot, double strike,
                       double rate, double dividend, int64_t expiry, int option_type, int nx, int nt,
                       double* price) {
  return guarded([&] {
    Surface sf = make_surface(*surface);
    *price = fd_american(sf, years_until(surface->times[0], valuation), spot, strike, rate, dividend,
                         years_until(valuation, expiry), option_type, nx, nt);
  });
}

static LvSetup lv_setup(const kg_american_terms& o, const kg_american_model& m) {
  if (!m.surface) throw Failure(INVALID_INPUT, "a local-volatility surface is required");
  LvSetup L;
  L.surface = make_surface(*m.surface);
  L.year0 = years_until(m.surface->times[0], m.valuation_time);
  L.spot = m.spot;
  L.strike = o.strike;
  L.sign = o.option_type;
  L.rate = m.rate;
  L.dividend = m.dividend;
  L.steps = o.step_count;
  L.time = years_until(m.valuation_time, o.expiry_time);
  L.dt = L.time / o.step_count;
  L.log_discount = -m.rate * L.dt;
  L.discount = std::exp(L.log_discount);
  return L;
}

// LSMC under local volatility: in-sample price, and the out-of-sample price
// when training_paths > 0 (training seed = seed ^ 0x9e3779b97f4a7c15).
int oracle_lv_lsmc(const kg_settings* s, const kg_american_terms* o, const kg_american_model* m, int basis,
                   int64_t training_paths, int workers, double* in_sample, double* out_of_sample) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    LvSetup L = lv_setup(*o, *m);
    std::vector<double> pol;
    if (training_paths > 0) {
      Settings tr = st;
      tr.path_count = training_paths;
      tr.seed = st.seed ^ 0x9e3779b97f4a7c15ULL;
      *in_sample = lv_lsmc(tr, L, basis, &pol);
      *out_of_sample = lv_price(st, L, basis, pol);
    } else {
      *in_sample = lv_lsmc(st, L, basis, &pol);
      *out_of_sample = *in_sample;
    }
  });
}

// European payoff under local volatility on the LSMC engine's log-Euler
// stepping (diagnostic of the scheme; kind 0 vanilla, 1 cash digital).
int oracle_lv_european(const kg_settings* s, const kg_american_terms* o, const kg_american_model* m, int kind,
                       double cash, int workers, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    write(lv_european(st, lv_setup(*o, *m), kind, cash), out);
  });
}

// Terminal spot and variance estimates of the Heston scheme (diagnostic).
int oracle_heston_terminal(const kg_settings* s, const kg_heston_terms* o, const kg_heston_market* m,
                           int workers, kg_estimate* out) {
  return guarded([&] {
    Settings st = to_settings(s, workers);
    Estimate e[2];
    heston_terminal(st, *o, *m, context(), e);
    write(e[0], &out[0]);
    write(e[1], &out[1]);
  });
}

// Plan of an autocallable note as the oracle builds it, for plan parity tests:
// per step (drift, diffusion, discount, event, observe), and bridge flag.
int oracle_autocallable_plan(const kg_autocall_terms* t, const kg_bsm_market* m, double* steps,
                             size_t cap, size_t* n_steps, int* bridge) {
  return guarded([&] {
    Note note{t};
    AutocallPlan p = autocall_plan(note, *m);
    *n_steps = p.steps.size();
    *bridge = p.bridge;
    for (size_t i = 0; i < p.steps.size() && i < cap; ++i) {
      steps[i * 5 + 0] = p.steps[i].drift;
      steps[i * 5 + 1] = p.steps[i].diffusion;
      steps[i * 5 + 2] = p.steps[i].discount;
      steps[i * 5 + 3] = p.steps[i].event;
      steps[i * 5 + 4] = p.steps[i].observe_knock_in;
    }
  });
}
}
