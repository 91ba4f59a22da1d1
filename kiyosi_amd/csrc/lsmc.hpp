// Shared declarations of the American (LSMC) engine: the device model, the
// host plan and the MLP regressor entry (lsmc.hip, mlp.hip).
#pragma once

#include <algorithm>
#include <vector>

#include "host_common.hpp"

namespace kg {

// How a sample's path is stored (american.rs:1149-1219): Brownian
// coordinates (normalised BSM, the antithetic leg negated), log-spot per
// path slot under local volatility, or the spot itself per path slot when
// ln S0 and the path's drift and diffusion could overflow exp of a
// coordinate (anchored BSM paths).
enum PathKind : int32_t { PATHS_BROWNIAN = 0, PATHS_LOCAL_VOL = 1, PATHS_ANCHORED = 2 };

// Paths above this many stored values are drawn in reverse (american.rs:22,
// 1218-1219): PREFERRED_PATH_WORKSPACE of 256 MiB in FP64 values.
constexpr double REVERSE_PATH_VALUES = 33554432.0;

// Scrambled Sobol directions of the sampler, [column * dims + dim], or
// dirs == nullptr for pseudorandom draws.
struct SobolRef {
  const uint64_t* dirs;
  const uint64_t* shifts;
  int32_t dims;
};

struct LsmcModel {
  int32_t lv, steps;
  int32_t anchored;  // PATHS_ANCHORED: spots stored per slot
  int32_t reverse;   // Brownian coordinates drawn terminal-first, one row kept
  double spot, strike, sign, rate, dividend, sigma, time, dt;
  double drift, diffusion, discount, log_discount;
  double year0;  // years from the surface's first row to valuation
  // BSM reference reconstruction per exercise date (spots_at, american.rs:1166-1215)
  const double* scale;
  const double* group_cutoff;
  const int32_t* screen;
  // local-volatility table (market/local_volatility.rs:61-359)
  int32_t nt, ns;
  const double* years;
  const double* spots;
  const double* vols;
  // Spot nodes equally spaced in log (to 1e-12): bracket index
  // floor((ln S - log_s0) inv_h), then corrected against the nodes.
  int32_t log_uniform;
  double log_s0, inv_h, log_spot;
};

struct LsmcHost {
  bool known = false;
  double known_value = 0.0;
  LsmcModel model{};
  Sampler pricing{}, training{};
  bool oos = false;
  double immediate = 0.0;
  // Host tables before upload: per-date scale, group cutoff and screen
  // flags, then (local volatility) surface years, spots and values from
  // lv_offset on. bind_tables points the model at their device copy.
  std::vector<double> tables;
  size_t lv_offset = 0;
};

inline void bind_tables(LsmcModel& m, const double* d, size_t lv_offset) {
  const int S = m.steps;
  m.scale = d;
  m.group_cutoff = d + S;
  m.screen = (const int32_t*)(d + 2 * S);
  if (m.lv) {
    m.years = d + lv_offset;
    m.spots = m.years + m.nt;
    m.vols = m.spots + m.ns;
  }
}

// Validates one request and builds its model and host tables (nothing on
// the device yet).
inline LsmcHost lsmc_plan_host(const kg_settings& settings, const kg_american_terms& o,
                               const kg_american_model& mk, const kg_lsmc_options& opt) {
  LsmcHost h;
  if (o.option_type != 1 && o.option_type != -1) fail(KG_INVALID_INPUT, "option_type must be 1 or -1");
  if (!(o.strike > 0.0) || !std::isfinite(o.strike)) fail(KG_INVALID_INPUT, "strike must be positive");
  if (o.step_count < 2 || o.step_count > 10000)
    fail(KG_INVALID_INPUT, "American Monte Carlo requires 2..=10,000 exercise intervals");
  if (opt.basis < KG_BASIS_QUADRATIC || opt.basis > KG_BASIS_MLP) fail(KG_INVALID_INPUT, "unknown basis");
  if (!(mk.spot > 0.0) || !std::isfinite(mk.spot)) fail(KG_INVALID_INPUT, "spot must be positive");
  if (mk.valuation_time < o.effective_time || mk.valuation_time > o.expiry_time)
    fail(KG_INVALID_INPUT, "valuation outside contract life");
  if (settings.sampling == KG_SOBOL && opt.basis == KG_BASIS_MLP)
    fail(KG_UNSUPPORTED_OPERATION, "the MLP regressor trains on pseudorandom paths");
  const double sign = o.option_type;
  h.immediate = std::max(sign * (mk.spot - o.strike), 0.0);
  const double time = (double)(o.expiry_time - mk.valuation_time) / (365.0 * 86400e6);
  h.pricing = validate_settings(&settings, (uint64_t)o.step_count);
  if (time == 0.0) {
    h.known = true;
    h.known_value = h.immediate;
    return h;
  }
  h.oos = opt.out_of_sample != 0;
  if (h.oos) {
    kg_settings t = settings;
    t.path_count = opt.training_path_count;
    t.seed = settings.seed ^ 0x9e3779b97f4a7c15ULL;  // american.rs:593
    h.training = validate_settings(&t, (uint64_t)o.step_count);
  }
  const bool lv = mk.surface != nullptr;
  if (!lv) {
    if (!(mk.volatility > 0.0) || !std::isfinite(mk.volatility))
      fail(KG_UNSUPPORTED_OPERATION, "the LSMC engine needs a positive volatility");
    if (o.option_type > 0 && mk.dividend <= 0.0 && mk.rate >= 0.0)
      fail(KG_UNSUPPORTED_OPERATION, "calls without dividends are priced by the European engine");
  }
  LsmcModel& m = h.model;
  m.lv = lv;
  m.steps = o.step_count;
  m.spot = mk.spot;
  m.strike = o.strike;
  m.sign = sign;
  m.rate = mk.rate;
  m.dividend = mk.dividend;
  m.sigma = mk.volatility;
  m.time = time;
  m.dt = time / o.step_count;
  m.drift = (mk.rate - mk.dividend - 0.5 * mk.volatility * mk.volatility) * m.dt;
  m.diffusion = mk.volatility * std::sqrt(m.dt);
  m.log_discount = -mk.rate * m.dt;
  m.discount = std::exp(m.log_discount);
  std::vector<double>& tables = h.tables;
  const int S = o.step_count;
  tables.resize(2 * (size_t)S + (size_t)S);  // scale, cutoff, screen (as int pairs)
  if (!lv) {
    const bool normalized =
        std::fabs(std::log(mk.spot)) + (std::fabs(m.drift) + 10.0 * std::fabs(m.diffusion)) * S < 600.0;
    if (!normalized && opt.basis == KG_BASIS_MLP)
      fail(KG_UNSUPPORTED_OPERATION, "the MLP regressor needs normalised (non-anchored) paths");
    m.anchored = !normalized;
    // The in-sample paths decide the mode, as in american_scenarios; the
    // MLP regressor (no reference counterpart) always draws forward.
    const Sampler& in_sample = h.oos ? h.training : h.pricing;
    m.reverse = normalized && opt.basis != KG_BASIS_MLP &&
                (double)in_sample.samples() * (double)S > REVERSE_PATH_VALUES;
    const double lg = log_ratio(o.strike, mk.spot);
    const double cutoff = sign * lg - 32.0 * std::numeric_limits<double>::epsilon() * (1.0 + std::fabs(lg));
    for (int k = 0; k < S; ++k) {
      const double elapsed = k + 1;
      const double drifts = m.drift * elapsed;
      const double gc = cutoff - sign * drifts;
      tables[k] = mk.spot * std::exp(drifts);
      tables[S + k] = gc;
      // Screen flags are packed as int32 in the third block of S doubles.
      reinterpret_cast<int32_t*>(&tables[2 * (size_t)S])[k] = gc > 0.5 * m.diffusion * std::sqrt(elapsed);
    }
  }
  const kg_local_vol_surface* sf = mk.surface;
  h.lv_offset = tables.size();
  if (lv) {
    if (sf->n_times < 2 || sf->n_spots < 2 || !sf->times || !sf->spots || !sf->values)
      fail(KG_INVALID_INPUT, "a local-volatility surface needs at least two rows and columns");
    for (int i = 1; i < sf->n_times; ++i)
      if (sf->times[i] <= sf->times[i - 1]) fail(KG_INVALID_INPUT, "surface times must increase");
    for (int i = 0; i < sf->n_spots; ++i)
      if (!(sf->spots[i] > 0.0) || (i && sf->spots[i] <= sf->spots[i - 1]))
        fail(KG_INVALID_INPUT, "surface spots must be positive and increasing");
    for (int i = 0; i < sf->n_times * sf->n_spots; ++i)
      if (!(sf->values[i] > 0.0) || !std::isfinite(sf->values[i]))
        fail(KG_INVALID_INPUT, "surface volatilities must be positive and finite");
    if (mk.valuation_time < sf->times[0] || o.expiry_time > sf->times[sf->n_times - 1])
      fail(KG_INVALID_INPUT, "time outside local-volatility coverage");
    m.nt = sf->n_times;
    m.ns = sf->n_spots;
    m.log_spot = std::log(mk.spot);
    m.log_s0 = std::log(sf->spots[0]);
    const double h = std::log(sf->spots[sf->n_spots - 1] / sf->spots[0]) / (sf->n_spots - 1);
    bool uniform = h > 0.0;
    for (int i = 1; i < sf->n_spots && uniform; ++i)
      uniform = std::fabs(std::log(sf->spots[i] / sf->spots[i - 1]) - h) <= 1e-12 * h + 1e-15;
    m.log_uniform = uniform;
    m.inv_h = uniform ? 1.0 / h : 0.0;
    m.year0 = (double)(mk.valuation_time - sf->times[0]) / (365.0 * 86400e6);
    for (int i = 0; i < sf->n_times; ++i)
      tables.push_back((double)(sf->times[i] - sf->times[0]) / (365.0 * 86400e6));
    for (int i = 0; i < sf->n_spots; ++i) tables.push_back(sf->spots[i]);
    for (int i = 0; i < sf->n_times * sf->n_spots; ++i) tables.push_back(sf->values[i]);
  }
  return h;
}

// lsmc_plan_host with its tables uploaded.
inline LsmcHost lsmc_plan(kg_context* ctx, const kg_settings& settings, const kg_american_terms& o,
                          const kg_american_model& mk, const kg_lsmc_options& opt) {
  LsmcHost h = lsmc_plan_host(settings, o, mk, opt);
  if (h.known) return h;
  double* d = (double*)ctx->lsmc_tables.reserve(std::max<size_t>(h.tables.size(), 1) * sizeof(double));
  KG_HIP(hipMemcpyAsync(d, h.tables.data(), h.tables.size() * sizeof(double), hipMemcpyHostToDevice,
                        ctx->stream));
  KG_HIP(hipStreamSynchronize(ctx->stream));
  bind_tables(h.model, d, h.lv_offset);
  return h;
}

// Fits and prices with the MLP continuation regressor (mlp.hip).
void american_mlp(kg_context* ctx, LsmcHost& h, const kg_lsmc_options& opt, kg_lsmc_report* out);

}  // namespace kg
