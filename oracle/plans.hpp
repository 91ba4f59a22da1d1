// Test infrastructure, not product code: plan building of the kiyosi 0.23.0
// European and autocallable Monte Carlo engines, restated on the CPU in FP64
// as the checker of the HIP engine.
#pragma once

#include "../include/kiyosi_amd.h"
#include "core.hpp"

namespace oracle {

constexpr int64_t MICROS_PER_DAY = 86400LL * 1000000LL;
inline double years_until(int64_t a, int64_t b) {  // time/date.rs:255-257
  return (double)(b - a) / (365.0 * (double)MICROS_PER_DAY);
}
inline int64_t floor_div(int64_t a, int64_t b) {
  int64_t q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}
inline int64_t day_of(int64_t t) { return floor_div(t, MICROS_PER_DAY); }
inline int64_t clock_of(int64_t t) { return t - day_of(t) * MICROS_PER_DAY; }
// TradingCalendar::is_trading_day for weekdays / all_days (time/calendar.rs:146-161, 317).
inline bool is_trading_day(int calendar, int64_t day) {
  if (calendar == KG_CALENDAR_ALL_DAYS) return true;
  int64_t w = ((day % 7) + 7 + 3) % 7;  // 1970-01-01 is a Thursday; Monday = 0
  return w < 5;
}

// transition (monte_carlo.rs:2123-2137)
inline void transition(double r, double q, double sigma, double time, double discount_time,
                       double* drift, double* diffusion, double* log_discount) {
  *drift = (r - q - 0.5 * sigma * sigma) * time;
  *diffusion = sigma * std::sqrt(time);
  *log_discount = -r * discount_time;
  if (!std::isfinite(*drift) || !std::isfinite(*diffusion) || !std::isfinite(*log_discount))
    unstable();
}

// ---------------------------------------------------------------- European
// EuropeanPayoff / TerminalPayoff (monte_carlo.rs:1165-1252).
struct TerminalPayoff {
  double sign, diffusion, half_variance, moneyness, log_scale;
  bool needs_peak_scaling() const { return sign > 0.0 && diffusion > 25.0; }
  double at_normal(double z) const {
    double distance = std::fma(diffusion, z, moneyness);
    double p;
    if (sign * distance <= 0.0)
      p = 0.0;
    else if (sign < 0.0)
      p = -std::expm1(distance);
    else
      p = -std::expm1(-distance) * std::exp(std::fma(diffusion, z, -half_variance));
    if (!std::isfinite(p)) unstable();
    return p;
  }
  double log_payoff(double z) const {  // :1107-1117
    double distance = std::fma(diffusion, z, moneyness);
    if (distance > 0.0)
      return std::fma(diffusion, z, log_scale - half_variance) + std::log(-std::expm1(-distance));
    return -std::numeric_limits<double>::infinity();
  }
};

// Returns true with *known set when no sampling is needed.
inline bool european_plan(const kg_european_terms& o, const kg_bsm_market& m, TerminalPayoff* pay,
                          double* known) {
  if (o.option_type != 1 && o.option_type != -1)
    throw Failure(INVALID_INPUT, "option_type must be 1 (call) or -1 (put)");
  // validate_life with EuropeanOption::valuation_start() == None
  // (instruments/monitoring.rs:470-481, instruments/vanilla.rs:198-200).
  if (m.valuation_time > o.expiry_time) throw Failure(INVALID_INPUT, "valuation outside contract life");
  double sign = o.option_type;
  double time = years_until(m.valuation_time, o.expiry_time);
  if (time == 0.0) {
    *known = std::max(sign * (m.spot - o.strike), 0.0);
    return true;
  }
  double sigma = m.volatility;
  if (sigma == 0.0) {
    // deterministic::vanilla_to_expiry (engines/deterministic.rs:54-80) for a
    // flat curve: the discounted intrinsic value of the forward.
    double fwd = m.spot * std::exp(-m.dividend * time) - o.strike * std::exp(-m.rate * time);
    *known = std::max(sign * fwd, 0.0);
    return true;
  }
  double diffusion = sigma * std::sqrt(time);
  double carry = (m.rate - m.dividend) * time;
  double half_variance = (0.5 * diffusion) * diffusion;
  double drift = carry - half_variance;
  if (!std::isfinite(drift) || !std::isfinite(diffusion)) unstable();
  double moneyness = log_ratio(m.spot, o.strike) + carry - half_variance;
  double log_scale =
      sign > 0.0 ? std::log(m.spot) - m.dividend * time : std::log(o.strike) - m.rate * time;
  *pay = TerminalPayoff{sign, diffusion, half_variance, moneyness, log_scale};
  return false;
}

// MonteCarloVanillaEngine::sample + price_with_statistics (monte_carlo.rs:1093-1141, 1307-1320).
inline Estimate european_price(const Settings& st, const kg_european_terms& o,
                               const kg_bsm_market& m, const Context& cx) {
  TerminalPayoff pay{};
  double known = 0.0;
  Moments total[1], pairs[1];
  if (european_plan(o, m, &pay, &known)) {
    Moments s;
    s.add(known, false);
    return estimate(st, s, Moments(), 0.0, true, false, 0);
  }
  double log_scale = pay.log_scale;
  if (pay.needs_peak_scaling()) {
    path_samples<1>(
        st, 1, false, cx,
        [&](const double* n, double* out) {
          double lv = pay.log_payoff(n[0]);
          if (lv > std::numeric_limits<double>::max() || std::isnan(lv)) unstable();
          out[0] = std::max(lv + 746.0, 0.0);
        },
        total, pairs);
    log_scale = total[0].scale - 746.0;
    path_samples<1>(
        st, 1, true, cx,
        [&](const double* n, double* out) { out[0] = std::exp(pay.log_payoff(n[0]) - log_scale); },
        total, pairs);
  } else {
    path_samples<1>(
        st, 1, true, cx, [&](const double* n, double* out) { out[0] = pay.at_normal(n[0]); },
        total, pairs);
  }
  bool is_known = m.valuation_time == o.expiry_time;
  return estimate(st, total[0], pairs[0], log_scale, is_known, true, st.seed);
}

// ---------------------------------------------------------------- autocallable
struct CashEvent {  // engines/monte_carlo/autocallable.rs:524-562
  bool has_knock_out;
  double knock_out, coupon_barrier, period_coupon, redemption;
  double discounted_payment(bool coupon_due, bool knocked_out, double discount) const {
    double coupon = coupon_due ? period_coupon : 0.0;
    double amount = knocked_out ? redemption + coupon : coupon;
    return scaled_value(1.0, amount, discount);
  }
};
struct PathStep {
  double drift, diffusion, discount;
  int event;  // index into events or -1
  bool observe_knock_in;
};
struct BridgeInterval {
  size_t end;
  double drift, diffusion;
};

struct Note {
  const kg_autocall_terms* t;
  double maturity_coupon_rate() const {
    return t->coupon_kind == KG_COUPON_SNOWBALL ? t->coupon : 0.0;
  }
  double period_coupon() const { return t->coupon_kind == KG_COUPON_PHOENIX ? t->coupon : 0.0; }
  bool has_ki() const { return t->knock_in_kind != KG_KNOCK_IN_NONE; }
  bool daily_knock_in() const { return t->knock_in_kind == KG_KNOCK_IN_DAILY; }
  double rebate_payment(int64_t time) const {  // instruments/autocallable.rs:1415-1418
    return t->rebate_rate + t->annual_rebate_rate * years_until(t->effective_time, time);
  }
  double knock_out_payment(const kg_autocall_event& e) const {  // :1419-1424
    return t->principal_redemption_ratio + rebate_payment(e.time) +
           e.coupon_rate * years_until(t->effective_time, e.time);
  }
  double terminal(double spot, bool knocked_in) const {  // :1396-1414
    double coupon;
    if (knocked_in && has_ki()) {
      if (t->redemption_kind == KG_REDEMPTION_DOWNSIDE) {
        double c = std::min(std::max(spot, t->lower_strike), t->upper_strike);
        if (std::isnan(spot)) c = spot;
        return t->principal_redemption_ratio + rebate_payment(t->expiry_time) +
               (c - t->upper_strike) / t->initial_spot;
      }
      coupon = t->fixed_coupon_rate;
    } else if (knocked_in) {
      coupon = maturity_coupon_rate();
    } else {
      coupon = maturity_coupon_rate();
    }
    return t->principal_redemption_ratio + rebate_payment(t->expiry_time) +
           coupon * years_until(t->effective_time, t->expiry_time);
  }
  bool knock_in_affects_payoff() const {  // :1307-1316
    if (!has_ki()) return false;
    if (t->redemption_kind == KG_REDEMPTION_FIXED_COUPON)
      return t->fixed_coupon_rate != maturity_coupon_rate();
    return t->lower_strike != t->upper_strike || maturity_coupon_rate() != 0.0;
  }
  bool daily_time(int64_t time) const {  // :1389-1394
    return daily_knock_in() && time >= t->effective_time && time <= t->expiry_time &&
           clock_of(time) == t->observation_clock_micros &&
           is_trading_day(t->calendar, day_of(time));
  }
  bool observes_knock_in(int64_t time) const {  // :1075-1080
    if (t->knock_in_kind == KG_KNOCK_IN_DAILY) return daily_time(time);
    if (t->knock_in_kind == KG_KNOCK_IN_AT_EXPIRY) return time == t->expiry_time;
    return false;
  }
  const kg_autocall_event* event_at(int64_t time) const {
    for (int i = 0; i < t->n_events; ++i)
      if (t->events[i].time == time) return &t->events[i];
    return nullptr;
  }
};

struct AutocallPlan {
  bool fixed = false;
  double fixed_value = 0.0;
  bool knocked_in = false;
  double coupon = 0.0;
  double initial_spot = 0.0;
  double safe_lo = 0.0, safe_hi = 0.0;
  bool has_ki = false;
  double knock_in = 0.0;
  bool expiry_knock_in = false;
  std::vector<PathStep> steps;
  std::vector<CashEvent> events;
  bool bridge = false;
  std::vector<BridgeInterval> intervals;
  std::vector<double> weight, fill;  // BridgePlan::daily
  const Note* note = nullptr;

  void validate_return(double lr) const {  // autocallable.rs:606-613
    if (!(lr >= safe_lo && lr <= safe_hi) && scaled_value(initial_spot, 1.0, lr) <= 0.0)
      unstable();
  }
  double discounted_terminal(double lr, bool ki, double discount) const {
    double spot = scaled_value(initial_spot, 1.0, lr);
    return scaled_value(1.0, note->terminal(spot, ki), discount);
  }
};

// DailyGrids::prepare (engines/monte_carlo/autocallable/portfolio.rs:75-186):
// whether price_many_on_shared_grid places the note on the shared daily grid
// of its expiry, calendar and clock. The grid holds the daily observation
// times after valuation, plus expiry when it is off the grid; it must fit
// the reference's 2 MiB slot (DAILY_GRID_BYTES; 72-byte PathStep, 8-byte
// Timestamp, the time vector grown from capacity four by doubling, before
// and after the expiry push) and contain every future event.
inline bool shared_daily_grid(const Note& note, int64_t valuation) {
  const kg_autocall_terms& t = *note.t;
  const uint64_t grid_bytes = 2u * 1024u * 1024u, step_bytes = 72, time_bytes = 8;
  const int64_t days = day_of(t.expiry_time) - day_of(valuation) + 1;
  if ((uint64_t)days >= grid_bytes / (time_bytes + step_bytes)) return false;
  for (int i = 0; i < t.n_events; ++i) {
    const int64_t time = t.events[i].time;
    if (time <= valuation) continue;
    if (time != t.expiry_time && !note.daily_time(time)) return false;
  }
  std::vector<int64_t> dates;
  for (int64_t d = day_of(valuation); d <= day_of(t.expiry_time); ++d) {
    const int64_t at = d * MICROS_PER_DAY + t.observation_clock_micros;
    if (at > valuation && at <= t.expiry_time && is_trading_day(t.calendar, d)) dates.push_back(at);
  }
  if (dates.empty()) return false;
  uint64_t capacity = 4;
  while (capacity < dates.size()) capacity *= 2;
  if (dates.back() != t.expiry_time) {
    dates.push_back(t.expiry_time);
    if (dates.size() > capacity) capacity *= 2;
  }
  return capacity * time_bytes + dates.size() * step_bytes <= grid_bytes;
}

// The plan of a note; with `shared_grid`, as price_many_on_shared_grid plans
// it: a note on a shared daily grid keeps its dates, transitions, events and
// knock-in fixings but is sampled step by step, without the Brownian bridge
// (autocallable.rs:341-374).
inline AutocallPlan autocall_plan(const Note& note, const kg_bsm_market& m, bool shared_grid = false) {
  const kg_autocall_terms& t = *note.t;
  AutocallPlan p;
  p.note = &note;
  if (m.volatility == 0.0)
    throw Failure(UNSUPPORTED, "zero volatility uses the deterministic engine");
  int64_t valuation = m.valuation_time;
  if (valuation < t.effective_time || valuation > t.expiry_time)
    throw Failure(INVALID_INPUT, "valuation outside contract life");
  // prior_with_work_check (instruments/autocallable.rs:1321-1370), the parts
  // that do not depend on custom calendars.
  int64_t first_out = INT64_MAX;
  for (int i = 0; i < t.n_events; ++i)
    if (t.events[i].has_knock_out) {
      first_out = t.events[i].time;
      break;
    }
  bool had_out = first_out < valuation;
  bool had_in = false;
  if (note.daily_knock_in() && valuation > t.effective_time) {
    int64_t end = std::min(valuation - 1, t.expiry_time);
    for (int64_t d = day_of(t.effective_time); d <= day_of(end) && !had_in; ++d) {
      int64_t at = d * MICROS_PER_DAY + t.observation_clock_micros;
      if (at >= t.effective_time && at <= end && is_trading_day(t.calendar, d)) had_in = true;
    }
  }
  if (t.barrier_state == KG_STATE_UNKNOWN && (had_out || had_in))
    throw Failure(6, "prior autocallable state is required");
  // pricing::initial (instruments/autocallable/pricing.rs:12-46)
  if (t.barrier_state == KG_STATE_KNOCKED_OUT) {
    p.fixed = true;
    p.fixed_value = 0.0;
    return p;
  }
  bool knocked_in = t.barrier_state == KG_STATE_KNOCKED_IN;
  if (note.daily_knock_in() && note.daily_time(valuation))
    knocked_in |= note.has_ki() && m.spot < t.knock_in_level;
  const kg_autocall_event* current = note.event_at(valuation);
  double coupon = 0.0;
  if (current && current->has_coupon_barrier && m.spot >= current->coupon_barrier_level)
    coupon = note.period_coupon();
  if (valuation == t.expiry_time) {
    // expiry_payment (pricing.rs:48-68)
    double v;
    if (current && current->has_knock_out && m.spot >= current->knock_out_level)
      v = note.knock_out_payment(*current);
    else
      v = note.terminal(m.spot, knocked_in || (note.observes_knock_in(t.expiry_time) &&
                                               note.has_ki() && m.spot < t.knock_in_level));
    p.fixed = true;
    p.fixed_value = coupon + v;
    return p;
  }
  if (current && current->has_knock_out && m.spot >= current->knock_out_level) {
    p.fixed = true;
    p.fixed_value = note.knock_out_payment(*current) + coupon;
    return p;
  }
  p.knocked_in = knocked_in;
  p.coupon = coupon;
  double r = m.rate, q = m.dividend, sigma = m.volatility;
  bool daily = note.daily_knock_in() && !knocked_in && note.knock_in_affects_payoff();
  p.initial_spot = m.spot;
  double log_spot = std::log(m.spot);
  double low = std::log(std::numeric_limits<double>::min()) + 1.0;
  double high = std::log(std::numeric_limits<double>::max()) - 1.0;
  p.safe_lo = std::max(low, low - log_spot);
  p.safe_hi = std::min(high, high - log_spot);
  p.has_ki = note.has_ki();
  if (p.has_ki) p.knock_in = log_ratio(t.knock_in_level, m.spot);
  p.expiry_knock_in = note.observes_knock_in(t.expiry_time);
  // future_event_times (pricing.rs:70-106)
  std::vector<int64_t> dates;
  for (int i = 0; i < t.n_events; ++i)
    if (t.events[i].time > valuation && t.events[i].time <= t.expiry_time)
      dates.push_back(t.events[i].time);
  if (daily) {
    for (int64_t d = day_of(valuation); d <= day_of(t.expiry_time); ++d) {
      int64_t at = d * MICROS_PER_DAY + t.observation_clock_micros;
      if (at > valuation && at <= t.expiry_time && is_trading_day(t.calendar, d))
        dates.push_back(at);
    }
  }
  dates.push_back(t.expiry_time);
  std::sort(dates.begin(), dates.end());
  dates.erase(std::unique(dates.begin(), dates.end()), dates.end());
  int64_t previous = valuation;
  int next = 0;
  while (next < t.n_events && t.events[next].time <= valuation) ++next;
  double excursion = 0.0;
  for (int64_t date : dates) {
    PathStep s{};
    transition(r, q, sigma, years_until(previous, date), years_until(valuation, date), &s.drift,
               &s.diffusion, &s.discount);
    excursion += std::fabs(s.drift) + 20.0 * s.diffusion;
    s.event = -1;
    if (next < t.n_events && t.events[next].time == date) {
      const kg_autocall_event& e = t.events[next];
      CashEvent c{};
      c.has_knock_out = e.has_knock_out != 0;
      c.knock_out = c.has_knock_out ? log_ratio(e.knock_out_level, m.spot) : 0.0;
      c.coupon_barrier =
          log_ratio(e.has_coupon_barrier ? e.coupon_barrier_level : 0.0, m.spot);
      c.period_coupon = e.has_coupon_barrier ? note.period_coupon() : 0.0;
      c.redemption = c.has_knock_out ? note.knock_out_payment(e) : 0.0;
      s.event = (int)p.events.size();
      p.events.push_back(c);
      ++next;
    }
    s.observe_knock_in = daily && note.daily_time(date);
    p.steps.push_back(s);
    previous = date;
  }
  // BridgePlan::new (autocallable.rs:463-515)
  bool safe = excursion >= p.safe_lo && excursion <= p.safe_hi && -excursion >= p.safe_lo &&
              -excursion <= p.safe_hi;
  if (daily && safe && !(shared_grid && shared_daily_grid(note, valuation))) {
    std::vector<BridgeInterval> iv;
    size_t first = 0;
    bool has_gap = false;
    for (size_t end = 0; end < p.steps.size(); ++end) {
      if (p.steps[end].event >= 0 || end + 1 == p.steps.size()) {
        has_gap |= end - first >= 8;
        int64_t start = first == 0 ? valuation : dates[first - 1];
        BridgeInterval b{end, 0, 0};
        double unused;
        transition(r, q, sigma, years_until(start, dates[end]), 0.0, &b.drift, &b.diffusion,
                   &unused);
        iv.push_back(b);
        first = end + 1;
      }
    }
    if (has_gap) {
      std::vector<double> w(p.steps.size(), 0.0), f(p.steps.size(), 0.0);
      bool ok = true;
      size_t first2 = 0;
      int64_t prev = valuation;
      for (auto& b : iv) {
        int64_t end_t = dates[b.end];
        for (size_t i = first2; i < b.end && ok; ++i) {
          double dt = years_until(prev, dates[i]);
          double fraction = dt / years_until(prev, end_t);
          if (!(0.0 < fraction && fraction < 1.0)) {
            ok = false;
            break;
          }
          w[i] = fraction;
          f[i] = sigma * std::sqrt(dt * (1.0 - fraction));
          prev = dates[i];
        }
        first2 = b.end + 1;
        prev = end_t;
      }
      if (ok) {
        p.bridge = true;
        p.intervals = iv;
        p.weight = w;
        p.fill = f;
      }
    }
  }
  return p;
}

// PathPlan::payoff (autocallable.rs:823-862), one lane.
template <class Path>
double autocall_payoff(const AutocallPlan& p, Path& normals) {
  double lr = 0.0;
  bool ki = p.knocked_in;
  double cash = p.coupon;
  double discount = 0.0;
  for (const PathStep& s : p.steps) {
    discount = s.discount;
    lr += std::fma(s.diffusion, normals.normal(), s.drift);
    p.validate_return(lr);
    if (s.observe_knock_in) ki |= p.has_ki && lr < p.knock_in;
    if (s.event >= 0) {
      const CashEvent& e = p.events[s.event];
      bool due = lr >= e.coupon_barrier;
      bool out = e.has_knock_out && lr >= e.knock_out;
      cash += e.discounted_payment(due, out, discount);
      if (out) {
        if (!std::isfinite(cash)) unstable();
        return cash;
      }
    }
  }
  ki |= p.expiry_knock_in && p.has_ki && lr < p.knock_in;
  cash += p.discounted_terminal(lr, ki, discount);
  if (!std::isfinite(cash)) unstable();
  return cash;
}

// PathPlan::bridge_payoffs::<1, true> (autocallable.rs:628-736).
inline double autocall_bridge_payoff(const AutocallPlan& p, SkippingPath& normals) {
  double cash = p.coupon;
  bool monitoring = !p.knocked_in && p.has_ki;
  double level = p.knock_in;
  bool ki = p.knocked_in;
  bool active = true;
  size_t start = 0;
  double lr = 0.0, discount = 0.0;
  for (const BridgeInterval& iv : p.intervals) {
    double z = normals.normal();
    double target = lr + std::fma(iv.diffusion, z, iv.drift);
    bool floor = monitoring;
    double minimum = std::numeric_limits<double>::infinity();
    for (size_t i = start; i < iv.end; ++i) {
      if (!floor || minimum < level) {
        normals.skip(iv.end - i);
        break;
      }
      double zi = normals.normal();
      lr += std::fma(p.weight[i], target - lr, p.fill[i] * zi);
      minimum = std::fmin(minimum, lr);
    }
    lr = target;
    start = iv.end + 1;
    if (active) {
      const PathStep& s = p.steps[iv.end];
      discount = s.discount;
      if (monitoring && (minimum < level || (s.observe_knock_in && lr < level))) {
        monitoring = false;
        ki = true;
      }
      if (s.event >= 0) {
        const CashEvent& e = p.events[s.event];
        bool hit = e.has_knock_out && lr >= e.knock_out;
        cash += e.discounted_payment(lr >= e.coupon_barrier, hit, discount);
        if (hit) {
          active = false;
          monitoring = false;
        }
      }
    }
    if (!active) {
      if (!std::isfinite(cash)) unstable();
      return cash;
    }
  }
  bool hit = ki || (p.expiry_knock_in && p.has_ki && lr < p.knock_in);
  cash += p.discounted_terminal(lr, hit, discount);
  if (!std::isfinite(cash)) unstable();
  return cash;
}

// MonteCarloAutocallableEngine::sample + price_with_statistics (autocallable.rs:154-174, 911-918).
inline Estimate autocall_price(const Settings& st, const kg_autocall_terms& t,
                               const kg_bsm_market& m, const Context& cx, bool shared_grid = false) {
  Note note{&t};
  AutocallPlan p = autocall_plan(note, m, shared_grid);
  if (p.fixed) {
    if (!std::isfinite(p.fixed_value)) unstable();
    Moments s;
    s.add(p.fixed_value, false);
    return estimate(st, s, s, 0.0, true, false, 0);
  }
  Moments total[1], pairs[1];
  size_t dims = p.steps.size();
  if (p.bridge) {
    path_samples_skipping<1>(
        st, dims, true, cx,
        [&](SkippingPath& path, double* out) { out[0] = autocall_bridge_payoff(p, path); }, total,
        pairs);
  } else {
    path_samples_lazy<1>(
        st, dims, true, cx, [&](LazyPath& path, double* out) { out[0] = autocall_payoff(p, path); },
        total, pairs);
  }
  return estimate(st, total[0], pairs[0], 0.0, false, true, st.seed);
}

}  // namespace oracle
