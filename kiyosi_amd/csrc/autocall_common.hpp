// Shared autocallable code: the device plan, the path loops and the host
// plan builder, used by the single-note engine (autocall.hip) and the book
// kernel (book.hip). See autocall.hip for the reference mapping.
#pragma once

#include <algorithm>
#include <type_traits>

#include "fp32_math.hpp"
#include "host_common.hpp"
#include "sobol_device.hpp"

namespace kg {


constexpr int64_t MICROS_PER_DAY = 86400LL * 1000000LL;

// ------------------------------------------------------------------ plan
struct AcStepD {  // per step, FP64
  double drift, diffusion;
};
struct AcEventD {
  double knock_out;       // log level, or +inf without knock-out
  double coupon_barrier;  // log level
  double pay_coupon;      // period coupon * discount factor
  double pay_out;         // redemption * discount factor
  double pay_both;        // (redemption + period coupon) * discount factor
  double pay_coupon_raw, pay_out_raw, pay_both_raw;  // undiscounted amounts
  double log_discount;
  int32_t has_knock_out;
  int32_t pad;
};
struct AcIntervalD {
  int32_t end;
  int32_t pad;
  double drift, diffusion;
};

struct AcPlanDev {
  int32_t n_steps;
  int32_t n_intervals;
  int32_t bridge;
  int32_t knocked_in;
  int32_t has_ki;
  int32_t expiry_knock_in;
  double knock_in;  // log level
  double coupon;    // current coupon from pricing::initial
  double initial_spot;
  double safe_lo, safe_hi;
  double final_log_discount;
  // terminal(spot, ki): principal + rebate + coupon*years or the downside
  double term_base;       // principal + rebate(expiry)
  double term_no_ki;      // coupon * years when not knocked in
  double term_ki_fixed;   // coupon * years when knocked in with a fixed coupon
  double lower, upper, inv_initial;
  int32_t downside;       // knock-in redemption is the downside
  int32_t pad;
  const AcStepD* steps;
  const int32_t* step_event;    // event index or -1
  const uint8_t* step_observe;  // observe_knock_in
  const AcEventD* events;
  const AcIntervalD* intervals;
  const double* weight;  // bridge fill weights per step
  const double* fill;    // bridge fill diffusions per step
  // FP32 copies of the per-step tables
  const float* steps32;     // drift, diffusion pairs
  const float* wf32;        // bridge fill 1 - weight, -diffusion pairs (one scalar load per step)
  const float* events32;    // knock_out, coupon_barrier pairs
  const float* intervals32; // drift, diffusion pairs
};

template <typename R>
struct Tab;
template <>
struct Tab<double> {
  static __device__ inline double drift(const AcPlanDev& p, int i) { return uload(&p.steps[i].drift); }
  static __device__ inline double diff(const AcPlanDev& p, int i) { return uload(&p.steps[i].diffusion); }
  static __device__ inline double weight(const AcPlanDev& p, int i) { return uload(p.weight + i); }
  static __device__ inline double f(const AcPlanDev& p, int i) { return uload(p.fill + i); }
  static __device__ inline double ko(const AcPlanDev& p, int e) { return uload(&p.events[e].knock_out); }
  static __device__ inline double cb(const AcPlanDev& p, int e) { return uload(&p.events[e].coupon_barrier); }
  static __device__ inline double idrift(const AcPlanDev& p, int k) { return uload(&p.intervals[k].drift); }
  static __device__ inline double idiff(const AcPlanDev& p, int k) { return uload(&p.intervals[k].diffusion); }
};
template <>
struct Tab<float> {
  static __device__ inline float drift(const AcPlanDev& p, int i) { return uload(p.steps32 + 2 * i); }
  static __device__ inline float diff(const AcPlanDev& p, int i) { return uload(p.steps32 + 2 * i + 1); }
  static __device__ inline float weight(const AcPlanDev& p, int i) { return uload(p.wf32 + 2 * i); }
  static __device__ inline float f(const AcPlanDev& p, int i) { return uload(p.wf32 + 2 * i + 1); }
  static __device__ inline float ko(const AcPlanDev& p, int e) { return uload(p.events32 + 2 * e); }
  static __device__ inline float cb(const AcPlanDev& p, int e) { return uload(p.events32 + 2 * e + 1); }
  static __device__ inline float idrift(const AcPlanDev& p, int k) { return p.intervals32[2 * k]; }
  static __device__ inline float idiff(const AcPlanDev& p, int k) { return p.intervals32[2 * k + 1]; }
};

// discounted_payment (autocallable.rs:552-561) with the discount factor
// folded in on the host; the log-space form covers non-normal products.
__device__ inline double event_payment(const AcEventD& e, bool due, bool out) {
  double v = out ? (due ? e.pay_both : e.pay_out) : (due ? e.pay_coupon : 0.0);
  if (v == 0.0 || isnormal(v)) return v;
  double raw = out ? (due ? e.pay_both_raw : e.pay_out_raw) : (due ? e.pay_coupon_raw : 0.0);
  return discounted_product(1.0, raw, e.log_discount);
}

// discounted_terminal (autocallable.rs:798-801) + AutocallableNote::terminal
// (instruments/autocallable.rs:1396-1414).
template <class Plan>
__device__ inline double terminal_value(const Plan& p, double lr, bool ki, uint32_t& bad) {
#pragma clang fp contract(off)
  double spot = discounted_product(p.initial_spot, 1.0, lr);
  if (!isfinite(spot)) bad = 1;
  double amount;
  if (ki && p.downside) {
    double c = fmin(fmax(spot, p.lower), p.upper);
    amount = p.term_base + (c - p.upper) * p.inv_initial;
  } else {
    amount = p.term_base + (ki ? p.term_ki_fixed : p.term_no_ki);
  }
  return discounted_product(1.0, amount, p.final_log_discount);
}

// One interval of daily bridge fill steps lr <- lr + w (target - lr) + f z
// and the knock-in test `minimum < level` over them. FP64 rounds as the
// reference writes it (lr += w.mul_add(target - lr, f * z),
// autocallable.rs:678) and tracks the minimum. FP32 steps the gap
// g = target - lr instead, g <- (1 - w) g - f z: two roundings and two
// instructions a step, and lr < level becomes g > target - level. The
// tables hold the coefficients each form takes: w and f for FP64,
// 1 - w and -f (each rounded once from FP64) for FP32. Every kernel that fills
// a bridge uses these, so lane and single-note prices agree bit for bit.
template <typename R>
struct FillTrack;
template <>
struct FillTrack<double> {
  double lr, lowest, level;
  __device__ inline void start(double lr0, double target, double lvl) {
    lr = lr0;
    lowest = INFINITY;
    level = lvl;
  }
  __device__ inline void step(double target, double w, double f, double z) {
    lr = lr + fma(w, target - lr, f * z);
    lowest = fmin(lowest, lr);
  }
  __device__ inline bool below() const { return lowest < level; }
};
template <>
struct FillTrack<float> {
  float gap, limit;
  bool under;
  __device__ inline void start(float lr0, float target, float lvl) {
    gap = target - lr0;
    limit = target - lvl;
    under = false;
  }
  __device__ inline void step(float, float keep, float neg_f, float z) {
    gap = fmaf(keep, gap, neg_f * z);
    under = under || gap > limit;
  }
  __device__ inline bool below() const { return under; }
};

// Source of the normals of one sample. DIRECT fixes at compile time where
// Sobol points come from: global directions (1), the block's LDS tables (0),
// or block-independent LDS tables XOR the sample block's base point, a row of
// global memory (2; the tables' `hi` then leaves the base out); -1 leaves
// the choice between 0 and 1 to the `direct` field.
template <bool SOBOL, bool FP32, int DIRECT = -1>
struct Draws {
  // Sobol: LDS tables and the lane's Gray-code digits
  const void* lo;
  const void* hi;
  const void* base;  // DIRECT 2: the sample block's base point, one word per dimension
  uint32_t glo, ghi;
  int dims;
  // Sobol fallback for long paths: point index and global directions
  const uint64_t* dirs;
  const uint64_t* shifts;
  uint64_t point;
  bool direct;
  // PCG. The FP32 normals of a lane queue up to three deep, so that whenever
  // some lane needs a new Box-Muller pair every lane draws its next pair on
  // the same call: lanes whose stream starts on an odd normal then stay in
  // step with the others instead of diverging on every call. The values and
  // their order are those of the lane's stream.
  NormalStream s;
  Pcg32 g32;
  float q32[3];
  int avail32;

  // Positions the FP32 stream of sample stream `block` at normal `first`.
  __device__ inline void start32(uint64_t seed, uint64_t block, uint64_t first) {
    g32.init(seed, block);
    g32.advance((first >> 1) * 4);
    avail32 = 0;
    if (first & 1) {
      float e;
      f32::box_muller_pair(g32, &e, &q32[0]);
      avail32 = 1;
    }
  }

  __device__ inline double normal64(int d) {
    if (SOBOL) {
      uint64_t bits = is_direct() ? sobol_point_direct(dirs, shifts, dims, point, d)
                                  : sobol_lds_bits64((const uint64_t*)lo, (const uint64_t*)hi, d, glo, ghi);
      if constexpr (DIRECT == 2) bits ^= ((const uint64_t*)base)[d];
      return normal_from_bits(bits);
    }
    return s.next();
  }
  __device__ inline bool is_direct() const { return DIRECT < 0 ? direct : DIRECT == 1; }
  // Top 32 of the 52 bits of Sobol dimension d of this sample's point.
  __device__ inline uint32_t bits32(int d) {
    if constexpr (DIRECT == 2)
      return sobol_lds_bits32((const uint32_t*)lo, (const uint32_t*)hi, d, glo, ghi) ^ ((const uint32_t*)base)[d];
    return is_direct() ? (uint32_t)(sobol_point_direct(dirs, shifts, dims, point, d) >> 20)
                       : sobol_lds_bits32((const uint32_t*)lo, (const uint32_t*)hi, d, glo, ghi);
  }
  __device__ inline float normal32(int d) {
    if (SOBOL) return f32::normal_from_bits(bits32(d));
    // Lanes consume in step, so at a refill every lane holds zero or one.
    if (__ballot(avail32 == 0) != 0 && avail32 <= 1) {
      float e, o;
      f32::box_muller_pair(g32, &e, &o);
      if (avail32 == 0) {
        q32[0] = e;
        q32[1] = o;
      } else {
        q32[1] = e;
        q32[2] = o;
      }
      avail32 += 2;
    }
    const float v = q32[0];
    q32[0] = q32[1];
    q32[1] = q32[2];
    --avail32;
    return v;
  }
  template <typename R>
  __device__ inline R normal(int d) {
    if constexpr (FP32)
      return normal32(d);
    else
      return normal64(d);
  }
  // The lane's next two normals (dimensions d and d + 1). PCG FP32: every
  // lane draws one Box-Muller pair on every call, without a branch, and
  // merges it with the zero, one or two normals it holds queued, so the
  // queue keeps its length and the values come in stream order.
  template <typename R>
  __device__ inline void normal_pair(int d, R& a, R& b) {
    if constexpr (FP32 && !SOBOL) {
      float e, o;
      f32::box_muller_pair(g32, &e, &o);
      const bool none = avail32 == 0, one = avail32 == 1, two = avail32 == 2;
      a = none ? e : q32[0];
      b = none ? o : (one ? e : q32[1]);
      q32[0] = one ? o : (two ? e : q32[0]);
      q32[1] = two ? o : q32[1];
    } else {
      a = normal<R>(d);
      b = normal<R>(d + 1);
    }
  }
  __device__ inline void skip(int count) {
    if (SOBOL || count <= 0) return;
    if (FP32) {
      for (; count > 0 && avail32 > 0; --count) {
        q32[0] = q32[1];
        q32[1] = q32[2];
        --avail32;
      }
      g32.advance((uint64_t)(count >> 1) * 4);
      if (count & 1) {  // draw the pair and keep its second normal
        float e;
        f32::box_muller_pair(g32, &e, &q32[0]);
        avail32 = 1;
      }
    } else {
      s.skip((uint64_t)count);
    }
  }
};

// Target of bridge interval k (the interval's end point), drawn from the
// interval's first dimension (bridge_payoffs::<1, true>,
// autocallable.rs:640-660); `draw` is that dimension's index in `dr`.
template <bool FP32, class D>
__device__ inline typename std::conditional<FP32, float, double>::type interval_target(
    const AcPlanDev& p, D& dr, int k, int draw, typename std::conditional<FP32, float, double>::type lr,
    float sign) {
  using R = typename std::conditional<FP32, float, double>::type;
  const R z = (R)sign * dr.template normal<R>(draw);
  return lr + (FP32 ? fmaf((float)Tab<R>::idiff(p, k), (float)z, (float)Tab<R>::idrift(p, k))
                    : fma((double)Tab<R>::idiff(p, k), (double)z, (double)Tab<R>::idrift(p, k)));
}

// The daily fill of one bridge interval, steps start..end-1 drawing
// dimensions start+1..end (autocallable.rs:662-690); returns whether a
// filling lane went under the knock-in level (the reference's
// `minimum < level`). Lanes that are not filling step along and ignore the
// result.
template <bool SOBOL, bool FP32, class D>
__device__ inline bool interval_fill(const AcPlanDev& p, D& dr, int start, int end,
                                     typename std::conditional<FP32, float, double>::type lr,
                                     typename std::conditional<FP32, float, double>::type target,
                                     typename std::conditional<FP32, float, double>::type level,
                                     bool filling, float sign) {
  using R = typename std::conditional<FP32, float, double>::type;
  FillTrack<R> track;
  track.start(lr, target, level);
  if constexpr (SOBOL) {
    // Sobol draws are stateless, so every lane steps the whole interval and
    // nothing in the loop branches: a lane that went under the level (where
    // the reference stops filling) stays below, and the lanes still filling
    // compute exactly what the reference computes.
    if (!wave_any(filling)) return false;
    if constexpr (FP32) {
      // The next step's point bits (LDS) and fill coefficients (scalar load)
      // are read one step ahead, so both overlap the current step's
      // arithmetic. The last coefficient read is step `end`, which exists
      // (`end` is a step of the path); the last point read is dimension
      // end + 1 <= dims, which the LDS tables' slack holds and the direct
      // form clamps to `end`. The loop runs in pairs of steps so that
      // consecutive LDS reads share one address and differ by an immediate
      // offset (the compiler does not unroll a loop holding wave votes).
      const float* wf = p.wf32 + 2 * start;
      uint32_t next = dr.bits32(start + 1);
      float next_w = wf[0], next_f = wf[1];
      auto fill_one = [&](int i) {
        const uint32_t bits = next;
        const float w = next_w, f = next_f;
        next = dr.bits32(dr.is_direct() && i + 2 > end ? end : i + 2);
        wf += 2;
        next_w = wf[0];
        next_f = wf[1];
        track.step(target, w, f, f32::normal_from_bits(bits, sign));
      };
      int i = start;
      for (; i + 1 < end; i += 2) {
        fill_one(i);
        fill_one(i + 1);
      }
      if (i < end) fill_one(i);
    } else {
      for (int i = start; i < end; ++i) {
        const R z = (R)sign * dr.template normal<R>(i + 1);
        track.step(target, Tab<double>::weight(p, i), Tab<double>::f(p, i), z);
      }
    }
  } else {
    for (int i = start; i < end; ++i) {
      if (__ballot(filling) == 0) break;
      if (filling && track.below()) {
        dr.skip(end - i);
        filling = false;
      }
      if (filling) {
        const R z = (R)sign * dr.template normal<R>(i + 1);
        track.step(target, Tab<R>::weight(p, i), Tab<R>::f(p, i), z);
      }
    }
  }
  return track.below();
}

// Bridge path (bridge_payoffs::<1, true>, autocallable.rs:628-736). `sign`
// is -1 for the antithetic leg. Returns the discounted cash of the path.
template <bool SOBOL, bool FP32, class D>
__device__ double bridge_path(const AcPlanDev& p, D& dr, bool live, float sign,
                              uint32_t& bad) {
  using R = typename std::conditional<FP32, float, double>::type;
  double cash = p.coupon;
  bool monitoring = !p.knocked_in && p.has_ki;
  const R level = (R)p.knock_in;
  bool ki = p.knocked_in;
  bool active = live;
  R lr = 0;
  int start = 0;
  for (int k = 0; k < p.n_intervals; ++k) {
    if (__ballot(active) == 0) break;
    const int end = uload(&p.intervals[k].end);
    R target = lr;
    bool filling = false;
    if (active) {
      target = interval_target<FP32>(p, dr, k, start, lr, sign);
      filling = monitoring;
      if (!filling) dr.skip(end - start);
    }
    const bool below = interval_fill<SOBOL, FP32>(p, dr, start, end, lr, target, level, filling, sign);
    lr = target;
    start = end + 1;
    if (active) {
      if (monitoring && (below || (uload_u8(p.step_observe, end) && lr < level))) {
        monitoring = false;
        ki = true;
      }
      const int e = uload(p.step_event + end);
      if (e >= 0) {
        const AcEventD& ev = p.events[e];
        const bool hit = ev.has_knock_out && lr >= Tab<R>::ko(p, e);
        cash += event_payment(ev, lr >= Tab<R>::cb(p, e), hit);
        if (hit) {
          active = false;
          monitoring = false;
        }
      }
    }
  }
  if (active) {
    const bool hit = ki || (p.expiry_knock_in && p.has_ki && lr < level);
    cash += terminal_value(p, (double)lr, hit, bad);
  }
  if (live && !isfinite(cash)) bad = 1;
  return cash;
}

// The same bridge path in two passes, for kernels that gather the paths
// still waiting on their knock-in. A path's daily fill decides only whether
// it knocks in, and knocking in changes only the payoff at expiry (event
// coupons and knock-outs read the interval end points alone), so the fill of
// a path that knocks out, or that is already knocked in, changes nothing.
// Pass one (bridge_events) walks the interval end points, events and
// observation-date knock-ins; pass two (bridge_knock_in) fills the paths
// that reach expiry still monitoring, until they knock in; the expiry payoff
// is added last, as bridge_path adds it. The sum is bridge_path's, bit for
// bit.
template <typename R>
struct BridgeEvents {
  double cash;     // coupon and event payments, expiry payoff excluded
  R lr;            // last end point drawn
  bool active;     // alive at expiry
  bool ki;         // knocked in without the daily fill
  bool need_fill;  // the expiry payoff waits on the daily fill
};

// What pass one reads of interval k (built on the host from the plan's
// tables and gathered per block in LDS, so that the interval loop waits on
// no global load): the end point's drift and diffusion (the plan's interval
// row), its observation flag and event (knock-out and coupon levels as the
// plan's precision holds them, discounted payments; the rare non-normal
// payments fall back to the event in global memory).
template <typename R>
struct IntervalRow {
  double pay_coupon, pay_out, pay_both;
  R drift, diffusion, knock_out, coupon_barrier;
  int32_t first, end, event, observe, has_knock_out, pad;  // first: the interval's first dimension
};

// Pass one over the rows of `rows` (one IntervalRow per interval). ROWS:
// the end point of interval k is dimension k of `dr` (tables that hold the
// intervals' first dimensions only); otherwise dimension `start` as in
// bridge_path. The arithmetic is bridge_path's, step for step.
template <bool SOBOL, bool FP32, bool ROWS, class D>
__device__ BridgeEvents<typename std::conditional<FP32, float, double>::type> bridge_events(
    const AcPlanDev& p, const IntervalRow<typename std::conditional<FP32, float, double>::type>* rows, D& dr,
    bool live, float sign) {
  using R = typename std::conditional<FP32, float, double>::type;
  BridgeEvents<R> out;
  double cash = p.coupon;
  bool monitoring = !p.knocked_in && p.has_ki;
  const R level = (R)p.knock_in;
  bool ki = p.knocked_in;
  bool active = live;
  R lr = 0;
  int start = 0;
  for (int k = 0; k < p.n_intervals; ++k) {
    if (__ballot(active) == 0) break;
    // The whole row is read up front and combined with non-short-circuit
    // `&` and `|`, so that no per-lane branch guards a read.
    const IntervalRow<R> row = rows[k];
    const int end = row.end;
    // Every lane steps: a lane that is not active (knocked out, or not a
    // sample) never has its lr, draws or flags read again, so only cash and
    // the knock-out itself need the `active` guard, through selects.
    const R z = (R)sign * dr.template normal<R>(ROWS ? k : start);
    lr = lr + (FP32 ? fmaf((float)row.diffusion, (float)z, (float)row.drift)
                    : fma((double)row.diffusion, (double)z, (double)row.drift));
    dr.skip(end - start);
    const bool observed_in = monitoring & (row.observe != 0) & (lr < level);
    ki = ki | observed_in;
    monitoring = monitoring & !observed_in;
    if (row.event >= 0) {
      const bool hit = active & (row.has_knock_out != 0) & (lr >= row.knock_out);
      const bool due = lr >= row.coupon_barrier;
      const double v = hit ? (due ? row.pay_both : row.pay_out) : (due ? row.pay_coupon : 0.0);
      double pay = v;
      const bool odd = active & !(v == 0.0 || isnormal(v));
      if (wave_any(odd)) {
        if (odd) pay = event_payment(p.events[row.event], due, hit);
      }
      cash = active ? cash + pay : cash;
      active = active & !hit;
      monitoring = monitoring & !hit;
    }
    start = end + 1;
  }
  out.cash = cash;
  out.lr = lr;
  out.active = active;
  out.ki = ki;
  // An expiry knock-in below the level settles the question too.
  out.need_fill = active && monitoring && !(p.expiry_knock_in && lr < level);
  return out;
}

// Pass two, for a path that bridge_events left monitoring at expiry (so it
// never knocks out): draws the same end points again and fills until the
// path knocks in. Returns whether it does; `lr_end` is the last end point.
template <bool SOBOL, bool FP32, class D>
__device__ bool bridge_knock_in(const AcPlanDev& p, D& dr, float sign,
                                typename std::conditional<FP32, float, double>::type& lr_end) {
  using R = typename std::conditional<FP32, float, double>::type;
  const R level = (R)p.knock_in;
  bool monitoring = true;
  R lr = 0;
  int start = 0;
  for (int k = 0; k < p.n_intervals; ++k) {
    const int end = p.intervals[k].end;
    const R target = interval_target<FP32>(p, dr, k, start, lr, sign);
    const bool filling = monitoring;
    if (!filling) dr.skip(end - start);
    if (interval_fill<SOBOL, FP32>(p, dr, start, end, lr, target, level, filling, sign) && filling)
      monitoring = false;
    lr = target;
    start = end + 1;
  }
  lr_end = lr;
  return !monitoring;
}

// The expiry payoff added to a pass-one result (bridge_path's last lines).
template <typename R>
__device__ double bridge_finish(const AcPlanDev& p, const BridgeEvents<R>& ev, bool late_ki, bool live,
                                uint32_t& bad) {
  double cash = ev.cash;
  if (ev.active) {
    const bool hit = ev.ki || late_ki || (p.expiry_knock_in && p.has_ki && ev.lr < (R)p.knock_in);
    cash += terminal_value(p, (double)ev.lr, hit, bad);
  }
  if (live && !isfinite(cash)) bad = 1;
  return cash;
}

// Daily or event-only path (payoff, autocallable.rs:823-862).
template <bool SOBOL, bool FP32, class D>
__device__ double step_path(const AcPlanDev& p, D& dr, bool live, float sign,
                            uint32_t& bad) {
  using R = typename std::conditional<FP32, float, double>::type;
  double cash = p.coupon;
  bool ki = p.knocked_in;
  const R level = (R)p.knock_in;
  bool active = live;
  R lr = 0;
  for (int i = 0; i < p.n_steps; ++i) {
    if (__ballot(active) == 0) break;
    if (!active) continue;
    R z = (R)sign * dr.template normal<R>(i);
    lr += FP32 ? fmaf((float)Tab<R>::diff(p, i), (float)z, (float)Tab<R>::drift(p, i))
               : fma((double)Tab<R>::diff(p, i), (double)z, (double)Tab<R>::drift(p, i));
    if (!((double)lr >= p.safe_lo && (double)lr <= p.safe_hi)) {
      // validate_return (autocallable.rs:606-613)
      double s = discounted_product(p.initial_spot, 1.0, (double)lr);
      if (!isfinite(s) || s <= 0.0) bad = 1;
    }
    if (uload_u8(p.step_observe, i)) ki = ki || (p.has_ki && lr < level);
    const int e = uload(p.step_event + i);
    if (e >= 0) {
      const AcEventD& ev = p.events[e];
      const bool out = ev.has_knock_out && lr >= Tab<R>::ko(p, e);
      cash += event_payment(ev, lr >= Tab<R>::cb(p, e), out);
      if (out) active = false;
    }
  }
  if (active) {
    ki = ki || (p.expiry_knock_in && p.has_ki && lr < level);
    cash += terminal_value(p, (double)lr, ki, bad);
  }
  if (live && !isfinite(cash)) bad = 1;
  return cash;
}

// ------------------------------------------------------------------ scenario lanes
// Up to AC_MAX_LANES market scenarios of one note on the same draws
// (PathPlan::sample with several plans, autocallable.rs:580-599, and the
// bridge_payoffs / spot_payoffs lane loops, :628-800). The lanes share the
// step grid (step_event, step_observe, interval ends; checked on the host)
// but may differ in every level, transition and payment. Each lane repeats
// the single-plan arithmetic of bridge_path / step_path on its own state, so
// its value equals that of pricing its scenario alone; a normal is drawn
// while any lane of the sample still needs it. Per-step tables interleave
// the lanes ([step][lane]) so that one base pointer serves every lane. The
// lane loops are unrolled to L (a power of two at least n) so that each
// lane's state stays in registers; more lanes than AC_MAX_LANES take
// further launches over the same draws.
constexpr int AC_MAX_LANES = 4;   // lanes per launch
constexpr int AC_MAX_SCENARIOS = 16;  // lanes per call, in launches of AC_MAX_LANES

struct AcLaneScalars {
  double knock_in, coupon, initial_spot, safe_lo, safe_hi, final_log_discount;
  double term_base, term_no_ki, term_ki_fixed, lower, upper, inv_initial;
  int32_t knocked_in, has_ki, expiry_knock_in, downside;
};

struct AcLanes {
  int32_t n, n_steps, n_intervals, bridge;
  int32_t slot_base, slots;  // lane l's moments: partials[block * slots + slot_base + 2l (+1)]
  AcLaneScalars s[AC_MAX_LANES];
  const int32_t* step_event;    // [step], shared
  const uint8_t* step_observe;  // [step], shared
  const int32_t* interval_end;  // [interval], shared
  const AcStepD* steps;         // [step][lane]
  const double* weight;         // [step][lane]
  const double* fill;           // [step][lane]
  const AcEventD* events;       // [event][lane]
  const AcStepD* intervals;     // [interval][lane] drift, diffusion
  const float* steps32;         // [step][lane] drift, diffusion
  const float* keep32;          // [step][lane] 1 - weight
  const float* fill32;          // [step][lane] -diffusion
  const float* events32;        // [event][lane] knock_out, coupon_barrier
  const float* intervals32;     // [interval][lane] drift, diffusion
};

template <typename R>
struct LaneTab;
template <>
struct LaneTab<double> {
  static __device__ inline double drift(const AcLanes& P, int i, int l) { return P.steps[i * P.n + l].drift; }
  static __device__ inline double diff(const AcLanes& P, int i, int l) { return P.steps[i * P.n + l].diffusion; }
  static __device__ inline double weight(const AcLanes& P, int i, int l) { return P.weight[i * P.n + l]; }
  static __device__ inline double f(const AcLanes& P, int i, int l) { return P.fill[i * P.n + l]; }
  static __device__ inline double ko(const AcLanes& P, int e, int l) { return P.events[e * P.n + l].knock_out; }
  static __device__ inline double cb(const AcLanes& P, int e, int l) { return P.events[e * P.n + l].coupon_barrier; }
  static __device__ inline double idrift(const AcLanes& P, int k, int l) { return P.intervals[k * P.n + l].drift; }
  static __device__ inline double idiff(const AcLanes& P, int k, int l) {
    return P.intervals[k * P.n + l].diffusion;
  }
};
template <>
struct LaneTab<float> {
  static __device__ inline float drift(const AcLanes& P, int i, int l) { return P.steps32[2 * (i * P.n + l)]; }
  static __device__ inline float diff(const AcLanes& P, int i, int l) { return P.steps32[2 * (i * P.n + l) + 1]; }
  static __device__ inline float weight(const AcLanes& P, int i, int l) { return P.keep32[i * P.n + l]; }
  static __device__ inline float f(const AcLanes& P, int i, int l) { return P.fill32[i * P.n + l]; }
  static __device__ inline float ko(const AcLanes& P, int e, int l) { return P.events32[2 * (e * P.n + l)]; }
  static __device__ inline float cb(const AcLanes& P, int e, int l) { return P.events32[2 * (e * P.n + l) + 1]; }
  static __device__ inline float idrift(const AcLanes& P, int k, int l) { return P.intervals32[2 * (k * P.n + l)]; }
  static __device__ inline float idiff(const AcLanes& P, int k, int l) {
    return P.intervals32[2 * (k * P.n + l) + 1];
  }
};

// Per-lane flags of one sample are bit masks (bit l for lane l), so that
// they occupy one register each rather than one lane mask per lane.
__device__ inline bool bit(uint32_t mask, int l) { return (mask >> l) & 1u; }

template <int L, bool SOBOL, bool FP32, class D>
__device__ void bridge_lanes(const AcLanes& P, D& dr, bool live, float sign, uint32_t& bad,
                             double (&cash)[L]) {
  using R = typename std::conditional<FP32, float, double>::type;
  using T = LaneTab<R>;
  const int n = P.n;
  uint32_t monitoring = 0, ki = 0, active = 0;
  R lr[L], target[L];
  FillTrack<R> track[L];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const bool on = l < n;
    cash[l] = on ? P.s[l].coupon : 0.0;
    if (on && !P.s[l].knocked_in && P.s[l].has_ki) monitoring |= 1u << l;
    if (on && P.s[l].knocked_in) ki |= 1u << l;
    if (on && live) active |= 1u << l;
    lr[l] = 0;
  }
  int start = 0;
  for (int k = 0; k < P.n_intervals; ++k) {
    if (__ballot(active != 0) == 0) break;
    const int end = P.interval_end[k];
    R z = 0;
    if (active) z = (R)sign * dr.template normal<R>(start);
#pragma unroll
    for (int l = 0; l < L; ++l) {
      target[l] = lr[l];
      if (bit(active, l))
        target[l] = lr[l] + (FP32 ? fmaf((float)T::idiff(P, k, l), (float)z, (float)T::idrift(P, k, l))
                                  : fma((double)T::idiff(P, k, l), (double)z, (double)T::idrift(P, k, l)));
      track[l].start(lr[l], target[l], (R)P.s[l].knock_in);
    }
    // monitoring implies active: a knock-out ends both.
    uint32_t filling = monitoring;
    if (active && !filling) dr.skip(end - start);
    for (int i = start; i < end; ++i) {
      if (__ballot(filling != 0) == 0) break;
      if (filling) {
#pragma unroll
        for (int l = 0; l < L; ++l)
          if (bit(filling, l) && track[l].below()) filling &= ~(1u << l);
        if (!filling) dr.skip(end - i);
      }
      if (filling) {
        const R zf = (R)sign * dr.template normal<R>(i + 1);
#pragma unroll
        for (int l = 0; l < L; ++l)
          if (bit(filling, l)) track[l].step(target[l], T::weight(P, i, l), T::f(P, i, l), zf);
      }
    }
    start = end + 1;
    const bool observe = P.step_observe[end];
    const int e = P.step_event[end];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      lr[l] = target[l];
      if (!bit(active, l)) continue;
      const R level = (R)P.s[l].knock_in;
      if (bit(monitoring, l) && (track[l].below() || (observe && lr[l] < level))) {
        monitoring &= ~(1u << l);
        ki |= 1u << l;
      }
      if (e >= 0) {
        const AcEventD& ev = P.events[e * n + l];
        const bool hit = ev.has_knock_out && lr[l] >= T::ko(P, e, l);
        cash[l] += event_payment(ev, lr[l] >= T::cb(P, e, l), hit);
        if (hit) {
          active &= ~(1u << l);
          monitoring &= ~(1u << l);
        }
      }
    }
  }
#pragma unroll
  for (int l = 0; l < L; ++l) {
    if (bit(active, l)) {
      const AcLaneScalars& p = P.s[l];
      const bool hit = bit(ki, l) || (p.expiry_knock_in && p.has_ki && lr[l] < (R)p.knock_in);
      cash[l] += terminal_value(p, (double)lr[l], hit, bad);
    }
    if (live && l < n && !isfinite(cash[l])) bad = 1;
  }
}

template <int L, bool SOBOL, bool FP32, class D>
__device__ void step_lanes(const AcLanes& P, D& dr, bool live, float sign, uint32_t& bad,
                           double (&cash)[L]) {
  using R = typename std::conditional<FP32, float, double>::type;
  using T = LaneTab<R>;
  const int n = P.n;
  uint32_t ki = 0, active = 0;
  R lr[L];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const bool on = l < n;
    cash[l] = on ? P.s[l].coupon : 0.0;
    if (on && P.s[l].knocked_in) ki |= 1u << l;
    if (on && live) active |= 1u << l;
    lr[l] = 0;
  }
  for (int i = 0; i < P.n_steps; ++i) {
    if (__ballot(active != 0) == 0) break;
    if (!active) continue;
    const R z = (R)sign * dr.template normal<R>(i);
    const bool observe = P.step_observe[i];
    const int e = P.step_event[i];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      if (!bit(active, l)) continue;
      const AcLaneScalars& p = P.s[l];
      lr[l] += FP32 ? fmaf((float)T::diff(P, i, l), (float)z, (float)T::drift(P, i, l))
                    : fma((double)T::diff(P, i, l), (double)z, (double)T::drift(P, i, l));
      if (!((double)lr[l] >= p.safe_lo && (double)lr[l] <= p.safe_hi)) {
        double s = discounted_product(p.initial_spot, 1.0, (double)lr[l]);
        if (!isfinite(s) || s <= 0.0) bad = 1;
      }
      if (observe && p.has_ki && lr[l] < (R)p.knock_in) ki |= 1u << l;
      if (e >= 0) {
        const AcEventD& ev = P.events[e * n + l];
        const bool out = ev.has_knock_out && lr[l] >= T::ko(P, e, l);
        cash[l] += event_payment(ev, lr[l] >= T::cb(P, e, l), out);
        if (out) active &= ~(1u << l);
      }
    }
  }
#pragma unroll
  for (int l = 0; l < L; ++l) {
    if (bit(active, l)) {
      const AcLaneScalars& p = P.s[l];
      const bool hit = bit(ki, l) || (p.expiry_knock_in && p.has_ki && lr[l] < (R)p.knock_in);
      cash[l] += terminal_value(p, (double)lr[l], hit, bad);
    }
    if (live && l < n && !isfinite(cash[l])) bad = 1;
  }
}

// ------------------------------------------------------------------ host plan
inline int64_t floor_div(int64_t a, int64_t b) {
  int64_t q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}
inline double years(int64_t a, int64_t b) { return (double)(b - a) / (365.0 * 86400e6); }
inline bool trading_day(int calendar, int64_t day) {
  if (calendar == KG_CALENDAR_ALL_DAYS) return true;
  return ((day % 7) + 7 + 3) % 7 < 5;  // 1970-01-01 was a Thursday
}
inline void transition(double r, double q, double sigma, double time, double discount_time,
                              double* drift, double* diffusion, double* log_discount) {
  *drift = (r - q - 0.5 * sigma * sigma) * time;  // monte_carlo.rs:2123-2137
  *diffusion = sigma * std::sqrt(time);
  *log_discount = -r * discount_time;
  if (!std::isfinite(*drift) || !std::isfinite(*diffusion) || !std::isfinite(*log_discount))
    unstable();
}

struct AcHostPlan {
  bool fixed = false;
  double fixed_value = 0.0;
  AcPlanDev dev{};
  std::vector<AcStepD> steps;
  std::vector<int32_t> step_event;
  std::vector<uint8_t> step_observe;
  std::vector<AcEventD> events;
  std::vector<AcIntervalD> intervals;
  std::vector<double> weight, fill;
  std::vector<float> steps32, keep32, fill32, wf32, events32, intervals32;
};

struct NoteView {
  const kg_autocall_terms& t;
  bool has_ki() const { return t.knock_in_kind != KG_KNOCK_IN_NONE; }
  bool daily() const { return t.knock_in_kind == KG_KNOCK_IN_DAILY; }
  double maturity_coupon() const { return t.coupon_kind == KG_COUPON_SNOWBALL ? t.coupon : 0.0; }
  double period_coupon() const { return t.coupon_kind == KG_COUPON_PHOENIX ? t.coupon : 0.0; }
  double rebate(int64_t time) const {
    return t.rebate_rate + t.annual_rebate_rate * years(t.effective_time, time);
  }
  double knock_out_payment(const kg_autocall_event& e) const {
    return t.principal_redemption_ratio + rebate(e.time) +
           e.coupon_rate * years(t.effective_time, e.time);
  }
  double terminal(double spot, bool ki) const {
    double coupon = maturity_coupon();
    if (ki && has_ki()) {
      if (t.redemption_kind == KG_REDEMPTION_DOWNSIDE)
        return t.principal_redemption_ratio + rebate(t.expiry_time) +
               (std::min(std::max(spot, t.lower_strike), t.upper_strike) - t.upper_strike) /
                   t.initial_spot;
      coupon = t.fixed_coupon_rate;
    }
    return t.principal_redemption_ratio + rebate(t.expiry_time) +
           coupon * years(t.effective_time, t.expiry_time);
  }
  bool affects() const {
    if (!has_ki()) return false;
    if (t.redemption_kind == KG_REDEMPTION_FIXED_COUPON) return t.fixed_coupon_rate != maturity_coupon();
    return t.lower_strike != t.upper_strike || maturity_coupon() != 0.0;
  }
  bool daily_time(int64_t time) const {
    return daily() && time >= t.effective_time && time <= t.expiry_time &&
           time - floor_div(time, MICROS_PER_DAY) * MICROS_PER_DAY == t.observation_clock_micros &&
           trading_day(t.calendar, floor_div(time, MICROS_PER_DAY));
  }
  bool observes(int64_t time) const {
    if (t.knock_in_kind == KG_KNOCK_IN_DAILY) return daily_time(time);
    return t.knock_in_kind == KG_KNOCK_IN_AT_EXPIRY && time == t.expiry_time;
  }
};

inline void validate_terms(const kg_autocall_terms& t) {
  if (t.n_events < 0 || (t.n_events > 0 && !t.events))
    fail(KG_INVALID_INPUT, "events must not be null");
  if (!(t.initial_spot > 0.0)) fail(KG_INVALID_INPUT, "initial_spot must be positive");
  if (t.expiry_time <= t.effective_time)
    fail(KG_INVALID_INPUT, "expiry must follow the effective time");
  for (int i = 0; i < t.n_events; ++i) {
    if (i && t.events[i].time <= t.events[i - 1].time)
      fail(KG_INVALID_INPUT, "observations must be strictly ordered");
    if (t.events[i].time < t.effective_time || t.events[i].time > t.expiry_time)
      fail(KG_INVALID_INPUT, "observations must lie within the contract life");
  }
  if (t.knock_in_kind != KG_KNOCK_IN_NONE && !(t.knock_in_level > 0.0))
    fail(KG_INVALID_INPUT, "knock_in_level must be positive");
  if (t.knock_in_kind != KG_KNOCK_IN_NONE && t.redemption_kind == KG_REDEMPTION_DOWNSIDE &&
      !(0.0 <= t.lower_strike && t.lower_strike <= t.upper_strike))
    fail(KG_INVALID_INPUT, "lower_strike must lie in [0, upper_strike]");
}

// Whether a batch priced on shared daily grids gives this note the daily
// grid of its expiry, calendar and clock (DailyGrids::prepare,
// engines/monte_carlo/autocallable/portfolio.rs:75-186): the grid of
// daily observation times after valuation (plus expiry when it falls off
// the grid) must be small enough for the reference's 2 MiB grid slot (a
// 72-byte path step and an 8-byte time per day, its time vector grown by
// doubling from four) and must hold every future event. A note on such a
// grid is sampled step by step, without the Brownian bridge
// (autocallable.rs:341-374 leaves `bridge` empty); its dates, transitions,
// events and knock-in fixings are those of its own plan.
inline bool shared_daily_grid(const kg_autocall_terms& t, const NoteView& note, int64_t valuation) {
  constexpr uint64_t GRID_BYTES = 2u * 1024u * 1024u, STEP_BYTES = 72, TIME_BYTES = 8;
  const int64_t days = floor_div(t.expiry_time, MICROS_PER_DAY) - floor_div(valuation, MICROS_PER_DAY) + 1;
  if ((uint64_t)days >= GRID_BYTES / (TIME_BYTES + STEP_BYTES)) return false;
  for (int i = 0; i < t.n_events; ++i) {
    const int64_t time = t.events[i].time;
    if (time > valuation && time != t.expiry_time && !note.daily_time(time)) return false;
  }
  uint64_t n = 0;
  for (int64_t d = floor_div(valuation, MICROS_PER_DAY); d <= floor_div(t.expiry_time, MICROS_PER_DAY); ++d) {
    const int64_t at = d * MICROS_PER_DAY + t.observation_clock_micros;
    if (at > valuation && at <= t.expiry_time && trading_day(t.calendar, d)) ++n;
  }
  if (n == 0) return false;
  uint64_t capacity = 4;
  while (capacity < n) capacity *= 2;
  const int64_t last = floor_div(t.expiry_time, MICROS_PER_DAY) * MICROS_PER_DAY + t.observation_clock_micros;
  const bool on_grid = last == t.expiry_time && trading_day(t.calendar, floor_div(t.expiry_time, MICROS_PER_DAY));
  if (!on_grid && ++n > capacity) capacity *= 2;
  return capacity * TIME_BYTES + n * STEP_BYTES <= GRID_BYTES;
}

// The sampled plan of a note (or its known value). With `shared_grid` the
// note is planned as a batch on shared daily grids plans it
// (price_many_on_shared_grid, autocallable/portfolio.rs:187-318).
inline AcHostPlan build_plan(const kg_autocall_terms& t, const kg_bsm_market& m, bool shared_grid = false) {
  validate_terms(t);
  NoteView note{t};
  AcHostPlan h;
  if (!(m.spot > 0.0) || !std::isfinite(m.spot)) fail(KG_INVALID_INPUT, "spot must be positive");
  if (!(m.volatility >= 0.0)) fail(KG_INVALID_INPUT, "volatility must be nonnegative");
  if (m.volatility == 0.0)
    fail(KG_UNSUPPORTED_OPERATION,
         "zero volatility selects the deterministic engine, which is not part of the GPU engine");
  const int64_t valuation = m.valuation_time;
  if (valuation < t.effective_time || valuation > t.expiry_time)
    fail(KG_INVALID_INPUT, "valuation outside contract life");
  // prior_with_work_check (instruments/autocallable.rs:1321-1370)
  int64_t first_out = INT64_MAX;
  for (int i = 0; i < t.n_events; ++i)
    if (t.events[i].has_knock_out) {
      first_out = t.events[i].time;
      break;
    }
  bool had_out = first_out < valuation;
  bool had_in = false;
  if (note.daily() && valuation > t.effective_time) {
    const int64_t end = std::min(valuation - 1, t.expiry_time);
    for (int64_t d = floor_div(t.effective_time, MICROS_PER_DAY);
         d <= floor_div(end, MICROS_PER_DAY) && !had_in; ++d) {
      const int64_t at = d * MICROS_PER_DAY + t.observation_clock_micros;
      had_in = at >= t.effective_time && at <= end && trading_day(t.calendar, d);
    }
  }
  if (t.barrier_state == KG_STATE_UNKNOWN && (had_out || had_in))
    fail(KG_MISSING_HISTORY, "prior autocallable state is required");
  if (t.barrier_state == KG_STATE_KNOCKED_OUT && !had_out)
    fail(KG_INVALID_INPUT, "knock-out predates observations");
  if (t.barrier_state == KG_STATE_KNOCKED_IN && !had_in)
    fail(KG_INVALID_INPUT, "knock-in predates monitoring");
  // pricing::initial (instruments/autocallable/pricing.rs:12-46)
  if (t.barrier_state == KG_STATE_KNOCKED_OUT) {
    h.fixed = true;
    return h;
  }
  bool knocked_in = t.barrier_state == KG_STATE_KNOCKED_IN;
  if (note.daily_time(valuation)) knocked_in = knocked_in || (note.has_ki() && m.spot < t.knock_in_level);
  const kg_autocall_event* current = nullptr;
  for (int i = 0; i < t.n_events; ++i)
    if (t.events[i].time == valuation) current = &t.events[i];
  double coupon = 0.0;
  if (current && current->has_coupon_barrier && m.spot >= current->coupon_barrier_level)
    coupon = note.period_coupon();
  if (valuation == t.expiry_time) {  // pricing::expiry_payment (:48-68)
    double v;
    if (current && current->has_knock_out && m.spot >= current->knock_out_level)
      v = note.knock_out_payment(*current);
    else
      v = note.terminal(m.spot, knocked_in || (note.observes(t.expiry_time) && note.has_ki() &&
                                               m.spot < t.knock_in_level));
    h.fixed = true;
    h.fixed_value = coupon + v;
    return h;
  }
  if (current && current->has_knock_out && m.spot >= current->knock_out_level) {
    h.fixed = true;
    h.fixed_value = note.knock_out_payment(*current) + coupon;
    return h;
  }
  const double r = m.rate, q = m.dividend, sigma = m.volatility;
  const bool daily = note.daily() && !knocked_in && note.affects();
  AcPlanDev& p = h.dev;
  p.knocked_in = knocked_in;
  p.coupon = coupon;
  p.initial_spot = m.spot;
  const double log_spot = std::log(m.spot);
  const double low = std::log(std::numeric_limits<double>::min()) + 1.0;
  const double high = std::log(std::numeric_limits<double>::max()) - 1.0;
  p.safe_lo = std::max(low, low - log_spot);  // safe_log_returns (monte_carlo.rs:2139-2145)
  p.safe_hi = std::min(high, high - log_spot);
  p.has_ki = note.has_ki();
  p.knock_in = p.has_ki ? log_ratio(t.knock_in_level, m.spot) : 0.0;
  p.expiry_knock_in = note.observes(t.expiry_time);
  // future_event_times (pricing.rs:70-106)
  std::vector<int64_t> dates;
  for (int i = 0; i < t.n_events; ++i)
    if (t.events[i].time > valuation) dates.push_back(t.events[i].time);
  if (daily)
    for (int64_t d = floor_div(valuation, MICROS_PER_DAY); d <= floor_div(t.expiry_time, MICROS_PER_DAY);
         ++d) {
      const int64_t at = d * MICROS_PER_DAY + t.observation_clock_micros;
      if (at > valuation && at <= t.expiry_time && trading_day(t.calendar, d)) dates.push_back(at);
    }
  dates.push_back(t.expiry_time);
  std::sort(dates.begin(), dates.end());
  dates.erase(std::unique(dates.begin(), dates.end()), dates.end());
  if (dates.size() > 100000) fail(KG_INVALID_INPUT, "the note has more than 100,000 path steps");
  int next = 0;
  while (next < t.n_events && t.events[next].time <= valuation) ++next;
  int64_t previous = valuation;
  double excursion = 0.0, last_discount = 0.0;
  for (int64_t date : dates) {
    AcStepD s;
    double disc;
    transition(r, q, sigma, years(previous, date), years(valuation, date), &s.drift, &s.diffusion,
               &disc);
    excursion += std::fabs(s.drift) + 20.0 * s.diffusion;
    int32_t ev = -1;
    if (next < t.n_events && t.events[next].time == date) {
      const kg_autocall_event& e = t.events[next++];
      AcEventD c{};
      c.has_knock_out = e.has_knock_out != 0;
      c.knock_out = c.has_knock_out ? log_ratio(e.knock_out_level, m.spot) : INFINITY;
      c.coupon_barrier = log_ratio(e.has_coupon_barrier ? e.coupon_barrier_level : 0.0, m.spot);
      const double pc = e.has_coupon_barrier ? note.period_coupon() : 0.0;
      const double red = c.has_knock_out ? note.knock_out_payment(e) : 0.0;
      c.log_discount = disc;
      c.pay_coupon_raw = pc;
      c.pay_out_raw = red;
      c.pay_both_raw = red + pc;
      c.pay_coupon = h_discounted_product(1.0, pc, disc);
      c.pay_out = h_discounted_product(1.0, red, disc);
      c.pay_both = h_discounted_product(1.0, red + pc, disc);
      ev = (int32_t)h.events.size();
      h.events.push_back(c);
    }
    h.steps.push_back(s);
    h.step_event.push_back(ev);
    h.step_observe.push_back(daily && note.daily_time(date));
    last_discount = disc;
    previous = date;
  }
  p.final_log_discount = last_discount;
  p.term_base = t.principal_redemption_ratio + note.rebate(t.expiry_time);
  const double life = years(t.effective_time, t.expiry_time);
  p.term_no_ki = note.maturity_coupon() * life;
  p.term_ki_fixed = (t.redemption_kind == KG_REDEMPTION_FIXED_COUPON || !note.has_ki())
                        ? (note.has_ki() ? t.fixed_coupon_rate : note.maturity_coupon()) * life
                        : 0.0;
  p.downside = note.has_ki() && t.redemption_kind == KG_REDEMPTION_DOWNSIDE;
  p.lower = t.lower_strike;
  p.upper = t.upper_strike;
  p.inv_initial = 1.0 / t.initial_spot;
  // BridgePlan::new (autocallable.rs:463-515)
  const bool safe = excursion >= p.safe_lo && excursion <= p.safe_hi && -excursion >= p.safe_lo &&
                    -excursion <= p.safe_hi;
  if (daily && safe && !(shared_grid && shared_daily_grid(t, note, valuation))) {
    std::vector<AcIntervalD> iv;
    size_t first = 0;
    bool gap = false;
    for (size_t end = 0; end < h.steps.size(); ++end) {
      if (h.step_event[end] >= 0 || end + 1 == h.steps.size()) {
        gap |= end - first >= 8;
        const int64_t start = first == 0 ? valuation : dates[first - 1];
        AcIntervalD b{};
        b.end = (int32_t)end;
        double unused;
        transition(r, q, sigma, years(start, dates[end]), 0.0, &b.drift, &b.diffusion, &unused);
        iv.push_back(b);
        first = end + 1;
      }
    }
    if (gap) {
      std::vector<double> w(h.steps.size(), 0.0), f(h.steps.size(), 0.0);
      bool ok = true;
      size_t from = 0;
      int64_t prev = valuation;
      for (const AcIntervalD& b : iv) {
        const int64_t end_t = dates[b.end];
        for (size_t i = from; i < (size_t)b.end; ++i) {
          const double dt = years(prev, dates[i]);
          const double fraction = dt / years(prev, end_t);
          if (!(0.0 < fraction && fraction < 1.0)) {
            ok = false;
            break;
          }
          w[i] = fraction;
          f[i] = sigma * std::sqrt(dt * (1.0 - fraction));
          prev = dates[i];
        }
        if (!ok) break;
        from = b.end + 1;
        prev = end_t;
      }
      if (ok) {
        p.bridge = 1;
        h.intervals = iv;
        h.weight = w;
        h.fill = f;
      }
    }
  }
  p.n_steps = (int32_t)h.steps.size();
  p.n_intervals = (int32_t)h.intervals.size();
  for (const AcStepD& s : h.steps) {
    h.steps32.push_back((float)s.drift);
    h.steps32.push_back((float)s.diffusion);
  }
  for (double v : h.weight) h.keep32.push_back((float)(1.0 - v));
  for (double v : h.fill) h.fill32.push_back((float)-v);
  for (size_t i = 0; i < h.keep32.size(); ++i) {
    h.wf32.push_back(h.keep32[i]);
    h.wf32.push_back(h.fill32[i]);
  }
  for (const AcEventD& e : h.events) {
    h.events32.push_back((float)e.knock_out);
    h.events32.push_back((float)e.coupon_barrier);
  }
  for (const AcIntervalD& b : h.intervals) {
    h.intervals32.push_back((float)b.drift);
    h.intervals32.push_back((float)b.diffusion);
  }
  return h;
}

// Copies the plan tables into one device buffer and points the plan at them.
template <class T>
inline void stage(std::vector<unsigned char>& blob, const std::vector<T>& v, size_t* offset) {
  *offset = (blob.size() + 255) & ~size_t(255);
  // Whole words, so that uload_u8 may read the word holding any entry.
  blob.resize(*offset + ((v.size() * sizeof(T) + 4) & ~size_t(3)));
  if (!v.empty()) std::memcpy(blob.data() + *offset, v.data(), v.size() * sizeof(T));
}

struct AcOffsets {
  size_t o[11];
};

inline AcOffsets stage_plan(std::vector<unsigned char>& blob, const AcHostPlan& h) {
  AcOffsets r;
  size_t* o = r.o;
  stage(blob, h.steps, &o[0]);
  stage(blob, h.step_event, &o[1]);
  stage(blob, h.step_observe, &o[2]);
  stage(blob, h.events, &o[3]);
  stage(blob, h.intervals, &o[4]);
  stage(blob, h.weight, &o[5]);
  stage(blob, h.fill, &o[6]);
  stage(blob, h.steps32, &o[7]);
  stage(blob, h.wf32, &o[8]);
  stage(blob, h.events32, &o[9]);
  stage(blob, h.intervals32, &o[10]);
  return r;
}

// The plan of `h` with its table pointers into device copy `d` of the blob.
inline AcPlanDev bind_plan(const AcHostPlan& h, unsigned char* d, const AcOffsets& off) {
  const size_t* o = off.o;
  AcPlanDev p = h.dev;
  p.steps = (const AcStepD*)(d + o[0]);
  p.step_event = (const int32_t*)(d + o[1]);
  p.step_observe = (const uint8_t*)(d + o[2]);
  p.events = (const AcEventD*)(d + o[3]);
  p.intervals = (const AcIntervalD*)(d + o[4]);
  p.weight = (const double*)(d + o[5]);
  p.fill = (const double*)(d + o[6]);
  p.steps32 = (const float*)(d + o[7]);
  p.wf32 = (const float*)(d + o[8]);
  p.events32 = (const float*)(d + o[9]);
  p.intervals32 = (const float*)(d + o[10]);
  return p;
}

inline AcPlanDev upload(kg_context* ctx, AcHostPlan& h) {
  std::vector<unsigned char> blob;
  const AcOffsets off = stage_plan(blob, h);
  unsigned char* d = (unsigned char*)ctx->scratch.reserve(blob.size());
  upload_async(ctx, d, blob.data(), blob.size());
  return bind_plan(h, d, off);
}

// Whether two sampled plans share the step grid the lane loops walk.
inline bool same_grid(const AcHostPlan& a, const AcHostPlan& b) {
  if (a.dev.n_steps != b.dev.n_steps || a.dev.bridge != b.dev.bridge || a.step_event != b.step_event ||
      a.step_observe != b.step_observe || a.intervals.size() != b.intervals.size())
    return false;
  for (size_t k = 0; k < a.intervals.size(); ++k)
    if (a.intervals[k].end != b.intervals[k].end) return false;
  return true;
}

}  // namespace kg
