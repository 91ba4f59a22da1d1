/*
 * kiyosi_amd — C ABI of the MI355X (gfx950) Monte Carlo pricing engine.
 *
 * The library replaces the CPU Monte Carlo engines of kiyosi 0.23.0
 * (crates/kiyosi/src/engines/monte_carlo.rs and engines/monte_carlo/). Each
 * entry point below corresponds to one `PricingEngine<I>` method of one
 * engine (crates/kiyosi/src/risk/analytics.rs:554-834) and names the
 * reference function it stands in for. Plain C types only: pointers, sizes,
 * integers and doubles. Times are signed microseconds since
 * 1970-01-01T00:00:00 UTC, the resolution of kiyosi's `Timestamp`
 * (crates/kiyosi/src/time/date.rs:255-257).
 *
 * Ownership follows the reference: inputs are borrowed and read-only, every
 * call owns its device buffers ("every call owns its buffers",
 * monte_carlo.rs:738-740), outputs are written to caller memory. A context is
 * bound to one GPU; calls on one context are serialised by an internal mutex,
 * so a context may be shared between threads like the reference's `Send +
 * Sync` engines (crates/kiyosi/src/lib.rs:103-125).
 */
#ifndef KIYOSI_AMD_H
#define KIYOSI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error kinds of crates/kiyosi/src/error.rs:18-57. */
typedef enum kg_status {
  KG_OK = 0,
  KG_INVALID_INPUT = 1,         /* Error::InvalidInput */
  KG_UNSUPPORTED_OPERATION = 2, /* Error::UnsupportedOperation */
  KG_COMPUTATION_FAILURE = 3,   /* Error::ComputationFailure (non-finite path or result) */
  KG_INTERRUPTED = 4,           /* Error::Io(Interrupted): the cancel flag was raised */
  KG_DEVICE_ERROR = 5,          /* Error::Io: a HIP runtime failure */
  KG_MISSING_HISTORY = 6,       /* Error::MissingHistory */
  KG_NON_CONVERGENCE = 7        /* Error::NonConvergence (a series or quadrature did not converge) */
} kg_status;

/* MonteCarloSampling (monte_carlo.rs:76-82). */
typedef enum kg_sampling { KG_PSEUDO_RANDOM = 0, KG_SOBOL = 1 } kg_sampling;

/* Arithmetic of the path loop. KG_FP64 is the plain form that restates the
 * reference arithmetic; KG_FP32_STEP steps paths in FP32 and accumulates
 * discounted payoffs in FP64. */
typedef enum kg_precision { KG_FP64 = 0, KG_FP32_STEP = 1 } kg_precision;

/* MonteCarloUncertaintyStatus (monte_carlo.rs:869-894). */
typedef enum kg_uncertainty {
  KG_AVAILABLE = 0,
  KG_UNSUPPORTED_SAMPLING = 1,
  KG_INSUFFICIENT_OBSERVATIONS = 2,
  KG_INCOMPLETE_ANTITHETIC_PAIR = 3,
  KG_ZERO_OBSERVED_VARIANCE = 4
} kg_uncertainty;

/* MonteCarloSettings (monte_carlo.rs:90-118) plus the GPU precision knob.
 * The seed must be resolved by the caller (`with_resolved_seed`,
 * monte_carlo.rs:163-177): has_seed == 0 is rejected. */
typedef struct kg_settings {
  int64_t path_count;         /* 1..=10,000,000 as in the reference, or up to 2^26 with allow_large */
  uint64_t seed;
  int32_t has_seed;
  int32_t sampling;           /* kg_sampling */
  int32_t antithetic;         /* pairs Z and -Z */
  int32_t precision;          /* kg_precision */
  int32_t max_active_workers; /* validated (>= 0) and otherwise unused on the GPU */
  int32_t allow_large;        /* lifts the 1e7-path and 1e9-path-step caps (GPU limit 2^26 paths) */
} kg_settings;

/* Moments (monte_carlo.rs:2155-2161): scale-normalised mean and M2. */
typedef struct kg_moments {
  uint64_t count;
  double scale;
  double mean;
  double m2;
} kg_moments;

/* MonteCarloEstimate (monte_carlo.rs:859-1001) plus GPU diagnostics. */
typedef struct kg_estimate {
  double price;
  double standard_error; /* valid iff has_standard_error */
  double ci_lower;       /* price -/+ 1.959963984540054 * standard_error */
  double ci_upper;
  int32_t has_standard_error;
  int32_t uncertainty_status; /* kg_uncertainty */
  int32_t has_seed;           /* 0 when no sampling occurred */
  int32_t reserved;
  uint64_t seed;
  /* Diagnostics, not part of the reference result: the standard error of
   * the per-path discounted values computed as if they were independent
   * (also for Sobol, where the reference reports none), the number of
   * simulated path-steps and the device time of the sampling kernels. */
  double path_standard_error;
  double path_steps;
  double device_ms;
} kg_estimate;

/* Flat-rate Black-Scholes-Merton market (market/parameters.rs:314-518,
 * BlackScholesMertonParameters with flat curves). */
typedef struct kg_bsm_market {
  double spot;
  double rate;
  double dividend;
  double volatility;
  int64_t valuation_time;
} kg_bsm_market;

typedef struct kg_context kg_context;

/* Creates a context on HIP device `device`. */
kg_status kg_context_create(int32_t device, kg_context** out);
void kg_context_destroy(kg_context* ctx);
/* Message of the last failed call on this context (empty if none). */
const char* kg_last_error(const kg_context* ctx);
/* Version string of the library. */
const char* kg_version(void);

/* ---- Cooperative cancellation: CancellationToken (crates/kiyosi/src/
 * cancellation.rs:33-68), whose flag the reference engines check per block,
 * per 64 paths and per 1,024 steps (engines/monte_carlo.rs:1749-1751,
 * 1813-1828). A token is one word of pinned, device-visible host memory
 * owned by the library; kg_cancel raises it from any thread and it stays
 * raised until kg_cancel_reset (a cancelled request returns no partial
 * result). A context bound to a token checks it when a call starts and
 * between the launches of multi-launch calls (LSMC dates, MLP training
 * steps, scenario groups, book replays and autotuning), and the persistent
 * book kernel reads it between tiles; a call that sees it raised returns
 * KG_INTERRUPTED (Error::Io with ErrorKind::Interrupted) with every kernel
 * ended normally, and the context stays usable. */
typedef struct kg_cancel_token kg_cancel_token;
kg_status kg_cancel_token_create(kg_cancel_token** out);
void kg_cancel_token_destroy(kg_cancel_token* token);
void kg_cancel(kg_cancel_token* token);
void kg_cancel_reset(kg_cancel_token* token);
int32_t kg_cancel_requested(const kg_cancel_token* token);
/* Binds `token` (NULL unbinds) to the context's later calls. */
kg_status kg_context_set_cancel_token(kg_context* ctx, kg_cancel_token* token);

/* ---- European vanilla: MonteCarloVanillaEngine::price_with_statistics
 * (monte_carlo.rs:1307-1320), sample (:1093-1141), EuropeanPayoff
 * (:1165-1252). option_type: 1 call, -1 put. */
typedef struct kg_european_terms {
  int32_t option_type;
  int32_t reserved;
  double strike;
  int64_t effective_time;
  int64_t expiry_time;
} kg_european_terms;

kg_status kg_vanilla_price(kg_context* ctx, const kg_settings* settings,
                           const kg_european_terms* option, const kg_bsm_market* market,
                           kg_estimate* out);

/* ---- kg_vanilla_price without device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc). Same kernels and results, bit for
 * bit; the per-block partials, the flag word and the merged moments live in
 * one fixed device array per GPU (about 2 MiB, held until the process ends)
 * instead of the context's buffers. At most 8,388,608 samples (32,768 blocks
 * of 256); larger requests return KG_INVALID_INPUT. Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (several
 * contexts or threads) race; serialise them. */
kg_status kg_vanilla_price_noalloc(kg_context* ctx, const kg_settings* settings,
                                   const kg_european_terms* option, const kg_bsm_market* market,
                                   kg_estimate* out);

/* ---- Shared-draw scenarios of a European option: the batched Greek stencil
 * of MonteCarloVanillaEngine::price_with_greeks_with_settings
 * (monte_carlo.rs:1334-1409, european_greek_prices :1143-1162). Prices the
 * option under n scenario markets on the same normals; out[i].price equals
 * kg_vanilla_price under markets[i] with the same settings, bit for bit.
 * Scenario estimates carry the price and path_standard_error only
 * (has_standard_error = 0). Peak-scaled payoffs (volatility * sqrt(T) > 25
 * for calls) return KG_UNSUPPORTED_OPERATION: price them one at a time. */
#define KG_MAX_SCENARIOS 64
kg_status kg_vanilla_scenarios(kg_context* ctx, const kg_settings* settings,
                               const kg_european_terms* option, const kg_bsm_market* markets,
                               int32_t n, kg_estimate* out);

/* ---- Autocallable notes (snowball, binary/ternary snowball, Phoenix):
 * MonteCarloAutocallableEngine::price_with_statistics
 * (engines/monte_carlo/autocallable.rs:874-918), plan building
 * prepare_sample_on_grid (:293-430), path payoffs bridge_payoffs
 * (:628-736) and payoff (:823-862), note terms AutocallableNote
 * (instruments/autocallable.rs:700-870, 1389-1424).
 *
 * `events` are the note's merged observation events
 * (AutocallableNote::new, instruments/autocallable.rs:972-1040): strictly
 * increasing times, each with an optional knock-out level, an optional
 * coupon barrier and the snowball coupon rate accrued on knock-out. */
typedef struct kg_autocall_event {
  int64_t time;
  int32_t has_knock_out;
  int32_t has_coupon_barrier;
  double knock_out_level;
  double coupon_barrier_level;
  double coupon_rate;
} kg_autocall_event;

enum {
  KG_CALENDAR_WEEKDAYS = 0, /* TradingCalendar::weekdays (time/calendar.rs:156-161) */
  KG_CALENDAR_ALL_DAYS = 1  /* TradingCalendar::all_days (time/calendar.rs:146-151) */
};
enum { KG_KNOCK_IN_NONE = 0, KG_KNOCK_IN_DAILY = 1, KG_KNOCK_IN_AT_EXPIRY = 2 };
enum { KG_REDEMPTION_DOWNSIDE = 0, KG_REDEMPTION_FIXED_COUPON = 1 };
enum { KG_COUPON_SNOWBALL = 0, KG_COUPON_PHOENIX = 1 };
enum {
  KG_STATE_UNKNOWN = -1, /* barrier_state None */
  KG_STATE_NO_EVENT = 0,
  KG_STATE_KNOCKED_OUT = 1,
  KG_STATE_KNOCKED_IN = 2
};

typedef struct kg_autocall_terms {
  int32_t calendar;
  int32_t knock_in_kind;   /* KG_KNOCK_IN_* */
  int32_t redemption_kind; /* KG_REDEMPTION_* (ignored without knock-in) */
  int32_t coupon_kind;     /* KG_COUPON_* */
  int32_t barrier_state;   /* KG_STATE_* */
  int32_t n_events;
  int64_t observation_clock_micros; /* time of day of daily knock-in checks */
  double knock_in_level;
  double initial_spot;
  double upper_strike;
  double lower_strike;
  double fixed_coupon_rate; /* ternary snowball minimum coupon */
  double coupon;            /* maturity coupon rate (snowball) or coupon per period (Phoenix) */
  double principal_redemption_ratio;
  double rebate_rate;
  double annual_rebate_rate;
  int64_t effective_time;
  int64_t expiry_time;
  const kg_autocall_event* events;
} kg_autocall_terms;

kg_status kg_autocallable_price(kg_context* ctx, const kg_settings* settings,
                                const kg_autocall_terms* note, const kg_bsm_market* market,
                                kg_estimate* out);

/* Prices one note in n market scenarios (1..KG_MAX_AUTOCALL_SCENARIOS) on the
 * same draws: the shared-draw batches of
 * MonteCarloAutocallableEngine::price_with_greeks_with_settings
 * (engines/monte_carlo/autocallable.rs:919-999), i.e. spot_prices (:174-197)
 * for spot bumps and market_prices (:199-241) for volatility and rate bumps,
 * both through PathPlan::sample with several plans (:580-599). out[i] is
 * exactly what kg_autocallable_price returns for markets[i] with the same
 * settings, device_ms aside (shared by the lanes). Scenarios with a known
 * value (knock-out or expiry at the valuation) take it. The sampled scenarios
 * must share the note's step grid (same valuation time, knock-in state and
 * bridge); otherwise the call returns KG_UNSUPPORTED_OPERATION and the
 * scenarios are priced one at a time. */
#define KG_MAX_AUTOCALL_SCENARIOS 16
kg_status kg_autocallable_scenarios(kg_context* ctx, const kg_settings* settings,
                                    const kg_autocall_terms* note, const kg_bsm_market* markets,
                                    int32_t n, kg_estimate* out);

/* ---- Autocallable portfolio: MonteCarloAutocallableEngine::price_many
 * (engines/monte_carlo/autocallable/portfolio.rs:187-318; Python binding
 * crates/kiyosi-python/src/engines/monte_carlo.rs:699-746). Every note is
 * priced with the one resolved seed of `settings`, and out[i] equals
 * kg_autocallable_price for notes[i] bit for bit, as the reference's batch
 * equals its scalar prices (crates/kiyosi/tests/autocallable_monte_carlo.rs
 * :52-99). Notes whose plans share the step grid (steps, events, knock-in
 * observations and bridge) run as lanes on one path set; the others run
 * alone. At most 1,048,576 notes (binding_support::validate_contract_batch_size). */
kg_status kg_autocallable_price_many(kg_context* ctx, const kg_settings* settings,
                                     const kg_autocall_terms* notes, size_t n,
                                     const kg_bsm_market* market, kg_estimate* out);

/* price_many_on_shared_grid (the same portfolio.rs:187-318 with
 * share_daily_grid): a note with daily knock-in monitoring whose future
 * events all fall on its daily observation times (or at expiry), and whose
 * daily grid fits the reference's grid slot (DailyGrids::prepare,
 * portfolio.rs:75-186), is sampled on that daily grid step by step, without
 * the Brownian bridge of its scalar plan (autocallable.rs:341-374); the
 * other notes keep their own plans. out[i] is the note's price on the
 * batch's draws, which the reference's coupon-family and threshold-strip
 * estimators (portfolio.rs:570-825) reproduce to rounding. */
kg_status kg_autocallable_price_many_on_shared_grid(kg_context* ctx, const kg_settings* settings,
                                                    const kg_autocall_terms* notes, size_t n,
                                                    const kg_bsm_market* market, kg_estimate* out);

/* ---- Heston path stepping. kiyosi has Heston only in its Fourier and FD
 * engines (market/heston.rs:58-227, engines/analytic/heston.rs,
 * engines/finite_difference/heston.rs); every reference MC engine rejects
 * Heston contexts with UnsupportedOperation (market/parameters.rs:271-278).
 * This entry point extends MonteCarloAsianEngine / MonteCarloVanillaEngine
 * to HestonParameters with a full-truncation Euler scheme in log-spot on
 * `step_count` equal steps: per step, normals 2k and 2k+1 of the sample
 * (one Box-Muller pair) drive variance and spot with correlation rho.
 * payoff: KG_HESTON_EUROPEAN pays on the terminal spot, KG_HESTON_ASIAN_*
 * on the arithmetic or geometric mean of the step_count post-step spots. */
enum {
  KG_HESTON_EUROPEAN = 0,
  KG_HESTON_ASIAN_ARITHMETIC = 1,
  KG_HESTON_ASIAN_GEOMETRIC = 2
};

typedef struct kg_heston_market {
  double spot;
  double rate;
  double dividend;
  double v0;    /* initial variance */
  double kappa; /* mean reversion speed */
  double theta; /* long-run variance */
  double xi;    /* volatility of variance */
  double rho;   /* spot-variance correlation */
  int64_t valuation_time;
} kg_heston_market;

typedef struct kg_heston_terms {
  int32_t payoff;      /* KG_HESTON_* */
  int32_t option_type; /* 1 call, -1 put */
  int32_t step_count;  /* 1..=10,000 equal steps from valuation to expiry */
  int32_t reserved;
  double strike;
  int64_t effective_time;
  int64_t expiry_time;
} kg_heston_terms;

kg_status kg_heston_price(kg_context* ctx, const kg_settings* settings,
                          const kg_heston_terms* option, const kg_heston_market* market,
                          kg_estimate* out);

/* ---- Single barrier options: MonteCarloBarrierEngine
 * (engines/monte_carlo/barrier.rs:160-541): scheduled fixings walk the
 * cumulative noise against the barrier on each observation date (knock-ins
 * with more than 8 fixings left are conditioned on the hit with a lognormal
 * price, :320-333); continuous monitoring draws the terminal value and
 * weights it by the Brownian-bridge survival (:402-440). A continuous
 * knock-out whose rebate is paid at the hit carries no rebate on its paths
 * and adds the analytic value of the hit payment (:486-531). touch_state is
 * the history before valuation (BarrierTouchState, instruments/barrier.rs:
 * 38-47; monitoring.rs:277-313): once monitoring began it must be known
 * (KG_MISSING_HISTORY), a touch before monitoring is invalid, and a prior
 * touch settles the contract (its at-hit rebate already paid). Delta and
 * gamma: kg_barrier_price_with_greeks below. */
enum { KG_DOWN_AND_OUT = 0, KG_DOWN_AND_IN = 1, KG_UP_AND_OUT = 2, KG_UP_AND_IN = 3 };
enum { KG_OBSERVE_CONTINUOUS = 0, KG_OBSERVE_SCHEDULED = 1 };
enum { KG_REBATE_AT_EXPIRY = 0, KG_REBATE_AT_HIT = 1 };
enum { KG_TOUCH_UNKNOWN = -1, KG_TOUCH_UNTOUCHED = 0, KG_TOUCH_TOUCHED = 1 };

typedef struct kg_barrier_terms {
  int32_t option_type;  /* 1 call, -1 put */
  int32_t barrier_type; /* KG_DOWN_AND_OUT .. KG_UP_AND_IN */
  int32_t observation;  /* KG_OBSERVE_* */
  int32_t rebate_timing;
  double strike;
  double barrier;
  double rebate;
  int64_t effective_time;
  int64_t expiry_time;
  int32_t n_observations; /* scheduled fixings, strictly increasing */
  int32_t touch_state;    /* KG_TOUCH_*: observations strictly before valuation */
  const int64_t* observation_times;
} kg_barrier_terms;

kg_status kg_barrier_price(kg_context* ctx, const kg_settings* settings,
                           const kg_barrier_terms* option, const kg_bsm_market* market,
                           kg_estimate* out);

/* ---- Spot Greeks of the barrier and digital engines: the likelihood-ratio
 * delta and gamma that MonteCarloBarrierEngine (barrier.rs:194-205, 440-472)
 * and MonteCarloDigitalEngine (digital.rs:116-136) compute inside the path
 * closure, on the draws of the price. `request` selects them; `spot_radius`
 * is the symmetric spot shift of the caller's stencil
 * (symmetric_shift(spot, spot_shift_at(spot)), barrier.rs:186-193): a
 * monitored barrier level within it makes the spot Greeks unavailable
 * (at_spot_discontinuity, instruments/monitoring.rs:430-445). Greeks are
 * never computed at expiry. `available` reports which were set; a barrier
 * resolved to its vanilla payoff reports the vanilla's analytic delta and
 * gamma, other resolved contracts zero (barrier.rs:206-229). */
enum { KG_GREEK_DELTA = 1, KG_GREEK_GAMMA = 2 };

typedef struct kg_spot_greeks {
  int32_t request;   /* in: KG_GREEK_* bits */
  int32_t available; /* out: KG_GREEK_* bits of the values set */
  double spot_radius;
  double delta;
  double gamma;
} kg_spot_greeks;

/* Single barrier with its spot Greeks:
 * MonteCarloBarrierEngine::price_with_greeks_with_settings (barrier.rs:618-640).
 * out->price is bit-identical to kg_barrier_price on the same inputs. */
kg_status kg_barrier_price_with_greeks(kg_context* ctx, const kg_settings* settings,
                                       const kg_barrier_terms* option, const kg_bsm_market* market,
                                       kg_spot_greeks* greeks, kg_estimate* out);

/* ---- Double barrier options: MonteCarloBarrierEngine for
 * DoubleBarrierOption (barrier.rs:14-18, 160-541, 648). The contract is its
 * lower (down-type) barrier plus the upper level
 * (instruments/double_barrier.rs:119-144). Scheduled fixings knock on
 * either level; continuous monitoring weights the terminal draw by the
 * image/spectral survival series of double_barrier.rs:6-112. A continuous
 * knock-out with an at-hit rebate adds the analytic first-exit value of the
 * rebate (engines/analytic/double_barrier.rs:423-520); touch_state as for
 * single barriers. */
enum { KG_DOUBLE_KNOCK_OUT = 0, KG_DOUBLE_KNOCK_IN = 1 };

typedef struct kg_double_barrier_terms {
  int32_t option_type;   /* 1 call, -1 put */
  int32_t knock_type;    /* KG_DOUBLE_KNOCK_* */
  int32_t observation;   /* KG_OBSERVE_* */
  int32_t rebate_timing; /* KG_REBATE_*; at-hit knock-in rebates are invalid */
  double strike;
  double lower_barrier;  /* 0 < lower < upper */
  double upper_barrier;
  double rebate;
  int64_t effective_time;
  int64_t expiry_time;
  int32_t n_observations; /* scheduled fixings, strictly increasing */
  int32_t touch_state;    /* KG_TOUCH_* */
  const int64_t* observation_times;
} kg_double_barrier_terms;

kg_status kg_double_barrier_price(kg_context* ctx, const kg_settings* settings,
                                  const kg_double_barrier_terms* option, const kg_bsm_market* market,
                                  kg_estimate* out);
kg_status kg_double_barrier_price_with_greeks(kg_context* ctx, const kg_settings* settings,
                                              const kg_double_barrier_terms* option,
                                              const kg_bsm_market* market, kg_spot_greeks* greeks,
                                              kg_estimate* out);

/* ---- European digitals: MonteCarloDigitalEngine (engines/monte_carlo/
 * digital.rs:55-150) for DigitalOption (instruments/vanilla.rs:236-366).
 * Cash-or-nothing pays cash_payout, asset-or-nothing the terminal spot,
 * when the terminal value is strictly in the money in standardised
 * coordinates. The valuation may precede the effective time. */
enum { KG_CASH_OR_NOTHING = 0, KG_ASSET_OR_NOTHING = 1 };

typedef struct kg_digital_terms {
  int32_t option_type; /* 1 call, -1 put */
  int32_t payout_kind; /* KG_CASH_OR_NOTHING or KG_ASSET_OR_NOTHING */
  double strike;
  double cash_payout;  /* nonnegative; ignored for asset settlement */
  int64_t effective_time;
  int64_t expiry_time;
} kg_digital_terms;

kg_status kg_digital_price(kg_context* ctx, const kg_settings* settings, const kg_digital_terms* option,
                           const kg_bsm_market* market, kg_estimate* out);
kg_status kg_digital_price_with_greeks(kg_context* ctx, const kg_settings* settings,
                                       const kg_digital_terms* option, const kg_bsm_market* market,
                                       kg_spot_greeks* greeks, kg_estimate* out);

/* ---- Average-price options: MonteCarloAsianEngine
 * (engines/monte_carlo/asian.rs:374-1257). Geometric averages sample the
 * exact one-dimensional lognormal (:741-784); arithmetic averages integrate
 * the continuous average by the trapezoid rule over step_count intervals
 * with a conditional last half interval and a geometric control variate
 * (:640-712, 1016-1061, 1100-1256). */
enum { KG_AVERAGE_ARITHMETIC = 0, KG_AVERAGE_GEOMETRIC = 1 };

typedef struct kg_asian_terms {
  int32_t option_type; /* 1 call, -1 put */
  int32_t averaging;   /* KG_AVERAGE_* */
  int32_t step_count;  /* integration intervals, 2..=10,000 */
  int32_t has_realized;
  double strike;
  double realized_average; /* average fixed before valuation, if any */
  int64_t effective_time;
  int64_t averaging_start_time;
  int64_t expiry_time;
} kg_asian_terms;

kg_status kg_asian_price(kg_context* ctx, const kg_settings* settings, const kg_asian_terms* option,
                         const kg_bsm_market* market, kg_estimate* out);

/* MonteCarloAsianEngine::price_with_greeks_with_settings (asian.rs:885-898
 * through engines.rs:138-162): the price with the engine's native delta and
 * gamma (greeks->request, KG_GREEK_* bits), the per-path conditional
 * estimators of GeometricParameters::geometric (asian.rs:741-784) and
 * ArithmeticParameters::arithmetic with its geometric control
 * (asian.rs:1063-1098, 1233-1255). greeks->available holds the Greeks
 * returned; at expiry (a known price) none are, and the caller's expiry
 * stencil supplies them. The price equals kg_asian_price's bit for bit. */
kg_status kg_asian_price_with_greeks(kg_context* ctx, const kg_settings* settings, const kg_asian_terms* option,
                                     const kg_bsm_market* market, kg_spot_greeks* greeks, kg_estimate* out);

/* Strike batch of an arithmetic average: MonteCarloAsianEngine::price_strikes
 * (engines/monte_carlo/asian.rs:197-345). prices[i] equals the price of
 * kg_asian_price for the option with strike strikes[i] and the same
 * settings, bit for bit, as the reference's batch equals its scalar prices
 * (crates/kiyosi/tests/asian_monte_carlo.rs:144-151). Distinct strikes are
 * priced once; up to 32 of them (the reference's STRIKE_TILE) share each
 * path of one launch. Strikes must be positive and finite. */
kg_status kg_asian_price_strikes(kg_context* ctx, const kg_settings* settings, const kg_asian_terms* option,
                                 const kg_bsm_market* market, const double* strikes, size_t n,
                                 double* prices);

/* ---- Accumulators and decumulators: MonteCarloAccumulatorEngine
 * (engines/monte_carlo/accumulator.rs:91-218; instrument
 * instruments/accumulator.rs:319-526). Steps at the trading-day fixings at
 * the observation clock from max(valuation, effective time), the explicit
 * knock-out times and expiry: knock-out settles the accumulated quantity at
 * the knock-out fixing, otherwise the daily quantity (times the acceleration
 * factor beyond the strike) accrues and settles at expiry. With an explicit
 * knock-out schedule and unit acceleration the purchases between knock-out
 * dates are fixed and grouped onto those dates (fixed_purchase_events,
 * :401-454). Valuation lies in [effective time, expiry]: the reference's
 * accumulator keeps the default valuation start (vanilla.rs:141-143,
 * monitoring.rs:470-481). The history follows validate_pricing (:319-358):
 * after accumulation or knock-out observations began, the settlement state
 * and the accumulated quantity must be known (KG_MISSING_HISTORY
 * otherwise). */
enum { KG_SETTLED_UNKNOWN = -1, KG_SETTLED_NO = 0, KG_SETTLED_YES = 1 };

typedef struct kg_accumulator_terms {
  int32_t option_type; /* 1 accumulator (call), -1 decumulator (put) */
  int32_t calendar;    /* KG_CALENDAR_* */
  double strike;
  double knock_out_level;
  double daily_quantity;
  double acceleration_factor;
  double accumulated_quantity; /* quantity accrued before valuation, if has_accumulated_quantity */
  int64_t observation_clock_micros;
  int64_t effective_time;
  int64_t expiry_time;
  int32_t has_accumulated_quantity;
  int32_t knock_out_settled;        /* KG_SETTLED_* */
  int32_t has_knock_out_schedule;   /* 0: daily knock-out at the observation clock */
  int32_t n_knock_out_observations; /* explicit knock-out times (the schedule may be empty) */
  const int64_t* knock_out_observation_times; /* strictly increasing trading days in the contract life */
} kg_accumulator_terms;

kg_status kg_accumulator_price(kg_context* ctx, const kg_settings* settings,
                               const kg_accumulator_terms* option, const kg_bsm_market* market,
                               kg_estimate* out);

/* ---- Heterogeneous book in one persistent launch. Every contract keeps
 * its own settings (path count, seed, sampling) and market, so each result
 * equals the single-contract entry point's on the same inputs. The kernel
 * pulls (contract, 256-sample block) tiles from a device queue, longest
 * contracts first; per-block partials are merged in block order. The
 * launch sequence is captured once in a hipGraph and replayed; tile shape
 * (samples per thread, workgroups per CU) comes from the autotuner unless
 * fixed in `tuning`. No counterpart in kiyosi, which prices one contract
 * per call (autocallable.rs:119-152 batches one note type only). */
enum {
  KG_BOOK_BARRIER = 0,
  KG_BOOK_ASIAN = 1,
  KG_BOOK_AUTOCALL = 2,
  KG_BOOK_ACCUMULATOR = 3,
  KG_BOOK_DOUBLE_BARRIER = 4,
  KG_BOOK_DIGITAL = 5
};

typedef struct kg_book_contract {
  int32_t kind; /* KG_BOOK_* */
  int32_t reserved;
  kg_settings settings;
  kg_bsm_market market;
  const void* terms; /* kg_barrier_terms, kg_asian_terms, kg_autocall_terms, kg_accumulator_terms,
                        kg_double_barrier_terms or kg_digital_terms, by kind */
} kg_book_contract;

typedef struct kg_book_tuning {
  int32_t samples_per_thread;  /* 1, 2 or 4; 0 = autotune */
  int32_t workgroups_per_cu;   /* 1..16; 0 = autotune */
  int32_t use_graph;           /* capture and replay the launch as a hipGraph */
  int32_t replays;             /* extra timed replays of the captured graph (benchmarking) */
  double device_ms;            /* out: mean device time of one book launch */
  double path_steps;           /* out: simulated path-steps of the book */
} kg_book_tuning;

kg_status kg_price_book(kg_context* ctx, const kg_book_contract* contracts, size_t n,
                        int32_t precision, kg_book_tuning* tuning, kg_estimate* out);

/* ---- American options by least-squares Monte Carlo:
 * MonteCarloVanillaEngine::price for AmericanOption (in-sample,
 * engines/monte_carlo/american.rs:955-1455) and
 * price_american_out_of_sample (:574-637).
 *
 * Model: flat BSM (surface == NULL) follows the reference exactly:
 * normalised Brownian coordinates per path, spot = S0 e^{drift (k+1)}
 * e^{diffusion W_k}, block-ordered normal equations. A local-volatility
 * surface (LocalVolatilitySurface, market/local_volatility.rs:61-359:
 * bilinear in Actual/365 years from the first row and in spot, spot clamped
 * to the grid) extends the engine: log-Euler steps with sigma(S_k, t_k)
 * taken at the start of each exercise interval.
 *
 * Regressors: KG_BASIS_QUADRATIC is the reference's (1, S/K, (S/K)^2) with
 * Gaussian elimination and degree fallback (:1693-1749); KG_BASIS_POLY4 uses
 * (1 .. (S/K)^4) with an FP64 Cholesky solve; KG_BASIS_MLP fits a
 * 3-64-64-1 GELU network on (S/K, t, sigma_loc) per exercise date with
 * bf16 MFMA forward/backward and Adam, warm-started from the later date. */
enum { KG_BASIS_QUADRATIC = 0, KG_BASIS_POLY4 = 1, KG_BASIS_MLP = 2 };

typedef struct kg_local_vol_surface {
  int32_t n_times; /* >= 2 rows */
  int32_t n_spots; /* >= 2 columns */
  const int64_t* times;  /* strictly increasing row times */
  const double* spots;   /* strictly increasing positive spot nodes */
  const double* values;  /* [n_times][n_spots] positive volatilities */
} kg_local_vol_surface;

typedef struct kg_american_terms {
  int32_t option_type; /* 1 call, -1 put */
  int32_t step_count;  /* exercise intervals, >= 2 */
  double strike;
  int64_t effective_time;
  int64_t expiry_time;
} kg_american_terms;

typedef struct kg_american_model {
  double spot;
  double rate;
  double dividend;
  double volatility; /* BSM volatility when surface is NULL */
  int64_t valuation_time;
  const kg_local_vol_surface* surface;
} kg_american_model;

typedef struct kg_lsmc_options {
  int32_t basis;               /* KG_BASIS_* */
  int32_t out_of_sample;       /* 0: in-sample price on the training paths */
  int64_t training_path_count; /* training paths for out-of-sample pricing */
  int32_t epochs;              /* MLP: passes over each date's samples (default 4) */
  int32_t batch;               /* MLP: minibatch size (default 65536) */
  double learning_rate;        /* MLP: Adam step (default 2e-3) */
} kg_lsmc_options;

typedef struct kg_lsmc_report {
  double price;          /* the engine's price: in-sample, or out-of-sample when requested */
  double standard_error; /* of the out-of-sample estimate (pseudorandom sampling) */
  double in_sample_price;
  double training_ms;    /* device time of path generation and regression */
  double pricing_ms;     /* device time of the out-of-sample pass */
  double path_steps;     /* simulated path-steps (training + pricing) */
  double mean_loss;      /* MLP: mean squared error of the last epoch over all dates */
  int32_t dates_fitted;  /* exercise dates with a fitted continuation value */
  int32_t reserved;
  double gradient_samples; /* MLP: training samples through the gradient kernel, over all Adam steps */
} kg_lsmc_report;

kg_status kg_american_price(kg_context* ctx, const kg_settings* settings,
                            const kg_american_terms* option, const kg_american_model* model,
                            const kg_lsmc_options* options, kg_lsmc_report* out);

/* Lanes of one American option on shared paths: lane l is `option` with
 * strike strikes[l] (option->strike when strikes is NULL) under models[l],
 * and out[l] equals kg_american_price of that lane alone, bit for bit.
 * Replaces american_scenarios::<N, S> (engines/monte_carlo/american.rs:955-1455)
 * as called by MonteCarloVanillaEngine::price_strikes (:677-771, groups of
 * 4/2/1 strikes) and by the American Greek stencil (greek_prices and
 * greek_scenario_group, :775-931, scenario groups at one valuation time).
 * In-sample polynomial fits share one set of stored paths when every lane
 * has the same valuation time and normalised paths, and either all lanes
 * are BSM (spot, rate, dividend, volatility and strike may differ: the
 * Brownian coordinates are model-free) or all share one local-volatility
 * market (strikes differ); one generation and one backward pass then fit
 * every lane, with a regression per lane. Other requests (MLP,
 * out-of-sample, anchored paths, mixed markets) price each lane alone.
 * Every lane is validated first; a failed request fills no lane. */
#define KG_MAX_AMERICAN_LANES 16
kg_status kg_american_price_lanes(kg_context* ctx, const kg_settings* settings,
                                  const kg_american_terms* option, const kg_american_model* models,
                                  const double* strikes, int32_t n_lanes, const kg_lsmc_options* options,
                                  kg_lsmc_report* out);

/* ---- Generator diagnostics, computed on the GPU with per-thread skip-ahead.
 * kg_pcg32_outputs: outputs first..first+count of Pcg32::new(seed, stream)
 * (monte_carlo.rs:2247-2267). kg_sobol_points: the 52-bit scrambled Sobol
 * point bits of points first..first+count for every dimension, point-major
 * (out[(p - first) * dims + d]), as SobolSequence::cursor yields them
 * (engines/monte_carlo/sobol.rs:136-243). kg_normals: normals of a block's
 * stream as NormalDraws yields them (monte_carlo.rs:1678-1703). */
kg_status kg_pcg32_outputs(kg_context* ctx, uint64_t seed, uint64_t stream, uint64_t first,
                           uint64_t count, uint32_t* out);
kg_status kg_sobol_points(kg_context* ctx, uint32_t dims, uint64_t seed, uint64_t first,
                          uint64_t count, uint64_t* out);
kg_status kg_pcg_normals(kg_context* ctx, uint64_t seed, uint64_t stream, uint64_t first,
                         uint64_t count, double* out);
/* The single-precision inverse normal of the FP32-step kernels applied to
 * (bits[i] + 1/2) / 2^32; no reference counterpart (the reference maps Sobol
 * bits through Cephes ndtri in FP64, engines/monte_carlo/sobol.rs:249-346). */
kg_status kg_fp32_normals_from_bits(kg_context* ctx, const uint32_t* bits, uint64_t count, float* out);

/* ---- MLP regressor diagnostics (no reference counterpart: they expose the
 * kernels of KG_BASIS_MLP so their gradients and Adam steps can be checked
 * against a CPU restatement). `params` holds the 4,481 FP32 parameters
 * W1[64][3], b1[64], W2[64 out][64 in], b2[64], W3[64], b3; `samples` holds n
 * rows (u0, u1, u2, y) of FP32.
 * kg_mlp_gradient: gradient of sum((out - y)^2) / n and that loss.
 * kg_mlp_train: `steps` Adam steps (beta 0.9/0.999, eps 1e-8, moments from
 * zero) over consecutive batches of `batch` rows, wrapping to row 0 when a
 * batch reaches the end; updates params in place.
 * kg_mlp_forward: the network output of every row. */
kg_status kg_mlp_gradient(kg_context* ctx, const float* params, const float* samples, int64_t n,
                          float* gradient, double* loss);
kg_status kg_mlp_train(kg_context* ctx, float* params, const float* samples, int64_t n, int32_t batch,
                       int32_t steps, double learning_rate);
kg_status kg_mlp_forward(kg_context* ctx, const float* params, const float* samples, int64_t n,
                         float* out);

#ifdef __cplusplus
}
#endif
#endif /* KIYOSI_AMD_H */
