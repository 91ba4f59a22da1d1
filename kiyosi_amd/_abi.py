"""ctypes mirror of the structs and enums in include/kiyosi_amd.h."""

import ctypes as C

OK, INVALID_INPUT, UNSUPPORTED_OPERATION, COMPUTATION_FAILURE = 0, 1, 2, 3
INTERRUPTED, DEVICE_ERROR, MISSING_HISTORY, NON_CONVERGENCE = 4, 5, 6, 7

PSEUDO_RANDOM, SOBOL = 0, 1
FP64, FP32_STEP = 0, 1

SETTLED_UNKNOWN, SETTLED_NO, SETTLED_YES = -1, 0, 1

STATUS_AVAILABLE, STATUS_UNSUPPORTED_SAMPLING, STATUS_INSUFFICIENT_OBSERVATIONS = 0, 1, 2
STATUS_INCOMPLETE_ANTITHETIC_PAIR, STATUS_ZERO_OBSERVED_VARIANCE = 3, 4

UNCERTAINTY_NAMES = {
    0: "available",
    1: "unsupported_sampling",
    2: "insufficient_observations",
    3: "incomplete_antithetic_pair",
    4: "zero_observed_variance",
}

CALENDAR_WEEKDAYS, CALENDAR_ALL_DAYS = 0, 1
KNOCK_IN_NONE, KNOCK_IN_DAILY, KNOCK_IN_AT_EXPIRY = 0, 1, 2
REDEMPTION_DOWNSIDE, REDEMPTION_FIXED_COUPON = 0, 1
COUPON_SNOWBALL, COUPON_PHOENIX = 0, 1
STATE_UNKNOWN, STATE_NO_EVENT, STATE_KNOCKED_OUT, STATE_KNOCKED_IN = -1, 0, 1, 2


class Settings(C.Structure):
    _fields_ = [
        ("path_count", C.c_int64),
        ("seed", C.c_uint64),
        ("has_seed", C.c_int32),
        ("sampling", C.c_int32),
        ("antithetic", C.c_int32),
        ("precision", C.c_int32),
        ("max_active_workers", C.c_int32),
        ("allow_large", C.c_int32),
    ]


class Moments(C.Structure):
    _fields_ = [
        ("count", C.c_uint64),
        ("scale", C.c_double),
        ("mean", C.c_double),
        ("m2", C.c_double),
    ]


class Estimate(C.Structure):
    _fields_ = [
        ("price", C.c_double),
        ("standard_error", C.c_double),
        ("ci_lower", C.c_double),
        ("ci_upper", C.c_double),
        ("has_standard_error", C.c_int32),
        ("uncertainty_status", C.c_int32),
        ("has_seed", C.c_int32),
        ("reserved", C.c_int32),
        ("seed", C.c_uint64),
        ("path_standard_error", C.c_double),
        ("path_steps", C.c_double),
        ("device_ms", C.c_double),
    ]


class BsmMarket(C.Structure):
    _fields_ = [
        ("spot", C.c_double),
        ("rate", C.c_double),
        ("dividend", C.c_double),
        ("volatility", C.c_double),
        ("valuation_time", C.c_int64),
    ]


class EuropeanTerms(C.Structure):
    _fields_ = [
        ("option_type", C.c_int32),
        ("reserved", C.c_int32),
        ("strike", C.c_double),
        ("effective_time", C.c_int64),
        ("expiry_time", C.c_int64),
    ]


class AutocallEvent(C.Structure):
    _fields_ = [
        ("time", C.c_int64),
        ("has_knock_out", C.c_int32),
        ("has_coupon_barrier", C.c_int32),
        ("knock_out_level", C.c_double),
        ("coupon_barrier_level", C.c_double),
        ("coupon_rate", C.c_double),
    ]


class AutocallTerms(C.Structure):
    _fields_ = [
        ("calendar", C.c_int32),
        ("knock_in_kind", C.c_int32),
        ("redemption_kind", C.c_int32),
        ("coupon_kind", C.c_int32),
        ("barrier_state", C.c_int32),
        ("n_events", C.c_int32),
        ("observation_clock_micros", C.c_int64),
        ("knock_in_level", C.c_double),
        ("initial_spot", C.c_double),
        ("upper_strike", C.c_double),
        ("lower_strike", C.c_double),
        ("fixed_coupon_rate", C.c_double),
        ("coupon", C.c_double),
        ("principal_redemption_ratio", C.c_double),
        ("rebate_rate", C.c_double),
        ("annual_rebate_rate", C.c_double),
        ("effective_time", C.c_int64),
        ("expiry_time", C.c_int64),
        ("events", C.POINTER(AutocallEvent)),
    ]


HESTON_EUROPEAN, HESTON_ASIAN_ARITHMETIC, HESTON_ASIAN_GEOMETRIC = 0, 1, 2


class HestonMarket(C.Structure):
    _fields_ = [
        ("spot", C.c_double),
        ("rate", C.c_double),
        ("dividend", C.c_double),
        ("v0", C.c_double),
        ("kappa", C.c_double),
        ("theta", C.c_double),
        ("xi", C.c_double),
        ("rho", C.c_double),
        ("valuation_time", C.c_int64),
    ]


class HestonTerms(C.Structure):
    _fields_ = [
        ("payoff", C.c_int32),
        ("option_type", C.c_int32),
        ("step_count", C.c_int32),
        ("reserved", C.c_int32),
        ("strike", C.c_double),
        ("effective_time", C.c_int64),
        ("expiry_time", C.c_int64),
    ]


DOWN_AND_OUT, DOWN_AND_IN, UP_AND_OUT, UP_AND_IN = 0, 1, 2, 3
OBSERVE_CONTINUOUS, OBSERVE_SCHEDULED = 0, 1
REBATE_AT_EXPIRY, REBATE_AT_HIT = 0, 1
AVERAGE_ARITHMETIC, AVERAGE_GEOMETRIC = 0, 1
BOOK_BARRIER, BOOK_ASIAN, BOOK_AUTOCALL, BOOK_ACCUMULATOR, BOOK_DOUBLE_BARRIER, BOOK_DIGITAL = 0, 1, 2, 3, 4, 5
GREEK_DELTA, GREEK_GAMMA = 1, 2
DOUBLE_KNOCK_OUT, DOUBLE_KNOCK_IN = 0, 1
TOUCH_UNKNOWN, TOUCH_UNTOUCHED, TOUCH_TOUCHED = -1, 0, 1
CASH_OR_NOTHING, ASSET_OR_NOTHING = 0, 1


class BarrierTerms(C.Structure):
    _fields_ = [
        ("option_type", C.c_int32),
        ("barrier_type", C.c_int32),
        ("observation", C.c_int32),
        ("rebate_timing", C.c_int32),
        ("strike", C.c_double),
        ("barrier", C.c_double),
        ("rebate", C.c_double),
        ("effective_time", C.c_int64),
        ("expiry_time", C.c_int64),
        ("n_observations", C.c_int32),
        ("touch_state", C.c_int32),
        ("observation_times", C.POINTER(C.c_int64)),
    ]


class SpotGreeks(C.Structure):
    _fields_ = [
        ("request", C.c_int32),
        ("available", C.c_int32),
        ("spot_radius", C.c_double),
        ("delta", C.c_double),
        ("gamma", C.c_double),
    ]


class DoubleBarrierTerms(C.Structure):
    _fields_ = [
        ("option_type", C.c_int32),
        ("knock_type", C.c_int32),
        ("observation", C.c_int32),
        ("rebate_timing", C.c_int32),
        ("strike", C.c_double),
        ("lower_barrier", C.c_double),
        ("upper_barrier", C.c_double),
        ("rebate", C.c_double),
        ("effective_time", C.c_int64),
        ("expiry_time", C.c_int64),
        ("n_observations", C.c_int32),
        ("touch_state", C.c_int32),
        ("observation_times", C.POINTER(C.c_int64)),
    ]


class DigitalTerms(C.Structure):
    _fields_ = [
        ("option_type", C.c_int32),
        ("payout_kind", C.c_int32),
        ("strike", C.c_double),
        ("cash_payout", C.c_double),
        ("effective_time", C.c_int64),
        ("expiry_time", C.c_int64),
    ]


class AsianTerms(C.Structure):
    _fields_ = [
        ("option_type", C.c_int32),
        ("averaging", C.c_int32),
        ("step_count", C.c_int32),
        ("has_realized", C.c_int32),
        ("strike", C.c_double),
        ("realized_average", C.c_double),
        ("effective_time", C.c_int64),
        ("averaging_start_time", C.c_int64),
        ("expiry_time", C.c_int64),
    ]


class AccumulatorTerms(C.Structure):
    _fields_ = [
        ("option_type", C.c_int32),
        ("calendar", C.c_int32),
        ("strike", C.c_double),
        ("knock_out_level", C.c_double),
        ("daily_quantity", C.c_double),
        ("acceleration_factor", C.c_double),
        ("accumulated_quantity", C.c_double),
        ("observation_clock_micros", C.c_int64),
        ("effective_time", C.c_int64),
        ("expiry_time", C.c_int64),
        ("has_accumulated_quantity", C.c_int32),
        ("knock_out_settled", C.c_int32),
        ("has_knock_out_schedule", C.c_int32),
        ("n_knock_out_observations", C.c_int32),
        ("knock_out_observation_times", C.POINTER(C.c_int64)),
    ]


class BookContract(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("reserved", C.c_int32),
        ("settings", Settings),
        ("market", BsmMarket),
        ("terms", C.c_void_p),
    ]


class BookTuning(C.Structure):
    _fields_ = [
        ("samples_per_thread", C.c_int32),
        ("workgroups_per_cu", C.c_int32),
        ("use_graph", C.c_int32),
        ("replays", C.c_int32),
        ("device_ms", C.c_double),
        ("path_steps", C.c_double),
    ]


BASIS_QUADRATIC, BASIS_POLY4, BASIS_MLP = 0, 1, 2
MAX_SCENARIOS = 64  # KG_MAX_SCENARIOS
MAX_AUTOCALL_SCENARIOS = 16  # KG_MAX_AUTOCALL_SCENARIOS
MAX_AMERICAN_LANES = 16  # KG_MAX_AMERICAN_LANES


class LocalVolSurface(C.Structure):
    _fields_ = [
        ("n_times", C.c_int32),
        ("n_spots", C.c_int32),
        ("times", C.POINTER(C.c_int64)),
        ("spots", C.POINTER(C.c_double)),
        ("values", C.POINTER(C.c_double)),
    ]


class AmericanTerms(C.Structure):
    _fields_ = [
        ("option_type", C.c_int32),
        ("step_count", C.c_int32),
        ("strike", C.c_double),
        ("effective_time", C.c_int64),
        ("expiry_time", C.c_int64),
    ]


class AmericanModel(C.Structure):
    _fields_ = [
        ("spot", C.c_double),
        ("rate", C.c_double),
        ("dividend", C.c_double),
        ("volatility", C.c_double),
        ("valuation_time", C.c_int64),
        ("surface", C.POINTER(LocalVolSurface)),
    ]


class LsmcOptions(C.Structure):
    _fields_ = [
        ("basis", C.c_int32),
        ("out_of_sample", C.c_int32),
        ("training_path_count", C.c_int64),
        ("epochs", C.c_int32),
        ("batch", C.c_int32),
        ("learning_rate", C.c_double),
    ]


class LsmcReport(C.Structure):
    _fields_ = [
        ("price", C.c_double),
        ("standard_error", C.c_double),
        ("in_sample_price", C.c_double),
        ("training_ms", C.c_double),
        ("pricing_ms", C.c_double),
        ("path_steps", C.c_double),
        ("mean_loss", C.c_double),
        ("dates_fitted", C.c_int32),
        ("reserved", C.c_int32),
        ("gradient_samples", C.c_double),
    ]
