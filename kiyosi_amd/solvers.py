"""Implied-parameter solvers on pinned Monte Carlo draws.

Restates kiyosi's implied_volatility (crates/kiyosi/src/risk/implied.rs
:84-248) and implied_snowball_coupon / implied_phoenix_coupon
(risk/coupon.rs:106-330) for the GPU engines. Every evaluation prices with
the engine's explicit seed, so all trials see the same draws and the price
is a smooth function of the parameter; each trial is one GPU pricing call,
and the solvers take the same trial sequence as the reference's.

The coupon solvers use the Monte Carlo autocallable engine's coupon hooks
(engines/monte_carlo/autocallable.rs:1016-1062): on fixed draws the price
is affine in the coupon, so the two bracket notes are priced as lanes of one
launch (price_many) and later trials interpolate between them; tolerances
tighter than 1e-10 of the price fall back to independent pricing of each
trial.
"""

import copy
import math

from .errors import (ComputationFailureError, InvalidInputError, NonConvergenceError, RootNotBracketedError,
                     UnsupportedOperationError)
from .market import BlackScholesMertonParameters, to_micros


class ImpliedVolatilitySettings:
    """Bracket and tolerances of implied_volatility (risk/implied.rs:8-40)."""

    DEFAULT_PRICE_TOLERANCE = 1e-8
    DEFAULT_PARAMETER_TOLERANCE = 1e-8
    DEFAULT_MAX_ITERATIONS = 100

    def __init__(self, lower_bound, upper_bound, price_tolerance=1e-8, parameter_tolerance=1e-8, max_iterations=100):
        self.lower_bound, self.upper_bound = lower_bound, upper_bound
        self.price_tolerance, self.parameter_tolerance = price_tolerance, parameter_tolerance
        self.max_iterations = max_iterations


class ImpliedCouponSettings:
    """Bracket and tolerances of the coupon solvers; defaults [0, 2]
    (risk/coupon.rs:32-50)."""

    def __init__(self, lower_bound=0.0, upper_bound=2.0, price_tolerance=1e-8, parameter_tolerance=1e-8,
                 max_iterations=100):
        self.lower_bound, self.upper_bound = lower_bound, upper_bound
        self.price_tolerance, self.parameter_tolerance = price_tolerance, parameter_tolerance
        self.max_iterations = max_iterations


class SnowballCouponQuote:
    """The first observation's coupon that reproduces the price, and the
    convention that moves the other coupons with it (risk/coupon.rs:15-30)."""

    def __init__(self, first_observation_coupon, quote_convention):
        self.first_observation_coupon = first_observation_coupon
        self.quote_convention = quote_convention

    def __repr__(self):
        return (f"SnowballCouponQuote(first_observation_coupon={self.first_observation_coupon!r}, "
                f"quote_convention={self.quote_convention!r})")


SHIFT_MATURITY_COUPON = "shift_maturity_coupon"
PRESERVE_MATURITY_COUPON = "preserve_maturity_coupon"


def _finite(name, value, allow_zero_negative=True):
    if not math.isfinite(value):
        raise InvalidInputError(f"{name} must be finite")


def _validate(observed, s):
    _finite("observed_price", observed)
    _finite("lower_bound", s.lower_bound)
    _finite("upper_bound", s.upper_bound)
    for name in ("price_tolerance", "parameter_tolerance"):
        value = getattr(s, name)
        if not (math.isfinite(value) and value >= 0.0):
            raise InvalidInputError(f"{name} must be finite and nonnegative")
    # validate_solver_bounds (crates/kiyosi/src/risk.rs:6-18)
    if s.lower_bound >= s.upper_bound:
        raise InvalidInputError("upper_bound must exceed lower_bound")
    if s.max_iterations <= 0:
        raise InvalidInputError("max_iterations must be positive")


def _validate_solver_seed(engine):
    """MonteCarloSettings::validate_solver_seed (monte_carlo.rs:154-161)."""
    if getattr(engine, "seed", 0) is None:
        raise InvalidInputError("implied-parameter solvers require an explicit seed; call with_resolved_seed() "
                                "before solving")


def _volatility_affects_price(engine, option, context):
    """OptionContract::volatility_affects_price, through the instrument's own
    hook where it has one (barriers); an autocallable note is unaffected when
    its value is settled on the host; vanilla, digital, Asian and accumulator
    contracts accepted by the GPU engines are affected."""
    hook = getattr(option, "volatility_affects_price", None)
    if hook is not None:
        return bool(hook(context))
    if hasattr(option, "_knock_in_affects_payoff"):
        return engine.price_with_statistics(option, context).seed is not None
    return True


def implied_volatility(engine, option, context, observed_price, settings):
    """The Black-Scholes volatility at which `engine` prices `option` at
    `observed_price`: Brent's method with the reference's safeguards
    (risk/implied.rs:121-248) on the engine's pinned seed."""
    if not isinstance(context.model_parameters, BlackScholesMertonParameters):
        raise UnsupportedOperationError("implied volatility needs Black-Scholes-Merton parameters")
    s = settings
    _validate(observed_price, s)
    if s.lower_bound < 0.0:
        raise InvalidInputError("lower_bound must be nonnegative")
    _validate_solver_seed(engine)
    model = context.model_parameters

    def evaluate(volatility):
        shifted = context.replace(model_parameters=BlackScholesMertonParameters(
            risk_free_rate=model.risk_free_rate, dividend_yield=model.dividend_yield, volatility=volatility))
        residual = engine.price(option, shifted) - observed_price
        if not math.isfinite(residual):
            raise ComputationFailureError("non-finite solver evaluation")
        return residual

    def require_identifiable():
        if (to_micros(context.valuation_time) == to_micros(option.expiry_time)
                or not _volatility_affects_price(engine, option, context)):
            raise UnsupportedOperationError("volatility does not affect the remaining payoff")

    zero = evaluate(0.0) if s.lower_bound == 0.0 else None
    if zero is not None and abs(zero) <= s.price_tolerance:
        require_identifiable()
        return 0.0
    a, b = s.lower_bound, s.upper_bound
    fa = zero if (zero is not None and a == 0.0) else evaluate(a)
    fb = fa if a == b else evaluate(b)
    require_identifiable()
    if abs(fa) <= s.price_tolerance:
        return a
    if abs(fb) <= s.price_tolerance:
        return b
    if (fa < 0.0) == (fb < 0.0):
        raise RootNotBracketedError("price is not bracketed by volatility bounds")
    if abs(fa) < abs(fb):
        a, b, fa, fb = b, a, fb, fa
    c, fc, d, bisected = a, fa, a, True
    for _ in range(s.max_iterations):
        scale = max(abs(fa), abs(fb), abs(fc))
        ya, yb, yc = fa / scale, fb / scale, fc / scale
        trial = _ieee(lambda: a + (b - a) * (ya / (ya - yb)))
        if fa != fc and fb != fc:
            trial = _ieee(lambda: a * (yb / (ya - yb)) * (yc / (ya - yc)) + b * (ya / (yb - ya)) * (yc / (yb - yc))
                          + c * (ya / (yc - ya)) * (yb / (yc - yb)))
        quarter = a + (b - a) * 0.25
        if (not math.isfinite(trial) or trial <= min(quarter, b) or trial >= max(quarter, b)
                or (bisected and abs(trial - b) >= 0.5 * abs(b - c))
                or (not bisected and abs(trial - b) >= 0.5 * abs(c - d))
                or (bisected and abs(b - c) <= s.parameter_tolerance)
                or (not bisected and abs(c - d) <= s.parameter_tolerance)):
            trial = a + (b - a) * 0.5
            bisected = True
        else:
            bisected = False
        if trial == a or trial == b:
            raise NonConvergenceError("volatility bracket has no representable interior trial")
        residual = evaluate(trial)
        if abs(residual) <= s.price_tolerance:
            return trial
        if abs(b - a) <= s.parameter_tolerance:
            raise NonConvergenceError("volatility bracket converged without matching the observed price")
        d, c, fc = c, b, fb
        if (fa < 0.0) == (residual < 0.0):
            a, fa = trial, residual
        else:
            b, fb = trial, residual
        if abs(fa) < abs(fb):
            a, b, fa, fb = b, a, fb, fa
    raise NonConvergenceError("implied-volatility solver did not converge")


def _ieee(formula):
    """A trial formula's value, or NaN where IEEE arithmetic (as in Rust)
    would divide by zero; the safeguards then bisect."""
    try:
        return formula()
    except ZeroDivisionError:
        return math.nan


def _midpoint(a, b):
    """f64::midpoint: (a + b) / 2, halving first where the sum could overflow."""
    lo, hi = 2.0 * 2.2250738585072014e-308, 1.7976931348623157e308 / 2.0
    if abs(a) <= hi and abs(b) <= hi:
        return (a + b) / 2.0
    if abs(a) < lo:
        return a + b / 2.0
    if abs(b) < lo:
        return a / 2.0 + b
    return a / 2.0 + b / 2.0


def _snowball_with_coupon(note, coupon, convention):
    """SnowballCoupon::with_coupon (instruments/autocallable.rs:647-660):
    every observation coupon moves by the shift of the first; with
    ShiftMaturityCoupon the maturity coupon moves too."""
    from .instruments import SnowballObservation
    out = copy.copy(note)
    shift = coupon - (note.autocall_observations[0].coupon_rate if note.autocall_observations else 0.0)
    out.autocall_observations = tuple(SnowballObservation(o.time, o.knock_out_level, o.coupon_rate + shift)
                                      for o in note.autocall_observations)
    if convention == SHIFT_MATURITY_COUPON:
        out.maturity_coupon_rate = note.maturity_coupon_rate + shift
    return out


def _minimum_quoted_coupon(note, convention):
    """AutocallableNote::minimum_quoted_coupon (instruments/autocallable.rs
    :667-693): under ShiftMaturityCoupon a fixed-coupon knock-in redemption
    bounds the quote from below, rounded into with_coupon's domain."""
    from . import _abi
    if convention != SHIFT_MATURITY_COUPON or note.knock_in_level is None:
        return None
    if note._redemption_kind != _abi.REDEMPTION_FIXED_COUPON:
        return None
    minimum, anchor = note.minimum_coupon_rate, note.autocall_observations[0].coupon_rate
    maturity = note.maturity_coupon_rate
    shift = minimum - maturity
    if maturity + shift < minimum:
        shift = math.nextafter(shift, math.inf)
    lower = anchor + shift
    return math.nextafter(lower, math.inf) if lower - anchor < shift else lower


def _phoenix_with_coupon(note, coupon):
    """PhoenixOption::with_coupon (instruments/autocallable.rs:1902-1907)."""
    out = copy.copy(note)
    out.coupon_per_period = coupon
    return out


def implied_snowball_coupon(engine, note, context, observed_price, convention=SHIFT_MATURITY_COUPON,
                            settings=None):
    """The first observation's coupon of a snowball (and of its variants)
    that `engine` prices at `observed_price` (risk/coupon.rs:106-138)."""
    if convention not in (SHIFT_MATURITY_COUPON, PRESERVE_MATURITY_COUPON):
        raise InvalidInputError("unknown coupon quote convention")
    coupon = _solve_coupon(engine, note, context, observed_price, settings or ImpliedCouponSettings(),
                           _minimum_quoted_coupon(note, convention),
                           lambda n, c: _snowball_with_coupon(n, c, convention))
    return SnowballCouponQuote(coupon, convention)


def implied_phoenix_coupon(engine, note, context, observed_price, settings=None):
    """The per-period coupon of a Phoenix note that `engine` prices at
    `observed_price` (risk/coupon.rs:169-189)."""
    return _solve_coupon(engine, note, context, observed_price, settings or ImpliedCouponSettings(), None,
                         _phoenix_with_coupon)


def _coupon_bounds(engine, lower_note, upper_note, context):
    """price_coupon_bounds and prepare_coupon_pricer of the Monte Carlo
    autocallable engine (autocallable.rs:1016-1062): both notes priced on the
    same draws as lanes of one launch. Returns (lower, upper, affine) where
    affine says whether later trials may interpolate (both notes sampled, and
    knock-in affects both payoffs alike)."""
    low, high = engine.price_many([lower_note, upper_note], context)
    sampled = all(engine.price_with_statistics(n, context).seed is not None for n in (lower_note, upper_note))
    affine = sampled and lower_note._knock_in_affects_payoff() == upper_note._knock_in_affects_payoff()
    return low, high, affine


def _solve_coupon(engine, note, context, observed, s, minimum, replace):
    """risk/coupon.rs:191-330."""
    _validate(observed, s)
    lo_bound = s.lower_bound
    if minimum is not None:
        lo_bound = max(lo_bound, minimum)
        if lo_bound >= s.upper_bound:
            raise RootNotBracketedError("coupon bounds exclude the contract's valid domain")
    _validate_solver_seed(engine)

    def checked(price):
        if not math.isfinite(price):
            raise ComputationFailureError("non-finite coupon price")
        return price

    def residual(value):
        r = value - observed
        if not math.isfinite(r):
            raise ComputationFailureError("non-finite coupon residual")
        return r

    lo, hi = lo_bound, s.upper_bound
    lower_note, upper_note = replace(note, lo), replace(note, hi)
    original_lo, original_hi = lo, hi
    lower, upper, affine = _coupon_bounds(engine, lower_note, upper_note, context)
    lower, upper = checked(lower), checked(upper)
    # Tight tolerances price every trial on its own (rounding of the rolled-back cash flows).
    curve = affine and s.price_tolerance >= 1e-10 * max(abs(observed), 1.0) and all(
        s.price_tolerance >= 1e-10 * max(abs(v), 1.0) for v in (lower, upper))
    if lower == upper:
        raise UnsupportedOperationError("coupon does not affect remaining cashflows")
    flo, fhi = residual(lower), residual(upper)
    if abs(flo) <= s.price_tolerance:
        return lo
    if abs(fhi) <= s.price_tolerance:
        return hi
    if (flo < 0.0) == (fhi < 0.0):
        raise RootNotBracketedError("price is not bracketed by coupon bounds")

    def fraction_of(value, low, high):
        span = high - low
        if math.isfinite(span):
            return (value - low) / span
        return (0.5 * value - 0.5 * low) / (0.5 * high - 0.5 * low)

    def price(coupon):
        trial = replace(note, coupon)
        if curve and trial._knock_in_affects_payoff() == lower_note._knock_in_affects_payoff():
            f = fraction_of(coupon, original_lo, original_hi)
            return checked(lower * (1.0 - f) + upper * f)
        return checked(engine.price(trial, context))

    f = fraction_of(observed, lower, upper)
    coupon = lo + (hi - lo) * f if (lo < 0.0) == (hi < 0.0) else lo * (1.0 - f) + hi * f
    if lo < coupon < hi and abs(residual(price(coupon))) <= s.price_tolerance:
        return coupon
    for _ in range(s.max_iterations):
        mid = _midpoint(lo, hi)
        if mid == lo or mid == hi:
            raise NonConvergenceError("coupon bracket has no representable midpoint")
        value = residual(price(mid))
        if abs(value) <= s.price_tolerance:
            return mid
        if hi - lo <= s.parameter_tolerance:
            raise NonConvergenceError("coupon bracket converged without matching the observed price")
        if (flo < 0.0) == (value < 0.0):
            lo, flo = mid, value
        else:
            hi = mid
    raise NonConvergenceError("implied-coupon solver did not converge")
