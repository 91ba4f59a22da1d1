"""Market inputs: time conversion, trading calendars and pricing contexts.

Mirrors the parts of crates/kiyosi/src/{time,market}/ that the Monte Carlo
engines read: Actual/365 Fixed year fractions at microsecond resolution
(time/date.rs:255-257), the weekdays and all-days calendars
(time/calendar.rs:146-161) and a flat-rate BSM ``PricingContext``
(market/parameters.rs:314-518).
"""

import datetime as _dt
import math

from .errors import InvalidInputError

_EPOCH = _dt.datetime(1970, 1, 1)
MICROS_PER_DAY = 86_400_000_000


def to_micros(value) -> int:
    """Microseconds since 1970-01-01T00:00 of a naive (UTC) datetime."""
    if isinstance(value, int):
        return value
    if isinstance(value, _dt.datetime):
        if value.tzinfo is not None:
            value = value.astimezone(_dt.timezone.utc).replace(tzinfo=None)
        delta = value - _EPOCH
        return (delta.days * 86_400 + delta.seconds) * 1_000_000 + delta.microseconds
    if isinstance(value, _dt.date):
        return to_micros(_dt.datetime(value.year, value.month, value.day))
    raise InvalidInputError(f"expected a datetime, got {type(value).__name__}")


def clock_micros(value) -> int:
    """Microseconds since midnight of a datetime.time."""
    if value is None:
        return 0
    return ((value.hour * 60 + value.minute) * 60 + value.second) * 1_000_000 + value.microsecond


def years_until(start, end) -> float:
    return (to_micros(end) - to_micros(start)) / (365.0 * MICROS_PER_DAY)


class TradingCalendar:
    """Trading-day predicate: ``weekdays()`` (basis 261) or ``all_days()`` (basis 365)."""

    WEEKDAYS, ALL_DAYS = 0, 1

    def __init__(self, kind: int = 0):
        self.kind = kind

    @classmethod
    def weekdays(cls):
        return cls(cls.WEEKDAYS)

    @classmethod
    def all_days(cls):
        return cls(cls.ALL_DAYS)

    def trading_days_per_year(self) -> int:
        return 261 if self.kind == self.WEEKDAYS else 365

    def is_trading_day(self, date) -> bool:
        return self.kind == self.ALL_DAYS or date.weekday() < 5

    def is_trading_day_number(self, day: int) -> bool:
        """Trading-day test of day `day` since 1970-01-01 (a Thursday)."""
        return self.kind == self.ALL_DAYS or (day + 3) % 7 < 5

    def add_trading_days(self, start, days: int):
        """The date `days` trading days after `start` (a datetime keeps its clock)."""
        current = start
        while days > 0:
            current += _dt.timedelta(days=1)
            if self.is_trading_day(current):
                days -= 1
        return current

    def __eq__(self, other):
        return isinstance(other, TradingCalendar) and other.kind == self.kind


def _finite(name, value, minimum=None, strict=False):
    value = float(value)
    if not math.isfinite(value):
        raise InvalidInputError(f"{name} must be finite")
    if minimum is not None and (value <= minimum if strict else value < minimum):
        raise InvalidInputError(f"{name} must be {'positive' if strict else 'nonnegative'}")
    return value


class BlackScholesMertonParameters:
    """Flat-rate BSM parameters (market/parameters.rs:217-224)."""

    def __init__(self, *, risk_free_rate, dividend_yield, volatility):
        self.risk_free_rate = _finite("risk_free_rate", risk_free_rate)
        self.dividend_yield = _finite("dividend_yield", dividend_yield)
        self.volatility = _finite("volatility", volatility, 0.0)

    def __eq__(self, other):
        return isinstance(other, BlackScholesMertonParameters) and vars(self) == vars(other)


class PricingContext:
    """Market snapshot for one valuation instant."""

    def __init__(self, *, model_parameters, spot_price, valuation_time):
        self.model_parameters = model_parameters
        self.spot_price = _finite("spot_price", spot_price, 0.0, strict=True)
        self.valuation_time = valuation_time

    def replace(self, **changes):
        fields = dict(model_parameters=self.model_parameters, spot_price=self.spot_price,
                      valuation_time=self.valuation_time)
        fields.update(changes)
        return PricingContext(**fields)

    def bsm_market(self):
        from . import _abi
        model = self.model_parameters
        if not isinstance(model, BlackScholesMertonParameters):
            from .errors import UnsupportedOperationError
            raise UnsupportedOperationError("this engine needs Black-Scholes-Merton parameters")
        return _abi.BsmMarket(self.spot_price, model.risk_free_rate, model.dividend_yield,
                              model.volatility, to_micros(self.valuation_time))

    # Deterministic-forward queries of PricingContext (market/parameters.rs:
    # 386-400, 460-484) for flat BSM curves, which quote no discount factors,
    # under Actual/365 Fixed variance time, which starts at the valuation.
    def deterministic_until(self, end) -> int:
        """End of the period without variance: `end` at zero volatility,
        else the valuation time (variance_start, time/convention.rs:119-135)."""
        end = to_micros(end)
        if self.model_parameters.volatility == 0.0:
            return end
        return min(to_micros(self.valuation_time), end)

    def forward_crosses(self, time, level, radius) -> bool:
        """Whether the forward at `time` of spots in [spot - radius, spot +
        radius] reaches `level` from both sides."""
        model = self.model_parameters
        carry = (model.risk_free_rate - model.dividend_yield) * years_until(self.valuation_time, time)

        def compare(spot):
            if spot > 0.0 and level > 0.0:
                x = log_ratio(spot, level) + carry
            else:
                x = spot - level
            return None if math.isnan(x) else (x > 0.0) - (x < 0.0)
        low, high = compare(max(self.spot_price - radius, 0.0)), compare(self.spot_price + radius)
        return low is not None and high is not None and low <= 0 and high >= 0

    def deterministic_crosses(self, time, level, radius) -> bool:
        time = to_micros(time)
        return (time >= to_micros(self.valuation_time) and self.deterministic_until(time) == time
                and self.forward_crosses(time, level, radius))


def log_ratio(numerator, denominator):
    """ln(numerator / denominator), accurate near 1 (numerics.rs:80-92)."""
    relative = (numerator - denominator) / denominator
    if abs(relative) < 0.5:
        return math.log1p(relative)
    ratio = numerator / denominator
    if ratio != 0.0 and math.isfinite(ratio) and abs(ratio) >= 2.2250738585072014e-308:
        return math.log(ratio)
    return math.log(numerator) - math.log(denominator)


class LocalVolatilitySurface:
    """Bilinear local-volatility table (market/local_volatility.rs:61-359):
    rows at `times`, columns at `spots`, values[time][spot]."""

    def __init__(self, times, spots, values):
        import numpy as np
        self.times = [to_micros(t) for t in times]
        self.spots = np.ascontiguousarray(spots, dtype=np.float64)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        if len(self.times) < 2 or self.spots.size < 2:
            raise InvalidInputError("a surface needs at least two times and two spots")
        if self.values.shape != (len(self.times), self.spots.size):
            raise InvalidInputError("values must have one row per time and one column per spot")

    def _struct(self):
        import ctypes as C
        from . import _abi
        times = (C.c_int64 * len(self.times))(*self.times)
        s = _abi.LocalVolSurface(len(self.times), self.spots.size, times,
                                 self.spots.ctypes.data_as(C.POINTER(C.c_double)),
                                 self.values.ctypes.data_as(C.POINTER(C.c_double)))
        s._keep = (times, self.spots, self.values)
        return s


class LocalVolatilityParameters:
    """Flat rates with a local-volatility surface (market/parameters.rs:217-224)."""

    def __init__(self, *, risk_free_rate, dividend_yield, surface):
        self.risk_free_rate = _finite("risk_free_rate", risk_free_rate)
        self.dividend_yield = _finite("dividend_yield", dividend_yield)
        self.surface = surface


def american_model(context):
    """The C model struct of a BSM or local-volatility pricing context."""
    import ctypes as C
    from . import _abi
    model = context.model_parameters
    if isinstance(model, BlackScholesMertonParameters):
        return _abi.AmericanModel(context.spot_price, model.risk_free_rate, model.dividend_yield,
                                  model.volatility, to_micros(context.valuation_time), None)
    if isinstance(model, LocalVolatilityParameters):
        surface = model.surface._struct()
        m = _abi.AmericanModel(context.spot_price, model.risk_free_rate, model.dividend_yield, 0.0,
                               to_micros(context.valuation_time), C.pointer(surface))
        m._keep = surface
        return m
    from .errors import UnsupportedOperationError
    raise UnsupportedOperationError("American Monte Carlo needs BSM or local-volatility parameters")
