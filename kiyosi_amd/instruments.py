"""Contracts priced by the Monte Carlo engines.

Field names and meanings follow kiyosi's Python classes
(python/src/kiyosi/_native.pyi) and the Rust terms they wrap
(crates/kiyosi/src/instruments/{vanilla,autocallable}.rs). Each contract
converts itself to the plain C structs of include/kiyosi_amd.h.
"""

import ctypes as C
import datetime as _dt
import math

from . import _abi
from .errors import InvalidInputError, KiyosiError, MissingHistoryError
from .market import MICROS_PER_DAY, TradingCalendar, clock_micros, to_micros


def _positive(name, value):
    value = float(value)
    if not (math.isfinite(value) and value > 0.0):
        raise InvalidInputError(f"{name} must be positive and finite")
    return value


def _finite(name, value):
    value = float(value)
    if not math.isfinite(value):
        raise InvalidInputError(f"{name} must be finite")
    return value


class EuropeanOption:
    """European vanilla call or put (instruments/vanilla.rs)."""

    def __init__(self, *, option_type, strike, effective_time, expiry_time):
        if option_type not in ("call", "put"):
            raise InvalidInputError("option_type must be 'call' or 'put'")
        self.option_type = option_type
        self.strike = _positive("strike", strike)
        self.effective_time = effective_time
        self.expiry_time = expiry_time
        if to_micros(expiry_time) < to_micros(effective_time):
            raise InvalidInputError("expiry precedes the effective time")

    def _terms(self):
        return _abi.EuropeanTerms(1 if self.option_type == "call" else -1, 0, self.strike,
                                  to_micros(self.effective_time), to_micros(self.expiry_time))

    def at_spot_discontinuity(self, context, radius):
        """The terminal kink at the strike (instruments/vanilla.rs:201-206)."""
        spot = context.spot_price
        return ((to_micros(context.valuation_time) == to_micros(self.expiry_time)
                 and spot - radius <= self.strike <= spot + radius)
                or context.deterministic_crosses(self.expiry_time, self.strike, radius))


class _Record:
    """An immutable record, as kiyosi's observation classes are
    (pyclass(frozen), crates/kiyosi-python/src/instruments/autocallable.rs):
    fields are set by the constructor only."""

    __slots__ = ()

    def __setattr__(self, name, value):
        raise AttributeError(f"{type(self).__name__} is immutable")

    def _set(self, **fields):
        for name, value in fields.items():
            object.__setattr__(self, name, value)


class KnockInObservation(_Record):
    """Daily knock-in at a clock time, or knock-in observed only at expiry."""

    __slots__ = ("_kind", "_clock")

    def __init__(self, kind, observation_clock_time=None):
        self._set(_kind=kind, _clock=observation_clock_time)

    @staticmethod
    def every_trading_day(observation_clock_time=_dt.time(0, 0)):
        return KnockInObservation("every_trading_day", observation_clock_time)

    @staticmethod
    def at_expiry():
        return KnockInObservation("at_expiry")

    def kind(self):
        return self._kind

    def observation_clock_time(self):
        return self._clock


class SnowballObservation(_Record):
    """Knock-out observation with an annual coupon accrued on knock-out."""

    __slots__ = ("time", "knock_out_level", "coupon_rate")

    def __init__(self, time, knock_out_level, coupon_rate):
        self._set(time=time, knock_out_level=_positive("knock_out_level", knock_out_level),
                  coupon_rate=_finite("coupon_rate", coupon_rate))


class KnockOutObservation(_Record):
    __slots__ = ("time", "knock_out_level")

    def __init__(self, time, knock_out_level):
        self._set(time=time, knock_out_level=_positive("knock_out_level", knock_out_level))


class CouponObservation(_Record):
    __slots__ = ("time", "coupon_barrier_level")

    def __init__(self, time, coupon_barrier_level):
        level = _finite("coupon_barrier_level", coupon_barrier_level)
        if level < 0:
            raise InvalidInputError("coupon_barrier_level must be nonnegative")
        self._set(time=time, coupon_barrier_level=level)


_STATES = {None: _abi.STATE_UNKNOWN, "no_event": _abi.STATE_NO_EVENT,
           "knocked_out": _abi.STATE_KNOCKED_OUT, "knocked_in": _abi.STATE_KNOCKED_IN}


class _Autocallable:
    """Shared terms of snowball variants and Phoenix notes
    (instruments/autocallable.rs:700-870).

    The C form of the terms is built on first use and kept: the observation
    sequences are tuples of immutable records, and assigning any attribute of
    the note drops the kept form."""

    _coupon_kind = _abi.COUPON_SNOWBALL
    _redemption_kind = _abi.REDEMPTION_DOWNSIDE

    def __setattr__(self, name, value):
        self.__dict__.pop("_c_terms", None)
        object.__setattr__(self, name, value)

    def _common(self, *, initial_spot, knock_in_level, upper_strike, lower_strike,
                knock_in_observation, effective_time, expiry_time, barrier_state,
                principal_redemption_ratio, rebate_rate, annual_rebate_rate, calendar):
        self.initial_spot = _positive("initial_spot", initial_spot)
        self.knock_in_level = None if knock_in_level is None else _positive("knock_in_level", knock_in_level)
        self.upper_strike = upper_strike
        self.lower_strike = lower_strike
        if upper_strike is not None and lower_strike is not None and not (0 <= lower_strike <= upper_strike):
            raise InvalidInputError("lower_strike must lie in [0, upper_strike]")
        self.knock_in_observation = knock_in_observation
        self.effective_time = effective_time
        self.expiry_time = expiry_time
        if to_micros(expiry_time) <= to_micros(effective_time):
            raise InvalidInputError("expiry must follow the effective time")
        if barrier_state not in _STATES:
            raise InvalidInputError("barrier_state must be None, 'no_event', 'knocked_in' or 'knocked_out'")
        self.barrier_state = barrier_state
        self.principal_redemption_ratio = _finite("principal_redemption_ratio", principal_redemption_ratio)
        self.rebate_rate = _finite("rebate_rate", rebate_rate)
        self.annual_rebate_rate = _finite("annual_rebate_rate", annual_rebate_rate)
        self.calendar = calendar if calendar is not None else TradingCalendar.weekdays()

    def _events(self):
        """Merged, sorted observation events (AutocallableNote::new,
        instruments/autocallable.rs:972-1040)."""
        events = {}
        for obs in getattr(self, "autocall_observations", ()):
            t = to_micros(obs.time)
            events.setdefault(t, [t, 0, 0, 0.0, 0.0, 0.0])
            e = events[t]
            e[1], e[3], e[5] = 1, obs.knock_out_level, obs.coupon_rate
        for obs in getattr(self, "knock_out_observations", ()):
            t = to_micros(obs.time)
            e = events.setdefault(t, [t, 0, 0, 0.0, 0.0, 0.0])
            if not e[1]:
                e[1], e[3] = 1, obs.knock_out_level
        for obs in getattr(self, "coupon_observations", ()):
            t = to_micros(obs.time)
            e = events.setdefault(t, [t, 0, 0, 0.0, 0.0, 0.0])
            if not e[2]:
                e[2], e[4] = 1, obs.coupon_barrier_level
        start, end = to_micros(self.effective_time), to_micros(self.expiry_time)
        ordered = [events[t] for t in sorted(events)]
        for e in ordered:
            if not (start <= e[0] <= end):
                raise InvalidInputError("observations must lie within the contract life")
        return ordered

    def _terms(self):
        terms = self.__dict__.get("_c_terms")
        if terms is None:
            terms = self._build_terms()
            self.__dict__["_c_terms"] = terms
        return terms

    def _build_terms(self):
        events = self._events()
        array = (_abi.AutocallEvent * max(len(events), 1))()
        for i, e in enumerate(events):
            array[i] = _abi.AutocallEvent(*e)
        ki = self.knock_in_observation
        if self.knock_in_level is None:
            kind, clock = _abi.KNOCK_IN_NONE, 0
        elif ki.kind() == "every_trading_day":
            kind, clock = _abi.KNOCK_IN_DAILY, clock_micros(ki.observation_clock_time())
        else:
            kind, clock = _abi.KNOCK_IN_AT_EXPIRY, 0
        terms = _abi.AutocallTerms(
            calendar=self.calendar.kind, knock_in_kind=kind, redemption_kind=self._redemption_kind,
            coupon_kind=self._coupon_kind, barrier_state=_STATES[self.barrier_state],
            n_events=len(events), observation_clock_micros=clock,
            knock_in_level=self.knock_in_level or 0.0, initial_spot=self.initial_spot,
            upper_strike=self.upper_strike or 0.0, lower_strike=self.lower_strike or 0.0,
            fixed_coupon_rate=getattr(self, "minimum_coupon_rate", 0.0), coupon=self._coupon(),
            principal_redemption_ratio=self.principal_redemption_ratio, rebate_rate=self.rebate_rate,
            annual_rebate_rate=self.annual_rebate_rate, effective_time=to_micros(self.effective_time),
            expiry_time=to_micros(self.expiry_time), events=C.cast(array, C.POINTER(_abi.AutocallEvent)))
        terms._keep = array
        return terms

    # Greek-stencil hooks of OptionContract for autocallable notes
    # (instruments/autocallable.rs:1426-1583).
    @property
    def valuation_start(self):
        return self.effective_time

    def _period_coupon(self):
        return self._coupon() if self._coupon_kind == _abi.COUPON_PHOENIX else 0.0

    def _daily_clock(self):
        ki = self.knock_in_observation
        if self.knock_in_level is None or ki is None or ki.kind() != "every_trading_day":
            return None
        return clock_micros(ki.observation_clock_time())

    def _knock_in_affects_payoff(self):
        """knock_in_affects_payoff (instruments/autocallable.rs:1075-1080)."""
        if self.knock_in_level is None:
            return False
        maturity = self._coupon() if self._coupon_kind == _abi.COUPON_SNOWBALL else 0.0
        if self._redemption_kind == _abi.REDEMPTION_FIXED_COUPON:
            return self.minimum_coupon_rate != maturity
        return self.lower_strike != self.upper_strike or maturity != 0.0

    def _has_daily_observation(self, start, end):
        """A daily knock-in observation in [start, end] (:1148-1159)."""
        clock = self._daily_clock()
        if clock is None:
            return False
        start = max(start, to_micros(self.effective_time))
        end = min(end, to_micros(self.expiry_time))
        for day in range(start // MICROS_PER_DAY, end // MICROS_PER_DAY + 1):
            at = day * MICROS_PER_DAY + clock
            if start <= at <= end and self.calendar.is_trading_day_number(day):
                return True
        return False

    def _observes_knock_in(self, time):
        clock = self._daily_clock()
        if clock is not None:
            return (to_micros(self.effective_time) <= time <= to_micros(self.expiry_time)
                    and time % MICROS_PER_DAY == clock
                    and self.calendar.is_trading_day_number(time // MICROS_PER_DAY))
        return (self.knock_in_level is not None and self.knock_in_observation is not None
                and self.knock_in_observation.kind() == "at_expiry" and time == to_micros(self.expiry_time))

    def _prior(self, time):
        """prior_with_work_check (:1320-1369): checks the stated barrier state
        against the observations before `time`."""
        outs = [e[0] for e in self._events() if e[1]]
        had_out = bool(outs) and outs[0] < time
        effective = to_micros(self.effective_time)
        had_in = (self._daily_clock() is not None and time > effective
                  and self._has_daily_observation(effective, min(time - 1, to_micros(self.expiry_time))))
        if self.barrier_state is None and (had_out or had_in):
            raise MissingHistoryError("prior autocallable state is required")
        if self.barrier_state == "knocked_out" and not had_out:
            raise InvalidInputError("knock-out predates observations")
        if self.barrier_state == "knocked_in" and not had_in:
            raise InvalidInputError("knock-in predates monitoring")
        return self.barrier_state == "knocked_in"

    def time_stencil_available(self, context, before, after):
        """No observation inside a time stencil [before, after] (:1556-1583)."""
        try:
            self._prior(before)
        except (InvalidInputError, MissingHistoryError):
            return False
        if self.barrier_state == "knocked_out":
            return True
        if any(before <= e[0] <= after for e in self._events()):
            return False
        return not (self._daily_clock() is not None and self.barrier_state != "knocked_in"
                    and self._knock_in_affects_payoff() and self._has_daily_observation(before, after))

    def at_spot_discontinuity(self, context, radius):
        """Whether a spot interval of `radius` crosses a level observed at the
        valuation or a deterministic one (:1438-1554)."""
        if self.barrier_state == "knocked_out":
            return False
        time = to_micros(context.valuation_time)
        until = context.deterministic_until(self.expiry_time)
        for e in self._events():
            if e[0] <= time or e[0] > until:
                continue
            if (e[1] and context.forward_crosses(e[0], e[3], radius)) or (
                    self._period_coupon() != 0.0 and context.forward_crosses(e[0], e[4] if e[2] else 0.0, radius)):
                return True
        if until > time and self._knock_in_affects_payoff():
            from .errors import UnsupportedOperationError
            raise UnsupportedOperationError("deterministic knock-in paths (zero volatility) are not supported")
        spot = context.spot_price

        def crosses(level):
            return spot - radius <= level <= spot + radius
        now = [e for e in self._events() if e[0] == time]
        if now:
            e = now[0]
            if self._period_coupon() != 0.0 and e[2] and e[4] > 0.0 and crosses(e[4]):
                return True
            if e[1] and crosses(e[3]):
                return True
            if e[1] and spot > e[3]:
                return False
        if self.knock_in_level is not None and time == to_micros(self.expiry_time):
            knocked = self.barrier_state == "knocked_in" or (self._observes_knock_in(time)
                                                             and spot < self.knock_in_level)
            if knocked and self._redemption_kind == _abi.REDEMPTION_DOWNSIDE and (
                    crosses(self.upper_strike) or crosses(self.lower_strike)):
                return True
        return (self.barrier_state != "knocked_in" and self._knock_in_affects_payoff()
                and self._observes_knock_in(time) and self.knock_in_level is not None
                and crosses(self.knock_in_level))


class SnowballOption(_Autocallable):
    """Snowball with knock-out coupons and capped knock-in downside loss."""

    def __init__(self, *, maturity_coupon_rate, initial_spot, knock_in_level, upper_strike,
                 lower_strike, knock_in_observation, effective_time, expiry_time,
                 autocall_observations, barrier_state=None, principal_redemption_ratio=1.0,
                 rebate_rate=0.0, annual_rebate_rate=0.0, calendar=None):
        self.maturity_coupon_rate = _finite("maturity_coupon_rate", maturity_coupon_rate)
        self.autocall_observations = tuple(autocall_observations)
        if not self.autocall_observations:
            raise InvalidInputError("autocall_observations must be nonempty")
        self._common(initial_spot=initial_spot, knock_in_level=knock_in_level,
                     upper_strike=upper_strike, lower_strike=lower_strike,
                     knock_in_observation=knock_in_observation, effective_time=effective_time,
                     expiry_time=expiry_time, barrier_state=barrier_state,
                     principal_redemption_ratio=principal_redemption_ratio,
                     rebate_rate=rebate_rate, annual_rebate_rate=annual_rebate_rate,
                     calendar=calendar)

    def _coupon(self):
        return self.maturity_coupon_rate


class PhoenixOption(_Autocallable):
    """Phoenix note: per-period coupons above a barrier, knock-out, knock-in downside."""

    _coupon_kind = _abi.COUPON_PHOENIX

    def __init__(self, *, coupon_per_period, initial_spot, knock_in_level, upper_strike,
                 lower_strike, knock_in_observation, effective_time, expiry_time,
                 knock_out_observations, coupon_observations, barrier_state=None,
                 principal_redemption_ratio=1.0, rebate_rate=0.0, annual_rebate_rate=0.0,
                 calendar=None):
        self.coupon_per_period = _finite("coupon_per_period", coupon_per_period)
        self.knock_out_observations = tuple(knock_out_observations)
        self.coupon_observations = tuple(coupon_observations)
        self._common(initial_spot=initial_spot, knock_in_level=knock_in_level,
                     upper_strike=upper_strike, lower_strike=lower_strike,
                     knock_in_observation=knock_in_observation, effective_time=effective_time,
                     expiry_time=expiry_time, barrier_state=barrier_state,
                     principal_redemption_ratio=principal_redemption_ratio,
                     rebate_rate=rebate_rate, annual_rebate_rate=annual_rebate_rate,
                     calendar=calendar)

    def _coupon(self):
        return self.coupon_per_period


class BinarySnowballOption(_Autocallable):
    """Snowball without knock-in: knock-out coupons, else the maturity coupon."""

    def __init__(self, *, maturity_coupon_rate, effective_time, expiry_time, autocall_observations,
                 barrier_state=None, principal_redemption_ratio=1.0, rebate_rate=0.0,
                 annual_rebate_rate=0.0, calendar=None, initial_spot=None):
        self.maturity_coupon_rate = _finite("maturity_coupon_rate", maturity_coupon_rate)
        self.autocall_observations = tuple(autocall_observations)
        spot = initial_spot or self.autocall_observations[0].knock_out_level
        self._common(initial_spot=spot, knock_in_level=None, upper_strike=None, lower_strike=None,
                     knock_in_observation=None, effective_time=effective_time,
                     expiry_time=expiry_time, barrier_state=barrier_state,
                     principal_redemption_ratio=principal_redemption_ratio,
                     rebate_rate=rebate_rate, annual_rebate_rate=annual_rebate_rate,
                     calendar=calendar)

    def _coupon(self):
        return self.maturity_coupon_rate


class TernarySnowballOption(_Autocallable):
    """Snowball whose knock-in pays a fixed minimum coupon instead of a loss."""

    _redemption_kind = _abi.REDEMPTION_FIXED_COUPON

    def __init__(self, *, maturity_coupon_rate, minimum_coupon_rate, knock_in_level,
                 effective_time, expiry_time, autocall_observations, knock_in_observation,
                 barrier_state=None, principal_redemption_ratio=1.0, rebate_rate=0.0,
                 annual_rebate_rate=0.0, calendar=None):
        self.maturity_coupon_rate = _finite("maturity_coupon_rate", maturity_coupon_rate)
        self.minimum_coupon_rate = _finite("minimum_coupon_rate", minimum_coupon_rate)
        self.autocall_observations = tuple(autocall_observations)
        self._common(initial_spot=knock_in_level, knock_in_level=knock_in_level,
                     upper_strike=None, lower_strike=None,
                     knock_in_observation=knock_in_observation, effective_time=effective_time,
                     expiry_time=expiry_time, barrier_state=barrier_state,
                     principal_redemption_ratio=principal_redemption_ratio,
                     rebate_rate=rebate_rate, annual_rebate_rate=annual_rebate_rate,
                     calendar=calendar)

    def _coupon(self):
        return self.maturity_coupon_rate


_BARRIER_TYPES = {"down_and_out": _abi.DOWN_AND_OUT, "down_and_in": _abi.DOWN_AND_IN,
                  "up_and_out": _abi.UP_AND_OUT, "up_and_in": _abi.UP_AND_IN}


class _Monitored:
    """Monitoring state shared by single and double barriers
    (instruments/monitoring.rs:272-321, 430-445), for Black-Scholes contexts
    with positive volatility, where no deterministic interval exists."""

    _upper = None

    def _times(self):
        return None if self.observation_schedule is None else [to_micros(t) for t in self.observation_schedule]

    def _monitored(self, time):
        times = self._times()
        return (to_micros(self.effective_time) <= time <= to_micros(self.expiry_time)
                and (times is None or time in times))

    def _levels(self):
        return [self.barrier_level] if self._upper is None else [self.barrier_level, self._upper]

    def _prior(self, time):
        """Monitoring::prior (monitoring.rs:277-292): whether an observation
        before `time` touched, from the supplied history."""
        times = self._times()
        had = time > to_micros(self.effective_time) if times is None else times[0] < time
        if not had and self.touch_state == "touched":
            raise InvalidInputError("touch predates monitoring")
        if had and self.touch_state is None:
            raise MissingHistoryError("prior barrier touch state is required")
        return self.touch_state == "touched"

    def _state(self, context):
        """(touched, finished) at the valuation (Monitoring::state)."""
        time, spot = to_micros(context.valuation_time), context.spot_price
        prior = self._prior(time)
        if self._upper is None:
            hit = spot >= self.barrier_level if self.barrier_type.startswith("up") else spot <= self.barrier_level
        else:
            hit = spot <= self.barrier_level or spot >= self._upper
        times = self._times()
        finished = time == to_micros(self.expiry_time) or (times is not None and times[-1] <= time)
        return prior or (self._monitored(time) and hit), finished

    def _knock_in(self):
        return self.barrier_type.endswith("_in")

    def _payoff_requires_touch(self):
        """payoff_requires_touch (monitoring.rs:314-321; double_barrier.rs:216-224)."""
        times = self._times()
        ends = times is None or times[-1] == to_micros(self.expiry_time)
        call = self.option_type == "call"
        if self._upper is not None:
            return ends and (self.strike >= self._upper if call else self.strike <= self.barrier_level)
        if self.barrier_type.startswith("up"):
            return ends and call and self.strike >= self.barrier_level
        return ends and not call and self.strike <= self.barrier_level

    def volatility_affects_price(self, context):
        """OptionContract::volatility_affects_price (instruments/barrier.rs:283-292,
        double_barrier.rs:226-235)."""
        try:
            touched, finished = self._state(context)
        except KiyosiError:
            return False
        knock_in = self._knock_in()
        return not ((touched and not knock_in) or (not touched and finished and knock_in)
                    or (not knock_in and self._payoff_requires_touch() and self.rebate == 0.0))

    def at_spot_discontinuity(self, context, radius):
        """A monitored level within the spot stencil, or the vanilla kink at
        expiry while the payoff is live (spot_discontinuity and
        live_payoff_discontinuity, monitoring.rs:322-340, 430-445)."""
        spot, time = context.spot_price, to_micros(context.valuation_time)
        if (self.touch_state != "touched" and self._monitored(time)
                and any(spot - radius <= level <= spot + radius for level in self._levels())):
            return True
        vanilla = EuropeanOption(option_type=self.option_type, strike=self.strike,
                                 effective_time=self.effective_time, expiry_time=self.expiry_time)
        try:
            touched, finished = self._state(context)
        except KiyosiError:
            return vanilla.at_spot_discontinuity(context, radius)
        if self._knock_in():
            return (touched or not finished) and vanilla.at_spot_discontinuity(context, radius)
        return not touched and vanilla.at_spot_discontinuity(context, radius)

    def _schedule(self):
        times = self.observation_schedule
        if times is None:
            return None, 0
        arr = (C.c_int64 * max(len(times), 1))(*[to_micros(t) for t in times])
        return arr, len(times)


# BarrierTouchState (instruments/barrier.rs:38-47); None is unknown history.
_TOUCH_STATES = {None: _abi.TOUCH_UNKNOWN, "untouched": _abi.TOUCH_UNTOUCHED, "touched": _abi.TOUCH_TOUCHED}


def _touch_state(value):
    if value not in _TOUCH_STATES:
        raise InvalidInputError("touch_state must be None, 'untouched' or 'touched'")
    return value


class BarrierOption(_Monitored):
    """Single barrier call or put (instruments/barrier.rs); continuous monitoring
    without an observation schedule, scheduled fixings otherwise."""

    def __init__(self, *, option_type, strike, effective_time, expiry_time, barrier_level, barrier_type,
                 rebate=0.0, rebate_timing="at_expiry", observation_schedule=None, touch_state=None,
                 calendar=None):
        if option_type not in ("call", "put"):
            raise InvalidInputError("option_type must be 'call' or 'put'")
        if barrier_type not in _BARRIER_TYPES:
            raise InvalidInputError("unknown barrier_type")
        if rebate_timing not in ("at_hit", "at_expiry"):
            raise InvalidInputError("rebate_timing must be 'at_hit' or 'at_expiry'")
        self.touch_state = _touch_state(touch_state)
        self.option_type = option_type
        self.strike = _positive("strike", strike)
        self.barrier_level = _positive("barrier_level", barrier_level)
        self.barrier_type = barrier_type
        self.rebate = _finite("rebate", rebate)
        self.rebate_timing = rebate_timing
        self.effective_time = effective_time
        self.expiry_time = expiry_time
        self.observation_schedule = None if observation_schedule is None else list(observation_schedule)
        self.calendar = calendar

    def _terms(self):
        arr, n = self._schedule()
        terms = _abi.BarrierTerms(
            1 if self.option_type == "call" else -1, _BARRIER_TYPES[self.barrier_type],
            _abi.OBSERVE_CONTINUOUS if arr is None else _abi.OBSERVE_SCHEDULED,
            _abi.REBATE_AT_HIT if self.rebate_timing == "at_hit" else _abi.REBATE_AT_EXPIRY,
            self.strike, self.barrier_level, self.rebate, to_micros(self.effective_time),
            to_micros(self.expiry_time), n, _TOUCH_STATES[self.touch_state],
            C.cast(arr, C.POINTER(C.c_int64)) if arr is not None else None)
        terms._keep = arr
        return terms


class DoubleBarrierOption(_Monitored):
    """Call or put knocked in or out by either of two levels
    (instruments/double_barrier.rs:74-144); equality touches. Continuous
    monitoring without an observation schedule, scheduled fixings otherwise."""

    def __init__(self, *, option_type, strike, effective_time, expiry_time, lower_barrier, upper_barrier,
                 barrier_type, rebate=0.0, rebate_timing="at_expiry", observation_schedule=None,
                 touch_state=None, calendar=None):
        if option_type not in ("call", "put"):
            raise InvalidInputError("option_type must be 'call' or 'put'")
        if barrier_type not in ("knock_in", "knock_out"):
            raise InvalidInputError("barrier_type must be 'knock_in' or 'knock_out'")
        if rebate_timing not in ("at_hit", "at_expiry"):
            raise InvalidInputError("rebate_timing must be 'at_hit' or 'at_expiry'")
        lower, upper = _positive("lower_barrier", lower_barrier), _finite("upper_barrier", upper_barrier)
        if not lower < upper:
            raise InvalidInputError("upper_barrier must exceed lower_barrier")
        if barrier_type == "knock_in" and rebate_timing == "at_hit":
            raise InvalidInputError("knock-in rebates are paid at expiry")
        self.touch_state = _touch_state(touch_state)
        self.option_type = option_type
        self.strike = _positive("strike", strike)
        self.barrier_level = lower
        self._upper = upper
        self.lower_barrier, self.upper_barrier = lower, upper
        # The lower barrier carries the contract (double_barrier.rs:124-135).
        self.barrier_type = "down_and_in" if barrier_type == "knock_in" else "down_and_out"
        self.double_barrier_type = barrier_type
        self.rebate = _finite("rebate", rebate)
        if self.rebate < 0.0:
            raise InvalidInputError("rebate must be nonnegative")
        self.rebate_timing = rebate_timing
        self.effective_time = effective_time
        self.expiry_time = expiry_time
        self.observation_schedule = None if observation_schedule is None else list(observation_schedule)
        self.calendar = calendar

    def _terms(self):
        arr, n = self._schedule()
        terms = _abi.DoubleBarrierTerms(
            1 if self.option_type == "call" else -1,
            _abi.DOUBLE_KNOCK_IN if self.double_barrier_type == "knock_in" else _abi.DOUBLE_KNOCK_OUT,
            _abi.OBSERVE_CONTINUOUS if arr is None else _abi.OBSERVE_SCHEDULED,
            _abi.REBATE_AT_HIT if self.rebate_timing == "at_hit" else _abi.REBATE_AT_EXPIRY,
            self.strike, self.lower_barrier, self.upper_barrier, self.rebate, to_micros(self.effective_time),
            to_micros(self.expiry_time), n, _TOUCH_STATES[self.touch_state],
            C.cast(arr, C.POINTER(C.c_int64)) if arr is not None else None)
        terms._keep = arr
        return terms


class DigitalOption:
    """European cash-or-nothing or asset-or-nothing option
    (instruments/vanilla.rs:236-366): pays `payout` (cash) or the terminal
    spot (asset) when strictly in the money."""

    def __init__(self, *, option_type, strike, effective_time, expiry_time, payout=None):
        if option_type not in ("call", "put"):
            raise InvalidInputError("option_type must be 'call' or 'put'")
        if payout is not None:
            payout = _finite("payout", payout)
            if payout < 0.0:
                raise InvalidInputError("payout must be nonnegative")
        self.option_type = option_type
        self.strike = _positive("strike", strike)
        self.effective_time = effective_time
        self.expiry_time = expiry_time
        if to_micros(expiry_time) < to_micros(effective_time):
            raise InvalidInputError("expiry precedes the effective time")
        self.payout = payout

    @classmethod
    def cash_or_nothing(cls, option_type, strike, payout, effective_time, expiry_time):
        return cls(option_type=option_type, strike=strike, payout=payout, effective_time=effective_time,
                   expiry_time=expiry_time)

    @classmethod
    def asset_or_nothing(cls, option_type, strike, effective_time, expiry_time):
        return cls(option_type=option_type, strike=strike, effective_time=effective_time, expiry_time=expiry_time)

    def _terms(self):
        return _abi.DigitalTerms(1 if self.option_type == "call" else -1,
                                 _abi.ASSET_OR_NOTHING if self.payout is None else _abi.CASH_OR_NOTHING,
                                 self.strike, float(self.payout or 0.0), to_micros(self.effective_time),
                                 to_micros(self.expiry_time))

    def at_spot_discontinuity(self, context, radius):
        """The terminal jump at the strike (instruments/vanilla.rs:378-380)."""
        return EuropeanOption(option_type=self.option_type, strike=self.strike, effective_time=self.effective_time,
                              expiry_time=self.expiry_time).at_spot_discontinuity(context, radius)


class _AveragePrice:
    _averaging = _abi.AVERAGE_ARITHMETIC

    def __init__(self, *, option_type, strike, effective_time, averaging_start_time, expiry_time,
                 realized_average=None):
        if option_type not in ("call", "put"):
            raise InvalidInputError("option_type must be 'call' or 'put'")
        self.option_type = option_type
        self.strike = _positive("strike", strike)
        self.effective_time = effective_time
        self.averaging_start_time = averaging_start_time
        self.expiry_time = expiry_time
        self.realized_average = realized_average

    def _terms(self, step_count=50):
        return _abi.AsianTerms(1 if self.option_type == "call" else -1, self._averaging, step_count,
                               int(self.realized_average is not None), self.strike,
                               float(self.realized_average or 0.0), to_micros(self.effective_time),
                               to_micros(self.averaging_start_time), to_micros(self.expiry_time))


class ArithmeticAveragePriceOption(_AveragePrice):
    """Continuously averaged arithmetic-average-price option (instruments/asian.rs)."""


class GeometricAveragePriceOption(_AveragePrice):
    """Continuously averaged geometric-average-price option (instruments/asian.rs)."""

    _averaging = _abi.AVERAGE_GEOMETRIC


class Accumulator:
    """Accumulator (call) or decumulator (put) with daily accrual and knock-out
    (instruments/accumulator.rs). ``knock_out_observation_schedule`` lists
    explicit knock-out times (None: daily at the observation clock);
    ``accumulated_quantity`` and ``knock_out_settled`` are the history before
    valuation (None: unknown)."""

    def __init__(self, *, strike, knock_out_level, daily_quantity, acceleration_factor, effective_time,
                 expiry_time, observation_clock_time=_dt.time(0, 0), accumulated_quantity=None,
                 option_type="call", calendar=None, knock_out_observation_schedule=None, knock_out_settled=None):
        if option_type not in ("call", "put"):
            raise InvalidInputError("option_type must be 'call' or 'put'")
        self.strike = _positive("strike", strike)
        self.knock_out_level = _positive("knock_out_level", knock_out_level)
        self.daily_quantity = _finite("daily_quantity", daily_quantity)
        self.acceleration_factor = _finite("acceleration_factor", acceleration_factor)
        self.effective_time = effective_time
        self.expiry_time = expiry_time
        self.observation_clock_time = observation_clock_time
        self.accumulated_quantity = None if accumulated_quantity is None else float(accumulated_quantity)
        self.option_type = option_type
        self.calendar = calendar if calendar is not None else TradingCalendar.weekdays()
        self.knock_out_observation_schedule = (None if knock_out_observation_schedule is None
                                               else list(knock_out_observation_schedule))
        if knock_out_settled not in (None, True, False):
            raise InvalidInputError("knock_out_settled must be None, True or False")
        self.knock_out_settled = knock_out_settled

    def _terms(self):
        times = self.knock_out_observation_schedule
        arr = (C.c_int64 * len(times))(*[to_micros(t) for t in times]) if times else None
        settled = {None: _abi.SETTLED_UNKNOWN, False: _abi.SETTLED_NO, True: _abi.SETTLED_YES}
        t = _abi.AccumulatorTerms(
            1 if self.option_type == "call" else -1, self.calendar.kind, self.strike, self.knock_out_level,
            self.daily_quantity, self.acceleration_factor, self.accumulated_quantity or 0.0,
            clock_micros(self.observation_clock_time), to_micros(self.effective_time),
            to_micros(self.expiry_time), int(self.accumulated_quantity is not None),
            settled[self.knock_out_settled], int(times is not None), len(times) if times else 0,
            C.cast(arr, C.POINTER(C.c_int64)) if arr else None)
        t._keep = arr
        return t


class AmericanOption(EuropeanOption):
    """American call or put (instruments/vanilla.rs), exercisable on the
    engine's exercise grid."""

    @property
    def valuation_start(self):
        """No valuation before the effective time (OptionContract's default,
        instruments/vanilla.rs:141-143, which AmericanOption keeps)."""
        return self.effective_time

    def _american_terms(self, step_count):
        return _abi.AmericanTerms(1 if self.option_type == "call" else -1, step_count, self.strike,
                                  to_micros(self.effective_time), to_micros(self.expiry_time))
