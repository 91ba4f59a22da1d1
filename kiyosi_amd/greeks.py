"""Numerical Greeks by central-difference stencils on shared Monte Carlo draws.

Restates, for flat-rate Black-Scholes-Merton contexts, the stencil planning
and completion of crates/kiyosi/src/risk/numerical.rs (symmetric_shift
:9-26, numerical_greeks :164-210, time_stencil :212-273,
complete_greeks_with_pair :331-640) and the shift defaults of
NumericalShiftSettings (risk/analytics.rs:382-421). Every scenario is priced
with the engine's resolved seed, so all scenarios see the same draws and the
differences carry little sampling noise. ``price_with_greeks`` of the
European engine prices the whole stencil in one shared-draw GPU launch
(MonteCarloVanillaEngine::price_with_greeks_with_settings,
engines/monte_carlo.rs:1334-1409); ``numerical_greeks`` prices the scenarios
one at a time and returns the same numbers.
"""

import datetime as _dt
import math

from .errors import InvalidInputError
from .market import BlackScholesMertonParameters, to_micros

GREEKS = ("delta", "gamma", "speed", "theta", "charm", "color", "vega", "volga", "vanna", "zomma", "rho",
          "dividend_rho")
_SPOT = ("delta", "gamma", "speed", "vanna", "zomma", "charm", "color")
_FAR_PAST = -(1 << 62)  # no valuation start


class NumericalShiftSettings:
    """Bump sizes: spot and volatility default to min(0.01, 1% of spot) and
    min(1e-4, volatility / 2); rates 1e-4; time one day (microseconds)."""

    def __init__(self, *, spot_shift=None, volatility_shift=None, rate_shift=1e-4, time_shift=None):
        for name, value in (("spot_shift", spot_shift), ("volatility_shift", volatility_shift),
                            ("rate_shift", rate_shift)):
            if value is not None and not (math.isfinite(value) and value > 0.0):
                raise InvalidInputError(f"{name} must be positive and finite")
        if time_shift is not None:
            micros = to_micros(_dt.datetime(1970, 1, 1) + time_shift)
            if micros <= 0:
                raise InvalidInputError("time_shift must be at least one microsecond")
        self.spot_shift, self.volatility_shift = spot_shift, volatility_shift
        self.rate_shift, self.time_shift = rate_shift, time_shift

    def spot_shift_at(self, spot):
        return self.spot_shift if self.spot_shift is not None else min(0.01, 0.01 * spot)

    def volatility_shift_at(self, volatility):
        return self.volatility_shift if self.volatility_shift is not None else min(1e-4, volatility / 2.0)

    def time_shift_micros(self):
        if self.time_shift is None:
            return 86_400_000_000
        return to_micros(_dt.datetime(1970, 1, 1) + self.time_shift)


class PricingResult:
    """Price, requested Greeks and the seed that reproduces them. A requested
    Greek without a value is "stencil_unavailable" when its stencil could not
    be formed (other bump sizes may recover it) and "undefined" when the
    derivative does not exist at this valuation (PricingResult::status,
    risk/analytics.rs:195-207)."""

    def __init__(self, price, requested, values, unavailable, seed, supported=GREEKS):
        self.price = price
        self.requested = frozenset(requested)
        self.supported = frozenset(supported)
        self._values = {g: v for g, v in values.items() if g in self.supported}
        self._unavailable = frozenset(unavailable) - set(self._values)
        self.seed = seed

    def get(self, greek):
        return self._values.get(greek)

    def require(self, greek):
        if greek not in self._values:
            raise InvalidInputError(f"{greek} is {self.status(greek)}")
        return self._values[greek]

    def status(self, greek):
        if greek not in self.requested:
            return "not_requested"
        if greek in self._values:
            return "available"
        if greek not in self.supported:
            return "unsupported"
        return "stencil_unavailable" if greek in self._unavailable else "undefined"

    def __eq__(self, other):
        return (isinstance(other, PricingResult) and self.price == other.price and self._values == other._values
                and self._unavailable == other._unavailable and self.seed == other.seed)

    def __repr__(self):
        return f"PricingResult(price={self.price!r}, greeks={self._values!r}, seed={self.seed!r})"


def _at_spot_discontinuity(option, context, radius):
    hook = getattr(option, "at_spot_discontinuity", None)
    return bool(hook(context, radius)) if hook else False


def _all_greeks_unavailable(option, context):
    """OptionContract::all_greeks_unavailable's default (instruments/vanilla.rs:161-166)."""
    return context.model_parameters.volatility == 0.0 and _at_spot_discontinuity(option, context, 0.0)


def _time_stencil_available(option, context, before, after):
    hook = getattr(option, "time_stencil_available", None)
    return bool(hook(context, before, after)) if hook else True


def normalize_request(request):
    if request is None or request == "all":
        return set(GREEKS)
    names = {request} if isinstance(request, str) else set(request)
    unknown = names - set(GREEKS)
    if unknown:
        raise InvalidInputError(f"unknown Greeks: {sorted(unknown)}")
    return names


def symmetric_shift(value, requested):
    """The bump both stencil legs can use exactly (numerical.rs:9-26)."""
    high = value + requested
    shift = high - value
    low = value - shift
    if math.isfinite(high) and math.isfinite(shift) and shift > 0.0 and math.isfinite(low) and value - low == shift:
        return shift
    return 0.0


def scenario(context, *, spot=None, volatility=None, rate_shift=0.0, dividend_shift=0.0, valuation=None):
    """`context` with a shifted spot, volatility, flat rates or valuation time (microseconds)."""
    model = context.model_parameters
    params = BlackScholesMertonParameters(risk_free_rate=model.risk_free_rate + rate_shift,
                                          dividend_yield=model.dividend_yield + dividend_shift,
                                          volatility=model.volatility if volatility is None else volatility)
    when = context.valuation_time if valuation is None else _dt.datetime(1970, 1, 1) + _dt.timedelta(
        microseconds=valuation)
    return context.replace(model_parameters=params, spot_price=context.spot_price if spot is None else spot,
                           valuation_time=when)


def key(context):
    m = context.model_parameters
    return (context.spot_price, m.volatility, m.risk_free_rate, m.dividend_yield, to_micros(context.valuation_time))


def time_stencil(option, context, settings):
    """(before, after) valuation times in microseconds, or None
    (numerical.rs:212-273 for flat curves, whose calendar variance time has
    no transitions). Automatic bumps halve until the contract accepts the
    stencil (no observation inside it)."""
    time = to_micros(context.valuation_time)
    start = getattr(option, "valuation_start", None)
    lower = to_micros(start) if start is not None else _FAR_PAST
    expiry = to_micros(option.expiry_time)
    upper = expiry
    radius = settings.time_shift_micros()
    before_width = time - lower
    after_width = upper - time
    shift = min(radius, max(upper - lower, 0))
    if settings.time_shift is None:
        # Near expiry the price is far from linear in time.
        shift = min(shift, max((expiry - time) // 8, 1))
    if before_width > 0 and after_width > 0:
        shift = min(shift, before_width, after_width)
    before = time - min(shift, before_width)
    after = time + min(shift, after_width)
    while settings.time_shift is None and shift > 1 and not _time_stencil_available(option, context, before, after):
        shift //= 2
        before = time - min(shift, before_width)
        after = time + min(shift, after_width)
    if before != after and _time_stencil_available(option, context, before, after):
        return before, after
    return None


def expiry_greeks(option, context, request, settings, p0, price):
    """Greeks at expiry from terminal-payoff stencils (numerical.rs:28-112):
    vega, volga, rho and dividend rho are zero; vanna and zomma are zero where
    the spot stencil exists; delta, gamma and speed come from terminal values
    at bumped spots; theta, charm and color stay undefined."""
    values, unavailable = {}, set()
    if not request:
        return values, unavailable
    for g in ("vega", "volga", "rho", "dividend_rho"):
        if g in request:
            values[g] = 0.0
    if not _at_spot_discontinuity(option, context, 0.0):
        unavailable |= {"delta", "gamma", "speed", "vanna", "zomma"}
    spot = context.spot_price
    h = symmetric_shift(spot, settings.spot_shift_at(spot))
    if h <= 0.0 or spot <= h or _at_spot_discontinuity(option, context, h):
        return values, unavailable
    for g in ("vanna", "zomma"):
        if g in request:
            values[g] = 0.0
    if request & {"delta", "gamma", "speed"}:
        up, down = price(context.replace(spot_price=spot + h)), price(context.replace(spot_price=spot - h))
        if "delta" in request:
            values["delta"] = (up - down) / h / 2.0
        if "gamma" in request:
            values["gamma"] = ((up - p0) - (p0 - down)) / h / h
        two_h = 2.0 * h
        if ("speed" in request and math.isfinite(two_h) and spot > two_h and symmetric_shift(spot, two_h) == two_h
                and not _at_spot_discontinuity(option, context, two_h)):
            up2 = price(context.replace(spot_price=spot + two_h))
            down2 = price(context.replace(spot_price=spot - two_h))
            values["speed"] = ((up2 - up) - (up - down) + (down - down2)) / h / h / two_h
    return values, unavailable


def complete(option, context, request, settings, p0, price):
    """Greeks of `request` from the stencil prices (complete_greeks_with_pair,
    numerical.rs:331-640); `price(context)` prices one scenario. Returns
    ({greek: value}, {greeks whose stencil is unavailable})."""
    request = normalize_request(request)
    model = context.model_parameters
    if not isinstance(model, BlackScholesMertonParameters):
        raise InvalidInputError("numerical Greeks need Black-Scholes-Merton parameters")
    time = to_micros(context.valuation_time)
    values = {}
    if time == to_micros(option.expiry_time):
        return expiry_greeks(option, context, request, settings, p0, price)
    if not request or _all_greeks_unavailable(option, context):
        return values, set()
    spot, vol = context.spot_price, model.volatility
    h = symmetric_shift(spot, settings.spot_shift_at(spot))
    v = symmetric_shift(vol, settings.volatility_shift_at(vol)) if vol > 0.0 else 0.0
    spot_ok = (any(g in request for g in _SPOT) and h > 0.0 and spot > h
               and not _at_spot_discontinuity(option, context, h))
    two_h = 2.0 * h
    speed_ok = ("speed" in request and spot_ok and math.isfinite(two_h) and spot > two_h
                and symmetric_shift(spot, two_h) == two_h and not _at_spot_discontinuity(option, context, two_h))

    def at(s=None, sigma=None, valuation=None):
        return price(scenario(context, spot=s, volatility=sigma, valuation=valuation))

    def derivatives(up, center, down):
        return (up - down) / h / 2.0, ((up - center) - (center - down)) / h / h

    def spot_derivatives(sigma, valuation, center):
        return derivatives(at(spot + h, sigma, valuation), center, at(spot - h, sigma, valuation))

    if spot_ok and ("delta" in request or "gamma" in request or speed_ok):
        up, down = at(spot + h), at(spot - h)
        delta, gamma = derivatives(up, p0, down)
        if "delta" in request:
            values["delta"] = delta
        if "gamma" in request:
            values["gamma"] = gamma
        if speed_ok:
            up2, down2 = at(spot + two_h), at(spot - two_h)
            values["speed"] = ((up2 - up) - (up - down) + (down - down2)) / h / h / two_h
    if (v > 0.0 and vol > v and math.isfinite(2.0 * v)
            and ("vega" in request or "volga" in request or (spot_ok and ("vanna" in request or "zomma" in request)))):
        prices = None
        if "vega" in request or "volga" in request:
            prices = (at(sigma=vol + v), at(sigma=vol - v))
            if "vega" in request:
                values["vega"] = (prices[0] - prices[1]) / (2.0 * v)
            if "volga" in request:
                values["volga"] = ((prices[0] - p0) - (p0 - prices[1])) / v / v
        if spot_ok and ("vanna" in request or "zomma" in request):
            if "zomma" in request:
                first, last = prices if prices else (at(sigma=vol + v), at(sigma=vol - v))
            else:
                first = last = p0
            high = spot_derivatives(vol + v, None, first)
            low = spot_derivatives(vol - v, None, last)
            if "vanna" in request:
                values["vanna"] = (high[0] - low[0]) / (2.0 * v)
            if "zomma" in request:
                values["zomma"] = (high[1] - low[1]) / (2.0 * v)
    if "rho" in request:
        r = symmetric_shift(model.risk_free_rate, settings.rate_shift)
        if r > 0.0 and math.isfinite(2.0 * r):
            high = price(scenario(context, rate_shift=r))
            low = price(scenario(context, rate_shift=-r))
            values["rho"] = (high - low) / (2.0 * r)
    if "dividend_rho" in request:
        q = symmetric_shift(model.dividend_yield, settings.rate_shift)
        if q > 0.0 and math.isfinite(2.0 * q):
            high = price(scenario(context, dividend_shift=q))
            low = price(scenario(context, dividend_shift=-q))
            values["dividend_rho"] = (high - low) / (2.0 * q)
    if any(g in request for g in ("theta", "charm", "color")):
        stencil = time_stencil(option, context, settings)
        if stencil is not None:
            before, after = stencil
            years = (after - before) / (365.0 * 86_400_000_000)

            def at_time(t):
                return p0 if t == time else at(valuation=t)
            prices = None
            if "theta" in request:
                prices = (at_time(after), at_time(before))
                values["theta"] = (prices[0] - prices[1]) / years
            if spot_ok and ("charm" in request or "color" in request):
                if "color" in request:
                    last, first = prices if prices else (at_time(after), at_time(before))
                else:
                    first = last = p0
                early = spot_derivatives(None, before, first)
                late = spot_derivatives(None, after, last)
                if "charm" in request:
                    values["charm"] = (late[0] - early[0]) / years
                if "color" in request:
                    values["color"] = (late[1] - early[1]) / years
    spot_undefined = _at_spot_discontinuity(option, context, 0.0)
    unavailable = {g for g in GREEKS
                   if not ((g in _SPOT and spot_undefined)
                           or (g in ("vega", "volga", "vanna", "zomma") and vol == 0.0))}
    return values, unavailable


def native_spot_result(option, context, request, settings, price, values, seed, reprice):
    """Completes an engine's native delta and gamma (engines.rs:138-162):
    only delta and gamma are supported; off expiry a missing one is
    "stencil_unavailable" unless the spot itself sits on a discontinuity;
    at expiry the terminal-payoff stencil supplies them
    (numerical.rs:28-112). `reprice(context)` prices one scenario."""
    request = normalize_request(request)
    supported = {"delta", "gamma"}
    values = {g: v for g, v in values.items() if g in request}
    if to_micros(context.valuation_time) != to_micros(option.expiry_time):
        unavailable = set() if _at_spot_discontinuity(option, context, 0.0) else supported
        return PricingResult(price, request, values, unavailable & request, seed, supported)
    extra, unavailable = expiry_greeks(option, context, request & supported, settings, price, reprice)
    values.update(extra)
    return PricingResult(price, request, values, unavailable & request, seed, supported)


def plan(option, context, request, settings):
    """The distinct scenario contexts the stencil prices, base first."""
    seen = {key(context): context}

    def record(ctx):
        seen.setdefault(key(ctx), ctx)
        return 0.0
    complete(option, context, request, settings, 0.0, record)
    return list(seen.values())


def numerical_greeks(engine, option, context, request=None, settings=None):
    """Greeks by pricing every stencil scenario with `engine.price_with_seed`
    on the engine's resolved seed (risk/numerical.rs:164-210); the result
    carries the seed when any scenario sampled."""
    settings = settings or NumericalShiftSettings()
    engine = engine if engine.seed is not None else engine.with_resolved_seed()
    cache, seed = {}, [None]

    def price(ctx):
        k = key(ctx)
        if k not in cache:
            value, used = engine.price_with_seed(option, ctx)
            cache[k] = value
            if used is not None:
                seed[0] = used
        return cache[k]
    p0 = price(context)
    values, unavailable = complete(option, context, request, settings, p0, price)
    return PricingResult(p0, normalize_request(request), values, unavailable, seed[0])
