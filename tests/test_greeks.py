"""Numerical Greeks on shared draws: the stencil (planning, bumps, time
stencil, finite-difference formulas) pinned against closed-form
Black-Scholes-Merton Greeks with the oracle as the pricer (CPU), and the
GPU's one-launch scenario batch checked bit for bit against pricing the
scenarios one at a time (the reference's numerical_greeks == price_with_greeks
property, python/tests/test_monte_carlo_statistics.py:141-179)."""

import datetime as dt
import math

import pytest
from scipy.stats import norm

from conftest import END, START
from kiyosi_amd import _abi, greeks
from kiyosi_amd.engines import MonteCarloVanillaEngine
from kiyosi_amd.instruments import EuropeanOption
from kiyosi_amd.market import BlackScholesMertonParameters, PricingContext, to_micros


def context(spot=100.0, r=0.05, q=0.02, sigma=0.2, valuation=START):
    return PricingContext(model_parameters=BlackScholesMertonParameters(risk_free_rate=r, dividend_yield=q,
                                                                        volatility=sigma),
                          spot_price=spot, valuation_time=valuation)


def option(kind="call", strike=100.0):
    return EuropeanOption(option_type=kind, strike=strike, effective_time=START, expiry_time=END)


def analytic(kind, s, k, r, q, sigma, t):
    """Closed-form BSM Greeks; theta is dV/d(valuation time) per year."""
    sign = 1.0 if kind == "call" else -1.0
    w = sigma * math.sqrt(t)
    d1 = (math.log(s / k) + (r - q) * t) / w + 0.5 * w
    d2 = d1 - w
    dq, dr = math.exp(-q * t), math.exp(-r * t)
    pdf = norm.pdf(d1)
    price = sign * (s * dq * norm.cdf(sign * d1) - k * dr * norm.cdf(sign * d2))
    theta = (-s * dq * pdf * sigma / (2 * math.sqrt(t)) - sign * r * k * dr * norm.cdf(sign * d2)
             + sign * q * s * dq * norm.cdf(sign * d1))
    return {
        "price": price,
        "delta": sign * dq * norm.cdf(sign * d1),
        "gamma": dq * pdf / (s * w),
        "vega": s * dq * pdf * math.sqrt(t),
        "rho": sign * k * t * dr * norm.cdf(sign * d2),
        "dividend_rho": -sign * s * t * dq * norm.cdf(sign * d1),
        "theta": theta,
        "vanna": -dq * pdf * d2 / sigma,
        "volga": s * dq * pdf * math.sqrt(t) * d1 * d2 / sigma,
    }


class OraclePricer:
    """The oracle behind the engine interface numerical_greeks uses."""

    def __init__(self, oracle, paths, seed, antithetic=True):
        self.oracle, self.paths, self.seed, self.antithetic = oracle, paths, seed, antithetic

    def with_resolved_seed(self):
        return self

    def price_with_seed(self, opt, ctx):
        s = _abi.Settings(self.paths, self.seed, 1, _abi.PSEUDO_RANDOM, int(self.antithetic), _abi.FP64, 0, 0)
        out = self.oracle.vanilla_price(s, opt._terms(), ctx.bsm_market(), workers=16)
        return out.price, (out.seed if out.has_seed else None)


def test_symmetric_shift_and_default_bumps():
    assert greeks.symmetric_shift(100.0, 0.01) == 100.01 - 100.0
    assert greeks.symmetric_shift(0.0, 1e-4) == 1e-4
    assert greeks.symmetric_shift(1.7e308, 1e307) == 0.0  # the high leg overflows
    assert greeks.symmetric_shift(1.0, 1e-17) == 0.0  # below the spacing of doubles at 1
    s = greeks.NumericalShiftSettings()
    assert s.spot_shift_at(100.0) == 0.01 and s.spot_shift_at(0.5) == 0.005
    assert s.volatility_shift_at(0.2) == 1e-4 and s.volatility_shift_at(1e-4) == 5e-5
    with pytest.raises(Exception):
        greeks.NumericalShiftSettings(spot_shift=0.0)


def test_time_stencil_is_centred_one_day_and_shrinks_near_expiry():
    day = 86_400_000_000
    t = to_micros(START)
    assert greeks.time_stencil(option(), context(), greeks.NumericalShiftSettings()) == (t - day, t + day)
    late = START + dt.timedelta(days=361)  # four days to expiry: one eighth of the remaining life
    before, after = greeks.time_stencil(option(), context(valuation=late), greeks.NumericalShiftSettings())
    assert after - to_micros(late) == (to_micros(END) - to_micros(late)) // 8 == to_micros(late) - before


def test_expiry_greeks_use_terminal_stencils():
    """At expiry: vega, volga, rho and dividend rho are zero; delta and gamma
    come from intrinsic values at bumped spots; at the strike the kink leaves
    the spot Greeks undefined; theta stays undefined."""
    ctx = context(spot=110.0, valuation=END)
    opt = option("call", 100.0)

    def intrinsic(c):
        return max(c.spot_price - 100.0, 0.0)
    values, unavailable = greeks.complete(opt, ctx, None, greeks.NumericalShiftSettings(), 10.0, intrinsic)
    assert values["vega"] == values["rho"] == values["volga"] == values["dividend_rho"] == 0.0
    assert values["delta"] == pytest.approx(1.0) and values["gamma"] == pytest.approx(0.0, abs=1e-9)
    assert values["vanna"] == 0.0 and "theta" not in values
    res = greeks.PricingResult(10.0, greeks.GREEKS, values, unavailable, None)
    assert res.status("theta") == "undefined"
    at_strike, unavailable = greeks.complete(opt, context(spot=100.0, valuation=END), None,
                                             greeks.NumericalShiftSettings(), 0.0, intrinsic)
    assert "delta" not in at_strike and "delta" not in unavailable  # undefined at the kink
    assert at_strike["vega"] == 0.0


def test_full_stencil_has_twenty_one_scenarios():
    scenarios = greeks.plan(option(), context(), None, greeks.NumericalShiftSettings())
    assert len(scenarios) == 21 and greeks.key(scenarios[0]) == greeks.key(context())
    assert len(greeks.plan(option(), context(), "delta", greeks.NumericalShiftSettings())) == 3


def snowball_at(valuation, clock=dt.time(15, 0), state="no_event"):
    from kiyosi_amd.instruments import KnockInObservation, SnowballObservation, SnowballOption
    from kiyosi_amd.market import TradingCalendar
    cal = TradingCalendar.weekdays()
    dates = [cal.add_trading_days(START, 21 * k) for k in range(1, 13)]
    return SnowballOption(maturity_coupon_rate=0.15, initial_spot=100.0, knock_in_level=75.0,
                          upper_strike=100.0, lower_strike=0.0,
                          knock_in_observation=KnockInObservation.every_trading_day(clock),
                          effective_time=START, expiry_time=dates[-1],
                          autocall_observations=[SnowballObservation(d, 100.0, 0.15) for d in dates],
                          barrier_state=state, calendar=cal)


def test_autocallable_time_stencil_halves_until_no_observation_lies_inside():
    """Monday 12:00 with daily knock-in at 15:00: one day, 12 h, 6 h and 3 h
    windows all hold a 15:00 observation (inclusive); 1.5 h does not
    (numerical.rs:258-265, instruments/autocallable.rs:1556-1583)."""
    t = dt.datetime(2025, 1, 6, 12, 0)
    note = snowball_at(t)
    before, after = greeks.time_stencil(note, context(valuation=t), greeks.NumericalShiftSettings())
    assert after - to_micros(t) == to_micros(t) - before == 5_400_000_000
    # An explicit bump is not shortened, and a stencil holding an observation is unavailable.
    fixed = greeks.NumericalShiftSettings(time_shift=dt.timedelta(hours=6))
    assert greeks.time_stencil(note, context(valuation=t), fixed) is None
    # Knock-in history that cannot hold before the stencil's start makes it unavailable.
    assert not snowball_at(t, state=None).time_stencil_available(context(valuation=t), to_micros(t) - 1,
                                                                 to_micros(t) + 1)


def test_autocallable_spot_discontinuities():
    """An observation at the valuation whose knock-out level lies within the
    spot stencil, and a daily knock-in observed now within it
    (instruments/autocallable.rs:1518-1554)."""
    from kiyosi_amd.market import TradingCalendar
    obs_day = TradingCalendar.weekdays().add_trading_days(START, 21)
    note = snowball_at(obs_day, clock=dt.time(0, 0))
    near = context(spot=100.005, valuation=obs_day)
    assert note.at_spot_discontinuity(near, 0.01) and not note.at_spot_discontinuity(near, 0.001)
    assert not note.at_spot_discontinuity(context(spot=101.0, valuation=obs_day), 0.01)
    day = dt.datetime(2025, 1, 7)
    assert note.at_spot_discontinuity(context(spot=75.004, valuation=day), 0.01)
    assert not note.at_spot_discontinuity(context(spot=75.004, valuation=day + dt.timedelta(hours=1)), 0.01)
    assert not snowball_at(day, state="knocked_out").at_spot_discontinuity(context(spot=75.0, valuation=day), 1.0)


@pytest.mark.parametrize("kind", ["call", "put"])
def test_stencil_greeks_on_the_oracle_match_closed_forms(oracle, kind):
    """2^20 antithetic paths, the same seed in every scenario: first-order
    Greeks within 1% of the closed forms; gamma and the second-order
    cross terms, noisier by construction, within 15%."""
    pricer = OraclePricer(oracle, 1 << 20, seed=1729)
    res = greeks.numerical_greeks(pricer, option(kind), context(),
                                  ["delta", "gamma", "vega", "rho", "dividend_rho", "theta", "vanna"])
    ref = analytic(kind, 100.0, 100.0, 0.05, 0.02, 0.2, 1.0)
    assert res.price == pytest.approx(ref["price"], rel=5e-3)
    for g in ("delta", "vega", "rho", "dividend_rho", "theta"):
        assert res.require(g) == pytest.approx(ref[g], rel=1e-2), g
    assert res.require("gamma") == pytest.approx(ref["gamma"], rel=0.15)
    assert res.require("vanna") == pytest.approx(ref["vanna"], rel=0.15)
    assert res.status("speed") == "not_requested"


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("precision,sampling,antithetic", [("fp64", "pseudo_random", True),
                                                           ("fp64", "sobol", False),
                                                           ("fp32", "pseudo_random", True),
                                                           ("fp32", "sobol", False)])
@pytest.mark.parametrize("kind", ["call", "put"])
def test_shared_draw_batch_equals_scenario_by_scenario(gpu, precision, sampling, antithetic, kind):
    engine = MonteCarloVanillaEngine(path_count=1 << 16, seed=4242, precision=precision, sampling=sampling,
                                     antithetic=antithetic)
    batched = engine.price_with_greeks(option(kind), context())
    scalar = greeks.numerical_greeks(engine, option(kind), context())
    assert batched == scalar
    assert batched.price == engine.price(option(kind), context())
    assert set(batched._values) == set(greeks.GREEKS)


@pytest.mark.gpu
def test_gpu_greeks_match_the_oracle_stencil(gpu, oracle):
    engine = MonteCarloVanillaEngine(path_count=1 << 18, seed=77, antithetic=True)
    got = engine.price_with_greeks(option("put", 95.0), context(sigma=0.3))
    ref = greeks.numerical_greeks(OraclePricer(oracle, 1 << 18, seed=77), option("put", 95.0), context(sigma=0.3))
    assert got.price == pytest.approx(ref.price, rel=1e-12)
    for g in ("delta", "vega", "rho", "dividend_rho", "theta"):
        assert got.require(g) == pytest.approx(ref.require(g), rel=1e-7, abs=1e-9), g
    for g in ("gamma", "volga", "vanna", "charm"):
        assert got.require(g) == pytest.approx(ref.require(g), rel=1e-4, abs=1e-6), g


@pytest.mark.gpu
def test_seeded_greeks_replay(gpu):
    engine = MonteCarloVanillaEngine(path_count=4096)
    first = engine.price_with_greeks(option(), context(), "delta")
    assert first.seed is not None
    replay = MonteCarloVanillaEngine(path_count=4096, seed=first.seed).price_with_greeks(option(), context(), "delta")
    assert replay == first
