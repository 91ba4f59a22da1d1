"""Synthetic benchmark contracts at their full sizes.

Each function returns (instrument, context, engine keyword arguments) for one
workload; bench.py and the tests use them.
"""

import datetime as dt

from .instruments import EuropeanOption, KnockInObservation, SnowballObservation, SnowballOption
from .market import BlackScholesMertonParameters, PricingContext, TradingCalendar

VALUATION = dt.datetime(2025, 1, 1)


def european_call_1m_pcg_antithetic():
    """European call under BSM: S0=K=100, r=5%, q=2%, sigma=20%, T=1y exactly;
    1,000,000 PCG paths in antithetic pairs, seed 1729."""
    option = EuropeanOption(option_type="call", strike=100.0, effective_time=VALUATION,
                            expiry_time=dt.datetime(2026, 1, 1))
    context = PricingContext(
        model_parameters=BlackScholesMertonParameters(risk_free_rate=0.05, dividend_yield=0.02,
                                                      volatility=0.2),
        spot_price=100.0, valuation_time=VALUATION)
    return option, context, dict(path_count=1_000_000, seed=1729, antithetic=True)


def snowball_note(trading_days=252, months=12, step_down=0.5, knock_in=75.0, coupon=0.15,
                  valuation=VALUATION, calendar=None):
    """Snowball on a weekdays calendar with daily knock-in monitoring at 00:00:
    expiry on the `trading_days`-th trading day after valuation, knock-out
    observations every trading_days/months trading days at 100% of S0
    stepping down by `step_down` per observation, coupon 15% p.a., knock-in
    at 75%, upper strike 100, lower strike 0, unit principal."""
    calendar = calendar or TradingCalendar.weekdays()
    spacing = trading_days // months
    observations, date = [], valuation
    for k in range(1, months + 1):
        date = calendar.add_trading_days(date, spacing)
        observations.append(SnowballObservation(date, 100.0 - step_down * (k - 1), coupon))
    expiry = calendar.add_trading_days(valuation, trading_days)
    return SnowballOption(
        maturity_coupon_rate=coupon, initial_spot=100.0, knock_in_level=knock_in,
        upper_strike=100.0, lower_strike=0.0,
        knock_in_observation=KnockInObservation.every_trading_day(dt.time(0, 0)),
        effective_time=valuation, expiry_time=expiry, autocall_observations=observations,
        calendar=calendar)


def snowball_252d_sobol_2p20():
    """Snowball under BSM, 252 daily steps, 12 monthly knock-out observations
    stepping down 0.5% from 100%, daily knock-in at 75%, coupon 15% p.a.;
    sigma=25%, r=2%, q=0; 2^20 Sobol paths, seed 1729."""
    note = snowball_note()
    context = PricingContext(
        model_parameters=BlackScholesMertonParameters(risk_free_rate=0.02, dividend_yield=0.0,
                                                      volatility=0.25),
        spot_price=100.0, valuation_time=VALUATION)
    return note, context, dict(path_count=1 << 20, seed=1729, sampling="sobol")
