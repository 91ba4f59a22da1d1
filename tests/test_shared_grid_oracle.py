"""Oracle pins of the shared daily grid rule (DailyGrids::prepare,
engines/monte_carlo/autocallable/portfolio.rs:75-186), CPU only: which notes
go on a shared daily grid, that such a note keeps its scalar plan's dates,
transitions and events and only loses the bridge, and that a note's price is
affine in its coupon on fixed draws, which is what lets the reference
interpolate coupon families (portfolio.rs:570-667) instead of simulating
them."""

import datetime as dt

import numpy as np
import pytest

from kiyosi_amd import _abi
from kiyosi_amd.instruments import KnockInObservation, SnowballObservation, SnowballOption
from kiyosi_amd.market import TradingCalendar

from test_gpu_shared_grid import COUPONS, END, market, snowball


def settings(paths, seed=42, sampling=_abi.PSEUDO_RANDOM, antithetic=0):
    return _abi.Settings(paths, seed, 1, sampling, antithetic, _abi.FP64, 0, 0)


def prices(oracle, note, s):
    m = market().bsm_market()
    return (oracle.autocallable_price(s, note._terms(), m, workers=8),
            oracle.autocallable_price_on_shared_grid(s, note._terms(), m, workers=8))


def test_a_bridge_note_steps_daily_on_the_shared_grid(oracle):
    scalar, shared = prices(oracle, snowball(0.06), settings(1 << 15))
    assert scalar.price != shared.price
    assert abs(scalar.price - shared.price) < 4 * np.hypot(scalar.standard_error, shared.standard_error)


def test_an_event_off_the_daily_times_keeps_the_own_plan(oracle):
    note = snowball(0.06, dates=[dt.datetime(2025, 2, 3, 12), dt.datetime(2025, 3, 3), END])
    scalar, shared = prices(oracle, note, settings(4096))
    assert shared.price == scalar.price


def test_a_note_without_a_bridge_is_priced_alike(oracle):
    """Knock-out fixings every fifth trading day leave no gap of eight steps,
    so the scalar plan steps daily as well: identical prices."""
    cal = TradingCalendar.weekdays()
    dates = [cal.add_trading_days(dt.datetime(2025, 1, 1), 5 * k) for k in range(1, 13)]
    note = SnowballOption(maturity_coupon_rate=0.06, initial_spot=100.0, knock_in_level=95.0, upper_strike=100.0,
                          lower_strike=60.0, knock_in_observation=KnockInObservation.every_trading_day(dt.time(0, 0)),
                          effective_time=dt.datetime(2025, 1, 1), expiry_time=dates[-1],
                          autocall_observations=[SnowballObservation(d, 110.0, 0.08) for d in dates],
                          barrier_state="no_event", calendar=cal)
    scalar, shared = prices(oracle, note, settings(4096))
    assert shared.price == scalar.price


def test_at_expiry_knock_in_has_no_daily_grid(oracle):
    note = snowball(0.06)
    note.knock_in_observation = KnockInObservation.at_expiry()
    scalar, shared = prices(oracle, note, settings(4096))
    assert shared.price == scalar.price


@pytest.mark.parametrize("sampling", [_abi.PSEUDO_RANDOM, _abi.SOBOL])
def test_shared_grid_prices_are_affine_in_the_coupon(oracle, sampling):
    """On fixed draws the snowball pays the maturity coupon linearly, so the
    family's prices lie on a line: the interpolation the reference's coupon
    families use reproduces each member to rounding."""
    s = settings(2048, sampling=sampling, antithetic=1)
    m = market().bsm_market()
    p = {c: oracle.autocallable_price_on_shared_grid(s, snowball(c)._terms(), m, workers=8).price
         for c in set(COUPONS) | {0.06}}
    low, high = min(p), max(p)
    for c, price in p.items():
        f = (c - low) / (high - low)
        assert price == pytest.approx(p[low] * (1 - f) + p[high] * f, rel=1e-12, abs=1e-14)
    assert abs(p[0.08] - p[-0.01]) > 1e-4
