"""Pins the oracle's accumulator estimator (CPU) with the reference's own
accumulator tests (crates/kiyosi/tests/accumulator.rs,
accumulator_extension.rs): exact settlements at the valuation date, the
near-deterministic prices at volatility 1e-8, the history and schedule
errors, and, for what the seed leaves open, the closed forms those tests
compare against (one remaining fixing, or two fixed-purchase events)."""

import datetime as dt
import math

import pytest
from scipy.stats import norm

from conftest import us
from kiyosi_amd import _abi
from kiyosi_amd.errors import ComputationFailureError, InvalidInputError, MissingHistoryError
from kiyosi_amd.instruments import Accumulator
from kiyosi_amd.market import TradingCalendar


def day(d, month=1, hour=0, micros=0):
    return dt.datetime(2025, month, d, hour) + dt.timedelta(microseconds=micros)


def settings(paths, seed=42, antithetic=False, sampling=_abi.PSEUDO_RANDOM):
    return _abi.Settings(paths, seed, 1, sampling, int(antithetic), 0, 0, 1)


def market(spot, at, sigma, rate=0.04, dividend=0.01):
    return _abi.BsmMarket(spot, rate, dividend, sigma, us(at))


def contract(direction="call", *, strike=100.0, level=None, daily=1.0, accel=2.0, start=day(1), end=day(6),
             held=3.0, settled=False, schedule=None, calendar=None):
    """accumulator_extension.rs:11-31 (terms): 100 strike, knock-out 110 for
    calls and 90 for puts, one unit a day doubled beyond the strike, three
    units held, not knocked out."""
    if level is None:
        level = 110.0 if direction == "call" else 90.0
    return Accumulator(strike=strike, knock_out_level=level, daily_quantity=daily, acceleration_factor=accel,
                       effective_time=start, expiry_time=end, accumulated_quantity=held, option_type=direction,
                       calendar=calendar, knock_out_observation_schedule=schedule, knock_out_settled=settled)


def price(oracle, option, mk, paths=1 << 15, seed=42):
    return oracle.accumulator_price(settings(paths, seed), option._terms(), mk, workers=8)


def years(a, b):
    return (b - a).total_seconds() / (365 * 86400)


class Lognormal:
    """Discounted band expectations of S_T under BSM from the valuation:
    band(lo, hi) = E[e^{-rT} (S_T - K) 1{lo <= S_T < hi}], from the
    asset-or-nothing and cash-or-nothing put closed forms."""

    def __init__(self, spot, strike, rate, dividend, sigma, T):
        self.s, self.k, self.r, self.q, self.w, self.T = spot, strike, rate, dividend, sigma * math.sqrt(T), T

    def below(self, level):  # (E[e^{-rT} S_T 1{S_T < L}], E[e^{-rT} 1{S_T < L}])
        if level == math.inf:
            return self.s * math.exp(-self.q * self.T), math.exp(-self.r * self.T)
        if level <= 0:
            return 0.0, 0.0
        d1 = (math.log(self.s / level) + (self.r - self.q) * self.T) / self.w + 0.5 * self.w
        return self.s * math.exp(-self.q * self.T) * norm.cdf(-d1), math.exp(-self.r * self.T) * norm.cdf(-(d1 - self.w))

    def band(self, lo, hi):
        (a_hi, c_hi), (a_lo, c_lo) = self.below(hi), self.below(lo)
        return (a_hi - a_lo) - self.k * (c_hi - c_lo)


def terminal_value(option, spot, rate, dividend, sigma, T, held, monitored):
    """One fixing remains, at expiry (accumulator_extension.rs:76-139): the
    held units settle as a forward; the last purchase (accelerated beyond the
    strike) settles unless the expiry observation knocks out."""
    ln = Lognormal(spot, option.strike, rate, dividend, sigma, T)
    k, lv, a, d = option.strike, option.knock_out_level if monitored else None, option.acceleration_factor, option.daily_quantity
    if option.option_type == "call":
        top = lv if monitored else math.inf
        purchase = a * ln.band(0.0, min(k, top)) + (ln.band(k, top) if top > k else 0.0)
        return held * ln.band(0.0, math.inf) + d * purchase
    bottom = lv if monitored else 0.0
    purchase = a * ln.band(max(k, bottom), math.inf) + (ln.band(bottom, k) if bottom < k else 0.0)
    return -(held * ln.band(0.0, math.inf) + d * purchase)


# ------------------------------------------------------------ known settlements
@pytest.mark.parametrize("spot,quantity", [
    (90.0, 5.0), (100.0, 4.0), (109.0, 4.0), (110.0, 3.0), (111.0, 3.0),
    (math.nextafter(100.0, 0.0), 5.0), (math.nextafter(100.0, 200.0), 4.0),
    (math.nextafter(110.0, 0.0), 4.0), (math.nextafter(110.0, 200.0), 3.0)])
def test_expiry_settlement_thresholds(oracle, spot, quantity):
    """accumulator.rs:291-311: at expiry the held units plus the day's
    purchase (none on a knock-out) settle, with exact strike and level
    comparisons."""
    option = contract(level=110.0)
    est = price(oracle, option, market(spot, day(6), 0.2))
    assert est.price == quantity * (spot - 100.0)
    assert est.has_seed == 0 and est.standard_error == 0.0


@pytest.mark.parametrize("direction,spot,monitored,expected", [
    ("call", 90.0, True, -50.0), ("call", 100.0, True, 0.0), ("call", 109.0, True, 36.0),
    ("call", 110.0, True, 30.0), ("call", 111.0, True, 33.0), ("call", 110.0, False, 40.0),
    ("put", 110.0, True, -50.0), ("put", 100.0, True, 0.0), ("put", 91.0, True, 36.0),
    ("put", 90.0, True, 30.0), ("put", 89.0, True, 33.0), ("put", 90.0, False, 40.0)])
def test_expiry_events_follow_direction_and_schedule(oracle, direction, spot, monitored, expected):
    """accumulator_extension.rs:177-223: an explicit schedule with expiry
    knocks out there; an empty schedule never knocks out."""
    option = contract(direction, schedule=[day(6)] if monitored else [])
    assert price(oracle, option, market(spot, day(6), 0.2)).price == expected


@pytest.mark.parametrize("direction,spot", [("call", 90.0), ("put", 110.0)])
def test_expiry_purchase_with_partial_multiplier(oracle, direction, spot):
    """accumulator_extension.rs:224-244: effective on Friday the 3rd, a
    quarter unit a day at half rate beyond the strike."""
    option = contract(direction, start=day(3), daily=0.25, accel=0.5, held=0.5, schedule=[])
    assert price(oracle, option, market(spot, day(6), 0.2)).price == -6.25


def test_knock_out_on_the_valuation_fixing_settles_now(oracle):
    """accumulator.rs:328-334: 110 on the 2nd knocks out three held units."""
    assert price(oracle, contract(level=110.0), market(110.0, day(2), 0.2)).price == 30.0


def test_outstanding_units_with_zero_curves_settle_as_a_forward(oracle):
    """accumulator.rs:19-52: valued at noon on Friday the 3rd, no fixing
    remains before the Saturday expiry and zero rates leave 3 x 5."""
    option = contract(start=day(3), end=day(4), level=130.0, held=3.0)
    est = price(oracle, option, market(105.0, day(3, hour=12), 0.2, 0.0, 0.0))
    assert est.price == 15.0 and est.standard_error == 0.0


@pytest.mark.parametrize("direction,level", [("call", 90.0), ("put", 110.0)])
@pytest.mark.parametrize("schedule", [None, [day(1), day(2), day(3)], [day(2), day(3)]])
def test_zero_multiplier_below_an_inverted_level_never_accrues(oracle, direction, level, schedule):
    """accumulator_extension.rs:420-479: every remaining fixing is a
    knock-out observation at which the strike itself knocks out, and a zero
    multiplier buys nothing below it: the value is known, with zero error."""
    option = contract(direction, level=level, accel=0.0, end=day(3), held=0.0, schedule=schedule)
    est = oracle.accumulator_price(settings(1), option._terms(), market(100.0, day(1, hour=12), 0.2), workers=1)
    assert est.price == 0.0 and est.standard_error == 0.0 and est.has_standard_error == 1
    holding = contract(direction, level=level, accel=0.0, end=day(3), held=3.0, schedule=schedule)
    spot = 85.0 if direction == "call" else 115.0
    est = oracle.accumulator_price(settings(1), holding._terms(), market(spot, day(1, hour=12), 0.2, 0.0, 0.0),
                                   workers=1)
    assert est.price == -45.0 and est.standard_error == 0.0


# ------------------------------------------------------ near-deterministic paths
@pytest.mark.parametrize("direction,spot", [("call", 120.0), ("put", 80.0)])
@pytest.mark.parametrize("schedule,expected", [(None, 60.0), ([day(3)], 80.0), ([day(6)], 100.0), ([], 120.0)])
def test_sparse_knock_out_dates_keep_daily_accrual(oracle, direction, spot, schedule, expected):
    """accumulator_extension.rs:246-272: volatility 1e-8 from noon on the
    1st; daily knock-out stops on the 2nd, explicit dates only on theirs."""
    option = contract(direction, schedule=schedule)
    got = price(oracle, option, market(spot, day(1, hour=12), 1e-8, 0.0, 0.0), paths=256).price
    assert got == pytest.approx(expected, abs=2e-7)


@pytest.mark.parametrize("direction,level,spot", [("call", 95.0, 120.0), ("put", 105.0, 80.0)])
def test_zero_multiplier_with_inverted_level_accrues_off_schedule(oracle, direction, level, spot):
    """accumulator_extension.rs:551-575: no knock-out dates, so the three
    remaining fixings each buy one unit."""
    option = contract(direction, level=level, accel=0.0, held=0.0, schedule=[])
    got = price(oracle, option, market(spot, day(1, hour=12), 1e-8, 0.0, 0.0), paths=256).price
    assert got == pytest.approx(60.0, abs=2e-7)


@pytest.mark.parametrize("accel", [1.0, 2.0])
@pytest.mark.parametrize("schedule,units,ko_day", [(None, 3.0, 2), ([day(15)], 12.0, 15)])
def test_knock_out_settles_with_its_own_date_discount(oracle, accel, schedule, units, ko_day):
    """At volatility 1e-8 and equal rates the spot stays at 120: the units
    held at the first knock-out date settle there, discounted from that
    date, whether purchases are grouped (unit multiplier) or daily."""
    at = day(1, hour=12)
    option = contract(accel=accel, end=dt.datetime(2025, 2, 3), schedule=schedule)
    got = price(oracle, option, market(120.0, at, 1e-8, 0.2, 0.2), paths=256).price
    assert got == pytest.approx(units * 20.0 * math.exp(-0.2 * years(at, day(ko_day))), abs=1e-5)


# --------------------------------------------------------------- equalities
@pytest.mark.parametrize("direction", ["call", "put"])
def test_explicit_daily_dates_equal_the_daily_rule(oracle, direction):
    """accumulator_extension.rs:577-602: listing every fixing as a knock-out
    date prices the same paths as the daily rule, bit for bit."""
    mk = market(100.25, day(1, hour=12), 0.4)
    daily = price(oracle, contract(direction), mk)
    explicit = price(oracle, contract(direction, schedule=[day(1), day(2), day(3), day(6)]), mk)
    assert daily.price == explicit.price and daily.standard_error == explicit.standard_error


@pytest.mark.parametrize("direction,spot,quantity", [
    ("call", math.nextafter(100.0, 0.0), 5.0), ("call", 100.0, 4.0), ("call", math.nextafter(100.0, 200.0), 4.0),
    ("call", 120.0, 4.0), ("put", math.nextafter(100.0, 0.0), 4.0), ("put", 100.0, 4.0),
    ("put", math.nextafter(100.0, 200.0), 5.0), ("put", 80.0, 4.0)])
def test_purchase_at_the_valuation_fixing_counts_once(oracle, direction, spot, quantity):
    """accumulator_extension.rs:727-773: valued on the 3rd's fixing the day's
    purchase joins the held units (strict strike comparisons); a microsecond
    later the same units are history."""
    at = price(oracle, contract(direction, schedule=[]), market(spot, day(3), 0.4))
    after = price(oracle, contract(direction, schedule=[], held=quantity), market(spot, day(3, micros=1), 0.4))
    assert at.price == pytest.approx(after.price, abs=1e-7)


# ------------------------------------------------------------- closed forms
@pytest.mark.parametrize("direction", ["call", "put"])
@pytest.mark.parametrize("level", [95.0, 100.0, 105.0])
@pytest.mark.parametrize("accel", [0.0, 0.5, 2.0])
@pytest.mark.parametrize("monitored", [False, True])
def test_one_remaining_fixing_matches_the_terminal_portfolio(oracle, direction, level, accel, monitored):
    """accumulator_extension.rs:775-808: from noon on Friday the 3rd only the
    expiry fixing remains, so the price is a forward plus strike and level
    digitals, with no grid error: the estimate meets it within 4 sigma."""
    option = contract(direction, level=level, accel=accel, schedule=[day(6)] if monitored else [])
    at = day(3, hour=12)
    est = price(oracle, option, market(100.25, at, 0.6), paths=200_000)
    expected = terminal_value(option, 100.25, 0.04, 0.01, 0.6, years(at, day(6)), 3.0, monitored)
    assert abs(est.price - expected) < 4 * est.standard_error


def test_fixed_purchases_before_a_terminal_knock_out(oracle):
    """accumulator_extension.rs:140-175: every day trades, unit
    acceleration and one knock-out date at expiry group the 3rd to 5th
    onto it (plus the purchase on the 2nd at valuation): seven units held,
    one more bought unless the 6th knocks out."""
    option = contract(accel=1.0, schedule=[day(6)], calendar=TradingCalendar.all_days())
    est = price(oracle, option, market(108.0, day(2), 0.2), paths=200_000)
    expected = terminal_value(option, 108.0, 0.04, 0.01, 0.2, years(day(2), day(6)), 7.0, True)
    assert abs(est.price - expected) < 4 * est.standard_error


@pytest.mark.parametrize("sigma", [0.0008, 0.2])
@pytest.mark.parametrize("direction,spot,level,rate,dividend", [("call", 110.05, 110.0, 0.04, 0.06),
                                                              ("put", 89.95, 90.0, 0.06, 0.04)])
def test_two_fixed_purchase_events_match_the_conditional_portfolio(oracle, direction, spot, level, rate, dividend,
                                                                   sigma):
    """accumulator_extension.rs:355-418: one knock-out date on the 15th of a
    month to February 3rd: twelve units settle at a knock-out there,
    otherwise all 26 at expiry, so only two dates are simulated and the
    price is exact in distribution."""
    expiry = dt.datetime(2025, 2, 3)
    option = contract(direction, level=level, accel=1.0, end=expiry, schedule=[day(15)])
    at = day(1, hour=12)
    est = price(oracle, option, market(spot, at, sigma, rate, dividend), paths=200_000)
    t1, T = years(at, day(15)), years(at, expiry)
    ln = Lognormal(spot, 100.0, rate, dividend, sigma, t1)
    alive = (ln.below(level) if direction == "call" else
             tuple(x - y for x, y in zip(ln.below(math.inf), ln.below(level))))
    # Alive on the 15th, the 26 units' forward there carries to expiry.
    alive_value = alive[0] * math.exp(-dividend * (T - t1)) - 100.0 * alive[1] * math.exp(-rate * (T - t1))
    knocked = ln.band(level, math.inf) if direction == "call" else ln.band(0.0, level)
    sign = 1.0 if direction == "call" else -1.0
    expected = sign * (12.0 * knocked + 26.0 * alive_value)
    assert abs(est.price - expected) < 4 * est.standard_error + 1e-9


def test_grouped_purchases_equal_the_daily_path_in_distribution(oracle):
    """A weekly knock-out schedule at unit acceleration: the grouped estimator
    (knock-out dates only) and the same contract with a 1.0000001 multiplier,
    simulated daily, agree within their errors."""
    fridays = [day(3) + dt.timedelta(days=7 * k) for k in range(8)]
    end = fridays[-1]
    grouped = contract(level=104.0, accel=1.0, end=end, schedule=fridays, held=0.0)
    daily = contract(level=104.0, accel=1.0000001, end=end, schedule=fridays, held=0.0)
    mk = market(100.0, day(1), 0.3)
    a = price(oracle, grouped, mk, paths=400_000, seed=5)
    b = price(oracle, daily, mk, paths=400_000, seed=6)
    assert abs(a.price - b.price) < 4 * math.hypot(a.standard_error, b.standard_error) + 1e-3


# ------------------------------------------------------------------- errors
def test_valuation_outside_the_life_and_overflowing_settlements(oracle):
    """accumulator.rs:343-388."""
    option = contract(level=110.0)
    for at in (dt.datetime(2024, 12, 31), day(6, micros=1)):
        with pytest.raises(InvalidInputError):
            price(oracle, option, market(100.0, at, 0.2))
    with pytest.raises(ComputationFailureError):
        price(oracle, contract(level=110.0, held=1.7976931348623157e308), market(111.0, day(6), 0.2))
    with pytest.raises(ComputationFailureError):
        price(oracle, option, market(100.0, day(3, hour=12), 1e308))


def test_knock_out_settlement_history(oracle):
    """accumulator.rs:55-118: after knock-out observations began the
    settlement state is required; a settled contract before any observation
    is invalid; an empty explicit schedule never needs it."""
    start, end = day(6), day(10)
    unknown = contract(accel=1.0, start=start, end=end, held=2.0, settled=None)
    mk = market(105.0, day(8), 0.2)
    with pytest.raises(MissingHistoryError):
        price(oracle, unknown, mk)
    price(oracle, contract(accel=1.0, start=start, end=end, held=None, settled=None), market(105.0, start, 0.2))
    with pytest.raises(InvalidInputError):
        price(oracle, contract(accel=1.0, start=start, end=end, held=2.0, settled=True), market(105.0, start, 0.2))
    assert price(oracle, contract(accel=1.0, start=start, end=end, held=2.0, settled=True), mk).price == 0.0
    price(oracle, contract(accel=1.0, start=start, end=end, held=2.0, settled=None, schedule=[]), mk)


@pytest.mark.parametrize("start", [day(2), day(4)])
def test_quantity_history(oracle, start):
    """accumulator.rs:917-1017: before the first fixing no quantity may be
    held; after it the quantity is required."""
    first = day(6) if start == day(4) else start
    base = dict(strike=90.0, level=150.0, start=start, end=day(7))
    price(oracle, contract(**base, held=None), market(100.0, first, 0.2, 0.0, 0.0))
    for at in (start, first):
        with pytest.raises(InvalidInputError):
            price(oracle, contract(**base, held=3.0), market(100.0, at, 0.2, 0.0, 0.0))
    noon = day(first.day, hour=12)
    with pytest.raises(MissingHistoryError):
        price(oracle, contract(**base, held=None), market(100.0, noon, 0.2, 0.0, 0.0))
    price(oracle, contract(**base, held=0.0), market(100.0, noon, 0.2, 0.0, 0.0))
    with pytest.raises(MissingHistoryError):
        price(oracle, contract(**base, held=None), market(100.0, day(7), 0.2, 0.0, 0.0))
    assert price(oracle, contract(**base, held=0.0), market(100.0, day(7), 0.2, 0.0, 0.0)).price == 10.0
    assert price(oracle, contract(**base, held=10.0), market(100.0, day(7), 0.2, 0.0, 0.0)).price == 110.0


@pytest.mark.parametrize("times", [[dt.datetime(2024, 12, 31)], [day(7)], [day(4)], [day(3), day(2)]])
def test_knock_out_dates_outside_life_off_calendar_or_unordered(oracle, times):
    """accumulator_extension.rs:665-676, 700-706."""
    with pytest.raises(InvalidInputError):
        price(oracle, contract("put", schedule=times), market(100.0, day(1, hour=12), 0.2))
