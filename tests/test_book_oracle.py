"""Pins the oracle's barrier, Asian and accumulator estimators (CPU): the
reference's own per-seed barrier prices, closed forms, and independent
brute-force simulations."""

import ctypes as C
import datetime as dt
import math

import numpy as np
import pytest
from scipy.stats import norm

from conftest import END, START, golden, us
from kiyosi_amd import _abi


def market(spot=100.0, rate=0.05, dividend=0.02, sigma=0.2):
    return _abi.BsmMarket(spot, rate, dividend, sigma, us(START))


def settings(paths, seed=1009, antithetic=False, sampling=_abi.PSEUDO_RANDOM):
    return _abi.Settings(paths, seed, 1, sampling, int(antithetic), 0, 0, 1)


def barrier(option_type, barrier_type, level, strike=100.0, rebate=0.0, observation=_abi.OBSERVE_CONTINUOUS,
            times=None, at_hit=False):
    arr = None
    n = 0
    if times is not None:
        arr = (C.c_int64 * len(times))(*times)
        n = len(times)
    t = _abi.BarrierTerms(option_type, barrier_type, observation, int(at_hit), strike, level, rebate,
                          us(START), us(END), n, 0, C.cast(arr, C.POINTER(C.c_int64)) if arr else None)
    t._keep = arr
    return t


@pytest.mark.parametrize("case_id", ["mc-barrier-down-out-call", "mc-barrier-up-in-put"])
def test_continuous_barrier_calibration_vectors(oracle, case_id):
    """Per-seed prices of kiyosi's continuous-bridge barrier estimator."""
    for run in [r for r in golden("calibration_monte_carlo.json")["runs"] if r["id"] == case_id]:
        i = run["inputs"]
        kind = 1 if i["option"] == "call" else -1
        btype = {("down", "out"): _abi.DOWN_AND_OUT, ("up", "in"): _abi.UP_AND_IN}[(i["direction"], i["activation"])]
        terms = barrier(kind, btype, float(i["barrier"]))
        for sample in run["samples"][:5] + [run["timing_seed"]]:
            est = oracle.barrier_price(settings(run["path_count"], sample["seed"]), terms, market(), workers=8)
            assert est.price == pytest.approx(sample["price"], rel=1e-13)


def daily(start=START, end=END):
    out, d = [], start
    while d < end:
        d += dt.timedelta(days=1)
        out.append(us(d))
    return out


def test_scheduled_barrier_approaches_the_continuous_one(oracle):
    """Daily fixings with the Broadie-Glasserman-Kou shift e^{-0.5826 sigma sqrt(dt)}
    of the barrier reproduce the continuous analytic down-and-out price."""
    ref = {c["id"]: c["value"] for c in golden("references.json")["monte_carlo"]}["mc-barrier-down-out-call"]
    times = daily()
    shifted = 90.0 * math.exp(0.5826 * 0.2 * math.sqrt(1 / 365))
    est = oracle.barrier_price(settings(200_000), barrier(1, _abi.DOWN_AND_OUT, shifted,
                                                          observation=_abi.OBSERVE_SCHEDULED, times=times),
                               market(), workers=8)
    assert abs(est.price - ref) < 4 * est.standard_error + 0.01


def test_scheduled_in_out_parity(oracle):
    """Knock-in (conditioned on the hit) plus knock-out equals the vanilla call."""
    times = daily()
    s = settings(200_000, seed=3)
    out = oracle.barrier_price(s, barrier(1, _abi.DOWN_AND_OUT, 92.0, observation=_abi.OBSERVE_SCHEDULED,
                                          times=times), market(), workers=8)
    kin = oracle.barrier_price(s, barrier(1, _abi.DOWN_AND_IN, 92.0, observation=_abi.OBSERVE_SCHEDULED,
                                          times=times), market(), workers=8)
    vanilla = {c["id"]: c["value"] for c in golden("references.json")["monte_carlo"]}["mc-european-call"]
    assert abs(out.price + kin.price - vanilla) < 4 * math.hypot(out.standard_error, kin.standard_error)


def test_knocked_out_at_valuation_pays_the_rebate(oracle):
    est = oracle.barrier_price(settings(64), barrier(1, _abi.UP_AND_OUT, 99.0, rebate=2.0), market())
    assert est.price == pytest.approx(2.0 * math.exp(-0.05), rel=1e-14)
    assert est.has_seed == 0


def test_geometric_asian_matches_the_continuous_closed_form(oracle):
    ref = {c["id"]: c for c in golden("references.json")["digitals_asian"]}["geometric-asian-bsm-call"]
    m = _abi.BsmMarket(100.0, 0.04, 0.01, 0.2, us(START))
    t = _abi.AsianTerms(1, _abi.AVERAGE_GEOMETRIC, 100, 0, 100.0, 0.0, us(START), us(START), us(END))
    est = oracle.asian_price(settings(400_000), t, m, workers=8)
    assert abs(est.price - ref["value"]) < 4 * est.standard_error


def test_arithmetic_asian_matches_a_brute_force_trapezoid_simulation(oracle):
    """The reference's arithmetic-Asian fixture is a two-moment lognormal fit
    (an approximation), so the pin is an independent trapezoid simulation with
    its own geometric control."""
    n, r, q, s = 50, 0.04, 0.01, 0.2
    m = _abi.BsmMarket(100.0, r, q, s, us(START))
    t = _abi.AsianTerms(1, _abi.AVERAGE_ARITHMETIC, n, 0, 100.0, 0.0, us(START), us(START), us(END))
    est = oracle.asian_price(settings(100_000), t, m, workers=8)
    rng = np.random.default_rng(11)
    dt_, P = 1 / n, 400_000
    lg = np.cumsum((r - q - 0.5 * s * s) * dt_ + s * math.sqrt(dt_) * rng.standard_normal((P, n)), axis=1)
    S = 100 * np.exp(lg)
    A = (50.0 + S[:, :-1].sum(1) + 0.5 * S[:, -1]) / n
    G = np.exp((0.5 * math.log(100) + np.log(S[:, :-1]).sum(1) + 0.5 * np.log(S[:, -1])) / n)
    a, g = np.maximum(A - 100, 0) * math.exp(-r), np.maximum(G - 100, 0) * math.exp(-r)
    mean = (r - q - 0.5 * s * s) * dt_ * n / 2
    var = s * s * dt_ * (n / 3 - 1 / (12 * n))
    lf, w = math.log(100) + mean + 0.5 * var, math.sqrt(var)
    d1 = (lf - math.log(100)) / w + 0.5 * w
    eg = math.exp(-r) * (math.exp(lf) * norm.cdf(d1) - 100 * norm.cdf(d1 - w))
    b = np.cov(a, g)[0, 1] / np.var(g)
    cv = a - b * (g - eg)
    brute, brute_se = cv.mean(), cv.std() / math.sqrt(P)
    assert abs(est.price - brute) < 4 * math.hypot(est.standard_error, brute_se)


def accumulator(kind=1, strike=95.0, ko=110.0, qty=1.0, accel=2.0, start=START, end=END):
    return _abi.AccumulatorTerms(kind, _abi.CALENDAR_WEEKDAYS, strike, ko, qty, accel, 0.0, 0, us(start), us(end))


def test_accumulator_without_knock_out_or_acceleration_is_a_forward_strip(oracle):
    """With an unreachable knock-out and unit acceleration every fixing adds
    one unit settled at expiry: value = N (S e^{-qT} - K e^{-rT})."""
    end = START + dt.timedelta(days=91)
    t = accumulator(strike=95.0, ko=1e9, accel=1.0, end=end)
    est = oracle.accumulator_price(settings(200_000), t, market(), workers=8)
    days = sum(1 for k in range(1, 92) if (START + dt.timedelta(days=k)).weekday() < 5)
    T = 91 / 365
    expected = days * (100 * math.exp(-0.02 * T) - 95 * math.exp(-0.05 * T))
    assert abs(est.price - expected) < 4 * est.standard_error


def test_accumulator_matches_a_brute_force_daily_simulation(oracle):
    end = START + dt.timedelta(days=120)
    t = accumulator(strike=97.0, ko=106.0, accel=2.0, end=end)
    est = oracle.accumulator_price(settings(200_000), t, market(sigma=0.25), workers=8)
    fix = [START + dt.timedelta(days=k) for k in range(1, 121) if (START + dt.timedelta(days=k)).weekday() < 5]
    times = np.array([(d - START).days / 365 for d in fix])
    dts = np.diff(np.concatenate([[0.0], times]))
    rng, P = np.random.default_rng(3), 200_000
    S = np.full(P, 100.0)
    q = np.zeros(P)
    alive = np.ones(P, bool)
    val = np.zeros(P)
    for k, d in enumerate(dts):
        S = S * np.exp((0.05 - 0.02 - 0.5 * 0.25**2) * d + 0.25 * math.sqrt(d) * rng.standard_normal(P))
        out = alive & (S >= 106.0)
        val[out] = (q * (S - 97.0) * math.exp(-0.05 * times[k]))[out]
        alive &= ~out
        q = np.where(alive, q + np.where(S < 97.0, 2.0, 1.0), q)
    val[alive] = (q * (S - 97.0) * math.exp(-0.05 * times[-1]))[alive]
    brute, brute_se = val.mean(), val.std() / math.sqrt(P)
    assert abs(est.price - brute) < 4 * math.hypot(est.standard_error, brute_se)


def geometric_closed_form(S, K, r, q, sig, T, kind):
    """Continuously sampled geometric average from 0 to T: log G is normal
    with mean log S + (r - q - sig^2/2) T/2 and variance sig^2 T/3."""
    w = sig * math.sqrt(T / 3)
    lf = math.log(S) + (r - q - 0.5 * sig * sig) * T / 2 + 0.5 * w * w
    d1 = (lf - math.log(K)) / w + 0.5 * w
    disc, F = math.exp(-r * T), math.exp(lf)
    price = disc * kind * (F * norm.cdf(kind * d1) - K * norm.cdf(kind * (d1 - w)))
    return price, disc * kind * F / S * norm.cdf(kind * d1), disc * F / S**2 * norm.pdf(d1) / w


@pytest.mark.parametrize("kind", [1, -1])
def test_geometric_asian_native_greeks(oracle, kind):
    """Delta is the per-path derivative of the sampled payoff, so it equals a
    common-random-number difference of prices; gamma (a likelihood-ratio
    form) meets the closed form's gamma within the sampling error."""
    t = _abi.AsianTerms(kind, _abi.AVERAGE_GEOMETRIC, 100, 0, 100.0, 0.0, us(START), us(START), us(END))
    s = settings(400_000, antithetic=True)
    mk = lambda spot: _abi.BsmMarket(spot, 0.04, 0.01, 0.2, us(START))
    est, g = oracle.asian_price_with_greeks(s, t, mk(100.0), workers=8)
    assert g.available == 3
    assert est.price == oracle.asian_price(s, t, mk(100.0), workers=8).price
    h = 0.01
    fd = (oracle.asian_price(s, t, mk(100 + h), workers=8).price
          - oracle.asian_price(s, t, mk(100 - h), workers=8).price) / (2 * h)
    assert abs(g.delta - fd) < 1e-4
    _, delta, gamma = geometric_closed_form(100.0, 100.0, 0.04, 0.01, 0.2, 1.0, kind)
    assert abs(g.delta - delta) < 3e-3
    assert g.gamma == pytest.approx(gamma, rel=0.02)
    only_gamma = oracle.asian_price_with_greeks(s, t, mk(100.0), request=2, workers=8)[1]
    assert only_gamma.available == 2 and only_gamma.gamma == g.gamma and only_gamma.delta == 0.0


@pytest.mark.parametrize("kind", [1, -1])
def test_arithmetic_asian_native_greeks_are_derivatives_of_the_conditional_price(oracle, kind):
    """Each path's arithmetic estimate is a smooth Black-type price of the
    last interval, and its delta and gamma are that price's first and second
    spot derivatives (with the geometric control's likewise), so both equal
    common-random-number differences of prices on the same draws."""
    t = _abi.AsianTerms(kind, _abi.AVERAGE_ARITHMETIC, 50, 0, 100.0, 0.0, us(START), us(START), us(END))
    s = settings(20_000, antithetic=True)
    mk = lambda spot: _abi.BsmMarket(spot, 0.04, 0.01, 0.2, us(START))
    price = lambda spot: oracle.asian_price(s, t, mk(spot), workers=8).price
    est, g = oracle.asian_price_with_greeks(s, t, mk(100.0), workers=8)
    assert g.available == 3 and est.price == price(100.0)
    h = 0.01
    assert g.delta == pytest.approx((price(100 + h) - price(100 - h)) / (2 * h), rel=1e-6)
    h = 0.05
    assert g.gamma == pytest.approx((price(100 + h) - 2 * est.price + price(100 - h)) / h**2, rel=1e-4)
    _, delta, gamma = geometric_closed_form(100.0, 100.0, 0.04, 0.01, 0.2, 1.0, kind)
    assert 0 < kind * g.delta < kind * delta + 0.05 and 0.8 * gamma < g.gamma < 1.3 * gamma

