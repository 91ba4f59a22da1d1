"""Pins of the oracle's double-barrier, digital and likelihood-ratio Greek
restatements (engines/monte_carlo/{barrier,double_barrier,digital}.rs).

- Double barriers: the reference's own golden price, delta and gamma of the
  continuous knock-out call and put (tests/double_barrier_monte_carlo.rs
  :121-157) and the analytic double-barrier prices of
  benchmarks/quantlib/references/barriers.json.
- The survival series: image and spectral forms agree where both converge,
  central differences give its derivatives, and a far upper level reduces it
  to the single-barrier bridge formula.
- Digitals: the analytic prices of references/digitals_asian.json and the
  Black-Scholes closed forms of delta and gamma.
- Single-barrier delta and gamma: common-random-number finite differences of
  the oracle's own prices (an independent estimator of the same derivative).
- Scheduled double barriers observed only at expiry: the closed-form
  corridor option.
"""

import ctypes as C
import datetime as dt
import math

import pytest

from conftest import END, START, golden, us
from kiyosi_amd import _abi


def market(spot=100.0, rate=0.04, dividend=0.01, sigma=0.2):
    return _abi.BsmMarket(spot, rate, dividend, sigma, us(START))


def settings(paths, seed=7, antithetic=False, sampling=_abi.PSEUDO_RANDOM):
    return _abi.Settings(paths, seed, 1, sampling, int(antithetic), _abi.FP64, 0, 0)


def double_terms(option_type, knock, lower=80.0, upper=120.0, strike=100.0, rebate=0.0, times=None):
    arr = (C.c_int64 * len(times))(*times) if times else None
    t = _abi.DoubleBarrierTerms(option_type, knock, _abi.OBSERVE_SCHEDULED if times else _abi.OBSERVE_CONTINUOUS,
                                _abi.REBATE_AT_EXPIRY, strike, lower, upper, rebate, us(START), us(END),
                                len(times) if times else 0, 0,
                                C.cast(arr, C.POINTER(C.c_int64)) if arr else None)
    t._keep = arr
    return t


def barrier_terms(option_type, barrier_type, level, strike=100.0, rebate=0.0, times=None):
    arr = (C.c_int64 * len(times))(*times) if times else None
    t = _abi.BarrierTerms(option_type, barrier_type, _abi.OBSERVE_SCHEDULED if times else _abi.OBSERVE_CONTINUOUS,
                          _abi.REBATE_AT_EXPIRY, strike, level, rebate, us(START), us(END),
                          len(times) if times else 0, 0, C.cast(arr, C.POINTER(C.c_int64)) if arr else None)
    t._keep = arr
    return t


def ncdf(x):
    return 0.5 * math.erfc(-x / math.sqrt(2.0))


def npdf(x):
    return math.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def reference(name, case):
    return next(c for c in golden("references.json")[name] if c["id"] == case)


@pytest.mark.parametrize("kind", ["call", "put"])
def test_double_knock_out_matches_the_reference_golden_greeks(oracle, kind):
    g = golden("python_golden.json")["double_barrier_knock_out_greeks"]
    want = g[kind]
    terms = double_terms(1 if kind == "call" else -1, _abi.DOUBLE_KNOCK_OUT)
    for seed in g["inputs"]["seeds"]:
        est, greeks = oracle.double_barrier_price(settings(32768, seed, sampling=_abi.SOBOL), terms, market(),
                                                  request=3, spot_radius=0.01, workers=8)
        assert greeks.available == 3
        assert abs(est.price - want["price"]) < g["tolerance"]
        assert abs(greeks.delta - want["delta"]) < g["tolerance"]
        assert abs(greeks.gamma - want["gamma"]) < g["tolerance"]


@pytest.mark.parametrize("case,knock,option_type", [("double-barrier-out-call", _abi.DOUBLE_KNOCK_OUT, 1),
                                                    ("double-barrier-in-put", _abi.DOUBLE_KNOCK_IN, -1)])
def test_double_barrier_prices_match_the_analytic_references(oracle, case, knock, option_type):
    ref = reference("barriers", case)
    est, _ = oracle.double_barrier_price(settings(1 << 19, 11, antithetic=True), double_terms(option_type, knock),
                                         market(), workers=8)
    assert abs(est.price - ref["value"]) < 4.0 * est.standard_error


def test_knock_in_plus_knock_out_is_the_vanilla_on_the_same_draws(oracle):
    s = settings(1 << 16, 5, sampling=_abi.SOBOL)
    out, _ = oracle.double_barrier_price(s, double_terms(1, _abi.DOUBLE_KNOCK_OUT), market(), workers=8)
    kin, _ = oracle.double_barrier_price(s, double_terms(1, _abi.DOUBLE_KNOCK_IN), market(), workers=8)
    w = 0.2
    d1 = (math.log(100 / 100) + (0.04 - 0.01 + 0.5 * w * w)) / w
    vanilla = 100 * math.exp(-0.01) * ncdf(d1) - 100 * math.exp(-0.04) * ncdf(d1 - w)
    # One terminal draw per path: survival and its complement split each path's payoff.
    assert out.price + kin.price == pytest.approx(vanilla, abs=5e-4)


@pytest.mark.parametrize("variance", [0.05, 0.2, 0.249, 0.25, 0.251, 0.4, 1.0])
@pytest.mark.parametrize("x,y", [(0.3, 0.6), (0.8, 0.7), (0.05, 0.9)])
def test_image_and_spectral_survival_series_agree(oracle, variance, x, y):
    dev = math.sqrt(variance)
    image = oracle.double_survival(x, y, 1.0, dev, True, 0)
    spectral = oracle.double_survival(x, y, 1.0, dev, True, 1)
    for a, b in zip(image, spectral):
        assert a == pytest.approx(b, rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("variance", [0.04, 0.3])
def test_survival_derivatives_are_its_initial_value_derivatives(oracle, variance):
    dev, band, x, y, h = math.sqrt(variance) * 1.7, 1.7, 0.45, 0.9, 1e-4
    f = lambda v: oracle.double_survival(v, y, band, dev, False)[0]
    _, first, second = oracle.double_survival(x, y, band, dev, True)
    assert first == pytest.approx((f(x + h) - f(x - h)) / (2 * h), rel=1e-6)
    assert second == pytest.approx((f(x + h) - 2 * f(x) + f(x - h)) / (h * h), rel=1e-4)


def test_far_upper_level_reduces_to_the_single_barrier_bridge(oracle):
    # P(min of the bridge > 0) = 1 - exp(-2 x y / s^2) for a lone lower level.
    x, y, s = 0.3, 0.5, 0.4
    got = oracle.double_survival(x, y, 50.0, s, False)[0]
    assert got == pytest.approx(-math.expm1(-2 * x * y / (s * s)), rel=1e-12)


def test_survival_outside_the_band_is_zero(oracle):
    assert oracle.double_survival(0.5, 1.2, 1.0, 0.3) == [0.0, 0.0, 0.0]
    assert oracle.double_survival(0.5, -0.1, 1.0, 0.3) == [0.0, 0.0, 0.0]


def test_scheduled_double_barrier_at_expiry_is_the_corridor_call(oracle):
    # Knock-out observed only at expiry: pays (S_T - K)+ on L < S_T < U.
    K, L, U, r, q, w = 100.0, 80.0, 120.0, 0.04, 0.01, 0.2
    terms = double_terms(1, _abi.DOUBLE_KNOCK_OUT, times=[us(END)])
    est, _ = oracle.double_barrier_price(settings(1 << 19, 3), terms, market(), workers=8)

    def d(level, add):
        return (math.log(100 / level) + r - q + add) / w
    # E[e^{-rT} (S_T - K) 1{K < S_T < U}]
    asset = 100 * math.exp(-q) * (ncdf(d(K, 0.5 * w * w)) - ncdf(d(U, 0.5 * w * w)))
    cash = K * math.exp(-r) * (ncdf(d(K, -0.5 * w * w)) - ncdf(d(U, -0.5 * w * w)))
    assert abs(est.price - (asset - cash)) < 4.0 * est.standard_error


@pytest.mark.parametrize("case,option_type,payout_kind,strike,payout,spot", [
    ("digital-bsm-cash-call", 1, _abi.CASH_OR_NOTHING, 100.0, 10.0, 100.0),
    ("digital-bsm-asset-put", -1, _abi.ASSET_OR_NOTHING, 105.0, 0.0, 110.0),
])
def test_digital_matches_the_analytic_references_and_closed_form_greeks(oracle, case, option_type, payout_kind,
                                                                        strike, payout, spot):
    ref = reference("digitals_asian", case)
    terms = _abi.DigitalTerms(option_type, payout_kind, strike, payout, us(START), us(END))
    m = market(spot=spot)
    est, greeks = oracle.digital_price(settings(1 << 20, 9), terms, m, request=3, workers=8)
    assert abs(est.price - ref["value"]) < 4.0 * est.standard_error
    w, s = 0.2, option_type
    d1 = (math.log(spot / strike) + 0.04 - 0.01 + 0.5 * w * w) / w
    d2 = d1 - w
    if payout_kind == _abi.CASH_OR_NOTHING:
        delta = s * payout * math.exp(-0.04) * npdf(d2) / (spot * w)
        gamma = -s * payout * math.exp(-0.04) * npdf(d2) * d1 / (spot * spot * w * w)
    else:
        delta = math.exp(-0.01) * (ncdf(s * d1) + s * npdf(d1) / w)
        gamma = s * math.exp(-0.01) * npdf(d1) / (spot * w) * (1.0 - d1 / w)
    assert greeks.available == 3
    assert greeks.delta == pytest.approx(delta, rel=0.02, abs=2e-4)
    assert greeks.gamma == pytest.approx(gamma, rel=0.05, abs=2e-4)


def test_digital_request_selects_the_greeks(oracle):
    terms = _abi.DigitalTerms(1, _abi.CASH_OR_NOTHING, 100.0, 1.0, us(START), us(END))
    est1, g1 = oracle.digital_price(settings(4096), terms, market(), request=1)
    est2, g2 = oracle.digital_price(settings(4096), terms, market(), request=0)
    assert est1.price == est2.price
    assert g1.available == 1 and g1.gamma == 0.0 and g1.delta != 0.0
    assert g2.available == 0


SINGLE = [
    ("continuous down-out call", lambda: barrier_terms(1, _abi.DOWN_AND_OUT, 85.0)),
    ("continuous up-in put rebate", lambda: barrier_terms(-1, _abi.UP_AND_IN, 115.0, rebate=1.5)),
    ("continuous up-out call", lambda: barrier_terms(1, _abi.UP_AND_OUT, 130.0)),
    ("monthly down-in put", lambda: barrier_terms(-1, _abi.DOWN_AND_IN, 90.0,
                                                 times=[us(START + dt.timedelta(days=30 * k)) for k in
                                                        range(1, 12)] + [us(END)])),
]


@pytest.mark.parametrize("name,make", SINGLE, ids=[c[0] for c in SINGLE])
def test_barrier_likelihood_delta_matches_common_random_number_differences(oracle, name, make):
    terms = make()
    s = settings(1 << 18, 13, antithetic=True)
    est, greeks = oracle.barrier_price_with_greeks(s, terms, market(), request=3, spot_radius=0.01, workers=8)
    assert greeks.available == 3
    h = 1.0
    up = oracle.barrier_price(s, terms, market(spot=100.0 + h), workers=8).price
    down = oracle.barrier_price(s, terms, market(spot=100.0 - h), workers=8).price
    fd_delta = (up - down) / (2 * h)
    assert greeks.delta == pytest.approx(fd_delta, abs=0.01 + 0.03 * abs(fd_delta))
    fd_gamma = (up - 2 * est.price + down) / (h * h)
    assert greeks.gamma == pytest.approx(fd_gamma, abs=0.004 + 0.15 * abs(fd_gamma))


def test_barrier_greeks_leave_the_price_unchanged_and_respect_the_stencil(oracle):
    terms = barrier_terms(1, _abi.DOWN_AND_OUT, 99.995)
    s = settings(4096, 3)
    plain = oracle.barrier_price(s, terms, market())
    est, g = oracle.barrier_price_with_greeks(s, terms, market(), request=3, spot_radius=0.01)
    assert est.price == plain.price
    assert g.available == 0  # the level lies inside the spot stencil
    est, g = oracle.barrier_price_with_greeks(s, terms, market(), request=3, spot_radius=0.001)
    assert g.available == 3


def test_resolved_knock_in_reports_the_vanilla_greeks(oracle):
    # Touched now: the knock-in is the vanilla, with its analytic delta and gamma.
    terms = barrier_terms(1, _abi.UP_AND_IN, 95.0)
    est, g = oracle.barrier_price_with_greeks(settings(1024), terms, market(), request=3, spot_radius=0.0)
    w = 0.2
    d1 = (0.04 - 0.01 + 0.5 * w * w) / w
    assert g.available == 3
    assert g.delta == pytest.approx(math.exp(-0.01) * ncdf(d1), rel=1e-12)
    assert g.gamma == pytest.approx(math.exp(-0.01) * npdf(d1) / (100 * w), rel=1e-12)
