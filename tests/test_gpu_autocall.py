"""GPU parity of the autocallable engine (config 2 and variants) with the
FP64 oracle, through the C ABI."""

import ctypes as C
import datetime as dt

import pytest

from kiyosi_amd import _abi, _native, workloads
from kiyosi_amd.instruments import (CouponObservation, KnockInObservation, KnockOutObservation,
                                    PhoenixOption)
from kiyosi_amd.market import BlackScholesMertonParameters, PricingContext, TradingCalendar

pytestmark = pytest.mark.gpu
V = workloads.VALUATION


def settings(paths, seed=1729, sampling=_abi.SOBOL, antithetic=False, precision=_abi.FP64):
    return _abi.Settings(paths, seed, 1, sampling, int(antithetic), precision, 0, 0)


def gpu_price(ctx, s, note, context):
    out = _abi.Estimate()
    terms, market = note._terms(), context.bsm_market()
    ctx.check(_native.lib().kg_autocallable_price(ctx.handle, C.byref(s), C.byref(terms),
                                                  C.byref(market), C.byref(out)))
    return out


def phoenix(ki=75.0, gap=21):
    cal = TradingCalendar.weekdays()
    dates = [cal.add_trading_days(V, gap * k) for k in range(1, 13)]
    return PhoenixOption(
        coupon_per_period=0.012, initial_spot=100.0, knock_in_level=ki, upper_strike=100.0,
        lower_strike=0.0, knock_in_observation=KnockInObservation.every_trading_day(dt.time(0, 0)),
        effective_time=V, expiry_time=dates[-1],
        knock_out_observations=[KnockOutObservation(d, 103.0) for d in dates[2:]],
        coupon_observations=[CouponObservation(d, 80.0) for d in dates], calendar=cal)


CASES = [
    ("snowball sobol", lambda: workloads.snowball_note(), settings(1 << 14)),
    ("snowball pcg", lambda: workloads.snowball_note(), settings(10_000, sampling=_abi.PSEUDO_RANDOM)),
    ("snowball pcg antithetic", lambda: workloads.snowball_note(),
     settings(10_001, sampling=_abi.PSEUDO_RANDOM, antithetic=True)),
    ("snowball sobol antithetic", lambda: workloads.snowball_note(), settings(1 << 13, antithetic=True)),
    ("snowball no bridge (weekly events)", lambda: workloads.snowball_note(months=50),
     settings(4096, sampling=_abi.PSEUDO_RANDOM)),
    ("phoenix pcg", phoenix, settings(8192, sampling=_abi.PSEUDO_RANDOM)),
    ("phoenix sobol no bridge", lambda: phoenix(gap=5), settings(4096)),
]


@pytest.mark.parametrize("name,make,s", CASES, ids=[c[0] for c in CASES])
def test_fp64_matches_the_oracle(gpu, oracle, name, make, s):
    note = make()
    context = PricingContext(model_parameters=BlackScholesMertonParameters(
        risk_free_rate=0.02, dividend_yield=0.0, volatility=0.25), spot_price=100.0, valuation_time=V)
    got = gpu_price(gpu, s, note, context)
    ref = oracle.autocallable_price(s, note._terms(), context.bsm_market(), workers=8)
    assert got.price == pytest.approx(ref.price, rel=1e-10), name
    assert got.uncertainty_status == ref.uncertainty_status
    if ref.has_standard_error:
        assert got.standard_error == pytest.approx(ref.standard_error, rel=1e-8)


def test_config2_fp64_full_size_matches_the_oracle(gpu, oracle):
    note, context, kw = workloads.snowball_252d_sobol_2p20()
    s = settings(kw["path_count"], kw["seed"])
    got = gpu_price(gpu, s, note, context)
    ref = oracle.autocallable_price(s, note._terms(), context.bsm_market(), workers=8)
    assert abs(got.price / ref.price - 1) < 1e-10


def test_config2_fp32_step_within_half_a_standard_error(gpu):
    note, context, kw = workloads.snowball_252d_sobol_2p20()
    plain = gpu_price(gpu, settings(kw["path_count"], kw["seed"]), note, context)
    fast = gpu_price(gpu, settings(kw["path_count"], kw["seed"], precision=_abi.FP32_STEP), note, context)
    assert abs(fast.price - plain.price) <= 0.5 * plain.path_standard_error


def test_fp32_pcg_within_half_a_standard_error(gpu):
    note = workloads.snowball_note()
    context = PricingContext(model_parameters=BlackScholesMertonParameters(
        risk_free_rate=0.02, dividend_yield=0.0, volatility=0.25), spot_price=100.0, valuation_time=V)
    plain = gpu_price(gpu, settings(1 << 18, sampling=_abi.PSEUDO_RANDOM), note, context)
    fast = gpu_price(gpu, settings(1 << 18, sampling=_abi.PSEUDO_RANDOM, precision=_abi.FP32_STEP),
                     note, context)
    assert abs(fast.price - plain.price) <= 0.5 * plain.standard_error


def test_python_engine_interface(gpu):
    import kiyosi_amd as k
    note, context, kw = workloads.snowball_252d_sobol_2p20()
    engine = k.MonteCarloAutocallableEngine(path_count=4096, seed=5, sampling="sobol")
    est = engine.price_with_statistics(note, context)
    assert est.uncertainty_status == "unsupported_sampling" and est.standard_error is None
    assert est.price == engine.price(note, context)
    knocked_out = workloads.snowball_note()
    knocked_out.barrier_state = "knocked_out"
    with pytest.raises(k.InvalidInputError):
        engine.price(knocked_out, context)


def daily_autocall(days, lockout=20):
    """A snowball that knocks out on any trading day after a lockout of
    `lockout` days, with daily knock-in: a bridge plan (the lockout leaves a
    gap of more than eight days) with one interval per knock-out day."""
    from kiyosi_amd.instruments import SnowballObservation, SnowballOption
    cal = TradingCalendar.weekdays()
    dates = [cal.add_trading_days(V, d) for d in range(lockout, days + 1)]
    return SnowballOption(
        maturity_coupon_rate=0.15, initial_spot=100.0, knock_in_level=75.0, upper_strike=100.0, lower_strike=0.0,
        knock_in_observation=KnockInObservation.every_trading_day(dt.time(0, 0)), effective_time=V,
        expiry_time=dates[-1], autocall_observations=[SnowballObservation(d, 103.0, 0.15) for d in dates],
        calendar=cal)


@pytest.mark.parametrize("days,sampling", [(1000, _abi.PSEUDO_RANDOM), (252, _abi.SOBOL)])
def test_daily_autocall_bridge_notes_beyond_the_pass_one_lds(gpu, oracle, days, sampling):
    """Bridge notes whose interval rows exceed pass one's 64 KiB of LDS (981
    knock-out days for pseudorandom FP64 rows of 80 bytes; 233 for Sobol,
    whose rows add 256 bytes of table words) price in one pass and match the
    oracle at 1e-10."""
    note = daily_autocall(days)
    context = PricingContext(model_parameters=BlackScholesMertonParameters(
        risk_free_rate=0.02, dividend_yield=0.0, volatility=0.25), spot_price=100.0, valuation_time=V)
    plan = oracle.autocallable_plan(note._terms(), context.bsm_market(), cap=4000)
    assert plan[1]  # a bridge plan
    s = settings(4096, sampling=sampling)
    got = gpu_price(gpu, s, note, context)
    ref = oracle.autocallable_price(s, note._terms(), context.bsm_market(), workers=8)
    assert got.price == pytest.approx(ref.price, rel=1e-10)
