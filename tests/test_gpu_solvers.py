"""Implied-parameter solvers over the GPU engines on pinned seeds
(risk/implied.rs, risk/coupon.rs): the oracle's price at a known parameter,
on the engine's seed, is inverted back to that parameter by trials priced on
the GPU, since every trial sees the same draws; the coupon solvers price
their bracket as two lanes of one launch."""

import datetime as dt

import pytest

import kiyosi_amd as k
from kiyosi_amd import solvers, workloads
from kiyosi_amd.engines import MonteCarloAutocallableEngine, MonteCarloBarrierEngine, MonteCarloVanillaEngine
from kiyosi_amd.market import BlackScholesMertonParameters, PricingContext

from test_gpu_autocall import phoenix

pytestmark = pytest.mark.gpu
START, END = dt.datetime(2025, 1, 1), dt.datetime(2026, 1, 1)
V = workloads.VALUATION


def context(sigma, valuation=START):
    return PricingContext(model_parameters=BlackScholesMertonParameters(risk_free_rate=0.02, dividend_yield=0.0,
                                                                        volatility=sigma),
                          spot_price=100.0, valuation_time=valuation)


@pytest.mark.parametrize("precision,tolerance", [("fp64", 1e-8), ("fp32", 1e-5)])
def test_implied_volatility_of_a_monte_carlo_european_price(gpu, oracle, precision, tolerance):
    """The observed price is the oracle's at sigma 0.27 on the engine's seed;
    the solver, pricing each trial on the GPU, recovers 0.27 (to the GPU's
    agreement with the oracle: 1e-10 relative in FP64, a small fraction of a
    standard error in FP32)."""
    option = k.EuropeanOption(option_type="put", strike=95.0, effective_time=START, expiry_time=END)
    engine = MonteCarloVanillaEngine(path_count=1 << 16, seed=5, antithetic=True, precision=precision)
    observed = oracle.vanilla_price(engine._settings(), option._terms(), context(0.27).bsm_market(), workers=8).price
    assert engine.price(option, context(0.27)) == pytest.approx(observed, rel=1e-10 if precision == "fp64" else 1e-5)
    # FP32-step prices move in steps near 1e-6 of the price as sigma moves, so
    # they can match the oracle's FP64 price to about 1e-5 only.
    price_tolerance = 1e-8 if precision == "fp64" else 1e-5
    got = k.implied_volatility(engine, option, context(0.2), observed,
                               k.ImpliedVolatilitySettings(0.05, 1.0, price_tolerance=price_tolerance))
    assert got == pytest.approx(0.27, abs=tolerance)


def test_implied_volatility_of_a_barrier_and_its_unidentifiable_case(gpu, oracle):
    option = k.BarrierOption(option_type="call", strike=100.0, effective_time=START, expiry_time=END,
                             barrier_level=80.0, barrier_type="down_and_out")
    engine = MonteCarloBarrierEngine(path_count=1 << 15, seed=9, sampling="sobol")
    observed = oracle.barrier_price(engine._settings(), option._terms(), context(0.31).bsm_market(), workers=8).price
    got = k.implied_volatility(engine, option, context(0.2), observed, k.ImpliedVolatilitySettings(0.05, 1.5))
    assert got == pytest.approx(0.31, abs=1e-6)
    # Knocked out at the valuation: the value is the rebate whatever the volatility.
    knocked = k.BarrierOption(option_type="call", strike=100.0, effective_time=START, expiry_time=END,
                              barrier_level=105.0, barrier_type="down_and_out")
    with pytest.raises(k.UnsupportedOperationError):
        k.implied_volatility(engine, knocked, context(0.2), 0.0, k.ImpliedVolatilitySettings(0.05, 1.5))


def test_implied_volatility_of_a_snowball(gpu, oracle):
    note = workloads.snowball_note()
    engine = MonteCarloAutocallableEngine(path_count=1 << 14, seed=3, sampling="sobol")
    observed = oracle.autocallable_price(engine._settings(), note._terms(), context(0.22, V).bsm_market(),
                                        workers=8).price
    got = k.implied_volatility(engine, note, context(0.3, V), observed, k.ImpliedVolatilitySettings(0.05, 0.8))
    assert got == pytest.approx(0.22, abs=1e-6)


@pytest.mark.parametrize("convention", [solvers.SHIFT_MATURITY_COUPON, solvers.PRESERVE_MATURITY_COUPON])
def test_implied_snowball_coupon(gpu, oracle, convention):
    note = workloads.snowball_note()
    engine = MonteCarloAutocallableEngine(path_count=1 << 14, seed=17, sampling="sobol")
    target = solvers._snowball_with_coupon(note, 0.118, convention)
    observed = oracle.autocallable_price(engine._settings(), target._terms(), context(0.25, V).bsm_market(),
                                        workers=8).price
    quote = k.implied_snowball_coupon(engine, note, context(0.25, V), observed, convention)
    assert quote.quote_convention == convention
    assert quote.first_observation_coupon == pytest.approx(0.118, abs=1e-8)
    # The quote reprices the note to the observed price.
    back = engine.price(solvers._snowball_with_coupon(note, quote.first_observation_coupon, convention),
                        context(0.25, V))
    assert back == pytest.approx(observed, abs=1e-8)


def test_implied_phoenix_coupon_with_tight_tolerance(gpu, oracle):
    note = phoenix()
    engine = MonteCarloAutocallableEngine(path_count=4096, seed=23, antithetic=True)
    observed = oracle.autocallable_price(engine._settings(), solvers._phoenix_with_coupon(note, 0.0085)._terms(),
                                        context(0.25, V).bsm_market(), workers=8).price
    for settings in (k.ImpliedCouponSettings(), k.ImpliedCouponSettings(price_tolerance=1e-13,
                                                                        parameter_tolerance=1e-15)):
        got = k.implied_phoenix_coupon(engine, note, context(0.25, V), observed, settings)
        assert engine.price(solvers._phoenix_with_coupon(note, got), context(0.25, V)) == pytest.approx(
            observed, abs=max(settings.price_tolerance, 1e-12))
        assert got == pytest.approx(0.0085, abs=1e-9)


def test_solvers_need_an_explicit_seed(gpu):
    engine = MonteCarloAutocallableEngine(path_count=1024)
    with pytest.raises(k.InvalidInputError, match="explicit seed"):
        k.implied_phoenix_coupon(engine, phoenix(), context(0.25, V), 1.0)
