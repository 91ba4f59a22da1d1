"""American lanes on shared paths (american_scenarios::<N, S>,
engines/monte_carlo/american.rs:955-1455) through kg_american_price_lanes:
strike groups (price_strikes, :677-771) and Greek scenario groups
(greek_prices, :775-931). The reference requires every lane to equal the
scalar price of that lane (tests/american_monte_carlo.rs:1018-1027); here
each lane is checked against kg_american_price bit for bit and against the
oracle's scalar LSMC at rel 1e-9."""

import ctypes as C

import pytest

from conftest import END, START, us
from kiyosi_amd import _abi, _native, workloads
from kiyosi_amd.market import american_model

pytestmark = pytest.mark.gpu

STRIKES = [100.0, 70.0, 110.0, 85.0, 100.0, 75.0, 140.0]


def settings(paths, seed=42, antithetic=True, sampling=_abi.PSEUDO_RANDOM, precision=_abi.FP64, large=False):
    return _abi.Settings(paths, seed, 1, sampling, int(antithetic), precision, 0, int(large))


def terms(steps=32, strike=100.0, option_type=-1):
    return _abi.AmericanTerms(option_type, steps, strike, us(START), us(END))


def model(spot=100.0, rate=0.05, dividend=0.02, sigma=0.25, at=START):
    return _abi.AmericanModel(spot, rate, dividend, sigma, us(at), None)


def options(basis=_abi.BASIS_QUADRATIC, training=0):
    return _abi.LsmcOptions(basis, int(training > 0), training, 4, 65536, 2e-3)


def scalar(ctx, s, t, m, basis=_abi.BASIS_QUADRATIC, training=0):
    out = _abi.LsmcReport()
    ctx.check(_native.lib().kg_american_price(ctx.handle, C.byref(s), C.byref(t), C.byref(m),
                                              C.byref(options(basis, training)), C.byref(out)))
    return out


def lanes(ctx, s, t, models, strikes=None, basis=_abi.BASIS_QUADRATIC, training=0):
    n = len(models)
    arr = (_abi.AmericanModel * n)(*models)
    ks = (C.c_double * n)(*strikes) if strikes is not None else None
    out = (_abi.LsmcReport * n)()
    status = _native.lib().kg_american_price_lanes(ctx.handle, C.byref(s), C.byref(t), arr, ks, n,
                                                   C.byref(options(basis, training)), out)
    return status, list(out)


def with_strike(t, k):
    return _abi.AmericanTerms(t.option_type, t.step_count, k, t.effective_time, t.expiry_time)


def bsm_market(m):
    return _abi.BsmMarket(m.spot, m.rate, m.dividend, m.volatility, m.valuation_time)


def oracle_price(oracle, s, t, m):
    put = _abi.EuropeanTerms(t.option_type, 0, t.strike, t.effective_time, t.expiry_time)
    price, _ = oracle.american_lsmc(s, put, bsm_market(m), t.step_count, workers=8)
    return price


@pytest.mark.parametrize("sampling,paths", [(_abi.PSEUDO_RANDOM, 4095), (_abi.SOBOL, 4096)])
@pytest.mark.parametrize("basis", [_abi.BASIS_QUADRATIC, _abi.BASIS_POLY4])
def test_strike_lanes_equal_scalar_prices_and_the_oracle(gpu, oracle, sampling, paths, basis):
    """The reference's strike-batch case: seven strikes (one repeated) on
    4,095 antithetic pseudorandom or 4,096 Sobol paths, 32 exercise dates."""
    s, t = settings(paths, sampling=sampling), terms()
    status, got = lanes(gpu, s, t, [model()] * len(STRIKES), STRIKES, basis)
    assert status == _abi.OK, _native.lib().kg_last_error(gpu.handle)
    for k, rep in zip(STRIKES, got):
        alone = scalar(gpu, s, with_strike(t, k), model(), basis)
        assert rep.price == alone.price, k
        assert rep.in_sample_price == alone.in_sample_price
        assert rep.dates_fitted == alone.dates_fitted == 31
        if basis == _abi.BASIS_QUADRATIC:
            assert rep.price == pytest.approx(oracle_price(oracle, s, with_strike(t, k), model()), rel=1e-9), k
    assert got[0].price == got[4].price  # the repeated strike


def scenario_models():
    """A Greek stencil's markets at one valuation time: bumped spot,
    volatility, rate and dividend around S = 100, sigma = 0.25."""
    h, v = 0.01, 1e-4
    return [model(), model(spot=100 + h), model(spot=100 - h), model(sigma=0.25 + v), model(sigma=0.25 - v),
            model(rate=0.05 + 1e-4), model(rate=0.05 - 1e-4), model(dividend=0.02 + 1e-4),
            model(spot=100 + h, sigma=0.25 + v)]


@pytest.mark.parametrize("antithetic", [False, True])
@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32_STEP])
def test_scenario_lanes_equal_scalar_prices(gpu, oracle, antithetic, precision):
    s, t = settings(20_001, seed=7, antithetic=antithetic, precision=precision), terms(50)
    models = scenario_models()
    status, got = lanes(gpu, s, t, models)
    assert status == _abi.OK, _native.lib().kg_last_error(gpu.handle)
    for m, rep in zip(models, got):
        assert rep.price == scalar(gpu, s, t, m).price
        if precision == _abi.FP64:
            assert rep.price == pytest.approx(oracle_price(oracle, s, t, m), rel=1e-9)


@pytest.mark.parametrize("sampling,paths", [(_abi.PSEUDO_RANDOM, 1 << 20), (_abi.SOBOL, 1 << 20)])
def test_reverse_mode_lanes(gpu, oracle, sampling, paths):
    """2^20 samples x 50 dates: reverse (Brownian bridge) paths shared by four
    strike lanes and a bumped-spot lane; two lanes against the oracle."""
    s, t = settings(paths, seed=1729, antithetic=False, sampling=sampling, large=True), terms(50)
    models = [model()] * 4 + [model(spot=101.0)]
    strikes = [100.0, 90.0, 110.0, 120.0, 100.0]
    status, got = lanes(gpu, s, t, models, strikes)
    assert status == _abi.OK, _native.lib().kg_last_error(gpu.handle)
    for m, k, rep in zip(models, strikes, got):
        assert rep.price == scalar(gpu, s, with_strike(t, k), m).price, k
    for i in (1, 4):
        ref = oracle_price(oracle, s, with_strike(t, strikes[i]), models[i])
        assert got[i].price == pytest.approx(ref, rel=1e-9)


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32_STEP])
@pytest.mark.parametrize("basis", [_abi.BASIS_QUADRATIC, _abi.BASIS_POLY4])
def test_local_vol_strike_lanes(gpu, oracle, precision, basis):
    """Strikes on one local-volatility market share the log-spot paths;
    a lane at another spot makes the call price each lane alone. Both equal
    the scalar engine; FP64 lanes meet the oracle's LV LSMC at 1e-8."""
    _, context, _, _ = workloads.american_put_local_vol()
    m = american_model(context)
    s, t = settings(20_000, seed=7, precision=precision), terms(50)
    strikes = [100.0, 90.0, 115.0]
    status, got = lanes(gpu, s, t, [m] * 3, strikes, basis)
    assert status == _abi.OK, _native.lib().kg_last_error(gpu.handle)
    for k, rep in zip(strikes, got):
        assert rep.price == scalar(gpu, s, with_strike(t, k), m, basis).price
        if precision == _abi.FP64:
            ins, _ = oracle.lv_lsmc(s, with_strike(t, k), m, basis, 0, workers=8)
            assert rep.price == pytest.approx(ins, rel=1e-8)
    bumped = _abi.AmericanModel(101.0, m.rate, m.dividend, 0.0, m.valuation_time, m.surface)
    status, got = lanes(gpu, s, t, [m, bumped], None, basis)
    assert status == _abi.OK
    assert got[1].price == scalar(gpu, s, t, bumped, basis).price


def test_lanes_that_cannot_share_paths_are_priced_alone(gpu):
    """Different valuation times, anchored paths, out-of-sample fits and the
    MLP regressor: each lane priced alone, equal to the scalar engine."""
    s, t = settings(8192, seed=3), terms(20)
    later = START.replace(month=3)
    status, got = lanes(gpu, s, t, [model(), model(at=later)])
    assert status == _abi.OK
    assert [r.price for r in got] == [scalar(gpu, s, t, model()).price, scalar(gpu, s, t, model(at=later)).price]
    big = _abi.AmericanModel(1e255, 0.05, 0.02, 0.2, us(START), None)
    status, got = lanes(gpu, s, with_strike(t, 1e255), [big, big], [1e255, 2e255])
    assert status == _abi.OK
    assert got[1].price == scalar(gpu, s, with_strike(t, 2e255), big).price
    status, got = lanes(gpu, s, t, [model(), model(spot=105.0)], training=16384)
    assert status == _abi.OK
    for m, rep in zip([model(), model(spot=105.0)], got):
        alone = scalar(gpu, s, t, m, training=16384)
        assert (rep.price, rep.in_sample_price) == (alone.price, alone.in_sample_price)
    status, got = lanes(gpu, s, t, [model(), model()], [100.0, 90.0], basis=_abi.BASIS_MLP)
    assert status == _abi.OK
    assert got[1].price == scalar(gpu, s, with_strike(t, 90.0), model(), _abi.BASIS_MLP).price


def test_a_failed_lane_fails_the_request(gpu):
    s, t = settings(4096), terms()
    status, got = lanes(gpu, s, t, [model()] * 3, [100.0, -1.0, 90.0])
    assert status == _abi.INVALID_INPUT
    assert all(r.price == 0.0 for r in got)
    status, _ = lanes(gpu, s, t, [model()] * 17)
    assert status == _abi.INVALID_INPUT
    call_no_dividend = _abi.AmericanTerms(1, 32, 100.0, us(START), us(END))
    status, _ = lanes(gpu, s, call_no_dividend, [model(dividend=0.0)] * 2)
    assert status == _abi.UNSUPPORTED_OPERATION


def test_lanes_at_expiry_are_the_immediate_payoffs(gpu):
    s, t = settings(4096), terms()
    at_expiry = [model(spot=90.0, at=END), model(spot=120.0, at=END)]
    status, got = lanes(gpu, s, t, at_expiry)
    assert status == _abi.OK
    assert [r.price for r in got] == [10.0, 0.0]


def test_engines_price_strikes_and_greeks_through_lanes(gpu):
    """MonteCarloVanillaEngine.price_strikes and price_with_greeks for an
    American put equal the per-strike prices and the scenario-by-scenario
    stencil (numerical_greeks) exactly."""
    import datetime as dt

    from kiyosi_amd import greeks
    from kiyosi_amd.engines import MonteCarloVanillaEngine
    from kiyosi_amd.instruments import AmericanOption
    from kiyosi_amd.market import BlackScholesMertonParameters, PricingContext
    ctx = PricingContext(model_parameters=BlackScholesMertonParameters(risk_free_rate=0.05, dividend_yield=0.02,
                                                                       volatility=0.25),
                         spot_price=100.0, valuation_time=START)
    put = AmericanOption(option_type="put", strike=100.0, effective_time=START, expiry_time=END)
    engine = MonteCarloVanillaEngine(path_count=8191, step_count=32, seed=42, antithetic=True)
    strikes = STRIKES + [float(k) for k in range(60, 141, 4)]
    prices = engine.price_strikes(put, ctx, strikes=strikes)
    expected = [engine.price(_copy_with_strike(put, k), ctx) for k in strikes]
    assert list(prices) == expected
    request = {"delta", "gamma", "vega", "rho", "theta", "vanna"}
    got = engine.price_with_greeks(put, ctx, request)
    ref = greeks.numerical_greeks(engine, put, ctx, request)
    assert got.price == ref.price
    for g in request:
        assert got.get(g) == ref.get(g), g
    assert got.seed == ref.seed == 42
    later = ctx.replace(valuation_time=START + dt.timedelta(days=30))
    assert engine.price_with_greeks(put, later, {"delta"}).get("delta") == greeks.numerical_greeks(
        engine, put, later, {"delta"}).get("delta")


def _copy_with_strike(option, strike):
    import copy
    out = copy.copy(option)
    out.strike = strike
    return out
