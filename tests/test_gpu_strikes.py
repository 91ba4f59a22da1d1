"""Strike batches on the GPU (price_strikes, engines/monte_carlo/asian.rs
:197-345 and american.rs:677-709): every price equals pricing that strike's
option alone with the same seed, bit for bit, as the reference's own tests
assert (crates/kiyosi/tests/asian_monte_carlo.rs:144-151,
american_monte_carlo.rs:1018-1023); repeated strikes repeat one quote and
the empty batch is empty. Also pins that an American option's Greeks come
from American prices."""

import datetime as dt

import pytest

import kiyosi_amd as k
from kiyosi_amd import greeks
from kiyosi_amd.engines import MonteCarloAsianEngine, MonteCarloVanillaEngine
from kiyosi_amd.instruments import AmericanOption, ArithmeticAveragePriceOption
from kiyosi_amd.market import BlackScholesMertonParameters, PricingContext

pytestmark = pytest.mark.gpu
START, END = dt.datetime(2025, 1, 1), dt.datetime(2026, 1, 1)


def context(sigma=0.2, spot=100.0):
    return PricingContext(model_parameters=BlackScholesMertonParameters(risk_free_rate=0.04, dividend_yield=0.01,
                                                                        volatility=sigma),
                          spot_price=spot, valuation_time=START)


def asian(kind="call", strike=100.0):
    return ArithmeticAveragePriceOption(option_type=kind, strike=strike, effective_time=START,
                                        averaging_start_time=START, expiry_time=END)


def alone(engine, template, strike, ctx):
    option = ArithmeticAveragePriceOption(option_type=template.option_type, strike=strike,
                                          effective_time=template.effective_time,
                                          averaging_start_time=template.averaging_start_time,
                                          expiry_time=template.expiry_time)
    return engine.price(option, ctx)


# 40 distinct strikes span two lane launches; deep in- and out-of-the-money
# strikes switch the geometric control off and on across lanes.
STRIKES = [60.0 + 2.0 * i for i in range(40)] + [100.0, 80.0, 100.0]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["call", "put"])
@pytest.mark.parametrize("sampling,antithetic,paths", [("pseudo_random", False, 4096),
                                                       ("pseudo_random", True, 3001),
                                                       ("sobol", False, 2048)])
def test_asian_strike_batch_equals_scalar_prices(gpu, oracle, kind, sampling, antithetic, paths, precision):
    """Each strike lane equals that strike priced alone on the GPU bit for bit,
    and the oracle's price of that strike on the same seed at 1e-10 (FP64)."""
    engine = MonteCarloAsianEngine(path_count=paths, seed=42, sampling=sampling, antithetic=antithetic,
                                   step_count=16, precision=precision)
    template = asian(kind)
    got = engine.price_strikes(template, context(), strikes=STRIKES)
    assert len(got) == len(STRIKES)
    for strike, price in zip(STRIKES, got):
        assert price == alone(engine, template, strike, context()), strike
        if precision == "fp64":
            ref = oracle.asian_price(engine._settings(), asian(kind, strike)._terms(16), context().bsm_market(),
                                     workers=8)
            assert price == pytest.approx(ref.price, rel=1e-10, abs=1e-12), strike
    assert got[40] == got[42]


def test_asian_strike_batch_edge_cases(gpu, oracle):
    engine = MonteCarloAsianEngine(path_count=1024, seed=7, step_count=8)
    assert engine.price_strikes(asian(), context(), strikes=[]) == ()
    with pytest.raises(k.InvalidInputError):
        engine.price_strikes(asian(), context(), strikes=[100.0, float("nan")])
    # A deep out-of-the-money put with a wide control band: a lane whose
    # controlled price turns negative is priced alone, without the control.
    low_vol = context(sigma=0.45)
    strikes = [40.0, 45.0, 50.0, 100.0]
    got = engine.price_strikes(asian("put"), low_vol, strikes=strikes)
    assert got == tuple(alone(engine, asian("put"), s, low_vol) for s in strikes)
    assert all(p >= 0.0 for p in got)
    for strike, price in zip(strikes, got):
        ref = oracle.asian_price(engine._settings(), asian("put", strike)._terms(8), low_vol.bsm_market(), workers=8)
        assert price == pytest.approx(ref.price, rel=1e-10, abs=1e-12), strike


def test_american_strike_batch_equals_scalar_prices(gpu, oracle):
    engine = MonteCarloVanillaEngine(path_count=1 << 14, seed=42, step_count=16, antithetic=True)
    template = AmericanOption(option_type="put", strike=100.0, effective_time=START, expiry_time=END)
    strikes = [90.0, 110.0, 100.0, 90.0]
    got = engine.price_strikes(template, context(), strikes=strikes)
    expected = tuple(engine.price(AmericanOption(option_type="put", strike=s, effective_time=START, expiry_time=END),
                                  context()) for s in strikes)
    assert got == expected
    assert got[0] == got[3] and got[0] < got[2] < got[1]
    import ctypes as C
    from kiyosi_amd import _abi
    from conftest import us
    for strike, price in zip(strikes, got):
        ref, _ = oracle.american_lsmc(engine._settings(), _abi.EuropeanTerms(-1, 0, strike, us(START), us(END)),
                                      context().bsm_market(), 16, workers=8)
        assert price == pytest.approx(ref, rel=1e-9), strike


def test_american_greeks_come_from_american_prices(gpu):
    engine = MonteCarloVanillaEngine(path_count=1 << 14, seed=11, step_count=16)
    put = AmericanOption(option_type="put", strike=110.0, effective_time=START, expiry_time=END)
    result = engine.price_with_greeks(put, context(), ["delta"])
    assert result.price == engine.price(put, context())
    reference = greeks.numerical_greeks(engine, put, context(), ["delta"])
    assert result.get("delta") == reference.get("delta")
    european = k.EuropeanOption(option_type="put", strike=110.0, effective_time=START, expiry_time=END)
    assert result.price > MonteCarloVanillaEngine(path_count=1 << 14, seed=11).price(european, context())
