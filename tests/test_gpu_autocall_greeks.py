"""Shared-draw scenario lanes of the autocallable engine (the Greek batches
of engines/monte_carlo/autocallable.rs:174-241, 919-999) through the C ABI:
every lane equals pricing its scenario alone, field for field, and
price_with_greeks equals numerical_greeks on the same seed."""

import ctypes as C
import datetime as dt

import pytest

from kiyosi_amd import _abi, _native, greeks, workloads
from kiyosi_amd.engines import MonteCarloAutocallableEngine
from kiyosi_amd.instruments import KnockInObservation, SnowballObservation, SnowballOption
from kiyosi_amd.market import BlackScholesMertonParameters, PricingContext, TradingCalendar

from test_gpu_autocall import gpu_price, phoenix, settings

pytestmark = pytest.mark.gpu
V = workloads.VALUATION
FIELDS = ("price", "standard_error", "ci_lower", "ci_upper", "has_standard_error", "uncertainty_status",
          "has_seed", "seed", "path_standard_error", "path_steps")


def context(spot=100.0, sigma=0.25, r=0.02, q=0.0, valuation=V):
    return PricingContext(model_parameters=BlackScholesMertonParameters(risk_free_rate=r, dividend_yield=q,
                                                                        volatility=sigma),
                          spot_price=spot, valuation_time=valuation)


def lanes(ctx, s, note, contexts):
    markets = (_abi.BsmMarket * len(contexts))(*[c.bsm_market() for c in contexts])
    out = (_abi.Estimate * len(contexts))()
    status = _native.lib().kg_autocallable_scenarios(ctx.handle, C.byref(s), C.byref(note._terms()), markets,
                                                     len(contexts), out)
    return status, out


def same_time_scenarios(note, ctx):
    """The full stencil's scenarios at the valuation time (spot, volatility,
    rate and dividend bumps and their crosses)."""
    plan = greeks.plan(note, ctx, None, greeks.NumericalShiftSettings())
    return [c for c in plan if c.valuation_time == ctx.valuation_time]


CASES = [
    ("snowball sobol", workloads.snowball_note, _abi.SOBOL, False, 1 << 13),
    ("snowball pcg antithetic", workloads.snowball_note, _abi.PSEUDO_RANDOM, True, 6001),
    ("snowball no bridge", lambda: workloads.snowball_note(months=50), _abi.PSEUDO_RANDOM, False, 4096),
    ("phoenix pcg antithetic", phoenix, _abi.PSEUDO_RANDOM, True, 4096),
    ("phoenix sobol", lambda: phoenix(gap=5), _abi.SOBOL, False, 4096),
]


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32_STEP], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name,make,sampling,antithetic,paths", CASES, ids=[c[0] for c in CASES])
def test_each_lane_equals_its_scenario_priced_alone(gpu, oracle, name, make, sampling, antithetic, paths,
                                                   precision):
    note = make()
    scenarios = same_time_scenarios(note, context())
    assert 9 <= len(scenarios) <= _abi.MAX_AUTOCALL_SCENARIOS  # spans several launch groups
    s = settings(paths, seed=4242, sampling=sampling, antithetic=antithetic, precision=precision)
    status, out = lanes(gpu, s, note, scenarios)
    assert status == _abi.OK, _native.lib().kg_last_error(gpu.handle)
    for i, scn in enumerate(scenarios):
        alone = gpu_price(gpu, s, note, scn)
        for f in FIELDS:
            assert getattr(out[i], f) == getattr(alone, f), (name, i, f)
        if precision == _abi.FP64:
            ref = oracle.autocallable_price(s, note._terms(), scn.bsm_market(), workers=8)
            assert out[i].price == pytest.approx(ref.price, rel=1e-10, abs=1e-13), (name, i)


def snowball_with_history():
    """A snowball valued after its effective date, with no barrier event so far."""
    cal = TradingCalendar.weekdays()
    dates = [cal.add_trading_days(V, 21 * k) for k in range(1, 13)]
    return SnowballOption(maturity_coupon_rate=0.15, initial_spot=100.0, knock_in_level=75.0,
                          upper_strike=100.0, lower_strike=0.0,
                          knock_in_observation=KnockInObservation.every_trading_day(dt.time(0, 0)),
                          effective_time=V, expiry_time=dates[-1],
                          autocall_observations=[SnowballObservation(d, 100.0, 0.15) for d in dates],
                          barrier_state="no_event", calendar=cal), dates


def test_scenarios_with_a_known_value_take_it(gpu, oracle):
    """On a knock-out date, scenarios at or above the level settle now; the
    others are sampled lanes."""
    note, dates = snowball_with_history()
    day = dates[0]
    contexts = [context(99.5, valuation=day), context(100.5, valuation=day), context(99.0, valuation=day),
                context(101.0, valuation=day)]
    s = settings(4096, sampling=_abi.PSEUDO_RANDOM)
    status, out = lanes(gpu, s, note, contexts)
    assert status == _abi.OK
    for i, c in enumerate(contexts):
        alone = gpu_price(gpu, s, note, c)
        for f in FIELDS:
            assert getattr(out[i], f) == getattr(alone, f), (i, f)
        ref = oracle.autocallable_price(s, note._terms(), c.bsm_market(), workers=8)
        assert out[i].price == pytest.approx(ref.price, rel=1e-10, abs=1e-13), i
    assert out[1].has_seed == 0 and out[0].has_seed == 1


def test_scenarios_on_different_grids_are_refused(gpu):
    """Below the knock-in level on a daily observation the note is knocked in
    and drops its daily grid: the batch returns UNSUPPORTED_OPERATION."""
    note, _ = snowball_with_history()
    day = dt.datetime(2025, 1, 7)
    s = settings(4096, sampling=_abi.PSEUDO_RANDOM)
    status, _ = lanes(gpu, s, note, [context(75.01, valuation=day), context(74.99, valuation=day)])
    assert status == _abi.UNSUPPORTED_OPERATION
    assert "path grid" in _native.lib().kg_last_error(gpu.handle).decode()


@pytest.mark.parametrize("precision,sampling,antithetic", [("fp64", "sobol", False), ("fp32", "sobol", False),
                                                           ("fp64", "pseudo_random", True)])
def test_price_with_greeks_equals_numerical_greeks(gpu, precision, sampling, antithetic):
    note = workloads.snowball_note()
    engine = MonteCarloAutocallableEngine(path_count=1 << 13, seed=99, sampling=sampling, antithetic=antithetic,
                                          precision=precision)
    batched = engine.price_with_greeks(note, context())
    scalar = greeks.numerical_greeks(engine, note, context())
    assert batched == scalar
    assert batched.price == engine.price(note, context())
    # Daily knock-in at the valuation clock leaves no observation-free time stencil.
    assert batched.status("theta") == "stencil_unavailable"
    for g in ("delta", "gamma", "vega", "rho", "vanna", "volga"):
        assert batched.status(g) == "available", g
    assert batched.require("delta") > 0.0 and batched.require("vega") < 0.0


def test_phoenix_greeks_include_theta_between_observations(gpu):
    """Phoenix valued mid-day between daily observations at 15:00: the time
    stencil shrinks until no observation lies inside, and theta is formed."""
    note = phoenix()
    note.knock_in_observation = KnockInObservation.every_trading_day(dt.time(15, 0))
    note.barrier_state = "no_event"
    t = dt.datetime(2025, 1, 6, 12, 0)
    engine = MonteCarloAutocallableEngine(path_count=4096, seed=5, antithetic=True)
    batched = engine.price_with_greeks(note, context(valuation=t))
    assert batched == greeks.numerical_greeks(engine, note, context(valuation=t))
    assert batched.status("theta") == "available"
