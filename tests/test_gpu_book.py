"""GPU parity of the barrier, Asian and accumulator engines and of the
persistent book kernel with the FP64 oracle."""

import ctypes as C
import datetime as dt
import math

import numpy as np
import pytest

from conftest import END, START, us
from kiyosi_amd import _abi, _native, workloads
from kiyosi_amd.engines import Book

pytestmark = pytest.mark.gpu


def market(spot=100.0, rate=0.05, dividend=0.02, sigma=0.2):
    return _abi.BsmMarket(spot, rate, dividend, sigma, us(START))


def settings(paths, seed=7, antithetic=False, sampling=_abi.PSEUDO_RANDOM, precision=_abi.FP64):
    return _abi.Settings(paths, seed, 1, sampling, int(antithetic), precision, 0, 0)


def daily(days=365):
    return [us(START + dt.timedelta(days=k)) for k in range(1, days + 1)]


def barrier_terms(option_type, barrier_type, level, strike=100.0, rebate=0.0, times=None, at_hit=False):
    arr = (C.c_int64 * len(times))(*times) if times else None
    t = _abi.BarrierTerms(option_type, barrier_type,
                          _abi.OBSERVE_SCHEDULED if times else _abi.OBSERVE_CONTINUOUS, int(at_hit), strike, level,
                          rebate, us(START), us(END), len(times) if times else 0, 0,
                          C.cast(arr, C.POINTER(C.c_int64)) if arr else None)
    t._keep = arr
    return t


def price_one(ctx, fn, s, terms, m):
    out = _abi.Estimate()
    ctx.check(getattr(_native.lib(), fn)(ctx.handle, C.byref(s), C.byref(terms), C.byref(m), C.byref(out)))
    return out


BARRIERS = [
    ("continuous down-out call", lambda: barrier_terms(1, _abi.DOWN_AND_OUT, 90.0), 65536, False),
    ("continuous up-in put", lambda: barrier_terms(-1, _abi.UP_AND_IN, 120.0), 4097, True),
    ("daily up-out call rebate at hit", lambda: barrier_terms(1, _abi.UP_AND_OUT, 125.0, rebate=3.0,
                                                              times=daily(), at_hit=True), 5000, False),
    ("daily down-in put conditioned", lambda: barrier_terms(-1, _abi.DOWN_AND_IN, 88.0, times=daily()), 6000, True),
    ("weekly down-in call", lambda: barrier_terms(1, _abi.DOWN_AND_IN, 90.0, times=daily()[6::7]), 3000, False),
]


@pytest.mark.parametrize("name,make,paths,antithetic", BARRIERS, ids=[b[0] for b in BARRIERS])
def test_barrier_fp64_matches_the_oracle(gpu, oracle, name, make, paths, antithetic):
    terms, s = make(), settings(paths, antithetic=antithetic)
    got = price_one(gpu, "kg_barrier_price", s, terms, market())
    ref = oracle.barrier_price(s, terms, market(), workers=8)
    assert got.price == pytest.approx(ref.price, rel=1e-10, abs=1e-13)
    assert got.uncertainty_status == ref.uncertainty_status
    if ref.has_standard_error:
        assert got.standard_error == pytest.approx(ref.standard_error, rel=1e-8)


def test_barrier_calibration_vector_on_the_gpu(gpu):
    got = price_one(gpu, "kg_barrier_price", settings(65536, seed=1009), barrier_terms(1, _abi.DOWN_AND_OUT, 90.0),
              market())
    assert got.price == pytest.approx(7.616185456813128, rel=1e-10)


@pytest.mark.parametrize("averaging", [_abi.AVERAGE_GEOMETRIC, _abi.AVERAGE_ARITHMETIC])
@pytest.mark.parametrize("kind", [1, -1])
@pytest.mark.parametrize("paths,antithetic,sampling", [(4096, False, _abi.PSEUDO_RANDOM),
                                                      (3001, True, _abi.PSEUDO_RANDOM),
                                                      (2048, False, _abi.SOBOL)])
def test_asian_fp64_matches_the_oracle(gpu, oracle, averaging, kind, paths, antithetic, sampling):
    t = _abi.AsianTerms(kind, averaging, 64, 0, 100.0, 0.0, us(START), us(START), us(END))
    s = settings(paths, antithetic=antithetic, sampling=sampling)
    m = market(rate=0.04, dividend=0.01)
    got, ref = price_one(gpu, "kg_asian_price", s, t, m), oracle.asian_price(s, t, m, workers=8)
    assert got.price == pytest.approx(ref.price, rel=1e-10)
    assert got.uncertainty_status == ref.uncertainty_status


@pytest.mark.parametrize("kind,strike,ko", [(1, 97.0, 106.0), (-1, 103.0, 94.0)])
def test_accumulator_fp64_matches_the_oracle(gpu, oracle, kind, strike, ko):
    t = _abi.AccumulatorTerms(kind, _abi.CALENDAR_WEEKDAYS, strike, ko, 1.0, 2.0, 0.0, 0, us(START),
                              us(START + dt.timedelta(days=180)))
    s = settings(5000, antithetic=True)
    m = market(sigma=0.25)
    got, ref = price_one(gpu, "kg_accumulator_price", s, t, m), oracle.accumulator_price(s, t, m, workers=8)
    assert got.price == pytest.approx(ref.price, rel=1e-10)
    assert got.standard_error == pytest.approx(ref.standard_error, rel=1e-8)


def oracle_entry(oracle, engine, instrument, context):
    s = engine._settings()
    m = context.bsm_market()
    name = type(engine).__name__
    if name == "MonteCarloBarrierEngine":
        return oracle.barrier_price(s, instrument._terms(), m, workers=16)
    if name == "MonteCarloAsianEngine":
        return oracle.asian_price(s, instrument._terms(engine.step_count), m, workers=16)
    if name == "MonteCarloAutocallableEngine":
        return oracle.autocallable_price(s, instrument._terms(), m, workers=16)
    return oracle.accumulator_price(s, instrument._terms(), m, workers=16)


def small_book(n=70):
    return workloads.book_4096(n=n, path_scale=1 / 64)


def test_book_fp64_matches_the_oracle_contract_by_contract(gpu, oracle):
    entries = small_book()
    estimates, tuning = Book(entries).price("fp64")
    for (engine, instrument, context), got in zip(entries, estimates):
        ref = oracle_entry(oracle, engine, instrument, context)
        assert got.price == pytest.approx(ref.price, rel=1e-10, abs=1e-12), type(instrument).__name__
        assert got.uncertainty_status == _abi.UNCERTAINTY_NAMES[ref.uncertainty_status]
    assert tuning.samples_per_thread in (1, 2, 4) and tuning.workgroups_per_cu in (2, 4, 8, 16)


def test_book_results_do_not_depend_on_tile_shape_or_graph(gpu):
    book = Book(small_book(40))
    base, _ = book.price("fp64", samples_per_thread=1, workgroups_per_cu=2, use_graph=False)
    for spt, wpc, graph in ((2, 4, True), (4, 8, True), (4, 2, False)):
        other, _ = book.price("fp64", samples_per_thread=spt, workgroups_per_cu=wpc, use_graph=graph, replays=3)
        assert [e.price for e in other] == [e.price for e in base]


def test_graph_replays_repeat_the_eager_result(gpu):
    book = Book(small_book(30))
    eager, _ = book.price("fp32", samples_per_thread=2, workgroups_per_cu=4, use_graph=False)
    for _ in range(3):
        replayed, _ = book.price("fp32", samples_per_thread=2, workgroups_per_cu=4, use_graph=True, replays=2)
        assert [e.price for e in replayed] == [e.price for e in eager]


def test_book_matches_single_contract_engines(gpu):
    entries = small_book(21)
    estimates, _ = Book(entries).price("fp64")
    for (engine, instrument, context), got in zip(entries, estimates):
        single = engine.price_with_statistics(instrument, context)
        assert got.price == pytest.approx(single.price, rel=1e-12, abs=1e-14)


def test_fp32_book_within_half_a_standard_error(gpu):
    book = Book(workloads.book_4096(n=280, path_scale=1 / 4))
    plain, _ = book.price("fp64")
    fast, _ = book.price("fp32")
    errors = np.array([abs(f.price - p.price) / p.path_standard_error for f, p in zip(fast, plain)
                       if p.path_standard_error > 0])
    assert np.mean(errors <= 0.5) >= 0.99, np.sort(errors)[-5:]
    assert errors.max() < 1.0


@pytest.mark.parametrize("averaging", [_abi.AVERAGE_GEOMETRIC, _abi.AVERAGE_ARITHMETIC])
@pytest.mark.parametrize("kind", [1, -1])
@pytest.mark.parametrize("paths,antithetic,sampling", [(4096, False, _abi.PSEUDO_RANDOM),
                                                      (3001, True, _abi.PSEUDO_RANDOM),
                                                      (2048, False, _abi.SOBOL)])
def test_asian_native_greeks_match_the_oracle(gpu, oracle, averaging, kind, paths, antithetic, sampling):
    t = _abi.AsianTerms(kind, averaging, 64, 0, 100.0, 0.0, us(START), us(START), us(END))
    s = settings(paths, antithetic=antithetic, sampling=sampling)
    m = market(rate=0.04, dividend=0.01)
    for request in (3, 1, 2):
        out, g = _abi.Estimate(), _abi.SpotGreeks(request, 0, 0.0, 0.0, 0.0)
        gpu.check(_native.lib().kg_asian_price_with_greeks(gpu.handle, C.byref(s), C.byref(t), C.byref(m),
                                                           C.byref(g), C.byref(out)))
        ref, rg = oracle.asian_price_with_greeks(s, t, m, request=request, workers=8)
        assert out.price == price_one(gpu, "kg_asian_price", s, t, m).price
        assert out.price == pytest.approx(ref.price, rel=1e-10)
        assert g.available == rg.available == request
        assert g.delta == pytest.approx(rg.delta, rel=1e-10, abs=1e-15)
        assert g.gamma == pytest.approx(rg.gamma, rel=1e-10, abs=1e-15)


def test_asian_engine_price_with_greeks(gpu, oracle):
    import kiyosi_amd as k
    opt = k.ArithmeticAveragePriceOption(option_type="put", strike=100, effective_time=START,
                                         averaging_start_time=START, expiry_time=END)
    ctx = k.PricingContext(model_parameters=k.BlackScholesMertonParameters(
        risk_free_rate=0.04, dividend_yield=0.01, volatility=0.2), spot_price=100, valuation_time=START)
    engine = k.MonteCarloAsianEngine(path_count=8192, seed=11, step_count=40)
    res = engine.price_with_greeks(opt, ctx, ["delta", "gamma"])
    assert res.price == engine.price(opt, ctx)
    _, rg = oracle.asian_price_with_greeks(engine._settings(), opt._terms(40), ctx.bsm_market(), workers=8)
    assert res.get("delta") == pytest.approx(rg.delta, rel=1e-10)
    assert res.get("gamma") == pytest.approx(rg.gamma, rel=1e-10)


@pytest.fixture(scope="module")
def full_book(gpu):
    """The 4,096-contract book at its full path counts (2^14..2^18 per
    contract), priced in one persistent launch in FP64 and in FP32 steps."""
    entries = workloads.book_4096()
    book = Book(entries)
    plain, _ = book.price("fp64")
    fast, tuning = book.price("fp32")
    return entries, plain, fast, tuning


def test_config4_full_book_fp32_within_half_a_standard_error(full_book):
    entries, plain, fast, tuning = full_book
    assert len(plain) == len(fast) == 4096
    errors = np.array([abs(f.price - p.price) / p.path_standard_error for f, p in zip(fast, plain)
                       if p.path_standard_error > 0])
    assert errors.size > 4000
    assert errors.max() <= 0.5, np.sort(errors)[-5:]


def test_config4_sampled_contracts_at_full_path_counts_match_the_oracle(full_book, oracle):
    """64 contracts drawn with a fixed seed from the full book: each FP64
    price from the one persistent launch equals the oracle's price of that
    contract alone, at its full path count, to 1e-10."""
    entries, plain, _, _ = full_book
    pick = np.random.default_rng(20251204).choice(len(entries), size=64, replace=False)
    kinds = set()
    for i in sorted(pick):
        engine, instrument, context = entries[i]
        ref = oracle_entry(oracle, engine, instrument, context)
        assert plain[i].price == pytest.approx(ref.price, rel=1e-10, abs=1e-12), (i, type(instrument).__name__)
        kinds.add(type(instrument).__name__)
    assert len(kinds) >= 5


@pytest.mark.parametrize("sampling", [_abi.PSEUDO_RANDOM, _abi.SOBOL])
def test_asian_native_greeks_redo_without_the_control(gpu, oracle, sampling):
    """Far out-of-the-money arithmetic puts at sigma 0.45 (the strikes whose
    controlled price turns negative in the strike-batch tests): the native
    Greeks entry resamples without the geometric control, as the reference
    does (asian.rs:618-630); price, delta and gamma equal the oracle's."""
    m = market(rate=0.04, dividend=0.01, sigma=0.45)
    s = settings(1024, sampling=sampling)
    for strike in (40.0, 45.0, 50.0):
        t = _abi.AsianTerms(-1, _abi.AVERAGE_ARITHMETIC, 8, 0, strike, 0.0, us(START), us(START), us(END))
        out, g = _abi.Estimate(), _abi.SpotGreeks(3, 0, 0.0, 0.0, 0.0)
        gpu.check(_native.lib().kg_asian_price_with_greeks(gpu.handle, C.byref(s), C.byref(t), C.byref(m),
                                                           C.byref(g), C.byref(out)))
        ref, rg = oracle.asian_price_with_greeks(s, t, m, request=3, workers=8)
        assert out.price >= 0.0
        assert out.price == pytest.approx(ref.price, rel=1e-10, abs=1e-14), strike
        assert g.delta == pytest.approx(rg.delta, rel=1e-10, abs=1e-14), strike
        assert g.gamma == pytest.approx(rg.gamma, rel=1e-10, abs=1e-14), strike


@pytest.mark.parametrize("averaging", [_abi.AVERAGE_GEOMETRIC, _abi.AVERAGE_ARITHMETIC])
@pytest.mark.parametrize("sampling", [_abi.PSEUDO_RANDOM, _abi.SOBOL])
def test_asian_native_greeks_fp32_within_half_a_standard_error(gpu, averaging, sampling):
    """The FP32-step spot-Greek kernels against the FP64 ones on the same
    draws: price within half a standard error of the FP64 run's per-path
    spread, delta and gamma within 1e-3 of their FP64 values."""
    t = _abi.AsianTerms(1, averaging, 64, 0, 100.0, 0.0, us(START), us(START), us(END))
    m = market(rate=0.04, dividend=0.01)
    res = {}
    for precision in (_abi.FP64, _abi.FP32_STEP):
        s = settings(1 << 16, sampling=sampling, precision=precision)
        out, g = _abi.Estimate(), _abi.SpotGreeks(3, 0, 0.0, 0.0, 0.0)
        gpu.check(_native.lib().kg_asian_price_with_greeks(gpu.handle, C.byref(s), C.byref(t), C.byref(m),
                                                           C.byref(g), C.byref(out)))
        res[precision] = (out, g)
    (plain, gp), (fast, gf) = res[_abi.FP64], res[_abi.FP32_STEP]
    assert abs(fast.price - plain.price) <= 0.5 * plain.path_standard_error
    assert gf.delta == pytest.approx(gp.delta, abs=1e-3)
    assert gf.gamma == pytest.approx(gp.gamma, abs=1e-3)


SEASON = START + dt.timedelta(days=120)
ASIAN_CASES = [
    # (name, kind, averaging, n, strike, realized, averaging_start, valuation)
    ("seasoned arithmetic call", 1, _abi.AVERAGE_ARITHMETIC, 40, 100.0, 92.0, START, SEASON),
    ("seasoned arithmetic put", -1, _abi.AVERAGE_ARITHMETIC, 40, 100.0, 104.0, START, SEASON),
    ("seasoned geometric call", 1, _abi.AVERAGE_GEOMETRIC, 40, 100.0, 97.0, START, SEASON),
    ("non-positive adjusted strike", 1, _abi.AVERAGE_ARITHMETIC, 40, 100.0, 400.0, START, SEASON),
    ("forward-start arithmetic call", 1, _abi.AVERAGE_ARITHMETIC, 40, 100.0, None,
     START + dt.timedelta(days=91), START),
    ("forward-start arithmetic put", -1, _abi.AVERAGE_ARITHMETIC, 33, 95.0, None,
     START + dt.timedelta(days=91), START),
    ("forward-start geometric", 1, _abi.AVERAGE_GEOMETRIC, 40, 100.0, None, START + dt.timedelta(days=91), START),
    ("single-fixing arithmetic", 1, _abi.AVERAGE_ARITHMETIC, 40, 100.0, None, END, START),
    ("single-fixing geometric put", -1, _abi.AVERAGE_GEOMETRIC, 40, 100.0, None, END, START),
]


@pytest.mark.parametrize("paths,antithetic,sampling", [(4096, False, _abi.PSEUDO_RANDOM),
                                                      (3001, True, _abi.PSEUDO_RANDOM),
                                                      (2048, False, _abi.SOBOL)])
@pytest.mark.parametrize("name,kind,averaging,n,strike,realized,avg_start,valuation", ASIAN_CASES,
                         ids=[c[0] for c in ASIAN_CASES])
def test_asian_carve_outs_match_the_oracle(gpu, oracle, name, kind, averaging, n, strike, realized, avg_start,
                                           valuation, paths, antithetic, sampling):
    """Seasoned averages, a non-positive adjusted strike, forward-start and
    single-fixing averages: price, delta and gamma of the GPU engine equal
    the oracle's at 1e-10 (asian.rs:374-560)."""
    t = _abi.AsianTerms(kind, averaging, n, int(realized is not None), strike, realized or 0.0, us(START),
                        us(avg_start), us(END))
    m = _abi.BsmMarket(100.0, 0.04, 0.01, 0.25, us(valuation))
    s = settings(paths, antithetic=antithetic, sampling=sampling)
    got, ref = price_one(gpu, "kg_asian_price", s, t, m), oracle.asian_price(s, t, m, workers=8)
    assert got.price == pytest.approx(ref.price, rel=1e-10, abs=1e-13), name
    assert got.uncertainty_status == ref.uncertainty_status
    out, g = _abi.Estimate(), _abi.SpotGreeks(3, 0, 0.0, 0.0, 0.0)
    gpu.check(_native.lib().kg_asian_price_with_greeks(gpu.handle, C.byref(s), C.byref(t), C.byref(m),
                                                       C.byref(g), C.byref(out)))
    ref_est, rg = oracle.asian_price_with_greeks(s, t, m, request=3, workers=8)
    assert out.price == got.price
    assert g.available == rg.available
    assert g.delta == pytest.approx(rg.delta, rel=1e-10, abs=1e-13), name
    assert g.gamma == pytest.approx(rg.gamma, rel=1e-10, abs=1e-13), name


@pytest.mark.parametrize("realized,avg_start,valuation", [(92.0, START, SEASON),
                                                          (None, START + dt.timedelta(days=91), START)])
def test_asian_strike_lanes_for_seasoned_and_forward_start_averages(gpu, oracle, realized, avg_start, valuation):
    strikes = [80.0, 95.0, 100.0, 110.0, 125.0]
    t = _abi.AsianTerms(1, _abi.AVERAGE_ARITHMETIC, 30, int(realized is not None), 100.0, realized or 0.0,
                        us(START), us(avg_start), us(END))
    m = _abi.BsmMarket(100.0, 0.04, 0.01, 0.25, us(valuation))
    s = settings(4096)
    arr = (C.c_double * len(strikes))(*strikes)
    out = (C.c_double * len(strikes))()
    gpu.check(_native.lib().kg_asian_price_strikes(gpu.handle, C.byref(s), C.byref(t), C.byref(m), arr,
                                                   len(strikes), out))
    for k, price in zip(strikes, out):
        t.strike = k
        assert price == pytest.approx(oracle.asian_price(s, t, m, workers=8).price, rel=1e-10, abs=1e-13), k


def test_book_prices_seasoned_and_forward_start_averages(gpu, oracle):
    from kiyosi_amd.engines import Book, MonteCarloAsianEngine
    from kiyosi_amd.instruments import ArithmeticAveragePriceOption, GeometricAveragePriceOption
    from kiyosi_amd.market import BlackScholesMertonParameters, PricingContext
    ctx_season = PricingContext(model_parameters=BlackScholesMertonParameters(
        risk_free_rate=0.04, dividend_yield=0.01, volatility=0.25), spot_price=100.0, valuation_time=SEASON)
    ctx_start = PricingContext(model_parameters=BlackScholesMertonParameters(
        risk_free_rate=0.04, dividend_yield=0.01, volatility=0.25), spot_price=100.0, valuation_time=START)
    engine = MonteCarloAsianEngine(path_count=4096, seed=77, step_count=30)
    entries = [
        (engine, ArithmeticAveragePriceOption(option_type="call", strike=100.0, effective_time=START,
                                              averaging_start_time=START, expiry_time=END, realized_average=92.0),
         ctx_season),
        (engine, GeometricAveragePriceOption(option_type="put", strike=100.0, effective_time=START,
                                             averaging_start_time=START, expiry_time=END, realized_average=101.0),
         ctx_season),
        (engine, ArithmeticAveragePriceOption(option_type="call", strike=100.0, effective_time=START,
                                              averaging_start_time=START + dt.timedelta(days=91), expiry_time=END),
         ctx_start),
    ]
    estimates, _ = Book(entries).price("fp64")
    for (eng, ins, ctx), got in zip(entries, estimates):
        ref = oracle.asian_price(eng._settings(), ins._terms(eng.step_count), ctx.bsm_market(), workers=8)
        assert got.price == pytest.approx(ref.price, rel=1e-10, abs=1e-13)


def test_a_kept_launch_graph_prices_the_next_book_of_the_same_shape(gpu):
    """The second book differs only in its seeds, so it has the same contract
    and tile counts and the call replays the graph kept from the first; the
    graph must read the new contracts."""
    first = small_book(30)
    second = [(type(e)(**dict(e._kwargs(), seed=e.seed + 1000)), i, c) for e, i, c in first]
    a, _ = Book(first).price("fp32")
    b, _ = Book(second).price("fp32")
    assert [e.price for e in a] != [e.price for e in b]
    assert [e.price for e in a] == [e.price for e in Book(first).price("fp32", use_graph=False)[0]]
    assert [e.price for e in b] == [e.price for e in Book(second).price("fp32", use_graph=False)[0]]
