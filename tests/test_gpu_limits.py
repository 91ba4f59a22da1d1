"""Input validation and size limits of the C ABI, mirroring the reference's
checks (MonteCarloSettings::validate, monte_carlo.rs:198-227; engine
`validate_*` functions), and one run at the GPU's largest path count."""

import ctypes as C
import datetime as dt

import pytest

from conftest import END, START, us
from kiyosi_amd import _abi, _native, workloads

pytestmark = pytest.mark.gpu


def settings(paths=4096, seed=7, sampling=_abi.PSEUDO_RANDOM, antithetic=0, precision=_abi.FP64, large=0,
             has_seed=1):
    return _abi.Settings(paths, seed, has_seed, sampling, antithetic, precision, 0, large)


def market(spot=100.0, sigma=0.2):
    return _abi.BsmMarket(spot, 0.05, 0.02, sigma, us(START))


def euro(kind=1, strike=100.0, expiry=END):
    return _abi.EuropeanTerms(kind, 0, strike, us(START), us(expiry))


def vanilla(gpu, s, o, m):
    out = _abi.Estimate()
    status = _native.lib().kg_vanilla_price(gpu.handle, C.byref(s), C.byref(o), C.byref(m), C.byref(out))
    return status, out


def message(gpu):
    return _native.lib().kg_last_error(gpu.handle).decode()


@pytest.mark.parametrize("kw,text", [
    (dict(paths=0), "path_count must be between 1 and 10000000"),
    (dict(paths=10_000_001), "path_count must be between 1 and 10000000"),
    (dict(paths=(1 << 26) + 1, large=1), "path_count must be between 1 and 67108864"),
    (dict(paths=3000, sampling=_abi.SOBOL), "power of two"),
    (dict(has_seed=0), "resolved seed"),
    (dict(precision=7), "unknown precision"),
    (dict(sampling=5), "unknown sampling"),
])
def test_settings_are_validated(gpu, kw, text):
    status, _ = vanilla(gpu, settings(**kw), euro(), market())
    assert status == _abi.INVALID_INPUT and text in message(gpu)


@pytest.mark.parametrize("terms,mkt,text", [
    (lambda: euro(kind=0), market, "option_type"),
    (lambda: euro(strike=0.0), market, "strike must be positive"),
    (lambda: euro(strike=-1.0), market, "strike must be positive"),
    (euro, lambda: market(spot=0.0), "spot must be positive"),
    (euro, lambda: market(spot=float("nan")), "spot must be positive"),
    (euro, lambda: market(sigma=-0.1), "volatility must be nonnegative"),
    (lambda: euro(expiry=START - dt.timedelta(days=1)), market, "valuation outside contract life"),
])
def test_european_terms_are_validated(gpu, terms, mkt, text):
    status, _ = vanilla(gpu, settings(), terms(), mkt())
    assert status == _abi.INVALID_INPUT and text in message(gpu)


def test_valuation_at_expiry_is_the_intrinsic_value(gpu):
    o = _abi.EuropeanTerms(1, 0, 90.0, us(START), us(START))
    status, out = vanilla(gpu, settings(), o, market())
    assert status == _abi.OK and out.price == 10.0 and out.has_seed == 0


def test_null_arguments_are_rejected(gpu):
    out = _abi.Estimate()
    s, o, m = settings(), euro(), market()
    lib = _native.lib()
    assert lib.kg_vanilla_price(gpu.handle, None, C.byref(o), C.byref(m), C.byref(out)) == _abi.INVALID_INPUT
    assert lib.kg_vanilla_price(gpu.handle, C.byref(s), None, C.byref(m), C.byref(out)) == _abi.INVALID_INPUT
    assert lib.kg_vanilla_price(None, C.byref(s), C.byref(o), C.byref(m), C.byref(out)) == _abi.INVALID_INPUT


def test_path_step_cap_follows_the_reference_and_allow_large_lifts_it(gpu):
    note, context, kw = workloads.snowball_252d_sobol_2p20()
    terms, m = note._terms(), context.bsm_market()
    out = _abi.Estimate()
    lib = _native.lib()
    big = settings(paths=1 << 22, sampling=_abi.SOBOL, precision=_abi.FP32_STEP)  # 4.2e6 x 252 > 1e9
    assert lib.kg_autocallable_price(gpu.handle, C.byref(big), C.byref(terms), C.byref(m),
                                     C.byref(out)) == _abi.INVALID_INPUT
    assert "one billion path steps" in message(gpu)
    big.allow_large = 1
    assert lib.kg_autocallable_price(gpu.handle, C.byref(big), C.byref(terms), C.byref(m), C.byref(out)) == _abi.OK
    assert abs(out.price - 0.9911) < 0.002


def test_largest_path_count_prices_near_the_analytic_value(gpu):
    """2^26 antithetic PCG paths, FP32 stepping: within 4 standard errors of
    Black-Scholes (9.227005508154047 for the reference's call)."""
    s = settings(paths=1 << 26, seed=1729, antithetic=1, precision=_abi.FP32_STEP, large=1)
    status, out = vanilla(gpu, s, euro(), market())
    assert status == _abi.OK
    assert out.standard_error < 0.002
    assert abs(out.price - 9.227005508154047) < 4 * out.standard_error


def test_empty_and_invalid_books(gpu):
    lib = _native.lib()
    tuning = _abi.BookTuning(0, 0, 0, 0, 0.0, 0.0)
    assert lib.kg_price_book(gpu.handle, None, 0, _abi.FP64, C.byref(tuning), None) == _abi.OK
    contracts = (_abi.BookContract * 1)()
    terms = euro()
    contracts[0] = _abi.BookContract(9, 0, settings(), market(), C.cast(C.pointer(terms), C.c_void_p))
    out = (_abi.Estimate * 1)()
    assert lib.kg_price_book(gpu.handle, contracts, 1, _abi.FP64, C.byref(tuning), out) == _abi.INVALID_INPUT
    assert "unknown contract kind" in message(gpu)
    tuning.samples_per_thread = 3
    asian = _abi.AsianTerms(1, 1, 64, 0, 100.0, 0.0, us(START), us(START), us(END))
    contracts[0] = _abi.BookContract(_abi.BOOK_ASIAN, 0, settings(), market(), C.cast(C.pointer(asian), C.c_void_p))
    assert lib.kg_price_book(gpu.handle, contracts, 1, _abi.FP64, C.byref(tuning), out) == _abi.INVALID_INPUT
    assert "samples_per_thread" in message(gpu)


def test_american_engine_limits(gpu):
    lib = _native.lib()
    model = _abi.AmericanModel(100.0, 0.05, 0.02, 0.2, us(START), None)
    report = _abi.LsmcReport()
    opts = _abi.LsmcOptions(_abi.BASIS_QUADRATIC, 0, 0, 4, 4096, 1e-3)
    terms = _abi.AmericanTerms(-1, 1, 100.0, us(START), us(END))
    s = settings()
    assert lib.kg_american_price(gpu.handle, C.byref(s), C.byref(terms), C.byref(model), C.byref(opts),
                                 C.byref(report)) == _abi.INVALID_INPUT
    terms.step_count = 50
    # Sobol paths price with the polynomial bases; the MLP regressor trains on
    # pseudorandom paths only.
    sob = settings(paths=4096, sampling=_abi.SOBOL)
    assert lib.kg_american_price(gpu.handle, C.byref(sob), C.byref(terms), C.byref(model), C.byref(opts),
                                 C.byref(report)) == _abi.OK
    mlp_default = _abi.LsmcOptions(_abi.BASIS_MLP, 0, 0, 4, 4096, 1e-3)
    assert lib.kg_american_price(gpu.handle, C.byref(sob), C.byref(terms), C.byref(model), C.byref(mlp_default),
                                 C.byref(report)) == _abi.UNSUPPORTED_OPERATION
    mlp = _abi.LsmcOptions(_abi.BASIS_MLP, 0, 0, 4, 16, 1e-3)
    assert lib.kg_american_price(gpu.handle, C.byref(s), C.byref(terms), C.byref(model), C.byref(mlp),
                                 C.byref(report)) == _abi.INVALID_INPUT
    assert "batch" in message(gpu)


def test_heston_limits(gpu):
    lib = _native.lib()
    out = _abi.Estimate()
    o = _abi.HestonTerms(_abi.HESTON_EUROPEAN, 1, 12, 0, 100.0, us(START), us(END))
    bad_rho = _abi.HestonMarket(100.0, 0.03, 0.0, 0.04, 2.0, 0.04, 0.5, -1.5, us(START))
    assert lib.kg_heston_price(gpu.handle, C.byref(settings()), C.byref(o), C.byref(bad_rho),
                               C.byref(out)) == _abi.INVALID_INPUT
    ok = _abi.HestonMarket(100.0, 0.03, 0.0, 0.04, 2.0, 0.04, 0.5, -0.7, us(START))
    assert lib.kg_heston_price(gpu.handle, C.byref(settings(sampling=_abi.SOBOL)), C.byref(o), C.byref(ok),
                               C.byref(out)) == _abi.UNSUPPORTED_OPERATION


def test_a_failed_call_leaves_the_context_usable(gpu):
    status, _ = vanilla(gpu, settings(paths=0), euro(), market())
    assert status == _abi.INVALID_INPUT
    status, out = vanilla(gpu, settings(paths=65536, seed=1009), euro(), market())
    assert status == _abi.OK and out.price == pytest.approx(9.249574403327761, rel=1e-10)
    assert message(gpu) == ""
