"""kg_vanilla_price_noalloc through the C ABI: the variant of kg_vanilla_price
that keeps its partials, flag word and merged moments in one fixed device
array per GPU. It equals kg_vanilla_price field for field on consecutive
calls (the second smaller, so stale partials of the first would show), for
pseudorandom and Sobol draws, FP64 and FP32 steps and the peak-scaled
two-pass payoff, and refuses requests beyond its workspace."""

import ctypes as C

import pytest

from conftest import END, START, us
from kiyosi_amd import _abi, _native

pytestmark = pytest.mark.gpu
FIELDS = ("price", "standard_error", "ci_lower", "ci_upper", "has_standard_error", "uncertainty_status",
          "has_seed", "seed", "path_standard_error", "path_steps")


def call(ctx, name, s, o, m):
    out = _abi.Estimate()
    status = getattr(_native.lib(), name)(ctx.handle, C.byref(s), C.byref(o), C.byref(m), C.byref(out))
    return status, out


def option(strike=100.0, kind=1):
    return _abi.EuropeanTerms(kind, 0, strike, us(START), us(END))


def market(spot=100.0, sigma=0.2, rate=0.05, dividend=0.02):
    return _abi.BsmMarket(spot, rate, dividend, sigma, us(START))


CASES = [
    # (paths, sampling, antithetic, precision, strike, market)
    ((1_000_000, 65537), _abi.PSEUDO_RANDOM, True, _abi.FP64, 100.0, market()),
    ((8_388_608, 4097), _abi.PSEUDO_RANDOM, False, _abi.FP64, 90.0, market()),
    ((1 << 20, 1 << 12), _abi.SOBOL, False, _abi.FP64, 100.0, market()),
    ((1 << 18, 1 << 10), _abi.SOBOL, True, _abi.FP32_STEP, 110.0, market()),
    ((500_000, 999), _abi.PSEUDO_RANDOM, True, _abi.FP32_STEP, 100.0, market()),
    ((65536, 256), _abi.PSEUDO_RANDOM, False, _abi.FP64, 1.0, market(spot=1e308, rate=800.0, dividend=0.0,
                                                                    sigma=40.0)),
]


@pytest.mark.parametrize("paths,sampling,antithetic,precision,strike,m", CASES)
def test_noalloc_equals_the_allocating_entry_point_on_consecutive_calls(gpu, paths, sampling, antithetic,
                                                                        precision, strike, m):
    for n in paths:  # the second call is the smaller one
        s = _abi.Settings(n, 1729, 1, sampling, int(antithetic), precision, 0, 0)
        st_a, ref = call(gpu, "kg_vanilla_price", s, option(strike), m)
        st_b, got = call(gpu, "kg_vanilla_price_noalloc", s, option(strike), m)
        assert st_a == st_b == _abi.OK, _native.lib().kg_last_error(gpu.handle)
        for f in FIELDS:
            assert getattr(got, f) == getattr(ref, f), (n, f)


def test_noalloc_refuses_requests_beyond_its_workspace(gpu):
    s = _abi.Settings(8_388_609, 1, 1, _abi.PSEUDO_RANDOM, 0, _abi.FP64, 0, 0)
    status, _ = call(gpu, "kg_vanilla_price_noalloc", s, option(), market())
    assert status == _abi.INVALID_INPUT
    assert b"8388608" in _native.lib().kg_last_error(gpu.handle)
    # Antithetic pairs halve the samples: 10,000,000 paths are 5,000,000 samples.
    s = _abi.Settings(10_000_000, 1, 1, _abi.PSEUDO_RANDOM, 1, _abi.FP64, 0, 0)
    status, _ = call(gpu, "kg_vanilla_price_noalloc", s, option(), market())
    assert status == _abi.OK


def test_noalloc_known_values_and_errors_match(gpu):
    at_expiry = _abi.EuropeanTerms(1, 0, 90.0, us(START), us(START))
    s = _abi.Settings(1000, 3, 1, _abi.PSEUDO_RANDOM, 0, _abi.FP64, 0, 0)
    a = call(gpu, "kg_vanilla_price", s, at_expiry, market())
    b = call(gpu, "kg_vanilla_price_noalloc", s, at_expiry, market())
    assert a[0] == b[0] == _abi.OK and a[1].price == b[1].price == 10.0
    bad = _abi.EuropeanTerms(3, 0, 90.0, us(START), us(END))
    assert call(gpu, "kg_vanilla_price_noalloc", s, bad, market())[0] == _abi.INVALID_INPUT
