"""Heston path stepping: the oracle pinned by closed forms and the reference's
Heston analytic price (CPU); the GPU engine checked against the oracle."""

import ctypes as C
import math

import pytest
from scipy.stats import norm

from conftest import END, START, golden, us
from kiyosi_amd import _abi, _native


def market(v0=0.04, kappa=2.0, theta=0.04, xi=0.5, rho=-0.7, r=0.03, q=0.0, spot=100.0):
    return _abi.HestonMarket(spot, r, q, v0, kappa, theta, xi, rho, us(START))


def terms(payoff=_abi.HESTON_ASIAN_ARITHMETIC, kind=1, steps=252, strike=100.0):
    return _abi.HestonTerms(payoff, kind, steps, 0, strike, us(START), us(END))


def settings(paths, seed=1729, antithetic=False, precision=_abi.FP64, large=False):
    return _abi.Settings(paths, seed, 1, _abi.PSEUDO_RANDOM, int(antithetic), precision, 0, int(large))


def bsm_call(s, k, r, q, sigma, t):
    d1 = (math.log(s / k) + (r - q + 0.5 * sigma**2) * t) / (sigma * math.sqrt(t))
    d2 = d1 - sigma * math.sqrt(t)
    return s * math.exp(-q * t) * norm.cdf(d1) - k * math.exp(-r * t) * norm.cdf(d2)


def test_zero_vol_of_vol_reduces_to_black_scholes(oracle):
    m = market(xi=0.0, kappa=1.3, r=0.05, q=0.02)
    est = oracle.heston_price(settings(400_000, antithetic=True), terms(_abi.HESTON_EUROPEAN, steps=12), m,
                              workers=8)
    assert abs(est.price - bsm_call(100, 100, 0.05, 0.02, 0.2, 1.0)) < 4 * est.standard_error


def test_zero_vol_of_vol_geometric_asian_closed_form(oracle):
    n, theta, r = 50, 0.04, 0.03
    m = market(xi=0.0, theta=theta, r=r)
    est = oracle.heston_price(settings(400_000), terms(_abi.HESTON_ASIAN_GEOMETRIC, steps=n), m, workers=8)
    dt = 1.0 / n
    mu = (r - 0.5 * theta) * dt * (n + 1) / 2
    var = theta * dt * (n + 1) * (2 * n + 1) / (6 * n)
    d1 = (math.log(100 / 100) + mu + var) / math.sqrt(var)
    d2 = d1 - math.sqrt(var)
    expected = math.exp(-r) * (100 * math.exp(mu + 0.5 * var) * norm.cdf(d1) - 100 * norm.cdf(d2))
    assert abs(est.price - expected) < 4 * est.standard_error


def test_heston_european_against_the_reference_analytic_price(oracle):
    """kiyosi's Fourier Heston call (references/heston.json) as a model-level
    cross-check; the Euler scheme's bias at 500 steps is allowed 0.02."""
    ref = {c["id"]: c for c in golden("references.json")["heston"]}["heston-analytic-call"]
    m = market(v0=0.04, kappa=1.5, theta=0.04, xi=0.3, rho=-0.7, r=0.05, q=0.02)
    est = oracle.heston_price(settings(300_000, antithetic=True), terms(_abi.HESTON_EUROPEAN, steps=500), m,
                              workers=8)
    assert abs(est.price - ref["value"]) < 4 * est.standard_error + 0.02


def test_put_call_parity_on_shared_paths(oracle):
    m = market()
    s = settings(20_000)
    call = oracle.heston_price(s, terms(_abi.HESTON_EUROPEAN, 1, steps=64), m, workers=8)
    put = oracle.heston_price(s, terms(_abi.HESTON_EUROPEAN, -1, steps=64), m, workers=8)
    # E[S_T] under the Euler log-spot scheme is not exactly the forward; the
    # difference of the two estimators is the sample mean of S_T - K on
    # identical paths, which must be close to the forward.
    forward = 100 * math.exp(0.03) - 100
    assert abs((call.price - put.price) * math.exp(0.03) - forward) < 0.15


def gpu_price(ctx, s, o, m):
    out = _abi.Estimate()
    ctx.check(_native.lib().kg_heston_price(ctx.handle, C.byref(s), C.byref(o), C.byref(m), C.byref(out)))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("payoff", [_abi.HESTON_EUROPEAN, _abi.HESTON_ASIAN_ARITHMETIC, _abi.HESTON_ASIAN_GEOMETRIC])
@pytest.mark.parametrize("paths,antithetic", [(1, False), (300, False), (4099, True), (20_000, False)])
def test_gpu_fp64_matches_the_oracle(gpu, oracle, payoff, paths, antithetic):
    s, o, m = settings(paths, antithetic=antithetic), terms(payoff, steps=63), market()
    got, ref = gpu_price(gpu, s, o, m), oracle.heston_price(s, o, m, workers=8)
    assert got.price == pytest.approx(ref.price, rel=1e-10, abs=1e-300)
    assert got.uncertainty_status == ref.uncertainty_status


@pytest.mark.gpu
def test_config3_full_size_fp32_within_half_a_standard_error(gpu, oracle):
    """2^22 PCG paths x 252 steps: FP32 stepping against the FP64 GPU engine on
    the same draws, and the FP64 GPU engine against the oracle on a
    2^16-path prefix of the same stream layout."""
    o, m = terms(), market()
    plain = gpu_price(gpu, settings(1 << 22, large=True), o, m)
    fast = gpu_price(gpu, settings(1 << 22, precision=_abi.FP32_STEP, large=True), o, m)
    assert abs(fast.price - plain.price) <= 0.5 * plain.standard_error
    s = settings(1 << 16)
    assert gpu_price(gpu, s, o, m).price == pytest.approx(oracle.heston_price(s, o, m, workers=16).price,
                                                          rel=1e-10)
