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


def test_fourier_checker_reproduces_the_reference_analytic_price():
    """The closed-form checker (oracle/heston_analytic.py) against the
    reference's analytic Heston call (references/heston.json, uncertainty 1e-11) and the
    Black-Scholes limit xi -> 0."""
    from oracle.heston_analytic import heston_call
    ref = {c["id"]: c for c in golden("references.json")["heston"]}
    for key in ("heston-analytic-call", "heston-fd-european-call"):
        i = ref[key]["inputs"]
        got = heston_call(i["spot"], i["strike"], i["rate"], i["dividend"], 1.0, i["initial_variance"],
                          i["long_run_variance"], i["mean_reversion_speed"], i["vol_of_vol"], i["correlation"])
        assert got == pytest.approx(ref[key]["value"], abs=5e-11)
    # Uncorrelated, so the price moves by O(xi^2) from Black-Scholes.
    assert heston_call(100, 100, 0.05, 0.02, 1.0, 0.04, 0.04, 1.5, 1e-4, 0.0) == pytest.approx(
        bsm_call(100, 100, 0.05, 0.02, 0.2, 1.0), abs=1e-6)


# Reference Heston parameters (references/heston.json) with either sign of the
# correlation. The 80-strike put moves from 1.348 (rho -0.7) to 0.898 (rho 0)
# and 0.319 (rho +0.7), the 120-strike call from 1.671 to 2.633 and 3.386, so
# a dropped, halved or sign-flipped correlation, or a misplaced vol of vol,
# fails by tens of standard errors. The full-truncation Euler bias at 64 steps
# (a quarter of the 16-step bias, which is about 0.01) is allowed 0.003.
@pytest.mark.parametrize("rho", [-0.7, 0.7])
@pytest.mark.parametrize("kind,strike", [(-1, 80.0), (1, 120.0), (1, 100.0)])
def test_euler_scheme_converges_to_the_fourier_price(oracle, rho, kind, strike):
    from oracle.heston_analytic import heston_call, heston_put
    m = market(v0=0.04, kappa=1.5, theta=0.04, xi=0.3, rho=rho, r=0.05, q=0.02)
    est = oracle.heston_price(settings(1 << 20, antithetic=True), terms(_abi.HESTON_EUROPEAN, kind, 64, strike), m,
                              workers=8)
    exact = (heston_call if kind == 1 else heston_put)(100, strike, 0.05, 0.02, 1.0, 0.04, 0.04, 1.5, 0.3, rho)
    assert abs(est.price - exact) < 4 * est.standard_error + 0.003, (est.price, exact, est.standard_error)


def test_scheme_keeps_the_spot_martingale_and_the_variance_mean(oracle):
    """Conditionally on the step's start, the log-spot increment is Gaussian
    with mean (r - q - v+/2) dt and variance v+ dt, so E[S_T] is exactly
    S0 e^{(r-q)T}; the variance recursion has mean
    E[v_{k+1}] = E[v_k] (1 - kappa dt) + kappa theta dt wherever v_k >= 0,
    i.e. theta + (v0 - theta)(1 - kappa dt)^n when truncation is rare (Feller
    holds here: 2 kappa theta = 0.12 > xi^2 = 0.09). A wrong drift correction,
    a missing kappa, theta or dt fails one of the two."""
    n = 64
    m = market(v0=0.09, kappa=1.5, theta=0.04, xi=0.3, rho=-0.7, r=0.05, q=0.02)
    spot, var = oracle.heston_terminal(settings(1 << 20, antithetic=True), terms(_abi.HESTON_EUROPEAN, steps=n), m,
                                       workers=8)
    assert abs(spot.price - 100 * math.exp(0.03)) < 4 * spot.standard_error
    expected = 0.04 + 0.05 * (1 - 1.5 / n) ** n
    assert abs(var.price - expected) < 4 * var.standard_error + 1e-5
    assert var.standard_error < 1e-4


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


@pytest.mark.gpu
def test_gpu_richardson_price_meets_the_reference_analytic_call(gpu, oracle):
    """The FP64 GPU engine at 2^26 antithetic paths, 64 and 128 steps: the
    Richardson combination 2 P(128) - P(64) meets kiyosi's analytic and FD
    Heston call (9.011278407541738, references/heston.json) within 3 of its
    standard errors, which are small enough that dropping the correlation
    (price 8.985, 0.026 lower) fails. The GPU equals the oracle on a 2^15-path
    prefix of the same stream layout at both step counts."""
    ref = {c["id"]: c["value"] for c in golden("references.json")["heston"]}
    m = market(v0=0.04, kappa=1.5, theta=0.04, xi=0.3, rho=-0.7, r=0.05, q=0.02)
    est = {}
    for n in (64, 128):
        o = terms(_abi.HESTON_EUROPEAN, 1, n)
        est[n] = gpu_price(gpu, settings(1 << 26, seed=20251204, antithetic=True, large=True), o, m)
        prefix = settings(1 << 15, seed=20251204, antithetic=True)
        assert gpu_price(gpu, prefix, o, m).price == pytest.approx(
            oracle.heston_price(prefix, o, m, workers=8).price, rel=1e-10)
    extrapolated = 2 * est[128].price - est[64].price
    se = (4 * est[128].standard_error ** 2 + est[64].standard_error ** 2) ** 0.5
    assert 3 * se < 0.026 / 2
    for key in ("heston-analytic-call", "heston-fd-european-call"):
        assert abs(extrapolated - ref[key]) < 3 * se, (extrapolated, ref[key], se)
