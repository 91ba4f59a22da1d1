"""Least-squares Monte Carlo: the oracle pinned by the reference's per-seed
LSMC prices and its finite-difference references (CPU); the GPU engine
checked against the oracle and the FD judge."""

import ctypes as C
import math

import numpy as np
import pytest

from conftest import END, START, fd_richardson, golden, us
from kiyosi_amd import _abi, _native, workloads
from kiyosi_amd.market import LocalVolatilitySurface, american_model


def bsm(spot=100.0, rate=0.05, dividend=0.02, sigma=0.2):
    return _abi.BsmMarket(spot, rate, dividend, sigma, us(START))


def put(strike=100.0):
    return _abi.EuropeanTerms(-1, 0, strike, us(START), us(END))


def settings(paths, seed=1009, antithetic=False, precision=_abi.FP64, large=False, sampling=_abi.PSEUDO_RANDOM):
    return _abi.Settings(paths, seed, 1, sampling, int(antithetic), precision, 0, int(large))


def test_lsmc_calibration_vectors(oracle):
    """Per-seed in-sample LSMC prices of kiyosi's American engine."""
    for run in golden("calibration_american.json")["runs"]:
        p = run["profile"]
        for sample in run["samples"][:4] + [run["timing_seed"]]:
            price, _ = oracle.american_lsmc(settings(p["path_count"], sample["seed"]), put(), bsm(), p["step_count"],
                                            workers=8)
            assert price == pytest.approx(sample["price"], rel=1e-13), (run["dimension"], sample)


def test_fd_judge_matches_the_reference_fd_prices(oracle):
    refs = {c["id"]: c["value"] for k in ("fd_local_vol", "fd_bsm") for c in golden("references.json")[k]}
    two = LocalVolatilitySurface([START, END], [80.0, 120.0], [[0.30, 0.15], [0.35, 0.20]])._struct()
    assert fd_richardson(oracle, two) == pytest.approx(refs["fd-local-vol-american-put"], abs=1e-4)
    flat = LocalVolatilitySurface([START, END], [80.0, 120.0], [[0.2, 0.2], [0.2, 0.2]])._struct()
    assert fd_richardson(oracle, flat) == pytest.approx(refs["fd-bsm-american-put"], abs=1e-4)


def test_local_vol_lsmc_with_a_flat_surface_is_the_reference_lsmc(oracle):
    """Log-Euler steps under a constant surface are the reference's exact BSM
    steps, so the local-vol LSMC reproduces kiyosi's quadratic LSMC."""
    flat = LocalVolatilitySurface([START, END], [80.0, 120.0], [[0.2, 0.2], [0.2, 0.2]])
    from kiyosi_amd.market import LocalVolatilityParameters, PricingContext
    ctx = PricingContext(model_parameters=LocalVolatilityParameters(risk_free_rate=0.05, dividend_yield=0.02,
                                                                    surface=flat),
                         spot_price=100.0, valuation_time=START)
    terms = _abi.AmericanTerms(-1, 50, 100.0, us(START), us(END))
    lv, _ = oracle.lv_lsmc(settings(65536), terms, american_model(ctx), _abi.BASIS_QUADRATIC, workers=8)
    ref = golden("calibration_american.json")["runs"][0]["samples"][0]
    assert lv == pytest.approx(ref["price"], rel=1e-9)


def two_by_two_model():
    """The reference's fd_local_vol surface (cases/fd_local_vol.json): sigma 0.30
    and 0.15 at spots 80 and 120 on the valuation date, 0.35 and 0.20 a year on."""
    from kiyosi_amd.market import LocalVolatilityParameters, PricingContext
    two = LocalVolatilitySurface([START, END], [80.0, 120.0], [[0.30, 0.15], [0.35, 0.20]])
    ctx = PricingContext(model_parameters=LocalVolatilityParameters(risk_free_rate=0.05, dividend_yield=0.02,
                                                                    surface=two),
                         spot_price=100.0, valuation_time=START)
    return american_model(ctx)


def test_local_vol_stepping_converges_to_the_reference_fd_prices(oracle):
    """The LSMC engine's log-Euler local-volatility steps (the oracle's lv_step,
    which the GPU engine matches at 1e-8) price a European call and a cash
    digital under the 2x2 surface; Richardson over 64 and 128 steps meets the
    reference's FD prices (references/fd_local_vol.json) within 3 standard
    errors of the extrapolated estimate. Reading the surface with time and
    spot swapped prices the call near 15.6, a sigma for sigma^2 in the drift
    moves it by about 5, and a surface read at the wrong time row by more than
    1: each fails by many standard errors."""
    refs = {c["id"]: c["value"] for c in golden("references.json")["fd_local_vol"]}
    model = two_by_two_model()
    s = settings(1 << 20, seed=1729, antithetic=True)
    for kind, cash, key in ((0, 1.0, "fd-local-vol-european-call"), (1, 10.0, "fd-local-vol-cash-digital-call")):
        coarse, fine = (oracle.lv_european(s, _abi.AmericanTerms(1, n, 100.0, us(START), us(END)), model, kind, cash,
                                           workers=8) for n in (64, 128))
        extrapolated = 2 * fine.price - coarse.price
        se = (4 * fine.standard_error**2 + coarse.standard_error**2) ** 0.5
        assert abs(extrapolated - refs[key]) < 3 * se, (key, extrapolated, refs[key], se)
        assert abs(fine.price - refs[key]) < 4 * fine.standard_error + 0.01


def test_local_vol_stepping_keeps_the_spot_martingale(oracle):
    """Each log-Euler step is Gaussian in log-spot given its start, with drift
    (r - q - sigma^2/2) dt, so E[S_T] = S0 e^{(r-q)T} exactly: a call struck at
    1e-9 is worth S0 e^{-qT} - K e^{-rT} within its standard error."""
    est = oracle.lv_european(settings(1 << 19, seed=5), _abi.AmericanTerms(1, 50, 1e-9, us(START), us(END)),
                             two_by_two_model(), 0, 1.0, workers=8)
    exact = 100 * math.exp(-0.02) - 1e-9 * math.exp(-0.05)
    assert abs(est.price - exact) < 4 * est.standard_error
    assert est.standard_error < 0.05


@pytest.mark.parametrize("sampling", [_abi.PSEUDO_RANDOM, _abi.SOBOL])
@pytest.mark.parametrize("reverse", [False, True])
def test_lsmc_paths_have_brownian_covariance(oracle, sampling, reverse):
    """The normalised coordinates after k intervals, drawn forward (running
    sums) or in reverse (terminal first, then the Brownian bridge back in
    time, american.rs:278-316), have E[W_k] = 0, E[W_k W_j] = min(k, j) and
    uncorrelated unit increments. A wrong bridge weight, a deviation without
    its square root or an off-by-one in the elapsed count breaks the
    variances by several percent."""
    steps = 20
    w = oracle.american_coordinates(settings(1 << 16, seed=42, sampling=sampling), steps, reverse, workers=8)
    n = w.shape[1]
    k = np.arange(1, steps + 1)
    for j in (0, 9, steps - 1):
        prod = w * w[j]
        se = prod.std(axis=1) / np.sqrt(n)
        assert np.all(np.abs(prod.mean(axis=1) - np.minimum(k, j + 1)) < 5 * se + 1e-12), j
    inc = np.diff(np.vstack([np.zeros(n), w]), axis=0)
    assert np.all(np.abs(inc.mean(axis=1)) < 5 / np.sqrt(n))
    lag = inc[1:] * inc[:-1]
    assert np.all(np.abs(lag.mean(axis=1)) < 5 * lag.std(axis=1) / np.sqrt(n))


def test_reverse_paths_reference_case(oracle):
    """The reference's reverse-mode test (tests/american_monte_carlo.rs:187-208):
    1,000,000 paths x 150 intervals (above 2^25 path values, so the engine
    draws in reverse), seed 42, within 0.04 of the lattice price; here the
    reference's FD price of the same put (references/fd_bsm.json) stands in
    for its 2000-step binomial price."""
    s = settings(1_000_000, seed=42)
    auto, _ = oracle.american_lsmc(s, put(), bsm(), 150, workers=8)
    forced, _ = oracle.american_lsmc(s, put(), bsm(), 150, workers=8, mode=1)
    assert auto == forced
    fd = {c["id"]: c["value"] for c in golden("references.json")["fd_bsm"]}["fd-bsm-american-put"]
    assert abs(auto - fd) < 0.04


@pytest.mark.parametrize("paths", [1 << 18, 1 << 20])
def test_sobol_lsmc_near_the_fd_price(oracle, paths):
    """Sobol draws in forward mode (2^18 x 50) and in reverse mode (2^20 x 50,
    dimension intervals - k for the coordinate after k intervals) price the
    reference's BSM put near its FD value."""
    fd = {c["id"]: c["value"] for c in golden("references.json")["fd_bsm"]}["fd-bsm-american-put"]
    price, _ = oracle.american_lsmc(settings(paths, seed=1729, sampling=_abi.SOBOL), put(), bsm(), 50, workers=8)
    assert abs(price - fd) < 0.04


def test_anchored_paths_are_scale_invariant(oracle):
    """Spot and strike at 1e255: ln S0 plus ten deviations over 50 steps passes
    600, so the reference steps each leg's spot directly (anchored paths,
    american.rs:1149-1157, 355-362). The put is homogeneous of degree one in
    (spot, strike), so on the same draws it is 1e253 times the price at 100,
    up to rounding."""
    for antithetic in (False, True):
        s = settings(1 << 16, seed=7, antithetic=antithetic)
        small, _ = oracle.american_lsmc(s, put(), bsm(), 50, workers=8)
        big, _ = oracle.american_lsmc(s, put(1e255), bsm(1e255), 50, workers=8)
        assert big / 1e253 == pytest.approx(small, rel=1e-11)


def test_out_of_sample_is_below_in_sample_and_near_fd(oracle):
    s = settings(131072, seed=3)
    ins, _ = oracle.american_lsmc(s, put(), bsm(), 50, workers=8)
    oos = oracle.american_out_of_sample(s, 131072, put(), bsm(), 50, workers=8)
    assert oos < ins + 0.02
    assert abs(oos - 6.660684545893456) < 0.06


# ---------------------------------------------------------------- GPU
def gpu_american(ctx, s, terms, model, basis=_abi.BASIS_QUADRATIC, training=0):
    options = _abi.LsmcOptions(basis, int(training > 0), training, 4, 4096, 1e-3)
    report = _abi.LsmcReport()
    ctx.check(_native.lib().kg_american_price(ctx.handle, C.byref(s), C.byref(terms), C.byref(model),
                                              C.byref(options), C.byref(report)))
    return report


def bsm_model(sigma=0.2):
    return _abi.AmericanModel(100.0, 0.05, 0.02, sigma, us(START), None)


def am_terms(steps=50):
    return _abi.AmericanTerms(-1, steps, 100.0, us(START), us(END))


@pytest.mark.gpu
def test_gpu_quadratic_lsmc_reproduces_the_calibration_vectors(gpu):
    for run in golden("calibration_american.json")["runs"]:
        p = run["profile"]
        for sample in run["samples"][:2]:
            got = gpu_american(gpu, settings(p["path_count"], sample["seed"]), am_terms(p["step_count"]), bsm_model())
            assert got.price == pytest.approx(sample["price"], rel=1e-9), (run["dimension"], sample)


@pytest.mark.gpu
@pytest.mark.parametrize("antithetic", [False, True])
def test_gpu_quadratic_lsmc_matches_the_oracle(gpu, oracle, antithetic):
    s = settings(40_001, seed=11, antithetic=antithetic)
    got = gpu_american(gpu, s, am_terms(), bsm_model())
    ref, _ = oracle.american_lsmc(s, put(), bsm(), 50, workers=8)
    assert got.price == pytest.approx(ref, rel=1e-9)


@pytest.mark.gpu
def test_gpu_out_of_sample_matches_the_oracle(gpu, oracle):
    s = settings(30_000, seed=5)
    got = gpu_american(gpu, s, am_terms(), bsm_model(), training=50_000)
    ref = oracle.american_out_of_sample(s, 50_000, put(), bsm(), 50, workers=8)
    assert got.price == pytest.approx(ref, rel=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("basis", [_abi.BASIS_QUADRATIC, _abi.BASIS_POLY4])
def test_gpu_local_vol_lsmc_matches_the_oracle(gpu, oracle, basis):
    option, context, engine, _ = workloads.american_put_local_vol()
    model = american_model(context)
    terms = option._american_terms(50)
    s = settings(20_000, seed=7, antithetic=True)
    got = gpu_american(gpu, s, terms, model, basis, training=30_000)
    ins, oos = oracle.lv_lsmc(s, terms, model, basis, 30_000, workers=8)
    assert got.in_sample_price == pytest.approx(ins, rel=1e-8)
    assert got.price == pytest.approx(oos, rel=1e-8)


@pytest.mark.gpu
@pytest.mark.parametrize("paths", [1 << 18, 1 << 20])
def test_gpu_fp32_lsmc_within_half_a_standard_error(gpu, paths):
    """FP32 steps (single-precision local-volatility interpolation included)
    against FP64 steps on the same draws, up to the full 2^20 training and
    2^20 pricing paths of the local-volatility put."""
    option, context, engine, _ = workloads.american_put_local_vol()
    model = american_model(context)
    terms = option._american_terms(50)
    plain = gpu_american(gpu, settings(paths, seed=9), terms, model, _abi.BASIS_POLY4, training=paths)
    fast = gpu_american(gpu, settings(paths, seed=9, precision=_abi.FP32_STEP), terms, model, _abi.BASIS_POLY4,
                        training=paths)
    assert abs(fast.price - plain.price) <= 0.5 * plain.standard_error


@pytest.mark.gpu
def test_config5_poly4_lower_bound_near_the_fd_judge(gpu, oracle):
    """2^20 training + 2^20 pricing paths: the poly-4 out-of-sample price is a
    lower bound of the 50-date Bermudan value, within 1 SE plus the
    Bermudan/Euler gap (0.02) of the FD American price."""
    option, context, engine, training = workloads.american_put_local_vol()
    model = american_model(context)
    terms = option._american_terms(50)
    rep = gpu_american(gpu, settings(1 << 20, seed=1729, precision=_abi.FP32_STEP), terms, model, _abi.BASIS_POLY4,
                       training=training)
    quad = gpu_american(gpu, settings(1 << 20, seed=1729, precision=_abi.FP32_STEP), terms, model,
                        _abi.BASIS_QUADRATIC, training=training)
    fd = fd_richardson(oracle, context.model_parameters.surface._struct(), dividend=0.0)
    assert rep.price <= fd + 3 * rep.standard_error
    assert abs(rep.price - fd) < rep.standard_error + 0.02
    assert rep.price >= quad.price - 2 * rep.standard_error


@pytest.mark.gpu
@pytest.mark.parametrize("paths,antithetic", [(1 << 20, False), (1_400_001, True)])
def test_gpu_reverse_paths_match_the_oracle(gpu, oracle, paths, antithetic):
    """Above 2^25 path values the engine draws the Brownian coordinates in
    reverse as the reference does (terminal first, the bridge back one date
    per regression): 2^20 x 50 (the size of the local-vol American workload)
    and 700,001 antithetic samples x 50 (a last block of 97 samples, an odd
    stream offset per row), FP64 against the oracle at 1e-9."""
    s = settings(paths, seed=1729, antithetic=antithetic, large=True)
    got = gpu_american(gpu, s, am_terms(), bsm_model())
    ref, _ = oracle.american_lsmc(s, put(), bsm(), 50, workers=16)
    forward, _ = oracle.american_lsmc(s, put(), bsm(), 50, workers=16, mode=0)
    assert ref != forward  # the reverse draws really differ from the forward ones
    assert got.price == pytest.approx(ref, rel=1e-9)


@pytest.mark.gpu
def test_gpu_reverse_paths_fp32_within_half_a_standard_error(gpu):
    """FP32 bridge and spots against the FP64 reverse-mode engine on the same
    draws, judged by the out-of-sample standard error of a 2^20-path pricing."""
    plain = gpu_american(gpu, settings(1 << 20, seed=5, large=True), am_terms(), bsm_model(), training=1 << 20)
    fast = gpu_american(gpu, settings(1 << 20, seed=5, precision=_abi.FP32_STEP, large=True), am_terms(), bsm_model(),
                        training=1 << 20)
    assert abs(fast.price - plain.price) <= 0.5 * plain.standard_error
    assert abs(fast.in_sample_price - plain.in_sample_price) <= 0.5 * plain.standard_error


@pytest.mark.gpu
@pytest.mark.parametrize("paths,antithetic", [(1 << 12, False), (1 << 16, True), (1 << 20, False)])
def test_gpu_sobol_lsmc_matches_the_oracle(gpu, oracle, paths, antithetic):
    """Sobol draws: forward (2^12, 2^16 antithetic) and reverse (2^20 x 50,
    dimension 50 - k for the coordinate after k intervals), FP64 at 1e-9."""
    s = settings(paths, seed=42, antithetic=antithetic, large=True, sampling=_abi.SOBOL)
    got = gpu_american(gpu, s, am_terms(), bsm_model())
    ref, _ = oracle.american_lsmc(s, put(), bsm(), 50, workers=16)
    assert got.price == pytest.approx(ref, rel=1e-9)


@pytest.mark.gpu
def test_gpu_sobol_out_of_sample_matches_the_oracle(gpu, oracle):
    s = settings(1 << 14, seed=5, sampling=_abi.SOBOL)
    got = gpu_american(gpu, s, am_terms(), bsm_model(), training=1 << 15)
    ref = oracle.american_out_of_sample(s, 1 << 15, put(), bsm(), 50, workers=16)
    assert got.price == pytest.approx(ref, rel=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32_STEP])
@pytest.mark.parametrize("antithetic", [False, True])
def test_gpu_anchored_paths_match_the_oracle(gpu, oracle, precision, antithetic):
    """Spot and strike at 1e255 take the anchored paths (each leg's spot
    stepped directly, in FP64 whatever the precision asked: such spots leave
    the FP32 range); GPU against the oracle at 1e-9."""
    s = settings(40_001, seed=11, antithetic=antithetic, precision=precision)
    terms = _abi.AmericanTerms(-1, 50, 1e255, us(START), us(END))
    model = _abi.AmericanModel(1e255, 0.05, 0.02, 0.2, us(START), None)
    got = gpu_american(gpu, s, terms, model)
    ref, _ = oracle.american_lsmc(s, put(1e255), bsm(1e255), 50, workers=16)
    assert got.price == pytest.approx(ref, rel=1e-9)
