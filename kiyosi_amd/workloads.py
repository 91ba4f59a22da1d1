"""Synthetic benchmark contracts at their full sizes.

Each function returns (instrument, context, engine keyword arguments) for one
workload; bench.py and the tests use them.
"""

import datetime as dt

from .instruments import EuropeanOption, KnockInObservation, SnowballObservation, SnowballOption
from .market import BlackScholesMertonParameters, PricingContext, TradingCalendar

VALUATION = dt.datetime(2025, 1, 1)


def european_call_1m_pcg_antithetic():
    """European call under BSM: S0=K=100, r=5%, q=2%, sigma=20%, T=1y exactly;
    1,000,000 PCG paths in antithetic pairs, seed 1729."""
    option = EuropeanOption(option_type="call", strike=100.0, effective_time=VALUATION,
                            expiry_time=dt.datetime(2026, 1, 1))
    context = PricingContext(
        model_parameters=BlackScholesMertonParameters(risk_free_rate=0.05, dividend_yield=0.02,
                                                      volatility=0.2),
        spot_price=100.0, valuation_time=VALUATION)
    return option, context, dict(path_count=1_000_000, seed=1729, antithetic=True)


def snowball_note(trading_days=252, months=12, step_down=0.5, knock_in=75.0, coupon=0.15,
                  valuation=VALUATION, calendar=None):
    """Snowball on a weekdays calendar with daily knock-in monitoring at 00:00:
    expiry on the `trading_days`-th trading day after valuation, knock-out
    observations every trading_days/months trading days at 100% of S0
    stepping down by `step_down` per observation, coupon 15% p.a., knock-in
    at 75%, upper strike 100, lower strike 0, unit principal."""
    calendar = calendar or TradingCalendar.weekdays()
    spacing = trading_days // months
    observations, date = [], valuation
    for k in range(1, months + 1):
        date = calendar.add_trading_days(date, spacing)
        observations.append(SnowballObservation(date, 100.0 - step_down * (k - 1), coupon))
    expiry = calendar.add_trading_days(valuation, trading_days)
    return SnowballOption(
        maturity_coupon_rate=coupon, initial_spot=100.0, knock_in_level=knock_in,
        upper_strike=100.0, lower_strike=0.0,
        knock_in_observation=KnockInObservation.every_trading_day(dt.time(0, 0)),
        effective_time=valuation, expiry_time=expiry, autocall_observations=observations,
        calendar=calendar)


def snowball_252d_sobol_2p20():
    """Snowball under BSM, 252 daily steps, 12 monthly knock-out observations
    stepping down 0.5% from 100%, daily knock-in at 75%, coupon 15% p.a.;
    sigma=25%, r=2%, q=0; 2^20 Sobol paths, seed 1729."""
    note = snowball_note()
    context = PricingContext(
        model_parameters=BlackScholesMertonParameters(risk_free_rate=0.02, dividend_yield=0.0,
                                                      volatility=0.25),
        spot_price=100.0, valuation_time=VALUATION)
    return note, context, dict(path_count=1 << 20, seed=1729, sampling="sobol")


def heston_asian_252_pcg_2p22(precision="fp64"):
    """Arithmetic-average Asian call under Heston: kappa=2, theta=0.04, xi=0.5,
    rho=-0.7, v0=0.04, S0=K=100, r=3%, q=0, T=1y; 252 equal steps, the average
    of the 252 post-step spots; 2^22 PCG paths, seed 1729 (beyond the
    reference's one-billion path-step cap, hence allow_large). Returns the C
    structs (settings, terms, market)."""
    from . import _abi
    from .market import to_micros
    settings = _abi.Settings(1 << 22, 1729, 1, _abi.PSEUDO_RANDOM, 0,
                             _abi.FP32_STEP if precision == "fp32" else _abi.FP64, 0, 1)
    terms = _abi.HestonTerms(_abi.HESTON_ASIAN_ARITHMETIC, 1, 252, 0, 100.0, to_micros(VALUATION),
                             to_micros(dt.datetime(2026, 1, 1)))
    market = _abi.HestonMarket(100.0, 0.03, 0.0, 0.04, 2.0, 0.04, 0.5, -0.7, to_micros(VALUATION))
    return settings, terms, market


BOOK_SEED = 20251203
BOOK_KINDS = ("up_and_out", "down_and_in", "geometric_asian", "arithmetic_asian", "snowball",
              "phoenix", "accumulator")


def book_4096(n=4096, seed=BOOK_SEED, path_scale=1.0):
    """Heterogeneous book drawn from `seed`: kinds uniform over up-and-out and
    down-and-in barriers (daily fixings), geometric and arithmetic Asians
    (daily grid), snowballs, Phoenix notes and accumulators; maturity
    U[0.25, 3] years as trading days (x252); K/S0 ~ U[0.8, 1.2]; barriers
    U[0.6, 0.9] (down) or U[1.05, 1.3] (up) of S0; sigma ~ U[0.15, 0.45],
    r ~ U[0, 0.05], q ~ U[0, 0.03]; paths 256 * round(2^U[14, 18] / 256),
    scaled by `path_scale`; contract i is seeded seed + i. Returns a list of
    (engine, instrument, context)."""
    import numpy as np

    from .engines import (MonteCarloAccumulatorEngine, MonteCarloAsianEngine,
                          MonteCarloAutocallableEngine, MonteCarloBarrierEngine)
    from .instruments import (Accumulator, ArithmeticAveragePriceOption, BarrierOption,
                              CouponObservation, GeometricAveragePriceOption, KnockOutObservation,
                              PhoenixOption)

    rng = np.random.default_rng(seed)
    calendar = TradingCalendar.weekdays()
    trading = [VALUATION]
    while len(trading) <= 3 * 252 + 1:
        trading.append(calendar.add_trading_days(trading[-1], 1))
    entries = []
    for i in range(n):
        kind = BOOK_KINDS[int(rng.integers(len(BOOK_KINDS)))]
        days = int(round(rng.uniform(0.25, 3.0) * 252))
        expiry = trading[days]
        moneyness = rng.uniform(0.8, 1.2)
        sigma, r, q = rng.uniform(0.15, 0.45), rng.uniform(0.0, 0.05), rng.uniform(0.0, 0.03)
        paths = max(256, int(256 * round(2 ** rng.uniform(14, 18) / 256) * path_scale) // 256 * 256)
        down, up = rng.uniform(0.6, 0.9) * 100, rng.uniform(1.05, 1.3) * 100
        option_type = "call" if rng.uniform() < 0.5 else "put"
        context = PricingContext(model_parameters=BlackScholesMertonParameters(
            risk_free_rate=r, dividend_yield=q, volatility=sigma), spot_price=100.0, valuation_time=VALUATION)
        kw = dict(path_count=paths, seed=seed + i)
        if kind in ("up_and_out", "down_and_in"):
            instrument = BarrierOption(option_type=option_type, strike=100 * moneyness, effective_time=VALUATION,
                                       expiry_time=expiry, barrier_level=up if kind == "up_and_out" else down,
                                       barrier_type=kind, observation_schedule=trading[1:days + 1])
            engine = MonteCarloBarrierEngine(**kw)
        elif kind.endswith("asian"):
            cls = GeometricAveragePriceOption if kind.startswith("geo") else ArithmeticAveragePriceOption
            instrument = cls(option_type=option_type, strike=100 * moneyness, effective_time=VALUATION,
                             averaging_start_time=VALUATION, expiry_time=expiry)
            engine = MonteCarloAsianEngine(step_count=max(2, days), **kw)
        elif kind == "snowball":
            months = max(1, days // 21)
            instrument = snowball_note(trading_days=21 * months, months=months, knock_in=down)
            engine = MonteCarloAutocallableEngine(**kw)
        elif kind == "phoenix":
            months = max(1, days // 21)
            dates = [trading[21 * k] for k in range(1, months + 1)]
            instrument = PhoenixOption(
                coupon_per_period=0.01, initial_spot=100.0, knock_in_level=down, upper_strike=100.0,
                lower_strike=0.0, knock_in_observation=KnockInObservation.every_trading_day(dt.time(0, 0)),
                effective_time=VALUATION, expiry_time=dates[-1],
                knock_out_observations=[KnockOutObservation(d, 100.0) for d in dates],
                coupon_observations=[CouponObservation(d, 80.0) for d in dates], calendar=calendar)
            engine = MonteCarloAutocallableEngine(**kw)
        else:
            instrument = Accumulator(strike=100 * moneyness, knock_out_level=up, daily_quantity=1.0,
                                     acceleration_factor=2.0, effective_time=VALUATION, expiry_time=expiry)
            engine = MonteCarloAccumulatorEngine(**kw)
        entries.append((engine, instrument, context))
    return entries


def config5_surface(times=51, spots=601):
    """sigma(S, t) = 0.2 + 0.1 exp(-(ln(S/100))^2 / 0.08) (1 + 0.5 t), tabulated
    for bilinear interpolation on `times` rows over one year from valuation and
    `spots` log-spaced columns on S in [100 e^-3, 100 e^3]."""
    import numpy as np

    from .market import LocalVolatilitySurface
    year = 365 * 86_400_000_000
    t = np.arange(times) / (times - 1)
    stamps = [VALUATION + dt.timedelta(microseconds=int(round(year * f))) for f in t]
    s = 100.0 * np.exp(np.linspace(-3.0, 3.0, spots))
    values = 0.2 + 0.1 * np.exp(-np.log(s / 100.0) ** 2 / 0.08)[None, :] * (1.0 + 0.5 * t)[:, None]
    return LocalVolatilitySurface(stamps, s, values)


def american_put_local_vol(paths=1 << 20, training=1 << 20, seed=1729):
    """American put under the local-volatility surface above: S0 = K = 100,
    r = 5%, q = 0, T = 1y, 50 exercise dates; `training` paths fit the
    exercise rule and `paths` independent paths price it."""
    from .engines import MonteCarloVanillaEngine
    from .instruments import AmericanOption
    from .market import LocalVolatilityParameters
    option = AmericanOption(option_type="put", strike=100.0, effective_time=VALUATION,
                            expiry_time=dt.datetime(2026, 1, 1))
    context = PricingContext(model_parameters=LocalVolatilityParameters(
        risk_free_rate=0.05, dividend_yield=0.0, surface=config5_surface()), spot_price=100.0,
        valuation_time=VALUATION)
    engine = MonteCarloVanillaEngine(path_count=paths, step_count=50, seed=seed, allow_large=paths > 10_000_000)
    return option, context, engine, training
