"""Monte Carlo engines with kiyosi's Python interface, running on the GPU.

Each engine mirrors the constructor keywords and pricing methods of the
kiyosi class of the same name (crates/kiyosi-python/src/engines/monte_carlo.rs
:212-342, 476-507): ``price``, ``price_with_seed`` and
``price_with_statistics``. Two keywords are added for the GPU: ``precision``
("fp64" restates the reference arithmetic, "fp32" steps paths in single
precision with FP64 accumulation) and ``device``. ``max_active_workers`` is
validated and otherwise unused.
"""

import ctypes as C
import os

from . import _abi, _native
from .errors import InvalidInputError, UnsupportedOperationError
from .instruments import EuropeanOption, _Autocallable

_SAMPLING = {"pseudo_random": _abi.PSEUDO_RANDOM, "sobol": _abi.SOBOL}
_PRECISION = {"fp64": _abi.FP64, "fp32": _abi.FP32_STEP}


class MonteCarloEstimate:
    """Price, standard error, 95% interval, uncertainty status and seed
    (monte_carlo.rs:859-1001), plus GPU diagnostics."""

    def __init__(self, raw: _abi.Estimate):
        self.price = raw.price
        self.standard_error = raw.standard_error if raw.has_standard_error else None
        self.confidence_interval = (raw.ci_lower, raw.ci_upper) if raw.has_standard_error else None
        self.uncertainty_status = _abi.UNCERTAINTY_NAMES[raw.uncertainty_status]
        self.seed = raw.seed if raw.has_seed else None
        self.path_standard_error = raw.path_standard_error
        self.path_steps = raw.path_steps
        self.device_ms = raw.device_ms

    def __repr__(self):
        return (f"MonteCarloEstimate(price={self.price!r}, standard_error={self.standard_error!r}, "
                f"uncertainty_status={self.uncertainty_status!r}, seed={self.seed!r})")


class _MonteCarloEngine:
    def __init__(self, *, path_count=131_072, seed=None, max_active_workers=0,
                 sampling="pseudo_random", antithetic=False, precision="fp64", device=0,
                 allow_large=False):
        if isinstance(path_count, bool) or not isinstance(path_count, int):
            raise TypeError("path_count must be an integer")
        limit = 1 << 26 if allow_large else 10_000_000
        if not 1 <= path_count <= limit:
            raise InvalidInputError(f"path_count must be between 1 and {limit}")
        if max_active_workers < 0:
            raise InvalidInputError("max_active_workers must be nonnegative")
        if sampling not in _SAMPLING:
            raise InvalidInputError("sampling must be 'pseudo_random' or 'sobol'")
        if sampling == "sobol" and path_count & (path_count - 1):
            raise InvalidInputError("path_count must be a power of two for Sobol sampling")
        if precision not in _PRECISION:
            raise InvalidInputError("precision must be 'fp64' or 'fp32'")
        if seed is not None and not 0 <= seed < 1 << 64:
            raise OverflowError("seed must fit in an unsigned 64-bit integer")
        self.path_count = path_count
        self.seed = seed
        self.max_active_workers = max_active_workers
        self.sampling = sampling
        self.antithetic = antithetic
        self.precision = precision
        self.device = device
        self.allow_large = allow_large

    def _kwargs(self):
        return dict(path_count=self.path_count, seed=self.seed,
                    max_active_workers=self.max_active_workers, sampling=self.sampling,
                    antithetic=self.antithetic, precision=self.precision, device=self.device,
                    allow_large=self.allow_large)

    def with_resolved_seed(self):
        """A replayable copy; draws a seed from system entropy only when none is set
        (monte_carlo.rs:163-177)."""
        kwargs = self._kwargs()
        if kwargs["seed"] is None:
            kwargs["seed"] = int.from_bytes(os.urandom(8), "little")
        return type(self)(**kwargs)

    def _settings(self) -> _abi.Settings:
        return _abi.Settings(self.path_count, self.seed or 0, int(self.seed is not None),
                             _SAMPLING[self.sampling], int(self.antithetic),
                             _PRECISION[self.precision], self.max_active_workers,
                             int(self.allow_large))

    def price_with_statistics(self, instrument, context) -> MonteCarloEstimate:
        engine = self if self.seed is not None else self.with_resolved_seed()
        ctx = _native.context(engine.device)
        out = _abi.Estimate()
        fn, terms, market = engine._entry(instrument, context)
        ctx.call(fn, C.byref(engine._settings()), C.byref(terms), C.byref(market), C.byref(out),
                 interruptible=engine.path_count > _native.LONG_CALL_PATHS)
        return MonteCarloEstimate(out)

    def price_with_seed(self, instrument, context):
        estimate = self.price_with_statistics(instrument, context)
        return estimate.price, estimate.seed

    def _scenarios(self, entry, terms, scenarios):
        """Prices of `scenarios` (pricing contexts) from one shared-draw batch
        entry point, or None when the batch does not apply to them."""
        markets = (_abi.BsmMarket * len(scenarios))(*[ctx.bsm_market() for ctx in scenarios])
        out = (_abi.Estimate * len(scenarios))()
        ctx = _native.context(self.device)
        status = entry(ctx.handle, C.byref(self._settings()), C.byref(terms), markets, len(scenarios), out)
        if status == _abi.UNSUPPORTED_OPERATION:
            return None
        ctx.check(status)
        return [(o.price, o.seed if o.has_seed else None) for o in out]

    def price(self, instrument, context) -> float:
        return self.price_with_statistics(instrument, context).price


class MonteCarloVanillaEngine(_MonteCarloEngine):
    """European vanilla Monte Carlo (monte_carlo.rs:832-1329)."""

    def __init__(self, *, step_count=50, **kwargs):
        if not 1 <= step_count <= 10_000:
            raise InvalidInputError("step_count is outside the supported Monte Carlo range")
        self.step_count = step_count
        super().__init__(**kwargs)

    def _kwargs(self):
        return dict(super()._kwargs(), step_count=self.step_count)

    def _entry(self, instrument, context):
        if not isinstance(instrument, EuropeanOption):
            raise TypeError("MonteCarloVanillaEngine prices EuropeanOption")
        return _native.lib().kg_vanilla_price, instrument._terms(), context.bsm_market()

    def price_with_statistics(self, instrument, context):
        from .instruments import AmericanOption
        if isinstance(instrument, AmericanOption):
            report = self.american_report(instrument, context)
            raw = _abi.Estimate(price=report.price, has_seed=1, seed=self._resolved().seed,
                                uncertainty_status=2)
            return MonteCarloEstimate(raw)
        return super().price_with_statistics(instrument, context)

    def _resolved(self):
        return self if self.seed is not None else self.with_resolved_seed()

    def price_with_greeks(self, option, context, request=None, settings=None):
        """Price and Greeks of a European option by central differences on
        shared draws: the stencil's scenarios are priced in one GPU launch on
        the same normals (price_with_greeks_with_settings,
        engines/monte_carlo.rs:1334-1409). Returns a PricingResult whose
        values equal numerical_greeks(engine, ...) with the same seed."""
        from . import greeks
        from .instruments import AmericanOption
        if isinstance(option, AmericanOption):
            # The exercise-grid estimator in every stencil scenario; the
            # scenarios at one valuation time are lanes on shared paths
            # (greek_prices / greek_scenario_group, american.rs:775-931).
            engine = self._resolved()

            def lanes(scenarios):
                return [(p, engine.seed) for p in engine._american_lanes(option, scenarios)]
            return _batched_greeks(engine, option, context, request, settings, _by_valuation_time, lanes)
        if not isinstance(option, EuropeanOption):
            return greeks.numerical_greeks(self, option, context, request, settings)
        engine = self._resolved()

        def batch(scenarios):
            if len(scenarios) > _abi.MAX_SCENARIOS:
                return None
            return engine._scenarios(_native.lib().kg_vanilla_scenarios, option._terms(), scenarios)
        return _batched_greeks(engine, option, context, request, settings, lambda scenarios: [scenarios], batch)

    def american_report(self, option, context, *, training_path_count=None, basis="quadratic",
                        epochs=4, batch=65536, learning_rate=2e-3):
        """Least-squares Monte Carlo on the exercise grid of step_count intervals:
        in-sample (kiyosi's price for AmericanOption) or, with training_path_count,
        fitted on separate training paths and priced out of sample. basis is
        "quadratic" (kiyosi's), "poly4" or "mlp". Returns the raw report."""
        from .market import american_model
        engine = self._resolved()
        ctx = _native.context(engine.device)
        bases = {"quadratic": _abi.BASIS_QUADRATIC, "poly4": _abi.BASIS_POLY4, "mlp": _abi.BASIS_MLP}
        options = _abi.LsmcOptions(bases[basis], int(training_path_count is not None),
                                   int(training_path_count or 0), epochs, batch, learning_rate)
        report = _abi.LsmcReport()
        model = american_model(context)
        ctx.call(_native.lib().kg_american_price, C.byref(engine._settings()),
                 C.byref(option._american_terms(engine.step_count)), C.byref(model), C.byref(options),
                 C.byref(report), interruptible=True)
        return report

    def _american_lanes(self, option, contexts, strikes=None):
        """In-sample American prices of `option` in each context (at
        strikes[i] when given) through kg_american_price_lanes: lanes at one
        valuation time share one set of paths, and each price equals
        ``price`` of that lane alone."""
        from .market import american_model
        engine = self._resolved()
        ctx = _native.context(engine.device)
        n = len(contexts)
        kept = [american_model(c) for c in contexts]  # the structs keep their surfaces alive
        models = (_abi.AmericanModel * n)(*kept)
        ks = (C.c_double * n)(*strikes) if strikes is not None else None
        reports = (_abi.LsmcReport * n)()
        options = _abi.LsmcOptions(_abi.BASIS_QUADRATIC, 0, 0, 4, 65536, 2e-3)
        ctx.call(_native.lib().kg_american_price_lanes, C.byref(engine._settings()),
                 C.byref(option._american_terms(engine.step_count)), models, ks, n, C.byref(options), reports,
                 interruptible=True)
        return [r.price for r in reports]

    def price_strikes(self, instrument, context, *, strikes):
        """American prices at each strike on one resolved seed, with a separate
        exercise rule fitted per distinct strike (MonteCarloVanillaEngine::
        price_strikes, american.rs:677-771): each equals ``price`` of that
        strike's option. The distinct strikes are lanes on one set of paths,
        up to MAX_AMERICAN_LANES per launch."""
        import math
        from .instruments import AmericanOption
        if not isinstance(instrument, AmericanOption):
            raise TypeError("price_strikes takes an AmericanOption")
        values = [float(k) for k in strikes]
        for k in values:
            if not (math.isfinite(k) and k > 0.0):
                raise InvalidInputError("strikes must be positive and finite")
        engine = self._resolved()
        distinct = sorted(set(values))
        quotes = {}
        for i in range(0, len(distinct), _abi.MAX_AMERICAN_LANES):
            chunk = distinct[i:i + _abi.MAX_AMERICAN_LANES]
            quotes.update(zip(chunk, engine._american_lanes(instrument, [context] * len(chunk), chunk)))
        return tuple(quotes[k] for k in values)

    def price_american_out_of_sample(self, option, context, *, training_path_count, basis="quadratic"):
        """Fits the exercise rule on training paths (seed ^ 0x9e3779b97f4a7c15) and
        prices it on this engine's paths (american.rs:574-637)."""
        return self.american_report(option, context, training_path_count=training_path_count,
                                    basis=basis).price


class MonteCarloAutocallableEngine(_MonteCarloEngine):
    """Snowball and Phoenix Monte Carlo (engines/monte_carlo/autocallable.rs)."""

    def _entry(self, instrument, context):
        if not isinstance(instrument, _Autocallable):
            raise TypeError("MonteCarloAutocallableEngine prices autocallable notes")
        return _native.lib().kg_autocallable_price, instrument._terms(), context.bsm_market()

    def _resolved(self):
        return self if self.seed is not None else self.with_resolved_seed()

    def price_many(self, instruments, context, shared_grid=False):
        """Prices of notes of one product type on one resolved seed
        (price_many / price_many_on_shared_grid, crates/kiyosi-python/src/
        engines/monte_carlo.rs:699-746). Without `shared_grid` each equals
        ``price`` of that note with the same seed, bit for bit. With it, a
        note with daily knock-in monitoring whose events fall on its daily
        observation times is sampled on its shared daily grid step by step,
        without the Brownian bridge (autocallable/portfolio.rs:75-186); its
        price is then that of the batch's draws, as the reference's shared
        grid estimators give it. Notes sharing a step grid run as lanes of
        one GPU path set."""
        notes = list(instruments)
        if len(notes) > 1 << 20:
            raise InvalidInputError("a batch holds at most 1,048,576 notes")
        if any(not isinstance(n, _Autocallable) for n in notes):
            raise TypeError("price_many prices autocallable notes")
        if len({type(n) for n in notes}) > 1:
            raise UnsupportedOperationError("Monte Carlo price_many requires one autocallable product type")
        if not notes:
            return ()
        engine = self._resolved()
        terms = [n._terms() for n in notes]
        array = (_abi.AutocallTerms * len(notes))(*terms)
        out = (_abi.Estimate * len(notes))()
        ctx = _native.context(engine.device)
        entry = (_native.lib().kg_autocallable_price_many_on_shared_grid if shared_grid
                 else _native.lib().kg_autocallable_price_many)
        ctx.check(entry(ctx.handle, C.byref(engine._settings()), array, len(notes), C.byref(context.bsm_market()),
                        out))
        return tuple(o.price for o in out)

    def price_with_greeks(self, note, context, request=None, settings=None):
        """Price and Greeks of an autocallable note by central differences on
        shared draws (price_with_greeks_with_settings,
        engines/monte_carlo/autocallable.rs:919-999): the stencil's scenarios
        at each valuation time are priced as lanes of one path set (spot_prices
        :174-197, market_prices :199-241). Each price equals pricing its
        scenario alone with the same seed, so the result equals
        numerical_greeks(engine, ...)."""
        from . import greeks
        if context.model_parameters.volatility == 0.0:
            raise UnsupportedOperationError(
                "zero volatility selects the deterministic engine, which is not part of the GPU engine")
        engine = self._resolved()
        terms = note._terms()

        def groups(scenarios):
            by_time = {}
            for scn in scenarios:
                by_time.setdefault(greeks.key(scn)[4], []).append(scn)
            out = []
            for same in by_time.values():
                out += [same[i:i + _abi.MAX_AUTOCALL_SCENARIOS]
                        for i in range(0, len(same), _abi.MAX_AUTOCALL_SCENARIOS)]
            return out
        return _batched_greeks(engine, note, context, request, settings, groups,
                               lambda scenarios: engine._scenarios(_native.lib().kg_autocallable_scenarios,
                                                                   terms, scenarios))


class _SpotGreekEngine(_MonteCarloEngine):
    """Engines whose Greeks are the likelihood-ratio delta and gamma computed
    on the draws of the price (engines.rs:138-162)."""

    def _greek_entry(self, instrument):
        raise NotImplementedError

    def price_with_greeks(self, option, context, request=None, settings=None):
        """Price with native delta and gamma from one GPU launch; other Greeks
        are unsupported, and at expiry the terminal-payoff stencil supplies
        delta and gamma. The price equals ``price`` with the same seed."""
        from . import greeks
        settings = settings or greeks.NumericalShiftSettings()
        requested = greeks.normalize_request(request)
        engine = self._resolved()
        fn, terms = engine._greek_entry(option)
        spot = context.spot_price
        radius = greeks.symmetric_shift(spot, settings.spot_shift_at(spot))
        bits = (_abi.GREEK_DELTA if "delta" in requested else 0) | (_abi.GREEK_GAMMA if "gamma" in requested else 0)
        spot_greeks = _abi.SpotGreeks(bits, 0, radius, 0.0, 0.0)
        out = _abi.Estimate()
        ctx = _native.context(engine.device)
        ctx.check(fn(ctx.handle, C.byref(engine._settings()), C.byref(terms), C.byref(context.bsm_market()),
                     C.byref(spot_greeks), C.byref(out)))
        values = {}
        if spot_greeks.available & _abi.GREEK_DELTA:
            values["delta"] = spot_greeks.delta
        if spot_greeks.available & _abi.GREEK_GAMMA:
            values["gamma"] = spot_greeks.gamma
        return greeks.native_spot_result(option, context, requested, settings, out.price, values,
                                         out.seed if out.has_seed else None,
                                         lambda scn: engine.price(option, scn))

    def _resolved(self):
        return self if self.seed is not None else self.with_resolved_seed()


class MonteCarloBarrierEngine(_SpotGreekEngine):
    """Single- and double-barrier Monte Carlo (engines/monte_carlo/barrier.rs)."""

    def _entry(self, instrument, context):
        from .instruments import BarrierOption, DoubleBarrierOption
        if isinstance(instrument, DoubleBarrierOption):
            return _native.lib().kg_double_barrier_price, instrument._terms(), context.bsm_market()
        if not isinstance(instrument, BarrierOption):
            raise TypeError("MonteCarloBarrierEngine prices BarrierOption and DoubleBarrierOption")
        return _native.lib().kg_barrier_price, instrument._terms(), context.bsm_market()

    def _greek_entry(self, instrument):
        from .instruments import BarrierOption, DoubleBarrierOption
        if isinstance(instrument, DoubleBarrierOption):
            return _native.lib().kg_double_barrier_price_with_greeks, instrument._terms()
        if not isinstance(instrument, BarrierOption):
            raise TypeError("MonteCarloBarrierEngine prices BarrierOption and DoubleBarrierOption")
        return _native.lib().kg_barrier_price_with_greeks, instrument._terms()


class MonteCarloDigitalEngine(_SpotGreekEngine):
    """Cash-or-nothing and asset-or-nothing Monte Carlo
    (engines/monte_carlo/digital.rs)."""

    def _entry(self, instrument, context):
        from .instruments import DigitalOption
        if not isinstance(instrument, DigitalOption):
            raise TypeError("MonteCarloDigitalEngine prices DigitalOption")
        return _native.lib().kg_digital_price, instrument._terms(), context.bsm_market()

    def _greek_entry(self, instrument):
        from .instruments import DigitalOption
        if not isinstance(instrument, DigitalOption):
            raise TypeError("MonteCarloDigitalEngine prices DigitalOption")
        return _native.lib().kg_digital_price_with_greeks, instrument._terms()


class MonteCarloAsianEngine(_SpotGreekEngine):
    """Average-price Monte Carlo (engines/monte_carlo/asian.rs). Greeks are
    the engine's native delta and gamma (asian.rs:885-898): the conditional
    estimators of the geometric and arithmetic payoffs on the price's
    draws."""

    def __init__(self, *, step_count=50, **kwargs):
        if not 2 <= step_count <= 10_000:
            raise InvalidInputError("step_count is outside the supported Monte Carlo range")
        self.step_count = step_count
        super().__init__(**kwargs)

    def _kwargs(self):
        return dict(super()._kwargs(), step_count=self.step_count)

    def _entry(self, instrument, context):
        from .instruments import _AveragePrice
        if not isinstance(instrument, _AveragePrice):
            raise TypeError("MonteCarloAsianEngine prices average-price options")
        return _native.lib().kg_asian_price, instrument._terms(self.step_count), context.bsm_market()

    def _greek_entry(self, instrument):
        from .instruments import _AveragePrice
        if not isinstance(instrument, _AveragePrice):
            raise TypeError("MonteCarloAsianEngine prices average-price options")
        return _native.lib().kg_asian_price_with_greeks, instrument._terms(self.step_count)

    def price_strikes(self, instrument, context, *, strikes):
        """Prices of the arithmetic-average option at each strike on one
        resolved seed (MonteCarloAsianEngine::price_strikes, asian.rs:197-345):
        each equals ``price`` of that strike's option bit for bit; distinct
        strikes share the GPU paths as lanes."""
        from .instruments import ArithmeticAveragePriceOption
        if not isinstance(instrument, ArithmeticAveragePriceOption):
            raise TypeError("price_strikes takes an ArithmeticAveragePriceOption")
        values = [float(k) for k in strikes]
        if not values:
            return ()
        engine = self._resolved()
        array = (C.c_double * len(values))(*values)
        out = (C.c_double * len(values))()
        ctx = _native.context(engine.device)
        ctx.check(_native.lib().kg_asian_price_strikes(ctx.handle, C.byref(engine._settings()),
                                                       C.byref(instrument._terms(engine.step_count)),
                                                       C.byref(context.bsm_market()), array, len(values), out))
        return tuple(out)

    def _resolved(self):
        return self if self.seed is not None else self.with_resolved_seed()


class MonteCarloAccumulatorEngine(_MonteCarloEngine):
    """Accumulator Monte Carlo (engines/monte_carlo/accumulator.rs)."""

    def _entry(self, instrument, context):
        from .instruments import Accumulator
        if not isinstance(instrument, Accumulator):
            raise TypeError("MonteCarloAccumulatorEngine prices Accumulator")
        return _native.lib().kg_accumulator_price, instrument._terms(), context.bsm_market()


_BOOK_KINDS = {
    "MonteCarloBarrierEngine": _abi.BOOK_BARRIER,
    "MonteCarloDigitalEngine": _abi.BOOK_DIGITAL,
    "MonteCarloAsianEngine": _abi.BOOK_ASIAN,
    "MonteCarloAutocallableEngine": _abi.BOOK_AUTOCALL,
    "MonteCarloAccumulatorEngine": _abi.BOOK_ACCUMULATOR,
}


def _by_valuation_time(scenarios):
    """Stencil scenarios grouped by valuation time, at most
    MAX_AMERICAN_LANES per group."""
    from .market import to_micros
    groups = {}
    for scn in scenarios:
        groups.setdefault(to_micros(scn.valuation_time), []).append(scn)
    return [g[i:i + _abi.MAX_AMERICAN_LANES] for g in groups.values()
            for i in range(0, len(g), _abi.MAX_AMERICAN_LANES)]


def _batched_greeks(engine, option, context, request, settings, groups, batch):
    """Plans the stencil, prices its scenarios group by group through
    `batch` (falling back to one call per scenario where a batch does not
    apply) and completes the Greeks from those prices."""
    from . import greeks
    settings = settings or greeks.NumericalShiftSettings()
    scenarios = greeks.plan(option, context, request, settings)
    prices, seed = {}, None
    for group in groups(scenarios):
        results = batch(group)
        if results is None:
            results = [engine.price_with_seed(option, scn) for scn in group]
        for scn, (price, used) in zip(group, results):
            prices[greeks.key(scn)] = price
            seed = used if used is not None else seed
    p0 = prices[greeks.key(context)]
    values, unavailable = greeks.complete(option, context, request, settings, p0,
                                          lambda scn: prices[greeks.key(scn)])
    return greeks.PricingResult(p0, greeks.normalize_request(request), values, unavailable, seed)


class Book:
    """A heterogeneous set of (engine, instrument, context) entries priced in one
    persistent GPU launch; each entry's result equals its engine's own price."""

    def __init__(self, entries):
        self.entries = list(entries)
        n = len(self.entries)
        self._array = (_abi.BookContract * max(n, 1))()
        self._keep = []
        for i, (engine, instrument, context) in enumerate(self.entries):
            kind = _BOOK_KINDS.get(type(engine).__name__)
            if kind is None:
                raise TypeError(f"{type(engine).__name__} cannot join a book")
            engine = engine if engine.seed is not None else engine.with_resolved_seed()
            if kind == _abi.BOOK_BARRIER and type(instrument).__name__ == "DoubleBarrierOption":
                kind = _abi.BOOK_DOUBLE_BARRIER
            if kind == _abi.BOOK_ASIAN:
                terms = instrument._terms(engine.step_count)
            else:
                terms = instrument._terms()
            self._keep.append(terms)
            self._array[i] = _abi.BookContract(kind, 0, engine._settings(), context.bsm_market(),
                                               C.cast(C.pointer(terms), C.c_void_p))

    def price(self, precision="fp64", device=0, samples_per_thread=0, workgroups_per_cu=0,
              use_graph=True, replays=0):
        """Prices every entry; returns (estimates, tuning) where tuning reports the
        tile shape used, the mean device time of one launch and the path-steps."""
        ctx = _native.context(device)
        n = len(self.entries)
        out = (_abi.Estimate * max(n, 1))()
        tuning = _abi.BookTuning(samples_per_thread, workgroups_per_cu, int(use_graph), replays, 0.0, 0.0)
        ctx.call(_native.lib().kg_price_book, self._array, n, _PRECISION[precision], C.byref(tuning), out,
                 interruptible=True)
        return [MonteCarloEstimate(out[i]) for i in range(n)], tuning
