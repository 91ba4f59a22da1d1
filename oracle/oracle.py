"""Test infrastructure, not product code.

ctypes loader for the FP64 CPU restatement of kiyosi's Monte Carlo engines
(``oracle/_build/liboracle.so``, built by ``make -C oracle``). Only tests/,
``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import it, as the
checker of the HIP engine; the library never does.
"""

import ctypes as C
import pathlib
import subprocess

import numpy as np

from kiyosi_amd import _abi
from kiyosi_amd.errors import raise_for_status

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
TABLE_PATH = HERE.parent / "kiyosi_amd" / "data" / "joe_kuo_6_21201.u32"

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        handle = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        handle.oracle_last_error.restype = C.c_char_p
        handle.oracle_set_sobol_table.argtypes = [P(C.c_uint32), C.c_size_t]
        handle.oracle_pcg32_outputs.argtypes = [C.c_uint64] * 4 + [P(C.c_uint32)]
        handle.oracle_pcg32_uniforms.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, P(C.c_double)]
        handle.oracle_pcg_normals.argtypes = [C.c_uint64] * 4 + [P(C.c_double)]
        handle.oracle_sin_cos_turns.argtypes = [P(C.c_double), C.c_size_t, P(C.c_double), P(C.c_double)]
        handle.oracle_inverse_normal_cdf.argtypes = [P(C.c_double), C.c_size_t, P(C.c_double)]
        handle.oracle_sobol_points.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, P(C.c_uint64)]
        handle.oracle_sobol_raw_directions.argtypes = [C.c_uint32, P(C.c_uint64)]
        handle.oracle_vanilla_price.argtypes = [
            P(_abi.Settings), P(_abi.EuropeanTerms), P(_abi.BsmMarket), C.c_int, P(_abi.Estimate)]
        handle.oracle_autocallable_price.argtypes = [
            P(_abi.Settings), P(_abi.AutocallTerms), P(_abi.BsmMarket), C.c_int, P(_abi.Estimate)]
        handle.oracle_autocallable_price_on_shared_grid.argtypes = [
            P(_abi.Settings), P(_abi.AutocallTerms), P(_abi.BsmMarket), C.c_int, P(_abi.Estimate)]
        handle.oracle_heston_price.argtypes = [
            P(_abi.Settings), P(_abi.HestonTerms), P(_abi.HestonMarket), C.c_int, P(_abi.Estimate)]
        for name, terms in (("barrier", _abi.BarrierTerms), ("asian", _abi.AsianTerms),
                            ("accumulator", _abi.AccumulatorTerms)):
            getattr(handle, f"oracle_{name}_price").argtypes = [
                P(_abi.Settings), P(terms), P(_abi.BsmMarket), C.c_int, P(_abi.Estimate)]
        handle.oracle_american_lsmc.argtypes = [P(_abi.Settings), P(_abi.EuropeanTerms), P(_abi.BsmMarket),
                                                 C.c_int, C.c_int, C.c_int, P(C.c_double), P(C.c_double)]
        handle.oracle_american_coordinates.argtypes = [P(_abi.Settings), C.c_int, C.c_int, C.c_int, P(C.c_double)]
        handle.oracle_american_out_of_sample.argtypes = [P(_abi.Settings), C.c_int64, P(_abi.EuropeanTerms),
                                                         P(_abi.BsmMarket), C.c_int, C.c_int, P(C.c_double)]
        handle.oracle_fd_american.argtypes = [P(_abi.LocalVolSurface), C.c_int64, C.c_double, C.c_double,
                                              C.c_double, C.c_double, C.c_int64, C.c_int, C.c_int, C.c_int,
                                              P(C.c_double)]
        handle.oracle_lv_lsmc.argtypes = [P(_abi.Settings), P(_abi.AmericanTerms), P(_abi.AmericanModel), C.c_int,
                                          C.c_int64, C.c_int, P(C.c_double), P(C.c_double)]
        handle.oracle_lv_european.argtypes = [P(_abi.Settings), P(_abi.AmericanTerms), P(_abi.AmericanModel),
                                              C.c_int, C.c_double, C.c_int, P(_abi.Estimate)]
        handle.oracle_heston_terminal.argtypes = [
            P(_abi.Settings), P(_abi.HestonTerms), P(_abi.HestonMarket), C.c_int, P(_abi.Estimate)]
        handle.oracle_autocallable_plan.argtypes = [
            P(_abi.AutocallTerms), P(_abi.BsmMarket), P(C.c_double), C.c_size_t,
            P(C.c_size_t), P(C.c_int)]
        table = np.fromfile(TABLE_PATH, dtype="<u4")
        handle.oracle_set_sobol_table(table.ctypes.data_as(P(C.c_uint32)), table.size)
        _lib = handle
    return _lib


def _check(status: int) -> None:
    if status:
        raise_for_status(status, lib().oracle_last_error().decode())


def _ptr(array, ctype):
    return array.ctypes.data_as(C.POINTER(ctype))


def pcg32_outputs(seed: int, stream: int, first: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    lib().oracle_pcg32_outputs(seed, stream, first, count, _ptr(out, C.c_uint32))
    return out


def pcg32_uniforms(seed: int, stream: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    lib().oracle_pcg32_uniforms(seed, stream, count, _ptr(out, C.c_double))
    return out


def pcg_normals(seed: int, stream: int, first: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    lib().oracle_pcg_normals(seed, stream, first, count, _ptr(out, C.c_double))
    return out


def sin_cos_turns(turns) -> tuple:
    turns = np.ascontiguousarray(turns, dtype=np.float64)
    s, c = np.empty_like(turns), np.empty_like(turns)
    lib().oracle_sin_cos_turns(_ptr(turns, C.c_double), turns.size, _ptr(s, C.c_double), _ptr(c, C.c_double))
    return s, c


def inverse_normal_cdf(p) -> np.ndarray:
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty_like(p)
    lib().oracle_inverse_normal_cdf(_ptr(p, C.c_double), p.size, _ptr(out, C.c_double))
    return out


def sobol_points(dims: int, seed: int, first: int, count: int) -> np.ndarray:
    out = np.empty((count, dims), dtype=np.uint64)
    _check(lib().oracle_sobol_points(dims, seed, first, count, _ptr(out, C.c_uint64)))
    return out


def sobol_raw_directions(dims: int) -> np.ndarray:
    out = np.empty((dims, 52), dtype=np.uint64)
    _check(lib().oracle_sobol_raw_directions(dims, _ptr(out, C.c_uint64)))
    return out


def vanilla_price(settings, terms, market, workers: int = 1):
    out = _abi.Estimate()
    _check(lib().oracle_vanilla_price(C.byref(settings), C.byref(terms), C.byref(market), workers, C.byref(out)))
    return out


def autocallable_price(settings, terms, market, workers: int = 1):
    out = _abi.Estimate()
    _check(lib().oracle_autocallable_price(C.byref(settings), C.byref(terms), C.byref(market), workers, C.byref(out)))
    return out


def autocallable_price_on_shared_grid(settings, terms, market, workers: int = 1):
    """The price of one note of a price_many_on_shared_grid batch."""
    out = _abi.Estimate()
    _check(lib().oracle_autocallable_price_on_shared_grid(C.byref(settings), C.byref(terms), C.byref(market),
                                                          workers, C.byref(out)))
    return out


def heston_price(settings, terms, market, workers: int = 1):
    out = _abi.Estimate()
    _check(lib().oracle_heston_price(C.byref(settings), C.byref(terms), C.byref(market), workers, C.byref(out)))
    return out


def autocallable_plan(terms, market, cap: int = 100000):
    steps = np.zeros((cap, 5), dtype=np.float64)
    n, bridge = C.c_size_t(), C.c_int()
    _check(lib().oracle_autocallable_plan(C.byref(terms), C.byref(market), _ptr(steps, C.c_double), cap,
                                          C.byref(n), C.byref(bridge)))
    return steps[: n.value].copy(), bool(bridge.value)


def barrier_price(settings, terms, market, workers: int = 1):
    out = _abi.Estimate()
    _check(lib().oracle_barrier_price(C.byref(settings), C.byref(terms), C.byref(market), workers, C.byref(out)))
    return out


def barrier_price_with_greeks(settings, terms, market, request=3, spot_radius=0.0, workers: int = 1):
    """(estimate, kg_spot_greeks) of MonteCarloBarrierEngine::native with delta/gamma."""
    out, greeks = _abi.Estimate(), _abi.SpotGreeks(request, 0, spot_radius, 0.0, 0.0)
    _check(lib().oracle_barrier_price_with_greeks(C.byref(settings), C.byref(terms), C.byref(market), workers,
                                                  C.byref(greeks), C.byref(out)))
    return out, greeks


def double_barrier_price(settings, terms, market, request=0, spot_radius=0.0, workers: int = 1):
    out, greeks = _abi.Estimate(), _abi.SpotGreeks(request, 0, spot_radius, 0.0, 0.0)
    _check(lib().oracle_double_barrier_price(C.byref(settings), C.byref(terms), C.byref(market), workers,
                                             C.byref(greeks), C.byref(out)))
    return out, greeks


def digital_price(settings, terms, market, request=0, workers: int = 1):
    out, greeks = _abi.Estimate(), _abi.SpotGreeks(request, 0, 0.0, 0.0, 0.0)
    _check(lib().oracle_digital_price(C.byref(settings), C.byref(terms), C.byref(market), workers,
                                      C.byref(greeks), C.byref(out)))
    return out, greeks


def double_survival(initial, terminal, band, deviation, derivatives=True, branch=-1):
    """[survival, d/d initial, d2/d initial2] of the double-barrier bridge series."""
    out = (C.c_double * 3)()
    _check(lib().oracle_double_survival(C.c_double(initial), C.c_double(terminal), C.c_double(band),
                                        C.c_double(deviation), int(derivatives), branch, out))
    return list(out)


def single_hit_value(amount, distance, up, drift, sigma, time, rate):
    """E[amount e^{-r tau} 1{tau <= T}] for a barrier at log distance
    `distance` (up or down) from the spot, with log drift `drift`."""
    out = C.c_double()
    _check(lib().oracle_single_hit_value(*[C.c_double(amount), C.c_double(distance)], int(up),
                                         *[C.c_double(v) for v in (drift, sigma, time, rate)], C.byref(out)))
    return out.value


def double_hit_value(amount, lower, upper, spot, rate, dividend, sigma, time):
    """The same for the first exit from (lower, upper)."""
    out = C.c_double()
    _check(lib().oracle_double_hit_value(*[C.c_double(v) for v in (amount, lower, upper, spot, rate, dividend, sigma,
                                                                   time)], C.byref(out)))
    return out.value


def double_exit_probability(lower, upper, spot, rate, dividend, sigma, time):
    """Probability that the spot leaves (lower, upper) before `time`."""
    out = C.c_double()
    _check(lib().oracle_double_exit_probability(*[C.c_double(v) for v in (lower, upper, spot, rate, dividend, sigma,
                                                                          time)], C.byref(out)))
    return out.value


def asian_price(settings, terms, market, workers: int = 1):
    out = _abi.Estimate()
    _check(lib().oracle_asian_price(C.byref(settings), C.byref(terms), C.byref(market), workers, C.byref(out)))
    return out


def asian_price_with_greeks(settings, terms, market, request=3, workers: int = 1):
    """(estimate, kg_spot_greeks) of MonteCarloAsianEngine::native with delta/gamma."""
    out, greeks = _abi.Estimate(), _abi.SpotGreeks(request, 0, 0.0, 0.0, 0.0)
    _check(lib().oracle_asian_price_with_greeks(C.byref(settings), C.byref(terms), C.byref(market), workers,
                                                C.byref(greeks), C.byref(out)))
    return out, greeks


def accumulator_price(settings, terms, market, workers: int = 1):
    out = _abi.Estimate()
    _check(lib().oracle_accumulator_price(C.byref(settings), C.byref(terms), C.byref(market), workers,
                                          C.byref(out)))
    return out


def american_lsmc(settings, terms, market, steps, workers: int = 1, mode: int = -1):
    """In-sample LSMC price and the fitted policy, shape (steps - 1, 4).
    mode -1 picks forward or reverse paths as the reference does; 0 and 1
    force them."""
    price = C.c_double()
    policy = np.zeros((steps - 1, 4))
    _check(lib().oracle_american_lsmc(C.byref(settings), C.byref(terms), C.byref(market), steps, workers, mode,
                                      C.byref(price), _ptr(policy, C.c_double)))
    return price.value, policy


def american_coordinates(settings, steps, reverse, workers: int = 1):
    """Normalised Brownian coordinates of the LSMC paths, shape (steps, samples)."""
    samples = (settings.path_count + 1) // 2 if settings.antithetic else settings.path_count
    out = np.zeros((steps, samples))
    _check(lib().oracle_american_coordinates(C.byref(settings), steps, int(reverse), workers,
                                             _ptr(out, C.c_double)))
    return out


def american_out_of_sample(settings, training_paths, terms, market, steps, workers: int = 1):
    price = C.c_double()
    _check(lib().oracle_american_out_of_sample(C.byref(settings), training_paths, C.byref(terms), C.byref(market),
                                               steps, workers, C.byref(price)))
    return price.value


def fd_american(surface, valuation, spot, strike, rate, dividend, expiry, option_type=-1, nx=1000, nt=1000):
    """Crank-Nicolson American price under a local-volatility surface (kg LocalVolSurface)."""
    price = C.c_double()
    _check(lib().oracle_fd_american(C.byref(surface), valuation, spot, strike, rate, dividend, expiry,
                                    option_type, nx, nt, C.byref(price)))
    return price.value


def lv_lsmc(settings, terms, model, basis, training_paths=0, workers: int = 1):
    """(in-sample, out-of-sample) LSMC prices under local volatility."""
    a, b = C.c_double(), C.c_double()
    _check(lib().oracle_lv_lsmc(C.byref(settings), C.byref(terms), C.byref(model), basis, training_paths,
                                workers, C.byref(a), C.byref(b)))
    return a.value, b.value


def lv_european(settings, terms, model, kind=0, cash=1.0, workers: int = 1):
    """European payoff (kind 0 vanilla, 1 cash digital paying `cash`) under
    local volatility on the LSMC engine's log-Euler stepping."""
    out = _abi.Estimate()
    _check(lib().oracle_lv_european(C.byref(settings), C.byref(terms), C.byref(model), kind, cash, workers,
                                    C.byref(out)))
    return out


def heston_terminal(settings, terms, market, workers: int = 1):
    """Estimates of E[S_T] and E[v_T] on the Heston scheme's paths."""
    out = (_abi.Estimate * 2)()
    _check(lib().oracle_heston_terminal(C.byref(settings), C.byref(terms), C.byref(market), workers, out))
    return out[0], out[1]
