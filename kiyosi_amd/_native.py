"""Loads libkiyosi_amd.so (the HIP engine) and declares its C ABI.

The library is built in-tree by ``python -m kiyosi_amd.build``. There is no
CPU fallback: pricing fails loudly when the library or a GPU is missing.
"""

import ctypes as C
import pathlib
import threading

from . import _abi
from .errors import DeviceError, raise_for_status

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libkiyosi_amd.so"

_lib = None
_lock = threading.Lock()
_contexts = {}

# Every symbol declared in include/kiyosi_amd.h, with (restype, argtypes).
P = C.POINTER
SIGNATURES = {
    "kg_version": (C.c_char_p, []),
    "kg_cancel_token_create": (C.c_int, [P(C.c_void_p)]),
    "kg_cancel_token_destroy": (None, [C.c_void_p]),
    "kg_cancel": (None, [C.c_void_p]),
    "kg_cancel_reset": (None, [C.c_void_p]),
    "kg_cancel_requested": (C.c_int32, [C.c_void_p]),
    "kg_context_set_cancel_token": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_context_create": (C.c_int, [C.c_int32, P(C.c_void_p)]),
    "kg_context_destroy": (None, [C.c_void_p]),
    "kg_last_error": (C.c_char_p, [C.c_void_p]),
    "kg_pcg32_outputs": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, P(C.c_uint32)]),
    "kg_pcg_normals": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, P(C.c_double)]),
    "kg_fp32_normals_from_bits": (C.c_int, [C.c_void_p, P(C.c_uint32), C.c_uint64, P(C.c_float)]),
    "kg_sobol_points": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, P(C.c_uint64)]),
    "kg_vanilla_price": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.EuropeanTerms), P(_abi.BsmMarket),
                                   P(_abi.Estimate)]),
    "kg_vanilla_price_noalloc": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.EuropeanTerms),
                                           P(_abi.BsmMarket), P(_abi.Estimate)]),
    "kg_vanilla_scenarios": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.EuropeanTerms), P(_abi.BsmMarket),
                                       C.c_int32, P(_abi.Estimate)]),
    "kg_autocallable_price": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.AutocallTerms), P(_abi.BsmMarket),
                                        P(_abi.Estimate)]),
    "kg_autocallable_scenarios": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.AutocallTerms),
                                            P(_abi.BsmMarket), C.c_int32, P(_abi.Estimate)]),
    "kg_autocallable_price_many": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.AutocallTerms), C.c_size_t,
                                             P(_abi.BsmMarket), P(_abi.Estimate)]),
    "kg_autocallable_price_many_on_shared_grid": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.AutocallTerms),
                                                            C.c_size_t, P(_abi.BsmMarket), P(_abi.Estimate)]),
    "kg_heston_price": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.HestonTerms), P(_abi.HestonMarket),
                                  P(_abi.Estimate)]),
    "kg_barrier_price": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.BarrierTerms), P(_abi.BsmMarket),
                                   P(_abi.Estimate)]),
    "kg_barrier_price_with_greeks": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.BarrierTerms),
                                               P(_abi.BsmMarket), P(_abi.SpotGreeks), P(_abi.Estimate)]),
    "kg_double_barrier_price": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.DoubleBarrierTerms),
                                          P(_abi.BsmMarket), P(_abi.Estimate)]),
    "kg_double_barrier_price_with_greeks": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.DoubleBarrierTerms),
                                                      P(_abi.BsmMarket), P(_abi.SpotGreeks), P(_abi.Estimate)]),
    "kg_digital_price": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.DigitalTerms), P(_abi.BsmMarket),
                                   P(_abi.Estimate)]),
    "kg_digital_price_with_greeks": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.DigitalTerms),
                                               P(_abi.BsmMarket), P(_abi.SpotGreeks), P(_abi.Estimate)]),
    "kg_asian_price_with_greeks": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.AsianTerms),
                                             P(_abi.BsmMarket), P(_abi.SpotGreeks), P(_abi.Estimate)]),
    "kg_asian_price": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.AsianTerms), P(_abi.BsmMarket),
                                 P(_abi.Estimate)]),
    "kg_asian_price_strikes": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.AsianTerms), P(_abi.BsmMarket),
                                         P(C.c_double), C.c_size_t, P(C.c_double)]),
    "kg_accumulator_price": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.AccumulatorTerms), P(_abi.BsmMarket),
                                       P(_abi.Estimate)]),
    "kg_american_price": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.AmericanTerms), P(_abi.AmericanModel),
                                    P(_abi.LsmcOptions), P(_abi.LsmcReport)]),
    "kg_american_price_lanes": (C.c_int, [C.c_void_p, P(_abi.Settings), P(_abi.AmericanTerms),
                                          P(_abi.AmericanModel), P(C.c_double), C.c_int32, P(_abi.LsmcOptions),
                                          P(_abi.LsmcReport)]),
    "kg_price_book": (C.c_int, [C.c_void_p, P(_abi.BookContract), C.c_size_t, C.c_int32, P(_abi.BookTuning),
                                P(_abi.Estimate)]),
    "kg_mlp_gradient": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int64, P(C.c_float),
                                  P(C.c_double)]),
    "kg_mlp_train": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int64, C.c_int32, C.c_int32,
                               C.c_double]),
    "kg_mlp_forward": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int64, P(C.c_float)]),
}


def lib():
    """Returns the loaded library, raising if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise DeviceError(
                    f"{LIB_PATH} is missing: build it with `python -m kiyosi_amd.build`")
            handle = C.CDLL(str(LIB_PATH))
            for name, (restype, argtypes) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = handle
        return _lib


# Calls expected to run longer than about one signal-polling period (10 ms
# of device time) run on a worker thread; see Context.call.
LONG_CALL_PATHS = 1 << 22
SIGNAL_POLL_S = 0.01


def run_interruptible(run, cancel, poll=SIGNAL_POLL_S):
    """Runs `run()` on a worker thread while the calling (main) thread waits in
    steps of `poll` seconds, so that Python handles signals between steps; on
    KeyboardInterrupt it calls `cancel()`, waits for the worker to end and
    re-raises, as kiyosi's Python binding does for every request
    (crates/kiyosi-python/src/cancellation.rs:5-47)."""
    done = threading.Event()
    result = {}

    def worker():
        try:
            result["value"] = run()
        except BaseException as error:  # re-raised on the caller's thread
            result["error"] = error
        finally:
            done.set()

    thread = threading.Thread(target=worker, name="kiyosi-request", daemon=True)
    thread.start()
    try:
        while not done.wait(poll):
            pass
    except KeyboardInterrupt:
        cancel()
        done.wait()
        thread.join()
        raise
    thread.join()
    if "error" in result:
        raise result["error"]
    return result["value"]


class Context:
    """One HIP device context (stream, workspace, Sobol cache) with its own
    cancellation token."""

    def __init__(self, device: int = 0):
        handle = C.c_void_p()
        status = lib().kg_context_create(device, C.byref(handle))
        if status:
            raise_for_status(status, f"cannot create a context on HIP device {device}")
        self.handle = handle
        self.device = device
        token = C.c_void_p()
        status = lib().kg_cancel_token_create(C.byref(token))
        if status:
            raise_for_status(status, "cannot create a cancellation token")
        self.token = token
        lib().kg_context_set_cancel_token(self.handle, self.token)

    def check(self, status: int) -> None:
        if status:
            raise_for_status(status, lib().kg_last_error(self.handle).decode())

    def cancel(self) -> None:
        """Raises the token: the running call (and later ones until the next
        call starts) return InterruptedError."""
        lib().kg_cancel(self.token)

    def call(self, fn, *args, interruptible=False):
        """Runs fn(handle, *args) for one request and raises on a failed status.
        The token is lowered first (each request starts uncancelled). Long
        requests (`interruptible`) run on a worker thread while the main
        thread polls signals every 10 ms, and a KeyboardInterrupt cancels
        them; shorter ones run on the calling thread, where a signal is
        handled as soon as they return."""
        lib().kg_cancel_reset(self.token)
        if interruptible and threading.current_thread() is threading.main_thread():
            status = run_interruptible(lambda: fn(self.handle, *args), self.cancel)
        else:
            status = fn(self.handle, *args)
        self.check(status)

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.kg_context_destroy(self.handle)
            self.handle = None
        if getattr(self, "token", None) and _lib is not None:
            _lib.kg_cancel_token_destroy(self.token)
            self.token = None


def context(device: int = 0) -> Context:
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            ctx = _contexts.setdefault(device, ctx)
    return ctx
