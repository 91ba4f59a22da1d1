"""Cooperative cancellation (crates/kiyosi/src/cancellation.rs:33-68 and the
Python binding's signal polling, crates/kiyosi-python/src/cancellation.rs
:5-47): the C tokens and their binding to a context, the Python worker-thread
runner on the CPU, and on the GPU calls stopped by a raised token with every
kernel ended normally and the context usable afterwards."""

import ctypes as C
import signal
import threading
import time

import pytest

from conftest import END, START, us
from kiyosi_amd import _abi, _native


def test_runner_returns_results_and_errors_on_the_calling_thread():
    assert _native.run_interruptible(lambda: 41 + 1, lambda: None) == 42
    with pytest.raises(ZeroDivisionError):
        _native.run_interruptible(lambda: 1 / 0, lambda: None)


def test_runner_cancels_the_request_on_keyboard_interrupt():
    """A SIGINT while the worker runs: the main thread's 10 ms poll sees it,
    calls cancel (which lets the stand-in request end), waits for the worker
    and re-raises KeyboardInterrupt."""
    release = threading.Event()
    finished = []

    def request():
        release.wait(5.0)
        finished.append(release.is_set())
        return 0

    timer = threading.Timer(0.05, lambda: signal.raise_signal(signal.SIGINT))
    timer.start()
    t0 = time.perf_counter()
    with pytest.raises(KeyboardInterrupt):
        _native.run_interruptible(request, release.set)
    timer.join()
    assert finished == [True]  # the worker ended through the cancel callback
    assert time.perf_counter() - t0 < 2.0


def test_tokens_raise_and_lower_without_a_gpu():
    lib = _native.lib()
    token = C.c_void_p()
    status = lib.kg_cancel_token_create(C.byref(token))
    if status == _abi.DEVICE_ERROR:
        pytest.skip("pinned host memory needs a HIP runtime with a device")
    assert status == _abi.OK
    assert lib.kg_cancel_requested(token) == 0
    lib.kg_cancel(token)
    assert lib.kg_cancel_requested(token) == 1
    lib.kg_cancel_reset(token)
    assert lib.kg_cancel_requested(token) == 0
    lib.kg_cancel_token_destroy(token)


# ---------------------------------------------------------------- GPU
def euro_call(ctx, paths=65536, seed=1009):
    s = _abi.Settings(paths, seed, 1, _abi.PSEUDO_RANDOM, 0, _abi.FP64, 0, 0)
    o = _abi.EuropeanTerms(1, 0, 100.0, us(START), us(END))
    m = _abi.BsmMarket(100.0, 0.05, 0.02, 0.2, us(START))
    out = _abi.Estimate()
    status = _native.lib().kg_vanilla_price(ctx.handle, C.byref(s), C.byref(o), C.byref(m), C.byref(out))
    return status, out


@pytest.mark.gpu
def test_a_raised_token_interrupts_calls_before_they_start(gpu):
    gpu.cancel()
    status, _ = euro_call(gpu)
    assert status == _abi.INTERRUPTED
    assert "interrupted" in _native.lib().kg_last_error(gpu.handle).decode()
    _native.lib().kg_cancel_reset(gpu.token)
    status, out = euro_call(gpu)
    assert status == _abi.OK and out.price == pytest.approx(9.249574403327761, rel=1e-10)


def priced_in_background(fn):
    box = {}

    def run():
        t0 = time.perf_counter()
        box["status"] = fn()
        box["seconds"] = time.perf_counter() - t0

    thread = threading.Thread(target=run)
    thread.start()
    return thread, box


@pytest.mark.gpu
def test_cancelling_the_full_book_mid_launch(gpu, oracle):
    """The 4,096-contract book in FP64 (about 0.75 s of device time in one
    persistent launch): the token raised 60 ms in stops every workgroup at
    its next poll, the call returns KG_INTERRUPTED well before the book would
    end, and the same context then prices a small book as the single-contract
    engines do."""
    from kiyosi_amd import workloads
    from kiyosi_amd.engines import Book
    book = Book(workloads.book_4096())
    book.price("fp64")  # autotune and warm-up, uncancelled
    out = (_abi.Estimate * len(book.entries))()
    tuning = _abi.BookTuning(0, 0, 0, 0, 0.0, 0.0)
    _native.lib().kg_cancel_reset(gpu.token)
    thread, box = priced_in_background(lambda: _native.lib().kg_price_book(
        gpu.handle, book._array, len(book.entries), _abi.FP64, C.byref(tuning), out))
    time.sleep(0.06)
    gpu.cancel()
    thread.join(30)
    assert not thread.is_alive()
    assert box["status"] == _abi.INTERRUPTED
    assert box["seconds"] < 0.5
    small = workloads.book_4096(n=24, path_scale=1 / 64)
    estimates, _ = Book(small).price("fp64")
    for (engine, instrument, context), got in zip(small, estimates):
        # The book's one-pass autocallable kernel and the single engine's
        # two-pass kernels agree to rounding (as in test_gpu_book.py).
        assert got.price == pytest.approx(engine.price(instrument, context), rel=1e-12, abs=1e-14)


@pytest.mark.gpu
def test_cancelling_an_mlp_fit_between_dates(gpu):
    """The MLP regressor's fit synchronises once per exercise date; a token
    raised 15 ms in ends the call at the next date."""
    from kiyosi_amd import workloads
    from kiyosi_amd.market import american_model
    option, context, engine, training = workloads.american_put_local_vol()
    s = _abi.Settings(1 << 20, 1729, 1, _abi.PSEUDO_RANDOM, 0, _abi.FP32_STEP, 0, 0)
    opts = _abi.LsmcOptions(_abi.BASIS_MLP, 1, training, 4, 16384, 1e-3)
    report = _abi.LsmcReport()
    terms, model = option._american_terms(50), american_model(context)
    _native.lib().kg_cancel_reset(gpu.token)
    thread, box = priced_in_background(lambda: _native.lib().kg_american_price(
        gpu.handle, C.byref(s), C.byref(terms), C.byref(model), C.byref(opts), C.byref(report)))
    time.sleep(0.015)
    gpu.cancel()
    thread.join(30)
    assert box["status"] == _abi.INTERRUPTED
    _native.lib().kg_cancel_reset(gpu.token)
    assert euro_call(gpu)[0] == _abi.OK


@pytest.mark.gpu
def test_keyboard_interrupt_cancels_a_python_book_call(gpu):
    """SIGINT during Book.price (an interruptible request): the main thread's
    poll raises KeyboardInterrupt after cancelling the token, so the call
    returns early instead of running the whole book."""
    from kiyosi_amd import workloads
    from kiyosi_amd.engines import Book
    book = Book(workloads.book_4096())
    book.price("fp64")
    timer = threading.Timer(0.06, lambda: signal.raise_signal(signal.SIGINT))
    timer.start()
    t0 = time.perf_counter()
    with pytest.raises(KeyboardInterrupt):
        book.price("fp64")
    timer.join()
    assert time.perf_counter() - t0 < 0.5
    estimates, _ = book.price("fp64", use_graph=False)
    assert len(estimates) == 4096
