"""Throughput of the Monte Carlo engine on one MI355X.

Headline workload (config 2): a snowball under BSM with 252 daily steps, 12
monthly knock-out observations stepping down 0.5% from 100%, daily knock-in
at 75%, coupon 15% p.a., sigma 25%, r 2%, q 0, priced with 2^20 scrambled
Sobol paths (seed 1729). One step = one complete pricing call through the C
ABI (plan build, sampling kernel, block merge, result copy) with the
FP32-stepping / FP64-accumulating engine. The plain FP64 engine runs the same
paths for the speed ratio and the price error, stated in units of the path
standard error.

Prints one JSON line. ``value`` is path-steps per second of the FP32-step
engine; ``roofline`` compares the sampling kernels' algorithmic FLOP rate (9
FLOP per nominal path-step: Sobol inverse normal 3, bridge fill 4, knock-in
compare 1, event/terminal share 1, counted over all 2^20 x 252 path-steps
although the kernels fill only the paths whose payoff depends on it) with
the FP32 vector peak, using the kernels' duration from HIP events on the
engine's stream; ``cpu_baseline`` times the
FP64 CPU restatement of the reference engine on the host cores. ``configs``
adds the other workloads: the European call (config 1), the Heston Asian
(config 3), the 4,096-contract book (config 4) and the local-volatility
American put with polynomial and MLP regressors (config 5).
"""

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FLOP_PER_PATH_STEP = 9.0
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
# HBM bytes per config-2 FP32 pricing of its two sampling kernels
# (bridge_events_kernel 13,438,765 + bridge_fill_kernel 16,685,122; the
# Sobol base rows are kept across calls): FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
# from the two counter passes of `bench.py --profile-pass`, scaled on a known
# 256 MiB read/write (profiles/hbm_traffic_fetch_write_counters.md).
C2_KERNEL_HBM_BYTES = 30_123_887


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_engine(engine, instrument, context, steps, warmup):
    for _ in range(warmup):
        engine.price_with_statistics(instrument, context)
    device_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        est = engine.price_with_statistics(instrument, context)
        device_ms.append(est.device_ms)
    wall = time.perf_counter() - t0
    return est, wall, sum(device_ms) / len(device_ms)


def pair(fp64, fp32, path_steps, reps64, reps32):
    """Speed and agreement of an FP32-step run against the FP64 plain run of
    the same paths: (est, wall, kernel_ms) tuples."""
    (e64, w64, d64), (e32, w32, d32) = fp64, fp32
    se = e64.path_standard_error
    return {"fp64_path_steps_per_s": path_steps * reps64 / w64, "fp32_path_steps_per_s": path_steps * reps32 / w32,
            "speedup": (path_steps * reps32 / w32) / (path_steps * reps64 / w64),
            "fp64_kernel_ms": d64, "fp32_kernel_ms": d32, "kernel_speedup": d64 / d32 if d32 else None,
            "price_fp64": e64.price, "price_fp32": e32.price,
            "price_error_se": abs(e32.price - e64.price) / se if se else None, "path_steps": path_steps}


def config1(reps):
    import kiyosi_amd as k
    from kiyosi_amd import workloads
    option, context, kw = workloads.european_call_1m_pcg_antithetic()
    r64 = run_engine(k.MonteCarloVanillaEngine(precision="fp64", **kw), option, context, reps, 2)
    r32 = run_engine(k.MonteCarloVanillaEngine(precision="fp32", **kw), option, context, reps, 2)
    out = pair(r64, r32, r32[0].path_steps, reps, reps)
    out["workload"] = "European call, 1e6 antithetic PCG paths, 1 step"
    return out


def config3(reps):
    import ctypes as C

    from kiyosi_amd import _abi, _native, workloads
    ctx = _native.context(0)

    def run(precision):
        s, o, m = workloads.heston_asian_252_pcg_2p22(precision)
        est = _abi.Estimate()
        ctx.check(_native.lib().kg_heston_price(ctx.handle, C.byref(s), C.byref(o), C.byref(m), C.byref(est)))
        t0, dev = time.perf_counter(), 0.0
        for _ in range(reps):
            ctx.check(_native.lib().kg_heston_price(ctx.handle, C.byref(s), C.byref(o), C.byref(m), C.byref(est)))
            dev += est.device_ms
        return est, time.perf_counter() - t0, dev / reps

    r64, r32 = run("fp64"), run("fp32")
    out = pair(r64, r32, r32[0].path_steps, reps, reps)
    out["workload"] = "Heston arithmetic Asian call, 2^22 PCG paths x 252 steps"
    return out


def config4(reps):
    from kiyosi_amd import workloads
    from kiyosi_amd.engines import Book
    t0 = time.perf_counter()
    book = Book(workloads.book_4096())
    plan_s = time.perf_counter() - t0
    out = {"workload": "4,096-contract book (barriers, Asians, snowballs, Phoenix, accumulators), one launch",
           "contracts": len(book.entries), "python_plan_s": plan_s}
    est = {}
    for precision in ("fp64", "fp32"):
        book.price(precision)  # autotune + warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            est[precision], tuning = book.price(precision, use_graph=True, replays=2)
        wall = (time.perf_counter() - t0) / reps
        out[precision] = {"path_steps": tuning.path_steps, "device_ms": tuning.device_ms,
                          "path_steps_per_s_device": tuning.path_steps / (tuning.device_ms * 1e-3),
                          "contracts_per_s_device": len(book.entries) / (tuning.device_ms * 1e-3),
                          "call_s": wall, "contracts_per_s_call": len(book.entries) / wall,
                          "samples_per_thread": tuning.samples_per_thread,
                          "workgroups_per_cu": tuning.workgroups_per_cu}
    errors = [abs(a.price - b.price) / b.path_standard_error for a, b in zip(est["fp32"], est["fp64"])
              if b.path_standard_error > 0]
    out["speedup_device"] = out["fp64"]["device_ms"] / out["fp32"]["device_ms"]
    out["fp32_error_se_max"] = max(errors)
    out["fp32_within_half_se"] = sum(e <= 0.5 for e in errors) / len(errors)
    return out


def config5():
    from kiyosi_amd import workloads
    from kiyosi_amd.engines import MonteCarloVanillaEngine
    option, context, engine, training = workloads.american_put_local_vol()
    out = {"workload": "American put, local vol, 50 dates, 2^20 training + 2^20 pricing paths, FP32 step"}
    eng = MonteCarloVanillaEngine(path_count=1 << 20, step_count=50, seed=1729, precision="fp32")
    eng.american_report(option, context, training_path_count=training, basis="poly4")  # warm-up
    for basis in ("quadratic", "poly4", "mlp"):
        t0 = time.perf_counter()
        r = eng.american_report(option, context, training_path_count=training, basis=basis)
        wall = time.perf_counter() - t0
        out[basis] = {"price": r.price, "standard_error": r.standard_error, "in_sample": r.in_sample_price,
                      "training_ms": r.training_ms, "pricing_ms": r.pricing_ms, "call_s": wall,
                      "path_steps": r.path_steps, "path_steps_per_s_call": r.path_steps / wall,
                      "mean_loss": r.mean_loss if basis == "mlp" else None}
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(note, context, kw, seconds=8.0):
    """The FP64 CPU restatement (oracle) on the host cores: repeated full
    2^20-path pricings on up to 16 threads, and 2^17-path pricings on one
    thread, each for about `seconds`."""
    from kiyosi_amd import _abi
    from oracle import oracle
    cores = min(os.cpu_count() or 1, 16)
    terms, market = note._terms(), context.bsm_market()

    def rate(paths, workers):
        s = _abi.Settings(paths, kw["seed"], 1, _abi.SOBOL, 0, 0, 0, 1)
        oracle.autocallable_price(s, terms, market, workers=workers)  # warm-up
        reps, t0 = 0, time.perf_counter()
        while True:
            oracle.autocallable_price(s, terms, market, workers=workers)
            reps += 1
            elapsed = time.perf_counter() - t0
            if elapsed >= seconds:
                break
        return paths * 252 * reps / elapsed, reps, elapsed

    multi, reps, elapsed = rate(kw["path_count"], cores)
    single, reps1, elapsed1 = rate(1 << 17, 1)
    return {"value": multi, "unit": "path-steps/s", "cores": cores, "kind": "port",
            "sample": f"{reps} full pricings of the 2^20-path snowball on {cores} threads ({elapsed:.1f} s) and "
                      f"{reps1} pricings of a 2^17-path prefix on 1 thread ({elapsed1:.1f} s); FP64 C++ "
                      f"restatement of kiyosi 0.23.0 MC",
            "single_core_value": single, "cpu": cpu_model(), "logical_cpus": os.cpu_count()}


def profile_pass():
    """A short run for hardware-counter passes: a torch elementwise multiply
    that reads and writes 256 MiB each (the calibration of FETCH_SIZE and
    WRITE_SIZE), three config-2 FP32 pricings and one config-5 poly-4 LSMC."""
    import torch

    import kiyosi_amd as k
    from kiyosi_amd import workloads
    n = 64 << 20
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    log(f"profile pass: calibration multiply of {n * 4} bytes read and written per launch")
    note, context, kw = workloads.snowball_252d_sobol_2p20()
    engine = k.MonteCarloAutocallableEngine(precision="fp32", **kw)
    for _ in range(3):
        engine.price_with_statistics(note, context)
    option, context5, _, training = workloads.american_put_local_vol()
    e5 = k.MonteCarloVanillaEngine(path_count=1 << 20, step_count=50, seed=1729, precision="fp32")
    e5.american_report(option, context5, training_path_count=training, basis="poly4")
    log("profile pass: done")


def main():
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    parser.add_argument("--headline-only", action="store_true", help="skip configs 1, 3, 4 and 5")
    parser.add_argument("--profile-pass", action="store_true",
                        help="run the kernels once for a hardware-counter pass; prints no result line")
    args = parser.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the engine shards paths by block across GPUs by design, not in this build)",
              file=sys.stderr)
        sys.exit(2)
    if args.profile_pass:
        profile_pass()
        return

    import kiyosi_amd as k
    from kiyosi_amd import workloads

    note, context, kw = workloads.snowball_252d_sobol_2p20()
    fp32 = k.MonteCarloAutocallableEngine(precision="fp32", **kw)
    fp64 = k.MonteCarloAutocallableEngine(precision="fp64", **kw)
    log("bench: config 2, FP64 plain engine")
    steps64 = max(3, args.steps // 4)
    est64, wall64, dev64 = run_engine(fp64, note, context, steps64, 1)
    log(f"bench: fp64 {wall64:.3f}s, kernel {dev64:.3f} ms")
    log("bench: config 2, FP32-step engine")
    est32, wall32, dev32 = run_engine(fp32, note, context, args.steps, args.warmup)
    log(f"bench: fp32 {wall32:.3f}s, kernel {dev32:.3f} ms")

    path_steps = est32.path_steps
    value = path_steps * args.steps / wall32
    value64 = path_steps * steps64 / wall64
    se = est64.path_standard_error
    achieved = FLOP_PER_PATH_STEP * path_steps / (dev32 * 1e-3) / 1e12
    result = {
        "metric": "path-steps/sec on 1 MI355X, FP32-step vs FP64 plain at |dPrice|<=0.5 SE",
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-step/fp64-accumulate",
        "data": "synthetic",
        "config": {"workload": "snowball, 252 daily steps, 12 monthly knock-outs, 2^20 Sobol paths",
                   "paths": kw["path_count"], "steps_per_path": 252, "sampling": "sobol",
                   "parallelism": "single GPU"},
        "fp64_plain": {"value": value64, "ms_per_step": 1e3 * wall64 / steps64,
                       "kernel_ms": dev64, "price": est64.price},
        "speedup_vs_fp64": value / value64,
        "price": est32.price,
        "price_error_se": abs(est32.price - est64.price) / se if se else None,
        "contracts_per_sec": args.steps / wall32,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": C2_KERNEL_HBM_BYTES,
                     "kernel": "bridge_events_kernel + bridge_fill_kernel <sobol, fp32>", "kernel_ms": dev32,
                     "basis": "9 FLOP per nominal path-step (2^20 x 252), the reference algorithm's count; "
                              "the kernels skip daily fills that cannot change the price"},
    }
    if not args.headline_only:
        configs = {}
        for name, fn in (("c1", lambda: config1(20)), ("c3", lambda: config3(3)), ("c4", lambda: config4(2)),
                         ("c5", config5)):
            log(f"bench: {name}")
            t0 = time.perf_counter()
            configs[name] = fn()
            log(f"bench: {name} done in {time.perf_counter() - t0:.1f}s")
        result["configs"] = configs
    if not args.no_cpu:
        log("bench: CPU baseline")
        result["cpu_baseline"] = cpu_baseline(note, context, kw)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
