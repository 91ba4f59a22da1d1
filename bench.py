"""Throughput of the Monte Carlo engine on one MI355X.

Headline workload (config 2): a snowball under BSM with 252 daily steps, 12
monthly knock-out observations stepping down 0.5% from 100%, daily knock-in
at 75%, coupon 15% p.a., sigma 25%, r 2%, q 0, priced with 2^20 scrambled
Sobol paths. One step = one complete pricing call through the C ABI (plan
build, Sobol scrambling for the call's seed, sampling kernels, block merge,
result copy) with the FP32-stepping / FP64-accumulating engine. Every timed
call takes a fresh seed, as an unseeded reference call draws a new scrambled
sequence, so nothing seed-dependent is reused between calls ("cold");
repeated calls on one seed ("warm") are reported beside. The plain FP64
engine runs cold calls too for the speed ratio, and both engines price seed
1729 for the price error, stated in units of the path standard error.

Prints one JSON line. ``value`` is cold path-steps per second of the
FP32-step engine; ``roofline`` compares the sampling kernels' algorithmic
FLOP rate (9 FLOP per nominal path-step: Sobol inverse normal 3, bridge fill
4, knock-in compare 1, event/terminal share 1, counted over all 2^20 x 252
path-steps although the kernels fill only the paths whose payoff depends on
it) with the FP32 vector peak, using the two sampling kernels' duration from
HIP events on the engine's stream; ``cpu_baseline`` times the FP64 CPU
restatement of the reference engine on the host cores. ``configs`` adds the
other workloads, each with its own roofline and CPU leg: the European call
(config 1), the Heston Asian (config 3), the 4,096-contract book (config 4)
and the local-volatility American put with polynomial and MLP regressors
(config 5).
"""

import argparse
import json
import os
import pathlib
import statistics
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0  # dense MFMA
HBM_PEAK_GBS = 8000.0
# Algorithmic FLOP per path-step (FMA 2, other arithmetic and elementary
# functions 1, integer RNG work 0).
C1_FLOP = 21.0   # per path: normal 6 per pair, payoffs, moments
C2_FLOP = 9.0    # Sobol normal 3, bridge fill 4, knock-in compare 1, events 1
C3_FLOP = 32.0   # two normals 12, correlation 3, truncation 2, variance 6, log-spot 6, exp 2, average 1
C4_FLOP = {"BarrierOption": 10.0, "GeometricAveragePriceOption": 10.0, "ArithmeticAveragePriceOption": 11.0,
           "SnowballOption": 10.0, "PhoenixOption": 10.0, "Accumulator": 13.0}
C5_FORWARD_FLOP = 25.0  # normal 6, local-vol bilinear 11, log-Euler 6, exp 2
C5_BYTES_PER_PATH_DATE = 24.0  # FP32 coordinate stored and reloaded, FP64 cash read and written
MLP_MFMA_FLOP = 3 * 2 * 64 * 64  # per sample and Adam step: layer-2 forward, input gradient, weight gradient
# HBM bytes per config-2 FP32 pricing of its two sampling kernels
# (bridge_events_kernel 13.45 MB + bridge_fill_kernel 16.54 MB): FETCH_SIZE
# and WRITE_SIZE from the two counter passes of `bench.py --profile-pass`,
# scaled on a known 256 MiB read/write.
C2_KERNEL_HBM_BYTES = 29_988_010
# Host threads for the CPU legs: the host share of one GPU on the box.
CPU_THREADS = 16
CPU_MIN_RUNS = 15


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def roofline(bound, achieved, peak, unit, traffic=None, **extra):
    out = {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
           "traffic": traffic}
    out.update(extra)
    return out


def timed_calls(engine, instrument, context, steps, seeds=None):
    """`steps` pricing calls; with `seeds`, call i uses seeds[i]. Returns the
    last estimate, wall seconds and the mean sampling-kernel milliseconds."""
    device_ms = []
    t0 = time.perf_counter()
    for i in range(steps):
        if seeds is not None:
            engine.seed = seeds[i]
        est = engine.price_with_statistics(instrument, context)
        device_ms.append(est.device_ms)
    return est, time.perf_counter() - t0, sum(device_ms) / len(device_ms)


def median_rate(work, fn, seconds):
    """Median of at least CPU_MIN_RUNS timed runs of fn() (more while under
    `seconds`): returns (work per second at the median run, runs, total s)."""
    fn()  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < CPU_MIN_RUNS or time.perf_counter() - t_start < seconds:
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return work / med, len(times), time.perf_counter() - t_start, med


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(value, unit, threads, runs, total, median_s, sample):
    return {"value": value, "unit": unit, "cores": threads, "kind": "port", "runs": runs,
            "median_run_s": median_s, "total_s": total, "sample": sample,
            "cpu": cpu_model(), "logical_cpus": os.cpu_count()}


# ---------------------------------------------------------------- config 1
def config1(reps, cpu):
    import kiyosi_amd as k
    from kiyosi_amd import workloads
    option, context, kw = workloads.european_call_1m_pcg_antithetic()
    out = {"workload": "European call, 1e6 antithetic PCG paths, 1 step"}
    res = {}
    for precision in ("fp64", "fp32"):
        engine = k.MonteCarloVanillaEngine(precision=precision, **kw)
        timed_calls(engine, option, context, 2)
        res[precision] = timed_calls(engine, option, context, reps, [50_000 + i for i in range(reps)])
        engine.seed = kw["seed"]
        res[precision] = (engine.price_with_statistics(option, context),) + res[precision][1:]
    (e64, w64, d64), (e32, w32, d32) = res["fp64"], res["fp32"]
    paths = kw["path_count"]
    out.update({"fp64_path_steps_per_s": paths * reps / w64, "fp32_path_steps_per_s": paths * reps / w32,
                "speedup": w64 / w32, "fp64_kernel_ms": d64, "fp32_kernel_ms": d32, "kernel_speedup": d64 / d32,
                "price_fp64": e64.price, "price_fp32": e32.price,
                "price_error_se": abs(e32.price - e64.price) / e64.standard_error})
    out["roofline"] = roofline("valu", C1_FLOP * paths / (d32 * 1e-3) / 1e12, FP32_PEAK_TFLOPS, "TFLOP/s",
                               kernel="european_kernel<fp32>", kernel_ms=d32,
                               basis="21 FLOP per path (normal 6 per antithetic pair, 2 payoffs x 8, moments)")
    if cpu:
        from kiyosi_amd import _abi
        from oracle import oracle
        s = _abi.Settings(paths, kw["seed"], 1, _abi.PSEUDO_RANDOM, 1, _abi.FP64, 0, 0)
        o, m = option._terms(), context.bsm_market()
        rate, runs, total, med = median_rate(paths, lambda: oracle.vanilla_price(s, o, m, workers=CPU_THREADS), 3.0)
        out["cpu_baseline"] = cpu_leg(rate, "path-steps/s", CPU_THREADS, runs, total, med,
                                      f"median of {runs} full 1e6-path pricings on {CPU_THREADS} threads")
    return out


# ---------------------------------------------------------------- config 3
def config3(reps, cpu):
    import ctypes as C

    from kiyosi_amd import _abi, _native, workloads
    ctx = _native.context(0)

    def run(precision, seeds):
        s, o, m = workloads.heston_asian_252_pcg_2p22(precision)
        est = _abi.Estimate()
        ctx.check(_native.lib().kg_heston_price(ctx.handle, C.byref(s), C.byref(o), C.byref(m), C.byref(est)))
        first = _abi.Estimate()
        first.price, first.path_standard_error = est.price, est.path_standard_error
        t0, dev = time.perf_counter(), 0.0
        for seed in seeds:
            s.seed = seed
            ctx.check(_native.lib().kg_heston_price(ctx.handle, C.byref(s), C.byref(o), C.byref(m), C.byref(est)))
            dev += est.device_ms
        return first, time.perf_counter() - t0, dev / len(seeds), est.path_steps

    e64, w64, d64, steps = run("fp64", [60_000 + i for i in range(reps)])
    e32, w32, d32, _ = run("fp32", [70_000 + i for i in range(reps)])
    out = {"workload": "Heston arithmetic Asian call, 2^22 PCG paths x 252 steps",
           "fp64_path_steps_per_s": steps * reps / w64, "fp32_path_steps_per_s": steps * reps / w32,
           "speedup": w64 / w32, "fp64_kernel_ms": d64, "fp32_kernel_ms": d32, "kernel_speedup": d64 / d32,
           "price_fp64": e64.price, "price_fp32": e32.price,
           "price_error_se": abs(e32.price - e64.price) / e64.path_standard_error, "path_steps": steps}
    out["roofline"] = roofline("valu", C3_FLOP * steps / (d32 * 1e-3) / 1e12, FP32_PEAK_TFLOPS, "TFLOP/s",
                               kernel="heston_kernel<fp32>", kernel_ms=d32,
                               basis="32 FLOP per path-step (2 normals 12, correlate 3, truncate 2, variance 6, "
                                     "log-spot 6, exp 2, average 1)")
    out["roofline_fp64"] = roofline("valu", C3_FLOP * steps / (d64 * 1e-3) / 1e12, FP64_PEAK_TFLOPS, "TFLOP/s",
                                    kernel="heston_kernel<fp64>", kernel_ms=d64)
    if cpu:
        from oracle import oracle
        s, o, m = workloads.heston_asian_252_pcg_2p22("fp64")
        s.path_count = 1 << 16
        work = s.path_count * 252
        rate, runs, total, med = median_rate(work, lambda: oracle.heston_price(s, o, m, workers=CPU_THREADS), 3.0)
        out["cpu_baseline"] = cpu_leg(rate, "path-steps/s", CPU_THREADS, runs, total, med,
                                      f"median of {runs} pricings of a 2^16-path prefix x 252 steps on "
                                      f"{CPU_THREADS} threads")
    return out


# ---------------------------------------------------------------- config 4
def oracle_entry(oracle, engine, instrument, context):
    s, m = engine._settings(), context.bsm_market()
    name = type(engine).__name__
    if name == "MonteCarloBarrierEngine":
        return oracle.barrier_price(s, instrument._terms(), m, workers=CPU_THREADS)
    if name == "MonteCarloAsianEngine":
        return oracle.asian_price(s, instrument._terms(engine.step_count), m, workers=CPU_THREADS)
    if name == "MonteCarloAutocallableEngine":
        return oracle.autocallable_price(s, instrument._terms(), m, workers=CPU_THREADS)
    return oracle.accumulator_price(s, instrument._terms(), m, workers=CPU_THREADS)


def config4(reps, cpu):
    import numpy as np

    from kiyosi_amd import workloads
    from kiyosi_amd.engines import Book
    t0 = time.perf_counter()
    entries = workloads.book_4096()
    book = Book(entries)
    plan_s = time.perf_counter() - t0
    n = len(entries)
    out = {"workload": "4,096-contract book (barriers, Asians, snowballs, Phoenix, accumulators), one launch",
           "contracts": n, "python_plan_s": plan_s}
    est = {}
    replays = 2
    for precision in ("fp64", "fp32"):
        book.price(precision)  # autotune + warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            est[precision], tuning = book.price(precision, use_graph=True, replays=replays)
        wall = (time.perf_counter() - t0) / (reps * (1 + replays))
        out[precision] = {"path_steps": tuning.path_steps, "device_ms": tuning.device_ms,
                          "path_steps_per_s_device": tuning.path_steps / (tuning.device_ms * 1e-3),
                          "contracts_per_s_device": n / (tuning.device_ms * 1e-3),
                          "execution_s": wall, "contracts_per_s_execution": n / wall,
                          "samples_per_thread": tuning.samples_per_thread,
                          "workgroups_per_cu": tuning.workgroups_per_cu}
    errors = [abs(a.price - b.price) / b.path_standard_error for a, b in zip(est["fp32"], est["fp64"])
              if b.path_standard_error > 0]
    out["speedup_device"] = out["fp64"]["device_ms"] / out["fp32"]["device_ms"]
    out["fp32_error_se_max"] = max(errors)
    out["fp32_within_half_se"] = sum(e <= 0.5 for e in errors) / len(errors)
    flop = sum(C4_FLOP[type(ins).__name__] * e.path_steps for (_, ins, _), e in zip(entries, est["fp32"]))
    dev = out["fp32"]["device_ms"]
    out["roofline"] = roofline("valu", flop / (dev * 1e-3) / 1e12, FP32_PEAK_TFLOPS, "TFLOP/s",
                               kernel="book_kernel<fp32, class> x 7 (graph: reset, one persistent kernel per path class, merge)", kernel_ms=dev,
                               basis="per path-step: barrier 10, geometric Asian 10, arithmetic Asian 11, "
                                     "snowball/Phoenix 10, accumulator 13 FLOP, weighted by each "
                                     f"contract's path-steps ({flop / out['fp32']['path_steps']:.2f} mean)")
    if cpu:
        from oracle import oracle
        pick = sorted(np.random.default_rng(20251204).choice(n, size=16, replace=False))
        work = sum(est["fp64"][i].path_steps for i in pick)
        rate, runs, total, med = median_rate(
            work, lambda: [oracle_entry(oracle, *entries[i]) for i in pick], 4.0)
        leg = cpu_leg(rate, "path-steps/s", CPU_THREADS, runs, total, med,
                      f"median of {runs} passes over 16 seeded-sample contracts of the book at full path counts "
                      f"({work:.3g} path-steps per pass), contract by contract on {CPU_THREADS} threads")
        leg["contracts_per_s"] = 16 / med
        out["cpu_baseline"] = leg
    return out


# ---------------------------------------------------------------- config 5
def config5(cpu):
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
                      "mean_loss": r.mean_loss if basis == "mlp" else None,
                      "gradient_samples": r.gradient_samples if basis == "mlp" else None}
    train_ms = out["poly4"]["training_ms"]
    algorithmic = training * 50 * C5_BYTES_PER_PATH_DATE
    out["roofline"] = roofline("hbm", algorithmic / (train_ms * 1e-3) / 1e9, HBM_PEAK_GBS, "GB/s",
                               kernel="poly-4 training (lsmc_generate + 49 x lsmc_backward + lsmc_solve)",
                               kernel_ms=train_ms,
                               basis="24 algorithmic bytes per training path and date: FP32 coordinate stored "
                                     "and reloaded, FP64 cash read and written")
    mlp = out["mlp"]
    out["roofline_mlp"] = roofline(
        "mfma", MLP_MFMA_FLOP * mlp["gradient_samples"] / (mlp["training_ms"] * 1e-3) / 1e12, BF16_PEAK_TFLOPS,
        "TFLOP/s", kernel="MLP training phase (mlp_grad + mlp_update per Adam step, per-date sample preparation)",
        kernel_ms=mlp["training_ms"],
        basis="24,576 bf16 MFMA FLOP per training sample and Adam step (forward 64x64, input gradient 64x64, "
              "weight gradient 64x64), over the whole training phase's device time")
    if cpu:
        from kiyosi_amd import _abi
        from kiyosi_amd.market import american_model
        from oracle import oracle
        s = _abi.Settings(1 << 16, 1729, 1, _abi.PSEUDO_RANDOM, 0, _abi.FP64, 0, 0)
        terms, model = option._american_terms(50), american_model(context)
        legs = {}
        for basis, code in (("quadratic", _abi.BASIS_QUADRATIC), ("poly4", _abi.BASIS_POLY4)):
            rate, runs, total, med = median_rate(
                s.path_count * 50, lambda: oracle.lv_lsmc(s, terms, model, code, workers=CPU_THREADS), 2.0)
            legs[basis] = cpu_leg(rate, "path-steps/s", CPU_THREADS, runs, total, med,
                                  f"median of {runs} in-sample {basis} LSMC fits on 2^16 paths x 50 dates, "
                                  f"{CPU_THREADS} threads")
        out["cpu_baseline"] = legs
    return out


# ---------------------------------------------------------------- headline
def headline_cpu_baseline(note, context, kw):
    """The FP64 CPU restatement (oracle) of the config-2 pricing on the host:
    the median of at least 15 full 2^20-path pricings on CPU_THREADS threads,
    and of 2^17-path pricings on one thread."""
    from kiyosi_amd import _abi
    from oracle import oracle
    terms, market = note._terms(), context.bsm_market()

    def leg(paths, workers, seconds):
        s = _abi.Settings(paths, kw["seed"], 1, _abi.SOBOL, 0, 0, 0, 1)
        return median_rate(paths * 252, lambda: oracle.autocallable_price(s, terms, market, workers=workers),
                           seconds)

    multi, runs, total, med = leg(kw["path_count"], CPU_THREADS, 4.0)
    single, runs1, total1, med1 = leg(1 << 17, 1, 3.0)
    out = cpu_leg(multi, "path-steps/s", CPU_THREADS, runs, total, med,
                  f"median of {runs} full pricings of the 2^20-path snowball on {CPU_THREADS} threads "
                  f"({total:.1f} s); single thread: median of {runs1} pricings of a 2^17-path prefix "
                  f"({total1:.1f} s); FP64 C++ restatement of kiyosi 0.23.0 MC")
    out["single_core_value"] = single
    return out


def profile_pass():
    """A short run for hardware-counter passes: a torch elementwise multiply
    that reads and writes 256 MiB each (the calibration of FETCH_SIZE and
    WRITE_SIZE), three config-2 FP32 pricings, one config-5 poly-4 LSMC and
    one config-5 MLP fit, and one FP32 pricing of the config-4 book."""
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
    e5.american_report(option, context5, training_path_count=training, basis="mlp")
    from kiyosi_amd.engines import Book
    Book(workloads.book_4096()).price("fp32", samples_per_thread=4, workgroups_per_cu=8)
    log("profile pass: done")


def main():
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
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
    # Prices on the workload's own seed (1729) for the FP32 error.
    est64 = fp64.price_with_statistics(note, context)
    est32 = fp32.price_with_statistics(note, context)
    log("bench: config 2, warm-up")
    timed_calls(fp32, note, context, args.warmup, [10_000 + i for i in range(args.warmup)])
    timed_calls(fp64, note, context, 1, [20_000])
    steps64 = max(3, args.steps // 4)
    log("bench: config 2, FP64 plain engine, fresh seed per call")
    _, wall64, dev64 = timed_calls(fp64, note, context, steps64, [30_000 + i for i in range(steps64)])
    log(f"bench: fp64 {wall64:.3f}s, kernels {dev64:.3f} ms")
    log("bench: config 2, FP32-step engine, fresh seed per call")
    _, wall32, dev32 = timed_calls(fp32, note, context, args.steps, [40_000 + i for i in range(args.steps)])
    log(f"bench: fp32 {wall32:.3f}s, kernels {dev32:.3f} ms")
    fp32.seed = kw["seed"]
    _, wall32w, dev32w = timed_calls(fp32, note, context, args.steps)
    fp64.seed = kw["seed"]
    _, wall64w, dev64w = timed_calls(fp64, note, context, steps64)

    path_steps = est32.path_steps
    value = path_steps * args.steps / wall32
    value64 = path_steps * steps64 / wall64
    se = est64.path_standard_error
    achieved = C2_FLOP * path_steps / (dev32 * 1e-3) / 1e12
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
        "config": {"workload": "snowball, 252 daily steps, 12 monthly knock-outs, 2^20 Sobol paths, "
                               "fresh seed per call",
                   "paths": kw["path_count"], "steps_per_path": 252, "sampling": "sobol",
                   "parallelism": "single GPU"},
        "fp64_plain": {"value": value64, "ms_per_step": 1e3 * wall64 / steps64, "kernel_ms": dev64,
                       "price": est64.price},
        "speedup_vs_fp64": value / value64,
        "warm_seed": {"fp32_value": path_steps * args.steps / wall32w, "fp32_ms_per_step": 1e3 * wall32w / args.steps,
                      "fp64_value": path_steps * steps64 / wall64w, "fp32_kernel_ms": dev32w,
                      "fp64_kernel_ms": dev64w},
        "price": est32.price,
        "price_error_se": abs(est32.price - est64.price) / se if se else None,
        "contracts_per_sec": args.steps / wall32,
        "roofline": roofline("valu", achieved, FP32_PEAK_TFLOPS, "TFLOP/s", C2_KERNEL_HBM_BYTES,
                             kernel="bridge_events_kernel + bridge_fill_kernel <sobol, fp32>", kernel_ms=dev32,
                             basis="9 FLOP per nominal path-step (2^20 x 252), the reference algorithm's count; "
                                   "the kernels skip daily fills that cannot change the price"),
    }
    if not args.headline_only:
        configs = {}
        cpu = not args.no_cpu
        for name, fn in (("c1", lambda: config1(20, cpu)), ("c3", lambda: config3(3, cpu)),
                         ("c4", lambda: config4(2, cpu)), ("c5", lambda: config5(cpu))):
            log(f"bench: {name}")
            t0 = time.perf_counter()
            configs[name] = fn()
            log(f"bench: {name} done in {time.perf_counter() - t0:.1f}s")
        result["configs"] = configs
    if not args.no_cpu:
        log("bench: CPU baseline")
        result["cpu_baseline"] = headline_cpu_baseline(note, context, kw)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
