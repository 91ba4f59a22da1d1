"""Throughput of the Monte Carlo engine on one MI355X.

Workload: a snowball under BSM with 252 daily steps, 12 monthly knock-out
observations stepping down 0.5% from 100%, daily knock-in at 75%, coupon 15%
p.a., sigma 25%, r 2%, q 0, priced with 2^20 scrambled Sobol paths (seed
1729). One step = one complete pricing call through the C ABI (plan build,
sampling kernel, block merge, result copy) with the FP32-stepping /
FP64-accumulating engine. The plain FP64 engine runs the same paths for the
speed ratio and the price error, which is stated in units of the path
standard error.

Prints one JSON line. ``value`` is path-steps per second of the FP32-step
engine; ``roofline`` compares the sampling kernel's algorithmic FLOP rate
(9 FLOP per path-step: Sobol inverse normal 3, bridge fill 4, knock-in
compare 1, event/terminal share 1) with the FP32 vector peak, using the
kernel duration from HIP events on the engine's stream; ``cpu_baseline``
times the FP64 CPU restatement of the reference engine on the host cores.
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_engine(engine, note, context, steps, warmup):
    for _ in range(warmup):
        engine.price_with_statistics(note, context)
    device_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        est = engine.price_with_statistics(note, context)
        device_ms.append(est.device_ms)
    wall = time.perf_counter() - t0
    return est, wall, sum(device_ms) / len(device_ms)


def cpu_baseline(note, context, kw, seconds=10.0):
    """The FP64 CPU restatement (oracle) on the host cores, repeated full
    pricings until about `seconds` have elapsed."""
    from kiyosi_amd import _abi
    from oracle import oracle
    cores = min(os.cpu_count() or 1, 16)
    s = _abi.Settings(kw["path_count"], kw["seed"], 1, _abi.SOBOL, 0, 0, 0, 1)
    terms, market = note._terms(), context.bsm_market()
    oracle.autocallable_price(s, terms, market, workers=cores)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.autocallable_price(s, terms, market, workers=cores)
        reps += 1
        elapsed = time.perf_counter() - t0
        if elapsed >= seconds:
            break
    path_steps = kw["path_count"] * 252 * reps
    return {"value": path_steps / elapsed, "unit": "path-steps/s", "cores": cores, "kind": "port",
            "sample": f"{reps} full pricings of the 2^20-path snowball (FP64 CPU restatement, "
                      f"{cores} threads, {elapsed:.1f} s)"}


def main():
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = parser.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the engine shards paths by block across GPUs by design, not in this build)",
              file=sys.stderr)
        sys.exit(2)

    import kiyosi_amd as k
    from kiyosi_amd import workloads

    note, context, kw = workloads.snowball_252d_sobol_2p20()
    fp32 = k.MonteCarloAutocallableEngine(precision="fp32", **kw)
    fp64 = k.MonteCarloAutocallableEngine(precision="fp64", **kw)
    log("bench: FP64 plain engine")
    est64, wall64, dev64 = run_engine(fp64, note, context, max(3, args.steps // 4), 1)
    log(f"bench: fp64 {wall64:.3f}s, kernel {dev64:.3f} ms")
    log("bench: FP32-step engine")
    est32, wall32, dev32 = run_engine(fp32, note, context, args.steps, args.warmup)
    log(f"bench: fp32 {wall32:.3f}s, kernel {dev32:.3f} ms")

    path_steps = est32.path_steps
    steps64 = max(3, args.steps // 4)
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
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "autocall_kernel<sobol, fp32>", "kernel_ms": dev32},
    }
    if not args.no_cpu:
        log("bench: CPU baseline")
        result["cpu_baseline"] = cpu_baseline(note, context, kw)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
