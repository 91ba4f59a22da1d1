"""Extracts the reference-held Monte Carlo vectors into tests/golden/*.json.

Reads kiyosi's benchmark calibration and reference files (data, not code):
  benchmarks/quantlib/calibration/monte_carlo.json          per-seed prices
  benchmarks/quantlib/calibration/american_monte_carlo.json per-seed LSMC prices
  benchmarks/quantlib/references/*.json                     analytic / FD values
and records the golden values of python/tests/test_monte_carlo.py (:46-52
and :761) and of crates/kiyosi/tests/double_barrier_monte_carlo.rs
(:121-157). Run once where the kiyosi sources are available:

    python tests/golden/extract_fixtures.py /path/to/kiyosi
"""

import json
import pathlib
import sys

OUT = pathlib.Path(__file__).resolve().parent


def main() -> None:
    root = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    bench = root / "benchmarks" / "quantlib"

    cases = {c["id"]: c for c in json.loads((bench / "cases" / "monte_carlo.json").read_text())["cases"]}
    calibration = json.loads((bench / "calibration" / "monte_carlo.json").read_text())
    runs = []
    for experiment in calibration["experiments"]:
        if experiment["backend"] != "kiyosi":
            continue
        case = cases[experiment["id"]]
        runs.append({
            "id": experiment["id"],
            "inputs": case["inputs"],
            "path_count": experiment["profile"]["path_count"],
            "antithetic": experiment["profile"]["antithetic"] == "true",
            "samples": [{"seed": s["seed"], "price": s["price"]}
                        for s in experiment["seed_ensemble"]["samples"]],
            "timing_seed": {"seed": experiment["profile"]["seed"],
                            "price": experiment["timing_seed_accuracy"]["price"]},
        })
    (OUT / "calibration_monte_carlo.json").write_text(json.dumps({
        "source": "benchmarks/quantlib/calibration/monte_carlo.json (kiyosi 0.19.0 build)",
        "runs": runs,
    }, indent=1))

    american_cases = {c["id"]: c for c in
                      json.loads((bench / "cases" / "american_monte_carlo.json").read_text())["cases"]}
    american = json.loads((bench / "calibration" / "american_monte_carlo.json").read_text())
    amer_runs = []
    for experiment in american.get("experiments", []):
        if experiment.get("backend") != "kiyosi":
            continue
        amer_runs.append({
            "id": "mc-american-put",
            "dimension": experiment.get("dimension"),
            "inputs": american_cases["mc-american-put"]["inputs"],
            "profile": experiment["profile"],
            "timing_seed": {"seed": experiment["profile"]["seed"],
                            "price": experiment["timing_seed_accuracy"]["price"]},
            "samples": [{"seed": s["seed"], "price": s["price"]}
                        for s in experiment["seed_ensemble"]["samples"]],
        })
    (OUT / "calibration_american.json").write_text(json.dumps({
        "source": "benchmarks/quantlib/calibration/american_monte_carlo.json (kiyosi 0.19.0 build)",
        "runs": amer_runs,
    }, indent=1))

    references = {}
    for name in ("monte_carlo", "american_monte_carlo", "heston", "fd_local_vol", "fd_bsm",
                 "barriers", "digitals_asian", "european"):
        path = bench / "references" / f"{name}.json"
        if path.exists():
            data = json.loads(path.read_text())
            case_path = bench / "cases" / f"{name}.json"
            inputs = {}
            if case_path.exists():
                inputs = {c["id"]: c.get("inputs") for c in json.loads(case_path.read_text())["cases"]}
            references[name] = [{"id": c["id"], "value": c["value"],
                                 "uncertainty": c.get("uncertainty"),
                                 "inputs": inputs.get(c["id"])} for c in data["cases"]]
    (OUT / "references.json").write_text(json.dumps(references, indent=1))

    (OUT / "python_golden.json").write_text(json.dumps({
        "sobol_seed42_dim0_point0_normal": {
            "value": -0.10364940758940244,
            "source": "python/tests/test_monte_carlo.py:761 (SciPy ndtri reference)",
        },
        "sobol_high_variance_tail": {
            "value": 5.805502904230645e-42, "rel": 1e-12,
            "inputs": {"spot": 1e308, "rate": 800.0, "dividend": 0.0, "sigma": 40.0,
                       "strike": 1.0, "path_count": 1, "seed": 42, "sampling": "sobol"},
            "source": "python/tests/test_monte_carlo.py:46-52; crates/kiyosi/tests/monte_carlo.rs:58-87",
        },
        "se_constants": {
            "plain": 13.831466750894657, "pair": 7.286842354,
            "source": "crates/kiyosi/tests/monte_carlo.rs:800-805; python/tests/test_monte_carlo.py:648",
        },
        "double_barrier_knock_out_greeks": {
            "inputs": {"strike": 100.0, "lower": 80.0, "upper": 120.0, "spot": 100.0, "rate": 0.04,
                       "dividend": 0.01, "sigma": 0.2, "valuation": "2025-01-01", "expiry": "2026-01-01",
                       "monitoring": "continuous", "sampling": "sobol", "path_count": 32768,
                       "seeds": [17, 42, 103]},
            "call": {"price": 1.0838814592583166, "delta": -0.009437716341942348,
                     "gamma": -0.006614552070516311},
            "put": {"price": 1.5760435147766145, "delta": -0.031725540191196444,
                    "gamma": -0.008683048910995694},
            "tolerance": 0.003,
            "crossover_sigmas": [0.20273255202675663, 0.20273255608140772, 0.3040988310811233],
            "source": "crates/kiyosi/tests/double_barrier_monte_carlo.rs:121-157, 232-276",
        },
    }, indent=1))
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
