"""Writes tests/golden/reference_values.json from the reference's own data files.

Sources (read as data, in a checkout of montepath):
  benchmarks/reference-fixtures.json -- closed-form / CRR / series fixtures
  benchmarks/release-results.json    -- published CPU benchmark rows

Each published row is a mean over seeds 42..44 (9303..9305 for Merton) of one
mc-core CPU pricer run on an 8-core Apple M2 with n_threads = 0 (all cores),
at MC_PATHS = 100_000 and MC_STEPS = 64 (crates/mc-bench/src/harness.rs:42-44
and the benchmark_mc_rust_cpu_* functions).  The configs below restate those
functions; the values come from the JSON files.

Usage: python tests/golden/make_reference_values.py [path/to/montepath]
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_values.json")

# benchmark row -> (pricer, config), from crates/mc-bench/src/harness.rs
ROWS = {
    "mc_cpu_european_call_rust": ("gbm_stepwise", {"payoff": "european_call"}),  # :1729-1771
    "mc_cpu_european_call_rust_terminal": ("european_terminal", {}),  # :2218-2260
    "mc_cpu_arithmetic_asian_call_rust": ("gbm_stepwise", {"payoff": "arithmetic_asian_call"}),  # :2522
    "mc_cpu_down_and_out_call_rust": ("gbm_stepwise", {"payoff": "down_and_out_call", "barrier": 80.0}),  # :3520
    "mc_cpu_lookback_call_rust": ("gbm_stepwise", {"payoff": "lookback_call"}),  # :3643
    "mc_cpu_heston_european_call_rust": (
        "heston", {"kappa": 1.5, "theta": 0.04, "v0": 0.04, "xi": 0.3, "rho": -0.6}),  # :3923
    "mc_cpu_merton_jump_diffusion_call_rust": (
        "merton", {"lam": 0.4, "mu_j": -0.08, "sigma_j": 0.25, "seed": 9303}),  # :4004
    "mc_cpu_basket_call_rust": (
        "basket2", {"s01": 100.0, "s02": 95.0, "sigma1": 0.2, "sigma2": 0.25, "rho": 0.35,
                    "weight1": 0.5, "weight2": 0.5}),  # :4456
    "mc_cpu_american_put_lsm_rust": ("american_lsm", {"basis_degree": 2}),  # :3730
    "mc_cpu_bermudan_put_lsm_rust": (
        "bermudan_lsm", {"basis_degree": 2, "exercise_steps": [16, 32, 48, 64]}),  # :3773
    # variance-reduction techniques (same defaults, technique set)
    "mc_cpu_european_call_rust_antithetic": (
        "gbm_stepwise", {"payoff": "european_call", "technique": "antithetic"}),  # :2387
    "mc_cpu_european_call_rust_control_variate": (
        "gbm_stepwise", {"payoff": "european_call", "technique": "control_variate"}),  # :2342
    "mc_cpu_european_call_rust_terminal_antithetic": (
        "european_terminal", {"technique": "antithetic"}),  # :2432
    "mc_cpu_european_call_rust_terminal_control_variate": (
        "european_terminal", {"technique": "control_variate"}),  # :2477
    "mc_cpu_arithmetic_asian_call_rust_control_variate": (
        "gbm_stepwise", {"payoff": "arithmetic_asian_call", "technique": "control_variate"}),  # :2566
    "mc_cpu_down_and_out_call_rust_control_variate": (
        "gbm_stepwise", {"payoff": "down_and_out_call", "barrier": 80.0,
                         "technique": "control_variate"}),  # :3563
    "mc_cpu_lookback_call_rust_control_variate": (
        "gbm_stepwise", {"payoff": "lookback_call", "technique": "control_variate"}),  # :3686
    # randomized Halton and Latin hypercube QMC (StructuredNormalSampler)
    "mc_cpu_european_call_rust_randomized_halton": (
        "gbm_structured", {"payoff": "european_call", "sampling": "randomized_halton"}),  # :3003
    "mc_cpu_european_call_rust_latin_hypercube": (
        "gbm_structured", {"payoff": "european_call", "sampling": "latin_hypercube"}),  # :3273
    "mc_cpu_arithmetic_asian_call_rust_randomized_halton": (
        "gbm_structured", {"payoff": "arithmetic_asian_call", "sampling": "randomized_halton"}),
    "mc_cpu_arithmetic_asian_call_rust_latin_hypercube": (
        "gbm_structured", {"payoff": "arithmetic_asian_call", "sampling": "latin_hypercube"}),
    "mc_cpu_down_and_out_call_rust_randomized_halton": (
        "gbm_structured", {"payoff": "down_and_out_call", "barrier": 80.0,
                           "sampling": "randomized_halton"}),
    "mc_cpu_down_and_out_call_rust_latin_hypercube": (
        "gbm_structured", {"payoff": "down_and_out_call", "barrier": 80.0,
                           "sampling": "latin_hypercube"}),
    "mc_cpu_basket_call_rust_randomized_halton": (
        "basket2_structured", {"s01": 100.0, "s02": 95.0, "sigma1": 0.2, "sigma2": 0.25, "rho": 0.35,
                               "weight1": 0.5, "weight2": 0.5, "sampling": "randomized_halton"}),
    "mc_cpu_basket_call_rust_latin_hypercube": (
        "basket2_structured", {"s01": 100.0, "s02": 95.0, "sigma1": 0.2, "sigma2": 0.25, "rho": 0.35,
                               "weight1": 0.5, "weight2": 0.5, "sampling": "latin_hypercube"}),
    # scrambled-Sobol rows: sobol_burley 0.5 bits, not reproducible without the crate
    "mc_cpu_european_call_rust_scrambled_sobol": ("gbm_qmc", {"bridge": False}),  # :3356
    "mc_cpu_european_call_rust_scrambled_sobol_brownian_bridge": ("gbm_qmc", {"bridge": True}),
    "mc_cpu_arithmetic_asian_call_rust_scrambled_sobol_brownian_bridge": (
        "gbm_qmc", {"bridge": True, "payoff": "arithmetic_asian_call"}),  # :218-223
    "mc_cpu_down_and_out_call_rust_scrambled_sobol_brownian_bridge": (
        "gbm_qmc", {"bridge": True, "payoff": "down_and_out_call", "barrier": 80.0}),  # :265-270
    "mc_cpu_basket_call_rust_scrambled_sobol": (
        "basket2_structured", {"s01": 100.0, "s02": 95.0, "sigma1": 0.2, "sigma2": 0.25, "rho": 0.35,
                               "weight1": 0.5, "weight2": 0.5, "sampling": "scrambled_sobol"}),  # :310-315
}

# stderr(technique) / stderr(standard) on one seed (harness.rs:2712-2891, :3607-3641)
QUALITY_ROWS = {
    "mc_cpu_european_call_rust_antithetic_quality": (
        "gbm_stepwise", {"payoff": "european_call", "technique": "antithetic", "seed": 7001}),
    "mc_cpu_european_call_rust_terminal_antithetic_quality": (
        "european_terminal", {"technique": "antithetic", "seed": 7002}),
    "mc_cpu_european_call_rust_control_variate_quality": (
        "gbm_stepwise", {"payoff": "european_call", "technique": "control_variate", "seed": 7003}),
    "mc_cpu_european_call_rust_terminal_control_variate_quality": (
        "european_terminal", {"technique": "control_variate", "seed": 7004}),
    "mc_cpu_arithmetic_asian_call_rust_control_variate_quality": (
        "gbm_stepwise", {"payoff": "arithmetic_asian_call", "technique": "control_variate",
                         "seed": 7111}),
    "mc_cpu_down_and_out_call_rust_control_variate_quality": (
        "gbm_stepwise", {"payoff": "down_and_out_call", "barrier": 80.0,
                         "technique": "control_variate", "seed": 7311}),
}

# |Delta - Black-Scholes Delta| (A&S CDF) at GREEK_PATHS = 50_000, GREEK_STEPS = 64
# (harness.rs:45-46, :4080-4184); the count row is 6 workloads' estimates
GREEK_ROWS = {
    "mc_cpu_european_call_greeks_bump_rust": (
        "european_greeks", {"estimator": "bump_and_revalue", "seed": 9401}),
    "mc_cpu_european_call_greeks_pathwise_rust": (
        "european_greeks", {"estimator": "pathwise", "seed": 9401}),
    "mc_cpu_european_call_greeks_likelihood_ratio_rust": (
        "european_greeks", {"estimator": "likelihood_ratio", "seed": 9401}),
    "mc_cpu_heston_greeks_black_scholes_limit_delta_quality": (
        "heston_greeks", {"estimator": "bump_and_revalue", "seed": 9402, "kappa": 1.5,
                          "theta": 0.04, "v0": 0.04, "xi": 0.0, "rho": 0.0}),
    "mc_cpu_all_workload_greeks_bump_rust": (
        "all_workload_greeks", {"estimator": "bump_and_revalue", "seed": 9403,
                                "n_paths": 25_000}),
}

# multilevel Asian (harness.rs:2613-2665, :2892-2980): the MLMC config is
# solved by solve_arithmetic_asian_mlmc_tolerance_cpu from a pilot at the row's
# seed (base 8 steps, 4 levels, refinement 2, 2048 pilot paths, target stderr
# 0.05, 1e5..2e6 step updates), then priced (seeds seed..seed+repeats-1); the
# quality and calibration rows compare with the standard Asian MC at
# 100_000 x 64 (n_threads 8) on the same seed
MLMC_ROWS = {
    "mc_cpu_arithmetic_asian_call_rust_mlmc": (
        "mlmc_price", {"sampling": "pseudorandom", "replicates": 1, "seed": 42, "repeats": 3}),
    "mc_cpu_arithmetic_asian_call_rust_mlmc_quality": (
        "mlmc_stderr_ratio", {"sampling": "pseudorandom", "replicates": 1, "seed": 7112}),
    "mc_cpu_arithmetic_asian_call_rust_mlmc_reference_calibration": (
        "mlmc_abs_error", {"sampling": "pseudorandom", "replicates": 1, "seed": 7114}),
    # scrambled-Sobol multilevel rows: sobol_burley bits, statistical only
    "mc_cpu_arithmetic_asian_call_rust_mlqmc": (
        "mlmc_price", {"sampling": "scrambled_sobol", "replicates": 4, "seed": 42, "repeats": 3}),
    "mc_cpu_arithmetic_asian_call_rust_mlqmc_quality": (
        "mlmc_stderr_ratio", {"sampling": "scrambled_sobol", "replicates": 4, "seed": 7113}),
    "mc_cpu_arithmetic_asian_call_rust_mlqmc_reference_calibration": (
        "mlmc_abs_error", {"sampling": "scrambled_sobol", "replicates": 4, "seed": 7115}),
}
MLMC_PLAN = {"base_steps": 8, "levels": 4, "refinement_factor": 2, "pilot_paths": 2_048,
             "target_stderr": 0.05, "min_step_updates": 100_000, "max_step_updates": 2_000_000}

COMMON = {"s0": 100.0, "k": 100.0, "r": 0.03, "sigma": 0.2, "t": 1.0, "n_paths": 100_000,
          "n_steps": 64, "seed": 42, "n_threads": 8, "repeats": 3}


def main() -> None:
    with open(os.path.join(REF, "benchmarks", "reference-fixtures.json")) as f:
        fixtures = json.load(f)
    with open(os.path.join(REF, "benchmarks", "release-results.json")) as f:
        release = json.load(f)
    fx = {}
    for item in fixtures["fixtures"]:
        fx[item["fixture_id"]] = {"parameters": item["parameters"], "expected": item["expected"]}
    rows = {}
    for r in release["results"]:
        name = r["benchmark_name"]
        if name in ROWS:
            kind, cfg = ROWS[name]
            full = dict(COMMON)
            full.update(cfg)
            rows[name] = {"pricer": kind, "config": full, "metric_name": r["metric_name"],
                          "published_value": r["metric_value"],
                          "per_iteration_us": r["per_iteration_us"]}
    quality = {}
    for r in release["results"]:
        name = r["benchmark_name"]
        if name in QUALITY_ROWS:
            kind, cfg = QUALITY_ROWS[name]
            full = dict(COMMON)
            full.update(cfg)
            full["repeats"] = 1
            quality[name] = {"pricer": kind, "config": full, "metric_name": r["metric_name"],
                             "published_value": r["metric_value"]}
    greeks = {}
    for r in release["results"]:
        name = r["benchmark_name"]
        if name in GREEK_ROWS:
            kind, cfg = GREEK_ROWS[name]
            full = dict(COMMON)
            full.update({"n_paths": 50_000, "repeats": 1})
            full.update(cfg)
            greeks[name] = {"pricer": kind, "config": full, "metric_name": r["metric_name"],
                            "published_value": r["metric_value"]}
    mlmc = {}
    for r in release["results"]:
        name = r["benchmark_name"]
        if name in MLMC_ROWS:
            kind, cfg = MLMC_ROWS[name]
            full = dict(MLMC_PLAN)
            full.update({"standard_paths": 100_000, "standard_steps": 64, "n_threads": 8,
                         "repeats": 1})
            full.update(cfg)
            mlmc[name] = {"pricer": kind, "config": full, "metric_name": r["metric_name"],
                          "published_value": r["metric_value"]}
    out = {
        "mlmc_rows": mlmc,
        "greek_rows": greeks,
        "source": "montepath benchmarks/reference-fixtures.json and benchmarks/release-results.json",
        "quality_rows": quality,
        "note": "published rows are means over seeds seed..seed+2 with n_threads=8",
        "fixtures": fx,
        "published_rows": rows,
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(OUT, len(fx), "fixtures", len(rows), "rows", len(quality), "quality rows", len(greeks),
          "Greek rows", len(mlmc), "MLMC rows")


if __name__ == "__main__":
    main()
