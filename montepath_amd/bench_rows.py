"""mc-bench rows for the rocm_hip backend.

Writes a ``BenchmarkReport`` (crates/mc-bench/src/result.rs:1-23:
``generated_at_unix_ms`` and ``results`` of ``BenchmarkResult``) whose rows
mirror the reference harness's CPU and Metal rows (crates/mc-bench/src/
harness.rs) at the same configurations -- 100,000 paths x 64 steps, three
repeats on seeds seed..seed+2, the same quality seeds -- named ``mc_rocm_*``
with ``backend: "rocm_hip"``, and checks gates analogous to
docs/benchmark-gates.md.  Every price comes from the GPU through the C-ABI;
the closed forms below (Black-Scholes with the reference's A&S normal CDF,
the CRR tree, the Merton series) are the reference values the quality rows
are measured against, as in the reference harness.

Usage: python -m montepath_amd.bench_rows [--out report.json]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from typing import Callable

from . import _abi as A
from .engine import Engine, make_mlmc_config, make_request

MC_PATHS, MC_STEPS, MC_REPEATS = 100_000, 64, 3
GREEK_PATHS, GREEK_STEPS = 50_000, 64
IMPL = "montepath_amd::rocm_hip"


# ----------------------------------------------------------- closed forms
def norm_cdf(x: float) -> float:
    """standard_normal_cdf (cpu.rs:898-910): Abramowitz & Stegun 7.1.26."""
    sign = -1.0 if x < 0.0 else 1.0
    z = abs(x) / math.sqrt(2.0)
    t = 1.0 / (1.0 + 0.3275911 * z)
    poly = (((((1.061405429 * t - 1.453152027) * t) + 1.421413741) * t - 0.284496736) * t
            + 0.254829592) * t
    return 0.5 * (1.0 + sign * (1.0 - poly * math.exp(-z * z)))


def bs_call(s0, k, r, sigma, t) -> float:
    """black_scholes_european_call_price (cpu.rs:839-859)."""
    if sigma == 0.0:
        return max(s0 - k * math.exp(-r * t), 0.0)
    d1 = (math.log(s0 / k) + (r + 0.5 * sigma * sigma) * t) / (sigma * math.sqrt(t))
    d2 = d1 - sigma * math.sqrt(t)
    return s0 * norm_cdf(d1) - k * math.exp(-r * t) * norm_cdf(d2)


def bs_delta(s0, k, r, sigma, t) -> float:
    """Delta of black_scholes_european_call_greeks (cpu.rs:866-892)."""
    d1 = (math.log(s0 / k) + (r + 0.5 * sigma * sigma) * t) / (sigma * math.sqrt(t))
    return norm_cdf(d1)


def crr_put(s0, k, r, sigma, t, steps, exercise) -> float:
    """binomial_put_reference_price (cpu.rs:5244-5305): CRR tree, exercise at
    the given tree steps and at maturity."""
    dt = t / steps
    u = math.exp(sigma * math.sqrt(dt))
    d = 1.0 / u
    p = (math.exp(r * dt) - d) / (u - d)
    disc = math.exp(-r * dt)
    mask = [False] * (steps + 1)
    for s in exercise:
        if s <= steps:
            mask[s] = True
    mask[steps] = True
    v = [max(k - s0 * u ** (steps - j) * d ** j, 0.0) for j in range(steps + 1)]
    for step in range(steps - 1, -1, -1):
        for j in range(step + 1):
            cont = disc * (p * v[j] + (1.0 - p) * v[j + 1])
            v[j] = max(cont, k - s0 * u ** (step - j) * d ** j) if mask[step] else cont
    return v[0]


def merton_series(s0, k, r, sigma, t, lam, mu_j, sigma_j, max_terms=96, tail=1e-12) -> float:
    """merton_jump_diffusion_call_reference_price (cpu.rs:3110-3149)."""
    lt = lam * t
    if lt == 0.0:
        return bs_call(s0, k, r, sigma, t)
    comp = lam * (math.exp(mu_j + 0.5 * sigma_j ** 2) - 1.0)
    base = math.log(s0) + (r - comp - 0.5 * sigma * sigma) * t
    w, cum, price = math.exp(-lt), 0.0, 0.0
    for n in range(max_terms):
        mean = base + n * mu_j
        var = sigma * sigma * t + n * sigma_j * sigma_j
        sd = math.sqrt(var)
        d2 = (mean - math.log(k)) / sd
        price += w * math.exp(-r * t) * (math.exp(mean + 0.5 * var) * norm_cdf(d2 + sd)
                                         - k * norm_cdf(d2))
        cum += w
        if max(1.0 - cum, 0.0) <= tail:
            break
        w *= lt / (n + 1)
    return price


# ------------------------------------------------------------------- rows
def result_row(name, implementation, methodology, iterations, total_ms, work, metric_name,
               metric_value) -> dict:
    """One BenchmarkResult (result.rs:3-17); throughput = work units / s."""
    per_ms = total_ms / iterations if iterations else 0.0
    return {
        "benchmark_name": name, "benchmark_version": "0.1", "implementation": implementation,
        "backend": "rocm_hip", "methodology": methodology, "planner_mode": "n/a",
        "iterations": iterations, "total_runtime_ms": total_ms,
        "per_iteration_us": per_ms * 1000.0,
        "throughput_per_sec": work if per_ms == 0.0 else work / (per_ms / 1000.0),
        "metric_name": metric_name, "metric_value": metric_value,
    }


def _timed(fn: Callable):
    t0 = time.perf_counter()
    out = fn()
    return out, (time.perf_counter() - t0) * 1000.0


class Rows:
    def __init__(self, eng: Engine, precision: int = 32):
        self.eng = eng
        self.precision = precision
        self.rows: list = []

    def _price(self, workload, **kw):
        kw.setdefault("precision", self.precision)
        return self.eng.price(make_request(workload, **kw))

    def repeated(self, name, methodology, workload, seed=42, repeats=MC_REPEATS, **kw):
        """Mean price over seeds seed..seed+repeats-1 (the harness's repeat loop),
        after one untimed call (code-object load and first allocations)."""
        self._price(workload, seed=seed + repeats, **kw)
        prices, ms = [], 0.0
        for i in range(repeats):
            res, dt = _timed(lambda: self._price(workload, seed=seed + i, **kw))
            prices.append(res.price)
            ms += dt
        n = kw.get("n_paths", MC_PATHS)
        self.rows.append(result_row(name, f"{IMPL}::mp_hip_price", methodology, repeats, ms, n,
                                    "price_estimate", sum(prices) / len(prices)))

    def quality(self, name, methodology, workload, seed, technique, **kw):
        """stderr(technique) / stderr(standard) on one seed."""
        std = self._price(workload, seed=seed, **kw)
        tech = self._price(workload, seed=seed, technique=technique, **kw)
        self.rows.append(result_row(name, f"{IMPL}::mp_hip_price", methodology, 1, 0.0, 0.0,
                                    "stderr_ratio_vs_standard",
                                    1.0 if std.stderr == 0.0 else tech.stderr / std.stderr))

    def abs_error(self, name, methodology, metric, fn, reference, n_paths):
        res, ms = _timed(fn)
        self.rows.append(result_row(name, f"{IMPL}::mp_hip_price", methodology, 1, ms, n_paths,
                                    metric, abs(res.price - reference)))


def pricing_rows(r: Rows) -> None:
    gbm = dict(n_paths=MC_PATHS, n_steps=MC_STEPS)
    r.repeated("mc_rocm_european_call_native", "stepwise_paths_native_rocm", "european_call", **gbm)
    for tech in ("antithetic", "control_variate"):
        r.repeated(f"mc_rocm_european_call_native_{tech}", f"stepwise_paths_native_rocm_{tech}",
                   "european_call", technique=tech, **gbm)
        r.repeated(f"mc_rocm_european_call_native_terminal_{tech}",
                   f"terminal_distribution_native_rocm_{tech}", "european_call", method="terminal",
                   technique=tech, **gbm)
    r.quality("mc_rocm_european_call_native_antithetic_quality",
              "stepwise_paths_native_rocm_antithetic", "european_call", 7001, "antithetic", **gbm)
    r.quality("mc_rocm_european_call_native_terminal_antithetic_quality",
              "terminal_distribution_native_rocm_antithetic", "european_call", 7002, "antithetic",
              method="terminal", **gbm)
    r.quality("mc_rocm_european_call_native_control_variate_quality",
              "stepwise_paths_native_rocm_control_variate", "european_call", 7003,
              "control_variate", **gbm)
    r.quality("mc_rocm_european_call_native_terminal_control_variate_quality",
              "terminal_distribution_native_rocm_control_variate", "european_call", 7004,
              "control_variate", method="terminal", **gbm)
    r.repeated("mc_rocm_european_call_native_terminal", "terminal_distribution_native_rocm",
               "european_call", method="terminal", **gbm)
    for smp in ("randomized_halton", "latin_hypercube", "scrambled_sobol",
                "scrambled_sobol_brownian_bridge"):
        r.repeated(f"mc_rocm_european_call_native_{smp}", f"stepwise_paths_native_rocm_{smp}",
                   "european_call", sampling=smp, **gbm)
    r.repeated("mc_rocm_arithmetic_asian_call_native", "arithmetic_asian_stepwise_native_rocm",
               "arithmetic_asian_call", **gbm)
    r.repeated("mc_rocm_arithmetic_asian_call_native_control_variate",
               "arithmetic_asian_stepwise_native_rocm_control_variate", "arithmetic_asian_call",
               technique="control_variate", **gbm)
    r.quality("mc_rocm_arithmetic_asian_call_native_control_variate_quality",
              "arithmetic_asian_stepwise_native_rocm_control_variate", "arithmetic_asian_call",
              7111, "control_variate", **gbm)
    r.repeated("mc_rocm_down_and_out_call_native", "down_and_out_stepwise_native_rocm",
               "down_and_out_call", barrier=80.0, **gbm)
    r.repeated("mc_rocm_down_and_out_call_native_control_variate",
               "down_and_out_stepwise_native_rocm_control_variate", "down_and_out_call",
               barrier=80.0, technique="control_variate", **gbm)
    r.quality("mc_rocm_down_and_out_call_native_control_variate_quality",
              "down_and_out_stepwise_native_rocm_control_variate", "down_and_out_call", 7311,
              "control_variate", barrier=80.0, **gbm)
    r.repeated("mc_rocm_lookback_call_native", "lookback_stepwise_native_rocm", "lookback_call",
               **gbm)
    r.repeated("mc_rocm_lookback_call_native_control_variate",
               "lookback_stepwise_native_rocm_control_variate", "lookback_call",
               technique="control_variate", **gbm)
    heston = dict(kappa=1.5, theta=0.04, v0=0.04, xi=0.3, rho=-0.6)
    r.repeated("mc_rocm_heston_european_call_native", "heston_stepwise_native_rocm",
               "heston_european_call", **gbm, **heston)
    limit = dict(kappa=1.5, theta=0.04, v0=0.04, xi=0.0, rho=0.0)
    r.abs_error("mc_rocm_heston_black_scholes_limit_quality",
                "heston_black_scholes_vol_of_vol_zero_limit", "abs_error_vs_black_scholes",
                lambda: r._price("heston_european_call", seed=9301, **gbm, **limit),
                bs_call(100.0, 100.0, 0.03, 0.2, 1.0), MC_PATHS)
    merton = dict(jump_intensity=0.4, jump_mean=-0.08, jump_volatility=0.25)
    r.repeated("mc_rocm_merton_jump_diffusion_call_native", "merton_jump_diffusion_native_rocm",
               "merton_jump_diffusion_call", seed=9303, n_paths=MC_PATHS, **merton)
    r.abs_error("mc_rocm_merton_jump_diffusion_reference_quality",
                "merton_jump_diffusion_merton_series_reference", "abs_error_vs_merton_series",
                lambda: r._price("merton_jump_diffusion_call", seed=9304, n_paths=MC_PATHS,
                                 **merton),
                merton_series(100.0, 100.0, 0.03, 0.2, 1.0, 0.4, -0.08, 0.25), MC_PATHS)
    basket = dict(n_paths=MC_PATHS, n_steps=1, asset_s0=[100.0, 95.0], asset_sigma=[0.2, 0.25],
                  asset_weight=[0.5, 0.5], correlation=[[1.0, 0.35], [0.35, 1.0]])
    r.repeated("mc_rocm_basket_call_native", "basket_terminal_native_rocm", "basket_call",
               **basket)
    for smp in ("randomized_halton", "latin_hypercube"):
        r.repeated(f"mc_rocm_basket_call_native_{smp}", f"basket_terminal_native_rocm_{smp}",
                   "basket_call", sampling=smp, **basket)
    lsm = dict(n_paths=MC_PATHS, n_steps=MC_STEPS, basis_degree=2, precision=64)
    r.repeated("mc_rocm_american_put_lsm_native", "american_put_longstaff_schwartz_laguerre_rocm",
               "american_put", **lsm)
    r.repeated("mc_rocm_bermudan_put_lsm_native",
               "bermudan_put_longstaff_schwartz_laguerre_rocm", "bermudan_put",
               exercise_steps=[16, 32, 48, 64], **lsm)
    r.abs_error("mc_rocm_american_put_lsm_binomial_reference_quality",
                "american_put_lsm_vs_crr_binomial_reference", "abs_error_vs_binomial_reference",
                lambda: r._price("american_put", seed=14_007, **lsm),
                crr_put(100.0, 100.0, 0.03, 0.2, 1.0, 512, range(513)), MC_PATHS)
    sched = sorted({min(max(int(math.floor(s / 64 * 512 + 0.5)), 1), 512)
                    for s in (16, 32, 48, 64)} | {512})
    r.abs_error("mc_rocm_bermudan_put_lsm_binomial_reference_quality",
                "bermudan_put_lsm_vs_crr_binomial_reference", "abs_error_vs_binomial_reference",
                lambda: r._price("bermudan_put", seed=14_008, exercise_steps=[16, 32, 48, 64],
                                 **lsm),
                crr_put(100.0, 100.0, 0.03, 0.2, 1.0, 512, sched), MC_PATHS)


def greek_rows(r: Rows) -> None:
    eng = r.eng
    delta = bs_delta(100.0, 100.0, 0.03, 0.2, 1.0)
    for est in ("bump_and_revalue", "pathwise", "likelihood_ratio"):
        req = make_request("european_call", n_paths=GREEK_PATHS, n_steps=GREEK_STEPS, seed=9401,
                           method="terminal", precision=r.precision)
        (base, ests), ms = _timed(lambda: eng.greeks(req, est))
        d = next(e["value"] for e in ests if e["greek"] == "delta")
        short = "bump" if est == "bump_and_revalue" else est
        r.rows.append(result_row(f"mc_rocm_european_call_greeks_{short}", f"{IMPL}::mp_hip_greeks",
                                 f"european_call_greeks_{short}", 1, ms, GREEK_PATHS,
                                 "abs_delta_error_vs_black_scholes", abs(d - delta)))
    req = make_request("heston_european_call", n_paths=GREEK_PATHS, n_steps=GREEK_STEPS,
                       seed=9402, kappa=1.5, theta=0.04, v0=0.04, xi=0.0, rho=0.0,
                       precision=r.precision)
    (base, ests), ms = _timed(lambda: eng.greeks(req))
    r.rows.append(result_row("mc_rocm_heston_greeks_black_scholes_limit_delta_quality",
                             f"{IMPL}::mp_hip_greeks", "heston_greeks_bump_black_scholes_limit", 1,
                             ms, GREEK_PATHS, "abs_delta_error_vs_black_scholes",
                             abs(ests[0]["value"] - delta)))
    from . import native

    reps, ms = _timed(lambda: native.price_all_current_greeks_bump_and_revalue_rocm(
        GREEK_PATHS // 2, GREEK_STEPS, 9403))
    count = sum(len(x["estimates"]) for x in reps)
    r.rows.append(result_row("mc_rocm_all_workload_greeks_bump", f"{IMPL}::mp_hip_greeks",
                             "all_current_workload_greeks_bump_and_revalue", 1, ms, count,
                             "greek_estimate_count", float(count)))


def _adaptive_mlmc(eng: Engine, sampling: str, replicates: int, seed: int):
    """adaptive_arithmetic_asian_mlmc_config (harness.rs:2635-2665), solved on the GPU."""
    cfg = make_mlmc_config(base_steps=8, levels=4, refinement_factor=2,
                           paths_per_level=[2048] * 4, seed=seed, sampling=sampling,
                           scramble_replicates=replicates)
    plan = eng.mlmc_solve_tolerance(cfg, 0.05, 2048, 100_000, 2_000_000)
    return plan["recommended_config"]


def mlmc_rows(r: Rows) -> None:
    eng = r.eng
    for smp, reps, tag in (("pseudorandom", 1, "mlmc"), ("scrambled_sobol", 4, "mlqmc")):
        cfg = _adaptive_mlmc(eng, smp, reps, 42)
        prices, ms, updates = [], 0.0, 0
        for i in range(MC_REPEATS):
            cfg.seed = 42 + i
            res, dt = _timed(lambda: eng.mlmc_price(cfg))
            prices.append(res["price"])
            ms += dt
            updates = res["total_step_updates"]
        r.rows.append(result_row(f"mc_rocm_arithmetic_asian_call_native_{tag}",
                                 f"{IMPL}::mp_hip_mlmc_price",
                                 "arithmetic_asian_multilevel_coupled_adaptive_tolerance",
                                 MC_REPEATS, ms, updates, "price_estimate",
                                 sum(prices) / len(prices)))
        qseed, cseed = (7112, 7114) if tag == "mlmc" else (7113, 7115)
        std = r._price("arithmetic_asian_call", n_paths=MC_PATHS, n_steps=MC_STEPS, seed=qseed)
        res = eng.mlmc_price(_adaptive_mlmc(eng, smp, reps, qseed))
        r.rows.append(result_row(f"mc_rocm_arithmetic_asian_call_native_{tag}_quality",
                                 f"{IMPL}::mp_hip_mlmc_price",
                                 "arithmetic_asian_multilevel_coupled_adaptive_tolerance", 1, 0.0,
                                 res["total_step_updates"], "stderr_ratio_vs_standard",
                                 res["stderr"] / std.stderr))
        ref = r._price("arithmetic_asian_call", n_paths=MC_PATHS, n_steps=MC_STEPS, seed=cseed)
        res, ms = _timed(lambda: eng.mlmc_price(_adaptive_mlmc(eng, smp, reps, cseed)))
        r.rows.append(result_row(f"mc_rocm_arithmetic_asian_call_native_{tag}_reference_calibration",
                                 f"{IMPL}::mp_hip_mlmc_price",
                                 "arithmetic_asian_mlmc_realized_error_vs_high_budget_standard_mc",
                                 1, ms, res["total_step_updates"] + MC_PATHS * MC_STEPS,
                                 "abs_error_vs_standard_reference", abs(res["price"] - ref.price)))


# ------------------------------------------------------------------ gates
REQUIRED = [
    "mc_rocm_european_call_native", "mc_rocm_down_and_out_call_native",
    "mc_rocm_lookback_call_native", "mc_rocm_american_put_lsm_native",
    "mc_rocm_bermudan_put_lsm_native", "mc_rocm_heston_european_call_native",
    "mc_rocm_merton_jump_diffusion_call_native",
]


def check_gates(results: list, cpu_results: list | None = None) -> list:
    """Gates analogous to docs/benchmark-gates.md items 4, 7-11 and 14, plus
    GPU-over-CPU throughput when the reference's CPU rows are given."""
    by = {x["benchmark_name"]: x for x in results}
    gates = []

    def gate(name, ok, detail):
        gates.append({"gate": name, "passed": bool(ok), "detail": detail})

    missing = [n for n in REQUIRED if n not in by]
    gate("required_rows_present", not missing, f"missing: {missing}")

    def metric_below(row, bound):
        v = by.get(row, {}).get("metric_value")
        ok = v is not None and math.isfinite(v) and 0.0 <= v < bound
        gate(f"{row} < {bound}", ok, f"value {v}")

    metric_below("mc_rocm_heston_black_scholes_limit_quality", 0.5)
    for est in ("bump", "pathwise", "likelihood_ratio"):
        metric_below(f"mc_rocm_european_call_greeks_{est}", 0.08)
    cnt = by.get("mc_rocm_all_workload_greeks_bump", {}).get("metric_value")
    gate("mc_rocm_all_workload_greeks_bump >= 24", cnt is not None and cnt >= 24, f"count {cnt}")
    metric_below("mc_rocm_american_put_lsm_binomial_reference_quality", 0.75)
    metric_below("mc_rocm_bermudan_put_lsm_binomial_reference_quality", 0.75)
    metric_below("mc_rocm_merton_jump_diffusion_reference_quality", 0.5)
    metric_below("mc_rocm_arithmetic_asian_call_native_mlmc_reference_calibration", 0.75)
    metric_below("mc_rocm_arithmetic_asian_call_native_mlqmc_reference_calibration", 0.75)
    if cpu_results:
        cpu = {x["benchmark_name"]: x for x in cpu_results}
        for g, c in (("mc_rocm_european_call_native", "mc_cpu_european_call_rust"),
                     ("mc_rocm_heston_european_call_native", "mc_cpu_heston_european_call_rust"),
                     ("mc_rocm_american_put_lsm_native", "mc_cpu_american_put_lsm_rust")):
            if g in by and c in cpu:
                gt, ct = by[g]["throughput_per_sec"], cpu[c]["throughput_per_sec"]
                gate(f"{g} throughput > {c}", gt > ct, f"{gt:.4g} vs {ct:.4g} paths/s")
    return gates


def run(eng: Engine | None = None, precision: int = 32) -> dict:
    eng = eng or Engine(0)
    r = Rows(eng, precision)
    pricing_rows(r)
    greek_rows(r)
    mlmc_rows(r)
    return {"generated_at_unix_ms": int(time.time() * 1000), "results": r.rows}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=None, help="write the BenchmarkReport JSON here")
    ap.add_argument("--cpu-report", default=None,
                    help="a reference BenchmarkReport (e.g. release-results.json) for the "
                         "GPU-over-CPU throughput gates")
    ap.add_argument("--precision", type=int, default=32, choices=[32, 64])
    args = ap.parse_args(argv)
    try:
        report = run(precision=args.precision)
    except A.MontepathHipError as err:
        print(f"rocm_hip benchmark failed: {err}", file=sys.stderr)
        return 2
    cpu = None
    if args.cpu_report:
        with open(args.cpu_report) as f:
            cpu = json.load(f)["results"]
    gates = check_gates(report["results"], cpu)
    report["gates"] = gates
    text = json.dumps(report, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    else:
        print(text)
    failed = [g for g in gates if not g["passed"]]
    for g in gates:
        print(f"[gate] {'PASS' if g['passed'] else 'FAIL'} {g['gate']}: {g['detail']}",
              file=sys.stderr)
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
