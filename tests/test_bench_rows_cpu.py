"""The mc-bench row writer's host parts: BenchmarkResult fields, gates, and
the closed forms its quality rows are measured against (pinned by the
oracle's restatements and the reference fixtures)."""
import math

import pytest

from montepath_amd import bench_rows as B

FIELDS = ["benchmark_name", "benchmark_version", "implementation", "backend", "methodology",
          "planner_mode", "iterations", "total_runtime_ms", "per_iteration_us",
          "throughput_per_sec", "metric_name", "metric_value"]


def test_result_row_has_the_reference_schema():
    r = B.result_row("mc_rocm_x", "impl", "meth", 3, 6.0, 100_000, "price_estimate", 9.4)
    assert list(r) == FIELDS  # result.rs:3-17, in order
    assert r["backend"] == "rocm_hip" and r["per_iteration_us"] == 2000.0
    assert r["throughput_per_sec"] == 100_000 / 0.002
    z = B.result_row("q", "impl", "meth", 1, 0.0, 0.0, "stderr_ratio_vs_standard", 0.4)
    assert z["throughput_per_sec"] == 0.0


def test_closed_forms_match_the_oracle(oracle):
    assert B.bs_call(100.0, 100.0, 0.03, 0.2, 1.0) == pytest.approx(
        oracle.lib().mpo_bs_call_ref(100.0, 100.0, 0.03, 0.2, 1.0), rel=1e-14)
    assert B.bs_delta(100.0, 100.0, 0.03, 0.2, 1.0) == pytest.approx(
        oracle.bs_greeks(100.0, 100.0, 0.03, 0.2, 1.0)["delta"], rel=1e-14)
    # the reference fixture: CRR American put at 512 steps (reference-fixtures.json)
    assert B.crr_put(100.0, 100.0, 0.03, 0.2, 1.0, 512, range(513)) == pytest.approx(
        6.741130672775123, rel=1e-12)
    want = oracle.merton_series(oracle.Merton(lam=0.4, mu_j=-0.08, sigma_j=0.25), exact=False)
    assert B.merton_series(100.0, 100.0, 0.03, 0.2, 1.0, 0.4, -0.08, 0.25) == pytest.approx(
        want, rel=1e-12)


def _rows(**over):
    vals = {
        "mc_rocm_heston_black_scholes_limit_quality": 0.04,
        "mc_rocm_european_call_greeks_bump": 0.001, "mc_rocm_european_call_greeks_pathwise": 0.001,
        "mc_rocm_european_call_greeks_likelihood_ratio": 0.003,
        "mc_rocm_all_workload_greeks_bump": 26.0,
        "mc_rocm_american_put_lsm_binomial_reference_quality": 0.05,
        "mc_rocm_bermudan_put_lsm_binomial_reference_quality": 0.05,
        "mc_rocm_merton_jump_diffusion_reference_quality": 0.01,
        "mc_rocm_arithmetic_asian_call_native_mlmc_reference_calibration": 0.02,
        "mc_rocm_arithmetic_asian_call_native_mlqmc_reference_calibration": 0.02,
    }
    vals.update(over)
    rows = [B.result_row(n, "i", "m", 1, 1.0, 1e6, "price_estimate", 1.0) for n in B.REQUIRED]
    rows += [B.result_row(n, "i", "m", 1, 1.0, 1.0, "x", v) for n, v in vals.items()]
    return rows


def test_gates_pass_and_fail():
    assert all(g["passed"] for g in B.check_gates(_rows()))
    bad = B.check_gates(_rows(**{"mc_rocm_all_workload_greeks_bump": 20.0,
                                 "mc_rocm_merton_jump_diffusion_reference_quality": math.nan}))
    failed = {g["gate"] for g in bad if not g["passed"]}
    assert failed == {"mc_rocm_all_workload_greeks_bump >= 24",
                      "mc_rocm_merton_jump_diffusion_reference_quality < 0.5"}
    rows = [r for r in _rows() if r["benchmark_name"] != "mc_rocm_lookback_call_native"]
    assert not B.check_gates(rows)[0]["passed"]
    cpu = [{"benchmark_name": "mc_cpu_european_call_rust", "throughput_per_sec": 2e9}]
    g = [x for x in B.check_gates(_rows(), cpu) if "throughput" in x["gate"]]
    assert len(g) == 1 and not g[0]["passed"]  # 1e9 paths/s GPU rows vs 2e9 CPU
