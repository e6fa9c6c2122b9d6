"""The mc-bench rows on the GPU: every row present and finite, the
benchmark-gates analogues pass, and the GPU throughput beats the reference's
published CPU rows."""
import json
import math
import os

import pytest

from montepath_amd import bench_rows as B

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_rows_and_gates(engine):
    report = B.run(engine)
    rows = report["results"]
    names = [r["benchmark_name"] for r in rows]
    assert len(names) == len(set(names)) and len(rows) >= 40
    for r in rows:
        assert r["backend"] == "rocm_hip" and math.isfinite(r["metric_value"]), r
    published = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))
    cpu = [{"benchmark_name": n, "throughput_per_sec": 1e6 / v["per_iteration_us"] * v["config"]["n_paths"]}
           for n, v in published["published_rows"].items() if "per_iteration_us" in v]
    gates = B.check_gates(rows, cpu)
    failed = [g for g in gates if not g["passed"]]
    assert not failed, failed
    by = {r["benchmark_name"]: r for r in rows}
    # price rows agree with the closed form / published CPU values statistically
    assert abs(by["mc_rocm_european_call_native"]["metric_value"] - 9.413403383853016) < 0.15
    ratio = by["mc_rocm_european_call_native_control_variate_quality"]["metric_value"]
    assert 0.35 < ratio < 0.47  # published 0.411
