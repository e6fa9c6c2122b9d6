"""One-launch Longstaff-Schwartz pricing against the per-date launches.

The one-launch kernel keeps each thread's cash flows in registers and meets
the other blocks through bounded waits on write-once slots per exercise
date; the
per-date path (one launch per date) is pinned to the oracle elsewhere.  The
two group the FP64 Gram sums differently, so the bar is: the same number of
regression dates, prices within 1e-10 relative, standard errors within
1e-8 relative, at least 99.99% identical exercise dates per path, and
bit-identical results when the one-launch pricing is repeated.
"""
import os

import numpy as np
import pytest

from montepath_amd.engine import Engine, make_request

pytestmark = pytest.mark.gpu

LSM = dict(basis_degree=3, r=0.05, sigma=0.25, k=100.0, t=1.0)


@pytest.fixture(scope="module")
def per_date_engine():
    """A context opened with the per-date launches selected."""
    os.environ["MONTEPATH_LSM_MULTI_LAUNCH"] = "1"
    try:
        eng = Engine(0)
    finally:
        del os.environ["MONTEPATH_LSM_MULTI_LAUNCH"]
    yield eng
    eng.close()


CASES = [
    # (workload, paths, dates, precision, extra) -- paths per thread 8, 4, 1 and a
    # Bermudan schedule
    ("american_put", 1 << 20, 50, 32, {}),
    ("american_put", 300_000, 50, 64, {}),
    ("american_put", 5_000, 16, 32, {}),
    ("bermudan_put", 100_000, 64, 64, {"exercise_steps": [8, 16, 24, 32, 40, 48, 56, 64]}),
]


@pytest.mark.parametrize("kind,n,m,prec,extra", CASES)
def test_one_launch_equals_per_date_launches(engine, per_date_engine, kind, n, m, prec, extra):
    print(f"[lsm-fused] {kind} n={n} m={m} fp{prec}", flush=True)
    req = make_request(kind, n_paths=n, n_steps=m, precision=prec, seed=5, **LSM, **extra)
    a = engine.price(req)
    launches_a = engine.last_timing()["main_launches"]
    b = per_date_engine.price(req)
    launches_b = per_date_engine.last_timing()["main_launches"]
    # one launch plus two fills, against a launch per date
    assert launches_a == 3 and launches_b > 3
    assert a.regression_steps == b.regression_steps
    assert abs(a.price - b.price) <= 1e-10 * b.price
    assert abs(a.stderr - b.stderr) <= 1e-8 * b.stderr
    assert abs(a.early_exercise_count - b.early_exercise_count) <= 1e-4 * n
    again = engine.price(req)
    assert (again.price, again.stderr, again.early_exercise_count) == (
        a.price, a.stderr, a.early_exercise_count)


def test_one_launch_exercise_dates_match_on_identical_normals(engine, per_date_engine):
    n, m = 1 << 18, 16
    print(f"[lsm-fused] normals n={n} m={m} fp64", flush=True)
    z = np.random.default_rng(11).standard_normal((m, n))
    req = make_request("american_put", n_paths=n, n_steps=m, precision=64, seed=1, **LSM)
    a, coeff_a, ex_a = engine.lsm_from_normals(req, z)
    assert engine.last_timing()["main_launches"] == 3
    b, coeff_b, ex_b = per_date_engine.lsm_from_normals(req, z)
    assert a.regression_steps == b.regression_steps == m - 1
    assert abs(a.price - b.price) <= 1e-10 * b.price
    assert np.mean(ex_a == ex_b) >= 0.9999
    assert np.all((ex_a >= 1) & (ex_a <= m))
