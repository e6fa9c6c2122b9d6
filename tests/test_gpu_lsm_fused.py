"""One-launch Longstaff-Schwartz pricing against the per-date launches.

The one-launch kernel keeps each thread's cash flows in registers and meets
the other blocks through bounded waits on write-once slots per exercise
date.  The matched per-date launches (what the server uses) run on the same
grid with the same per-path arithmetic and reduce the block sums in the
same order, so they must give the same bits: price, standard error,
counts, coefficients and every path's exercise date.  The plain per-date
launches (the only form for requests too large for one launch) group the
FP64 Gram sums differently: same regression count, prices within 1e-10
relative, at least 99.99% identical exercise dates, and the same agreement
with the oracle as the parity tests ask of the one-launch form.
"""
import os

import numpy as np
import pytest

from montepath_amd.engine import Engine, make_request

pytestmark = pytest.mark.gpu

LSM = dict(basis_degree=3, r=0.05, sigma=0.25, k=100.0, t=1.0)


def _engine_with(mode: str) -> Engine:
    os.environ["MONTEPATH_LSM_MULTI_LAUNCH"] = mode
    try:
        return Engine(0)
    finally:
        del os.environ["MONTEPATH_LSM_MULTI_LAUNCH"]


@pytest.fixture(scope="module")
def matched_engine():
    """A context opened with the matched per-date launches."""
    eng = _engine_with("1")
    yield eng
    eng.close()


@pytest.fixture(scope="module")
def plain_engine():
    """A context opened with the plain grid-stride per-date launches."""
    eng = _engine_with("2")
    yield eng
    eng.close()


CASES = [
    # (workload, paths, dates, precision, extra) -- paths per thread 4, 2, 1 and a
    # Bermudan schedule
    ("american_put", 1 << 20, 50, 32, {}),
    ("american_put", 300_000, 50, 64, {}),
    ("american_put", 5_000, 16, 32, {}),
    ("bermudan_put", 100_000, 64, 64, {"exercise_steps": [8, 16, 24, 32, 40, 48, 56, 64]}),
    # FP32 with path counts that leave idle lanes and strikes that FP32 does
    # not represent (float(k) < k): idle lanes must stay out of the regression
    ("american_put", 10_000, 50, 32, {"k": 116.10130597591788}),
    ("american_put", 22_725, 50, 32, {"k": 119.84186}),
    ("american_put", 1_047_001, 50, 32, {"k": 84.8445}),
    ("american_put", 153_242, 50, 32, {"k": 97.218}),
    # the weighted Laguerre basis e^{-x/2} L_i(x)
    ("american_put", 1 << 20, 50, 32, {"basis_weighted": True}),
    ("bermudan_put", 100_000, 64, 64, {"exercise_steps": [8, 16, 24, 32, 40, 48, 56, 64],
                                       "basis_weighted": True}),
]


@pytest.mark.parametrize("kind,n,m,prec,extra", CASES)
def test_one_launch_equals_matched_per_date_launches(engine, matched_engine, plain_engine, kind, n, m,
                                                     prec, extra):
    print(f"[lsm] {kind} n={n} m={m} fp{prec} {sorted(extra)}", flush=True)
    req = make_request(kind, n_paths=n, n_steps=m, precision=prec, seed=5, **{**LSM, **extra})
    a = engine.price(req)
    launches_a = engine.last_timing()["main_launches"]
    b = matched_engine.price(req)
    launches_b = matched_engine.last_timing()["main_launches"]
    # one launch plus two fills, against two launches per date
    assert launches_a == 3 and launches_b > 3
    assert (a.price, a.stderr, a.early_exercise_count, a.maturity_exercise_count, a.regression_steps) == (
        b.price, b.stderr, b.early_exercise_count, b.maturity_exercise_count, b.regression_steps)
    again = engine.price(req)
    assert (again.price, again.stderr, again.early_exercise_count) == (
        a.price, a.stderr, a.early_exercise_count)
    c = plain_engine.price(req)
    assert c.regression_steps == a.regression_steps
    assert abs(c.price - a.price) <= 1e-10 * a.price
    assert abs(c.stderr - a.stderr) <= 1e-8 * a.stderr
    assert abs(c.early_exercise_count - a.early_exercise_count) <= 1e-4 * n


def test_one_launch_and_matched_agree_path_by_path_on_identical_normals(engine, matched_engine, plain_engine):
    n, m = 1 << 18, 16
    print(f"[lsm] normals n={n} m={m} fp64", flush=True)
    z = np.random.default_rng(11).standard_normal((m, n))
    req = make_request("american_put", n_paths=n, n_steps=m, precision=64, seed=1, **LSM)
    a, coeff_a, ex_a = engine.lsm_from_normals(req, z)
    assert engine.last_timing()["main_launches"] == 3
    b, coeff_b, ex_b = matched_engine.lsm_from_normals(req, z)
    assert a.price == b.price and a.stderr == b.stderr
    assert np.array_equal(coeff_a, coeff_b)
    assert np.array_equal(ex_a, ex_b)
    assert a.regression_steps == m - 1
    assert np.all((ex_a >= 1) & (ex_a <= m))
    c, _, ex_c = plain_engine.lsm_from_normals(req, z)
    assert c.regression_steps == a.regression_steps
    assert abs(c.price - a.price) <= 1e-10 * a.price
    assert np.mean(ex_c == ex_a) >= 0.9999


@pytest.mark.parametrize("weighted", [False, True])
def test_plain_per_date_launches_match_oracle(plain_engine, oracle, weighted):
    n, m = 20000, 16
    cfg = oracle.Lsm(n_paths=n, n_steps=m, seed=42, basis_degree=3, r=0.05, sigma=0.25,
                     basis_weighted=weighted)
    z = oracle.rng_normals(42, n * m).reshape(n, m)
    ref = oracle.lsm_from_paths(cfg, oracle.lsm_paths_from_normals(cfg, z), want_detail=True)
    req = make_request("american_put", n_paths=n, n_steps=m, precision=64, basis_degree=3,
                       r=0.05, sigma=0.25, basis_weighted=weighted)
    res, _, ex = plain_engine.lsm_from_normals(req, z.T.copy())
    assert plain_engine.last_timing()["main_launches"] > 3
    assert res.regression_steps == ref["regression_steps"]
    assert abs(res.price - ref["price"]) <= 1e-10 * ref["price"]
    assert abs(res.stderr - ref["stderr"]) <= 1e-8 * ref["stderr"]
    assert np.mean(ex == ref["exercise_steps"]) >= 0.9999


def test_one_launch_timeout_falls_back_to_per_date_launches(engine):
    # A context whose one-launch waits give up after a single poll (forced
    # from the host; every block records the timeout and runs to its end):
    # the pricing is repeated with the matched per-date launches, which give
    # the same bits as an undisturbed one-launch pricing.
    os.environ["MONTEPATH_LSM_SPIN_LIMIT"] = "1"
    try:
        eng = Engine(0)
    finally:
        del os.environ["MONTEPATH_LSM_SPIN_LIMIT"]
    try:
        for n, prec in ((1 << 18, 32), (50_000, 64)):
            print(f"[lsm] spin-limited n={n} fp{prec}", flush=True)
            req = make_request("american_put", n_paths=n, n_steps=50, precision=prec, seed=8, **LSM)
            want = engine.price(req)
            assert engine.last_timing()["main_launches"] == 3
            got = eng.price(req)
            assert eng.last_timing()["main_launches"] > 3  # the per-date launches ran
            assert (got.price, got.stderr, got.early_exercise_count, got.maturity_exercise_count,
                    got.regression_steps) == (want.price, want.stderr, want.early_exercise_count,
                                              want.maturity_exercise_count, want.regression_steps)
    finally:
        eng.close()
