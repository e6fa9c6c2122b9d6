"""GPU tests of the per-call host path.

A one-request call receives its request record as a kernel argument, stages
small inputs through a pinned buffer and reads its result row from pinned
host memory; a batch group of one request takes the one-request path; batch
inputs and result rows live in pinned buffers that grow with the batch.
None of this may change a result: every price below must equal the same
request priced alone, bit for bit, whatever was priced before it.
"""
import numpy as np
import pytest

from montepath_amd.engine import make_request

pytestmark = pytest.mark.gpu

HESTON = dict(kappa=2.0, theta=0.04, v0=0.04, xi=0.5, rho=-0.7)

SINGLES = [
    ("european_call", dict(n_paths=50_000, n_steps=64, precision=32)),
    ("arithmetic_asian_call", dict(n_paths=20_000, n_steps=48, precision=64)),
    ("down_and_out_call", dict(n_paths=30_000, n_steps=32, precision=32, barrier=85.0)),
    ("heston_european_call", dict(n_paths=40_000, n_steps=32, precision=32,
                                  heston_scheme="full_truncation", strikes=[90.0, 100.0, 110.0],
                                  **HESTON)),
    ("merton_jump_diffusion_call", dict(n_paths=30_000, precision=64, jump_intensity=1.0,
                                        jump_mean=-0.1, jump_volatility=0.15)),
    ("european_call", dict(n_paths=25_000, n_steps=16, precision=64,
                           sampling="randomized_halton")),
    ("lookback_call", dict(n_paths=25_000, n_steps=40, precision=32, technique="antithetic")),
]


def key(r):
    return (r.price, r.stderr, tuple(r.strike_prices), tuple(r.strike_stderrs))


@pytest.mark.parametrize("kind,kw", SINGLES)
def test_one_request_batch_equals_single(engine, kind, kw):
    print(f"[host-path] {kind} {sorted(kw)}", flush=True)
    req = make_request(kind, seed=21, contract_id=4, **kw)
    alone = engine.price(req)
    (batched,) = engine.price_batch([req])
    assert key(batched) == key(alone)
    # and again after other work has reused the staging and result buffers
    engine.price(make_request("american_put", n_paths=8_192, n_steps=16, basis_degree=3, seed=2))
    assert key(engine.price(req)) == key(alone)


def test_result_rows_survive_batch_growth_and_shrink(engine):
    rng = np.random.default_rng(7)
    pool = []
    for i in range(80):
        kind = ("european_call", "arithmetic_asian_call")[i % 2]
        pool.append(make_request(kind, n_paths=int(2 ** rng.uniform(11, 15)),
                                 n_steps=int(rng.integers(8, 80)), k=float(rng.uniform(85, 115)),
                                 seed=5, contract_id=i, precision=32 if i % 3 else 64))
    alone = [key(engine.price(r)) for r in pool]
    print("[host-path] batches of 3, 80, 5", flush=True)
    for lo, hi in ((0, 3), (0, 80), (10, 15)):
        got = [key(r) for r in engine.price_batch(pool[lo:hi])]
        assert got == alone[lo:hi]


def test_lsm_result_row_between_other_calls(engine):
    req = make_request("american_put", n_paths=50_000, n_steps=25, basis_degree=3, seed=9,
                       r=0.05, sigma=0.25, k=100.0, t=1.0)
    first = engine.price(req)
    engine.price(make_request("european_call", n_paths=10_000, n_steps=8, seed=1))
    engine.price_batch([make_request("european_call", n_paths=4_096, n_steps=8, seed=1, contract_id=j)
                        for j in range(4)])
    again = engine.price(req)
    assert (again.price, again.stderr, again.early_exercise_count, again.regression_steps) == (
        first.price, first.stderr, first.early_exercise_count, first.regression_steps)
    assert first.regression_steps == 24
