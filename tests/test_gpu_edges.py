"""Edge cases on the GPU: ragged path and step counts (one path, partial
waves, step counts off the four-step draw blocks, odd Brownian-bridge
lengths, one regression date), degenerate contracts (deep out of the money,
zero volatility) and invalid requests, each against the oracle or a closed
form.  FP64 per-path bar as in the parity tests: 1e-12 relative.
"""
import math

import numpy as np
import pytest

from montepath_amd.engine import make_request

pytestmark = pytest.mark.gpu

GBM_PAYOFFS = ["european_call", "arithmetic_asian_call", "down_and_out_call", "lookback_call"]
RAGGED = [(1, 1), (1, 5), (63, 3), (65, 7), (130, 1), (257, 6)]


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("n,m", RAGGED)
@pytest.mark.parametrize("payoff", GBM_PAYOFFS)
def test_ragged_paths_on_reference_normals(engine, oracle, payoff, n, m):
    print(f"[edges] {payoff} n={n} m={m}", flush=True)
    z = oracle.rng_normals(77, n * m).reshape(n, m)
    cfg = oracle.Gbm(n_paths=n, n_steps=m, seed=77, barrier=90.0)
    term_ref, pay_ref = oracle.gbm_paths_from_normals(cfg, z, payoff)
    req = make_request(payoff, n_paths=n, n_steps=m, precision=64, barrier=90.0)
    term, pay = engine.paths_from_normals(req, z.T)
    assert relmax(term, term_ref) <= 1e-12
    assert np.all(np.abs(pay * math.exp(-0.03) - pay_ref) <= 1e-12 * (np.abs(pay_ref) + 100.0))
    res = engine.price_from_normals(req, z.T)
    mean = float(np.mean(pay_ref))
    assert abs(res.price - mean) <= 1e-12 * (abs(mean) + 1.0)
    if n == 1:
        assert res.stderr == 0.0


@pytest.mark.parametrize("n,m", [(1, 1), (3, 2), (65, 3), (100, 5), (333, 12)])
@pytest.mark.parametrize("bridge", [False, True])
def test_ragged_sobol_paths_match_oracle(engine, oracle, n, m, bridge):
    print(f"[edges] sobol bridge={bridge} n={n} m={m}", flush=True)
    cfg = oracle.Gbm(n_paths=n, n_steps=m, seed=5)
    p_ref, _, term_ref = oracle.gbm_qmc_price(cfg, "european_call", bridge=bridge)
    sampling = "scrambled_sobol_brownian_bridge" if bridge else "scrambled_sobol"
    req = make_request("european_call", n_paths=n, n_steps=m, seed=5, precision=64, sampling=sampling)
    res, term, _ = engine.price_with_paths(req)
    assert relmax(term, term_ref) <= 1e-12
    assert abs(res.price - p_ref) <= 1e-12 * (abs(p_ref) + 1.0)


@pytest.mark.parametrize("n", [1, 2, 63, 65, 4097])
@pytest.mark.parametrize("m", [1, 3, 5])
def test_ragged_philox_requests_alone_and_batched(engine, n, m):
    print(f"[edges] philox n={n} m={m}", flush=True)
    reqs = [make_request(kind, n_paths=n, n_steps=m, seed=13, contract_id=i, precision=prec)
            for i, (kind, prec) in enumerate([("european_call", 32), ("arithmetic_asian_call", 32),
                                              ("lookback_call", 64), ("european_call", 64)])]
    alone = [engine.price(r) for r in reqs]
    batched = engine.price_batch(reqs)
    for a, b in zip(alone, batched):
        assert (a.price, a.stderr) == (b.price, b.stderr)
        assert math.isfinite(a.price) and a.price >= 0.0
        if n == 1:
            assert a.stderr == 0.0


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("sampling", ["pseudorandom", "scrambled_sobol_brownian_bridge"])
def test_deep_out_of_the_money_prices_zero(engine, precision, sampling):
    req = make_request("european_call", n_paths=20_000, n_steps=16, k=1e6, precision=precision,
                       sampling=sampling)
    res = engine.price(req)
    assert res.price == 0.0 and res.stderr == 0.0


@pytest.mark.parametrize("precision", [32, 64])
def test_zero_volatility_asian_is_deterministic(engine, precision):
    # sigma = 0 (allowed for the path-dependent payoffs, as in the reference's
    # validation): S_i = S0 exp(r i dt) on every path
    s0, k, r, t, m = 100.0, 95.0, 0.03, 1.0, 16
    req = make_request("arithmetic_asian_call", n_paths=10_000, n_steps=m, s0=s0, k=k, r=r, t=t,
                       sigma=0.0, precision=precision)
    res = engine.price(req)
    avg = sum(s0 * math.exp(r * t * i / m) for i in range(1, m + 1)) / m
    exact = math.exp(-r * t) * (avg - k)
    assert abs(res.price - exact) <= (1e-12 if precision == 64 else 1e-5) * exact
    # identical payoffs: the population variance is rounding noise only
    assert res.stderr <= 1e-6


@pytest.mark.parametrize("n,m", [(777, 2), (1500, 3), (4097, 7)])
def test_ragged_lsm_on_reference_normals(engine, oracle, n, m):
    print(f"[edges] lsm n={n} m={m}", flush=True)
    cfg = oracle.Lsm(n_paths=n, n_steps=m, seed=8, basis_degree=3, r=0.05, sigma=0.25)
    z = oracle.rng_normals(8, n * m).reshape(n, m)
    ref = oracle.lsm_from_paths(cfg, oracle.lsm_paths_from_normals(cfg, z), want_detail=True)
    req = make_request("american_put", n_paths=n, n_steps=m, precision=64, basis_degree=3, r=0.05,
                       sigma=0.25)
    res, _, ex = engine.lsm_from_normals(req, z.T.copy())
    assert res.regression_steps == ref["regression_steps"]
    assert abs(res.price - ref["price"]) <= 1e-10 * ref["price"]
    assert int(np.sum(ex != ref["exercise_steps"])) <= 2


@pytest.mark.parametrize("bad", [dict(n_paths=0), dict(n_steps=0, kind="arithmetic_asian_call"),
                                 dict(sigma=-0.2), dict(sigma=0.0), dict(t=0.0), dict(s0=0.0),
                                 dict(sigma=-0.1, kind="lookback_call"),
                                 # validate_american_put_config cpu.rs:5127-5135: n_steps >= 2
                                 dict(n_steps=1, kind="american_put")])
def test_invalid_requests_are_errors(engine, bad):
    kw = dict(bad)
    kind = kw.pop("kind", "european_call")
    req = make_request(kind, **kw)
    with pytest.raises(Exception):
        engine.price(req)


def _corr(n, rho):
    c = np.full((n, n), rho)
    np.fill_diagonal(c, 1.0)
    return c


MODEL_EDGES = [
    # nearly singular correlation, the most assets, negative correlation,
    # one asset; Merton with 40 jumps a year (Poisson counts past the cached
    # terms), terminal and stepped
    ("basket_call", dict(n_steps=64, asset_s0=[100.0] * 5, asset_sigma=[0.2, 0.22, 0.24, 0.26, 0.28],
                         correlation=_corr(5, 0.99))),
    ("basket_call", dict(n_steps=64, asset_s0=[100.0] * 8, asset_sigma=[0.2] * 8, correlation=_corr(8, 0.0))),
    ("basket_call", dict(n_steps=64, asset_s0=[100.0] * 3, asset_sigma=[0.2, 0.22, 0.24],
                         correlation=_corr(3, -0.45))),
    ("basket_call", dict(n_steps=64, asset_s0=[100.0], asset_sigma=[0.2], correlation=_corr(1, 0.0))),
    ("merton_jump_diffusion_call", dict(n_steps=0, jump_intensity=40.0)),
    ("merton_jump_diffusion_call", dict(n_steps=64, jump_intensity=40.0)),
]


@pytest.mark.parametrize("kind,kw", MODEL_EDGES)
def test_model_edges_fp32_against_fp64(engine, kind, kw):
    # FP32 per path against FP64 on the same draws (measured worst 2.9e-6)
    # and prices (measured within 1e-4 standard errors)
    print(f"[edges] {kind} {sorted(kw)}", flush=True)
    a, ta, _ = engine.price_with_paths(make_request(kind, n_paths=1 << 16, seed=4, precision=32, **kw))
    b, tb, _ = engine.price_with_paths(make_request(kind, n_paths=1 << 16, seed=4, precision=64, **kw))
    assert not np.isnan(ta).any()
    assert float(np.max(np.abs(ta - tb) / np.abs(tb))) <= 1e-5
    assert abs(a.price - b.price) <= 1e-3 * b.stderr, (a.price, b.price, b.stderr)
