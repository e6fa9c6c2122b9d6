"""Seeded random market parameters and shapes, FP64 on the GPU against the
oracle on identical draws (per-path bar 1e-12 relative, as the parity tests).

The parity tests fix the market (S0 = K = 100, r = 0.03, sigma = 0.2, T = 1);
here every case draws its own rate (negative included), volatility, maturity,
strike, barrier, path count and step count, so that a constant folded in the
wrong place or a parameter read from the wrong field fails.  Heston cases keep
xi <= 0.6: far outside the Feller condition the full-truncation scheme is ill
conditioned in its draws, and the GPU's fused multiply-adds (the oracle rounds
every product) would be amplified past 1e-12 there
(tests/test_heston_conditioning_cpu.py).
"""
import math

import numpy as np
import pytest

from montepath_amd.engine import make_request

pytestmark = pytest.mark.gpu

GBM_PAYOFFS = ["european_call", "arithmetic_asian_call", "down_and_out_call", "lookback_call"]


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _gbm_cases(k=24, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        out.append(dict(payoff=GBM_PAYOFFS[i % 4], n=int(rng.integers(1, 3000)), m=int(rng.integers(1, 130)),
                        r=float(rng.uniform(-0.02, 0.12)), sigma=float(rng.uniform(0.01, 0.9)),
                        t=float(rng.uniform(0.05, 3.0)), k=float(rng.uniform(60.0, 150.0)),
                        barrier=float(rng.uniform(50.0, 99.0)), seed=1000 + i))
    return out


@pytest.mark.parametrize("case", _gbm_cases(), ids=lambda c: f"{c['payoff']}-n{c['n']}-m{c['m']}")
def test_gbm_random_parameters(engine, oracle, case):
    c = case
    print(f"[random] {c}", flush=True)
    n, m = c["n"], c["m"]
    z = oracle.rng_normals(c["seed"], n * m).reshape(n, m)
    cfg = oracle.Gbm(n_paths=n, n_steps=m, seed=c["seed"], r=c["r"], sigma=c["sigma"], t=c["t"], k=c["k"],
                     barrier=c["barrier"])
    term_ref, pay_ref = oracle.gbm_paths_from_normals(cfg, z, c["payoff"])
    req = make_request(c["payoff"], n_paths=n, n_steps=m, precision=64, r=c["r"], sigma=c["sigma"], t=c["t"],
                       k=c["k"], barrier=c["barrier"])
    term, pay = engine.paths_from_normals(req, z.T)
    assert relmax(term, term_ref) <= 1e-12
    disc = math.exp(-c["r"] * c["t"])
    assert np.all(np.abs(pay * disc - pay_ref) <= 1e-12 * (np.abs(pay_ref) + 100.0))


def _heston_cases(k=12, seed=77):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        out.append(dict(n=int(rng.integers(1, 2000)), m=int(rng.integers(1, 100)), r=float(rng.uniform(-0.01, 0.1)),
                        t=float(rng.uniform(0.1, 2.0)), k=float(rng.uniform(70.0, 130.0)),
                        v0=float(rng.uniform(0.01, 0.2)), kappa=float(rng.uniform(0.3, 5.0)),
                        theta=float(rng.uniform(0.01, 0.15)), xi=float(rng.uniform(0.05, 0.6)),
                        rho=float(rng.uniform(-0.95, 0.5)), scheme=["reference_floor", "full_truncation"][i % 2],
                        seed=2000 + i))
    return out


@pytest.mark.parametrize("case", _heston_cases(), ids=lambda c: f"{c['scheme']}-n{c['n']}-m{c['m']}")
def test_heston_random_parameters(engine, oracle, case):
    c = case
    print(f"[random] heston {c}", flush=True)
    n, m = c["n"], c["m"]
    z = oracle.rng_normals(c["seed"], n * m * 2).reshape(n, m, 2)
    cfg = oracle.Heston(n_paths=n, n_steps=m, r=c["r"], t=c["t"], k=c["k"], v0=c["v0"], kappa=c["kappa"],
                        theta=c["theta"], xi=c["xi"], rho=c["rho"])
    term_ref, _ = oracle.heston_paths_from_normals(cfg, z, 0 if c["scheme"] == "reference_floor" else 1)
    req = make_request("heston_european_call", n_paths=n, n_steps=m, precision=64, r=c["r"], t=c["t"], k=c["k"],
                       v0=c["v0"], kappa=c["kappa"], theta=c["theta"], xi=c["xi"], rho=c["rho"],
                       heston_scheme=c["scheme"])
    term, _ = engine.paths_from_normals(req, np.ascontiguousarray(z.transpose(1, 2, 0)))
    assert relmax(term, term_ref) <= 1e-12


def _merton_cases(k=8, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        out.append(dict(n=int(rng.integers(1, 3000)), m=[1, 4, 16, 33][i % 4], r=float(rng.uniform(-0.01, 0.1)),
                        sigma=float(rng.uniform(0.05, 0.5)), t=float(rng.uniform(0.1, 2.0)),
                        lam=float(rng.uniform(0.0, 8.0)), mu_j=float(rng.uniform(-0.3, 0.2)),
                        sigma_j=float(rng.uniform(0.0, 0.4)), seed=3000 + i))
    return out


@pytest.mark.parametrize("case", _merton_cases(), ids=lambda c: f"n{c['n']}-m{c['m']}")
def test_merton_random_parameters(engine, oracle, case):
    c = case
    print(f"[random] merton {c}", flush=True)
    n, m = c["n"], c["m"]
    rng = np.random.default_rng(c["seed"])
    z = rng.standard_normal((n, m))
    u = (rng.integers(0, 1 << 23, size=(n, m)) + 0.5) / float(1 << 23)
    zj = rng.standard_normal((n, m))
    cfg = oracle.Merton(n_paths=n, r=c["r"], sigma=c["sigma"], t=c["t"], lam=c["lam"], mu_j=c["mu_j"],
                        sigma_j=c["sigma_j"])
    term_ref, _ = oracle.merton_paths_from_draws(cfg, m, z, u, zj)
    req = make_request("merton_jump_diffusion_call", n_paths=n, n_steps=m, precision=64, r=c["r"],
                       sigma=c["sigma"], t=c["t"], jump_intensity=c["lam"], jump_mean=c["mu_j"],
                       jump_volatility=c["sigma_j"])
    term, _ = engine.paths_from_normals(req, z.T.copy(), u.T.copy(), zj.T.copy())
    assert relmax(term, term_ref) <= 1e-12


def _qmc_cases(k=16, seed=99):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        m = int(rng.choice([1, 7, 16, 64, 100, 256, 300]))
        n = int(min(rng.integers(1, 70_000), 3_000_000 // m))
        out.append(dict(payoff=GBM_PAYOFFS[i % 4], bridge=bool(i % 2), n=n, m=m,
                        seed=int(rng.integers(0, 1 << 32)), r=float(rng.uniform(-0.02, 0.12)),
                        sigma=float(rng.uniform(0.05, 0.8)), t=float(rng.uniform(0.1, 2.5)),
                        k=float(rng.uniform(70.0, 140.0)), barrier=float(rng.uniform(50.0, 99.0))))
    return out


@pytest.mark.parametrize("case", _qmc_cases(),
                         ids=lambda c: f"{c['payoff']}-bb{int(c['bridge'])}-n{c['n']}-m{c['m']}")
def test_gbm_sobol_random_parameters(engine, oracle, case):
    # scrambled Sobol (with and without the Brownian bridge) across the
    # 2^16-point block and the 256-dimension wrap, random seeds and markets
    c = case
    print(f"[random] sobol {c}", flush=True)
    cfg = oracle.Gbm(n_paths=c["n"], n_steps=c["m"], seed=c["seed"], r=c["r"], sigma=c["sigma"], t=c["t"],
                     k=c["k"], barrier=c["barrier"])
    p_ref, _, term_ref = oracle.gbm_qmc_price(cfg, c["payoff"], bridge=c["bridge"])
    sampling = "scrambled_sobol_brownian_bridge" if c["bridge"] else "scrambled_sobol"
    req = make_request(c["payoff"], n_paths=c["n"], n_steps=c["m"], seed=c["seed"], precision=64, r=c["r"],
                       sigma=c["sigma"], t=c["t"], k=c["k"], barrier=c["barrier"], sampling=sampling)
    res, term, _ = engine.price_with_paths(req)
    assert relmax(term, term_ref) <= 1e-12
    assert abs(res.price - p_ref) <= 1e-12 * (abs(p_ref) + 1.0)


def _heston_qmc_cases(k=8, seed=123):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        m = int(rng.choice([16, 32, 64, 100, 128, 256]))
        out.append(dict(n=int(rng.integers(1, 4000)), m=m, seed=int(rng.integers(0, 1 << 32)),
                        r=float(rng.uniform(-0.01, 0.1)), t=float(rng.uniform(0.1, 2.0)),
                        k=float(rng.uniform(70.0, 130.0)), v0=float(rng.uniform(0.01, 0.2)),
                        kappa=float(rng.uniform(0.3, 5.0)), theta=float(rng.uniform(0.01, 0.15)),
                        xi=float(rng.uniform(0.05, 0.6)), rho=float(rng.uniform(-0.95, 0.5))))
    return out


@pytest.mark.parametrize("case", _heston_qmc_cases(), ids=lambda c: f"n{c['n']}-m{c['m']}")
def test_heston_sobol_bridge_random_parameters(engine, oracle, case):
    # two-factor Sobol + bridge (register bridge for 16 * 2^J steps, LDS
    # stack otherwise), full truncation, random seeds and markets
    c = case
    print(f"[random] heston sobol {c}", flush=True)
    cfg = oracle.Heston(n_paths=c["n"], n_steps=c["m"], seed=c["seed"], r=c["r"], t=c["t"], k=c["k"],
                        v0=c["v0"], kappa=c["kappa"], theta=c["theta"], xi=c["xi"], rho=c["rho"])
    term_ref = oracle.heston_qmc_paths(cfg, 1, path0=0)
    req = make_request("heston_european_call", n_paths=c["n"], n_steps=c["m"], seed=c["seed"], precision=64,
                       r=c["r"], t=c["t"], k=c["k"], v0=c["v0"], kappa=c["kappa"], theta=c["theta"],
                       xi=c["xi"], rho=c["rho"], heston_scheme="full_truncation",
                       sampling="scrambled_sobol_brownian_bridge")
    _, term, _ = engine.price_with_paths(req)
    assert relmax(term, term_ref) <= 1e-12


def _lsm_cases(k=10, seed=31):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        m = int(rng.integers(2, 41))
        sched = None
        if i % 3 == 2:
            pick = rng.choice(np.arange(1, m), size=min(m - 1, int(rng.integers(1, 8))), replace=False)
            sched = sorted(int(s) for s in pick) + [m]
        out.append(dict(n=int(rng.integers(500, 20_000)), m=m, seed=int(rng.integers(1, 1 << 31)),
                        r=float(rng.uniform(0.0, 0.1)), sigma=float(rng.uniform(0.1, 0.5)),
                        t=float(rng.uniform(0.25, 2.0)), k=float(rng.uniform(80.0, 120.0)),
                        degree=int(rng.integers(1, 4)), weighted=bool(i % 4 == 3), sched=sched))
    return out


@pytest.mark.parametrize("case", _lsm_cases(),
                         ids=lambda c: f"n{c['n']}-m{c['m']}-d{c['degree']}-{'berm' if c['sched'] else 'amer'}")
def test_lsm_random_parameters(engine, oracle, case):
    # FP64 Longstaff-Schwartz on identical normals: random markets, basis
    # degree 1-3, both bases, American and Bermudan schedules.  The Gram
    # solves round differently from the oracle's (pivot reciprocal, FMA), so
    # a path whose exercise test is a tie at that rounding may take another
    # date; each such path moves the price by at most K / n.
    c = case
    print(f"[random] lsm {c}", flush=True)
    n, m = c["n"], c["m"]
    cfg = oracle.Lsm(n_paths=n, n_steps=m, seed=c["seed"], r=c["r"], sigma=c["sigma"], t=c["t"], k=c["k"],
                     basis_degree=c["degree"], basis_weighted=c["weighted"])
    z = oracle.rng_normals(c["seed"], n * m).reshape(n, m)
    paths = oracle.lsm_paths_from_normals(cfg, z)
    ref = oracle.lsm_from_paths(cfg, paths, schedule=c["sched"], want_detail=True)
    kind = "bermudan_put" if c["sched"] else "american_put"
    extra = {"exercise_steps": c["sched"]} if c["sched"] else {}
    req = make_request(kind, n_paths=n, n_steps=m, precision=64, r=c["r"], sigma=c["sigma"], t=c["t"], k=c["k"],
                       basis_degree=c["degree"], basis_weighted=c["weighted"], **extra)
    res, _, ex = engine.lsm_from_normals(req, z.T.copy())
    assert res.regression_steps == ref["regression_steps"]
    flips = int(np.sum(ex != ref["exercise_steps"]))
    assert flips <= max(2, int(1e-4 * n)), flips
    assert abs(res.price - ref["price"]) <= 1e-10 * ref["price"] + flips * c["k"] / n


def _basket_cases(k=10, seed=17):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        na = int(rng.integers(1, 9))
        f = rng.standard_normal((na, na + 2))
        c = f @ f.T
        d = np.sqrt(np.diag(c))
        corr = c / np.outer(d, d)
        np.fill_diagonal(corr, 1.0)
        out.append(dict(na=na, n=int(rng.integers(1, 3000)), m=int(rng.integers(1, 41)),
                        s0=[float(x) for x in rng.uniform(60.0, 140.0, na)],
                        sigma=[float(x) for x in rng.uniform(0.05, 0.6, na)],
                        w=[float(x) for x in rng.uniform(0.05, 1.0, na)], corr=corr,
                        r=float(rng.uniform(-0.01, 0.1)), t=float(rng.uniform(0.1, 2.5)),
                        k=float(rng.uniform(70.0, 130.0)), seed=4000 + i))
    return out


@pytest.mark.parametrize("case", _basket_cases(), ids=lambda c: f"a{c['na']}-n{c['n']}-m{c['m']}")
def test_basket_random_parameters(engine, oracle, case):
    # random assets (1-8), weights, correlation matrices (random factor
    # loadings) and markets; FP64 basket values per path against the oracle
    c = case
    print(f"[random] basket a={c['na']} n={c['n']} m={c['m']}", flush=True)
    n, m, na = c["n"], c["m"], c["na"]
    z = oracle.rng_normals(c["seed"], n * m * na).reshape(n, m, na)
    ref = oracle.basket_paths_from_normals(c["s0"], c["sigma"], c["w"], c["corr"], c["r"], c["t"], m, n, z)
    req = make_request("basket_call", n_paths=n, n_steps=m, precision=64, r=c["r"], t=c["t"], k=c["k"],
                       asset_s0=c["s0"], asset_sigma=c["sigma"], asset_weight=c["w"], correlation=c["corr"])
    term, _ = engine.paths_from_normals(req, np.ascontiguousarray(z.transpose(1, 2, 0)))
    assert relmax(term, ref) <= 1e-12


SAMPLINGS = ["randomized_halton", "latin_hypercube", "scrambled_sobol", "scrambled_sobol_brownian_bridge"]
TECHNIQUES = ["standard", "antithetic", "control_variate"]


def _structured_cases(k=16, seed=8):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        out.append(dict(payoff=GBM_PAYOFFS[i % 4], sampling=SAMPLINGS[(i // 4) % 4], technique=TECHNIQUES[i % 3],
                        n=int(rng.integers(2, 6000)), m=int(rng.integers(1, 60)), seed=int(rng.integers(1, 1 << 31)),
                        r=float(rng.uniform(-0.01, 0.1)), sigma=float(rng.uniform(0.05, 0.7)),
                        t=float(rng.uniform(0.1, 2.5)), k=float(rng.uniform(75.0, 130.0)),
                        barrier=float(rng.uniform(55.0, 95.0))))
    return out


@pytest.mark.parametrize("case", _structured_cases(),
                         ids=lambda c: f"{c['payoff']}-{c['sampling']}-{c['technique']}-n{c['n']}-m{c['m']}")
def test_structured_samplers_random_parameters(engine, oracle, case):
    # every structured sampler with every technique at random markets and
    # shapes: FP64 price and standard error against the oracle's pricers
    c = case
    print(f"[random] structured {c}", flush=True)
    cfg = oracle.Gbm(n_paths=c["n"], n_steps=c["m"], seed=c["seed"], r=c["r"], sigma=c["sigma"], t=c["t"],
                     k=c["k"], barrier=c["barrier"])
    p_ref, se_ref, _ = oracle.gbm_structured_price(cfg, c["payoff"], c["sampling"], c["technique"])
    res = engine.price(make_request(c["payoff"], n_paths=c["n"], n_steps=c["m"], seed=c["seed"], precision=64,
                                    r=c["r"], sigma=c["sigma"], t=c["t"], k=c["k"], barrier=c["barrier"],
                                    sampling=c["sampling"], technique=c["technique"]))
    assert abs(res.price - p_ref) <= 1e-12 * (abs(p_ref) + 1e-3), (res.price, p_ref)
    assert abs(res.stderr - se_ref) <= 1e-10 * (abs(se_ref) + 1e-6), (res.stderr, se_ref)


def _greek_cases(k=8, seed=41):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        out.append(dict(estimator=["pathwise", "likelihood_ratio"][i % 2], n=int(rng.integers(1000, 60_000)),
                        s0=float(rng.uniform(70.0, 130.0)), k=float(rng.uniform(70.0, 130.0)),
                        sigma=float(rng.uniform(0.05, 0.6)), t=float(rng.uniform(0.1, 3.0)),
                        r=float(rng.uniform(-0.01, 0.1)), seed=5000 + i))
    return out


@pytest.mark.parametrize("case", _greek_cases(), ids=lambda c: f"{c['estimator']}-n{c['n']}")
def test_single_pass_greeks_random_parameters(engine, oracle, case):
    # pathwise and likelihood-ratio Greeks of the terminal European call at
    # random markets, FP64 on identical normals against the oracle
    c = case
    print(f"[random] greeks {c}", flush=True)
    z = oracle.rng_normals(c["seed"], c["n"])
    mk = dict(s0=c["s0"], k=c["k"], sigma=c["sigma"], t=c["t"], r=c["r"])
    want = oracle.terminal_greeks(oracle.Gbm(n_paths=c["n"], n_steps=1, **mk), c["estimator"], z)
    req = make_request("european_call", n_paths=c["n"], n_steps=64, precision=64, **mk)
    _, est = engine.greeks_from_normals(req, c["estimator"], z)
    assert [e["greek"] for e in est] == list(want)
    for e in est:
        v, se = want[e["greek"]]
        assert abs(e["value"] - v) <= 1e-10 * abs(v) + 1e-12, (e["greek"], e["value"], v)
        assert abs(e["stderr"] - se) <= 1e-9 * se + 1e-15, (e["greek"], e["stderr"], se)


def _mlmc_cases(k=8, seed=61):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        base, ref, level = int(rng.integers(2, 17)), int(rng.integers(2, 5)), int(rng.integers(0, 4))
        while base * ref ** level > 400:
            level -= 1
        out.append(dict(base=base, ref=ref, level=level, paths=int(rng.integers(1, 3000)),
                        s0=float(rng.uniform(80.0, 120.0)), k=float(rng.uniform(80.0, 120.0)),
                        r=float(rng.uniform(-0.01, 0.1)), sigma=float(rng.uniform(0.05, 0.6)),
                        t=float(rng.uniform(0.2, 2.0)), seed=6000 + i))
    return out


@pytest.mark.parametrize("case", _mlmc_cases(), ids=lambda c: f"b{c['base']}-r{c['ref']}-l{c['level']}-n{c['paths']}")
def test_mlmc_level_terms_random_parameters(engine, oracle, case):
    # multilevel Asian level terms (fine path and coupled coarse path) at
    # random markets, base steps, refinement factors and levels, FP64 on
    # identical normals; level differences are compared on the scale of the
    # fine payoff sum, as they cancel
    from montepath_amd.engine import make_mlmc_config

    c = case
    print(f"[random] mlmc {c}", flush=True)
    fine = c["base"] * c["ref"] ** c["level"]
    z = oracle.rng_normals(c["seed"], c["paths"] * fine).reshape(c["paths"], fine)
    mk = dict(s0=c["s0"], k=c["k"], r=c["r"], sigma=c["sigma"], t=c["t"])
    cfg_o = oracle.mlmc_config(base_steps=c["base"], levels=4, refinement_factor=c["ref"], **mk)
    want = oracle.mlmc_level_moments(cfg_o, c["level"], c["paths"], z)
    cfg = make_mlmc_config(base_steps=c["base"], levels=4, refinement_factor=c["ref"], precision=64, **mk)
    got = engine.mlmc_level_from_normals(cfg, c["level"], z)
    scale = abs(want[0]) if c["level"] == 0 else max(abs(want[0]), math.sqrt(want[1] * c["paths"]))
    assert abs(got[0] - want[0]) <= 1e-12 * scale + 1e-300, (got, want)
    assert abs(got[1] - want[1]) <= 1e-11 * abs(want[1]) + 1e-300, (got, want)
