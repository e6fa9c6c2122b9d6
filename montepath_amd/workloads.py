"""Synthetic workloads of the benchmark configurations, shared by bench.py
and the full-size GPU tests.

* the Heston precision workload: 2^22 paths x 256 full-truncation Euler
  steps, scrambled Sobol with Brownian-bridge ordering, strikes 80/100/120;
* the batching workload: 4,096 GBM contracts (half European, half
  arithmetic-average Asian calls) with drawn parameters;
* the serving workload: a Poisson stream of mixed requests whose path
  counts follow from a target standard error.
"""
from __future__ import annotations

import math

import numpy as np

N_PATHS = 1 << 22
N_STEPS = 256
STRIKES = [80.0, 100.0, 120.0]
HESTON = dict(s0=100.0, v0=0.04, kappa=2.0, theta=0.04, xi=0.5, rho=-0.7, r=0.03, t=1.0)


def heston_request(precision: int, n_paths: int = N_PATHS, n_steps: int = N_STEPS):
    from montepath_amd.engine import make_request

    return make_request("heston_european_call", n_paths=n_paths, n_steps=n_steps, seed=42,
                        precision=precision, heston_scheme="full_truncation",
                        sampling="scrambled_sobol_brownian_bridge", strikes=STRIKES, **HESTON)


def batch_contracts(n: int = 4096, seed: int = 7):
    """4,096 contracts: half European, half arithmetic-average Asian calls under
    GBM; K~U[80,120], sigma~U[0.1,0.6], T~U[0.1,3], r~U[0,0.05], paths
    log-uniform in [2^12, 2^20], steps uniform in {16..365}; parameter draws
    from one seeded host generator; Philox keyed by (seed 7, contract id)."""
    from montepath_amd.engine import make_request

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        reqs.append(make_request(
            "european_call" if i < n // 2 else "arithmetic_asian_call",
            n_paths=int(2 ** rng.uniform(12, 20)), n_steps=int(rng.integers(16, 366)),
            k=float(rng.uniform(80, 120)), sigma=float(rng.uniform(0.1, 0.6)),
            t=float(rng.uniform(0.1, 3.0)), r=float(rng.uniform(0.0, 0.05)),
            seed=7, contract_id=i, precision=32))
    return reqs


SERVE_KINDS = ("european_call", "arithmetic_asian_call", "heston_european_call",
               "merton_jump_diffusion_call", "american_put")


def serve_request(kind: str, k: float, paths: int, contract: int, precision: int = 32):
    from montepath_amd.engine import make_request

    common = dict(seed=11, contract_id=contract, precision=precision)
    if kind in ("european_call", "arithmetic_asian_call"):
        return make_request(kind, n_paths=paths, n_steps=64, k=k, **common)
    if kind == "heston_european_call":
        return make_request(kind, n_paths=paths, n_steps=64, k=k, heston_scheme="full_truncation",
                            **HESTON, **common)
    if kind == "merton_jump_diffusion_call":
        return make_request(kind, n_paths=paths, n_steps=0, k=k, jump_intensity=1.0,
                            jump_mean=-0.1, jump_volatility=0.15, sigma=0.2, **common)
    return make_request("american_put", n_paths=paths, n_steps=50, k=k, basis_degree=3,
                        r=0.05, sigma=0.25, **common)


def serve_pilot(eng) -> dict:
    """Coefficient of variation (payoff sd / price) per request kind and
    strike, from 2^16-path pilots; interpolated in K."""
    grid = np.linspace(80.0, 120.0, 9)
    cv = {}
    for kind in SERVE_KINDS:
        vals = []
        for k in grid:
            r = eng.price(serve_request(kind, float(k), 1 << 16, 0))
            vals.append(r.stderr * math.sqrt(1 << 16) / max(r.price, 1e-12))
        cv[kind] = (grid, np.array(vals))
    return cv


def serve_mix(n: int, rate: float, cv: dict, seed: int = 11):
    """Poisson arrivals of mixed requests -- 40% GBM European/Asian, 25% Heston,
    20% Merton, 15% LSM American put; K~U[80,120] -- each with a target
    standard error of 1e-3 to 1e-2 (log-uniform) relative to its price, which
    the pilot's coefficient of variation turns into a path count clipped to
    [1e4, 1e7]."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    reqs = []
    for i in range(n):
        u = rng.random()
        kind = (SERVE_KINDS[0] if rng.random() < 0.5 else SERVE_KINDS[1]) if u < 0.40 else \
            SERVE_KINDS[2] if u < 0.65 else SERVE_KINDS[3] if u < 0.85 else SERVE_KINDS[4]
        k = float(rng.uniform(80, 120))
        rel_se = 10 ** rng.uniform(-3, -2)
        g, v = cv[kind]
        paths = int(min(max((float(np.interp(k, g, v)) / rel_se) ** 2, 1e4), 1e7))
        reqs.append(serve_request(kind, k, paths, i))
    return reqs, arrivals


