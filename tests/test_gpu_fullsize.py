"""GPU tests at the benchmark configurations' full sizes.

* the batching configuration: 4,096 GBM contracts (half European, half
  arithmetic-average Asian calls, 2^12..2^20 paths, 16..365 steps, 1.4e11
  path-steps) in one prepared, graph-captured persistent launch.  At this
  size the work queue hands out several tiles per claim.  Replays are
  identical; a fixed sample of contracts spanning the path and step ranges
  equals the same contracts priced alone (one-request launches, which skip
  the queue), bit for bit; every European price agrees with Black-Scholes
  (log-Euler GBM is exact in distribution).
* the serving configuration: a Poisson stream at 2,000 requests/s of the
  mixed requests (40% GBM European/Asian, 25% Heston, 20% Merton, 15%
  American put with 50 dates, r = 0.05, sigma = 0.25, degree 3), path counts
  from a target standard error (1e4..1e7 paths).  Served results equal the
  same requests priced alone, for a sample of every kind that includes the
  largest requests (at least 1e6 paths).
* the LSM configuration: the American put S0 = K = 100, r = 0.05,
  sigma = 0.25, T = 1 on 50 dates with 2^20 paths and degree-3 Laguerre
  regression, unweighted and weighted, near the CRR Bermudan-50 value.
"""
import math

import numpy as np
import pytest

from montepath_amd.engine import make_request
from montepath_amd.workloads import SERVE_KINDS, batch_contracts, serve_mix, serve_pilot

pytestmark = pytest.mark.gpu


def black_scholes_call(s0, k, r, sigma, t):
    sd = sigma * math.sqrt(t)
    d1 = (math.log(s0 / k) + (r + 0.5 * sigma * sigma) * t) / sd
    d2 = d1 - sd
    ncdf = lambda x: 0.5 * math.erfc(-x / math.sqrt(2.0))  # noqa: E731
    return s0 * ncdf(d1) - k * math.exp(-r * t) * ncdf(d2)


@pytest.mark.parametrize("weighted", [False, True])
def test_lsm_full_size_near_crr_bermudan(engine, weighted):
    print(f"[fullsize] lsm 2^20 x 50 weighted={weighted}", flush=True)
    req = make_request("american_put", n_paths=1 << 20, n_steps=50, k=100.0, r=0.05, sigma=0.25,
                       t=1.0, basis_degree=3, precision=32, seed=42, basis_weighted=weighted)
    res = engine.price(req)
    assert res.regression_steps == 49
    # CRR, 5000 tree steps exercising every 100th: 7.962057; LSM's
    # suboptimal exercise rule biases it low by a few cents at most
    assert abs(res.price - 7.962057) <= 4 * res.stderr + 0.03, (res.price, res.stderr)


def test_4096_contract_batch_full_size(engine):
    reqs = batch_contracts()
    pb = engine.prepare_batch(reqs, use_graph=True)
    try:
        info = pb.info()
        print(f"[c3] {len(reqs)} contracts, {info['n_tiles']} tiles, "
              f"{info['path_steps']:.3e} path-steps, grid {info['grid']} x {info['block']}",
              flush=True)
        assert info["path_steps"] > 1.3e11
        # at least 32 tiles per wave: the work queue claims several tiles at once
        waves = info["grid"] * info["block"] // 64
        assert info["n_tiles"] >= 32 * waves
        pb.launch()
        pb.sync()
        first = pb.results()
        pb.launch()
        pb.sync()
        second = pb.results()
    finally:
        pb.free()
    assert [(a.price, a.stderr) for a in first] == [(b.price, b.stderr) for b in second]

    # the same batch under another launch tuning (two paths per lane, four
    # 4-step draw blocks per step-loop iteration): the same bits
    default = engine.get_tuning(reqs[0])
    engine.set_tuning(reqs[0], 128, 2, 4)
    try:
        pb = engine.prepare_batch(reqs, use_graph=False)
        try:
            pb.launch()
            pb.sync()
            retuned = pb.results()
        finally:
            pb.free()
    finally:
        engine.set_tuning(reqs[0], default["block_size"], default["paths_per_lane"],
                          default["step_unroll"])
    assert [(a.price, a.stderr) for a in first] == [(b.price, b.stderr) for b in retuned]

    # a fixed sample spanning the path and step ranges, both payoffs
    paths = np.array([r.n_paths for r in reqs])
    steps = np.array([r.n_steps for r in reqs])
    pick = set()
    for half in (range(0, 2048), range(2048, 4096)):
        idx = np.array(list(half))
        for key in (paths, steps):
            order = idx[np.argsort(key[idx], kind="stable")]
            pick.update(int(order[j]) for j in np.linspace(0, len(order) - 1, 8).astype(int))
    pick = sorted(pick)
    assert len(pick) >= 24
    for i in pick:
        alone = engine.price(reqs[i])
        assert (alone.price, alone.stderr) == (first[i].price, first[i].stderr), \
            (i, reqs[i].n_paths, reqs[i].n_steps, alone.price, first[i].price)

    # every European contract against Black-Scholes
    z = []
    for i in range(2048):
        r = reqs[i]
        bs = black_scholes_call(r.s0, r.k, r.r, r.sigma, r.t)
        z.append((first[i].price - bs) / first[i].stderr)
    z = np.array(z)
    assert np.all(np.isfinite(z))
    assert abs(z.mean()) < 4.0 / math.sqrt(len(z)), z.mean()
    assert np.mean(np.abs(z) < 3.0) > 0.99, np.sort(np.abs(z))[-10:]
    assert np.max(np.abs(z)) < 5.0, np.max(np.abs(z))


def test_serving_stream_full_mix(engine):
    rate, seconds = 2000.0, 5.0
    cv = serve_pilot(engine)
    reqs, arrivals = serve_mix(int(rate * seconds), rate, cv)
    paths = np.array([r.n_paths for r in reqs])
    print(f"[c5] {len(reqs)} requests over {arrivals[-1]:.2f} s, paths p50 "
          f"{np.median(paths):.0f} max {paths.max()}, {int(np.sum(paths >= 1e6))} with >= 1e6",
          flush=True)
    assert paths.max() == 10_000_000 and paths.min() == 10_000
    res, lat, stats = engine.serve_trace(reqs, arrivals, max_inflight=2)
    print(f"[c5] {stats['requests_per_s']:.0f} req/s, p50 {stats['p50_ms']:.3f} ms, "
          f"p99 {stats['p99_ms']:.3f} ms, max {stats['max_ms']:.3f} ms, "
          f"{stats['launches']} launches", flush=True)
    assert stats["completed"] == len(reqs)
    assert all(math.isfinite(r.price) and r.n_paths > 0 for r in res)
    # LSM requests carry the configuration's parameters
    lsm = [r for r in reqs if r.workload == make_request("american_put").workload]
    assert lsm and all(r.n_steps == 50 and r.r == 0.05 and r.sigma == 0.25 for r in lsm)
    kinds = {k: make_request(k).workload for k in SERVE_KINDS}
    checked = 0
    for kind, wl in kinds.items():
        idx = [i for i, r in enumerate(reqs) if r.workload == wl]
        assert idx, kind
        by_size = sorted(idx, key=lambda i: reqs[i].n_paths)
        sample = set(by_size[-3:]) | set(by_size[:: max(1, len(by_size) // 4)])
        for i in sorted(sample):
            alone = engine.price(reqs[i])
            assert (alone.price, alone.stderr) == (res[i].price, res[i].stderr), \
                (kind, i, reqs[i].n_paths, alone.price, res[i].price)
            checked += 1
    assert checked >= 20
    big = [i for i in range(len(reqs)) if reqs[i].n_paths >= 1_000_000]
    assert big
