"""GPU tests of batching, serving and tuning.

Continuous batching must not change results: every request priced through
the server -- admitted while others are on the device, packed with whatever
else was pending -- equals the same request priced alone, bit for bit.
"""
import numpy as np
import pytest

from montepath_amd import _abi as A
from montepath_amd.engine import make_request

pytestmark = pytest.mark.gpu


def mixed_requests(n, seed=3):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        kind = i % 5
        k = float(rng.uniform(80, 120))
        paths = int(2 ** rng.uniform(12, 16))
        if kind == 0:
            reqs.append(make_request("european_call", n_paths=paths, n_steps=64, k=k, seed=11,
                                     contract_id=i))
        elif kind == 1:
            reqs.append(make_request("arithmetic_asian_call", n_paths=paths, n_steps=48, k=k,
                                     seed=11, contract_id=i))
        elif kind == 2:
            reqs.append(make_request("heston_european_call", n_paths=paths, n_steps=32, k=k,
                                     seed=11, contract_id=i, kappa=2.0, xi=0.5, rho=-0.7,
                                     heston_scheme="full_truncation"))
        elif kind == 3:
            reqs.append(make_request("merton_jump_diffusion_call", n_paths=paths, k=k, seed=11,
                                     contract_id=i, jump_intensity=1.0, jump_mean=-0.1,
                                     jump_volatility=0.15))
        else:
            reqs.append(make_request("american_put", n_paths=min(paths, 1 << 14), n_steps=20,
                                     basis_degree=3, seed=11, contract_id=i))
    return reqs


def test_served_results_equal_results_alone(engine):
    reqs = mixed_requests(40)
    alone = [engine.price(r) for r in reqs]
    arrivals = np.linspace(0.0, 0.02, len(reqs))  # arrive while earlier work runs
    res, lat, stats = engine.serve_trace(reqs, arrivals, max_inflight=2, max_batch=8)
    assert stats["completed"] == len(reqs)
    assert stats["launches"] >= 2
    for a, b in zip(alone, res):
        assert a.price == b.price and a.stderr == b.stderr
    assert np.all(lat > 0)


def test_admission_keeps_device_bytes_within_budget(engine):
    """Two workers, a budget B and American puts whose reservations are about
    0.6 B each: no two of them may be on the device together, the bytes the
    workers hold or reserve never exceed B, and every result equals the same
    request priced alone."""
    fp = engine.footprint(make_request("american_put", n_paths=1 << 18, n_steps=50))
    budget = int((fp + fp // 4 + (1 << 20)) / 0.6)
    reqs = []
    for i in range(12):
        if i % 3 == 2:
            reqs.append(make_request("european_call", n_paths=1 << 14, n_steps=32, seed=5,
                                     contract_id=i))
        else:
            n = (1 << 18) - 4096 * (i % 4)
            reqs.append(make_request("american_put", n_paths=n, n_steps=50, r=0.05, sigma=0.25,
                                     basis_degree=3, seed=5, contract_id=i))
    alone = [engine.price(r) for r in reqs]
    with engine.server(hbm_budget_bytes=budget, max_inflight=2, max_batch=4) as srv:
        tickets = [srv.submit(r) for r in reqs]
        served = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs)
    assert 0 < st["max_admitted_bytes"] <= budget, (st, budget)
    for a, b in zip(alone, served):
        assert a.price == b.price and a.stderr == b.stderr


def test_budget_admission_and_errors():
    from montepath_amd.engine import Engine

    import ctypes as C

    eng = Engine(0)
    lib = eng.lib
    srv = C.c_void_p()
    cfg = A.ServeConfig(4 << 20, 1, 4)  # 4 MiB budget
    A.check(lib.mp_hip_server_open(eng._ctx, C.byref(cfg), C.byref(srv)))
    try:
        big = make_request("american_put", n_paths=1 << 20, n_steps=50)
        t = C.c_uint64()
        rc = lib.mp_hip_submit(srv, C.byref(big), C.byref(t))
        assert rc == A.MP_E_OOM
        small = make_request("european_call", n_paths=4096, n_steps=16)
        A.check(lib.mp_hip_submit(srv, C.byref(small), C.byref(t)))
        out = A.Result()
        A.check(lib.mp_hip_wait(srv, t.value, C.byref(out)))
        assert out.status == A.MP_OK and out.price > 0
        assert lib.mp_hip_poll(srv, 12345, C.byref(out), C.byref(C.c_int())) == A.MP_E_NOT_FOUND
    finally:
        lib.mp_hip_server_close(srv)
        eng.close()


def test_autotune_keeps_results(engine):
    req = make_request("european_call", n_paths=1 << 18, n_steps=64, seed=4)
    before = engine.price(req)
    best, ms = engine.autotune(req)
    assert best["block_size"] in (128, 256) and best["paths_per_lane"] in (1, 2, 4)
    assert best["step_unroll"] in (1, 2, 4)
    assert engine.get_tuning(req) == best
    assert ms > 0
    after = engine.price(req)
    assert after.price == before.price and after.stderr == before.stderr
    engine.set_tuning(req, 256, 1)


@pytest.mark.parametrize("kind,extra", [
    ("european_call", {}),
    ("arithmetic_asian_call", {}),
    ("down_and_out_call", {"barrier": 85.0}),
    ("heston_european_call", {"kappa": 2.0, "xi": 0.5, "rho": -0.7, "heston_scheme": "full_truncation"}),
    ("heston_european_call", {"kappa": 2.0, "xi": 0.5, "rho": -0.7}),
    ("merton_jump_diffusion_call", {"jump_intensity": 1.0, "jump_mean": -0.1, "jump_volatility": 0.15}),
])
def test_every_tuning_gives_identical_results(engine, kind, extra):
    # block size and paths interleaved per lane select different compiled
    # kernels; the per-path arithmetic and the summation order must not change
    req = make_request(kind, n_paths=(1 << 16) + 77, n_steps=37, seed=4, **extra)
    seen = set()
    for bs in (128, 256):
        for ppl in (1, 2, 4):
            engine.set_tuning(req, bs, ppl)
            r = engine.price(req)
            seen.add((r.price, r.stderr))
    engine.set_tuning(req, 256, 1)
    assert len(seen) == 1, seen


def test_terminal_method_matches_black_scholes(engine):
    req = make_request("european_call", n_paths=1 << 20, n_steps=64, method="terminal", seed=9)
    res = engine.price(req)
    assert res.n_steps == 64
    assert abs(res.price - 9.413403383853016) <= 3 * res.stderr


def test_bermudan_and_american_against_crr(engine, oracle):
    am = engine.price(make_request("american_put", n_paths=1 << 18, n_steps=64, basis_degree=2,
                                   precision=64, seed=5))
    crr = oracle.crr_put(100, 100, 0.03, 0.2, 1.0, 512, range(0, 513))
    assert abs(am.price - crr) <= 5 * am.stderr + 0.15
    assert am.regression_steps == 63
    be = engine.price(make_request("bermudan_put", n_paths=1 << 18, n_steps=64, basis_degree=2,
                                   exercise_steps=[16, 32, 48, 64], precision=64, seed=5))
    crr_b = oracle.crr_put(100, 100, 0.03, 0.2, 1.0, 512, [128, 256, 384, 512])
    assert abs(be.price - crr_b) <= 5 * be.stderr + 0.15
    assert be.regression_steps == 3


def test_five_asset_basket_bounds(engine):
    from oracle import closed_forms as cf

    na = 5
    corr = np.full((na, na), 0.5) + 0.5 * np.eye(na)
    req = make_request("basket_call", n_paths=1 << 20, n_steps=64, asset_s0=[100.0] * na,
                       asset_sigma=[0.2] * na, asset_weight=[0.2] * na, correlation=corr, seed=2)
    res = engine.price(req)
    geo = cf.geometric_basket_call([100.0] * na, [0.2] * na, [0.2] * na, corr, 0.03, 1.0, 100.0)
    assert res.price >= geo - 3 * res.stderr
    # perfectly correlated identical assets are one asset: Black-Scholes
    req1 = make_request("basket_call", n_paths=1 << 20, n_steps=16, asset_s0=[100.0] * 3,
                        asset_sigma=[0.2] * 3, asset_weight=[1 / 3] * 3,
                        correlation=np.full((3, 3), 1.0) - 1e-9 * (1 - np.eye(3)), seed=2)
    r1 = engine.price(req1)
    assert abs(r1.price - 9.413403383853016) <= 3 * r1.stderr
