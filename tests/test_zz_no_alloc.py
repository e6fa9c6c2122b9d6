"""The allocation-free pricing entry point against mp_hip_price.

mp_hip_price_noalloc prices into a static device array instead of the
context's arenas; its results must equal mp_hip_price's bit for bit.  Two
calls in a row, the second with a smaller request (the array keeps the first
call's contents), and a second context on the same device, which shares the
array; requests beyond the array or of other kinds are rejected.
"""
import ctypes as C

import pytest

from montepath_amd import _abi as A
from montepath_amd.engine import Engine, make_request

pytestmark = pytest.mark.gpu


def _noalloc(eng, req):
    out = A.Result()
    rc = eng.lib.mp_hip_price_noalloc(eng._ctx, C.byref(req), C.byref(out))
    return rc, out


def _price(eng, req):
    out = A.Result()
    A.check(eng.lib.mp_hip_price(eng._ctx, C.byref(req), C.byref(out)))
    return out


def _same(a, b):
    return (a.price, a.std_error, a.n_paths, a.n_steps) == (b.price, b.std_error, b.n_paths, b.n_steps) and \
        [a.strike_price[i] for i in range(4)] == [b.strike_price[i] for i in range(4)]


CASES = [
    # (larger, smaller) request pairs, priced in that order
    (dict(workload="european_call", n_paths=1_000_000, n_steps=64, precision=32),
     dict(workload="european_call", n_paths=10_000, n_steps=16, precision=32)),
    (dict(workload="arithmetic_asian_call", n_paths=300_001, n_steps=100, precision=64),
     dict(workload="down_and_out_call", n_paths=4099, n_steps=37, precision=64, barrier=90.0)),
    (dict(workload="heston_european_call", n_paths=200_000, n_steps=128, precision=32,
          kappa=2.0, xi=0.5, rho=-0.7, heston_scheme="full_truncation",
          strikes=[80.0, 100.0, 120.0]),
     dict(workload="merton_jump_diffusion_call", n_paths=50_000, n_steps=0, precision=32,
          jump_intensity=1.0, jump_mean=-0.1, jump_volatility=0.15)),
    (dict(workload="lookback_call", n_paths=500_000, n_steps=52, precision=32,
          technique="antithetic", fp64_check=True),
     dict(workload="european_call", n_paths=1000, n_steps=1, precision=32,
          technique="control_variate")),
]


@pytest.mark.parametrize("big,small", CASES)
def test_noalloc_equals_price_twice_in_a_row(engine, big, small):
    reqs = []
    for kw in (big, small):
        kw = dict(kw)
        wl = kw.pop("workload")
        reqs.append(make_request(wl, seed=17, contract_id=3, **kw))
    print(f"[noalloc] {big['workload']} {big['n_paths']} -> {small['workload']} {small['n_paths']}",
          flush=True)
    want = [_price(engine, r) for r in reqs]
    for r, w in zip(reqs, want):
        rc, got = _noalloc(engine, r)
        assert rc == A.MP_OK, engine.lib.mp_hip_last_error()
        assert got.status == A.MP_OK
        assert _same(got, w), (got.price, w.price)


def test_noalloc_from_a_second_context_shares_the_device_array(engine):
    other = Engine(0)
    try:
        a = make_request("european_call", n_paths=250_000, n_steps=64, seed=5)
        b = make_request("heston_european_call", n_paths=20_000, n_steps=32, seed=6,
                         kappa=2.0, xi=0.5, rho=-0.7)
        for eng, r in ((engine, a), (other, b), (engine, b), (other, a)):
            rc, got = _noalloc(eng, r)
            assert rc == A.MP_OK, eng.lib.mp_hip_last_error()
            assert _same(got, _price(eng, r))
    finally:
        other.close()


def test_noalloc_rejects_what_it_cannot_price(engine):
    # beyond the device array: 1e8 paths x 365 steps needs ~400k tiles
    big = make_request("european_call", n_paths=100_000_000, n_steps=365)
    rc, _ = _noalloc(engine, big)
    assert rc == A.MP_E_OOM
    assert b"scratch" in engine.lib.mp_hip_last_error()
    for kind, kw in (("american_put", dict(n_steps=50)), ("european_call", dict(sampling="scrambled_sobol")),
                     ("european_call", dict(sampling="randomized_halton"))):
        rc, _ = _noalloc(engine, make_request(kind, n_paths=10_000, **kw))
        assert rc == A.MP_E_UNSUPPORTED, kind
    # and the entry point still prices correctly afterwards
    r = make_request("european_call", n_paths=20_000, n_steps=8)
    rc, got = _noalloc(engine, r)
    assert rc == A.MP_OK and _same(got, _price(engine, r))
