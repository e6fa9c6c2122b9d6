"""Sobol + Brownian-bridge pricing (FP32 and FP64) with the bridge in
registers (step counts 16 * 2^J) against the LDS-stack kernel.

Both consume the bridge plan in the same preorder.  FP64 uses the same
arithmetic, so the bar is bit-identical per-path terminal values and prices.
The FP32 register bridge carries increments scaled by 2^depth and rounds
each half of a split once (the stack kernel forms the right half as the
interval minus the left), and its European kernels add the per-step drift
once per path; each FP32 form is within about 1.5e-6 of FP64 on the same
draws (the parity tests hold both to 1e-5), so the bar between them is
5e-6 relative per path (measured 2.2e-6) and 1e-4 standard errors on
prices (measured 4e-5).  The stack kernel is pinned to the oracle by the
parity tests (FP64 within 1e-12, FP32 within 1e-5).
"""
import os

import numpy as np
import pytest

from montepath_amd.engine import Engine, make_request

pytestmark = pytest.mark.gpu

HESTON = dict(kappa=2.0, theta=0.04, v0=0.04, xi=0.5, rho=-0.7)


@pytest.fixture(scope="module")
def stack_engine():
    """A context opened with the LDS-stack bridge selected."""
    os.environ["MONTEPATH_QMC_STACK_BRIDGE"] = "1"
    try:
        eng = Engine(0)
    finally:
        del os.environ["MONTEPATH_QMC_STACK_BRIDGE"]
    yield eng
    eng.close()


CASES = [
    # (workload, steps, extra): European specialisations (both Heston schemes),
    # the generic body (Asian, antithetic) and one to five outer levels
    ("heston_european_call", 256, dict(heston_scheme="full_truncation", strikes=[80.0, 100.0, 120.0], **HESTON)),
    ("heston_european_call", 64, dict(heston_scheme="reference_floor", **HESTON)),
    ("heston_european_call", 16, dict(heston_scheme="full_truncation", **HESTON)),
    ("heston_european_call", 512, dict(heston_scheme="full_truncation", technique="antithetic", **HESTON)),
    ("european_call", 128, {}),
    ("arithmetic_asian_call", 32, {}),
    ("lookback_call", 256, dict(technique="control_variate")),
]


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("kind,m,extra", CASES)
def test_register_bridge_equals_stack_bridge(engine, stack_engine, kind, m, extra, precision):
    n = 5000
    print(f"[qmc-tree] {kind} m={m} fp{precision} {sorted(extra)}", flush=True)
    req = make_request(kind, n_paths=n, n_steps=m, precision=precision, seed=9,
                       sampling="scrambled_sobol_brownian_bridge", **extra)
    a, term_a, pay_a = engine.price_with_paths(req)
    b, term_b, pay_b = stack_engine.price_with_paths(req)
    if precision == 64:
        assert np.array_equal(term_a, term_b)
        assert np.array_equal(pay_a, pay_b)
        assert a.strike_prices == b.strike_prices
        assert a.strike_stderrs == b.strike_stderrs
        return
    used = term_b > 0  # antithetic requests leave the entries past their pairs at 0
    assert np.array_equal(term_a > 0, used)
    rel = np.abs(term_a[used] - term_b[used]) / term_b[used]
    assert float(rel.max()) <= 5e-6, float(rel.max())
    assert float(np.max(np.abs(pay_a - pay_b) / (np.abs(pay_b) + 100.0))) <= 5e-6
    for pa, pb, sa, sb in zip(a.strike_prices, b.strike_prices, a.strike_stderrs, b.strike_stderrs):
        assert abs(pa - pb) <= 1e-4 * sb + 1e-12, (pa, pb, sb)
        assert abs(sa - sb) <= 1e-4 * sb + 1e-12, (sa, sb)


def test_heston_xi_zero_on_the_register_bridge(engine):
    # The European register-bridge kernels scale the variance factor by xi,
    # so xi = 0 is priced by the generic kernel: constant variance v0 = theta
    # makes the model GBM with sigma = sqrt(v0)
    import math

    kw = dict(n_paths=1 << 16, n_steps=256, seed=3, kappa=2.0, theta=0.04, v0=0.04, xi=0.0, rho=-0.7,
              heston_scheme="full_truncation", sampling="scrambled_sobol_brownian_bridge")
    a, t32, _ = engine.price_with_paths(make_request("heston_european_call", precision=32, **kw))
    b, t64, _ = engine.price_with_paths(make_request("heston_european_call", precision=64, **kw))
    assert float(np.max(np.abs(t32 - t64) / t64)) <= 1e-5
    sd = 0.2
    d1 = (math.log(100.0 / 100.0) + (0.03 + 0.5 * sd * sd)) / sd
    ncdf = lambda x: 0.5 * math.erfc(-x / math.sqrt(2.0))  # noqa: E731
    bs = 100.0 * ncdf(d1) - 100.0 * math.exp(-0.03) * ncdf(d1 - sd)
    assert abs(b.price - bs) <= 4 * b.stderr + 1e-3, (b.price, bs, b.stderr)
    assert abs(a.price - b.price) <= 1e-4 * b.stderr + 1e-9


@pytest.mark.parametrize("rho,xi", [(-1.0, 0.5), (1.0, 0.3), (0.2, 0.01)])
def test_heston_factor_scaled_bridge_edges(engine, rho, xi):
    # The European register-bridge kernels carry xi dW on the variance
    # factor and sqrt(1 - rho^2) dW on the independent one: rho = +-1 (the
    # independent factor's scales are zero) and a small xi (rho / xi
    # large).  FP32 per path against FP64 on the same draws (measured worst
    # 4.0e-6, 1.2e-6, 7.8e-7).
    kw = dict(n_paths=1 << 16, n_steps=256, seed=11, kappa=2.0, theta=0.04, v0=0.04, xi=xi, rho=rho,
              heston_scheme="full_truncation", sampling="scrambled_sobol_brownian_bridge")
    print(f"[qmc-tree] heston rho={rho} xi={xi}", flush=True)
    a, t32, _ = engine.price_with_paths(make_request("heston_european_call", precision=32, **kw))
    b, t64, _ = engine.price_with_paths(make_request("heston_european_call", precision=64, **kw))
    err = np.abs(t32 - t64) / t64
    print(f"[qmc-tree] max {float(err.max()):.3g}", flush=True)
    assert float(err.max()) <= 1e-5, float(err.max())
    assert abs(a.price - b.price) <= 1e-4 * b.stderr + 1e-9, (a.price, b.price, b.stderr)


def test_heston_ill_conditioned_regime_agrees_on_prices(engine):
    # xi = 1 (2 kappa theta / xi^2 = 0.16): paths spend steps near v = 0,
    # where the full-truncation update's derivative xi sqrt(dt) Z / (2 sqrt(v))
    # is unbounded, so the scheme amplifies any perturbation of the draws;
    # FP64 arithmetic fed the same increments rounded to FP32 diverges as
    # much (tests/test_heston_conditioning_cpu.py).  No per-path bar can hold
    # between the FP32 and FP64 runs here (measured: 18% of paths beyond 1e-5,
    # worst 0.26); the prices agree (measured 0.04 standard errors apart).
    kw = dict(n_paths=1 << 16, n_steps=256, seed=11, kappa=2.0, theta=0.04, v0=0.04, xi=1.0, rho=-0.7,
              heston_scheme="full_truncation", sampling="scrambled_sobol_brownian_bridge")
    print("[qmc-tree] heston xi=1.0 prices", flush=True)
    a = engine.price(make_request("heston_european_call", precision=32, **kw))
    b = engine.price(make_request("heston_european_call", precision=64, **kw))
    assert abs(a.price - b.price) <= 0.2 * b.stderr, (a.price, b.price, b.stderr)
    assert abs(a.stderr - b.stderr) <= 0.02 * b.stderr, (a.stderr, b.stderr)


@pytest.mark.parametrize("kind,n,m", [("heston_european_call", 1 << 14, 4096), ("heston_european_call", 1 << 12, 8192),
                                      ("european_call", 1 << 15, 2048), ("arithmetic_asian_call", 1 << 15, 1024)])
def test_long_bridges_fp32_against_fp64(engine, kind, n, m):
    # Register bridges with 6 to 9 outer levels (16 * 2^J steps) and Sobol
    # dimensions far past the 256-dimension wrap (up to 16,384 for Heston):
    # FP32 per path against FP64 on the same draws (measured worst 3.8e-6)
    # and prices (measured within 4e-5 standard errors)
    extra = dict(HESTON, heston_scheme="full_truncation") if kind.startswith("heston") else {}
    print(f"[qmc-tree] long bridge {kind} n={n} m={m}", flush=True)
    kw = dict(n_paths=n, n_steps=m, seed=3, sampling="scrambled_sobol_brownian_bridge", **extra)
    a, ta, _ = engine.price_with_paths(make_request(kind, precision=32, **kw))
    b, tb, _ = engine.price_with_paths(make_request(kind, precision=64, **kw))
    assert float(np.max(np.abs(ta - tb) / tb)) <= 1e-5
    assert abs(a.price - b.price) <= 1e-3 * b.stderr, (a.price, b.price, b.stderr)
