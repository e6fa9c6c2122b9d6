"""FP32 GBM pricing from the engine's own Philox draws against FP64 on the
same draws.

The FP32 Philox body of the work-queue kernel folds the Box-Muller radius
and the volatility into one multiply per pair, accumulates a European
log-return without per-step drift and carries Asian paths in log2 units; the
FP64 body computes every normal and step as the reference does.  Both see
the same uniforms, so per path they must agree to FP32 accuracy: terminal
values within 1e-5 relative, payoffs within 1e-5 of (payoff + 100), barrier
flags flipping on at most 0.1% of paths (a path grazing the barrier within
FP32 rounding).  Step counts cover the full 4-step blocks and every tail.
"""
import numpy as np
import pytest

from montepath_amd.engine import make_request

pytestmark = pytest.mark.gpu

PAYOFFS = [("european_call", {}), ("arithmetic_asian_call", {}),
           ("down_and_out_call", {"barrier": 90.0}), ("lookback_call", {})]


@pytest.mark.parametrize("n_steps", [1, 2, 3, 37, 64])
@pytest.mark.parametrize("kind,extra", PAYOFFS)
def test_fp32_philox_paths_match_fp64_on_the_same_draws(engine, kind, extra, n_steps):
    n = 5003  # a partial last wave in the last tile
    kw = dict(n_paths=n, n_steps=n_steps, seed=23, contract_id=9, sigma=0.35, t=2.0, **extra)
    r32, t32, p32 = engine.price_with_paths(make_request(kind, precision=32, **kw))
    r64, t64, p64 = engine.price_with_paths(make_request(kind, precision=64, **kw))
    rel = np.abs(t32 - t64) / t64
    assert float(rel.max()) <= 1e-5, float(rel.max())
    bad = np.abs(p32 - p64) > 1e-5 * (np.abs(p64) + 100.0)
    if kind == "down_and_out_call":
        assert bad.mean() <= 1e-3
    else:
        assert not bad.any(), float(np.max(np.abs(p32 - p64) / (np.abs(p64) + 100.0)))
    assert abs(r32.price - r64.price) <= 3e-3 * r64.stderr + 1e-9


@pytest.mark.parametrize("kind,extra", PAYOFFS)
def test_fp32_philox_tunings_and_batch_give_identical_bits(engine, kind, extra):
    req = make_request(kind, n_paths=70_001, n_steps=45, seed=3, contract_id=1, **extra)
    alone = engine.price(req)
    for block in (64, 128, 256):
        for ppl in (1, 2, 4):
            for unroll in (1, 2, 4):
                engine.set_tuning(req, block, ppl, unroll)
                got = engine.price(req)
                assert (got.price, got.stderr) == (alone.price, alone.stderr), (block, ppl, unroll)
    engine.set_tuning(req, 256, 1)
    others = [make_request("european_call", n_paths=9000 + i, n_steps=20 + i, seed=3, contract_id=2 + i)
              for i in range(5)]
    batch = engine.price_batch(others[:2] + [req] + others[2:])
    assert (batch[2].price, batch[2].stderr) == (alone.price, alone.stderr)
