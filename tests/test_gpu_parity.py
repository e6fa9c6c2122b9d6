"""GPU parity: the HIP engine against the CPU oracle on identical draws.

Bars (written per test): FP64 path arithmetic within 1e-12 relative of the
oracle's f64 restatement of the reference; FP32 path arithmetic within 1e-5
relative; integer draws (Philox words, Sobol points) bit-exact.
"""
import math

import numpy as np
import pytest

from montepath_amd.engine import make_request

pytestmark = pytest.mark.gpu

GBM_PAYOFFS = ["european_call", "arithmetic_asian_call", "down_and_out_call", "lookback_call"]
TOL = {64: 1e-12, 32: 1e-5}


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


# ----------------------------------------------------------------- draws
def test_philox_uniform_words_bit_exact(engine, oracle):
    req = make_request("european_call", seed=0x1234_5678_9ABC_DEF0, contract_id=77)
    got = engine.draws(req, 1, 1, 1000, 64, 12)
    ref = oracle.philox_uniforms(0x1234_5678_9ABC_DEF0, 77, 1, 1000, 64, 12)
    assert np.array_equal(got, ref)


def test_philox_fp64_normals_match_oracle(engine, oracle):
    req = make_request("european_call", seed=7, contract_id=3)
    got = engine.draws(req, 0, 0, 0, 256, 20)
    ref = oracle.philox_normals(7, 3, 0, 0, 256, 20)
    assert float(np.max(np.abs(got - ref))) <= 1e-12


@pytest.mark.parametrize("stream,path0,n,m", [(2, 5, 1024, 16), (0, 0, 65536, 64)])
def test_philox_fp32_hardware_normals_close_to_fp64(engine, oracle, stream, path0, n, m):
    # v_log_f32 / v_sqrt_f32 / v_sin_f32 / v_cos_f32 path against the f64
    # restatement on the same uniforms (measured: max 7.1e-7 absolute over
    # 4.2e6 normals, max 7.7e-7 relative where |z| > 1e-3)
    req = make_request("european_call", seed=99, contract_id=0)
    got = engine.draws(req, 4, stream, path0, n, m)
    ref = oracle.philox_normals(99, 0, stream, path0, n, m)
    err = np.abs(got - ref)
    assert float(np.max(err / np.maximum(np.abs(ref), 1.0))) <= 2e-6
    assert float(np.max(err[np.abs(ref) > 1e-3] / np.abs(ref[np.abs(ref) > 1e-3]))) <= 3e-6
    assert float(np.median(err)) <= 1e-7


def _sobol_raw_ref(oracle, seed, point, dims):
    block, local = point >> 16, point & 0xFFFF
    bseed = (seed + block * 0x9E3779B9) & 0xFFFFFFFF
    return [oracle.sobol_scrambled_u32(local, d % 256, (bseed + d // 256) & 0xFFFFFFFF)
            for d in range(dims)]


def test_sobol_scrambled_integers_bit_exact_across_block_and_dim_wrap(engine, oracle):
    seed = 42
    req = make_request("european_call", seed=seed)
    path0, n, dims = 65530, 12, 520  # crosses the 2^16-point block and the 256-dim wrap
    got = engine.draws(req, 3, 0, path0, n, dims).astype(np.uint64)
    for i in range(n):
        ref = np.array(_sobol_raw_ref(oracle, seed, path0 + i, dims), dtype=np.uint64)
        assert np.array_equal(got[i], ref), f"point {path0 + i}"


def test_sobol_normals_match_oracle(engine, oracle):
    req = make_request("european_call", seed=11)
    got = engine.draws(req, 2, 0, 0, 64, 300)
    for i in range(64):
        ref = oracle.sobol_point_normals(11, i, 300)
        assert float(np.max(np.abs(got[i] - ref))) <= 1e-12


def test_sobol_fp64_normals_tight_on_the_tails(engine, oracle):
    # The FP64 inverse normal's own log, sqrt and divide against the
    # oracle's libm log and sqrt and IEEE divide, on 4,096 points x 256
    # dimensions (about 50,000 tail draws: p < 0.02425 or p > 0.97575,
    # |z| > 1.97296).  The tail rational barely cancels, so what remains
    # is a few roundings: the GPU fuses the Horner steps into FMAs and the
    # oracle does not (at most 1.1e-15 relative apart, by exact arithmetic
    # on the CPU in test_fp64_inverse_normal_cpu.py), and the log is within
    # 1 ulp of libm's.  The central rational
    # cancels by a factor of ~200 next to the branch points, where fused
    # and unfused Horner steps alone differ by up to 3.5e-13 relative, so
    # the whole range keeps the 1e-12 absolute bar.  Measured: tail 9.1e-16
    # relative (50,857 draws), whole range 6.7e-13 absolute.
    req = make_request("european_call", seed=5)
    got = engine.draws(req, 2, 0, 0, 4096, 256)
    ref = np.stack([oracle.sobol_point_normals(5, i, 256) for i in range(4096)])
    tail = np.abs(ref) > 1.973
    rel_tail = float(np.max(np.abs(got[tail] - ref[tail]) / np.abs(ref[tail])))
    abs_all = float(np.max(np.abs(got - ref)))
    print(f"[fp64 normals] tail draws {int(tail.sum())}, max rel {rel_tail:.3g}; all max abs {abs_all:.3g}",
          flush=True)
    assert int(tail.sum()) > 40000
    assert rel_tail <= 4e-15
    assert abs_all <= 1e-12


# ------------------------------------------------- per-path arithmetic (GBM)
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("payoff", GBM_PAYOFFS)
def test_gbm_paths_from_reference_normals(engine, oracle, payoff, precision):
    n, m = 4096, 37  # 37 steps: exercises the partial 4-step Philox block
    z = oracle.rng_normals(404, n * m).reshape(n, m)
    cfg = oracle.Gbm(n_paths=n, n_steps=m, seed=404, barrier=90.0)
    term_ref, pay_ref = oracle.gbm_paths_from_normals(cfg, z, payoff)
    req = make_request(payoff, n_paths=n, n_steps=m, precision=precision, barrier=90.0)
    term, pay = engine.paths_from_normals(req, z.T)
    assert relmax(term, term_ref) <= TOL[precision]
    disc = math.exp(-0.03)
    diff = np.abs(pay * disc - pay_ref)
    scale = np.abs(pay_ref) + 100.0
    bad = diff > TOL[precision] * scale
    if payoff == "down_and_out_call" and precision == 32:
        # a path grazing the barrier within FP32 rounding may flip its knock flag
        assert bad.mean() <= 1e-3
    else:
        assert not bad.any(), float(np.max(diff / scale))


# ------------------------------------------------------------- Heston
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("scheme", ["reference_floor", "full_truncation"])
def test_heston_paths_from_reference_normals(engine, oracle, scheme, precision):
    n, m = 2048, 64
    z = oracle.rng_normals(9, n * m * 2).reshape(n, m, 2)
    cfg = oracle.Heston(n_paths=n, n_steps=m, kappa=2.0, xi=0.5, rho=-0.7)
    sc = 0 if scheme == "reference_floor" else 1
    term_ref, _ = oracle.heston_paths_from_normals(cfg, z, sc)
    req = make_request("heston_european_call", n_paths=n, n_steps=m, precision=precision,
                       kappa=2.0, theta=0.04, v0=0.04, xi=0.5, rho=-0.7, heston_scheme=scheme)
    term, _ = engine.paths_from_normals(req, np.ascontiguousarray(z.transpose(1, 2, 0)))
    err = np.abs(term - term_ref) / term_ref
    if precision == 64:
        assert float(err.max()) <= 1e-12
    else:
        # every path within 1e-5 relative (measured worst: 3.1e-7)
        assert float(err.max()) <= 1e-5, float(err.max())


# ------------------------------------------------------------- Merton
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("n_steps", [1, 16])
def test_merton_paths_from_draws(engine, oracle, n_steps, precision):
    n = 4096
    rng = np.random.default_rng(5)
    z = rng.standard_normal((n, n_steps))
    # uniforms on the engine's 23-bit grid so FP32 and FP64 see them exactly
    u = (rng.integers(0, 1 << 23, size=(n, n_steps)) + 0.5) / float(1 << 23)
    zj = rng.standard_normal((n, n_steps))
    cfg = oracle.Merton(n_paths=n, lam=1.0, mu_j=-0.1, sigma_j=0.15, sigma=0.2)
    term_ref, jumps = oracle.merton_paths_from_draws(cfg, n_steps, z, u, zj)
    assert jumps.sum() > 0
    req = make_request("merton_jump_diffusion_call", n_paths=n, n_steps=n_steps,
                       precision=precision, jump_intensity=1.0, jump_mean=-0.1,
                       jump_volatility=0.15)
    term, _ = engine.paths_from_normals(req, z.T.copy(), u.T.copy(), zj.T.copy())
    assert relmax(term, term_ref) <= TOL[precision]


# ------------------------------------------------------------- basket
@pytest.mark.parametrize("precision", [64, 32])
def test_basket_paths_from_normals(engine, oracle, precision):
    n, m, na = 2048, 16, 5
    corr = np.full((na, na), 0.5) + 0.5 * np.eye(na)
    s0 = [100.0, 95.0, 105.0, 90.0, 110.0]
    sig = [0.2, 0.25, 0.3, 0.15, 0.35]
    w = [0.2] * 5
    z = oracle.rng_normals(21, n * m * na).reshape(n, m, na)
    ref = oracle.basket_paths_from_normals(s0, sig, w, corr, 0.03, 1.0, m, n, z)
    req = make_request("basket_call", n_paths=n, n_steps=m, precision=precision, asset_s0=s0,
                       asset_sigma=sig, asset_weight=w, correlation=corr)
    term, _ = engine.paths_from_normals(req, np.ascontiguousarray(z.transpose(1, 2, 0)))
    assert relmax(term, ref) <= TOL[precision]


# ------------------------------------------------------------- LSM
@pytest.mark.parametrize("weighted", [False, True])
def test_lsm_fp64_matches_oracle_on_reference_normals(engine, oracle, weighted):
    n, m = 20000, 16
    cfg = oracle.Lsm(n_paths=n, n_steps=m, seed=42, basis_degree=3, r=0.05, sigma=0.25,
                     basis_weighted=weighted)
    z = oracle.rng_normals(42, n * m).reshape(n, m)
    paths = oracle.lsm_paths_from_normals(cfg, z)
    ref = oracle.lsm_from_paths(cfg, paths, want_detail=True)
    req = make_request("american_put", n_paths=n, n_steps=m, precision=64, basis_degree=3,
                       r=0.05, sigma=0.25, basis_weighted=weighted)
    res, coeff, ex = engine.lsm_from_normals(req, z.T.copy())
    assert res.regression_steps == ref["regression_steps"]
    assert abs(res.price - ref["price"]) <= 1e-10 * ref["price"]
    assert abs(res.stderr - ref["stderr"]) <= 1e-8 * ref["stderr"]
    assert np.mean(ex == ref["exercise_steps"]) >= 0.9999
    # degree-3 Laguerre Gram matrices are ill-conditioned, so coefficients
    # carry the summation order; the fitted continuation values at the
    # in-the-money paths the regression saw do not
    for d in range(1, m):
        if not np.any(ref["coefficients"][d]):
            continue
        s = paths[:, d]
        x = s[s < 100.0] / 100.0
        basis = np.stack([oracle.laguerre_basis(v, 3, weighted) for v in x], axis=1)
        cont_ref = ref["coefficients"][d] @ basis
        cont = coeff[d] @ basis
        assert np.max(np.abs(cont - cont_ref)) <= 2e-6 * np.max(np.abs(cont_ref)), d


def test_lsm_fp32_within_noise_of_fp64(engine):
    req64 = make_request("american_put", n_paths=1 << 16, n_steps=50, precision=64,
                         basis_degree=3, r=0.05, sigma=0.25, seed=3)
    req32 = make_request("american_put", n_paths=1 << 16, n_steps=50, precision=32,
                         basis_degree=3, r=0.05, sigma=0.25, seed=3)
    a, b = engine.price(req64), engine.price(req32)
    assert abs(a.price - b.price) <= a.stderr
    assert a.regression_steps == 49


# ------------------------------------------------------------- QMC per path
@pytest.mark.parametrize("bridge", [False, True])
def test_gbm_qmc_terminals_fp64_match_oracle(engine, oracle, bridge):
    n, m = 3000, 64
    cfg = oracle.Gbm(n_paths=n, n_steps=m, seed=42)
    p_ref, se_ref, term_ref = oracle.gbm_qmc_price(cfg, "european_call", bridge=bridge)
    sampling = "scrambled_sobol_brownian_bridge" if bridge else "scrambled_sobol"
    req = make_request("european_call", n_paths=n, n_steps=m, seed=42, precision=64,
                       sampling=sampling)
    res, term, _ = engine.price_with_paths(req)
    assert relmax(term, term_ref) <= 1e-12
    assert abs(res.price - p_ref) <= 1e-12 * p_ref


@pytest.mark.parametrize("precision", [64, 32])
def test_heston_qmc_bridge_terminals_match_oracle(engine, oracle, precision):
    n, m = 2048, 256
    cfg = oracle.Heston(n_paths=n, n_steps=m, kappa=2.0, xi=0.5, rho=-0.7, seed=42)
    term_ref = oracle.heston_qmc_paths(cfg, 1, path0=0)
    req = make_request("heston_european_call", n_paths=n, n_steps=m, seed=42, precision=precision,
                       kappa=2.0, theta=0.04, v0=0.04, xi=0.5, rho=-0.7,
                       heston_scheme="full_truncation", sampling="scrambled_sobol_brownian_bridge")
    res, term, _ = engine.price_with_paths(req)
    err = np.abs(term - term_ref) / term_ref
    if precision == 64:
        assert float(err.max()) <= 1e-12
    else:
        # every path within 1e-5 relative (measured worst: 1.3e-6; the FP32
        # run also draws its normals in FP32)
        assert float(err.max()) <= 1e-5, float(err.max())
        req.precision = 64
        ref64 = engine.price(req)
        assert abs(res.price - ref64.price) <= 0.05 * ref64.stderr


# ------------------------------------------------------------- prices
@pytest.mark.parametrize("precision", [64, 32])
def test_c1_price_within_3se_of_black_scholes(engine, oracle, precision):
    req = make_request("european_call", n_paths=100_000, n_steps=64, seed=42, precision=precision)
    res = engine.price(req)
    bs = 9.413403383853016  # reference-fixtures.json black_scholes_european_call_atm_1y
    assert abs(res.price - bs) <= 3 * res.stderr
    assert 0.03 < res.stderr < 0.06


def test_fp32_and_fp64_runs_share_draws(engine):
    req = make_request("european_call", n_paths=200_000, n_steps=64, seed=5, fp64_check=True)
    res = engine.price(req)
    assert res.fp64_ref_error <= 1e-5 * res.price


def test_plain_fp64_path_equals_engine_fp64(engine):
    req = make_request("european_call", n_paths=50_000, n_steps=32, seed=8, precision=64)
    a = engine.price(req)
    b = engine.price_plain_fp64(req)
    assert abs(a.price - b.price) <= 1e-12 * a.price
    assert abs(a.stderr - b.stderr) <= 1e-9 * a.stderr


def _contract_scalars(cfg):
    dt = cfg.t / cfg.n_steps
    return dict(log_s0=math.log(cfg.s0), strike=cfg.k,
                drift_dt=(cfg.r - 0.5 * cfg.sigma * cfg.sigma) * dt,
                vol_dt=cfg.sigma * math.sqrt(dt), discount=math.exp(-cfg.r * cfg.t))


@pytest.mark.parametrize("n,m", [(1, 1), (4099, 37), (1000, 16), (257, 200)])
def test_cuda_contract_kernel_matches_oracle_path_by_path(engine, oracle, n, m):
    # the reference's stepwise European contract on the reference's own
    # normal stream, path-major: log S_T is formed in the reference's order
    # with every product rounded, so a payoff differs from the oracle's by
    # the last bits of exp only
    cfg = oracle.Gbm(n_paths=n, n_steps=m, seed=42)
    z = oracle.rng_normals(42, n * m).reshape(n, m)
    _, ref = oracle.gbm_paths_from_normals(cfg, z)
    got = engine.european_call_stepwise_v1(z, **_contract_scalars(cfg))
    assert engine.last_timing()["main_launches"] == 1
    err = np.abs(got - ref)
    assert float(err.max()) <= 4 * np.finfo(float).eps * 400.0, float(err.max())
    assert np.array_equal(got > 0, ref > 0)


def test_cuda_contract_kernel_prices_the_cpu_runnable_case(engine, oracle):
    # 100,000 paths x 64 steps, seed 42, one thread: the contract's payoffs
    # summarised as the reference does equal the oracle's price; the price is
    # within 3 standard errors of Black-Scholes
    n, m = 100_000, 64
    cfg = oracle.Gbm(n_paths=n, n_steps=m, seed=42, n_threads=1)
    z = oracle.rng_normals(42, n * m).reshape(n, m)
    pay = engine.european_call_stepwise_v1(z, **_contract_scalars(cfg))
    price, stderr = oracle.gbm_price(cfg)
    mean = pay.mean()
    se = math.sqrt(max((pay * pay).mean() - mean * mean, 0.0)) / math.sqrt(n)
    assert abs(mean - price) <= 1e-12 * price, (mean, price)
    assert abs(se - stderr) <= 1e-9 * stderr
    assert abs(mean - 9.413403383853016) <= 3 * se


def test_cuda_contract_kernel_rejects_invalid_shapes(engine):
    from montepath_amd import _abi as A

    with pytest.raises(A.MontepathHipError):
        engine.european_call_stepwise_v1(np.zeros((0, 4)), 0.0, 1.0, 0.0, 0.1, 1.0)
    with pytest.raises(A.MontepathHipError):
        engine.european_call_stepwise_v1(np.zeros((4, 0)), 0.0, 1.0, 0.0, 0.1, 1.0)


def test_heston_full_truncation_matches_characteristic_function(engine):
    from oracle import closed_forms as cf

    req = make_request("heston_european_call", n_paths=1 << 19, n_steps=128, seed=1,
                       kappa=2.0, theta=0.04, v0=0.04, xi=0.5, rho=-0.7,
                       heston_scheme="full_truncation", strikes=[80.0, 100.0, 120.0])
    res = engine.price(req)
    for j, k in enumerate((80.0, 100.0, 120.0)):
        ref = cf.heston_call_lewis(100.0, k, 0.03, 0.04, 2.0, 0.04, 0.5, -0.7, 1.0)
        assert abs(res.strike_prices[j] - ref) <= 3 * res.strike_stderrs[j], (k, res.strike_prices[j], ref)


def test_merton_matches_series(engine, oracle):
    req = make_request("merton_jump_diffusion_call", n_paths=1 << 20, n_steps=0, seed=3)
    res = engine.price(req)
    ref = oracle.merton_series(oracle.Merton())
    assert abs(res.price - ref) <= 3 * res.stderr


def test_batch_results_identical_alone_and_in_mix(engine):
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(24):
        reqs.append(make_request(
            "european_call" if i % 2 == 0 else "arithmetic_asian_call",
            n_paths=int(2 ** rng.uniform(12, 16)), n_steps=int(rng.integers(16, 100)),
            k=float(rng.uniform(80, 120)), sigma=float(rng.uniform(0.1, 0.6)),
            t=float(rng.uniform(0.1, 3)), r=float(rng.uniform(0, 0.05)), seed=7, contract_id=i))
    mix = engine.price_batch(reqs)
    for i in (0, 5, 13, 23):
        alone = engine.price(reqs[i])
        assert alone.price == mix[i].price and alone.stderr == mix[i].stderr


def test_prepared_graph_batch_replays_identically(engine):
    reqs = [make_request("european_call", n_paths=1 << 14, n_steps=32, seed=7, contract_id=i,
                         k=90.0 + i) for i in range(8)]
    eager = engine.price_batch(reqs)
    pb = engine.prepare_batch(reqs, use_graph=True)
    for _ in range(3):
        pb.launch()
        pb.sync()
        got = pb.results()
        assert [g.price for g in got] == [e.price for e in eager]
        engine.price(reqs[0])  # eager work between replays
    pb.free()
