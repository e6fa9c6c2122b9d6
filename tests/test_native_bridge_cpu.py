"""The price_*_rocm bridge contract on the CPU, with an engine double.

Checks the parts that mirror crates/montepath-python/src/lib.rs: key aliases
and defaults (fill_gbm_fields :652-678, heston/merton/basket/put parsers),
the response schema of metal_pricing_response (:730-768) with backend
rocm_hip, and ValueError on bad input -- the same style of test double the
reference uses for its Python surface (python/tests/test_python_surface.py).
"""
import math

import pytest

from montepath_amd import _abi as A
from montepath_amd import native
from montepath_amd.engine import PriceResult


class FakeEngine:
    def __init__(self, fail_code=None):
        self.requests = []
        self.fail_code = fail_code

    def price(self, req):
        self.requests.append(req)
        if self.fail_code:
            raise A.MontepathHipError(self.fail_code, "device said no")
        ns = max(req.n_strikes, 1)
        return PriceResult(price=1.25, stderr=0.01, runtime_ms=0.5, fp64_ref_error=math.nan,
                           strike_prices=[1.25] * ns, strike_stderrs=[0.01] * ns,
                           n_paths=req.n_paths, n_steps=req.n_steps, regression_steps=7)

    def greeks(self, reqs, estimator="bump_and_revalue"):
        single = isinstance(reqs, A.Request)
        reqs = [reqs] if single else list(reqs)
        self.requests.extend(reqs)
        self.greek_calls = getattr(self, "greek_calls", []) + [(len(reqs), estimator)]
        out = []
        for req in reqs:
            base = self.price(req)
            names = ["delta", "vega", "rho", "theta"]
            out.append((base, [{"greek": g, "estimator": estimator, "value": 0.5, "stderr": 0.01}
                               for g in names]))
        return out[0] if single else out

    def mlmc_price(self, cfg):
        self.mlmc_cfgs = getattr(self, "mlmc_cfgs", []) + [cfg]
        return {"price": 5.3, "stderr": 0.05, "total_paths": 1, "total_step_updates": 1,
                "scramble_replicates": 1, "replicate_estimates": [5.3],
                "levels": [{"level": 0, "coarse_steps": None}], "runtime_ms": 0.1}


@pytest.fixture
def fake():
    eng = FakeEngine()
    native.set_engine_factory(lambda: eng)
    yield eng
    native.set_engine_factory(lambda: None)


def test_aliases_and_defaults(fake):
    out = native.price_european_call_rocm({"spot": 90.0, "strike": 95.0, "volatility": 0.3,
                                           "maturity": 2.0, "rate": 0.01, "paths": 1000,
                                           "steps": 16})
    r = fake.requests[-1]
    assert (r.s0, r.k, r.sigma, r.t, r.r, r.n_paths, r.n_steps) == (90.0, 95.0, 0.3, 2.0, 0.01, 1000, 16)
    assert r.seed == 42 and r.workload == A.EUROPEAN_CALL and r.method == A.METHOD_STEPWISE
    # primary keys win over aliases, as in lookup() (lib.rs:813-820)
    native.price_european_call_rocm({"s0": 80.0, "spot": 90.0})
    assert fake.requests[-1].s0 == 80.0
    assert out["price"] == 1.25 and out["stderr"] == 0.01


def test_response_schema(fake):
    out = native.price_arithmetic_asian_call_rocm({"n_paths": 10})
    assert set(out) == {"price", "stderr", "values", "manifest", "warnings"}
    assert out["values"] == {"price": 1.25, "stderr": 0.01, "runtime_ms": 0.5}
    m = out["manifest"]
    assert m["schema_version"] == "python-native-runtime.v1"
    assert m["backend"] == "rocm_hip" and m["execution_mode"] == "native_extension_rocm"
    assert m["workload"] == "arithmetic_asian_call"
    assert m["function"] == "price_arithmetic_asian_call_rocm"
    assert m["config"]["n_paths"] == 10 and m["config"]["technique"] == "standard"


def test_terminal_method_and_multi_strike(fake):
    native.price_european_call_rocm({"method": "terminal-distribution"})
    assert fake.requests[-1].method == A.METHOD_TERMINAL
    out = native.price_heston_european_call_rocm({"strikes": [80, 100, 120], "vol_of_vol": 0.5,
                                                  "scheme": "full_truncation"})
    r = fake.requests[-1]
    assert r.n_strikes == 3 and list(r.strikes)[:3] == [80.0, 100.0, 120.0]
    assert r.heston_scheme == A.HESTON_FULL_TRUNCATION and r.xi == 0.5
    assert len(out["values"]["strike_prices"]) == 3


def test_put_details_and_schedule(fake):
    out = native.price_bermudan_put_rocm({"exercise_steps": [8, 16], "steps": 16})
    r = fake.requests[-1]
    assert r.n_exercise == 2 and [r.exercise_steps[i] for i in range(2)] == [8, 16]
    assert out["values"]["regression_steps"] == 7
    native.price_american_put_rocm({})
    r = fake.requests[-1]
    assert r.workload == A.AMERICAN_PUT and r.n_steps == 64 and r.basis_degree == 2


def test_basket_reference_keys(fake):
    native.price_basket_call_rocm({"spot1": 100.0, "spot2": 90.0, "correlation": 0.5})
    r = fake.requests[-1]
    assert r.n_assets == 2 and r.asset_s0[1] == 90.0
    assert r.correlation[1] == 0.5 and r.correlation[2] == 0.5 and r.correlation[0] == 1.0
    assert r.asset_weight[0] == 0.5 and r.n_steps == 1


def test_merton_defaults_are_terminal(fake):
    native.price_merton_jump_diffusion_call_rocm({})
    r = fake.requests[-1]
    assert r.n_steps == 0 and r.jump_intensity == 0.4 and r.jump_mean == -0.08


@pytest.mark.parametrize("fn,cfg,msg", [
    (native.price_european_call_rocm, {"technique": "bogus"}, "unsupported Monte Carlo technique"),
    (native.price_basket_call_rocm, {"sampling": "scrambled_sobol"}, "randomized Halton or Latin"),
    (native.price_european_call_rocm, {"sampling": "bogus"}, "unsupported sampling method"),
    (native.price_european_call_rocm, {"method": "bogus"}, "unsupported European call method"),
    (native.price_european_call_rocm, {"n_paths": -5}, "non-negative integer"),
    (native.price_european_call_rocm, {"precision": "fp16"}, "unsupported precision"),
])
def test_bad_configs_raise_value_error(fake, fn, cfg, msg):
    with pytest.raises(ValueError, match=msg):
        fn(cfg)


def test_technique_reaches_the_request(fake):
    native.price_european_call_rocm({"technique": "antithetic"})
    assert fake.requests[-1].technique == 1
    native.price_heston_european_call_rocm({"technique": "control-variate"})
    assert fake.requests[-1].technique == 2
    native.price_basket_call_rocm({"technique": "antithetic"})
    assert fake.requests[-1].technique == 1


def test_sampling_reaches_the_request(fake):
    native.price_european_call_rocm({"sampling": "latin-hypercube"})
    assert fake.requests[-1].sampling == A.LATIN_HYPERCUBE
    native.price_heston_european_call_rocm({"sampling": "randomized_halton"})
    assert fake.requests[-1].sampling == A.RANDOMIZED_HALTON
    native.price_basket_call_rocm({"sampling": "randomized_halton", "technique": "antithetic"})
    assert fake.requests[-1].sampling == A.RANDOMIZED_HALTON and fake.requests[-1].technique == 1


def test_engine_errors_become_value_errors():
    native.set_engine_factory(lambda: FakeEngine(fail_code=A.MP_E_INVALID_CONFIG))
    try:
        with pytest.raises(ValueError, match="native ROCm execution failed: MP_E_INVALID_CONFIG"):
            native.price_european_call_rocm({})
    finally:
        native.set_engine_factory(lambda: None)


def test_known_function_list_is_complete():
    for name in native.KNOWN_ROCM_FUNCTIONS:
        assert callable(getattr(native, name))


def test_greeks_report_shape_and_european_default_method(fake):
    rep = native.greeks_rocm("european_call", {"n_paths": 5000, "seed": 9})
    r = fake.requests[-1]
    # the reference's Greek functions price European calls by the terminal distribution
    assert r.method == A.METHOD_TERMINAL and r.seed == 9 and r.n_paths == 5000
    assert rep["workload"] == "european_call" and rep["estimator"] == "bump_and_revalue"
    assert rep["backend"] == "rocm_hip" and rep["base_price"] == 1.25
    assert [e["greek"] for e in rep["estimates"]] == ["delta", "vega", "rho", "theta"]
    native.greeks_rocm("european_call", {"method": "stepwise_paths"}, "pathwise")
    assert fake.requests[-1].method == A.METHOD_STEPWISE
    assert fake.greek_calls[-1] == (1, "pathwise")


def test_all_workload_greeks_run_as_one_batch(fake):
    reps = native.price_all_current_greeks_bump_and_revalue_rocm(25_000, 64, 9403)
    assert [r["workload"] for r in reps] == ["european_call", "arithmetic_asian_call",
                                             "down_and_out_call", "lookback_call", "basket_call",
                                             "heston_european_call"]
    assert fake.greek_calls == [(6, "bump_and_revalue")]
    seeds = [r.seed for r in fake.requests[:6]]
    assert seeds == [9403 + i for i in range(6)]
    heston = fake.requests[5]
    assert (heston.kappa, heston.xi, heston.rho) == (1.5, 0.3, -0.6)


@pytest.mark.parametrize("args,msg", [
    (("merton_jump_diffusion_call", {}), "Greeks are offered for"),
    (("european_call", {}, "adjoint"), "unsupported Greek estimator"),
])
def test_bad_greek_requests_raise_value_error(fake, args, msg):
    with pytest.raises(ValueError, match=msg):
        native.greeks_rocm(*args)


def test_mlmc_config_keys_and_defaults(fake):
    out = native.arithmetic_asian_mlmc_rocm({"spot": 95.0, "levels": 3, "base_steps": 8,
                                             "paths_per_level": [400, 200, 100],
                                             "sampling": "latin-hypercube",
                                             "scramble_replicates": 4})
    c = fake.mlmc_cfgs[-1]
    assert (c.s0, c.k, c.levels, c.base_steps, c.refinement_factor) == (95.0, 100.0, 3, 8, 2)
    assert list(c.paths_per_level[:3]) == [400, 200, 100]
    assert c.sampling == A.LATIN_HYPERCUBE and c.scramble_replicates == 4
    # the reference prices multilevel runs in FP64
    assert c.precision == 64 and c.seed == 42
    assert out["stderr"] == 0.05 and out["manifest"]["workload"] == "arithmetic_asian_mlmc"
    native.arithmetic_asian_mlmc_rocm({})
    c = fake.mlmc_cfgs[-1]
    assert list(c.paths_per_level[:4]) == [50_000, 25_000, 12_500, 6_250] and c.base_steps == 16
    with pytest.raises(ValueError, match="one positive path count per level"):
        native.arithmetic_asian_mlmc_rocm({"levels": 3})
