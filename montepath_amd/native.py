"""``price_*_rocm`` native-bridge functions: the rocm_hip counterparts of the
reference's ``_native.price_*`` / ``price_*_metal`` PyO3 functions.

Same contract as crates/montepath-python/src/lib.rs:
  * input: any mapping with ``.get``; keys and aliases as the reference's
    config parsers (fill_gbm_fields / fill_path_fields lib.rs:652-678,
    heston/merton/basket/american/bermudan parsers lib.rs:380-480), defaults
    from the reference configs (runtime/cpu.rs:1296-1860);
  * output: ``{price, stderr, values, manifest, warnings}`` shaped like
    metal_pricing_response (lib.rs:730-768) with backend ``rocm_hip``;
  * errors: ``ValueError`` (the reference raises PyValueError), including for
    configs the reference's validators reject with a panic.

Every call runs on the GPU through the C-ABI; there is no CPU fallback.
"""
from __future__ import annotations

import math
import threading
from typing import Any, Callable, Mapping

from . import _abi as A
from .engine import MLMC_DEFAULTS, Engine, make_mlmc_config, make_request

KNOWN_ROCM_FUNCTIONS = (
    "price_european_call_rocm",
    "price_arithmetic_asian_call_rocm",
    "price_down_and_out_call_rocm",
    "price_lookback_call_rocm",
    "price_heston_european_call_rocm",
    "price_merton_jump_diffusion_call_rocm",
    "price_basket_call_rocm",
    "price_american_put_rocm",
    "price_bermudan_put_rocm",
    "greeks_rocm",
    "price_all_current_greeks_bump_and_revalue_rocm",
    "arithmetic_asian_mlmc_rocm",
    "solve_arithmetic_asian_mlmc_tolerance_rocm",
    "compare_arithmetic_asian_mlmc_reference_rocm",
)

NATIVE_MODULE = "montepath_amd._lib.libmontepath_hip"
REPRODUCIBILITY = (
    "bit-reproducible for identical config, seed, contract id, library version and "
    "MI355X (gfx950): counter-based draws and fixed-order FP64 reductions"
)

_local = threading.local()


def _engine() -> Engine:
    """One engine (HIP context) per host thread, like the reference's
    thread-local Metal runtime cache (backend/metal.rs:52-55)."""
    eng = getattr(_local, "engine", None)
    if eng is None:
        eng = Engine(0)
        _local.engine = eng
    return eng


def set_engine_factory(factory: Callable[[], Any]) -> None:
    """Install the engine used by this thread (tests inject doubles here)."""
    _local.engine = factory()


# ----------------------------------------------------------------- parsing
def _lookup(cfg: Mapping, keys):
    for k in keys:
        v = cfg.get(k)
        if v is not None:
            return v
    return None


def _f(cfg, keys, default):
    v = _lookup(cfg, keys)
    return float(v) if v is not None else default


def _u(cfg, keys, default):
    v = _lookup(cfg, keys)
    if v is None:
        return default
    if isinstance(v, bool) or int(v) != v or int(v) < 0:
        raise ValueError(f"'{keys[0]}' must be a non-negative integer")
    return int(v)


def _normalize(s: str) -> str:
    return s.strip().lower().replace("-", "_")


def _technique(cfg) -> str:
    v = _lookup(cfg, ["technique"])
    t = "standard" if v is None else _normalize(str(v))
    if t not in ("standard", "antithetic", "control_variate"):
        raise ValueError(f"unsupported Monte Carlo technique '{v}'")
    return t


def _sampling(cfg) -> str:
    v = _lookup(cfg, ["sampling"])
    s = "pseudorandom" if v is None else _normalize(str(v))
    if s not in ("pseudorandom", "randomized_halton", "latin_hypercube", "scrambled_sobol",
                 "scrambled_sobol_brownian_bridge"):
        raise ValueError(f"unsupported sampling method '{v}'")
    return s


def _gbm(cfg, steps_default=252):
    return {
        "s0": _f(cfg, ["s0", "spot"], 100.0), "k": _f(cfg, ["k", "strike"], 100.0),
        "r": _f(cfg, ["r", "rate"], 0.03), "sigma": _f(cfg, ["sigma", "volatility"], 0.2),
        "t": _f(cfg, ["t", "maturity"], 1.0),
        "n_paths": _u(cfg, ["n_paths", "paths"], 100_000),
        "n_steps": _u(cfg, ["n_steps", "steps"], steps_default),
        "seed": _u(cfg, ["seed"], 42), "n_threads": _u(cfg, ["n_threads"], 0),
    }


def _engine_keys(cfg: Mapping) -> dict:
    """rocm_hip-only keys: precision (32/64), contract_id, fp64_check, strikes."""
    out = {}
    p = cfg.get("precision")
    if p is not None:
        p = str(p).lower()
        out["precision"] = {"32": 32, "fp32": 32, "float32": 32, "64": 64, "fp64": 64,
                            "float64": 64}.get(p)
        if out["precision"] is None:
            raise ValueError(f"unsupported precision '{cfg.get('precision')}'")
    if cfg.get("contract_id") is not None:
        out["contract_id"] = _u(cfg, ["contract_id"], 0)
    if cfg.get("fp64_check") is not None:
        out["fp64_check"] = bool(cfg.get("fp64_check"))
    if cfg.get("strikes") is not None:
        out["strikes"] = [float(x) for x in cfg.get("strikes")]
    return out




# ---------------------------------------------------------------- response
def _response(workload: str, function: str, config: dict, res, extra_values=None,
              warnings=None) -> dict:
    values = {"price": res.price, "stderr": res.stderr, "runtime_ms": res.runtime_ms}
    if len(res.strike_prices) > 1:
        values["strike_prices"] = list(res.strike_prices)
        values["strike_stderrs"] = list(res.strike_stderrs)
    if res.fp64_ref_error == res.fp64_ref_error:  # not NaN
        values["fp64_ref_error"] = res.fp64_ref_error
    if extra_values:
        values.update(extra_values)
    return {
        "price": res.price,
        "stderr": res.stderr,
        "values": values,
        "manifest": {
            "schema_version": "python-native-runtime.v1",
            "workload": workload,
            "backend": "rocm_hip",
            "execution_mode": "native_extension_rocm",
            "function": function,
            "native_module": NATIVE_MODULE,
            "config": config,
            "reproducibility": REPRODUCIBILITY,
            "performance_claim": "use benchmark artifacts for timing claims",
        },
        "warnings": list(warnings or []) + [
            "rocm_hip prices with counter-based Philox or scrambled-Sobol draws; results agree "
            "with the CPU reference statistically, not path by path",
        ],
    }


def _run(workload: str, function: str, config: dict, req_kw: dict, **resp):
    try:
        req = make_request(workload, **req_kw)
        res = _engine().price(req)
    except A.MontepathHipError as err:
        raise ValueError(f"native ROCm execution failed: {err}") from err
    extra = resp.pop("extra", None)
    if extra:
        extra = {k: getattr(res, k) for k in extra}
    return _response(workload, function, config, res, extra_values=extra, **resp)


def _gbm_family(workload: str, function: str, cfg: Mapping, barrier: bool = False):
    c = _gbm(cfg)
    c["technique"] = _technique(cfg)
    c["sampling"] = _sampling(cfg)
    if barrier:
        c["barrier"] = _f(cfg, ["barrier"], 80.0)
    kw = dict(n_paths=c["n_paths"], n_steps=c["n_steps"], seed=c["seed"], s0=c["s0"], k=c["k"],
              r=c["r"], sigma=c["sigma"], t=c["t"], sampling=c["sampling"],
              technique=c["technique"])
    if barrier:
        kw["barrier"] = c["barrier"]
    ek = _engine_keys(cfg)
    kw.update(ek)
    c.update(ek)
    return c, kw


# ------------------------------------------------------------- functions
def _european_call_parts(config: Mapping):
    """European call.  The reference's bridge prices by the terminal
    distribution (Auto, runtime/cpu.rs:3020-3023); ``method`` may ask for
    ``stepwise_paths`` (the default here, the path engine's workload)."""
    c, kw = _gbm_family("european_call", "price_european_call_rocm", config)
    m = config.get("method")
    method = "stepwise_paths" if m is None else _normalize(str(m))
    if method not in ("auto", "terminal_distribution", "stepwise_paths"):
        raise ValueError(f"unsupported European call method '{m}'")
    c["method"] = method
    kw["method"] = "terminal" if method in ("auto", "terminal_distribution") else "stepwise"
    return c, kw


def _arithmetic_asian_call_parts(config: Mapping):
    c, kw = _gbm_family("arithmetic_asian_call", "price_arithmetic_asian_call_rocm", config)
    return c, kw


def _down_and_out_call_parts(config: Mapping):
    c, kw = _gbm_family("down_and_out_call", "price_down_and_out_call_rocm", config, barrier=True)
    return c, kw


def _lookback_call_parts(config: Mapping):
    c, kw = _gbm_family("lookback_call", "price_lookback_call_rocm", config)
    return c, kw


def _heston_european_call_parts(config: Mapping):
    c = {
        "s0": _f(config, ["s0", "spot"], 100.0), "k": _f(config, ["k", "strike"], 100.0),
        "r": _f(config, ["r", "rate"], 0.03), "v0": _f(config, ["v0", "initial_variance"], 0.04),
        "kappa": _f(config, ["kappa", "mean_reversion"], 1.5),
        "theta": _f(config, ["theta", "long_run_variance"], 0.04),
        "vol_of_vol": _f(config, ["vol_of_vol"], 0.3),
        "rho": _f(config, ["rho", "correlation"], -0.6), "t": _f(config, ["t", "maturity"], 1.0),
        "n_paths": _u(config, ["n_paths", "paths"], 100_000),
        "n_steps": _u(config, ["n_steps", "steps"], 252), "seed": _u(config, ["seed"], 42),
        "n_threads": _u(config, ["n_threads"], 0), "technique": _technique(config),
    }
    scheme = _normalize(str(config.get("scheme", "reference_floor")))
    if scheme not in ("reference_floor", "full_truncation"):
        raise ValueError(f"unsupported Heston scheme '{scheme}'")
    c["scheme"] = scheme
    sampling = _sampling(config)
    c["sampling"] = sampling
    kw = dict(n_paths=c["n_paths"], n_steps=c["n_steps"], seed=c["seed"], s0=c["s0"], k=c["k"],
              r=c["r"], t=c["t"], v0=c["v0"], kappa=c["kappa"], theta=c["theta"],
              xi=c["vol_of_vol"], rho=c["rho"], heston_scheme=scheme, sampling=sampling,
              technique=c["technique"])
    ek = _engine_keys(config)
    kw.update(ek)
    c.update(ek)
    return c, kw


def _merton_jump_diffusion_call_parts(config: Mapping):
    c = {
        "s0": _f(config, ["s0", "spot"], 100.0), "k": _f(config, ["k", "strike"], 100.0),
        "r": _f(config, ["r", "rate"], 0.03), "sigma": _f(config, ["sigma", "volatility"], 0.2),
        "jump_intensity": _f(config, ["jump_intensity"], 0.4),
        "jump_mean": _f(config, ["jump_mean"], -0.08),
        "jump_volatility": _f(config, ["jump_volatility"], 0.25),
        "t": _f(config, ["t", "maturity"], 1.0),
        "n_paths": _u(config, ["n_paths", "paths"], 100_000), "seed": _u(config, ["seed"], 42),
        "n_threads": _u(config, ["n_threads"], 0),
        # the reference's Merton is terminal; n_steps > 1 walks compound-Poisson steps
        "n_steps": _u(config, ["n_steps", "steps"], 0),
    }
    kw = dict(n_paths=c["n_paths"], n_steps=c["n_steps"], seed=c["seed"], s0=c["s0"], k=c["k"],
              r=c["r"], sigma=c["sigma"], t=c["t"], jump_intensity=c["jump_intensity"],
              jump_mean=c["jump_mean"], jump_volatility=c["jump_volatility"])
    ek = _engine_keys(config)
    kw.update(ek)
    c.update(ek)
    return c, kw


def _basket_call_parts(config: Mapping):
    """Two-asset basket with the reference's keys (s01, s02, sigma1, sigma2,
    rho, weight1, weight2), or n assets via asset_s0 / asset_sigma /
    asset_weight / correlation lists.  n_steps (default 1) walks the assets
    stepwise; one step is the reference's terminal basket."""
    technique = _technique(config)
    sampling = _sampling(config)
    if sampling.startswith("scrambled_sobol"):
        raise ValueError("rocm_hip prices baskets with pseudorandom, randomized Halton or Latin "
                         "hypercube draws")
    if config.get("asset_s0") is not None:
        s0 = [float(x) for x in config.get("asset_s0")]
        sig = [float(x) for x in config.get("asset_sigma")]
        n = len(s0)
        w = [float(x) for x in config.get("asset_weight", [1.0 / n] * n)]
        corr = config.get("correlation")
        if corr is None or isinstance(corr, (int, float)):
            rho = 0.0 if corr is None else float(corr)
            corr = [[1.0 if i == j else rho for j in range(n)] for i in range(n)]
    else:
        s0 = [_f(config, ["s01", "spot_1", "spot1"], 100.0), _f(config, ["s02", "spot_2", "spot2"], 95.0)]
        sig = [_f(config, ["sigma1", "volatility_1", "volatility1"], 0.2),
               _f(config, ["sigma2", "volatility_2", "volatility2"], 0.25)]
        w = [_f(config, ["weight1", "weight_1"], 0.5), _f(config, ["weight2", "weight_2"], 0.5)]
        rho = _f(config, ["rho", "correlation"], 0.35)
        corr = [[1.0, rho], [rho, 1.0]]
    c = {
        "asset_s0": s0, "asset_sigma": sig, "asset_weight": w, "correlation": corr,
        "k": _f(config, ["k", "strike"], 100.0), "r": _f(config, ["r", "rate"], 0.03),
        "t": _f(config, ["t", "maturity"], 1.0),
        "n_paths": _u(config, ["n_paths", "paths"], 100_000),
        "n_steps": _u(config, ["n_steps", "steps"], 1), "seed": _u(config, ["seed"], 42),
        "n_threads": _u(config, ["n_threads"], 0), "technique": technique, "sampling": sampling,
    }
    kw = dict(n_paths=c["n_paths"], n_steps=c["n_steps"], seed=c["seed"], k=c["k"], r=c["r"],
              t=c["t"], asset_s0=s0, asset_sigma=sig, asset_weight=w, correlation=corr,
              technique=technique, sampling=sampling)
    ek = _engine_keys(config)
    kw.update(ek)
    c.update(ek)
    return c, kw


_LSM_DETAILS = ["early_exercise_count", "maturity_exercise_count", "regression_steps"]


def _american_put_parts(config: Mapping):
    c = _gbm(config, steps_default=64)
    c["basis_degree"] = _u(config, ["basis_degree"], 2)
    c["basis_weighted"] = bool(config.get("basis_weighted", False))
    kw = dict(n_paths=c["n_paths"], n_steps=c["n_steps"], seed=c["seed"], s0=c["s0"], k=c["k"],
              r=c["r"], sigma=c["sigma"], t=c["t"], basis_degree=c["basis_degree"],
              basis_weighted=c["basis_weighted"])
    ek = _engine_keys(config)
    kw.update(ek)
    c.update(ek)
    return c, kw


def _bermudan_put_parts(config: Mapping):
    c = _gbm(config, steps_default=64)
    c["basis_degree"] = _u(config, ["basis_degree"], 2)
    c["basis_weighted"] = bool(config.get("basis_weighted", False))
    sched = config.get("exercise_steps")
    c["exercise_steps"] = [16, 32, 48, 64] if sched is None else [int(x) for x in sched]
    kw = dict(n_paths=c["n_paths"], n_steps=c["n_steps"], seed=c["seed"], s0=c["s0"], k=c["k"],
              r=c["r"], sigma=c["sigma"], t=c["t"], basis_degree=c["basis_degree"],
              basis_weighted=c["basis_weighted"], exercise_steps=c["exercise_steps"])
    ek = _engine_keys(config)
    kw.update(ek)
    c.update(ek)
    return c, kw


# ------------------------------------------------- the price_*_rocm surface
def _priced(workload: str, parts, extra=None):
    fn = f"price_{workload}_rocm"

    def price(config: Mapping) -> dict:
        c, kw = parts(config)
        return _run(workload, fn, c, kw, extra=extra)

    price.__name__ = fn
    price.__qualname__ = fn
    price.__doc__ = parts.__doc__
    return price


price_european_call_rocm = _priced("european_call", _european_call_parts)
price_arithmetic_asian_call_rocm = _priced("arithmetic_asian_call", _arithmetic_asian_call_parts)
price_down_and_out_call_rocm = _priced("down_and_out_call", _down_and_out_call_parts)
price_lookback_call_rocm = _priced("lookback_call", _lookback_call_parts)
price_heston_european_call_rocm = _priced("heston_european_call", _heston_european_call_parts)
price_merton_jump_diffusion_call_rocm = _priced("merton_jump_diffusion_call",
                                                _merton_jump_diffusion_call_parts)
price_basket_call_rocm = _priced("basket_call", _basket_call_parts)
price_american_put_rocm = _priced("american_put", _american_put_parts, extra=_LSM_DETAILS)
price_bermudan_put_rocm = _priced("bermudan_put", _bermudan_put_parts, extra=_LSM_DETAILS)


# ------------------------------------------------------------------- Greeks
_GREEK_PARTS = {
    "european_call": _european_call_parts,
    "arithmetic_asian_call": _arithmetic_asian_call_parts,
    "down_and_out_call": _down_and_out_call_parts,
    "lookback_call": _lookback_call_parts,
    "basket_call": _basket_call_parts,
    "heston_european_call": _heston_european_call_parts,
}
_SINGLE_PASS_NOTE = (
    "single-pass Greeks use the terminal draws of the request's pseudorandom stream; the "
    "configured technique and sampling apply to the base price only")


def _greek_request(workload: str, config: Mapping):
    parts = _GREEK_PARTS.get(workload)
    if parts is None:
        raise ValueError(f"Greeks are offered for {', '.join(_GREEK_PARTS)}; not '{workload}'")
    cfg = dict(config)
    if workload == "european_call" and cfg.get("method") is None:
        cfg["method"] = "auto"  # european_call_price_mc_cpu's default (terminal distribution)
    c, kw = parts(cfg)
    return c, make_request(workload, **kw)


def _greek_report(workload: str, estimator: str, c: dict, base, estimates) -> dict:
    """GreekReport (runtime/cpu.rs:171-183) with backend rocm_hip."""
    warnings = []
    if estimator != "bump_and_revalue":
        warnings.append(_SINGLE_PASS_NOTE)
    return {
        "workload": workload, "estimator": estimator, "backend": "rocm_hip",
        "paths": c["n_paths"], "steps": c["n_steps"], "seed": c["seed"],
        "base_price": base.price, "base_stderr": base.stderr, "estimated_runtime_ms": None,
        "estimates": estimates, "warnings": warnings,
    }


def greeks_rocm(workload: str, config: Mapping, estimator: str = "bump_and_revalue") -> dict:
    """Monte Carlo Greeks of one workload on the GPU (european_call_greeks_cpu,
    arithmetic_asian_call_greeks_cpu, down_and_out_call_greeks_cpu,
    lookback_call_greeks_cpu, basket_call_greeks_cpu,
    heston_european_call_greeks_cpu; runtime/cpu.rs:4036-4356).  Bump and
    revalue prices every leg with common random numbers in one batch;
    pathwise and likelihood ratio are European-call single-pass estimators.
    The European default method is the reference's (terminal distribution)."""
    est = _normalize(str(estimator))
    if est not in ("bump_and_revalue", "pathwise", "likelihood_ratio"):
        raise ValueError(f"unsupported Greek estimator '{estimator}'")
    c, req = _greek_request(workload, config)
    try:
        base, estimates = _engine().greeks(req, est)
    except A.MontepathHipError as err:
        raise ValueError(f"native ROCm execution failed: {err}") from err
    return _greek_report(workload, est, c, base, estimates)


def price_all_current_greeks_bump_and_revalue_rocm(n_paths: int, n_steps: int,
                                                   seed: int) -> list:
    """price_all_current_greeks_bump_and_revalue_cpu (runtime/cpu.rs:4358-4407):
    default configs of six workloads with seeds seed..seed+5, every bumped leg
    of all six in one GPU batch."""
    cfgs = [
        ("european_call", {"n_paths": n_paths, "n_steps": n_steps, "seed": seed}),
        ("arithmetic_asian_call", {"n_paths": n_paths, "n_steps": n_steps, "seed": seed + 1}),
        ("down_and_out_call", {"n_paths": n_paths, "n_steps": n_steps, "seed": seed + 2}),
        ("lookback_call", {"n_paths": n_paths, "n_steps": n_steps, "seed": seed + 3}),
        ("basket_call", {"n_paths": n_paths, "seed": seed + 4}),
        ("heston_european_call", {"n_paths": n_paths, "n_steps": n_steps, "seed": seed + 5}),
    ]
    built = [(w,) + _greek_request(w, cfg) for w, cfg in cfgs]
    try:
        out = _engine().greeks([req for _, _, req in built], "bump_and_revalue")
    except A.MontepathHipError as err:
        raise ValueError(f"native ROCm execution failed: {err}") from err
    return [_greek_report(w, "bump_and_revalue", c, base, est)
            for (w, c, _), (base, est) in zip(built, out)]


# --------------------------------------------------------- multilevel Asian
def _mlmc_parts(config: Mapping):
    """asian_mlmc_config (lib.rs:490-510): GBM keys and aliases, base_steps,
    levels, refinement_factor, paths_per_level, seed, sampling,
    scramble_replicates (defaults cpu.rs:1551-1568)."""
    d = MLMC_DEFAULTS
    c = {
        "s0": _f(config, ["s0", "spot"], d["s0"]), "k": _f(config, ["k", "strike"], d["k"]),
        "r": _f(config, ["r", "rate"], d["r"]), "sigma": _f(config, ["sigma", "volatility"], d["sigma"]),
        "t": _f(config, ["t", "maturity"], d["t"]),
        "base_steps": _u(config, ["base_steps"], d["base_steps"]),
        "levels": _u(config, ["levels"], d["levels"]),
        "refinement_factor": _u(config, ["refinement_factor"], d["refinement_factor"]),
        "seed": _u(config, ["seed"], d["seed"]), "sampling": _sampling(config),
        "scramble_replicates": _u(config, ["scramble_replicates"], d["scramble_replicates"]),
    }
    ppl = config.get("paths_per_level")
    c["paths_per_level"] = list(d["paths_per_level"]) if ppl is None else [int(x) for x in ppl]
    if len(c["paths_per_level"]) != c["levels"]:
        raise ValueError("paths_per_level must contain one positive path count per level")
    ek = _engine_keys(config)
    kw = dict(c)
    kw["precision"] = ek.get("precision", 64)
    if "contract_id" in ek:
        kw["contract_id"] = ek["contract_id"]
    c.update(ek)
    return c, kw


def _mlmc_call(fn):
    try:
        return fn()
    except A.MontepathHipError as err:
        raise ValueError(f"native ROCm execution failed: {err}") from err


def arithmetic_asian_mlmc_rocm(config: Mapping) -> dict:
    """Multilevel arithmetic Asian (the reference's arithmetic_asian_mlmc,
    lib.rs:218-230 -> arithmetic_asian_call_price_mlmc_cpu): every level of
    every replicate in one GPU launch.  Response as workload_response
    (lib.rs:709-728): {values: ArithmeticAsianMlmcResult, manifest, warnings,
    stderr}."""
    c, kw = _mlmc_parts(config)
    res = _mlmc_call(lambda: _engine().mlmc_price(make_mlmc_config(**kw)))
    return {
        "values": res,
        "manifest": {
            "schema_version": "python-native-runtime.v1", "workload": "arithmetic_asian_mlmc",
            "backend": "rocm_hip", "execution_mode": "native_extension_rocm",
            "function": "arithmetic_asian_mlmc_rocm", "native_module": NATIVE_MODULE,
            "config": c, "reproducibility": REPRODUCIBILITY,
            "performance_claim": "use benchmark artifacts for timing claims",
        },
        "warnings": [],
        "stderr": res["stderr"],
    }


def solve_arithmetic_asian_mlmc_tolerance_rocm(config: Mapping, tolerance: Mapping) -> dict:
    """solve_arithmetic_asian_mlmc_tolerance_cpu (cpu.rs:3352-3423): paths per
    level for a target standard error (defaults cpu.rs:1620-1629)."""
    c, kw = _mlmc_parts(config)
    plan = _mlmc_call(lambda: _engine().mlmc_solve_tolerance(
        make_mlmc_config(**kw), float(tolerance.get("target_stderr", 0.05)),
        int(tolerance.get("pilot_paths_per_level", 2_048)),
        int(tolerance.get("min_step_updates", 100_000)),
        int(tolerance.get("max_step_updates", 10_000_000))))
    rec = dict(c)
    rec["paths_per_level"] = list(plan["paths_per_level"])
    if rec["sampling"] == "pseudorandom":
        rec["scramble_replicates"] = 1
    plan["recommended_config"] = rec
    return plan


def compare_arithmetic_asian_mlmc_reference_rocm(config: Mapping, reference: Mapping) -> dict:
    """compare_arithmetic_asian_mlmc_reference_cpu (cpu.rs:3425-3491): the
    multilevel estimate against a standard Asian run (both on the GPU)."""
    c, kw = _mlmc_parts(config)
    est = _mlmc_call(lambda: _engine().mlmc_price(make_mlmc_config(**kw)))
    rc, rkw = _arithmetic_asian_call_parts(reference)
    ref = _mlmc_call(lambda: _engine().price(make_request("arithmetic_asian_call", **rkw)))
    error = est["price"] - ref.price
    combined = math.sqrt(est["stderr"] ** 2 + ref.stderr ** 2)
    if combined > 0.0:
        units = error / combined
    else:
        units = 0.0 if error == 0.0 else math.copysign(math.inf, error)
    eps = 2.220446049250313e-16
    warnings = []
    if any(abs(c[key] - rc[key]) > eps for key in ("s0", "k", "r", "sigma", "t")):
        warnings.append("MLMC config and reference config do not share the same market parameters")
    finest = c["base_steps"] * c["refinement_factor"] ** (c["levels"] - 1)
    if rc["n_steps"] < finest:
        warnings.append("reference step count is below the finest MLMC level; comparison includes "
                        "time-discretization mismatch")
    if abs(units) > 4.0:
        warnings.append(f"MLMC estimate is {abs(units):.3f} combined standard errors from the "
                        "reference run")
    return {
        "workload": "arithmetic_asian_call", "backend": "rocm_hip",
        "reference_name": "high_budget_arithmetic_asian_standard_mc",
        "sampling": c["sampling"], "levels": c["levels"], "estimate_price": est["price"],
        "estimate_stderr": est["stderr"], "estimate_step_updates": est["total_step_updates"],
        "reference_price": ref.price, "reference_stderr": ref.stderr,
        "reference_paths": rc["n_paths"], "reference_steps": rc["n_steps"], "error": error,
        "abs_error": abs(error), "combined_stderr": combined,
        "error_combined_stderr_units": units, "warnings": warnings,
    }
