"""Python handle on the HIP engine: one ``Engine`` per device and thread.

Thin: every call goes straight to the C-ABI (include/montepath_hip.h).
Requests are built with :func:`make_request` from the same field names the
reference configs use (crates/mc-core/src/runtime/cpu.rs:1296-1860).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A

WORKLOADS = {
    "european_call": A.EUROPEAN_CALL,
    "arithmetic_asian_call": A.ARITHMETIC_ASIAN_CALL,
    "down_and_out_call": A.DOWN_AND_OUT_CALL,
    "lookback_call": A.LOOKBACK_CALL,
    "heston_european_call": A.HESTON_EUROPEAN_CALL,
    "merton_jump_diffusion_call": A.MERTON_JUMP_DIFFUSION_CALL,
    "basket_call": A.BASKET_CALL,
    "american_put": A.AMERICAN_PUT,
    "bermudan_put": A.BERMUDAN_PUT,
}
SAMPLING = {
    "pseudorandom": A.PSEUDORANDOM,
    "randomized_halton": A.RANDOMIZED_HALTON,
    "latin_hypercube": A.LATIN_HYPERCUBE,
    "scrambled_sobol": A.SCRAMBLED_SOBOL,
    "scrambled_sobol_brownian_bridge": A.SCRAMBLED_SOBOL_BRIDGE,
}


# MonteCarloTechnique (runtime/cpu.rs:24-28)
TECHNIQUES = {"standard": 0, "antithetic": 1, "control_variate": 2}


def make_request(workload: str, **kw) -> A.Request:
    """Build an mp_request.  Unknown keys raise; omitted keys keep the
    reference defaults where the reference has one."""
    r = A.Request()
    if workload not in WORKLOADS:
        raise ValueError(f"unknown workload '{workload}'")
    r.workload = WORKLOADS[workload]
    r.precision = int(kw.pop("precision", 32))
    sampling = kw.pop("sampling", "pseudorandom")
    r.sampling = SAMPLING[sampling] if isinstance(sampling, str) else int(sampling)
    technique = kw.pop("technique", "standard")
    r.technique = TECHNIQUES[technique] if isinstance(technique, str) else int(technique)
    method = kw.pop("method", "stepwise")
    r.method = A.METHOD_TERMINAL if method in ("terminal", "terminal_distribution", 1) else A.METHOD_STEPWISE
    scheme = kw.pop("heston_scheme", "reference_floor")
    r.heston_scheme = (A.HESTON_FULL_TRUNCATION if scheme in ("full_truncation", 1)
                       else A.HESTON_REFERENCE_FLOOR)
    strikes = kw.pop("strikes", None)
    r.basis_degree = int(kw.pop("basis_degree", 2))
    r.basis_weighted = int(bool(kw.pop("basis_weighted", False)))
    r.fp64_check = int(bool(kw.pop("fp64_check", False)))
    r.allow_intrinsic_now = int(bool(kw.pop("allow_intrinsic_now", workload == "american_put")))
    r.n_paths = int(kw.pop("n_paths", 100_000))
    r.n_steps = int(kw.pop("n_steps", 0 if workload == "merton_jump_diffusion_call" else 252))
    r.seed = int(kw.pop("seed", 42)) & 0xFFFFFFFFFFFFFFFF
    r.contract_id = int(kw.pop("contract_id", 0))
    r.s0 = float(kw.pop("s0", 100.0))
    r.k = float(kw.pop("k", 100.0))
    r.r = float(kw.pop("r", 0.03))
    r.sigma = float(kw.pop("sigma", 0.2))
    r.t = float(kw.pop("t", 1.0))
    r.barrier = float(kw.pop("barrier", 80.0))
    r.v0 = float(kw.pop("v0", 0.04))
    r.kappa = float(kw.pop("kappa", 1.5))
    r.theta = float(kw.pop("theta", 0.04))
    r.xi = float(kw.pop("xi", kw.pop("vol_of_vol", 0.3)))
    r.rho = float(kw.pop("rho", -0.6))
    r.jump_intensity = float(kw.pop("jump_intensity", 0.4))
    r.jump_mean = float(kw.pop("jump_mean", -0.08))
    r.jump_volatility = float(kw.pop("jump_volatility", 0.25))
    if strikes is not None:
        strikes = list(strikes)
        if not 1 <= len(strikes) <= A.MAX_STRIKES:
            raise ValueError("1..4 strikes")
        r.n_strikes = len(strikes)
        for i, s in enumerate(strikes):
            r.strikes[i] = float(s)
        if len(strikes) == 1:
            r.k = float(strikes[0])
    if workload == "basket_call":
        s0s = list(kw.pop("asset_s0"))
        n = len(s0s)
        sig = list(kw.pop("asset_sigma"))
        w = list(kw.pop("asset_weight", [1.0 / n] * n))
        corr = np.asarray(kw.pop("correlation", np.eye(n)), dtype=np.float64)
        r.n_assets = n
        for i in range(n):
            r.asset_s0[i] = float(s0s[i])
            r.asset_sigma[i] = float(sig[i])
            r.asset_weight[i] = float(w[i])
            for j in range(n):
                r.correlation[i * n + j] = float(corr[i, j])
    sched = kw.pop("exercise_steps", None)
    if sched is not None:
        arr = (C.c_uint32 * len(sched))(*[int(x) for x in sched])
        r.n_exercise = len(sched)
        r.exercise_steps = C.cast(arr, C.POINTER(C.c_uint32))
        r._sched_keepalive = arr  # keep the buffer alive with the request
    if kw:
        raise ValueError(f"unknown request keys: {sorted(kw)}")
    return r


@dataclass
class PriceResult:
    price: float
    stderr: float
    runtime_ms: float
    fp64_ref_error: float
    strike_prices: list = field(default_factory=list)
    strike_stderrs: list = field(default_factory=list)
    n_paths: int = 0
    n_steps: int = 0
    early_exercise_count: int = 0
    maturity_exercise_count: int = 0
    regression_steps: int = 0

    @classmethod
    def of(cls, r: A.Result) -> "PriceResult":
        ns = max(int(r.n_strikes), 1)
        return cls(
            price=r.price, stderr=r.std_error, runtime_ms=r.runtime_ms,
            fp64_ref_error=r.fp64_ref_error,
            strike_prices=[r.strike_price[i] for i in range(ns)],
            strike_stderrs=[r.strike_std_error[i] for i in range(ns)],
            n_paths=int(r.n_paths), n_steps=int(r.n_steps),
            early_exercise_count=int(r.early_exercise_count),
            maturity_exercise_count=int(r.maturity_exercise_count),
            regression_steps=int(r.regression_steps),
        )


GREEK_ESTIMATORS = {
    "bump_and_revalue": A.GREEKS_BUMP_AND_REVALUE,
    "pathwise": A.GREEKS_PATHWISE,
    "likelihood_ratio": A.GREEKS_LIKELIHOOD_RATIO,
}
_ESTIMATOR_NAMES = {v: k for k, v in GREEK_ESTIMATORS.items()}


MLMC_DEFAULTS = {"s0": 100.0, "k": 100.0, "r": 0.03, "sigma": 0.2, "t": 1.0, "base_steps": 16,
                 "levels": 4, "refinement_factor": 2,
                 "paths_per_level": (50_000, 25_000, 12_500, 6_250), "seed": 42,
                 "sampling": "pseudorandom", "scramble_replicates": 1, "precision": 64,
                 "contract_id": 0}


def make_mlmc_config(**kw) -> A.MlmcConfig:
    """mp_mlmc_config from ArithmeticAsianMlmcConfig fields (defaults
    cpu.rs:1551-1568); precision defaults to FP64 like the reference."""
    unknown = set(kw) - set(MLMC_DEFAULTS)
    if unknown:
        raise TypeError(f"unknown MLMC config keys: {sorted(unknown)}")
    v = dict(MLMC_DEFAULTS)
    v.update(kw)
    c = A.MlmcConfig()
    for f in ("s0", "k", "r", "sigma", "t"):
        setattr(c, f, float(v[f]))
    c.base_steps, c.levels = int(v["base_steps"]), int(v["levels"])
    c.refinement_factor = int(v["refinement_factor"])
    c.scramble_replicates = int(v["scramble_replicates"])
    smp = v["sampling"]
    c.sampling = SAMPLING[smp] if isinstance(smp, str) else int(smp)
    c.precision = int(v["precision"])
    c.seed, c.contract_id = int(v["seed"]), int(v["contract_id"])
    paths = list(v["paths_per_level"])
    if len(paths) > A.MLMC_MAX_LEVELS:
        raise ValueError("at most 16 levels")
    for i, p in enumerate(paths):
        c.paths_per_level[i] = int(p)
    return c


def _mlmc_dict(r: A.MlmcResult) -> dict:
    """ArithmeticAsianMlmcResult (cpu.rs:1570-1591) as a dict."""
    lv = []
    for i in range(r.n_levels):
        L = r.levels[i]
        lv.append({"level": int(L.level), "paths": int(L.paths), "fine_steps": int(L.fine_steps),
                   "coarse_steps": int(L.coarse_steps) or None, "mean": L.mean,
                   "variance": L.variance, "stderr": L.std_error,
                   "cost_step_updates": int(L.cost_step_updates)})
    return {"price": r.price, "stderr": r.std_error, "total_paths": int(r.total_paths),
            "total_step_updates": int(r.total_step_updates),
            "scramble_replicates": int(r.scramble_replicates),
            "replicate_estimates": [r.replicate_estimates[i] for i in range(r.scramble_replicates)],
            "levels": lv, "runtime_ms": r.runtime_ms}


def _allocation_dict(a: A.MlmcAllocation) -> dict:
    lv = []
    for i in range(a.n_levels):
        L = a.levels[i]
        lv.append({"level": int(L.level), "fine_steps": int(L.fine_steps),
                   "coarse_steps": int(L.coarse_steps) or None, "pilot_paths": int(L.pilot_paths),
                   "pilot_variance": L.pilot_variance, "cost_per_path": int(L.cost_per_path),
                   "recommended_paths": int(L.recommended_paths)})
    return {"target_step_updates": int(a.target_step_updates),
            "estimated_step_updates": int(a.estimated_step_updates),
            "paths_per_level": [x["recommended_paths"] for x in lv], "levels": lv}


def _greek_dict(e: A.GreekEstimate) -> dict:
    """GreekEstimate (cpu.rs:157-169) as a dict; None where the estimator has no value."""
    opt = lambda x: None if x != x else x
    return {
        "greek": A.GREEK_NAMES[e.greek], "estimator": _ESTIMATOR_NAMES[e.estimator],
        "value": e.value, "stderr": e.std_error, "bump_size": opt(e.bump_size),
        "base_price": e.base_price, "bumped_up_price": opt(e.bumped_up_price),
        "bumped_down_price": opt(e.bumped_down_price),
    }


def _estimator_code(estimator) -> int:
    return GREEK_ESTIMATORS[estimator] if isinstance(estimator, str) else int(estimator)


def _pd(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Engine:
    """One HIP context (stream + resident arenas) on one device."""

    def __init__(self, device: int = 0):
        self.lib = A.load()
        self._ctx = C.c_void_p()
        A.check(self.lib.mp_hip_open(device, C.byref(self._ctx)))
        self.device = device

    def close(self) -> None:
        if self._ctx:
            self.lib.mp_hip_close(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- info
    def device_info(self) -> dict:
        info = A.DeviceInfo()
        A.check(self.lib.mp_hip_device_info(self._ctx, C.byref(info)))
        return {
            "device_id": info.device_id.decode(), "name": info.name.decode(),
            "arch": info.arch.decode(), "memory_total_mb": info.memory_total_mb,
            "memory_free_mb": info.memory_free_mb, "compute_units": info.compute_units,
            "supports_float64": bool(info.supports_float64),
            "max_threads_hint": info.max_threads_hint, "wavefront_size": info.wavefront_size,
        }

    def last_timing(self) -> dict:
        t = A.Timing()
        A.check(self.lib.mp_hip_last_timing(self._ctx, C.byref(t)))
        return {"total_ms": t.total_ms, "main_kernel_ms": t.main_kernel_ms,
                "finalize_ms": t.finalize_ms, "main_launches": t.main_launches}

    # ------------------------------------------------------------- pricing
    def price(self, req: A.Request) -> PriceResult:
        res = A.Result()
        A.check(self.lib.mp_hip_price(self._ctx, C.byref(req), C.byref(res)))
        return PriceResult.of(res)

    def footprint(self, req: A.Request) -> int:
        """HBM bytes the request needs while in flight."""
        b = C.c_uint64()
        A.check(self.lib.mp_hip_request_footprint(C.byref(req), C.byref(b)))
        return b.value

    def price_batch(self, reqs: list) -> list:
        n = len(reqs)
        arr = (A.Request * n)(*reqs)
        res = (A.Result * n)()
        A.check(self.lib.mp_hip_price_batch(self._ctx, arr, n, res))
        return [PriceResult.of(r) for r in res]

    def price_plain_fp64(self, req: A.Request) -> PriceResult:
        res = A.Result()
        A.check(self.lib.mp_hip_price_plain_fp64(self._ctx, C.byref(req), C.byref(res)))
        return PriceResult.of(res)

    def european_call_stepwise_v1(self, normals: np.ndarray, log_s0: float, strike: float,
                                  drift_dt: float, vol_dt: float, discount: float) -> np.ndarray:
        """The reference's CUDA kernel contract: normals [n_paths, n_steps]
        (path-major, FP64) in, each path's discounted payoff out."""
        z = np.ascontiguousarray(normals, dtype=np.float64)
        if z.ndim != 2:
            raise ValueError("normals must be a [n_paths, n_steps] array")
        n_paths, n_steps = z.shape
        out = np.empty(n_paths)
        A.check(self.lib.mp_hip_european_call_stepwise_v1(
            self._ctx, _pd(z), _pd(out), n_paths, n_steps, log_s0, strike, drift_dt, vol_dt,
            discount))
        return out

    def prepare_batch(self, reqs: list, use_graph: bool = True) -> "PreparedBatch":
        return PreparedBatch(self, reqs, use_graph)

    # --------------------------------------------------------- parity hooks
    def paths_from_normals(self, req: A.Request, normals, uniforms=None, jump_normals=None):
        """normals step-major; returns (terminal, undiscounted payoff) per path."""
        n = int(req.n_paths)
        normals = np.ascontiguousarray(normals, dtype=np.float64)
        term = np.empty(n)
        pay = np.empty(n)
        u = None if uniforms is None else np.ascontiguousarray(uniforms, dtype=np.float64)
        j = None if jump_normals is None else np.ascontiguousarray(jump_normals, dtype=np.float64)
        A.check(self.lib.mp_hip_paths_from_normals(
            self._ctx, C.byref(req), _pd(normals), None if u is None else _pd(u),
            None if j is None else _pd(j), _pd(term), _pd(pay)))
        return term, pay

    def price_from_normals(self, req: A.Request, normals, uniforms=None, jump_normals=None):
        """Price from caller draws (step-major, n_paths units: pairs for an
        antithetic request), any technique."""
        normals = np.ascontiguousarray(normals, dtype=np.float64)
        u = None if uniforms is None else np.ascontiguousarray(uniforms, dtype=np.float64)
        j = None if jump_normals is None else np.ascontiguousarray(jump_normals, dtype=np.float64)
        res = A.Result()
        A.check(self.lib.mp_hip_price_from_normals(
            self._ctx, C.byref(req), _pd(normals), None if u is None else _pd(u),
            None if j is None else _pd(j), C.byref(res)))
        return PriceResult.of(res)

    def greeks(self, reqs, estimator="bump_and_revalue"):
        """Greeks of one request or a list: [(base PriceResult, [estimate dict])].
        Bump-and-revalue legs of every request run as one batch."""
        single = isinstance(reqs, A.Request)
        reqs = [reqs] if single else list(reqs)
        n = len(reqs)
        arr = (A.Request * n)(*reqs)
        base = (A.Result * n)()
        est = (A.GreekEstimate * (n * A.MAX_GREEKS))()
        counts = (C.c_int32 * n)()
        A.check(self.lib.mp_hip_greeks(self._ctx, arr, n, _estimator_code(estimator), base, est,
                                       counts))
        out = [(PriceResult.of(base[i]),
                [_greek_dict(est[i * A.MAX_GREEKS + j]) for j in range(counts[i])])
               for i in range(n)]
        return out[0] if single else out

    def greeks_from_normals(self, req: A.Request, estimator, normals):
        """Pathwise / likelihood-ratio Greeks of a European call from one
        terminal normal per path; returns (base PriceResult, [estimate dict])."""
        z = np.ascontiguousarray(normals, dtype=np.float64)
        assert z.size == int(req.n_paths)
        base = A.Result()
        est = (A.GreekEstimate * A.MAX_GREEKS)()
        count = C.c_int32()
        A.check(self.lib.mp_hip_greeks_from_normals(self._ctx, C.byref(req),
                                                    _estimator_code(estimator), _pd(z),
                                                    C.byref(base), est, C.byref(count)))
        return PriceResult.of(base), [_greek_dict(est[j]) for j in range(count.value)]

    # ------------------------------------------------------------ multilevel
    def mlmc_price(self, cfg: A.MlmcConfig) -> dict:
        out = A.MlmcResult()
        A.check(self.lib.mp_hip_mlmc_price(self._ctx, C.byref(cfg), C.byref(out)))
        return _mlmc_dict(out)

    def mlmc_tune_allocation(self, cfg: A.MlmcConfig, target_step_updates: int,
                             pilot_paths: int) -> dict:
        out = A.MlmcAllocation()
        A.check(self.lib.mp_hip_mlmc_tune_allocation(self._ctx, C.byref(cfg), target_step_updates,
                                                     pilot_paths, C.byref(out)))
        return _allocation_dict(out)

    def mlmc_solve_tolerance(self, cfg: A.MlmcConfig, target_stderr: float,
                             pilot_paths_per_level: int = 2_048, min_step_updates: int = 100_000,
                             max_step_updates: int = 10_000_000) -> dict:
        tol = A.MlmcTolerance(target_stderr, pilot_paths_per_level, min_step_updates,
                              max_step_updates)
        out = A.MlmcTolerancePlan()
        A.check(self.lib.mp_hip_mlmc_solve_tolerance(self._ctx, C.byref(cfg), C.byref(tol),
                                                     C.byref(out)))
        alloc = _allocation_dict(out.allocation)
        return {"target_stderr": out.target_stderr, "estimated_stderr": out.estimated_stderr,
                "target_met": bool(out.target_met),
                "max_step_updates_hit": bool(out.max_step_updates_hit),
                "scramble_replicates": int(out.scramble_replicates),
                "estimated_step_updates": int(out.estimated_step_updates),
                "paths_per_level": alloc["paths_per_level"], "allocation": alloc,
                "recommended_config": out.recommended}

    def mlmc_level_from_normals(self, cfg: A.MlmcConfig, level: int, normals) -> tuple:
        """(sum, sum of squares) of one level's discounted terms; normals
        [path][fine step] (transposed to the engine's step-major layout)."""
        z = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).T)
        out = np.zeros(2)
        A.check(self.lib.mp_hip_mlmc_level_from_normals(self._ctx, C.byref(cfg), level, z.shape[1],
                                                        _pd(z), _pd(out)))
        return out[0], out[1]

    def price_with_paths(self, req: A.Request):
        """Price and return (result, terminal values, undiscounted payoffs)."""
        n = int(req.n_paths)
        term = np.empty(n)
        pay = np.empty(n)
        res = A.Result()
        A.check(self.lib.mp_hip_price_with_paths(self._ctx, C.byref(req), C.byref(res), _pd(term),
                                                 _pd(pay)))
        return PriceResult.of(res), term, pay

    def lsm_from_normals(self, req: A.Request, normals):
        n, m = int(req.n_paths), int(req.n_steps)
        normals = np.ascontiguousarray(normals, dtype=np.float64)
        assert normals.size == n * m
        res = A.Result()
        coeff = np.zeros((m, 4))
        ex = np.zeros(n, dtype=np.uint32)
        A.check(self.lib.mp_hip_lsm_from_normals(
            self._ctx, C.byref(req), _pd(normals), C.byref(res), _pd(coeff),
            ex.ctypes.data_as(C.POINTER(C.c_uint32))))
        return PriceResult.of(res), coeff, ex

    def draws(self, req: A.Request, kind: int, stream: int, path0: int, n_paths: int,
              n_draws: int) -> np.ndarray:
        out = np.empty((n_paths, n_draws))
        A.check(self.lib.mp_hip_draws(self._ctx, C.byref(req), kind, stream, path0, n_paths,
                                      n_draws, _pd(out)))
        return out

    # -------------------------------------------------------------- tuning
    def autotune(self, req: A.Request):
        best = A.Tuning()
        ms = C.c_double()
        A.check(self.lib.mp_hip_autotune(self._ctx, C.byref(req), C.byref(best), C.byref(ms)))
        return {"block_size": best.block_size, "paths_per_lane": best.paths_per_lane,
                "tile_paths": best.tile_paths, "step_unroll": best.step_unroll}, ms.value

    def get_tuning(self, req: A.Request) -> dict:
        t = A.Tuning()
        A.check(self.lib.mp_hip_get_tuning(self._ctx, C.byref(req), C.byref(t)))
        return {"block_size": t.block_size, "paths_per_lane": t.paths_per_lane,
                "tile_paths": t.tile_paths, "step_unroll": t.step_unroll}

    def set_tuning(self, req: A.Request, block_size: int, paths_per_lane: int,
                   step_unroll: int = 1) -> None:
        t = A.Tuning(block_size, paths_per_lane, 0, step_unroll)
        A.check(self.lib.mp_hip_set_tuning(self._ctx, C.byref(req), C.byref(t)))

    # ------------------------------------------------------------- serving
    def serve_trace(self, reqs: list, arrival_s, hbm_budget_bytes: int = 0,
                    max_inflight: int = 2, max_batch: int = 4096):
        n = len(reqs)
        arr = (A.Request * n)(*reqs)
        res = (A.Result * n)()
        arrival = np.ascontiguousarray(arrival_s, dtype=np.float64)
        lat = np.empty(n)
        cfg = A.ServeConfig(hbm_budget_bytes, max_inflight, max_batch)
        st = A.ServeStats()
        A.check(self.lib.mp_hip_serve_trace(self._ctx, C.byref(cfg), arr, _pd(arrival), n, res,
                                            _pd(lat), C.byref(st)))
        stats = {f: getattr(st, f) for f, _ in A.ServeStats._fields_}
        return [PriceResult.of(r) for r in res], lat, stats

    def server(self, hbm_budget_bytes: int = 0, max_inflight: int = 2,
               max_batch: int = 4096) -> "Server":
        return Server(self, hbm_budget_bytes, max_inflight, max_batch)


class Server:
    """A live continuous-batching server on the engine's device."""

    def __init__(self, eng: Engine, hbm_budget_bytes: int, max_inflight: int, max_batch: int):
        self.lib = eng.lib
        self._s = C.c_void_p()
        cfg = A.ServeConfig(hbm_budget_bytes, max_inflight, max_batch)
        A.check(self.lib.mp_hip_server_open(eng._ctx, C.byref(cfg), C.byref(self._s)))

    def submit(self, req: A.Request) -> int:
        t = C.c_uint64()
        A.check(self.lib.mp_hip_submit(self._s, C.byref(req), C.byref(t)))
        return t.value

    def wait(self, ticket: int) -> "PriceResult":
        out = A.Result()
        A.check(self.lib.mp_hip_wait(self._s, ticket, C.byref(out)))
        return PriceResult.of(out)

    def stats(self) -> dict:
        st = A.ServeStats()
        A.check(self.lib.mp_hip_server_stats(self._s, C.byref(st)))
        return {"completed": st.completed, "launches": st.launches,
                "max_admitted_bytes": st.max_admitted_bytes}

    def close(self) -> None:
        if self._s:
            self.lib.mp_hip_server_close(self._s)
            self._s = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class PreparedBatch:
    """Requests resident in HBM, replayed as one hipGraph launch."""

    def __init__(self, eng: Engine, reqs: list, use_graph: bool = True):
        self.eng = eng
        self.lib = eng.lib
        self.n = len(reqs)
        arr = (A.Request * self.n)(*reqs)
        self._b = C.c_void_p()
        A.check(self.lib.mp_hip_batch_prepare(eng._ctx, arr, self.n, int(use_graph),
                                              C.byref(self._b)))

    def info(self) -> dict:
        nt, ps = C.c_uint64(), C.c_uint64()
        g, b = C.c_int(), C.c_int()
        A.check(self.lib.mp_hip_batch_info(self._b, C.byref(nt), C.byref(ps), C.byref(g),
                                           C.byref(b)))
        return {"n_tiles": nt.value, "path_steps": ps.value, "grid": g.value, "block": b.value}

    def launch(self) -> None:
        A.check(self.lib.mp_hip_batch_launch(self._b))

    def sync(self) -> None:
        A.check(self.lib.mp_hip_batch_sync(self._b))

    def results(self) -> list:
        res = (A.Result * self.n)()
        A.check(self.lib.mp_hip_batch_results(self._b, res))
        return [PriceResult.of(r) for r in res]

    def free(self) -> None:
        if self._b:
            self.lib.mp_hip_batch_free(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def is_nan(x: float) -> bool:
    return isinstance(x, float) and math.isnan(x)
