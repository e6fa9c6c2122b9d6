"""ctypes bindings to the C oracle ``oracle/_build/libmp_oracle.so``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Each wrapper names
the reference function it restates; the restatements themselves, with
file:line citations into ``crates/mc-core/src/runtime/cpu.rs``, live in
``oracle/mp_oracle.c``.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libmp_oracle.so")

EUROPEAN, ASIAN, DOWN_AND_OUT, LOOKBACK = 0, 1, 2, 3
PAYOFF_CODES = {
    "european_call": EUROPEAN,
    "arithmetic_asian_call": ASIAN,
    "down_and_out_call": DOWN_AND_OUT,
    "lookback_call": LOOKBACK,
}
HESTON_REFERENCE_FLOOR, HESTON_FULL_TRUNCATION = 0, 1


class Result(C.Structure):
    _fields_ = [("price", C.c_double), ("stderr", C.c_double)]


class GbmCfg(C.Structure):
    _fields_ = [
        ("s0", C.c_double), ("k", C.c_double), ("r", C.c_double), ("sigma", C.c_double),
        ("t", C.c_double), ("barrier", C.c_double),
        ("n_paths", C.c_uint64), ("n_steps", C.c_uint64), ("seed", C.c_uint64),
        ("n_threads", C.c_uint64),
    ]


class HestonCfg(C.Structure):
    _fields_ = [
        ("s0", C.c_double), ("k", C.c_double), ("r", C.c_double), ("v0", C.c_double),
        ("kappa", C.c_double), ("theta", C.c_double), ("xi", C.c_double), ("rho", C.c_double),
        ("t", C.c_double),
        ("n_paths", C.c_uint64), ("n_steps", C.c_uint64), ("seed", C.c_uint64),
        ("n_threads", C.c_uint64),
    ]


class MertonCfg(C.Structure):
    _fields_ = [
        ("s0", C.c_double), ("k", C.c_double), ("r", C.c_double), ("sigma", C.c_double),
        ("lam", C.c_double), ("mu_j", C.c_double), ("sigma_j", C.c_double), ("t", C.c_double),
        ("n_paths", C.c_uint64), ("seed", C.c_uint64), ("n_threads", C.c_uint64),
    ]


class MlmcCfg(C.Structure):
    _fields_ = [("s0", C.c_double), ("k", C.c_double), ("r", C.c_double), ("sigma", C.c_double),
                ("t", C.c_double), ("refinement", C.c_uint64), ("seed", C.c_uint64),
                ("sampling", C.c_int64)]


class BsGreeks(C.Structure):
    _fields_ = [("price", C.c_double), ("delta", C.c_double), ("vega", C.c_double),
                ("rho", C.c_double), ("theta", C.c_double)]


class Basket2Cfg(C.Structure):
    _fields_ = [
        ("s01", C.c_double), ("s02", C.c_double), ("k", C.c_double), ("r", C.c_double),
        ("sigma1", C.c_double), ("sigma2", C.c_double), ("rho", C.c_double),
        ("weight1", C.c_double), ("weight2", C.c_double), ("t", C.c_double),
        ("n_paths", C.c_uint64), ("seed", C.c_uint64),
    ]


class LsmCfg(C.Structure):
    _fields_ = [
        ("s0", C.c_double), ("k", C.c_double), ("r", C.c_double), ("sigma", C.c_double),
        ("t", C.c_double),
        ("n_paths", C.c_uint64), ("n_steps", C.c_uint64), ("seed", C.c_uint64),
        ("basis_degree", C.c_uint64), ("allow_intrinsic_now", C.c_int),
        ("basis_weighted", C.c_int),
    ]


class LsmResult(C.Structure):
    _fields_ = [
        ("price", C.c_double), ("stderr", C.c_double),
        ("early_exercise_count", C.c_uint64), ("maturity_exercise_count", C.c_uint64),
        ("regression_steps", C.c_uint64),
    ]


class BbOp(C.Structure):
    _fields_ = [
        ("left", C.c_uint64), ("mid", C.c_uint64), ("right", C.c_uint64),
        ("normal_index", C.c_uint64),
        ("lw", C.c_double), ("rw", C.c_double), ("sd", C.c_double),
    ]


_lib = None

_D = C.POINTER(C.c_double)
_U64 = C.POINTER(C.c_uint64)
_U32 = C.POINTER(C.c_uint32)


def lib() -> C.CDLL:
    """Load the oracle library (built by ``oracle/Makefile`` via build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"oracle library missing at {LIB_PATH}; run __graft_entry__.build() first"
        )
    L = C.CDLL(LIB_PATH)
    sig = {
        "mpo_splitmix64": (C.c_uint64, [C.c_uint64]),
        "mpo_derive_chunk_seed": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "mpo_rng_normals": (None, [C.c_uint64, C.c_uint64, _D]),
        "mpo_rng_u64": (None, [C.c_uint64, C.c_uint64, _U64]),
        "mpo_summarize": (Result, [C.c_uint64, C.c_double, C.c_double]),
        "mpo_gbm_stepwise_price": (Result, [C.POINTER(GbmCfg), C.c_int]),
        "mpo_european_terminal_price": (Result, [C.POINTER(GbmCfg)]),
        "mpo_gbm_stepwise_tech_price": (Result, [C.POINTER(GbmCfg), C.c_int, C.c_int]),
        "mpo_cv_record": (None, [C.c_void_p, C.c_double, C.c_double]),
        "mpo_summarize_cv": (Result, [C.c_void_p, C.c_double]),
        "mpo_gbm_tech_from_normals": (Result, [C.POINTER(GbmCfg), C.c_int, C.c_int, _D]),
        "mpo_european_terminal_tech_price": (Result, [C.POINTER(GbmCfg), C.c_int]),
        "mpo_terminal_tech_from_normals": (Result, [C.POINTER(GbmCfg), C.c_int, _D]),
        "mpo_gbm_paths_from_normals": (None, [C.POINTER(GbmCfg), C.c_int, _D, _D, _D]),
        "mpo_heston_price": (Result, [C.POINTER(HestonCfg), C.c_int]),
        "mpo_heston_tech_price": (Result, [C.POINTER(HestonCfg), C.c_int, C.c_int]),
        "mpo_structured_normals": (None, [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                          C.c_uint64, _D]),
        "mpo_gbm_structured_price": (Result, [C.POINTER(GbmCfg), C.c_int, C.c_int, C.c_int, _D]),
        "mpo_terminal_structured_price": (Result, [C.POINTER(GbmCfg), C.c_int, C.c_int]),
        "mpo_basket2_structured_price": (Result, [C.c_void_p, C.c_int, C.c_int]),
        "mpo_heston_tech_from_normals": (Result, [C.POINTER(HestonCfg), C.c_int, C.c_int, _D]),
        "mpo_basket2_tech_price": (Result, [C.c_void_p, C.c_int]),
        "mpo_basket_tech_from_normals": (Result, [C.c_uint64, _D, _D, _D, _D, C.c_double, C.c_double,
                                                  C.c_double, C.c_uint64, C.c_uint64, C.c_int, _D]),
        "mpo_heston_paths_from_normals": (None, [C.POINTER(HestonCfg), C.c_int, _D, _D, _D]),
        "mpo_merton_price": (Result, [C.POINTER(MertonCfg)]),
        "mpo_poisson_from_uniform": (C.c_uint64, [C.c_double, C.c_double]),
        "mpo_merton_paths_from_draws": (
            None, [C.POINTER(MertonCfg), C.c_uint64, _D, _D, _D, _D, _U64]),
        "mpo_basket2_price": (Result, [C.POINTER(Basket2Cfg)]),
        "mpo_basket_paths_from_normals": (
            None, [C.c_uint64, _D, _D, _D, _D, C.c_double, C.c_double, C.c_uint64,
                   C.c_uint64, _D, _D]),
        "mpo_cholesky": (C.c_int, [C.c_uint64, _D, _D]),
        "mpo_lsm_solve": (C.c_int, [_D, _D, C.c_uint64, _D]),
        "mpo_laguerre_basis": (None, [C.c_double, C.c_uint64, C.c_int, _D]),
        "mpo_lsm_from_paths": (
            None, [C.POINTER(LsmCfg), _D, _U64, C.c_uint64, C.POINTER(LsmResult), _D, _U64]),
        "mpo_lsm_paths": (None, [C.POINTER(LsmCfg), _D]),
        "mpo_lsm_paths_from_normals": (None, [C.POINTER(LsmCfg), _D, _D]),
        "mpo_lsm_price": (None, [C.POINTER(LsmCfg), _U64, C.c_uint64, C.POINTER(LsmResult)]),
        "mpo_norm_cdf_as": (C.c_double, [C.c_double]),
        "mpo_norm_cdf_exact": (C.c_double, [C.c_double]),
        "mpo_bs_call": (C.c_double, [C.c_double] * 5),
        "mpo_bs_call_ref": (C.c_double, [C.c_double] * 5),
        "mpo_bs_put": (C.c_double, [C.c_double] * 5),
        "mpo_merton_series": (
            C.c_double, [C.POINTER(MertonCfg), C.c_uint64, C.c_double, C.c_int]),
        "mpo_crr_put": (
            C.c_double, [C.c_double] * 5 + [C.c_uint64, C.POINTER(C.c_uint8)]),
        "mpo_inverse_normal": (C.c_double, [C.c_double]),
        "mpo_bb_plan": (C.c_uint64, [C.c_uint64, C.c_double, C.POINTER(BbOp)]),
        "mpo_bb_increments": (None, [C.c_uint64, C.c_double, _D, _D]),
        "mpo_sobol_init": (C.c_int, [_U32, _U32, C.c_uint64, C.c_uint64]),
        "mpo_sobol_directions": (None, [_U32]),
        "mpo_sobol_u32": (C.c_uint32, [C.c_uint32, C.c_uint32]),
        "mpo_sobol_scrambled_u32": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
        "mpo_sobol_sample": (C.c_double, [C.c_uint32, C.c_uint32, C.c_uint32]),
        "mpo_sobol_point_normals": (None, [C.c_uint32, C.c_uint64, C.c_uint64, _D]),
        "mpo_gbm_qmc_price": (Result, [C.POINTER(GbmCfg), C.c_int, C.c_int, _D]),
        "mpo_heston_qmc_paths": (None, [C.POINTER(HestonCfg), C.c_int, C.c_uint64, _D]),
        "mpo_bs_greeks": (BsGreeks, [C.c_double] * 5),
        "mpo_terminal_greeks": (None, [C.POINTER(GbmCfg), C.c_int, _D, _D]),
        "mpo_mlmc_level": (None, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, _D]),
        "mpo_mlmc_level_from_normals": (None, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, _D,
                                               _D]),
        "mpo_philox4x32_10": (None, [_U32, _U32, _U32]),
        "mpo_philox_normals": (
            None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, _D]),
        "mpo_philox_uniforms": (
            None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, _D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    _init_sobol(L)
    return L


def _init_sobol(L) -> None:
    """Feed the Joe-Kuo (new-joe-kuo-6.21201) polynomials and initial
    direction values shipped with scipy to the oracle's own recurrence."""
    import scipy.stats as st

    path = os.path.join(os.path.dirname(st.__file__), "_sobol_direction_numbers.npz")
    with np.load(path) as d:
        poly = np.ascontiguousarray(d["poly"][:256], dtype=np.uint32)
        vinit = np.ascontiguousarray(d["vinit"][:256], dtype=np.uint32)
    rc = L.mpo_sobol_init(_p32(poly), _p32(vinit), vinit.shape[1], 256)
    if rc != 0:
        raise RuntimeError("oracle Sobol initialisation failed")


def _pd(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _p64(a: np.ndarray):
    return a.ctypes.data_as(_U64)


def _p32(a: np.ndarray):
    return a.ctypes.data_as(_U32)


# ------------------------------------------------------------------ configs
@dataclass
class Gbm:
    s0: float = 100.0
    k: float = 100.0
    r: float = 0.03
    sigma: float = 0.2
    t: float = 1.0
    n_paths: int = 100_000
    n_steps: int = 64
    seed: int = 42
    n_threads: int = 1
    barrier: float = 80.0

    def c(self) -> GbmCfg:
        return GbmCfg(self.s0, self.k, self.r, self.sigma, self.t, self.barrier,
                      self.n_paths, self.n_steps, self.seed, self.n_threads)


@dataclass
class Heston:
    s0: float = 100.0
    k: float = 100.0
    r: float = 0.03
    v0: float = 0.04
    kappa: float = 2.0
    theta: float = 0.04
    xi: float = 0.5
    rho: float = -0.7
    t: float = 1.0
    n_paths: int = 10_000
    n_steps: int = 64
    seed: int = 42
    n_threads: int = 1

    def c(self) -> HestonCfg:
        return HestonCfg(self.s0, self.k, self.r, self.v0, self.kappa, self.theta, self.xi,
                         self.rho, self.t, self.n_paths, self.n_steps, self.seed, self.n_threads)


@dataclass
class Merton:
    s0: float = 100.0
    k: float = 100.0
    r: float = 0.03
    sigma: float = 0.2
    lam: float = 0.4
    mu_j: float = -0.08
    sigma_j: float = 0.25
    t: float = 1.0
    n_paths: int = 100_000
    seed: int = 42
    n_threads: int = 1

    def c(self) -> MertonCfg:
        return MertonCfg(self.s0, self.k, self.r, self.sigma, self.lam, self.mu_j, self.sigma_j,
                         self.t, self.n_paths, self.seed, self.n_threads)


@dataclass
class Lsm:
    s0: float = 100.0
    k: float = 100.0
    r: float = 0.03
    sigma: float = 0.2
    t: float = 1.0
    n_paths: int = 20_000
    n_steps: int = 64
    seed: int = 42
    basis_degree: int = 2
    allow_intrinsic_now: bool = True
    basis_weighted: bool = False  # extension: e^{-x/2} L_i(x)

    def c(self) -> LsmCfg:
        return LsmCfg(self.s0, self.k, self.r, self.sigma, self.t, self.n_paths, self.n_steps,
                      self.seed, self.basis_degree, int(self.allow_intrinsic_now),
                      int(self.basis_weighted))


# ------------------------------------------------------------------ wrappers
def rng_normals(seed: int, n: int) -> np.ndarray:
    """MonteCarloRng::standard_normal stream (cpu.rs:2992-3006)."""
    out = np.empty(n, dtype=np.float64)
    lib().mpo_rng_normals(seed, n, _pd(out))
    return out


def rng_u64(seed: int, n: int) -> np.ndarray:
    """MonteCarloRng::next_u64 stream (cpu.rs:2975-2983)."""
    out = np.empty(n, dtype=np.uint64)
    lib().mpo_rng_u64(seed, n, _p64(out))
    return out


def splitmix64(x: int) -> int:
    return int(lib().mpo_splitmix64(x))


def derive_chunk_seed(base: int, idx: int) -> int:
    return int(lib().mpo_derive_chunk_seed(base, idx))


TECHNIQUE_CODES = {"standard": 0, "antithetic": 1, "control_variate": 2}


class CvMoments(C.Structure):
    """ControlVariateMoments (cpu.rs:6202-6230)."""
    _fields_ = [("count", C.c_uint64), ("payoff_sum", C.c_double), ("payoff_sq_sum", C.c_double),
                ("control_sum", C.c_double), ("control_sq_sum", C.c_double),
                ("cross_sum", C.c_double)]


def cv_record(m: CvMoments, payoff: float, control: float) -> None:
    lib().mpo_cv_record(C.byref(m), payoff, control)


def summarize_cv(m: CvMoments, expectation: float) -> tuple[float, float]:
    """summarize_control_variate (cpu.rs:7738-7771)."""
    r = lib().mpo_summarize_cv(C.byref(m), expectation)
    return r.price, r.stderr


def summarize(n: int, total: float, sq_total: float) -> tuple[float, float]:
    """summarize_payoffs (cpu.rs:7641-7652): population variance."""
    r = lib().mpo_summarize(n, total, sq_total)
    return r.price, r.stderr


def gbm_price(cfg: Gbm, payoff: str = "european_call",
              technique: str = "standard") -> tuple[float, float]:
    """Stepwise GBM family pricer (cpu.rs:3228-3262 and siblings), with the
    antithetic and control-variate variants (cpu.rs:7807-7909 and siblings)."""
    c = cfg.c()
    if technique == "standard":
        r = lib().mpo_gbm_stepwise_price(C.byref(c), PAYOFF_CODES[payoff])
    else:
        r = lib().mpo_gbm_stepwise_tech_price(C.byref(c), PAYOFF_CODES[payoff],
                                              TECHNIQUE_CODES[technique])
    return r.price, r.stderr


def gbm_tech_from_normals(cfg: Gbm, normals: np.ndarray, payoff: str,
                          technique: str) -> tuple[float, float]:
    """The technique pricer on a [unit][step] normal buffer (units = pairs for
    antithetic, paths for the control variate), one sequential pass."""
    normals = np.ascontiguousarray(normals, dtype=np.float64)
    units = (cfg.n_paths + 1) // 2 if technique == "antithetic" else cfg.n_paths
    assert normals.size == units * cfg.n_steps
    c = cfg.c()
    r = lib().mpo_gbm_tech_from_normals(C.byref(c), PAYOFF_CODES[payoff],
                                        TECHNIQUE_CODES[technique], _pd(normals))
    return r.price, r.stderr


def european_terminal_price(cfg: Gbm, technique: str = "standard") -> tuple[float, float]:
    """european_call_price_mc_cpu_terminal (cpu.rs:3191-3226), with the
    antithetic (cpu.rs:7785-7805) and control-variate (cpu.rs:7837-7858) variants."""
    c = cfg.c()
    if technique == "standard":
        r = lib().mpo_european_terminal_price(C.byref(c))
    else:
        r = lib().mpo_european_terminal_tech_price(C.byref(c), TECHNIQUE_CODES[technique])
    return r.price, r.stderr


def terminal_tech_from_normals(cfg: Gbm, normals: np.ndarray, technique: str) -> tuple[float, float]:
    """The terminal technique pricer on a caller buffer, one normal per unit."""
    normals = np.ascontiguousarray(normals, dtype=np.float64)
    units = (cfg.n_paths + 1) // 2 if technique == "antithetic" else cfg.n_paths
    assert normals.size == units
    c = cfg.c()
    r = lib().mpo_terminal_tech_from_normals(C.byref(c), TECHNIQUE_CODES[technique], _pd(normals))
    return r.price, r.stderr


def gbm_paths_from_normals(cfg: Gbm, normals: np.ndarray, payoff: str = "european_call"):
    """Per-path (S_T, discounted payoff) from a [path][step] normal buffer."""
    normals = np.ascontiguousarray(normals, dtype=np.float64)
    assert normals.size == cfg.n_paths * cfg.n_steps
    term = np.empty(cfg.n_paths)
    pay = np.empty(cfg.n_paths)
    c = cfg.c()
    lib().mpo_gbm_paths_from_normals(C.byref(c), PAYOFF_CODES[payoff], _pd(normals), _pd(term),
                                     _pd(pay))
    return term, pay


def heston_price(cfg: Heston, scheme: int = HESTON_REFERENCE_FLOOR,
                 technique: str = "standard") -> tuple[float, float]:
    """heston_european_call_price_mc_standard (cpu.rs:8869-8881) and the
    antithetic / control-variate drivers (cpu.rs:8883-8910)."""
    c = cfg.c()
    if technique == "standard":
        r = lib().mpo_heston_price(C.byref(c), scheme)
    else:
        r = lib().mpo_heston_tech_price(C.byref(c), scheme, TECHNIQUE_CODES[technique])
    return r.price, r.stderr


def heston_tech_from_normals(cfg: Heston, normals: np.ndarray, scheme: int,
                             technique: str) -> tuple[float, float]:
    """Heston technique pricer on a [unit][step][2] buffer of (z_var, z_ind)."""
    normals = np.ascontiguousarray(normals, dtype=np.float64)
    units = (cfg.n_paths + 1) // 2 if technique == "antithetic" else cfg.n_paths
    assert normals.size == units * cfg.n_steps * 2
    c = cfg.c()
    r = lib().mpo_heston_tech_from_normals(C.byref(c), scheme, TECHNIQUE_CODES[technique],
                                           _pd(normals))
    return r.price, r.stderr


def heston_paths_from_normals(cfg: Heston, normals: np.ndarray,
                              scheme: int = HESTON_REFERENCE_FLOOR):
    """Per-path S_T from [path][step][2] normals (z_var, z_ind)."""
    normals = np.ascontiguousarray(normals, dtype=np.float64)
    assert normals.size == cfg.n_paths * cfg.n_steps * 2
    term = np.empty(cfg.n_paths)
    var = np.empty(cfg.n_paths)
    c = cfg.c()
    lib().mpo_heston_paths_from_normals(C.byref(c), scheme, _pd(normals), _pd(term), _pd(var))
    return term, var


def merton_price(cfg: Merton) -> tuple[float, float]:
    """merton_jump_diffusion_call_price_mc_cpu (cpu.rs:3151-3189)."""
    c = cfg.c()
    r = lib().mpo_merton_price(C.byref(c))
    return r.price, r.stderr


def merton_paths_from_draws(cfg: Merton, n_steps: int, z, u, zj):
    z, u, zj = (np.ascontiguousarray(a, dtype=np.float64) for a in (z, u, zj))
    term = np.empty(cfg.n_paths)
    jumps = np.empty(cfg.n_paths, dtype=np.uint64)
    c = cfg.c()
    lib().mpo_merton_paths_from_draws(C.byref(c), n_steps, _pd(z), _pd(u), _pd(zj), _pd(term),
                                      _p64(jumps))
    return term, jumps


def poisson_from_uniform(u: float, mean: float) -> int:
    return int(lib().mpo_poisson_from_uniform(u, mean))


def merton_series(cfg: Merton, max_terms: int = 96, tail_tol: float = 1e-12,
                  exact: bool = True) -> float:
    """merton_jump_diffusion_call_reference_price (cpu.rs:3110-3149)."""
    c = cfg.c()
    return float(lib().mpo_merton_series(C.byref(c), max_terms, tail_tol, int(exact)))


def basket2_price(s01=100.0, s02=100.0, k=100.0, r=0.03, sigma1=0.2, sigma2=0.25, rho=0.5,
                  weight1=0.5, weight2=0.5, t=1.0, n_paths=100_000, seed=42,
                  technique="standard"):
    """simulate_basket_pseudorandom_standard (cpu.rs:7503-7518) and its
    antithetic / control-variate variants (cpu.rs:7520-7553)."""
    c = Basket2Cfg(s01, s02, k, r, sigma1, sigma2, rho, weight1, weight2, t, n_paths, seed)
    if technique == "standard":
        res = lib().mpo_basket2_price(C.byref(c))
    else:
        res = lib().mpo_basket2_tech_price(C.byref(c), TECHNIQUE_CODES[technique])
    return res.price, res.stderr


def basket_tech_from_normals(s0, sigma, weight, corr, r, t, k, n_steps, n_units, technique,
                             normals):
    """n-asset stepped basket priced with a technique from [unit][step][asset]
    normals (antithetic units are pairs)."""
    s0, sigma, weight = (np.ascontiguousarray(x, dtype=np.float64) for x in (s0, sigma, weight))
    chol = cholesky(np.asarray(corr, dtype=np.float64))
    normals = np.ascontiguousarray(normals, dtype=np.float64)
    assert normals.size == n_units * n_steps * len(s0)
    res = lib().mpo_basket_tech_from_normals(len(s0), _pd(s0), _pd(sigma), _pd(weight), _pd(chol),
                                             r, t, k, n_steps, n_units, TECHNIQUE_CODES[technique],
                                             _pd(normals))
    return res.price, res.stderr


def cholesky(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    out = np.empty_like(a)
    if lib().mpo_cholesky(n, _pd(a), _pd(out)) != 0:
        raise ValueError("matrix is not positive definite")
    return out


def basket_paths_from_normals(s0, sigma, weight, corr, r, t, n_steps, n_paths, normals):
    s0, sigma, weight = (np.ascontiguousarray(x, dtype=np.float64) for x in (s0, sigma, weight))
    chol = cholesky(np.asarray(corr, dtype=np.float64))
    normals = np.ascontiguousarray(normals, dtype=np.float64)
    n_assets = len(s0)
    assert normals.size == n_paths * n_steps * n_assets
    out = np.empty(n_paths)
    lib().mpo_basket_paths_from_normals(n_assets, _pd(s0), _pd(sigma), _pd(weight), _pd(chol),
                                        r, t, n_steps, n_paths, _pd(normals), _pd(out))
    return out


def laguerre_basis(x: float, degree: int, weighted: bool = False) -> np.ndarray:
    """laguerre_lsm_basis (cpu.rs:7667-7682); weighted: times e^{-x/2} (extension)."""
    out = np.zeros(4)
    lib().mpo_laguerre_basis(float(x), int(degree), int(bool(weighted)), _pd(out))
    return out


def lsm_solve(mat, rhs, dim):
    mat = np.ascontiguousarray(mat, dtype=np.float64).reshape(16)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64).reshape(4)
    out = np.zeros(4)
    rc = lib().mpo_lsm_solve(_pd(mat), _pd(rhs), dim, _pd(out))
    return None if rc != 0 else out


def lsm_paths(cfg: Lsm) -> np.ndarray:
    out = np.empty((cfg.n_paths, cfg.n_steps + 1))
    c = cfg.c()
    lib().mpo_lsm_paths(C.byref(c), _pd(out))
    return out


def lsm_paths_from_normals(cfg: Lsm, normals) -> np.ndarray:
    normals = np.ascontiguousarray(normals, dtype=np.float64)
    out = np.empty((cfg.n_paths, cfg.n_steps + 1))
    c = cfg.c()
    lib().mpo_lsm_paths_from_normals(C.byref(c), _pd(normals), _pd(out))
    return out


def lsm_from_paths(cfg: Lsm, paths: np.ndarray, schedule=None, want_detail=False):
    """Backward induction of price_put_lsm_cpu (cpu.rs:3835-3924)."""
    if schedule is None:
        schedule = np.arange(1, cfg.n_steps + 1, dtype=np.uint64)
    schedule = np.ascontiguousarray(schedule, dtype=np.uint64)
    paths = np.ascontiguousarray(paths, dtype=np.float64)
    res = LsmResult()
    coeff = np.zeros((cfg.n_steps, 4))
    ex = np.zeros(cfg.n_paths, dtype=np.uint64)
    c = cfg.c()
    lib().mpo_lsm_from_paths(C.byref(c), _pd(paths), _p64(schedule), len(schedule), C.byref(res),
                             _pd(coeff), _p64(ex))
    out = {
        "price": res.price, "stderr": res.stderr,
        "early_exercise_count": res.early_exercise_count,
        "maturity_exercise_count": res.maturity_exercise_count,
        "regression_steps": res.regression_steps,
    }
    if want_detail:
        out["coefficients"] = coeff
        out["exercise_steps"] = ex
    return out


def lsm_price(cfg: Lsm, schedule=None):
    """american_put_price_lsm_cpu / bermudan_put_price_lsm_cpu (cpu.rs:3947-4020)."""
    if schedule is None:
        schedule = np.arange(1, cfg.n_steps + 1, dtype=np.uint64)
    schedule = np.ascontiguousarray(schedule, dtype=np.uint64)
    res = LsmResult()
    c = cfg.c()
    lib().mpo_lsm_price(C.byref(c), _p64(schedule), len(schedule), C.byref(res))
    return {
        "price": res.price, "stderr": res.stderr,
        "early_exercise_count": res.early_exercise_count,
        "maturity_exercise_count": res.maturity_exercise_count,
        "regression_steps": res.regression_steps,
    }


def bs_call(s0, k, r, sigma, t, exact=True) -> float:
    """black_scholes_european_call_price (cpu.rs:839-859); exact=False uses
    the reference's A&S CDF."""
    f = lib().mpo_bs_call if exact else lib().mpo_bs_call_ref
    return float(f(s0, k, r, sigma, t))


def bs_put(s0, k, r, sigma, t) -> float:
    return float(lib().mpo_bs_put(s0, k, r, sigma, t))


def norm_cdf_as(x: float) -> float:
    return float(lib().mpo_norm_cdf_as(x))


def crr_put(s0, k, r, sigma, t, steps, exercise_steps) -> float:
    """binomial_put_reference_price (cpu.rs:5244-5305)."""
    mask = np.zeros(steps + 1, dtype=np.uint8)
    for s in exercise_steps:
        if s <= steps:
            mask[s] = 1
    return float(lib().mpo_crr_put(s0, k, r, sigma, t, steps,
                                   mask.ctypes.data_as(C.POINTER(C.c_uint8))))


def inverse_normal(p: float) -> float:
    """inverse_standard_normal (cpu.rs:6012-6061)."""
    return float(lib().mpo_inverse_normal(p))


def bb_plan(n_steps: int, maturity: float):
    """BrownianBridgePlan::new (cpu.rs:5923-5961) as a list of dicts."""
    ops = (BbOp * max(n_steps, 1))()
    n = lib().mpo_bb_plan(n_steps, maturity, ops)
    return [
        {"left": o.left, "mid": o.mid, "right": o.right, "normal_index": o.normal_index,
         "lw": o.lw, "rw": o.rw, "sd": o.sd}
        for o in ops[:n]
    ]


def bb_increments(n_steps: int, maturity: float, inputs) -> np.ndarray:
    inputs = np.ascontiguousarray(inputs, dtype=np.float64)
    out = np.empty(n_steps)
    lib().mpo_bb_increments(n_steps, maturity, _pd(inputs), _pd(out))
    return out


def sobol_directions() -> np.ndarray:
    out = np.empty((256, 32), dtype=np.uint32)
    lib().mpo_sobol_directions(_p32(out))
    return out


def sobol_u32(index: int, dim: int) -> int:
    return int(lib().mpo_sobol_u32(index, dim))


def sobol_scrambled_u32(index: int, dim: int, seed: int) -> int:
    return int(lib().mpo_sobol_scrambled_u32(index, dim, seed))


def sobol_sample(index: int, dim: int, seed: int) -> float:
    return float(lib().mpo_sobol_sample(index, dim, seed))


def sobol_point_normals(seed: int, point: int, dims: int) -> np.ndarray:
    """fill_scrambled_sobol_point_standard_normals (cpu.rs:5393-5432)."""
    out = np.empty(dims)
    lib().mpo_sobol_point_normals(seed & 0xFFFFFFFF, point, dims, _pd(out))
    return out


def gbm_qmc_price(cfg: Gbm, payoff: str = "european_call", bridge: bool = True):
    """simulate_stepwise_qmc (cpu.rs:7113-7197) and siblings."""
    term = np.empty(cfg.n_paths)
    c = cfg.c()
    r = lib().mpo_gbm_qmc_price(C.byref(c), PAYOFF_CODES[payoff], int(bridge), _pd(term))
    return r.price, r.stderr, term


SAMPLING_CODES = {"pseudorandom": 0, "randomized_halton": 1, "latin_hypercube": 2,
                  "scrambled_sobol": 3, "scrambled_sobol_brownian_bridge": 4}


def structured_normals(sampling: str, seed: int, n_paths: int, dims: int, point0: int,
                       n_points: int) -> np.ndarray:
    """StructuredNormalSampler::standard_normal (cpu.rs:5790-5840) for points
    [point0, point0 + n_points), all dims: randomized Halton, Latin hypercube
    (built over n_paths points) or scrambled Sobol."""
    out = np.empty((n_points, dims))
    lib().mpo_structured_normals(SAMPLING_CODES[sampling], seed, n_paths, dims, point0, n_points,
                                 _pd(out))
    return out


def gbm_structured_price(cfg: Gbm, payoff: str, sampling: str, technique: str = "standard"):
    """QMC GBM pricers (cpu.rs:7113-7501) with any structured sampler and
    technique; returns (price, stderr, S_T per unit)."""
    units = (cfg.n_paths + 1) // 2 if technique == "antithetic" else cfg.n_paths
    term = np.empty(units)
    c = cfg.c()
    r = lib().mpo_gbm_structured_price(C.byref(c), PAYOFF_CODES[payoff], SAMPLING_CODES[sampling],
                                       TECHNIQUE_CODES[technique], _pd(term))
    return r.price, r.stderr, term


def terminal_structured_price(cfg: Gbm, sampling: str, technique: str = "standard"):
    """simulate_terminal_qmc (cpu.rs:7064-7111)."""
    c = cfg.c()
    r = lib().mpo_terminal_structured_price(C.byref(c), SAMPLING_CODES[sampling],
                                            TECHNIQUE_CODES[technique])
    return r.price, r.stderr


def basket2_structured_price(sampling: str, technique: str = "standard", s01=100.0, s02=100.0,
                             k=100.0, r=0.03, sigma1=0.2, sigma2=0.25, rho=0.5, weight1=0.5,
                             weight2=0.5, t=1.0, n_paths=100_000, seed=42):
    """simulate_basket_qmc (cpu.rs:7555-7597)."""
    c = Basket2Cfg(s01, s02, k, r, sigma1, sigma2, rho, weight1, weight2, t, n_paths, seed)
    res = lib().mpo_basket2_structured_price(C.byref(c), SAMPLING_CODES[sampling],
                                             TECHNIQUE_CODES[technique])
    return res.price, res.stderr


def heston_qmc_paths(cfg: Heston, scheme: int, path0: int = 0) -> np.ndarray:
    out = np.empty(cfg.n_paths)
    c = cfg.c()
    lib().mpo_heston_qmc_paths(C.byref(c), scheme, path0, _pd(out))
    return out


def philox4x32_10(ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().mpo_philox4x32_10(c, k, o)
    return list(o)


def philox_normals(seed: int, contract: int, stream: int, path0: int, n_paths: int,
                   n_normals: int) -> np.ndarray:
    out = np.empty((n_paths, n_normals))
    lib().mpo_philox_normals(seed, contract, stream, path0, n_paths, n_normals, _pd(out))
    return out


def philox_uniforms(seed: int, contract: int, stream: int, path0: int, n_paths: int,
                    n: int) -> np.ndarray:
    out = np.empty((n_paths, n))
    lib().mpo_philox_uniforms(seed, contract, stream, path0, n_paths, n, _pd(out))
    return out


# ------------------------------------------------------------------ Greeks
GREEK_NAMES = ("delta", "delta2", "vega", "vega2", "rho", "theta")
_DBL_MIN = 2.2250738585072014e-308


def bs_greeks(s0: float, k: float, r: float, sigma: float, t: float) -> dict:
    """black_scholes_european_call_greeks (cpu.rs:866-892), A&S normal CDF."""
    g = lib().mpo_bs_greeks(s0, k, r, sigma, t)
    return {"price": g.price, "delta": g.delta, "vega": g.vega, "rho": g.rho, "theta": g.theta}


def terminal_greeks(cfg: Gbm, estimator: str, normals: np.ndarray | None = None) -> dict:
    """Pathwise (Delta, Vega, Rho, Theta) or likelihood-ratio (Delta, Vega,
    Rho) terminal-GBM Greeks (cpu.rs:4670-4857): {greek: (value, stderr)}.
    Draws: the single reference stream of cfg.seed, or normals[path]."""
    code = {"pathwise": 1, "likelihood_ratio": 2}[estimator]
    out = np.zeros(8)
    buf = None
    if normals is not None:
        buf = np.ascontiguousarray(normals, dtype=np.float64)
        assert buf.size == cfg.n_paths
    c = cfg.c()
    lib().mpo_terminal_greeks(C.byref(c), code, _pd(buf) if buf is not None else None, _pd(out))
    names = ("delta", "vega", "rho", "theta") if code == 1 else ("delta", "vega", "rho")
    return {n: (out[2 * i], out[2 * i + 1]) for i, n in enumerate(names)}


def _bump_spot(spot):
    return max(abs(spot) * 0.01, 0.01)


def _bump_volatility(vol):
    return max(abs(vol) * 0.05, 0.005)


def _bump_maturity(t):
    return min(1.0 / 365.0, t * 0.25)


def _bump_estimate(greek, base_price, center, bump, price_at, reverse=False):
    """push_bump_estimate (cpu.rs:4938-4992): central common-random-number
    difference, down leg floored (Rho at -1, the others at f64::MIN_POSITIVE),
    independent-leg standard error."""
    floor = -1.0 if greek == "rho" else _DBL_MIN
    lo = max(center - bump, floor)
    up = price_at(center + bump)
    down = price_at(lo)
    eff = ((center + bump) - lo) * 0.5
    raw = (up[0] - down[0]) / (2.0 * eff) if eff > 0.0 else 0.0
    se = math.sqrt(up[1] * up[1] + down[1] * down[1]) / (2.0 * eff) if eff > 0.0 else 0.0
    return {"greek": greek, "value": -raw if reverse else raw, "stderr": se, "bump_size": eff,
            "base_price": base_price, "bumped_up_price": up[0], "bumped_down_price": down[0]}


def bump_greeks(workload: str, cfg, scheme: int = HESTON_REFERENCE_FLOOR) -> list:
    """Bump-and-revalue Greeks (european_call_bump_greeks_cpu cpu.rs:4409-4478,
    the Asian / down-and-out / lookback arms :4047-4324, basket :4484-4590,
    Heston :4592-4668) on the oracle pricers.  cfg: Gbm (GBM family), Heston,
    or a dict of basket2_price keywords."""
    from dataclasses import replace

    if workload == "basket_call":
        price = lambda **kw: basket2_price(**{**cfg, **kw})
        base = price()
        b = cfg
        return [
            _bump_estimate("delta", base[0], b["s01"], _bump_spot(b["s01"]), lambda v: price(s01=v)),
            _bump_estimate("delta2", base[0], b["s02"], _bump_spot(b["s02"]), lambda v: price(s02=v)),
            _bump_estimate("vega", base[0], b["sigma1"], _bump_volatility(b["sigma1"]),
                           lambda v: price(sigma1=v)),
            _bump_estimate("vega2", base[0], b["sigma2"], _bump_volatility(b["sigma2"]),
                           lambda v: price(sigma2=v)),
            _bump_estimate("rho", base[0], b["r"], 0.001, lambda v: price(r=v)),
            _bump_estimate("theta", base[0], b["t"], _bump_maturity(b["t"]), lambda v: price(t=v),
                           reverse=True),
        ]
    if workload == "heston_european_call":
        price = lambda c: heston_price(c, scheme)
        base = price(cfg)
        vol = math.sqrt(max(cfg.v0, 0.0))
        return [
            _bump_estimate("delta", base[0], cfg.s0, _bump_spot(cfg.s0),
                           lambda v: price(replace(cfg, s0=v))),
            _bump_estimate("vega", base[0], vol, _bump_volatility(vol),
                           lambda v: price(replace(cfg, v0=max(v, 0.0) * max(v, 0.0)))),
            _bump_estimate("rho", base[0], cfg.r, 0.001, lambda v: price(replace(cfg, r=v))),
            _bump_estimate("theta", base[0], cfg.t, _bump_maturity(cfg.t),
                           lambda v: price(replace(cfg, t=v)), reverse=True),
        ]
    if workload == "european_call":
        price = european_terminal_price
    else:
        price = lambda c: gbm_price(c, workload)
    base = price(cfg)
    return [
        _bump_estimate("delta", base[0], cfg.s0, _bump_spot(cfg.s0),
                       lambda v: price(replace(cfg, s0=v))),
        _bump_estimate("vega", base[0], cfg.sigma, _bump_volatility(cfg.sigma),
                       lambda v: price(replace(cfg, sigma=v))),
        _bump_estimate("rho", base[0], cfg.r, 0.001, lambda v: price(replace(cfg, r=v))),
        _bump_estimate("theta", base[0], cfg.t, _bump_maturity(cfg.t),
                       lambda v: price(replace(cfg, t=v)), reverse=True),
    ]


# ---------------------------------------------------------------- MLMC Asian
MLMC_DEFAULTS = {"s0": 100.0, "k": 100.0, "r": 0.03, "sigma": 0.2, "t": 1.0, "base_steps": 16,
                 "levels": 4, "refinement_factor": 2,
                 "paths_per_level": [50_000, 25_000, 12_500, 6_250], "seed": 42,
                 "sampling": "pseudorandom", "scramble_replicates": 1}


def mlmc_config(**kw) -> dict:
    """ArithmeticAsianMlmcConfig with the reference defaults (cpu.rs:1536-1568)."""
    c = dict(MLMC_DEFAULTS)
    c.update(kw)
    return c


def mlmc_fine_steps(cfg: dict, level: int) -> int:
    return cfg["base_steps"] * cfg["refinement_factor"] ** level


def _mlmc_c(cfg: dict, seed: int) -> MlmcCfg:
    return MlmcCfg(cfg["s0"], cfg["k"], cfg["r"], cfg["sigma"], cfg["t"], cfg["refinement_factor"],
                   seed, SAMPLING_CODES[cfg["sampling"]])


def mlmc_level_moments(cfg: dict, level: int, paths: int, normals=None) -> tuple[float, float]:
    """simulate_asian_mlmc_level (cpu.rs:6639-6773): (sum, sum of squares)."""
    out = np.zeros(2)
    c = _mlmc_c(cfg, cfg["seed"])
    fine = mlmc_fine_steps(cfg, level)
    if normals is None:
        lib().mpo_mlmc_level(C.byref(c), level, paths, fine, _pd(out))
    else:
        z = np.ascontiguousarray(normals, dtype=np.float64)
        assert z.size == paths * fine
        lib().mpo_mlmc_level_from_normals(C.byref(c), level, paths, fine, _pd(z), _pd(out))
    return out[0], out[1]


def _mlmc_single(cfg: dict) -> dict:
    """arithmetic_asian_call_price_mlmc_single_cpu (cpu.rs:3612-3661)."""
    price = 0.0
    var_contrib = 0.0
    total_paths = 0
    total_updates = 0
    levels = []
    for level in range(cfg["levels"]):
        paths = cfg["paths_per_level"][level]
        fine = mlmc_fine_steps(cfg, level)
        coarse = None if level == 0 else fine // cfg["refinement_factor"]
        s, sq = mlmc_level_moments(cfg, level, paths)
        n = float(paths)
        mean = s / n
        variance = max((sq / n) - mean * mean, 0.0)
        stderr = math.sqrt(variance / n)
        cost = paths * (fine + (coarse or 0))
        price += mean
        var_contrib += variance / n
        total_paths += paths
        total_updates += cost
        levels.append({"level": level, "paths": paths, "fine_steps": fine, "coarse_steps": coarse,
                       "mean": mean, "variance": variance, "stderr": stderr,
                       "cost_step_updates": cost})
    return {"price": price, "stderr": math.sqrt(var_contrib), "total_paths": total_paths,
            "total_step_updates": total_updates, "scramble_replicates": 1,
            "replicate_estimates": [price], "levels": levels}


def mlmc_price(cfg: dict) -> dict:
    """arithmetic_asian_call_price_mlmc_cpu (cpu.rs:3283-3293), with the
    replicated structured estimator (cpu.rs:3663-3751)."""
    if cfg["sampling"] == "pseudorandom" or cfg["scramble_replicates"] <= 1:
        return _mlmc_single(cfg)
    reps = []
    for r in range(cfg["scramble_replicates"]):
        rc = dict(cfg)
        rc["seed"] = lib().mpo_derive_chunk_seed(cfg["seed"], r)
        rc["scramble_replicates"] = 1
        reps.append(_mlmc_single(rc))
    R = len(reps)
    est = [x["price"] for x in reps]
    price = sum(est) / R
    stderr = math.sqrt(sum((e - price) * (e - price) for e in est) / (R - 1) / R)
    levels = []
    for li in range(cfg["levels"]):
        first = reps[0]["levels"][li]
        mean = sum(x["levels"][li]["mean"] for x in reps) / R
        variance = sum((x["levels"][li]["mean"] - mean) ** 2 for x in reps) / (R - 1)
        levels.append({**first, "paths": first["paths"] * R, "mean": mean, "variance": variance,
                       "stderr": math.sqrt(variance / R),
                       "cost_step_updates": first["cost_step_updates"] * R})
    return {"price": price, "stderr": stderr, "total_paths": sum(x["total_paths"] for x in reps),
            "total_step_updates": sum(x["total_step_updates"] for x in reps),
            "scramble_replicates": R, "replicate_estimates": est, "levels": levels}


def _pilot_levels(cfg: dict, pilot_paths: int) -> list:
    """measure_arithmetic_asian_mlmc_pilot_levels (cpu.rs:3567-3598)."""
    out = []
    for level in range(cfg["levels"]):
        fine = mlmc_fine_steps(cfg, level)
        coarse = None if level == 0 else fine // cfg["refinement_factor"]
        s, sq = mlmc_level_moments(cfg, level, pilot_paths)
        n = float(pilot_paths)
        mean = s / n
        out.append({"level": level, "fine_steps": fine, "coarse_steps": coarse,
                    "pilot_paths": pilot_paths,
                    "pilot_variance": max((sq / n) - mean * mean, 1.0e-18),
                    "cost_per_path": fine + (coarse or 0), "recommended_paths": 1})
    return out


def _allocate(levels: list, numer: float, by_variance: bool) -> dict:
    denom = sum(math.sqrt(l["pilot_variance"] * l["cost_per_path"]) for l in levels)
    updates = 0
    for l in levels:
        w = math.sqrt(l["pilot_variance"] / l["cost_per_path"])
        raw = math.ceil(denom * w / numer) if by_variance else math.ceil(numer * w / denom)
        l["recommended_paths"] = max(int(raw), 2)
        updates += l["recommended_paths"] * l["cost_per_path"]
    return {"target_step_updates": updates if by_variance else int(numer),
            "estimated_step_updates": updates,
            "paths_per_level": [l["recommended_paths"] for l in levels], "levels": levels}


def mlmc_tune_allocation(cfg: dict, target_step_updates: int, pilot_paths: int) -> dict:
    """tune_arithmetic_asian_mlmc_allocation_cpu (cpu.rs:3295-3350)."""
    return _allocate(_pilot_levels(cfg, pilot_paths), float(target_step_updates), False)


def _estimate_stderr(plan: dict, reps: int) -> float:
    var = sum(l["pilot_variance"] / l["recommended_paths"] for l in plan["levels"])
    return math.sqrt(var / reps)


def mlmc_solve_tolerance(cfg: dict, target_stderr: float, pilot_paths: int = 2_048,
                         min_step_updates: int = 100_000,
                         max_step_updates: int = 10_000_000) -> dict:
    """solve_arithmetic_asian_mlmc_tolerance_cpu (cpu.rs:3352-3423)."""
    reps = 1 if cfg["sampling"] == "pseudorandom" else max(cfg["scramble_replicates"], 1)
    pilot = dict(cfg)
    pilot["scramble_replicates"] = 1
    plan = _allocate(_pilot_levels(pilot, pilot_paths), target_stderr * target_stderr * reps, True)
    est = _estimate_stderr(plan, reps)
    updates = plan["estimated_step_updates"] * reps
    if updates < min_step_updates:
        per = max(-(-min_step_updates // reps), 1)
        plan = mlmc_tune_allocation(pilot, per, pilot_paths)
        est = _estimate_stderr(plan, reps)
        updates = plan["estimated_step_updates"] * reps
    hit = False
    if updates > max_step_updates:
        hit = True
        per = max(max_step_updates // reps, 1)
        plan = mlmc_tune_allocation(pilot, per, pilot_paths)
        est = _estimate_stderr(plan, reps)
        updates = plan["estimated_step_updates"] * reps
    rec = dict(cfg)
    rec["paths_per_level"] = list(plan["paths_per_level"])
    if rec["sampling"] == "pseudorandom":
        rec["scramble_replicates"] = 1
    return {"target_stderr": target_stderr, "estimated_stderr": est,
            "target_met": est <= target_stderr and updates <= max_step_updates,
            "max_step_updates_hit": hit, "scramble_replicates": reps,
            "estimated_step_updates": updates, "paths_per_level": list(plan["paths_per_level"]),
            "allocation": plan, "recommended_config": rec}
