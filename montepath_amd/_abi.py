"""ctypes mirror of include/montepath_hip.h and the loader of the in-tree
HIP library.  Loading fails loudly when the library is missing: the
``rocm_hip`` backend never falls back to a CPU path."""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "libmontepath_hip.so")

MP_OK = 0
MP_E_UNKNOWN_DEVICE = 1
MP_E_INCOMPATIBLE_INPUT = 2
MP_E_UNSUPPORTED = 3
MP_E_OOM = 4
MP_E_HIP = 5
MP_E_INVALID_CONFIG = 6
MP_E_NOT_FOUND = 7
MP_E_CANCELLED = 8
STATUS_NAMES = {
    0: "MP_OK", 1: "MP_E_UNKNOWN_DEVICE", 2: "MP_E_INCOMPATIBLE_INPUT", 3: "MP_E_UNSUPPORTED",
    4: "MP_E_OOM", 5: "MP_E_HIP", 6: "MP_E_INVALID_CONFIG", 7: "MP_E_NOT_FOUND",
    8: "MP_E_CANCELLED",
}

EUROPEAN_CALL = 1
ARITHMETIC_ASIAN_CALL = 2
DOWN_AND_OUT_CALL = 3
LOOKBACK_CALL = 4
HESTON_EUROPEAN_CALL = 5
MERTON_JUMP_DIFFUSION_CALL = 6
BASKET_CALL = 7
AMERICAN_PUT = 8
BERMUDAN_PUT = 9

PSEUDORANDOM = 0
RANDOMIZED_HALTON = 1
LATIN_HYPERCUBE = 2
SCRAMBLED_SOBOL = 3
SCRAMBLED_SOBOL_BRIDGE = 4

METHOD_STEPWISE = 0
METHOD_TERMINAL = 1

HESTON_REFERENCE_FLOOR = 0
HESTON_FULL_TRUNCATION = 1

MAX_STRIKES = 4
MAX_ASSETS = 8
MAX_GREEKS = 6

GREEK_DELTA, GREEK_DELTA2, GREEK_VEGA, GREEK_VEGA2, GREEK_RHO, GREEK_THETA = range(6)
GREEK_NAMES = ("delta", "delta2", "vega", "vega2", "rho", "theta")
GREEKS_BUMP_AND_REVALUE, GREEKS_PATHWISE, GREEKS_LIKELIHOOD_RATIO = 0, 1, 2


class Request(C.Structure):
    _fields_ = [
        ("workload", C.c_int32), ("precision", C.c_int32), ("sampling", C.c_int32),
        ("technique", C.c_int32), ("method", C.c_int32), ("heston_scheme", C.c_int32),
        ("n_strikes", C.c_int32), ("basis_degree", C.c_int32), ("fp64_check", C.c_int32),
        ("allow_intrinsic_now", C.c_int32), ("n_assets", C.c_int32), ("n_exercise", C.c_int32),
        ("basis_weighted", C.c_int32),
        ("n_paths", C.c_uint64), ("n_steps", C.c_uint64), ("seed", C.c_uint64),
        ("contract_id", C.c_uint64),
        ("s0", C.c_double), ("k", C.c_double), ("r", C.c_double), ("sigma", C.c_double),
        ("t", C.c_double), ("barrier", C.c_double), ("strikes", C.c_double * MAX_STRIKES),
        ("v0", C.c_double), ("kappa", C.c_double), ("theta", C.c_double), ("xi", C.c_double),
        ("rho", C.c_double),
        ("jump_intensity", C.c_double), ("jump_mean", C.c_double),
        ("jump_volatility", C.c_double),
        ("asset_s0", C.c_double * MAX_ASSETS), ("asset_sigma", C.c_double * MAX_ASSETS),
        ("asset_weight", C.c_double * MAX_ASSETS),
        ("correlation", C.c_double * (MAX_ASSETS * MAX_ASSETS)),
        ("exercise_steps", C.POINTER(C.c_uint32)),
    ]


class Result(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("n_strikes", C.c_int32),
        ("price", C.c_double), ("std_error", C.c_double), ("runtime_ms", C.c_double),
        ("fp64_ref_error", C.c_double),
        ("strike_price", C.c_double * MAX_STRIKES),
        ("strike_std_error", C.c_double * MAX_STRIKES),
        ("n_paths", C.c_uint64), ("n_steps", C.c_uint64),
        ("early_exercise_count", C.c_uint64), ("maturity_exercise_count", C.c_uint64),
        ("regression_steps", C.c_uint64),
    ]


class GreekEstimate(C.Structure):
    _fields_ = [
        ("greek", C.c_int32), ("estimator", C.c_int32), ("value", C.c_double),
        ("std_error", C.c_double), ("bump_size", C.c_double), ("base_price", C.c_double),
        ("bumped_up_price", C.c_double), ("bumped_down_price", C.c_double),
    ]


MLMC_MAX_LEVELS = 16
MLMC_MAX_REPLICATES = 64


class MlmcConfig(C.Structure):
    _fields_ = [
        ("s0", C.c_double), ("k", C.c_double), ("r", C.c_double), ("sigma", C.c_double),
        ("t", C.c_double), ("base_steps", C.c_uint32), ("levels", C.c_uint32),
        ("refinement_factor", C.c_uint32), ("scramble_replicates", C.c_uint32),
        ("sampling", C.c_int32), ("precision", C.c_int32), ("seed", C.c_uint64),
        ("contract_id", C.c_uint64), ("paths_per_level", C.c_uint64 * MLMC_MAX_LEVELS),
    ]


class MlmcLevel(C.Structure):
    _fields_ = [
        ("level", C.c_uint64), ("paths", C.c_uint64), ("fine_steps", C.c_uint64),
        ("coarse_steps", C.c_uint64), ("mean", C.c_double), ("variance", C.c_double),
        ("std_error", C.c_double), ("cost_step_updates", C.c_uint64),
    ]


class MlmcResult(C.Structure):
    _fields_ = [
        ("price", C.c_double), ("std_error", C.c_double), ("total_paths", C.c_uint64),
        ("total_step_updates", C.c_uint64), ("scramble_replicates", C.c_uint32),
        ("n_levels", C.c_uint32), ("runtime_ms", C.c_double),
        ("levels", MlmcLevel * MLMC_MAX_LEVELS),
        ("replicate_estimates", C.c_double * MLMC_MAX_REPLICATES),
    ]


class MlmcAllocationLevel(C.Structure):
    _fields_ = [
        ("level", C.c_uint64), ("fine_steps", C.c_uint64), ("coarse_steps", C.c_uint64),
        ("pilot_paths", C.c_uint64), ("pilot_variance", C.c_double),
        ("cost_per_path", C.c_uint64), ("recommended_paths", C.c_uint64),
    ]


class MlmcAllocation(C.Structure):
    _fields_ = [
        ("target_step_updates", C.c_uint64), ("estimated_step_updates", C.c_uint64),
        ("n_levels", C.c_uint32), ("reserved", C.c_uint32),
        ("levels", MlmcAllocationLevel * MLMC_MAX_LEVELS),
    ]


class MlmcTolerance(C.Structure):
    _fields_ = [("target_stderr", C.c_double), ("pilot_paths_per_level", C.c_uint64),
                ("min_step_updates", C.c_uint64), ("max_step_updates", C.c_uint64)]


class MlmcTolerancePlan(C.Structure):
    _fields_ = [
        ("target_stderr", C.c_double), ("estimated_stderr", C.c_double),
        ("target_met", C.c_int32), ("max_step_updates_hit", C.c_int32),
        ("scramble_replicates", C.c_uint32), ("reserved", C.c_uint32),
        ("estimated_step_updates", C.c_uint64), ("allocation", MlmcAllocation),
        ("recommended", MlmcConfig),
    ]


class DeviceInfo(C.Structure):
    _fields_ = [
        ("device_id", C.c_char * 32), ("name", C.c_char * 96), ("arch", C.c_char * 32),
        ("memory_total_mb", C.c_uint64), ("memory_free_mb", C.c_uint64),
        ("compute_units", C.c_int32), ("supports_float64", C.c_int32),
        ("max_threads_hint", C.c_int32), ("wavefront_size", C.c_int32),
    ]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_inflight", C.c_int32),
                ("max_batch", C.c_int32)]


class ServeStats(C.Structure):
    _fields_ = [
        ("completed", C.c_uint64), ("wall_s", C.c_double), ("requests_per_s", C.c_double),
        ("path_steps_per_s", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
        ("max_ms", C.c_double), ("launches", C.c_uint64), ("max_admitted_bytes", C.c_uint64),
    ]


class Tuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("paths_per_lane", C.c_int32),
                ("tile_paths", C.c_int32), ("step_unroll", C.c_int32)]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("main_kernel_ms", C.c_double),
                ("finalize_ms", C.c_double), ("main_launches", C.c_uint64)]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)
_U64 = C.POINTER(C.c_uint64)

# Every entry point declared in include/montepath_hip.h, with its ctypes signature.
SIGNATURES = {
    "mp_hip_version": (C.c_char_p, []),
    "mp_hip_last_error": (C.c_char_p, []),
    "mp_hip_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mp_hip_open": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "mp_hip_close": (None, [_P]),
    "mp_hip_device_info": (C.c_int, [_P, C.POINTER(DeviceInfo)]),
    "mp_hip_price": (C.c_int, [_P, C.POINTER(Request), C.POINTER(Result)]),
    "mp_hip_price_noalloc": (C.c_int, [_P, C.POINTER(Request), C.POINTER(Result)]),
    "mp_hip_price_batch": (C.c_int, [_P, C.POINTER(Request), C.c_size_t, C.POINTER(Result)]),
    "mp_hip_batch_prepare": (C.c_int, [_P, C.POINTER(Request), C.c_size_t, C.c_int,
                                       C.POINTER(_P)]),
    "mp_hip_batch_launch": (C.c_int, [_P]),
    "mp_hip_batch_sync": (C.c_int, [_P]),
    "mp_hip_batch_results": (C.c_int, [_P, C.POINTER(Result)]),
    "mp_hip_batch_info": (C.c_int, [_P, _U64, _U64, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mp_hip_batch_free": (None, [_P]),
    "mp_hip_price_plain_fp64": (C.c_int, [_P, C.POINTER(Request), C.POINTER(Result)]),
    "mp_hip_european_call_stepwise_v1": (C.c_int, [_P, _D, _D, C.c_int32, C.c_int32, C.c_double,
                                                   C.c_double, C.c_double, C.c_double, C.c_double]),
    "mp_hip_paths_from_normals": (C.c_int, [_P, C.POINTER(Request), _D, _D, _D, _D, _D]),
    "mp_hip_price_from_normals": (C.c_int, [_P, C.POINTER(Request), _D, _D, _D, C.POINTER(Result)]),
    "mp_hip_greeks": (C.c_int, [_P, C.POINTER(Request), C.c_size_t, C.c_int32, C.POINTER(Result),
                                C.POINTER(GreekEstimate), C.POINTER(C.c_int32)]),
    "mp_hip_greeks_from_normals": (C.c_int, [_P, C.POINTER(Request), C.c_int32, _D,
                                             C.POINTER(Result), C.POINTER(GreekEstimate),
                                             C.POINTER(C.c_int32)]),
    "mp_hip_mlmc_price": (C.c_int, [_P, C.POINTER(MlmcConfig), C.POINTER(MlmcResult)]),
    "mp_hip_mlmc_tune_allocation": (C.c_int, [_P, C.POINTER(MlmcConfig), C.c_uint64, C.c_uint64,
                                              C.POINTER(MlmcAllocation)]),
    "mp_hip_mlmc_solve_tolerance": (C.c_int, [_P, C.POINTER(MlmcConfig), C.POINTER(MlmcTolerance),
                                              C.POINTER(MlmcTolerancePlan)]),
    "mp_hip_mlmc_level_from_normals": (C.c_int, [_P, C.POINTER(MlmcConfig), C.c_uint32, C.c_uint64,
                                                 _D, _D]),
    "mp_hip_price_with_paths": (C.c_int, [_P, C.POINTER(Request), C.POINTER(Result), _D, _D]),
    "mp_hip_lsm_from_normals": (C.c_int, [_P, C.POINTER(Request), _D, C.POINTER(Result), _D,
                                          _U32]),
    "mp_hip_draws": (C.c_int, [_P, C.POINTER(Request), C.c_int, C.c_uint32, C.c_uint64,
                               C.c_uint64, C.c_uint64, _D]),
    "mp_hip_request_footprint": (C.c_int, [C.POINTER(Request), _U64]),
    "mp_hip_server_open": (C.c_int, [_P, C.POINTER(ServeConfig), C.POINTER(_P)]),
    "mp_hip_server_close": (None, [_P]),
    "mp_hip_submit": (C.c_int, [_P, C.POINTER(Request), _U64]),
    "mp_hip_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(Result), C.POINTER(C.c_int)]),
    "mp_hip_wait": (C.c_int, [_P, C.c_uint64, C.POINTER(Result)]),
    "mp_hip_cancel": (C.c_int, [_P, C.c_uint64]),
    "mp_hip_server_stats": (C.c_int, [_P, C.POINTER(ServeStats)]),
    "mp_hip_serve_trace": (C.c_int, [_P, C.POINTER(ServeConfig), C.POINTER(Request), _D,
                                     C.c_size_t, C.POINTER(Result), _D, C.POINTER(ServeStats)]),
    "mp_hip_autotune": (C.c_int, [_P, C.POINTER(Request), C.POINTER(Tuning), _D]),
    "mp_hip_get_tuning": (C.c_int, [_P, C.POINTER(Request), C.POINTER(Tuning)]),
    "mp_hip_set_tuning": (C.c_int, [_P, C.POINTER(Request), C.POINTER(Tuning)]),
    "mp_hip_last_timing": (C.c_int, [_P, C.POINTER(Timing)]),
}

_lib = None


class MontepathHipError(RuntimeError):
    """Raised for a non-zero mp_status; ``code`` carries the status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


def load() -> C.CDLL:
    """Load libmontepath_hip.so from the package tree (built by build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"rocm_hip native library not found at {LIB_PATH}; build it with "
            "`python -m montepath_amd.build` (there is no CPU fallback)"
        )
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != MP_OK:
        msg = load().mp_hip_last_error()
        raise MontepathHipError(rc, msg.decode() if msg else "")
