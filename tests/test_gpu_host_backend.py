"""The C++ host layer (montepath::RocmHipBackend) on the GPU: discovery,
support, cost, compile and execution of the reference's three execution
inputs (backend/mod.rs:179-184) and of the additive ones for the families
the reference prices on the CPU only, through the C-ABI, in FP32 and FP64.  Each
execution must equal the same request priced through the Python engine bit
for bit, repeat identically, and the European price must sit within 3
standard errors of Black-Scholes.
"""
import json
import os
import subprocess

import pytest

from montepath_amd import build
from montepath_amd.engine import make_request

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rows():
    exe = build.HOST_SELFTEST
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built: run montepath_amd/build.py")
    print("[host] backend_selftest gpu", flush=True)
    out = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    return [json.loads(line) for line in out.stdout.splitlines() if line.strip()]


def test_discovery_and_plan(rows):
    devs = [r for r in rows if r["check"] == "device"]
    assert devs and devs[0]["device_id"] == "rocm:0"
    assert devs[0]["backend_id"] == "rocm_hip" and devs[0]["vendor"] == "amd"
    assert devs[0]["supports_float64"] and devs[0]["memory_total_mb"] > 100_000
    p = [r for r in rows if r["check"] == "plan"][0]
    assert p["support"] == "supported"
    assert p["artifact_id"] == "rocm-native:100000:64:64" and p["artifact_device"] == "rocm:0"
    assert p["peak_mb"] >= 1 and p["runtime_ms"] > 0


# the same requests through the engine: the reference configs' defaults
# (runtime/cpu.rs:1296-1818) at 100,000 paths and 64 steps
ENGINE_KW = {
    "european_call": dict(n_steps=64),
    "arithmetic_asian_call": dict(n_steps=64),
    "down_and_out_call": dict(n_steps=64, barrier=80.0),
    "lookback_call": dict(n_steps=64),
    "heston_european_call": dict(n_steps=64, v0=0.04, kappa=1.5, theta=0.04, xi=0.3, rho=-0.6,
                                 heston_scheme="reference_floor"),
    "merton_jump_diffusion_call": dict(n_steps=0, jump_intensity=0.4, jump_mean=-0.08, jump_volatility=0.25),
    "basket_call": dict(n_steps=1, asset_s0=[100.0, 95.0], asset_sigma=[0.2, 0.25], asset_weight=[0.5, 0.5],
                        correlation=[[1.0, 0.35], [0.35, 1.0]]),
    "american_put": dict(n_steps=64, basis_degree=2),
    "bermudan_put": dict(n_steps=64, basis_degree=2, exercise_steps=[16, 32, 48, 64]),
}


def test_executions_equal_the_engine(rows, engine):
    ex = [r for r in rows if r["check"] == "execute"]
    assert len(ex) == 2 * len(ENGINE_KW)
    for r in ex:
        res = engine.price(make_request(r["workload"], n_paths=100_000, seed=42, precision=r["precision"],
                                        **ENGINE_KW[r["workload"]]))
        assert (r["price"], r["std_error"]) == (res.price, res.stderr), r
        assert r["repeat_identical"]
    eu = [r for r in ex if r["workload"] == "european_call"]
    bs = 9.413403383853016  # S0 = K = 100, r = 0.03, sigma = 0.2, T = 1
    for r in eu:
        assert abs(r["price"] - bs) <= 3 * r["std_error"]


def test_execution_errors(rows):
    e = [r for r in rows if r["check"] == "execute_errors"][0]
    assert e["shape_mismatch"] == "IncompatibleExecutionInput"


def test_threads_each_open_their_own_context(rows):
    t = [r for r in rows if r["check"] == "threads"][0]
    assert t["identical"], t
