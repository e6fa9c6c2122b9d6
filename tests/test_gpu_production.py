"""rocm_hip production layer on the GPU: selection succeeds, and
execute_workload returns the bridge function's own result (same draws, same
bits) in the reference's execution payload."""
import pytest

from montepath_amd import native
from montepath_amd import production as P

pytestmark = pytest.mark.gpu

CASES = [
    ("european_call", {"n_paths": 200_000, "n_steps": 64, "seed": 42}),
    ("heston_european_call", {"n_paths": 100_000, "n_steps": 64, "seed": 7, "sampling": "scrambled_sobol"}),
    ("american_put", {"n_paths": 50_000, "n_steps": 20, "seed": 3}),
    ("arithmetic_asian_mlmc", {"levels": 3, "base_steps": 4, "seed": 11,
                               "paths_per_level": [20_000, 10_000, 5_000]}),
]


@pytest.mark.parametrize("workload,config", CASES)
def test_execute_workload_equals_bridge(workload, config):
    print(f"[production] {workload}", flush=True)
    sel = P.select_backend(workload, backend="rocm")
    assert sel.ok and sel.backend_id == "rocm_hip" and sel.execution_mode == "native_extension_rocm"
    out = P.execute_workload(workload, config, backend="rocm_hip")
    assert out["ok"] and out["schema_version"] == "montepath-execution.v1"
    assert out["manifest"]["backend"] == "rocm_hip"
    direct = getattr(native, P.ROCM_FUNCTION_BY_WORKLOAD[workload])(dict(config))
    price = direct["price"] if "price" in direct else direct["values"]["price"]
    assert out["result"]["price"] == price
    assert out["result"]["stderr"] == direct["stderr"]
    assert P.rocm_capability().status == "available"


def test_policy_rejects_before_execution():
    with pytest.raises(P.ProductionCapabilityError) as err:
        P.execute_workload("american_put", {"sampling": "sobol"}, backend="rocm")
    assert err.value.selection.diagnostics[0]["code"] == "MC_BACKEND_ROCM_SAMPLING_UNSUPPORTED"
