"""CPU tests of the rocm_hip production layer: aliases, selection and
validation diagnostics, and the strict no-fallback contract when no device
is visible (this container has none).  Mirrors the reference's selection
tests for the Metal bridge (python/tests/test_production.py)."""
import pytest

from montepath_amd import production as P


def test_aliases_and_maps():
    for alias in ("rocm", "rocm_hip", "hip"):
        assert P._normalize_backend(alias) == "rocm_hip"
    # the reference's aliases are kept
    for alias, target in (("cpu", "cpu_native"), ("metal", "apple_metal"), ("cuda", "nvidia_cuda")):
        assert P._normalize_backend(alias) == target
    from montepath_amd import native
    for fn in P.ROCM_FUNCTION_BY_WORKLOAD.values():
        assert callable(getattr(native, fn))
        assert fn in native.KNOWN_ROCM_FUNCTIONS


def test_unsupported_workload_and_unknown_backend():
    sel = P.select_backend("inventory_policy", backend="rocm")
    assert not sel.ok and sel.diagnostics[0]["code"] == "MC_BACKEND_ROCM_WORKLOAD_UNSUPPORTED"
    sel = P.select_backend("european_call", backend="tpu")
    assert not sel.ok and sel.diagnostics[0]["code"] == "MC_BACKEND_UNKNOWN"
    sel = P.select_backend("european_call", backend="cpu_native")
    assert not sel.ok and sel.diagnostics[0]["code"] == "MC_BACKEND_OTHER_MODULE"
    assert sel.as_dict()["diagnostics"][0]["suggestion"]


@pytest.mark.skipif(P.runtime_status()[0], reason="a HIP device is visible")
def test_no_device_is_an_error_not_a_fallback():
    sel = P.select_backend("european_call", backend="rocm_hip")
    assert not sel.ok and sel.diagnostics[0]["code"] == "MC_BACKEND_ROCM_UNAVAILABLE"
    cap = P.rocm_capability()
    assert cap.status == "unavailable" and not cap.agent_tool_ready
    with pytest.raises(P.ProductionCapabilityError) as err:
        P.execute_workload("european_call", {"n_paths": 1000}, backend="rocm")
    assert err.value.selection.diagnostics[0]["code"] == "MC_BACKEND_ROCM_UNAVAILABLE"


def test_validation_payload_shape():
    v = P.validate_workload_request("american_put", {"sampling": "sobol"}, backend="rocm")
    assert v["schema_version"] == "montepath-validation.v1"
    assert not v["ok"] and v["workload"] == "american_put"
    assert set(v["selection"]) == {"ok", "backend_id", "execution_mode", "workload",
                                   "agent_tool_ready", "diagnostics", "warnings", "native_module"}
