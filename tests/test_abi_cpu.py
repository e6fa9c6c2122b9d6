"""CPU checks of the C-ABI surface and host-side validation (no GPU compute).

* the in-tree library loads and exports every entry point the header
  declares, and the ctypes mirror declares exactly those;
* struct layouts agree between the header and the ctypes mirror;
* invalid requests are rejected with status codes (never an abort),
  mirroring the reference's config validators (runtime/cpu.rs:5010-5153);
* opening a missing device returns MP_E_UNKNOWN_DEVICE.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

from montepath_amd import _abi as A
from montepath_amd.engine import make_request

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "montepath_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mp_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = A.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(A.SIGNATURES) == names


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "montepath_hip.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(mp_request),'
        ' sizeof(mp_result), sizeof(mp_device_info), sizeof(mp_serve_config),'
        ' sizeof(mp_serve_stats), offsetof(mp_request, exercise_steps),'
        ' offsetof(mp_result, regression_steps), offsetof(mp_request, correlation));}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = [C.sizeof(A.Request), C.sizeof(A.Result), C.sizeof(A.DeviceInfo),
            C.sizeof(A.ServeConfig), C.sizeof(A.ServeStats), A.Request.exercise_steps.offset,
            A.Result.regression_steps.offset, A.Request.correlation.offset]
    assert out == want


def footprint(req):
    lib = A.load()
    b = C.c_uint64()
    rc = lib.mp_hip_request_footprint(C.byref(req), C.byref(b))
    return rc, b.value, lib.mp_hip_last_error().decode()


def test_valid_requests_report_footprints():
    rc, b, _ = footprint(make_request("european_call", n_paths=1 << 20, n_steps=64))
    assert rc == A.MP_OK and b > 0
    rc, b, _ = footprint(make_request("american_put", n_paths=1 << 20, n_steps=50))
    assert rc == A.MP_OK
    # FP32 path matrix [step][path] for steps 1..49 dominates: ~196 MiB
    assert 49 * 4 * (1 << 20) <= b <= 49 * 4 * (1 << 20) + 16 * (1 << 20)
    rc, b64, _ = footprint(make_request("american_put", n_paths=1 << 20, n_steps=50, precision=64))
    assert b64 > b


@pytest.mark.parametrize("workload,kw,code,msg", [
    ("european_call", {"n_paths": 0}, A.MP_E_INVALID_CONFIG, "n_paths must be > 0"),
    ("european_call", {"sigma": 0.0}, A.MP_E_INVALID_CONFIG, "sigma must be > 0"),
    ("european_call", {"k": 0.0}, A.MP_E_INVALID_CONFIG, "k must be > 0"),
    ("european_call", {"n_steps": 0}, A.MP_E_INVALID_CONFIG, "n_steps must be > 0"),
    ("european_call", {"technique": 3}, A.MP_E_INVALID_CONFIG, "technique must be"),
    ("merton_jump_diffusion_call", {"technique": "antithetic"}, A.MP_E_UNSUPPORTED,
     "standard technique only"),
    ("american_put", {"technique": "control_variate"}, A.MP_E_UNSUPPORTED, "standard technique only"),
    ("heston_european_call", {"technique": "control_variate", "strikes": [80.0, 100.0, 120.0]},
     A.MP_E_INVALID_CONFIG,
     "one strike per request"),
    ("european_call", {"sampling": 7}, A.MP_E_INVALID_CONFIG, "unknown sampling method"),
    ("merton_jump_diffusion_call", {"sampling": "latin_hypercube"}, A.MP_E_UNSUPPORTED, "pseudorandom"),
    ("american_put", {"sampling": "randomized_halton"}, A.MP_E_UNSUPPORTED, "pseudorandom"),
    ("heston_european_call", {"sampling": "randomized_halton", "n_steps": 40000}, A.MP_E_UNSUPPORTED,
     "at most 65536 dimensions"),
    ("european_call", {"precision": 16}, A.MP_E_INVALID_CONFIG, "precision"),
    ("european_call", {"n_paths": 1 << 33}, A.MP_E_UNSUPPORTED, "2^32"),
    ("down_and_out_call", {"barrier": 0.0}, A.MP_E_INVALID_CONFIG, "barrier must be > 0"),
    ("heston_european_call", {"rho": 1.5}, A.MP_E_INVALID_CONFIG, "rho must be in [-1, 1]"),
    ("heston_european_call", {"v0": -0.1}, A.MP_E_INVALID_CONFIG, "v0 must be >= 0"),
    ("merton_jump_diffusion_call", {"jump_intensity": -1.0}, A.MP_E_INVALID_CONFIG, "jump_intensity"),
    ("merton_jump_diffusion_call", {"sampling": "scrambled_sobol"}, A.MP_E_UNSUPPORTED, "pseudorandom"),
    ("american_put", {"n_steps": 1}, A.MP_E_INVALID_CONFIG, "n_steps must be >= 2"),
    ("american_put", {"basis_degree": 0}, A.MP_E_INVALID_CONFIG, "basis_degree must be > 0"),
    ("european_call", {"basis_weighted": True}, A.MP_E_INVALID_CONFIG, "American and Bermudan puts only"),
    ("bermudan_put", {"n_steps": 64, "exercise_steps": [16, 65]}, A.MP_E_INVALID_CONFIG,
     "exercise_steps must be in 1..=n_steps"),
    ("bermudan_put", {"n_steps": 64}, A.MP_E_INVALID_CONFIG, "at least one exercise date"),
])
def test_invalid_requests_return_status_codes(workload, kw, code, msg):
    rc, _, err = footprint(make_request(workload, **kw))
    assert rc == code
    assert msg in err


def test_techniques_are_accepted_where_the_reference_offers_them():
    basket = dict(asset_s0=[100.0, 95.0], asset_sigma=[0.2, 0.25], asset_weight=[0.5, 0.5],
                  correlation=[[1.0, 0.35], [0.35, 1.0]])
    for workload in ("european_call", "arithmetic_asian_call", "down_and_out_call", "lookback_call",
                     "heston_european_call", "basket_call"):
        for tech in ("antithetic", "control_variate"):
            kw = basket if workload == "basket_call" else {}
            rc, b, err = footprint(make_request(workload, technique=tech, **kw))
            assert rc == A.MP_OK, (workload, tech, err)


def test_structured_samplers_are_accepted_where_offered():
    # the GBM family, terminal European and the basket as in the reference;
    # Heston as an extension
    basket = dict(asset_s0=[100.0, 95.0], asset_sigma=[0.2, 0.25], asset_weight=[0.5, 0.5],
                  correlation=[[1.0, 0.35], [0.35, 1.0]])
    for workload in ("european_call", "arithmetic_asian_call", "down_and_out_call", "lookback_call",
                     "heston_european_call", "basket_call"):
        for sampling in ("randomized_halton", "latin_hypercube"):
            for tech in ("standard", "antithetic", "control_variate"):
                kw = basket if workload == "basket_call" else {}
                rc, _, err = footprint(make_request(workload, sampling=sampling, technique=tech, **kw))
                assert rc == A.MP_OK, (workload, sampling, tech, err)
    rc, _, err = footprint(make_request("european_call", method="terminal", sampling="latin_hypercube"))
    assert rc == A.MP_OK, err


def test_basket_validation():
    req = make_request("basket_call", asset_s0=[100.0] * 2, asset_sigma=[0.2, 0.2])
    req.n_assets = 9
    rc, _, err = footprint(req)
    assert rc == A.MP_E_INVALID_CONFIG and "n_assets" in err


def test_open_missing_device_is_an_error_not_an_abort():
    lib = A.load()
    ctx = C.c_void_p()
    rc = lib.mp_hip_open(99, C.byref(ctx))
    assert rc == A.MP_E_UNKNOWN_DEVICE
    assert not ctx.value
    assert b"not available" in lib.mp_hip_last_error()


def test_null_arguments_are_rejected():
    lib = A.load()
    assert lib.mp_hip_price(None, None, None) == A.MP_E_INVALID_CONFIG
    assert lib.mp_hip_request_footprint(None, None) == A.MP_E_INVALID_CONFIG


def test_version_string():
    assert b"montepath-hip" in A.load().mp_hip_version()
