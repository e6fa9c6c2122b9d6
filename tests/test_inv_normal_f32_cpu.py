"""The FP32 inverse normal's central polynomial (mp_inv_normal_f32_coeffs.h,
generated by montepath_amd/csrc/gen_inv_normal_f32.py) against
scipy.special.ndtri on every uniform of the kernel's 2^-23 grid, evaluated
the way the kernel evaluates it (FP32, one rounding per fused multiply-add).
The GPU tests check the device function itself through the FP32-vs-FP64
per-path bars."""
import importlib.util
import os
import re

import numpy as np
from scipy.special import ndtri

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "montepath_amd", "csrc")


def _gen():
    spec = importlib.util.spec_from_file_location("gen_inv_normal_f32",
                                                  os.path.join(CSRC, "gen_inv_normal_f32.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _header_coeffs():
    text = open(os.path.join(CSRC, "mp_inv_normal_f32_coeffs.h")).read()
    vals = dict(re.findall(r"#define MP_INVN_C(\d+) (\S+)f", text))
    return np.array([float(vals[str(i)]) for i in range(len(vals))], dtype=np.float32)


def test_header_matches_the_fit():
    gen = _gen()
    assert np.array_equal(_header_coeffs(), gen.fit())


def test_fp32_inverse_normal_on_the_whole_uniform_grid():
    gen = _gen()
    k = np.arange(0, 1 << 23, dtype=np.int64)
    z = gen.simulate(_header_coeffs(), k).astype(np.float64)
    p = k * 2.0 ** -23
    ref = ndtri(np.maximum(p, np.finfo(float).tiny))
    big = np.abs(ref) > 1e-3
    rel = np.abs(z[big] - ref[big]) / np.abs(ref[big])
    # measured: 3.3e-7 (central branch, 99.954% of the grid), 2.0e-7 (tail)
    assert float(rel.max()) <= 5e-7, float(rel.max())
    assert float(np.max(np.abs(z[~big] - ref[~big]))) <= 1e-9
    # antisymmetry about p = 1/2 (x = 2p - 1 is exact)
    assert np.array_equal(z[1:(1 << 22)], -z[(1 << 23) - 1:(1 << 22):-1])
