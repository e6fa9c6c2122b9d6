"""Writes mp_inv_normal_f32_coeffs.h: the central-branch polynomial of the
FP32 inverse normal (mp_inv_normal_f32 in mp_device.h).

Phi^-1(p) = sqrt(2) erfinv(x) with x = 2p - 1, written as x * R(l) with
l = log2(1 - x^2), the value the kernel's v_log_f32 returns.  For
w = -ln(1 - x^2) < 7 (all but 0.046% of the uniforms) R is a degree-9
polynomial in l itself (no shift or scaling of l: one instruction fewer per
draw), fitted here by iteratively reweighted Chebyshev least squares on the
relative error of sqrt(2) erfinv(x) / x (scipy.special.erfinv as the
reference function): maximum relative error 2.0e-7 in exact arithmetic.  The
tail w >= 7 keeps the w >= 5 branch of Giles' single-precision erfinv
("Approximating the erfinv function", GPU Computing Gems, 2011).
`simulate()` evaluates both branches the way the kernel does (FP32, one
rounding per fused multiply-add) on the whole 2^-23 grid of uniforms;
tests/test_inv_normal_f32_cpu.py checks it against scipy.special.ndtri.
"""
import os

import numpy as np
from numpy.polynomial import chebyshev as Ch
from scipy.special import erfinv

W_CENTRAL = 7.0   # central branch for w < W_CENTRAL
DEGREE = 9
# Giles' tail branch (w >= 5 there), coefficients times sqrt(2), highest first,
# in s = sqrt(w) - 3
TAIL = [-0.000283145724, 0.000142765653, 0.00190825947, -0.00519501232, 0.00811688974,
        -0.0107797878, 0.0133485785, 1.41658103, 4.00643444]
LN2 = 0.69314718


def fit() -> np.ndarray:
    """Coefficients of R(l), ascending powers of l, as FP32 values."""
    w = np.linspace(0.0, W_CENTRAL, 700001)[:-1]
    x = np.sqrt(-np.expm1(-w))
    x[0] = 1.0
    f = np.sqrt(2.0) * erfinv(x) / x
    f[0] = np.sqrt(np.pi / 2.0)  # limit as x -> 0
    l = -w / np.log(2.0)
    half = W_CENTRAL / np.log(2.0) / 2.0
    u = (l + half) / half  # [-1, 1) for the fit only
    wts = 1.0 / f
    best = None
    for _ in range(60):
        c = Ch.chebfit(u, f, DEGREE, w=wts)
        e = np.abs(Ch.chebval(u, c) - f) / f
        if best is None or e.max() < best[0]:
            best = (e.max(), c.copy())
        wts = wts * (1.0 + 3.0 * e / e.max())
    p = np.polynomial.Polynomial(Ch.cheb2poly(best[1]))
    q = p(np.polynomial.Polynomial([1.0, 1.0 / half]))
    return q.coef.astype(np.float32)


def l_central() -> np.float32:
    """The central branch's bound on l = log2(1 - x^2): w < W_CENTRAL."""
    return np.float32(-W_CENTRAL / np.log(2.0))


def simulate(coef: np.ndarray, k: np.ndarray) -> np.ndarray:
    """The kernel's FP32 normal for uniforms p = k * 2^-23 (k < 2^23)."""
    f32 = np.float32
    p = k.astype(np.float64) * 2.0 ** -23
    x = (2.0 * (1.0 + p) - 3.0).astype(f32)                                  # exact
    om = (1.0 - x.astype(np.float64) ** 2).astype(f32)                       # one rounding
    with np.errstate(divide="ignore"):
        l = np.log2(om.astype(np.float64)).astype(f32)
    central = l > l_central()
    r = np.full(l.shape, coef[-1], dtype=f32)
    for c in coef[-2::-1]:
        r = (np.float64(c) + r.astype(np.float64) * l.astype(np.float64)).astype(f32)
    with np.errstate(invalid="ignore"):
        s = (np.sqrt((-LN2 * l[~central].astype(np.float64)).astype(f32).astype(np.float64)).astype(f32) - f32(3.0))
    rt = np.full(s.shape, f32(TAIL[0]))
    for c in TAIL[1:]:
        rt = (np.float64(f32(c)) + rt.astype(np.float64) * s.astype(np.float64)).astype(f32)
    r[~central] = rt
    r[k == 0] = f32(37.51937936814848)
    return (r.astype(np.float64) * x.astype(np.float64)).astype(f32)


def main() -> None:
    coef = fit()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mp_inv_normal_f32_coeffs.h")
    with open(out, "w") as fh:
        fh.write("// Generated by gen_inv_normal_f32.py: central branch of mp_inv_normal_f32,\n")
        fh.write(f"// R(l) = sum c_i l^i, l = log2(1 - x^2), for -ln(1 - x^2) < {W_CENTRAL}.\n")
        fh.write("#pragma once\n")
        fh.write(f"#define MP_INVN_L_CENTRAL {float(l_central())!r}f  // log2(1 - x^2) above this: w < {W_CENTRAL}\n")
        for i, c in enumerate(coef):
            fh.write(f"#define MP_INVN_C{i} {float(c)!r}f\n")
    print(out)


if __name__ == "__main__":
    main()
