"""Conditioning of the full-truncation Heston Euler scheme (CPU, numpy).

The per-path FP32 bar (terminal values within 1e-5 relative of FP64 on the
same draws) rests on the scheme being well conditioned in its draws.  It is
for the parameters the engine is measured on (xi = 0.5, 2 kappa theta /
xi^2 = 0.64).  Where the Feller condition fails badly, paths spend steps
near v = 0, where the variance update's derivative
xi sqrt(dt) Z / (2 sqrt(v)) is unbounded, and FP64 arithmetic fed the same
Brownian increments rounded to FP32 (a relative change of at most 6e-8)
ends far from FP64 on unrounded increments: the divergence the GPU's FP32
runs show there comes from the scheme, not from FP32 arithmetic.
"""
import numpy as np


def _terminals(xi, round_inputs, n=4096, m=256, seed=5):
    kappa, theta, v0, rho, r, t = 2.0, 0.04, 0.04, -0.7, 0.03, 1.0
    dt = t / m
    sdt = np.sqrt(dt)
    rp = np.sqrt(1.0 - rho * rho)
    rng = np.random.default_rng(seed)
    x = np.zeros(n)
    v = np.full(n, v0)
    for _ in range(m):
        dwv = rng.standard_normal(n) * sdt
        dwi = rng.standard_normal(n) * sdt
        if round_inputs:
            dwv = dwv.astype(np.float32).astype(np.float64)
            dwi = dwi.astype(np.float32).astype(np.float64)
        vp = np.maximum(v, 0.0)
        sv = np.sqrt(vp)
        x = x + (r * dt - 0.5 * dt * vp) + sv * (rho * dwv + rp * dwi)
        v = v + kappa * dt * (theta - vp) + xi * sv * dwv
    return 100.0 * np.exp(x)


def _worst(xi):
    a = _terminals(xi, False)
    b = _terminals(xi, True)
    return float(np.max(np.abs(a - b) / a))


def test_well_conditioned_at_the_measured_parameters():
    assert _worst(0.5) <= 1e-5


def test_ill_conditioned_where_the_feller_condition_fails_badly():
    # FP64 throughout; only the draws' last bits differ
    assert _worst(1.0) >= 1e-3
