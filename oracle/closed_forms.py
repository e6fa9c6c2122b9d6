"""Closed-form and semi-analytic prices used to pin the oracle and the engine.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* Black-Scholes with the exact normal CDF (the reference's fixtures use
  exact-Phi values; its own function ``black_scholes_european_call_price``,
  cpu.rs:839-859, uses the A&S 7.1.26 approximation and lives in the C
  oracle as ``bs_call(..., exact=False)``).
* Heston (1993) call by two independent quadratures of the characteristic
  function -- the P1/P2 form in the "little Heston trap" parametrisation
  (Albrecher, Mayer, Schoutens & Tistaert 2007) and the Lewis (2000)
  single-integral form.  The reference has no Heston closed form.
* Discrete geometric-average Asian call and geometric basket call
  (lognormal closed forms; textbook, e.g. Kemna & Vorst 1990).
"""
from __future__ import annotations

import math
import warnings

import numpy as np
from scipy import integrate, special

warnings.filterwarnings("ignore", category=integrate.IntegrationWarning)


def norm_cdf(x: float) -> float:
    return 0.5 * special.erfc(-x / math.sqrt(2.0))


def bs_call(s0: float, k: float, r: float, sigma: float, t: float) -> float:
    if sigma == 0.0 or t == 0.0:
        return max(s0 - k * math.exp(-r * t), 0.0)
    st = math.sqrt(t)
    d1 = (math.log(s0 / k) + (r + 0.5 * sigma * sigma) * t) / (sigma * st)
    d2 = d1 - sigma * st
    return s0 * norm_cdf(d1) - k * math.exp(-r * t) * norm_cdf(d2)


def bs_put(s0, k, r, sigma, t) -> float:
    return bs_call(s0, k, r, sigma, t) + k * math.exp(-r * t) - s0


def _heston_cf_logfwd(u: complex, v0, kappa, theta, xi, rho, t) -> complex:
    """E[exp(i u ln(S_T / F_T))], F_T the forward, little-trap form."""
    iu = 1j * u
    beta = kappa - rho * xi * iu
    d = np.sqrt(beta * beta + xi * xi * (iu + u * u))
    g = (beta - d) / (beta + d)
    edt = np.exp(-d * t)
    c = kappa * theta / (xi * xi) * ((beta - d) * t - 2.0 * np.log((1.0 - g * edt) / (1.0 - g)))
    dd = (beta - d) / (xi * xi) * (1.0 - edt) / (1.0 - g * edt)
    return np.exp(c + dd * v0)


def heston_call_p1p2(s0, k, r, v0, kappa, theta, xi, rho, t) -> float:
    """Heston (1993) P1/P2 form: C = S0 P1 - K e^{-rT} P2, with
    P1 from the share-measure CF phi(u - i)/phi(-i) and P2 from phi(u)."""
    fwd = s0 * math.exp(r * t)
    x = math.log(fwd / k)

    def phi(u):
        return _heston_cf_logfwd(u, v0, kappa, theta, xi, rho, t)

    def integrand_p2(u):
        return (np.exp(1j * u * x) * phi(u) / (1j * u)).real

    def integrand_p1(u):
        return (np.exp(1j * u * x) * phi(u - 1j) / (1j * u)).real

    opts = dict(limit=1000, epsabs=1e-13, epsrel=1e-13)
    p1 = 0.5 + integrate.quad(integrand_p1, 0.0, np.inf, **opts)[0] / math.pi
    p2 = 0.5 + integrate.quad(integrand_p2, 0.0, np.inf, **opts)[0] / math.pi
    return s0 * p1 - k * math.exp(-r * t) * p2


def heston_call_lewis(s0, k, r, v0, kappa, theta, xi, rho, t) -> float:
    """Lewis (2000): C = S0 - sqrt(S0 K e^{-rT})/pi * int Re[e^{iux} phi(u - i/2)]/(u^2+1/4)."""
    fwd = s0 * math.exp(r * t)
    x = math.log(fwd / k)

    def integrand(u):
        val = np.exp(1j * u * x) * _heston_cf_logfwd(u - 0.5j, v0, kappa, theta, xi, rho, t)
        return val.real / (u * u + 0.25)

    opts = dict(limit=1000, epsabs=1e-13, epsrel=1e-13)
    integral = integrate.quad(integrand, 0.0, np.inf, **opts)[0]
    return s0 - math.sqrt(s0 * k * math.exp(-r * t)) * integral / math.pi


def geometric_asian_call(s0, k, r, sigma, t, n_steps) -> float:
    """Discrete geometric average over monitoring dates 1..M (S0 excluded),
    the same dates the reference's arithmetic Asian uses (cpu.rs:6578-6584)."""
    m = n_steps
    dt = t / m
    mu = math.log(s0) + (r - 0.5 * sigma * sigma) * dt * (m + 1) / 2.0
    var = sigma * sigma * dt * (m + 1) * (2 * m + 1) / (6.0 * m)
    sd = math.sqrt(var)
    d2 = (mu - math.log(k)) / sd
    d1 = d2 + sd
    return math.exp(-r * t) * (math.exp(mu + 0.5 * var) * norm_cdf(d1) - k * norm_cdf(d2))


def geometric_basket_call(s0, sigma, weight, corr, r, t, k) -> float:
    """Call on prod S_i^{w_i} (weights normalised to one): a lower bound for
    the arithmetic basket call with the same normalised weights."""
    s0, sigma, w = (np.asarray(a, dtype=np.float64) for a in (s0, sigma, weight))
    w = w / w.sum()
    corr = np.asarray(corr, dtype=np.float64)
    mu = float(np.sum(w * (np.log(s0) + (r - 0.5 * sigma * sigma) * t)))
    cov = np.outer(sigma, sigma) * corr * t
    var = float(w @ cov @ w)
    sd = math.sqrt(var)
    d2 = (mu - math.log(k)) / sd
    d1 = d2 + sd
    return math.exp(-r * t) * (math.exp(mu + 0.5 * var) * norm_cdf(d1) - k * norm_cdf(d2))
