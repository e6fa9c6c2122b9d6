"""The FP64 inverse normal's own log, square root and divide
(mp_log_pos_f64, mp_sqrt_pos_f64, mp_div_f64 in mp_device.h), evaluated the
way the kernel evaluates them (one rounding per fused multiply-add, done
here in exact rational arithmetic), against the host's libm over their
argument range, with the constants read from the header.  The hardware
seeds (v_rcp_f64, v_rsq_f64) are modelled as the correctly rounded value
perturbed by 2^-20, worse than the instructions' own error, so the Newton
steps are shown to recover full precision from it.  The GPU tests check the
device functions themselves through the tail normals against the oracle."""
import math
import os
import re
from fractions import Fraction

import numpy as np

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "montepath_amd", "csrc", "mp_device.h")
SEED_ERR = 2.0 ** -20


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _log_body():
    text = open(HEADER).read()
    start = text.index("double mp_log_pos_f64(double x)")
    return text[start:text.index("\n}\n", start)]


def _constants():
    body = _log_body()
    low = float(re.search(r"m < ([0-9.e+-]+);", body).group(1))
    dens = [int(d) for d in re.findall(r"2\.0 / (\d+)\.0", body)]
    hi, lo = re.search(r"fma\(ed, ([0-9.e+-]+), __builtin_fma\(ed, ([0-9.e+-]+), lm\)\)", body).groups()
    return low, dens, float(hi), float(lo)


def div(n, d):
    r = (1.0 / d) * (1.0 + SEED_ERR)
    e = fma(-d, r, 1.0)
    r = fma(e, r, r)
    e = fma(-d, r, 1.0)
    r = fma(e, r, r)
    q = n * r
    return fma(fma(-d, q, n), r, q)


def sqrt_pos(x):
    y = (1.0 / math.sqrt(x)) * (1.0 + SEED_ERR)
    g, h = x * y, 0.5 * y
    r = fma(-g, h, 0.5)
    g, h = fma(g, r, g), fma(h, r, h)
    d = fma(-g, g, x)
    g = fma(d, h, g)
    d = fma(-g, g, x)
    return fma(d, h, g)


def log_pos(x, consts):
    low, dens, ln2_hi, ln2_lo = consts
    m, e = math.frexp(x)
    if m < low:
        m, e = m + m, e - 1
    f = m - 1.0
    s = div(f, 2.0 + f)
    z = s * s
    p = 2.0 / dens[0]
    for k in dens[1:]:
        p = fma(p, z, 2.0 / k)
    lm = fma(s * z, p, s + s)
    ed = float(e)
    return fma(ed, ln2_hi, fma(ed, ln2_lo, lm))


def _ulps(a, b):
    return abs(a - b) / math.ulp(b)


def test_header_constants():
    low, dens, hi, lo = _constants()
    assert low == 0.70710678118654752
    assert dens == list(range(21, 2, -2))  # 2/(2k+1), k = 10 .. 1
    # ln 2 in two parts: the first has 42 significant bits (e * hi exact for
    # |e| < 2^11), the sum is ln 2 to well below an ulp
    assert Fraction(hi).denominator <= 2 ** 42
    ln2 = Fraction(6931471805599453094172321214581765680755, 10 ** 40)
    assert abs(Fraction(hi) + Fraction(lo) - ln2) < Fraction(1, 2 ** 100)


def test_log_within_one_ulp_of_libm():
    consts = _constants()
    rng = np.random.default_rng(1)
    tiny = 2.2250738585072014e-308
    xs = list(np.exp(rng.uniform(math.log(tiny), 0.0, 1500)))
    xs += list(rng.uniform(1e-7, 0.02425, 800))            # the tail's uniforms
    xs += list(rng.integers(1, 1 << 23, 400) * 2.0 ** -23)  # the Sobol grid
    xs += [tiny, 2.0 ** -23, 0.02425, 0.5, 0.70710678118654752, 0.7071067811865476,
           np.nextafter(0.70710678118654752, 0.0), 0.97575, 1.0 - 2.0 ** -23, 1.0]
    worst = max(_ulps(log_pos(float(x), consts), math.log(x)) for x in xs if x != 1.0)
    assert worst <= 1.0, worst
    assert log_pos(1.0, consts) == 0.0


def test_sqrt_and_divide_round_correctly_from_a_coarse_seed():
    rng = np.random.default_rng(2)
    # sqrt arguments: -2 ln(tail) for tail in [2^-1022, 0.97575]
    for v in list(rng.uniform(0.049, 1417.0, 1500)) + [0.0491, 1.0, 2.0, 1416.79]:
        assert sqrt_pos(float(v)) == math.sqrt(v)
    # quotients of the rationals (|n| < 70, d in [0.003, 90]) and s = f / (2 + f)
    for n, d in zip(rng.uniform(-70.0, 70.0, 1500), np.exp(rng.uniform(math.log(0.003), math.log(90.0), 1500))):
        assert div(float(n), float(d)) == n / d
    for f in rng.uniform(-0.2929, 0.4143, 500):
        assert div(float(f), 2.0 + float(f)) == f / (2.0 + f)


C = [-7.784894002430293e-3, -3.223964580411365e-1, -2.400758277161838,
     -2.549732539343734, 4.374664141464968, 2.938163982698783]
D = [7.784695709041462e-3, 3.224671290700398e-1, 2.445134137142996, 3.754408661907416]


def _tail(q, fused):
    step = fma if fused else (lambda a, b, c: a * b + c)
    n = C[0]
    for c in C[1:]:
        n = step(n, q, c)
    d = D[0]
    for c in D[1:]:
        d = step(d, q, c)
    return n / step(d, q, 1.0)


def test_tail_rational_fused_and_unfused_agree_to_a_few_ulp():
    # The GPU fuses the Horner steps and the oracle does not; on the tail
    # (p < 0.02425 on the 23-bit grid) that alone stays within 5 ulp
    # (1.1e-15 relative; the deepest points, p = 2^-23, reach it), which is
    # what the GPU's tail-normal bar against the oracle rests on.
    rng = np.random.default_rng(3)
    ks = list(rng.integers(1, 203424, 2000)) + [1, 2, 203423]
    worst = 0.0
    for k in ks:
        q = math.sqrt(-2.0 * math.log(k * 2.0 ** -23))
        fused, unfused = _tail(q, True), _tail(q, False)
        worst = max(worst, abs(fused - unfused) / abs(unfused))
    assert worst <= 1.2e-15, worst
