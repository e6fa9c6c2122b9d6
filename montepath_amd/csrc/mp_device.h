// mp_device.h -- device helpers of the montepath HIP engine (gfx950, wave64).
#pragma once
#include "mp_inv_normal_f32_coeffs.h"
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MP_WAVE 64

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC 2011).  Counter words:
// (path, step block, stream, seed high word); key: (seed low word, contract).
// ---------------------------------------------------------------------------
struct mp_u4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ mp_u4 mp_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                           uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        // one v_mad_u64_u32 per 32x32 -> 64-bit product gives both halves
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        // three-input XOR in one v_bitop3_b32 (truth table 0x96)
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return mp_u4{c0, c1, c2, c3};
}

// mp_philox for a counter whose words c1..c3 and key are the same in every
// lane of the wave (step block, stream, seed high word; seed low word,
// contract), with only c0 (the path) per lane: the first round's c2 product
// and the second round's c0 product are then wave-uniform and, when the
// caller holds c1..c3 and the key in scalar registers (values the compiler
// knows to be uniform), are formed on the scalar unit.  Same output as
// mp_philox.
__device__ __forceinline__ mp_u4 mp_philox_wave(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                uint32_t k0, uint32_t k1) {
    // round 1: M1 * c2 is uniform
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint32_t p1_hi = __umulhi(0xCD9E8D57u, c2);
    const uint32_t p1_lo = 0xCD9E8D57u * c2;
    uint32_t x0 = p1_hi ^ c1 ^ k0;  // uniform
    uint32_t x1 = p1_lo;            // uniform
    uint32_t x2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
    uint32_t x3 = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
    // round 2: M0 * x0 is uniform.  A vector instruction reads at most one
    // scalar register, so a pair of uniform XOR operands that changes with
    // c1 is combined on the scalar unit first (readfirstlane pins the value
    // there; it is uniform, so the lane read is exact) and the vector side is
    // one XOR.  Terms that do not depend on c1 stay plain, so a caller's loop
    // over c1 can hoist them (readfirstlane is not moved out of loops).
    {
        const uint32_t q0_hi = __umulhi(0xD2511F53u, x0);
        const uint32_t q0_lo = 0xD2511F53u * x0;
        const uint64_t q1 = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(q1 >> 32), x1, k0, 0x96);
        const uint32_t n2 = x3 ^ __builtin_amdgcn_readfirstlane(q0_hi ^ k1);
        x0 = n0;
        x1 = (uint32_t)q1;
        x2 = n2;
        x3 = q0_lo;  // uniform
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    // round 3: x3 is still uniform
    {
        const uint64_t p0r = (uint64_t)0xD2511F53u * x0;
        const uint64_t p1r = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1r >> 32), x1, k0, 0x96);
        const uint32_t n2 = (uint32_t)(p0r >> 32) ^ __builtin_amdgcn_readfirstlane(x3 ^ k1);
        x0 = n0;
        x1 = (uint32_t)p1r;
        x2 = n2;
        x3 = (uint32_t)p0r;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
#pragma unroll
    for (int round = 3; round < 10; ++round) {
        const uint64_t p0r = (uint64_t)0xD2511F53u * x0;
        const uint64_t p1r = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1r >> 32), x1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0r >> 32), x3, k1, 0x96);
        x0 = n0;
        x1 = (uint32_t)p1r;
        x2 = n2;
        x3 = (uint32_t)p0r;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return mp_u4{x0, x1, x2, x3};
}

// Uniform in [2^-24, 1 - 2^-24] with 23 significant bits plus a half step:
// exact in both float and double, so FP32 and FP64 runs see the same draws.
__device__ __forceinline__ float mp_u01f(uint32_t x) {
    // (0x7f:x) >> 9 is the float 1 + k 2^-23 (k = x >> 9); subtracting
    // 1 - 2^-24 is exact (Sterbenz) and leaves (k + 1/2) 2^-23
    return __uint_as_float(__builtin_amdgcn_alignbit(0x7fu, x, 9)) - 0.99999994039535522f;
}
__device__ __forceinline__ double mp_u01d(uint32_t x) {
    return ((double)(x >> 9) + 0.5) * 0x1p-23;
}

// Box-Muller pair.  FP32 uses the CDNA transcendental units directly:
// v_log_f32 is log2, v_sin/v_cos take their argument in revolutions, so
// sin(2*pi*u) is one instruction.
__device__ __forceinline__ void mp_box_muller(uint32_t a, uint32_t b, float& z0, float& z1) {
    const float u1 = mp_u01f(a);
    const float u2 = mp_u01f(b);
    const float rad = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
    z0 = rad * __builtin_amdgcn_cosf(u2);
    z1 = rad * __builtin_amdgcn_sinf(u2);
}
__device__ __forceinline__ void mp_box_muller(uint32_t a, uint32_t b, double& z0, double& z1) {
    const double u1 = mp_u01d(a);
    const double u2 = mp_u01d(b);
    const double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    z0 = rad * c;
    z1 = rad * s;
}

template <class R>
__device__ __forceinline__ void mp_normals4(uint32_t path, uint32_t block, uint32_t stream,
                                            uint32_t seed_hi, uint32_t k0, uint32_t k1, R z[4]) {
    const mp_u4 r = mp_philox(path, block, stream, seed_hi, k0, k1);
    mp_box_muller(r.x, r.y, z[0], z[1]);
    mp_box_muller(r.z, r.w, z[2], z[3]);
}

template <class R>
__device__ __forceinline__ void mp_uniforms4(uint32_t path, uint32_t block, uint32_t stream,
                                             uint32_t seed_hi, uint32_t k0, uint32_t k1, R u[4]) {
    const mp_u4 r = mp_philox(path, block, stream, seed_hi, k0, k1);
    if constexpr (sizeof(R) == 4) {
        u[0] = mp_u01f(r.x);
        u[1] = mp_u01f(r.y);
        u[2] = mp_u01f(r.z);
        u[3] = mp_u01f(r.w);
    } else {
        u[0] = mp_u01d(r.x);
        u[1] = mp_u01d(r.y);
        u[2] = mp_u01d(r.z);
        u[3] = mp_u01d(r.w);
    }
}

// ---------------------------------------------------------------------------
// Precision helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float mp_exp(float x) { return __expf(x); }
__device__ __forceinline__ double mp_exp(double x) { return exp(x); }
__device__ __forceinline__ float mp_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double mp_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float mp_log(float x) {
    return 0.69314718055994531f * __builtin_amdgcn_logf(x);
}
__device__ __forceinline__ double mp_log(double x) { return log(x); }
__device__ __forceinline__ float mp_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double mp_rcp(double x) { return 1.0 / x; }

// FP64 arithmetic of the FP64 inverse normal, for positive normal
// arguments of moderate size only (its callers clamp).  The library's log,
// sqrt and divide spend about 75, 17 and 12 instructions on double-double
// compensation, range scaling and special cases that these arguments do
// not need; the rational approximation that follows (Acklam, 1.15e-9
// relative) could not use the last half-ulp either.
//
// n / d: v_rcp_f64 seed, two Newton steps on the reciprocal, then one
// residual correction of the quotient (Markstein), which leaves it within
// one rounding of n / d.
__device__ __forceinline__ double mp_div_f64(double n, double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(e, r, r);
    e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(e, r, r);
    const double q = n * r;
    return __builtin_fma(__builtin_fma(-d, q, n), r, q);
}

// sqrt x: v_rsq_f64 seed y, then coupled Newton steps on g ~ sqrt x and
// h ~ 1 / (2 sqrt x), the last two on the residual x - g^2.
__device__ __forceinline__ double mp_sqrt_pos_f64(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

// log x: x = 2^e m with m in [sqrt(1/2), sqrt(2)) (m - 1 is exact);
// log m = 2 atanh(s), s = (m - 1) / (m + 1), |s| < 0.1716, summed through
// s^21 with the Taylor coefficients 2 / (2k + 1) (truncation below
// 2.5e-17 relative); e ln 2 is added in two parts, the first with 42
// significant bits so that e * ln2_hi is exact.
__device__ __forceinline__ double mp_log_pos_f64(double x) {
    double m = __builtin_amdgcn_frexp_mant(x);  // [1/2, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
    const bool low = m < 0.70710678118654752;
    m = low ? m + m : m;
    e = low ? e - 1 : e;
    const double f = m - 1.0;
    const double s = mp_div_f64(f, 2.0 + f);
    const double z = s * s;
    double p = 2.0 / 21.0;
    p = __builtin_fma(p, z, 2.0 / 19.0);
    p = __builtin_fma(p, z, 2.0 / 17.0);
    p = __builtin_fma(p, z, 2.0 / 15.0);
    p = __builtin_fma(p, z, 2.0 / 13.0);
    p = __builtin_fma(p, z, 2.0 / 11.0);
    p = __builtin_fma(p, z, 2.0 / 9.0);
    p = __builtin_fma(p, z, 2.0 / 7.0);
    p = __builtin_fma(p, z, 2.0 / 5.0);
    p = __builtin_fma(p, z, 2.0 / 3.0);
    const double lm = __builtin_fma(s * z, p, s + s);
    const double ed = static_cast<double>(e);
    return __builtin_fma(ed, 0.6931471805598903, __builtin_fma(ed, 5.497923018708371e-14, lm));
}

// Compensated (Kahan) running sum in the path precision.
template <class R>
struct mp_kahan {
    R s, c;
    __device__ __forceinline__ void init() { s = R(0); c = R(0); }
    __device__ __forceinline__ void add(R x) {
        const R y = x - c;
        const R t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// ---------------------------------------------------------------------------
// Wave reductions (DPP-backed shuffles on wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mp_wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, MP_WAVE);
    return v;
}

// A double moved between lanes by a DPP pattern within 16-lane rows: 0 swaps
// adjacent lanes, 1 adjacent pairs (quad_perm [1,0,3,2] / [2,3,0,1]), 2
// mirrors each half-row, 3 each row.
__device__ __forceinline__ double mp_dpp_f64(double v, int ctrl_id) {
    // every pattern reads inside the lane's own row, so the bound-control
    // form (no old value to keep) moves the same data
    const long long b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
    switch (ctrl_id) {
        case 0:
            lo = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xf, 0xf, true);
            hi = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xf, 0xf, true);
            break;
        case 1:
            lo = __builtin_amdgcn_mov_dpp(lo, 0x4E, 0xf, 0xf, true);
            hi = __builtin_amdgcn_mov_dpp(hi, 0x4E, 0xf, 0xf, true);
            break;
        case 2:
            lo = __builtin_amdgcn_mov_dpp(lo, 0x141, 0xf, 0xf, true);
            hi = __builtin_amdgcn_mov_dpp(hi, 0x141, 0xf, 0xf, true);
            break;
        default:
            lo = __builtin_amdgcn_mov_dpp(lo, 0x140, 0xf, 0xf, true);
            hi = __builtin_amdgcn_mov_dpp(hi, 0x140, 0xf, 0xf, true);
            break;
    }
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// Sum over each 16-lane row of a wave in a fixed order (pairs, quads,
// half-rows, rows) with DPP moves, no LDS traffic; every lane of a row
// returns that row's sum.
__device__ __forceinline__ double mp_row_sum_dpp(double v) {
    v += mp_dpp_f64(v, 0);
    v += mp_dpp_f64(v, 1);
    v += mp_dpp_f64(v, 2);
    v += mp_dpp_f64(v, 3);
    return v;
}

// Neumaier summation of a sequence of doubles, used for every cross-tile and
// cross-block total so that the global sums are compensated and in a fixed
// order.
struct mp_neumaier {
    double s, c;
    __device__ __host__ void init() { s = 0.0; c = 0.0; }
    __device__ __host__ void add(double x) {
        const double t = s + x;
        if (fabs(s) >= fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    __device__ __host__ double value() const { return s + c; }
};

// ---------------------------------------------------------------------------
// Acklam inverse normal (the reference's inverse_standard_normal,
// crates/mc-core/src/runtime/cpu.rs:6012-6061, including its clamp of p to
// [f64::MIN_POSITIVE, 1 - f64::EPSILON]).  For p == 0 the FP32 form uses
// ln(f64::MIN_POSITIVE) directly so both precisions give the same tail value.
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ R mp_acklam(R p) {
    const R a0 = R(-3.969683028665376e1), a1 = R(2.209460984245205e2),
            a2 = R(-2.759285104469687e2), a3 = R(1.383577518672690e2),
            a4 = R(-3.066479806614716e1), a5 = R(2.506628277459239);
    const R b0 = R(-5.447609879822406e1), b1 = R(1.615858368580409e2),
            b2 = R(-1.556989798598866e2), b3 = R(6.680131188771972e1),
            b4 = R(-1.328068155288572e1);
    const R c0 = R(-7.784894002430293e-3), c1 = R(-3.223964580411365e-1),
            c2 = R(-2.400758277161838), c3 = R(-2.549732539343734), c4 = R(4.374664141464968),
            c5 = R(2.938163982698783);
    const R d0 = R(7.784695709041462e-3), d1 = R(3.224671290700398e-1),
            d2 = R(2.445134137142996), d3 = R(3.754408661907416);
    const R p_low = R(0.02425);
    const R p_high = R(1) - p_low;
    if constexpr (sizeof(R) == 8) {
        // Straight-line FP64: the tail branch (4.85% of uniforms) is taken
        // by some lane in almost every wave, so every lane evaluates both
        // rationals, each with its branch's operations in its branch's
        // order, and divides the pair its branch selects.  Two Horner
        // chains (19 FMAs) issue fewer instructions than one chain with
        // per-lane coefficient selects (10 FMAs, 22 selects).
        const bool central = p >= p_low && p <= p_high;
        const bool lower = p < p_low;
        const R tail = lower ? p : R(1) - p;
        const R q = mp_sqrt_pos_f64(R(-2) * mp_log_pos_f64(tail > R(2.2250738585072014e-308)
                                                                ? tail
                                                                : R(2.2250738585072014e-308)));
        const R qc = p - R(0.5);
        const R r = qc * qc;
        const R nc = (((((a0 * r + a1) * r + a2) * r + a3) * r + a4) * r + a5) * qc;
        const R dc = ((((b0 * r + b1) * r + b2) * r + b3) * r + b4) * r + R(1);
        const R nt = ((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5;
        const R dt = (((d0 * q + d1) * q + d2) * q + d3) * q + R(1);
        const R z = mp_div_f64(central ? nc : nt, central ? dc : dt);
        return central || lower ? z : -z;
    }
    if (p >= p_low && p <= p_high) {
        const R q = p - R(0.5);
        const R r = q * q;
        const R num = (((((a0 * r + a1) * r + a2) * r + a3) * r + a4) * r + a5) * q;
        const R den = ((((b0 * r + b1) * r + b2) * r + b3) * r + b4) * r + R(1);
        return num * mp_rcp(den);
    }
    // tails: q = sqrt(-2 ln(min(p, 1-p))); p == 0 is clamped to f64::MIN_POSITIVE
    const bool lower = p < p_low;
    const R tail = lower ? p : R(1) - p;
    R lnp;
    if constexpr (sizeof(R) == 4)
        lnp = tail > R(0) ? mp_log(tail) : R(-708.3964185322641);
    else
        lnp = mp_log(tail > R(2.2250738585072014e-308) ? tail : R(2.2250738585072014e-308));
    const R q = mp_sqrt(R(-2) * lnp);
    const R num = ((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5;
    const R den = (((d0 * q + d1) * q + d2) * q + d3) * q + R(1);
    const R z = num * mp_rcp(den);
    return lower ? z : -z;
}

// ---------------------------------------------------------------------------
// FP32 inverse normal for the FP32 paths: Phi^-1(p) = x R(w), x = 2p - 1,
// w = -ln(1 - x^2).  Central branch (w < 7, all but 0.046% of the
// uniforms): R is a degree-9 polynomial in l = log2(1 - x^2), the log's
// own output, fitted for this build (gen_inv_normal_f32.py: reweighted
// Chebyshev least squares on the relative error of sqrt(2) erfinv(x) / x;
// 3.3e-7 relative evaluated in FP32 over the whole uniform grid).  Tail (w >= 7): the w >= 5 branch of
// Giles' single-precision erfinv ("Approximating the erfinv function", GPU
// Computing Gems, 2011), coefficients times sqrt(2).  The wide central
// branch keeps the divergent tail out of most waves: with w < 5 as in
// Giles' own split, some lane of a 64-lane wave took the tail on about one
// draw in five.  Acklam's rational form, which the FP64 paths keep because
// it is the reference's, loses up to 1e-4 relative in FP32 to cancellation
// between its alternating coefficients, and its tail branch (4.85% of
// draws) diverges in almost every wave.  p lies on the 2^-23 grid in [0,1)
// and arrives as the bits of the float 1 + p; 4p(1-p) is formed exactly up
// to one rounding.  p == 0 returns the FP64 reference value at
// f64::MIN_POSITIVE; it lies in the tail (log2 0 = -inf), so the test for
// it costs nothing on the central branch.
// ---------------------------------------------------------------------------
// The parts of mp_inv_normal_f32, for callers that form the central values
// of several draws before one shared check for the rare tail:
// mp_invn_central(fbits, x, l) returns R on the central branch and sets
// x = 2p - 1 and l = log2(1 - x^2); the draw lies in the tail when
// !(l > MP_INVN_L_CENTRAL), where mp_invn_tail(fbits, l) gives R instead;
// the normal is R * x.
__device__ __forceinline__ float mp_invn_central(uint32_t fbits, float& x, float& l) {
    // fbits = 0x3f800000 | t23 is the float 1 + p, p = t23 * 2^-23
    const float f = __uint_as_float(fbits);
    x = __builtin_fmaf(f, 2.0f, -3.0f);  // 2p - 1, exact
    // 1 - x^2 = 4p(1-p) exactly; the FMA rounds it once.  l = log2 of it,
    // w = -ln(4p(1-p)) = -ln2 * l (the tail's variable).
    l = __builtin_amdgcn_logf(__builtin_fmaf(-x, x, 1.0f));
    float r = MP_INVN_C9;  // w < 7: a polynomial in l itself
    r = MP_INVN_C8 + r * l;
    r = MP_INVN_C7 + r * l;
    r = MP_INVN_C6 + r * l;
    r = MP_INVN_C5 + r * l;
    r = MP_INVN_C4 + r * l;
    r = MP_INVN_C3 + r * l;
    r = MP_INVN_C2 + r * l;
    r = MP_INVN_C1 + r * l;
    r = MP_INVN_C0 + r * l;
    return r;
}

__device__ __forceinline__ float mp_invn_tail(uint32_t fbits, float l) {
    const float w = __builtin_amdgcn_sqrtf(-0.69314718f * l) - 3.0f;
    float r = -0.000283145724f;
    r = 0.000142765653f + r * w;
    r = 0.00190825947f + r * w;
    r = -0.00519501232f + r * w;
    r = 0.00811688974f + r * w;
    r = -0.0107797878f + r * w;
    r = 0.0133485785f + r * w;
    r = 1.41658103f + r * w;
    r = 4.00643444f + r * w;
    // p == 0: x = -1, so r * x is exactly the tail value
    return fbits == 0x3f800000u ? 37.51937936814848f : r;
}

__device__ __forceinline__ float mp_inv_normal_f32(uint32_t fbits) {
    float x, l;
    float r = mp_invn_central(fbits, x, l);
    if (!(l > MP_INVN_L_CENTRAL)) r = mp_invn_tail(fbits, l);
    return r * x;
}

// max(v, 0) for finite v in one instruction (v_med3 with FLT_MAX; a plain
// max also quiets signalling NaNs first).
__device__ __forceinline__ float mp_relu(float v) { return __builtin_amdgcn_fmed3f(v, 0.0f, 3.40282347e38f); }
__device__ __forceinline__ double mp_relu(double v) { return v > 0.0 ? v : 0.0; }

// ---------------------------------------------------------------------------
// Owen-scrambled Sobol (hash-based nested uniform scrambling, Burley 2020).
// lk_scramble lets lower bits influence only higher bits; applied to
// bit-reversed values it is a nested uniform scramble of the digits.
// ---------------------------------------------------------------------------
__device__ __host__ __forceinline__ uint32_t mp_hash32(uint32_t x) {
    x ^= 0x79c68e4au;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// scramble with an already-hashed seed
__device__ __host__ __forceinline__ uint32_t mp_lk(uint32_t n, uint32_t hs) {
    n ^= n * 0x3d20adeau;
    n += hs;
    n *= (hs >> 16) | 1u;
    n ^= n * 0x05526c56u;
    n ^= n * 0x53a22864u;
    return n;
}

// The same permutation, exact in bits 0..23 only, with full-rate 24-bit
// multiplies: every step maps bits 0..23 to bits 0..23 (products and sums
// carry upward), and a uniform needs only 23 of those bits.
__device__ __forceinline__ uint32_t mp_lk_low24(uint32_t n, uint32_t hs) {
    n ^= __umul24(n, 0x3d20adeau);
    n += hs;
    n = __umul24(n, (hs >> 16) | 1u);
    n ^= __umul24(n, 0x05526c56u);
    n ^= __umul24(n, 0x53a22864u);
    return n;
}

// v_mul_u32_u24 with a wave-uniform first operand: the low 32 bits of
// (a mod 2^24) * (b mod 2^24) at full rate.  Inline asm because, seeing that
// only low product bits are used, the compiler drops the operand masks of
// __umul24 and falls back to the quarter-rate 32-bit multiply.
__device__ __forceinline__ uint32_t mp_mul24_su(uint32_t a_uniform, uint32_t b) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "s"(a_uniform), "v"(b));
    return r;
}

__device__ __forceinline__ uint32_t mp_mul24_vv(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// mp_lk_low24 with the hashed seed hs and its multiplier hm = (hs >> 16) | 1
// in vector registers (read from a table).
__device__ __forceinline__ uint32_t mp_lk_low24_vv(uint32_t n, uint32_t hs, uint32_t hm) {
    n ^= mp_mul24_su(0x3d20adeau, n);
    n += hs;
    n = mp_mul24_vv(hm, n);
    n ^= mp_mul24_su(0x05526c56u, n);
    n ^= mp_mul24_su(0x53a22864u, n);
    return n;
}

// mp_lk_low24 for a wave-uniform hashed seed.
__device__ __forceinline__ uint32_t mp_lk_low24_su(uint32_t n, uint32_t hs_uniform) {
    n ^= mp_mul24_su(0x3d20adeau, n);
    n += hs_uniform;
    n = mp_mul24_su((hs_uniform >> 16) | 1u, n);
    n ^= mp_mul24_su(0x05526c56u, n);
    n ^= mp_mul24_su(0x53a22864u, n);
    return n;
}

__device__ __forceinline__ uint32_t mp_rev32(uint32_t x) { return __builtin_bitreverse32(x); }

// Shuffled 16-bit point index for a local index (< 2^16) under a seed.
__device__ __forceinline__ uint32_t mp_sobol_shuffle(uint32_t local, uint32_t hashed_seed) {
    return mp_rev32(mp_lk(mp_rev32(local & 0xFFFFu), hashed_seed)) & 0xFFFFu;
}

// Cross-block publication without agent-scope fences.  On gfx950 each XCD
// has its own L2 and an agent-scope fence writes back and invalidates it,
// which costs microseconds per wave.  Values meant for another block are
// instead stored and loaded as relaxed agent-scope atomics (coherent across
// XCDs, not held in a non-coherent L2), and only their order relative to
// the signalling atomic is enforced: a workgroup fence plus a wait for all
// outstanding vector-memory operations of the wave.
__device__ __forceinline__ void mp_store_coherent(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), __double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double mp_load_coherent(const double* p) {
    return __longlong_as_double((long long)__hip_atomic_load(
        reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// s_waitcnt immediate for gfx9: vmcnt = 0 (bits 3:0 and 15:14), expcnt and
// lgkmcnt at their maxima (no wait).
__device__ __forceinline__ void mp_vmem_drain_release() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));
}

__device__ __forceinline__ void mp_vmem_drain_acquire() {
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
