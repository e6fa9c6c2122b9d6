// mp_engine.hip -- HIP kernels of the montepath Monte Carlo path engine for
// MI355X (gfx950, CDNA4, wave64).
//
// Layout and execution model
//   * one lane per path; a wave owns a tile of tile_paths consecutive paths
//     of one request and loops over it 64*PPL paths at a time;
//   * waves pull tiles from a device work queue (one atomic per tile), so a
//     launch can carry thousands of uneven requests (persistent kernel);
//   * each tile writes (sum, sum of squares) per strike in FP64 to its own
//     slot; a finalize kernel sums a request's tiles in tile order with
//     Neumaier compensation, so results are bit-identical however tiles were
//     scheduled and whatever else shared the launch;
//   * path arithmetic in R = float or double; normals from Philox4x32-10 in
//     registers, from a caller buffer (parity), or from Owen-scrambled Sobol
//     with a streaming Brownian bridge (LDS).
#include <algorithm>

#include "mp_device.h"
#include "mp_engine.h"

namespace {

constexpr int kNibWords = 256 * 128 / 2;  // Sobol index tables: [dim][128] 16-bit words (64 KiB)

// ---------------------------------------------------------------------------
// Per-path model state
// ---------------------------------------------------------------------------
template <class R>
struct path_state {
    R x;             // log-return since t0
    R v;             // Heston variance
    mp_kahan<R> acc; // Asian running sum of S/S0
    R xmax;          // lookback: max log-return (>= 0, S0 included)
    bool knocked;    // down-and-out
};

template <class R>
struct model_params {
    R drift_dt, vol_dt, xb;
    R kappa_dt, theta, xi_sqdt, rho, rho_perp, rate_dt, half_dt, sqrt_dt;
    R sigma, xi;  // per unit of Brownian increment (increment-form steps)
    R mu_j, sigma_j;
    int payoff, scheme, model;
};

// A request-wide parameter in the path precision.  FP32 values are moved to
// a scalar register (the double-to-float conversion is a vector op), so they
// hold no vector register for the kernel's lifetime; the value is unchanged.
template <class R>
__device__ __forceinline__ R mp_param(double x) {
    if constexpr (sizeof(R) == 4)
        return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, (float)x)));
    else
        return x;
}

template <class R>
__device__ __forceinline__ model_params<R> load_params(const mp_dreq& q) {
    model_params<R> m;
    m.drift_dt = mp_param<R>(q.drift_dt);
    m.vol_dt = mp_param<R>(q.vol_dt);
    m.xb = mp_param<R>(q.log_barrier_rel);
    m.kappa_dt = mp_param<R>(q.kappa * q.dt);
    m.theta = mp_param<R>(q.theta);
    m.xi_sqdt = mp_param<R>(q.xi * q.sqrt_dt);
    m.rho = mp_param<R>(q.rho);
    m.rho_perp = mp_param<R>(q.rho_perp);
    m.rate_dt = mp_param<R>(q.rate * q.dt);
    m.half_dt = mp_param<R>(0.5 * q.dt);
    m.sqrt_dt = mp_param<R>(q.sqrt_dt);
    m.sigma = mp_param<R>(q.vol_dt / q.sqrt_dt);
    m.xi = mp_param<R>(q.xi);
    m.mu_j = mp_param<R>(q.mu_j);
    m.sigma_j = mp_param<R>(q.sigma_j);
    m.payoff = q.payoff;
    m.scheme = q.scheme;
    m.model = q.model;
    return m;
}

template <class R>
__device__ __forceinline__ void state_init(path_state<R>& st, const mp_dreq& q) {
    st.x = R(0);
    st.v = R(q.v0);
    st.acc.init();
    st.xmax = R(0);
    st.knocked = q.knocked_at_start != 0;
}

// Path-dependent observation after a step (payoff is wave-uniform).
template <class R>
__device__ __forceinline__ void observe(path_state<R>& st, const model_params<R>& m) {
    if (m.payoff == MP_PAY_ASIAN) {
        st.acc.add(mp_exp(st.x));
    } else if (m.payoff == MP_PAY_DOWN_AND_OUT) {
        st.knocked = st.knocked || (st.x <= m.xb);
    } else if (m.payoff == MP_PAY_LOOKBACK) {
        st.xmax = st.x > st.xmax ? st.x : st.xmax;
    }
}

// GBM log-Euler step (reference cpu.rs:6464: log_s += drift_dt + vol_dt * z).
template <class R>
__device__ __forceinline__ void gbm_step(path_state<R>& st, const model_params<R>& m, R z) {
    st.x += m.drift_dt + m.vol_dt * z;
    observe(st, m);
}

// Heston Euler step (reference heston_path_terminal cpu.rs:8674-8683; scheme 1
// keeps the unfloored variance -- full truncation).
template <class R>
__device__ __forceinline__ void heston_step(path_state<R>& st, const model_params<R>& m, R z_var,
                                            R z_ind) {
    const R vp = st.v > R(0) ? st.v : R(0);
    const R sv = mp_sqrt(vp);
    const R z_price = m.rho * z_var + m.rho_perp * z_ind;
    st.x += (m.rate_dt - m.half_dt * vp) + sv * m.sqrt_dt * z_price;
    R v = st.v + m.kappa_dt * (m.theta - vp) + m.xi_sqdt * sv * z_var;
    if (m.scheme == 0) v = v > R(0) ? v : R(0);
    st.v = v;
    observe(st, m);
}

// The same steps driven by Brownian increments dW = sqrt(dt) z (the FP32
// Sobol/bridge kernel, whose bridge produces increments): one multiply fewer
// per factor, equal to the steps above up to rounding.
template <class R>
__device__ __forceinline__ void gbm_step_dw(path_state<R>& st, const model_params<R>& m, R dw) {
    st.x += m.drift_dt + m.sigma * dw;
    observe(st, m);
}

template <class R>
__device__ __forceinline__ void heston_step_dw(path_state<R>& st, const model_params<R>& m, R dw_var,
                                               R dw_ind) {
    const R vp = mp_relu(st.v);
    const R sv = mp_sqrt(vp);
    const R dw_price = m.rho * dw_var + m.rho_perp * dw_ind;
    st.x += (m.rate_dt - m.half_dt * vp) + sv * dw_price;
    R v = st.v + m.kappa_dt * (m.theta - vp) + m.xi * sv * dw_var;
    if (m.scheme == 0) v = mp_relu(v);
    st.v = v;
    observe(st, m);
}

// The European Heston step of the FP32 register bridge, whose leaves carry
// the variance factor's increment times xi (u_var = xi dW_var) and the
// independent factor's times rho_perp (u_ind = rho_perp dW_ind), so the
// price increment is one FMA and the variance noise needs no multiply by
// xi.  x_T is linear in the per-step rate term, so the caller adds
// n rate_dt once after the last step.
__device__ __forceinline__ void heston_step_u(path_state<float>& st, const model_params<float>& m,
                                              float rho_over_xi, float u_var, float u_ind) {
    const float vp = mp_relu(st.v);
    const float sv = mp_sqrt(vp);
    const float dw_price = rho_over_xi * u_var + u_ind;  // rho dW_var + rho_perp dW_ind
    st.x = sv * dw_price + (st.x - m.half_dt * vp);
    float v = (st.v + m.kappa_dt * (m.theta - vp)) + sv * u_var;
    if (m.scheme == 0) v = mp_relu(v);
    st.v = v;
}

// Poisson count by inversion of u against the request's cumulative
// probabilities (reference sample_poisson cpu.rs:932-953; the same recursion
// in FP64, so both precisions count identical jumps).
__device__ __forceinline__ int poisson_count(double u, const mp_dreq& q) {
    if (q.jump_mean_dt == 0.0) return 0;
    int n = 0;
    const int nc = q.n_cum;
    while (n < nc && u > q.cum[n]) ++n;
    if (n < nc) return n;
    // beyond the cached terms: continue the reference recursion
    double prob = q.cum[nc - 1] - q.cum[nc - 2];
    double cum = q.cum[nc - 1];
    int k = nc - 1;
    while (u > cum && k < 4095) {
        k += 1;
        prob *= q.jump_mean_dt / (double)k;
        cum += prob;
    }
    return k;
}

// Merton stepped jump diffusion: compensated drift, diffusion and a
// compound-Poisson jump N*mu_j + sigma_j*sqrt(N)*zj.
template <class R>
__device__ __forceinline__ void merton_apply(path_state<R>& st, const model_params<R>& m, R z,
                                             int n_jumps, R zj) {
    R shock = m.drift_dt + m.vol_dt * z;
    if (n_jumps != 0) shock += R(n_jumps) * m.mu_j + m.sigma_j * mp_sqrt(R(n_jumps)) * zj;
    st.x += shock;
    observe(st, m);
}

// Coarse path of a multilevel level (MlmcCoupled, cpu.rs:6812-6849): the sum
// of `refine` fine normals over sqrt(refine) drives one coarse step.
template <class R>
struct mlmc_coarse {
    R x;             // coarse log-return
    mp_kahan<R> acc; // running sum of coarse S/S0
    R zsum;          // fine normals of the current coarse step
    int cnt;
};

template <class R>
__device__ __forceinline__ void mlmc_step(path_state<R>& st, mlmc_coarse<R>& c, const model_params<R>& m,
                                          R z, int refine, R coarse_drift, R coarse_vol, R sqrt_refine) {
    st.x += m.drift_dt + m.vol_dt * z;
    st.acc.add(mp_exp(st.x));
    if (refine > 0) {
        c.zsum += z;
        if (++c.cnt == refine) {
            c.x += coarse_drift + coarse_vol * (c.zsum / sqrt_refine);
            c.acc.add(mp_exp(c.x));
            c.zsum = R(0);
            c.cnt = 0;
        }
    }
}

// Undiscounted payoff for strike k (discounting happens once, in FP64, in the
// finalize kernel).
template <class R>
__device__ __forceinline__ R payoff_of(const path_state<R>& st, const model_params<R>& m,
                                       const mp_dreq& q, int n_steps, R s0, R k) {
    if (m.payoff == MP_PAY_ASIAN) {
        const R avg = s0 * (st.acc.s * (R(1) / R(n_steps)));
        return avg > k ? avg - k : R(0);
    }
    if (m.payoff == MP_PAY_LOOKBACK) {
        const R smax = s0 * mp_exp(st.xmax);
        return smax > k ? smax - k : R(0);
    }
    const R sT = s0 * mp_exp(st.x);
    if (m.payoff == MP_PAY_DOWN_AND_OUT && st.knocked) return R(0);
    return sT > k ? sT - k : R(0);
}

// ---------------------------------------------------------------------------
// Work-queue tile kernel (GBM / Heston / Merton).  SRC 0: Philox, 1: buffer.
// ---------------------------------------------------------------------------
// Claim the next queue position for the whole wave.  Lane 0 takes it with one
// atomic and parks it in the wave's LDS slot; every lane then reads the slot,
// so all active lanes hold the same value by data flow, whatever exec mask the
// compiler's loop structurisation leaves at this point.
__device__ __forceinline__ unsigned int next_queue_position(unsigned int* counter,
                                                            volatile unsigned int* slot,
                                                            int lane, unsigned int k) {
    if (lane == 0) *slot = atomicAdd(counter, k);
    __builtin_amdgcn_wave_barrier();
    return __builtin_amdgcn_readfirstlane(*slot);
}

// Standard normal of a structured sampler at (point, dimension), in FP64 for
// both path precisions.  Randomized Halton (RandomizedHalton::standard_normal,
// cpu.rs:5706-5713 with halton_with_shift :6063-6077): the radical inverse of
// point + 1 in the dimension's prime base plus the dimension's shift, mod 1.
// Latin hypercube (cpu.rs:5764-5778): stratum (multiplier * point + offset)
// mod n, jitter from the dimension's seed hashed with the point, clamped to
// [f64::MIN_POSITIVE, 1 - f64::EPSILON].  Then Acklam's inverse normal.  The
// Halton digits are accumulated without multiply-add fusion, as the
// reference rounds them.
__device__ __noinline__ double structured_normal(const mp_sampler_dim* tab, const mp_dreq& q,
                                                 uint64_t point, uint32_t dim) {
#pragma clang fp contract(off)
    const mp_sampler_dim sd = tab[q.sampler_off + dim];
    double u;
    if (q.sampling == MP_SAMPLER_HALTON) {
        // point + 1 < 2^32 (requests hold at most 2^32 - 1 paths): 32-bit digits
        const uint32_t base = (uint32_t)sd.a;
        uint32_t index = (uint32_t)point + 1u;
        double f = 1.0 / (double)base, value = 0.0;
        while (index > 0) {
            const uint32_t digit = index % base;
            value = value + f * (double)digit;
            index /= base;
            f = f / (double)base;
        }
        value = value > 2.2250738585072014e-308 ? value : 2.2250738585072014e-308;
        const double shifted = value + sd.shift;
        u = shifted - floor(shifted);
    } else {
        const uint64_t n = q.sampler_n;
        const uint64_t stratum = n == 0 ? 0 : (sd.a * point + sd.b) % n;
        uint64_t h = (sd.c ^ (point * 0xBF58476D1CE4E5B9ull)) + 0x9E3779B97F4A7C15ull;
        h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
        h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
        h ^= h >> 31;  // splitmix64
        double jitter = (double)(h >> 11) * 0x1p-53;
        jitter = jitter > 2.2250738585072014e-308 ? jitter : 2.2250738585072014e-308;
        u = ((double)stratum + jitter) / (double)n;
        u = u < 2.2250738585072014e-308 ? 2.2250738585072014e-308 : u;
        u = u > 1.0 - 2.220446049250313e-16 ? 1.0 - 2.220446049250313e-16 : u;
    }
    return mp_acklam<double>(u);
}

template <class R, bool U, bool HT = false>
struct sobol_cursor;

// Normal of a structured source at (point, dimension): randomized Halton and
// Latin hypercube from their tables; scrambled Sobol (multilevel levels) from
// the index tables, block seed and 256-dimension wrap of the Sobol kernels.
template <class R>
__device__ __forceinline__ R structured_draw(const mp::tile_launch& L, const mp_dreq& q, uint64_t point,
                                             uint32_t dim);

// Adds one unit's contribution to the FP64 sums.  Standard: the payoff per
// strike.  Antithetic: the pair estimate (payoff(+z) + payoff(-z)) / 2 per
// strike (the reference's block estimate, cpu.rs:7807-7835).  Control
// variate (one strike): discounted payoff P and control C = discount * S_T
// into sum[0] (P), sq[0] (P^2), sum[1] (C), sq[1] (C^2), sum[2] (P*C), which
// land in partial slots 0..4 (moments of cpu.rs:6202-6230).
template <class R>
__device__ __forceinline__ void accumulate_unit(const path_state<R>& st, const path_state<R>& sa,
                                                const model_params<R>& m, const mp_dreq& q,
                                                int tech, int n_steps, R s0, const R* kk, int ns,
                                                double* sum, double* sq) {
    if (tech == MP_TECH_CONTROL_VARIATE) {
        const double pd = q.discount * (double)payoff_of(st, m, q, n_steps, s0, kk[0]);
        const double cd = q.discount * (double)(s0 * mp_exp(st.x));
        sum[0] += pd;
        sq[0] += pd * pd;
        sum[1] += cd;
        sq[1] += cd * cd;
        sum[2] += pd * cd;
        return;
    }
#pragma unroll
    for (int sidx = 0; sidx < MP_NS_MAX; ++sidx) {
        if (sidx < ns) {
            double pay = (double)payoff_of(st, m, q, n_steps, s0, kk[sidx]);
            if (tech == MP_TECH_ANTITHETIC) pay = 0.5 * (pay + (double)payoff_of(sa, m, q, n_steps, s0, kk[sidx]));
            sum[sidx] += pay;
            sq[sidx] += pay * pay;
        }
    }
}

// One tile of one request.  MODEL / PAY / SCH >= 0 compile the model,
// payoff and Heston scheme in (the request must hold those values); -1 reads
// them from the request at run time.
template <class R, int SRC, int PPL, int MODEL, int PAY, int SCH>
__device__ __forceinline__ void tile_body(const mp::tile_launch& L, uint32_t tile, const mp_dreq& q,
                                          int lane) {
    model_params<R> m = load_params<R>(q);
    if constexpr (MODEL >= 0) m.model = MODEL;
    if constexpr (PAY >= 0) m.payoff = PAY;
    if constexpr (SCH >= 0) m.scheme = SCH;
    // specialised bodies run the standard technique only
    const int tech = MODEL < 0 ? q.technique : MP_TECH_STANDARD;
    const int n_steps = q.n_steps;
    const int ns = q.n_strikes;
    const R s0 = R(q.s0);
    R kk[MP_NS_MAX];
#pragma unroll
    for (int j = 0; j < MP_NS_MAX; ++j) kk[j] = R(q.strikes[j]);
    // multilevel coupling (PAY == MP_PAY_ASIAN_MLMC)
    const int refine = q.mlmc_refine;
    const R c_drift = R(q.mlmc_coarse_drift), c_vol = R(q.mlmc_coarse_vol);
    const R sqrt_ref = R(q.mlmc_sqrt_refine);
    // unit = path, or antithetic pair (the pair's second path takes -z)
    const uint64_t pbeg = (uint64_t)(tile - q.tile0) * q.tile_paths;
    const uint64_t pend = min(pbeg + (uint64_t)q.tile_paths, q.n_paths);

    double sum[MP_NS_MAX], sq[MP_NS_MAX];
#pragma unroll
    for (int j = 0; j < MP_NS_MAX; ++j) sum[j] = sq[j] = 0.0;

    for (uint64_t base = pbeg; base < pend; base += (uint64_t)MP_WAVE * PPL) {
        path_state<R> st[PPL], sa[PPL];
        mlmc_coarse<R> cs[PPL];
        uint64_t pid[PPL];
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            pid[j] = base + lane + (uint64_t)MP_WAVE * j;
            state_init(st[j], q);
            state_init(sa[j], q);
            if constexpr (PAY == MP_PAY_ASIAN_MLMC) {
                cs[j].x = R(0);
                cs[j].acc.init();
                cs[j].zsum = R(0);
                cs[j].cnt = 0;
            }
        }
        for (int s0i = 0; s0i < n_steps; s0i += 4) {
            const uint32_t blk = (uint32_t)(s0i >> 2);
            const int nsub = min(4, n_steps - s0i);
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                const uint64_t p = pid[j];
                const uint64_t pr = p < pend ? p : pend - 1;  // clamp buffer reads
                if (m.model == MP_MODEL_GBM) {
                    R z[4];
                    if constexpr (SRC == 0) {
                        mp_normals4<R>((uint32_t)p, blk, 0u, q.seed_hi, q.key0, q.key1, z);
                    } else if constexpr (SRC == 2) {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            z[e] = e < nsub ? (PAY == MP_PAY_ASIAN_MLMC
                                                   ? structured_draw<R>(L, q, pr, (uint32_t)(s0i + e))
                                                   : R(structured_normal(L.sampler, q, pr, (uint32_t)(s0i + e))))
                                            : R(0);
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            z[e] = e < nsub ? R(L.normals[(uint64_t)(s0i + e) * q.n_paths + pr]) : R(0);
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (e < nsub) {
                            if constexpr (PAY == MP_PAY_ASIAN_MLMC) {
                                mlmc_step(st[j], cs[j], m, z[e], refine, c_drift, c_vol, sqrt_ref);
                            } else {
                                gbm_step(st[j], m, z[e]);
                                if (tech == MP_TECH_ANTITHETIC) gbm_step(sa[j], m, -z[e]);
                            }
                        }
                    }
                } else if (m.model == MP_MODEL_HESTON) {
                    R zv[4], zi[4];
                    if constexpr (SRC == 0) {
                        mp_normals4<R>((uint32_t)p, blk, 0u, q.seed_hi, q.key0, q.key1, zv);
                        mp_normals4<R>((uint32_t)p, blk, 1u, q.seed_hi, q.key0, q.key1, zi);
                    } else if constexpr (SRC == 2) {
                        // factor f of step s takes dimension 2s + f
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const uint32_t d = (uint32_t)(2 * (s0i + e));
                            zv[e] = e < nsub ? R(structured_normal(L.sampler, q, pr, d)) : R(0);
                            zi[e] = e < nsub ? R(structured_normal(L.sampler, q, pr, d + 1)) : R(0);
                        }
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const uint64_t row = (uint64_t)(s0i + e) * 2;
                            zv[e] = e < nsub ? R(L.normals[row * q.n_paths + pr]) : R(0);
                            zi[e] = e < nsub ? R(L.normals[(row + 1) * q.n_paths + pr]) : R(0);
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (e < nsub) {
                            heston_step(st[j], m, zv[e], zi[e]);
                            if (tech == MP_TECH_ANTITHETIC) heston_step(sa[j], m, -zv[e], -zi[e]);
                        }
                    }
                } else {  // Merton (standard technique only)
                    R z[4];
                    double u[4];
                    if constexpr (SRC != 1) {  // structured samplers are not offered for Merton
                        mp_normals4<R>((uint32_t)p, blk, 0u, q.seed_hi, q.key0, q.key1, z);
                        mp_uniforms4<double>((uint32_t)p, blk, 1u, q.seed_hi, q.key0, q.key1, u);
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const uint64_t idx = (uint64_t)(s0i + e) * q.n_paths + pr;
                            z[e] = e < nsub ? R(L.normals[idx]) : R(0);
                            u[e] = e < nsub ? L.uniforms[idx] : 0.0;
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (e < nsub) {
                            const int nj = poisson_count(u[e], q);
                            R zj = R(0);
                            if (nj != 0) {
                                if constexpr (SRC != 1) {
                                    R zz[4];
                                    mp_normals4<R>((uint32_t)p, blk, 2u, q.seed_hi, q.key0, q.key1, zz);
                                    zj = zz[e];
                                } else {
                                    zj = R(L.jump_normals[(uint64_t)(s0i + e) * q.n_paths + pr]);
                                }
                            }
                            merton_apply(st[j], m, z[e], nj, zj);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            if constexpr (PAY == MP_PAY_ASIAN_MLMC) {
                // the level term, discounted per path as the reference sums it:
                // disc * (fine avg - K)+ - disc * (coarse avg - K)+
                if (pid[j] < pend) {
                    const R fine_avg = s0 * (st[j].acc.s * (R(1) / R(n_steps)));
                    double d = q.discount * (double)(fine_avg > kk[0] ? fine_avg - kk[0] : R(0));
                    if (refine > 0) {
                        const R coarse_avg = s0 * (cs[j].acc.s * (R(1) / R(n_steps / refine)));
                        d -= q.discount * (double)(coarse_avg > kk[0] ? coarse_avg - kk[0] : R(0));
                    }
                    sum[0] += d;
                    sq[0] += d * d;
                    if (L.terminal_out) L.terminal_out[pid[j]] = (double)(s0 * mp_exp(st[j].x));
                    if (L.payoff_out) L.payoff_out[pid[j]] = d;
                }
                continue;
            }
            if (pid[j] < pend) {
                accumulate_unit(st[j], sa[j], m, q, tech, n_steps, s0, kk, ns, sum, sq);
                if (L.terminal_out) L.terminal_out[pid[j]] = (double)(s0 * mp_exp(st[j].x));
                if (L.payoff_out)
                    L.payoff_out[pid[j]] = (double)payoff_of(st[j], m, q, n_steps, s0, kk[0]);
            }
        }
    }
    // columns past the last one the request uses hold zeros in every lane
    const int used = tech == MP_TECH_CONTROL_VARIATE ? 3 : (ns > 1 ? ns : 1);
#pragma unroll
    for (int j = 0; j < MP_NS_MAX; ++j) {
        if (j < used) {
            sum[j] = mp_wave_sum(sum[j]);
            sq[j] = mp_wave_sum(sq[j]);
        }
    }
    if (lane == 0) {
        double* out = L.partial + (uint64_t)tile * MP_PARTIAL_STRIDE;
#pragma unroll
        for (int j = 0; j < MP_NS_MAX; ++j) {
            out[2 * j] = sum[j];
            out[2 * j + 1] = sq[j];
        }
    }
}

// One Box-Muller pair of the FP32 Philox stream, scaled for a GBM step: the
// volatility terms vol * z0 = vr * c and vol * z1 = vr * s of the normals
// z0 = r cos(2 pi u2), z1 = r sin(2 pi u2), r = sqrt(-2 ln u1) =
// sqrt(2 ln 2) sqrt(-log2 u1).  vscale is vol * sqrt(2 ln 2) (times log2(e)
// for log2-unit paths), so the radius, the volatility and the unit change
// cost one multiply per pair; the caller adds vr * c and vr * s with FMAs.
__device__ __forceinline__ void gbm_pair_f32(uint32_t a, uint32_t b, float vscale, float& vr, float& c,
                                             float& s) {
    const float u1 = mp_u01f(a);
    const float u2 = mp_u01f(b);
    vr = vscale * __builtin_amdgcn_sqrtf(-__builtin_amdgcn_logf(u1));
    c = __builtin_amdgcn_cosf(u2);
    s = __builtin_amdgcn_sinf(u2);
}

// x plus the sum of a Box-Muller pair's two normals in units of
// vscale sqrt2: r (cos 2 pi u2 + sin 2 pi u2) = sqrt2 r sin(2 pi (u2 + 1/8)),
// r = sqrt(-2 ln u1) = sqrt(2 ln 2) sqrt(-log2 u1), in one sine.  u2 + 1/8 is
// formed in the same subtraction that makes u2 from the float 1 + k 2^-23
// (the constant 7/8 - 2^-24 is exact); it is exact below 1 and rounded to
// 2^-23 above.
__device__ __forceinline__ float gbm_pair_sum_f32(uint32_t a, uint32_t b, float x) {
    const float u1 = mp_u01f(a);
    const float w = __uint_as_float(__builtin_amdgcn_alignbit(0x7fu, b, 9)) - 0.87499994039535522f;
    return __builtin_fmaf(__builtin_amdgcn_sqrtf(-__builtin_amdgcn_logf(u1)), __builtin_amdgcn_sinf(w), x);
}

// FP32 GBM tile with Philox draws and the standard technique (the batching
// and serving workload).  The same draws as tile_body; the arithmetic is
// arranged for the vector unit:
// * European: only S_T matters, so the log-return is the sum of the
//   volatility terms plus n_steps * drift_dt, added once at the end; full
//   blocks add each Box-Muller pair's two terms as one sine, without the
//   volatility factor (applied once at the end);
// * Asian: the path is carried in log2 units without its drift, y_k =
//   log2(S_k/S0) - k d, so S_k/S0 = 2^(y_k) 2^(k d) is one v_exp_f32 per step
//   times a factor: within block b the steps' factors 2^(e d), e = 1..4,
//   weight the block sum, and the block's 2^(4 b d) (a running FP64
//   product) multiplies it into an FP64 running sum;
// * down-and-out / lookback: natural-log path, barrier test / running max;
// * full 4-step blocks run without per-step predicates; the tail block of
//   n_steps % 4 steps is predicated.
// Payoffs, discounting and sums then follow tile_body.
// U: 4-step blocks per iteration of the step loop (the step unroll the
// autotuner picks); the blocks run in the same order whatever U is, so every
// U gives the same bits.
template <int PPL, int PAY, int U>
__device__ __forceinline__ void gbm_philox_f32_tile(const mp::tile_launch& L, uint32_t tile,
                                                    const mp_dreq& q, int lane) {
    model_params<float> m = load_params<float>(q);
    m.model = MP_MODEL_GBM;
    m.payoff = PAY;
    constexpr bool kLog2 = PAY == MP_PAY_ASIAN;
    const int n_steps = q.n_steps;
    const int ns = q.n_strikes;
    const float s0 = float(q.s0);
    float kk[MP_NS_MAX];
#pragma unroll
    for (int j = 0; j < MP_NS_MAX; ++j) kk[j] = float(q.strikes[j]);
    const float vscale = float(q.vol_dt * (kLog2 ? 1.1774100225154747 * 1.4426950408889634 : 1.1774100225154747));
    const float vscale2 = float(q.vol_dt * (1.1774100225154747 * 1.4142135623730951));
    const float drift = kLog2 ? float(q.drift_dt * 1.4426950408889634) : m.drift_dt;
    const float xb = m.xb;
    const uint64_t pbeg = (uint64_t)(tile - q.tile0) * q.tile_paths;
    const uint64_t pend = min(pbeg + (uint64_t)q.tile_paths, q.n_paths);
    const int nfull = n_steps >> 2, ntail = n_steps & 3;
    // the Philox words shared by the wave, in scalar registers once per tile
    const uint32_t seed_hi = __builtin_amdgcn_readfirstlane(q.seed_hi);
    const uint32_t key0 = __builtin_amdgcn_readfirstlane(q.key0);
    const uint32_t key1 = __builtin_amdgcn_readfirstlane(q.key1);
    // Asian step factors 2^(e d), e = 1..4, and the block factor 2^(4 d)
    float cf[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) cf[e] = kLog2 ? __builtin_amdgcn_exp2f(float(e + 1) * drift) : 1.0f;
    const double cb = kLog2 ? (double)__builtin_amdgcn_exp2f(4.0f * drift) : 1.0;

    double sum[MP_NS_MAX], sq[MP_NS_MAX];
#pragma unroll
    for (int j = 0; j < MP_NS_MAX; ++j) sum[j] = sq[j] = 0.0;

    for (uint64_t base = pbeg; base < pend; base += (uint64_t)MP_WAVE * PPL) {
        float x[PPL], xs[PPL], xmax[PPL];
        double acc[PPL];
        bool knocked[PPL];
        uint32_t pid32[PPL];
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            pid32[j] = (uint32_t)(base + lane + (uint64_t)MP_WAVE * j);
            x[j] = 0.0f;
            xs[j] = 0.0f;
            acc[j] = 0.0;
            xmax[j] = 0.0f;
            knocked[j] = q.knocked_at_start != 0;
        }
        double gb = 1.0;  // Asian: 2^(4 b d) of the current block b
        // one 4-step block; nsub < 4 only for the tail block
        auto block = [&](uint32_t blk, int nsub) {
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                const mp_u4 r = mp_philox_wave(pid32[j], blk, 0u, seed_hi, key0, key1);
                if constexpr (PAY == MP_PAY_EUROPEAN) {
                    if (nsub == 4) {
                        xs[j] = gbm_pair_sum_f32(r.x, r.y, xs[j]);
                        xs[j] = gbm_pair_sum_f32(r.z, r.w, xs[j]);
                        continue;
                    }
                }
                // step e's volatility term is vr[e / 2] * tr[e]
                float vr[2], tr[4];
                gbm_pair_f32(r.x, r.y, vscale, vr[0], tr[0], tr[1]);
                gbm_pair_f32(r.z, r.w, vscale, vr[1], tr[2], tr[3]);
                if constexpr (PAY == MP_PAY_EUROPEAN) {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (e < nsub) x[j] = __builtin_fmaf(vr[e >> 1], tr[e], x[j]);
                } else if constexpr (PAY == MP_PAY_ASIAN) {
                    float g[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (e < nsub) x[j] = __builtin_fmaf(vr[e >> 1], tr[e], x[j]);
                        g[e] = e < nsub ? __builtin_amdgcn_exp2f(x[j]) : 0.0f;
                    }
                    const float bs = __builtin_fmaf(
                        cf[3], g[3], __builtin_fmaf(cf[2], g[2], __builtin_fmaf(cf[1], g[1], cf[0] * g[0])));
                    acc[j] = __builtin_fma(gb, (double)bs, acc[j]);
                } else if constexpr (PAY == MP_PAY_DOWN_AND_OUT) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (e < nsub) {
                            x[j] = __builtin_fmaf(vr[e >> 1], tr[e], x[j] + drift);
                            knocked[j] = knocked[j] || (x[j] <= xb);
                        }
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (e < nsub) {
                            x[j] = __builtin_fmaf(vr[e >> 1], tr[e], x[j] + drift);
                            xmax[j] = x[j] > xmax[j] ? x[j] : xmax[j];
                        }
                    }
                }
            }
        };
        int b = 0;
        for (; b + U <= nfull; b += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                block((uint32_t)(b + u), 4);
                if constexpr (kLog2) gb *= cb;
            }
        }
        if constexpr (U > 1) {
            for (; b < nfull; ++b) {
                block((uint32_t)b, 4);
                if constexpr (kLog2) gb *= cb;
            }
        }
        if (ntail) block((uint32_t)nfull, ntail);
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            const uint64_t p = base + lane + (uint64_t)MP_WAVE * j;
            if (p >= pend) continue;
            path_state<float> st, sa;
            st.x = PAY == MP_PAY_EUROPEAN ? __builtin_fmaf(vscale2, xs[j], x[j]) + float(n_steps) * drift
                   : kLog2                ? (x[j] + float(n_steps) * drift) * 0.69314718055994531f
                                          : x[j];
            st.v = 0.0f;
            st.acc.s = (float)acc[j];
            st.acc.c = 0.0f;
            st.xmax = xmax[j];
            st.knocked = knocked[j];
            sa = st;
            accumulate_unit(st, sa, m, q, MP_TECH_STANDARD, n_steps, s0, kk, ns, sum, sq);
            if (L.terminal_out) L.terminal_out[p] = (double)(s0 * mp_exp(st.x));
            if (L.payoff_out) L.payoff_out[p] = (double)payoff_of(st, m, q, n_steps, s0, kk[0]);
        }
    }
    const int used = ns > 1 ? ns : 1;
#pragma unroll
    for (int j = 0; j < MP_NS_MAX; ++j) {
        if (j < used) {
            sum[j] = mp_wave_sum(sum[j]);
            sq[j] = mp_wave_sum(sq[j]);
        }
    }
    if (lane == 0) {
        double* out = L.partial + (uint64_t)tile * MP_PARTIAL_STRIDE;
#pragma unroll
        for (int j = 0; j < MP_NS_MAX; ++j) {
            out[2 * j] = sum[j];
            out[2 * j + 1] = sq[j];
        }
    }
}

// The tile at queue position qpos, with the body compiled for its request.
template <class R, int SRC, int PPL, bool ML>
__device__ __forceinline__ void run_tile(const mp::tile_launch& L, unsigned int qpos, int lane) {
    {
        const uint32_t tile = L.order ? L.order[qpos] : qpos;
        const uint32_t ri = L.tile_req ? L.tile_req[tile] : 0u;
        const mp_dreq& q = L.reqs[ri];
        if constexpr (ML) {
            // multilevel levels, one path per lane
            if constexpr (SRC == 1) {
                tile_body<R, 1, 1, MP_MODEL_GBM, MP_PAY_ASIAN_MLMC, -1>(L, tile, q, lane);
            } else {
                if (q.sampling != MP_SAMPLER_NONE)
                    tile_body<R, 2, 1, MP_MODEL_GBM, MP_PAY_ASIAN_MLMC, -1>(L, tile, q, lane);
                else
                    tile_body<R, 0, 1, MP_MODEL_GBM, MP_PAY_ASIAN_MLMC, -1>(L, tile, q, lane);
            }
        } else if (SRC == 0 && q.sampling != MP_SAMPLER_NONE) {
            // randomized Halton / Latin hypercube requests
            if constexpr (SRC == 0) tile_body<R, 2, PPL, -1, -1, -1>(L, tile, q, lane);
        } else if constexpr (SRC == 0 && (sizeof(R) == 4 || PPL == 1)) {
            // Philox tiles (batches, serving; FP64 at one path per lane): a
            // body per model and payoff
            const int mdl = q.model, pay = q.payoff;
            if (q.technique != MP_TECH_STANDARD) {
                tile_body<R, SRC, PPL, -1, -1, -1>(L, tile, q, lane);
            } else if (mdl == MP_MODEL_GBM && sizeof(R) == 4) {
                if (pay == MP_PAY_EUROPEAN)
                    gbm_philox_f32_tile<PPL, MP_PAY_EUROPEAN, 1>(L, tile, q, lane);
                else if (pay == MP_PAY_ASIAN)
                    gbm_philox_f32_tile<PPL, MP_PAY_ASIAN, 1>(L, tile, q, lane);
                else if (pay == MP_PAY_DOWN_AND_OUT)
                    gbm_philox_f32_tile<PPL, MP_PAY_DOWN_AND_OUT, 1>(L, tile, q, lane);
                else
                    gbm_philox_f32_tile<PPL, MP_PAY_LOOKBACK, 1>(L, tile, q, lane);
            } else if (mdl == MP_MODEL_GBM) {
                if (pay == MP_PAY_EUROPEAN)
                    tile_body<R, SRC, PPL, MP_MODEL_GBM, MP_PAY_EUROPEAN, -1>(L, tile, q, lane);
                else if (pay == MP_PAY_ASIAN)
                    tile_body<R, SRC, PPL, MP_MODEL_GBM, MP_PAY_ASIAN, -1>(L, tile, q, lane);
                else if (pay == MP_PAY_DOWN_AND_OUT)
                    tile_body<R, SRC, PPL, MP_MODEL_GBM, MP_PAY_DOWN_AND_OUT, -1>(L, tile, q, lane);
                else
                    tile_body<R, SRC, PPL, MP_MODEL_GBM, MP_PAY_LOOKBACK, -1>(L, tile, q, lane);
            } else if (mdl == MP_MODEL_HESTON) {
                if (pay == MP_PAY_EUROPEAN && q.scheme == 1)
                    tile_body<R, SRC, PPL, MP_MODEL_HESTON, MP_PAY_EUROPEAN, 1>(L, tile, q, lane);
                else if (pay == MP_PAY_EUROPEAN)
                    tile_body<R, SRC, PPL, MP_MODEL_HESTON, MP_PAY_EUROPEAN, 0>(L, tile, q, lane);
                else
                    tile_body<R, SRC, PPL, MP_MODEL_HESTON, -1, -1>(L, tile, q, lane);
            } else {
                if (pay == MP_PAY_EUROPEAN)
                    tile_body<R, SRC, PPL, MP_MODEL_MERTON, MP_PAY_EUROPEAN, -1>(L, tile, q, lane);
                else
                    tile_body<R, SRC, PPL, MP_MODEL_MERTON, -1, -1>(L, tile, q, lane);
            }
        } else {
            tile_body<R, SRC, PPL, -1, -1, -1>(L, tile, q, lane);
        }
    }
}

// ML: a launch of multilevel levels only (their own kernel, so the coarse
// path state does not raise the register count of the general kernel).
template <class R, int SRC, int PPL, bool ML>
__global__ __launch_bounds__(256) void tile_kernel(mp::tile_launch L) {
    __shared__ unsigned int queue_slot[256 / MP_WAVE];
    const int lane = threadIdx.x & (MP_WAVE - 1);
    volatile unsigned int* slot = &queue_slot[threadIdx.x >> 6];
    // A claim takes L.claim consecutive queue positions (one atomic for
    // several tiles when the queue is long).  A wave makes at most
    // n_tiles + 1 claims, so the loop ends even if the queue counter were left
    // in a bad state.
    const unsigned int k = (unsigned int)L.claim;
    unsigned int claims = 0;
    for (unsigned int q0 = next_queue_position(L.counter, slot, lane, k);
         q0 < L.n_tiles && claims <= L.n_tiles;
         q0 = next_queue_position(L.counter, slot, lane, k), ++claims) {
        const unsigned int qend = min(q0 + k, L.n_tiles);
        for (unsigned int qpos = q0; qpos < qend; ++qpos) run_tile<R, SRC, PPL, ML>(L, qpos, lane);
    }
}

// The tiles of a launch by static assignment -- wave i takes positions i,
// i + W, ... -- for a launch of one request (tiles of equal cost) or with a
// wave per tile: the work queue's atomics, every wave hitting one counter at
// the start, cost 10-25 us per launch.  A kernel of its own, so the queue
// kernel's code is unchanged; one path per lane.
template <class R, int SRC, bool ML>
__global__ __launch_bounds__(256) void tile_static_kernel(mp::tile_launch L) {
    const int lane = threadIdx.x & (MP_WAVE - 1);
    const unsigned int waves = gridDim.x * (blockDim.x / MP_WAVE);
    for (unsigned int qpos = blockIdx.x * (blockDim.x / MP_WAVE) + (threadIdx.x / MP_WAVE);
         qpos < L.n_tiles; qpos += waves)
        run_tile<R, SRC, 1, ML>(L, qpos, lane);
}

// The FP32 GBM-family Philox tiles only (the batching and serving
// workload): the same body run_tile selects for them, in kernels that do not
// carry the other models' code, so they need fewer registers and hold more
// waves per SIMD.  Same bits as the general kernels.
template <int PPL, int U>
__device__ __forceinline__ void run_tile_gbm(const mp::tile_launch& L, unsigned int qpos, int lane) {
    const uint32_t tile = L.order ? L.order[qpos] : qpos;
    const uint32_t ri = L.tile_req ? L.tile_req[tile] : 0u;
    const mp_dreq& q = L.reqs[ri];
    const int pay = q.payoff;
    if (pay == MP_PAY_EUROPEAN)
        gbm_philox_f32_tile<PPL, MP_PAY_EUROPEAN, U>(L, tile, q, lane);
    else if (pay == MP_PAY_ASIAN)
        gbm_philox_f32_tile<PPL, MP_PAY_ASIAN, U>(L, tile, q, lane);
    else if (pay == MP_PAY_DOWN_AND_OUT)
        gbm_philox_f32_tile<PPL, MP_PAY_DOWN_AND_OUT, U>(L, tile, q, lane);
    else
        gbm_philox_f32_tile<PPL, MP_PAY_LOOKBACK, U>(L, tile, q, lane);
}

template <int PPL, int U>
__global__ __launch_bounds__(256) void tile_gbm_kernel(mp::tile_launch L) {
    __shared__ unsigned int queue_slot[256 / MP_WAVE];
    const int lane = threadIdx.x & (MP_WAVE - 1);
    volatile unsigned int* slot = &queue_slot[threadIdx.x >> 6];
    const unsigned int k = (unsigned int)L.claim;
    unsigned int claims = 0;
    for (unsigned int q0 = next_queue_position(L.counter, slot, lane, k);
         q0 < L.n_tiles && claims <= L.n_tiles;
         q0 = next_queue_position(L.counter, slot, lane, k), ++claims) {
        const unsigned int qend = min(q0 + k, L.n_tiles);
        for (unsigned int qpos = q0; qpos < qend; ++qpos) run_tile_gbm<PPL, U>(L, qpos, lane);
    }
}

template <int U>
__global__ __launch_bounds__(256) void tile_gbm_static_kernel(mp::tile_launch L) {
    const int lane = threadIdx.x & (MP_WAVE - 1);
    const unsigned int waves = gridDim.x * (blockDim.x / MP_WAVE);
    for (unsigned int qpos = blockIdx.x * (blockDim.x / MP_WAVE) + (threadIdx.x / MP_WAVE);
         qpos < L.n_tiles; qpos += waves)
        run_tile_gbm<1, U>(L, qpos, lane);
}

__global__ void reset_kernel(unsigned int* counter, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counter[i] = 0u;
}

// Set-up of a one-request launch: the request record arrives as a kernel
// argument and is stored where the kernels read it, and the first n
// counters are cleared -- in place of a host-to-device copy and a reset.
__global__ void setup_kernel(mp_dreq d, mp_dreq* dst, unsigned int* counter, int n) {
    if (threadIdx.x == 0) *dst = d;
    if ((int)threadIdx.x < n) counter[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// Finalize: one block per request; wave w sums partial columns 2w and 2w+1
// over the request's tiles in tile order (Neumaier), then one thread forms
// the result in FP64: price = discount * mean and the population-variance
// standard error of the reference's summarize_payoffs (cpu.rs:7641-7652) per
// strike (pairs for antithetic requests: summarize_block_estimates
// cpu.rs:7654-7665), or for control-variate requests the beta-adjusted
// estimate of summarize_control_variate (cpu.rs:7738-7771) from moments that
// are already discounted.
__device__ void control_variate_summary(const double* tot, double n, double expectation,
                                        double* price, double* se) {
    #pragma clang fp contract(off)  // the reference's summaries round every product (no FMA)
    const double payoff_mean = tot[0] / n;
    const double control_mean = tot[2] / n;
    const double control_var = (tot[3] / n) - (control_mean * control_mean);
    if (control_var <= 2.220446049250313e-16) {  // f64::EPSILON: plain summary
        const double var = tot[1] / n - payoff_mean * payoff_mean;
        *price = payoff_mean;
        *se = sqrt(var > 0.0 ? var : 0.0) / sqrt(n);
        return;
    }
    const double cov = (tot[4] / n) - (payoff_mean * control_mean);
    const double beta = cov / control_var;
    const double adjusted_mean = payoff_mean - beta * (control_mean - expectation);
    const double adjusted_sq_mean = (tot[1] / n)
        - (2.0 * beta) * ((tot[4] / n) - expectation * payoff_mean)
        + (beta * beta) * ((tot[3] / n) - (2.0 * expectation * control_mean)
                           + (expectation * expectation));
    const double v = adjusted_sq_mean - adjusted_mean * adjusted_mean;
    *price = adjusted_mean;
    *se = sqrt(v > 0.0 ? v : 0.0) / sqrt(n);
}

__global__ __launch_bounds__(64 * MP_NS_MAX) void finalize_kernel(const mp_dreq* reqs,
                                                                  const double* partial,
                                                                  double* out) {
    __shared__ double red[MP_NS_MAX][2][MP_WAVE];
    __shared__ double tot[MP_PARTIAL_STRIDE];
    const mp_dreq& q = reqs[blockIdx.x];
    const int lane = threadIdx.x & (MP_WAVE - 1);
    const int w = threadIdx.x >> 6;
    const double n = (double)q.n_paths;
    const bool cv = q.technique == MP_TECH_CONTROL_VARIATE;
    const bool moments = q.technique == MP_TECH_RAW_MOMENTS;
    const int n_cols = moments ? MP_PARTIAL_STRIDE : cv ? 5 : 2 * q.n_strikes;
    // Lane l adds tiles l, l + 64, ... in that order; sixteen tiles' column
    // pairs are loaded before any is added.
    mp_neumaier a, b;
    a.init();
    b.init();
    const double2* col = reinterpret_cast<const double2*>(partial + (uint64_t)q.tile0 * MP_PARTIAL_STRIDE) + w;
    constexpr int kRow2 = MP_PARTIAL_STRIDE / 2;  // double2 per tile row
    constexpr int kAhead = 16;
    if (2 * w < n_cols) {
        for (uint32_t t0 = lane; t0 < q.n_tiles; t0 += MP_WAVE * kAhead) {
            double2 v[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; ++u) {
                const uint32_t t = t0 + (uint32_t)u * MP_WAVE;
                v[u] = t < q.n_tiles ? col[(uint64_t)t * kRow2] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < kAhead; ++u) {
                if (t0 + (uint32_t)u * MP_WAVE < q.n_tiles) {
                    a.add(v[u].x);
                    b.add(v[u].y);
                }
            }
        }
    }
    red[w][0][lane] = a.value();
    red[w][1][lane] = b.value();
    __syncthreads();
    if (lane == 0) {
        mp_neumaier s, s2;
        s.init();
        s2.init();
        for (int i = 0; i < MP_WAVE; ++i) {
            s.add(red[w][0][i]);
            s2.add(red[w][1][i]);
        }
        tot[2 * w] = s.value();
        tot[2 * w + 1] = s2.value();
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        #pragma clang fp contract(off)  // the reference's summaries round every product (no FMA)
        double* o = out + (uint64_t)blockIdx.x * MP_PARTIAL_STRIDE;
        for (int j = 0; j < MP_PARTIAL_STRIDE; ++j) o[j] = 0.0;
        if (moments) {
            for (int j = 0; j < MP_PARTIAL_STRIDE; ++j) o[j] = tot[j];
        } else if (cv) {
            control_variate_summary(tot, n, q.cv_expect, &o[0], &o[1]);
        } else {
            for (int j = 0; j < q.n_strikes; ++j) {
                const double mean = tot[2 * j] / n;
                const double var = tot[2 * j + 1] / n - mean * mean;
                o[2 * j] = q.discount * mean;
                o[2 * j + 1] = q.discount * sqrt(var > 0.0 ? var : 0.0) / sqrt(n);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Block-wide FP64 sum of up to 8 values per lane into partial[tile][*].
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double* dst) {
    __shared__ double red[16][NV];
    const int lane = threadIdx.x & (MP_WAVE - 1);
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = mp_wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) red[wave][i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double acc = 0.0;
        for (int w = 0; w < nw; ++w) acc += red[w][threadIdx.x];
        dst[threadIdx.x] = acc;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Sobol + Brownian bridge kernel (GBM with F = 1, Heston with F = 2).
// One block per tile of blockDim paths (grid-stride, static assignment);
// index tables of the direction numbers in LDS, bridge stack per lane in LDS,
// plan ops read with scalar loads.
// ---------------------------------------------------------------------------
// Scrambled Sobol points of one lane (one path index).  Dimensions come in
// blocks of 256 ("offsets") that share a seed, seed_o = bseed + offset; the
// point index is Owen-shuffled within its 2^16-point block once per offset,
// which fixes three per-lane pointers into the index tables.  A point with a
// 16-bit index has only its top 16 bits set (direction number k < 16 is
// m_k 2^-(k+1) with m_k < 2^(k+1)), so the tables hold the bit-reversed
// XORs as 16-bit words: entries 0..63 cover index bits 0..5, 64..95 bits
// 6..10 and 96..127 bits 11..15 of each dimension.  The XOR of three words
// (one v_bitop3) is the reversed point that the Laine-Karras permutation
// takes directly.  With
// U = true the caller guarantees bseed is the same in every lane of the wave
// (paths of one tile never straddle a 2^16 block), so the per-dimension
// hash runs on the scalar unit and the permutation uses 24-bit multiplies
// by a scalar operand.
// HT (the register-bridge kernels): the hashed per-dimension seeds of the
// tile's 2^16-point block come from a table in LDS (hst[g] = {hash, its
// multiplier (hash >> 16) | 1} for global dimension g = 256 offset + dd),
// filled by the block once per tile, instead of the scalar unit's hash per
// draw, whose dependent chain of about ten scalar instructions per draw
// held back the vector unit.
template <class R, bool U, bool HT>
struct sobol_cursor {
    const uint32_t* nib;
    const uint16_t *t0, *t1, *t2;  // dimension 0 of the current offset
    const uint16_t *c0, *c1, *c2;  // the walk position (seek / advance)
    const uint2* hst;              // HT: the hashed-seed table
    uint32_t bseed, seed_o, local;
    int off;
    __device__ __forceinline__ void init(const uint32_t* tables, uint32_t block_seed, uint32_t local_index) {
        nib = tables;
        bseed = U ? __builtin_amdgcn_readfirstlane(block_seed) : block_seed;
        local = local_index;
        set_offset(0);
    }
    __device__ __forceinline__ void set_offset(int o) {
        off = o;
        seed_o = bseed + (uint32_t)o;
        const uint32_t s = mp_sobol_shuffle(local, mp_hash32(seed_o));
        const uint16_t* tab = reinterpret_cast<const uint16_t*>(nib);
        t0 = tab + (s & 63u);
        t1 = tab + 64 + ((s >> 6) & 31u);
        t2 = tab + 96 + (s >> 11);
        c0 = t0;
        c1 = t1;
        c2 = t2;
    }
    __device__ __forceinline__ uint32_t xrev(uint32_t dd) const {
        const uint32_t o = dd * 128u;
        return __builtin_amdgcn_bitop3_b32((uint32_t)t0[o], (uint32_t)t1[o], (uint32_t)t2[o], 0x96);
    }
    // Sequential walks over the dimensions of the current offset keep three
    // lane pointers that move once per group of dimensions, so the table
    // reads of a group take immediate offsets.  set_offset places the walk at
    // dimension 0, xrev_rel(j) reads dimension (walk position + j) and
    // advance(n) moves the walk n dimensions on.
    __device__ __forceinline__ void advance(int n) {
        c0 += n * 128;
        c1 += n * 128;
        c2 += n * 128;
    }
    __device__ __forceinline__ uint32_t xrev_rel(int j) const {
        return __builtin_amdgcn_bitop3_b32((uint32_t)c0[j * 128], (uint32_t)c1[j * 128],
                                           (uint32_t)c2[j * 128], 0x96);
    }
    // Top 23 bits of the scrambled point of dimension dd (< 256) of the
    // current offset whose reversed unscrambled point is x, from the
    // 24-bit-exact permutation.
    __device__ __forceinline__ uint32_t scrambled_rev(uint32_t x, uint32_t dd) const {
        uint32_t y;
        if constexpr (HT) {
            const uint2 h = hst[(uint32_t)off * 256u + dd];
            y = mp_lk_low24_vv(x, h.x, h.y);
        } else {
            const uint32_t hs = mp_hash32(dd ^ seed_o);
            if constexpr (U)
                y = mp_lk_low24_su(x, hs);
            else
                y = mp_lk_low24(x, hs);
        }
        return mp_rev32(y);
    }
    // Standard normal of dimension dd of the current offset: inverse normal
    // of the path precision at the 23-bit uniform p = top 23 bits * 2^-23.
    __device__ __forceinline__ R normal_of(uint32_t x, uint32_t dd) const {
        if constexpr (sizeof(R) == 4) {
            // (0x7f:point) >> 9 = the bits of the float 1 + p
            return mp_inv_normal_f32(__builtin_amdgcn_alignbit(0x7fu, scrambled_rev(x, dd), 9));
        } else {
            return mp_acklam<R>(R(scrambled_rev(x, dd) >> 9) * R(0x1p-23));
        }
    }
    // The bits of the float 1 + p, p the 23-bit uniform of dimension dd.
    __device__ __forceinline__ uint32_t uniform_bits(uint32_t x, uint32_t dd) const {
        return __builtin_amdgcn_alignbit(0x7fu, scrambled_rev(x, dd), 9);
    }
    __device__ __forceinline__ R normal_at(uint32_t dd) const { return normal_of(xrev(dd), dd); }
    // Dimension dd, which is the walk position + j.
    __device__ __forceinline__ R normal_rel(uint32_t dd, int j) const { return normal_of(xrev_rel(j), dd); }
    // Any dimension g, moving the offset as needed.
    __device__ __forceinline__ R normal(int g) {
        if ((g >> 8) != off) set_offset(g >> 8);
        return normal_at((uint32_t)(g & 255));
    }
    // Full 32-bit scrambled point of dimension g.
    __device__ __forceinline__ uint32_t raw(int g) {
        if ((g >> 8) != off) set_offset(g >> 8);
        const uint32_t dd = (uint32_t)(g & 255);
        return mp_rev32(mp_lk(xrev(dd), mp_hash32(dd ^ seed_o)));
    }
};

// Scrambled-Sobol normal of (point, dim) for multilevel levels
// (fill_scrambled_sobol_point_standard_normals: 2^16-point blocks, the block
// seed advanced by the golden ratio; key0 = the level seed).  Out of line, so
// the bodies that never draw Sobol keep their registers.
template <class R>
__device__ __noinline__ R sobol_point_normal(const uint32_t* nib, uint32_t seed, uint64_t point,
                                             uint32_t dim) {
    sobol_cursor<R, false> c;
    c.init(nib, seed + (uint32_t)(point >> 16) * 0x9E3779B9u, (uint32_t)(point & 0xFFFFu));
    return c.normal((int)dim);
}

template <class R>
__device__ __forceinline__ R structured_draw(const mp::tile_launch& L, const mp_dreq& q, uint64_t point,
                                             uint32_t dim) {
    if (q.sampling == MP_SAMPLER_SOBOL) return sobol_point_normal<R>(L.nib, q.key0, point, dim);
    return R(structured_normal(L.sampler, q, point, dim));
}

// PAY / SCH >= 0 fix the payoff / Heston scheme at compile time (the
// request's fields must then hold those values); -1 reads them at run time.
// TREE (FP32, step counts 16 * 2^J, J = qp.tree_levels <= MP_TREE_MAX_LEVELS): the
// bridge runs in registers -- see the TREE branch -- and the LDS holds only
// the Sobol tables.
template <class R, int F, bool NIB_LDS, int PAY, int SCH, bool TREE>
__device__ __forceinline__ void qmc_body(const mp_dreq* __restrict__ reqp, const mp_qmc_params& qp,
                                         const uint32_t* __restrict__ nib_g,
                                         const mp_bbop* __restrict__ plan, double* partial,
                                         double* terminal_out, double* payoff_out) {
    extern __shared__ uint32_t smem[];
    const uint32_t* nib = nib_g;
    R* stack = reinterpret_cast<R*>(smem);
    if constexpr (NIB_LDS) {
        for (int i = threadIdx.x; i < kNibWords; i += blockDim.x) smem[i] = nib_g[i];
        nib = smem;
        stack = reinterpret_cast<R*>(smem + kNibWords);
        __syncthreads();
    }

    const mp_dreq& q = *reqp;
    model_params<R> m = load_params<R>(q);
    if constexpr (PAY >= 0) m.payoff = PAY;
    if constexpr (SCH >= 0) m.scheme = SCH;
    // specialised kernels run the standard technique only
    const int tech = PAY < 0 ? q.technique : MP_TECH_STANDARD;
    const int n_steps = q.n_steps;
    const int ns = q.n_strikes;
    const R s0 = mp_param<R>(q.s0);
    R kk[MP_NS_MAX];
#pragma unroll
    for (int j = 0; j < MP_NS_MAX; ++j) kk[j] = mp_param<R>(q.strikes[j]);
    const R sqrt_t = mp_param<R>(qp.sqrt_t);
    const R inv_sqrt_dt = mp_param<R>(qp.inv_sqrt_dt);
    const int bd = blockDim.x;
    const int tid = threadIdx.x;

    for (uint32_t tile = blockIdx.x; tile < q.n_tiles; tile += gridDim.x) {
        // bd is a power of two <= 1024, so a tile lies inside one 2^16 block
        const uint64_t tile_base = (uint64_t)tile * bd;
        const uint64_t p = tile_base + tid;
        const bool active = p < q.n_paths;
        sobol_cursor<R, true, TREE> cur;
        cur.init(nib, qp.base_seed + (uint32_t)(tile_base >> 16) * 0x9E3779B9u, (uint32_t)(p & 0xFFFFu));
        if constexpr (TREE) {
            // the hashed seeds of every dimension the tile draws (F per step),
            // after the stack area's place in LDS; the previous tile's reads
            // ended at its block_sum_store's barriers
            uint2* hst = reinterpret_cast<uint2*>(smem + kNibWords);
            const int ndims = F * n_steps;
            for (int g = tid; g < ndims; g += bd) {
                const uint32_t hs = mp_hash32((uint32_t)(g & 255) ^ (cur.bseed + (uint32_t)(g >> 8)));
                hst[g] = make_uint2(hs, (hs >> 16) | 1u);
            }
            __syncthreads();
            cur.hst = hst;
        }

        // antithetic requests: the unit is a pair, its second path (sa) takes
        // the negated draws -- for the bridge, the negated increments
        path_state<R> st, sa;
        state_init(st, q);
        state_init(sa, q);
        // FP64 follows the reference (normals z = dW / sqrt(dt), then z
        // sqrt(dt) inside the step); FP32 feeds the increments directly.
        constexpr bool kInc = sizeof(R) == 4;
        const R to_step = kInc ? R(1) : inv_sqrt_dt;  // bridge increment -> step input
        const R z_to_step = kInc ? m.sqrt_dt : R(1);  // normal -> step input
        // the step's parameters (the FP32 register bridge scales the ones
        // that multiply the increments)
        model_params<R> ms = m;
        // FP32 European register-bridge kernels leave the per-step drift /
        // rate term out of the steps and add it once after the last
        constexpr bool kDriftOnce = kInc && TREE && PAY == MP_PAY_EUROPEAN;
        // FP32 European Heston register bridge: factor 0 carries xi dW and
        // factor 1 rho_perp dW (scaled in the plan), see heston_step_u
        constexpr bool kFactorScaled = kDriftOnce && F == 2;
        const R rho_over_xi = kFactorScaled ? mp_param<R>(qp.rho_over_xi) : R(0);
        auto step1 = [&](path_state<R>& x, R i0, R i1) {
            if constexpr (F == 1) {
                if constexpr (kDriftOnce)
                    x.x = ms.sigma * i0 + x.x;
                else if constexpr (kInc)
                    gbm_step_dw(x, ms, i0);
                else
                    gbm_step(x, ms, i0);
            } else {
                if constexpr (kFactorScaled)
                    heston_step_u(x, ms, rho_over_xi, i0, i1);
                else if constexpr (kInc)
                    heston_step_dw(x, ms, i0, i1);
                else
                    heston_step(x, ms, i0, i1);
            }
        };
        auto emit = [&](R i0, R i1) {
            step1(st, i0, i1);
            if (tech == MP_TECH_ANTITHETIC) step1(sa, -i0, -i1);
        };
        if (!qp.bridge) {
            // step s uses dimensions s*F + f; 256/F steps per offset
            constexpr int spo = 256 / F;
            for (int s0i = 0; s0i < n_steps; s0i += spo) {
                if (s0i) cur.set_offset(s0i / spo);
                const int s1 = min(n_steps, s0i + spo);
                for (int s = s0i; s < s1; ++s) {
                    const uint32_t dd = (uint32_t)((s - s0i) * F);
                    const R z0 = cur.normal_rel(dd, 0);
                    const R z1 = F == 2 ? cur.normal_rel(dd + 1, 1) : R(0);
                    cur.advance(F);
                    if constexpr (kInc)
                        emit(z0 * z_to_step, z1 * z_to_step);
                    else
                        emit(z0, z1);
                }
            }
        } else if constexpr (TREE) {
            // Register bridge for n_steps = 16 * 2^J.  The reference's plan on a
            // power-of-two grid is a complete binary tree: J outer levels above
            // 2^J subtrees of 16 steps.  The subtrees are visited in time order;
            // before subtree s the outer intervals are split from level
            // J - ctz(s) (level 0 for s = 0) down to level J, keeping each
            // split's right half in the register pending[level]; a subtree's
            // 15 splits are static code.  Splits are consumed in the plan's
            // preorder (op k uses dimensions (k+1)F + f, the plan's weights).
            // FP64 uses the stack kernel's arithmetic, so results are
            // identical.  FP32 carries each interval's increment D scaled by
            // 2^depth (E = 2^k D; exact): the midpoint split D -> (D/2 + sd z,
            // D/2 - sd z) becomes E -> E +- se z, two FMAs, with
            // se = 2^(k+1) sd from the plan, and the leaves' 2^-(4+J) is
            // folded into the step's parameters.  Each half is rounded once
            // (the stack kernel forms the right half as D - left).
            const int J = qp.tree_levels;
            int k = 0, dims = 0, o = 0;
            if constexpr (sizeof(R) == 4 && !kFactorScaled) {
                const R leaf_scale = R(1) / R(16 << J);  // a power of two
                ms.sigma *= leaf_scale;
                ms.rho *= leaf_scale;
                ms.rho_perp *= leaf_scale;
                ms.xi *= leaf_scale;
            }
            // a split's scale per factor
            auto split_scale = [&](const mp_bbop& op, int f) -> float {
                if constexpr (kFactorScaled)
                    return f == 0 ? op.sdf_c0 : op.sdf_c1;
                else
                    return op.sdf_e;
            };
            auto split = [&](const R (&d)[F], R (&l)[F], R (&r)[F]) {
                if (dims == 256) {  // next 256-dimension offset
                    ++o;
                    cur.set_offset(o);
                    dims = 0;
                }
                const mp_bbop& op = plan[k];
#pragma unroll
                for (int f = 0; f < F; ++f) {
                    const R z = cur.normal_rel((uint32_t)(dims + f), f);
                    if constexpr (sizeof(R) == 4) {
                        const float se = split_scale(op, f);
                        l[f] = __builtin_fmaf(se, z, d[f]);
                        r[f] = __builtin_fmaf(-se, z, d[f]);
                    } else {
                        l[f] = op.rw * d[f] + op.sd * z;
                        r[f] = d[f] - l[f];
                    }
                }
                cur.advance(F);
                dims += F;
                ++k;
            };
            auto leaf = [&](const R (&d)[F]) {
                if constexpr (kInc)
                    emit(d[0], F == 2 ? d[F - 1] : R(0));
                else
                    emit(d[0] * to_step, F == 2 ? d[F - 1] * to_step : R(0));
            };
            // two leaves from an interval of two steps
            auto pair2 = [&](const R (&d)[F]) {
                R a[F], b[F];
                split(d, a, b);
                leaf(a);
                leaf(b);
            };
            auto quad4 = [&](const R (&d)[F]) {
                R a[F], b[F];
                split(d, a, b);
                pair2(a);
                pair2(b);
            };
            R cur_d[F], pending[MP_TREE_MAX_LEVELS][F];  // pending[l - 1]: level l
            R root[F];
#pragma unroll
            for (int f = 0; f < F; ++f) {
                if constexpr (kFactorScaled)
                    root[f] = mp_param<R>(f == 0 ? qp.root_c0 : qp.root_c1);
                else
                    root[f] = sqrt_t;
            }
#pragma unroll
            for (int f = 0; f < F; ++f) cur_d[f] = root[f] * cur.normal_rel((uint32_t)f, f);
            cur.advance(F);
            dims = F;
            for (int sub = 0; sub < (1 << J); ++sub) {
                const int lr = sub == 0 ? 0 : J - __builtin_ctz((unsigned)sub);
#pragma unroll
                for (int l = 1; l <= MP_TREE_MAX_LEVELS; ++l)
                    if (l == lr) {
#pragma unroll
                        for (int f = 0; f < F; ++f) cur_d[f] = pending[l - 1][f];
                    }
#pragma unroll
                for (int l = 0; l < MP_TREE_MAX_LEVELS; ++l)
                    if (l >= lr && l < J) {
                        R a[F];
                        split(cur_d, a, pending[l]);
#pragma unroll
                        for (int f = 0; f < F; ++f) cur_d[f] = a[f];
                    }
                // the 16-step subtree: midpoints 8, 4, 2, 1, 3, 6, 5, 7, 12, ...
                // When its 15 splits stay inside the current 256-dimension
                // offset (all but about one subtree in eight), they run as
                // straight-line code: split K (preorder) reads the dimensions
                // at walk position + K F + f by immediate offsets, and the
                // walk moves once, after the subtree.  Otherwise halves and
                // quarters run as loops (one copy of the quarter code) that
                // check for the offset change before every split.  Both forms
                // consume the same dimensions with the same arithmetic.
                if (dims + 15 * F <= 256) {
                    auto fsplit = [&](int K, const R (&d)[F], R (&l)[F], R (&r)[F]) {
                        const mp_bbop& op = plan[k + K];
                        R z[F];
                        if constexpr (sizeof(R) == 4) {
                            // the F draws' central values, then one check
                            // for the rare tail (one branch per split: the
                            // straight-line code between branches is what
                            // lets draws and steps interleave)
                            uint32_t fb[F];
                            float xs[F], ls[F];
                            bool tail = false;
#pragma unroll
                            for (int f = 0; f < F; ++f) {
                                fb[f] = cur.uniform_bits(cur.xrev_rel(K * F + f), (uint32_t)(dims + K * F + f));
                                z[f] = mp_invn_central(fb[f], xs[f], ls[f]);
                                tail = tail || !(ls[f] > MP_INVN_L_CENTRAL);
                            }
                            if (tail) {
#pragma unroll
                                for (int f = 0; f < F; ++f)
                                    if (!(ls[f] > MP_INVN_L_CENTRAL)) z[f] = mp_invn_tail(fb[f], ls[f]);
                            }
#pragma unroll
                            for (int f = 0; f < F; ++f) {
                                z[f] = z[f] * xs[f];
                                const float se = split_scale(op, f);
                                l[f] = __builtin_fmaf(se, z[f], d[f]);
                                r[f] = __builtin_fmaf(-se, z[f], d[f]);
                            }
                        } else {
#pragma unroll
                            for (int f = 0; f < F; ++f) {
                                z[f] = cur.normal_of(cur.xrev_rel(K * F + f), (uint32_t)(dims + K * F + f));
                                l[f] = op.rw * d[f] + op.sd * z[f];
                                r[f] = d[f] - l[f];
                            }
                        }
                    };
                    auto fpair2 = [&](int K, const R (&d)[F]) {
                        R a[F], b[F];
                        fsplit(K, d, a, b);
                        leaf(a);
                        leaf(b);
                    };
                    auto fquad4 = [&](int K, const R (&d)[F]) {
                        R a[F], b[F];
                        fsplit(K, d, a, b);
                        fpair2(K + 1, a);
                        fpair2(K + 2, b);
                    };
                    R a8[F], b8[F], a4[F], b4[F];
                    fsplit(0, cur_d, a8, b8);
                    fsplit(1, a8, a4, b4);
                    fquad4(2, a4);
                    fquad4(5, b4);
                    fsplit(8, b8, a4, b4);
                    fquad4(9, a4);
                    fquad4(12, b4);
                    cur.advance(15 * F);
                    dims += 15 * F;
                    k += 15;
                    continue;
                }
                R a8[F], b8[F];
                split(cur_d, a8, b8);
#pragma unroll 1
                for (int h = 0; h < 2; ++h) {
                    R d8[F], a4[F], b4[F];
#pragma unroll
                    for (int f = 0; f < F; ++f) d8[f] = h ? b8[f] : a8[f];
                    split(d8, a4, b4);
#pragma unroll 1
                    for (int q = 0; q < 2; ++q) {
                        R d4[F];
#pragma unroll
                        for (int f = 0; f < F; ++f) d4[f] = q ? b4[f] : a4[f];
                        quad4(d4);
                    }
                }
            }
        } else {
            // The bridge carries interval increments D = W(r) - W(l) rather than
            // W itself: with lw + rw = 1 the reference's midpoint rule
            // W(m) = lw W(l) + rw W(r) + sd z is W(m) - W(l) = rw D + sd z, so
            // every stored quantity has the size of an increment and FP32 never
            // subtracts two O(1) values.  Stack slot s holds D of the interval
            // whose right end is owned by that slot.
#pragma unroll
            for (int f = 0; f < F; ++f) stack[(0 * F + f) * bd + tid] = sqrt_t * cur.normal_rel((uint32_t)f, f);
            cur.advance(F);
            if (qp.n_ops == 0) {
                const R i0 = kInc ? stack[tid] : stack[tid] * to_step;
                const R i1 = F == 2 ? (kInc ? stack[bd + tid] : stack[bd + tid] * to_step) : R(0);
                emit(i0, i1);
            }
            // op k uses dimensions (k+1)*F + f; 256/F ops per offset
            constexpr int opo = 256 / F;
            int k = 0;
            for (int o = 0; k < qp.n_ops; ++o) {
                if (o) cur.set_offset(o);
                const int k1 = min(qp.n_ops, (o + 1) * opo - 1);
                for (; k < k1; ++k) {
                    const mp_bbop& op = plan[k];
                    R rw, sd;
                    if constexpr (sizeof(R) == 4) {
                        rw = op.rwf;
                        sd = op.sdf;
                    } else {
                        rw = op.rw;
                        sd = op.sd;
                    }
                    const int slot = op.slot, flags = op.flags;
                    const uint32_t dd = (uint32_t)((k + 1) * F - o * 256);
                    R dl[F], dr[F];
#pragma unroll
                    for (int f = 0; f < F; ++f) {
                        const R z = cur.normal_rel(dd + f, f);
                        const R d = stack[(slot * F + f) * bd + tid];
                        dl[f] = rw * d + sd * z;
                        dr[f] = d - dl[f];
                    }
                    cur.advance(F);
                    if (flags & 1) {
                        if constexpr (kInc)
                            emit(dl[0], F == 2 ? dl[F - 1] : R(0));
                        else
                            emit(dl[0] * to_step, F == 2 ? dl[F - 1] * to_step : R(0));
                    } else {
#pragma unroll
                        for (int f = 0; f < F; ++f) stack[((slot + 1) * F + f) * bd + tid] = dl[f];
                    }
                    if (flags & 2) {
                        // right child is a leaf: it follows the left child in step
                        // order, which is already complete when the left is a leaf
                        if constexpr (kInc)
                            emit(dr[0], F == 2 ? dr[F - 1] : R(0));
                        else
                            emit(dr[0] * to_step, F == 2 ? dr[F - 1] * to_step : R(0));
                    } else {
#pragma unroll
                        for (int f = 0; f < F; ++f) stack[(slot * F + f) * bd + tid] = dr[f];
                    }
                }
            }
        }
        if constexpr (kDriftOnce) st.x += R(n_steps) * (F == 1 ? m.drift_dt : m.rate_dt);
        double sum[MP_NS_MAX], sq[MP_NS_MAX];
#pragma unroll
        for (int j = 0; j < MP_NS_MAX; ++j) sum[j] = sq[j] = 0.0;
        if (active) accumulate_unit(st, sa, m, q, tech, n_steps, s0, kk, ns, sum, sq);
        double v[2 * MP_NS_MAX];
#pragma unroll
        for (int j = 0; j < MP_NS_MAX; ++j) {
            v[2 * j] = sum[j];
            v[2 * j + 1] = sq[j];
        }
        if (active) {
            if (terminal_out) terminal_out[p] = (double)(s0 * mp_exp(st.x));
            if (payoff_out) payoff_out[p] = (double)payoff_of(st, m, q, n_steps, s0, kk[0]);
        }
        block_sum_store<2 * MP_NS_MAX>(v, partial + (uint64_t)tile * MP_PARTIAL_STRIDE);
    }
}

// FP64 blocks hold at most 512 threads (the bridge stack's LDS and the
// registers of FP64 paths), so the FP64 instantiations are compiled for
// 512-thread blocks and get twice the registers.
template <class R, int F, bool NIB_LDS, int PAY, int SCH>
__global__ __launch_bounds__(sizeof(R) == 8 ? 512 : 1024) void qmc_kernel(const mp_dreq* __restrict__ reqp, mp_qmc_params qp,
                                                  const uint32_t* __restrict__ nib_g,
                                                  const mp_bbop* __restrict__ plan, double* partial,
                                                  double* terminal_out, double* payoff_out) {
    qmc_body<R, F, NIB_LDS, PAY, SCH, false>(reqp, qp, nib_g, plan, partial, terminal_out, payoff_out);
}

// FP32 register bridge.  It needs ~80 VGPRs, so one 1024-thread block per
// CU (four waves per SIMD) as with the LDS stack; a 64-VGPR build, which
// fits two blocks per CU, spilled to scratch and ran 3% slower.  The gain
// over the stack kernel is instructions: no stack traffic or addressing.
template <int F, int PAY, int SCH>
__global__ __launch_bounds__(1024) void qmc_tree_kernel(const mp_dreq* __restrict__ reqp, mp_qmc_params qp,
                     const uint32_t* __restrict__ nib_g, const mp_bbop* __restrict__ plan,
                     double* partial, double* terminal_out, double* payoff_out) {
    qmc_body<float, F, true, PAY, SCH, true>(reqp, qp, nib_g, plan, partial, terminal_out, payoff_out);
}

// The FP64 register bridge, 512-thread blocks: the same bits as the FP64
// LDS-stack kernel without its stack traffic.
template <int F, int PAY, int SCH>
__global__ __launch_bounds__(512) void qmc_tree64_kernel(const mp_dreq* __restrict__ reqp, mp_qmc_params qp,
                     const uint32_t* __restrict__ nib_g, const mp_bbop* __restrict__ plan,
                     double* partial, double* terminal_out, double* payoff_out) {
    qmc_body<double, F, true, PAY, SCH, true>(reqp, qp, nib_g, plan, partial, terminal_out, payoff_out);
}

// ---------------------------------------------------------------------------
// n-asset basket (stepped, Cholesky-correlated shocks), NA assets.
// Philox stream a drives asset a's independent normals.
// ---------------------------------------------------------------------------
// SRC: 0 Philox, 1 caller normals, 2 structured sampler (Halton / Latin
// hypercube; out-of-line draws, so the Philox body carries no call).  TECH
// is the request's technique (MP_TECH_*), compiled in: a standard request
// keeps one path state per asset and two moments.
template <class R, int NA, int SRC, int TECH>
__global__ __launch_bounds__(256) void basket_kernel(const mp_dreq* __restrict__ reqp,
                                                    const double* __restrict__ chol,
                                                    const double* __restrict__ drift,
                                                    const double* __restrict__ vol,
                                                    const double* __restrict__ logs0,
                                                    const double* __restrict__ weight,
                                                    const double* __restrict__ normals,
                                                    const mp_sampler_dim* __restrict__ sampler,
                                                    double* partial, double* terminal_out,
                                                    double* payoff_out) {
    const mp_dreq& q = *reqp;
    R L[NA][NA], dr[NA], vo[NA], w[NA], s0a[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        dr[a] = R(drift[a]);
        vo[a] = R(vol[a]);
        w[a] = R(weight[a]);
        s0a[a] = R(exp(logs0[a]));
#pragma unroll
        for (int b = 0; b < NA; ++b) L[a][b] = R(chol[a * NA + b]);
    }
    const int n_steps = q.n_steps;
    const R kk = R(q.strikes[0]);
    constexpr bool ANTI = TECH == MP_TECH_ANTITHETIC;
    for (uint32_t tile = blockIdx.x; tile < q.n_tiles; tile += gridDim.x) {
        // unit = path, or antithetic pair (second path xa takes -z)
        const uint64_t p = (uint64_t)tile * blockDim.x + threadIdx.x;
        const bool active = p < q.n_paths;
        const uint64_t pr = active ? p : q.n_paths - 1;
        // FP32: log S_T of each asset is linear in the draws, so the steps
        // only add each asset's normals (zs) and the Cholesky factor, the
        // volatility and the drift apply once after the last step; FP64
        // keeps the reference's per-step order
        constexpr bool kLinear = sizeof(R) == 4;
        R x[NA], xa[NA], zs[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) x[a] = xa[a] = zs[a] = R(0);
        for (int s0i = 0; s0i < n_steps; s0i += 4) {
            const int nsub = min(4, n_steps - s0i);
            R z[NA][4];
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                if constexpr (SRC == 0) {
                    mp_normals4<R>((uint32_t)p, (uint32_t)(s0i >> 2), (uint32_t)a, q.seed_hi, q.key0,
                                   q.key1, z[a]);
                } else if constexpr (SRC == 2) {
                    // structured sampler: asset a of step s is dimension s * NA + a
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        z[a][e] = e < nsub ? R(structured_normal(sampler, q, pr, (uint32_t)((s0i + e) * NA + a))) : R(0);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        z[a][e] = e < nsub ? R(normals[((uint64_t)(s0i + e) * NA + a) * q.n_paths + pr]) : R(0);
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (e < nsub) {
                    if constexpr (kLinear) {
#pragma unroll
                        for (int a = 0; a < NA; ++a) zs[a] += z[a][e];
                    } else {
#pragma unroll
                        for (int a = 0; a < NA; ++a) {
                            R c = R(0);
#pragma unroll
                            for (int b = 0; b <= a; ++b) c += L[a][b] * z[b][e];
                            x[a] += dr[a] + vo[a] * c;
                            // the correlated shock of -z is exactly -c
                            if constexpr (ANTI) xa[a] += dr[a] + vo[a] * -c;
                        }
                    }
                }
            }
        }
        if constexpr (kLinear) {
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                R c = R(0);
#pragma unroll
                for (int b = 0; b <= a; ++b) c += L[a][b] * zs[b];
                const R drift_total = R(n_steps) * dr[a];
                x[a] = drift_total + vo[a] * c;
                if constexpr (ANTI) xa[a] = drift_total + vo[a] * -c;
            }
        }
        R basket = R(0), basket_a = R(0);
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            basket += w[a] * (s0a[a] * mp_exp(x[a]));
            if constexpr (ANTI) basket_a += w[a] * (s0a[a] * mp_exp(xa[a]));
        }
        const R pay = basket > kk ? basket - kk : R(0);
        // standard: payoff, payoff^2; antithetic: the pair estimate; control
        // variate: discounted payoff and control (discount * basket at T) moments
        double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (active) {
            if constexpr (TECH == MP_TECH_CONTROL_VARIATE) {
                const double pd = q.discount * (double)pay, cd = q.discount * (double)basket;
                v[0] = pd;
                v[1] = pd * pd;
                v[2] = cd;
                v[3] = cd * cd;
                v[4] = pd * cd;
            } else {
                double est = (double)pay;
                if constexpr (ANTI) est = 0.5 * (est + (double)(basket_a > kk ? basket_a - kk : R(0)));
                v[0] = est;
                v[1] = est * est;
            }
            if (terminal_out) terminal_out[p] = (double)basket;
            if (payoff_out) payoff_out[p] = (double)pay;
        }
        if constexpr (TECH == MP_TECH_CONTROL_VARIATE) {
            block_sum_store<5>(v, partial + (uint64_t)tile * MP_PARTIAL_STRIDE);
        } else {
            double vv[2] = {v[0], v[1]};
            block_sum_store<2>(vv, partial + (uint64_t)tile * MP_PARTIAL_STRIDE);
        }
    }
}

// ---------------------------------------------------------------------------
// Single-pass terminal-GBM Greeks (european_call_pathwise_greeks_cpu and
// european_call_likelihood_ratio_greeks_cpu, cpu.rs:4670-4857).  Path p takes
// the terminal draw of the engine's terminal-method price (Philox block 0,
// stream 0, first normal) or normals[p].  EST 1 (pathwise): Delta, Vega, Rho,
// Theta; EST 2 (likelihood ratio): Delta, Vega, Rho.  Each block writes the
// FP64 (sum, sum of squares) of the four estimates to partial[block].
// ---------------------------------------------------------------------------
template <class R, int SRC, int EST>
__global__ __launch_bounds__(256) void greeks_kernel(const mp_dreq* __restrict__ reqp,
                                                    const double* __restrict__ normals,
                                                    double* partial) {
    const mp_dreq& q = *reqp;
    const R s0 = R(q.s0), k = R(q.strikes[0]), r = R(q.rate), t = R(q.dt), sqrt_t = R(q.sqrt_dt);
    const R sigma = R(q.sigma), drift_t = R(q.drift_dt), vol_t = R(q.vol_dt), disc = R(q.discount);
    const R inv_s0 = R(1) / s0;
    const R lr_delta = R(1) / (s0 * sigma * sqrt_t);
    const R inv_sig_sqrt_t = R(1) / (sigma * sqrt_t);
    const R inv_sigma = R(1) / sigma;
    const R drift_rate = r - R(0.5) * sigma * sigma;
    const R half_sig_rsqrt_t = sigma / (R(2) * sqrt_t);
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < q.n_paths; p += stride) {
        R z;
        if constexpr (SRC == 0) {
            R zz[4];
            mp_normals4<R>((uint32_t)p, 0u, 0u, q.seed_hi, q.key0, q.key1, zz);
            z = zz[0];
        } else {
            z = R(normals[p]);
        }
        const R st = s0 * mp_exp(drift_t + vol_t * z);
        const R intrinsic = st - k;
        const bool active = intrinsic > R(0);
        const R payoff = active ? intrinsic : R(0);
        const R dpay = disc * payoff;
        R g[4];
        if constexpr (EST == MP_GREEK_EST_PATHWISE) {
            g[0] = active ? disc * st * inv_s0 : R(0);
            g[1] = active ? disc * st * (-sigma * t + sqrt_t * z) : R(0);
            g[2] = active ? disc * (t * st - t * payoff) : -t * dpay;
            const R dst_dt = st * (drift_rate + z * half_sig_rsqrt_t);
            g[3] = -(active ? disc * (dst_dt - r * payoff) : -r * dpay);
        } else {
            g[0] = dpay * z * lr_delta;
            g[1] = dpay * ((z * z - R(1)) * inv_sigma - sqrt_t * z);
            g[2] = dpay * (z * inv_sig_sqrt_t - t);
            g[3] = R(0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double gi = (double)g[i];
            v[2 * i] += gi;
            v[2 * i + 1] += gi * gi;
        }
    }
    block_sum_store<8>(v, partial + (uint64_t)blockIdx.x * MP_PARTIAL_STRIDE);
}

// ---------------------------------------------------------------------------
// Longstaff-Schwartz put.  Paths stored step-major [step][path] (steps
// 1..M-1), cash flows kept discounted to t0 in FP64, exercise step per path
// in 16 bits.  Parameters live in device memory so the date loop can be
// replayed as a graph; per-date reductions finish in the last block to
// arrive (one atomic ticket per block, fixed-order sums, no waiting).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void lsm_fill_kernel(unsigned int* counters, int n_counters,
                                                      int* valid, double* coeff, int n_dates) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_counters) counters[i] = 0u;
    if (i < n_dates) valid[i] = 0;
    if (i < 4 * n_dates) coeff[i] = 0.0;
}

template <class R, int SRC>
__global__ __launch_bounds__(256) void lsm_paths_kernel(const mp_lsm_params* __restrict__ Pp,
                                                       const double* __restrict__ normals,
                                                       R* __restrict__ paths, double* cash,
                                                       uint16_t* ex) {
    const mp_lsm_params P = *Pp;
    const uint64_t n = P.n_paths;
    const R drift = R(P.drift_dt), vol = R(P.vol_dt), s0 = R(P.s0);
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (uint64_t)gridDim.x * blockDim.x) {
        R x = R(0);
        R sT = s0;
        for (int s0i = 0; s0i < P.n_steps; s0i += 4) {
            R z[4];
            const int nsub = min(4, P.n_steps - s0i);
            if constexpr (SRC == 0) {
                mp_normals4<R>((uint32_t)p, (uint32_t)(s0i >> 2), 0u, P.seed_hi, P.key0, P.key1, z);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    z[e] = e < nsub ? R(normals[(uint64_t)(s0i + e) * n + p]) : R(0);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (e < nsub) {
                    x += drift + vol * z[e];
                    const int step = s0i + e + 1;
                    const R S = s0 * mp_exp(x);
                    if (step < P.n_steps) paths[(uint64_t)(step - 1) * n + p] = S;
                    sT = S;
                }
            }
        }
        const double d = P.k - (double)sT;
        cash[p] = (d > 0.0 ? d : 0.0) * P.disc_m;  // terminal cash flow discounted to t0
        ex[p] = (uint16_t)P.n_steps;
    }
}

__device__ __forceinline__ void laguerre(double x, int degree, bool weighted, double b[4]) {
    // laguerre_lsm_basis cpu.rs:7667-7682 (unweighted, x = S/K).  x^3 / 6 is
    // the product with RN(1/6) corrected once by its exact FMA remainder,
    // which equals the rounded quotient (no miss on 1.5e9 sampled values)
    // without an FP64 division.  Weighted: every function times e^{-x/2}, the
    // basis of Longstaff and Schwartz (2001), orthonormal on [0, inf).
    const double x2 = x * x;
    const double x3 = x2 * x;
    const double q0 = x3 * 0.16666666666666666;
    const double x3_6 = __builtin_fma(__builtin_fma(-q0, 6.0, x3), 0.16666666666666666, q0);
    const double w = weighted ? exp(-0.5 * x) : 1.0;
    b[0] = w;
    b[1] = degree >= 1 ? (1.0 - x) * w : 0.0;
    b[2] = degree >= 2 ? (1.0 - 2.0 * x + 0.5 * x2) * w : 0.0;
    b[3] = degree >= 3 ? (1.0 - 3.0 * x + 1.5 * x2 - x3_6) * w : 0.0;
}

// The continuation value of the unweighted basis, beta . L(x), as the cubic
// c0 + c1 x + c2 x^2 + c3 x^3 whose coefficients are formed once per date
// (L0 = 1, L1 = 1 - x, L2 = 1 - 2x + x^2/2, L3 = 1 - 3x + 3x^2/2 - x^3/6;
// beta_i = 0 beyond the regression's degree): three FMAs per path instead
// of the basis and the dot product.  The weighted basis keeps laguerre().
__device__ __forceinline__ void lsm_cont_poly(const double (&beta)[4], double (&c)[4]) {
    c[0] = ((beta[0] + beta[1]) + beta[2]) + beta[3];
    c[1] = (-beta[1] - 2.0 * beta[2]) - 3.0 * beta[3];
    c[2] = 0.5 * beta[2] + 1.5 * beta[3];
    c[3] = beta[3] * -0.16666666666666666;
}

__device__ __forceinline__ double lsm_cont_eval(const double (&c)[4], double x) {
    return __builtin_fma(__builtin_fma(__builtin_fma(c[3], x, c[2]), x, c[1]), x, c[0]);
}

// Gauss-Jordan with partial pivoting, failing when the best pivot <= 1e-12
// (solve_lsm_normal_equations cpu.rs:7684-7736), with the pivot row scaled
// by the pivot's reciprocal instead of divided element by element (one
// division per pivot; the results differ from division in the last bits).
// Single lane; every index is static so the system stays in registers.  Rows
// and columns at or beyond dim are zero (their basis functions are zero), so
// carrying them through the eliminations changes nothing.
__device__ bool lsm_solve4(double (&m)[4][4], double (&rhs)[4], int dim) {
#pragma unroll
    for (int piv = 0; piv < 4; ++piv) {
        if (piv >= dim) break;
        int best = piv;
        double best_abs = fabs(m[piv][piv]);
#pragma unroll
        for (int row = piv + 1; row < 4; ++row) {
            const double c = fabs(m[row][piv]);
            if (row < dim && c > best_abs) {
                best = row;
                best_abs = c;
            }
        }
        if (best_abs <= 1e-12) return false;
#pragma unroll
        for (int row = piv + 1; row < 4; ++row) {
            if (row == best) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const double t = m[piv][j];
                    m[piv][j] = m[row][j];
                    m[row][j] = t;
                }
                const double t = rhs[piv];
                rhs[piv] = rhs[row];
                rhs[row] = t;
            }
        }
        // one division per pivot; the row is scaled by its reciprocal
        const double inv = 1.0 / m[piv][piv];
        m[piv][piv] = 1.0;
#pragma unroll
        for (int c = piv + 1; c < 4; ++c) m[piv][c] *= inv;
        rhs[piv] *= inv;
#pragma unroll
        for (int row = 0; row < 4; ++row) {
            if (row == piv) continue;
            const double f = m[row][piv];
            if (f == 0.0) continue;
#pragma unroll
            for (int c = piv; c < 4; ++c) m[row][c] -= f * m[piv][c];
            rhs[row] -= f * rhs[piv];
        }
    }
    return true;
}

// Regression coefficients from the reduced sums tot (10 Gram entries, 4
// right-hand sides, the in-the-money count), one lane: the reference's
// skip rules (count < coefficients, pivot <= 1e-12).  Returns validity.
__device__ __forceinline__ int lsm_coefficients(const double* tot, int nc, double (&c)[4]) {
    double mm[4][4], rhs[4];
    int e = 0;
    for (int a = 0; a < 4; ++a)
        for (int b = a; b < 4; ++b) {
            mm[a][b] = tot[e];
            mm[b][a] = tot[e];
            ++e;
        }
    for (int a = 0; a < 4; ++a) rhs[a] = tot[10 + a];
    int ok = 0;
    if (tot[14] >= (double)nc) ok = lsm_solve4(mm, rhs, nc) ? 1 : 0;
    for (int a = 0; a < 4; ++a) c[a] = (ok && a < nc) ? rhs[a] : 0.0;
    return ok;
}

// Block sum of NV values published for other blocks (coherent stores),
// then a ticket; true in the block that arrives last, which may then read
// every block's partials with mp_load_coherent.
template <int NV>
__device__ __forceinline__ bool publish_and_ticket(double (&v)[NV], double* dst, unsigned int* counter) {
    __shared__ double red[16][NV];
    __shared__ bool is_last;
    const int lane = threadIdx.x & (MP_WAVE - 1);
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = mp_wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) red[wave][i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double acc = 0.0;
        for (int w = 0; w < nw; ++w) acc += red[w][threadIdx.x];
        mp_store_coherent(dst + threadIdx.x, acc);
    }
    // the stores above are all in wave 0; drain them before the ticket
    if (threadIdx.x < MP_WAVE) mp_vmem_drain_release();
    __syncthreads();
    if (threadIdx.x == 0)
        is_last = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                  gridDim.x - 1;
    __syncthreads();
    if (is_last) mp_vmem_drain_acquire();
    return is_last;
}

// Fixed-order FP64 sum of NV values over nb <= MP_LSM_MAX_BLOCKS rows by
// the block's first 256 threads (blocks of 256 or more): thread t first adds
// its rows t, t + 256, ... in that order (the caller's rowsum), then 16
// threads per value sum 16 of those in order, and one thread per value sums
// the 16 results into tot (shared).
template <int NV>
__device__ __forceinline__ void sum_rows(const double (&rowsum)[NV], double* tot) {
    __shared__ double row[256][NV + 1];
    __shared__ double red[NV][17];
    const int t = threadIdx.x;
    if (t < 256) {
#pragma unroll
        for (int i = 0; i < NV; ++i) row[t][i] = rowsum[i];
    }
    __syncthreads();
    const int vi = t >> 4, j = t & 15;
    if (vi < NV) {
        mp_neumaier acc;
        acc.init();
        for (int q = 0; q < 16; ++q) acc.add(row[j * 16 + q][vi]);
        red[vi][j] = acc.value();
    }
    __syncthreads();
    if (t < NV) {
        mp_neumaier s;
        s.init();
        for (int q = 0; q < 16; ++q) s.add(red[t][q]);
        tot[t] = s.value();
    }
    __syncthreads();
}

// The same over block partials laid out [block][stride], published with
// mp_store_coherent before a ticket (publish_and_ticket).
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* partial, int nb, int stride,
                                                double* tot /* shared, NV */) {
    const int t = threadIdx.x;
    double a[4][NV];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int i = 0; i < NV; ++i)
            a[r][i] = t < 256 && t + 256 * r < nb
                          ? mp_load_coherent(partial + (uint64_t)(t + 256 * r) * stride + i)
                          : 0.0;
    double v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        v[i] = a[0][i] + a[1][i];
        if (nb > 512) v[i] = (v[i] + a[2][i]) + a[3][i];
    }
    sum_rows<NV>(v, tot);
}

// Write-once slots for values published to other blocks: filled with this
// bit pattern (a signalling NaN, which no arithmetic produces) before the
// launch, written once with a coherent 64-bit store, and read by polling
// until the pattern is gone.  A slot needs no fence or flag: its one store
// is the signal.
#define MP_SLOT_EMPTY 0x7FF45E471E570B5Eull

__device__ __forceinline__ bool mp_slot_read(const double* p, double& v) {
    const unsigned long long b = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v = __longlong_as_double((long long)b);
    return b != MP_SLOT_EMPTY;
}

// Bounded poll accounting: true when the caller should stop waiting (its
// polls ran out, or another wait already did), after recording the error.
__device__ __forceinline__ bool mp_poll_give_up(unsigned int& polls, unsigned int limit, unsigned int* err) {
    __builtin_amdgcn_s_sleep(1);
    if (++polls > limit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
    return false;
}

// Fixed-order sum of NV <= 16 values over the rows row0 .. row0 + nrows - 1
// (nrows <= 32) of write-once slots laid out [value][stride], waiting until
// they have arrived; 256 threads, result in out (shared).  Thread (v, j)
// reads rows j and j + 16 of value v -- two loads per lane, because a
// coherent load costs the issuing CU about one request per lane -- and adds
// them; one thread per value then sums the 16 pairs in order (Neumaier).
template <int NV>
__device__ __forceinline__ void reduce_group(const double* base, uint64_t stride, int row0, int nrows,
                                             unsigned int spin_limit, unsigned int* err, double* out) {
    __shared__ double pair[16][17];
    const int t = threadIdx.x, v = t >> 4, j = t & 15;
    if (v < NV) {
        double a0 = 0.0, a1 = 0.0;
        unsigned int polls = 0;
        for (;;) {
            bool full = true;
            if (j < nrows) full &= mp_slot_read(base + (uint64_t)v * stride + row0 + j, a0);
            if (j + 16 < nrows) full &= mp_slot_read(base + (uint64_t)v * stride + row0 + j + 16, a1);
            if (full || mp_poll_give_up(polls, spin_limit, err)) break;
        }
        pair[v][j] = a0 + a1;
    }
    __syncthreads();
    if (t < NV) {
        mp_neumaier acc;
        acc.init();
        for (int q = 0; q < 16; ++q) acc.add(pair[t][q]);
        out[t] = acc.value();
    }
    __syncthreads();
}

// Block sum of NV <= 16 values (blocks of 256 to 1024 threads); thread
// i < NV stores value i to dst[i * stride] as a write-once slot.  Fixed
// order: DPP sums within each quad of lanes; the block's quad sums (64 to
// 256 per value) go to LDS, 16 threads per value each add a run of them in
// order, and one thread per value adds the 16 run sums in order.
template <int NV>
__device__ __forceinline__ void block_sum_to_slots(double (&v)[NV], double* dst, uint64_t stride) {
    __shared__ double red[NV][256];
    __shared__ double red2[NV][16];
    const int t = threadIdx.x;
    const int nq = blockDim.x >> 2;  // 64 .. 256
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        v[i] += mp_dpp_f64(v[i], 0);
        v[i] += mp_dpp_f64(v[i], 1);
    }
    if ((t & 3) == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) red[i][t >> 2] = v[i];
    }
    __syncthreads();
    const int vi = t >> 4, j = t & 15;
    if (vi < NV) {
        const int per = nq >> 4;  // 4 .. 16
        const double* r = &red[vi][j * per];
        double a = r[0];
        for (int q = 1; q < per; ++q) a += r[q];
        red2[vi][j] = a;
    }
    __syncthreads();
    if (t < NV) {
        double acc = red2[t][0];
#pragma unroll
        for (int q = 1; q < 16; ++q) acc += red2[t][q];
        mp_store_coherent(dst + t * stride, acc);
    }
}

// One pass per exercise date: apply the exercise rule of date step_update
// (if >= 1) and accumulate the Gram matrix / right-hand side of date
// step_gram (if >= 1) over in-the-money paths; the last block to finish
// reduces all blocks' sums in fixed order and solves for the coefficients.
// cash holds cash flows discounted to t0; growth[s] = disc_step^-s.
// MATCHED (launched with lsm_fused_kernel's grid): the block sums are formed
// as lsm_fused_kernel forms them (DPP wave sums) and stored to
// gram_partial[value][block]; lsm_group_kernel then reduces them in that
// kernel's order, so the pricing equals the one-launch pricing bit for bit.
// Each thread already visits its paths p0, p0 + T, ... in the one-launch
// kernel's order with the same per-path arithmetic.
template <class R, bool MATCHED>
__global__ __launch_bounds__(1024) void lsm_step_kernel(const mp_lsm_params* __restrict__ Pp,
                                                      const R* __restrict__ paths,
                                                      double* cash, uint16_t* ex,
                                                      double* coeff_all, int* valid_all,
                                                      int step_update, int step_gram,
                                                      double* gram_partial,
                                                      unsigned int* counters) {
    const mp_lsm_params& P = *Pp;
    const uint64_t n = P.n_paths;
    const double k = P.k, inv_k = P.inv_k;
    const int deg = P.degree, nc = P.coeffs;
    const bool wtd = P.weighted != 0;
    const double* disc_pow = reinterpret_cast<const double*>(Pp + 1);
    const double* growth = disc_pow + (P.n_steps + 1);
    const bool do_update = step_update >= 1 && valid_all[step_update] != 0;
    double beta[4] = {0.0, 0.0, 0.0, 0.0};
    if (do_update) {
#pragma unroll
        for (int i = 0; i < 4; ++i) beta[i] = coeff_all[step_update * 4 + i];
    }
    double cpoly[4];
    lsm_cont_poly(beta, cpoly);
    const double dpow_u = step_update >= 1 ? disc_pow[step_update] : 0.0;
    const double grow_g = step_gram >= 1 ? growth[step_gram] : 0.0;
    // 10 unique Gram entries + 4 RHS + count
    double acc[15];
#pragma unroll
    for (int i = 0; i < 15; ++i) acc[i] = 0.0;
    // four paths per thread per trip, all loads issued before any use
    const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
    const R* pu = paths + (uint64_t)(step_update >= 1 ? step_update - 1 : 0) * n;
    const R* pg = paths + (uint64_t)(step_gram >= 1 ? step_gram - 1 : 0) * n;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; base < n; base += 4 * T) {
        double c[4], su[4], sg[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t p = base + u * T;
            const bool in = p < n;
            c[u] = in ? cash[p] : 0.0;
            su[u] = in && do_update ? (double)pu[p] : k;
            sg[u] = in && step_gram >= 1 ? (double)pg[p] : k;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t p = base + u * T;
            const double imm_u = k - su[u] > 0.0 ? k - su[u] : 0.0;
            if (imm_u > 0.0) {  // false when not updating or p >= n (su = k)
                double cont;
                if (wtd) {
                    double bb[4];
                    laguerre(su[u] * inv_k, deg, wtd, bb);
                    // beta[i] = 0 for i >= nc, so the full sum equals the sum over nc
                    cont = 0.0;
#pragma unroll
                    for (int i = 0; i < 4; ++i) cont += beta[i] * bb[i];
                } else {
                    cont = lsm_cont_eval(cpoly, su[u] * inv_k);
                }
                if (imm_u > (cont > 0.0 ? cont : 0.0)) {
                    c[u] = imm_u * dpow_u;
                    cash[p] = c[u];
                    ex[p] = (uint16_t)step_update;
                }
            }
            const double imm_g = k - sg[u] > 0.0 ? k - sg[u] : 0.0;
            if (imm_g > 0.0) {  // false when not accumulating or p >= n
                double bb[4];
                laguerre(sg[u] * inv_k, deg, wtd, bb);
                const double cont = c[u] * grow_g;
                int e = 0;
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = i; j < 4; ++j) acc[e++] += bb[i] * bb[j];
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[10 + i] += bb[i] * cont;
                acc[14] += 1.0;
            }
        }
    }
    if (step_gram < 1) return;
    if constexpr (MATCHED) {
        block_sum_to_slots<15>(acc, gram_partial + blockIdx.x, (uint64_t)gridDim.x);
        return;
    }
    if (!publish_and_ticket<15>(acc, gram_partial + (uint64_t)blockIdx.x * 16, &counters[step_gram]))
        return;
    __shared__ double tot[15];
    reduce_partials<15>(gram_partial, gridDim.x, 16, tot);
    if (threadIdx.x == 0) {
        double m[4][4], rhs[4];
        int e = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = i; j < 4; ++j) {
                m[i][j] = tot[e];
                m[j][i] = tot[e];
                ++e;
            }
        for (int i = 0; i < 4; ++i) rhs[i] = tot[10 + i];
        int ok = 0;
        if (tot[14] >= (double)nc) ok = lsm_solve4(m, rhs, nc) ? 1 : 0;
        for (int i = 0; i < 4; ++i) coeff_all[step_gram * 4 + i] = (ok && i < nc) ? rhs[i] : 0.0;
        valid_all[step_gram] = ok;
        counters[step_gram] = 0u;  // ready for the next pricing
    }
}

// Reduction of one date's matched block sums (partial[value][block], nb
// blocks) in lsm_fused_kernel's order: block g sums the 32 blocks of its
// group, the last block to finish sums the group totals and solves.
__global__ __launch_bounds__(256) void lsm_group_kernel(const mp_lsm_params* __restrict__ Pp,
                                                       const double* partial, int nb,
                                                       double* gpart, unsigned int* counter,
                                                       double* coeff_all, int* valid_all,
                                                       int step_gram, unsigned int* err) {
    __shared__ double gsum[15], tot[15];
    const int g = blockIdx.x;
    const int ng = gridDim.x;
    reduce_group<15>(partial, (uint64_t)nb, 32 * g, min(32, nb - 32 * g), 1u, err, gsum);
    if (ng > 1) {
        if (threadIdx.x < 15) mp_store_coherent(gpart + threadIdx.x * 32 + g, gsum[threadIdx.x]);
        if (threadIdx.x < MP_WAVE) mp_vmem_drain_release();
        __shared__ bool is_last;
        __syncthreads();
        if (threadIdx.x == 0)
            is_last = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                      (unsigned int)ng - 1;
        __syncthreads();
        if (!is_last) return;
        mp_vmem_drain_acquire();
        reduce_group<15>(gpart, 32, 0, ng, 1u, err, tot);
    } else if (threadIdx.x < 15) {
        tot[threadIdx.x] = gsum[threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double c[4];
        const int ok = lsm_coefficients(tot, Pp->coeffs, c);
        for (int a = 0; a < 4; ++a) coeff_all[step_gram * 4 + a] = c[a];
        valid_all[step_gram] = ok;
        *counter = 0u;  // ready for the next pricing
    }
}

// Sums of discounted cash flows and exercise counts; the last block writes
// price, standard error, counts and the number of regression dates.
__global__ __launch_bounds__(1024) void lsm_final_kernel(const mp_lsm_params* __restrict__ Pp,
                                                       const double* cash, const uint16_t* ex,
                                                       const int* valid_all, double* partial,
                                                       unsigned int* counter, double* out) {
    const mp_lsm_params& P = *Pp;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P.n_paths;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const double c = cash[p];
        v[0] += c;
        v[1] += c * c;
        const int e = ex[p];
        if (e < P.n_steps)
            v[2] += 1.0;
        else if (c > 0.0)
            v[3] += 1.0;
    }
    if (!publish_and_ticket<4>(v, partial + (uint64_t)blockIdx.x * 4, counter)) return;
    __shared__ double tot[4];
    reduce_partials<4>(partial, gridDim.x, 4, tot);
    if (threadIdx.x == 0) {
        #pragma clang fp contract(off)  // the reference's summaries round every product (no FMA)
        int rs = 0;
        for (int s = 1; s <= P.n_steps; ++s) rs += valid_all[s] ? 1 : 0;
        const double n = (double)P.n_paths;
        const double mean = tot[0] / n;
        const double var = tot[1] / n - mean * mean;
        double price = mean, se = sqrt(var > 0.0 ? var : 0.0) / sqrt(n);
        if (P.allow_intrinsic) {
            const double intrinsic = P.k - P.s0 > 0.0 ? P.k - P.s0 : 0.0;
            if (intrinsic > price) {
                price = intrinsic;
                se = 0.0;
            }
        }
        out[0] = price;
        out[1] = se;
        out[2] = tot[2];
        out[3] = tot[3];
        out[4] = (double)rs;
        *counter = 0u;
    }
}

// ---------------------------------------------------------------------------
// The whole Longstaff-Schwartz pricing in one launch (requests up to
// 16 x 256 x MP_LSM_MAX_BLOCKS paths).  Thread t of block b owns paths
// b*256 + t + j*G*256 (j < J, G blocks): it generates them forward, writes
// S at dates 1..M-1 step-major, and keeps their discounted cash flows and
// exercise dates in registers for the whole backward pass.  Per regression
// date i every block stores its Gram/RHS sums into write-once slots
// slots[i][value][block]; block g < NG waits for the sums of blocks
// 32g .. 32g + 31 and stores their fixed-order total into gslots[i][value][g];
// every block then waits for the NG group totals, adds them in group order
// and solves the 4x4 system itself (the same bits in every block), so a date
// costs two cross-block hops.  (A coherent load costs the issuing CU about
// one request per lane, so the host picks the fewest blocks -- one per CU
// when the paths allow -- which keeps NG, and every block's loads of the
// group totals, small.)  Every wait is bounded: after spin_limit polls a
// block records an error and carries on, so every wave reaches the end.
// The block count must not exceed what the device holds at once (checked
// by the host).  Arithmetic and its order per path equal lsm_paths_kernel,
// lsm_step_kernel and lsm_final_kernel; only the grouping of the FP64 Gram
// sums differs (paths per thread, then wave, block and fixed block order),
// and lsm_group_kernel adds the group totals in the same order.
// ---------------------------------------------------------------------------
template <class R, int SRC, int J, int BS>
__global__ __launch_bounds__(BS) void lsm_fused_kernel(const mp_lsm_params* __restrict__ Pp,
                                                       const double* __restrict__ normals,
                                                       R* __restrict__ paths,
                                                       const int* __restrict__ reg_dates, int n_reg,
                                                       double* coeff_all, int* valid_all,
                                                       double* slots, double* gslots,
                                                       double* fin_partial, unsigned int* ticket,
                                                       uint16_t* ex_out, double* out,
                                                       unsigned int* err, unsigned int spin_limit) {
    const mp_lsm_params& P = *Pp;
    const uint64_t n = P.n_paths;
    const int m = P.n_steps;
    const double k = P.k, inv_k = P.inv_k;
    const int deg = P.degree, nc = P.coeffs;
    const bool wtd = P.weighted != 0;
    const double* disc_pow = reinterpret_cast<const double*>(Pp + 1);
    const double* growth = disc_pow + (m + 1);
    const int G = gridDim.x;
    const uint64_t T = (uint64_t)G * blockDim.x;
    const uint64_t p0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;

    // forward paths (lsm_paths_kernel's arithmetic)
    double cash[J];
    int exd[J];
    {
        const R drift = R(P.drift_dt), vol = R(P.vol_dt), s0 = R(P.s0);
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const uint64_t p = p0 + (uint64_t)j * T;
            cash[j] = 0.0;
            exd[j] = m;
            if (p >= n) continue;
            R x = R(0);
            R sT = s0;
            for (int s0i = 0; s0i < m; s0i += 4) {
                R z[4];
                const int nsub = min(4, m - s0i);
                if constexpr (SRC == 0) {
                    mp_normals4<R>((uint32_t)p, (uint32_t)(s0i >> 2), 0u, P.seed_hi, P.key0, P.key1, z);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) z[e] = e < nsub ? R(normals[(uint64_t)(s0i + e) * n + p]) : R(0);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (e < nsub) {
                        x += drift + vol * z[e];
                        const int step = s0i + e + 1;
                        const R S = s0 * mp_exp(x);
                        if (step < m) paths[(uint64_t)(step - 1) * n + p] = S;
                        sT = S;
                    }
                }
            }
            const double d = k - (double)sT;
            cash[j] = (d > 0.0 ? d : 0.0) * P.disc_m;
        }
    }

    __shared__ double tot[15], gsum[15];
    __shared__ double beta_s[4], cpoly_s[4];
    __shared__ int valid_s;
    // Lanes past the last path take no part in the regression.  (A padded
    // price of R(k) is not enough: in FP32, k - float(k) can be positive.)
    bool live[J];
#pragma unroll
    for (int j = 0; j < J; ++j) live[j] = p0 + (uint64_t)j * T < n;
    // Prices of the current regression date (sd) and of the next one (sn),
    // loaded one exchange ahead.  A date's basis values, Gram sums and count
    // depend on its prices only, not on exercise decisions, so they are
    // formed while the previous date's sums travel between blocks; the basis
    // values (bb) are kept for that date's right-hand side, which needs the
    // cash flows after the later dates' exercise and is formed after the
    // previous solve.  The exercise test evaluates the basis once more.
    R sd[J], sn[J];
    double bb[J][4], gacc[11];
    auto load_date = [&](int i, R (&dst)[J]) {
        const R* pd = paths + (uint64_t)(reg_dates[i] - 1) * n;
#pragma unroll
        for (int j = 0; j < J; ++j) dst[j] = live[j] ? pd[(uint64_t)(p0 + (uint64_t)j * T)] : R(0);
    };
    auto basis_of = [&](const R (&sx)[J]) {
#pragma unroll
        for (int q = 0; q < 11; ++q) gacc[q] = 0.0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const double s = (double)sx[j];
            laguerre(s * inv_k, deg, wtd, bb[j]);
            if (live[j] && k - s > 0.0) {
                int e = 0;
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = a; b < 4; ++b) gacc[e++] += bb[j][a] * bb[j][b];
                gacc[10] += 1.0;
            }
        }
    };
    if (n_reg > 0) {
        load_date(0, sd);
        basis_of(sd);
        if (n_reg > 1) load_date(1, sn);
    }
    const int NG = (G + 31) >> 5;
    // a date's index and discount factors are loaded one date ahead, so the
    // dependent global loads stay off the per-date critical path
    int date_n = n_reg > 0 ? reg_dates[0] : 0;
    double grow_n = n_reg > 0 ? growth[date_n] : 0.0, dpow_n = n_reg > 0 ? disc_pow[date_n] : 0.0;
    for (int i = 0; i < n_reg; ++i) {
        const int date = date_n;
        const double grow = grow_n, dpow = dpow_n;
        if (i + 1 < n_reg) {
            date_n = reg_dates[i + 1];
            grow_n = growth[date_n];
            dpow_n = disc_pow[date_n];
        }
        double acc[15];
#pragma unroll
        for (int q = 0; q < 10; ++q) acc[q] = gacc[q];
        acc[14] = gacc[10];
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[10 + a] = 0.0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const double s = (double)sd[j];
            if (live[j] && k - s > 0.0) {
                const double cont = cash[j] * grow;
#pragma unroll
                for (int a = 0; a < 4; ++a) acc[10 + a] += bb[j][a] * cont;
            }
        }
        double* slot_i = slots + (uint64_t)i * 16 * G;
        double* gslot_i = gslots + (uint64_t)i * 16 * 32;
        block_sum_to_slots<15>(acc, slot_i + blockIdx.x, (uint64_t)G);
        // blocks 0 .. NG-1 each reduce a group of 32 blocks' sums and publish
        // the group total
        if ((int)blockIdx.x < NG) {
            const int g = blockIdx.x;
            reduce_group<15>(slot_i, (uint64_t)G, 32 * g, min(32, G - 32 * g), spin_limit, err, gsum);
            if (threadIdx.x < 15) mp_store_coherent(gslot_i + threadIdx.x * 32 + g, gsum[threadIdx.x]);
        }
        // off the critical path: the next date's basis and Gram sums, and the
        // prices of the date after it
        R s2[J];
#pragma unroll
        for (int j = 0; j < J; ++j) s2[j] = R(0);
        if (i + 1 < n_reg) basis_of(sn);
        if (i + 2 < n_reg) load_date(i + 2, s2);
        // every block adds the group totals in group order and solves
        reduce_group<15>(gslot_i, 32, 0, NG, spin_limit, err, tot);
        if (threadIdx.x == 0) {
            double c[4];
            const int ok = lsm_coefficients(tot, nc, c);
            double cp[4];
            lsm_cont_poly(c, cp);
            for (int a = 0; a < 4; ++a) {
                beta_s[a] = c[a];
                cpoly_s[a] = cp[a];
            }
            valid_s = ok;
            if (blockIdx.x == 0) {
                // host outputs and the regression count (read after the final ticket)
                for (int a = 0; a < 4; ++a) mp_store_coherent(coeff_all + date * 4 + a, c[a]);
                __hip_atomic_store(valid_all + date, ok, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
        if (valid_s) {
            double beta[4], cp[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                beta[a] = beta_s[a];
                cp[a] = cpoly_s[a];
            }
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const double s = (double)sd[j];
                const double imm = k - s > 0.0 ? k - s : 0.0;
                if (live[j] && imm > 0.0) {
                    double cont;
                    if (wtd) {
                        double bx[4];
                        laguerre(s * inv_k, deg, wtd, bx);
                        cont = 0.0;
#pragma unroll
                        for (int a = 0; a < 4; ++a) cont += beta[a] * bx[a];
                    } else {
                        cont = lsm_cont_eval(cp, s * inv_k);
                    }
                    if (imm > (cont > 0.0 ? cont : 0.0)) {
                        cash[j] = imm * dpow;
                        exd[j] = date;
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            sd[j] = sn[j];
            sn[j] = s2[j];
        }
        __syncthreads();  // beta_s / valid_s are rewritten at the next date
    }

    // price, standard error and exercise counts (lsm_final_kernel)
    double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const uint64_t p = p0 + (uint64_t)j * T;
        if (p >= n) continue;
        const double c = cash[j];
        v[0] += c;
        v[1] += c * c;
        if (exd[j] < m)
            v[2] += 1.0;
        else if (c > 0.0)
            v[3] += 1.0;
        if (ex_out) ex_out[p] = (uint16_t)exd[j];
    }
    if (!publish_and_ticket<4>(v, fin_partial + (uint64_t)blockIdx.x * 4, ticket)) return;
    reduce_partials<4>(fin_partial, gridDim.x, 4, tot);
    if (threadIdx.x == 0) {
        #pragma clang fp contract(off)  // the reference's summaries round every product (no FMA)
        int rs = 0;
        for (int s = 1; s <= m; ++s) rs += __hip_atomic_load(valid_all + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 1 : 0;
        const double nn = (double)n;
        const double mean = tot[0] / nn;
        const double var = tot[1] / nn - mean * mean;
        double price = mean, se = sqrt(var > 0.0 ? var : 0.0) / sqrt(nn);
        if (P.allow_intrinsic) {
            const double intrinsic = P.k - P.s0 > 0.0 ? P.k - P.s0 : 0.0;
            if (intrinsic > price) {
                price = intrinsic;
                se = 0.0;
            }
        }
        out[0] = price;
        out[1] = se;
        out[2] = tot[2];
        out[3] = tot[3];
        out[4] = (double)rs;
    }
}

// ---------------------------------------------------------------------------
// Plain FP64 comparator (shape of the reference's CUDA contract): one lane per
// path, FP64 Philox normals and path, discounted payoff per path to HBM.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void plain_fp64_kernel(const mp_dreq* __restrict__ reqp,
                                                        double* payoffs) {
    const mp_dreq& q = *reqp;
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= q.n_paths) return;
    const model_params<double> m = load_params<double>(q);
    path_state<double> st;
    state_init(st, q);
    for (int s0i = 0; s0i < q.n_steps; s0i += 4) {
        double z[4];
        mp_normals4<double>((uint32_t)p, (uint32_t)(s0i >> 2), 0u, q.seed_hi, q.key0, q.key1, z);
        const int nsub = min(4, q.n_steps - s0i);
        for (int e = 0; e < nsub; ++e) gbm_step(st, m, z[e]);
    }
    payoffs[p] = q.discount * payoff_of(st, m, q, q.n_steps, q.s0, q.strikes[0]);
}

// ---------------------------------------------------------------------------
// The reference's CUDA kernel contract for the stepwise European call
// (crates/mc-core/src/backend/cuda.rs:247-310: buffer 0 normals, path-major
// [path][step], FP64; buffer 1 payoffs [path], FP64; scalars n_paths,
// n_steps, log_s0, strike, drift_dt, vol_dt, discount; one logical thread
// per path, 256 per group; body kernels/european_call_stepwise_v1.cu:1-28).
// One lane per path as the contract fixes, but the block's 256 rows are
// read as tiles of 256 paths x 16 steps: consecutive lanes load consecutive
// addresses of a row (128-byte runs), four tiles (64 steps) are loaded into
// registers at once so that one HBM round trip covers them, and each tile
// then goes through LDS (rows padded by one word against bank conflicts)
// for each lane to walk its own row.  Per path: log S += drift_dt +
// vol_dt * z in step order, then (S_T - K) * discount if S_T > K, else 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void european_call_stepwise_v1_kernel(
    const double* __restrict__ normals, double* __restrict__ payoffs, int n_paths, int n_steps,
    double log_s0, double strike, double drift_dt, double vol_dt, double discount) {
    constexpr int C = 16;  // steps per tile
    constexpr int P = 4;   // tiles in flight
    __shared__ double tile[256][C + 1];
    const int t = threadIdx.x;
    const int64_t row0 = (int64_t)blockIdx.x * 256;
    const int64_t path = row0 + t;
    double log_s_t = log_s0;
    for (int s0 = 0; s0 < n_steps; s0 += C * P) {
        // element (r, c) of tile p is normals[(row0 + r) * n_steps + s0 + p C + c],
        // thread t holding e = t + 256 i (r = e / C, c = e % C)
        double v[P][C];
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int i = 0; i < C; ++i) {
                const int e = t + 256 * i;
                const int r = e / C, c = e % C;
                const int64_t pr = row0 + r;
                const int step = s0 + p * C + c;
                v[p][i] = step < n_steps && pr < n_paths ? normals[pr * n_steps + step] : 0.0;
            }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int base = s0 + p * C;
            if (base >= n_steps) break;
            const int cols = min(C, n_steps - base);
#pragma unroll
            for (int i = 0; i < C; ++i) {
                const int e = t + 256 * i;
                tile[e / C][e % C] = v[p][i];
            }
            __syncthreads();
            {
#pragma clang fp contract(off)  // each product rounded, as the reference's CPU path computes it
                for (int c = 0; c < cols; ++c) log_s_t += drift_dt + vol_dt * tile[t][c];
            }
            __syncthreads();
        }
    }
    if (path >= n_paths) return;
    const double s_t = exp(log_s_t);
    payoffs[path] = s_t > strike ? (s_t - strike) * discount : 0.0;
}

__global__ __launch_bounds__(256) void plain_reduce_kernel(const double* payoffs, uint64_t n,
                                                          double* partial) {
    double v[2] = {0.0, 0.0};
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const double x = payoffs[p];
        v[0] += x;
        v[1] += x * x;
    }
    block_sum_store<2>(v, partial + (uint64_t)blockIdx.x * MP_PARTIAL_STRIDE);
}

// ---------------------------------------------------------------------------
// Draw emission for parity tests.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void draws_kernel(const mp_dreq* __restrict__ reqp,
                                                   mp_qmc_params qp,
                                                   const uint32_t* __restrict__ nib, int kind,
                                                   uint32_t stream, uint64_t path0,
                                                   uint64_t n_paths, uint64_t n_draws,
                                                   double* out) {
    const mp_dreq& q = *reqp;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_paths) return;
    const uint64_t p = path0 + i;
    if (kind <= 1) {
        for (uint64_t j = 0; j < n_draws; j += 4) {
            double z[4];
            if (kind == 0)
                mp_normals4<double>((uint32_t)p, (uint32_t)(j >> 2), stream, q.seed_hi, q.key0, q.key1, z);
            else
                mp_uniforms4<double>((uint32_t)p, (uint32_t)(j >> 2), stream, q.seed_hi, q.key0, q.key1, z);
            for (uint64_t e = 0; e < 4 && j + e < n_draws; ++e) out[i * n_draws + j + e] = z[e];
        }
    } else {
        sobol_cursor<double, false> cs;
        cs.init(nib, qp.base_seed + (uint32_t)(p >> 16) * 0x9E3779B9u, (uint32_t)(p & 0xFFFFu));
        for (uint64_t j = 0; j < n_draws; ++j) {
            if (kind == 2)
                out[i * n_draws + j] = cs.normal((int)j);
            else
                out[i * n_draws + j] = (double)cs.raw((int)j);
        }
    }
}

// FP32 draws (normals of the FP32 generator) for parity of the hardware
// transcendental path: kind 4.
__global__ __launch_bounds__(256) void draws_f32_kernel(const mp_dreq* __restrict__ reqp,
                                                       uint32_t stream, uint64_t path0,
                                                       uint64_t n_paths, uint64_t n_draws,
                                                       double* out) {
    const mp_dreq& q = *reqp;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_paths) return;
    const uint64_t p = path0 + i;
    for (uint64_t j = 0; j < n_draws; j += 4) {
        float z[4];
        mp_normals4<float>((uint32_t)p, (uint32_t)(j >> 2), stream, q.seed_hi, q.key0, q.key1, z);
        for (uint64_t e = 0; e < 4 && j + e < n_draws; ++e) out[i * n_draws + j + e] = (double)z[e];
    }
}

}  // namespace

// ===========================================================================
// Launchers
// ===========================================================================
namespace mp {

hipError_t launch_reset(unsigned int* counter, int n, hipStream_t s) {
    hipLaunchKernelGGL(reset_kernel, dim3((n + 63) / 64), dim3(64), 0, s, counter, n);
    return hipGetLastError();
}

hipError_t launch_setup(const mp_dreq& d, mp_dreq* dst, unsigned int* counter, int n, hipStream_t s) {
    if (n < 0 || n > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(setup_kernel, dim3(1), dim3(64), 0, s, d, dst, counter, n);
    return hipGetLastError();
}

template <int U>
static hipError_t tiles_gbm(const tile_launch& L, hipStream_t s) {
    if (L.claim == 0)
        hipLaunchKernelGGL((tile_gbm_static_kernel<U>), dim3(L.grid), dim3(L.block), 0, s, L);
    else if (L.ppl == 2)
        hipLaunchKernelGGL((tile_gbm_kernel<2, U>), dim3(L.grid), dim3(L.block), 0, s, L);
    else if (L.ppl == 4)
        hipLaunchKernelGGL((tile_gbm_kernel<4, U>), dim3(L.grid), dim3(L.block), 0, s, L);
    else
        hipLaunchKernelGGL((tile_gbm_kernel<1, U>), dim3(L.grid), dim3(L.block), 0, s, L);
    return hipGetLastError();
}

template <class R, int SRC>
static hipError_t tiles_ppl(const tile_launch& L, hipStream_t s) {
    if (sizeof(R) == 4 && SRC == 0 && L.gbm_f32 && !L.mlmc) {
        if (L.unroll == 4) return tiles_gbm<4>(L, s);
        if (L.unroll == 2) return tiles_gbm<2>(L, s);
        return tiles_gbm<1>(L, s);
    }
    if (L.claim == 0) {
        if (L.mlmc)
            hipLaunchKernelGGL((tile_static_kernel<R, SRC, true>), dim3(L.grid), dim3(L.block), 0, s, L);
        else
            hipLaunchKernelGGL((tile_static_kernel<R, SRC, false>), dim3(L.grid), dim3(L.block), 0, s, L);
        return hipGetLastError();
    }
    if (L.mlmc) {
        hipLaunchKernelGGL((tile_kernel<R, SRC, 1, true>), dim3(L.grid), dim3(L.block), 0, s, L);
        return hipGetLastError();
    }
    switch (L.ppl) {
        case 2:
            hipLaunchKernelGGL((tile_kernel<R, SRC, 2, false>), dim3(L.grid), dim3(L.block), 0, s, L);
            break;
        case 4:
            hipLaunchKernelGGL((tile_kernel<R, SRC, 4, false>), dim3(L.grid), dim3(L.block), 0, s, L);
            break;
        default:
            hipLaunchKernelGGL((tile_kernel<R, SRC, 1, false>), dim3(L.grid), dim3(L.block), 0, s, L);
    }
    return hipGetLastError();
}

hipError_t launch_tiles(const tile_launch& L0, int precision, hipStream_t s) {
    // Positions per claim.  None -- tile_static_kernel, wave i taking
    // positions i, i + W, ... -- when every tile has a wave of its own or all
    // tiles belong to one request (equal cost), at one path per lane: the
    // queue's atomics, all waves hitting one counter at the start, cost
    // 10-25 us per launch.
    // Otherwise 1 while the queue holds fewer than 16 tiles per wave (every
    // wave should get work), up to 4 on long queues, where the counter's
    // atomic rate would bound the launch.
    tile_launch L = L0;
    const uint64_t waves = (uint64_t)L.grid * (uint64_t)(L.block / MP_WAVE);
    L.claim = (waves >= L.n_tiles || L.tile_req == nullptr) && (L.ppl == 1 || L.mlmc)
                  ? 0
                  : (int)std::min<uint64_t>(4, std::max<uint64_t>(1, L.n_tiles / std::max<uint64_t>(1, 16 * waves)));
    const bool buf = L.normals != nullptr;
    if (precision == 64) return buf ? tiles_ppl<double, 1>(L, s) : tiles_ppl<double, 0>(L, s);
    return buf ? tiles_ppl<float, 1>(L, s) : tiles_ppl<float, 0>(L, s);
}

hipError_t launch_finalize(const mp_dreq* reqs, int n_reqs, const double* partial, double* out,
                           hipStream_t s) {
    hipLaunchKernelGGL(finalize_kernel, dim3(n_reqs), dim3(64 * MP_NS_MAX), 0, s, reqs, partial, out);
    return hipGetLastError();
}

int qmc_tree_levels(uint64_t n_steps) {
    for (int j = 0; j <= MP_TREE_MAX_LEVELS; ++j)
        if (n_steps == (16ull << j)) return j;
    return -1;
}

size_t qmc_lds_bytes(int precision, int factors, int n_slots, int block, bool nib_lds) {
    const size_t rb = precision == 64 ? 8 : 4;
    return (nib_lds ? (size_t)kNibWords * 4 : 0) + (size_t)n_slots * factors * block * rb;
}

template <class R, int F, int PAY, int SCH>
static void qmc_launch_nl(bool nib_lds, dim3 g, dim3 b, size_t lds, hipStream_t s,
                          const mp_dreq* req, const mp_qmc_params& qp, const uint32_t* nib,
                          const mp_bbop* plan, double* partial, double* term, double* pay) {
    if (nib_lds && qp.bridge && qp.tree_levels >= 0) {
        if constexpr (sizeof(R) == 4)
            hipLaunchKernelGGL((qmc_tree_kernel<F, PAY, SCH>), g, b, lds, s, req, qp, nib, plan, partial,
                               term, pay);
        else
            hipLaunchKernelGGL((qmc_tree64_kernel<F, PAY, SCH>), g, b, lds, s, req, qp, nib, plan, partial,
                               term, pay);
        return;
    }
    if (nib_lds)
        hipLaunchKernelGGL((qmc_kernel<R, F, true, PAY, SCH>), g, b, lds, s, req, qp, nib, plan,
                           partial, term, pay);
    else
        hipLaunchKernelGGL((qmc_kernel<R, F, false, PAY, SCH>), g, b, lds, s, req, qp, nib, plan,
                           partial, term, pay);
}

hipError_t launch_qmc(const mp_dreq* req, const mp_qmc_params& qp, int payoff, int scheme,
                      const uint32_t* nib, const mp_bbop* plan, double* partial,
                      double* terminal_out, double* payoff_out, int precision, int grid, int block,
                      bool nib_lds, hipStream_t s) {
    size_t lds = qmc_lds_bytes(precision, qp.factors, qp.n_slots > 0 ? qp.n_slots : 1, block, nib_lds);
    // register-bridge kernels: the hashed-seed table (8 bytes per dimension)
    // after the Sobol tables
    if (nib_lds && qp.bridge && qp.tree_levels >= 0)
        lds = std::max(lds, (size_t)kNibWords * 4 + (size_t)qp.factors * (16u << qp.tree_levels) * 8);
    const dim3 g(grid), b(block);
    const bool euro = payoff == MP_PAY_EUROPEAN;
    if (precision == 64) {
        // payoff and scheme compiled in as for FP32; FP64 arithmetic is the
        // same in every specialisation (the same bits as the generic body)
        if (qp.factors == 2 && euro && scheme == 1)
            qmc_launch_nl<double, 2, MP_PAY_EUROPEAN, 1>(nib_lds, g, b, lds, s, req, qp, nib, plan, partial, terminal_out, payoff_out);
        else if (qp.factors == 2 && euro && scheme == 0)
            qmc_launch_nl<double, 2, MP_PAY_EUROPEAN, 0>(nib_lds, g, b, lds, s, req, qp, nib, plan, partial, terminal_out, payoff_out);
        else if (qp.factors == 2)
            qmc_launch_nl<double, 2, -1, -1>(nib_lds, g, b, lds, s, req, qp, nib, plan, partial, terminal_out, payoff_out);
        else if (euro)
            qmc_launch_nl<double, 1, MP_PAY_EUROPEAN, -1>(nib_lds, g, b, lds, s, req, qp, nib, plan, partial, terminal_out, payoff_out);
        else
            qmc_launch_nl<double, 1, -1, -1>(nib_lds, g, b, lds, s, req, qp, nib, plan, partial, terminal_out, payoff_out);
    } else if (qp.factors == 2) {
        // FP32 European Heston: payoff and scheme compiled in
        if (euro && scheme == 1)
            qmc_launch_nl<float, 2, MP_PAY_EUROPEAN, 1>(nib_lds, g, b, lds, s, req, qp, nib, plan, partial, terminal_out, payoff_out);
        else if (euro && scheme == 0)
            qmc_launch_nl<float, 2, MP_PAY_EUROPEAN, 0>(nib_lds, g, b, lds, s, req, qp, nib, plan, partial, terminal_out, payoff_out);
        else
            qmc_launch_nl<float, 2, -1, -1>(nib_lds, g, b, lds, s, req, qp, nib, plan, partial, terminal_out, payoff_out);
    } else {
        if (euro)
            qmc_launch_nl<float, 1, MP_PAY_EUROPEAN, -1>(nib_lds, g, b, lds, s, req, qp, nib, plan, partial, terminal_out, payoff_out);
        else
            qmc_launch_nl<float, 1, -1, -1>(nib_lds, g, b, lds, s, req, qp, nib, plan, partial, terminal_out, payoff_out);
    }
    return hipGetLastError();
}

template <class R, int NA, int TECH>
static void basket_src(int src, const mp_dreq* req, const double* chol, const double* drift,
                       const double* vol, const double* logs0, const double* weight,
                       const double* normals, const mp_sampler_dim* sampler, double* partial,
                       double* term, double* pay, int grid, int block, hipStream_t s) {
    if (src == 1)
        hipLaunchKernelGGL((basket_kernel<R, NA, 1, TECH>), dim3(grid), dim3(block), 0, s, req, chol,
                           drift, vol, logs0, weight, normals, sampler, partial, term, pay);
    else if (src == 2)
        hipLaunchKernelGGL((basket_kernel<R, NA, 2, TECH>), dim3(grid), dim3(block), 0, s, req, chol,
                           drift, vol, logs0, weight, normals, sampler, partial, term, pay);
    else
        hipLaunchKernelGGL((basket_kernel<R, NA, 0, TECH>), dim3(grid), dim3(block), 0, s, req, chol,
                           drift, vol, logs0, weight, normals, sampler, partial, term, pay);
}

template <class R, int NA>
static hipError_t basket_na(int src, int tech, const mp_dreq* req, const double* chol,
                            const double* drift, const double* vol, const double* logs0,
                            const double* weight, const double* normals, const mp_sampler_dim* sampler,
                            double* partial, double* term, double* pay, int grid, int block,
                            hipStream_t s) {
    if (tech == MP_TECH_ANTITHETIC)
        basket_src<R, NA, MP_TECH_ANTITHETIC>(src, req, chol, drift, vol, logs0, weight, normals,
                                              sampler, partial, term, pay, grid, block, s);
    else if (tech == MP_TECH_CONTROL_VARIATE)
        basket_src<R, NA, MP_TECH_CONTROL_VARIATE>(src, req, chol, drift, vol, logs0, weight, normals,
                                                   sampler, partial, term, pay, grid, block, s);
    else
        basket_src<R, NA, MP_TECH_STANDARD>(src, req, chol, drift, vol, logs0, weight, normals,
                                            sampler, partial, term, pay, grid, block, s);
    return hipGetLastError();
}

template <class R>
static hipError_t basket_r(int na, int src, int tech, const mp_dreq* req, const double* chol,
                           const double* drift, const double* vol, const double* logs0,
                           const double* weight, const double* normals, const mp_sampler_dim* sampler,
                           double* partial, double* term, double* pay, int grid, int block,
                           hipStream_t s) {
#define MP_BASKET_CASE(N)                                                                      \
    case N:                                                                                    \
        return basket_na<R, N>(src, tech, req, chol, drift, vol, logs0, weight, normals, sampler, \
                               partial, term, pay, grid, block, s);
    switch (na) {
        MP_BASKET_CASE(1)
        MP_BASKET_CASE(2)
        MP_BASKET_CASE(3)
        MP_BASKET_CASE(4)
        MP_BASKET_CASE(5)
        MP_BASKET_CASE(6)
        MP_BASKET_CASE(7)
        MP_BASKET_CASE(8)
        default:
            return hipErrorInvalidValue;
    }
#undef MP_BASKET_CASE
}

hipError_t launch_basket(const mp_dreq* req, const mp_sampler_dim* sampler, int n_assets,
                         bool structured, int technique, const double* chol, const double* drift,
                         const double* vol, const double* log_s0, const double* weight,
                         const double* normals, double* partial, double* terminal_out,
                         double* payoff_out, int precision, int grid, int block, hipStream_t s) {
    const int src = normals ? 1 : structured ? 2 : 0;
    if (precision == 64)
        return basket_r<double>(n_assets, src, technique, req, chol, drift, vol, log_s0, weight,
                                normals, sampler, partial, terminal_out, payoff_out, grid, block, s);
    return basket_r<float>(n_assets, src, technique, req, chol, drift, vol, log_s0, weight, normals,
                           sampler, partial, terminal_out, payoff_out, grid, block, s);
}

template <class R>
static hipError_t greeks_r(const mp_dreq* req, int est, const double* normals, double* partial,
                           int grid, int block, hipStream_t s) {
    if (est == MP_GREEK_EST_PATHWISE) {
        if (normals)
            hipLaunchKernelGGL((greeks_kernel<R, 1, MP_GREEK_EST_PATHWISE>), dim3(grid), dim3(block), 0, s,
                               req, normals, partial);
        else
            hipLaunchKernelGGL((greeks_kernel<R, 0, MP_GREEK_EST_PATHWISE>), dim3(grid), dim3(block), 0, s,
                               req, normals, partial);
    } else {
        if (normals)
            hipLaunchKernelGGL((greeks_kernel<R, 1, MP_GREEK_EST_LIKELIHOOD_RATIO>), dim3(grid), dim3(block),
                               0, s, req, normals, partial);
        else
            hipLaunchKernelGGL((greeks_kernel<R, 0, MP_GREEK_EST_LIKELIHOOD_RATIO>), dim3(grid), dim3(block),
                               0, s, req, normals, partial);
    }
    return hipGetLastError();
}

hipError_t launch_greeks(const mp_dreq* req, int estimator, const double* normals, double* partial,
                         int precision, int grid, int block, hipStream_t s) {
    if (block != 256 || grid < 1) return hipErrorInvalidValue;
    if (estimator != MP_GREEK_EST_PATHWISE && estimator != MP_GREEK_EST_LIKELIHOOD_RATIO)
        return hipErrorInvalidValue;
    if (precision == 64) return greeks_r<double>(req, estimator, normals, partial, grid, block, s);
    return greeks_r<float>(req, estimator, normals, partial, grid, block, s);
}

hipError_t launch_lsm_fill(unsigned int* counters, int n_counters, int* valid, double* coeff,
                           int n_dates, hipStream_t s) {
    const int n = std::max(n_counters, 4 * n_dates);
    hipLaunchKernelGGL(lsm_fill_kernel, dim3((n + 255) / 256), dim3(256), 0, s, counters,
                       n_counters, valid, coeff, n_dates);
    return hipGetLastError();
}

hipError_t launch_lsm_paths(const mp_lsm_params* p, const double* normals, void* paths,
                            double* cash, uint16_t* ex, int precision, int grid, int block,
                            hipStream_t s) {
    if (precision == 64) {
        if (normals)
            hipLaunchKernelGGL((lsm_paths_kernel<double, 1>), dim3(grid), dim3(block), 0, s, p,
                               normals, (double*)paths, cash, ex);
        else
            hipLaunchKernelGGL((lsm_paths_kernel<double, 0>), dim3(grid), dim3(block), 0, s, p,
                               normals, (double*)paths, cash, ex);
    } else {
        if (normals)
            hipLaunchKernelGGL((lsm_paths_kernel<float, 1>), dim3(grid), dim3(block), 0, s, p,
                               normals, (float*)paths, cash, ex);
        else
            hipLaunchKernelGGL((lsm_paths_kernel<float, 0>), dim3(grid), dim3(block), 0, s, p,
                               normals, (float*)paths, cash, ex);
    }
    return hipGetLastError();
}

// Blocks of 1024 threads take up to 4 paths per thread (the register budget
// of four waves per SIMD), blocks of 512 up to 8, blocks of 256 up to 16.
template <class R, int SRC>
static const void* lsm_fused_fn(int J, int bs) {
    if (bs == 1024) {
        switch (J) {
            case 1: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 1, 1024>);
            case 2: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 2, 1024>);
            case 4: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 4, 1024>);
            default: return nullptr;
        }
    }
    if (bs == 512) {
        switch (J) {
            case 1: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 1, 512>);
            case 2: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 2, 512>);
            case 4: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 4, 512>);
            case 8: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 8, 512>);
            default: return nullptr;
        }
    }
    if (bs != 256) return nullptr;
    switch (J) {
        case 1: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 1, 256>);
        case 2: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 2, 256>);
        case 4: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 4, 256>);
        case 8: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 8, 256>);
        case 16: return reinterpret_cast<const void*>(&lsm_fused_kernel<R, SRC, 16, 256>);
        default: return nullptr;
    }
}

static const void* lsm_fused_select(int precision, bool normals, int J, int bs) {
    if (precision == 64) return normals ? lsm_fused_fn<double, 1>(J, bs) : lsm_fused_fn<double, 0>(J, bs);
    return normals ? lsm_fused_fn<float, 1>(J, bs) : lsm_fused_fn<float, 0>(J, bs);
}

int lsm_fused_blocks_per_cu(int precision, bool normals, int J, int bs) {
    const void* f = lsm_fused_select(precision, normals, J, bs);
    int nb = 0;
    if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, bs, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

__global__ __launch_bounds__(256) void slot_fill_kernel(unsigned long long* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = MP_SLOT_EMPTY;
}

hipError_t launch_slot_fill(double* slots, uint64_t n, hipStream_t s) {
    const uint64_t blocks = std::min<uint64_t>(1024, (n + 255) / 256);
    hipLaunchKernelGGL(slot_fill_kernel, dim3((unsigned)std::max<uint64_t>(1, blocks)), dim3(256), 0, s,
                       reinterpret_cast<unsigned long long*>(slots), n);
    return hipGetLastError();
}

hipError_t launch_lsm_fused(const mp_lsm_params* p, const double* normals, void* paths,
                            const int* reg_dates, int n_reg, double* coeff_all, int* valid_all,
                            double* slots, double* gslots, double* fin_partial,
                            unsigned int* ticket, uint16_t* ex, double* out, unsigned int* err,
                            unsigned int spin_limit, int J, int bs, int grid, int precision,
                            hipStream_t s) {
    const void* f = lsm_fused_select(precision, normals != nullptr, J, bs);
    if (!f || grid < 1 || grid > MP_LSM_MAX_BLOCKS) return hipErrorInvalidValue;
    void* args[] = {&p, &normals, &paths, &reg_dates, &n_reg, &coeff_all, &valid_all, &slots,
                    &gslots, &fin_partial, &ticket, &ex, &out, &err, &spin_limit};
    return hipLaunchKernel(f, dim3(grid), dim3(bs), args, 0, s);
}

template <bool MATCHED>
static void lsm_step_m(const mp_lsm_params* p, const void* paths, double* cash, uint16_t* ex,
                       double* coeff_all, int* valid_all, int step_update, int step_gram,
                       double* gram_partial, unsigned int* counters, int precision, int grid,
                       int block, hipStream_t s) {
    if (precision == 64)
        hipLaunchKernelGGL((lsm_step_kernel<double, MATCHED>), dim3(grid), dim3(block), 0, s, p,
                           (const double*)paths, cash, ex, coeff_all, valid_all, step_update,
                           step_gram, gram_partial, counters);
    else
        hipLaunchKernelGGL((lsm_step_kernel<float, MATCHED>), dim3(grid), dim3(block), 0, s, p,
                           (const float*)paths, cash, ex, coeff_all, valid_all, step_update,
                           step_gram, gram_partial, counters);
}

hipError_t launch_lsm_step(const mp_lsm_params* p, const void* paths, double* cash,
                           uint16_t* ex, double* coeff_all, int* valid_all, int step_update,
                           int step_gram, double* gram_partial, unsigned int* counters,
                           int precision, int grid, int block, bool matched, hipStream_t s) {
    if (grid < 1 || grid > MP_LSM_MAX_BLOCKS || (block != 256 && block != 512 && block != 1024))
        return hipErrorInvalidValue;
    if (matched)
        lsm_step_m<true>(p, paths, cash, ex, coeff_all, valid_all, step_update, step_gram,
                         gram_partial, counters, precision, grid, block, s);
    else
        lsm_step_m<false>(p, paths, cash, ex, coeff_all, valid_all, step_update, step_gram,
                          gram_partial, counters, precision, grid, block, s);
    return hipGetLastError();
}

hipError_t launch_lsm_group(const mp_lsm_params* p, const double* partial, int nb, double* gpart,
                            unsigned int* counter, double* coeff_all, int* valid_all, int step_gram,
                            unsigned int* err, hipStream_t s) {
    if (nb < 1 || nb > MP_LSM_MAX_BLOCKS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(lsm_group_kernel, dim3((nb + 31) / 32), dim3(256), 0, s, p, partial, nb, gpart,
                       counter, coeff_all, valid_all, step_gram, err);
    return hipGetLastError();
}

hipError_t launch_lsm_final(const mp_lsm_params* p, const double* cash, const uint16_t* ex,
                            const int* valid_all, double* partial, unsigned int* counter,
                            double* out, int grid, int block, hipStream_t s) {
    if (grid < 1 || grid > MP_LSM_MAX_BLOCKS || (block != 256 && block != 512 && block != 1024))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(lsm_final_kernel, dim3(grid), dim3(block), 0, s, p, cash, ex, valid_all,
                       partial, counter, out);
    return hipGetLastError();
}

hipError_t launch_plain_fp64(const mp_dreq* req, double* payoffs, int grid, int block,
                             hipStream_t s) {
    hipLaunchKernelGGL(plain_fp64_kernel, dim3(grid), dim3(block), 0, s, req, payoffs);
    return hipGetLastError();
}

hipError_t launch_european_call_stepwise_v1(const double* normals, double* payoffs, int n_paths,
                                           int n_steps, double log_s0, double strike, double drift_dt,
                                           double vol_dt, double discount, hipStream_t s) {
    const int grid = (n_paths + 255) / 256;
    hipLaunchKernelGGL(european_call_stepwise_v1_kernel, dim3(grid), dim3(256), 0, s, normals, payoffs,
                       n_paths, n_steps, log_s0, strike, drift_dt, vol_dt, discount);
    return hipGetLastError();
}

hipError_t launch_plain_reduce(const double* payoffs, uint64_t n, double* partial, int n_blocks,
                               hipStream_t s) {
    hipLaunchKernelGGL(plain_reduce_kernel, dim3(n_blocks), dim3(256), 0, s, payoffs, n, partial);
    return hipGetLastError();
}

hipError_t launch_draws(const mp_dreq* req, const mp_qmc_params& qp, const uint32_t* nib,
                        int kind, uint32_t stream, uint64_t path0, uint64_t n_paths,
                        uint64_t n_draws, double* out, hipStream_t s) {
    const int grid = (int)((n_paths + 255) / 256);
    if (kind == 4)
        hipLaunchKernelGGL(draws_f32_kernel, dim3(grid), dim3(256), 0, s, req, stream, path0,
                           n_paths, n_draws, out);
    else
        hipLaunchKernelGGL(draws_kernel, dim3(grid), dim3(256), 0, s, req, qp, nib, kind, stream,
                           path0, n_paths, n_draws, out);
    return hipGetLastError();
}

}  // namespace mp
