// mp_engine.h -- device-side request descriptors and kernel launchers shared
// by the kernels (mp_engine.hip) and the C-ABI host layer (mp_api.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MP_NS_MAX 4            // strikes priced from one simulation
#define MP_POISSON_TERMS 24    // cached cumulative Poisson probabilities per request
#define MP_PARTIAL_STRIDE (2 * MP_NS_MAX)  // doubles per tile partial: (sum, sumsq) per strike
// Control-variate tiles use the first five slots of a partial row for the
// discounted moments sum P, sum P^2, sum C, sum C^2, sum P*C (one strike).
enum mp_technique { MP_TECH_STANDARD = 0, MP_TECH_ANTITHETIC = 1, MP_TECH_CONTROL_VARIATE = 2 };
// Finalize mode for single-pass Greek estimators and multilevel levels
// (never a caller's technique): the eight columns of a partial row are
// (sum, sum of squares) of up to four per-path estimates, written out raw.
#define MP_TECH_RAW_MOMENTS 3
enum mp_greek_est { MP_GREEK_EST_PATHWISE = 1, MP_GREEK_EST_LIKELIHOOD_RATIO = 2 };
// Structured samplers drawn per (point, dimension) in the work-queue and
// basket kernels (the ABI's sampling codes).
enum mp_structured { MP_SAMPLER_NONE = 0, MP_SAMPLER_HALTON = 1, MP_SAMPLER_LHS = 2,
                     MP_SAMPLER_SOBOL = 3 /* multilevel levels only: one Sobol point per path */ };

// One dimension of a structured sampler.  Randomized Halton: a = prime base,
// shift = Cranley-Patterson shift.  Latin hypercube: a = stride multiplier,
// b = offset, c = jitter seed.
struct mp_sampler_dim {
    uint64_t a, b, c;
    double shift;
};
// Largest structured-sampler dimension count per request (32 B per dimension).
#define MP_SAMPLER_MAX_DIMS 65536
// LSM kernels run at most this many blocks: the last block reads every
// block's partial sums in one round trip (two rows per thread).
#define MP_LSM_MAX_BLOCKS 1024
#define MP_TREE_MAX_LEVELS 4

enum mp_model { MP_MODEL_GBM = 0, MP_MODEL_HESTON = 1, MP_MODEL_MERTON = 2 };
enum mp_payoff { MP_PAY_EUROPEAN = 0, MP_PAY_ASIAN = 1, MP_PAY_DOWN_AND_OUT = 2, MP_PAY_LOOKBACK = 3,
                 // one level of a multilevel arithmetic Asian: the discounted
                 // fine payoff minus the coupled coarse payoff (level 0: fine only)
                 MP_PAY_ASIAN_MLMC = 4 };

// Per-request descriptor resident in HBM.  All fields are uniform across a
// wave, so the kernels read them with scalar loads.
struct mp_dreq {
    int32_t model;
    int32_t payoff;
    int32_t n_steps;      // steps actually walked (1 for the terminal method)
    int32_t n_strikes;
    int32_t scheme;       // Heston scheme
    int32_t n_cum;        // Merton: cached cumulative Poisson terms
    uint32_t key0, key1;  // Philox key (seed low word, contract id)
    uint32_t seed_hi;     // Philox counter word 3
    uint32_t tile0;       // first tile id of this request
    uint32_t n_tiles;
    uint32_t tile_paths;  // paths per tile (multiple of 64)
    uint64_t n_paths;
    double log_s0, s0;
    double drift_dt, vol_dt;  // GBM/Merton per-step log drift (compensated for Merton) and vol
    double log_barrier_rel;   // ln(B) - ln(S0): knock-out level in log-return units
    int32_t knocked_at_start; // S0 <= B
    int32_t technique;        // 0 standard, 1 antithetic pairs, 2 control variate
    double strikes[MP_NS_MAX];
    double discount;
    double cv_expect;         // E[discounted control] (control variate)
    int32_t sampling;         // mp_structured: 0 = Philox (or caller buffers)
    uint32_t sampler_off;     // first mp_sampler_dim of this request
    uint64_t sampler_n;       // points the sampler is built over (requested n_paths)
    double v0, kappa, theta, xi, rho, rho_perp, rate, dt, sqrt_dt;
    double sigma;             // GBM volatility (single-pass Greeks)
    int32_t mlmc_refine;      // multilevel: fine steps per coarse step (0 at level 0)
    int32_t pad_mlmc;
    double mlmc_coarse_drift, mlmc_coarse_vol, mlmc_sqrt_refine;
    double mu_j, sigma_j, jump_mean_dt;
    double cum[MP_POISSON_TERMS];
};

// Brownian-bridge op in streaming (preorder) form on interval increments:
// stack[slot] holds D = W(r) - W(l); the left part is dl = rw*D + sd*z and the
// right part dr = D - dl.  flags bit0: left child is a leaf (emit dl, else push
// it to stack[slot + 1]); bit1: right child is a leaf (emit dr, else store it
// in stack[slot]).  rwf/sdf are rw/sd rounded to FP32 for the FP32 kernel.
struct mp_bbop {
    float rwf, sdf;
    float sdf_e;    // FP32 register bridge: 2^(depth + 1) sd, depth of the split interval (root 0)
    float sdf_c0;   // FP32 European Heston register bridge: factor f carries c_f 2^(depth - K) D,
    float sdf_c1;   // so a split adds +- c_f 2^(depth + 1 - K) sd z (c_0 = xi, c_1 = rho_perp, 2^K steps)
    float pad_;
    double rw, sd;
    int32_t slot;
    int32_t flags;
};

struct mp_qmc_params {
    uint32_t base_seed;   // low 32 bits of the request seed
    int32_t factors;      // 1 (GBM) or 2 (Heston)
    int32_t bridge;       // 0: dims in step order, 1: Brownian bridge
    int32_t n_ops;
    int32_t n_slots;
    int32_t tree_levels;  // J when the bridge runs in registers (n_steps = 16 * 2^J), else -1
    int32_t pad;
    double sqrt_t;
    double inv_sqrt_dt;
    // FP32 European Heston register bridge: the root increments' scales
    // c_f 2^-K sqrt(T) and rho / xi (xi != 0; the host sends xi == 0 to the
    // generic kernel)
    double root_c0, root_c1, rho_over_xi;
};

// Followed in device memory by disc_pow[n_steps+1] (disc_step^s) and
// growth[n_steps+1] (disc_step^-s); the struct is padded to 128 bytes.
struct alignas(128) mp_lsm_params {
    int32_t n_steps;
    int32_t degree;
    int32_t coeffs;
    int32_t allow_intrinsic;
    uint64_t n_paths;
    double log_s0, s0, k, inv_k, drift_dt, vol_dt, disc_step, disc_m;
    uint32_t key0, key1, seed_hi;
    uint32_t weighted;          // basis e^{-x/2} L_i(x) instead of L_i(x)
};

// Launch helpers (defined in mp_engine.hip).
namespace mp {

struct tile_launch {
    const mp_dreq* reqs;
    const uint32_t* tile_req;   // tile id -> request index (null: every tile is request 0's)
    const uint32_t* order;      // queue position -> tile id
    uint32_t n_tiles;
    unsigned int* counter;
    double* partial;            // [tile][MP_PARTIAL_STRIDE]
    const double* normals;      // buffer source (step-major) or null
    const double* uniforms;
    const double* jump_normals;
    double* terminal_out;
    double* payoff_out;
    const mp_sampler_dim* sampler;  // structured-sampler tables of every request
    const uint32_t* nib;            // Sobol index tables (multilevel Sobol levels)
    int grid;
    int block;
    int ppl;                    // paths interleaved per lane
    int mlmc;                   // every request is a multilevel level (own kernel)
    int claim;                  // queue positions per claim, 0 = one tile per wave (launch_tiles sets it)
    int gbm_f32;                // every request is FP32 GBM-family with Philox draws and the standard
                                // technique: kernels holding that body only (fewer registers)
    int unroll;                 // FP32 GBM-family body: 4-step draw blocks per step-loop iteration (1, 2, 4)
};

hipError_t launch_reset(unsigned int* counter, int n, hipStream_t s);
// Stores one request record (a kernel argument) at dst and clears
// counter[0 .. n-1] (n <= 64): a launch in place of two copy commands.
hipError_t launch_setup(const mp_dreq& d, mp_dreq* dst, unsigned int* counter, int n, hipStream_t s);
hipError_t launch_tiles(const tile_launch& L, int precision, hipStream_t s);
hipError_t launch_finalize(const mp_dreq* reqs, int n_reqs, const double* partial, double* out,
                           hipStream_t s);
// payoff / scheme: the request's (they select compiled-in specialisations)
hipError_t launch_qmc(const mp_dreq* req, const mp_qmc_params& qp, int payoff, int scheme,
                      const uint32_t* nib, const mp_bbop* plan, double* partial,
                      double* terminal_out, double* payoff_out, int precision, int grid, int block,
                      bool nib_lds, hipStream_t s);
size_t qmc_lds_bytes(int precision, int factors, int n_slots, int block, bool nib_lds);
// J with n_steps = 16 * 2^J (J <= MP_TREE_MAX_LEVELS), else -1: the step
// counts whose Brownian bridge runs in registers (no LDS stack)
int qmc_tree_levels(uint64_t n_steps);
// structured: Halton / Latin hypercube draws; technique: MP_TECH_*
hipError_t launch_basket(const mp_dreq* req, const mp_sampler_dim* sampler, int n_assets,
                         bool structured, int technique, const double* chol, const double* drift,
                         const double* vol, const double* log_s0, const double* weight,
                         const double* normals, double* partial, double* terminal_out,
                         double* payoff_out, int precision, int grid, int block, hipStream_t s);
// Single-pass terminal-GBM Greeks (estimator mp_greek_est): per-block
// partial rows of four estimates' (sum, sum of squares).
hipError_t launch_greeks(const mp_dreq* req, int estimator, const double* normals, double* partial,
                         int precision, int grid, int block, hipStream_t s);
hipError_t launch_lsm_fill(unsigned int* counters, int n_counters, int* valid, double* coeff,
                           int n_dates, hipStream_t s);
hipError_t launch_lsm_paths(const mp_lsm_params* p, const double* normals, void* paths,
                            double* cash, uint16_t* ex, int precision, int grid, int block,
                            hipStream_t s);
// matched: block sums to gram_partial[value][block] for lsm_group (the
// one-launch kernel's reduction order); else the last block reduces and solves.
hipError_t launch_lsm_step(const mp_lsm_params* p, const void* paths, double* cash,
                           uint16_t* ex, double* coeff_all, int* valid_all, int step_update,
                           int step_gram, double* gram_partial, unsigned int* counters,
                           int precision, int grid, int block, bool matched, hipStream_t s);
// One date's reduction of matched block sums: ceil(nb / 32) blocks; gpart
// holds 16 x 32 doubles of group totals.
hipError_t launch_lsm_group(const mp_lsm_params* p, const double* partial, int nb, double* gpart,
                            unsigned int* counter, double* coeff_all, int* valid_all, int step_gram,
                            unsigned int* err, hipStream_t s);
// Whole pricing in one launch (grid blocks of bs threads, all resident at
// once; J paths per thread, J in {1,2,4} for bs = 1024 and {1,2,4,8,16} for
// bs = 256).  slots: n_reg * 16 * grid
// doubles and gslots: n_reg * 16 * 32, both filled by launch_slot_fill
// first; ticket and err zeroed; err is set when a wait timed out.
int lsm_fused_blocks_per_cu(int precision, bool normals, int J, int bs);
hipError_t launch_slot_fill(double* slots, uint64_t n, hipStream_t s);
hipError_t launch_lsm_fused(const mp_lsm_params* p, const double* normals, void* paths,
                            const int* reg_dates, int n_reg, double* coeff_all, int* valid_all,
                            double* slots, double* gslots, double* fin_partial,
                            unsigned int* ticket, uint16_t* ex, double* out, unsigned int* err,
                            unsigned int spin_limit, int J, int bs, int grid, int precision,
                            hipStream_t s);
hipError_t launch_lsm_final(const mp_lsm_params* p, const double* cash, const uint16_t* ex,
                            const int* valid_all, double* partial, unsigned int* counter,
                            double* out, int grid, int block, hipStream_t s);
hipError_t launch_plain_fp64(const mp_dreq* req, double* payoffs, int grid, int block,
                             hipStream_t s);
// The reference's stepwise European contract: normals [path][step] in, the
// discounted payoff of each path out; one thread per path, 256 per block.
hipError_t launch_european_call_stepwise_v1(const double* normals, double* payoffs, int n_paths,
                                           int n_steps, double log_s0, double strike, double drift_dt,
                                           double vol_dt, double discount, hipStream_t s);
hipError_t launch_plain_reduce(const double* payoffs, uint64_t n, double* partial,
                               int n_blocks, hipStream_t s);
hipError_t launch_draws(const mp_dreq* req, const mp_qmc_params& qp, const uint32_t* nib,
                        int kind, uint32_t stream, uint64_t path0, uint64_t n_paths,
                        uint64_t n_draws, double* out, hipStream_t s);

}  // namespace mp
