/*
 * montepath_hip.h -- C-ABI of the MI355X (gfx950) Monte Carlo path engine,
 * the `rocm_hip` backend of montepath.
 *
 * Plain C: integers, doubles and caller-owned pointers only.  Every entry
 * point returns an mp_status; no exception or abort crosses the boundary.
 * The library never keeps a caller pointer past return, except that
 * mp_hip_submit copies the request (and its exercise schedule) it is given.
 *
 * One mp_ctx owns one HIP stream and the device arenas resident on it (RNG
 * descriptors, LSM path/cash-flow buffers, reduction scratch).  A context
 * is single-thread-affine, like the reference's thread-local Metal runtime
 * cache (crates/mc-core/src/backend/metal.rs:52-55): use one context per
 * host thread.
 *
 * Reference interfaces this ABI replaces are cited per entry point
 * (paths relative to the montepath repository).
 */
#ifndef MONTEPATH_HIP_H
#define MONTEPATH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MP_ABI_VERSION 1

/* Status codes.  Mapping onto the reference's BackendError
 * (crates/mc-core/src/backend/mod.rs:186-194):
 *   MP_E_UNKNOWN_DEVICE      -> BackendError::UnknownDevice
 *   MP_E_INCOMPATIBLE_INPUT  -> BackendError::IncompatibleExecutionInput
 *   MP_E_UNSUPPORTED, MP_E_OOM, MP_E_HIP -> BackendError::UnsupportedFeature(msg)
 *   MP_E_INVALID_CONFIG      -> the assert! panics of the reference's config
 *                               validators (crates/mc-core/src/runtime/cpu.rs:5010-5153)
 *   MP_E_NOT_FOUND           -> unknown serving ticket
 *   MP_E_CANCELLED           -> a serving request cancelled before it ran */
typedef enum mp_status {
    MP_OK = 0,
    MP_E_UNKNOWN_DEVICE = 1,
    MP_E_INCOMPATIBLE_INPUT = 2,
    MP_E_UNSUPPORTED = 3,
    MP_E_OOM = 4,
    MP_E_HIP = 5,
    MP_E_INVALID_CONFIG = 6,
    MP_E_NOT_FOUND = 7,
    MP_E_CANCELLED = 8
} mp_status;

/* Workload families (reference: PricingWorkloadFamily and the
 * BackendExecutionInput variants, backend/mod.rs:179-184, extended with the
 * families the reference prices only on CPU). */
typedef enum mp_workload {
    MP_EUROPEAN_CALL = 1,
    MP_ARITHMETIC_ASIAN_CALL = 2,
    MP_DOWN_AND_OUT_CALL = 3,
    MP_LOOKBACK_CALL = 4,
    MP_HESTON_EUROPEAN_CALL = 5,
    MP_MERTON_JUMP_DIFFUSION_CALL = 6,
    MP_BASKET_CALL = 7,
    MP_AMERICAN_PUT = 8,
    MP_BERMUDAN_PUT = 9
} mp_workload;

/* Sampling (reference SamplingMethod, runtime/cpu.rs:5796-5830).  Randomized
 * Halton and Latin hypercube draw one point per path (per antithetic pair)
 * over at most 65536 dimensions: the GBM family, Heston (extension: factor f
 * of step s is dimension 2s + f) and the basket (asset a of step s is
 * dimension s * n_assets + a).  Merton and LSM take pseudorandom draws only. */
typedef enum mp_sampling {
    MP_PSEUDORANDOM = 0,
    MP_RANDOMIZED_HALTON = 1,
    MP_LATIN_HYPERCUBE = 2,
    MP_SCRAMBLED_SOBOL = 3,
    MP_SCRAMBLED_SOBOL_BRIDGE = 4
} mp_sampling;

/* European method (reference EuropeanCallMethod): the reference's Auto and
 * TerminalDistribution sample S_T directly; StepwisePaths walks n_steps. */
typedef enum mp_method {
    MP_METHOD_STEPWISE = 0,
    MP_METHOD_TERMINAL = 1
} mp_method;

/* MonteCarloTechnique (runtime/cpu.rs:24-28) */
typedef enum mp_technique_code {
    MP_TECHNIQUE_STANDARD = 0,
    MP_TECHNIQUE_ANTITHETIC = 1,
    MP_TECHNIQUE_CONTROL_VARIATE = 2
} mp_technique_code;

typedef enum mp_heston_scheme {
    MP_HESTON_REFERENCE_FLOOR = 0,  /* variance floored after each update, cpu.rs:8682 */
    MP_HESTON_FULL_TRUNCATION = 1   /* Lord et al. (2010): only v+ enters drift/diffusion */
} mp_heston_scheme;

#define MP_MAX_STRIKES 4
#define MP_MAX_ASSETS 8

/* One pricing request.  Zero-initialise, then fill what the workload uses.
 * Field meanings follow the reference configs (EuropeanCallConfig,
 * HestonEuropeanCallConfig, MertonJumpDiffusionCallConfig, BasketCallConfig,
 * AmericanPutConfig, BermudanPutConfig; runtime/cpu.rs:1296-1860). */
typedef struct mp_request {
    int32_t workload;        /* mp_workload */
    int32_t precision;       /* 32: FP32 paths + FP64 accumulation; 64: FP64 throughout */
    int32_t sampling;        /* mp_sampling */
    int32_t technique;       /* 0 standard, 1 antithetic pairs, 2 control variate (discounted
                                S_T, one strike); MonteCarloTechnique (runtime/cpu.rs:24-28).
                                Not offered for Merton and LSM, like the reference. */
    int32_t method;          /* mp_method (European only) */
    int32_t heston_scheme;   /* mp_heston_scheme */
    int32_t n_strikes;       /* 0 or 1: price strike k; 2..4: price strikes[] from one simulation */
    int32_t basis_degree;    /* LSM Laguerre degree, clamped to [1,3] like the reference */
    int32_t fp64_check;      /* nonzero: also run FP64 on the same draws, report fp64_ref_error */
    int32_t allow_intrinsic_now; /* American put floor at max(K - S0, 0) (cpu.rs:3911-3917) */
    int32_t n_assets;        /* basket */
    int32_t n_exercise;      /* Bermudan: entries in exercise_steps */
    int32_t basis_weighted;  /* LSM: nonzero regresses on the weighted Laguerre functions
                                e^{-x/2} L_i(x) of Longstaff-Schwartz (2001) instead of the
                                reference's unweighted L_i(x) (cpu.rs:7667-7682, the default) */
    uint64_t n_paths;
    uint64_t n_steps;
    uint64_t seed;
    uint64_t contract_id;    /* Philox key word: (seed, contract_id) names the stream */
    double s0, k, r, sigma, t;
    double barrier;                     /* down-and-out */
    double strikes[MP_MAX_STRIKES];     /* multi-strike European/Heston */
    double v0, kappa, theta, xi, rho;   /* Heston (xi = vol_of_vol) */
    double jump_intensity, jump_mean, jump_volatility; /* Merton */
    double asset_s0[MP_MAX_ASSETS];
    double asset_sigma[MP_MAX_ASSETS];
    double asset_weight[MP_MAX_ASSETS];
    double correlation[MP_MAX_ASSETS * MP_MAX_ASSETS]; /* row-major n_assets x n_assets */
    const uint32_t* exercise_steps;     /* Bermudan schedule, 1..n_steps */
} mp_request;

/* Result of one request (reference RunOutput {price, stderr, runtime_ms},
 * backend/mod.rs:172-177, plus the per-strike and LSM details the
 * reference returns in its typed results). */
typedef struct mp_result {
    int32_t status;          /* mp_status of this request */
    int32_t n_strikes;
    double price;
    double std_error;        /* population-variance standard error (cpu.rs:7641-7652) */
    double runtime_ms;
    double fp64_ref_error;   /* |price - price of an FP64 run on the same draws|; NaN if not run */
    double strike_price[MP_MAX_STRIKES];
    double strike_std_error[MP_MAX_STRIKES];
    uint64_t n_paths;
    uint64_t n_steps;
    uint64_t early_exercise_count;    /* LSM (AmericanPutResult fields) */
    uint64_t maturity_exercise_count;
    uint64_t regression_steps;
} mp_result;

/* Greeks (reference Greek / GreekEstimator / GreekEstimate, runtime/cpu.rs:
 * 140-170).  Up to MP_MAX_GREEKS estimates per request. */
#define MP_MAX_GREEKS 6
typedef enum mp_greek {
    MP_GREEK_DELTA = 0,
    MP_GREEK_DELTA2 = 1,     /* basket: second asset's spot */
    MP_GREEK_VEGA = 2,
    MP_GREEK_VEGA2 = 3,      /* basket: second asset's volatility */
    MP_GREEK_RHO = 4,
    MP_GREEK_THETA = 5       /* calendar convention: -dV/dT */
} mp_greek;

typedef enum mp_greek_estimator {
    MP_GREEKS_BUMP_AND_REVALUE = 0,
    MP_GREEKS_PATHWISE = 1,
    MP_GREEKS_LIKELIHOOD_RATIO = 2
} mp_greek_estimator;

typedef struct mp_greek_estimate {
    int32_t greek;             /* mp_greek */
    int32_t estimator;         /* mp_greek_estimator */
    double value;
    double std_error;          /* bump: independent-leg propagation; single pass: sample SE */
    double bump_size;          /* effective half-width of the central difference; NaN single pass */
    double base_price;
    double bumped_up_price;    /* NaN for single-pass estimators */
    double bumped_down_price;
} mp_greek_estimate;

/* Multilevel Monte Carlo arithmetic Asian call (ArithmeticAsianMlmcConfig,
 * runtime/cpu.rs:1536-1568; defaults s0 = k = 100, r = 0.03, sigma = 0.2,
 * t = 1, base_steps 16, 4 levels, refinement 2).  Level l walks
 * base_steps * refinement^l fine steps; levels >= 1 price the coupled
 * difference fine - coarse, the coarse path driven by sums of `refinement`
 * fine normals / sqrt(refinement).  Level l draws from the seed
 * derive_chunk_seed(seed, l) (replicate r of a structured run first takes
 * derive_chunk_seed(seed, r)): Philox keyed by it, or a structured sampler
 * built from it over (paths, fine steps). */
#define MP_MLMC_MAX_LEVELS 16
#define MP_MLMC_MAX_REPLICATES 64
typedef struct mp_mlmc_config {
    double s0, k, r, sigma, t;
    uint32_t base_steps;
    uint32_t levels;                /* 1..MP_MLMC_MAX_LEVELS */
    uint32_t refinement_factor;     /* >= 2 */
    uint32_t scramble_replicates;   /* structured sampling only; 1..MP_MLMC_MAX_REPLICATES */
    int32_t sampling;               /* mp_sampling (both Sobol codes draw plain Sobol points) */
    int32_t precision;              /* 32 or 64 */
    uint64_t seed;
    uint64_t contract_id;
    uint64_t paths_per_level[MP_MLMC_MAX_LEVELS];
} mp_mlmc_config;

/* ArithmeticAsianMlmcLevelResult (cpu.rs:1570-1580); coarse_steps 0 = none. */
typedef struct mp_mlmc_level {
    uint64_t level, paths, fine_steps, coarse_steps;
    double mean, variance, std_error;
    uint64_t cost_step_updates;
} mp_mlmc_level;

/* ArithmeticAsianMlmcResult (cpu.rs:1582-1591). */
typedef struct mp_mlmc_result {
    double price, std_error;
    uint64_t total_paths, total_step_updates;
    uint32_t scramble_replicates, n_levels;
    double runtime_ms;
    mp_mlmc_level levels[MP_MLMC_MAX_LEVELS];
    double replicate_estimates[MP_MLMC_MAX_REPLICATES];
} mp_mlmc_result;

/* ArithmeticAsianMlmcAllocationLevel / Plan (cpu.rs:1593-1610). */
typedef struct mp_mlmc_allocation_level {
    uint64_t level, fine_steps, coarse_steps, pilot_paths;
    double pilot_variance;
    uint64_t cost_per_path, recommended_paths;
} mp_mlmc_allocation_level;

typedef struct mp_mlmc_allocation {
    uint64_t target_step_updates, estimated_step_updates;
    uint32_t n_levels, reserved;
    mp_mlmc_allocation_level levels[MP_MLMC_MAX_LEVELS];
} mp_mlmc_allocation;

/* ArithmeticAsianMlmcToleranceConfig / TolerancePlan (cpu.rs:1612-1642). */
typedef struct mp_mlmc_tolerance {
    double target_stderr;
    uint64_t pilot_paths_per_level, min_step_updates, max_step_updates;
} mp_mlmc_tolerance;

typedef struct mp_mlmc_tolerance_plan {
    double target_stderr, estimated_stderr;
    int32_t target_met, max_step_updates_hit;
    uint32_t scramble_replicates, reserved;
    uint64_t estimated_step_updates;
    mp_mlmc_allocation allocation;
    mp_mlmc_config recommended;     /* the input with paths_per_level from the plan */
} mp_mlmc_tolerance_plan;

/* Device description (reference DeviceInfo, backend/mod.rs:32-43). */
typedef struct mp_device_info {
    char device_id[32];      /* "rocm:<ordinal>" */
    char name[96];
    char arch[32];           /* e.g. "gfx950" */
    uint64_t memory_total_mb;
    uint64_t memory_free_mb;
    int32_t compute_units;
    int32_t supports_float64;
    int32_t max_threads_hint;
    int32_t wavefront_size;
} mp_device_info;

typedef struct mp_ctx mp_ctx;

/* Library version string. */
const char* mp_hip_version(void);

/* Number of visible HIP devices (reference RuntimeBackend::discover_devices,
 * backend/mod.rs:200; CUDA's nvidia-smi probe, backend/cuda.rs:205-241). */
int mp_hip_device_count(int* count);

/* Open a context on one device: stream, arenas, Sobol tables (reference
 * RuntimeBackend::compile, backend/mod.rs:203-207, and the Metal runtime
 * cache set-up, backend/metal.rs:686-751).  Environment read here:
 * MONTEPATH_LSM_MULTI_LAUNCH=1 prices American/Bermudan puts with per-date
 * launches that give the one-launch pricing's bits and never wait on other
 * blocks (for contexts that share the device with concurrent work; the
 * one-launch form needs its whole grid resident), =2 with the plain
 * per-date launches; MONTEPATH_QMC_STACK_BRIDGE=1 keeps Brownian bridges
 * on 16 * 2^J steps in LDS (same bits as the register form). */
int mp_hip_open(int device_ordinal, mp_ctx** out);

/* Release the context and every device buffer it owns. */
void mp_hip_close(mp_ctx* ctx);

/* Fill mp_device_info (reference DeviceInfo, backend/mod.rs:32-43). */
int mp_hip_device_info(mp_ctx* ctx, mp_device_info* out);

/* Price one request on the GPU, blocking (reference RuntimeBackend::execute,
 * backend/mod.rs:208-212, and the strict native entry
 * execute_apple_metal_native, backend/metal.rs:142-174: never falls back to
 * the CPU). */
int mp_hip_price(mp_ctx* ctx, const mp_request* req, mp_result* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code that must not call hipMalloc, such as work recorded under HIP
 * graph capture or tight serving loops): mp_hip_price for one request of the
 * GBM family (European, Asian, down-and-out, lookback; stepwise or terminal),
 * Heston or Merton with pseudorandom (Philox) draws, any technique, with the
 * same result bits.  Its device buffers -- the request record, the queue
 * counter and the per-tile FP64 partial sums -- live in a static device
 * array of MP_NOALLOC_SCRATCH_BYTES per device, set up by mp_hip_open, so the
 * call allocates nothing (the context's pinned result row receives the
 * result).  A request needing more than MP_NOALLOC_SCRATCH_BYTES (under 1 KiB
 * plus 64 bytes per path tile: about 131,000 tiles, e.g. 3e7 paths x 365
 * steps) fails with MP_E_OOM; Sobol, Halton/LHS, basket, American/Bermudan and
 * multilevel requests fail with MP_E_UNSUPPORTED.  Every call on a device
 * uses that device's one array: calls on one device from several threads of
 * the process are serialized by the library. */
#define MP_NOALLOC_SCRATCH_BYTES (8u << 20)
int mp_hip_price_noalloc(mp_ctx* ctx, const mp_request* req, mp_result* out);

/* Price n independent requests in one persistent launch (work queue of
 * (request, path-tile) items); results are bit-identical to pricing each
 * request alone.  No reference counterpart: the reference executes one
 * input per call (backend/mod.rs:208-212). */
int mp_hip_price_batch(mp_ctx* ctx, const mp_request* reqs, size_t n, mp_result* out);

/* Prepared batch: requests validated, packed and uploaded once (descriptors,
 * tile map and queue order resident in HBM), optionally captured as a
 * hipGraph (queue reset + persistent kernel + finalize) and replayed on its
 * own stream.  launch only enqueues; sync waits; results waits and reads the
 * result rows (which the finalize kernel writes to pinned host memory).  GBM,
 * Heston and Merton pseudorandom requests of one precision.  No reference
 * counterpart (see mp_hip_price_batch). */
typedef struct mp_batch mp_batch;
int mp_hip_batch_prepare(mp_ctx* ctx, const mp_request* reqs, size_t n, int use_graph,
                         mp_batch** out);
int mp_hip_batch_launch(mp_batch* batch);
int mp_hip_batch_sync(mp_batch* batch);
int mp_hip_batch_results(mp_batch* batch, mp_result* out);
int mp_hip_batch_info(mp_batch* batch, uint64_t* n_tiles, uint64_t* path_steps, int* grid,
                      int* block);
void mp_hip_batch_free(mp_batch* batch);

/* Plain FP64 comparator: one launch per request, one lane per path, FP64
 * normals and arithmetic, per-path payoffs written to HBM and reduced by a
 * second kernel -- the shape of the reference's CUDA kernel contract
 * (backend/cuda.rs:247-310, kernels/european_call_stepwise_v1.cu).
 * GBM European/Asian requests only. */
int mp_hip_price_plain_fp64(mp_ctx* ctx, const mp_request* req, mp_result* out);

/* The reference's CUDA kernel contract for the stepwise European call,
 * binding for binding (crates/mc-core/src/backend/cuda.rs:247-310,
 * first_cuda_kernel_contract; kernel body
 * crates/mc-core/src/backend/kernels/european_call_stepwise_v1.cu:1-28):
 * normals (binding 0, FP64, n_paths * n_steps values, path-major: path p's
 * step s at p * n_steps + s), payoffs (binding 1, FP64, n_paths values: the
 * discounted payoff (S_T - K) * discount if S_T > K, else 0, with
 * log S_T = log_s0 + sum over steps of (drift_dt + vol_dt * z) in step
 * order, each product rounded), and the scalars n_paths, n_steps, log_s0,
 * strike, drift_dt, vol_dt, discount; one logical thread per path, 256 per
 * workgroup.  Both buffers are host memory: normals are copied to the
 * device, payoffs back.  At most 2^31 normals.  mp_hip_last_timing reports
 * the kernel's time (main_kernel_ms). */
int mp_hip_european_call_stepwise_v1(mp_ctx* ctx, const double* normals, double* payoffs,
                                     int32_t n_paths, int32_t n_steps, double log_s0, double strike,
                                     double drift_dt, double vol_dt, double discount);

/* Parity hook: run the request's path arithmetic on caller-supplied normal
 * draws instead of the engine's generator ("identical normal buffers").
 * normals are step-major: normals[(step * F + f) * n_paths + path] with F
 * factors (1 for GBM, 2 for Heston as (z_var, z_ind), n_assets for a
 * basket).  For Merton, uniforms and jump_normals (same layout, F = 1) carry
 * the Poisson uniforms and jump-size normals.  Writes terminal values (basket
 * value for a basket) and undiscounted payoffs per path for strike 0.
 * Reference counterpart: the same-normals test hooks
 * *_from_f32_normals (runtime/cpu.rs:5434-5652). */
/* Parity hook: price a request (any technique) from caller-supplied draws
 * laid out as for mp_hip_paths_from_normals, with n_paths units (pairs for an
 * antithetic request: ceil(n_paths / 2); the pair's second path uses -z).
 * Reference counterpart: the technique pricers of runtime/cpu.rs:7785-7909
 * run on one normal stream. */
int mp_hip_price_from_normals(mp_ctx* ctx, const mp_request* req, const double* normals,
                              const double* uniforms, const double* jump_normals, mp_result* out);

int mp_hip_paths_from_normals(mp_ctx* ctx, const mp_request* req, const double* normals,
                              const double* uniforms, const double* jump_normals,
                              double* terminal_out, double* payoff_out);

/* Greeks of n requests.  out holds MP_MAX_GREEKS estimates per request
 * (row i starts at out + i * MP_MAX_GREEKS), n_estimates[i] how many were
 * written, base[i] the request priced as given.
 *
 * Bump-and-revalue (european_call_bump_greeks_cpu and the Asian, barrier,
 * lookback, basket and Heston arms, runtime/cpu.rs:4036-4668, with
 * push_bump_estimate :4938-4992 and bump_spot/volatility/rate/maturity
 * :4994-5008): every request's up and down legs are priced with its own seed
 * and contract id -- common random numbers -- and all legs of all requests
 * run in one persistent work-queue launch (baskets in their own kernel).
 * GBM family: Delta, Vega, Rho, Theta; Heston: the same with Vega on
 * sqrt(v0); basket: Delta, Delta2, Vega, Vega2 (assets 0 and 1), Rho, Theta.
 *
 * Pathwise / likelihood ratio (european_call_pathwise_greeks_cpu,
 * european_call_likelihood_ratio_greeks_cpu, :4670-4857): European calls
 * only, from the terminal draws of the request's Philox stream whatever its
 * method, technique or sampling (the reference likewise uses plain terminal
 * shocks); pathwise gives Delta, Vega, Rho, Theta, likelihood ratio Delta,
 * Vega, Rho.  Other workloads: MP_E_UNSUPPORTED (as the reference reports). */
int mp_hip_greeks(mp_ctx* ctx, const mp_request* reqs, size_t n, int32_t estimator,
                  mp_result* base, mp_greek_estimate* out, int32_t* n_estimates);

/* Parity hook: the pathwise / likelihood-ratio Greeks of one European call
 * from caller normals (one terminal draw per path, n_paths doubles). */
int mp_hip_greeks_from_normals(mp_ctx* ctx, const mp_request* req, int32_t estimator,
                               const double* normals, mp_result* base, mp_greek_estimate* out,
                               int32_t* n_estimates);

/* Multilevel Asian price (arithmetic_asian_call_price_mlmc_cpu,
 * runtime/cpu.rs:3283-3293, single estimator :3612-3661, replicated
 * structured estimator :3663-3751).  Every (replicate, level) runs in one
 * persistent work-queue launch; level terms are summed in FP64. */
int mp_hip_mlmc_price(mp_ctx* ctx, const mp_mlmc_config* cfg, mp_mlmc_result* out);

/* tune_arithmetic_asian_mlmc_allocation_cpu (cpu.rs:3295-3350): pilot every
 * level with pilot_paths paths (one launch), then split target_step_updates
 * in proportion to sqrt(variance / cost), at least two paths per level. */
int mp_hip_mlmc_tune_allocation(mp_ctx* ctx, const mp_mlmc_config* cfg, uint64_t target_step_updates,
                                uint64_t pilot_paths, mp_mlmc_allocation* out);

/* solve_arithmetic_asian_mlmc_tolerance_cpu (cpu.rs:3352-3423): paths per
 * level for a target standard error, clamped to [min, max] step updates. */
int mp_hip_mlmc_solve_tolerance(mp_ctx* ctx, const mp_mlmc_config* cfg,
                                const mp_mlmc_tolerance* tolerance, mp_mlmc_tolerance_plan* out);

/* Parity hook: (sum, sum of squares) of one level's discounted terms from
 * caller normals, step-major (fine_steps x paths). */
int mp_hip_mlmc_level_from_normals(mp_ctx* ctx, const mp_mlmc_config* cfg, uint32_t level,
                                   uint64_t paths, const double* normals, double* sum_sq);

/* Price one request and also return, per path, the terminal value (basket
 * value for a basket) and the undiscounted payoff for strike 0, from the
 * engine's own draws (Philox or Sobol).  Same kernels as mp_hip_price.  An
 * antithetic request fills the first ceil(n/2) entries (the first path of
 * each pair) and zeroes the rest. */
int mp_hip_price_with_paths(mp_ctx* ctx, const mp_request* req, mp_result* out,
                            double* terminal_out, double* payoff_out);

/* Parity hook for LSM: price an American/Bermudan put from caller-supplied
 * normals (step-major, n_steps x n_paths); coefficients_out (n_steps x 4,
 * may be NULL) receives the regression coefficients per date (zero where a
 * date was skipped) and exercise_out (n_paths, may be NULL) the exercise
 * step per path.  Reference: price_put_lsm_cpu, runtime/cpu.rs:3813-3945. */
int mp_hip_lsm_from_normals(mp_ctx* ctx, const mp_request* req, const double* normals,
                            mp_result* out, double* coefficients_out, uint32_t* exercise_out);

/* Debug/parity hook: the engine's own draws for paths [path0, path0+n_paths)
 * of a request, written path-major ([path][n_draws]).  kind 0: Philox
 * normals of stream `stream`; kind 1: Philox uniforms; kind 2: scrambled
 * Sobol standard normals of the point (n_draws dims, reference block/seed
 * rules); kind 3: raw scrambled Sobol 32-bit integers as doubles. */
int mp_hip_draws(mp_ctx* ctx, const mp_request* req, int kind, uint32_t stream, uint64_t path0,
                 uint64_t n_paths, uint64_t n_draws, double* out);

/* HBM bytes a request needs while in flight (reference
 * estimate_gpu_bytes_per_path / plan_gpu_chunking, backend/mod.rs:372-405). */
int mp_hip_request_footprint(const mp_request* req, uint64_t* bytes);

/* ---- serving: continuous batching on one device -------------------------
 * Requests are admitted by HBM footprint as slots free up; admitted work is
 * packed into persistent launches on up to `max_inflight` concurrent
 * streams while later requests keep arriving.  No reference counterpart
 * (the reference is synchronous; production.py:440-481).
 *
 * Admission: every request, the first of a launch included, is admitted
 * only while the device bytes the workers hold in their arenas plus the
 * reservations of the launches in flight (footprint + 25% + 1 MiB each) stay
 * within hbm_budget_bytes.  When the oldest pending request does not fit,
 * idle workers free their arenas and the request waits for running launches
 * to finish; a request whose reservation alone exceeds the budget is
 * rejected by mp_hip_submit with MP_E_OOM. */
typedef struct mp_serve_config {
    uint64_t hbm_budget_bytes;   /* admission budget (0: 75% of free memory) */
    int32_t max_inflight;        /* concurrent launches in flight (streams), 1..8 */
    int32_t max_batch;           /* requests per launch, upper bound */
} mp_serve_config;

typedef struct mp_server mp_server;

int mp_hip_server_open(mp_ctx* ctx, const mp_serve_config* cfg, mp_server** out);
void mp_hip_server_close(mp_server* srv);
/* Queue a request (copied); returns a ticket. */
int mp_hip_submit(mp_server* srv, const mp_request* req, uint64_t* ticket);
/* Non-blocking: *done = 1 and *out filled once the ticket has finished. */
int mp_hip_poll(mp_server* srv, uint64_t ticket, mp_result* out, int* done);
/* Blocking wait for a ticket. */
int mp_hip_wait(mp_server* srv, uint64_t ticket, mp_result* out);
/* Cancel a ticket that has not been launched yet. */
int mp_hip_cancel(mp_server* srv, uint64_t ticket);

/* Replay an arrival trace in real time through a server and report latency
 * (arrival -> result) per request.  arrival_s[i] is seconds from start. */
typedef struct mp_serve_stats {
    uint64_t completed;
    double wall_s;
    double requests_per_s;
    double path_steps_per_s;
    double p50_ms;
    double p99_ms;
    double max_ms;
    uint64_t launches;
    uint64_t max_admitted_bytes;
} mp_serve_stats;

int mp_hip_serve_trace(mp_ctx* ctx, const mp_serve_config* cfg, const mp_request* reqs,
                       const double* arrival_s, size_t n, mp_result* results,
                       double* latency_ms, mp_serve_stats* stats);

/* Counters of a live server: requests completed, launches, and the largest
 * number of device bytes held or reserved by its workers at any admission
 * or completion (max_admitted_bytes; never above hbm_budget_bytes).  The
 * latency and rate fields are left 0. */
int mp_hip_server_stats(mp_server* srv, mp_serve_stats* stats);

/* ---- tuning --------------------------------------------------------------
 * Kernel launch parameters: block size, paths interleaved per lane, paths
 * per work-queue tile (fixed by the request, reported only) and the step
 * unroll: 4-step draw blocks per iteration of the step loop (1, 2 or 4; 0
 * in mp_hip_set_tuning means 1), which selects compiled variants of the
 * FP32 GBM-family Philox body (the batching workload) and is ignored by the
 * other bodies.  mp_hip_autotune times the candidates on a sample of the
 * request and keeps the fastest per (model, payoff, steps bucket,
 * precision).  No choice changes a result: tiles and the order of every
 * sum are fixed by the request. */
typedef struct mp_tuning {
    int32_t block_size;
    int32_t paths_per_lane;
    int32_t tile_paths;
    int32_t step_unroll;
} mp_tuning;

int mp_hip_autotune(mp_ctx* ctx, const mp_request* req, mp_tuning* best, double* best_ms);
int mp_hip_get_tuning(mp_ctx* ctx, const mp_request* req, mp_tuning* out);
int mp_hip_set_tuning(mp_ctx* ctx, const mp_request* req, const mp_tuning* t);

/* Time of the last mp_hip_price / mp_hip_price_batch kernel region measured
 * with HIP events on the context stream, per kernel role. */
typedef struct mp_timing {
    double total_ms;        /* whole request on the stream */
    double main_kernel_ms;  /* the path-simulation kernel(s) */
    double finalize_ms;
    uint64_t main_launches;
} mp_timing;

int mp_hip_last_timing(mp_ctx* ctx, mp_timing* out);

/* Thread-local message for the last failing call on this thread; valid until
 * the next call. */
const char* mp_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MONTEPATH_HIP_H */
