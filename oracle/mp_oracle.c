/*
 * mp_oracle.c -- CPU oracle for the montepath Monte Carlo path engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the reference
 * CPU algorithms of montepath's mc-core (crates/mc-core/src/runtime/cpu.rs)
 * so that the HIP engine can be checked against them.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Bit-exactness notes:
 *   - Rust's f64 ln/exp/sin/cos call the platform libm; on Linux that is
 *     glibc, which this file also calls, so the scalar stream is the same.
 *   - Rust never contracts a*b+c into an FMA; this file must be compiled
 *     with -ffp-contract=off (oracle/Makefile does) to keep that rounding.
 *   - Rust's f64::powi lowers to compiler-rt's __powidf2 (binary
 *     exponentiation); powi_ref() below restates that algorithm.
 *
 * Functions marked "extension" have no counterpart in the reference
 * (stepped Merton, n-asset basket, two-factor Sobol/Brownian-bridge Heston,
 * Philox draws, the weighted Laguerre basis); they define the engine's semantics for those cases and
 * are pinned by closed forms and invariants in tests/.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <unistd.h>

#define MPO_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* RNG: MonteCarloRng, cpu.rs:2955-3007                                */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t state;
    int has_cached;
    double cached;
} mpo_rng;

static void rng_init(mpo_rng* g, uint64_t seed) {
    /* cpu.rs:2962-2967 -- a zero seed is replaced by the golden-ratio word */
    g->state = seed == 0 ? 0x9E3779B97F4A7C15ULL : seed;
    g->has_cached = 0;
    g->cached = 0.0;
}

static uint64_t rng_next_u64(mpo_rng* g) {
    /* cpu.rs:2975-2983 -- xorshift64* */
    uint64_t x = g->state;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    g->state = x;
    return x * 0x2545F4914F6CDD1DULL;
}

static double rng_open01(mpo_rng* g) {
    /* cpu.rs:2985-2990 -- top 53 bits, floored at f64::MIN_POSITIVE */
    uint64_t raw = rng_next_u64(g) >> 11;
    double v = (double)raw * (1.0 / (double)(1ULL << 53));
    return v > DBL_MIN ? v : DBL_MIN;
}

static double rng_normal(mpo_rng* g) {
    /* cpu.rs:2992-3006 -- Box-Muller, returns r*cos first, caches r*sin */
    if (g->has_cached) {
        g->has_cached = 0;
        return g->cached;
    }
    double u1 = rng_open01(g);
    double u2 = rng_open01(g);
    double radius = sqrt(-2.0 * log(u1));
    double theta = 2.0 * M_PI * u2;
    double z0 = radius * cos(theta);
    double z1 = radius * sin(theta);
    g->cached = z1;
    g->has_cached = 1;
    return z0;
}

/* cpu.rs:8928-8932 */
MPO_EXPORT uint64_t mpo_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* cpu.rs:8922-8924 (and backend/mod.rs:713-715, stable_fallback_seed) */
MPO_EXPORT uint64_t mpo_derive_chunk_seed(uint64_t base, uint64_t idx) {
    return mpo_splitmix64(base + idx * 0x9E3779B97F4A7C15ULL);
}

/* Raw stream access for tests: n successive standard_normal() draws. */
MPO_EXPORT void mpo_rng_normals(uint64_t seed, uint64_t n, double* out) {
    mpo_rng g;
    rng_init(&g, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = rng_normal(&g);
}

MPO_EXPORT void mpo_rng_u64(uint64_t seed, uint64_t n, uint64_t* out) {
    mpo_rng g;
    rng_init(&g, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = rng_next_u64(&g);
}

/* compiler-rt __powidf2, the lowering of Rust's f64::powi */
static double powi_ref(double a, int b) {
    const int recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

static double fmax_ref(double a, double b) { return a > b ? a : (b == b ? b : a); }

/* ------------------------------------------------------------------ */
/* Summaries: cpu.rs:7641-7665                                         */
/* ------------------------------------------------------------------ */
typedef struct {
    double price;
    double stderr_;
} mpo_result;

MPO_EXPORT mpo_result mpo_summarize(uint64_t n_paths, double sum, double sq_sum) {
    /* population variance, one pass, clamped at zero (cpu.rs:7641-7652) */
    double n = (double)n_paths;
    double price = sum / n;
    double variance = (sq_sum / n) - (price * price);
    mpo_result r;
    r.price = price;
    r.stderr_ = sqrt(variance > 0.0 ? variance : 0.0) / sqrt(n);
    return r;
}

/* ------------------------------------------------------------------ */
/* GBM path family (standard technique, pseudorandom)                  */
/* ------------------------------------------------------------------ */
enum { MPO_EUROPEAN = 0, MPO_ASIAN = 1, MPO_DOWN_AND_OUT = 2, MPO_LOOKBACK = 3 };

typedef struct {
    double s0, k, r, sigma, t, barrier;
    uint64_t n_paths, n_steps, seed, n_threads;
} mpo_gbm_cfg;

typedef struct {
    int payoff;
    uint64_t seed, n_paths, n_steps;
    double s0, k, barrier, drift_dt, vol_dt, discount;
    double sum, sq_sum;
} gbm_chunk;

/* One chunk of the stepwise GBM family: European cpu.rs:6443-6471,
 * Asian cpu.rs:6561-6591, down-and-out cpu.rs:7973-8009,
 * lookback cpu.rs:8324-8353. */
static void* gbm_chunk_run(void* arg) {
    gbm_chunk* c = (gbm_chunk*)arg;
    mpo_rng g;
    rng_init(&g, c->seed);
    double sum = 0.0, sq = 0.0;
    const double log_s0 = log(c->s0);
    for (uint64_t p = 0; p < c->n_paths; ++p) {
        double log_s = log_s0;
        double payoff;
        if (c->payoff == MPO_EUROPEAN) {
            for (uint64_t s = 0; s < c->n_steps; ++s) {
                double z = rng_normal(&g);
                log_s += c->drift_dt + c->vol_dt * z;
            }
            payoff = fmax_ref(exp(log_s) - c->k, 0.0) * c->discount;
        } else if (c->payoff == MPO_ASIAN) {
            double acc = 0.0;
            for (uint64_t s = 0; s < c->n_steps; ++s) {
                double z = rng_normal(&g);
                log_s += c->drift_dt + c->vol_dt * z;
                acc += exp(log_s);
            }
            double avg = acc / (double)c->n_steps;
            payoff = fmax_ref(avg - c->k, 0.0) * c->discount;
        } else if (c->payoff == MPO_DOWN_AND_OUT) {
            int knocked = c->s0 <= c->barrier;
            for (uint64_t s = 0; s < c->n_steps; ++s) {
                double z = rng_normal(&g);
                log_s += c->drift_dt + c->vol_dt * z;
                if (exp(log_s) <= c->barrier) knocked = 1;
            }
            payoff = knocked ? 0.0 : fmax_ref(exp(log_s) - c->k, 0.0) * c->discount;
        } else {
            double max_s = c->s0;
            for (uint64_t s = 0; s < c->n_steps; ++s) {
                double z = rng_normal(&g);
                log_s += c->drift_dt + c->vol_dt * z;
                max_s = fmax_ref(max_s, exp(log_s));
            }
            payoff = fmax_ref(max_s - c->k, 0.0) * c->discount;
        }
        sum += payoff;
        sq += payoff * payoff;
    }
    c->sum = sum;
    c->sq_sum = sq;
    return NULL;
}

static uint64_t resolved_threads(uint64_t requested) {
    /* cpu.rs:8912-8920 -- 0 means all cores */
    if (requested > 0) return requested;
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (uint64_t)n : 1;
}

/* Chunked execution exactly as the reference: chunk i has
 * base + (i < rem) paths, seeded derive_chunk_seed(seed, i), and is
 * reduced in spawn order (cpu.rs:6232-6276, single-chunk rule :3245). */
typedef void* (*chunk_fn)(void*);

static void run_chunks(void* chunks, size_t stride, uint64_t count, chunk_fn fn) {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * count);
    for (uint64_t i = 0; i < count; ++i)
        pthread_create(&th[i], NULL, fn, (char*)chunks + i * stride);
    for (uint64_t i = 0; i < count; ++i) pthread_join(th[i], NULL);
    free(th);
}

/* european_call_price_mc_cpu_stepwise cpu.rs:3228-3262 and siblings
 * (Asian cpu.rs:7885-7909, down-and-out, lookback): standard technique. */
MPO_EXPORT mpo_result mpo_gbm_stepwise_price(const mpo_gbm_cfg* cfg, int payoff) {
    const double dt = cfg->t / (double)cfg->n_steps;
    const double drift_dt = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * dt;
    const double vol_dt = cfg->sigma * sqrt(dt);
    const double discount = exp(-cfg->r * cfg->t);
    uint64_t threads = resolved_threads(cfg->n_threads);
    uint64_t count = (threads <= 1 || cfg->n_paths < threads * 2000) ? 1 : threads;
    gbm_chunk* ch = (gbm_chunk*)calloc(count, sizeof(gbm_chunk));
    uint64_t base = cfg->n_paths / count, rem = cfg->n_paths % count;
    for (uint64_t i = 0; i < count; ++i) {
        ch[i].payoff = payoff;
        ch[i].seed = count == 1 ? cfg->seed : mpo_derive_chunk_seed(cfg->seed, i);
        ch[i].n_paths = count == 1 ? cfg->n_paths : base + (i < rem ? 1 : 0);
        ch[i].n_steps = cfg->n_steps;
        ch[i].s0 = cfg->s0;
        ch[i].k = cfg->k;
        ch[i].barrier = cfg->barrier;
        ch[i].drift_dt = drift_dt;
        ch[i].vol_dt = vol_dt;
        ch[i].discount = discount;
    }
    if (count == 1)
        gbm_chunk_run(&ch[0]);
    else
        run_chunks(ch, sizeof(gbm_chunk), count, gbm_chunk_run);
    double sum = 0.0, sq = 0.0;
    for (uint64_t i = 0; i < count; ++i) {
        sum += ch[i].sum;
        sq += ch[i].sq_sum;
    }
    free(ch);
    return mpo_summarize(cfg->n_paths, sum, sq);
}

typedef struct {
    uint64_t seed, n_paths;
    double s0, k, drift_t, vol_t, discount, sum, sq_sum;
} term_chunk;

/* simulate_terminal_chunk cpu.rs:6312-6353 (standard) with
 * european_call_payoff_from_shock cpu.rs:7773-7783 */
static void* term_chunk_run(void* arg) {
    term_chunk* c = (term_chunk*)arg;
    mpo_rng g;
    rng_init(&g, c->seed);
    double sum = 0.0, sq = 0.0;
    for (uint64_t p = 0; p < c->n_paths; ++p) {
        double z = rng_normal(&g);
        double s_t = c->s0 * exp(c->drift_t + c->vol_t * z);
        double payoff = fmax_ref(s_t - c->k, 0.0) * c->discount;
        sum += payoff;
        sq += payoff * payoff;
    }
    c->sum = sum;
    c->sq_sum = sq;
    return NULL;
}

/* european_call_price_mc_cpu_terminal cpu.rs:3191-3226 */
MPO_EXPORT mpo_result mpo_european_terminal_price(const mpo_gbm_cfg* cfg) {
    const double drift_t = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * cfg->t;
    const double vol_t = cfg->sigma * sqrt(cfg->t);
    const double discount = exp(-cfg->r * cfg->t);
    uint64_t threads = resolved_threads(cfg->n_threads);
    uint64_t count = (threads <= 1 || cfg->n_paths < threads * 2000) ? 1 : threads;
    term_chunk* ch = (term_chunk*)calloc(count, sizeof(term_chunk));
    uint64_t base = cfg->n_paths / count, rem = cfg->n_paths % count;
    for (uint64_t i = 0; i < count; ++i) {
        ch[i].seed = count == 1 ? cfg->seed : mpo_derive_chunk_seed(cfg->seed, i);
        ch[i].n_paths = count == 1 ? cfg->n_paths : base + (i < rem ? 1 : 0);
        ch[i].s0 = cfg->s0;
        ch[i].k = cfg->k;
        ch[i].drift_t = drift_t;
        ch[i].vol_t = vol_t;
        ch[i].discount = discount;
    }
    if (count == 1)
        term_chunk_run(&ch[0]);
    else
        run_chunks(ch, sizeof(term_chunk), count, term_chunk_run);
    double sum = 0.0, sq = 0.0;
    for (uint64_t i = 0; i < count; ++i) {
        sum += ch[i].sum;
        sq += ch[i].sq_sum;
    }
    free(ch);
    return mpo_summarize(cfg->n_paths, sum, sq);
}

/* Per-path GBM arithmetic from a caller-supplied normal buffer laid out
 * path-major ([path][step], the order the reference stream is consumed in).
 * Restates the inner loops of cpu.rs:6459-6471 / 6578-6587 / 7990-8000 /
 * 8340-8349.  Writes S_T and the discounted payoff per path. */
MPO_EXPORT void mpo_gbm_paths_from_normals(const mpo_gbm_cfg* cfg, int payoff,
                                           const double* normals, double* terminal_out,
                                           double* payoff_out) {
    const double dt = cfg->t / (double)cfg->n_steps;
    const double drift_dt = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * dt;
    const double vol_dt = cfg->sigma * sqrt(dt);
    const double discount = exp(-cfg->r * cfg->t);
    const double log_s0 = log(cfg->s0);
    for (uint64_t p = 0; p < cfg->n_paths; ++p) {
        const double* z = normals + p * cfg->n_steps;
        double log_s = log_s0, acc = 0.0, max_s = cfg->s0;
        int knocked = cfg->s0 <= cfg->barrier;
        for (uint64_t s = 0; s < cfg->n_steps; ++s) {
            log_s += drift_dt + vol_dt * z[s];
            if (payoff == MPO_ASIAN) acc += exp(log_s);
            if (payoff == MPO_DOWN_AND_OUT && exp(log_s) <= cfg->barrier) knocked = 1;
            if (payoff == MPO_LOOKBACK) max_s = fmax_ref(max_s, exp(log_s));
        }
        double s_t = exp(log_s);
        double pay;
        if (payoff == MPO_ASIAN)
            pay = fmax_ref(acc / (double)cfg->n_steps - cfg->k, 0.0) * discount;
        else if (payoff == MPO_DOWN_AND_OUT)
            pay = knocked ? 0.0 : fmax_ref(s_t - cfg->k, 0.0) * discount;
        else if (payoff == MPO_LOOKBACK)
            pay = fmax_ref(max_s - cfg->k, 0.0) * discount;
        else
            pay = fmax_ref(s_t - cfg->k, 0.0) * discount;
        if (terminal_out) terminal_out[p] = s_t;
        if (payoff_out) payoff_out[p] = pay;
    }
}

/* ------------------------------------------------------------------ */
/* Variance-reduction techniques for the GBM family                    */
/* Antithetic pairs: stepwise European cpu.rs:7807-7835 (one stream),  */
/* Asian :6982-7017 / :7912-7938, down-and-out :8048-8100 / :8255-8288,*/
/* lookback :8355-8390 / :8569-8600; control variate: moments          */
/* cpu.rs:6202-6230, summary :7738-7771, chunks European :6532-6559,   */
/* Asian :6593-6623, down-and-out :8011-8046, lookback :8392-8420.      */
/* ------------------------------------------------------------------ */
enum { MPO_STANDARD = 0, MPO_ANTITHETIC = 1, MPO_CONTROL_VARIATE = 2 };

typedef struct {
    uint64_t count;
    double payoff_sum, payoff_sq_sum, control_sum, control_sq_sum, cross_sum;
} mpo_cv_moments;

static void cv_record(mpo_cv_moments* m, double payoff, double control) {
    m->count += 1;
    m->payoff_sum += payoff;
    m->payoff_sq_sum += payoff * payoff;
    m->control_sum += control;
    m->control_sq_sum += control * control;
    m->cross_sum += payoff * control;
}

static void cv_merge(mpo_cv_moments* m, const mpo_cv_moments* o) {
    m->count += o->count;
    m->payoff_sum += o->payoff_sum;
    m->payoff_sq_sum += o->payoff_sq_sum;
    m->control_sum += o->control_sum;
    m->control_sq_sum += o->control_sq_sum;
    m->cross_sum += o->cross_sum;
}

MPO_EXPORT void mpo_cv_record(mpo_cv_moments* m, double payoff, double control) {
    cv_record(m, payoff, control);
}

/* summarize_control_variate cpu.rs:7738-7771 */
MPO_EXPORT mpo_result mpo_summarize_cv(const mpo_cv_moments* m, double expectation) {
    const double n = (double)m->count;
    const double payoff_mean = m->payoff_sum / n;
    const double control_mean = m->control_sum / n;
    const double control_var = (m->control_sq_sum / n) - (control_mean * control_mean);
    if (control_var <= DBL_EPSILON) return mpo_summarize(m->count, m->payoff_sum, m->payoff_sq_sum);
    const double cov = (m->cross_sum / n) - (payoff_mean * control_mean);
    const double beta = cov / control_var;
    const double adjusted_mean = payoff_mean - beta * (control_mean - expectation);
    const double adjusted_sq_mean = (m->payoff_sq_sum / n)
        - (2.0 * beta) * ((m->cross_sum / n) - expectation * payoff_mean)
        + (beta * beta) * ((m->control_sq_sum / n) - (2.0 * expectation * control_mean)
                           + (expectation * expectation));
    double adjusted_var = adjusted_sq_mean - (adjusted_mean * adjusted_mean);
    if (!(adjusted_var > 0.0)) adjusted_var = 0.0; /* f64::max(0.0) also maps NaN to 0 */
    mpo_result r;
    r.price = adjusted_mean;
    r.stderr_ = sqrt(adjusted_var) / sqrt(n);
    return r;
}

/* A normal source: the reference stream, or a caller buffer read in order. */
typedef struct {
    mpo_rng g;
    const double* buf;
    uint64_t pos;
} normal_src;

static double src_normal(normal_src* s) {
    if (s->buf) return s->buf[s->pos++];
    return rng_normal(&s->g);
}

typedef struct {
    double log_s, acc, max_s;
    int knocked;
} gbm_state;

static void gbm_state_init(gbm_state* st, double log_s0, double s0, double barrier) {
    st->log_s = log_s0;
    st->acc = 0.0;
    st->max_s = s0;
    st->knocked = s0 <= barrier;
}

/* log_s += shock, then the payoff's observation of exp(log_s) */
static void gbm_state_step(gbm_state* st, int payoff, double shock, double barrier) {
    st->log_s += shock;
    if (payoff == MPO_ASIAN) st->acc += exp(st->log_s);
    if (payoff == MPO_DOWN_AND_OUT && exp(st->log_s) <= barrier) st->knocked = 1;
    if (payoff == MPO_LOOKBACK) st->max_s = fmax_ref(st->max_s, exp(st->log_s));
}

static double gbm_state_payoff(const gbm_state* st, int payoff, double k, uint64_t n_steps,
                               double discount) {
    if (payoff == MPO_ASIAN) return fmax_ref(st->acc / (double)n_steps - k, 0.0) * discount;
    if (payoff == MPO_LOOKBACK) return fmax_ref(st->max_s - k, 0.0) * discount;
    if (payoff == MPO_DOWN_AND_OUT && st->knocked) return 0.0;
    return fmax_ref(exp(st->log_s) - k, 0.0) * discount;
}

typedef struct {
    int payoff, technique;
    uint64_t seed, n_units, n_steps;
    double s0, k, barrier, drift_dt, vol_dt, discount;
    const double* normals; /* [unit][step] or NULL for the reference stream */
    double sum, sq_sum;
    mpo_cv_moments m;
} gbm_tech_chunk;

static void* gbm_tech_chunk_run(void* arg) {
    gbm_tech_chunk* c = (gbm_tech_chunk*)arg;
    normal_src src;
    memset(&src, 0, sizeof(src));
    src.buf = c->normals;
    if (!src.buf) rng_init(&src.g, c->seed);
    const double log_s0 = log(c->s0);
    double sum = 0.0, sq = 0.0;
    mpo_cv_moments m;
    memset(&m, 0, sizeof(m));
    for (uint64_t u = 0; u < c->n_units; ++u) {
        if (c->technique == MPO_ANTITHETIC) {
            gbm_state a, b;
            gbm_state_init(&a, log_s0, c->s0, c->barrier);
            gbm_state_init(&b, log_s0, c->s0, c->barrier);
            for (uint64_t s = 0; s < c->n_steps; ++s) {
                double z = src_normal(&src);
                gbm_state_step(&a, c->payoff, c->drift_dt + c->vol_dt * z, c->barrier);
                gbm_state_step(&b, c->payoff, c->drift_dt - c->vol_dt * z, c->barrier);
            }
            double pa = gbm_state_payoff(&a, c->payoff, c->k, c->n_steps, c->discount);
            double pb = gbm_state_payoff(&b, c->payoff, c->k, c->n_steps, c->discount);
            double block = 0.5 * (pa + pb);
            sum += block;
            sq += block * block;
        } else {
            gbm_state st;
            gbm_state_init(&st, log_s0, c->s0, c->barrier);
            for (uint64_t s = 0; s < c->n_steps; ++s) {
                double z = src_normal(&src);
                gbm_state_step(&st, c->payoff, c->drift_dt + c->vol_dt * z, c->barrier);
            }
            double s_t = exp(st.log_s);
            double pay = gbm_state_payoff(&st, c->payoff, c->k, c->n_steps, c->discount);
            cv_record(&m, pay, c->discount * s_t);
        }
    }
    c->sum = sum;
    c->sq_sum = sq;
    c->m = m;
    return NULL;
}

/* Stepwise GBM family with a technique.  European antithetic runs one
 * stream (cpu.rs:7807-7835); every other combination is chunked like the
 * standard pricers, over pairs (antithetic) or paths (control variate). */
static mpo_result gbm_tech(const mpo_gbm_cfg* cfg, int payoff, int technique,
                           const double* normals) {
    const double dt = cfg->t / (double)cfg->n_steps;
    const double drift_dt = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * dt;
    const double vol_dt = cfg->sigma * sqrt(dt);
    const double discount = exp(-cfg->r * cfg->t);
    const uint64_t units = technique == MPO_ANTITHETIC ? (cfg->n_paths + 1) / 2 : cfg->n_paths;
    uint64_t threads = resolved_threads(cfg->n_threads);
    uint64_t count = (threads <= 1 || units < threads * 2000) ? 1 : threads;
    if (technique == MPO_ANTITHETIC && payoff == MPO_EUROPEAN) count = 1;
    if (normals) count = 1;
    gbm_tech_chunk* ch = (gbm_tech_chunk*)calloc(count, sizeof(gbm_tech_chunk));
    uint64_t base = units / count, rem = units % count;
    for (uint64_t i = 0; i < count; ++i) {
        ch[i].payoff = payoff;
        ch[i].technique = technique;
        ch[i].seed = count == 1 ? cfg->seed : mpo_derive_chunk_seed(cfg->seed, i);
        ch[i].n_units = count == 1 ? units : base + (i < rem ? 1 : 0);
        ch[i].n_steps = cfg->n_steps;
        ch[i].s0 = cfg->s0;
        ch[i].k = cfg->k;
        ch[i].barrier = cfg->barrier;
        ch[i].drift_dt = drift_dt;
        ch[i].vol_dt = vol_dt;
        ch[i].discount = discount;
        ch[i].normals = normals;
    }
    if (count == 1)
        gbm_tech_chunk_run(&ch[0]);
    else
        run_chunks(ch, sizeof(gbm_tech_chunk), count, gbm_tech_chunk_run);
    mpo_result r;
    if (technique == MPO_ANTITHETIC) {
        double sum = 0.0, sq = 0.0;
        for (uint64_t i = 0; i < count; ++i) {
            sum += ch[i].sum;
            sq += ch[i].sq_sum;
        }
        r = mpo_summarize(units, sum, sq); /* summarize_block_estimates cpu.rs:7654-7665 */
    } else {
        mpo_cv_moments m;
        memset(&m, 0, sizeof(m));
        for (uint64_t i = 0; i < count; ++i) cv_merge(&m, &ch[i].m);
        /* standard (from caller normals only): the discounted payoff moments */
        r = technique == MPO_STANDARD ? mpo_summarize(units, m.payoff_sum, m.payoff_sq_sum)
                                      : mpo_summarize_cv(&m, cfg->s0);
    }
    free(ch);
    return r;
}

MPO_EXPORT mpo_result mpo_gbm_stepwise_tech_price(const mpo_gbm_cfg* cfg, int payoff, int technique) {
    if (technique == MPO_STANDARD) return mpo_gbm_stepwise_price(cfg, payoff);
    return gbm_tech(cfg, payoff, technique, NULL);
}

/* The same pricer on a caller buffer of normals laid out [unit][step]
 * (units = pairs for antithetic, paths for the control variate). */
MPO_EXPORT mpo_result mpo_gbm_tech_from_normals(const mpo_gbm_cfg* cfg, int payoff, int technique,
                                                const double* normals) {
    return gbm_tech(cfg, payoff, technique, normals);
}

typedef struct {
    int technique;
    uint64_t seed, n_units;
    double s0, k, drift_t, vol_t, discount;
    const double* normals;
    double sum, sq_sum;
    mpo_cv_moments m;
} term_tech_chunk;

/* terminal antithetic cpu.rs:7785-7805 (one stream) and control variate
 * chunk cpu.rs:6509-6530, with european_call_payoff_from_shock :7773-7783 */
static void* term_tech_chunk_run(void* arg) {
    term_tech_chunk* c = (term_tech_chunk*)arg;
    normal_src src;
    memset(&src, 0, sizeof(src));
    src.buf = c->normals;
    if (!src.buf) rng_init(&src.g, c->seed);
    double sum = 0.0, sq = 0.0;
    mpo_cv_moments m;
    memset(&m, 0, sizeof(m));
    for (uint64_t u = 0; u < c->n_units; ++u) {
        double z = src_normal(&src);
        if (c->technique == MPO_ANTITHETIC) {
            double pa = fmax_ref(c->s0 * exp(c->drift_t + c->vol_t * z) - c->k, 0.0) * c->discount;
            double pb = fmax_ref(c->s0 * exp(c->drift_t + c->vol_t * -z) - c->k, 0.0) * c->discount;
            double block = 0.5 * (pa + pb);
            sum += block;
            sq += block * block;
        } else {
            double s_t = c->s0 * exp(c->drift_t + c->vol_t * z);
            cv_record(&m, fmax_ref(s_t - c->k, 0.0) * c->discount, c->discount * s_t);
        }
    }
    c->sum = sum;
    c->sq_sum = sq;
    c->m = m;
    return NULL;
}

static mpo_result term_tech(const mpo_gbm_cfg* cfg, int technique, const double* normals) {
    const double drift_t = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * cfg->t;
    const double vol_t = cfg->sigma * sqrt(cfg->t);
    const double discount = exp(-cfg->r * cfg->t);
    const uint64_t units = technique == MPO_ANTITHETIC ? (cfg->n_paths + 1) / 2 : cfg->n_paths;
    uint64_t threads = resolved_threads(cfg->n_threads);
    uint64_t count = (threads <= 1 || units < threads * 2000) ? 1 : threads;
    if (technique == MPO_ANTITHETIC || normals) count = 1;
    term_tech_chunk* ch = (term_tech_chunk*)calloc(count, sizeof(term_tech_chunk));
    uint64_t base = units / count, rem = units % count;
    for (uint64_t i = 0; i < count; ++i) {
        ch[i].technique = technique;
        ch[i].seed = count == 1 ? cfg->seed : mpo_derive_chunk_seed(cfg->seed, i);
        ch[i].n_units = count == 1 ? units : base + (i < rem ? 1 : 0);
        ch[i].s0 = cfg->s0;
        ch[i].k = cfg->k;
        ch[i].drift_t = drift_t;
        ch[i].vol_t = vol_t;
        ch[i].discount = discount;
        ch[i].normals = normals;
    }
    if (count == 1)
        term_tech_chunk_run(&ch[0]);
    else
        run_chunks(ch, sizeof(term_tech_chunk), count, term_tech_chunk_run);
    mpo_result r;
    if (technique == MPO_ANTITHETIC) {
        r = mpo_summarize(units, ch[0].sum, ch[0].sq_sum);
    } else {
        mpo_cv_moments m;
        memset(&m, 0, sizeof(m));
        for (uint64_t i = 0; i < count; ++i) cv_merge(&m, &ch[i].m);
        /* standard (from caller normals only): the discounted payoff moments */
        r = technique == MPO_STANDARD ? mpo_summarize(units, m.payoff_sum, m.payoff_sq_sum)
                                      : mpo_summarize_cv(&m, cfg->s0);
    }
    free(ch);
    return r;
}

MPO_EXPORT mpo_result mpo_european_terminal_tech_price(const mpo_gbm_cfg* cfg, int technique) {
    if (technique == MPO_STANDARD) return mpo_european_terminal_price(cfg);
    return term_tech(cfg, technique, NULL);
}

MPO_EXPORT mpo_result mpo_terminal_tech_from_normals(const mpo_gbm_cfg* cfg, int technique,
                                                     const double* normals) {
    return term_tech(cfg, technique, normals);
}

/* ------------------------------------------------------------------ */
/* Heston: cpu.rs:8632-8910                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    double s0, k, r, v0, kappa, theta, xi, rho, t;
    uint64_t n_paths, n_steps, seed, n_threads;
} mpo_heston_cfg;

enum { MPO_HESTON_REFERENCE_FLOOR = 0, MPO_HESTON_FULL_TRUNCATION = 1 };

typedef struct {
    double log_s0, v0, kappa, theta, xi, rho, rho_perp, dt, sqrt_dt, rate, discount, strike;
    uint64_t n_steps;
    int scheme;
} heston_params;

static heston_params make_heston_params(const mpo_heston_cfg* cfg, int scheme) {
    /* heston_params cpu.rs:8649-8665 */
    heston_params h;
    h.log_s0 = log(cfg->s0);
    h.strike = cfg->k;
    h.v0 = cfg->v0;
    h.kappa = cfg->kappa;
    h.theta = cfg->theta;
    h.xi = cfg->xi;
    h.rho = cfg->rho;
    double rp = 1.0 - cfg->rho * cfg->rho;
    h.rho_perp = sqrt(rp > 0.0 ? rp : 0.0);
    h.dt = cfg->t / (double)cfg->n_steps;
    h.sqrt_dt = sqrt(cfg->t / (double)cfg->n_steps);
    h.n_steps = cfg->n_steps;
    h.rate = cfg->r;
    h.discount = exp(-cfg->r * cfg->t);
    h.scheme = scheme;
    return h;
}

/* One Euler step.  scheme 0 restates heston_path_terminal cpu.rs:8674-8683
 * (variance floored at zero after every update); scheme 1 is the
 * full-truncation scheme of Lord, Koekkoek & van Dijk (2010): identical
 * except that the stored variance is not floored, only its positive part
 * enters drift and diffusion (extension). */
static inline void heston_step(const heston_params* h, double z_var, double z_ind, double* log_s,
                               double* variance) {
    double v = *variance;
    double v_pos = v > 0.0 ? v : 0.0;
    double z_price = h->rho * z_var + h->rho_perp * z_ind;
    *log_s += (h->rate - 0.5 * v_pos) * h->dt + sqrt(v_pos) * h->sqrt_dt * z_price;
    v = v + h->kappa * (h->theta - v_pos) * h->dt + h->xi * sqrt(v_pos) * h->sqrt_dt * z_var;
    if (h->scheme == MPO_HESTON_REFERENCE_FLOOR) v = v > 0.0 ? v : 0.0;
    *variance = v;
}

typedef struct {
    heston_params h;
    uint64_t seed, n_paths;
    double sum, sq_sum;
} heston_chunk;

/* simulate_heston_chunk cpu.rs:8705-8718 */
static void* heston_chunk_run(void* arg) {
    heston_chunk* c = (heston_chunk*)arg;
    mpo_rng g;
    rng_init(&g, c->seed);
    double sum = 0.0, sq = 0.0;
    for (uint64_t p = 0; p < c->n_paths; ++p) {
        double log_s = c->h.log_s0, v = c->h.v0;
        for (uint64_t s = 0; s < c->h.n_steps; ++s) {
            double z_var = 1.0 * rng_normal(&g);
            double z_ind = 1.0 * rng_normal(&g);
            heston_step(&c->h, z_var, z_ind, &log_s, &v);
        }
        double s_t = exp(log_s);
        double payoff = fmax_ref(s_t - c->h.strike, 0.0) * c->h.discount;
        sum += payoff;
        sq += payoff * payoff;
    }
    c->sum = sum;
    c->sq_sum = sq;
    return NULL;
}

/* heston_european_call_price_mc_standard cpu.rs:8869-8881 */
MPO_EXPORT mpo_result mpo_heston_price(const mpo_heston_cfg* cfg, int scheme) {
    heston_params h = make_heston_params(cfg, scheme);
    uint64_t threads = resolved_threads(cfg->n_threads);
    uint64_t count = (threads <= 1 || cfg->n_paths < threads * 2000) ? 1 : threads;
    heston_chunk* ch = (heston_chunk*)calloc(count, sizeof(heston_chunk));
    uint64_t base = cfg->n_paths / count, rem = cfg->n_paths % count;
    for (uint64_t i = 0; i < count; ++i) {
        ch[i].h = h;
        ch[i].seed = count == 1 ? cfg->seed : mpo_derive_chunk_seed(cfg->seed, i);
        ch[i].n_paths = count == 1 ? cfg->n_paths : base + (i < rem ? 1 : 0);
    }
    if (count == 1)
        heston_chunk_run(&ch[0]);
    else
        run_chunks(ch, sizeof(heston_chunk), count, heston_chunk_run);
    double sum = 0.0, sq = 0.0;
    for (uint64_t i = 0; i < count; ++i) {
        sum += ch[i].sum;
        sq += ch[i].sq_sum;
    }
    free(ch);
    return mpo_summarize(cfg->n_paths, sum, sq);
}

/* Per-path Heston from normals laid out [path][step][2] with (z_var, z_ind)
 * in the reference's draw order (cpu.rs:8672-8673). */
MPO_EXPORT void mpo_heston_paths_from_normals(const mpo_heston_cfg* cfg, int scheme,
                                              const double* normals, double* terminal_out,
                                              double* variance_out) {
    heston_params h = make_heston_params(cfg, scheme);
    for (uint64_t p = 0; p < cfg->n_paths; ++p) {
        const double* z = normals + p * cfg->n_steps * 2;
        double log_s = h.log_s0, v = h.v0;
        for (uint64_t s = 0; s < cfg->n_steps; ++s) heston_step(&h, z[2 * s], z[2 * s + 1], &log_s, &v);
        terminal_out[p] = exp(log_s);
        if (variance_out) variance_out[p] = v;
    }
}

/* Heston with a technique: antithetic chunk cpu.rs:8714-8762 (pairs, the
 * second path takes -z_var, -z_ind), control variate chunk cpu.rs:8764-8776
 * with heston_discounted_payoff :8688-8698 (control = S_T * discount);
 * drivers :8883-8910. */
typedef struct {
    heston_params h;
    int technique;
    uint64_t seed, n_units;
    const double* normals; /* [unit][step][2] or NULL for the reference stream */
    double sum, sq_sum;
    mpo_cv_moments m;
} heston_tech_chunk;

static void* heston_tech_chunk_run(void* arg) {
    heston_tech_chunk* c = (heston_tech_chunk*)arg;
    const heston_params* h = &c->h;
    normal_src src;
    memset(&src, 0, sizeof(src));
    src.buf = c->normals;
    if (!src.buf) rng_init(&src.g, c->seed);
    double sum = 0.0, sq = 0.0;
    mpo_cv_moments m;
    memset(&m, 0, sizeof(m));
    for (uint64_t u = 0; u < c->n_units; ++u) {
        if (c->technique == MPO_ANTITHETIC) {
            double log_a = h->log_s0, var_a = h->v0, log_b = h->log_s0, var_b = h->v0;
            for (uint64_t s = 0; s < h->n_steps; ++s) {
                double z_var = src_normal(&src);
                double z_ind = src_normal(&src);
                heston_step(h, z_var, z_ind, &log_a, &var_a);
                heston_step(h, -z_var, -z_ind, &log_b, &var_b);
            }
            double pa = fmax_ref(exp(log_a) - h->strike, 0.0) * h->discount;
            double pb = fmax_ref(exp(log_b) - h->strike, 0.0) * h->discount;
            double block = 0.5 * (pa + pb);
            sum += block;
            sq += block * block;
        } else {
            double log_s = h->log_s0, v = h->v0;
            for (uint64_t s = 0; s < h->n_steps; ++s) {
                double z_var = 1.0 * src_normal(&src);
                double z_ind = 1.0 * src_normal(&src);
                heston_step(h, z_var, z_ind, &log_s, &v);
            }
            double s_t = exp(log_s);
            cv_record(&m, fmax_ref(s_t - h->strike, 0.0) * h->discount, s_t * h->discount);
        }
    }
    c->sum = sum;
    c->sq_sum = sq;
    c->m = m;
    return NULL;
}

static mpo_result heston_tech(const mpo_heston_cfg* cfg, int scheme, int technique,
                              const double* normals) {
    heston_params h = make_heston_params(cfg, scheme);
    const uint64_t units = technique == MPO_ANTITHETIC ? (cfg->n_paths + 1) / 2 : cfg->n_paths;
    uint64_t threads = resolved_threads(cfg->n_threads);
    uint64_t count = (threads <= 1 || units < threads * 2000) ? 1 : threads;
    if (normals) count = 1;
    heston_tech_chunk* ch = (heston_tech_chunk*)calloc(count, sizeof(heston_tech_chunk));
    uint64_t base = units / count, rem = units % count;
    for (uint64_t i = 0; i < count; ++i) {
        ch[i].h = h;
        ch[i].technique = technique;
        ch[i].seed = count == 1 ? cfg->seed : mpo_derive_chunk_seed(cfg->seed, i);
        ch[i].n_units = count == 1 ? units : base + (i < rem ? 1 : 0);
        ch[i].normals = normals;
    }
    if (count == 1)
        heston_tech_chunk_run(&ch[0]);
    else
        run_chunks(ch, sizeof(heston_tech_chunk), count, heston_tech_chunk_run);
    mpo_result r;
    if (technique == MPO_ANTITHETIC) {
        double sum = 0.0, sq = 0.0;
        for (uint64_t i = 0; i < count; ++i) {
            sum += ch[i].sum;
            sq += ch[i].sq_sum;
        }
        r = mpo_summarize(units, sum, sq);
    } else {
        mpo_cv_moments m;
        memset(&m, 0, sizeof(m));
        for (uint64_t i = 0; i < count; ++i) cv_merge(&m, &ch[i].m);
        /* standard (from caller normals only): the discounted payoff moments */
        r = technique == MPO_STANDARD ? mpo_summarize(units, m.payoff_sum, m.payoff_sq_sum)
                                      : mpo_summarize_cv(&m, cfg->s0);
    }
    free(ch);
    return r;
}

MPO_EXPORT mpo_result mpo_heston_tech_price(const mpo_heston_cfg* cfg, int scheme, int technique) {
    if (technique == MPO_STANDARD) return mpo_heston_price(cfg, scheme);
    return heston_tech(cfg, scheme, technique, NULL);
}

MPO_EXPORT mpo_result mpo_heston_tech_from_normals(const mpo_heston_cfg* cfg, int scheme,
                                                   int technique, const double* normals) {
    return heston_tech(cfg, scheme, technique, normals);
}

/* ------------------------------------------------------------------ */
/* Merton jump diffusion: cpu.rs:932-953, 3151-3189, 6408-6441         */
/* ------------------------------------------------------------------ */
typedef struct {
    double s0, k, r, sigma, lambda, mu_j, sigma_j, t;
    uint64_t n_paths, seed, n_threads;
} mpo_merton_cfg;

/* sample_poisson cpu.rs:932-953: inversion from a fresh uniform that does
 * not touch the cached Box-Muller normal; capped at 4096 terms. */
static uint64_t sample_poisson(mpo_rng* g, double mean) {
    if (mean == 0.0) return 0;
    double u = rng_open01(g);
    double prob = exp(-mean);
    double cum = prob;
    uint64_t k = 0;
    while (u > cum) {
        k += 1;
        prob *= mean / (double)k;
        cum += prob;
        if (k >= 4096) break; /* the reference asserts here */
    }
    return k;
}

typedef struct {
    uint64_t seed, n_paths;
    double s0, k, drift_t, diff_vol_t, jump_mean_t, mu_j, sigma_j, discount, sum, sq_sum;
} merton_chunk;

/* simulate_merton_jump_diffusion_chunk cpu.rs:6408-6441 */
static void* merton_chunk_run(void* arg) {
    merton_chunk* c = (merton_chunk*)arg;
    mpo_rng g;
    rng_init(&g, c->seed);
    double sum = 0.0, sq = 0.0;
    const double log_s0 = log(c->s0);
    for (uint64_t p = 0; p < c->n_paths; ++p) {
        double diffusion = c->diff_vol_t * rng_normal(&g);
        uint64_t n = sample_poisson(&g, c->jump_mean_t);
        double jump = 0.0;
        if (n != 0) jump = (double)n * c->mu_j + c->sigma_j * sqrt((double)n) * rng_normal(&g);
        double s_t = exp(log_s0 + c->drift_t + diffusion + jump);
        double payoff = fmax_ref(s_t - c->k, 0.0) * c->discount;
        sum += payoff;
        sq += payoff * payoff;
    }
    c->sum = sum;
    c->sq_sum = sq;
    return NULL;
}

/* merton_jump_diffusion_call_price_mc_cpu cpu.rs:3151-3189 */
MPO_EXPORT mpo_result mpo_merton_price(const mpo_merton_cfg* cfg) {
    double comp = cfg->lambda * (exp(cfg->mu_j + 0.5 * (cfg->sigma_j * cfg->sigma_j)) - 1.0);
    double drift_t = (cfg->r - comp - 0.5 * cfg->sigma * cfg->sigma) * cfg->t;
    double diff_vol_t = cfg->sigma * sqrt(cfg->t);
    double jump_mean_t = cfg->lambda * cfg->t;
    double discount = exp(-cfg->r * cfg->t);
    uint64_t threads = resolved_threads(cfg->n_threads);
    uint64_t count = (threads <= 1 || cfg->n_paths < threads * 2000) ? 1 : threads;
    merton_chunk* ch = (merton_chunk*)calloc(count, sizeof(merton_chunk));
    uint64_t base = cfg->n_paths / count, rem = cfg->n_paths % count;
    for (uint64_t i = 0; i < count; ++i) {
        ch[i].seed = count == 1 ? cfg->seed : mpo_derive_chunk_seed(cfg->seed, i);
        ch[i].n_paths = count == 1 ? cfg->n_paths : base + (i < rem ? 1 : 0);
        ch[i].s0 = cfg->s0;
        ch[i].k = cfg->k;
        ch[i].drift_t = drift_t;
        ch[i].diff_vol_t = diff_vol_t;
        ch[i].jump_mean_t = jump_mean_t;
        ch[i].mu_j = cfg->mu_j;
        ch[i].sigma_j = cfg->sigma_j;
        ch[i].discount = discount;
    }
    if (count == 1)
        merton_chunk_run(&ch[0]);
    else
        run_chunks(ch, sizeof(merton_chunk), count, merton_chunk_run);
    double sum = 0.0, sq = 0.0;
    for (uint64_t i = 0; i < count; ++i) {
        sum += ch[i].sum;
        sq += ch[i].sq_sum;
    }
    free(ch);
    return mpo_summarize(cfg->n_paths, sum, sq);
}

/* Poisson inversion from a given uniform, same recursion as sample_poisson. */
MPO_EXPORT uint64_t mpo_poisson_from_uniform(double u, double mean) {
    if (mean == 0.0) return 0;
    double prob = exp(-mean);
    double cum = prob;
    uint64_t k = 0;
    while (u > cum) {
        k += 1;
        prob *= mean / (double)k;
        cum += prob;
        if (k >= 4096) break;
    }
    return k;
}

/* Stepped Merton (extension): per step the log-price moves by the
 * compensated diffusion drift, sigma*sqrt(dt)*z, and a compound-Poisson
 * jump with N ~ Poisson(lambda*dt) drawn by inversion of u and jump size
 * N*mu_j + sigma_j*sqrt(N)*zj.  With n_steps == 1 this is exactly the
 * reference's terminal formula (cpu.rs:6423-6434).  Draw buffers are
 * [path][step]. */
MPO_EXPORT void mpo_merton_paths_from_draws(const mpo_merton_cfg* cfg, uint64_t n_steps,
                                            const double* z, const double* u, const double* zj,
                                            double* terminal_out, uint64_t* jumps_out) {
    double comp = cfg->lambda * (exp(cfg->mu_j + 0.5 * (cfg->sigma_j * cfg->sigma_j)) - 1.0);
    double dt = cfg->t / (double)n_steps;
    double drift_dt = (cfg->r - comp - 0.5 * cfg->sigma * cfg->sigma) * dt;
    double vol_dt = cfg->sigma * sqrt(dt);
    double jump_mean_dt = cfg->lambda * dt;
    const double log_s0 = log(cfg->s0);
    for (uint64_t p = 0; p < cfg->n_paths; ++p) {
        double x = log_s0;
        uint64_t total = 0;
        for (uint64_t s = 0; s < n_steps; ++s) {
            uint64_t i = p * n_steps + s;
            double shock = drift_dt + vol_dt * z[i];
            uint64_t n = mpo_poisson_from_uniform(u[i], jump_mean_dt);
            if (n != 0) shock += (double)n * cfg->mu_j + cfg->sigma_j * sqrt((double)n) * zj[i];
            x += shock;
            total += n;
        }
        terminal_out[p] = exp(x);
        if (jumps_out) jumps_out[p] = total;
    }
}

/* ------------------------------------------------------------------ */
/* Basket: 2-asset terminal, cpu.rs:7503-7518, 7599-7639               */
/* ------------------------------------------------------------------ */
typedef struct {
    double s01, s02, k, r, sigma1, sigma2, rho, weight1, weight2, t;
    uint64_t n_paths, seed;
} mpo_basket2_cfg;

/* simulate_basket_pseudorandom_standard cpu.rs:7503-7518 (single stream) */
MPO_EXPORT mpo_result mpo_basket2_price(const mpo_basket2_cfg* cfg) {
    double drift1 = (cfg->r - 0.5 * cfg->sigma1 * cfg->sigma1) * cfg->t;
    double drift2 = (cfg->r - 0.5 * cfg->sigma2 * cfg->sigma2) * cfg->t;
    double vol1 = cfg->sigma1 * sqrt(cfg->t);
    double vol2 = cfg->sigma2 * sqrt(cfg->t);
    double rp = 1.0 - cfg->rho * cfg->rho;
    double ortho = sqrt(rp > 0.0 ? rp : 0.0);
    double discount = exp(-cfg->r * cfg->t);
    mpo_rng g;
    rng_init(&g, cfg->seed);
    double sum = 0.0, sq = 0.0;
    for (uint64_t p = 0; p < cfg->n_paths; ++p) {
        double z1 = rng_normal(&g);
        double z2i = rng_normal(&g);
        double z2 = cfg->rho * z1 + ortho * z2i;
        double s1 = cfg->s01 * exp(drift1 + vol1 * z1);
        double s2 = cfg->s02 * exp(drift2 + vol2 * z2);
        double basket = cfg->weight1 * s1 + cfg->weight2 * s2;
        double payoff = fmax_ref(basket - cfg->k, 0.0) * discount;
        sum += payoff;
        sq += payoff * payoff;
    }
    return mpo_summarize(cfg->n_paths, sum, sq);
}

/* Two-asset terminal basket with a technique, one stream: antithetic
 * cpu.rs:7520-7538 (pair (z1, z2) and (-z1, -z2)), control variate
 * :7540-7553 with control = basket * discount and expectation
 * w1 S01 + w2 S02 (basket_control_expectation :7637-7639). */
MPO_EXPORT mpo_result mpo_basket2_tech_price(const mpo_basket2_cfg* cfg, int technique) {
    if (technique == MPO_STANDARD) return mpo_basket2_price(cfg);
    double drift1 = (cfg->r - 0.5 * cfg->sigma1 * cfg->sigma1) * cfg->t;
    double drift2 = (cfg->r - 0.5 * cfg->sigma2 * cfg->sigma2) * cfg->t;
    double vol1 = cfg->sigma1 * sqrt(cfg->t);
    double vol2 = cfg->sigma2 * sqrt(cfg->t);
    double rp = 1.0 - cfg->rho * cfg->rho;
    double ortho = sqrt(rp > 0.0 ? rp : 0.0);
    double discount = exp(-cfg->r * cfg->t);
    mpo_rng g;
    rng_init(&g, cfg->seed);
    const uint64_t units = technique == MPO_ANTITHETIC ? (cfg->n_paths + 1) / 2 : cfg->n_paths;
    double sum = 0.0, sq = 0.0;
    mpo_cv_moments m;
    memset(&m, 0, sizeof(m));
    for (uint64_t u = 0; u < units; ++u) {
        double z1 = rng_normal(&g);
        double z2i = rng_normal(&g);
        double pay[2], ctrl[2];
        for (int side = 0; side < (technique == MPO_ANTITHETIC ? 2 : 1); ++side) {
            double a = side ? -z1 : z1, b = side ? -z2i : z2i;
            double z2 = cfg->rho * a + ortho * b;
            double s1 = cfg->s01 * exp(drift1 + vol1 * a);
            double s2 = cfg->s02 * exp(drift2 + vol2 * z2);
            double basket = cfg->weight1 * s1 + cfg->weight2 * s2;
            pay[side] = fmax_ref(basket - cfg->k, 0.0) * discount;
            ctrl[side] = basket * discount;
        }
        if (technique == MPO_ANTITHETIC) {
            double block = 0.5 * (pay[0] + pay[1]);
            sum += block;
            sq += block * block;
        } else {
            cv_record(&m, pay[0], ctrl[0]);
        }
    }
    if (technique == MPO_ANTITHETIC) return mpo_summarize(units, sum, sq);
    return mpo_summarize_cv(&m, cfg->weight1 * cfg->s01 + cfg->weight2 * cfg->s02);
}

/* The n-asset stepped basket (below) with a technique on a caller buffer
 * [unit][step][asset]: antithetic pairs use -z for the second path, the
 * control variate records (payoff, basket * discount) with expectation
 * sum_i w_i S0_i. */
static double basket_value(uint64_t n_assets, const double* s0, const double* drift,
                           const double* vol, const double* weight, const double* chol,
                           uint64_t n_steps, const double* z, double sign) {
    double x[16];
    for (uint64_t a = 0; a < n_assets; ++a) x[a] = log(s0[a]);
    for (uint64_t s = 0; s < n_steps; ++s) {
        const double* zs = z + s * n_assets;
        for (uint64_t a = 0; a < n_assets; ++a) {
            double c = 0.0;
            for (uint64_t b = 0; b <= a; ++b) c += chol[a * n_assets + b] * (sign * zs[b]);
            x[a] += drift[a] + vol[a] * c;
        }
    }
    double basket = 0.0;
    for (uint64_t a = 0; a < n_assets; ++a) basket += weight[a] * exp(x[a]);
    return basket;
}

MPO_EXPORT mpo_result mpo_basket_tech_from_normals(uint64_t n_assets, const double* s0,
                                                   const double* sigma, const double* weight,
                                                   const double* chol, double r, double t,
                                                   double k, uint64_t n_steps, uint64_t n_units,
                                                   int technique, const double* normals) {
    double dt = t / (double)n_steps;
    double drift[16], vol[16];
    for (uint64_t a = 0; a < n_assets; ++a) {
        drift[a] = (r - 0.5 * sigma[a] * sigma[a]) * dt;
        vol[a] = sigma[a] * sqrt(dt);
    }
    const double discount = exp(-r * t);
    double sum = 0.0, sq = 0.0, expect = 0.0;
    for (uint64_t a = 0; a < n_assets; ++a) expect += weight[a] * s0[a];
    mpo_cv_moments m;
    memset(&m, 0, sizeof(m));
    for (uint64_t u = 0; u < n_units; ++u) {
        const double* z = normals + u * n_steps * n_assets;
        double basket = basket_value(n_assets, s0, drift, vol, weight, chol, n_steps, z, 1.0);
        double pay = fmax_ref(basket - k, 0.0) * discount;
        if (technique == MPO_ANTITHETIC) {
            double bb = basket_value(n_assets, s0, drift, vol, weight, chol, n_steps, z, -1.0);
            double block = 0.5 * (pay + fmax_ref(bb - k, 0.0) * discount);
            sum += block;
            sq += block * block;
        } else if (technique == MPO_CONTROL_VARIATE) {
            cv_record(&m, pay, basket * discount);
        } else {
            sum += pay;
            sq += pay * pay;
        }
    }
    if (technique == MPO_CONTROL_VARIATE) return mpo_summarize_cv(&m, expect);
    return mpo_summarize(n_units, sum, sq);
}

/* n-asset stepped basket (extension generalising cpu.rs:7622-7635):
 * correlated shocks x = L z with L the lower Cholesky factor (row-major
 * n x n), per-asset log-Euler stepping, terminal basket sum_i w_i S_i(T).
 * normals are [path][step][asset]. */
MPO_EXPORT void mpo_basket_paths_from_normals(uint64_t n_assets, const double* s0,
                                              const double* sigma, const double* weight,
                                              const double* chol, double r, double t,
                                              uint64_t n_steps, uint64_t n_paths,
                                              const double* normals, double* basket_out) {
    double dt = t / (double)n_steps;
    double drift[16], vol[16], x[16];
    for (uint64_t a = 0; a < n_assets; ++a) {
        drift[a] = (r - 0.5 * sigma[a] * sigma[a]) * dt;
        vol[a] = sigma[a] * sqrt(dt);
    }
    for (uint64_t p = 0; p < n_paths; ++p) {
        for (uint64_t a = 0; a < n_assets; ++a) x[a] = log(s0[a]);
        for (uint64_t s = 0; s < n_steps; ++s) {
            const double* z = normals + (p * n_steps + s) * n_assets;
            for (uint64_t a = 0; a < n_assets; ++a) {
                double c = 0.0;
                for (uint64_t b = 0; b <= a; ++b) c += chol[a * n_assets + b] * z[b];
                x[a] += drift[a] + vol[a] * c;
            }
        }
        double basket = 0.0;
        for (uint64_t a = 0; a < n_assets; ++a) basket += weight[a] * exp(x[a]);
        basket_out[p] = basket;
    }
}

/* Lower Cholesky factor of an n x n symmetric positive definite matrix
 * (row-major).  Returns 0 on success, -1 if not positive definite. */
MPO_EXPORT int mpo_cholesky(uint64_t n, const double* a, double* l) {
    memset(l, 0, sizeof(double) * n * n);
    for (uint64_t i = 0; i < n; ++i) {
        for (uint64_t j = 0; j <= i; ++j) {
            double s = a[i * n + j];
            for (uint64_t k = 0; k < j; ++k) s -= l[i * n + k] * l[j * n + k];
            if (i == j) {
                if (s <= 0.0) return -1;
                l[i * n + i] = sqrt(s);
            } else {
                l[i * n + j] = s / l[j * n + j];
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Longstaff-Schwartz put: cpu.rs:3813-3945, 7667-7736                 */
/* ------------------------------------------------------------------ */
typedef struct {
    double s0, k, r, sigma, t;
    uint64_t n_paths, n_steps, seed;
    uint64_t basis_degree;
    int allow_intrinsic_now;
    int basis_weighted; /* extension: e^{-x/2} L_i(x) */
} mpo_lsm_cfg;

typedef struct {
    double price, stderr_;
    uint64_t early_exercise_count, maturity_exercise_count, regression_steps;
} mpo_lsm_result;

/* laguerre_lsm_basis cpu.rs:7667-7682 (unweighted Laguerre in x=S/K).
 * Extension: weighted != 0 multiplies every function by e^{-x/2}, the
 * weighted Laguerre functions of Longstaff and Schwartz (2001, eq. 2). */
static void laguerre_basis(double x, uint64_t degree, int weighted, double b[4]) {
    double x2 = x * x;
    double x3 = x2 * x;
    double w = weighted ? exp(-0.5 * x) : 1.0;
    b[0] = 1.0;
    b[1] = 0.0;
    b[2] = 0.0;
    b[3] = 0.0;
    if (degree >= 1) b[1] = 1.0 - x;
    if (degree >= 2) b[2] = 1.0 - 2.0 * x + 0.5 * x2;
    if (degree >= 3) b[3] = 1.0 - 3.0 * x + 1.5 * x2 - x3 / 6.0;
    for (int i = 0; i < 4; ++i) b[i] *= w;
}

MPO_EXPORT void mpo_laguerre_basis(double x, uint64_t degree, int weighted, double* out) {
    uint64_t d = degree < 1 ? 1 : (degree > 3 ? 3 : degree);
    laguerre_basis(x, d, weighted, out);
}

/* solve_lsm_normal_equations cpu.rs:7684-7736: Gauss-Jordan with partial
 * pivoting; fails if the best pivot magnitude is <= 1e-12. */
MPO_EXPORT int mpo_lsm_solve(const double* mat_in, const double* rhs_in, uint64_t dim,
                             double* out) {
    double m[4][4], rhs[4];
    for (int i = 0; i < 4; ++i) {
        rhs[i] = rhs_in[i];
        for (int j = 0; j < 4; ++j) m[i][j] = mat_in[i * 4 + j];
    }
    for (uint64_t piv = 0; piv < dim; ++piv) {
        uint64_t best = piv;
        double best_abs = fabs(m[piv][piv]);
        for (uint64_t row = piv + 1; row < dim; ++row) {
            double c = fabs(m[row][piv]);
            if (c > best_abs) {
                best = row;
                best_abs = c;
            }
        }
        if (best_abs <= 1e-12) return -1;
        if (best != piv) {
            for (int j = 0; j < 4; ++j) {
                double tmp = m[piv][j];
                m[piv][j] = m[best][j];
                m[best][j] = tmp;
            }
            double tr = rhs[piv];
            rhs[piv] = rhs[best];
            rhs[best] = tr;
        }
        double pivot = m[piv][piv];
        for (uint64_t c = piv; c < dim; ++c) m[piv][c] /= pivot;
        rhs[piv] /= pivot;
        for (uint64_t row = 0; row < dim; ++row) {
            if (row == piv) continue;
            double f = m[row][piv];
            if (f == 0.0) continue;
            for (uint64_t c = piv; c < dim; ++c) m[row][c] -= f * m[piv][c];
            rhs[row] -= f * rhs[piv];
        }
    }
    for (int i = 0; i < 4; ++i) out[i] = rhs[i];
    return 0;
}

/* The backward induction of price_put_lsm_cpu (cpu.rs:3835-3924) over a
 * given path matrix laid out [path][step] with n_steps+1 columns (column 0
 * is S0).  Optional outputs: per-date coefficients (n_steps x 4, zero when
 * the date was skipped) and per-path exercise step / cashflow. */
MPO_EXPORT void mpo_lsm_from_paths(const mpo_lsm_cfg* cfg, const double* paths,
                                   const uint64_t* schedule, uint64_t n_sched,
                                   mpo_lsm_result* res, double* coeff_out,
                                   uint64_t* exercise_out) {
    uint64_t degree = cfg->basis_degree < 1 ? 1 : (cfg->basis_degree > 3 ? 3 : cfg->basis_degree);
    uint64_t coeffs = degree + 1;
    uint64_t n = cfg->n_paths, m = cfg->n_steps, len = m + 1;
    double dt = cfg->t / (double)m;
    double disc_step = exp(-cfg->r * dt);
    double kk = cfg->k > DBL_MIN ? cfg->k : DBL_MIN;
    double* cash = (double*)malloc(sizeof(double) * n);
    uint64_t* ex = (uint64_t*)malloc(sizeof(uint64_t) * n);
    for (uint64_t p = 0; p < n; ++p) {
        double term = paths[p * len + m];
        cash[p] = fmax_ref(cfg->k - term, 0.0);
        ex[p] = m;
    }
    if (coeff_out) memset(coeff_out, 0, sizeof(double) * m * 4);
    uint64_t regression_steps = 0;
    /* schedule iterated in reverse, skipping the last date (cpu.rs:3843) */
    for (int64_t si = (int64_t)n_sched - 2; si >= 0; --si) {
        uint64_t step = schedule[si];
        double ata[16] = {0}, atb[4] = {0};
        uint64_t itm = 0;
        for (uint64_t p = 0; p < n; ++p) {
            double s = paths[p * len + step];
            double imm = fmax_ref(cfg->k - s, 0.0);
            if (imm <= 0.0) continue;
            double cont = cash[p] * powi_ref(disc_step, (int)(ex[p] - step));
            double b[4];
            laguerre_basis(s / kk, degree, cfg->basis_weighted, b);
            for (uint64_t i = 0; i < coeffs; ++i) {
                atb[i] += b[i] * cont;
                for (uint64_t j = 0; j < coeffs; ++j) ata[i * 4 + j] += b[i] * b[j];
            }
            itm += 1;
        }
        if (itm < coeffs) continue;
        double beta[4];
        if (mpo_lsm_solve(ata, atb, coeffs, beta) != 0) continue;
        regression_steps += 1;
        if (coeff_out)
            for (int i = 0; i < 4; ++i) coeff_out[step * 4 + i] = i < (int)coeffs ? beta[i] : 0.0;
        for (uint64_t p = 0; p < n; ++p) {
            double s = paths[p * len + step];
            double imm = fmax_ref(cfg->k - s, 0.0);
            if (imm <= 0.0) continue;
            double b[4];
            laguerre_basis(s / kk, degree, cfg->basis_weighted, b);
            double cont = 0.0;
            for (uint64_t i = 0; i < coeffs; ++i) cont += beta[i] * b[i];
            if (imm > fmax_ref(cont, 0.0)) {
                cash[p] = imm;
                ex[p] = step;
            }
        }
    }
    double sum = 0.0, sq = 0.0;
    uint64_t early = 0, mature = 0;
    for (uint64_t p = 0; p < n; ++p) {
        double d = cash[p] * powi_ref(disc_step, (int)ex[p]);
        sum += d;
        sq += d * d;
        if (ex[p] < m)
            early += 1;
        else if (cash[p] > 0.0)
            mature += 1;
        if (exercise_out) exercise_out[p] = ex[p];
    }
    mpo_result s = mpo_summarize(n, sum, sq);
    if (cfg->allow_intrinsic_now) {
        double intrinsic = fmax_ref(cfg->k - cfg->s0, 0.0);
        if (intrinsic > s.price) {
            s.price = intrinsic;
            s.stderr_ = 0.0;
        }
    }
    res->price = s.price;
    res->stderr_ = s.stderr_;
    res->early_exercise_count = early;
    res->maturity_exercise_count = mature;
    res->regression_steps = regression_steps;
    free(cash);
    free(ex);
}

/* GBM path matrix of price_put_lsm_cpu (cpu.rs:3820-3833): one stream,
 * multiplicative update S *= exp(drift + vol*z). */
MPO_EXPORT void mpo_lsm_paths(const mpo_lsm_cfg* cfg, double* paths) {
    double dt = cfg->t / (double)cfg->n_steps;
    double drift = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * dt;
    double vol_dt = cfg->sigma * sqrt(dt);
    uint64_t len = cfg->n_steps + 1;
    mpo_rng g;
    rng_init(&g, cfg->seed);
    for (uint64_t p = 0; p < cfg->n_paths; ++p) {
        double s = cfg->s0;
        paths[p * len] = s;
        for (uint64_t st = 1; st <= cfg->n_steps; ++st) {
            double z = rng_normal(&g);
            s *= exp(drift + vol_dt * z);
            paths[p * len + st] = s;
        }
    }
}

/* Same path matrix from a normal buffer [path][step]. */
MPO_EXPORT void mpo_lsm_paths_from_normals(const mpo_lsm_cfg* cfg, const double* normals,
                                           double* paths) {
    double dt = cfg->t / (double)cfg->n_steps;
    double drift = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * dt;
    double vol_dt = cfg->sigma * sqrt(dt);
    uint64_t len = cfg->n_steps + 1;
    for (uint64_t p = 0; p < cfg->n_paths; ++p) {
        double s = cfg->s0;
        paths[p * len] = s;
        for (uint64_t st = 1; st <= cfg->n_steps; ++st) {
            s *= exp(drift + vol_dt * normals[p * cfg->n_steps + st - 1]);
            paths[p * len + st] = s;
        }
    }
}

/* american_put_price_lsm_cpu cpu.rs:3947-3981 / bermudan :3983-4020 */
MPO_EXPORT void mpo_lsm_price(const mpo_lsm_cfg* cfg, const uint64_t* schedule, uint64_t n_sched,
                              mpo_lsm_result* res) {
    double* paths = (double*)malloc(sizeof(double) * cfg->n_paths * (cfg->n_steps + 1));
    mpo_lsm_paths(cfg, paths);
    mpo_lsm_from_paths(cfg, paths, schedule, n_sched, res, NULL, NULL);
    free(paths);
}

/* ------------------------------------------------------------------ */
/* Closed forms: cpu.rs:839-930, 3110-3149, 5244-5305                  */
/* ------------------------------------------------------------------ */
/* standard_normal_cdf cpu.rs:898-910: Abramowitz & Stegun 7.1.26 */
MPO_EXPORT double mpo_norm_cdf_as(double x) {
    double sign = x < 0.0 ? -1.0 : 1.0;
    double z = fabs(x) / sqrt(2.0);
    double t = 1.0 / (1.0 + 0.3275911 * z);
    double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027,
           a5 = 1.061405429;
    double poly = (((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t;
    double erf_approx = 1.0 - poly * exp(-z * z);
    return 0.5 * (1.0 + sign * erf_approx);
}

MPO_EXPORT double mpo_norm_cdf_exact(double x) { return 0.5 * erfc(-x / sqrt(2.0)); }

static double bs_call_with(double s0, double k, double r, double sigma, double t,
                           double (*cdf)(double)) {
    /* black_scholes_european_call_price cpu.rs:839-859 */
    if (k == 0.0) return s0;
    if (t == 0.0) return fmax_ref(s0 - k, 0.0);
    if (sigma == 0.0) return fmax_ref(s0 - k * exp(-r * t), 0.0);
    double sqrt_t = sqrt(t);
    double d1 = (log(s0 / k) + (r + 0.5 * sigma * sigma) * t) / (sigma * sqrt_t);
    double d2 = d1 - sigma * sqrt_t;
    return s0 * cdf(d1) - k * exp(-r * t) * cdf(d2);
}

/* exact-Phi variant: the fixtures in reference-fixtures.json use it */
MPO_EXPORT double mpo_bs_call(double s0, double k, double r, double sigma, double t) {
    return bs_call_with(s0, k, r, sigma, t, mpo_norm_cdf_exact);
}
/* the reference function's own output (A&S CDF) */
MPO_EXPORT double mpo_bs_call_ref(double s0, double k, double r, double sigma, double t) {
    return bs_call_with(s0, k, r, sigma, t, mpo_norm_cdf_as);
}
/* cpu.rs:861-864 put-call parity */
MPO_EXPORT double mpo_bs_put(double s0, double k, double r, double sigma, double t) {
    return mpo_bs_call(s0, k, r, sigma, t) + k * exp(-r * t) - s0;
}

/* discounted_lognormal_call_price cpu.rs:912-930 */
static double lognormal_call(double log_mean, double log_var, double k, double r, double t,
                             double (*cdf)(double)) {
    if (log_var == 0.0) return exp(-r * t) * fmax_ref(exp(log_mean) - k, 0.0);
    double sd = sqrt(log_var);
    double d2 = (log_mean - log(k)) / sd;
    double d1 = d2 + sd;
    return exp(-r * t) * (exp(log_mean + 0.5 * log_var) * cdf(d1) - k * cdf(d2));
}

/* merton_jump_diffusion_call_reference_price cpu.rs:3110-3149;
 * exact != 0 selects the exact normal CDF. */
MPO_EXPORT double mpo_merton_series(const mpo_merton_cfg* cfg, uint64_t max_terms, double tail_tol,
                                    int exact) {
    double (*cdf)(double) = exact ? mpo_norm_cdf_exact : mpo_norm_cdf_as;
    double lt = cfg->lambda * cfg->t;
    if (lt == 0.0) return bs_call_with(cfg->s0, cfg->k, cfg->r, cfg->sigma, cfg->t, cdf);
    double comp = cfg->lambda * (exp(cfg->mu_j + 0.5 * (cfg->sigma_j * cfg->sigma_j)) - 1.0);
    double base = log(cfg->s0) + (cfg->r - comp - 0.5 * cfg->sigma * cfg->sigma) * cfg->t;
    double w = exp(-lt), cum = 0.0, price = 0.0;
    for (uint64_t n = 0; n < max_terms; ++n) {
        double lm = base + (double)n * cfg->mu_j;
        double lv = cfg->sigma * cfg->sigma * cfg->t + (double)n * cfg->sigma_j * cfg->sigma_j;
        price += w * lognormal_call(lm, lv, cfg->k, cfg->r, cfg->t, cdf);
        cum += w;
        if (fmax_ref(1.0 - cum, 0.0) <= tail_tol) break;
        w *= lt / (double)(n + 1);
    }
    return price;
}

/* binomial_put_reference_price cpu.rs:5244-5305 (CRR tree); exercise_mask
 * has reference_steps+1 entries (the last is forced true). */
MPO_EXPORT double mpo_crr_put(double s0, double k, double r, double sigma, double t,
                              uint64_t steps, const uint8_t* exercise_mask_in) {
    double dt = t / (double)steps;
    double u = exp(sigma * sqrt(dt));
    double d = 1.0 / u;
    double growth = exp(r * dt);
    double p = (growth - d) / (u - d);
    double disc = exp(-r * dt);
    double* v = (double*)malloc(sizeof(double) * (steps + 1));
    uint8_t* mask = (uint8_t*)malloc(steps + 1);
    memcpy(mask, exercise_mask_in, steps + 1);
    mask[steps] = 1;
    for (uint64_t dm = 0; dm <= steps; ++dm) {
        uint64_t um = steps - dm;
        double s = s0 * powi_ref(u, (int)um) * powi_ref(d, (int)dm);
        v[dm] = fmax_ref(k - s, 0.0);
    }
    for (int64_t st = (int64_t)steps - 1; st >= 0; --st) {
        for (uint64_t dm = 0; dm <= (uint64_t)st; ++dm) {
            double cont = disc * (p * v[dm] + (1.0 - p) * v[dm + 1]);
            if (mask[st]) {
                uint64_t um = (uint64_t)st - dm;
                double s = s0 * powi_ref(u, (int)um) * powi_ref(d, (int)dm);
                v[dm] = fmax_ref(cont, k - s);
            } else {
                v[dm] = cont;
            }
        }
    }
    double out = v[0];
    free(v);
    free(mask);
    return out;
}

/* ------------------------------------------------------------------ */
/* Inverse normal (Acklam), cpu.rs:6012-6061                           */
/* ------------------------------------------------------------------ */
MPO_EXPORT double mpo_inverse_normal(double p) {
    static const double A[6] = {-3.969683028665376e1, 2.209460984245205e2, -2.759285104469687e2,
                                1.383577518672690e2, -3.066479806614716e1, 2.506628277459239};
    static const double B[5] = {-5.447609879822406e1, 1.615858368580409e2, -1.556989798598866e2,
                                6.680131188771972e1, -1.328068155288572e1};
    static const double C[6] = {-7.784894002430293e-3, -3.223964580411365e-1, -2.400758277161838,
                                -2.549732539343734, 4.374664141464968, 2.938163982698783};
    static const double D[4] = {7.784695709041462e-3, 3.224671290700398e-1, 2.445134137142996,
                                3.754408661907416};
    const double P_LOW = 0.02425, P_HIGH = 1.0 - P_LOW;
    if (p < DBL_MIN) p = DBL_MIN;
    if (p > 1.0 - DBL_EPSILON) p = 1.0 - DBL_EPSILON;
    if (p < P_LOW) {
        double q = sqrt(-2.0 * log(p));
        return (((((C[0] * q + C[1]) * q + C[2]) * q + C[3]) * q + C[4]) * q + C[5]) /
               ((((D[0] * q + D[1]) * q + D[2]) * q + D[3]) * q + 1.0);
    }
    if (p <= P_HIGH) {
        double q = p - 0.5;
        double r = q * q;
        return (((((A[0] * r + A[1]) * r + A[2]) * r + A[3]) * r + A[4]) * r + A[5]) * q /
               (((((B[0] * r + B[1]) * r + B[2]) * r + B[3]) * r + B[4]) * r + 1.0);
    }
    double q = sqrt(-2.0 * log(1.0 - p));
    return -(((((C[0] * q + C[1]) * q + C[2]) * q + C[3]) * q + C[4]) * q + C[5]) /
           ((((D[0] * q + D[1]) * q + D[2]) * q + D[3]) * q + 1.0);
}

/* ------------------------------------------------------------------ */
/* Brownian bridge, cpu.rs:5923-5980                                   */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t left, mid, right, normal_index;
    double lw, rw, sd;
} mpo_bb_op;

/* BrownianBridgePlan::new: DFS preorder over (left, right) intervals,
 * normal_index = creation order starting at 1.  Returns the op count. */
MPO_EXPORT uint64_t mpo_bb_plan(uint64_t n_steps, double maturity, mpo_bb_op* ops) {
    double dt = maturity / (double)n_steps;
    uint64_t* stack = (uint64_t*)malloc(sizeof(uint64_t) * 2 * (n_steps + 2));
    uint64_t top = 0, count = 0, idx = 1;
    stack[top++] = 0;
    stack[top++] = n_steps;
    while (top > 0) {
        uint64_t right = stack[--top];
        uint64_t left = stack[--top];
        if (right <= left + 1) continue;
        uint64_t mid = (left + right) / 2;
        double tl = (double)left * dt, tm = (double)mid * dt, tr = (double)right * dt;
        double denom = tr - tl;
        mpo_bb_op* op = &ops[count++];
        op->left = left;
        op->mid = mid;
        op->right = right;
        op->lw = (tr - tm) / denom;
        op->rw = (tm - tl) / denom;
        op->sd = sqrt((tm - tl) * (tr - tm) / denom);
        op->normal_index = idx++;
        stack[top++] = mid;
        stack[top++] = right;
        stack[top++] = left;
        stack[top++] = mid;
    }
    free(stack);
    return count;
}

/* fill_step_normals_from_inputs cpu.rs:5963-5980 */
MPO_EXPORT void mpo_bb_increments(uint64_t n_steps, double maturity, const double* inputs,
                                  double* increments) {
    mpo_bb_op* ops = (mpo_bb_op*)malloc(sizeof(mpo_bb_op) * (n_steps + 1));
    uint64_t n_ops = mpo_bb_plan(n_steps, maturity, ops);
    double* w = (double*)calloc(n_steps + 1, sizeof(double));
    double inv_sqrt_dt = 1.0 / sqrt(maturity / (double)n_steps);
    w[n_steps] = sqrt(maturity) * inputs[0];
    for (uint64_t i = 0; i < n_ops; ++i) {
        mpo_bb_op* op = &ops[i];
        w[op->mid] = op->lw * w[op->left] + op->rw * w[op->right] + op->sd * inputs[op->normal_index];
    }
    for (uint64_t s = 1; s <= n_steps; ++s) increments[s - 1] = (w[s] - w[s - 1]) * inv_sqrt_dt;
    free(w);
    free(ops);
}

/* ------------------------------------------------------------------ */
/* Scrambled Sobol                                                     */
/* ------------------------------------------------------------------ */
/* The reference draws Sobol points from the sobol_burley 0.5 crate
 * (cpu.rs:5393-5432, 5654-5684), whose source is not available here.
 * This restates the published construction it implements -- Joe-Kuo
 * (new-joe-kuo-6.21201) direction numbers, the index shuffled and the
 * point Owen-scrambled by hash-based nested uniform scrambling (Burley,
 * "Practical Hash-based Owen Scrambling", JCGT 2020) -- over 2^16-point
 * sequences and 256 dimensions, with the reference's own block/seed/wrap
 * arithmetic on top.  Unscrambled integer points are pinned against
 * scipy; the scrambled bits are not pinned against the crate. */
#define MPO_SOBOL_DIMS 256
#define MPO_SOBOL_BITS 32

static uint32_t g_sobol_dirs[MPO_SOBOL_DIMS][MPO_SOBOL_BITS];
static int g_sobol_ready = 0;

/* Direction numbers from the primitive polynomials (poly, with leading and
 * constant terms) and initial values m_1..m_s by the Bratley-Fox recurrence
 * m_k = 2a_1 m_{k-1} ^ 4a_2 m_{k-2} ^ ... ^ 2^s m_{k-s} ^ m_{k-s}. */
MPO_EXPORT int mpo_sobol_init(const uint32_t* poly, const uint32_t* vinit, uint64_t vinit_cols,
                              uint64_t n_dims) {
    if (n_dims < MPO_SOBOL_DIMS) return -1;
    for (int d = 0; d < MPO_SOBOL_DIMS; ++d) {
        uint32_t pl = poly[d];
        int s = 0;
        while ((pl >> (s + 1)) != 0) ++s;
        uint64_t m[MPO_SOBOL_BITS + 1];
        if (d == 0) {
            for (int k = 1; k <= MPO_SOBOL_BITS; ++k) m[k] = 1;
        } else {
            for (int k = 1; k <= s; ++k) m[k] = vinit[d * vinit_cols + (k - 1)];
            for (int k = s + 1; k <= MPO_SOBOL_BITS; ++k) {
                uint64_t v = m[k - s] ^ (m[k - s] << s);
                for (int i = 1; i < s; ++i) {
                    uint32_t a_i = (pl >> (s - i)) & 1u;
                    if (a_i) v ^= m[k - i] << i;
                }
                m[k] = v;
            }
        }
        for (int k = 1; k <= MPO_SOBOL_BITS; ++k)
            g_sobol_dirs[d][k - 1] = (uint32_t)(m[k] << (MPO_SOBOL_BITS - k));
    }
    g_sobol_ready = 1;
    return 0;
}

MPO_EXPORT void mpo_sobol_directions(uint32_t* out) {
    memcpy(out, g_sobol_dirs, sizeof(g_sobol_dirs));
}

static uint32_t rev32(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
    return (x >> 16) | (x << 16);
}

static uint32_t hash32(uint32_t x) {
    x ^= 0x79c68e4au;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

/* Laine-Karras style permutation on bit-reversed values: every step only
 * lets lower bits influence higher bits, so on reversed digits it is a
 * nested uniform (Owen) scramble. */
static uint32_t lk_scramble(uint32_t n, uint32_t seed) {
    uint32_t s = hash32(seed);
    n ^= n * 0x3d20adeau;
    n += s;
    n *= (s >> 16) | 1u;
    n ^= n * 0x05526c56u;
    n ^= n * 0x53a22864u;
    return n;
}

/* Unscrambled Sobol integer point (index < 2^32, dimension < 256). */
MPO_EXPORT uint32_t mpo_sobol_u32(uint32_t index, uint32_t dim) {
    uint32_t x = 0;
    for (int b = 0; b < 32 && index; ++b, index >>= 1)
        if (index & 1u) x ^= g_sobol_dirs[dim][b];
    return x;
}

/* Scrambled sample as a 32-bit integer: shuffled index (16 significant
 * digits), Sobol point, Owen scramble seeded by dim ^ seed. */
MPO_EXPORT uint32_t mpo_sobol_scrambled_u32(uint32_t index, uint32_t dim, uint32_t seed) {
    uint32_t shuffled = rev32(lk_scramble(rev32(index & 0xFFFFu), seed)) & 0xFFFFu;
    uint32_t x = mpo_sobol_u32(shuffled, dim);
    return rev32(lk_scramble(rev32(x), dim ^ seed));
}

/* f32 output in [0,1) with 23 significant bits, as the crate returns. */
MPO_EXPORT double mpo_sobol_sample(uint32_t index, uint32_t dim, uint32_t seed) {
    return (double)(mpo_sobol_scrambled_u32(index, dim, seed) >> 9) * (1.0 / 8388608.0);
}

/* fill_scrambled_sobol_point_standard_normals cpu.rs:5393-5432: blocks of
 * 2^16 points re-seeded with block*0x9E3779B9, dimensions >= 256 wrap with
 * seed + dim/256.  (The crate's 4-d fast path yields the same values as the
 * scalar path in this restatement.) */
MPO_EXPORT void mpo_sobol_point_normals(uint32_t base_seed, uint64_t point, uint64_t dims,
                                        double* out) {
    uint64_t block = point >> 16;
    uint32_t local = (uint32_t)(point & 0xFFFFu);
    uint32_t block_seed = base_seed + (uint32_t)block * 0x9E3779B9u;
    for (uint64_t d = 0; d < dims; ++d) {
        uint32_t off = (uint32_t)(d / MPO_SOBOL_DIMS);
        uint32_t dim = (uint32_t)(d % MPO_SOBOL_DIMS);
        double u = mpo_sobol_sample(local, dim, block_seed + off);
        out[d] = mpo_inverse_normal(u);
    }
}

/* simulate_stepwise_qmc cpu.rs:7113-7197 (standard technique) and its
 * Asian / down-and-out / lookback siblings: one Sobol point per path,
 * dims = n_steps, optional Brownian bridge. */
MPO_EXPORT mpo_result mpo_gbm_qmc_price(const mpo_gbm_cfg* cfg, int payoff, int bridge,
                                        double* terminal_out) {
    const double dt = cfg->t / (double)cfg->n_steps;
    const double drift_dt = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * dt;
    const double vol_dt = cfg->sigma * sqrt(dt);
    const double discount = exp(-cfg->r * cfg->t);
    const double log_s0 = log(cfg->s0);
    uint64_t m = cfg->n_steps;
    double* in = (double*)malloc(sizeof(double) * m);
    double* z = (double*)malloc(sizeof(double) * m);
    double sum = 0.0, sq = 0.0;
    for (uint64_t p = 0; p < cfg->n_paths; ++p) {
        mpo_sobol_point_normals((uint32_t)cfg->seed, p, m, in);
        if (bridge)
            mpo_bb_increments(m, cfg->t, in, z);
        else
            memcpy(z, in, sizeof(double) * m);
        double log_s = log_s0, acc = 0.0, max_s = cfg->s0;
        int knocked = cfg->s0 <= cfg->barrier;
        for (uint64_t s = 0; s < m; ++s) {
            log_s += drift_dt + vol_dt * z[s];
            if (payoff == MPO_ASIAN) acc += exp(log_s);
            if (payoff == MPO_DOWN_AND_OUT && exp(log_s) <= cfg->barrier) knocked = 1;
            if (payoff == MPO_LOOKBACK) max_s = fmax_ref(max_s, exp(log_s));
        }
        double s_t = exp(log_s), pay;
        if (payoff == MPO_ASIAN)
            pay = fmax_ref(acc / (double)m - cfg->k, 0.0) * discount;
        else if (payoff == MPO_DOWN_AND_OUT)
            pay = knocked ? 0.0 : fmax_ref(s_t - cfg->k, 0.0) * discount;
        else if (payoff == MPO_LOOKBACK)
            pay = fmax_ref(max_s - cfg->k, 0.0) * discount;
        else
            pay = fmax_ref(s_t - cfg->k, 0.0) * discount;
        if (terminal_out) terminal_out[p] = s_t;
        sum += pay;
        sq += pay * pay;
    }
    free(in);
    free(z);
    return mpo_summarize(cfg->n_paths, sum, sq);
}

/* ------------------------------------------------------------------ */
/* Structured normal samplers (StructuredNormalSampler cpu.rs:5790-5840)*/
/* RandomizedHalton cpu.rs:5686-5721 (halton_with_shift :6063-6077,    */
/* first_n_primes :6079-6102); LatinHypercube cpu.rs:5723-5788          */
/* (coprime_multiplier :5983-6010, open01_from_u64 :6104-6107); scrambled*/
/* Sobol as mpo_sobol_point_normals.                                    */
/* ------------------------------------------------------------------ */
enum { MPO_PSEUDORANDOM = 0, MPO_HALTON = 1, MPO_LHS = 2, MPO_SOBOL = 3, MPO_SOBOL_BB = 4 };

static double open01_from_u64(uint64_t v) {
    double x = (double)(v >> 11) * (1.0 / 9007199254740992.0);
    return x > DBL_MIN ? x : DBL_MIN;
}

static uint64_t gcd_u64(uint64_t a, uint64_t b) {
    while (b != 0) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

static uint64_t coprime_multiplier(uint64_t seed, uint64_t modulus) {
    if (modulus <= 1) return 1;
    uint64_t c = seed % modulus;
    if (c < 1) c = 1;
    if (c % 2 == 0) c = c == UINT64_MAX ? c : c + 1;
    while (gcd_u64(c, modulus) != 1) {
        c = c > UINT64_MAX - 2 ? UINT64_MAX : c + 2;
        if (c >= modulus) c = 1;
    }
    return c;
}

static void first_n_primes(uint64_t n, uint32_t* out) {
    uint64_t count = 0;
    uint32_t cand = 2;
    while (count < n) {
        uint32_t limit = (uint32_t)sqrt((double)cand);
        int prime = 1;
        for (uint64_t i = 0; i < count; ++i) {
            if (out[i] > limit) break;
            if (cand % out[i] == 0) {
                prime = 0;
                break;
            }
        }
        if (prime) out[count++] = cand;
        cand = cand == UINT32_MAX ? cand : cand + 1;
    }
}

static double halton(uint64_t index, uint32_t base) {
    double f = 1.0 / (double)base, value = 0.0;
    while (index > 0) {
        value += f * (double)(index % (uint64_t)base);
        index /= (uint64_t)base;
        f /= (double)base;
    }
    return value > DBL_MIN ? value : DBL_MIN;
}

typedef struct {
    int kind;
    uint64_t seed, n_paths, dims;
    uint32_t* primes;
    double* shifts;
    uint64_t *mult, *off, *jit;
} mpo_sampler;

static void sampler_init(mpo_sampler* s, int kind, uint64_t seed, uint64_t n_paths, uint64_t dims) {
    memset(s, 0, sizeof(*s));
    s->kind = kind;
    s->seed = seed;
    s->n_paths = n_paths;
    s->dims = dims;
    if (kind == MPO_HALTON) {
        uint64_t nd = dims > 2 ? dims : 2;
        s->primes = (uint32_t*)malloc(sizeof(uint32_t) * nd);
        s->shifts = (double*)malloc(sizeof(double) * nd);
        first_n_primes(nd, s->primes);
        for (uint64_t d = 0; d < nd; ++d)
            s->shifts[d] = open01_from_u64(mpo_splitmix64(seed + d * 0x9E3779B97F4A7C15ull));
    } else if (kind == MPO_LHS) {
        uint64_t nd = dims > 1 ? dims : 1;
        s->mult = (uint64_t*)malloc(sizeof(uint64_t) * nd);
        s->off = (uint64_t*)malloc(sizeof(uint64_t) * nd);
        s->jit = (uint64_t*)malloc(sizeof(uint64_t) * nd);
        for (uint64_t d = 0; d < nd; ++d) {
            uint64_t mixed = mpo_splitmix64(seed ^ (d * 0x9E3779B97F4A7C15ull) ^ 0xD1B54A32D192ED03ull);
            s->mult[d] = coprime_multiplier(mixed, n_paths > 1 ? n_paths : 1);
            s->off[d] = n_paths == 0 ? 0 : mpo_splitmix64(mixed ^ 0xA24BAED4963EE407ull) % n_paths;
            s->jit[d] = mpo_splitmix64(mixed ^ 0x9FB21C651E98DF25ull);
        }
    }
}

static void sampler_free(mpo_sampler* s) {
    free(s->primes);
    free(s->shifts);
    free(s->mult);
    free(s->off);
    free(s->jit);
}

static double sampler_normal(const mpo_sampler* s, uint64_t point, uint64_t dim) {
    if (s->kind == MPO_HALTON) {
        double shifted = halton(point + 1, s->primes[dim]) + s->shifts[dim];
        return mpo_inverse_normal(shifted - floor(shifted));
    }
    if (s->kind == MPO_LHS) {
        uint64_t stratum = s->n_paths == 0 ? 0 : (s->mult[dim] * point + s->off[dim]) % s->n_paths;
        double jitter = open01_from_u64(mpo_splitmix64(s->jit[dim] ^ (point * 0xBF58476D1CE4E5B9ull)));
        double u = ((double)stratum + jitter) / (double)s->n_paths;
        if (u < DBL_MIN) u = DBL_MIN;
        if (u > 1.0 - DBL_EPSILON) u = 1.0 - DBL_EPSILON;
        return mpo_inverse_normal(u);
    }
    /* scrambled Sobol, fill_scrambled_sobol_point_standard_normals */
    uint64_t block = point >> 16;
    uint32_t local = (uint32_t)(point & 0xFFFFu);
    uint32_t block_seed = (uint32_t)s->seed + (uint32_t)block * 0x9E3779B9u;
    uint32_t off = (uint32_t)(dim / MPO_SOBOL_DIMS);
    return mpo_inverse_normal(mpo_sobol_sample(local, (uint32_t)(dim % MPO_SOBOL_DIMS), block_seed + off));
}

/* Normals of one point in step order: the sampler's dims 0..m-1, or for
 * Sobol with a Brownian bridge the bridge's step normals (cpu.rs:5963-5980). */
static void point_step_normals(const mpo_sampler* s, uint64_t point, uint64_t m, double t, double* in,
                               double* z) {
    for (uint64_t d = 0; d < m; ++d) in[d] = sampler_normal(s, point, d);
    if (s->kind == MPO_SOBOL_BB)
        mpo_bb_increments(m, t, in, z);
    else
        memcpy(z, in, sizeof(double) * m);
}

MPO_EXPORT void mpo_structured_normals(int kind, uint64_t seed, uint64_t n_paths, uint64_t dims,
                                       uint64_t point0, uint64_t n_points, double* out) {
    mpo_sampler s;
    sampler_init(&s, kind, seed, n_paths, dims);
    for (uint64_t p = 0; p < n_points; ++p)
        for (uint64_t d = 0; d < dims; ++d) out[p * dims + d] = sampler_normal(&s, point0 + p, d);
    sampler_free(&s);
}

/* QMC GBM family with any technique: simulate_stepwise_qmc cpu.rs:7113-7197,
 * Asian :7199-7290, down-and-out :7292-7407, lookback :7409-7501; one point
 * per path (or pair), dims = n_steps, the sampler built over n_paths. */
MPO_EXPORT mpo_result mpo_gbm_structured_price(const mpo_gbm_cfg* cfg, int payoff, int sampling,
                                               int technique, double* terminal_out) {
    const uint64_t m = cfg->n_steps;
    const double dt = cfg->t / (double)m;
    const double drift_dt = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * dt;
    const double vol_dt = cfg->sigma * sqrt(dt);
    const double discount = exp(-cfg->r * cfg->t);
    const double log_s0 = log(cfg->s0);
    mpo_sampler smp;
    sampler_init(&smp, sampling, cfg->seed, cfg->n_paths, m);
    const uint64_t units = technique == MPO_ANTITHETIC ? (cfg->n_paths + 1) / 2 : cfg->n_paths;
    double* in = (double*)malloc(sizeof(double) * m);
    double* z = (double*)malloc(sizeof(double) * m);
    double sum = 0.0, sq = 0.0;
    mpo_cv_moments mo;
    memset(&mo, 0, sizeof(mo));
    for (uint64_t u = 0; u < units; ++u) {
        point_step_normals(&smp, u, m, cfg->t, in, z);
        gbm_state a, b;
        gbm_state_init(&a, log_s0, cfg->s0, cfg->barrier);
        gbm_state_init(&b, log_s0, cfg->s0, cfg->barrier);
        for (uint64_t st = 0; st < m; ++st) {
            gbm_state_step(&a, payoff, drift_dt + vol_dt * z[st], cfg->barrier);
            if (technique == MPO_ANTITHETIC) gbm_state_step(&b, payoff, drift_dt - vol_dt * z[st], cfg->barrier);
        }
        double pa = gbm_state_payoff(&a, payoff, cfg->k, m, discount);
        if (terminal_out) terminal_out[u] = exp(a.log_s);
        if (technique == MPO_ANTITHETIC) {
            double block = 0.5 * (pa + gbm_state_payoff(&b, payoff, cfg->k, m, discount));
            sum += block;
            sq += block * block;
        } else if (technique == MPO_CONTROL_VARIATE) {
            cv_record(&mo, pa, discount * exp(a.log_s));
        } else {
            sum += pa;
            sq += pa * pa;
        }
    }
    free(in);
    free(z);
    sampler_free(&smp);
    if (technique == MPO_CONTROL_VARIATE) return mpo_summarize_cv(&mo, cfg->s0);
    return mpo_summarize(units, sum, sq);
}

/* QMC terminal European: simulate_terminal_qmc cpu.rs:7064-7111 (one dim). */
MPO_EXPORT mpo_result mpo_terminal_structured_price(const mpo_gbm_cfg* cfg, int sampling, int technique) {
    const double drift_t = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * cfg->t;
    const double vol_t = cfg->sigma * sqrt(cfg->t);
    const double discount = exp(-cfg->r * cfg->t);
    mpo_sampler smp;
    sampler_init(&smp, sampling == MPO_SOBOL_BB ? MPO_SOBOL : sampling, cfg->seed, cfg->n_paths, 1);
    const uint64_t units = technique == MPO_ANTITHETIC ? (cfg->n_paths + 1) / 2 : cfg->n_paths;
    double* z = (double*)malloc(sizeof(double) * (units ? units : 1));
    for (uint64_t u = 0; u < units; ++u) z[u] = sampler_normal(&smp, u, 0);
    sampler_free(&smp);
    mpo_result r;
    if (technique == MPO_STANDARD) {
        double sum = 0.0, sq = 0.0;
        for (uint64_t u = 0; u < units; ++u) {
            double pay = fmax_ref(cfg->s0 * exp(drift_t + vol_t * z[u]) - cfg->k, 0.0) * discount;
            sum += pay;
            sq += pay * pay;
        }
        r = mpo_summarize(units, sum, sq);
    } else {
        r = term_tech(cfg, technique, z);
    }
    free(z);
    return r;
}

/* QMC two-asset basket: simulate_basket_qmc cpu.rs:7555-7597 (dims 2). */
MPO_EXPORT mpo_result mpo_basket2_structured_price(const mpo_basket2_cfg* cfg, int sampling,
                                                   int technique) {
    double drift1 = (cfg->r - 0.5 * cfg->sigma1 * cfg->sigma1) * cfg->t;
    double drift2 = (cfg->r - 0.5 * cfg->sigma2 * cfg->sigma2) * cfg->t;
    double vol1 = cfg->sigma1 * sqrt(cfg->t);
    double vol2 = cfg->sigma2 * sqrt(cfg->t);
    double rp = 1.0 - cfg->rho * cfg->rho;
    double ortho = sqrt(rp > 0.0 ? rp : 0.0);
    double discount = exp(-cfg->r * cfg->t);
    mpo_sampler smp;
    sampler_init(&smp, sampling == MPO_SOBOL_BB ? MPO_SOBOL : sampling, cfg->seed, cfg->n_paths, 2);
    const uint64_t units = technique == MPO_ANTITHETIC ? (cfg->n_paths + 1) / 2 : cfg->n_paths;
    double sum = 0.0, sq = 0.0;
    mpo_cv_moments mo;
    memset(&mo, 0, sizeof(mo));
    for (uint64_t u = 0; u < units; ++u) {
        double z1 = sampler_normal(&smp, u, 0), z2i = sampler_normal(&smp, u, 1);
        double pay[2], ctrl[2];
        for (int side = 0; side < (technique == MPO_ANTITHETIC ? 2 : 1); ++side) {
            double a = side ? -z1 : z1, b = side ? -z2i : z2i;
            double z2 = cfg->rho * a + ortho * b;
            double s1 = cfg->s01 * exp(drift1 + vol1 * a);
            double s2 = cfg->s02 * exp(drift2 + vol2 * z2);
            double basket = cfg->weight1 * s1 + cfg->weight2 * s2;
            pay[side] = fmax_ref(basket - cfg->k, 0.0) * discount;
            ctrl[side] = basket * discount;
        }
        if (technique == MPO_ANTITHETIC) {
            double block = 0.5 * (pay[0] + pay[1]);
            sum += block;
            sq += block * block;
        } else if (technique == MPO_CONTROL_VARIATE) {
            cv_record(&mo, pay[0], ctrl[0]);
        } else {
            sum += pay[0];
            sq += pay[0] * pay[0];
        }
    }
    sampler_free(&smp);
    if (technique == MPO_CONTROL_VARIATE)
        return mpo_summarize_cv(&mo, cfg->weight1 * cfg->s01 + cfg->weight2 * cfg->s02);
    return mpo_summarize(units, sum, sq);
}

/* Two-factor Sobol + Brownian-bridge Heston (extension; the reference has
 * no QMC Heston).  One Sobol point of 2*n_steps dims per path; bridge input
 * k of factor f (0 = variance shock, 1 = independent shock) is dimension
 * 2k+f, so both factors' terminal values take the first two dims.  Writes
 * S_T for paths [path0, path0 + n_paths). */
MPO_EXPORT void mpo_heston_qmc_paths(const mpo_heston_cfg* cfg, int scheme, uint64_t path0,
                                     double* terminal_out) {
    heston_params h = make_heston_params(cfg, scheme);
    uint64_t m = cfg->n_steps;
    double* pt = (double*)malloc(sizeof(double) * 2 * m);
    double* in0 = (double*)malloc(sizeof(double) * m);
    double* in1 = (double*)malloc(sizeof(double) * m);
    double* z0 = (double*)malloc(sizeof(double) * m);
    double* z1 = (double*)malloc(sizeof(double) * m);
    for (uint64_t p = 0; p < cfg->n_paths; ++p) {
        mpo_sobol_point_normals((uint32_t)cfg->seed, path0 + p, 2 * m, pt);
        for (uint64_t k = 0; k < m; ++k) {
            in0[k] = pt[2 * k];
            in1[k] = pt[2 * k + 1];
        }
        mpo_bb_increments(m, cfg->t, in0, z0);
        mpo_bb_increments(m, cfg->t, in1, z1);
        double log_s = h.log_s0, v = h.v0;
        for (uint64_t s = 0; s < m; ++s) heston_step(&h, z0[s], z1[s], &log_s, &v);
        terminal_out[p] = exp(log_s);
    }
    free(pt);
    free(in0);
    free(in1);
    free(z0);
    free(z1);
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 draws (extension: the GPU engine's pseudorandom source) */
/* ------------------------------------------------------------------ */
/* Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as easy as 1,
 * 2, 3" (SC 2011): 10 rounds of the Philox-4x32 S-P network. */
MPO_EXPORT void mpo_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                                  uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Engine mapping of Philox output to uniforms: 23 significant bits plus a
 * half-ulp offset, u in [2^-24, 1 - 2^-24], exact in both f32 and f64. */
static double philox_uniform(uint32_t x) { return ((double)(x >> 9) + 0.5) * (1.0 / 8388608.0); }

/* Engine Philox normals in f64: counter (path, block, stream, seed_hi),
 * key (seed_lo, contract); each call yields two Box-Muller pairs
 * (x0,x1)->(r cos, r sin), (x2,x3)->(r cos, r sin).  Normal j of a path
 * is element j%4 of block j/4.  out is [path][n_normals]. */
MPO_EXPORT void mpo_philox_normals(uint64_t seed, uint32_t contract, uint32_t stream,
                                   uint64_t path0, uint64_t n_paths, uint64_t n_normals,
                                   double* out) {
    uint32_t key[2] = {(uint32_t)seed, contract};
    for (uint64_t p = 0; p < n_paths; ++p) {
        for (uint64_t j = 0; j < n_normals; j += 4) {
            uint32_t ctr[4] = {(uint32_t)(path0 + p), (uint32_t)(j / 4), stream,
                               (uint32_t)(seed >> 32)};
            uint32_t r[4];
            mpo_philox4x32_10(ctr, key, r);
            double z[4];
            for (int q = 0; q < 2; ++q) {
                double u1 = philox_uniform(r[2 * q]);
                double u2 = philox_uniform(r[2 * q + 1]);
                double rad = sqrt(-2.0 * log(u1));
                z[2 * q] = rad * cos(2.0 * M_PI * u2);
                z[2 * q + 1] = rad * sin(2.0 * M_PI * u2);
            }
            for (uint64_t q = 0; q < 4 && j + q < n_normals; ++q) out[p * n_normals + j + q] = z[q];
        }
    }
}

/* Raw Philox uniforms with the engine mapping; out is [path][n]. */
MPO_EXPORT void mpo_philox_uniforms(uint64_t seed, uint32_t contract, uint32_t stream,
                                    uint64_t path0, uint64_t n_paths, uint64_t n, double* out) {
    uint32_t key[2] = {(uint32_t)seed, contract};
    for (uint64_t p = 0; p < n_paths; ++p) {
        for (uint64_t j = 0; j < n; j += 4) {
            uint32_t ctr[4] = {(uint32_t)(path0 + p), (uint32_t)(j / 4), stream,
                               (uint32_t)(seed >> 32)};
            uint32_t r[4];
            mpo_philox4x32_10(ctr, key, r);
            for (uint64_t q = 0; q < 4 && j + q < n; ++q) out[p * n + j + q] = philox_uniform(r[q]);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Greeks: black_scholes_european_call_greeks cpu.rs:866-896; pathwise */
/* and likelihood-ratio single-pass estimators cpu.rs:4670-4892 with   */
/* MomentAccumulator (sample variance n-1).  Bump-and-revalue is the   */
/* composition of the pricers above (oracle/cpu.py).                   */
/* ------------------------------------------------------------------ */
typedef struct {
    double price, delta, vega, rho, theta;
} mpo_bs_greeks_t;

MPO_EXPORT mpo_bs_greeks_t mpo_bs_greeks(double s0, double k, double r, double sigma, double t) {
    double sqrt_t = sqrt(t);
    double d1 = (log(s0 / k) + (r + 0.5 * sigma * sigma) * t) / (sigma * sqrt_t);
    double d2 = d1 - sigma * sqrt_t;
    double discount = exp(-r * t);
    double pdf_d1 = exp(-0.5 * d1 * d1) / sqrt(2.0 * M_PI);
    mpo_bs_greeks_t g;
    g.price = s0 * mpo_norm_cdf_as(d1) - k * discount * mpo_norm_cdf_as(d2);
    g.delta = mpo_norm_cdf_as(d1);
    g.vega = s0 * pdf_d1 * sqrt_t;
    g.rho = k * t * discount * mpo_norm_cdf_as(d2);
    g.theta = -(s0 * pdf_d1 * sigma) / (2.0 * sqrt_t) - r * k * discount * mpo_norm_cdf_as(d2);
    return g;
}

/* MomentAccumulator::mean_stderr cpu.rs:4861-4876 */
static void moment_mean_stderr(uint64_t count, double sum, double sq_sum, double* mean,
                               double* stderr_out) {
    if (count == 0) {
        *mean = 0.0;
        *stderr_out = 0.0;
        return;
    }
    double n = (double)count;
    *mean = sum / n;
    double variance = 0.0;
    if (count > 1) {
        variance = (sq_sum - sum * sum / n) / (n - 1.0);
        if (!(variance > 0.0)) variance = 0.0;
    }
    *stderr_out = sqrt(variance / n);
}

/* Single-pass terminal-GBM Greeks.  estimator 1 = pathwise (Delta, Vega,
 * Rho, Theta), 2 = likelihood ratio (Delta, Vega, Rho; slot 3 zero).  The
 * draws are the reference stream MonteCarloRng::new(seed) (one stream for
 * every path, whatever n_threads), or the caller's normals[path] when
 * normals is not NULL.  out[2g] = value, out[2g+1] = stderr. */
MPO_EXPORT void mpo_terminal_greeks(const mpo_gbm_cfg* cfg, int estimator, const double* normals,
                                    double* out) {
    mpo_rng g;
    rng_init(&g, cfg->seed);
    double sqrt_t = sqrt(cfg->t);
    double drift_t = (cfg->r - 0.5 * cfg->sigma * cfg->sigma) * cfg->t;
    double vol_t = cfg->sigma * sqrt_t;
    double discount = exp(-cfg->r * cfg->t);
    double sum[4] = {0, 0, 0, 0}, sq[4] = {0, 0, 0, 0};
    for (uint64_t p = 0; p < cfg->n_paths; ++p) {
        double z = normals ? normals[p] : rng_normal(&g);
        double st = cfg->s0 * exp(drift_t + vol_t * z);
        double v[4] = {0, 0, 0, 0};
        if (estimator == 1) {
            double intrinsic = st - cfg->k;
            int active = intrinsic > 0.0;
            double payoff = fmax_ref(intrinsic, 0.0);
            double discounted = discount * payoff;
            v[0] = active ? discount * st / cfg->s0 : 0.0;
            v[1] = active ? discount * st * (-cfg->sigma * cfg->t + sqrt_t * z) : 0.0;
            v[2] = active ? discount * (cfg->t * st - cfg->t * payoff) : -cfg->t * discounted;
            double dst_dt =
                st * (cfg->r - 0.5 * cfg->sigma * cfg->sigma + cfg->sigma * z / (2.0 * sqrt_t));
            double dp = active ? discount * (dst_dt - cfg->r * payoff) : -cfg->r * discounted;
            v[3] = -dp;
        } else {
            double discounted = discount * fmax_ref(st - cfg->k, 0.0);
            v[0] = discounted * z / (cfg->s0 * cfg->sigma * sqrt_t);
            double vega_score = (z * z - 1.0) / cfg->sigma - sqrt_t * z;
            v[1] = discounted * vega_score;
            v[2] = discounted * (z / (cfg->sigma * sqrt_t) - cfg->t);
        }
        int n_g = estimator == 1 ? 4 : 3;
        for (int i = 0; i < n_g; ++i) {
            sum[i] += v[i];
            sq[i] += v[i] * v[i];
        }
    }
    int n_g = estimator == 1 ? 4 : 3;
    for (int i = 0; i < 4; ++i) {
        if (i < n_g)
            moment_mean_stderr(cfg->n_paths, sum[i], sq[i], &out[2 * i], &out[2 * i + 1]);
        else
            out[2 * i] = out[2 * i + 1] = 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* Multilevel Monte Carlo arithmetic Asian: simulate_asian_mlmc_level  */
/* cpu.rs:6639-6773 with the level-0 payoff :6775-6795 and the coupled */
/* fine/coarse difference :6812-6849 (fine increments summed in groups */
/* of refinement_factor, divided by sqrt(refinement)).  Each level     */
/* draws from MonteCarloRng(derive_chunk_seed(seed, level)) or from a  */
/* structured sampler built with that seed over (paths, fine_steps).   */
/* out[0] = sum, out[1] = sum of squares of the discounted level term. */
/* ------------------------------------------------------------------ */
typedef struct {
    double s0, k, r, sigma, t;
    uint64_t refinement, seed;
    int64_t sampling;
} mpo_mlmc_cfg;

static double mlmc_term(const mpo_mlmc_cfg* c, uint64_t fine_steps, const double* z, double fine_drift,
                        double fine_vol, double coarse_drift, double coarse_vol, double sqrt_ref,
                        double discount, int level0) {
    if (level0) {
        double log_s = log(c->s0), sum = 0.0;
        for (uint64_t i = 0; i < fine_steps; ++i) {
            log_s += fine_drift + fine_vol * z[i];
            sum += exp(log_s);
        }
        return fmax_ref(sum / (double)fine_steps - c->k, 0.0) * discount;
    }
    uint64_t coarse_steps = fine_steps / c->refinement;
    double fine_log = log(c->s0), coarse_log = log(c->s0), fine_sum = 0.0, coarse_sum = 0.0;
    uint64_t idx = 0;
    for (uint64_t cs = 0; cs < coarse_steps; ++cs) {
        double zsum = 0.0;
        for (uint64_t j = 0; j < c->refinement; ++j) {
            double zz = z[idx++];
            zsum += zz;
            fine_log += fine_drift + fine_vol * zz;
            fine_sum += exp(fine_log);
        }
        double coarse_z = zsum / sqrt_ref;
        coarse_log += coarse_drift + coarse_vol * coarse_z;
        coarse_sum += exp(coarse_log);
    }
    double fine_pay = fmax_ref(fine_sum / (double)(coarse_steps * c->refinement) - c->k, 0.0) * discount;
    double coarse_pay = fmax_ref(coarse_sum / (double)coarse_steps - c->k, 0.0) * discount;
    return fine_pay - coarse_pay;
}

MPO_EXPORT void mpo_mlmc_level(const mpo_mlmc_cfg* c, uint64_t level, uint64_t paths,
                               uint64_t fine_steps, double* out) {
    const uint64_t lseed = mpo_derive_chunk_seed(c->seed, level);
    const double discount = exp(-c->r * c->t);
    const double fine_dt = c->t / (double)fine_steps;
    const double fine_drift = (c->r - 0.5 * c->sigma * c->sigma) * fine_dt;
    const double fine_vol = c->sigma * sqrt(fine_dt);
    double coarse_drift = 0.0, coarse_vol = 0.0;
    if (level > 0) {
        const double coarse_dt = c->t / (double)(fine_steps / c->refinement);
        coarse_drift = (c->r - 0.5 * c->sigma * c->sigma) * coarse_dt;
        coarse_vol = c->sigma * sqrt(coarse_dt);
    }
    const double sqrt_ref = sqrt((double)c->refinement);
    double* z = (double*)malloc(sizeof(double) * (fine_steps ? fine_steps : 1));
    mpo_rng g;
    rng_init(&g, lseed);
    mpo_sampler smp;
    const int structured = c->sampling != MPO_PSEUDORANDOM;
    if (structured)
        sampler_init(&smp, c->sampling == MPO_SOBOL_BB ? MPO_SOBOL : (int)c->sampling, lseed, paths,
                     fine_steps);
    double sum = 0.0, sq = 0.0;
    for (uint64_t p = 0; p < paths; ++p) {
        for (uint64_t i = 0; i < fine_steps; ++i)
            z[i] = structured ? sampler_normal(&smp, p, i) : rng_normal(&g);
        double v = mlmc_term(c, fine_steps, z, fine_drift, fine_vol, coarse_drift, coarse_vol, sqrt_ref,
                             discount, level == 0);
        sum += v;
        sq += v * v;
    }
    if (structured) sampler_free(&smp);
    free(z);
    out[0] = sum;
    out[1] = sq;
}

/* The same level term from caller normals [path][fine_steps] (parity). */
MPO_EXPORT void mpo_mlmc_level_from_normals(const mpo_mlmc_cfg* c, uint64_t level, uint64_t paths,
                                            uint64_t fine_steps, const double* normals, double* out) {
    const double discount = exp(-c->r * c->t);
    const double fine_dt = c->t / (double)fine_steps;
    const double fine_drift = (c->r - 0.5 * c->sigma * c->sigma) * fine_dt;
    const double fine_vol = c->sigma * sqrt(fine_dt);
    double coarse_drift = 0.0, coarse_vol = 0.0;
    if (level > 0) {
        const double coarse_dt = c->t / (double)(fine_steps / c->refinement);
        coarse_drift = (c->r - 0.5 * c->sigma * c->sigma) * coarse_dt;
        coarse_vol = c->sigma * sqrt(coarse_dt);
    }
    const double sqrt_ref = sqrt((double)c->refinement);
    double sum = 0.0, sq = 0.0;
    for (uint64_t p = 0; p < paths; ++p) {
        double v = mlmc_term(c, fine_steps, normals + p * fine_steps, fine_drift, fine_vol, coarse_drift,
                             coarse_vol, sqrt_ref, discount, level == 0);
        sum += v;
        sq += v * v;
    }
    out[0] = sum;
    out[1] = sq;
}
