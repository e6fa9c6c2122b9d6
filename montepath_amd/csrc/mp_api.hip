// mp_api.hip -- C-ABI host layer of the montepath HIP engine.
//
// Validates requests the way the reference's config validators do
// (crates/mc-core/src/runtime/cpu.rs:5010-5153) but returns status codes
// instead of panicking, builds per-request device descriptors, packs
// requests into work-queue launches, drives the LSM date loop, and runs the
// continuous-batching server.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "montepath_hip.h"
#include "mp_device.h"
#include "mp_engine.h"
#include "sobol_joe_kuo_table.h"

#define MP_EXPORT extern "C" __attribute__((visibility("default")))

// Device scratch of the allocation-free one-request pricing
// (mp_hip_price_noalloc): the request record, the queue counter and the
// per-tile FP64 partial sums.  Each device holds its own copy of the array
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (g_scratch_ptr), never
// through the array's name.
static __device__ __align__(256) unsigned char scratch[MP_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(MP_E_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

// Device address of `scratch` per device ordinal (0..63), resolved once per
// device by scratch_setup with that device current; calls that price into a
// device's scratch hold that device's g_scratch_call_mu.
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};
std::mutex g_scratch_call_mu[64];

int scratch_setup(int device) {
    if (device < 0 || device >= 64) return fail(MP_E_UNSUPPORTED, "device ordinals above 63 are not supported");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return MP_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        return fail(MP_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    g_scratch_ptr[device] = p;
    return MP_OK;
}

// compiler-rt __powidf2: the lowering of Rust's f64::powi used by the reference
double powi_ref(double a, int b) {
    const bool recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

uint64_t pow2_floor(uint64_t x) {
    uint64_t p = 1;
    while (p * 2 <= x) p *= 2;
    return p;
}

// Device buffer that grows on demand and is reused across calls.
struct dbuf {
    void* p = nullptr;
    size_t cap = 0;
    ~dbuf() {
        if (p) (void)hipFree(p);
    }
    int ensure(size_t bytes) {
        if (bytes == 0) bytes = 8;
        if (cap >= bytes) return MP_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = bytes + bytes / 4;
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return fail(MP_E_OOM, "device allocation of " + std::to_string(want) + " bytes failed");
        }
        cap = want;
        return MP_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Pinned host memory, mapped for the device and coherent, grown on demand
// (only between uses that the stream has completed).
struct hbuf {
    void* p = nullptr;
    void* dev = nullptr;  // the device's address of p
    size_t cap = 0;
    hbuf() = default;
    hbuf(const hbuf&) = delete;
    hbuf& operator=(const hbuf&) = delete;
    ~hbuf() {
        if (p) (void)hipHostFree(p);
    }
    int ensure(size_t bytes) {
        if (bytes == 0) bytes = 8;
        if (cap >= bytes) return MP_OK;
        if (p) (void)hipHostFree(p);
        p = dev = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
        if (hipHostMalloc(&p, want, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
            hipHostGetDevicePointer(&dev, p, 0) != hipSuccess) {
            (void)hipGetLastError();
            if (p) (void)hipHostFree(p);
            p = dev = nullptr;
            return fail(MP_E_OOM, "pinned host allocation of " + std::to_string(want) + " bytes failed");
        }
        cap = want;
        return MP_OK;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

#define RETCHK(expr)                  \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != MP_OK) return rc_; \
    } while (0)

#define ENSURE(buf, bytes)                      \
    do {                                        \
        int rc_ = (buf).ensure(bytes);          \
        if (rc_ != MP_OK) return rc_;           \
    } while (0)

enum req_kind { KIND_TILE = 0, KIND_QMC = 1, KIND_BASKET = 2, KIND_LSM = 3 };

struct prepared {
    mp_dreq d;
    int kind;
    int precision;
    int factors;       // QMC
    int bridge;        // QMC
    int n_assets;      // basket
    std::vector<uint32_t> schedule;  // LSM dates
    int allow_intrinsic;
    int degree;
    int weighted = 0;  // LSM basis e^{-x/2} L_i(x)
    uint64_t n_steps_cfg;  // steps as configured (terminal method walks 1)
    uint64_t n_paths_req;  // paths as requested (antithetic units are pairs)
    std::vector<mp_sampler_dim> sampler;  // Halton / Latin hypercube dimensions
};

}  // namespace

struct mp_batch;

struct mp_ctx {
    int device = 0;
    mp_batch* bat[2] = {nullptr, nullptr};  // reusable FP32 / FP64 batches on the context stream
    hipStream_t stream = nullptr;
    int n_cu = 0;
    size_t max_lds = 0;
    std::string name, arch;
    uint32_t dirs[256][32];
    uint32_t* d_nib = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    mp_timing last{};
    std::map<uint64_t, mp_tuning> tuning;
    // arenas
    dbuf reqs, partial, counter;
    dbuf normals, uniforms, jumps, term, pay;
    dbuf plan, basket_par;
    dbuf lsm_paths, lsm_cash, lsm_ex, lsm_coeff, lsm_valid, lsm_gram, lsm_par, lsm_cnt,
        lsm_part, lsm_dates, plain_pay;
    // one-launch LSM pricing when the request's grid fits on the device at
    // once; MONTEPATH_LSM_MULTI_LAUNCH at open selects per-date launches:
    // 1 matched to the one-launch pricing (same bits), 2 the plain grid-stride
    // passes (the only per-date form for requests too large for one launch)
    bool lsm_fused = true;
    bool lsm_matched = true;
    // polls before a one-launch pricing's wait gives up (MONTEPATH_LSM_SPIN_LIMIT
    // at open, for tests of the fallback), and how often it fell back
    unsigned int lsm_spin_limit = 1u << 20;
    uint64_t lsm_fallbacks = 0;
    int lsm_blocks_per_cu[2][2][5][3] = {};  // [fp64][normals][log2 J][block shape], 0 = not yet queried
    // FP32 Sobol-bridge requests on 16 * 2^J steps keep the bridge in
    // registers; MONTEPATH_QMC_STACK_BRIDGE=1 at open selects the LDS stack
    bool qmc_tree = true;
    dbuf draws;
    dbuf sampler;
    // result row of one-request pricings: coherent pinned host memory that
    // finalize_kernel writes directly and the host reads once the stream has
    // synchronised (no copy command)
    double* h_res = nullptr;      // 32 doubles: the result row, then a wait-error word
    double* h_res_dev = nullptr;
    // small host-to-device inputs of one call (request parameters, plans,
    // tables) go through this pinned buffer as asynchronous copies; a call
    // synchronises its stream before it returns, so the next one reuses it
    // from the start
    hbuf h_stage;
    size_t stage_off = 0;
};

namespace {

// Host-to-device copy of a call's small input through the context's pinned
// staging buffer; inputs that do not fit are copied from where they are.
int upload(mp_ctx* c, void* dst, const void* src, size_t bytes) {
    const size_t at = (c->stage_off + 255) & ~(size_t)255;
    if (at + bytes <= c->h_stage.cap) {
        char* st = c->h_stage.as<char>() + at;
        std::memcpy(st, src, bytes);
        c->stage_off = at + bytes;
        HIPCHK(hipMemcpyAsync(dst, st, bytes, hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    }
    return MP_OK;
}

// 32-bit direction numbers from the Joe-Kuo primitive polynomials and
// initial values (Bratley-Fox recurrence).
void sobol_directions(uint32_t dirs[256][32]) {
    for (int d = 0; d < 256; ++d) {
        const uint32_t poly = mp_sobol_poly[d];
        int s = 0;
        while ((poly >> (s + 1)) != 0) ++s;
        uint64_t m[33];
        if (d == 0) {
            for (int k = 1; k <= 32; ++k) m[k] = 1;
        } else {
            for (int k = 1; k <= s; ++k) m[k] = mp_sobol_vinit[d][k - 1];
            for (int k = s + 1; k <= 32; ++k) {
                uint64_t v = m[k - s] ^ (m[k - s] << s);
                for (int i = 1; i < s; ++i)
                    if ((poly >> (s - i)) & 1u) v ^= m[k - i] << i;
                m[k] = v;
            }
        }
        for (int k = 1; k <= 32; ++k) dirs[d][k - 1] = (uint32_t)(m[k] << (32 - k));
    }
}

// Index tables: for each dimension, 128 16-bit words -- entries 0..63 are
// the XORs of direction numbers 0..5 selected by a 6-bit value, 64..95 of
// numbers 6..10 and 96..127 of numbers 11..15 (5-bit values) -- so a 16-bit
// index needs three lookups.  Stored bit-reversed (the kernels scramble the
// reversed point); the first 16 direction numbers have their low 16 bits
// clear, so the reversed XORs fit 16 bits.  Returns false if they do not.
bool sobol_index_tables(const uint32_t dirs[256][32], std::vector<uint32_t>& words) {
    static const int lo[3] = {0, 6, 11}, nb[3] = {6, 5, 5}, at[3] = {0, 64, 96};
    std::vector<uint16_t> tab(256 * 128, 0);
    for (int d = 0; d < 256; ++d) {
        for (int k = 0; k < 16; ++k)
            if (dirs[d][k] & 0xFFFFu) return false;
        for (int j = 0; j < 3; ++j)
            for (int n = 0; n < (1 << nb[j]); ++n) {
                uint32_t x = 0;
                for (int b = 0; b < nb[j]; ++b)
                    if ((n >> b) & 1) x ^= dirs[d][lo[j] + b];
                tab[d * 128 + at[j] + n] = (uint16_t)__builtin_bitreverse32(x);
            }
    }
    words.assign(tab.size() / 2, 0u);
    std::memcpy(words.data(), tab.data(), tab.size() * sizeof(uint16_t));
    return true;
}

// Streaming Brownian-bridge plan (preorder of the reference's
// BrownianBridgePlan::new, cpu.rs:5923-5961) with stack slots.
// c0, c1: the factor scales of the FP32 European Heston register bridge
// (xi and rho_perp); K = log2(n_steps) for power-of-two grids (0 otherwise)
int bridge_plan(uint64_t n_steps, double maturity, std::vector<mp_bbop>& ops, int& n_slots,
                double c0 = 1.0, double c1 = 1.0, int K = 0) {
    ops.clear();
    n_slots = 1;
    const double dt = maturity / (double)n_steps;
    struct item {
        uint64_t l, r;
        int slot;
        int depth;
    };
    std::vector<item> st;
    st.push_back({0, n_steps, 0, 0});
    while (!st.empty()) {
        item it = st.back();
        st.pop_back();
        if (it.r <= it.l + 1) continue;
        const uint64_t mid = (it.l + it.r) / 2;
        const double tl = (double)it.l * dt, tm = (double)mid * dt, tr = (double)it.r * dt;
        const double denom = tr - tl;
        mp_bbop op;
        op.rw = (tm - tl) / denom;
        op.sd = std::sqrt((tm - tl) * (tr - tm) / denom);
        op.rwf = (float)op.rw;
        op.sdf = (float)op.sd;
        op.sdf_e = (float)std::ldexp(op.sd, it.depth + 1);
        op.sdf_c0 = (float)(c0 * std::ldexp(op.sd, it.depth + 1 - K));
        op.sdf_c1 = (float)(c1 * std::ldexp(op.sd, it.depth + 1 - K));
        op.pad_ = 0.f;
        op.slot = it.slot;
        const bool left_leaf = mid <= it.l + 1;
        const bool right_leaf = it.r <= mid + 1;
        op.flags = (left_leaf ? 1 : 0) | (right_leaf ? 2 : 0);
        ops.push_back(op);
        n_slots = std::max(n_slots, it.slot + (left_leaf ? 1 : 2));
        st.push_back({mid, it.r, it.slot, it.depth + 1});
        st.push_back({it.l, mid, it.slot + 1, it.depth + 1});
    }
    return MP_OK;
}

bool finite(double x) { return std::isfinite(x); }

// Paths per work-queue tile: a pure function of the request, so a request's
// tiles -- and hence its FP64 partial sums -- are the same alone and in a mix.
// At most 2^17 path-steps per tile, and at least ~1,024 tiles (one wave
// each) when the request has the paths, so a request priced alone spreads
// over the chip; 64 to 1,024 paths.
uint32_t tile_paths_for(uint64_t n_paths, uint64_t n_steps) {
    const uint64_t by_work = (1ull << 17) / std::max<uint64_t>(n_steps, 1);
    const uint64_t by_count = n_paths / 1024;
    uint64_t t = pow2_floor(std::max<uint64_t>(1, std::min(by_work, by_count)));
    t = std::min<uint64_t>(std::max<uint64_t>(t, 64), 1024);
    return (uint32_t)t;
}

// ---------------------------------------------------------------------------
// Structured-sampler tables, built on the host once per request.
// ---------------------------------------------------------------------------
uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double open01(uint64_t v) {
    const double x = (double)(v >> 11) * (1.0 / 9007199254740992.0);
    return x > DBL_MIN ? x : DBL_MIN;
}

uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b != 0) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// Smallest odd candidate at or after seed mod n (stepping by 2, wrapping to 1)
// that is coprime to n (coprime_multiplier, cpu.rs:5983-6001).
uint64_t coprime_multiplier(uint64_t seed, uint64_t n) {
    if (n <= 1) return 1;
    uint64_t c = std::max<uint64_t>(seed % n, 1);
    if (c % 2 == 0) c += 1;
    while (gcd64(c, n) != 1) {
        c += 2;
        if (c >= n) c = 1;
    }
    return c;
}

// The first n primes by a sieve of Eratosthenes over a bound from Rosser's
// theorem (p_n < n (ln n + ln ln n) for n >= 6).
void first_primes(uint64_t n, std::vector<uint32_t>& out) {
    out.clear();
    const double nd = (double)std::max<uint64_t>(n, 6);
    const uint64_t lim = (uint64_t)(nd * (std::log(nd) + std::log(std::log(nd)))) + 16;
    std::vector<uint8_t> comp(lim + 1, 0);
    for (uint64_t i = 2; i <= lim && out.size() < n; ++i) {
        if (comp[i]) continue;
        out.push_back((uint32_t)i);
        for (uint64_t j = i * i; j <= lim; j += i) comp[j] = 1;
    }
}

// Randomized Halton (RandomizedHalton::new, cpu.rs:5692-5704): dimension d
// takes the d-th prime and the shift open01(splitmix64(seed + d * golden)),
// over max(dims, 2) dimensions.  Latin hypercube (LatinHypercube::new,
// cpu.rs:5737-5762): per dimension a seed mixed from (seed, d), a stride
// multiplier coprime to n_paths, an offset and a jitter seed.
void build_sampler(int kind, uint64_t seed, uint64_t n_paths, uint64_t dims,
                   std::vector<mp_sampler_dim>& out) {
    out.clear();
    if (kind == MP_SAMPLER_HALTON) {
        const uint64_t nd = std::max<uint64_t>(dims, 2);
        std::vector<uint32_t> primes;
        first_primes(nd, primes);
        out.resize(nd);
        for (uint64_t d = 0; d < nd; ++d) {
            out[d].a = primes[d];
            out[d].b = out[d].c = 0;
            out[d].shift = open01(splitmix64(seed + d * 0x9E3779B97F4A7C15ull));
        }
    } else {
        const uint64_t nd = std::max<uint64_t>(dims, 1);
        out.resize(nd);
        for (uint64_t d = 0; d < nd; ++d) {
            const uint64_t mixed = splitmix64(seed ^ (d * 0x9E3779B97F4A7C15ull) ^ 0xD1B54A32D192ED03ull);
            out[d].a = coprime_multiplier(mixed, std::max<uint64_t>(n_paths, 1));
            out[d].b = n_paths == 0 ? 0 : splitmix64(mixed ^ 0xA24BAED4963EE407ull) % n_paths;
            out[d].c = splitmix64(mixed ^ 0x9FB21C651E98DF25ull);
            out[d].shift = 0.0;
        }
    }
}

int prepare_request(const mp_request& r, prepared& P) {
    std::memset(&P.d, 0, sizeof(P.d));
    P.schedule.clear();
    P.weighted = 0;
    P.precision = r.precision == 64 ? 64 : 32;
    if (r.precision != 0 && r.precision != 32 && r.precision != 64)
        return fail(MP_E_INVALID_CONFIG, "precision must be 32 or 64");
    if (r.technique < MP_TECH_STANDARD || r.technique > MP_TECH_CONTROL_VARIATE)
        return fail(MP_E_INVALID_CONFIG, "technique must be 0 (standard), 1 (antithetic) or 2 (control variate)");
    if (r.sampling != MP_PSEUDORANDOM && r.sampling != MP_RANDOMIZED_HALTON &&
        r.sampling != MP_LATIN_HYPERCUBE && r.sampling != MP_SCRAMBLED_SOBOL &&
        r.sampling != MP_SCRAMBLED_SOBOL_BRIDGE)
        return fail(MP_E_INVALID_CONFIG, "unknown sampling method " + std::to_string(r.sampling));
    if (r.n_paths == 0) return fail(MP_E_INVALID_CONFIG, "n_paths must be > 0");
    if (r.n_paths > 0xFFFFFFFFull)
        return fail(MP_E_UNSUPPORTED, "n_paths above 2^32 per request is not supported");
    const double s0 = r.s0, k = r.k, rr = r.r, sig = r.sigma, t = r.t;
    if (!finite(s0) || !finite(k) || !finite(rr) || !finite(sig) || !finite(t))
        return fail(MP_E_INVALID_CONFIG, "non-finite market parameter");
    mp_dreq& d = P.d;
    d.n_paths = r.n_paths;
    d.key0 = (uint32_t)r.seed;
    d.key1 = (uint32_t)r.contract_id;
    d.seed_hi = (uint32_t)(r.seed >> 32);
    d.s0 = s0;
    d.log_s0 = std::log(s0);
    d.rate = rr;
    d.discount = std::exp(-rr * t);
    d.log_barrier_rel = -1e300;
    // strikes
    if (r.n_strikes < 0 || r.n_strikes > MP_MAX_STRIKES)
        return fail(MP_E_INVALID_CONFIG, "n_strikes must be in 0..4");
    if (r.n_strikes <= 1) {
        d.n_strikes = 1;
        d.strikes[0] = k;
    } else {
        d.n_strikes = r.n_strikes;
        for (int i = 0; i < r.n_strikes; ++i) {
            if (!(r.strikes[i] >= 0.0)) return fail(MP_E_INVALID_CONFIG, "strikes must be >= 0");
            d.strikes[i] = r.strikes[i];
        }
    }
    P.kind = KIND_TILE;
    P.factors = 1;
    P.bridge = r.sampling == MP_SCRAMBLED_SOBOL_BRIDGE;
    P.n_steps_cfg = r.n_steps;
    P.sampler.clear();
    const bool qmc = r.sampling == MP_SCRAMBLED_SOBOL || r.sampling == MP_SCRAMBLED_SOBOL_BRIDGE;
    const bool structured = r.sampling == MP_RANDOMIZED_HALTON || r.sampling == MP_LATIN_HYPERCUBE;

    switch (r.workload) {
        case MP_EUROPEAN_CALL:
        case MP_ARITHMETIC_ASIAN_CALL:
        case MP_DOWN_AND_OUT_CALL:
        case MP_LOOKBACK_CALL: {
            if (r.n_steps == 0) return fail(MP_E_INVALID_CONFIG, "n_steps must be > 0");
            if (r.workload == MP_EUROPEAN_CALL) {
                // validate_european_call_config cpu.rs:5010-5017
                if (!(s0 > 0.0)) return fail(MP_E_INVALID_CONFIG, "s0 must be > 0");
                if (!(k > 0.0)) return fail(MP_E_INVALID_CONFIG, "k must be > 0");
                if (!(sig > 0.0)) return fail(MP_E_INVALID_CONFIG, "sigma must be > 0");
                if (!(t > 0.0)) return fail(MP_E_INVALID_CONFIG, "t must be > 0");
            } else if (r.workload == MP_LOOKBACK_CALL) {
                // validate_lookback_call_config
                if (!(s0 > 0.0)) return fail(MP_E_INVALID_CONFIG, "s0 must be > 0");
                if (!(k >= 0.0)) return fail(MP_E_INVALID_CONFIG, "k must be >= 0");
                if (!(sig >= 0.0)) return fail(MP_E_INVALID_CONFIG, "sigma must be >= 0");
                if (!(t > 0.0)) return fail(MP_E_INVALID_CONFIG, "t must be > 0");
            } else {
                if (!(s0 > 0.0)) return fail(MP_E_INVALID_CONFIG, "s0 must be > 0");
                if (!(t > 0.0)) return fail(MP_E_INVALID_CONFIG, "t must be > 0");
                if (!(sig >= 0.0)) return fail(MP_E_INVALID_CONFIG, "sigma must be >= 0");
            }
            if (r.workload == MP_DOWN_AND_OUT_CALL) {
                if (!(r.barrier > 0.0)) return fail(MP_E_INVALID_CONFIG, "barrier must be > 0");
                d.log_barrier_rel = std::log(r.barrier) - d.log_s0;
                d.knocked_at_start = s0 <= r.barrier;
            }
            d.model = MP_MODEL_GBM;
            d.payoff = r.workload == MP_EUROPEAN_CALL ? MP_PAY_EUROPEAN
                       : r.workload == MP_ARITHMETIC_ASIAN_CALL ? MP_PAY_ASIAN
                       : r.workload == MP_DOWN_AND_OUT_CALL ? MP_PAY_DOWN_AND_OUT
                                                              : MP_PAY_LOOKBACK;
            uint64_t m = r.n_steps;
            if (r.workload == MP_EUROPEAN_CALL && r.method == MP_METHOD_TERMINAL) m = 1;
            if (m > 0x7FFFFFFFull) return fail(MP_E_UNSUPPORTED, "n_steps too large");
            d.n_steps = (int32_t)m;
            d.dt = t / (double)m;
            d.sqrt_dt = std::sqrt(d.dt);
            d.drift_dt = (rr - 0.5 * sig * sig) * d.dt;
            d.vol_dt = sig * std::sqrt(d.dt);
            d.sigma = sig;
            if (qmc) {
                P.kind = KIND_QMC;
                P.factors = 1;
            }
            break;
        }
        case MP_HESTON_EUROPEAN_CALL: {
            // validate_heston_european_call_config cpu.rs (Heston block)
            if (r.n_steps == 0) return fail(MP_E_INVALID_CONFIG, "n_steps must be > 0");
            if (!(s0 > 0.0)) return fail(MP_E_INVALID_CONFIG, "s0 must be > 0");
            if (!(k >= 0.0)) return fail(MP_E_INVALID_CONFIG, "k must be >= 0");
            if (!(r.v0 >= 0.0)) return fail(MP_E_INVALID_CONFIG, "v0 must be >= 0");
            if (!(r.kappa >= 0.0)) return fail(MP_E_INVALID_CONFIG, "kappa must be >= 0");
            if (!(r.theta >= 0.0)) return fail(MP_E_INVALID_CONFIG, "theta must be >= 0");
            if (!(r.xi >= 0.0)) return fail(MP_E_INVALID_CONFIG, "vol_of_vol must be >= 0");
            if (!(r.rho >= -1.0 && r.rho <= 1.0))
                return fail(MP_E_INVALID_CONFIG, "rho must be in [-1, 1]");
            if (!(t > 0.0)) return fail(MP_E_INVALID_CONFIG, "t must be > 0");
            if (r.heston_scheme != 0 && r.heston_scheme != 1)
                return fail(MP_E_INVALID_CONFIG, "heston_scheme must be 0 or 1");
            if (r.n_steps > 0x7FFFFFFFull) return fail(MP_E_UNSUPPORTED, "n_steps too large");
            d.model = MP_MODEL_HESTON;
            d.payoff = MP_PAY_EUROPEAN;
            d.scheme = r.heston_scheme;
            d.n_steps = (int32_t)r.n_steps;
            d.v0 = r.v0;
            d.kappa = r.kappa;
            d.theta = r.theta;
            d.xi = r.xi;
            d.rho = r.rho;
            const double rp = 1.0 - r.rho * r.rho;
            d.rho_perp = std::sqrt(rp > 0.0 ? rp : 0.0);
            d.dt = t / (double)r.n_steps;
            d.sqrt_dt = std::sqrt(t / (double)r.n_steps);
            if (qmc) {
                P.kind = KIND_QMC;
                P.factors = 2;
            }
            break;
        }
        case MP_MERTON_JUMP_DIFFUSION_CALL: {
            // validate_merton_jump_diffusion_call_config
            if (!(s0 > 0.0)) return fail(MP_E_INVALID_CONFIG, "s0 must be > 0");
            if (!(k > 0.0)) return fail(MP_E_INVALID_CONFIG, "k must be > 0");
            if (!(sig >= 0.0)) return fail(MP_E_INVALID_CONFIG, "sigma must be >= 0");
            if (!(r.jump_intensity >= 0.0))
                return fail(MP_E_INVALID_CONFIG, "jump_intensity must be >= 0");
            if (!(r.jump_volatility >= 0.0))
                return fail(MP_E_INVALID_CONFIG, "jump_volatility must be >= 0");
            if (!(t > 0.0)) return fail(MP_E_INVALID_CONFIG, "t must be > 0");
            if (qmc || structured)
                return fail(MP_E_UNSUPPORTED, "Merton is priced with pseudorandom draws only");
            const uint64_t m = r.n_steps == 0 ? 1 : r.n_steps;  // reference Merton is terminal
            if (m > 0x7FFFFFFFull) return fail(MP_E_UNSUPPORTED, "n_steps too large");
            d.model = MP_MODEL_MERTON;
            d.payoff = MP_PAY_EUROPEAN;
            d.n_steps = (int32_t)m;
            d.dt = t / (double)m;
            d.sqrt_dt = std::sqrt(d.dt);
            const double comp = r.jump_intensity *
                                (std::exp(r.jump_mean + 0.5 * (r.jump_volatility * r.jump_volatility)) - 1.0);
            d.drift_dt = (rr - comp - 0.5 * sig * sig) * d.dt;
            d.vol_dt = sig * std::sqrt(d.dt);
            d.mu_j = r.jump_mean;
            d.sigma_j = r.jump_volatility;
            d.jump_mean_dt = r.jump_intensity * d.dt;
            // cumulative Poisson probabilities by the reference recursion
            double prob = std::exp(-d.jump_mean_dt), cum = prob;
            d.n_cum = MP_POISSON_TERMS;
            for (int i = 0; i < MP_POISSON_TERMS; ++i) {
                d.cum[i] = cum;
                prob *= d.jump_mean_dt / (double)(i + 1);
                cum += prob;
            }
            P.n_steps_cfg = m;
            break;
        }
        case MP_BASKET_CALL: {
            if (r.n_assets < 1 || r.n_assets > MP_MAX_ASSETS)
                return fail(MP_E_INVALID_CONFIG, "n_assets must be in 1..8");
            if (!(k >= 0.0)) return fail(MP_E_INVALID_CONFIG, "k must be >= 0");
            if (!(t > 0.0)) return fail(MP_E_INVALID_CONFIG, "t must be > 0");
            if (qmc) return fail(MP_E_UNSUPPORTED, "basket is priced with pseudorandom draws only");
            for (int a = 0; a < r.n_assets; ++a) {
                if (!(r.asset_s0[a] > 0.0)) return fail(MP_E_INVALID_CONFIG, "asset s0 must be > 0");
                if (!(r.asset_sigma[a] >= 0.0))
                    return fail(MP_E_INVALID_CONFIG, "asset sigma must be >= 0");
            }
            const uint64_t m = r.n_steps == 0 ? 1 : r.n_steps;
            d.model = -1;
            d.n_steps = (int32_t)m;
            d.dt = t / (double)m;
            d.n_strikes = 1;
            d.strikes[0] = k;
            P.kind = KIND_BASKET;
            P.n_assets = r.n_assets;
            P.n_steps_cfg = m;
            break;
        }
        case MP_AMERICAN_PUT:
        case MP_BERMUDAN_PUT: {
            // validate_american_put_config / validate_bermudan_put_config
            if (r.n_steps < 2) return fail(MP_E_INVALID_CONFIG, "n_steps must be >= 2 for LSM regression");
            if (r.n_steps > 65535) return fail(MP_E_UNSUPPORTED, "LSM supports at most 65535 steps");
            if (!(s0 > 0.0)) return fail(MP_E_INVALID_CONFIG, "s0 must be > 0");
            if (!(k > 0.0)) return fail(MP_E_INVALID_CONFIG, "k must be > 0");
            if (!(sig >= 0.0)) return fail(MP_E_INVALID_CONFIG, "sigma must be >= 0");
            if (!(t > 0.0)) return fail(MP_E_INVALID_CONFIG, "t must be > 0");
            if (r.basis_degree <= 0) return fail(MP_E_INVALID_CONFIG, "basis_degree must be > 0");
            if (qmc || structured)
                return fail(MP_E_UNSUPPORTED, "LSM is priced with pseudorandom draws only");
            P.kind = KIND_LSM;
            P.degree = std::min(std::max(r.basis_degree, 1), 3);
            P.weighted = r.basis_weighted != 0;
            if (r.workload == MP_AMERICAN_PUT) {
                for (uint32_t s = 1; s <= r.n_steps; ++s) P.schedule.push_back(s);
                P.allow_intrinsic = 1;
            } else {
                if (r.n_exercise <= 0 || !r.exercise_steps)
                    return fail(MP_E_INVALID_CONFIG, "exercise_steps must include at least one exercise date");
                for (int i = 0; i < r.n_exercise; ++i) {
                    const uint32_t s = r.exercise_steps[i];
                    if (s == 0 || s > r.n_steps)
                        return fail(MP_E_INVALID_CONFIG, "exercise_steps must be in 1..=n_steps");
                    P.schedule.push_back(s);
                }
                // normalize_bermudan_exercise_schedule: sort and deduplicate
                std::sort(P.schedule.begin(), P.schedule.end());
                P.schedule.erase(std::unique(P.schedule.begin(), P.schedule.end()), P.schedule.end());
                P.allow_intrinsic = 0;
            }
            if (r.workload == MP_AMERICAN_PUT && r.allow_intrinsic_now == 0) P.allow_intrinsic = 1;
            d.n_steps = (int32_t)r.n_steps;
            break;
        }
        default:
            return fail(MP_E_UNSUPPORTED, "unknown workload " + std::to_string(r.workload));
    }
    if (r.basis_weighted && P.kind != KIND_LSM)
        return fail(MP_E_INVALID_CONFIG, "basis_weighted applies to American and Bermudan puts only");
    P.n_paths_req = r.n_paths;
    if (structured) {
        // one point per path (or antithetic pair); dimensions: step s of the
        // GBM family, factor f of Heston step s at 2s + f, asset a of basket
        // step s at s * n_assets + a
        const uint64_t dims = d.model == MP_MODEL_HESTON ? 2ull * (uint64_t)d.n_steps
                              : P.kind == KIND_BASKET  ? (uint64_t)d.n_steps * (uint64_t)P.n_assets
                                                       : (uint64_t)d.n_steps;
        if (dims > MP_SAMPLER_MAX_DIMS)
            return fail(MP_E_UNSUPPORTED, "Halton and Latin hypercube sampling take at most " +
                                              std::to_string(MP_SAMPLER_MAX_DIMS) + " dimensions");
        d.sampling = r.sampling == MP_RANDOMIZED_HALTON ? MP_SAMPLER_HALTON : MP_SAMPLER_LHS;
        d.sampler_n = r.n_paths;
        build_sampler(d.sampling, r.seed, r.n_paths, dims, P.sampler);
    }
    if (r.technique != MP_TECH_STANDARD) {
        // the reference offers techniques for the GBM family, Heston and the
        // basket (MonteCarloTechnique, cpu.rs:24-28); Merton and LSM have none
        if (d.model == MP_MODEL_MERTON || P.kind == KIND_LSM)
            return fail(MP_E_UNSUPPORTED, "this workload is priced with the standard technique only");
        if (r.technique == MP_TECH_CONTROL_VARIATE && d.n_strikes > 1)
            return fail(MP_E_INVALID_CONFIG, "the control variate prices one strike per request");
        d.technique = r.technique;
        if (P.kind == KIND_BASKET) {
            // basket_control_expectation cpu.rs:7637-7639: sum of w_i S0_i
            double e = 0.0;
            for (int a = 0; a < r.n_assets; ++a) e += r.asset_weight[a] * r.asset_s0[a];
            d.cv_expect = e;
        } else {
            d.cv_expect = s0;  // E[discount * S_T] = S0 (summarize_control_variate(.., cfg.s0))
        }
        // antithetic: a unit is a pair, ceil(n / 2) of them (cpu.rs:7812)
        if (r.technique == MP_TECH_ANTITHETIC) d.n_paths = (r.n_paths + 1) / 2;
    }
    if (P.kind == KIND_TILE || P.kind == KIND_BASKET) {
        d.tile_paths = tile_paths_for(d.n_paths, (uint64_t)d.n_steps);
        d.n_tiles = (uint32_t)((d.n_paths + d.tile_paths - 1) / d.tile_paths);
    }
    return MP_OK;
}

double cost_of(const mp_dreq& d) {
    const double w = d.model == MP_MODEL_HESTON ? 2.5 : d.model == MP_MODEL_MERTON ? 1.4
                     : d.payoff == MP_PAY_ASIAN ? 1.2 : 1.0;
    return w * (double)d.n_steps;
}

void fill_result(mp_result* out, const double* fin, const prepared& P) {
    out->status = MP_OK;
    out->n_strikes = P.d.n_strikes;
    for (int j = 0; j < MP_MAX_STRIKES; ++j) {
        out->strike_price[j] = j < P.d.n_strikes ? fin[2 * j] : 0.0;
        out->strike_std_error[j] = j < P.d.n_strikes ? fin[2 * j + 1] : 0.0;
    }
    out->price = fin[0];
    out->std_error = fin[1];
    out->n_paths = P.n_paths_req;
    out->n_steps = P.n_steps_cfg;
    out->fp64_ref_error = NAN;
}

int grid_for_tiles(const mp_ctx* c, uint32_t n_tiles, int block) {
    const int wpb = block / 64;
    const int want = (int)((n_tiles + wpb - 1) / wpb);
    return std::max(1, std::min(want, c->n_cu * 8));
}

uint64_t tuning_key(const mp_dreq& d, int precision) {
    const uint64_t bucket = d.n_steps <= 32 ? 0 : d.n_steps <= 128 ? 1 : d.n_steps <= 512 ? 2 : 3;
    return ((uint64_t)(d.model + 1) << 24) | ((uint64_t)d.payoff << 16) | (bucket << 8) |
           (uint64_t)(precision == 64);
}

mp_tuning tuning_for(const mp_ctx* c, const mp_dreq& d, int precision) {
    auto it = c->tuning.find(tuning_key(d, precision));
    if (it != c->tuning.end()) return it->second;
    mp_tuning t{};
    t.block_size = 256;
    t.paths_per_lane = 1;
    t.tile_paths = (int32_t)d.tile_paths;
    t.step_unroll = 1;
    return t;
}

}  // namespace

// ===========================================================================
// Batches: a set of tile-kernel requests packed into one persistent launch.
// ===========================================================================
struct mp_batch {
    mp_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int precision = 32;
    std::vector<prepared> reqs;     // tile-kernel requests of this batch
    std::vector<size_t> req_index;  // position in the caller's array
    uint32_t n_tiles = 0;
    int grid = 1, block = 256, ppl = 1, unroll = 1;
    bool mlmc = false;
    bool gbm_f32 = false;  // every request takes the FP32 GBM-family Philox body
    dbuf d_reqs, d_tile_req, d_order, d_partial, d_counter, d_sampler;
    hbuf h_stage;  // request records, tile map, queue order, sampler tables
    hbuf h_out;    // result rows, written by finalize_kernel
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = false;
    ~mp_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (e2) (void)hipEventDestroy(e2);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// A tile the FP32 GBM-family Philox body prices (tile_gbm_kernel).
bool gbm_philox_tile(const mp_dreq& d) {
    return d.model == MP_MODEL_GBM && d.technique == MP_TECH_STANDARD && d.sampling == MP_SAMPLER_NONE &&
           d.payoff >= MP_PAY_EUROPEAN && d.payoff <= MP_PAY_LOOKBACK;
}

int batch_build(mp_batch* b) {
    const size_t n = b->reqs.size();
    std::vector<mp_dreq> dreqs(n);
    uint32_t tiles = 0;
    std::vector<mp_sampler_dim> sampler;  // every structured request's dimensions
    for (size_t i = 0; i < n; ++i) {
        dreqs[i] = b->reqs[i].d;
        dreqs[i].tile0 = tiles;
        if (dreqs[i].sampling != MP_SAMPLER_NONE) {
            dreqs[i].sampler_off = (uint32_t)sampler.size();
            sampler.insert(sampler.end(), b->reqs[i].sampler.begin(), b->reqs[i].sampler.end());
        }
        tiles += dreqs[i].n_tiles;
    }
    b->n_tiles = tiles;
    std::vector<uint32_t> tile_req(tiles), order(tiles);
    for (size_t i = 0; i < n; ++i)
        for (uint32_t t = 0; t < dreqs[i].n_tiles; ++t) tile_req[dreqs[i].tile0 + t] = (uint32_t)i;
    // Longest-processing-time-first queue order: costly tiles are pulled first.
    for (uint32_t t = 0; t < tiles; ++t) order[t] = t;
    std::vector<double> cost(n);
    for (size_t i = 0; i < n; ++i) cost[i] = cost_of(dreqs[i]) * dreqs[i].tile_paths;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
        return cost[tile_req[a]] > cost[tile_req[c]];
    });
    ENSURE(b->d_reqs, n * sizeof(mp_dreq));
    ENSURE(b->d_tile_req, (size_t)tiles * 4);
    ENSURE(b->d_order, (size_t)tiles * 4);
    ENSURE(b->d_partial, (size_t)tiles * MP_PARTIAL_STRIDE * 8);
    ENSURE(b->d_counter, 64);
    if (!sampler.empty()) ENSURE(b->d_sampler, sampler.size() * sizeof(mp_sampler_dim));
    // Stage the inputs in pinned memory and copy them asynchronously: the
    // staging outlives the copies (a batch is rebuilt only after its stream
    // has completed), and the launches follow them on the same stream.
    const size_t o_req = 0;
    const size_t o_tile = (o_req + n * sizeof(mp_dreq) + 255) & ~(size_t)255;
    const size_t o_order = (o_tile + (size_t)tiles * 4 + 255) & ~(size_t)255;
    const size_t o_samp = (o_order + (size_t)tiles * 4 + 255) & ~(size_t)255;
    ENSURE(b->h_stage, o_samp + sampler.size() * sizeof(mp_sampler_dim));
    char* st = b->h_stage.as<char>();
    std::memcpy(st + o_req, dreqs.data(), n * sizeof(mp_dreq));
    std::memcpy(st + o_tile, tile_req.data(), (size_t)tiles * 4);
    std::memcpy(st + o_order, order.data(), (size_t)tiles * 4);
    if (!sampler.empty()) std::memcpy(st + o_samp, sampler.data(), sampler.size() * sizeof(mp_sampler_dim));
    HIPCHK(hipMemcpyAsync(b->d_reqs.p, st + o_req, n * sizeof(mp_dreq), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_tile_req.p, st + o_tile, (size_t)tiles * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_order.p, st + o_order, (size_t)tiles * 4, hipMemcpyHostToDevice, b->stream));
    if (!sampler.empty())
        HIPCHK(hipMemcpyAsync(b->d_sampler.p, st + o_samp, sampler.size() * sizeof(mp_sampler_dim),
                              hipMemcpyHostToDevice, b->stream));
    ENSURE(b->h_out, n * MP_PARTIAL_STRIDE * 8);
    // launch shape: tuned for the dominant request class
    mp_tuning tu = tuning_for(b->ctx, dreqs.empty() ? mp_dreq{} : dreqs[0], b->precision);
    b->block = tu.block_size > 0 ? tu.block_size : 256;
    b->ppl = tu.paths_per_lane > 0 ? tu.paths_per_lane : 1;
    b->unroll = tu.step_unroll > 0 ? tu.step_unroll : 1;
    size_t n_mlmc = 0;
    for (const mp_dreq& d : dreqs) n_mlmc += d.payoff == MP_PAY_ASIAN_MLMC;
    if (n_mlmc != 0 && n_mlmc != dreqs.size())
        return fail(MP_E_INVALID_CONFIG, "multilevel levels run in batches of their own");
    b->mlmc = n_mlmc != 0;
    if (b->mlmc) b->ppl = 1;  // the multilevel kernel runs one path per lane
    b->gbm_f32 = b->precision == 32 && std::all_of(dreqs.begin(), dreqs.end(), gbm_philox_tile);
    b->grid = grid_for_tiles(b->ctx, tiles, b->block);
    return MP_OK;
}

int batch_enqueue(mp_batch* b) {
    if (b->reqs.empty()) return MP_OK;
    if (b->use_graph && b->exec) {
        HIPCHK(hipGraphLaunch(b->exec, b->stream));
        return MP_OK;
    }
    mp::tile_launch L{};
    L.reqs = b->d_reqs.as<mp_dreq>();
    L.tile_req = b->d_tile_req.as<uint32_t>();
    L.order = b->d_order.as<uint32_t>();
    L.n_tiles = b->n_tiles;
    L.counter = b->d_counter.as<unsigned int>();
    L.partial = b->d_partial.as<double>();
    L.sampler = b->d_sampler.as<mp_sampler_dim>();
    L.nib = b->ctx->d_nib;
    L.mlmc = b->mlmc ? 1 : 0;
    L.gbm_f32 = b->gbm_f32 ? 1 : 0;
    L.grid = b->grid;
    L.block = b->block;
    L.ppl = b->ppl;
    L.unroll = b->unroll;
    HIPCHK(mp::launch_reset(L.counter, 1, b->stream));
    HIPCHK(hipEventRecord(b->e0, b->stream));
    HIPCHK(mp::launch_tiles(L, b->precision, b->stream));
    HIPCHK(hipEventRecord(b->e1, b->stream));
    HIPCHK(mp::launch_finalize(L.reqs, (int)b->reqs.size(), L.partial, static_cast<double*>(b->h_out.dev), b->stream));
    HIPCHK(hipEventRecord(b->e2, b->stream));
    return MP_OK;
}

// Capture reset + tiles + finalize into a graph (no memset nodes: the queue
// counter is cleared by a kernel).
int batch_capture(mp_batch* b) {
    if (b->reqs.empty()) return MP_OK;
    mp::tile_launch L{};
    L.reqs = b->d_reqs.as<mp_dreq>();
    L.tile_req = b->d_tile_req.as<uint32_t>();
    L.order = b->d_order.as<uint32_t>();
    L.n_tiles = b->n_tiles;
    L.counter = b->d_counter.as<unsigned int>();
    L.partial = b->d_partial.as<double>();
    L.sampler = b->d_sampler.as<mp_sampler_dim>();
    L.nib = b->ctx->d_nib;
    L.mlmc = b->mlmc ? 1 : 0;
    L.gbm_f32 = b->gbm_f32 ? 1 : 0;
    L.grid = b->grid;
    L.block = b->block;
    L.ppl = b->ppl;
    L.unroll = b->unroll;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = mp::launch_reset(L.counter, 1, b->stream);
    hipError_t e2 = mp::launch_tiles(L, b->precision, b->stream);
    hipError_t e3 = mp::launch_finalize(L.reqs, (int)b->reqs.size(), L.partial, static_cast<double*>(b->h_out.dev), b->stream);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(b->stream, &g));
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(MP_E_HIP, "kernel launch failed during graph capture");
    }
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->use_graph = true;
    return MP_OK;
}

int batch_collect(mp_batch* b, mp_result* results, bool timed) {
    if (b->reqs.empty()) return MP_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    float main_ms = 0.f, fin_ms = 0.f;
    if (timed && !b->use_graph) {
        (void)hipEventElapsedTime(&main_ms, b->e0, b->e1);
        (void)hipEventElapsedTime(&fin_ms, b->e1, b->e2);
        b->ctx->last.main_kernel_ms = main_ms;
        b->ctx->last.finalize_ms = fin_ms;
        b->ctx->last.main_launches = 1;
    }
    for (size_t i = 0; i < b->reqs.size(); ++i) {
        fill_result(&results[b->req_index[i]], b->h_out.as<double>() + i * MP_PARTIAL_STRIDE, b->reqs[i]);
        results[b->req_index[i]].runtime_ms = main_ms + fin_ms;
    }
    return MP_OK;
}

int batch_init(mp_ctx* ctx, mp_batch* b, hipStream_t stream, int precision) {
    b->ctx = ctx;
    b->precision = precision;
    if (stream) {
        b->stream = stream;
    } else {
        HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->own_stream = true;
    }
    HIPCHK(hipEventCreate(&b->e0));
    HIPCHK(hipEventCreate(&b->e1));
    HIPCHK(hipEventCreate(&b->e2));
    return MP_OK;
}

// ---------------------------------------------------------------------------
// Single-request paths that are not work-queue tiles
// ---------------------------------------------------------------------------
int run_qmc(mp_ctx* c, const mp_request& r, prepared& P, int precision, mp_result* out,
            double* term_host, double* pay_host) {
    c->stage_off = 0;  // the previous call has synchronised
    mp_dreq d = P.d;
    std::vector<mp_bbop> ops;
    int n_slots = 1;
    mp_qmc_params qp{};
    qp.base_seed = (uint32_t)r.seed;
    qp.factors = P.factors;
    qp.bridge = P.bridge;
    qp.sqrt_t = std::sqrt(r.t);
    qp.inv_sqrt_dt = 1.0 / std::sqrt(r.t / (double)d.n_steps);
    int log2_steps = 0;
    while ((1ull << (log2_steps + 1)) <= (uint64_t)d.n_steps) ++log2_steps;
    const int K = (1ull << log2_steps) == (uint64_t)d.n_steps ? log2_steps : 0;
    qp.root_c0 = r.xi * std::ldexp(std::sqrt(r.t), -K);
    qp.root_c1 = std::sqrt(std::max(0.0, 1.0 - r.rho * r.rho)) * std::ldexp(std::sqrt(r.t), -K);
    qp.rho_over_xi = r.xi != 0.0 ? r.rho / r.xi : 0.0;
    if (P.bridge)
        bridge_plan((uint64_t)d.n_steps, r.t, ops, n_slots, r.xi,
                    std::sqrt(std::max(0.0, 1.0 - r.rho * r.rho)), K);
    qp.n_ops = (int)ops.size();
    qp.n_slots = n_slots;
    // bridges on 16 * 2^J steps run in registers when the Sobol tables sit
    // in LDS (no LDS stack)
    const int tree = P.bridge && c->qmc_tree ? mp::qmc_tree_levels((uint64_t)d.n_steps) : -1;
    qp.tree_levels = -1;
    // block size: largest power of two <= 1024 whose LDS fits
    // block size: largest power of two <= 1024 whose LDS (plus 2 KB of static
    // reduction space) fits; the Sobol tables move to global memory if the
    // per-block LDS cannot hold them
    const size_t reserve = 2048;
    bool nib_lds = true;
    // FP64 kernels are compiled for blocks of at most 512 threads
    const int max_block = precision == 64 ? 512 : 1024;
    int block = max_block;
    while (block > 64 && mp::qmc_lds_bytes(precision, P.factors, n_slots, block, true) + reserve > c->max_lds)
        block /= 2;
    if (mp::qmc_lds_bytes(precision, P.factors, n_slots, block, true) + reserve > c->max_lds) {
        nib_lds = false;
        block = max_block;
        while (block > 64 && mp::qmc_lds_bytes(precision, P.factors, n_slots, block, false) + reserve > c->max_lds)
            block /= 2;
        if (mp::qmc_lds_bytes(precision, P.factors, n_slots, block, false) + reserve > c->max_lds)
            return fail(MP_E_UNSUPPORTED, "Brownian-bridge stack does not fit in LDS for this step count");
    }
    if (nib_lds && tree >= 0) {
        qp.tree_levels = tree;
        n_slots = 0;
        block = max_block;
    }
    const size_t lds = mp::qmc_lds_bytes(precision, P.factors, n_slots, block, nib_lds) + reserve;
    const int per_cu = std::max<int>(1, (int)(c->max_lds / std::max<size_t>(lds, 1)));
    d.tile_paths = (uint32_t)block;
    d.n_tiles = (uint32_t)((d.n_paths + block - 1) / block);
    d.tile0 = 0;
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(d.n_tiles, (uint64_t)c->n_cu * per_cu));
    ENSURE(c->reqs, sizeof(mp_dreq));
    ENSURE(c->plan, std::max<size_t>(ops.size(), 1) * sizeof(mp_bbop));
    ENSURE(c->partial, (size_t)d.n_tiles * MP_PARTIAL_STRIDE * 8);
    if (term_host) ENSURE(c->term, d.n_paths * 8);
    if (pay_host) ENSURE(c->pay, d.n_paths * 8);
    HIPCHK(mp::launch_setup(d, c->reqs.as<mp_dreq>(), nullptr, 0, c->stream));
    if (!ops.empty())
        RETCHK(upload(c, c->plan.p, ops.data(), ops.size() * sizeof(mp_bbop)));
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    // techniques run in the generic kernel (payoff -1 selects it)
    // (a Heston request with xi = 0 too: the European register-bridge
    // kernels scale the variance factor by xi)
    const int spec_payoff =
        d.technique == MP_TECH_STANDARD && !(P.factors == 2 && r.xi == 0.0) ? d.payoff : -1;
    HIPCHK(mp::launch_qmc(c->reqs.as<mp_dreq>(), qp, spec_payoff, d.scheme, c->d_nib, c->plan.as<mp_bbop>(),
                          c->partial.as<double>(), term_host ? c->term.as<double>() : nullptr,
                          pay_host ? c->pay.as<double>() : nullptr, precision, grid, block, nib_lds,
                          c->stream));
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(mp::launch_finalize(c->reqs.as<mp_dreq>(), 1, c->partial.as<double>(), c->h_res_dev, c->stream));
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    if (term_host) HIPCHK(hipMemcpyAsync(term_host, c->term.p, d.n_paths * 8, hipMemcpyDeviceToHost, c->stream));
    if (pay_host) HIPCHK(hipMemcpyAsync(pay_host, c->pay.p, d.n_paths * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double fin[MP_PARTIAL_STRIDE];
    std::memcpy(fin, c->h_res, sizeof(fin));
    float main_ms = 0.f, fin_ms = 0.f;
    (void)hipEventElapsedTime(&main_ms, c->ev[0], c->ev[1]);
    (void)hipEventElapsedTime(&fin_ms, c->ev[1], c->ev[2]);
    c->last.main_kernel_ms = main_ms;
    c->last.finalize_ms = fin_ms;
    c->last.main_launches = 1;
    fill_result(out, fin, P);
    out->runtime_ms = main_ms + fin_ms;
    return MP_OK;
}

int run_basket(mp_ctx* c, const mp_request& r, prepared& P, int precision, const double* normals_host,
               mp_result* out, double* term_host, double* pay_host) {
    c->stage_off = 0;  // the previous call has synchronised
    const int na = P.n_assets;
    // Cholesky of the correlation matrix
    std::vector<double> corr(na * na), L(na * na, 0.0);
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < na; ++j) {
            double v = r.correlation[i * na + j];
            if (i == j && v == 0.0) v = 1.0;  // zero-initialised requests mean identity diagonal
            corr[i * na + j] = v;
        }
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < na; ++j)
            if (std::fabs(corr[i * na + j] - corr[j * na + i]) > 1e-12)
                return fail(MP_E_INVALID_CONFIG, "correlation matrix must be symmetric");
    for (int i = 0; i < na; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = corr[i * na + j];
            for (int q = 0; q < j; ++q) s -= L[i * na + q] * L[j * na + q];
            if (i == j) {
                if (!(s > 0.0)) return fail(MP_E_INVALID_CONFIG, "correlation matrix is not positive definite");
                L[i * na + i] = std::sqrt(s);
            } else {
                L[i * na + j] = s / L[j * na + j];
            }
        }
    const double dt = r.t / (double)P.d.n_steps;
    std::vector<double> par(na * na + 4 * na);
    for (int i = 0; i < na * na; ++i) par[i] = L[i];
    for (int a = 0; a < na; ++a) {
        par[na * na + a] = (r.r - 0.5 * r.asset_sigma[a] * r.asset_sigma[a]) * dt;
        par[na * na + na + a] = r.asset_sigma[a] * std::sqrt(dt);
        par[na * na + 2 * na + a] = std::log(r.asset_s0[a]);
        par[na * na + 3 * na + a] = r.asset_weight[a];
    }
    mp_dreq d = P.d;
    const int block = 256;
    d.tile_paths = block;
    d.n_tiles = (uint32_t)((d.n_paths + block - 1) / block);
    d.tile0 = 0;
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(d.n_tiles, (uint64_t)c->n_cu * 8));
    ENSURE(c->reqs, sizeof(mp_dreq));
    ENSURE(c->basket_par, par.size() * 8);
    ENSURE(c->partial, (size_t)d.n_tiles * MP_PARTIAL_STRIDE * 8);
    const size_t nn = (size_t)d.n_paths * d.n_steps * na;
    if (normals_host) ENSURE(c->normals, nn * 8);
    if (term_host) ENSURE(c->term, d.n_paths * 8);
    if (pay_host) ENSURE(c->pay, d.n_paths * 8);
    if (normals_host) d.sampling = MP_SAMPLER_NONE;  // caller buffers replace every draw
    if (d.sampling != MP_SAMPLER_NONE) {
        d.sampler_off = 0;
        ENSURE(c->sampler, P.sampler.size() * sizeof(mp_sampler_dim));
        RETCHK(upload(c, c->sampler.p, P.sampler.data(), P.sampler.size() * sizeof(mp_sampler_dim)));
    }
    HIPCHK(mp::launch_setup(d, c->reqs.as<mp_dreq>(), nullptr, 0, c->stream));
    RETCHK(upload(c, c->basket_par.p, par.data(), par.size() * 8));
    if (normals_host) HIPCHK(hipMemcpyAsync(c->normals.p, normals_host, nn * 8, hipMemcpyHostToDevice, c->stream));
    const double* bp = c->basket_par.as<double>();
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    HIPCHK(mp::launch_basket(c->reqs.as<mp_dreq>(), c->sampler.as<mp_sampler_dim>(), na,
                             d.sampling != MP_SAMPLER_NONE, d.technique, bp,
                             bp + na * na, bp + na * na + na,
                             bp + na * na + 2 * na, bp + na * na + 3 * na,
                             normals_host ? c->normals.as<double>() : nullptr, c->partial.as<double>(),
                             term_host ? c->term.as<double>() : nullptr,
                             pay_host ? c->pay.as<double>() : nullptr, precision, grid, block, c->stream));
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(mp::launch_finalize(c->reqs.as<mp_dreq>(), 1, c->partial.as<double>(), c->h_res_dev, c->stream));
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    if (term_host) HIPCHK(hipMemcpyAsync(term_host, c->term.p, d.n_paths * 8, hipMemcpyDeviceToHost, c->stream));
    if (pay_host) HIPCHK(hipMemcpyAsync(pay_host, c->pay.p, d.n_paths * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double fin[MP_PARTIAL_STRIDE];
    std::memcpy(fin, c->h_res, sizeof(fin));
    float main_ms = 0.f, fin_ms = 0.f;
    (void)hipEventElapsedTime(&main_ms, c->ev[0], c->ev[1]);
    (void)hipEventElapsedTime(&fin_ms, c->ev[1], c->ev[2]);
    c->last.main_kernel_ms = main_ms;
    c->last.finalize_ms = fin_ms;
    c->last.main_launches = 1;
    fill_result(out, fin, P);
    out->runtime_ms = main_ms + fin_ms;
    return MP_OK;
}

// Block shapes of the one-launch LSM pricing, in order of preference:
// threads per block and the largest log2 of paths per thread.
struct lsm_shape {
    int bs, max_lj;
};
constexpr lsm_shape kLsmShapes[] = {{512, 3}, {256, 4}, {1024, 2}};
inline int lsm_shape_index(int bs) { return bs == 1024 ? 2 : (bs == 512 ? 1 : 0); }

// One LSM pricing.  per_date forces the matched per-date launches; *gave_up
// reports a one-launch pricing whose bounded waits ran out (its result is
// then not valid).
int run_lsm_once(mp_ctx* c, const mp_request& r, prepared& P, int precision, const double* normals_host,
                 mp_result* out, double* coeff_host, uint32_t* ex_host, bool per_date, bool* gave_up) {
    *gave_up = false;
    c->stage_off = 0;  // the previous call has synchronised
    const uint64_t n = r.n_paths;
    const int m = (int)r.n_steps;
    mp_lsm_params p{};
    p.n_steps = m;
    p.degree = P.degree;
    p.coeffs = P.degree + 1;
    p.weighted = (uint32_t)P.weighted;
    p.allow_intrinsic = P.allow_intrinsic;
    p.n_paths = n;
    p.s0 = r.s0;
    p.log_s0 = std::log(r.s0);
    p.k = r.k;
    // the reference divides by k.max(f64::MIN_POSITIVE); k > 0 is validated
    p.inv_k = 1.0 / r.k;
    const double dt = r.t / (double)m;
    p.drift_dt = (r.r - 0.5 * r.sigma * r.sigma) * dt;
    p.vol_dt = r.sigma * std::sqrt(dt);
    p.disc_step = std::exp(-r.r * dt);
    p.disc_m = powi_ref(p.disc_step, m);
    p.key0 = (uint32_t)r.seed;
    p.key1 = (uint32_t)r.contract_id;
    p.seed_hi = (uint32_t)(r.seed >> 32);
    // parameters followed by disc_step^s and disc_step^-s, one upload
    const size_t blob_bytes = sizeof(mp_lsm_params) + 2 * (size_t)(m + 1) * 8;
    std::vector<unsigned char> blob(blob_bytes);
    std::memcpy(blob.data(), &p, sizeof(p));
    double* disc_pow = reinterpret_cast<double*>(blob.data() + sizeof(mp_lsm_params));
    double* growth = disc_pow + (m + 1);
    for (int s = 0; s <= m; ++s) {
        disc_pow[s] = powi_ref(p.disc_step, s);
        growth[s] = powi_ref(p.disc_step, -s);
    }
    const size_t rb = precision == 64 ? 8 : 4;
    const int block = 256;
    // regression dates: schedule without its last element, latest first
    const int ns = (int)P.schedule.size();
    std::vector<int> reg_dates;
    for (int i = ns - 2; i >= 0; --i) reg_dates.push_back((int)P.schedule[i]);
    const int n_reg = (int)reg_dates.size();
    // The one-launch decomposition: blocks of fused_bs threads with J paths
    // per thread.  Fewer blocks make each date's exchange cheaper, so the
    // first fit with one block per CU wins, in the order of kLsmShapes; if
    // none fits, the fewest paths per thread that keep every 256-thread block
    // resident.  It is priced in one launch, or (contexts opened with
    // per-date launches, e.g. the server's) as matched per-date launches on
    // the same grid and block size that reproduce the one-launch pricing bit
    // for bit.
    int fused_j = 0, fused_grid = 0, fused_bs = block;
    auto try_shape = [&](int bs, int lj, uint64_t per_cu) {
        const int J = 1 << lj;
        const uint64_t g = (n + (uint64_t)bs * J - 1) / ((uint64_t)bs * J);
        int& bpc = c->lsm_blocks_per_cu[precision == 64][normals_host != nullptr][lj][lsm_shape_index(bs)];
        if (bpc == 0) bpc = std::max(-1, mp::lsm_fused_blocks_per_cu(precision, normals_host != nullptr, J, bs));
        if (bpc <= 0) return false;
        const uint64_t resident = std::min<uint64_t>(per_cu, (uint64_t)bpc) * (uint64_t)c->n_cu;
        // write-once slots of every date: at most 64 MiB
        const uint64_t slot_bytes = (uint64_t)n_reg * (16 * g + 16 * 32) * 8;
        if (g > (uint64_t)MP_LSM_MAX_BLOCKS || g > resident || slot_bytes > (64ull << 20)) return false;
        fused_j = J;
        fused_bs = bs;
        fused_grid = (int)std::max<uint64_t>(1, g);
        return true;
    };
    for (const lsm_shape& sh : kLsmShapes) {
        if (fused_j) break;
        for (int lj = 0; lj <= sh.max_lj && !fused_j; ++lj) try_shape(sh.bs, lj, 1);
    }
    for (int lj = 0; lj < 5 && !fused_j; ++lj) try_shape(256, lj, ~0ull);
    const bool matched = fused_j && ((!c->lsm_fused && c->lsm_matched) || per_date);
    if (!c->lsm_fused && !c->lsm_matched) fused_j = 0;
    if (matched) fused_j = 0;  // per-date launches on the one-launch grid
    const int grid = (fused_j || matched)
                         ? fused_grid
                         : (int)std::max<uint64_t>(1, std::min<uint64_t>({(n + block - 1) / block,
                                                                          (uint64_t)c->n_cu * 2,
                                                                          (uint64_t)MP_LSM_MAX_BLOCKS}));
    ENSURE(c->lsm_paths, (size_t)(m - 1) * n * rb);
    ENSURE(c->lsm_cash, n * 8);
    ENSURE(c->lsm_ex, n * 2);
    ENSURE(c->lsm_coeff, (size_t)(m + 1) * 4 * 8);
    ENSURE(c->lsm_valid, (size_t)(m + 1) * 4);
    // per-date launches: block partials [grid][16]; one launch: write-once
    // slots [n_reg][16][grid] and group totals [n_reg][16][32]
    const size_t n_slots = fused_j ? (size_t)n_reg * (16 * grid + 16 * 32) : 0;
    // matched per-date launches: block sums [16][grid], then group totals [16][32]
    ENSURE(c->lsm_gram, std::max((size_t)grid * 16 + 16 * 32, n_slots) * 8);
    ENSURE(c->lsm_par, blob_bytes);
    // counters: per-date launches use tickets [m + 2] and (matched) the error
    // flag [m + 2]; one launch uses the final ticket [0] and the error flag [1]
    ENSURE(c->lsm_cnt, (size_t)(m + 3) * 4);
    ENSURE(c->lsm_part, (size_t)grid * 4 * 8);
    ENSURE(c->lsm_dates, (size_t)std::max(n_reg, 1) * 4);
    if (normals_host) ENSURE(c->normals, (size_t)m * n * 8);
    RETCHK(upload(c, c->lsm_par.p, blob.data(), blob_bytes));
    if (n_reg) RETCHK(upload(c, c->lsm_dates.p, reg_dates.data(), (size_t)n_reg * 4));
    if (normals_host)
        HIPCHK(hipMemcpyAsync(c->normals.p, normals_host, (size_t)m * n * 8, hipMemcpyHostToDevice, c->stream));
    const mp_lsm_params* dp = c->lsm_par.as<mp_lsm_params>();
    unsigned int* cnt = c->lsm_cnt.as<unsigned int>();
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    HIPCHK(mp::launch_lsm_fill(cnt, m + 3, c->lsm_valid.as<int>(), c->lsm_coeff.as<double>(), m + 1,
                               c->stream));
    int launches = 2;
    // One-launch pricings need every block of their grid resident at once:
    // two of them overlapping on one device could each hold part of it and
    // wait (boundedly) on the rest, so they are serialized per device
    // within the process, from launch to completion.
    static std::mutex fused_mu[64];
    std::unique_lock<std::mutex> fused_lock;
    if (fused_j) fused_lock = std::unique_lock<std::mutex>(fused_mu[c->device & 63]);
    if (fused_j) {
        double* slots = c->lsm_gram.as<double>();
        if (n_slots) HIPCHK(mp::launch_slot_fill(slots, n_slots, c->stream));
        // 2^20 polls of at least 64 cycles and one memory round trip each:
        // far beyond any legitimate wait; cnt[0] is the final ticket, cnt[1] the error flag
        HIPCHK(mp::launch_lsm_fused(dp, normals_host ? c->normals.as<double>() : nullptr, c->lsm_paths.p,
                                    c->lsm_dates.as<int>(), n_reg, c->lsm_coeff.as<double>(),
                                    c->lsm_valid.as<int>(), slots, slots + (size_t)n_reg * 16 * grid,
                                    c->lsm_part.as<double>(), cnt, c->lsm_ex.as<uint16_t>(),
                                    c->h_res_dev, cnt + 1, c->lsm_spin_limit, fused_j, fused_bs, grid,
                                    precision, c->stream));
        ++launches;
    } else {
    // matched launches use the one-launch block size (same per-thread paths
    // and block sums); plain ones 256-thread blocks
    const int sblock = matched ? fused_bs : block;
    // forward paths with a grid that fills the device (grid-stride, so any
    // grid gives the same paths)
    const int fwd_grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((n + block - 1) / block,
                                                                      (uint64_t)c->n_cu * 16));
    HIPCHK(mp::launch_lsm_paths(dp, normals_host ? c->normals.as<double>() : nullptr, c->lsm_paths.p,
                                c->lsm_cash.as<double>(), c->lsm_ex.as<uint16_t>(), precision, fwd_grid,
                                block, c->stream));
    // each pass also applies the exercise rule of the date regressed before it
    int prev = 0;
    launches = 3;
    for (int i = 0; i < n_reg; ++i) {
        const int date = reg_dates[i];
        HIPCHK(mp::launch_lsm_step(dp, c->lsm_paths.p, c->lsm_cash.as<double>(), c->lsm_ex.as<uint16_t>(),
                                   c->lsm_coeff.as<double>(), c->lsm_valid.as<int>(), prev, date,
                                   c->lsm_gram.as<double>(), cnt, precision, grid, sblock, matched, c->stream));
        ++launches;
        if (matched) {
            HIPCHK(mp::launch_lsm_group(dp, c->lsm_gram.as<double>(), grid, c->lsm_gram.as<double>() + 16 * grid,
                                        cnt + date, c->lsm_coeff.as<double>(), c->lsm_valid.as<int>(), date,
                                        cnt + (m + 2), c->stream));
            ++launches;
        }
        prev = date;
    }
    if (prev >= 1) {
        HIPCHK(mp::launch_lsm_step(dp, c->lsm_paths.p, c->lsm_cash.as<double>(), c->lsm_ex.as<uint16_t>(),
                                   c->lsm_coeff.as<double>(), c->lsm_valid.as<int>(), prev, 0,
                                   c->lsm_gram.as<double>(), cnt, precision, grid, sblock, matched, c->stream));
        ++launches;
    }
    HIPCHK(mp::launch_lsm_final(dp, c->lsm_cash.as<double>(), c->lsm_ex.as<uint16_t>(), c->lsm_valid.as<int>(),
                                c->lsm_part.as<double>(), cnt + (m + 1), c->h_res_dev, grid, sblock,
                                c->stream));
    }
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    // the result row is written to pinned host memory by the last block; the
    // wait-error word is copied next to it
    unsigned int* h_err = reinterpret_cast<unsigned int*>(c->h_res + 16);
    *h_err = 0u;
    if (fused_j) HIPCHK(hipMemcpyAsync(h_err, cnt + 1, 4, hipMemcpyDeviceToHost, c->stream));
    if (matched) HIPCHK(hipMemcpyAsync(h_err, cnt + (m + 2), 4, hipMemcpyDeviceToHost, c->stream));
    std::vector<double> coeff;
    if (coeff_host) {
        coeff.resize((size_t)(m + 1) * 4);
        HIPCHK(hipMemcpyAsync(coeff.data(), c->lsm_coeff.p, coeff.size() * 8, hipMemcpyDeviceToHost, c->stream));
    }
    std::vector<uint16_t> ex;
    if (ex_host) {
        ex.resize(n);
        HIPCHK(hipMemcpyAsync(ex.data(), c->lsm_ex.p, n * 2, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (*h_err && fused_j) {
        *gave_up = true;
        return MP_OK;
    }
    if (*h_err)
        return fail(MP_E_HIP, "LSM: a per-date reduction found a block's sums missing");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    c->last.main_kernel_ms = ms;
    c->last.finalize_ms = 0.0;
    c->last.main_launches = (uint64_t)launches;
    if (coeff_host)
        for (int s = 0; s < m; ++s)
            for (int i = 0; i < 4; ++i) coeff_host[s * 4 + i] = coeff[(size_t)s * 4 + i];
    if (ex_host)
        for (uint64_t i = 0; i < n; ++i) ex_host[i] = ex[i];
    std::memset(out, 0, sizeof(*out));
    out->status = MP_OK;
    out->n_strikes = 1;
    double res[8];
    std::memcpy(res, c->h_res, 5 * 8);
    out->price = res[0];
    out->std_error = res[1];
    out->strike_price[0] = res[0];
    out->strike_std_error[0] = res[1];
    out->early_exercise_count = (uint64_t)res[2];
    out->maturity_exercise_count = (uint64_t)res[3];
    out->regression_steps = (uint64_t)res[4];
    out->n_paths = n;
    out->n_steps = (uint64_t)m;
    out->runtime_ms = ms;
    out->fp64_ref_error = NAN;
    return MP_OK;
}

// LSM pricing.  A one-launch pricing needs every block of its grid resident
// at once; when other work on the device (another process) keeps some of
// them out, the others' bounded waits run out and the pricing is repeated
// with the matched per-date launches, which need no co-residency and give
// the same bits.
int run_lsm(mp_ctx* c, const mp_request& r, prepared& P, int precision, const double* normals_host,
            mp_result* out, double* coeff_host, uint32_t* ex_host) {
    bool gave_up = false;
    int rc = run_lsm_once(c, r, P, precision, normals_host, out, coeff_host, ex_host, false, &gave_up);
    if (rc == MP_OK && gave_up) {
        c->lsm_fallbacks += 1;
        rc = run_lsm_once(c, r, P, precision, normals_host, out, coeff_host, ex_host, true, &gave_up);
        if (rc == MP_OK && gave_up) rc = fail(MP_E_HIP, "LSM: per-date launches reported a timed-out wait");
    }
    return rc;
}

// Device buffers of a one-request launch given by the caller instead of the
// context's arenas (the allocation-free pricing's scratch).
struct tile_bufs {
    mp_dreq* reqs;
    unsigned int* counter;
    double* partial;
};

int run_tiles_single(mp_ctx* c, prepared& P, int precision, mp_result* out, const double* normals,
                     const double* uniforms, const double* jumps, double* term_host,
                     double* pay_host, const tile_bufs* fixed = nullptr) {
    c->stage_off = 0;  // the previous call has synchronised
    mp_dreq d = P.d;
    d.tile0 = 0;
    const bool buffer = normals != nullptr;
    const uint64_t n = d.n_paths;
    const int F = d.model == MP_MODEL_HESTON ? 2 : 1;
    const size_t nn = (size_t)n * d.n_steps * F;
    if (!fixed) {
        ENSURE(c->reqs, sizeof(mp_dreq));
        ENSURE(c->partial, (size_t)d.n_tiles * MP_PARTIAL_STRIDE * 8);
        ENSURE(c->counter, 64);
    }
    mp_dreq* d_reqs = fixed ? fixed->reqs : c->reqs.as<mp_dreq>();
    unsigned int* d_counter = fixed ? fixed->counter : c->counter.as<unsigned int>();
    double* d_partial = fixed ? fixed->partial : c->partial.as<double>();
    const bool want_paths = term_host != nullptr || pay_host != nullptr;
    if (want_paths) {
        ENSURE(c->term, n * 8);
        ENSURE(c->pay, n * 8);
    }
    if (buffer) {
        ENSURE(c->normals, nn * 8);
        if (d.model == MP_MODEL_MERTON) {
            ENSURE(c->uniforms, (size_t)n * d.n_steps * 8);
            ENSURE(c->jumps, (size_t)n * d.n_steps * 8);
        }
    }
    if (!buffer && !P.sampler.empty()) {
        d.sampler_off = 0;
        ENSURE(c->sampler, P.sampler.size() * sizeof(mp_sampler_dim));
        RETCHK(upload(c, c->sampler.p, P.sampler.data(), P.sampler.size() * sizeof(mp_sampler_dim)));
    }
    if (buffer) d.sampling = MP_SAMPLER_NONE;  // caller buffers replace every draw
    if (buffer) {
        HIPCHK(hipMemcpyAsync(c->normals.p, normals, nn * 8, hipMemcpyHostToDevice, c->stream));
        if (d.model == MP_MODEL_MERTON) {
            HIPCHK(hipMemcpyAsync(c->uniforms.p, uniforms, (size_t)n * d.n_steps * 8, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(c->jumps.p, jumps, (size_t)n * d.n_steps * 8, hipMemcpyHostToDevice, c->stream));
        }
    }
    const mp_tuning tu = tuning_for(c, d, precision);
    mp::tile_launch L{};
    L.reqs = d_reqs;
    L.tile_req = nullptr;  // one request: every tile is request 0's
    L.order = nullptr;
    L.n_tiles = d.n_tiles;
    L.counter = d_counter;
    L.partial = d_partial;
    L.normals = buffer ? c->normals.as<double>() : nullptr;
    L.uniforms = buffer && d.model == MP_MODEL_MERTON ? c->uniforms.as<double>() : nullptr;
    L.jump_normals = buffer && d.model == MP_MODEL_MERTON ? c->jumps.as<double>() : nullptr;
    L.sampler = c->sampler.as<mp_sampler_dim>();
    L.nib = c->d_nib;
    L.mlmc = d.payoff == MP_PAY_ASIAN_MLMC ? 1 : 0;
    L.gbm_f32 = precision == 32 && !buffer && gbm_philox_tile(d) ? 1 : 0;
    L.terminal_out = want_paths ? c->term.as<double>() : nullptr;
    L.payoff_out = want_paths ? c->pay.as<double>() : nullptr;
    L.block = tu.block_size > 0 ? tu.block_size : 256;
    L.ppl = tu.paths_per_lane > 0 && d.payoff != MP_PAY_ASIAN_MLMC ? tu.paths_per_lane : 1;
    L.unroll = tu.step_unroll > 0 ? tu.step_unroll : 1;
    L.grid = grid_for_tiles(c, d.n_tiles, L.block);
    HIPCHK(mp::launch_setup(d, d_reqs, L.counter, 1, c->stream));
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    HIPCHK(mp::launch_tiles(L, precision, c->stream));
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(mp::launch_finalize(L.reqs, 1, L.partial, c->h_res_dev, c->stream));
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    if (term_host) HIPCHK(hipMemcpyAsync(term_host, c->term.p, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (pay_host) HIPCHK(hipMemcpyAsync(pay_host, c->pay.p, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double fin[MP_PARTIAL_STRIDE];
    std::memcpy(fin, c->h_res, sizeof(fin));
    float main_ms = 0.f, fin_ms = 0.f;
    (void)hipEventElapsedTime(&main_ms, c->ev[0], c->ev[1]);
    (void)hipEventElapsedTime(&fin_ms, c->ev[1], c->ev[2]);
    c->last.main_kernel_ms = main_ms;
    c->last.finalize_ms = fin_ms;
    c->last.main_launches = 1;
    fill_result(out, fin, P);
    out->runtime_ms = main_ms + fin_ms;
    return MP_OK;
}

int price_one(mp_ctx* c, const mp_request& r, int precision, mp_result* out) {
    prepared P;
    int rc = prepare_request(r, P);
    if (rc != MP_OK) return rc;
    switch (P.kind) {
        case KIND_QMC:
            return run_qmc(c, r, P, precision, out, nullptr, nullptr);
        case KIND_BASKET:
            return run_basket(c, r, P, precision, nullptr, out, nullptr, nullptr);
        case KIND_LSM:
            return run_lsm(c, r, P, precision, nullptr, out, nullptr, nullptr);
        default:
            return run_tiles_single(c, P, precision, out, nullptr, nullptr, nullptr, nullptr, nullptr);
    }
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
MP_EXPORT const char* mp_hip_version(void) { return "montepath-hip 0.1.0 (gfx950)"; }

MP_EXPORT const char* mp_hip_last_error(void) { return g_err.c_str(); }

MP_EXPORT int mp_hip_device_count(int* count) {
    if (!count) return fail(MP_E_INVALID_CONFIG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        return fail(MP_E_UNKNOWN_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
    }
    *count = n;
    return MP_OK;
}

MP_EXPORT int mp_hip_open(int device_ordinal, mp_ctx** out) {
    if (!out) return fail(MP_E_INVALID_CONFIG, "null context pointer");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device_ordinal < 0 || device_ordinal >= n) {
        (void)hipGetLastError();
        return fail(MP_E_UNKNOWN_DEVICE, "rocm:" + std::to_string(device_ordinal) + " is not available");
    }
    // a context that fails part-way is closed, which frees what it holds
    struct open_guard {
        mp_ctx* c;
        ~open_guard() {
            if (c) mp_hip_close(c);
        }
    } guard{new mp_ctx()};
    mp_ctx* c = guard.c;
    c->device = device_ordinal;
    HIPCHK(hipSetDevice(device_ordinal));
    RETCHK(scratch_setup(device_ordinal));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device_ordinal));
    c->n_cu = prop.multiProcessorCount;
    c->max_lds = prop.sharedMemPerBlock;
    c->name = prop.name;
    c->arch = prop.gcnArchName;
    if (const char* v = std::getenv("MONTEPATH_LSM_MULTI_LAUNCH")) {
        c->lsm_fused = !(v[0] == '1' || v[0] == '2');
        c->lsm_matched = v[0] != '2';
    }
    if (const char* v = std::getenv("MONTEPATH_QMC_STACK_BRIDGE")) c->qmc_tree = !(v[0] == '1');
    if (const char* v = std::getenv("MONTEPATH_LSM_SPIN_LIMIT")) {
        const long lim = std::strtol(v, nullptr, 10);
        if (lim >= 1 && lim <= (1l << 24)) c->lsm_spin_limit = (unsigned int)lim;
    }
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
    sobol_directions(c->dirs);
    std::vector<uint32_t> nib;
    if (!sobol_index_tables(c->dirs, nib))
        return fail(MP_E_UNSUPPORTED, "Sobol direction numbers 0..15 have low bits set");
    HIPCHK(hipMalloc(&c->d_nib, nib.size() * 4));
    HIPCHK(hipMemcpy(c->d_nib, nib.data(), nib.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_res), 32 * 8,
                         hipHostMallocMapped | hipHostMallocCoherent));
    ENSURE(c->h_stage, 1u << 20);
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->h_res_dev), c->h_res, 0));
    guard.c = nullptr;
    *out = c;
    return MP_OK;
}

namespace {

// Device bytes a context's arenas hold (its reusable batches included; the
// Sobol index tables and pinned host buffers are not counted).
uint64_t ctx_device_bytes(const mp_ctx* c) {
    const dbuf* bufs[] = {&c->reqs,      &c->partial,   &c->counter,  &c->normals,   &c->uniforms,
                          &c->jumps,     &c->term,      &c->pay,      &c->plan,      &c->basket_par,
                          &c->lsm_paths, &c->lsm_cash,  &c->lsm_ex,   &c->lsm_coeff, &c->lsm_valid,
                          &c->lsm_gram,  &c->lsm_par,   &c->lsm_cnt,  &c->lsm_part,  &c->lsm_dates,
                          &c->plain_pay, &c->draws,     &c->sampler};
    uint64_t b = 0;
    for (const dbuf* d : bufs) b += d->cap;
    for (const mp_batch* bt : c->bat)
        if (bt)
            b += bt->d_reqs.cap + bt->d_tile_req.cap + bt->d_order.cap + bt->d_partial.cap +
                 bt->d_counter.cap + bt->d_sampler.cap;
    return b;
}

// Frees a context's arenas and reusable batches (after its stream has
// drained); the next pricing allocates what it needs again.
void ctx_trim(mp_ctx* c) {
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dbuf* bufs[] = {&c->reqs,      &c->partial,  &c->counter,  &c->normals,   &c->uniforms,
                    &c->jumps,     &c->term,     &c->pay,      &c->plan,      &c->basket_par,
                    &c->lsm_paths, &c->lsm_cash, &c->lsm_ex,   &c->lsm_coeff, &c->lsm_valid,
                    &c->lsm_gram,  &c->lsm_par,  &c->lsm_cnt,  &c->lsm_part,  &c->lsm_dates,
                    &c->plain_pay, &c->draws,    &c->sampler};
    for (dbuf* d : bufs) d->release();
    for (auto& b : c->bat) {
        delete b;
        b = nullptr;
    }
}

}  // namespace

MP_EXPORT void mp_hip_close(mp_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& b : c->bat) delete b;
    if (c->d_nib) (void)hipFree(c->d_nib);
    if (c->h_res) (void)hipHostFree(c->h_res);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

MP_EXPORT int mp_hip_device_info(mp_ctx* c, mp_device_info* out) {
    if (!c || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->device_id, sizeof(out->device_id), "rocm:%d", c->device);
    std::snprintf(out->name, sizeof(out->name), "%s", c->name.c_str());
    std::snprintf(out->arch, sizeof(out->arch), "%s", c->arch.c_str());
    size_t fr = 0, tot = 0;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemGetInfo(&fr, &tot));
    out->memory_total_mb = tot >> 20;
    out->memory_free_mb = fr >> 20;
    out->compute_units = c->n_cu;
    out->supports_float64 = 1;
    out->max_threads_hint = 1024;
    out->wavefront_size = 64;
    return MP_OK;
}

MP_EXPORT int mp_hip_price(mp_ctx* c, const mp_request* req, mp_result* out) {
    if (!c || !req || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    const int prec = req->precision == 64 ? 64 : 32;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = price_one(c, *req, prec, out);
    if (rc != MP_OK) {
        out->status = rc;
        return rc;
    }
    if (req->fp64_check && prec == 32) {
        mp_result ref;
        const mp_timing keep = c->last;
        rc = price_one(c, *req, 64, &ref);
        c->last = keep;
        if (rc != MP_OK) return rc;
        out->fp64_ref_error = std::fabs(out->price - ref.price);
    }
    const auto t1 = std::chrono::steady_clock::now();
    c->last.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return MP_OK;
}

// mp_hip_price for one work-queue request with Philox draws, priced into the
// device's static scratch: the request record at offset 0, the queue counter
// and then the tile partials after it (256-byte aligned).  The set-up kernel writes the record
// and clears the counter, and every tile writes its partial slot before the
// finalize kernel reads it, so nothing a previous call left in the scratch is
// read.  Calls on one device are serialized by g_scratch_call_mu.
MP_EXPORT int mp_hip_price_noalloc(mp_ctx* c, const mp_request* req, mp_result* out) {
    if (!c || !req || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    if (c->device < 0 || c->device >= 64 || g_scratch_ptr[c->device] == nullptr)
        return fail(MP_E_HIP, "the device scratch of this context's device is not set up");
    HIPCHK(hipSetDevice(c->device));
    prepared P;
    int rc = prepare_request(*req, P);
    if (rc != MP_OK) return rc;
    if (P.kind != KIND_TILE || P.d.payoff == MP_PAY_ASIAN_MLMC)
        return fail(MP_E_UNSUPPORTED,
                    "mp_hip_price_noalloc prices GBM European/Asian/down-and-out/lookback, Heston and "
                    "Merton requests; price Sobol, basket and American/Bermudan requests with mp_hip_price");
    if (!P.sampler.empty() || P.d.sampling != MP_SAMPLER_NONE)
        return fail(MP_E_UNSUPPORTED, "mp_hip_price_noalloc takes pseudorandom (Philox) draws only");
    // request record, then the counter, then the partials, each 256-byte aligned
    constexpr uint64_t rec_bytes = (sizeof(mp_dreq) + 255) / 256 * 256;
    const uint64_t need = rec_bytes + 256 + (uint64_t)P.d.n_tiles * MP_PARTIAL_STRIDE * 8;
    if (need > MP_NOALLOC_SCRATCH_BYTES)
        return fail(MP_E_OOM, "request needs " + std::to_string(need) + " bytes of device scratch (" +
                                  std::to_string(P.d.n_tiles) + " tiles); mp_hip_price_noalloc has " +
                                  std::to_string(MP_NOALLOC_SCRATCH_BYTES) + " per device");
    unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[c->device]);
    const tile_bufs fixed{reinterpret_cast<mp_dreq*>(base), reinterpret_cast<unsigned int*>(base + rec_bytes),
                          reinterpret_cast<double*>(base + rec_bytes + 256)};
    std::lock_guard<std::mutex> lk(g_scratch_call_mu[c->device]);
    const int prec = req->precision == 64 ? 64 : 32;
    const auto t0 = std::chrono::steady_clock::now();
    rc = run_tiles_single(c, P, prec, out, nullptr, nullptr, nullptr, nullptr, nullptr, &fixed);
    if (rc != MP_OK) {
        out->status = rc;
        return rc;
    }
    if (req->fp64_check && prec == 32) {
        mp_result ref;
        const mp_timing keep = c->last;
        rc = run_tiles_single(c, P, 64, &ref, nullptr, nullptr, nullptr, nullptr, nullptr, &fixed);
        c->last = keep;
        if (rc != MP_OK) return rc;
        out->fp64_ref_error = std::fabs(out->price - ref.price);
    }
    const auto t1 = std::chrono::steady_clock::now();
    c->last.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return MP_OK;
}

MP_EXPORT int mp_hip_price_batch(mp_ctx* c, const mp_request* reqs, size_t n, mp_result* out) {
    if (!c || (!reqs && n) || (!out && n)) return fail(MP_E_INVALID_CONFIG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    const auto t0 = std::chrono::steady_clock::now();
    int rc = MP_OK;
    for (int i = 0; i < 2; ++i) {
        if (!c->bat[i]) {
            c->bat[i] = new mp_batch();
            rc = batch_init(c, c->bat[i], c->stream, i ? 64 : 32);
            if (rc != MP_OK) return rc;
        }
        c->bat[i]->reqs.clear();
        c->bat[i]->req_index.clear();
    }
    mp_batch& b32 = *c->bat[0];
    mp_batch& b64 = *c->bat[1];
    std::vector<size_t> others;
    for (size_t i = 0; i < n; ++i) {
        prepared P;
        rc = prepare_request(reqs[i], P);
        if (rc != MP_OK) {
            out[i].status = rc;
            return fail(rc, "request " + std::to_string(i) + ": " + g_err);
        }
        if (P.kind == KIND_TILE && !reqs[i].fp64_check) {
            mp_batch& b = P.precision == 64 ? b64 : b32;
            b.reqs.push_back(P);
            b.req_index.push_back(i);
        } else {
            others.push_back(i);
        }
    }
    for (mp_batch* b : {&b32, &b64}) {
        if (b->reqs.empty()) continue;
        if (b->reqs.size() == 1) {
            // one request: the one-request path (the same tiles, hence the
            // same bits, without staging the tile map and queue order)
            rc = run_tiles_single(c, b->reqs[0], b->precision, &out[b->req_index[0]], nullptr, nullptr,
                                  nullptr, nullptr, nullptr);
            if (rc != MP_OK) return rc;
            continue;
        }
        rc = batch_build(b);
        if (rc != MP_OK) return rc;
        rc = batch_enqueue(b);
        if (rc != MP_OK) return rc;
        rc = batch_collect(b, out, true);
        if (rc != MP_OK) return rc;
    }
    for (size_t i : others) {
        rc = mp_hip_price(c, &reqs[i], &out[i]);
        if (rc != MP_OK) return rc;
    }
    const auto t1 = std::chrono::steady_clock::now();
    c->last.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return MP_OK;
}

// ---- prepared batches (resident inputs, graph replay) ----------------------
MP_EXPORT int mp_hip_batch_prepare(mp_ctx* c, const mp_request* reqs, size_t n, int use_graph,
                                   mp_batch** out) {
    if (!c || !reqs || !out || n == 0) return fail(MP_E_INVALID_CONFIG, "null argument or empty batch");
    HIPCHK(hipSetDevice(c->device));
    std::unique_ptr<mp_batch> b(new mp_batch());
    int prec = reqs[0].precision == 64 ? 64 : 32;
    int rc = batch_init(c, b.get(), nullptr, prec);
    if (rc != MP_OK) return rc;
    for (size_t i = 0; i < n; ++i) {
        prepared P;
        rc = prepare_request(reqs[i], P);
        if (rc != MP_OK) return fail(rc, "request " + std::to_string(i) + ": " + g_err);
        if (P.kind != KIND_TILE)
            return fail(MP_E_UNSUPPORTED, "prepared batches carry GBM/Heston/Merton pseudorandom requests only");
        if (P.precision != prec) return fail(MP_E_INVALID_CONFIG, "a prepared batch has one precision");
        b->reqs.push_back(P);
        b->req_index.push_back(i);
    }
    rc = batch_build(b.get());
    if (rc != MP_OK) return rc;
    if (use_graph) {
        rc = batch_capture(b.get());
        if (rc != MP_OK) return rc;
    }
    *out = b.release();
    return MP_OK;
}

MP_EXPORT int mp_hip_batch_launch(mp_batch* b) {
    if (!b) return fail(MP_E_INVALID_CONFIG, "null batch");
    return batch_enqueue(b);
}

MP_EXPORT int mp_hip_batch_sync(mp_batch* b) {
    if (!b) return fail(MP_E_INVALID_CONFIG, "null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    return MP_OK;
}

MP_EXPORT int mp_hip_batch_results(mp_batch* b, mp_result* out) {
    if (!b || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    return batch_collect(b, out, true);
}

MP_EXPORT int mp_hip_batch_info(mp_batch* b, uint64_t* n_tiles, uint64_t* path_steps, int* grid,
                                int* block) {
    if (!b) return fail(MP_E_INVALID_CONFIG, "null batch");
    uint64_t ps = 0;
    for (auto& p : b->reqs) ps += p.d.n_paths * (uint64_t)p.d.n_steps;
    if (n_tiles) *n_tiles = b->n_tiles;
    if (path_steps) *path_steps = ps;
    if (grid) *grid = b->grid;
    if (block) *block = b->block;
    return MP_OK;
}

MP_EXPORT void mp_hip_batch_free(mp_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    delete b;
}

MP_EXPORT int mp_hip_price_plain_fp64(mp_ctx* c, const mp_request* req, mp_result* out) {
    if (!c || !req || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    prepared P;
    int rc = prepare_request(*req, P);
    if (rc != MP_OK) return rc;
    if (P.kind != KIND_TILE || P.d.model != MP_MODEL_GBM ||
        (P.d.payoff != MP_PAY_EUROPEAN && P.d.payoff != MP_PAY_ASIAN) || P.d.n_strikes != 1 ||
        P.d.technique != MP_TECH_STANDARD || P.d.sampling != MP_SAMPLER_NONE)
        return fail(MP_E_UNSUPPORTED, "the plain FP64 path prices single-strike GBM European/Asian calls (standard technique, pseudorandom draws)");
    mp_dreq d = P.d;
    d.discount = std::exp(-req->r * req->t);
    const uint64_t n = d.n_paths;
    const int block = 256;
    const int grid = (int)((n + block - 1) / block);
    const int rblocks = std::max(1, std::min(grid, c->n_cu * 4));
    ENSURE(c->reqs, sizeof(mp_dreq));
    ENSURE(c->plain_pay, n * 8);
    ENSURE(c->partial, (size_t)rblocks * MP_PARTIAL_STRIDE * 8);
    // finalize sums payoffs that are already discounted
    mp_dreq fin_d = d;
    fin_d.tile0 = 0;
    fin_d.n_tiles = (uint32_t)rblocks;
    fin_d.n_strikes = 1;
    fin_d.discount = 1.0;
    HIPCHK(mp::launch_setup(d, c->reqs.as<mp_dreq>(), nullptr, 0, c->stream));
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    HIPCHK(mp::launch_plain_fp64(c->reqs.as<mp_dreq>(), c->plain_pay.as<double>(), grid, block, c->stream));
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(mp::launch_plain_reduce(c->plain_pay.as<double>(), n, c->partial.as<double>(), rblocks, c->stream));
    HIPCHK(mp::launch_setup(fin_d, c->reqs.as<mp_dreq>(), nullptr, 0, c->stream));
    HIPCHK(mp::launch_finalize(c->reqs.as<mp_dreq>(), 1, c->partial.as<double>(), c->h_res_dev, c->stream));
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double fin[MP_PARTIAL_STRIDE];
    std::memcpy(fin, c->h_res, sizeof(fin));
    float main_ms = 0.f, fin_ms = 0.f;
    (void)hipEventElapsedTime(&main_ms, c->ev[0], c->ev[1]);
    (void)hipEventElapsedTime(&fin_ms, c->ev[1], c->ev[2]);
    c->last.main_kernel_ms = main_ms;
    c->last.finalize_ms = fin_ms;
    c->last.main_launches = 1;
    fill_result(out, fin, P);
    out->runtime_ms = main_ms + fin_ms;
    return MP_OK;
}

MP_EXPORT int mp_hip_european_call_stepwise_v1(mp_ctx* c, const double* normals, double* payoffs,
                                               int32_t n_paths, int32_t n_steps, double log_s0,
                                               double strike, double drift_dt, double vol_dt,
                                               double discount) {
    if (!c || !normals || !payoffs) return fail(MP_E_INVALID_CONFIG, "null argument");
    if (n_paths < 1 || n_steps < 1) return fail(MP_E_INVALID_CONFIG, "n_paths and n_steps must be at least 1");
    if ((uint64_t)n_paths * (uint64_t)n_steps > (1ull << 31))
        return fail(MP_E_INVALID_CONFIG, "the contract's normals buffer holds at most 2^31 values");
    HIPCHK(hipSetDevice(c->device));
    const size_t nn = (size_t)n_paths * (size_t)n_steps;
    ENSURE(c->normals, nn * 8);
    ENSURE(c->plain_pay, (size_t)n_paths * 8);
    HIPCHK(hipMemcpyAsync(c->normals.p, normals, nn * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    HIPCHK(mp::launch_european_call_stepwise_v1(c->normals.as<double>(), c->plain_pay.as<double>(), n_paths,
                                                n_steps, log_s0, strike, drift_dt, vol_dt, discount,
                                                c->stream));
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(hipMemcpyAsync(payoffs, c->plain_pay.p, (size_t)n_paths * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    c->last.main_kernel_ms = ms;
    c->last.finalize_ms = 0.0;
    c->last.total_ms = ms;
    c->last.main_launches = 1;
    return MP_OK;
}

MP_EXPORT int mp_hip_paths_from_normals(mp_ctx* c, const mp_request* req, const double* normals,
                                        const double* uniforms, const double* jump_normals,
                                        double* terminal_out, double* payoff_out) {
    if (!c || !req || !normals || !terminal_out || !payoff_out)
        return fail(MP_E_INVALID_CONFIG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    prepared P;
    int rc = prepare_request(*req, P);
    if (rc != MP_OK) return rc;
    if (req->sampling != MP_PSEUDORANDOM)
        return fail(MP_E_INCOMPATIBLE_INPUT, "normal buffers replace pseudorandom draws only");
    const int prec = P.precision;
    mp_result res;
    if (P.kind == KIND_BASKET)
        return run_basket(c, *req, P, prec, normals, &res, terminal_out, payoff_out);
    if (P.kind != KIND_TILE)
        return fail(MP_E_INCOMPATIBLE_INPUT, "use mp_hip_lsm_from_normals for LSM requests");
    if (P.d.model == MP_MODEL_MERTON && (!uniforms || !jump_normals))
        return fail(MP_E_INVALID_CONFIG, "Merton needs uniforms and jump_normals buffers");
    return run_tiles_single(c, P, prec, &res, normals, uniforms, jump_normals, terminal_out, payoff_out);
}

MP_EXPORT int mp_hip_price_from_normals(mp_ctx* c, const mp_request* req, const double* normals,
                                        const double* uniforms, const double* jump_normals,
                                        mp_result* out) {
    if (!c || !req || !normals || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    prepared P;
    int rc = prepare_request(*req, P);
    if (rc != MP_OK) return rc;
    if (req->sampling != MP_PSEUDORANDOM)
        return fail(MP_E_INCOMPATIBLE_INPUT, "normal buffers replace pseudorandom draws only");
    if (P.kind == KIND_BASKET) return run_basket(c, *req, P, P.precision, normals, out, nullptr, nullptr);
    if (P.kind != KIND_TILE)
        return fail(MP_E_INCOMPATIBLE_INPUT, "use mp_hip_lsm_from_normals for LSM requests");
    if (P.d.model == MP_MODEL_MERTON && (!uniforms || !jump_normals))
        return fail(MP_E_INVALID_CONFIG, "Merton needs uniforms and jump_normals buffers");
    return run_tiles_single(c, P, P.precision, out, normals, uniforms, jump_normals, nullptr, nullptr);
}

MP_EXPORT int mp_hip_price_with_paths(mp_ctx* c, const mp_request* req, mp_result* out,
                                      double* terminal_out, double* payoff_out) {
    if (!c || !req || !out || !terminal_out || !payoff_out)
        return fail(MP_E_INVALID_CONFIG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    prepared P;
    int rc = prepare_request(*req, P);
    if (rc != MP_OK) return rc;
    switch (P.kind) {
        case KIND_QMC:
            rc = run_qmc(c, *req, P, P.precision, out, terminal_out, payoff_out);
            break;
        case KIND_BASKET:
            rc = run_basket(c, *req, P, P.precision, nullptr, out, terminal_out, payoff_out);
            break;
        case KIND_LSM:
            return fail(MP_E_INCOMPATIBLE_INPUT, "LSM per-path output: use mp_hip_lsm_from_normals");
        default:
            rc = run_tiles_single(c, P, P.precision, out, nullptr, nullptr, nullptr, terminal_out,
                                  payoff_out);
    }
    // antithetic requests price ceil(n/2) pairs: entries past the pairs are zero
    if (rc == MP_OK && P.d.n_paths < req->n_paths) {
        const size_t tail = (size_t)(req->n_paths - P.d.n_paths);
        std::memset(terminal_out + P.d.n_paths, 0, tail * sizeof(double));
        std::memset(payoff_out + P.d.n_paths, 0, tail * sizeof(double));
    }
    return rc;
}

// ===========================================================================
// Greeks
// ===========================================================================
namespace {

// bump_spot / bump_volatility / bump_rate / bump_maturity, cpu.rs:4994-5008
double bump_spot(double s) { return std::max(std::fabs(s) * 0.01, 0.01); }
double bump_volatility(double v) { return std::max(std::fabs(v) * 0.05, 0.005); }
double bump_maturity(double t) { return std::min(1.0 / 365.0, t * 0.25); }

enum greek_param { GP_S0, GP_SIGMA, GP_RATE, GP_MATURITY, GP_HESTON_VOL, GP_ASSET_S0, GP_ASSET_SIGMA };

struct greek_spec {
    int greek;
    int param;
    int asset;
    double center, bump;
    bool reverse;  // Theta: the calendar convention negates dV/dT
};

void set_param(mp_request& r, const greek_spec& g, double v) {
    switch (g.param) {
        case GP_S0: r.s0 = v; break;
        case GP_SIGMA: r.sigma = v; break;
        case GP_RATE: r.r = v; break;
        case GP_MATURITY: r.t = v; break;
        case GP_HESTON_VOL: {
            const double x = std::max(v, 0.0);  // Vega in volatility units: v0 = vol^2
            r.v0 = x * x;
            break;
        }
        case GP_ASSET_S0: r.asset_s0[g.asset] = v; break;
        default: r.asset_sigma[g.asset] = v; break;
    }
}

// parameter_floor, cpu.rs:4985-4992: Rho may go to -1, the rest stay positive
double param_floor(int greek) { return greek == MP_GREEK_RHO ? -1.0 : DBL_MIN; }

// The bumped parameters of one request, in the reference's order.
int bump_specs(const mp_request& r, std::vector<greek_spec>& out) {
    out.clear();
    auto add = [&](int greek, int param, int asset, double center, double bump, bool rev) {
        out.push_back(greek_spec{greek, param, asset, center, bump, rev});
    };
    switch (r.workload) {
        case MP_EUROPEAN_CALL:
        case MP_ARITHMETIC_ASIAN_CALL:
        case MP_DOWN_AND_OUT_CALL:
        case MP_LOOKBACK_CALL:
            add(MP_GREEK_DELTA, GP_S0, 0, r.s0, bump_spot(r.s0), false);
            add(MP_GREEK_VEGA, GP_SIGMA, 0, r.sigma, bump_volatility(r.sigma), false);
            add(MP_GREEK_RHO, GP_RATE, 0, r.r, 0.001, false);
            add(MP_GREEK_THETA, GP_MATURITY, 0, r.t, bump_maturity(r.t), true);
            return MP_OK;
        case MP_HESTON_EUROPEAN_CALL: {  // heston_call_bump_greeks_cpu cpu.rs:4592-4668
            const double vol = std::sqrt(std::max(r.v0, 0.0));
            add(MP_GREEK_DELTA, GP_S0, 0, r.s0, bump_spot(r.s0), false);
            add(MP_GREEK_VEGA, GP_HESTON_VOL, 0, vol, bump_volatility(vol), false);
            add(MP_GREEK_RHO, GP_RATE, 0, r.r, 0.001, false);
            add(MP_GREEK_THETA, GP_MATURITY, 0, r.t, bump_maturity(r.t), true);
            return MP_OK;
        }
        case MP_BASKET_CALL:  // basket_call_bump_greeks_cpu cpu.rs:4484-4590
            if (r.n_assets < 1 || r.n_assets > MP_MAX_ASSETS)
                return fail(MP_E_INVALID_CONFIG, "n_assets must be in 1..8");
            add(MP_GREEK_DELTA, GP_ASSET_S0, 0, r.asset_s0[0], bump_spot(r.asset_s0[0]), false);
            if (r.n_assets > 1)
                add(MP_GREEK_DELTA2, GP_ASSET_S0, 1, r.asset_s0[1], bump_spot(r.asset_s0[1]), false);
            add(MP_GREEK_VEGA, GP_ASSET_SIGMA, 0, r.asset_sigma[0], bump_volatility(r.asset_sigma[0]), false);
            if (r.n_assets > 1)
                add(MP_GREEK_VEGA2, GP_ASSET_SIGMA, 1, r.asset_sigma[1],
                    bump_volatility(r.asset_sigma[1]), false);
            add(MP_GREEK_RHO, GP_RATE, 0, r.r, 0.001, false);
            add(MP_GREEK_THETA, GP_MATURITY, 0, r.t, bump_maturity(r.t), true);
            return MP_OK;
        default:
            return fail(MP_E_UNSUPPORTED,
                        "bump-and-revalue Greeks cover the GBM family, Heston and the basket");
    }
}

void nan_estimate(mp_greek_estimate& e) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    e.value = e.std_error = e.bump_size = e.base_price = e.bumped_up_price = e.bumped_down_price = nan;
}

// Every request's base and bumped legs in one batch (push_bump_estimate,
// cpu.rs:4938-4983: central difference over the effective half-width, the
// down leg floored; standard error from the two legs as if independent).
int greeks_bump(mp_ctx* c, const mp_request* reqs, size_t n, mp_result* base, mp_greek_estimate* out,
                int32_t* n_est) {
    std::vector<std::vector<greek_spec>> specs(n);
    std::vector<mp_request> legs;
    std::vector<size_t> first(n);
    for (size_t i = 0; i < n; ++i) {
        int rc = bump_specs(reqs[i], specs[i]);
        if (rc != MP_OK) return fail(rc, "request " + std::to_string(i) + ": " + g_err);
        first[i] = legs.size();
        legs.push_back(reqs[i]);
        for (const greek_spec& g : specs[i]) {
            mp_request up = reqs[i], down = reqs[i];
            set_param(up, g, g.center + g.bump);
            set_param(down, g, std::max(g.center - g.bump, param_floor(g.greek)));
            legs.push_back(up);
            legs.push_back(down);
        }
    }
    std::vector<mp_result> res(legs.size());
    int rc = mp_hip_price_batch(c, legs.data(), legs.size(), res.data());
    if (rc != MP_OK) return rc;
    for (size_t i = 0; i < n; ++i) {
        const mp_result& b = res[first[i]];
        base[i] = b;
        n_est[i] = (int32_t)specs[i].size();
        for (size_t j = 0; j < specs[i].size(); ++j) {
            const greek_spec& g = specs[i][j];
            const mp_result& up = res[first[i] + 1 + 2 * j];
            const mp_result& down = res[first[i] + 2 + 2 * j];
            const double lo = std::max(g.center - g.bump, param_floor(g.greek));
            const double eff = ((g.center + g.bump) - lo) * 0.5;
            const double raw = eff > 0.0 ? (up.price - down.price) / (2.0 * eff) : 0.0;
            mp_greek_estimate& e = out[i * MP_MAX_GREEKS + j];
            e.greek = g.greek;
            e.estimator = MP_GREEKS_BUMP_AND_REVALUE;
            e.value = g.reverse ? -raw : raw;
            e.std_error = eff > 0.0 ? std::sqrt(up.std_error * up.std_error + down.std_error * down.std_error) /
                                          (2.0 * eff)
                                    : 0.0;
            e.bump_size = eff;
            e.base_price = b.price;
            e.bumped_up_price = up.price;
            e.bumped_down_price = down.price;
        }
    }
    return MP_OK;
}

// Pathwise / likelihood-ratio Greeks of one European call: the base price as
// requested (or from the caller's normals, terminal method), then one pass
// of greeks_kernel over the terminal draws and MomentAccumulator summaries
// (sample variance, cpu.rs:4859-4876).
int greeks_single(mp_ctx* c, const mp_request& r, int estimator, const double* normals,
                  mp_result* base, mp_greek_estimate* out, int32_t* n_est) {
    if (r.workload != MP_EUROPEAN_CALL)
        return fail(MP_E_UNSUPPORTED,
                    "pathwise and likelihood-ratio Greeks are offered for European calls only; use "
                    "bump-and-revalue");
    if (!(r.sigma > 0.0)) return fail(MP_E_INVALID_CONFIG, "sigma must be > 0 for single-pass Greeks");
    mp_request rt = r;  // terminal draws, plain Philox stream, strike k
    rt.method = MP_METHOD_TERMINAL;
    rt.technique = MP_TECH_STANDARD;
    rt.sampling = MP_PSEUDORANDOM;
    rt.n_strikes = 0;
    rt.fp64_check = 0;
    prepared P;
    int rc = prepare_request(rt, P);
    if (rc != MP_OK) return rc;
    const int precision = P.precision;
    if (normals) {
        rc = run_tiles_single(c, P, precision, base, normals, nullptr, nullptr, nullptr, nullptr);
    } else {
        rc = price_one(c, r, r.precision == 64 ? 64 : 32, base);
    }
    if (rc != MP_OK) return rc;
    mp_dreq d = P.d;
    const int block = 256;
    const uint64_t want = (d.n_paths + block - 1) / block;
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)c->n_cu * 8));
    d.tile0 = 0;
    d.n_tiles = (uint32_t)grid;
    d.technique = MP_TECH_RAW_MOMENTS;
    ENSURE(c->reqs, sizeof(mp_dreq));
    ENSURE(c->partial, (size_t)grid * MP_PARTIAL_STRIDE * 8);
    if (normals) ENSURE(c->normals, d.n_paths * 8);
    HIPCHK(mp::launch_setup(d, c->reqs.as<mp_dreq>(), nullptr, 0, c->stream));
    if (normals) HIPCHK(hipMemcpyAsync(c->normals.p, normals, d.n_paths * 8, hipMemcpyHostToDevice, c->stream));
    const int est = estimator == MP_GREEKS_PATHWISE ? MP_GREEK_EST_PATHWISE : MP_GREEK_EST_LIKELIHOOD_RATIO;
    HIPCHK(mp::launch_greeks(c->reqs.as<mp_dreq>(), est, normals ? c->normals.as<double>() : nullptr,
                             c->partial.as<double>(), precision, grid, block, c->stream));
    HIPCHK(mp::launch_finalize(c->reqs.as<mp_dreq>(), 1, c->partial.as<double>(), c->h_res_dev, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double tot[MP_PARTIAL_STRIDE];
    std::memcpy(tot, c->h_res, sizeof(tot));
    static const int pathwise[4] = {MP_GREEK_DELTA, MP_GREEK_VEGA, MP_GREEK_RHO, MP_GREEK_THETA};
    const int count = est == MP_GREEK_EST_PATHWISE ? 4 : 3;
    const double nn = (double)d.n_paths;
    for (int j = 0; j < count; ++j) {
        const double sum = tot[2 * j], sq = tot[2 * j + 1];
        const double mean = sum / nn;
        double var = d.n_paths > 1 ? (sq - sum * sum / nn) / (nn - 1.0) : 0.0;
        if (!(var > 0.0)) var = 0.0;
        mp_greek_estimate& e = out[j];
        nan_estimate(e);
        e.greek = pathwise[j];
        e.estimator = estimator;
        e.value = mean;
        e.std_error = std::sqrt(var / nn);
        e.base_price = base->price;
    }
    *n_est = count;
    return MP_OK;
}

}  // namespace

MP_EXPORT int mp_hip_greeks(mp_ctx* c, const mp_request* reqs, size_t n, int32_t estimator,
                            mp_result* base, mp_greek_estimate* out, int32_t* n_estimates) {
    if (!c || (!reqs && n) || ((!base || !out || !n_estimates) && n))
        return fail(MP_E_INVALID_CONFIG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    for (size_t i = 0; i < n; ++i) {
        n_estimates[i] = 0;
        for (int j = 0; j < MP_MAX_GREEKS; ++j) nan_estimate(out[i * MP_MAX_GREEKS + j]);
    }
    if (estimator == MP_GREEKS_BUMP_AND_REVALUE) return greeks_bump(c, reqs, n, base, out, n_estimates);
    if (estimator != MP_GREEKS_PATHWISE && estimator != MP_GREEKS_LIKELIHOOD_RATIO)
        return fail(MP_E_INVALID_CONFIG, "estimator must be 0 (bump and revalue), 1 (pathwise) or 2 "
                                         "(likelihood ratio)");
    for (size_t i = 0; i < n; ++i) {
        int rc = greeks_single(c, reqs[i], estimator, nullptr, &base[i], out + i * MP_MAX_GREEKS,
                               &n_estimates[i]);
        if (rc != MP_OK) return rc;
    }
    return MP_OK;
}

MP_EXPORT int mp_hip_greeks_from_normals(mp_ctx* c, const mp_request* req, int32_t estimator,
                                         const double* normals, mp_result* base,
                                         mp_greek_estimate* out, int32_t* n_estimates) {
    if (!c || !req || !normals || !base || !out || !n_estimates)
        return fail(MP_E_INVALID_CONFIG, "null argument");
    if (estimator != MP_GREEKS_PATHWISE && estimator != MP_GREEKS_LIKELIHOOD_RATIO)
        return fail(MP_E_INVALID_CONFIG, "caller normals drive the single-pass estimators only");
    HIPCHK(hipSetDevice(c->device));
    *n_estimates = 0;
    for (int j = 0; j < MP_MAX_GREEKS; ++j) nan_estimate(out[j]);
    return greeks_single(c, *req, estimator, normals, base, out, n_estimates);
}

// ===========================================================================
// Multilevel Monte Carlo arithmetic Asian
// ===========================================================================
namespace {

uint64_t derive_chunk_seed(uint64_t base, uint64_t idx) {
    return splitmix64(base + idx * 0x9E3779B97F4A7C15ull);  // cpu.rs:8922-8924
}

uint64_t mlmc_fine(const mp_mlmc_config& c, uint32_t level) {
    uint64_t s = c.base_steps;
    for (uint32_t i = 0; i < level; ++i) s *= c.refinement_factor;
    return s;
}

bool mlmc_structured(const mp_mlmc_config& c) { return c.sampling != MP_PSEUDORANDOM; }

// validate_arithmetic_asian_mlmc_config (cpu.rs:3493-3510) plus this
// engine's limits (levels, replicates, steps per request).
int mlmc_validate(const mp_mlmc_config& c, bool need_paths) {
    if (c.base_steps == 0) return fail(MP_E_INVALID_CONFIG, "base_steps must be > 0");
    if (c.levels == 0) return fail(MP_E_INVALID_CONFIG, "levels must be > 0");
    if (c.levels > MP_MLMC_MAX_LEVELS) return fail(MP_E_UNSUPPORTED, "at most 16 levels");
    if (c.refinement_factor < 2) return fail(MP_E_INVALID_CONFIG, "refinement_factor must be >= 2");
    if (c.scramble_replicates == 0) return fail(MP_E_INVALID_CONFIG, "scramble_replicates must be > 0");
    if (c.scramble_replicates > MP_MLMC_MAX_REPLICATES)
        return fail(MP_E_UNSUPPORTED, "at most 64 scramble replicates");
    if (c.sampling != MP_PSEUDORANDOM && c.sampling != MP_RANDOMIZED_HALTON &&
        c.sampling != MP_LATIN_HYPERCUBE && c.sampling != MP_SCRAMBLED_SOBOL &&
        c.sampling != MP_SCRAMBLED_SOBOL_BRIDGE)
        return fail(MP_E_INVALID_CONFIG, "unknown sampling method " + std::to_string(c.sampling));
    if (c.precision != 0 && c.precision != 32 && c.precision != 64)
        return fail(MP_E_INVALID_CONFIG, "precision must be 32 or 64");
    if (!(c.s0 > 0.0) || !(c.t > 0.0) || !(c.sigma >= 0.0) || !finite(c.k) || !finite(c.r))
        return fail(MP_E_INVALID_CONFIG, "s0 and t must be > 0, sigma >= 0, k and r finite");
    double steps = (double)c.base_steps;
    for (uint32_t l = 1; l < c.levels; ++l) steps *= (double)c.refinement_factor;
    if (steps > 2147483647.0) return fail(MP_E_UNSUPPORTED, "finest level exceeds 2^31 - 1 steps");
    if (mlmc_structured(c) && c.sampling != MP_SCRAMBLED_SOBOL && c.sampling != MP_SCRAMBLED_SOBOL_BRIDGE &&
        steps > (double)MP_SAMPLER_MAX_DIMS)
        return fail(MP_E_UNSUPPORTED, "Halton and Latin hypercube levels take at most 65536 steps");
    if (need_paths)
        for (uint32_t l = 0; l < c.levels; ++l) {
            if (c.paths_per_level[l] == 0) return fail(MP_E_INVALID_CONFIG, "paths_per_level values must be > 0");
            if (c.paths_per_level[l] > 0xFFFFFFFFull)
                return fail(MP_E_UNSUPPORTED, "n_paths above 2^32 per level is not supported");
        }
    return MP_OK;
}

struct mlmc_unit {
    uint64_t seed;    // the estimator's seed (a replicate's derived seed)
    uint32_t level;
    uint64_t paths;
};

// One level of one estimator as a work-queue request whose tiles sum the
// discounted level terms (finalize writes the raw sum and sum of squares).
int mlmc_prepare(const mp_mlmc_config& c, const mlmc_unit& u, prepared& P) {
    std::memset(&P.d, 0, sizeof(P.d));
    P.schedule.clear();
    P.sampler.clear();
    P.kind = KIND_TILE;
    P.precision = c.precision == 64 ? 64 : 32;
    mp_dreq& d = P.d;
    const uint64_t fine = mlmc_fine(c, u.level);
    const uint64_t lseed = derive_chunk_seed(u.seed, u.level);
    d.model = MP_MODEL_GBM;
    d.payoff = MP_PAY_ASIAN_MLMC;
    d.n_steps = (int32_t)fine;
    d.n_strikes = 1;
    d.strikes[0] = c.k;
    d.key0 = (uint32_t)lseed;
    d.key1 = (uint32_t)c.contract_id;
    d.seed_hi = (uint32_t)(lseed >> 32);
    d.n_paths = u.paths;
    d.s0 = c.s0;
    d.log_s0 = std::log(c.s0);
    d.rate = c.r;
    d.discount = std::exp(-c.r * c.t);
    d.log_barrier_rel = -1e300;
    d.dt = c.t / (double)fine;
    d.sqrt_dt = std::sqrt(d.dt);
    d.drift_dt = (c.r - 0.5 * c.sigma * c.sigma) * d.dt;
    d.vol_dt = c.sigma * std::sqrt(d.dt);
    d.sigma = c.sigma;
    d.technique = MP_TECH_RAW_MOMENTS;
    if (u.level > 0) {
        const double coarse_dt = c.t / (double)(fine / c.refinement_factor);
        d.mlmc_refine = (int32_t)c.refinement_factor;
        d.mlmc_coarse_drift = (c.r - 0.5 * c.sigma * c.sigma) * coarse_dt;
        d.mlmc_coarse_vol = c.sigma * std::sqrt(coarse_dt);
    }
    d.mlmc_sqrt_refine = std::sqrt((double)c.refinement_factor);
    if (c.sampling == MP_RANDOMIZED_HALTON || c.sampling == MP_LATIN_HYPERCUBE) {
        d.sampling = c.sampling == MP_RANDOMIZED_HALTON ? MP_SAMPLER_HALTON : MP_SAMPLER_LHS;
        d.sampler_n = u.paths;
        build_sampler(d.sampling, lseed, u.paths, fine, P.sampler);
    } else if (c.sampling != MP_PSEUDORANDOM) {
        d.sampling = MP_SAMPLER_SOBOL;  // ScrambledSobol::new(level seed): key0 is its low word
    }
    P.n_paths_req = u.paths;
    P.n_steps_cfg = fine;
    d.tile_paths = tile_paths_for(d.n_paths, fine);
    d.n_tiles = (uint32_t)((d.n_paths + d.tile_paths - 1) / d.tile_paths);
    return MP_OK;
}

// Raw (sum, sum of squares) of every unit, all units in one launch.
int mlmc_moments(mp_ctx* c, const mp_mlmc_config& cfg, const std::vector<mlmc_unit>& units,
                 std::vector<double>& sums, std::vector<double>& sqs, double* ms) {
    const int prec = cfg.precision == 64 ? 64 : 32;
    mp_batch*& bp = c->bat[prec == 64 ? 1 : 0];
    if (!bp) {
        bp = new mp_batch();
        int rc = batch_init(c, bp, c->stream, prec);
        if (rc != MP_OK) return rc;
    }
    mp_batch* b = bp;
    b->reqs.clear();
    b->req_index.clear();
    for (size_t i = 0; i < units.size(); ++i) {
        prepared P;
        int rc = mlmc_prepare(cfg, units[i], P);
        if (rc != MP_OK) return rc;
        b->reqs.push_back(std::move(P));
        b->req_index.push_back(i);
    }
    int rc = batch_build(b);
    if (rc != MP_OK) return rc;
    rc = batch_enqueue(b);
    if (rc != MP_OK) return rc;
    std::vector<mp_result> res(units.size());
    rc = batch_collect(b, res.data(), true);
    if (rc != MP_OK) return rc;
    sums.resize(units.size());
    sqs.resize(units.size());
    for (size_t i = 0; i < units.size(); ++i) {
        sums[i] = res[i].price;      // raw-moment finalize: column 0 = sum
        sqs[i] = res[i].std_error;   // column 1 = sum of squares
    }
    if (ms) *ms = units.empty() ? 0.0 : res[0].runtime_ms;
    return MP_OK;
}

std::vector<mlmc_unit> mlmc_units(const mp_mlmc_config& c, uint64_t seed, const uint64_t* paths,
                                  uint64_t pilot) {
    std::vector<mlmc_unit> u;
    for (uint32_t l = 0; l < c.levels; ++l) u.push_back(mlmc_unit{seed, l, paths ? paths[l] : pilot});
    return u;
}

// measure_arithmetic_asian_mlmc_pilot_levels (cpu.rs:3567-3598)
int mlmc_pilot(mp_ctx* c, const mp_mlmc_config& cfg, uint64_t pilot, mp_mlmc_allocation& a) {
    std::vector<double> s, q;
    int rc = mlmc_moments(c, cfg, mlmc_units(cfg, cfg.seed, nullptr, pilot), s, q, nullptr);
    if (rc != MP_OK) return rc;
    std::memset(&a, 0, sizeof(a));
    a.n_levels = cfg.levels;
    const double n = (double)pilot;
    for (uint32_t l = 0; l < cfg.levels; ++l) {
        mp_mlmc_allocation_level& L = a.levels[l];
        L.level = l;
        L.fine_steps = mlmc_fine(cfg, l);
        L.coarse_steps = l == 0 ? 0 : L.fine_steps / cfg.refinement_factor;
        L.pilot_paths = pilot;
        const double mean = s[l] / n;
        L.pilot_variance = std::max((q[l] / n) - mean * mean, 1.0e-18);
        L.cost_per_path = L.fine_steps + L.coarse_steps;
        L.recommended_paths = 1;
    }
    return MP_OK;
}

uint64_t sat_add(uint64_t a, uint64_t b) { return a > UINT64_MAX - b ? UINT64_MAX : a + b; }
uint64_t sat_mul(uint64_t a, uint64_t b) { return (b != 0 && a > UINT64_MAX / b) ? UINT64_MAX : a * b; }
uint64_t f64_to_u64_sat(double x) {  // Rust's `as usize` (saturating, NaN -> 0)
    if (!(x > 0.0)) return 0;
    if (x >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)x;
}

// Recommended paths: by a step budget (tune) or a target variance (solve).
void mlmc_allocate(mp_mlmc_allocation& a, double numer, bool by_variance) {
    double denom = 0.0;
    for (uint32_t l = 0; l < a.n_levels; ++l)
        denom += std::sqrt(a.levels[l].pilot_variance * (double)a.levels[l].cost_per_path);
    uint64_t updates = 0;
    for (uint32_t l = 0; l < a.n_levels; ++l) {
        mp_mlmc_allocation_level& L = a.levels[l];
        const double w = std::sqrt(L.pilot_variance / (double)L.cost_per_path);
        const double raw = by_variance ? std::ceil(denom * w / numer) : std::ceil(numer * w / denom);
        L.recommended_paths = std::max<uint64_t>(f64_to_u64_sat(raw), 2);
        updates = sat_add(updates, sat_mul(L.recommended_paths, L.cost_per_path));
    }
    a.estimated_step_updates = updates;
    a.target_step_updates = by_variance ? updates : (uint64_t)numer;
}

double mlmc_plan_stderr(const mp_mlmc_allocation& a, uint32_t reps) {
    double v = 0.0;
    for (uint32_t l = 0; l < a.n_levels; ++l) v += a.levels[l].pilot_variance / (double)a.levels[l].recommended_paths;
    return std::sqrt(v / (double)reps);
}

void mlmc_fill_level(const mp_mlmc_config& cfg, uint32_t l, uint64_t paths, double sum, double sq,
                     mp_mlmc_level& L) {
    L.level = l;
    L.paths = paths;
    L.fine_steps = mlmc_fine(cfg, l);
    L.coarse_steps = l == 0 ? 0 : L.fine_steps / cfg.refinement_factor;
    const double n = (double)paths;
    L.mean = sum / n;
    L.variance = std::max((sq / n) - L.mean * L.mean, 0.0);
    L.std_error = std::sqrt(L.variance / n);
    L.cost_step_updates = sat_mul(paths, L.fine_steps + L.coarse_steps);
}

}  // namespace

MP_EXPORT int mp_hip_mlmc_price(mp_ctx* c, const mp_mlmc_config* cfgp, mp_mlmc_result* out) {
    if (!c || !cfgp || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    const mp_mlmc_config& cfg = *cfgp;
    int rc = mlmc_validate(cfg, true);
    if (rc != MP_OK) return rc;
    HIPCHK(hipSetDevice(c->device));
    std::memset(out, 0, sizeof(*out));
    const uint32_t R = (mlmc_structured(cfg) && cfg.scramble_replicates > 1) ? cfg.scramble_replicates : 1;
    std::vector<mlmc_unit> units;
    for (uint32_t r = 0; r < R; ++r) {
        const uint64_t seed = R == 1 ? cfg.seed : derive_chunk_seed(cfg.seed, r);
        auto u = mlmc_units(cfg, seed, cfg.paths_per_level, 0);
        units.insert(units.end(), u.begin(), u.end());
    }
    std::vector<double> s, q;
    double ms = 0.0;
    rc = mlmc_moments(c, cfg, units, s, q, &ms);
    if (rc != MP_OK) return rc;
    const uint32_t nl = cfg.levels;
    out->n_levels = nl;
    out->scramble_replicates = R;
    out->runtime_ms = ms;
    // one estimator per replicate (arithmetic_asian_call_price_mlmc_single_cpu)
    std::vector<mp_mlmc_level> lv((size_t)R * nl);
    std::vector<double> est(R), se(R);
    for (uint32_t r = 0; r < R; ++r) {
        double price = 0.0, var_contrib = 0.0;
        for (uint32_t l = 0; l < nl; ++l) {
            mp_mlmc_level& L = lv[(size_t)r * nl + l];
            mlmc_fill_level(cfg, l, cfg.paths_per_level[l], s[(size_t)r * nl + l], q[(size_t)r * nl + l], L);
            price += L.mean;
            var_contrib += L.variance / (double)L.paths;
            out->total_paths = sat_add(out->total_paths, L.paths);
            out->total_step_updates = sat_add(out->total_step_updates, L.cost_step_updates);
        }
        est[r] = price;
        se[r] = std::sqrt(var_contrib);
        out->replicate_estimates[r] = price;
    }
    if (R == 1) {
        out->price = est[0];
        out->std_error = se[0];
        for (uint32_t l = 0; l < nl; ++l) out->levels[l] = lv[l];
        return MP_OK;
    }
    // replicated estimator (cpu.rs:3663-3751): mean and sample spread of the
    // replicate estimates, per level the spread of the level means
    double sum = 0.0;
    for (uint32_t r = 0; r < R; ++r) sum += est[r];
    out->price = sum / (double)R;
    double ss = 0.0;
    for (uint32_t r = 0; r < R; ++r) ss += (est[r] - out->price) * (est[r] - out->price);
    out->std_error = std::sqrt(ss / (double)(R - 1) / (double)R);
    for (uint32_t l = 0; l < nl; ++l) {
        mp_mlmc_level L = lv[l];
        double m = 0.0;
        for (uint32_t r = 0; r < R; ++r) m += lv[(size_t)r * nl + l].mean;
        m /= (double)R;
        double v = 0.0;
        for (uint32_t r = 0; r < R; ++r) {
            const double dd = lv[(size_t)r * nl + l].mean - m;
            v += dd * dd;
        }
        v /= (double)(R - 1);
        L.paths = sat_mul(L.paths, R);
        L.mean = m;
        L.variance = v;
        L.std_error = std::sqrt(v / (double)R);
        L.cost_step_updates = sat_mul(L.cost_step_updates, R);
        out->levels[l] = L;
    }
    return MP_OK;
}

MP_EXPORT int mp_hip_mlmc_tune_allocation(mp_ctx* c, const mp_mlmc_config* cfgp, uint64_t target_step_updates,
                                          uint64_t pilot_paths, mp_mlmc_allocation* out) {
    if (!c || !cfgp || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    int rc = mlmc_validate(*cfgp, false);
    if (rc != MP_OK) return rc;
    if (target_step_updates == 0) return fail(MP_E_INVALID_CONFIG, "target_step_updates must be > 0");
    if (pilot_paths <= 1) return fail(MP_E_INVALID_CONFIG, "pilot_paths_per_level must be > 1");
    HIPCHK(hipSetDevice(c->device));
    rc = mlmc_pilot(c, *cfgp, pilot_paths, *out);
    if (rc != MP_OK) return rc;
    mlmc_allocate(*out, (double)target_step_updates, false);
    out->target_step_updates = target_step_updates;
    return MP_OK;
}

MP_EXPORT int mp_hip_mlmc_solve_tolerance(mp_ctx* c, const mp_mlmc_config* cfgp,
                                          const mp_mlmc_tolerance* tol, mp_mlmc_tolerance_plan* out) {
    if (!c || !cfgp || !tol || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    int rc = mlmc_validate(*cfgp, false);
    if (rc != MP_OK) return rc;
    // validate_arithmetic_asian_mlmc_tolerance_config (cpu.rs:3512-3529)
    if (!(std::isfinite(tol->target_stderr) && tol->target_stderr > 0.0))
        return fail(MP_E_INVALID_CONFIG, "target_stderr must be finite and > 0");
    if (tol->pilot_paths_per_level <= 1) return fail(MP_E_INVALID_CONFIG, "pilot_paths_per_level must be > 1");
    if (tol->min_step_updates == 0) return fail(MP_E_INVALID_CONFIG, "min_step_updates must be > 0");
    if (tol->max_step_updates < tol->min_step_updates)
        return fail(MP_E_INVALID_CONFIG, "max_step_updates must be >= min_step_updates");
    HIPCHK(hipSetDevice(c->device));
    const mp_mlmc_config& cfg = *cfgp;
    const uint32_t reps = mlmc_structured(cfg) ? std::max<uint32_t>(cfg.scramble_replicates, 1) : 1;
    mp_mlmc_config pilot = cfg;
    pilot.scramble_replicates = 1;
    std::memset(out, 0, sizeof(*out));
    mp_mlmc_allocation a;
    rc = mlmc_pilot(c, pilot, tol->pilot_paths_per_level, a);
    if (rc != MP_OK) return rc;
    mp_mlmc_allocation pilot_levels = a;  // the tune passes re-measure with the same seed
    mlmc_allocate(a, tol->target_stderr * tol->target_stderr * (double)reps, true);
    double est = mlmc_plan_stderr(a, reps);
    uint64_t updates = sat_mul(a.estimated_step_updates, reps);
    if (updates < tol->min_step_updates) {
        const uint64_t per = std::max<uint64_t>((tol->min_step_updates + reps - 1) / reps, 1);
        a = pilot_levels;
        mlmc_allocate(a, (double)per, false);
        a.target_step_updates = per;
        est = mlmc_plan_stderr(a, reps);
        updates = sat_mul(a.estimated_step_updates, reps);
    }
    int hit = 0;
    if (updates > tol->max_step_updates) {
        hit = 1;
        const uint64_t per = std::max<uint64_t>(tol->max_step_updates / reps, 1);
        a = pilot_levels;
        mlmc_allocate(a, (double)per, false);
        a.target_step_updates = per;
        est = mlmc_plan_stderr(a, reps);
        updates = sat_mul(a.estimated_step_updates, reps);
    }
    out->target_stderr = tol->target_stderr;
    out->estimated_stderr = est;
    out->target_met = est <= tol->target_stderr && updates <= tol->max_step_updates;
    out->max_step_updates_hit = hit;
    out->scramble_replicates = reps;
    out->estimated_step_updates = updates;
    out->allocation = a;
    out->recommended = cfg;
    for (uint32_t l = 0; l < cfg.levels; ++l) out->recommended.paths_per_level[l] = a.levels[l].recommended_paths;
    if (cfg.sampling == MP_PSEUDORANDOM) out->recommended.scramble_replicates = 1;
    return MP_OK;
}

MP_EXPORT int mp_hip_mlmc_level_from_normals(mp_ctx* c, const mp_mlmc_config* cfgp, uint32_t level,
                                             uint64_t paths, const double* normals, double* sum_sq) {
    if (!c || !cfgp || !normals || !sum_sq) return fail(MP_E_INVALID_CONFIG, "null argument");
    int rc = mlmc_validate(*cfgp, false);
    if (rc != MP_OK) return rc;
    if (level >= cfgp->levels) return fail(MP_E_INVALID_CONFIG, "level out of range");
    if (paths == 0 || paths > 0xFFFFFFFFull) return fail(MP_E_INVALID_CONFIG, "paths must be in 1..2^32-1");
    HIPCHK(hipSetDevice(c->device));
    mp_mlmc_config cfg = *cfgp;
    cfg.sampling = MP_PSEUDORANDOM;
    prepared P;
    rc = mlmc_prepare(cfg, mlmc_unit{cfg.seed, level, paths}, P);
    if (rc != MP_OK) return rc;
    mp_result r;
    rc = run_tiles_single(c, P, P.precision, &r, normals, nullptr, nullptr, nullptr, nullptr);
    if (rc != MP_OK) return rc;
    sum_sq[0] = r.price;
    sum_sq[1] = r.std_error;
    return MP_OK;
}

MP_EXPORT int mp_hip_lsm_from_normals(mp_ctx* c, const mp_request* req, const double* normals,
                                      mp_result* out, double* coefficients_out,
                                      uint32_t* exercise_out) {
    if (!c || !req || !normals || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    prepared P;
    int rc = prepare_request(*req, P);
    if (rc != MP_OK) return rc;
    if (P.kind != KIND_LSM) return fail(MP_E_INCOMPATIBLE_INPUT, "not an American/Bermudan put request");
    return run_lsm(c, *req, P, P.precision, normals, out, coefficients_out, exercise_out);
}

MP_EXPORT int mp_hip_draws(mp_ctx* c, const mp_request* req, int kind, uint32_t stream,
                           uint64_t path0, uint64_t n_paths, uint64_t n_draws, double* out) {
    if (!c || !req || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    if (kind < 0 || kind > 4) return fail(MP_E_INVALID_CONFIG, "draw kind must be 0..4");
    if (n_paths == 0 || n_draws == 0) return MP_OK;
    if (path0 + n_paths > 0xFFFFFFFFull) return fail(MP_E_UNSUPPORTED, "path index beyond 2^32");
    HIPCHK(hipSetDevice(c->device));
    mp_dreq d{};
    d.key0 = (uint32_t)req->seed;
    d.key1 = (uint32_t)req->contract_id;
    d.seed_hi = (uint32_t)(req->seed >> 32);
    mp_qmc_params qp{};
    qp.base_seed = (uint32_t)req->seed;
    const size_t bytes = (size_t)n_paths * n_draws * 8;
    ENSURE(c->reqs, sizeof(mp_dreq));
    ENSURE(c->draws, bytes);
    HIPCHK(hipMemcpyAsync(c->reqs.p, &d, sizeof(d), hipMemcpyHostToDevice, c->stream));
    HIPCHK(mp::launch_draws(c->reqs.as<mp_dreq>(), qp, c->d_nib, kind, stream, path0, n_paths, n_draws,
                            c->draws.as<double>(), c->stream));
    HIPCHK(hipMemcpyAsync(out, c->draws.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MP_OK;
}

MP_EXPORT int mp_hip_request_footprint(const mp_request* req, uint64_t* bytes) {
    if (!req || !bytes) return fail(MP_E_INVALID_CONFIG, "null argument");
    prepared P;
    int rc = prepare_request(*req, P);
    if (rc != MP_OK) return rc;
    uint64_t b = sizeof(mp_dreq) + 64;
    switch (P.kind) {
        case KIND_LSM: {
            const uint64_t rb = P.precision == 64 ? 8 : 4;
            b += (uint64_t)(req->n_steps - 1) * req->n_paths * rb + req->n_paths * (8 + 2) +
                 (req->n_steps + 1) * 64 + 4096 * 16 * 8;
            break;
        }
        case KIND_QMC:
            b += (uint64_t)((P.d.n_paths + 63) / 64) * MP_PARTIAL_STRIDE * 8 + 64 * 1024;
            break;
        default:
            b += (uint64_t)P.d.n_tiles * (MP_PARTIAL_STRIDE * 8 + 8);
    }
    *bytes = b;
    return MP_OK;
}

MP_EXPORT int mp_hip_last_timing(mp_ctx* c, mp_timing* out) {
    if (!c || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    *out = c->last;
    return MP_OK;
}

MP_EXPORT int mp_hip_get_tuning(mp_ctx* c, const mp_request* req, mp_tuning* out) {
    if (!c || !req || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    prepared P;
    int rc = prepare_request(*req, P);
    if (rc != MP_OK) return rc;
    *out = tuning_for(c, P.d, P.precision);
    return MP_OK;
}

MP_EXPORT int mp_hip_set_tuning(mp_ctx* c, const mp_request* req, const mp_tuning* t) {
    if (!c || !req || !t) return fail(MP_E_INVALID_CONFIG, "null argument");
    if (t->block_size != 64 && t->block_size != 128 && t->block_size != 256)
        return fail(MP_E_INVALID_CONFIG, "block_size must be 64, 128 or 256");
    if (t->paths_per_lane != 1 && t->paths_per_lane != 2 && t->paths_per_lane != 4)
        return fail(MP_E_INVALID_CONFIG, "paths_per_lane must be 1, 2 or 4");
    if (t->step_unroll != 0 && t->step_unroll != 1 && t->step_unroll != 2 && t->step_unroll != 4)
        return fail(MP_E_INVALID_CONFIG, "step_unroll must be 1, 2 or 4 (0 = 1)");
    prepared P;
    int rc = prepare_request(*req, P);
    if (rc != MP_OK) return rc;
    mp_tuning v = *t;
    v.tile_paths = (int32_t)P.d.tile_paths;
    if (v.step_unroll == 0) v.step_unroll = 1;
    c->tuning[tuning_key(P.d, P.precision)] = v;
    return MP_OK;
}

// Times each (block size, paths per lane) candidate on the request (or a
// bounded sample of it) and keeps the fastest for its class.  Results do not
// depend on the choice: tiles, hence partial sums, are fixed by the request.
MP_EXPORT int mp_hip_autotune(mp_ctx* c, const mp_request* req, mp_tuning* best, double* best_ms) {
    if (!c || !req || !best) return fail(MP_E_INVALID_CONFIG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    prepared P;
    int rc = prepare_request(*req, P);
    if (rc != MP_OK) return rc;
    if (P.kind != KIND_TILE) return fail(MP_E_UNSUPPORTED, "autotuning covers work-queue requests");
    mp_request sample = *req;
    const uint64_t budget = 1ull << 28;  // path-steps per trial
    if (sample.n_paths * std::max<uint64_t>(P.d.n_steps, 1) > budget)
        sample.n_paths = std::max<uint64_t>(budget / std::max<uint64_t>(P.d.n_steps, 1), 4096);
    const uint64_t key = tuning_key(P.d, P.precision);
    const int blocks[] = {128, 256};
    const int ppls[] = {1, 2, 4};
    // the step unroll selects compiled variants of the FP32 GBM-family
    // Philox body only; the other bodies have one step loop
    const bool unrolls = P.precision == 32 && gbm_philox_tile(P.d);
    const int n_unroll = unrolls ? 3 : 1;
    const int unroll_of[] = {1, 2, 4};
    double best_t = 1e300;
    mp_tuning bt{};
    for (int bs : blocks)
        for (int ppl : ppls)
            for (int ui = 0; ui < n_unroll; ++ui) {
                mp_tuning t{bs, ppl, (int32_t)P.d.tile_paths, unroll_of[ui]};
                c->tuning[key] = t;
                mp_result r;
                double tmin = 1e300;
                for (int rep = 0; rep < 3; ++rep) {
                    rc = price_one(c, sample, P.precision, &r);
                    if (rc != MP_OK) return rc;
                    tmin = std::min(tmin, c->last.main_kernel_ms);
                }
                if (tmin < best_t) {
                    best_t = tmin;
                    bt = t;
                }
            }
    c->tuning[key] = bt;
    *best = bt;
    if (best_ms) *best_ms = best_t;
    return MP_OK;
}

// ===========================================================================
// Continuous-batching server
// ===========================================================================
namespace {

struct ticket_state {
    mp_request req;
    std::vector<uint32_t> schedule;
    mp_result res;
    int done = 0;
    uint64_t footprint = 0;
};

}  // namespace

struct mp_server {
    mp_ctx* ctx = nullptr;
    mp_serve_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_start;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, std::unique_ptr<ticket_state>> tickets;
    uint64_t next_ticket = 1;
    // per worker: device bytes its context's arenas hold after its last
    // pricing, and the bytes reserved for the batch it is pricing now
    struct slot_bytes {
        uint64_t held = 0, reserved = 0;
    };
    std::vector<slot_bytes> slots;
    uint64_t max_admitted = 0;
    uint64_t launches = 0;
    uint64_t completed = 0;
    uint64_t trims = 0;
    bool trim_wanted = false;  // a request waits for memory that idle workers hold
    bool stop = false;
    int started = 0, start_failed = 0;
    std::vector<std::thread> workers;
};

namespace {

// Bytes a request may add to a worker's arenas: its footprint, the arenas'
// 25% growth slack and the small per-call buffers.
uint64_t reserve_bytes(uint64_t footprint) { return footprint + footprint / 4 + (1u << 20); }

// Device bytes the server's workers hold or have reserved, except worker
// `skip` (-1: all of them).
uint64_t server_bytes(const mp_server* s, int skip) {
    uint64_t b = 0;
    for (size_t v = 0; v < s->slots.size(); ++v)
        if ((int)v != skip) b += s->slots[v].held + s->slots[v].reserved;
    return b;
}

// A worker owns one context and stream (one in-flight launch slot).  It takes
// pending requests in arrival order while the server's device bytes -- every
// worker's arenas plus the reservations of the batches in flight -- stay
// within the HBM budget, the first request of a batch included, packs the
// work-queue requests into one persistent launch and prices the others one by
// one; requests that arrive meanwhile wait for the next free worker.  When
// the head request does not fit, the worker frees its own arenas if that
// makes room, and otherwise asks idle workers to free theirs and waits for
// running batches to finish.  A request over the budget on its own is
// rejected at submission, so the head request always fits eventually.
void server_worker(mp_server* s, int slot) {
    (void)hipSetDevice(s->ctx->device);
    mp_ctx* own = nullptr;
    const int orc = mp_hip_open(s->ctx->device, &own);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (orc == MP_OK) {
            s->started += 1;
            s->slots[slot].held = ctx_device_bytes(own);
        } else {
            s->start_failed += 1;
        }
    }
    s->cv_start.notify_all();
    if (orc != MP_OK) return;
    // other workers' launches share the device, so a one-launch LSM pricing
    // (which needs its whole grid resident) could wait on them: price LSM
    // requests with the matched per-date launches, which give the same bits
    own->lsm_fused = false;
    const uint64_t budget = s->cfg.hbm_budget_bytes;
    for (;;) {
        std::vector<uint64_t> take;
        uint64_t bytes = 0;
        bool trim_self = false, quit = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop && s->pending.empty()) {
                    quit = true;
                    break;
                }
                if (s->trim_wanted && s->slots[slot].held > 0) {
                    trim_self = true;  // an idle worker gives its arenas back
                    break;
                }
                uint64_t head_fp = 0;
                while (!s->pending.empty() && (int)take.size() < s->cfg.max_batch) {
                    const uint64_t id = s->pending.front();
                    auto it = s->tickets.find(id);
                    if (it == s->tickets.end()) {
                        s->pending.pop_front();
                        continue;
                    }
                    const uint64_t rb = reserve_bytes(it->second->footprint);
                    if (server_bytes(s, slot) + s->slots[slot].held + bytes + rb > budget) {
                        head_fp = rb;
                        break;
                    }
                    s->pending.pop_front();
                    bytes += rb;
                    take.push_back(id);
                }
                if (!take.empty()) break;
                if (head_fp) {
                    if (s->slots[slot].held > 0 && server_bytes(s, slot) + head_fp <= budget) {
                        trim_self = true;  // this worker's own arenas are in the way
                        break;
                    }
                    s->trim_wanted = true;
                    s->cv_work.notify_all();
                }
                s->cv_work.wait(lk);
            }
            if (!take.empty()) {
                s->slots[slot].reserved = bytes;
                s->max_admitted = std::max(s->max_admitted, server_bytes(s, -1));
                s->launches += 1;
            }
        }
        if (quit) break;
        if (trim_self) {
            ctx_trim(own);
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->slots[slot].held = ctx_device_bytes(own);
                s->trims += 1;
                bool idle_holding = false;
                for (const auto& v : s->slots) idle_holding |= (v.reserved == 0 && v.held > 0);
                if (!idle_holding) s->trim_wanted = false;
            }
            s->cv_work.notify_all();
            continue;
        }
        // split into work-queue requests (one persistent launch) and the rest
        std::vector<mp_request> wq;
        std::vector<uint64_t> wq_ids, rest_ids;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (uint64_t id : take) {
                ticket_state* t = s->tickets[id].get();
                prepared P;
                if (prepare_request(t->req, P) == MP_OK && P.kind == KIND_TILE && !t->req.fp64_check) {
                    wq.push_back(t->req);
                    wq_ids.push_back(id);
                } else {
                    rest_ids.push_back(id);
                }
            }
        }
        std::vector<mp_result> res(wq.size());
        int rc = wq.empty() ? MP_OK : mp_hip_price_batch(own, wq.data(), wq.size(), res.data());
        std::vector<mp_result> rres(rest_ids.size());
        std::vector<int> rrc(rest_ids.size(), MP_OK);
        for (size_t i = 0; i < rest_ids.size(); ++i) {
            mp_request rq;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                rq = s->tickets[rest_ids[i]]->req;
            }
            rrc[i] = mp_hip_price(own, &rq, &rres[i]);
        }
        const uint64_t held = ctx_device_bytes(own);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (size_t i = 0; i < wq_ids.size(); ++i) {
                ticket_state* t = s->tickets[wq_ids[i]].get();
                t->res = res[i];
                if (rc != MP_OK) t->res.status = rc;
                t->done = 1;
            }
            for (size_t i = 0; i < rest_ids.size(); ++i) {
                ticket_state* t = s->tickets[rest_ids[i]].get();
                t->res = rres[i];
                if (rrc[i] != MP_OK) t->res.status = rrc[i];
                t->done = 1;
            }
            s->completed += take.size();
            // the batch's allocations are now part of the arenas
            s->slots[slot].held = held;
            s->slots[slot].reserved = 0;
            s->max_admitted = std::max(s->max_admitted, server_bytes(s, -1));
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    mp_hip_close(own);
}

}  // namespace

MP_EXPORT int mp_hip_server_open(mp_ctx* c, const mp_serve_config* cfg, mp_server** out) {
    if (!c || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    std::unique_ptr<mp_server> s(new mp_server());
    s->ctx = c;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 2;
    if (s->cfg.max_inflight > 8) s->cfg.max_inflight = 8;
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 4096;
    if (s->cfg.hbm_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipMemGetInfo(&fr, &tot));
        s->cfg.hbm_budget_bytes = (uint64_t)(fr * 0.75);
    }
    s->slots.resize(s->cfg.max_inflight);
    for (int i = 0; i < s->cfg.max_inflight; ++i) s->workers.emplace_back(server_worker, s.get(), i);
    int failed = 0;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_start.wait(lk, [&] { return s->started + s->start_failed == s->cfg.max_inflight; });
        failed = s->start_failed;
        if (failed) s->stop = true;
    }
    if (failed) {
        s->cv_work.notify_all();
        for (auto& t : s->workers) t.join();
        return fail(MP_E_HIP, "a server worker could not open its device context");
    }
    *out = s.release();
    return MP_OK;
}

MP_EXPORT int mp_hip_server_stats(mp_server* s, mp_serve_stats* stats) {
    if (!s || !stats) return fail(MP_E_INVALID_CONFIG, "null argument");
    std::memset(stats, 0, sizeof(*stats));
    std::lock_guard<std::mutex> lk(s->mu);
    stats->completed = s->completed;
    stats->launches = s->launches;
    stats->max_admitted_bytes = s->max_admitted;
    return MP_OK;
}

MP_EXPORT void mp_hip_server_close(mp_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    delete s;
}

MP_EXPORT int mp_hip_submit(mp_server* s, const mp_request* req, uint64_t* ticket) {
    if (!s || !req || !ticket) return fail(MP_E_INVALID_CONFIG, "null argument");
    std::unique_ptr<ticket_state> t(new ticket_state());
    t->req = *req;
    if (req->exercise_steps && req->n_exercise > 0) {
        t->schedule.assign(req->exercise_steps, req->exercise_steps + req->n_exercise);
        t->req.exercise_steps = t->schedule.data();
    }
    int rc = mp_hip_request_footprint(&t->req, &t->footprint);
    if (rc != MP_OK) return rc;
    if (reserve_bytes(t->footprint) > s->cfg.hbm_budget_bytes)
        return fail(MP_E_OOM, "request footprint exceeds the server's HBM budget");
    std::memset(&t->res, 0, sizeof(t->res));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const uint64_t id = s->next_ticket++;
        s->tickets[id] = std::move(t);
        s->pending.push_back(id);
        *ticket = id;
    }
    s->cv_work.notify_one();
    return MP_OK;
}

MP_EXPORT int mp_hip_poll(mp_server* s, uint64_t ticket, mp_result* out, int* done) {
    if (!s || !out || !done) return fail(MP_E_INVALID_CONFIG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(MP_E_NOT_FOUND, "unknown ticket");
    *done = it->second->done;
    if (*done) {
        *out = it->second->res;
        s->tickets.erase(it);
    }
    return MP_OK;
}

MP_EXPORT int mp_hip_wait(mp_server* s, uint64_t ticket, mp_result* out) {
    if (!s || !out) return fail(MP_E_INVALID_CONFIG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(MP_E_NOT_FOUND, "unknown ticket");
    ticket_state* t = it->second.get();
    s->cv_done.wait(lk, [&] { return t->done != 0; });
    *out = t->res;
    s->tickets.erase(ticket);
    return out->status == MP_OK ? MP_OK : fail(out->status, "request failed");
}

MP_EXPORT int mp_hip_cancel(mp_server* s, uint64_t ticket) {
    if (!s) return fail(MP_E_INVALID_CONFIG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = std::find(s->pending.begin(), s->pending.end(), ticket);
    if (it == s->pending.end()) return fail(MP_E_NOT_FOUND, "ticket is not pending");
    s->pending.erase(it);
    ticket_state* t = s->tickets[ticket].get();
    t->res.status = MP_E_CANCELLED;
    t->done = 1;
    s->cv_done.notify_all();
    return MP_OK;
}

MP_EXPORT int mp_hip_serve_trace(mp_ctx* c, const mp_serve_config* cfg, const mp_request* reqs,
                                 const double* arrival_s, size_t n, mp_result* results,
                                 double* latency_ms, mp_serve_stats* stats) {
    if (!c || !reqs || !arrival_s || !results || !latency_ms || !stats)
        return fail(MP_E_INVALID_CONFIG, "null argument");
    mp_server* s = nullptr;
    int rc = mp_hip_server_open(c, cfg, &s);
    if (rc != MP_OK) return rc;
    std::vector<uint64_t> tickets(n);
    std::vector<std::chrono::steady_clock::time_point> t_arr(n);
    std::vector<size_t> live;  // submitted, not yet completed (polled in submission order)
    const auto start = std::chrono::steady_clock::now();
    size_t next = 0, finished = 0;
    uint64_t path_steps = 0;
    while (finished < n) {
        const double now = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        while (next < n && arrival_s[next] <= now) {
            t_arr[next] = start + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                      std::chrono::duration<double>(arrival_s[next]));
            rc = mp_hip_submit(s, &reqs[next], &tickets[next]);
            if (rc != MP_OK) {
                mp_hip_server_close(s);
                return rc;
            }
            live.push_back(next);
            ++next;
        }
        size_t keep = 0;
        for (size_t j = 0; j < live.size(); ++j) {
            const size_t i = live[j];
            int done = 0;
            rc = mp_hip_poll(s, tickets[i], &results[i], &done);
            if (rc != MP_OK) {
                mp_hip_server_close(s);
                return rc;
            }
            if (done) {
                ++finished;
                latency_ms[i] = std::chrono::duration<double, std::milli>(
                                    std::chrono::steady_clock::now() - t_arr[i]).count();
                path_steps += reqs[i].n_paths * std::max<uint64_t>(results[i].n_steps, 1);
            } else {
                live[keep++] = i;
            }
        }
        live.resize(keep);
        if (finished < n) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::vector<double> lat(latency_ms, latency_ms + n);
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) { return lat.empty() ? 0.0 : lat[std::min(lat.size() - 1, (size_t)(q * (lat.size() - 1) + 0.5))]; };
    std::memset(stats, 0, sizeof(*stats));
    stats->completed = n;
    stats->wall_s = wall;
    stats->requests_per_s = n / wall;
    stats->path_steps_per_s = path_steps / wall;
    stats->p50_ms = pct(0.50);
    stats->p99_ms = pct(0.99);
    stats->max_ms = lat.empty() ? 0.0 : lat.back();
    stats->launches = s->launches;
    stats->max_admitted_bytes = s->max_admitted;
    mp_hip_server_close(s);
    return MP_OK;
}
