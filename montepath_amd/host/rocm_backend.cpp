// RocmHipBackend and the backend helpers over the C-ABI (montepath_hip.h).
#include "rocm_backend.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "montepath_hip.h"

namespace montepath {

const char* backend_id_name(BackendId id) {
    switch (id) {
        case BackendId::CpuNative: return "cpu_native";
        case BackendId::NvidiaCuda: return "nvidia_cuda";
        case BackendId::AppleMetal: return "apple_metal";
        case BackendId::RocmHip: return "rocm_hip";
    }
    return "unknown";
}

const char* support_level_name(SupportLevel level) {
    switch (level) {
        case SupportLevel::Supported: return "supported";
        case SupportLevel::SupportedWithFallbacks: return "supported_with_fallbacks";
        case SupportLevel::Unsupported: return "unsupported";
    }
    return "unknown";
}

BackendError BackendError::unknown_device(const std::string& device_id) {
    return BackendError(Kind::UnknownDevice, "device '" + device_id + "' is not available for this backend");
}

BackendError BackendError::incompatible_execution_input() {
    return BackendError(Kind::IncompatibleExecutionInput, "execution input is not compatible with compiled artifact");
}

BackendError BackendError::unsupported_feature(const std::string& what) {
    return BackendError(Kind::UnsupportedFeature, "unsupported feature: " + what);
}

namespace {

// Engine contexts of this thread, one per device, closed when the thread ends.
struct ThreadContexts {
    std::map<int, mp_ctx*> ctx;
    ~ThreadContexts() {
        for (auto& kv : ctx) mp_hip_close(kv.second);
    }
};
thread_local ThreadContexts t_contexts;

std::string last_error() {
    const char* m = mp_hip_last_error();
    return m ? std::string(m) : std::string();
}

// mp_status -> BackendError (the header's mapping)
void check(int rc, const std::string& device_id) {
    if (rc == MP_OK) return;
    if (rc == MP_E_UNKNOWN_DEVICE) throw BackendError::unknown_device(device_id);
    if (rc == MP_E_INCOMPATIBLE_INPUT) throw BackendError::incompatible_execution_input();
    throw BackendError::unsupported_feature(last_error());
}

mp_ctx* context_for(int device, const std::string& device_id) {
    auto it = t_contexts.ctx.find(device);
    if (it != t_contexts.ctx.end()) return it->second;
    mp_ctx* c = nullptr;
    check(mp_hip_open(device, &c), device_id);
    t_contexts.ctx[device] = c;
    return c;
}

int device_count() {
    int n = 0;
    if (mp_hip_device_count(&n) != MP_OK) return 0;
    return n;
}

int32_t sampling_code(SamplingMethod s) {
    switch (s) {
        case SamplingMethod::Pseudorandom: return MP_PSEUDORANDOM;
        case SamplingMethod::RandomizedHalton: return MP_RANDOMIZED_HALTON;
        case SamplingMethod::LatinHypercube: return MP_LATIN_HYPERCUBE;
        case SamplingMethod::ScrambledSobol: return MP_SCRAMBLED_SOBOL;
        case SamplingMethod::ScrambledSobolBrownianBridge: return MP_SCRAMBLED_SOBOL_BRIDGE;
    }
    return MP_PSEUDORANDOM;
}

int32_t technique_code(MonteCarloTechnique t) {
    switch (t) {
        case MonteCarloTechnique::Standard: return MP_TECHNIQUE_STANDARD;
        case MonteCarloTechnique::Antithetic: return MP_TECHNIQUE_ANTITHETIC;
        case MonteCarloTechnique::ControlVariate: return MP_TECHNIQUE_CONTROL_VARIATE;
    }
    return MP_TECHNIQUE_STANDARD;
}

// The GBM-family request of one execution input (stepwise paths, as the
// GPU backends' execution of these inputs is, backend/mod.rs:484-540).
template <class Cfg>
mp_request gbm_request(const Cfg& c, int32_t workload, int precision) {
    mp_request r;
    std::memset(&r, 0, sizeof(r));
    r.workload = workload;
    r.precision = precision;
    r.sampling = sampling_code(c.sampling);
    r.technique = technique_code(c.technique);
    r.method = MP_METHOD_STEPWISE;
    r.n_paths = c.n_paths;
    r.n_steps = c.n_steps;
    r.seed = c.seed;
    r.s0 = c.s0;
    r.k = c.k;
    r.r = c.r;
    r.sigma = c.sigma;
    r.t = c.t;
    r.barrier = 80.0;
    return r;
}

// The request of one execution input.  schedule receives a Bermudan
// schedule, which the request points into.
mp_request to_request(const BackendExecutionInput& input, int precision, std::vector<uint32_t>& schedule) {
    if (const auto* e = std::get_if<EuropeanCallConfig>(&input)) return gbm_request(*e, MP_EUROPEAN_CALL, precision);
    if (const auto* a = std::get_if<ArithmeticAsianCallConfig>(&input))
        return gbm_request(*a, MP_ARITHMETIC_ASIAN_CALL, precision);
    if (const auto* d = std::get_if<DownAndOutCallConfig>(&input)) {
        mp_request r = gbm_request(*d, MP_DOWN_AND_OUT_CALL, precision);
        r.barrier = d->barrier;
        return r;
    }
    if (const auto* l = std::get_if<LookbackCallConfig>(&input)) return gbm_request(*l, MP_LOOKBACK_CALL, precision);
    mp_request r;
    std::memset(&r, 0, sizeof(r));
    r.precision = precision;
    r.sampling = MP_PSEUDORANDOM;
    r.technique = MP_TECHNIQUE_STANDARD;
    r.method = MP_METHOD_STEPWISE;
    if (const auto* h = std::get_if<HestonEuropeanCallConfig>(&input)) {
        r.workload = MP_HESTON_EUROPEAN_CALL;
        r.heston_scheme = MP_HESTON_REFERENCE_FLOOR;
        r.technique = technique_code(h->technique);
        r.n_paths = h->n_paths;
        r.n_steps = h->n_steps;
        r.seed = h->seed;
        r.s0 = h->s0;
        r.k = h->k;
        r.r = h->r;
        r.t = h->t;
        r.v0 = h->v0;
        r.kappa = h->kappa;
        r.theta = h->theta;
        r.xi = h->vol_of_vol;
        r.rho = h->rho;
    } else if (const auto* m = std::get_if<MertonJumpDiffusionCallConfig>(&input)) {
        r.workload = MP_MERTON_JUMP_DIFFUSION_CALL;
        r.n_paths = m->n_paths;
        r.n_steps = 0;  // one step to maturity, as the reference's config
        r.seed = m->seed;
        r.s0 = m->s0;
        r.k = m->k;
        r.r = m->r;
        r.sigma = m->sigma;
        r.t = m->t;
        r.jump_intensity = m->jump_intensity;
        r.jump_mean = m->jump_mean;
        r.jump_volatility = m->jump_volatility;
    } else if (const auto* b = std::get_if<BasketCallConfig>(&input)) {
        r.workload = MP_BASKET_CALL;
        r.sampling = sampling_code(b->sampling);
        r.technique = technique_code(b->technique);
        r.n_paths = b->n_paths;
        r.n_steps = 1;  // terminal, as the reference's two-asset basket
        r.seed = b->seed;
        r.k = b->k;
        r.r = b->r;
        r.t = b->t;
        r.n_assets = 2;
        r.asset_s0[0] = b->s01;
        r.asset_s0[1] = b->s02;
        r.asset_sigma[0] = b->sigma1;
        r.asset_sigma[1] = b->sigma2;
        r.asset_weight[0] = b->weight1;
        r.asset_weight[1] = b->weight2;
        r.correlation[0] = 1.0;
        r.correlation[1] = b->rho;
        r.correlation[2] = b->rho;
        r.correlation[3] = 1.0;
    } else {
        const bool bermudan = std::holds_alternative<BermudanPutConfig>(input);
        const AmericanPutConfig* am = std::get_if<AmericanPutConfig>(&input);
        const BermudanPutConfig* be = std::get_if<BermudanPutConfig>(&input);
        r.workload = bermudan ? MP_BERMUDAN_PUT : MP_AMERICAN_PUT;
        r.allow_intrinsic_now = bermudan ? 0 : 1;
        r.n_paths = bermudan ? be->n_paths : am->n_paths;
        r.n_steps = bermudan ? be->n_steps : am->n_steps;
        r.seed = bermudan ? be->seed : am->seed;
        r.s0 = bermudan ? be->s0 : am->s0;
        r.k = bermudan ? be->k : am->k;
        r.r = bermudan ? be->r : am->r;
        r.sigma = bermudan ? be->sigma : am->sigma;
        r.t = bermudan ? be->t : am->t;
        r.basis_degree = static_cast<int32_t>(bermudan ? be->basis_degree : am->basis_degree);
        if (bermudan) {
            schedule.clear();
            for (std::size_t s : be->exercise_steps) schedule.push_back(static_cast<uint32_t>(s));
            r.n_exercise = static_cast<int32_t>(schedule.size());
            r.exercise_steps = schedule.data();
        }
    }
    return r;
}

// (n_paths, n_steps) of an input; n_steps is nullopt for the step-less
// Merton and basket configs
std::pair<std::size_t, std::optional<std::size_t>> shape_of(const BackendExecutionInput& input) {
    return std::visit(
        [](const auto& c) -> std::pair<std::size_t, std::optional<std::size_t>> {
            using C = std::decay_t<decltype(c)>;
            if constexpr (std::is_same_v<C, MertonJumpDiffusionCallConfig> || std::is_same_v<C, BasketCallConfig>)
                return {c.n_paths, std::nullopt};
            else
                return {c.n_paths, c.n_steps};
        },
        input);
}

std::uint64_t splitmix64(std::uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    std::uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

RocmHipBackend::RocmHipBackend(int precision) : precision_(precision) {
    if (precision != 32 && precision != 64) throw BackendError::unsupported_feature("precision must be 32 or 64");
}

BackendInfo RocmHipBackend::describe_backend() const {
    return BackendInfo{BackendId::RocmHip,
                       "AMD ROCm HIP",
                       mp_hip_version() ? mp_hip_version() : "",
                       "rocm",
                       {"float32", "float64"},
                       {"philox4x32-10", "scrambled_sobol", "randomized_halton", "latin_hypercube"},
                       {"iid", "qmc"},
                       {"sum", "mean", "variance"}};
}

std::vector<DeviceInfo> RocmHipBackend::discover_devices() const {
    std::vector<DeviceInfo> out;
    const int n = device_count();
    for (int d = 0; d < n; ++d) {
        const std::string id = "rocm:" + std::to_string(d);
        try {
            mp_device_info info;
            std::memset(&info, 0, sizeof(info));
            check(mp_hip_device_info(context_for(d, id), &info), id);
            out.push_back(DeviceInfo{id, BackendId::RocmHip, info.name, "amd",
                                     static_cast<std::size_t>(info.memory_total_mb),
                                     static_cast<std::size_t>(info.memory_free_mb), info.supports_float64 != 0, false,
                                     static_cast<std::size_t>(info.max_threads_hint)});
        } catch (const BackendError&) {
            // a device the runtime lists but cannot open is not offered
        }
    }
    return out;
}

int RocmHipBackend::device_index(const std::string& device_id) const {
    if (device_id.rfind("rocm:", 0) != 0 || device_id.size() == 5) return -1;
    int d = 0;
    for (std::size_t i = 5; i < device_id.size(); ++i) {
        if (device_id[i] < '0' || device_id[i] > '9' || d > 1000000) return -1;
        d = d * 10 + (device_id[i] - '0');
    }
    return d;
}

void RocmHipBackend::validate_device(const DeviceInfo& device) const {
    const int d = device_index(device.device_id);
    if (device.backend_id != BackendId::RocmHip || d < 0 || d >= device_count())
        throw BackendError::unknown_device(device.device_id);
}

SupportReport RocmHipBackend::supports(const ExecutionPlan& /*plan*/, const DeviceInfo& device) const {
    try {
        validate_device(device);
    } catch (const BackendError&) {
        return SupportReport{BackendId::RocmHip, device.device_id, SupportLevel::Unsupported, {"unknown_device"}, {}};
    }
    return SupportReport{BackendId::RocmHip, device.device_id, SupportLevel::Supported, {}, {}};
}

CostEstimate RocmHipBackend::estimate_cost(const ExecutionPlan& plan, const DeviceInfo& device) const {
    // Measured on one MI355X: about 1e12 path-steps/s for FP32 GBM-family
    // batches and 2e11 in FP64, plus tens of microseconds per request.
    const double path_steps = static_cast<double>(plan.n_paths) * static_cast<double>(plan.n_steps);
    const double rate_per_ms = precision_ == 64 ? 2.0e8 : 1.0e9;
    const double runtime_ms = std::max(0.02, path_steps / rate_per_ms);
    // peak device memory of the request (the engine's footprint model), or
    // the reference's GPU model when the request is not valid
    double peak_mb = 0.0;
    EuropeanCallConfig c;
    c.n_paths = plan.n_paths;
    c.n_steps = plan.n_steps;
    std::vector<uint32_t> schedule;
    const mp_request r = to_request(BackendExecutionInput{c}, precision_, schedule);
    std::uint64_t bytes = 0;
    if (mp_hip_request_footprint(&r, &bytes) == MP_OK) {
        peak_mb = std::ceil(static_cast<double>(bytes) / (1024.0 * 1024.0));
    } else {
        const GpuChunkingPlan ch = plan_gpu_chunking(plan.n_paths, device.memory_total_mb,
                                                     GpuChunkingConfig{estimate_gpu_bytes_per_path(plan), 0.75, 32768, 8192});
        peak_mb = static_cast<double>(ch.estimated_peak_memory_mb);
    }
    return CostEstimate{BackendId::RocmHip, device.device_id, 0.0, runtime_ms, runtime_ms, std::max(peak_mb, 1.0),
                        "medium"};
}

CompiledArtifact RocmHipBackend::compile(const ExecutionPlan& plan, const DeviceInfo& device) const {
    validate_device(device);
    // the kernels are built ahead of time; compiling opens this thread's
    // context on the device, so a device that cannot run fails here
    (void)context_for(device_index(device.device_id), device.device_id);
    return CompiledArtifact{"rocm-native:" + std::to_string(plan.n_paths) + ":" + std::to_string(plan.n_steps) + ":" +
                                std::to_string(plan.features.step_count),
                            BackendId::RocmHip,
                            device.device_id,
                            plan.n_paths,
                            plan.n_steps,
                            plan.planner_mode,
                            ArtifactExecutionMode::GpuNative};
}

RunOutput RocmHipBackend::execute(const CompiledArtifact& artifact, const BackendExecutionInput& input) const {
    if (artifact.backend_id != BackendId::RocmHip) throw BackendError::incompatible_execution_input();
    // the shape check of the GPU backends' executors (backend/mod.rs:488-490);
    // step-less inputs match on their path count
    const auto shape = shape_of(input);
    if (shape.first != artifact.n_paths || (shape.second && *shape.second != artifact.n_steps))
        throw BackendError::incompatible_execution_input();
    const int d = device_index(artifact.device_id);
    if (d < 0) throw BackendError::unknown_device(artifact.device_id);
    std::vector<uint32_t> schedule;
    const mp_request r = to_request(input, precision_, schedule);
    const auto started = std::chrono::steady_clock::now();
    mp_result out;
    std::memset(&out, 0, sizeof(out));
    check(mp_hip_price(context_for(d, artifact.device_id), &r, &out), artifact.device_id);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - started).count();
    return RunOutput{out.price, out.std_error, ms};
}

ReproSupport RocmHipBackend::reproducibility_capabilities(const DeviceInfo& /*device*/) const {
    // the same request on the same device gives the same bits (fixed-order
    // reductions, counter-based draws), also batched and under any tuning
    return ReproSupport{true, true, true, true};
}

GpuChunkingPlan plan_gpu_chunking(std::size_t total_paths, std::optional<std::size_t> device_memory_mb,
                                  const GpuChunkingConfig& config) {
    total_paths = std::max<std::size_t>(total_paths, 1);
    const std::size_t bytes_per_path = std::max<std::size_t>(config.bytes_per_path, 1);
    const double utilization = std::clamp(config.target_utilization, 0.10, 0.95);
    const std::size_t budget_mb = std::max<std::size_t>(device_memory_mb.value_or(config.fallback_budget_mb), 1);
    const std::size_t budget_bytes =
        static_cast<std::size_t>(std::floor(static_cast<double>(budget_mb) * 1024.0 * 1024.0 * utilization));
    const std::size_t max_paths_fit = std::max<std::size_t>(budget_bytes / bytes_per_path, 1);
    const std::size_t paths_per_chunk = std::min(std::max(max_paths_fit, config.minimum_paths_per_chunk), total_paths);
    const std::size_t chunk_count = (total_paths + paths_per_chunk - 1) / paths_per_chunk;
    // saturating product, then a ceiling division by 1 MiB
    const std::size_t bytes = paths_per_chunk > SIZE_MAX / bytes_per_path ? SIZE_MAX : paths_per_chunk * bytes_per_path;
    const std::size_t mib = std::size_t(1024) * 1024;
    const std::size_t peak = std::max<std::size_t>(bytes / mib + (bytes % mib ? 1 : 0), 1);
    return GpuChunkingPlan{total_paths, paths_per_chunk, chunk_count, budget_mb, peak};
}

std::size_t estimate_gpu_bytes_per_path(const ExecutionPlan& plan) {
    // state, payoff and RNG state, plus four bytes per step up to 256 steps
    return 8 + 8 + 16 + std::min<std::size_t>(plan.n_steps, 256) * 4;
}

std::uint64_t stable_fallback_seed(std::uint64_t base_seed, std::uint64_t chunk_index) {
    return splitmix64(base_seed + chunk_index * 0x9E3779B97F4A7C15ull);
}

}  // namespace montepath
