// C++ host layer above the C-ABI: the reference's backend interface
// (crates/mc-core/src/backend/mod.rs) with a RocmHipBackend that prices
// through libmontepath_hip.  Names, fields and error kinds follow the
// reference's Rust types so that a caller ports line for line; types the
// reference keeps in serde structs are plain structs here.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <optional>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

namespace montepath {

// BackendId (planner/mod.rs:22-28), with the rocm_hip backend added.
enum class BackendId { CpuNative, NvidiaCuda, AppleMetal, RocmHip };
const char* backend_id_name(BackendId id);  // serde names: "cpu_native", ..., "rocm_hip"

// PlannerMode (planner/mod.rs:7-14) and SupportLevel (planner/mod.rs:77-83).
enum class PlannerMode { Safe, Balanced, Aggressive, Explain };
enum class SupportLevel { Supported, SupportedWithFallbacks, Unsupported };
const char* support_level_name(SupportLevel level);

// SamplingMethod (runtime/cpu.rs:5796-5830) and MonteCarloTechnique
// (runtime/cpu.rs:24-28).
enum class SamplingMethod { Pseudorandom, RandomizedHalton, LatinHypercube, ScrambledSobol, ScrambledSobolBrownianBridge };
enum class MonteCarloTechnique { Standard, Antithetic, ControlVariate };

// backend/mod.rs:20-30
struct BackendInfo {
    BackendId backend_id;
    std::string display_name;
    std::string version;
    std::string platform;
    std::vector<std::string> supported_precisions;
    std::vector<std::string> supported_rngs;
    std::vector<std::string> supported_sampling_modes;
    std::vector<std::string> supported_reduction_ops;
};

// backend/mod.rs:32-43
struct DeviceInfo {
    std::string device_id;
    BackendId backend_id;
    std::string name;
    std::string vendor;
    std::optional<std::size_t> memory_total_mb;
    std::optional<std::size_t> memory_free_mb;
    bool supports_float64;
    bool supports_unified_memory;
    std::size_t max_threads_hint;
};

// backend/mod.rs:45-52
struct SupportReport {
    BackendId backend_id;
    std::string device_id;
    SupportLevel support_level;
    std::vector<std::string> unsupported_features;
    std::vector<std::string> warnings;
};

// backend/mod.rs:54-63
struct CostEstimate {
    BackendId backend_id;
    std::string device_id;
    double estimated_compile_ms;
    double estimated_runtime_ms;
    double estimated_total_ms;
    double estimated_peak_memory_mb;
    std::string confidence;
};

// backend/mod.rs:65-80
struct GpuChunkingConfig {
    std::size_t bytes_per_path;
    double target_utilization;
    std::size_t minimum_paths_per_chunk;
    std::size_t fallback_budget_mb;
};

struct GpuChunkingPlan {
    std::size_t total_paths;
    std::size_t paths_per_chunk;
    std::size_t chunk_count;
    std::size_t target_budget_mb;
    std::size_t estimated_peak_memory_mb;
};

// backend/mod.rs:82-88
struct ReproSupport {
    bool supports_same_backend_exact;
    bool supports_same_backend_deterministic;
    bool supports_cross_backend_statistical;
    bool supports_stable_chunking;
};

// FeatureSummary and ExecutionPlan (planner/mod.rs:118-148); the decision
// report is the planner's and is not consulted by a backend.
struct FeatureSummary {
    std::size_t random_variable_count = 0;
    std::size_t state_variable_count = 0;
    std::size_t step_count = 0;
    std::size_t observation_count = 0;
    std::size_t reduction_count = 0;
    std::size_t conditional_expression_count = 0;
};

struct ExecutionPlan {
    BackendId backend = BackendId::RocmHip;
    PlannerMode planner_mode = PlannerMode::Balanced;
    std::size_t n_paths = 0;
    std::size_t n_steps = 0;
    FeatureSummary features;
};

// ArtifactExecutionMode (backend/mod.rs:139-144), with native GPU execution.
enum class ArtifactExecutionMode { CpuNative, GpuFallback, GpuNative };

// backend/mod.rs:160-170 (the native-artifact manifest is the CUDA and
// Metal backends' staging record; this backend's kernels are built ahead)
struct CompiledArtifact {
    std::string artifact_id;
    BackendId backend_id;
    std::string device_id;
    std::size_t n_paths;
    std::size_t n_steps;
    PlannerMode planner_mode;
    ArtifactExecutionMode execution_mode;
};

// RunOutput (backend/mod.rs:172-177); the reference's `stderr` field is
// `std_error` here (stderr is a C macro).
struct RunOutput {
    double price;
    double std_error;
    double runtime_ms;
};

// EuropeanCallConfig, ArithmeticAsianCallConfig, DownAndOutCallConfig
// (runtime/cpu.rs:1296-1327, 1502-1533, 1644-1677), reference defaults.
struct EuropeanCallConfig {
    double s0 = 100.0, k = 100.0, r = 0.03, sigma = 0.2, t = 1.0;
    std::size_t n_paths = 100000, n_steps = 252;
    std::uint64_t seed = 42;
    std::size_t n_threads = 0;
    MonteCarloTechnique technique = MonteCarloTechnique::Standard;
    SamplingMethod sampling = SamplingMethod::Pseudorandom;
};

struct ArithmeticAsianCallConfig {
    double s0 = 100.0, k = 100.0, r = 0.03, sigma = 0.2, t = 1.0;
    std::size_t n_paths = 100000, n_steps = 252;
    std::uint64_t seed = 42;
    std::size_t n_threads = 0;
    MonteCarloTechnique technique = MonteCarloTechnique::Standard;
    SamplingMethod sampling = SamplingMethod::Pseudorandom;
};

struct DownAndOutCallConfig {
    double s0 = 100.0, k = 100.0, barrier = 80.0, r = 0.03, sigma = 0.2, t = 1.0;
    std::size_t n_paths = 100000, n_steps = 252;
    std::uint64_t seed = 42;
    std::size_t n_threads = 0;
    MonteCarloTechnique technique = MonteCarloTechnique::Standard;
    SamplingMethod sampling = SamplingMethod::Pseudorandom;
};

// The families the reference prices on the CPU only, as additive execution
// inputs: HestonEuropeanCallConfig, MertonJumpDiffusionCallConfig,
// LookbackCallConfig, BasketCallConfig, AmericanPutConfig, BermudanPutConfig
// (runtime/cpu.rs:1430-1818), reference defaults.  The reference's Heston
// CPU path floors the variance (the engine's reference_floor scheme); its
// Merton and basket configs have no step count (one step to maturity).
struct HestonEuropeanCallConfig {
    double s0 = 100.0, k = 100.0, r = 0.03, v0 = 0.04, kappa = 1.5, theta = 0.04, vol_of_vol = 0.3, rho = -0.6,
           t = 1.0;
    std::size_t n_paths = 100000, n_steps = 252;
    std::uint64_t seed = 42;
    std::size_t n_threads = 0;
    MonteCarloTechnique technique = MonteCarloTechnique::Standard;
};

struct MertonJumpDiffusionCallConfig {
    double s0 = 100.0, k = 100.0, r = 0.03, sigma = 0.2, jump_intensity = 0.4, jump_mean = -0.08,
           jump_volatility = 0.25, t = 1.0;
    std::size_t n_paths = 100000;
    std::uint64_t seed = 42;
    std::size_t n_threads = 0;
};

struct LookbackCallConfig {
    double s0 = 100.0, k = 100.0, r = 0.03, sigma = 0.2, t = 1.0;
    std::size_t n_paths = 100000, n_steps = 252;
    std::uint64_t seed = 42;
    std::size_t n_threads = 0;
    MonteCarloTechnique technique = MonteCarloTechnique::Standard;
    SamplingMethod sampling = SamplingMethod::Pseudorandom;
};

struct BasketCallConfig {
    double s01 = 100.0, s02 = 95.0, k = 100.0, r = 0.03, sigma1 = 0.2, sigma2 = 0.25, rho = 0.35, weight1 = 0.5,
           weight2 = 0.5, t = 1.0;
    std::size_t n_paths = 100000;
    std::uint64_t seed = 42;
    std::size_t n_threads = 0;
    MonteCarloTechnique technique = MonteCarloTechnique::Standard;
    SamplingMethod sampling = SamplingMethod::Pseudorandom;
};

struct AmericanPutConfig {
    double s0 = 100.0, k = 100.0, r = 0.03, sigma = 0.2, t = 1.0;
    std::size_t n_paths = 100000, n_steps = 64;
    std::uint64_t seed = 42;
    std::size_t n_threads = 0;
    std::size_t basis_degree = 2;
};

struct BermudanPutConfig {
    double s0 = 100.0, k = 100.0, r = 0.03, sigma = 0.2, t = 1.0;
    std::size_t n_paths = 100000, n_steps = 64;
    std::uint64_t seed = 42;
    std::size_t n_threads = 0;
    std::size_t basis_degree = 2;
    std::vector<std::size_t> exercise_steps{16, 32, 48, 64};
};

// BackendExecutionInput (backend/mod.rs:179-184), the reference's three
// variants first, then the additive ones
using BackendExecutionInput =
    std::variant<EuropeanCallConfig, ArithmeticAsianCallConfig, DownAndOutCallConfig, HestonEuropeanCallConfig,
                 MertonJumpDiffusionCallConfig, LookbackCallConfig, BasketCallConfig, AmericanPutConfig,
                 BermudanPutConfig>;

// BackendError (backend/mod.rs:186-194): the three variants and their messages.
class BackendError : public std::runtime_error {
  public:
    enum class Kind { UnknownDevice, IncompatibleExecutionInput, UnsupportedFeature };
    static BackendError unknown_device(const std::string& device_id);
    static BackendError incompatible_execution_input();
    static BackendError unsupported_feature(const std::string& what);
    Kind kind() const { return kind_; }

  private:
    BackendError(Kind kind, const std::string& message) : std::runtime_error(message), kind_(kind) {}
    Kind kind_;
};

// RuntimeBackend (backend/mod.rs:196-213)
class RuntimeBackend {
  public:
    virtual ~RuntimeBackend() = default;
    virtual BackendId backend_id() const = 0;
    virtual BackendInfo describe_backend() const = 0;
    virtual std::vector<DeviceInfo> discover_devices() const = 0;
    virtual SupportReport supports(const ExecutionPlan& plan, const DeviceInfo& device) const = 0;
    virtual CostEstimate estimate_cost(const ExecutionPlan& plan, const DeviceInfo& device) const = 0;
    virtual CompiledArtifact compile(const ExecutionPlan& plan, const DeviceInfo& device) const = 0;
    virtual RunOutput execute(const CompiledArtifact& artifact, const BackendExecutionInput& input) const = 0;
    virtual ReproSupport reproducibility_capabilities(const DeviceInfo& device) const = 0;
};

// The rocm_hip backend: devices "rocm:<n>", native execution of every
// execution input through mp_hip_price (FP32 paths with FP64 sums, or FP64
// throughout), one engine context per thread and device (contexts are
// single-thread-affine, as the Metal backend's runtime cache is
// thread-local, backend/metal.rs:52-55).
class RocmHipBackend final : public RuntimeBackend {
  public:
    explicit RocmHipBackend(int precision = 32);
    BackendId backend_id() const override { return BackendId::RocmHip; }
    BackendInfo describe_backend() const override;
    std::vector<DeviceInfo> discover_devices() const override;
    SupportReport supports(const ExecutionPlan& plan, const DeviceInfo& device) const override;
    CostEstimate estimate_cost(const ExecutionPlan& plan, const DeviceInfo& device) const override;
    CompiledArtifact compile(const ExecutionPlan& plan, const DeviceInfo& device) const override;
    RunOutput execute(const CompiledArtifact& artifact, const BackendExecutionInput& input) const override;
    ReproSupport reproducibility_capabilities(const DeviceInfo& device) const override;

  private:
    int device_index(const std::string& device_id) const;  // -1 when not a rocm device
    void validate_device(const DeviceInfo& device) const;
    int precision_;
};

// plan_gpu_chunking and estimate_gpu_bytes_per_path (backend/mod.rs:372-405)
GpuChunkingPlan plan_gpu_chunking(std::size_t total_paths, std::optional<std::size_t> device_memory_mb,
                                  const GpuChunkingConfig& config);
std::size_t estimate_gpu_bytes_per_path(const ExecutionPlan& plan);
// stable_fallback_seed (backend/mod.rs:713-723): splitmix64 of
// base + index * 0x9E3779B97F4A7C15
std::uint64_t stable_fallback_seed(std::uint64_t base_seed, std::uint64_t chunk_index);

}  // namespace montepath
