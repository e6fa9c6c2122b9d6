// Drives montepath::RocmHipBackend and prints one JSON object per line.
//   backend_selftest cpu   host-only checks (no device needed)
//   backend_selftest gpu   discovery, compile and execution on device rocm:0
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "rocm_backend.hpp"

using namespace montepath;

namespace {

std::string q(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o + "\"";
}

std::string list(const std::vector<std::string>& v) {
    std::string o = "[";
    for (std::size_t i = 0; i < v.size(); ++i) o += (i ? "," : "") + q(v[i]);
    return o + "]";
}

const char* kind_name(BackendError::Kind k) {
    switch (k) {
        case BackendError::Kind::UnknownDevice: return "UnknownDevice";
        case BackendError::Kind::IncompatibleExecutionInput: return "IncompatibleExecutionInput";
        case BackendError::Kind::UnsupportedFeature: return "UnsupportedFeature";
    }
    return "?";
}

template <class F>
std::string error_of(F&& f) {
    try {
        f();
    } catch (const BackendError& e) {
        return kind_name(e.kind());
    }
    return "none";
}

int cpu_checks() {
    RocmHipBackend b;
    const BackendInfo info = b.describe_backend();
    std::printf("{\"check\":\"describe\",\"backend_id\":%s,\"display_name\":%s,\"platform\":%s,\"precisions\":%s,"
                "\"rngs\":%s,\"sampling\":%s,\"reductions\":%s}\n",
                q(backend_id_name(info.backend_id)).c_str(), q(info.display_name).c_str(), q(info.platform).c_str(),
                list(info.supported_precisions).c_str(), list(info.supported_rngs).c_str(),
                list(info.supported_sampling_modes).c_str(), list(info.supported_reduction_ops).c_str());
    struct Case {
        std::size_t total;
        long long mem;  // < 0: no device memory known
        std::size_t bpp;
        double util;
        std::size_t min_paths, fallback;
    };
    const Case cases[] = {{1000000, 8192, 288, 0.75, 32768, 8192},     {1000000000, 1024, 1056, 0.5, 32768, 8192},
                          {0, 8192, 288, 0.75, 32768, 8192},           {5000, -1, 0, 0.01, 1, 16},
                          {50000000, 1, 1056, 0.99, 32768, 8192},      {123456789, -1, 100, 0.75, 1000, 64}};
    for (const Case& c : cases) {
        const std::optional<std::size_t> mem =
            c.mem < 0 ? std::nullopt : std::optional<std::size_t>(static_cast<std::size_t>(c.mem));
        const GpuChunkingPlan p = plan_gpu_chunking(c.total, mem, GpuChunkingConfig{c.bpp, c.util, c.min_paths, c.fallback});
        std::printf("{\"check\":\"chunking\",\"total\":%zu,\"mem\":%lld,\"bpp\":%zu,\"util\":%.17g,\"min\":%zu,"
                    "\"fallback\":%zu,\"total_paths\":%zu,\"paths_per_chunk\":%zu,\"chunk_count\":%zu,"
                    "\"budget_mb\":%zu,\"peak_mb\":%zu}\n",
                    c.total, c.mem, c.bpp, c.util, c.min_paths, c.fallback, p.total_paths, p.paths_per_chunk,
                    p.chunk_count, p.target_budget_mb, p.estimated_peak_memory_mb);
    }
    for (std::size_t s : {0, 1, 64, 256, 1000}) {
        ExecutionPlan plan;
        plan.n_steps = s;
        std::printf("{\"check\":\"bytes_per_path\",\"n_steps\":%zu,\"bytes\":%zu}\n", s, estimate_gpu_bytes_per_path(plan));
    }
    const std::uint64_t seeds[][2] = {{0, 0}, {42, 0}, {42, 1}, {42, 7}, {0xFFFFFFFFFFFFFFFFull, 3}};
    for (const auto& s : seeds)
        std::printf("{\"check\":\"seed\",\"base\":%llu,\"index\":%llu,\"seed\":%llu}\n", (unsigned long long)s[0],
                    (unsigned long long)s[1], (unsigned long long)stable_fallback_seed(s[0], s[1]));
    // error behaviour that needs no device
    ExecutionPlan plan;
    plan.n_paths = 1000;
    plan.n_steps = 16;
    const DeviceInfo cuda{"cuda:0", BackendId::NvidiaCuda, "x", "nvidia", 1024, std::nullopt, true, false, 1024};
    const DeviceInfo bad{"rocm:x", BackendId::RocmHip, "x", "amd", 1024, std::nullopt, true, false, 1024};
    const SupportReport r = b.supports(plan, cuda);
    const CompiledArtifact cpu_art{"cpu-native:1000:16:0", BackendId::CpuNative, "cpu:host", 1000, 16,
                                   PlannerMode::Balanced, ArtifactExecutionMode::CpuNative};
    EuropeanCallConfig e;
    e.n_paths = 1000;
    e.n_steps = 16;
    std::printf("{\"check\":\"errors\",\"supports_foreign\":%s,\"supports_foreign_features\":%s,"
                "\"compile_foreign\":%s,\"compile_malformed\":%s,\"execute_foreign_artifact\":%s,"
                "\"precision_16\":%s}\n",
                q(support_level_name(r.support_level)).c_str(), list(r.unsupported_features).c_str(),
                q(error_of([&] { b.compile(plan, cuda); })).c_str(), q(error_of([&] { b.compile(plan, bad); })).c_str(),
                q(error_of([&] { b.execute(cpu_art, BackendExecutionInput{e}); })).c_str(),
                q(error_of([&] { RocmHipBackend x(16); })).c_str());
    return 0;
}

int gpu_checks() {
    RocmHipBackend b32(32), b64(64);
    const std::vector<DeviceInfo> devs = b32.discover_devices();
    for (const DeviceInfo& d : devs)
        std::printf("{\"check\":\"device\",\"device_id\":%s,\"backend_id\":%s,\"name\":%s,\"vendor\":%s,"
                    "\"memory_total_mb\":%zu,\"supports_float64\":%s,\"max_threads_hint\":%zu}\n",
                    q(d.device_id).c_str(), q(backend_id_name(d.backend_id)).c_str(), q(d.name).c_str(),
                    q(d.vendor).c_str(), d.memory_total_mb.value_or(0), d.supports_float64 ? "true" : "false",
                    d.max_threads_hint);
    if (devs.empty()) {
        std::printf("{\"check\":\"no_device\"}\n");
        return 1;
    }
    const DeviceInfo& dev = devs[0];
    ExecutionPlan plan;
    plan.n_paths = 100000;
    plan.n_steps = 64;
    plan.features.step_count = 64;
    const SupportReport s = b32.supports(plan, dev);
    const CostEstimate c = b32.estimate_cost(plan, dev);
    const CompiledArtifact art = b32.compile(plan, dev);
    std::printf("{\"check\":\"plan\",\"support\":%s,\"runtime_ms\":%.17g,\"peak_mb\":%.17g,\"confidence\":%s,"
                "\"artifact_id\":%s,\"artifact_device\":%s}\n",
                q(support_level_name(s.support_level)).c_str(), c.estimated_runtime_ms, c.estimated_peak_memory_mb,
                q(c.confidence).c_str(), q(art.artifact_id).c_str(), q(art.device_id).c_str());
    EuropeanCallConfig e;
    e.n_paths = 100000;
    e.n_steps = 64;
    ArithmeticAsianCallConfig a;
    a.n_paths = 100000;
    a.n_steps = 64;
    DownAndOutCallConfig d;
    d.n_paths = 100000;
    d.n_steps = 64;
    HestonEuropeanCallConfig h;
    h.n_paths = 100000;
    h.n_steps = 64;
    MertonJumpDiffusionCallConfig m;
    m.n_paths = 100000;
    LookbackCallConfig l;
    l.n_paths = 100000;
    l.n_steps = 64;
    BasketCallConfig bk;
    bk.n_paths = 100000;
    AmericanPutConfig am;
    am.n_paths = 100000;
    BermudanPutConfig be;
    be.n_paths = 100000;
    const std::pair<const char*, BackendExecutionInput> inputs[] = {
        {"european_call", e},   {"arithmetic_asian_call", a}, {"down_and_out_call", d},
        {"heston_european_call", h}, {"merton_jump_diffusion_call", m}, {"lookback_call", l},
        {"basket_call", bk},    {"american_put", am},         {"bermudan_put", be}};
    for (const auto& kv : inputs) {
        // an artifact of the input's own shape (step-less inputs match any)
        ExecutionPlan p2 = plan;
        p2.n_steps = 64;
        for (const RocmHipBackend* b : {&b32, &b64}) {
            const CompiledArtifact art2 = b->compile(p2, dev);
            const RunOutput o = b->execute(art2, kv.second);
            const RunOutput o2 = b->execute(art2, kv.second);
            std::printf("{\"check\":\"execute\",\"workload\":%s,\"precision\":%d,\"price\":%.17g,\"std_error\":%.17g,"
                        "\"runtime_ms\":%.6g,\"repeat_identical\":%s}\n",
                        q(kv.first).c_str(), b == &b32 ? 32 : 64, o.price, o.std_error, o.runtime_ms,
                        (o.price == o2.price && o.std_error == o2.std_error) ? "true" : "false");
        }
    }
    // one context per thread: two threads executing at once each open their
    // own and give the single-thread result
    {
        const RunOutput ref = b32.execute(art, BackendExecutionInput{e});
        RunOutput got[2] = {};
        std::string err[2];
        std::vector<std::thread> th;
        for (int i = 0; i < 2; ++i)
            th.emplace_back([&, i] {
                try {
                    got[i] = b32.execute(art, BackendExecutionInput{e});
                } catch (const BackendError& x) {
                    err[i] = x.what();
                }
            });
        for (auto& t : th) t.join();
        const bool same = err[0].empty() && err[1].empty() && got[0].price == ref.price &&
                          got[1].price == ref.price && got[0].std_error == ref.std_error &&
                          got[1].std_error == ref.std_error;
        std::printf("{\"check\":\"threads\",\"identical\":%s,\"errors\":%s}\n", same ? "true" : "false",
                    list({err[0], err[1]}).c_str());
    }
    EuropeanCallConfig other = e;
    other.n_paths = 99999;
    std::printf("{\"check\":\"execute_errors\",\"shape_mismatch\":%s}\n",
                q(error_of([&] { b32.execute(art, BackendExecutionInput{other}); })).c_str());
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 2 && std::strcmp(argv[1], "cpu") == 0) return cpu_checks();
    if (argc == 2 && std::strcmp(argv[1], "gpu") == 0) {
        try {
            return gpu_checks();
        } catch (const BackendError& e) {
            std::printf("{\"check\":\"error\",\"kind\":\"%s\",\"message\":%s}\n", kind_name(e.kind()), q(e.what()).c_str());
            return 1;
        }
    }
    std::fprintf(stderr, "usage: backend_selftest cpu|gpu\n");
    return 2;
}
