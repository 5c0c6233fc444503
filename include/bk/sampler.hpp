// bk/sampler.hpp -- C++17 drop-in for the reference's src/sampling/sampler.cuh
// interface, implemented over the C ABI of libbk.so (bk_capi.h).
//
// Same type and member names, defaults and meaning as the reference:
//   SamplerConfig ......... sampler.cuh:235-260
//   SeedSamplingStats ..... sampler.cuh:262-270
//   SamplingResult ........ sampler.cuh:272-297
//   ISampler .............. sampler.cuh:303-310
//   create_sampler(n) ..... sampler.cuh:737-743
// Errors from the library become std::runtime_error, as the reference's
// CHECK_CUDA / CHECK_CUSTATEVEC macros do (sampler.cuh:31-47).
//
// Differences a caller can observe:
//   * create_sampler(n) serves n <= 32 (the reference: n <= 30, above that a
//     stub that throws); n > 32 still returns a sampler whose sample() throws.
//   * unique_basis is returned sorted ascending (the reference returns the
//     iteration order of a std::unordered_set, which is unspecified).
//   * SamplerConfig gains two fields, precision and order, defaulting to the
//     reference's behaviour (complex128, first-order phase-then-mixer).
#pragma once

#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../bk_capi.h"

// ---- SpinState (labs_energy.hpp:8-84), the part the sampler interface uses ----
#ifndef BK_HAVE_REFERENCE_SPINSTATE
struct SpinState {
    uint64_t bits = 0;
    constexpr SpinState() = default;
    constexpr explicit SpinState(uint64_t b) : bits(b) {}
    int hamming_distance(const SpinState& o) const { return __builtin_popcountll(bits ^ o.bits); }
    bool operator==(const SpinState& o) const { return bits == o.bits; }
    bool operator!=(const SpinState& o) const { return bits != o.bits; }
    bool operator<(const SpinState& o) const { return bits < o.bits; }
};
#endif

struct SamplerConfig {
    int n = 16;
    double g = 0.5;
    double dt = 0.05;
    double t_max = 60.0;
    int num_sample_points = 20;
    int shots_per_point = 100;
    uint64_t seed = 42;
    int num_sampling_seeds = 3;
    uint64_t seed_stride = 10007;
    double t_start = 0.0;
    double dt_sample = 0.0;
    bool track_p_less = true;
    bool verbose = true;
    bool record_trajectory = false;
    double dt_trajectory = 0.0;
    int precision = BK_C128;  // engine extension: BK_C64 | BK_C128
    int order = BK_LIE;       // engine extension: BK_LIE | BK_STRANG
};

struct SeedSamplingStats {
    uint64_t seed = 0;
    size_t total_shots_collected = 0;
    size_t improved_shots = 0;
    double success_rate = 0.0;
    double per_eval_rate = 0.0;
    double mean_hamming_flips = 0.0;
    int64_t best_energy = std::numeric_limits<int64_t>::max();
};

struct SamplingResult {
    std::vector<SpinState> unique_basis;
    size_t total_shots_collected = 0;
    size_t improved_shots = 0;
    double success_rate = 0.0;
    double per_eval_rate = 0.0;
    size_t total_evals = 0;
    double mean_hamming_flips = 0.0;
    std::vector<SeedSamplingStats> seed_stats;
    std::vector<double> time_points;
    std::vector<double> p_less_history;
    std::vector<double> dh_history;
    double window_avg_dh = 0.0;
    double initial_rise_rate_avg = 0.0;
    double initial_rise_rate_max = 0.0;
    int d_min = 0;
    int64_t start_energy = 0;
    double p_blind = 0.0;
    double peak_p_less = 0.0;
    double peak_time = 0.0;
};

class ISampler {
public:
    virtual ~ISampler() = default;
    virtual SamplingResult sample(const SpinState& start_state, const SamplerConfig& config) = 0;
};

namespace bk {

inline void check(int rc, const char* what)
{
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + bk_last_error());
}

// Owns one libbk context (device 0 by default).
class Context {
public:
    explicit Context(int device = 0) { check(bk_ctx_create(device, &ctx_), "bk_ctx_create"); }
    ~Context()
    {
        if (ctx_) bk_ctx_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    bk_ctx* get() const { return ctx_; }

private:
    bk_ctx* ctx_ = nullptr;
};

// A library-owned bk_sampling_result as the reference's SamplingResult.
inline SamplingResult to_sampling_result(const bk_sampling_result& r)
{
    SamplingResult out;
    out.unique_basis.reserve(static_cast<size_t>(r.unique_basis_len));
    for (int64_t i = 0; i < r.unique_basis_len; ++i) out.unique_basis.emplace_back(r.unique_basis[i]);
    out.total_shots_collected = r.total_shots_collected;
    out.improved_shots = r.improved_shots;
    out.success_rate = r.success_rate;
    out.per_eval_rate = r.per_eval_rate;
    out.total_evals = r.total_evals;
    out.mean_hamming_flips = r.mean_hamming_flips;
    for (int64_t i = 0; i < r.seed_stats_len; ++i) {
        const bk_seed_stats& s = r.seed_stats[i];
        SeedSamplingStats o;
        o.seed = s.seed;
        o.total_shots_collected = s.total_shots_collected;
        o.improved_shots = s.improved_shots;
        o.success_rate = s.success_rate;
        o.per_eval_rate = s.per_eval_rate;
        o.mean_hamming_flips = s.mean_hamming_flips;
        o.best_energy = s.best_energy;
        out.seed_stats.push_back(o);
    }
    out.time_points.assign(r.time_points, r.time_points + r.history_len);
    out.p_less_history.assign(r.p_less_history, r.p_less_history + r.history_len);
    out.dh_history.assign(r.dh_history, r.dh_history + r.history_len);
    out.window_avg_dh = r.window_avg_dh;
    out.initial_rise_rate_avg = r.initial_rise_rate_avg;
    out.initial_rise_rate_max = r.initial_rise_rate_max;
    out.d_min = r.d_min;
    out.start_energy = r.start_energy;
    out.p_blind = r.p_blind;
    out.peak_p_less = r.peak_p_less;
    out.peak_time = r.peak_time;
    return out;
}

// The GPU backend behind create_sampler (the role of CuStateVecSampler).
class HipSampler : public ISampler {
public:
    explicit HipSampler(int device = 0) : ctx_(device) {}

    SamplingResult sample(const SpinState& start_state, const SamplerConfig& c) override
    {
        bk_sampler_config cfg;
        bk_sampler_config_defaults(&cfg);
        cfg.n = c.n;
        cfg.g = c.g;
        cfg.dt = c.dt;
        cfg.t_max = c.t_max;
        cfg.num_sample_points = c.num_sample_points;
        cfg.shots_per_point = c.shots_per_point;
        cfg.seed = c.seed;
        cfg.num_sampling_seeds = c.num_sampling_seeds;
        cfg.seed_stride = c.seed_stride;
        cfg.t_start = c.t_start;
        cfg.dt_sample = c.dt_sample;
        cfg.track_p_less = c.track_p_less ? 1 : 0;
        cfg.verbose = c.verbose ? 1 : 0;
        cfg.record_trajectory = c.record_trajectory ? 1 : 0;
        cfg.dt_trajectory = c.dt_trajectory;
        cfg.precision = c.precision;
        cfg.order = c.order;
        bk_sampling_result* r = nullptr;
        check(bk_sampler_run(ctx_.get(), &cfg, start_state.bits, &r), "bk_sampler_run");
        std::unique_ptr<bk_sampling_result, void (*)(bk_sampling_result*)> guard(r, bk_result_free);
        return to_sampling_result(*r);
    }

    bk_ctx* context() const { return ctx_.get(); }

private:
    Context ctx_;
};

class TooLargeSampler : public ISampler {
public:
    SamplingResult sample(const SpinState&, const SamplerConfig&) override
    {
        throw std::runtime_error("statevector backend supports N <= 32");
    }
};

}  // namespace bk

inline std::unique_ptr<ISampler> create_sampler(int n)
{
    if (n <= BK_MAX_N) return std::make_unique<bk::HipSampler>();
    return std::make_unique<bk::TooLargeSampler>();
}
