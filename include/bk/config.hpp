// bk/config.hpp -- C++17 drop-in for the reference's benchmark configuration
// (src/config/config.hpp): the same sections, fields, defaults, keys and
// aliases, so the reference's .cfg / .ini files drive this engine unchanged.
//
//   SystemConfig / TrapsConfig / QuantumConfig / ClassicalConfig / OutputConfig
//   / BenchmarkConfig ........................... config.hpp:82-237
//   BenchmarkConfig::load_from_file / _string ... config.hpp:237-275
//   apply_key_value (every key and alias) ....... config.hpp:277-389
//   QuantumConfig::t_max / sample_times /
//     trotter_steps_per_sample .................. config.hpp:119-156
//   OutputConfig path templates ({N}, {n}) ...... config.hpp:180-224
//   to_ini_string ............................... config.hpp:452-500
// plus to_sampler_config(), the mapping BenchmarkConnector::run_quantum_simulation
// applies (connector.cuh:628-643).
//
// Parsing rules (as the reference): '#' or ';' outside quotes starts a comment;
// [section] headers are accepted and ignored (keys are global); key = value or
// key : value, keys case-insensitive; a malformed line or unknown key warns on
// std::cerr and is skipped; booleans are 1/true/yes/on or 0/false/no/off (else
// std::invalid_argument); numbers go through std::stoul / stoull / stoi / stod
// (which throw on garbage); list items that do not parse are dropped.  Setting
// kick_trials_per_d also sets restart_trials; "g" sets a single field; verbose
// sets both the output and the quantum flag.
#pragma once

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <functional>
#include <initializer_list>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "sampler.hpp"

namespace bk {
namespace config {

namespace text {
inline std::string trim(const std::string& s)
{
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}
inline std::string lower(std::string s)
{
    for (char& ch : s) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    return s;
}
// the line up to its first '#' or ';' outside quotes (either quote toggles)
inline std::string uncomment(const std::string& line)
{
    bool quoted = false;
    for (size_t i = 0; i < line.size(); ++i) {
        const char ch = line[i];
        if (ch == '"' || ch == '\'') quoted = !quoted;
        else if (!quoted && (ch == '#' || ch == ';')) return line.substr(0, i);
    }
    return line;
}
inline bool boolean(const std::string& v)
{
    const std::string s = lower(trim(v));
    for (const char* t : {"1", "true", "yes", "on"})
        if (s == t) return true;
    for (const char* f : {"0", "false", "no", "off"})
        if (s == f) return false;
    throw std::invalid_argument("Cannot parse boolean value: '" + v + "'");
}
template <typename T>
std::vector<T> list(const std::string& v)
{
    std::vector<T> out;
    std::stringstream all(v);
    std::string item;
    while (std::getline(all, item, ',')) {
        item = trim(item);
        if (item.empty()) continue;
        std::stringstream one(item);
        T x;
        if (one >> x) out.push_back(x);
    }
    return out;
}
template <typename T>
std::string join(const std::vector<T>& v)
{
    std::ostringstream os;
    for (size_t i = 0; i < v.size(); ++i) os << (i ? ", " : "") << v[i];
    return os.str();
}
}  // namespace text

struct SystemConfig {
    std::vector<int> n_values = {16, 20, 24};
    uint64_t seed_base = 42;
    std::string device = "auto";
};

struct TrapsConfig {
    size_t pool_minima_count = 100;
    size_t benchmark_trap_count = 5;
    size_t max_energy_ranks = 2;
    std::string minima_cache_dir = "seeds";
    bool force_regenerate_traps = false;
};

struct QuantumConfig {
    std::vector<double> g_values = {0.25, 0.5, 0.75};
    double dt = 0.05;
    double t_start = 10.0;
    double dt_sample = 1.5;
    int k_sample_points = 20;
    int m_shots_per_point = 100;
    int num_sampling_seeds = 3;
    uint64_t seed_stride = 10007;
    bool track_p_less = true;
    bool verbose = true;
    bool record_trajectory = false;
    double dt_trajectory = 0.0;
    double t_max_val = 0.0;
    bool run_classical = true;

    // t_max_val when set, else the last sampling time t_start + (K - 1) * step
    double t_max() const
    {
        if (t_max_val > 0.0) return t_max_val;
        if (k_sample_points <= 1) return t_start;
        return t_start + (k_sample_points - 1) * (dt_sample > 0.0 ? dt_sample : dt);
    }
    // trajectory: k * step for k = 0 .. round(t_max / step) (step = dt_trajectory,
    // else dt, else 0.05); sampling: t_start + k * (dt_sample or dt), k < K
    std::vector<double> sample_times() const
    {
        std::vector<double> t;
        if (record_trajectory) {
            double step = dt_trajectory > 0.0 ? dt_trajectory : dt;
            if (step <= 0.0) step = 0.05;
            const int count = static_cast<int>(std::round(t_max() / step)) + 1;
            for (int k = 0; k < count; ++k) t.push_back(k * step);
            return t;
        }
        const double step = dt_sample > 0.0 ? dt_sample : dt;
        for (int k = 0; k < k_sample_points; ++k) t.push_back(t_start + k * step);
        return t;
    }
    size_t trotter_steps_per_sample() const
    {
        if (dt <= 0.0) return 1;
        const double step = dt_sample > 0.0 ? dt_sample : dt;
        return std::max<size_t>(1, static_cast<size_t>(std::round(step / dt)));
    }
};

struct ClassicalConfig {
    size_t kick_trials_per_d = 50000;
    std::vector<int> kick_d_values = {1, 2, 3, 4, 5, 6};
    size_t restart_trials = 50000;
    size_t max_evals_per_start = 0;
    size_t batch_size = 10000;
    bool enable_tabu = true;
    size_t tabu_trials = 50000;
    size_t tabu_max_steps = 100;
    std::vector<int> tabu_kick_d_values = {2, 4, 6, 8, 10, 12};
    int tabu_tenure = 0;
};

struct OutputConfig {
    std::string output_dir = "data_raw";
    std::string results_csv = "benchmark_results_n{N}.csv";
    std::string time_series_csv = "p_less_dynamics_n{N}.csv";
    std::string summary_csv = "benchmark_summary.csv";
    std::string structured_json = "structured_benchmark_n{N}.json";
    std::string all_structured_json = "all_seeds_benchmark.json";
    bool verbose = true;

    // {N} / {n} -> n; a bare relative file name goes under output_dir
    std::string resolve_path(const std::string& pattern, int n = 0) const
    {
        std::string s = pattern;
        const std::string ns = std::to_string(n);
        for (const char* tag : {"{N}", "{n}"})
            for (size_t at = s.find(tag); at != std::string::npos; at = s.find(tag)) s.replace(at, 3, ns);
        const std::filesystem::path p(s);
        if (p.is_relative() && !output_dir.empty() && p.parent_path().empty())
            return (std::filesystem::path(output_dir) / p).string();
        return p.string();
    }
    std::string get_results_path(int n) const { return resolve_path(results_csv, n); }
    std::string get_time_series_path(int n) const { return resolve_path(time_series_csv, n); }
    std::string get_summary_path() const { return resolve_path(summary_csv, 0); }
    std::string get_structured_json_path(int n) const { return resolve_path(structured_json, n); }
    std::string get_all_structured_json_path() const { return resolve_path(all_structured_json, 0); }
};

struct BenchmarkConfig {
    SystemConfig system;
    TrapsConfig traps;
    QuantumConfig quantum;
    ClassicalConfig classical;
    OutputConfig output;

    static BenchmarkConfig load_from_file(const std::string& path)
    {
        std::ifstream f(path);
        if (!f.is_open()) throw std::runtime_error("BenchmarkConfig: Cannot open file: " + path);
        std::stringstream ss;
        ss << f.rdbuf();
        return load_from_string(ss.str());
    }

    static BenchmarkConfig load_from_string(const std::string& content)
    {
        BenchmarkConfig cfg;
        std::stringstream in(content);
        std::string raw, section;
        int line_no = 0;
        while (std::getline(in, raw)) {
            ++line_no;
            const std::string line = text::trim(text::uncomment(raw));
            if (line.empty()) continue;
            if (line.front() == '[' && line.back() == ']') {
                section = text::lower(text::trim(line.substr(1, line.size() - 2)));
                continue;
            }
            size_t eq = line.find('=');
            if (eq == std::string::npos) eq = line.find(':');
            if (eq == std::string::npos) {
                std::cerr << "[Config Warning] Malformed line " << line_no << ": " << raw << "\n";
                continue;
            }
            cfg.apply_key_value(section, text::lower(text::trim(line.substr(0, eq))), text::trim(line.substr(eq + 1)),
                                line_no);
        }
        return cfg;
    }

    void apply_key_value(const std::string& section, const std::string& key, const std::string& val, int line_no)
    {
        (void)section;  // keys are global; sections only group the file
        const auto& table = setters();
        const auto it = table.find(key);
        if (it == table.end()) {
            std::cerr << "[Config Warning] Unknown parameter on line " << line_no << ": '" << key << "'\n";
            return;
        }
        it->second(*this, val);
    }

    std::string to_ini_string() const
    {
        std::ostringstream s;
        s << "# Bernasconi-Krylov Benchmark Configuration\n\n";
        s << "[system]\nn_values = " << text::join(system.n_values) << "\nseed_base = " << system.seed_base
          << "\ndevice = " << system.device << "\n\n";
        s << "[traps]\npool_minima_count = " << traps.pool_minima_count << "\nbenchmark_trap_count = "
          << traps.benchmark_trap_count << "\nmax_energy_ranks = " << traps.max_energy_ranks
          << "\nminima_cache_dir = " << traps.minima_cache_dir
          << "\nforce_regenerate_traps = " << (traps.force_regenerate_traps ? "true" : "false") << "\n\n";
        s << "[quantum]\ng_values = " << text::join(quantum.g_values) << "\ndt = " << quantum.dt
          << "\nt_start = " << quantum.t_start << "\ndt_sample = " << quantum.dt_sample
          << "\nk_sample_points = " << quantum.k_sample_points << "\nm_shots_per_point = " << quantum.m_shots_per_point
          << "\ntrack_p_less = " << (quantum.track_p_less ? "true" : "false")
          << "\nrecord_trajectory = " << (quantum.record_trajectory ? "true" : "false") << "\n";
        if (quantum.record_trajectory) s << "dt_trajectory = " << quantum.dt_trajectory << "\n";
        if (quantum.t_max_val > 0.0) s << "t_max = " << quantum.t_max_val << "\n";
        s << "run_classical = " << (quantum.run_classical ? "true" : "false") << "\n\n";
        s << "[classical]\nkick_trials_per_d = " << classical.kick_trials_per_d
          << "\nkick_d_values = " << text::join(classical.kick_d_values)
          << "\nrestart_trials = " << classical.restart_trials
          << "\nmax_evals_per_start = " << classical.max_evals_per_start << "\nbatch_size = " << classical.batch_size
          << "\n\n";
        s << "[output]\noutput_dir = " << output.output_dir << "\nresults_csv = " << output.results_csv
          << "\ntime_series_csv = " << output.time_series_csv << "\nsummary_csv = " << output.summary_csv
          << "\nverbose = " << (output.verbose ? "true" : "false") << "\n";
        return s.str();
    }

private:
    using Setter = std::function<void(BenchmarkConfig&, const std::string&)>;

    // every key the reference accepts, aliases on one line
    static const std::unordered_map<std::string, Setter>& setters()
    {
        static const std::unordered_map<std::string, Setter> table = [] {
            std::unordered_map<std::string, Setter> t;
            const auto add = [&t](std::initializer_list<const char*> keys, Setter f) {
                for (const char* k : keys) t.emplace(k, f);
            };
            using C = BenchmarkConfig;
            using S = const std::string&;
            add({"n_values", "n", "n_spins"}, [](C& c, S v) { c.system.n_values = text::list<int>(v); });
            add({"seed_base", "seed", "master_seed"}, [](C& c, S v) { c.system.seed_base = std::stoull(v); });
            add({"device", "backend"}, [](C& c, S v) { c.system.device = text::lower(v); });
            add({"pool_minima_count", "target_minima_count", "minima_pool_size"},
                [](C& c, S v) { c.traps.pool_minima_count = std::stoul(v); });
            add({"benchmark_trap_count", "top_minima_count", "num_traps"},
                [](C& c, S v) { c.traps.benchmark_trap_count = std::stoul(v); });
            add({"max_energy_ranks", "max_ranks", "energy_ranks", "num_ranks"},
                [](C& c, S v) { c.traps.max_energy_ranks = std::stoul(v); });
            add({"minima_cache_dir", "seeds_dir", "seed_dir"}, [](C& c, S v) { c.traps.minima_cache_dir = v; });
            add({"force_regenerate_traps", "force_regenerate_seeds", "regenerate_seeds"},
                [](C& c, S v) { c.traps.force_regenerate_traps = text::boolean(v); });
            add({"g_values"}, [](C& c, S v) { c.quantum.g_values = text::list<double>(v); });
            add({"g"}, [](C& c, S v) { c.quantum.g_values = {std::stod(v)}; });
            add({"dt", "simulation_dt"}, [](C& c, S v) { c.quantum.dt = std::stod(v); });
            add({"t_start", "sampling_start_time", "burn_in_time"}, [](C& c, S v) { c.quantum.t_start = std::stod(v); });
            add({"dt_sample", "sample_dt", "sampling_delta_t"}, [](C& c, S v) { c.quantum.dt_sample = std::stod(v); });
            add({"k_sample_points", "k_steps", "num_sample_points", "k"},
                [](C& c, S v) { c.quantum.k_sample_points = std::stoi(v); });
            add({"m_shots_per_point", "m_samples", "shots_per_point", "m"},
                [](C& c, S v) { c.quantum.m_shots_per_point = std::stoi(v); });
            add({"num_sampling_seeds", "sampling_seeds", "num_seeds"},
                [](C& c, S v) { c.quantum.num_sampling_seeds = std::stoi(v); });
            add({"seed_stride", "sampling_seed_stride"}, [](C& c, S v) { c.quantum.seed_stride = std::stoull(v); });
            add({"track_p_less"}, [](C& c, S v) { c.quantum.track_p_less = text::boolean(v); });
            add({"record_trajectory", "track_trajectory", "continuous_trajectory"},
                [](C& c, S v) { c.quantum.record_trajectory = text::boolean(v); });
            add({"dt_trajectory", "dt_log", "trajectory_dt"}, [](C& c, S v) { c.quantum.dt_trajectory = std::stod(v); });
            add({"t_max", "simulation_t_max"}, [](C& c, S v) { c.quantum.t_max_val = std::stod(v); });
            add({"run_classical", "enable_classical"}, [](C& c, S v) { c.quantum.run_classical = text::boolean(v); });
            add({"compute_transverse", "eval_transverse"}, [](C&, S) {});  // accepted, no effect
            add({"kick_trials_per_d", "classical_trials_per_d", "num_classical_trials"}, [](C& c, S v) {
                c.classical.kick_trials_per_d = std::stoul(v);
                c.classical.restart_trials = c.classical.kick_trials_per_d;
            });
            add({"kick_d_values", "kick_scan_d", "d_scan", "kick_distances"},
                [](C& c, S v) { c.classical.kick_d_values = text::list<int>(v); });
            add({"restart_trials", "num_restart_trials"}, [](C& c, S v) { c.classical.restart_trials = std::stoul(v); });
            add({"max_evals_per_start", "evaluation_budget"},
                [](C& c, S v) { c.classical.max_evals_per_start = std::stoul(v); });
            add({"batch_size"}, [](C& c, S v) { c.classical.batch_size = std::stoul(v); });
            add({"enable_tabu", "run_tabu"}, [](C& c, S v) { c.classical.enable_tabu = text::boolean(v); });
            add({"tabu_trials", "num_tabu_trials"}, [](C& c, S v) { c.classical.tabu_trials = std::stoul(v); });
            add({"tabu_max_steps", "tabu_steps"}, [](C& c, S v) { c.classical.tabu_max_steps = std::stoul(v); });
            add({"tabu_kick_d_values", "tabu_kick_d", "tabu_d_values"},
                [](C& c, S v) { c.classical.tabu_kick_d_values = text::list<int>(v); });
            add({"tabu_tenure", "tenure"}, [](C& c, S v) { c.classical.tabu_tenure = std::stoi(v); });
            add({"output_dir"}, [](C& c, S v) { c.output.output_dir = v; });
            add({"results_csv", "csv_results", "csv_output"}, [](C& c, S v) { c.output.results_csv = v; });
            add({"time_series_csv", "csv_time_series", "csv_trajectory"}, [](C& c, S v) { c.output.time_series_csv = v; });
            add({"structured_json", "json_output", "structured_output", "json_benchmark"},
                [](C& c, S v) { c.output.structured_json = v; });
            add({"all_structured_json", "all_json_output", "overall_json"},
                [](C& c, S v) { c.output.all_structured_json = v; });
            add({"summary_csv", "csv_summary"}, [](C& c, S v) { c.output.summary_csv = v; });
            add({"verbose"}, [](C& c, S v) {
                const bool b = text::boolean(v);
                c.output.verbose = b;
                c.quantum.verbose = b;
            });
            return t;
        }();
        return table;
    }
};

// The sampler configuration of one (n, g) run of a benchmark config, as the
// connector builds it (connector.cuh:628-643); seed is the trap's sampling seed.
inline SamplerConfig to_sampler_config(const QuantumConfig& q, int n, double g, uint64_t seed)
{
    SamplerConfig s;
    s.n = n;
    s.g = g;
    s.dt = q.dt;
    s.t_max = q.t_max();
    s.num_sample_points = q.k_sample_points;
    s.shots_per_point = q.m_shots_per_point;
    s.seed = seed;
    s.t_start = q.t_start;
    s.dt_sample = q.dt_sample;
    s.track_p_less = q.track_p_less;
    s.verbose = q.verbose;
    s.record_trajectory = q.record_trajectory;
    s.dt_trajectory = q.dt_trajectory;
    s.num_sampling_seeds = q.num_sampling_seeds;
    s.seed_stride = q.seed_stride;
    return s;
}

}  // namespace config
}  // namespace bk

// the reference's global names
#ifndef BK_NO_GLOBAL_CONFIG_NAMES
using bk::config::BenchmarkConfig;
using bk::config::ClassicalConfig;
using bk::config::OutputConfig;
using bk::config::QuantumConfig;
using bk::config::SystemConfig;
using bk::config::TrapsConfig;
#endif
