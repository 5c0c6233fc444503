// bk/wire.hpp -- the benchmark connector's output files for sampler results
// (C++17, header-only, host code): the results CSV, the time-series CSV and the
// structured JSON in the reference's wire format, so that outputs diff directly
// against the reference's data files and feed its post-processing unchanged.
//
//   TrapBenchmark ............ SingleTrapBenchmark, connector.cuh:26-117
//   fill_quantum ............. connector.cuh:647-680
//   window_avg_p_less ........ connector.cuh:816-830
//   set_canonical ............ connector.cuh:1284-1292
//   export_results_csv ....... connector.cuh:839-904
//   export_time_series_csv ... connector.cuh:906-933
//   export_structured_json ... connector.cuh:935-1155
//   export_exhaustive_json ... exhaustive_minima.cpp:328-372 (the explorer's file)
//
// Byte for byte, stream state included.  The reference writes each file
// through one std::ofstream and sets a float format only on some columns, so
// a column written without one takes the format the previous column left:
// from the second row of the time-series CSV on, g prints as 3.000000e-01
// (data_processed/probability_dynamics_n*.csv show it).  std::uppercase, set
// for the hex trap bits, is never cleared: scientific values written after it
// print 'E'.  `Col` below names, per column, which of these the reference does.
//
// The classical-baseline fields (kick + steepest descent, restart, tabu) stay
// at the zeros the connector writes when it skips the classical runs; a caller
// that has them can fill them in.
#pragma once

#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <map>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sampler.hpp"

namespace bk {
namespace wire {

// One trap at one field g (field names and defaults of SingleTrapBenchmark).
struct TrapBenchmark {
    int n = 0;
    double g = 0.0;
    size_t trap_rank = 0;
    SpinState trap_state;
    int64_t trap_energy = 0;
    int d_min = 0;

    std::string canonical_id = "";
    std::string canonical_hex = "";
    std::string canonical_bitstring = "";
    double basin_hit_fraction = 0.0;

    double window_avg_p_less = 0.0;
    double window_avg_dh = 0.0;
    double initial_rise_rate_avg = 0.0;
    double initial_rise_rate_max = 0.0;
    size_t quantum_improved_shots = 0;
    size_t quantum_total_shots = 0;
    double quantum_success_rate = 0.0;
    double quantum_per_eval_rate = 0.0;
    size_t quantum_basis_size = 0;
    double quantum_time_sec = 0.0;
    std::vector<double> time_points;
    std::vector<double> p_less_history;
    std::vector<double> dh_history;
    double p_blind = 0.0;
    double peak_p_less = 0.0;
    double peak_time = 0.0;

    int num_sampling_seeds = 1;
    std::vector<uint64_t> sample_seeds;
    std::vector<double> sample_success_rates;
    std::vector<size_t> sample_improved_shots;
    std::vector<size_t> sample_total_shots;
    std::vector<int64_t> sample_best_energies;
    std::vector<double> sample_mean_hamming_flips;

    int best_d = 0;
    double best_d_rate = 0.0;
    size_t best_d_successes = 0;
    size_t best_d_trials = 0;
    size_t best_d_evals = 0;
    std::map<int, double> rate_by_d;
    std::map<int, size_t> successes_by_d;
    std::map<int, size_t> trials_by_d;
    std::map<int, size_t> evals_by_d;
    double kick_time_sec = 0.0;

    double restart_rate = 0.0;
    size_t restart_successes = 0;
    size_t restart_trials = 0;
    size_t restart_evals = 0;
    double restart_time_sec = 0.0;

    bool tabu_enabled = false;
    double tabu_direct_rate = 0.0;
    size_t tabu_direct_successes = 0;
    size_t tabu_direct_trials = 0;
    size_t tabu_direct_evals = 0;
    double tabu_direct_time_sec = 0.0;

    int tabu_kick_d = 0;
    double tabu_kick_rate = 0.0;
    size_t tabu_kick_successes = 0;
    size_t tabu_kick_trials = 0;
    size_t tabu_kick_evals = 0;
    std::map<int, double> tabu_rate_by_d;
    std::map<int, size_t> tabu_successes_by_d;
    std::map<int, size_t> tabu_trials_by_d;
    std::map<int, size_t> tabu_evals_by_d;
    double tabu_kick_time_sec = 0.0;

    double tabu_restart_rate = 0.0;
    size_t tabu_restart_successes = 0;
    size_t tabu_restart_trials = 0;
    size_t tabu_restart_evals = 0;
    double tabu_restart_time_sec = 0.0;

    double tabu_best_rate = 0.0;
    std::string tabu_best_method = "";

    size_t sd_total_evaluations = 0;
    size_t tabu_total_evaluations = 0;
    size_t total_evaluations = 0;
};

// Copy a sampler result into the record (what the connector takes from
// CuStateVecSampler::sample).  d_min is kept when already set.
inline void fill_quantum(TrapBenchmark& b, const SamplingResult& res)
{
    b.time_points = res.time_points;
    b.p_less_history = res.p_less_history;
    b.dh_history = res.dh_history;
    b.window_avg_dh = res.window_avg_dh;
    b.initial_rise_rate_avg = res.initial_rise_rate_avg;
    b.initial_rise_rate_max = res.initial_rise_rate_max;
    if (b.d_min == 0) b.d_min = res.d_min;
    b.quantum_improved_shots = res.improved_shots;
    b.quantum_total_shots = res.total_shots_collected;
    b.quantum_success_rate = res.success_rate;
    b.quantum_per_eval_rate = res.per_eval_rate;
    b.quantum_basis_size = res.unique_basis.size();
    b.p_blind = res.p_blind;
    b.peak_p_less = res.peak_p_less;
    b.peak_time = res.peak_time;
    b.num_sampling_seeds = static_cast<int>(res.seed_stats.size());
    b.sample_seeds.clear();
    b.sample_success_rates.clear();
    b.sample_improved_shots.clear();
    b.sample_total_shots.clear();
    b.sample_mean_hamming_flips.clear();
    b.sample_best_energies.clear();
    for (const SeedSamplingStats& s : res.seed_stats) {
        b.sample_seeds.push_back(s.seed);
        b.sample_success_rates.push_back(s.success_rate);
        b.sample_improved_shots.push_back(s.improved_shots);
        b.sample_total_shots.push_back(s.total_shots_collected);
        b.sample_mean_hamming_flips.push_back(s.mean_hamming_flips);
        b.sample_best_energies.push_back(s.best_energy);
    }
}

// Mean of P_< over the sampling window [t_start, t_start + (k - 1) dt_sample]
// (the whole run, [t_start, t_max], when there is no such window), with the
// connector's 1e-9 tolerance on the ends.
inline double window_avg_p_less(const TrapBenchmark& b, double t_start, int k_sample_points, double dt_sample,
                                double t_max)
{
    const double t_end = (k_sample_points > 1 && dt_sample > 0.0) ? t_start + (k_sample_points - 1) * dt_sample : t_max;
    double sum = 0.0;
    size_t count = 0;
    for (size_t i = 0; i < b.time_points.size() && i < b.p_less_history.size(); ++i) {
        const double t = b.time_points[i];
        if (t >= t_start - 1e-9 && t <= t_end + 1e-9) {
            sum += b.p_less_history[i];
            ++count;
        }
    }
    return count > 0 ? sum / static_cast<double>(count) : 0.0;
}

// canonical_id / canonical_hex ("0x" + upper-case hex) and the n-character
// bit string, most significant spin first.
inline void set_canonical(TrapBenchmark& b, uint64_t canonical_bits, int n)
{
    std::ostringstream h;
    h << "0x" << std::hex << std::uppercase << canonical_bits;
    b.canonical_hex = h.str();
    b.canonical_id = h.str();
    b.canonical_bitstring.assign(static_cast<size_t>(n), '0');
    for (int i = 0; i < n; ++i)
        if ((canonical_bits >> i) & 1u) b.canonical_bitstring[static_cast<size_t>(n - 1 - i)] = '1';
}

// Column writer over one stream.  keep(): the format the stream is in (what
// the reference does for a column it writes without a manipulator); sci()/
// fix(): set std::scientific / std::fixed and the precision first; bits():
// "0x", upper-case hex, back to decimal (std::uppercase stays set).
class Col {
public:
    explicit Col(std::ostream& os) : os_(os) {}
    template <typename T>
    Col& keep(const T& v)
    {
        os_ << v;
        return *this;
    }
    Col& sci(double v)
    {
        os_ << std::scientific << std::setprecision(6) << v;
        return *this;
    }
    Col& fix(double v, int prec = 4)
    {
        os_ << std::fixed << std::setprecision(prec) << v;
        return *this;
    }
    Col& bits(uint64_t v)
    {
        os_ << "0x" << std::hex << std::uppercase << v << std::dec;
        return *this;
    }
    Col& txt(const char* s)
    {
        os_ << s;
        return *this;
    }
    Col& txt(const std::string& s)
    {
        os_ << s;
        return *this;
    }
    // "[a, b, ...]" with each element written by f
    template <typename V, typename F>
    Col& list(const std::vector<V>& v, F f)
    {
        os_ << "[";
        for (size_t i = 0; i < v.size(); ++i) {
            if (i) os_ << ", ";
            f(*this, v[i]);
        }
        os_ << "]";
        return *this;
    }
    // "{"k": v, ...}" with each value written by f
    template <typename V, typename F>
    Col& dict(const std::map<int, V>& m, F f)
    {
        os_ << "{";
        size_t i = 0;
        for (const auto& kv : m) {
            os_ << (i++ ? ", " : "") << "\"" << kv.first << "\": ";
            f(*this, kv.second);
        }
        os_ << "}";
        return *this;
    }

private:
    std::ostream& os_;
};

namespace detail {
inline std::ofstream open_out(const std::string& path, const char* what)
{
    const std::filesystem::path p(path);
    if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path());
    std::ofstream f(path);
    if (!f.is_open()) throw std::runtime_error(std::string("BenchmarkConnector: ") + what + path);
    return f;
}
inline double seed_rate(const TrapBenchmark& r, size_t i)
{
    return r.sample_success_rates.size() > i ? r.sample_success_rates[i] : 0.0;
}
}  // namespace detail

// One row per record: trap metadata, quantum metrics, classical baselines,
// timings, the first three per-seed sampling rates and the trap bits.
inline void export_results_csv(const std::string& path, const std::vector<TrapBenchmark>& results)
{
    std::ofstream file = detail::open_out(path, "Cannot open results CSV for writing: ");
    file << "N,g,trap_rank,canonical_id,canonical_hex,trap_energy,d_min,window_avg_p_less,window_avg_dh,"
            "initial_rise_rate_avg,initial_rise_rate_max,peak_p_less,peak_time,p_blind,"
            "best_d,best_d_rate,best_d_successes,best_d_trials,best_d_evals,"
            "restart_rate,restart_successes,restart_trials,restart_evals,"
            "tabu_best_method,tabu_best_rate,tabu_direct_rate,tabu_kick_d,tabu_kick_rate,tabu_restart_rate,"
            "sd_total_evaluations,tabu_total_evaluations,total_evaluations,"
            "quantum_time_sec,kick_time_sec,restart_time_sec,tabu_kick_time_sec,"
            "sample_seed1_rate,sample_seed2_rate,sample_seed3_rate,trap_bits\n";
    Col c(file);
    for (const TrapBenchmark& r : results) {
        c.keep(r.n).txt(",").keep(r.g).txt(",").keep(r.trap_rank).txt(",");
        c.txt(r.canonical_id).txt(",").txt(r.canonical_hex).txt(",").keep(r.trap_energy).txt(",").keep(r.d_min).txt(",");
        c.sci(r.window_avg_p_less).txt(",").fix(r.window_avg_dh).txt(",");
        c.sci(r.initial_rise_rate_avg).txt(",").sci(r.initial_rise_rate_max).txt(",");
        c.keep(r.peak_p_less).txt(",").fix(r.peak_time).txt(",").sci(r.p_blind).txt(",");
        c.keep(r.best_d).txt(",").keep(r.best_d_rate).txt(",").keep(r.best_d_successes).txt(",");
        c.keep(r.best_d_trials).txt(",").keep(r.best_d_evals).txt(",");
        c.keep(r.restart_rate).txt(",").keep(r.restart_successes).txt(",").keep(r.restart_trials).txt(",");
        c.keep(r.restart_evals).txt(",");
        c.txt("\"").txt(r.tabu_best_method).txt("\",").keep(r.tabu_best_rate).txt(",").keep(r.tabu_direct_rate).txt(",");
        c.keep(r.tabu_kick_d).txt(",").keep(r.tabu_kick_rate).txt(",").keep(r.tabu_restart_rate).txt(",");
        c.keep(r.sd_total_evaluations).txt(",").keep(r.tabu_total_evaluations).txt(",").keep(r.total_evaluations).txt(",");
        c.fix(r.quantum_time_sec).txt(",").keep(r.kick_time_sec).txt(",").keep(r.restart_time_sec).txt(",");
        c.keep(r.tabu_kick_time_sec).txt(",");
        c.sci(detail::seed_rate(r, 0)).txt(",").sci(detail::seed_rate(r, 1)).txt(",").sci(detail::seed_rate(r, 2)).txt(",");
        c.bits(r.trap_state.bits).txt("\n");
    }
}

// One row per recorded time of every record: N, g, rank, E0, t, P_<, <d_H>, P_blind.
inline void export_time_series_csv(const std::string& path, const std::vector<TrapBenchmark>& results)
{
    std::ofstream file = detail::open_out(path, "Cannot open time series CSV: ");
    file << "N,g,trap_rank,trap_energy,time,p_less,dH,p_blind\n";
    Col c(file);
    for (const TrapBenchmark& r : results) {
        for (size_t i = 0; i < r.time_points.size() && i < r.p_less_history.size(); ++i) {
            const double dh = i < r.dh_history.size() ? r.dh_history[i] : 0.0;
            c.keep(r.n).txt(",").keep(r.g).txt(",").keep(r.trap_rank).txt(",").keep(r.trap_energy).txt(",");
            c.fix(r.time_points[i]).txt(",").sci(r.p_less_history[i]).txt(",").fix(dh).txt(",").sci(r.p_blind).txt("\n");
        }
    }
}

// {"system": {N, num_seeds}, "seeds": [per (N, trap_rank): metadata, classical
// baselines of its first record, quantum dynamics keyed by g (two decimals)]}.
// n_override = 0 takes N from the first record.
inline void export_structured_json(const std::string& path, const std::vector<TrapBenchmark>& results,
                                   int n_override = 0)
{
    std::ofstream file = detail::open_out(path, "Cannot open structured JSON for writing: ");
    std::map<std::pair<int, size_t>, std::vector<const TrapBenchmark*>> traps;
    for (const TrapBenchmark& r : results) traps[{r.n, r.trap_rank}].push_back(&r);
    const int sys_n = (n_override == 0 && !results.empty()) ? results.front().n : n_override;
    Col c(file);
    c.txt("{\n  \"system\": {\n    \"N\": ").keep(sys_n).txt(",\n    \"num_seeds\": ").keep(traps.size());
    c.txt("\n  },\n  \"seeds\": [\n");
    const auto sci_v = [](Col& o, double v) { o.sci(v); };
    const auto fix_v = [](Col& o, double v) { o.fix(v); };
    const auto keep_v = [](Col& o, auto v) { o.keep(v); };
    size_t ti = 0;
    for (auto it = traps.begin(); it != traps.end(); ++it, ++ti) {
        const TrapBenchmark& f = *it->second.front();
        c.txt("    {\n      \"N\": ").keep(f.n).txt(",\n      \"trap_rank\": ").keep(f.trap_rank);
        c.txt(",\n      \"canonical_id\": \"").txt(f.canonical_id);
        c.txt("\",\n      \"canonical_hex\": \"").txt(f.canonical_hex);
        c.txt("\",\n      \"canonical_bitstring\": \"").txt(f.canonical_bitstring);
        c.txt("\",\n      \"trap_state_hex\": \"").bits(f.trap_state.bits);
        c.txt("\",\n      \"trap_energy_E0\": ").keep(f.trap_energy).txt(",\n      \"d_min\": ").keep(f.d_min).txt(",\n");

        c.txt("      \"classical_baselines\": {\n");
        c.txt("        \"best_d\": ").keep(f.best_d).txt(",\n        \"best_d_rate\": ").sci(f.best_d_rate);
        c.txt(",\n        \"best_d_successes\": ").keep(f.best_d_successes);
        c.txt(",\n        \"best_d_trials\": ").keep(f.best_d_trials);
        c.txt(",\n        \"best_d_evals\": ").keep(f.best_d_evals).txt(",\n");
        c.txt("        \"rates_by_d\": ").dict(f.rate_by_d, sci_v).txt(",\n");
        c.txt("        \"trials_by_d\": ").dict(f.trials_by_d, keep_v).txt(",\n");
        c.txt("        \"evals_by_d\": ").dict(f.evals_by_d, keep_v).txt(",\n");
        c.txt("        \"restart_rate\": ").sci(f.restart_rate);
        c.txt(",\n        \"restart_successes\": ").keep(f.restart_successes);
        c.txt(",\n        \"restart_trials\": ").keep(f.restart_trials);
        c.txt(",\n        \"restart_evals\": ").keep(f.restart_evals);
        c.txt(",\n        \"sd_total_evaluations\": ").keep(f.sd_total_evaluations).txt(",\n");
        c.txt("        \"tabu_direct_rate\": ").sci(f.tabu_direct_rate);
        c.txt(",\n        \"tabu_direct_successes\": ").keep(f.tabu_direct_successes);
        c.txt(",\n        \"tabu_direct_trials\": ").keep(f.tabu_direct_trials);
        c.txt(",\n        \"tabu_direct_evals\": ").keep(f.tabu_direct_evals).txt(",\n");
        c.txt("        \"tabu_kick_d\": ").keep(f.tabu_kick_d).txt(",\n        \"tabu_kick_rate\": ").sci(f.tabu_kick_rate);
        c.txt(",\n        \"tabu_kick_successes\": ").keep(f.tabu_kick_successes);
        c.txt(",\n        \"tabu_kick_trials\": ").keep(f.tabu_kick_trials);
        c.txt(",\n        \"tabu_kick_evals\": ").keep(f.tabu_kick_evals).txt(",\n");
        c.txt("        \"tabu_rates_by_d\": ").dict(f.tabu_rate_by_d, sci_v).txt(",\n");
        c.txt("        \"tabu_trials_by_d\": ").dict(f.tabu_trials_by_d, keep_v).txt(",\n");
        c.txt("        \"tabu_evals_by_d\": ").dict(f.tabu_evals_by_d, keep_v).txt(",\n");
        c.txt("        \"tabu_restart_rate\": ").sci(f.tabu_restart_rate);
        c.txt(",\n        \"tabu_restart_successes\": ").keep(f.tabu_restart_successes);
        c.txt(",\n        \"tabu_restart_trials\": ").keep(f.tabu_restart_trials);
        c.txt(",\n        \"tabu_restart_evals\": ").keep(f.tabu_restart_evals).txt(",\n");
        c.txt("        \"tabu_best_rate\": ").sci(f.tabu_best_rate);
        c.txt(",\n        \"tabu_best_method\": \"").txt(f.tabu_best_method);
        c.txt("\",\n        \"tabu_total_evaluations\": ").keep(f.tabu_total_evaluations);
        c.txt(",\n        \"total_evaluations\": ").keep(f.total_evaluations).txt(",\n");
        c.txt("        \"kick_time_sec\": ").fix(f.kick_time_sec);
        c.txt(",\n        \"restart_time_sec\": ").fix(f.restart_time_sec);
        c.txt(",\n        \"tabu_direct_time_sec\": ").fix(f.tabu_direct_time_sec);
        c.txt(",\n        \"tabu_kick_time_sec\": ").fix(f.tabu_kick_time_sec);
        c.txt(",\n        \"tabu_restart_time_sec\": ").fix(f.tabu_restart_time_sec).txt("\n      },\n");

        c.txt("      \"quantum_dynamics\": {\n");
        const std::vector<const TrapBenchmark*>& gs = it->second;
        for (size_t gi = 0; gi < gs.size(); ++gi) {
            const TrapBenchmark& q = *gs[gi];
            std::ostringstream key;
            key << std::fixed << std::setprecision(2) << q.g;
            c.txt("        \"").txt(key.str()).txt("\": {\n          \"g\": ").fix(q.g, 2);
            c.txt(",\n          \"window_avg_p_less\": ").sci(q.window_avg_p_less);
            c.txt(",\n          \"window_avg_dH\": ").fix(q.window_avg_dh);
            c.txt(",\n          \"initial_rise_rate_avg\": ").sci(q.initial_rise_rate_avg);
            c.txt(",\n          \"initial_rise_rate_max\": ").sci(q.initial_rise_rate_max);
            c.txt(",\n          \"peak_p_less\": ").sci(q.peak_p_less);
            c.txt(",\n          \"peak_time\": ").fix(q.peak_time);
            c.txt(",\n          \"p_blind\": ").sci(q.p_blind).txt(",\n");
            c.txt("          \"time_array\": ").list(q.time_points, fix_v).txt(",\n");
            c.txt("          \"p_less_array\": ").list(q.p_less_history, sci_v).txt(",\n");
            c.txt("          \"dH_array\": ").list(q.dh_history, fix_v).txt(",\n");
            c.txt("          \"quantum_time_sec\": ").fix(q.quantum_time_sec).txt(",\n");
            c.txt("          \"sampling_validation\": {\n");
            c.txt("            \"num_sampling_seeds\": ").keep(q.num_sampling_seeds).txt(",\n");
            c.txt("            \"seeds\": ").list(q.sample_seeds, keep_v).txt(",\n");
            c.txt("            \"sample_success_rates\": ").list(q.sample_success_rates, sci_v).txt(",\n");
            c.txt("            \"sample_improved_shots\": ").list(q.sample_improved_shots, keep_v).txt(",\n");
            c.txt("            \"sample_total_shots\": ").list(q.sample_total_shots, keep_v).txt(",\n");
            c.txt("            \"sample_best_energies\": ").list(q.sample_best_energies, keep_v).txt(",\n");
            c.txt("            \"sample_mean_hamming_flips\": ").list(q.sample_mean_hamming_flips, fix_v).txt("\n");
            c.txt("          }\n        }").txt(gi + 1 < gs.size() ? "," : "").txt("\n");
        }
        c.txt("      }\n    }").txt(ti + 1 < traps.size() ? "," : "").txt("\n");
    }
    c.txt("  ]\n}\n");
}

// The local-minima explorer's result (ExhaustiveResult / LevelInfo /
// OrbitInfo, exhaustive_minima.cpp:29-50): the trap levels above E* and their
// orbits, each by its canonical state.
struct MinimaOrbit {
    uint64_t canonical_bits = 0;
    int d_min = 0;
    double sd_hit_fraction = 0.0;
};
struct MinimaLevel {
    int64_t energy = 0;
    std::vector<MinimaOrbit> orbits;
};
struct ExhaustiveMinima {
    int n = 0;
    int64_t e_star = 0;
    std::vector<MinimaLevel> levels;
};

// save_to_json (exhaustive_minima.cpp:328-372): canonical_id is "0x" and
// upper-case hex, the bit string is n characters with the most significant
// spin first, sd_hit_fraction is fixed with 6 decimals.
inline void export_exhaustive_json(const std::string& path, const ExhaustiveMinima& r)
{
    std::ofstream f = detail::open_out(path, "cannot open exhaustive-minima JSON: ");
    Col c(f);
    c.txt("{\n  \"N\": ").keep(r.n).txt(",\n  \"E_star\": ").keep(r.e_star).txt(",\n  \"levels\": [\n");
    for (size_t i = 0; i < r.levels.size(); ++i) {
        const MinimaLevel& lv = r.levels[i];
        c.txt("    {\n      \"E0\": ").keep(lv.energy).txt(",\n      \"orbits\": [\n");
        for (size_t j = 0; j < lv.orbits.size(); ++j) {
            const MinimaOrbit& o = lv.orbits[j];
            TrapBenchmark ids;
            set_canonical(ids, o.canonical_bits, r.n);
            c.txt("        {\n          \"canonical_id\": \"").txt(ids.canonical_id);
            c.txt("\",\n          \"bitstring\": \"").txt(ids.canonical_bitstring);
            c.txt("\",\n          \"d_min\": ").keep(o.d_min);
            c.txt(",\n          \"sd_hit_fraction\": ").fix(o.sd_hit_fraction, 6);
            c.txt("\n        }").txt(j + 1 < lv.orbits.size() ? "," : "").txt("\n");
        }
        c.txt("      ]\n    }").txt(i + 1 < r.levels.size() ? "," : "").txt("\n");
    }
    c.txt("  ]\n}\n");
}

}  // namespace wire
}  // namespace bk
