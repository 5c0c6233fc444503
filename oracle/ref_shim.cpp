// ref_shim.cpp -- C entry points over the reference's own CPU code, compiled
// from the reference sources where they lie (see oracle/Makefile target "ref").
// TEST INFRASTRUCTURE ONLY: used by the CPU tests to pin the oracle's
// restatement of labs_energy.hpp and classical.cuh against the original.
#include <cstdint>

#include <labs_energy/labs_energy.hpp>
#include <classical/classical.cuh>

extern "C" {

int64_t bkref_energy(int n, uint64_t s) { return LabsEvaluator(n).calcEnergy(SpinState{s}); }

void bkref_ck(int n, uint64_t s, int64_t* out)
{
    std::vector<int64_t> c = LabsEvaluator(n).calcCk(SpinState{s});
    for (int k = 0; k < n; ++k) out[k] = c[k];
}

uint64_t bkref_canonical(int n, uint64_t s) { return SpinState{s}.canonical(n).bits; }

int64_t bkref_known_optimal_energy(int n) { return LabsEvaluator::known_optimal_energy(n); }

int bkref_d_min(int n, uint64_t s0, int64_t target, int max_d)
{
    return LabsEvaluator(n).compute_d_min(SpinState{s0}, target, max_d);
}

// seed -> first draw of the optimizer's generator (classical.cuh:602-604)
uint64_t bkref_random_state(int n, uint64_t seed)
{
    LabsEvaluator ev(n);
    ClassicalOptimizer opt(ev, seed);
    return opt.random_state().bits;
}

// classical.cuh:625-713
void bkref_steepest_descent(int n, uint64_t start, uint64_t* final_state, int64_t* final_energy,
                            uint64_t* evals, uint64_t* steps)
{
    LabsEvaluator ev(n);
    ClassicalOptimizer opt(ev, 0);
    DescentResult r = opt.steepest_descent(SpinState{start});
    *final_state = r.final_state.bits;
    *final_energy = r.final_energy;
    *evals = r.evals_used;
    *steps = r.steps_taken;
}

}  // extern "C"

// The reference's exhaustive local-minima explorer (src/tools/exhaustive_minima.cpp),
// compiled from its own source file; its command-line main is renamed so the
// shared library exports none.
#define main bkref_exhaustive_minima_cli_main
#include <tools/exhaustive_minima.cpp>
#undef main

extern "C" {

// run_exploration (exhaustive_minima.cpp:109-216) with the optional descent
// sampling (run_sd_benchmarks, :218-271); the levels it keeps (at most 2)
// are written level by level: energies, orbit counts, then per orbit its
// canonical state, d_min and sd_hit_fraction (at most cap orbits).  Returns
// the number of levels.  With json_path non-null the result is also saved by
// the reference's own writer (save_to_json, :328-372).
int bkref_exhaustive_minima(int n, int with_sd, uint64_t num_sd, int64_t* e_star, int64_t* total_minima,
                            int64_t* total_orbits, int64_t* level_energy, int64_t* level_orbits, uint64_t* canon,
                            int* d_min, double* sd_hit, int64_t cap, const char* json_path)
{
    ExhaustiveResult r = run_exploration(n, 0, with_sd != 0, static_cast<size_t>(num_sd));
    *e_star = r.e_star;
    *total_minima = r.total_minima_found;
    *total_orbits = r.total_orbits_found;
    int64_t w = 0;
    for (size_t l = 0; l < r.levels.size(); ++l) {
        level_energy[l] = r.levels[l].energy;
        level_orbits[l] = static_cast<int64_t>(r.levels[l].orbits.size());
        for (const OrbitInfo& o : r.levels[l].orbits) {
            if (w >= cap) break;
            canon[w] = o.canonical_state.bits;
            d_min[w] = o.d_min;
            sd_hit[w] = o.sd_hit_fraction;
            ++w;
        }
    }
    if (json_path) save_to_json(r, json_path);
    return static_cast<int>(r.levels.size());
}

}  // extern "C"
