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
