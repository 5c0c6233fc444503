// bk_host.hpp -- host-side math of the engine (exact integer energies, d_min,
// trap construction, tridiagonal eigensolver, schedule and derived metrics).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bk {

// labs_energy.hpp:121-132
int64_t host_energy(int n, uint64_t s);
// SpinState::canonical (labs_energy.hpp:51-78): smallest image under the order-8 symmetry group
uint64_t host_canonical(int n, uint64_t s);
// labs_energy.hpp:156-168
int64_t host_known_optimal(int n);
// labs_energy.hpp:175-207
int host_d_min(int n, uint64_t s0, int64_t target, int max_d);
// classical.cuh:602-604, 625-713
void host_make_trap(int n, uint64_t seed, uint64_t* start, uint64_t* s0, int64_t* e0);
void host_steepest_descent(int n, uint64_t start, uint64_t* fin, int64_t* fe);

// Symmetric tridiagonal eigen-decomposition (implicit QR, Wilkinson shift).
// d: diagonal (m), e: off-diagonal (m-1).  On success d holds ascending
// eigenvalues and, if z != nullptr, z[k*m + i] = component i of eigenvector k.
bool tridiag_eigen(std::vector<double>& d, std::vector<double> e, std::vector<double>* z);

// sampler.cuh:496-603: per step k = 0..total, bit0 = measure, bit1 = sample point
struct ScheduleIn {
    double t_max, dt, t_start, dt_sample, dt_trajectory;
    int num_sample_points, shots_per_point;
    bool track_p_less, record_trajectory;
};
int build_schedule(const ScheduleIn& in, std::vector<unsigned char>& flags);

// sampler.cuh:642-683
struct Derived {
    double rise_avg, rise_max, peak_p, peak_t, window_dh;
};
Derived derived_metrics(const std::vector<double>& t, const std::vector<double>& p,
                        const std::vector<double>& dh, double avg_start, double avg_end);

}  // namespace bk
