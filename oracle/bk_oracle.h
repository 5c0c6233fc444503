/*
 * bk_oracle.h -- CPU restatement of the reference's Trotter / ED / Lanczos path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only to compute expected values or to time the CPU
 * baseline.  The product (bernasconi-krylov_amd/) never links or calls it.
 *
 * Each function cites the reference file:line whose algorithm it restates.
 * Parity pins (what checks each function against something other than itself):
 *   bko_energy / bko_ck        exhaustive E(s), C_k tables (data_raw/classical/exhaustive),
 *                              test_labs_energy.cpp known answers, oracle/_ref build
 *   bko_known_optimal_energy   brute-force min over all 2^N (N <= 24 in tests)
 *   bko_d_min                  brute force over all states (tests), oracle/_ref,
 *                              exhaustive_minima_n*.json d_min values
 *   bko_exhaustive_minima      exhaustive_minima_n{20,22,24}.json (E*, trap energies,
 *                              canonical orbit ids), brute force over all states and
 *                              their neighbours at n <= 14 (tests)
 *   bko_random_state /
 *   bko_steepest_descent       oracle/_ref (reference classical.cuh compiled here),
 *                              seed fixtures in tests/golden
 *   bko_adjoint_grad           central finite differences of bko_trotter (tests)
 *   bko_trotter                probability_dynamics_n*.csv golden trajectories,
 *                              closed form <d_H>(dt) = N sin^2(g dt), mixer-only
 *                              closed form, dense exp(-iHt) via scipy at N <= 10
 *   bko_hv_*                   dense H built independently in numpy (tests)
 *   bko_syev / bko_dense_*     scipy.linalg.eigh (tests), closed spectra
 *   bko_lanczos                dense ED ground energy (test_lanczos.cu:38 bound 1e-6)
 *   bko_krylov_propagate       scipy expm_multiply at N <= 10
 *   bko_sample_inverse_cdf     brute-force CDF search in numpy (tests)
 *   bko_derived_metrics        hand-computed series in tests (first drop, slopes)
 *   bko_schedule               reference time-grid rules, test_sampler.cu /
 *                              test_config.cpp counts
 */
#ifndef BK_ORACLE_H
#define BK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- L0 math core: labs_energy.hpp ---- */
int64_t bko_energy(int n, uint64_t s);
void bko_ck(int n, uint64_t s, int64_t *ck /* length n, ck[0]=0 */);
uint64_t bko_canonical(int n, uint64_t s);
int64_t bko_known_optimal_energy(int n);
int bko_d_min(int n, uint64_t s0, int64_t target_energy, int max_d);
/* energies of all 2^n states (int32), OpenMP */
void bko_all_energies(int n, int32_t *out);
/* count of states with E < e0 (P_blind numerator) */
int64_t bko_count_below(int n, int64_t e0);

/* ---- mt19937_64 and trap construction: classical.cuh:602-713 ---- */
typedef struct { uint64_t mt[312]; int idx; } bko_mt64;
void bko_mt64_seed(bko_mt64 *g, uint64_t seed);
uint64_t bko_mt64_next(bko_mt64 *g);
/* libstdc++ std::uniform_real_distribution<double>(a,b) on mt19937_64 */
double bko_mt64_uniform(bko_mt64 *g, double a, double b);
uint64_t bko_random_state(int n, uint64_t seed);
void bko_steepest_descent(int n, uint64_t start, uint64_t *final_state,
                          int64_t *final_energy, uint64_t *evals, uint64_t *steps);

/* ---- exhaustive one-flip local minima: exhaustive_minima.cpp:74-208 ----
 * Scans s in [0, 2^(n-1)) for one-flip minima and groups their canonical
 * images by energy.  Writes E* (lowest energy with a minimum; -1 if none),
 * total_minima (2 x the half-space count, as the reference reports it),
 * total_orbits (distinct canonical states over all energies), num_energies,
 * and for the trap levels l = 0..levels-1 (the energies after E*)
 * level_energy[l], level_orbits[l] and their canonical states in ascending
 * order, level by level, into orbits (at most cap).  Returns the number of
 * trap levels found (<= levels), -1 on bad arguments or allocation failure. */
int bko_exhaustive_minima(int n, int levels, int64_t *e_star, int64_t *total_minima, int64_t *total_orbits,
                          int64_t *num_energies, int64_t *level_energy, int64_t *level_orbits, uint64_t *orbits,
                          int64_t cap);

/* ---- FP64 statevector Trotter loop: sampler.cuh:54-227, 412-619 ----
 * order: 1 = Lie, phase first (U = X P, the reference's loop);
 *        2 = Strang with the phase outside (U = P(dt/2) X P(dt/2)).
 * psi: 2*2^n doubles (interleaved re, im).  If init_from_s0, psi is set to |s0>.
 * g_steps: per-step field (length steps) or NULL for constant g.
 * Observables of the state before step k are written at index k (k = 0..steps);
 * any output pointer may be NULL.  threads <= 0: OpenMP default. */
int bko_trotter(int n, uint64_t s0, int64_t e0, double g, const double *g_steps,
                double dt, int steps, int order, int init_from_s0, double *psi,
                double *p_less, double *dh, double *norm, int threads);
/* Timed CPU baseline: allocate |s0>, precompute energies, then time only the
 * 'steps' Trotter steps (no observables).  Returns 0 and the step-loop seconds. */
int bko_trotter_bench(int n, uint64_t s0, double g, double dt, int steps, int order,
                      int threads, double *seconds);
/* Gradient of L = wp * P_<(T) + wd * <d_H>(T) with respect to the field of each
 * step (g_steps[k], k = 0..steps-1) for the evolution of bko_trotter from |s0>
 * (no reference counterpart: the reference has no differentiable path).
 * Plain reverse mode in the unrotated frame: psi_T from a forward run,
 * lambda_T = W psi_T; then for k from the last step down, psi_{k-1} = U_k^+ psi_k,
 * grad[k] = 2 Re <lambda_k| dU_k/dg |psi_{k-1}> with dU/dg applied literally
 * (the step's factors with i dt sum_q X_q inserted after the mixer), and
 * lambda_{k-1} = U_k^+ lambda_k.  obs2 (may be NULL) = {P_<(T), <d_H>(T)}. */
int bko_adjoint_grad(int n, uint64_t s0, int64_t e0, const double *g_steps, double dt,
                     int steps, int order, double wp, double wd, double *grad, double *obs2,
                     int threads);

/* observables of one state */
void bko_observables(int n, const double *psi, uint64_t s0, int64_t e0,
                     double *p_less, double *dh, double *norm);

/* ---- Matrix-free H.v: diagonalizer.cuh:60-77 (diag = E/N when normalize) ---- */
void bko_hv_real(int n, double g, int normalize_by_n, const double *x, double *y);
void bko_hv_complex(int n, double g, int normalize_by_n, const double *x, double *y);

/* ---- Dense ED: diagonalizer.cuh:250-294 and connector.cuh:683-748 ---- */
void bko_dense_h(int n, double g, int normalize_by_n, double *h /* dim*dim row-major */);
/* symmetric eigensolver: a (dim*dim row-major) is overwritten with the
 * eigenvectors stored as ROWS (a[j*dim+i] = component i of eigenvector j);
 * w gets ascending eigenvalues.  Returns 0 on success. */
int bko_syev(int dim, double *a, double *w);
/* tridiagonal eigensolver: d (dim), e (dim-1 off-diagonal).  zt (dim*dim) holds
 * row vectors that the QL rotations act on (identity for the eigenvectors of T;
 * Q^T for those of A = Q T Q^T); may be NULL.  On return d is ascending and
 * row j of zt is the matching eigenvector. */
int bko_tridiag_eig(int dim, double *d, double *e, double *zt);
/* spectral evolution of |s0> at times t[0..nt): P_<, <d_H> (connector.cuh:709-748) */
int bko_ed_evolve(int n, double g, uint64_t s0, int64_t e0, const double *t, int nt,
                  double *p_less, double *dh, int inner_loop_trig, int threads);

/* ---- Lanczos: diagonalizer.cuh:420-554 (no reorthogonalisation, v0 from
 * mt19937_64(42) uniform(-0.5,0.5), breakdown tol) ---- */
int bko_lanczos(int n, double g, int normalize_by_n, int max_krylov_dim, double tol,
                int k_pairs, double *evals, double *evecs /* k_pairs * 2^n, may be NULL */);

/* ---- Krylov propagator (full reorthogonalisation): psi <- exp(-i H tau) psi ---- */
int bko_krylov_propagate(int n, double g, int normalize_by_n, double *psi, double tau, int m);

/* ---- Projective sampling: inverse CDF over |psi|^2, random-number order ----
 * index = min{ i : cdf_i > r * cdf_last }  (cdf_i inclusive prefix sum) */
void bko_sample_inverse_cdf(int n, const double *psi, const double *randnums, int shots,
                            int64_t *out);

/* ---- Derived metrics: sampler.cuh:642-683 ---- */
void bko_derived_metrics(const double *t, const double *p, const double *dh, int len,
                         double avg_start, double avg_end, double *rise_avg,
                         double *rise_max, double *peak_p, double *peak_t,
                         double *window_dh, double *window_p);

/* ---- Measurement / sampling schedule: sampler.cuh:496-603 ----
 * For every step k = 0..steps writes flags[k]: bit0 = P_</<dH> measured,
 * bit1 = projective sample point taken.  Returns total_trotter_steps. */
int bko_schedule(double t_max, double dt, int num_sample_points, int shots_per_point,
                 double t_start, double dt_sample, int track_p_less,
                 int record_trajectory, double dt_trajectory, unsigned char *flags,
                 int flags_len);

#ifdef __cplusplus
}
#endif
#endif
