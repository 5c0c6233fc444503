/*
 * bk_capi.h -- C ABI of the MI355X-native Trotter engine (libbk.so).
 *
 * Drop-in boundary for the reference's src/sampling and src/subspace paths.
 * Plain C: POD structs, pointers and sizes, int error codes (0 = success,
 * negative = error, message in bk_last_error()); no exceptions or C++/torch
 * types cross the ABI.  A context owns one device, one HIP stream and one
 * statevector; calls on one context are not thread-safe (the reference's
 * CuStateVecSampler is not either, sampler.cuh:318-338).
 *
 * Reference interfaces replaced (file:line relative to the reference repo):
 *   ISampler::sample / create_sampler ........ sampler.cuh:303-310, 737-743, 412-707
 *   precompute / phase / mixer / reductions .. sampler.cuh:54-227, 604-617
 *   custatevecSampler{Create,Preprocess,Sample} sampler.cuh:464-471, 564-573
 *   matrix-free H.v MatVecFunc ............... diagonalizer.cuh:60-77, 400-406, 420-424
 *   run_gpu_lanczos_internal ................. diagonalizer.cuh:420-554
 *   LabsEvaluator::calcEnergy / compute_d_min  labs_energy.hpp:121-132, 175-207
 *   ClassicalOptimizer random_state + steepest_descent classical.cuh:602-713
 */
#ifndef BK_CAPI_H
#define BK_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BK_OK 0
#define BK_E_INVALID (-1)   /* bad argument (n out of range, NaN field, null pointer, ...) */
#define BK_E_HIP (-2)       /* HIP runtime error (message names the call) */
#define BK_E_NOMEM (-3)     /* device or host allocation failed */
#define BK_E_NOSTATE (-4)   /* the context holds no statevector yet */
#define BK_E_NODEVICE (-5)  /* no HIP device */

#define BK_C64 1  /* complex64: float2 amplitudes, FP32 butterflies */
#define BK_C128 2 /* complex128: double2 amplitudes, FP64 throughout */

#define BK_LIE 1    /* first order, phase first: U = X(g dt) P(dt)  (sampler.cuh:604-617) */
#define BK_STRANG 2 /* second order: U = P(dt/2) X(g dt) P(dt/2) */

#define BK_MAX_N 32

typedef struct bk_ctx bk_ctx;

/* Last error message of the calling thread ("" when none). */
const char* bk_last_error(void);
int bk_version(void);
/* Number of HIP devices (0 when none), without creating a context. */
int bk_device_count(void);

/* Device memory the library holds in this process (all contexts), and a cap on
 * it (0 = none, the default).  An allocation that would pass the cap fails
 * with BK_E_NOMEM, as it would on a full device, and leaves the context usable.
 * No counterpart in the reference (its CHECK_CUDA throws on a failed cudaMalloc,
 * sampler.cuh:31-47, 429-431); for callers that share a device. */
int bk_set_device_memory_limit(uint64_t bytes);
uint64_t bk_device_memory_held(void);

/* One context per device (replaces custatevecCreate/Destroy, sampler.cuh:448-449, 703). */
int bk_ctx_create(int device, bk_ctx** out);
void bk_ctx_destroy(bk_ctx* ctx);

/* ---- L0 math core (host, exact integers) ---- */
/* labs_energy.hpp:121-132 */
int64_t bk_energy(int n, uint64_t s);
/* labs_energy.hpp:175-207 (target < 0 -> E(s0); returns the reference's not-found value) */
int bk_d_min(int n, uint64_t s0, int64_t target, int max_d);
/* labs_energy.hpp:156-168 (-1 outside 3..32) */
int64_t bk_known_optimal_energy(int n);
/* classical.cuh:602-604 (first draw of mt19937_64(seed)) + :625-713 (steepest descent) */
int bk_make_trap(int n, uint64_t seed, uint64_t* start, uint64_t* s0, int64_t* e0);

/* SpinState::canonical (labs_energy.hpp:51-78): the smallest of the 8 images of s
 * under reversal, alternating-spin flip and global negation */
uint64_t bk_canonical(int n, uint64_t s);

/* ---- trap supply at scale (classical.cuh) ----
 * batch_steepest_descent_kernel (classical.cuh:459-535) / steepest_descent
 * (:625-713) on the GPU: best-improvement single flips, lowest index on ties,
 * from each start until no flip lowers E or the next pass would exceed
 * max_evals_per_start evaluations (0: no budget).  Host arrays of count
 * entries; eval_counts (may be NULL) counts 1 + n per pass as the reference
 * does.  Bit-identical to the host descent of bk_make_trap.  2 <= n <= 64. */
int bk_descent_batch(bk_ctx* ctx, int n, const uint64_t* starts, int64_t count, uint64_t max_evals_per_start,
                     uint64_t* final_states, int64_t* final_energies, int64_t* eval_counts);

/* MinimumFamily (classical.cuh:60-67) */
typedef struct {
    int64_t id;                   /* 1-based rank after sorting */
    uint64_t canonical_state;     /* canonical representative of the symmetry orbit */
    uint64_t sample_state;        /* the first converged state found in it */
    int64_t energy;
    uint64_t times_visited;       /* descents that ended in the orbit */
    uint64_t evals_to_first_find; /* evaluations spent until its first discovery */
} bk_minimum_family;

/* collect_distinct_minima_gpu (classical.cuh:861-950): batches of batch_size
 * starts drawn from mt19937_64(seed) (random_state, :602-604), descended on the
 * GPU and grouped by canonical state, until target_families orbits, max_starts
 * starts or 4 batches without a new orbit; sorted by energy, then visits (then
 * canonical bits, which the reference leaves unspecified).  Writes up to cap
 * families to out, their number to *count (total_starts / total_evals may be
 * NULL). */
int bk_collect_minima(bk_ctx* ctx, int n, uint64_t seed, int64_t target_families, int64_t batch_size,
                      int64_t max_starts, bk_minimum_family* out, int64_t cap, int64_t* count, int64_t* total_starts,
                      int64_t* total_evals);

/* compute_d_min (labs_energy.hpp:175-207) for count states on the device:
 * out[i] = the smallest Hamming radius d <= max_d (all radii when max_d <= 0)
 * at which some state has E < targets[i] (targets NULL or targets[i] < 0:
 * E(states[i])); max_d + 1 (n when max_d <= 0) when there is none or the
 * target is at most E*(n).  The value of bk_d_min for every state.
 * 2 <= n <= 32, count <= 2^24. */
int bk_d_min_batch(bk_ctx* ctx, int n, const uint64_t* states, const int64_t* targets, int64_t count, int max_d,
                   int* out);

/* ---- exhaustive local-minima explorer (src/tools/exhaustive_minima.cpp) ----
 * run_exploration (:109-216): scans all 2^(n-1) states with the top spin up
 * (global negation gives the rest) for one-flip minima (no single flip lowers
 * E, :74-107), finds E* (the lowest energy of a minimum) and the next
 * 'levels' energies that hold minima (the trap levels; the reference keeps
 * 2), and lists every symmetry orbit of minima at those levels by its
 * canonical state with its exact d_min at the level energy (:200-203).  With
 * num_sd_starts > 0, steepest descents (bk_descent_batch) from the first
 * num_sd_starts draws of mt19937_64(sd_seed) & (2^n - 1) give each listed
 * orbit's sd_hit_fraction, the share of descents ending in it
 * (run_sd_benchmarks, :218-271; the reference's CLI uses seed 42 and 100000
 * starts).  3 <= n <= 32. */
#define BK_MINIMA_MAX_LEVELS 8
typedef struct {
    uint64_t canonical_state;  /* SpinState::canonical of the orbit */
    int64_t energy;            /* the level energy E0 */
    int level;                 /* 1-based trap level */
    int d_min;                 /* compute_d_min(canonical_state, E0) */
    double sd_hit_fraction;    /* hits / num_sd_starts (0 without descents) */
} bk_minima_orbit;

typedef struct {               /* ExhaustiveResult (:43-50) */
    int n;
    int64_t e_star;              /* -1 when no minimum exists */
    int64_t total_minima_found;  /* minima over all 2^n states: twice the half-space count (:151) */
    int64_t total_orbits_found;  /* distinct canonical orbits over all energies */
    int64_t num_energies;        /* energies that hold a minimum */
    int num_levels;              /* trap levels found (<= levels) */
    int64_t level_energy[BK_MINIMA_MAX_LEVELS];
    int64_t level_orbits[BK_MINIMA_MAX_LEVELS];
    int64_t orbits_written;      /* entries written to orbits: level by level, ascending canonical state */
    double search_time_sec;      /* device time of the scan */
} bk_exhaustive_result;

int bk_exhaustive_minima(bk_ctx* ctx, int n, int levels, uint64_t num_sd_starts, uint64_t sd_seed,
                         bk_exhaustive_result* res, bk_minima_orbit* orbits, int64_t cap);

/* save_to_json (exhaustive_minima.cpp:328-372): the explorer's JSON file, byte
 * for byte the reference writer's for the same values.  Host-only. */
int bk_export_exhaustive_json(const char* path, const bk_exhaustive_result* res, const bk_minima_orbit* orbits,
                              int64_t count);

/* ---- statevector held by the context ---- */
/* allocate 2^n amplitudes in 'precision' and set |start_bits> (sampler.cuh:428-438) */
int bk_state_init(bk_ctx* ctx, int n, int precision, uint64_t start_bits);
/* copy a host statevector in (interleaved re,im; float for C64, double for C128) */
int bk_state_upload(bk_ctx* ctx, int n, int precision, const void* host_amps);
/* copy the statevector out (any pending closing Strang half-phase is applied first) */
int bk_state_download(bk_ctx* ctx, void* host_amps);
/* device pointer and stream of the held state (for zero-copy interop) */
int bk_state_device(bk_ctx* ctx, void** d_amps, void** hip_stream, int* n, int* precision);

/* ---- evolution (sampler.cuh:509-619 without the sampling) ---- */
typedef struct {
    int n;              /* 1..32 */
    int precision;      /* BK_C64 | BK_C128 */
    int order;          /* BK_LIE | BK_STRANG */
    double dt;          /* Trotter step */
    int steps;          /* number of steps */
    double g;           /* constant field (used when g_steps == NULL) */
    const double* g_steps; /* optional per-step field, length 'steps' (host memory) */
    uint64_t start_bits;   /* |s0>, also the reference for <d_H> */
    int64_t e0;            /* P_< threshold; < 0 -> E(start_bits) */
    int init_from_start;   /* 1: state <- |start_bits>; 0: continue from the held state */
    int measure_every;     /* record observables every k steps (0 -> only after the last) */
} bk_evolve_params;

/* Observables recorded by bk_evolve: slot i holds the state after i*measure_every
 * steps (the last slot the final state).  Caller-owned host arrays of 'capacity'
 * entries; 'count' is set to the number written. */
typedef struct {
    double* p_less;  /* sum_{E(s)<E0} |psi_s|^2          (sampler.cuh:149-195) */
    double* dh;      /* sum popcount(s ^ s0) |psi_s|^2 */
    double* norm;    /* sum |psi_s|^2 (drift diagnostic) */
    int64_t capacity;
    int64_t count;
    int64_t e0;      /* threshold used */
    double p_blind;  /* #{E < E0} / 2^n  (sampler.cuh:198-227, 383-399); -1 when
                        init_from_start == 0 (not recomputed for a continuation) */
} bk_series;

int bk_evolve(bk_ctx* ctx, const bk_evolve_params* p, bk_series* out);

/* ---- adjoint gradient of the evolution (no reference counterpart: the
 * reference has no differentiable path) ----
 * Objective L = w_pless * P_<(T) + w_dh * <d_H>(T) of the state after
 * ev.steps Trotter steps from |ev.start_bits> with field g_k in step k
 * (ev.g_steps, or the constant ev.g).
 * bk_adjoint_forward runs the steps like bk_evolve (init_from_start must be 1;
 * measure_every is ignored), keeps the state after every checkpoint_every-th
 * step on the device (the tape: floor((steps-1)/checkpoint_every) states) and
 * writes {P_<, <d_H>, norm} of the final state to obs3 (may be NULL).  The
 * final state stays held.
 * bk_adjoint_backward steps both the state and the adjoint state backwards
 * over the tape of the last bk_adjoint_forward on this context (BK_E_NOSTATE
 * if the held state changed since): grad_steps[k] = dL/dg_k, k = 0..steps-1
 * (host, caller-owned).  At every checkpoint the reversed state is replaced by
 * the taped one; drift (may be NULL) receives the largest sum |reversed -
 * taped|^2 seen there.  The held state is left unchanged. */
typedef struct {
    bk_evolve_params ev;
    int checkpoint_every; /* >= 1 */
} bk_adjoint_params;

int bk_adjoint_forward(bk_ctx* ctx, const bk_adjoint_params* p, double* obs3);
int bk_adjoint_backward(bk_ctx* ctx, double w_pless, double w_dh, double* grad_steps, double* drift);
/* forward then backward */
int bk_adjoint_grad(bk_ctx* ctx, const bk_adjoint_params* p, double w_pless, double w_dh, double* obs3,
                    double* grad_steps, double* drift);

/* ---- batched evolution of many traps (no reference counterpart: the reference
 * runs one trap per ISampler::sample call, sampler.cuh:412-707) ----
 * Every instance starts from |start_bits> and takes 'steps' Trotter steps of
 * its own field g with the common dt / order / precision; out[i] receives
 * P_<, <d_H>, norm of instance i after the last step (input order).  The
 * states live in their own device arena (the context's held state is not
 * touched).  Each instance runs the same schedule, kernels and arithmetic as
 * a bk_evolve of that instance alone (fused runs where bk_evolve fuses them,
 * under the same tuning), so results are bit-identical to serial runs.  n must
 * be at least the LOW tile (13 for BK_C64, 12 for BK_C128). */
typedef struct {
    int n;
    uint64_t start_bits;
    int64_t e0;  /* P_< threshold; < 0 -> E(start_bits) */
    double g;
} bk_instance;

typedef struct {
    int precision;  /* BK_C64 | BK_C128 */
    int order;      /* BK_LIE | BK_STRANG */
    double dt;
    int steps;      /* >= 1 */
    int use_graph;  /* 1: capture the whole run (all steps, all instances) as one hipGraph and launch it once */
    int autotune;   /* 1: choose grid / row bits / cache-block size by timing short trial runs
                       (the choice is cached per workload shape on the context) */
    int64_t block_bytes; /* autotune = 0: cache block size (0 = every state advances each step) */
    int out_of_place;    /* autotune = 0, block_bytes > 0: a block's passes ping-pong between its
                            states and a second block-sized buffer instead of updating in place */
} bk_batch_params;

typedef struct {
    double p_less, dh, norm;
} bk_batch_result;

typedef struct {
    double device_ms;     /* stream time of the evolution (from the first step to the results) */
    int grid;             /* workgroups per sweep launch */
    int row_bits;         /* row bits of the high-qubit sweeps */
    int64_t block_bytes;  /* cache block: states stepped together through all steps (0 = all each step) */
    int64_t launches;     /* kernel launches of the run (graph nodes with use_graph) */
    int tuned;            /* 1: the parameters come from the autotuner */
    int out_of_place;     /* 1: blocks ran out of place (ping-pong through a block-sized buffer) */
} bk_batch_stats;

int bk_batch_run(bk_ctx* ctx, const bk_instance* inst, int count, const bk_batch_params* p, bk_batch_result* out,
                 bk_batch_stats* stats);

/* P_<, <d_H>, norm of the held state (out[3]) */
int bk_observables(bk_ctx* ctx, uint64_t s0, int64_t e0, double* out3);
/* <psi_a|psi_b> = sum_s conj(a_s) b_s (out2 = re, im) of the states two
 * contexts on one device hold (same n, either precision), in FP64 with a
 * fixed summation order.  No counterpart in the reference; the state-overlap
 * error 1 - |<psi64|psi128>| of a complex64 run against complex128. */
int bk_state_overlap(bk_ctx* a, bk_ctx* b, double* out2);
/* exact #{ s : E(s) < e0 } over 2^n states, on the device */
int bk_count_below(bk_ctx* ctx, int n, int64_t e0, uint64_t* count);

/* ---- projective sampling (custatevecSamplerPreprocess + Sample with
 * CUSTATEVEC_SAMPLER_OUTPUT_RANDNUM_ORDER, sampler.cuh:564-573):
 * out_idx[i] = min{ s : cdf(s) > randnums[i] * cdf(2^n - 1) } on the held state. */
int bk_sample(bk_ctx* ctx, const double* randnums, int64_t shots, int64_t* out_idx);

/* No counterpart in the reference: bk_sample for callers that cannot allocate
 * device memory (work under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Same kernels and results as bk_sample, on a caller-held
 * device state d_amps (2^n amplitudes in 'precision', interleaved re,im), with
 * device random numbers d_randnums[shots] and device output d_out_idx[shots];
 * launched on hip_stream (NULL: the null stream), which must belong to the
 * current device, and not synchronised.  No device memory is allocated: the
 * block sums live in a library-held static array of 16 MiB per device (n <= 32;
 * larger n, shots > 2^30 or null pointers return BK_E_INVALID).  Every call on
 * one device shares that array, so calls that overlap on one device race. */
int bk_sample_noalloc(int n, int precision, const void* d_amps, const double* d_randnums, int64_t shots,
                      int64_t* d_out_idx, void* hip_stream);

/* ---- matrix-free H.v (diagonalizer.cuh:60-77; the MatVecFunc of :400-406) ----
 * y = (E/N or E) * x - g * sum_j X_j x on DEVICE pointers; is_complex selects
 * complex128 (interleaved) over FP64 real.  stream may be NULL (context stream). */
int bk_hv(bk_ctx* ctx, int n, double g, int normalize_by_n, int is_complex, const double* d_x,
          double* d_y, void* hip_stream);

/* ---- Lanczos (diagonalizer.cuh:420-554: v0 from mt19937_64(42) uniform(-0.5,0.5),
 * no reorthogonalisation, breakdown tol, Ritz vectors) on the full 2^n space. ----
 * evals: k_pairs ascending; evecs (host, k_pairs*2^n, may be NULL). Returns the
 * number of pairs (>= 0) or an error code. */
int bk_lanczos(bk_ctx* ctx, int n, double g, int normalize_by_n, int max_krylov_dim, double tol,
               int k_pairs, double* evals, double* evecs);

/* ---- the subspace ("sampled-Krylov") solvers and dense ED of diagonalizer.cuh ----
 * H restricted to the states of basis (host array of dim distinct states < 2^n,
 * any order, n <= 64): E(s)/n (or E(s)) on the diagonal, -g between states one
 * spin flip apart.  evals: k_pairs ascending; evecs (host, k_pairs * dim, may be
 * NULL): each eigenvector in the order of basis.  Return the number of pairs.
 *
 * Diagonalizer::solve(basis, g, n, cfg), DiagMethod::Lanczos
 * (solve_subspace_lanczos, diagonalizer.cuh:299-372): the Lanczos of bk_lanczos
 * over H.v of the sorted basis (subspace_matvec_kernel_d, :79-114). */
int bk_subspace_lanczos(bk_ctx* ctx, int n, double g, int normalize_by_n, const uint64_t* basis, int64_t dim,
                        int max_krylov_dim, double tol, int k_pairs, double* evals, double* evecs);
/* DiagMethod::ExactDense on a basis (solve_subspace_dense, diagonalizer.cuh:188-245):
 * all eigenpairs of the dense H on the device (dim <= 32768). */
int bk_subspace_dense(bk_ctx* ctx, int n, double g, int normalize_by_n, const uint64_t* basis, int64_t dim,
                      int k_pairs, double* evals, double* evecs);
/* DiagMethod::ExactDense on the full 2^n space (solve_fullspace_dense,
 * diagonalizer.cuh:250-294): n <= 14 as in the reference; evecs k_pairs * 2^n. */
int bk_fullspace_dense(bk_ctx* ctx, int n, double g, int normalize_by_n, int k_pairs, double* evals,
                       double* evecs);

/* ---- Krylov propagator: held complex128 state <- exp(-i H tau) state, m-step
 * Lanczos with full reorthogonalisation, restarted every tau_restart. ---- */
int bk_krylov_propagate(bk_ctx* ctx, double g, int normalize_by_n, double tau, double tau_restart,
                        int m);

/* ---- ISampler::sample as one call (sampler.cuh:412-707) ---- */
typedef struct {            /* SamplerConfig, sampler.cuh:235-260 */
    int n;
    double g;
    double dt;
    double t_max;
    int num_sample_points;
    int shots_per_point;
    uint64_t seed;
    int num_sampling_seeds;
    uint64_t seed_stride;
    double t_start;
    double dt_sample;
    int track_p_less;
    int verbose;
    int record_trajectory;
    double dt_trajectory;
    /* engine extensions (defaults: BK_C128, BK_LIE reproduce the reference) */
    int precision;
    int order;
} bk_sampler_config;

/* the values of SamplerConfig's member initialisers (sampler.cuh:235-260) */
void bk_sampler_config_defaults(bk_sampler_config* cfg);

typedef struct {            /* SeedSamplingStats, sampler.cuh:262-270 */
    uint64_t seed;
    uint64_t total_shots_collected;
    uint64_t improved_shots;
    double success_rate;
    double per_eval_rate;
    double mean_hamming_flips;
    int64_t best_energy;
} bk_seed_stats;

typedef struct {            /* SamplingResult, sampler.cuh:272-297; library-owned arrays */
    uint64_t* unique_basis;
    int64_t unique_basis_len;
    uint64_t total_shots_collected;
    uint64_t improved_shots;
    double success_rate;
    double per_eval_rate;
    uint64_t total_evals;
    double mean_hamming_flips;
    bk_seed_stats* seed_stats;
    int64_t seed_stats_len;
    double* time_points;
    double* p_less_history;
    double* dh_history;
    int64_t history_len;
    double window_avg_dh;
    double initial_rise_rate_avg;
    double initial_rise_rate_max;
    int d_min;
    int64_t start_energy;
    double p_blind;
    double peak_p_less;
    double peak_time;
    /* engine extension: sampled indices in draw order (point-major, seed, shot) */
    int64_t* samples;
    int64_t samples_len;
} bk_sampling_result;

/* Throws nothing; n > BK_MAX_N returns BK_E_INVALID (the reference throws for
 * n > 30, sampler.cuh:416-418, this engine runs to 32). */
int bk_sampler_run(bk_ctx* ctx, const bk_sampler_config* cfg, uint64_t start_bits,
                   bk_sampling_result** out);
void bk_result_free(bk_sampling_result* r);

/* ---- how the first sweep of a step obtains E(s) (the reference caches E as FP64,
 * 8 B per amplitude, sampler.cuh:54-70, 320-338) ----
 * BK_ENERGY_CACHE16 (default): an exact E(s) cache filled once per n and read by
 *   the sweep, 1.5 B/amplitude (12-bit codes E >> 2; E mod 4 is fixed by n; the
 *   name is kept from its first, 16-bit layout);
 * BK_ENERGY_ONTHEFLY: recomputed every step from popcounts (no extra memory). */
#define BK_ENERGY_ONTHEFLY 0
#define BK_ENERGY_CACHE16 1
int bk_set_energy_mode(bk_ctx* ctx, int mode);

/* ---- sweep tuning knobs (no reference counterpart; used by the autotuner) ----
 * "grid"     : workgroups per sweep launch (persistent loop over tiles)
 * "row_bits" : contiguous low bits carried by the high-qubit sweeps (3..8)
 * "pingpong" : 1 = out-of-place sweeps between two state buffers (2x memory)
 * "page_bits": > 0 deals the qubits >= page_bits (whose partner amplitudes lie
 *              in other 2^page_bits-amplitude pages) evenly over the high-qubit
 *              sweeps instead of giving them all to the last one (0 = contiguous)
 * "prefetch_low" / "prefetch_high" : 1 = issue the next tile's loads before the
 *              current tile's stores (software pipelining) in that sweep kind
 * "fused"    : 1 (default) = a run of K steps takes one pass per step when the
 *              high qubits form one sweep group (complex64 n = 15..23, complex128
 *              13..21) and 2K+1 passes when they form two (complex64 24..32,
 *              complex128 22..32): the pass that applies step k's phase also
 *              rotates its qubits for step k+1; 0 = a step at a time
 * "wide_low" : 1 (default) = a complex64 fused run whose two high groups would
 *              need 16-amplitude rows (n = 32) takes a 14-qubit LOW tile instead,
 *              so both groups keep 32-amplitude rows; 0 = the 13-qubit tile
 * "narrow_phase" : 1 (default) = in a complex64 fused run with two high groups,
 *              the phase-carrying group moves on 2^13-amplitude tiles (two
 *              workgroups per CU) when its rows stay at least 64 amplitudes
 *              long (n = 26..28; smaller states stay in the last-level cache);
 *              0 = 2^14-amplitude tiles.  The state is the
 *              same bit for bit either way.
 * "lds_tab13" : phase passes on 2^13-amplitude tiles (complex64: the LOW
 *              tile and the narrow phase-carrying group; complex128: the high
 *              groups) read the phase table from LDS (complex64 n <= 29;
 *              complex128 the whole table to n = 29, its first 2032 entries
 *              beyond; default on at complex64 n = 26..27 and complex128
 *              n >= 22) instead of global memory; results are the same bit for
 *              bit.
 * "nt_plain" : 1 = in a complex64 fused run, the plain high-group pass loads
 *              the state with nontemporal loads (default on from n = 27, where
 *              it shortens the step by up to 1.8 %); results are the same bit
 *              for bit.
 * "grid_bphase" / "grid_lphase" : workgroups of a fused run's one-state phase
 *              passes over the high group / the LOW tile (0 = "grid"; default
 *              256 / 16384 for complex64 from n = 30, else 0).
 * "rev_grid_low" / "rev_grid_bphase" : workgroups of a fused reverse run's
 *              passes over the LOW tile / the phase-carrying group (0 = "grid",
 *              at most one workgroup per 2^14 amplitudes for the group;
 *              rev_grid_low defaults to 512 for complex64 at n = 26, else 0).
 * "fold_finish" : 1 (default) = in a fused run, the pass after a measured
 *              phase pass sums that pass's partials in its first workgroup (and
 *              the reverse run's A pass the previous step's gradient partials);
 *              0 = a finishing kernel after every measured pass.  Results are
 *              the same bit for bit.
 * A key keeps its value for the context until set again; every other key takes
 * the defaults of each new state's n and precision. */
int bk_set_tuning(bk_ctx* ctx, const char* key, int value);

/* The HIGH sweeps one Trotter step runs for (n, precision, row_bits, page_bits):
 * writes {h1, a1, h2, a2, R} per sweep into out (5*cap ints) and returns the
 * sweep count.  Qubits [0, 13) (c64) / [0, 12) (c128) belong to the final LOW
 * sweep; a HIGH sweep rotates qubits [h1, h1+a1) and [h2, h2+a2) on tiles of
 * 2^R-amplitude rows.  Host-only (no device needed). */
int bk_sweep_plan(int n, int precision, int row_bits, int page_bits, int* out, int cap);

/* ---- instrumentation (no reference counterpart; the reference only wall-times
 * the whole sample() call, connector.cuh:627, 832-833) ----
 * When enabled, every sweep launch of bk_evolve / bk_sampler_run is bracketed
 * by HIP events on the context stream.  bk_timing_read synchronises, returns
 * one row per kernel name (launch count, summed device ms, algorithmic HBM
 * bytes per launch) and clears the accumulated rows. */
typedef struct {
    char name[48];
    int64_t launches;
    double total_ms;
    double bytes_per_launch;
} bk_kernel_timing;

int bk_timing_enable(bk_ctx* ctx, int enable);
/* Plain float4 copy of 'bytes' (read + write), reps times: returns GB/s of
 * read+write traffic.  The practical HBM ceiling for the sweeps' access mix,
 * and the known-size kernel that calibrates FETCH_SIZE / WRITE_SIZE. */
int bk_copy_probe(bk_ctx* ctx, uint64_t bytes, int reps, double* gbs);
/* The same measurement in another access pattern (no counterpart in the
 * reference; a measurement helper).  pattern 0: the grid-stride copy of
 * bk_copy_probe.  1: one 4 KB chunk per 256-thread workgroup, a grid of
 * bytes/4096 workgroups, nontemporal loads and stores: the fastest
 * read+write stream of the part.  2: one 128 KB chunk per 512-thread
 * workgroup, the footprint of a complex64 14-qubit sweep tile.  bytes must
 * be a positive multiple of 16, 4096 or 131072 respectively. */
int bk_copy_probe_pattern(bk_ctx* ctx, uint64_t bytes, int reps, int pattern, double* gbs);
int bk_timing_read(bk_ctx* ctx, bk_kernel_timing* out, int cap, int* count);

/* ---- the benchmark connector's output files (connector.cuh:839-1155) ----
 * One record per (trap, g): what BenchmarkConnector::run keeps of a trap and
 * its sampler result (SingleTrapBenchmark, connector.cuh:26-117, quantum part).
 * Host-only; the files are byte-for-byte the reference writer's for the same
 * values (include/bk/wire.hpp holds the C++ writers these call). */
typedef struct bk_trap_record {
    int n;
    double g;
    int64_t trap_rank;                 /* 1-based rank of the trap */
    uint64_t trap_bits;                /* trap state s0 */
    uint64_t canonical_bits;           /* canonical representative of its symmetry orbit */
    int64_t trap_energy;               /* E0 */
    int d_min;                         /* 0: taken from the result */
    double quantum_time_sec;           /* wall time of the sampler call */
    /* window of window_avg_p_less (connector.cuh:816-830): [t_start, t_start +
     * (window_points - 1) window_dt], or [t_start, t_max] when window_points < 2
     * or window_dt <= 0 */
    double window_t_start;
    int window_points;
    double window_dt;
    double t_max;
    const bk_sampling_result* result;  /* the sampler's result (bk_sampler_run); null: zeros */
} bk_trap_record;

/* export_results_csv (connector.cuh:839-904): header + one row per record */
int bk_export_results_csv(const char* path, const bk_trap_record* recs, int count);
/* export_time_series_csv (connector.cuh:906-933): one row per recorded time */
int bk_export_time_series_csv(const char* path, const bk_trap_record* recs, int count);
/* export_structured_json (connector.cuh:935-1155): records grouped by
 * (n, trap_rank); n_override = 0 takes N from the first record */
int bk_export_structured_json(const char* path, const bk_trap_record* recs, int count, int n_override);

#ifdef __cplusplus
}
#endif
#endif
