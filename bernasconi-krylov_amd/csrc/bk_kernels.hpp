// bk_kernels.hpp -- host-side interface of the gfx950 kernels (launchers).
//
// Device memory layout: the statevector is one contiguous array indexed by
// the spin bit pattern s (bit i <-> spin i, bit 1 <-> spin -1;
// labs_energy.hpp:20-23), amplitudes interleaved (re, im) as float2 (c64) or
// double2 (c128) -- the layout of sampler.cuh:426 in either precision.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bk_sweeps.hpp"

namespace bk {

enum Precision : int { kC64 = 1, kC128 = 2 };

// One Trotter step's kernel sequence for a statevector of 2^n amplitudes.
struct StepSpec {
    int n;
    int precision;
    // mixer: per-qubit [[c, i s],[i s, c]], c = cos(g dt), s = sin(g dt)
    double gdt;
    // diagonal phase applied after the step's rotations (the next step's
    // opening phase): exp(-i dtp E(s)/n)
    double dtp;
    const void* phase_tab;  // device table indexed by E >> 2 (float2 or double2)
    // observables of the stepped state reduced into slot
    double* series_slot;    // device pointer to 3 doubles {P_<, <d_H>, norm} or null
    uint32_t s0;
    uint32_t e0;
    // optional energy cache (12-bit codes E(s) >> 2, packed eight to three
    // words: launch_energy_fill); null -> energies are computed on the fly
    // with popcounts in the sweep
    const uint32_t* ecache = nullptr;
    // Optional error-feedback phase table (fb_carry non-null): before the step's
    // phase multiply, entry m of phase_tab (writable, fb_count entries) is set
    // to round(sd * fb_master[m] * fb_carry[m]) and fb_carry[m] to the exact
    // quotient target / rounded, in FP64; sd is the step's norm-exact mixer
    // scale.  The rounding residual of every entry is carried into the next
    // step, so the product of the applied factors follows the exact product to
    // one rounding instead of drifting by one rounding per step.  Tiled steps
    // (n >= the LOW tile bits) in tan form need it: the sweeps take the
    // step's c^n from the table (hipErrorInvalidValue without one).
    const double* fb_master = nullptr;  // double2 exp(-i dtp E_m/n) per entry
    double* fb_carry = nullptr;         // double2 per entry
    int fb_count = 0;
};

// Optional per-launch timing hook (HIP events recorded on the launch stream).
struct KernelTimer {
    virtual ~KernelTimer() = default;
    virtual void begin(hipStream_t s) = 0;
    // name of the kernel just launched and its algorithmic HBM bytes
    virtual void end(hipStream_t s, const char* name, double bytes) = 0;
};

struct SweepTuning {
    int tile_bits;   // TB of the LOW sweep (13 for c64, 12 for c128 by default)
    int row_bits;    // R of HIGH sweeps
    int grid;        // workgroups per sweep launch (persistent loop over tiles)
    bool pingpong;   // out-of-place sweeps between two state buffers
    bool prefetch_low;   // software-pipelined tile loads in the LOW sweep
    bool prefetch_high;  // ... and in the HIGH sweeps
    int page_bits;       // > 0: deal qubits >= page_bits evenly over the HIGH sweeps
    bool fused = true;   // runs of steps use the two-pass schedule when the plan allows
    bool wide_low = true;  // complex64 fused runs may take a 14-qubit LOW tile (n = 32: full rows in both high groups)
    // complex64 fused runs with two high groups: the phase-carrying group B
    // moves on 2^13-amplitude tiles (one row bit fewer, 256 threads, 64 KB of
    // LDS, phase table in global memory), two workgroups per CU, when its rows
    // keep at least 64 amplitudes (default_tuning: on)
    bool narrow_phase = false;
    // fused runs: the pass after a measured phase pass finishes its partials
    // (false: a k_finish3 kernel after each measured pass)
    bool fold_finish = true;
    // complex64 phase passes on 2^13-amplitude tiles take the phase table
    // into LDS for n <= 29 (it fits beside the 64 KB exchange at two
    // workgroups per CU; default_tuning: n = 26..27)
    bool lds_tab13 = false;
    // complex64 fused runs: the plain high-group pass loads nontemporally
    // (default_tuning: n >= 27)
    bool nt_plain = false;
    // workgroups of a fused run's one-state phase passes over the high group
    // and over the LOW tile (0: grid)
    int grid_bphase = 0;
    int grid_lphase = 0;
    // workgroups of a fused reverse run's passes over the LOW tile and over
    // the phase-carrying group (0: grid)
    int rev_grid_low = 0;
    int rev_grid_bphase = 0;
    KernelTimer* timer = nullptr;
};

SweepTuning default_tuning(int n, int precision);

// scratch for reductions (partials): bytes needed for a given grid
size_t partials_bytes(int grid);
// the largest workgroup count any pass of these settings launches with
int max_grid(const SweepTuning& t);

// Launch one Trotter step (phase + all rotations) on 'stream'.  With psi_alt
// non-null the sweeps run out of place, alternating psi -> psi_alt -> psi ...;
// *result_in_alt tells which buffer holds the stepped state.
hipError_t launch_step(void* psi, void* psi_alt, int* result_in_alt, const StepSpec& spec,
                       const SweepTuning& tun, double* partials, hipStream_t stream);

// ---- a run of Trotter steps ----
// A phase-carrying pass rotates its tile's qubits for step k, applies P(k)
// (and measures), then rotates them again for step k+1; the X_q of one step
// commute, so only P has to sit between the steps.  With the high qubits in
// one sweep group H (complex64 n = 15..23, complex128 n = 13..21) a run of K
// steps takes K + 1 passes:
//   H(1), then for k = 1..K a phase pass on LOW (odd k) or H (even k);
// with two groups A, B (complex64 n = 24..32, complex128 n = 22..32) it takes
// 2K + 1 passes:
//   LOW(1), then for k = 1..K: A(k) and a phase pass on B (odd k) / LOW (even k);
// the last phase pass stops after P(K).  Other n (and the (c, s) butterfly)
// run step by step (launch_step).  The error-feedback table and the per-step
// observables are passed as in StepSpec.
struct RunSpec {
    int n;
    int precision;
    int steps;
    const double* gdt;       // host: g dt of each step
    double* const* slots;    // host: device slot for each step's observables (entries may be null; array may be null)
    const void* phase_tab;   // error-feedback step table (see StepSpec)
    // optional second table buffer (fb_count entries): the phase pass of step k
    // then computes step k+1's table into the other buffer, steps alternating
    // between phase_tab and phase_tab2, instead of a table kernel per step
    void* phase_tab2 = nullptr;
    const double* fb_master;
    double* fb_carry;
    int fb_count;
    uint32_t s0, e0;
    const uint32_t* ecache;       // LOW-layout E(s) cache, filled for low_bits
    const uint32_t* ecache_high;  // E(s) in the phase-carrying high group's layout (launch_energy_fill_high);
                                  // null: step by step
    int low_bits = 0;             // LOW tile of the fused run (fused_low_bits; 0: the default tile)
};
hipError_t launch_run(void* psi, const RunSpec& rs, const SweepTuning& tun, double* partials, hipStream_t stream);
// the high-qubit groups of a fused run for n: one group (one pass a step),
// two (two passes a step) or none (runs go step by step); the last group is
// the phase-carrying one
// low_bits: the LOW tile the groups are planned around (0: the default tile;
// fused_low_bits gives the forward run's choice)
std::vector<HighSweep> fused_sweeps(int n, int precision, const SweepTuning& tun, int low_bits = 0);
int fused_low_bits(int n, int precision, const SweepTuning& tun);
hipError_t launch_energy_fill_high(uint32_t* out, int n, int precision, const HighSweep& hs, hipStream_t stream);

// ---- reverse pass of the adjoint gradient ----
// One reverse step of phi (reversed forward state) and chi (adjoint state),
// both in place: per-qubit rotation exp(i spec.gdt X) (spec.gdt = -g dt undoes
// a step of field g), then exp(-i spec.dtp E/n) from spec.phase_tab (and
// spec.ecache if set).  *grad_slot <- Im <chi| sum_q X_q |phi> of the states
// as they enter.  gpart holds reverse_partials_count() doubles.
hipError_t launch_reverse_step(void* phi, void* chi, const StepSpec& spec, const SweepTuning& tun, double* gpart,
                               double* grad_slot, hipStream_t stream);
size_t reverse_partials_count(int n, int precision, const SweepTuning& tun);
// K reverse steps in one fused run (the two-state form of launch_run's
// schedule, K + 1 or 2K + 1 passes): step i rotates by exp(i gdt[i] X), then
// applies phase_tab; *grad_slots[i] <- Im <chi| sum_q X_q |phi> as step i
// begins.  The states leave the run in the plain frame, so a run can stop at a
// checkpoint.  Without ecache_high, or for n that launch_run also takes step
// by step, the steps run one by one (launch_reverse_step).
struct RevRunSpec {
    int n;
    int precision;
    int steps;
    const double* gdt;           // host: per reverse step
    double* const* grad_slots;   // host: device slot per reverse step
    const void* phase_tab;       // exp(-i dtp E/n) per E >> 2
    const uint32_t* ecache;      // LOW-layout E(s)
    const uint32_t* ecache_high; // E(s) in the phase-carrying high group's layout
    // complex64: an error-feedback table for dtp (as StepSpec's); when given
    // and every step allows it, the runs use the tan butterfly (the table
    // carries each step's c^n) instead of (c, s) with phase_tab
    void* fb_applied = nullptr;
    void* fb_applied2 = nullptr;  // optional second buffer (as RunSpec::phase_tab2)
    const double* fb_master = nullptr;
    double* fb_carry = nullptr;
    int fb_count = 0;
};
hipError_t launch_reverse_run(void* phi, void* chi, const RevRunSpec& rs, const SweepTuning& tun, double* gpart,
                              hipStream_t stream);
// phi = tab[E] * src; chi = (wp [E < e0] + wd popcount(s ^ s0)) * phi
hipError_t launch_adj_init(const void* src, void* phi, void* chi, int n, int precision, const void* tab, uint32_t s0,
                           uint32_t e0, double wp, double wd, hipStream_t stream);
// dst = tab[E] * src; *out = sum |new dst - old dst|^2 (partials: 4096 doubles)
hipError_t launch_phase_from(void* dst, const void* src, int n, int precision, const void* tab, double* partials,
                             double* out, hipStream_t stream);

// ---- batched evolution: many states of one n, one launch per sweep ----
// Mixer coefficients as step_impl uses them, and the norm-exact scale that the
// error-feedback table folds in (sd).
struct MixInfo {
    bool tan_form;
    double bc, bs, sd;
};
MixInfo mix_info(int n, double gdt, int precision);
int low_tile_bits(int precision);  // TB of the LOW sweep: tiles per state = 2^(n - TB)

struct BatchGroup {
    int n;
    int precision;
    int nstates;
    void* psi;                // nstates * 2^n amplitudes, state i at i * 2^n
    void* psi_out;            // launch_batch_pass: output states in the same layout (null: in place)
    const void* tab;          // error-feedback tables, tab_stride entries per state
    int tab_stride;
    const uint32_t* ecache;   // E(s) cache of n (shared by the states)
    const double* bc;         // per state (device)
    const double* bs;         // per state (device)
    const uint32_t* s0;       // per state (device)
    const uint32_t* e0;       // per state (device)
    bool tan_form;            // every state of the group uses the same butterfly form
    // phase passes: when tab_next is non-null the pass also sets the next
    // step's tables (tab_stride entries per state, the layout of tab) from
    // fb_master (shared), the per-state carries fb_carry and scales fb_sd
    void* tab_next = nullptr;
    const double* fb_master = nullptr;
    double* fb_carry = nullptr;
    const double* fb_sd = nullptr;
};
// One Trotter step of every state in g: HIGH sweeps then the LOW sweep with the
// phase; with partials non-null the LOW sweep also writes 3 partials per tile
// (finish them with launch_batch_finish).  Same arithmetic as launch_step on
// each state alone.
hipError_t launch_batch_step(const BatchGroup& g, const SweepTuning& tun, double* partials, hipStream_t stream);
// One pass of a fused batched run (the schedule of launch_run, applied to every
// state of g; a state's field is the same in every step): LOW or high group
// without the phase, or a phase-carrying pass (tan butterflies) with optional
// measurement (3 partials per tile) and rotations for the next step.
struct BatchPass {
    enum Kind { kLowPlain, kHighPlain, kLowPhase, kHighPhase } kind;
    HighSweep hs;                  // high passes
    const uint32_t* ecache_high;   // kHighPhase: E(s) in hs's layout (launch_energy_fill_high)
    bool measure, post;
};
hipError_t launch_batch_pass(const BatchGroup& g, const BatchPass& bp, const SweepTuning& tun, double* partials,
                             hipStream_t stream);
// k_phase_feedback for nstates tables (sd per state); master shared
hipError_t launch_batch_feedback(const double* master, double* carry, void* applied, int count, int nstates,
                                 const double* sd, int precision, hipStream_t stream);
hipError_t launch_batch_finish(const double* partials, int tiles_per_state, int nstates, double* out,
                               hipStream_t stream);
// state i (zeroed beforehand) gets amplitude re_i + i im_i at s0_i
hipError_t launch_batch_init(void* psi, int n, int nstates, int precision, const uint32_t* s0, const double* re,
                             const double* im, hipStream_t stream);

// Read-only measurement of the state into series_slot (3 doubles).
hipError_t launch_measure(const void* psi, int n, int precision, uint32_t s0, uint32_t e0,
                          double* series_slot, double* partials, int grid, hipStream_t stream);

// Elementwise phase psi_s *= tab[E(s)]; with series_slot non-null the
// observables of the resulting state are reduced into it in the same pass.
hipError_t launch_phase(void* psi, int n, int precision, const void* phase_tab, uint32_t s0, uint32_t e0,
                        double* series_slot, double* partials, hipStream_t stream);

// psi = (re + i im) |s0>
hipError_t launch_init_basis(void* psi, int n, int precision, uint64_t s0, double re, double im,
                             hipStream_t stream);

// fb_carry[m] = 1 for m < count (start of a new error-feedback history)
hipError_t launch_carry_reset(double* carry, int count, hipStream_t stream);

// Bytes of an energy cache of n qubits: a 12-bit code m = E(s) >> 2 per state
// (E(s) = 4m + (n/2 mod 4) exactly; m <= 2604 for n <= 32), eight codes in
// three 32-bit words
inline size_t energy_cache_bytes(int n) { return (static_cast<size_t>(3) << n) / 2; }
// The energy codes of all 2^n states (n >= 3) in the LOW sweep's tile layout
// for 'precision' (the layout differs between c64 and c128)
// low_bits: the LOW tile whose layout the cache follows (0: the default tile)
hipError_t launch_energy_fill(uint32_t* out, int n, int precision, hipStream_t stream, int low_bits = 0);

// count of s with E(s) < e0 over all 2^n states (exact, integer)
hipError_t launch_count_below(int n, uint32_t e0, unsigned long long* d_count,
                              unsigned long long* partials, int grid, hipStream_t stream);

// ---- sampling (inverse CDF over |psi|^2) ----
// block sums of |psi|^2 (FP64) for blocks of 2^sample_block_bits amplitudes
constexpr int kSampleBlockBits = 12;
hipError_t launch_sample(const void* psi, int n, int precision, const double* d_rand, int shots,
                         int64_t* d_out, double* d_block_sums, double* d_block_prefix,
                         hipStream_t stream);
size_t sample_scratch_doubles(int n);

// <a|b> = sum_s conj(a_s) b_s over two 2^n states of any precisions, in FP64:
// partials[2 i], partials[2 i + 1] = re, im of workgroup i's contiguous chunk
hipError_t launch_overlap(const void* a, int prec_a, const void* b, int prec_b, int n, double* partials, int grid,
                          hipStream_t stream);

// ---- H.v (matrix-free, diagonalizer.cuh:60-77) ----
// y = (E/n or E) * x - g * sum_j X_j x ; real FP64 (complex=0) or complex128 (complex=1).
// ecache: E(s) as uint16 in state order (launch_energy_natural), required from
// n = 13 (complex128) / 14 (real) up.  scratch: hv_scratch_doubles(n, is_complex)
// doubles that let every high-qubit pass write one array instead of updating
// y (null: they update y).
hipError_t launch_hv(int n, double g, int normalize_by_n, int is_complex, const double* x,
                     double* y, hipStream_t stream, const uint16_t* ecache = nullptr, double* scratch = nullptr);
size_t hv_scratch_doubles(int n, int is_complex);
hipError_t launch_energy_natural(uint16_t* out, int n, hipStream_t stream);

// ---- BLAS-1 with deterministic reductions (FP64) ----
// out[0] = sum_i x_i * y_i over len elements (complex: sum conj(x_i) y_i -> out[0..1])
hipError_t launch_dot(int64_t len, int is_complex, const double* x, const double* y, double* out,
                      double* partials, hipStream_t stream);
// y += a * x  (complex a = (ar, ai) when is_complex)
hipError_t launch_axpy(int64_t len, int is_complex, double ar, double ai, const double* x,
                       double* y, hipStream_t stream);
// x *= a (real)
hipError_t launch_scal(int64_t len, double a, double* x, hipStream_t stream);
// y = sum_j c_j * V_j  (m basis vectors of length len, stride len; complex coefficients)
hipError_t launch_combine(int64_t len, int is_complex, int m, const double* V,
                          const double* coef, double* y, hipStream_t stream);
size_t blas_partials_doubles();
// One classical Gram-Schmidt sweep of complex128 w against m <= kMdMax basis
// vectors V_0..V_{m-1} (stride len): coef = V^H w (device, 2m doubles, fixed-order
// reductions), then w -= V coef.  partials: multidot_partials_doubles() doubles.
constexpr int kMdMax = 32;
hipError_t launch_gs_block(int64_t len, int m, const double* V, double* w, double* coef, double* partials,
                           hipStream_t stream);
size_t multidot_partials_doubles();
// Both sweeps of classical Gram-Schmidt (CGS2) of complex128 w against the
// m <= kMdMax vectors of V in three passes: c1 = V^H w; w -= V c1 with
// c2 = V^H w on the new w; next = w - V c2 with |next|^2.  coef: 4 kMdMax + 2
// doubles, c1 at [0, 2m), c2 at [2 kMdMax, 2 kMdMax + 2m), |next|^2 at
// [4 kMdMax].  next = nullptr stops after c2 (w then holds the first sweep's
// result).  next must not alias w or V's m vectors.
hipError_t launch_cgs2(int64_t len, int m, const double* V, double* w, double* next, double* coef, double* partials,
                       hipStream_t stream);

// ---- subspace Hamiltonian over a sorted basis (diagonalizer.cuh:79-114, 188-245) ----
// nbr[i * n + j] = index of basis[i] ^ 2^j in the sorted basis, or -1
hipError_t launch_sub_neighbors(const uint64_t* basis, int64_t dim, int n, int32_t* nbr, hipStream_t stream);
// y = diag .* x - g * (sum over neighbours)
hipError_t launch_sub_hv(const double* diag, const int32_t* nbr, int64_t dim, int n, double g, const double* x,
                         double* y, hipStream_t stream);
// H (column-major, zero on entry) = diag on the diagonal, -g between neighbours
hipError_t launch_sub_dense(const double* diag, const int32_t* nbr, int64_t dim, int n, double g, double* H,
                            hipStream_t stream);

// ---- trap supply (classical.cuh:459-535): single-flip steepest descent of
// count starts, one wave per start; max_evals = 0: no budget ----
hipError_t launch_descent(int n, int64_t count, const uint64_t* starts, uint64_t max_evals, uint64_t* fin, int64_t* fe,
                          int64_t* evc, hipStream_t stream);

// plain float4 copy of 'bytes' (multiple of 16): HBM ceiling probe
hipError_t launch_copy_probe(const void* src, void* dst, uint64_t bytes, hipStream_t stream);
// copy of 'bytes' in access pattern 0 (the grid-stride copy above), 1 (4 KB
// per workgroup, one-shot grid, nontemporal) or 2 (128 KB per workgroup, the
// footprint of a complex64 sweep tile); bytes a multiple of the pattern's chunk
hipError_t launch_copy_pattern(int pattern, const void* src, void* dst, uint64_t bytes, hipStream_t stream);
uint64_t copy_pattern_chunk(int pattern);

}  // namespace bk
