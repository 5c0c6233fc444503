// bk_kernels.hip -- gfx950 kernels for the Trotterised transverse-field evolution
// under H = E(s)/N - g sum_i sigma^x_i (sampler.cuh:54-227, 604-617) and the
// matrix-free H.v of diagonalizer.cuh:60-77.
//
// Hot path (one Trotter step, bandwidth-bound):
//   LOW sweep   : one HBM pass over contiguous 2^TB-amplitude tiles.  Per
//                 amplitude it computes E(s) with popcounts (labs_energy.hpp:
//                 121-132), accumulates P_<, <d_H> and the norm of the incoming
//                 state in FP64 (sampler.cuh:149-195), multiplies by the
//                 tabulated phase exp(-i dt E/N) (sampler.cuh:73-91), then
//                 applies the sigma_x rotations of qubits 0..TB-1 (registers +
//                 LDS exchanges).
//   HIGH sweeps : one HBM pass each, rotating the next TB-R qubits on tiles
//                 made of 2^R-amplitude contiguous rows.
// The per-qubit rotation [[c, i s],[i s, c]] (sampler.cuh:451-457) is applied
// as c * [[1, i t],[i t, 1]] with t = tan(g dt): every butterfly is 4 FMAs and
// the product of the n factors c is folded into the phase multiply.  When
// |cos(g dt)| < 0.5 the full (c, s) butterfly is used instead.
#include "bk_kernels.hpp"
#include "bk_plan.hpp"
#include "bk_sweeps.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <type_traits>
#include <utility>
#include <vector>

namespace bk {

template <typename V>
struct RealOf;
template <>
struct RealOf<float2> {
    using T = float;
};
template <>
struct RealOf<double2> {
    using T = double;
};

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

// labs_energy.hpp:121-132 on 32-bit words (n <= 32).
__device__ __forceinline__ uint32_t labs_energy_u32(uint32_t s, int n)
{
    uint32_t e = 0;
#pragma unroll 4
    for (int k = 1; k < n; ++k) {
        const uint32_t pairs = (s ^ (s >> k)) & (0xFFFFFFFFu >> (32 - (n - k)));
        const int c = (n - k) - 2 * static_cast<int>(__popc(pairs));
        e += static_cast<uint32_t>(c * c);
    }
    return e;
}

// Amplitude store of a sweep.  NT: a nontemporal store, which streams the
// tile back without keeping it in the caches; measured faster for passes
// without the phase and for LOW-tile passes over one large state, slower for
// the phase passes of high-qubit groups and for batched passes, whose small
// states are re-read from the last-level cache by the next pass.
template <bool NT, typename V>
__device__ __forceinline__ void store_amp(V* p, const V& v)
{
    if constexpr (NT) {
        using R = typename RealOf<V>::T;
        typedef R N __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(*reinterpret_cast<const N*>(&v), reinterpret_cast<N*>(p));
    } else {
        *p = v;
    }
}

// Amplitude load of a sweep.  NT: a nontemporal load (streamed through the
// caches without keeping the line).  Measured per pass kind (same box, same
// results): faster for the LOW passes (N=30 plain 3.32 -> 3.16 ms, phase pass
// with the table in global memory 3.645 -> 3.612 ms), the N=26 narrow phase
// pass with its table in LDS (0.221 -> 0.202 ms) and every complex128 pass
// (N=28 step 3.40 -> 3.37 ms); slower for the complex64 plain high-group pass
// (N=30 3.220 -> 3.230 ms, N=26 0.188 -> 0.196 ms) and the two-state passes
// (N=26 LOW phase 0.560 -> 0.575 ms), which keep cached loads.
template <bool NT, typename V>
__device__ __forceinline__ V load_amp(const V* p)
{
    if constexpr (NT) {
        using R = typename RealOf<V>::T;
        typedef R N __attribute__((ext_vector_type(2)));
        const N w = __builtin_nontemporal_load(reinterpret_cast<const N*>(p));
        V v;
        v.x = w.x;
        v.y = w.y;
        return v;
    } else {
        return *p;
    }
}

template <typename V, typename R>
__device__ __forceinline__ void bfly_tan(V& a, V& b, R t)
{
    const R ax = a.x, ay = a.y;
    a.x = fma(-t, b.y, ax);
    a.y = fma(t, b.x, ay);
    b.x = fma(-t, ay, b.x);
    b.y = fma(t, ax, b.y);
}

template <typename V, typename R>
__device__ __forceinline__ void bfly_cs(V& a, V& b, R c, R s)
{
    const R ax = a.x, ay = a.y;
    a.x = c * ax - s * b.y;
    a.y = c * ay + s * b.x;
    b.x = c * b.x - s * ay;
    b.y = c * b.y + s * ax;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Deterministic block reduction of NV doubles over the first NW waves
// (pairwise tree over the waves; later waves take part in the barrier only).
// 'scratch' must hold NW*NV doubles and be free for use.
template <int NV, int NW = 4>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* scratch, double* out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    if (lane == 0 && wave < NW)
#pragma unroll
        for (int i = 0; i < NV; ++i) scratch[wave * NV + i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double t[NW];
#pragma unroll
            for (int w = 0; w < NW; ++w) t[w] = scratch[w * NV + i];
#pragma unroll
            for (int h = 1; h < NW; h <<= 1)
#pragma unroll
                for (int w = 0; w + h < NW; w += 2 * h) t[w] += t[w + h];
            out[i] = t[0];
        }
    }
}

// The finishing sums load kFinishBatch strided partials per thread before
// adding them in index order: the loads overlap instead of one L2 round trip
// per partial (a 12288-partial gradient sum took 19 us as a dependent chain).
constexpr int kFinishBatch = 8;

// The first kThreads threads of the workgroup sum (any further threads take
// part in the barrier only), so a sweep workgroup finishing a previous pass's
// partials (SweepArgs::fin_out) gives the same bits as k_finish3.
__device__ __forceinline__ void finish3_body(const double* __restrict__ partials, int count, double* __restrict__ out,
                                             double* scratch)
{
    double acc[3] = {0.0, 0.0, 0.0};
    constexpr int bd = kThreads;
    for (int i = threadIdx.x < kThreads ? static_cast<int>(threadIdx.x) : count; i < count; i += kFinishBatch * bd) {
        double v[kFinishBatch][3];
#pragma unroll
        for (int k = 0; k < kFinishBatch; ++k) {
            const int ik = i + k * bd;
#pragma unroll
            for (int c = 0; c < 3; ++c) v[k][c] = ik < count ? partials[3 * ik + c] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kFinishBatch; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] += v[k][c];
    }
    block_reduce_store<3>(acc, scratch, out);
}

// 'count' doubles in a fixed order into *out by the first kThreads threads
// (as finish3_body: a sweep workgroup folding it gives k_finish_sum's bits)
__device__ __forceinline__ void finish_sum_body(const double* __restrict__ partials, int count, double* __restrict__ out,
                                                double* scratch)
{
    double acc[1] = {0.0};
    constexpr int bd = kThreads;
    for (int i = threadIdx.x < kThreads ? static_cast<int>(threadIdx.x) : count; i < count; i += kFinishBatch * bd) {
        double v[kFinishBatch];
#pragma unroll
        for (int k = 0; k < kFinishBatch; ++k) v[k] = i + k * bd < count ? partials[i + k * bd] : 0.0;
#pragma unroll
        for (int k = 0; k < kFinishBatch; ++k) acc[0] += v[k];
    }
    block_reduce_store<1>(acc, scratch, out);
}

// an LDS element by byte offset
template <typename V>
__device__ __forceinline__ V& lds_at(unsigned char* base, uint32_t byte_off)
{
    return *reinterpret_cast<V*>(base + byte_off);
}

// LDS swizzle: XOR the low SW bits of the element index with the next SW bits.
template <int SW>
__device__ __forceinline__ uint32_t swz(uint32_t u)
{
    return u ^ ((u >> SW) & ((1u << SW) - 1u));
}

// ---------------------------------------------------------------------------
// Tiled sweep kernel
// ---------------------------------------------------------------------------
// F_GRAD (two-state sweeps only): before each qubit's rotation, accumulate
// Im sum_s conj(chi_s) phi_{s ^ 2^q} over the tile (chi = second state); the
// sum over all rotated qubits is Im <chi| sum_q X_q |phi>, unchanged by the
// rotations because both states receive the same ones and every X_q commutes
// with them.  One double per workgroup.
// F_BATCH: the launch covers many states of the same n stored one after another
// (state i at psi + i * 2^n); tile t belongs to state t >> inst_shift, whose
// phase table, butterfly coefficients and P_< / <d_H> references come from
// per-state arrays.  With F_MEASURE every tile writes its own three partials.
// F_POSTROT (with F_PHASE): after the phase -- the end of one Trotter step --
// the tile's qubits are rotated once more, with the next step's coefficients
// (bc2, bs2), walking the phases back to the load layout.  A pass that closes
// step k then also opens step k+1 on its own qubits, which lets a run cover
// n <= 32 qubits in two HBM passes per step instead of three or four.  With
// F_GRAD (two states) the rotations after the phase add their gradient terms
// to partials2: they belong to the next step.
enum : int {
    F_PHASE = 1, F_MEASURE = 2, F_TAN = 4, F_ECACHE = 8, F_PREFETCH = 16, F_GRAD = 32, F_BATCH = 64,
    F_POSTROT = 128, F_LDSTAB = 256, F_LDSTABP = 512, F_NTLOAD = 1024
};

struct SweepArgs {
    uint64_t ntiles;
    const void* tab;   // phase table (V per energy) for F_PHASE
    double* partials;  // gridDim.x * 3 doubles for F_MEASURE, gridDim.x for F_GRAD
    double* partials2; // F_GRAD with F_POSTROT: gridDim.x doubles, terms of the post-phase rotations
    const uint32_t* ecache;  // F_ECACHE: packed 12-bit codes E(s) >> 2 (k_energy_fill)
    void* dst;               // output state (null: in place)
    void* psi2;              // two-state sweeps: the second state (same tiles, same rotations)
    void* dst2;              // its output (null: in place)
    double bc, bs;     // butterfly coefficients (F_TAN: bs = tan, bc unused)
    double bc2, bs2;   // F_POSTROT: coefficients of the rotations after the phase
    // F_GRAD with F_TAN: a tan-form rotation leaves both states scaled by 1/c, so
    // the term of the m-th rotation of a step taken so far is c^-2m too large:
    // terms taken after m rotations of the step are weighted c^2m.  A tile
    // sums its terms as T = T * gic + g (gic = 1 / c^2 = 1 + t^2) in rotation
    // order and adds gw * T, gw = c^2 (m0 + K - 1) for K rotations after m0;
    // gw2 / gic2 the same for the rotations after the phase.  Horner's form
    // keeps no power of c live across the tile loop.
    double gw, gic, gw2, gic2;
    int n;
    // HIGH: the active local bits [R, TB) map to two global segments,
    // [h, h+a1) then [h2, h2+(TB-R-a1)); the tile index fills the remaining
    // bits [R, h), [h+a1, h2), [h2+a2, n) in that order.
    int h;
    int a1;
    int h2;
    int lo_bits;       // = h - R
    int mid_bits;      // = h2 - (h + a1)
    uint32_t s0, e0;
    // F_BATCH
    int inst_shift;             // log2 of the tiles per state
    int tab_stride;             // phase-table entries per state
    const double* bc_arr;       // per state (F_TAN: unused)
    const double* bs_arr;       // per state
    const uint32_t* s0_arr;     // per state
    const uint32_t* e0_arr;     // per state
    // next step's error-feedback table (fb_next non-null): the first
    // workgroups set fb_next[m] = round(fb_scale * fb_master[m] * fb_carry[m])
    // and update the carry (feedback_entry), one entry per thread
    void* fb_next;
    const double* fb_master;
    double* fb_carry;
    int fb_count;
    double fb_scale;
    const double* fb_sd_arr;  // F_BATCH: a scale per state, fb_count entries per state
    int fb_nstates;
    // fin_out non-null: the first workgroup sums fin_count partial triples of
    // the previous pass (fin_partials) into fin_out, as k_finish3 does
    // (fin_single: fin_count doubles into one, as k_finish_sum does)
    const double* fin_partials;
    int fin_count;
    double* fin_out;
    int fin_single;
};

// entries of a phase table for n spins: E = 4m + c with E <= sum_{j<n} j^2
__host__ __device__ constexpr int phase_tab_count(int n)
{
    return static_cast<int>(((static_cast<int64_t>(n - 1) * n * (2 * n - 1) / 6) >> 2) + 1);
}

// energy-cache words per load of a sweep thread holding NR codes (the cache
// layout is described at pack8)
__host__ __device__ constexpr int ec_vec_words(int NR) { return (3 * NR / 8) % 4 == 0 ? 4 : (3 * NR / 8) % 2 == 0 ? 2 : 1; }

// complex64 passes over 2^14-amplitude tiles (the high-qubit groups and the
// wide LOW tile) that apply the phase copy the phase table into LDS behind
// the exchange buffer (at most 2605 entries, 20.8 KB at n = 32): the
// per-amplitude table reads become LDS gathers instead of scattered global
// loads through the vector memory pipeline the state streams use.  The
// 2^13-amplitude tiles (LOW, and the narrow phase-carrying high group) keep
// the table in global memory, where the extra LDS would halve their
// workgroups per CU.  Batched passes keep per-state tables in global memory.
// F_LDSTAB: a 2^13-amplitude complex64 phase pass of n <= kLdsTabMaxN keeps
// the table in LDS too: its 1929 entries at most (15.4 KB) beside the 64 KB
// exchange still leave two workgroups per CU.
// Complex128 high groups (2^13 amplitudes, 128 KB of exchange, one workgroup
// per CU) fit the table's 1929 x 16 B (30.9 KB) beside it up to the same n;
// their rows there are 2^4 .. 2^8 amplitudes.
// F_LDSTABP (complex128 high groups, n >= 30): the first kLdsTabPart entries
// go to LDS (E < 4 kLdsTabPart + c_n, every state but those within a few
// domain walls of the two aligned ones); a code beyond them is read from
// global memory by the lanes that hold one.
constexpr int kLdsTabMaxN = 29, kLdsTab128RMin = 4, kLdsTab128RMax = 8, kLdsTabPart = 2032;
constexpr int kLdsTabPartRMin = 3, kLdsTabPartRMax = 5;  // complex128 rows at n = 30..32
template <typename V, bool LOW, int TB, int FLAGS>
constexpr bool lds_phase_table()
{
    return ((TB >= 14 && sizeof(V) == 8) || (FLAGS & (F_LDSTAB | F_LDSTABP)) != 0) && (FLAGS & F_BATCH) == 0 &&
           (FLAGS & F_PHASE) != 0;
}
// entries the LDS table holds
template <int TB, int FLAGS>
constexpr int lds_phase_table_cap()
{
    return (FLAGS & F_LDSTABP) != 0 ? kLdsTabPart : TB >= 14 ? phase_tab_count(32) : phase_tab_count(kLdsTabMaxN);
}

// Error-feedback phase table: applied = round(sd * master * carry),
// carry <- exact target / applied (FP64 complex).  One entry.
template <typename V>
__device__ __forceinline__ void feedback_entry(const double2 w, double2* __restrict__ carry, V* __restrict__ applied,
                                               double sd)
{
    using R = typename RealOf<V>::T;
    const double2 c = *carry;
    const double tr = sd * (w.x * c.x - w.y * c.y), ti = sd * (w.x * c.y + w.y * c.x);
    V a;
    a.x = static_cast<R>(tr);
    a.y = static_cast<R>(ti);
    const double ar = static_cast<double>(a.x), ai = static_cast<double>(a.y);
    const double inv = 1.0 / (ar * ar + ai * ai);
    double2 nc;
    nc.x = (tr * ar + ti * ai) * inv;  // target / a
    nc.y = (ti * ar - tr * ai) * inv;
    *carry = nc;
    *applied = a;
}

template <int TB_, int RB_, int R_>
struct HighPlan {
    static constexpr Plan plan = make_high_plan(TB_, RB_, R_);
};
template <int TB_, int RB_>
struct LowPlan {
    static constexpr Plan plan = make_low_plan(TB_, RB_);
};

// NS = 2 moves two states through the same tiles and rotations (the reverse
// pass of the adjoint gradient); they share the LDS buffer one after the other.
template <typename V, class PH, int FLAGS, int SW, int NS = 1>
__global__ __launch_bounds__(1 << PH::plan.TT()) void k_sweep(V* psi, const SweepArgs a)
{
    using R = typename RealOf<V>::T;
    constexpr Plan P = PH::plan;
    constexpr int TB = P.TB, RB = P.RB, NR = 1 << RB, NP = P.NP, RR = P.R, TT = P.TT();
    constexpr int NW = (1 << TT) / 64;  // waves per workgroup
    static_assert(NS == 1 || NS == 2, "one or two states");
    static_assert((1 << TT) >= kThreads, "a sweep workgroup holds the kThreads threads of a folded finish");
    static_assert(NS == 1 || (FLAGS & (F_MEASURE | F_PREFETCH)) == 0,
                  "two-state sweeps run without measurement or prefetch");
    static_assert((FLAGS & F_GRAD) == 0 || NS == 2, "the gradient pairs two states");
    constexpr bool BATCH = (FLAGS & F_BATCH) != 0;
    constexpr bool POST = (FLAGS & F_POSTROT) != 0;
    static_assert(!POST || ((NS == 1 || (FLAGS & F_GRAD) != 0) && (FLAGS & F_PREFETCH) == 0 && (FLAGS & F_PHASE) != 0),
                  "post-phase rotations: one state (or the gradient's two), with the phase, no prefetch");
    static_assert(!BATCH || (NS == 1 && (FLAGS & F_PREFETCH) == 0), "batched sweeps: one state per tile, no prefetch");
    constexpr bool LOW = P.low;
    // nontemporal stores in plain passes and one-state LOW phase passes (the
    // two-state LOW phase pass at N=26 stores cached: 0.511 -> 0.493 ms)
    constexpr bool NTSTORE = !BATCH && ((FLAGS & F_PHASE) == 0 || (LOW && NS == 1));
    // nontemporal loads in one-state passes, except the complex64 plain high-group
    // pass, and in complex64 two-state passes but the LOW phase pass.  Two-state
    // passes at N=26: the high-group phase pass runs 0.447 -> 0.425 ms when the
    // plain pass before it loads nontemporally too (0.447 with its own loads
    // alone), the plain pass 0.397 -> 0.402 ms, the LOW phase pass 0.516 -> 0.519
    // F_NTLOAD: the complex64 plain high-group pass too (fused runs at n >= 27)
    constexpr bool NTLOAD = !BATCH && (NS == 1 ? !(sizeof(V) == 8 && !LOW && (FLAGS & (F_PHASE | F_NTLOAD)) == 0)
                                               : sizeof(V) == 8 && !(LOW && (FLAGS & F_PHASE) != 0));
    V* sp[NS];
    V* dp[NS];
    sp[0] = psi;
    dp[0] = a.dst ? static_cast<V*>(a.dst) : psi;
    if constexpr (NS == 2) {
        sp[1] = static_cast<V*>(a.psi2);
        dp[1] = a.dst2 ? static_cast<V*>(a.dst2) : sp[1];
    }
    constexpr uint32_t ROWMASK = (1u << RR) - 1u;
    // Static LDS: its offset is a compile-time constant that the DS
    // instructions absorb, where a dynamic buffer's link-time base costs one
    // more add per access.  Phase passes of complex64 high-qubit groups keep
    // the phase table (up to n = 32) behind the exchange buffer.
    // A one-phase plan exchanges nothing: it only needs the reduction
    // scratch (NW x 3 doubles), which leaves its occupancy to the registers.
    constexpr bool LDS_TAB = lds_phase_table<V, P.low, TB, FLAGS>();
    constexpr int XCH = NP > 1 ? 1 << TB : 0;
    constexpr int SCRATCH = (3 * NW * 8 + static_cast<int>(sizeof(V)) - 1) / static_cast<int>(sizeof(V));
    constexpr int TABCAP = LDS_TAB ? lds_phase_table_cap<TB, FLAGS>() : 0;
    static_assert((FLAGS & (F_LDSTAB | F_LDSTABP)) == 0 || TB == 13, "the 2^13-amplitude tiles' LDS table");
    static_assert((FLAGS & F_LDSTABP) == 0 || sizeof(V) == 16, "the partial LDS table: complex128 high groups");
    constexpr int LDS_ELEMS = XCH + TABCAP > SCRATCH
                                  ? XCH + TABCAP
                                  : SCRATCH;
    __shared__ __align__(16) V lds_buf[LDS_ELEMS];
    unsigned char* smem_raw = reinterpret_cast<unsigned char*>(lds_buf);
    V* lds = lds_buf;
    const uint32_t t = threadIdx.x;

    uint32_t tu[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        uint32_t u = 0;
#pragma unroll
        for (int b = 0; b < TT; ++b) u |= ((t >> b) & 1u) << P.ph[p].thr[b];
        tu[p] = u;
    }
    // offset of a local index from the tile base (HIGH: rows stay, active bits
    // move to h and h2)
    auto gl = [&](uint32_t u) -> uint64_t {
        if constexpr (LOW) return u;
        else {
            const uint32_t ua = u >> RR;
            return static_cast<uint64_t>(u & ROWMASK) + (static_cast<uint64_t>(ua & ((1u << a.a1) - 1u)) << a.h) +
                   (static_cast<uint64_t>(ua >> a.a1) << a.h2);
        }
    };
    const uint64_t tg0 = gl(tu[0]);
    const uint64_t tgL = gl(tu[NP - 1]);
    // HIGH sweeps address register j as the sum of the (uniform) strides of
    // its register bits, rebuilt per tile: kept across the loop, the NR (two
    // states: 2 x NR) 64-bit offsets overflow the scalar registers, and their
    // spill reloads cost more vector instructions than the additions.
    // Batched sweeps keep the per-register offsets: their per-tile state
    // parameters leave no vector registers for the rebuilt addresses.
    constexpr bool STRIDED = !LOW && !BATCH;
    uint64_t lstr[RB], sstr[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
        lstr[b] = gl(1u << P.ph[0].reg[b]);
        sstr[b] = gl(1u << P.ph[NP - 1].reg[b]);
    }
    auto reg_off = [&](const uint64_t(&str)[RB], int j) -> uint64_t {
        uint64_t o = 0;
#pragma unroll
        for (int b = 0; b < RB; ++b)
            if ((j >> b) & 1) o += str[b];
        return o;
    };
    const R bc = static_cast<R>(a.bc), bs = static_cast<R>(a.bs);
    [[maybe_unused]] const V* ltab = nullptr;
    if constexpr (LDS_TAB) {
        V* tl = lds + XCH;
        const V* gt = static_cast<const V*>(a.tab);
        const int cnt = (FLAGS & F_LDSTABP) != 0 ? min(phase_tab_count(a.n), TABCAP) : phase_tab_count(a.n);
        for (int i = static_cast<int>(t); i < cnt; i += 1 << TT) tl[i] = gt[i];
        __syncthreads();
        ltab = tl;
    }
    double acc[3] = {0.0, 0.0, 0.0};
    double gacc[1] = {0.0};
    double gacc2[1] = {0.0};
    constexpr bool PF = (FLAGS & F_PREFETCH) != 0;
    // energy cache words: the cache is stored tile by tile in the last phase's
    // register layout (thread t's NR codes contiguous), eight 12-bit codes per
    // three words (k_energy_fill)
    constexpr int NEC = (FLAGS & F_ECACHE) ? 3 * NR / 8 : 1;
    static_assert(!(FLAGS & F_ECACHE) || (NR % 8 == 0), "cache rows are whole groups of eight codes");

    auto tile_base = [&](uint64_t tile) -> uint64_t {
        if constexpr (LOW) {
            return tile << TB;
        } else {
            const uint64_t lo = tile & ((1ull << a.lo_bits) - 1ull);
            const uint64_t mid = (tile >> a.lo_bits) & ((1ull << a.mid_bits) - 1ull);
            const uint64_t hi = tile >> (a.lo_bits + a.mid_bits);
            return (lo << RR) | (mid << (a.h + a.a1)) | (hi << (a.h2 + (TB - RR - a.a1)));
        }
    };
    // phase-0 layout loads of one tile (amplitudes and, with the cache, E(s))
    // energy-cache position of a tile: the LOW cache is in state order, tile by
    // tile; a HIGH pass's cache holds its tiles one after another
    auto ec_base = [&](uint64_t local_tile, uint64_t base) -> uint64_t {
        if constexpr (LOW) return base;
        else return local_tile << TB;
    };
    auto load = [&](uint64_t base, uint64_t ecb, V(&xv)[NS][NR], uint32_t(&ev)[NEC]) {
        if constexpr (STRIDED) {
            uint64_t str[RB];
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                str[b] = lstr[b];
                asm volatile("" : "+v"(str[b]));  // opaque per tile: no hoisting
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const V* tp = sp[s] + base + tg0;
#pragma unroll
                for (int j = 0; j < NR; ++j) xv[s][j] = load_amp<NTLOAD>(tp + reg_off(str, j));
            }
        } else {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const V* tp = sp[s] + base;
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const uint32_t pj = reg_pattern(P.ph[0], RB, j);
                    xv[s][j] = load_amp<NTLOAD>(tp + tg0 + gl(pj));
                }
            }
        }
        if constexpr ((FLAGS & F_ECACHE) != 0) {
            // ecb counts amplitudes: 3/8 of a word each (tile bases are multiples
            // of 8); vector q of this thread's words sits at q * VW << TT (store8)
            constexpr int VW = ec_vec_words(NR);
            const uint32_t* ew = reinterpret_cast<const uint32_t*>(a.ecache) + (ecb / 8) * 3 +
                                 static_cast<uint64_t>(t) * VW;
            if constexpr (VW == 4) {
#pragma unroll
                for (int q = 0; q < NEC / 4; ++q) {
                    const uint4 w = *reinterpret_cast<const uint4*>(ew + (q << (TT + 2)));
                    ev[4 * q + 0] = w.x;
                    ev[4 * q + 1] = w.y;
                    ev[4 * q + 2] = w.z;
                    ev[4 * q + 3] = w.w;
                }
            } else if constexpr (VW == 2) {
#pragma unroll
                for (int q = 0; q < NEC / 2; ++q) {
                    const uint2 w = *reinterpret_cast<const uint2*>(ew + (q << (TT + 1)));
                    ev[2 * q + 0] = w.x;
                    ev[2 * q + 1] = w.y;
                }
            } else {
#pragma unroll
                for (int q = 0; q < NEC; ++q) ev[q] = ew[q << TT];
            }
        }
    };

    V x[NS][NR];
    uint32_t ec[NEC];
    // F_GRAD: Im conj(chi_j) phi_j' + Im conj(chi_j') phi_j over the pairs of bit b.
    // Phase passes (POST): two FMA chains, (sum chi.x phi'.y) - (sum chi.y phi'.x),
    // four FMAs per pair (N=26 two-state LOW phase pass 0.550 -> 0.533 ms).
    // Plain passes: one multiply and three FMAs per pair, independent across
    // pairs, then a pairwise tree (the chains slow them: 0.388 -> 0.421 ms).
    auto grad_terms = [&](int b) -> R {
        if constexpr (NS == 2 && POST) {
            R gp = R(0), gm = R(0);
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                if ((j >> b) & 1) continue;
                const V c0 = x[1][j], c1 = x[1][j | (1 << b)];
                const V f0 = x[0][j], f1 = x[0][j | (1 << b)];
                gp = fma(c0.x, f1.y, gp);
                gp = fma(c1.x, f0.y, gp);
                gm = fma(c0.y, f1.x, gm);
                gm = fma(c1.y, f0.x, gm);
            }
            return gp - gm;
        } else if constexpr (NS == 2) {
            R d[NR / 2];
            int k = 0;
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                if ((j >> b) & 1) continue;
                const V c0 = x[1][j], c1 = x[1][j | (1 << b)];
                const V f0 = x[0][j], f1 = x[0][j | (1 << b)];
                d[k++] = fma(-c1.y, f0.x, fma(-c0.y, f1.x, fma(c1.x, f0.y, c0.x * f1.y)));
            }
#pragma unroll
            for (int w = NR / 4; w >= 1; w /= 2)
#pragma unroll
                for (int i = 0; i < w; ++i) d[i] += d[i + w];
            return d[0];
        } else {
            return R(0);
        }
    };
    // Two-state sweeps load each tile at the top of its iteration.  Loading
    // the next tile at the bottom (as the one-state sweeps do) makes the loop
    // header merge two loads in a phi, which the compiler rewrites as a load
    // of a phi of pointers: 2 x NR 64-bit addresses carried around the loop,
    // more than the register file holds beside 2 x NR amplitudes.
    constexpr bool TOPLOAD = NS == 2 || BATCH;
    uint64_t tile = blockIdx.x;
    if constexpr (!TOPLOAD) {
        if (tile < a.ntiles) load(tile_base(tile), ec_base(tile, tile_base(tile)), x, ec);
    }
    // the next step's error-feedback table, one entry per thread of the first
    // workgroups, while the first tile's loads are in flight (this pass reads
    // the other table buffer; the pass that read this one has finished)
    if constexpr ((FLAGS & F_PHASE) != 0) {
        if (a.fb_next != nullptr) {
            const double2* w = reinterpret_cast<const double2*>(a.fb_master);
            double2* cr = reinterpret_cast<double2*>(a.fb_carry);
            V* nx = static_cast<V*>(a.fb_next);
            if constexpr (!BATCH) {
                for (int i = static_cast<int>(blockIdx.x * (1u << TT) + t); i < a.fb_count;
                     i += static_cast<int>(gridDim.x) << TT)
                    feedback_entry<V>(w[i], cr + i, nx + i, a.fb_scale);
            } else {
                // the tables of all states, k_phase_feedback_batch's indexing
                const int64_t total = static_cast<int64_t>(a.fb_count) * a.fb_nstates;
                for (int64_t k = static_cast<int64_t>(blockIdx.x) * (1 << TT) + t; k < total;
                     k += static_cast<int64_t>(gridDim.x) << TT) {
                    const int i = static_cast<int>(k / a.fb_count);
                    const int m = static_cast<int>(k - static_cast<int64_t>(i) * a.fb_count);
                    feedback_entry<V>(w[m], cr + k, nx + k, a.fb_sd_arr[i]);
                }
            }
        }
    }
    // the previous pass's observables, finished while the first tile loads
    if (a.fin_out != nullptr && blockIdx.x == 0) {
        __shared__ double fin_scratch[12];
        if (a.fin_single) finish_sum_body(a.fin_partials, a.fin_count, a.fin_out, fin_scratch);
        else finish3_body(a.fin_partials, a.fin_count, a.fin_out, fin_scratch);
    }
    for (; tile < a.ntiles; tile += gridDim.x) {
        uint64_t lt = tile;
        uint32_t inst = 0;
        R bst = bs, bct = bc;
        if constexpr (BATCH) {
            inst = static_cast<uint32_t>(tile >> a.inst_shift);
            lt = tile & ((1ull << a.inst_shift) - 1ull);
            sp[0] = psi + (static_cast<uint64_t>(inst) << a.n);
            dp[0] = a.dst ? static_cast<V*>(a.dst) + (static_cast<uint64_t>(inst) << a.n) : sp[0];
            bst = static_cast<R>(a.bs_arr[inst]);
            if constexpr ((FLAGS & F_TAN) == 0) bct = static_cast<R>(a.bc_arr[inst]);
        }
        const uint64_t base = tile_base(lt);
        const uint64_t nxt = tile + gridDim.x;
        if constexpr (TOPLOAD) load(base, ec_base(lt, base), x, ec);
        // software pipelining: the next tile's loads are issued before this
        // tile's stores, so waiting for them never waits on the stores
        V xn[NS][PF ? NR : 1];
        uint32_t ecn[PF ? NEC : 1];
        if constexpr (PF) {
            if (nxt < a.ntiles) load(tile_base(nxt), ec_base(nxt, tile_base(nxt)), xn, ecn);
        }
        R tg = R(0);
        [[maybe_unused]] const R gic = static_cast<R>(a.gic);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            if (p > 0) {
                // swz is linear over XOR and the thread and register bits are
                // disjoint: swz(tu | pat) = swz(tu) ^ swz(pat), and so are the
                // byte offsets (sizeof(V) is a power of two).  The thread part
                // is made opaque per tile, so each access is one XOR with a
                // constant instead of an address held across the loop.
                uint32_t wb = swz<SW>(tu[p - 1]) * sizeof(V), rb = swz<SW>(tu[p]) * sizeof(V);
                asm volatile("" : "+v"(wb), "+v"(rb));
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < NR; ++j)
                        lds_at<V>(smem_raw, wb ^ (swz<SW>(reg_pattern(P.ph[p - 1], RB, j)) * sizeof(V))) = x[s][j];
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < NR; ++j)
                        x[s][j] = lds_at<V>(smem_raw, rb ^ (swz<SW>(reg_pattern(P.ph[p], RB, j)) * sizeof(V)));
                }
            }
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                if (!((P.ph[p].act >> b) & 1u)) continue;
                if constexpr ((FLAGS & F_GRAD) != 0 && (FLAGS & F_TAN) != 0) {
                    tg = fma(tg, gic, grad_terms(b));
                } else if constexpr ((FLAGS & F_GRAD) != 0) {
                    tg += grad_terms(b);
                }
#pragma unroll
                for (int s = 0; s < NS; ++s)
#pragma unroll
                    for (int j = 0; j < NR; ++j) {
                        if ((j >> b) & 1) continue;
                        if constexpr ((FLAGS & F_TAN) != 0) bfly_tan<V, R>(x[s][j], x[s][j | (1 << b)], bst);
                        else bfly_cs<V, R>(x[s][j], x[s][j | (1 << b)], bct, bst);
                    }
            }
        }
        if constexpr ((FLAGS & F_GRAD) != 0 && (FLAGS & F_TAN) != 0) gacc[0] += a.gw * static_cast<double>(tg);
        else if constexpr ((FLAGS & F_GRAD) != 0) gacc[0] += static_cast<double>(tg);
        // after the last rotation of the step: scale (c^n in tan form) and the
        // next step's phase exp(-i dt E/n), then P_< / <d_H> / norm of the
        // stepped state (|psi|^2 does not see the phase)
        if constexpr ((FLAGS & (F_PHASE | F_MEASURE)) != 0) {
            static_assert(LOW || (FLAGS & F_ECACHE) != 0, "a high-qubit pass reads E(s) from its own cache");
            const V* tab = static_cast<const V*>(a.tab);
            uint32_t e0v = a.e0, s0v = a.s0;
            if constexpr (BATCH) {
                tab += static_cast<uint64_t>(inst) * a.tab_stride;
                e0v = a.e0_arr[inst];
                s0v = a.s0_arr[inst];
            }
            // with the cache the energy comes as its code m (E = 4m + c_n):
            // E < e0 exactly when m < ceil((e0 - c_n) / 4)
            [[maybe_unused]] const uint32_t cres = (static_cast<uint32_t>(a.n) >> 1) & 3u;
            [[maybe_unused]] const uint32_t m0 = e0v > cres ? (e0v - cres + 3u) >> 2 : 0u;
            R tacc[3] = {R(0), R(0), R(0)};
            // d_H weight of register j: s = T | off_j with T = base + tgL (thread and
            // tile bits) and off_j on the register bits M, disjoint, so
            // popcount(s ^ s0) = popcount((T ^ s0) & ~M) + popcount(j ^ d), where d
            // gathers the register-bit positions of T ^ s0 into a register index
            [[maybe_unused]] int pcT = 0;
            [[maybe_unused]] uint32_t dj = 0;
            if constexpr ((FLAGS & F_MEASURE) != 0) {
                const uint32_t D = static_cast<uint32_t>(base + tgL) ^ s0v;
                uint32_t M = 0;
#pragma unroll
                for (int b = 0; b < RB; ++b) {
                    const uint32_t gb = static_cast<uint32_t>(gl(1u << P.ph[NP - 1].reg[b]));
                    M |= gb;
                    dj |= ((D & gb) ? 1u : 0u) << b;
                }
                pcT = __popc(D & ~M);
            }
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                uint32_t mcode;  // the phase table's index, E >> 2
                bool below;      // E(s) < e0
                if constexpr ((FLAGS & F_ECACHE) != 0) {
                    // j is a constant once the loop is unrolled: one shift or one
                    // funnel shift per code
                    const int bit = 12 * (j & 7), w = 3 * (j >> 3) + (bit >> 5), sh = bit & 31;
                    if (sh <= 20) mcode = (ec[w] >> sh) & 0xFFFu;
                    else mcode = ((ec[w] >> sh) | (ec[w + 1] << (32 - sh))) & 0xFFFu;
                    below = mcode < m0;
                } else {
                    const uint32_t E =
                        labs_energy_u32(static_cast<uint32_t>(base + tgL + gl(reg_pattern(P.ph[NP - 1], RB, j))), a.n);
                    mcode = E >> 2;
                    below = E < e0v;
                }
                if constexpr ((FLAGS & F_PHASE) != 0) {
                    // E = 4m + c_n: one entry per m; the table carries the step's scale
                    V ph;
                    if constexpr ((FLAGS & F_LDSTABP) != 0) {
                        if (mcode < static_cast<uint32_t>(TABCAP)) ph = ltab[mcode];
                        else ph = tab[mcode];
                    } else {
                        ph = LDS_TAB ? ltab[mcode] : tab[mcode];
                    }
                    const R pr = ph.x, pi = ph.y;
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        const R xr = x[s][j].x, xi = x[s][j].y;
                        x[s][j].x = xr * pr - xi * pi;
                        x[s][j].y = xr * pi + xi * pr;
                    }
                }
                if constexpr ((FLAGS & F_MEASURE) != 0) {
                    // per-tile partials in the state's precision (a thread's
                    // NR terms), folded into FP64 once per tile
                    const R pr = fma(x[0][j].x, x[0][j].x, x[0][j].y * x[0][j].y);
                    tacc[2] += pr;
                    tacc[0] += below ? pr : R(0);
                    tacc[1] = fma(pr, static_cast<R>(pcT + __popc(static_cast<uint32_t>(j) ^ dj)), tacc[1]);
                }
            }
            if constexpr ((FLAGS & F_MEASURE) != 0) {
#pragma unroll
                for (int i = 0; i < 3; ++i) acc[i] += static_cast<double>(tacc[i]);
            }
        }
        if constexpr (POST) {
            // the next step's rotations of this tile's qubits: phases in
            // reverse order, ending in the load layout
            // a batched state keeps its field: the next step's coefficients are its own
            const R bs2 = BATCH ? bst : static_cast<R>(a.bs2), bc2 = BATCH ? bct : static_cast<R>(a.bc2);
            R tg2 = R(0);
            [[maybe_unused]] const R gic2 = static_cast<R>(a.gic2);
#pragma unroll
            for (int p = NP - 1; p >= 0; --p) {
                if (p < NP - 1) {
                    uint32_t wb = swz<SW>(tu[p + 1]) * sizeof(V), rb = swz<SW>(tu[p]) * sizeof(V);
                    asm volatile("" : "+v"(wb), "+v"(rb));
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        __syncthreads();
#pragma unroll
                        for (int j = 0; j < NR; ++j)
                            lds_at<V>(smem_raw, wb ^ (swz<SW>(reg_pattern(P.ph[p + 1], RB, j)) * sizeof(V))) = x[s][j];
                        __syncthreads();
#pragma unroll
                        for (int j = 0; j < NR; ++j)
                            x[s][j] = lds_at<V>(smem_raw, rb ^ (swz<SW>(reg_pattern(P.ph[p], RB, j)) * sizeof(V)));
                    }
                }
#pragma unroll
                for (int b = 0; b < RB; ++b) {
                    if (!((P.ph[p].act >> b) & 1u)) continue;
                    if constexpr ((FLAGS & F_GRAD) != 0 && (FLAGS & F_TAN) != 0) {
                        tg2 = fma(tg2, gic2, grad_terms(b));
                    } else if constexpr ((FLAGS & F_GRAD) != 0) {
                        tg2 += grad_terms(b);
                    }
#pragma unroll
                    for (int s = 0; s < NS; ++s)
#pragma unroll
                        for (int j = 0; j < NR; ++j) {
                            if ((j >> b) & 1) continue;
                            if constexpr ((FLAGS & F_TAN) != 0) bfly_tan<V, R>(x[s][j], x[s][j | (1 << b)], bs2);
                            else bfly_cs<V, R>(x[s][j], x[s][j | (1 << b)], bc2, bs2);
                        }
                }
            }
            if constexpr ((FLAGS & F_GRAD) != 0 && (FLAGS & F_TAN) != 0) gacc2[0] += a.gw2 * static_cast<double>(tg2);
            else if constexpr ((FLAGS & F_GRAD) != 0) gacc2[0] += static_cast<double>(tg2);
        }
        if constexpr (STRIDED) {
            // registers are in the last phase's layout, or (POST) in the load layout
            uint64_t str[RB];
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                str[b] = POST ? lstr[b] : sstr[b];
                asm volatile("" : "+v"(str[b]));
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                V* op = dp[s] + base + (POST ? tg0 : tgL);
#pragma unroll
                for (int j = 0; j < NR; ++j) store_amp<NTSTORE>(op + reg_off(str, j), x[s][j]);
            }
        } else {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                V* op = dp[s] + base;
                constexpr int SP = POST ? 0 : NP - 1;  // layout of the registers at the store
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const uint32_t pj = reg_pattern(P.ph[SP], RB, j);
                    store_amp<NTSTORE>(op + (POST ? tg0 : tgL) + gl(pj), x[s][j]);
                }
            }
        }
        if constexpr (BATCH && (FLAGS & F_MEASURE) != 0) {
            // one set of partials per tile (tiles of different states never share a sum)
            __syncthreads();
            block_reduce_store<3, NW>(acc, reinterpret_cast<double*>(smem_raw), a.partials + 3 * tile);
            acc[0] = acc[1] = acc[2] = 0.0;
        }
        if (nxt < a.ntiles) {
            if constexpr (PF) {
#pragma unroll
                for (int j = 0; j < NR; ++j) x[0][j] = xn[0][j];
                if constexpr ((FLAGS & F_ECACHE) != 0)
#pragma unroll
                    for (int j = 0; j < NEC; ++j) ec[j] = ecn[j];
            } else if constexpr (!TOPLOAD) {
                load(tile_base(nxt), ec_base(nxt, tile_base(nxt)), x, ec);
            }
        }
    }
    if constexpr ((FLAGS & F_MEASURE) != 0 && !BATCH) {
        __syncthreads();
        block_reduce_store<3, NW>(acc, reinterpret_cast<double*>(smem_raw), a.partials + 3 * blockIdx.x);
    }
    if constexpr ((FLAGS & F_GRAD) != 0) {
        __syncthreads();
        block_reduce_store<1, NW>(gacc, reinterpret_cast<double*>(smem_raw), a.partials + blockIdx.x);
        if constexpr (POST) {
            __syncthreads();  // thread 0 has read the scratch words before they are reused
            block_reduce_store<1, NW>(gacc2, reinterpret_cast<double*>(smem_raw), a.partials2 + blockIdx.x);
        }
    }
}

// ---------------------------------------------------------------------------
// Small-n / generic kernels
// ---------------------------------------------------------------------------

// one qubit's rotation over all pairs (used when n is below the tile size)
template <typename V, bool TAN>
__global__ __launch_bounds__(kThreads) void k_qubit(V* __restrict__ psi, int q, uint64_t npairs,
                                                   double bc_, double bs_)
{
    using R = typename RealOf<V>::T;
    const R bc = static_cast<R>(bc_), bs = static_cast<R>(bs_);
    const uint64_t st = 1ull << q;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < npairs;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t lo = ((i >> q) << (q + 1)) | (i & (st - 1));
        V x0 = psi[lo], x1 = psi[lo | st];
        if (TAN) bfly_tan<V, R>(x0, x1, bs);
        else bfly_cs<V, R>(x0, x1, bc, bs);
        psi[lo] = x0;
        psi[lo | st] = x1;
    }
}

// elementwise measurement (F_MEASURE) and/or phase*scale (F_PHASE) over the
// whole state; contiguous chunk per block so the partials are deterministic.
template <typename V, int FLAGS>
__global__ __launch_bounds__(kThreads) void k_elementwise(V* __restrict__ psi, uint64_t dim,
                                                         uint64_t chunk, const void* tab_,
                                                         double scale, int n, uint32_t s0,
                                                         uint32_t e0, double* partials)
{
    using R = typename RealOf<V>::T;
    __shared__ double scratch[12];
    const V* tab = static_cast<const V*>(tab_);
    const R sc = static_cast<R>(scale);
    double acc[3] = {0.0, 0.0, 0.0};
    const uint64_t beg = blockIdx.x * chunk;
    const uint64_t end = (beg + chunk < dim) ? beg + chunk : dim;
    for (uint64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        V x = psi[i];
        const uint32_t s = static_cast<uint32_t>(i);
        const uint32_t E = labs_energy_u32(s, n);
        if constexpr ((FLAGS & F_PHASE) != 0) {
            const V ph = tab[E >> 2];
            const R pr = ph.x * sc, pi = ph.y * sc;
            const R xr = x.x, xi = x.y;
            x.x = xr * pr - xi * pi;
            x.y = xr * pi + xi * pr;
            psi[i] = x;
        }
        if constexpr ((FLAGS & F_MEASURE) != 0) {
            const double xr = static_cast<double>(x.x), xi = static_cast<double>(x.y);
            const double pr = fma(xr, xr, xi * xi);
            acc[2] += pr;
            acc[0] += (E < e0) ? pr : 0.0;
            acc[1] = fma(pr, static_cast<double>(__popc(s ^ s0)), acc[1]);
        }
    }
    if constexpr ((FLAGS & F_MEASURE) != 0) block_reduce_store<3>(acc, scratch, partials + 3 * blockIdx.x);
}

// sum 'count' triples of partials in a fixed order into out[0..2]
__global__ __launch_bounds__(kThreads) void k_finish3(const double* __restrict__ partials, int count,
                                                     double* __restrict__ out)
{
    __shared__ double scratch[12];
    finish3_body(partials, count, out, scratch);
}

// one workgroup per state: its 'count' partial triples into out[3 i .. 3 i + 2]
__global__ __launch_bounds__(kThreads) void k_finish3_batch(const double* __restrict__ partials, int count,
                                                           double* __restrict__ out)
{
    __shared__ double scratch[12];
    finish3_body(partials + 3 * static_cast<size_t>(blockIdx.x) * count, count, out + 3 * blockIdx.x, scratch);
}

// state i of a batch (at psi + i * 2^n) gets amplitude (re_i, im_i) at s0_i
template <typename V>
__global__ __launch_bounds__(kThreads) void k_set_basis_batch(V* psi, int n, int nstates, const uint32_t* __restrict__ s0,
                                                             const double* __restrict__ re, const double* __restrict__ im)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nstates; i += gridDim.x * blockDim.x) {
        V v;
        v.x = static_cast<typename RealOf<V>::T>(re[i]);
        v.y = static_cast<typename RealOf<V>::T>(im[i]);
        psi[(static_cast<uint64_t>(i) << n) + s0[i]] = v;
    }
}

template <typename V>
__global__ void k_set_basis(V* psi, uint64_t s0, double re, double im)
{
    V v;
    v.x = static_cast<typename RealOf<V>::T>(re);
    v.y = static_cast<typename RealOf<V>::T>(im);
    psi[s0] = v;
}

// The energy caches hold m = E(s) >> 2 for every s: every E(s) is 4m + c_n with
// c_n = floor(n/2) mod 4 (C_k = (n-k) mod 2), and m <= 2604 at n = 32 fits 12
// bits.  Eight codes share three 32-bit words (1.5 B per amplitude):
//   w0 = m0 | m1 << 12 | m2 << 24, w1 = m2 >> 8 | m3 << 4 | m4 << 16 | m5 << 28,
//   w2 = m5 >> 4 | m6 << 8 | m7 << 20.
// A cache is laid out tile by tile.  Thread t of a sweep holds the NR codes
// of its last-phase registers in 3 NR / 8 words, which it loads as vectors of
// VW words; vector q of every thread of the tile precedes vector q + 1, so
// each of a wave's cache loads reads one contiguous run of 64 VW words.
__device__ __forceinline__ void pack8(uint32_t* __restrict__ w, const uint32_t (&m)[8])
{
    w[0] = m[0] | (m[1] << 12) | (m[2] << 24);
    w[1] = (m[2] >> 8) | (m[3] << 4) | (m[4] << 16) | (m[5] << 28);
    w[2] = (m[5] >> 4) | (m[6] << 8) | (m[7] << 20);
}

// the three words of codes 8g .. 8g+7 (in the tile order of a TB/RB plan)
template <int TB, int RB>
__device__ __forceinline__ void store8(uint32_t* __restrict__ out, uint64_t g, const uint32_t (&m)[8])
{
    constexpr int NR = 1 << RB, TT = TB - RB, VW = ec_vec_words(NR);
    uint32_t w[3];
    pack8(w, m);
    const uint64_t i0 = 8 * g;
    const uint64_t tile_words = (i0 >> TB) * (3ull << (TB - 3));
    const uint32_t rem = static_cast<uint32_t>(i0) & ((1u << TB) - 1u);
    const uint32_t t = rem >> RB, k = (rem & (NR - 1u)) >> 3;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const uint32_t tw = 3 * k + q;  // word of thread t
        out[tile_words + (tw / VW) * (static_cast<uint64_t>(VW) << TT) + t * VW + tw % VW] = w[q];
    }
}

// LOW-sweep layout (plan PH)
template <class PH>
__global__ __launch_bounds__(kThreads) void k_energy_fill(uint32_t* __restrict__ out, int n, uint64_t dim)
{
    constexpr Plan P = PH::plan;
    constexpr int TB = P.TB, RB = P.RB, NR = 1 << RB, NP = P.NP;
    static_assert(NR % 8 == 0, "codes are packed eight at a time");
    for (uint64_t g = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; g < dim / 8;
         g += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        uint32_t m[8];
        for (int q = 0; q < 8; ++q) {
            const uint64_t i = 8 * g + q;
            const uint64_t tile = i >> TB;
            const uint32_t t = static_cast<uint32_t>(i >> RB) & ((1u << (TB - RB)) - 1u);
            const uint32_t j = static_cast<uint32_t>(i) & (NR - 1u);
            uint32_t u = 0;
            for (int b = 0; b < TB - RB; ++b) u |= ((t >> b) & 1u) << P.ph[NP - 1].thr[b];
            for (int b = 0; b < RB; ++b) u |= ((j >> b) & 1u) << P.ph[NP - 1].reg[b];
            m[q] = labs_energy_u32(static_cast<uint32_t>((tile << TB) | u), n) >> 2;
        }
        store8<TB, RB>(out, g, m);
    }
}

// the layout in which a HIGH pass (plan PH, geometry h/a1/h2) holds its
// amplitudes at the phase: the same addressing as k_sweep's tile_base and gl
template <class PH>
__global__ __launch_bounds__(kThreads) void k_energy_fill_high(uint32_t* __restrict__ out, int n, int h, int a1, int h2,
                                                              int lo_bits, int mid_bits, uint64_t dim)
{
    constexpr Plan P = PH::plan;
    constexpr int TB = P.TB, RB = P.RB, NR = 1 << RB, NP = P.NP, RR = P.R, TT = P.TT();
    constexpr uint32_t ROWMASK = (1u << RR) - 1u;
    static_assert(NR % 8 == 0, "codes are packed eight at a time");
    for (uint64_t g = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; g < dim / 8;
         g += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        uint32_t m[8];
        for (int q = 0; q < 8; ++q) {
            const uint64_t i = 8 * g + q;
            const uint64_t tile = i >> TB;
            const uint32_t rem = static_cast<uint32_t>(i) & ((1u << TB) - 1u);
            const uint32_t t = rem >> RB, j = rem & (NR - 1u);
            uint32_t u = 0;
            for (int b = 0; b < TT; ++b) u |= ((t >> b) & 1u) << P.ph[NP - 1].thr[b];
            for (int b = 0; b < RB; ++b) u |= ((j >> b) & 1u) << P.ph[NP - 1].reg[b];
            const uint32_t ua = u >> RR;
            const uint64_t off = static_cast<uint64_t>(u & ROWMASK) +
                                 (static_cast<uint64_t>(ua & ((1u << a1) - 1u)) << h) +
                                 (static_cast<uint64_t>(ua >> a1) << h2);
            const uint64_t lo = tile & ((1ull << lo_bits) - 1ull);
            const uint64_t mid = (tile >> lo_bits) & ((1ull << mid_bits) - 1ull);
            const uint64_t hi = tile >> (lo_bits + mid_bits);
            const uint64_t base = (lo << RR) | (mid << (h + a1)) | (hi << (h2 + (TB - RR - a1)));
            m[q] = labs_energy_u32(static_cast<uint32_t>(base + off), n) >> 2;
        }
        store8<TB, RB>(out, g, m);
    }
}

// number of s in [0, 2^n) with E(s) < e0
__global__ __launch_bounds__(kThreads) void k_count_below(int n, uint32_t e0, uint64_t dim,
                                                         uint64_t chunk,
                                                         unsigned long long* partials)
{
    __shared__ unsigned long long scratch[4];
    unsigned long long c = 0;
    const uint64_t beg = blockIdx.x * chunk;
    const uint64_t end = (beg + chunk < dim) ? beg + chunk : dim;
    for (uint64_t i = beg + threadIdx.x; i < end; i += blockDim.x)
        c += (labs_energy_u32(static_cast<uint32_t>(i), n) < e0) ? 1ull : 0ull;
    c = wave_sum_u64(c);
    if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = (scratch[0] + scratch[1]) + (scratch[2] + scratch[3]);
}

__global__ __launch_bounds__(kThreads) void k_finish_u64(const unsigned long long* __restrict__ partials,
                                                        int count, unsigned long long* out)
{
    __shared__ unsigned long long scratch[4];
    unsigned long long c = 0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) c += partials[i];
    c = wave_sum_u64(c);
    if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) *out = (scratch[0] + scratch[1]) + (scratch[2] + scratch[3]);
}

template <typename V>
__global__ __launch_bounds__(kThreads) void k_phase_feedback(const double2* __restrict__ master,
                                                            double2* __restrict__ carry, V* __restrict__ applied,
                                                            int count, double sd)
{
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < count; m += gridDim.x * blockDim.x)
        feedback_entry<V>(master[m], carry + m, applied + m, sd);
}

// the same update for nstates tables of 'count' entries (state i: carry/applied
// entries [i*count, (i+1)*count), scale sd[i]); master is shared
template <typename V>
__global__ __launch_bounds__(kThreads) void k_phase_feedback_batch(const double2* __restrict__ master,
                                                                  double2* __restrict__ carry, V* __restrict__ applied,
                                                                  int count, int nstates, const double* __restrict__ sd)
{
    const int64_t total = static_cast<int64_t>(count) * nstates;
    for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < total;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int i = static_cast<int>(k / count);
        const int m = static_cast<int>(k - static_cast<int64_t>(i) * count);
        feedback_entry<V>(master[m], carry + k, applied + k, sd[i]);
    }
}

__global__ __launch_bounds__(kThreads) void k_carry_reset(double2* __restrict__ carry, int count)
{
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < count; m += gridDim.x * blockDim.x)
        carry[m] = make_double2(1.0, 0.0);
}

// sum 'count' doubles in a fixed order into *out
__global__ __launch_bounds__(kThreads) void k_finish_sum(const double* __restrict__ partials, int count,
                                                        double* __restrict__ out)
{
    __shared__ double scratch[4];
    finish_sum_body(partials, count, out, scratch);
}

// ---------------------------------------------------------------------------
// Adjoint-pass kernels (two states: phi = the reversed forward state, chi =
// the adjoint state)
// ---------------------------------------------------------------------------

// one qubit's rotation of both states (states below the tile size), with the
// pair term Im conj(chi_s) phi_{s^q} + Im conj(chi_{s^q}) phi_s taken before it
template <typename V>
__global__ __launch_bounds__(kThreads) void k_qubit2(V* __restrict__ phi, V* __restrict__ chi, int q,
                                                    uint64_t npairs, double bc_, double bs_,
                                                    double* __restrict__ partials)
{
    using R = typename RealOf<V>::T;
    __shared__ double scratch[4];
    const R bc = static_cast<R>(bc_), bs = static_cast<R>(bs_);
    const uint64_t st = 1ull << q;
    double acc[1] = {0.0};
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < npairs;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t lo = ((i >> q) << (q + 1)) | (i & (st - 1));
        V f0 = phi[lo], f1 = phi[lo | st], c0 = chi[lo], c1 = chi[lo | st];
        acc[0] += static_cast<double>(fma(c0.x, f1.y, -c0.y * f1.x)) +
                  static_cast<double>(fma(c1.x, f0.y, -c1.y * f0.x));
        bfly_cs<V, R>(f0, f1, bc, bs);
        bfly_cs<V, R>(c0, c1, bc, bs);
        phi[lo] = f0;
        phi[lo | st] = f1;
        chi[lo] = c0;
        chi[lo | st] = c1;
    }
    block_reduce_store<1>(acc, scratch, partials + blockIdx.x);
}

// both states times exp(-i dtp E/n) (the table entry of E)
template <typename V>
__global__ __launch_bounds__(kThreads) void k_phase2(V* __restrict__ phi, V* __restrict__ chi, uint64_t dim,
                                                    const void* tab_, int n)
{
    using R = typename RealOf<V>::T;
    const V* tab = static_cast<const V*>(tab_);
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < dim;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const V ph = tab[labs_energy_u32(static_cast<uint32_t>(i), n) >> 2];
        V f = phi[i], c = chi[i];
        const R fr = f.x, cr = c.x;
        f.x = fr * ph.x - f.y * ph.y;
        f.y = fr * ph.y + f.y * ph.x;
        c.x = cr * ph.x - c.y * ph.y;
        c.y = cr * ph.y + c.y * ph.x;
        phi[i] = f;
        chi[i] = c;
    }
}

// Start of the reverse pass: phi = P * src, chi = W * phi with the diagonal
// objective weight W(s) = wp [E(s) < e0] + wd popcount(s ^ s0).
template <typename V>
__global__ __launch_bounds__(kThreads) void k_adj_init(const V* __restrict__ src, V* __restrict__ phi,
                                                      V* __restrict__ chi, uint64_t dim, const void* tab_,
                                                      int n, uint32_t s0, uint32_t e0, double wp, double wd)
{
    using R = typename RealOf<V>::T;
    const V* tab = static_cast<const V*>(tab_);
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < dim;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t s = static_cast<uint32_t>(i);
        const uint32_t E = labs_energy_u32(s, n);
        const V ph = tab[E >> 2];
        const V x = src[i];
        V f;
        f.x = x.x * ph.x - x.y * ph.y;
        f.y = x.x * ph.y + x.y * ph.x;
        const R w = static_cast<R>((E < e0 ? wp : 0.0) + wd * static_cast<double>(__popc(s ^ s0)));
        V c;
        c.x = w * f.x;
        c.y = w * f.y;
        phi[i] = f;
        chi[i] = c;
    }
}

// Checkpoint reset: dst <- P * src, with sum |new - old|^2 per block (the
// drift of the reversed state from the stored forward state).
template <typename V>
__global__ __launch_bounds__(kThreads) void k_phase_from(V* __restrict__ dst, const V* __restrict__ src,
                                                        uint64_t dim, uint64_t chunk, const void* tab_, int n,
                                                        double* __restrict__ partials)
{
    __shared__ double scratch[4];
    const V* tab = static_cast<const V*>(tab_);
    double acc[1] = {0.0};
    const uint64_t beg = blockIdx.x * chunk;
    const uint64_t end = (beg + chunk < dim) ? beg + chunk : dim;
    for (uint64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const V ph = tab[labs_energy_u32(static_cast<uint32_t>(i), n) >> 2];
        const V x = src[i];
        V f;
        f.x = x.x * ph.x - x.y * ph.y;
        f.y = x.x * ph.y + x.y * ph.x;
        const V o = dst[i];
        const double dr = static_cast<double>(f.x) - static_cast<double>(o.x);
        const double di = static_cast<double>(f.y) - static_cast<double>(o.y);
        acc[0] += dr * dr + di * di;
        dst[i] = f;
    }
    block_reduce_store<1>(acc, scratch, partials + blockIdx.x);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace {

constexpr int kLowTB64 = 13, kLowRB64 = 5;
constexpr int kWideLowTB = 14;  // complex64 LOW tile of fused runs with 19 high qubits (512 threads, 128 KB LDS)
constexpr int kLowTB128 = 12, kLowRB128 = 4;
constexpr int kSW64 = 5, kSW128 = 4;

// LOW tiles: 2^TB amplitudes, 256 threads.  HIGH tiles: 2^HTB amplitudes,
// 2^(HTB-HRB) threads and 2^HTB * sizeof(V) bytes of LDS (128 KB), so their
// contiguous rows can be 256-512 B while each sweep still rotates 8-9 qubits.
template <typename V>
struct Cfg;
template <>
struct Cfg<float2> {
    static constexpr int TB = kLowTB64, RB = kLowRB64, SW = kSW64;
    static constexpr int HTB = 14, HRB = 5;
};
template <>
struct Cfg<double2> {
    static constexpr int TB = kLowTB128, RB = kLowRB128, SW = kSW128;
    static constexpr int HTB = 13, HRB = 4;
};

static_assert(LowPlan<Cfg<float2>::TB, Cfg<float2>::RB>::plan.TT() == kThreadBits &&
                  LowPlan<Cfg<double2>::TB, Cfg<double2>::RB>::plan.TT() == kThreadBits,
              "the default LOW tiles run kThreads threads (the reduction scratch and grids assume it)");

struct MixCoef {
    bool tan_form;
    double bc, bs, scale;
};

MixCoef mix_coef(int n, double gdt)
{
    // reference matrix [[cos(rx/2), -i sin(rx/2)], ...] with rx = -2 g dt
    const double rx = -2.0 * gdt;
    const double c = std::cos(rx / 2.0), s = -std::sin(rx / 2.0);
    MixCoef m{};
    if (std::fabs(c) >= 0.5) {
        m.tan_form = true;
        m.bc = 1.0;
        m.bs = s / c;
        m.scale = std::pow(c, n);
    } else {
        m.tan_form = false;
        m.bc = c;
        m.bs = s;
        m.scale = 1.0;
    }
    return m;
}

// norm-exact mixer scale for the coefficients as a kernel of real type Rt sees
// them: every tan-form butterfly scales the pair norm by 1 + t^2, every (c, s)
// butterfly by c^2 + s^2
template <typename Rt>
double exact_scale(const MixCoef& mc, int n)
{
    const double b = static_cast<double>(static_cast<Rt>(mc.bs));
    if (mc.tan_form) return std::pow(1.0 + b * b, -0.5 * n);
    const double c = static_cast<double>(static_cast<Rt>(mc.bc));
    return std::pow(c * c + b * b, -0.5 * n);
}

template <typename V, int FLAGS>
hipError_t launch_low(V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    using C = Cfg<V>;
    using PL = LowPlan<C::TB, C::RB>;
    auto k = k_sweep<V, PL, FLAGS, C::SW>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1 << PL::plan.TT()), 0, st, psi, a);
    return hipGetLastError();
}

// runtime flag set -> template instantiation of the LOW sweep
template <typename V, int FL = 0>
hipError_t launch_low_dyn(int fl, V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    if constexpr (FL >= 32) {
        return hipErrorInvalidValue;
    } else if constexpr ((FL & F_PHASE) == 0) {
        return launch_low_dyn<V, FL + 1>(fl, psi, a, grid, st);  // LOW always carries the phase
    } else {
        if (fl == FL) return launch_low<V, FL>(psi, a, grid, st);
        return launch_low_dyn<V, FL + 1>(fl, psi, a, grid, st);
    }
}

template <typename V, int R, bool TAN, bool PF, bool NT = false>
hipError_t launch_high_r(V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    using C = Cfg<V>;
    using PH = HighPlan<C::HTB, C::HRB, R>;
    auto k = k_sweep<V, PH, (TAN ? F_TAN : 0) | (PF ? F_PREFETCH : 0) | (NT ? F_NTLOAD : 0), C::SW>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1 << PH::plan.TT()), 0, st, psi, a);
    return hipGetLastError();
}

template <typename V, bool TAN, bool PF, int R = 3, bool NT = false>
hipError_t launch_high(V* psi, int rbits, const SweepArgs& a, int grid, hipStream_t st)
{
    if constexpr (R > Cfg<V>::HTB - 1) {
        return hipErrorInvalidValue;
    } else {
        if (rbits == R) return launch_high_r<V, R, TAN, PF, NT>(psi, a, grid, st);
        return launch_high<V, TAN, PF, R + 1, NT>(psi, rbits, a, grid, st);
    }
}

// nt: nontemporal amplitude loads (complex64 tan-form prefetching passes only)
template <typename V>
hipError_t launch_high_dyn(bool tan, bool pf, V* psi, int rbits, const SweepArgs& a, int grid, hipStream_t st,
                           bool nt = false)
{
    if constexpr (std::is_same_v<V, float2>) {
        if (nt) return tan && pf ? launch_high<V, true, true, 3, true>(psi, rbits, a, grid, st) : hipErrorInvalidValue;
    }
    if (nt) return hipErrorInvalidValue;
    if (tan) return pf ? launch_high<V, true, true>(psi, rbits, a, grid, st)
                       : launch_high<V, true, false>(psi, rbits, a, grid, st);
    return pf ? launch_high<V, false, true>(psi, rbits, a, grid, st)
              : launch_high<V, false, false>(psi, rbits, a, grid, st);
}

int clamp_grid(uint64_t ntiles, int grid)
{
    const uint64_t g = std::min<uint64_t>(ntiles, static_cast<uint64_t>(std::max(grid, 1)));
    return static_cast<int>(g);
}

template <typename V>
hipError_t step_impl(V* psi, V* alt, int* in_alt, const StepSpec& sp, const SweepTuning& tun,
                     double* partials, hipStream_t st)
{
    // Step order: HIGH sweeps (qubits [TB, n)), then the LOW sweep (qubits
    // [0, TB)) which ends with scale * exp(-i dtp E/n) and the reductions.
    *in_alt = 0;
    using C = Cfg<V>;
    using Rt = typename RealOf<V>::T;
    const int n = sp.n;
    MixCoef mc = mix_coef(n, sp.gdt);
    const uint64_t dim = 1ull << n;
    const bool measure = sp.series_slot != nullptr;
    hipError_t e = hipSuccess;
    if (sp.fb_carry) {
        // norm-exact scale for the coefficients as the kernels see them, folded
        // into the error-feedback table (the sweeps then apply scale 1)
        const double sd = exact_scale<Rt>(mc, n);
        const int fg = std::max(1, std::min((sp.fb_count + kThreads - 1) / kThreads, 64));
        hipLaunchKernelGGL(k_phase_feedback<V>, dim3(fg), dim3(kThreads), 0, st,
                           reinterpret_cast<const double2*>(sp.fb_master), reinterpret_cast<double2*>(sp.fb_carry),
                           static_cast<V*>(const_cast<void*>(sp.phase_tab)), sp.fb_count, sd);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        mc.scale = 1.0;
    }
    if (n < C::TB) {
        // small states: one kernel per qubit, then elementwise scale*phase (+measure)
        const uint64_t np = dim / 2;
        const int qg = static_cast<int>(std::min<uint64_t>((np + kThreads - 1) / kThreads, 4096));
        for (int q = 0; q < n; ++q) {
            if (mc.tan_form)
                hipLaunchKernelGGL((k_qubit<V, true>), dim3(qg), dim3(kThreads), 0, st, psi, q, np,
                                   mc.bc, mc.bs);
            else
                hipLaunchKernelGGL((k_qubit<V, false>), dim3(qg), dim3(kThreads), 0, st, psi, q, np,
                                   mc.bc, mc.bs);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        const uint64_t chunk = 4096;
        const int grid = static_cast<int>((dim + chunk - 1) / chunk);
        if (measure)
            hipLaunchKernelGGL((k_elementwise<V, F_PHASE | F_MEASURE>), dim3(grid), dim3(kThreads), 0,
                               st, psi, dim, chunk, sp.phase_tab, mc.scale, n, sp.s0, sp.e0, partials);
        else
            hipLaunchKernelGGL((k_elementwise<V, F_PHASE>), dim3(grid), dim3(kThreads), 0, st, psi,
                               dim, chunk, sp.phase_tab, mc.scale, n, sp.s0, sp.e0, partials);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (measure) {
            hipLaunchKernelGGL(k_finish3, dim3(1), dim3(kThreads), 0, st, partials, grid, sp.series_slot);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        return hipSuccess;
    }
    SweepArgs a{};
    a.tab = sp.phase_tab;
    a.partials = partials;
    a.bc = mc.bc;
    a.bs = mc.bs;
    // the tiled sweeps take the step's overall factor (c^n in tan form) from the
    // error-feedback table
    if (mc.scale != 1.0) return hipErrorInvalidValue;
    a.n = n;
    a.s0 = sp.s0;
    a.e0 = sp.e0;
    a.ecache = sp.ecache;
    a.ntiles = dim >> C::TB;
    const double sweep_bytes = 2.0 * static_cast<double>(sizeof(V)) * static_cast<double>(dim);
    // out-of-place sweeps alternate between psi and alt when alt is given
    V* cur = psi;
    V* oth = alt;
    // HIGH sweeps over qubits [TB, n)
    const std::vector<HighSweep> plan = plan_high_sweeps(n, C::TB, C::HTB, std::max(tun.row_bits, 3),
                                                         tun.page_bits);
    for (const HighSweep& hs : plan) {
        SweepArgs b = a;
        b.dst = alt ? oth : nullptr;
        b.h = hs.h1;
        b.a1 = hs.a1;
        b.h2 = hs.h2;
        b.lo_bits = hs.h1 - hs.R;
        b.mid_bits = hs.h2 - (hs.h1 + hs.a1);
        b.ntiles = dim >> C::HTB;
        const int g = clamp_grid(b.ntiles, tun.grid);
        if (tun.timer) tun.timer->begin(st);
        e = launch_high_dyn<V>(mc.tan_form, tun.prefetch_high, cur, hs.R, b, g, st);
        if (e != hipSuccess) return e;
        if (alt) std::swap(cur, oth);
        if (tun.timer) {
            char name[48];
            if (hs.a2 > 0)
                std::snprintf(name, sizeof(name), "sweep_high_q%d_%d+q%d_%d", hs.h1, hs.h1 + hs.a1 - 1, hs.h2,
                              hs.h2 + hs.a2 - 1);
            else
                std::snprintf(name, sizeof(name), "sweep_high_q%d_%d", hs.h1, hs.h1 + hs.a1 - 1);
            tun.timer->end(st, name, sweep_bytes);
        }
    }
    // LOW sweep: qubits [0, TB), then scale*phase and the reductions
    const int g1 = clamp_grid(a.ntiles, tun.grid);
    a.dst = alt ? oth : nullptr;
    const int fl = F_PHASE | (measure ? F_MEASURE : 0) | (mc.tan_form ? F_TAN : 0) |
                   (sp.ecache ? F_ECACHE : 0) | (tun.prefetch_low ? F_PREFETCH : 0);
    if (tun.timer) tun.timer->begin(st);
    e = launch_low_dyn<V>(fl, cur, a, g1, st);
    if (e != hipSuccess) return e;
    if (alt) std::swap(cur, oth);
    if (tun.timer) tun.timer->end(st, "sweep_low", sweep_bytes);
    if (measure) {
        hipLaunchKernelGGL(k_finish3, dim3(1), dim3(kThreads), 0, st, partials, g1, sp.series_slot);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    *in_alt = (cur != psi);
    return hipSuccess;
}

// ---- reverse (adjoint) sweeps: two states, gradient terms ----
template <typename V, int R, int FL>
hipError_t launch_high_rev_r(V* phi, const SweepArgs& a, int grid, hipStream_t st)
{
    using C = Cfg<V>;
    using PH = HighPlan<C::HTB, C::HRB, R>;
    auto k = k_sweep<V, PH, FL, C::SW, 2>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1 << PH::plan.TT()), 0, st, phi, a);
    return hipGetLastError();
}

template <typename V, int FL = F_GRAD, int R = 3>
hipError_t launch_high_rev(V* phi, int rbits, const SweepArgs& a, int grid, hipStream_t st)
{
    if constexpr (R > Cfg<V>::HTB - 1) {
        return hipErrorInvalidValue;
    } else {
        if (rbits == R) return launch_high_rev_r<V, R, FL>(phi, a, grid, st);
        return launch_high_rev<V, FL, R + 1>(phi, rbits, a, grid, st);
    }
}

template <typename V, int FL = F_PHASE | F_GRAD | F_ECACHE>
hipError_t launch_low_rev(V* phi, const SweepArgs& a, int grid, hipStream_t st)
{
    using C = Cfg<V>;
    using PL = LowPlan<C::TB, C::RB>;
    auto k = k_sweep<V, PL, FL, C::SW, 2>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1 << PL::plan.TT()), 0, st, phi, a);
    return hipGetLastError();
}

// One reverse step of both states: the rotations exp(i gdt X) per qubit
// (sp.gdt = -g dt undoes a forward step of field g), the gradient terms of
// the states as they enter, then exp(-i dtp E/n) from sp.phase_tab.
// grad_slot <- Im <chi| sum_q X_q |phi> (FP64, deterministic order).
template <typename V>
hipError_t reverse_step_impl(V* phi, V* chi, const StepSpec& sp, const SweepTuning& tun, double* gpart,
                             double* grad_slot, hipStream_t st)
{
    using C = Cfg<V>;
    const int n = sp.n;
    const double bc = std::cos(sp.gdt), bs = std::sin(sp.gdt);
    const uint64_t dim = 1ull << n;
    hipError_t e = hipSuccess;
    if (n < C::TB) {
        const uint64_t np = dim / 2;
        const int qg = static_cast<int>(std::min<uint64_t>((np + kThreads - 1) / kThreads, 64));
        for (int q = 0; q < n; ++q) {
            hipLaunchKernelGGL(k_qubit2<V>, dim3(qg), dim3(kThreads), 0, st, phi, chi, q, np, bc, bs,
                               gpart + static_cast<size_t>(q) * qg);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        const int pg = static_cast<int>(std::min<uint64_t>((dim + kThreads - 1) / kThreads, 1024));
        hipLaunchKernelGGL(k_phase2<V>, dim3(pg), dim3(kThreads), 0, st, phi, chi, dim, sp.phase_tab, n);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kThreads), 0, st, gpart, n * qg, grad_slot);
        return hipGetLastError();
    }
    SweepArgs a{};
    a.tab = sp.phase_tab;
    a.bc = bc;
    a.bs = bs;
    a.n = n;
    a.ecache = sp.ecache;
    a.psi2 = chi;
    a.ntiles = dim >> C::TB;
    const double sweep_bytes = 4.0 * static_cast<double>(sizeof(V)) * static_cast<double>(dim);
    int off = 0;
    const std::vector<HighSweep> plan = plan_high_sweeps(n, C::TB, C::HTB, std::max(tun.row_bits, 3),
                                                         tun.page_bits);
    for (const HighSweep& hs : plan) {
        SweepArgs b = a;
        b.h = hs.h1;
        b.a1 = hs.a1;
        b.h2 = hs.h2;
        b.lo_bits = hs.h1 - hs.R;
        b.mid_bits = hs.h2 - (hs.h1 + hs.a1);
        b.ntiles = dim >> C::HTB;
        b.partials = gpart + off;
        const int g = clamp_grid(b.ntiles, tun.grid);
        if (tun.timer) tun.timer->begin(st);
        e = launch_high_rev<V>(phi, hs.R, b, g, st);
        if (e != hipSuccess) return e;
        if (tun.timer) tun.timer->end(st, "adjoint_sweep_high", sweep_bytes);
        off += g;
    }
    const int g1 = clamp_grid(a.ntiles, tun.grid);
    a.partials = gpart + off;
    if (tun.timer) tun.timer->begin(st);
    if (!sp.ecache) return hipErrorInvalidValue;  // the reverse LOW sweep reads E(s) from the cache
    e = launch_low_rev<V>(phi, a, g1, st);
    if (e != hipSuccess) return e;
    if (tun.timer) tun.timer->end(st, "adjoint_sweep_low", sweep_bytes);
    off += g1;
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kThreads), 0, st, gpart, off, grad_slot);
    return hipGetLastError();
}

template <typename V>
hipError_t adj_init_impl(const V* src, V* phi, V* chi, int n, const void* tab, uint32_t s0, uint32_t e0, double wp,
                         double wd, hipStream_t st)
{
    const uint64_t dim = 1ull << n;
    const int grid = static_cast<int>(std::min<uint64_t>((dim + kThreads - 1) / kThreads, 16384));
    hipLaunchKernelGGL(k_adj_init<V>, dim3(grid), dim3(kThreads), 0, st, src, phi, chi, dim, tab, n, s0, e0, wp,
                       wd);
    return hipGetLastError();
}

// ---- batched sweeps: many states of one n in one launch per sweep ----
template <typename V, int R, bool TAN>
hipError_t launch_high_batch_r(V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    using C = Cfg<V>;
    using PH = HighPlan<C::HTB, C::HRB, R>;
    auto k = k_sweep<V, PH, F_BATCH | (TAN ? F_TAN : 0), C::SW>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1 << PH::plan.TT()), 0, st, psi, a);
    return hipGetLastError();
}

template <typename V, bool TAN, int R = 3>
hipError_t launch_high_batch(V* psi, int rbits, const SweepArgs& a, int grid, hipStream_t st)
{
    if constexpr (R > Cfg<V>::HTB - 1) {
        return hipErrorInvalidValue;
    } else {
        if (rbits == R) return launch_high_batch_r<V, R, TAN>(psi, a, grid, st);
        return launch_high_batch<V, TAN, R + 1>(psi, rbits, a, grid, st);
    }
}

template <typename V, int FL>
hipError_t launch_low_batch(V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    using C = Cfg<V>;
    using PL = LowPlan<C::TB, C::RB>;
    auto k = k_sweep<V, PL, FL, C::SW>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1 << PL::plan.TT()), 0, st, psi, a);
    return hipGetLastError();
}

template <typename V>
hipError_t batch_step_impl(const BatchGroup& g, const SweepTuning& tun, double* partials, hipStream_t st)
{
    using C = Cfg<V>;
    const int n = g.n;
    if (n < C::TB || !g.ecache || g.psi_out) return hipErrorInvalidValue;  // step-by-step sweeps run in place
    SweepArgs a{};
    a.tab = g.tab;
    a.tab_stride = g.tab_stride;
    a.ecache = g.ecache;
    a.n = n;
    a.bc_arr = g.bc;
    a.bs_arr = g.bs;
    a.s0_arr = g.s0;
    a.e0_arr = g.e0;
    a.partials = partials;
    V* psi = static_cast<V*>(g.psi);
    const double sweep_bytes = 2.0 * static_cast<double>(sizeof(V)) * static_cast<double>(g.nstates) *
                               static_cast<double>(1ull << n);
    hipError_t e = hipSuccess;
    for (const HighSweep& hs : plan_high_sweeps(n, C::TB, C::HTB, std::max(tun.row_bits, 3), tun.page_bits)) {
        SweepArgs b = a;
        b.h = hs.h1;
        b.a1 = hs.a1;
        b.h2 = hs.h2;
        b.lo_bits = hs.h1 - hs.R;
        b.mid_bits = hs.h2 - (hs.h1 + hs.a1);
        b.inst_shift = n - C::HTB;
        b.ntiles = static_cast<uint64_t>(g.nstates) << (n - C::HTB);
        const int grid = clamp_grid(b.ntiles, tun.grid);
        if (tun.timer) tun.timer->begin(st);
        e = g.tan_form ? launch_high_batch<V, true>(psi, hs.R, b, grid, st)
                       : launch_high_batch<V, false>(psi, hs.R, b, grid, st);
        if (e != hipSuccess) return e;
        if (tun.timer) tun.timer->end(st, "batch_sweep_high", sweep_bytes);
    }
    a.inst_shift = n - C::TB;
    a.ntiles = static_cast<uint64_t>(g.nstates) << (n - C::TB);
    const int grid = clamp_grid(a.ntiles, tun.grid);
    constexpr int B = F_BATCH | F_PHASE | F_ECACHE;
    if (tun.timer) tun.timer->begin(st);
    if (g.tan_form)
        e = partials ? launch_low_batch<V, B | F_TAN | F_MEASURE>(psi, a, grid, st)
                     : launch_low_batch<V, B | F_TAN>(psi, a, grid, st);
    else
        e = partials ? launch_low_batch<V, B | F_MEASURE>(psi, a, grid, st) : launch_low_batch<V, B>(psi, a, grid, st);
    if (e != hipSuccess) return e;
    if (tun.timer) tun.timer->end(st, "batch_sweep_low", sweep_bytes);
    return hipSuccess;
}

// ---- fused runs: two HBM passes per Trotter step ----
// row bits of a phase-carrying high pass: 3..12 covers every n whose high
// qubits fit one or two sweep groups (n <= 32)
constexpr int kFusedRMin = 3, kFusedRMax = 12;
// the narrow phase-carrying high group (SweepTuning::narrow_phase): complex64
// tiles of 2^kNarrowTB amplitudes with rows of 2^6 .. 2^8 amplitudes
// (n >= kNarrowMinN: states beyond the last-level cache; the batched passes,
// whose states are smaller, keep 2^14-amplitude tiles, so a batched state of
// n < kNarrowMinN follows its serial run bit for bit, observables included)
constexpr int kNarrowTB = 13, kNarrowRMin = 6, kNarrowRMax = 8, kNarrowMinN = 26;

// LTB: tile bits of the LOW pass (the default tile, or the complex64 wide tile
// of 2^kWideLowTB amplitudes, 512 threads, that a fused run can take)
template <typename V, int FL, int LTB = Cfg<V>::TB>
hipError_t launch_low_fl(V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    using C = Cfg<V>;
    using PL = LowPlan<LTB, C::RB>;
    auto k = k_sweep<V, PL, FL, C::SW>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1 << PL::plan.TT()), 0, st, psi, a);
    return hipGetLastError();
}

// a LOW pass of a fused run with ltb tile bits
template <typename V, int FL>
hipError_t launch_low_run(int ltb, V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    if constexpr (std::is_same_v<V, float2> && (FL & F_LDSTAB) == 0) {
        if (ltb == kWideLowTB) return launch_low_fl<V, FL, kWideLowTB>(psi, a, grid, st);
    }
    if (ltb != Cfg<V>::TB) return hipErrorInvalidValue;
    return launch_low_fl<V, FL>(psi, a, grid, st);
}

template <typename V, int R, int FL, int HTB = Cfg<V>::HTB>
hipError_t launch_high_fl(V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    using C = Cfg<V>;
    using PH = HighPlan<HTB, C::HRB, R>;
    auto k = k_sweep<V, PH, FL, C::SW>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1 << PH::plan.TT()), 0, st, psi, a);
    return hipGetLastError();
}

template <typename V, int FL, int R = kFusedRMin>
hipError_t launch_highp_r(int rbits, V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    if constexpr (R > ((FL & F_LDSTAB) ? kLdsTab128RMax : (FL & F_LDSTABP) ? kLdsTabPartRMax : kFusedRMax)) {
        return hipErrorInvalidValue;
    } else {
        if (rbits == R) return launch_high_fl<V, R, FL>(psi, a, grid, st);
        return launch_highp_r<V, FL, R + 1>(rbits, psi, a, grid, st);
    }
}

// the narrow phase-carrying group: 2^kNarrowTB-amplitude tiles (complex64)
template <typename V, int FL, int R = kNarrowRMin>
hipError_t launch_highn_r(int rbits, V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    if constexpr (R > kNarrowRMax || !std::is_same_v<V, float2>) {
        return hipErrorInvalidValue;
    } else {
        if (rbits == R) return launch_high_fl<V, R, FL, kNarrowTB>(psi, a, grid, st);
        return launch_highn_r<V, FL, R + 1>(rbits, psi, a, grid, st);
    }
}

// a phase-carrying high pass of a group on htb tile bits
template <typename V, int FL>
hipError_t launch_highp(int htb, int rbits, V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    if (htb == Cfg<V>::HTB) return launch_highp_r<V, FL>(rbits, psi, a, grid, st);
    if (htb == kNarrowTB) return launch_highn_r<V, FL>(rbits, psi, a, grid, st);
    return hipErrorInvalidValue;
}

// phase-carrying passes (tan butterflies): F_PHASE | F_ECACHE | F_TAN | any of
// F_MEASURE, F_POSTROT
template <typename V>
hipError_t launch_ppass(bool low, int ltb, int fl, int htb, int rbits, V* psi, const SweepArgs& a, int grid,
                        hipStream_t st)
{
    constexpr int B = F_PHASE | F_ECACHE | F_TAN;
    switch (fl) {
    case B:
        return low ? launch_low_run<V, B>(ltb, psi, a, grid, st) : launch_highp<V, B>(htb, rbits, psi, a, grid, st);
    case B | F_MEASURE:
        return low ? launch_low_run<V, B | F_MEASURE>(ltb, psi, a, grid, st)
                   : launch_highp<V, B | F_MEASURE>(htb, rbits, psi, a, grid, st);
    case B | F_POSTROT:
        return low ? launch_low_run<V, B | F_POSTROT>(ltb, psi, a, grid, st)
                   : launch_highp<V, B | F_POSTROT>(htb, rbits, psi, a, grid, st);
    case B | F_MEASURE | F_POSTROT:
        return low ? launch_low_run<V, B | F_MEASURE | F_POSTROT>(ltb, psi, a, grid, st)
                   : launch_highp<V, B | F_MEASURE | F_POSTROT>(htb, rbits, psi, a, grid, st);
    }
    if constexpr (std::is_same_v<V, double2>) {
        // complex128 high groups (2^13-amplitude tiles) with the phase table in LDS
        constexpr int BL = B | F_LDSTAB;
        if (low) return hipErrorInvalidValue;
        switch (fl) {
        case BL: return launch_highp_r<V, BL, kLdsTab128RMin>(rbits, psi, a, grid, st);
        case BL | F_MEASURE: return launch_highp_r<V, BL | F_MEASURE, kLdsTab128RMin>(rbits, psi, a, grid, st);
        case BL | F_POSTROT: return launch_highp_r<V, BL | F_POSTROT, kLdsTab128RMin>(rbits, psi, a, grid, st);
        case BL | F_MEASURE | F_POSTROT:
            return launch_highp_r<V, BL | F_MEASURE | F_POSTROT, kLdsTab128RMin>(rbits, psi, a, grid, st);
        }
        constexpr int BP = B | F_LDSTABP;
        switch (fl) {
        case BP: return launch_highp_r<V, BP, kLdsTabPartRMin>(rbits, psi, a, grid, st);
        case BP | F_MEASURE: return launch_highp_r<V, BP | F_MEASURE, kLdsTabPartRMin>(rbits, psi, a, grid, st);
        case BP | F_POSTROT: return launch_highp_r<V, BP | F_POSTROT, kLdsTabPartRMin>(rbits, psi, a, grid, st);
        case BP | F_MEASURE | F_POSTROT:
            return launch_highp_r<V, BP | F_MEASURE | F_POSTROT, kLdsTabPartRMin>(rbits, psi, a, grid, st);
        }
    }
    if constexpr (std::is_same_v<V, float2>) {
        // 2^13-amplitude tiles with the phase table in LDS (n <= kLdsTabMaxN)
        constexpr int BL = B | F_LDSTAB;
        switch (fl) {
        case BL:
            return low ? launch_low_run<V, BL>(ltb, psi, a, grid, st) : launch_highn_r<V, BL>(rbits, psi, a, grid, st);
        case BL | F_MEASURE:
            return low ? launch_low_run<V, BL | F_MEASURE>(ltb, psi, a, grid, st)
                       : launch_highn_r<V, BL | F_MEASURE>(rbits, psi, a, grid, st);
        case BL | F_POSTROT:
            return low ? launch_low_run<V, BL | F_POSTROT>(ltb, psi, a, grid, st)
                       : launch_highn_r<V, BL | F_POSTROT>(rbits, psi, a, grid, st);
        case BL | F_MEASURE | F_POSTROT:
            return low ? launch_low_run<V, BL | F_MEASURE | F_POSTROT>(ltb, psi, a, grid, st)
                       : launch_highn_r<V, BL | F_MEASURE | F_POSTROT>(rbits, psi, a, grid, st);
        }
    }
    return hipErrorInvalidValue;
}

// The high-qubit groups of a fused run whose LOW pass covers qubits [0, ltb)
// (0: the default tile): the fewest groups (one or two) that the planner makes
// with rows of at least 2^kFusedRMin amplitudes, the last group (the
// phase-carrying one) with R in [kFusedRMin, kFusedRMax].
template <typename V>
std::vector<HighSweep> planned_groups(int n, const SweepTuning& tun, int ltb)
{
    using C = Cfg<V>;
    const int tb = ltb > 0 ? ltb : C::TB;
    if (!tun.fused || n <= tb) return {};
    for (int rmin = kFusedRMin + 1; rmin >= kFusedRMin; --rmin) {  // prefer 16-amplitude rows
        std::vector<HighSweep> plan = plan_high_sweeps(n, tb, C::HTB, rmin, tun.page_bits);
        if (plan.empty() || plan.size() > 2) continue;
        if (plan.back().R < kFusedRMin || plan.back().R > kFusedRMax) continue;
        if (plan.size() == 1) return plan;
        // two groups: both fit with the caller's row preference as well
        std::vector<HighSweep> pref = plan_high_sweeps(n, tb, C::HTB, std::max(tun.row_bits, 3), tun.page_bits);
        if (pref.size() == 2 && pref.back().R >= kFusedRMin && pref.back().R <= kFusedRMax) return pref;
        return plan;
    }
    return {};
}

// tile bits of a high group's passes: R row bits and its rotated qubits
inline int group_tb(const HighSweep& hs) { return hs.R + hs.a1 + hs.a2; }

// planned_groups, with the phase-carrying group of a two-group complex64 run
// moved to 2^kNarrowTB-amplitude tiles when SweepTuning::narrow_phase asks
// for it, n >= kNarrowMinN and its rows stay at least 2^kNarrowRMin
// amplitudes long; or to 1024-thread passes when SweepTuning::phase_rb asks
template <typename V>
std::vector<HighSweep> fused_groups(int n, const SweepTuning& tun, int ltb = 0)
{
    std::vector<HighSweep> g = planned_groups<V>(n, tun, ltb);
    if constexpr (std::is_same_v<V, float2>) {
        if (tun.narrow_phase && g.size() == 2 && n >= kNarrowMinN) {
            HighSweep& b = g.back();
            const int r = b.R - (group_tb(b) - kNarrowTB);
            if (group_tb(b) > kNarrowTB && r >= kNarrowRMin && r <= kNarrowRMax) b.R = r;
        }
    }
    return g;
}

// LOW tile bits of a fused forward run.  complex64 at n = 32 has 19 high
// qubits: beside a 13-qubit LOW tile two groups of at most 14 - R qubits need
// 16-amplitude rows (R = 4, 128 B) in one of them, and short rows cost HBM
// efficiency.  The wide LOW tile (14 qubits) leaves 18 high qubits, two
// groups of 9 with 32-amplitude rows.
template <typename V>
int fused_ltb(int n, const SweepTuning& tun)
{
    using C = Cfg<V>;
    constexpr int kFullRows = 5;  // 32 amplitudes, 256 B
    if constexpr (std::is_same_v<V, float2>) {
        if (tun.wide_low) {
            const std::vector<HighSweep> narrow = planned_groups<V>(n, tun, 0);
            bool short_rows = false;
            for (const HighSweep& h : narrow) short_rows |= h.R < kFullRows;
            if (short_rows && narrow.size() == 2) {
                const std::vector<HighSweep> wide = planned_groups<V>(n, tun, kWideLowTB);
                bool ok = wide.size() == 2;
                for (const HighSweep& h : wide) ok &= h.R >= kFullRows;
                if (ok) return kWideLowTB;
            }
        }
    }
    return C::TB;
}

SweepArgs high_args(const SweepArgs& a, const HighSweep& hs, uint64_t dim)
{
    SweepArgs b = a;
    b.h = hs.h1;
    b.a1 = hs.a1;
    b.h2 = hs.h2;
    b.lo_bits = hs.h1 - hs.R;
    b.mid_bits = hs.h2 - (hs.h1 + hs.a1);
    b.ntiles = dim >> (hs.R + hs.a1 + hs.a2);
    return b;
}

template <typename V>
hipError_t run_impl(V* psi, const RunSpec& rs, const SweepTuning& tun, double* partials, hipStream_t st)
{
    using C = Cfg<V>;
    using Rt = typename RealOf<V>::T;
    const int n = rs.n, K = rs.steps;
    const uint64_t dim = 1ull << n;
    hipError_t e = hipSuccess;
    const int ltb = rs.low_bits > 0 ? rs.low_bits : C::TB;  // the LOW cache's tile
    const std::vector<HighSweep> grp = (K >= 1 && rs.ecache && rs.ecache_high) ? fused_groups<V>(n, tun, ltb)
                                                                               : std::vector<HighSweep>{};
    bool fused = !grp.empty();
    for (int k = 0; fused && k < K; ++k)
        if (!mix_coef(n, rs.gdt[k]).tan_form) fused = false;  // phase-carrying passes use the tan butterfly
    // step by step needs the default LOW cache layout
    if (!fused && ltb != C::TB && K >= 1) return hipErrorInvalidValue;
    const bool two = grp.size() == 2;
    const HighSweep A = fused && two ? grp[0] : HighSweep{};
    const HighSweep B = fused ? grp.back() : HighSweep{};
    if (!fused) {
        for (int k = 0; k < K; ++k) {
            StepSpec sp{};
            sp.n = n;
            sp.precision = rs.precision;
            sp.gdt = rs.gdt[k];
            sp.phase_tab = rs.phase_tab;
            sp.series_slot = rs.slots ? rs.slots[k] : nullptr;
            sp.s0 = rs.s0;
            sp.e0 = rs.e0;
            sp.ecache = rs.ecache;
            sp.fb_master = rs.fb_master;
            sp.fb_carry = rs.fb_carry;
            sp.fb_count = rs.fb_count;
            int in_alt = 0;
            if ((e = step_impl(psi, static_cast<V*>(nullptr), &in_alt, sp, tun, partials, st)) != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const bool tan = true;
    const bool fold = rs.phase_tab2 != nullptr;
    // measured pass k writes partials region k & 1; the next pass of the run
    // finishes it into the step's slot (the last one, k_finish3)
    const size_t region_doubles = 3 * static_cast<size_t>(std::max(max_grid(tun), 4096));
    struct Fin {
        const double* part = nullptr;
        int count = 0;
        double* slot = nullptr;
    } fin;
    auto take_fin = [&](SweepArgs& b) {
        b.fin_partials = fin.part;
        b.fin_count = fin.count;
        b.fin_out = fin.slot;
        fin = Fin{};
    };
    SweepArgs a{};
    a.tab = rs.phase_tab;
    a.partials = partials;
    a.n = n;
    a.s0 = rs.s0;
    a.e0 = rs.e0;
    auto coef = [&](int k, double* bc, double* bs) {
        const MixCoef m = mix_coef(n, rs.gdt[k]);
        *bc = m.bc;
        *bs = m.bs;
    };
    const double pass_bytes = 2.0 * static_cast<double>(sizeof(V)) * static_cast<double>(dim);
    // Two high groups A, B: LOW(1), then per step A(k) and a phase pass that
    // alternates B (even k) / LOW (odd k).  One high group H: H(1), then per
    // step a phase pass alternating LOW (even k) / H (odd k) -- one pass a step.
    if (two) {
        SweepArgs b = a;
        coef(0, &b.bc, &b.bs);
        b.ntiles = dim >> ltb;
        const int g = clamp_grid(b.ntiles, tun.grid);
        if (tun.timer) tun.timer->begin(st);
        e = launch_low_run<V, F_TAN>(ltb, psi, b, g, st);
        if (e != hipSuccess) return e;
        if (tun.timer) tun.timer->end(st, "fused_low_open", pass_bytes);
    } else {
        SweepArgs b = high_args(a, B, dim);
        coef(0, &b.bc, &b.bs);
        const int g = clamp_grid(b.ntiles, tun.grid);
        if (tun.timer) tun.timer->begin(st);
        e = launch_high_dyn<V>(tan, false, psi, B.R, b, g, st);
        if (e != hipSuccess) return e;
        if (tun.timer) tun.timer->end(st, "fused_high_open", pass_bytes);
    }
    for (int k = 0; k < K; ++k) {
        // A: the first high-qubit group of step k (two groups only)
        if (two) {
            SweepArgs b = high_args(a, A, dim);
            coef(k, &b.bc, &b.bs);
            take_fin(b);
            const int g = clamp_grid(b.ntiles, tun.grid);
            if (tun.timer) tun.timer->begin(st);
            // the next tile's loads go out before this tile's rotations (N=30: 3.242 -> 3.225 ms)
            e = launch_high_dyn<V>(tan, true, psi, A.R, b, g, st, std::is_same_v<V, float2> && tun.nt_plain);
            if (e != hipSuccess) return e;
            if (tun.timer) tun.timer->end(st, "fused_high_a", pass_bytes);
        }
        // the step's error-feedback phase table: its own kernel, or (two table
        // buffers) written by the previous step's phase pass
        V* tab_k = static_cast<V*>(fold && (k & 1) ? rs.phase_tab2 : const_cast<void*>(rs.phase_tab));
        if (!fold || k == 0) {
            const MixCoef m = mix_coef(n, rs.gdt[k]);
            const int fg = std::max(1, std::min((rs.fb_count + kThreads - 1) / kThreads, 64));
            hipLaunchKernelGGL(k_phase_feedback<V>, dim3(fg), dim3(kThreads), 0, st,
                               reinterpret_cast<const double2*>(rs.fb_master), reinterpret_cast<double2*>(rs.fb_carry),
                               tab_k, rs.fb_count, exact_scale<Rt>(m, n));
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        // phase-carrying pass
        const bool low = two ? (k & 1) != 0 : (k & 1) == 0;
        const bool post = k + 1 < K;
        double* slot = rs.slots ? rs.slots[k] : nullptr;
        SweepArgs b = low ? a : high_args(a, B, dim);
        if (low) b.ntiles = dim >> ltb;
        b.ecache = low ? rs.ecache : rs.ecache_high;
        b.tab = tab_k;
        if (fold && post) {
            b.fb_next = (k & 1) ? const_cast<void*>(rs.phase_tab) : rs.phase_tab2;
            b.fb_master = rs.fb_master;
            b.fb_carry = rs.fb_carry;
            b.fb_count = rs.fb_count;
            b.fb_scale = exact_scale<Rt>(mix_coef(n, rs.gdt[k + 1]), n);
        }
        coef(k, &b.bc, &b.bs);
        if (post) coef(k + 1, &b.bc2, &b.bs2);
        take_fin(b);
        b.partials = partials + static_cast<size_t>(k & 1) * region_doubles;
        // complex64 2^13-amplitude tiles (the default LOW tile, the narrow
        // group) take the phase table into LDS while it fits
        const bool ldstab = tun.lds_tab13 && n <= kLdsTabMaxN &&
                            (std::is_same_v<V, float2> ? (low ? ltb == kNarrowTB : group_tb(B) == kNarrowTB)
                                                       : (!low && group_tb(B) == Cfg<V>::HTB));
        // complex128 high groups beyond kLdsTabMaxN: the table's first entries in LDS
        const bool ldstabp = std::is_same_v<V, double2> && tun.lds_tab13 && n > kLdsTabMaxN && !low &&
                             group_tb(B) == Cfg<V>::HTB;
        const int fl = F_PHASE | F_ECACHE | F_TAN | (slot ? F_MEASURE : 0) | (post ? F_POSTROT : 0) |
                       (ldstab ? F_LDSTAB : 0) | (ldstabp ? F_LDSTABP : 0);
        const int gp = low ? tun.grid_lphase : tun.grid_bphase;
        const int g = clamp_grid(b.ntiles, gp > 0 ? gp : tun.grid);
        if (tun.timer) tun.timer->begin(st);
        e = launch_ppass<V>(low, ltb, fl, group_tb(B), B.R, psi, b, g, st);
        if (e != hipSuccess) return e;
        if (tun.timer) tun.timer->end(st, low ? "fused_low_phase" : "fused_high_phase", pass_bytes);
        if (slot) fin = Fin{b.partials, g, slot};
        if (fin.slot && !tun.fold_finish) {
            hipLaunchKernelGGL(k_finish3, dim3(1), dim3(kThreads), 0, st, fin.part, fin.count, fin.slot);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            fin = Fin{};
        }
    }
    if (fin.slot) {
        hipLaunchKernelGGL(k_finish3, dim3(1), dim3(kThreads), 0, st, fin.part, fin.count, fin.slot);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---- fused reverse runs (adjoint): the schedule of run_impl on two states ----
constexpr int kRevPass = F_PHASE | F_GRAD | F_ECACHE | F_POSTROT;

template <typename V, int FL = kRevPass, int R = kFusedRMin, int HTB = Cfg<V>::HTB>
hipError_t launch_high_rev_ppass_r(V* phi, int rbits, const SweepArgs& a, int grid, hipStream_t st)
{
    using C = Cfg<V>;
    constexpr int RMAX = HTB == Cfg<V>::HTB ? kFusedRMax : kNarrowRMax;
    if constexpr (R > RMAX || R >= HTB) {
        return hipErrorInvalidValue;
    } else {
        if (rbits != R) return launch_high_rev_ppass_r<V, FL, R + 1, HTB>(phi, rbits, a, grid, st);
        using PH = HighPlan<HTB, C::HRB, R>;
        auto k = k_sweep<V, PH, FL, C::SW, 2>;
        hipLaunchKernelGGL(k, dim3(grid), dim3(1 << PH::plan.TT()), 0, st, phi, a);
        return hipGetLastError();
    }
}

// reverse phase pass of a high group on htb tile bits
template <typename V, int FL = kRevPass>
hipError_t launch_high_rev_ppass(V* phi, int htb, int rbits, const SweepArgs& a, int grid, hipStream_t st)
{
    if constexpr ((FL & F_LDSTAB) == 0) {
        if (htb == Cfg<V>::HTB) return launch_high_rev_ppass_r<V, FL>(phi, rbits, a, grid, st);
    }
    if constexpr (std::is_same_v<V, float2>) {
        if (htb == kNarrowTB) return launch_high_rev_ppass_r<V, FL, kNarrowRMin, kNarrowTB>(phi, rbits, a, grid, st);
    }
    return hipErrorInvalidValue;
}

// Reverse steps i = 0..K-1 of phi and chi, each the rotations exp(i gdt_i X)
// then the phase from rs.phase_tab; step i's gradient sum goes to
// grad_slots[i].  Two high groups A, B: LOW(0), then per step A(i) and a
// phase pass alternating B (even i) / LOW (odd i); one group H: H(0), then a
// phase pass alternating LOW (even i) / H (odd i).  A phase pass adds the terms
// of its first rotations to step i and those after the phase to step i + 1;
// the last pass rotates by the identity (c = 1, s = 0, exact) after its phase.
// Every qubit group of step i has written its partials into region i & 1 when
// step i's phase pass ends, which is when the region is summed.
// workgroups of a fused reverse run's passes over the LOW tile (tb bits),
// group A and group B (the phase-carrying group), on 2^htb-amplitude tiles
void reverse_grids(uint64_t dim, int tb, int htb, const SweepTuning& tun, int* gl, int* ga, int* gb)
{
    *gl = clamp_grid(dim >> tb, tun.rev_grid_low > 0 ? tun.rev_grid_low : tun.grid);
    *ga = clamp_grid(dim >> htb, tun.grid);
    *gb = clamp_grid(dim >> htb, tun.rev_grid_bphase > 0 ? tun.rev_grid_bphase : tun.grid);
}

template <typename V>
hipError_t reverse_run_impl(V* phi, V* chi, const RevRunSpec& rs, const SweepTuning& tun, double* gpart,
                            hipStream_t st)
{
    using C = Cfg<V>;
    const int n = rs.n, K = rs.steps;
    const uint64_t dim = 1ull << n;
    hipError_t e = hipSuccess;
    const std::vector<HighSweep> grp = (K >= 1 && rs.ecache && rs.ecache_high) ? fused_groups<V>(n, tun)
                                                                               : std::vector<HighSweep>{};
    if (grp.empty()) {
        for (int i = 0; i < K; ++i) {
            StepSpec sp{};
            sp.n = n;
            sp.precision = rs.precision;
            sp.gdt = rs.gdt[i];
            sp.phase_tab = rs.phase_tab;
            sp.ecache = rs.ecache;
            if ((e = reverse_step_impl(phi, chi, sp, tun, gpart, rs.grad_slots[i], st)) != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const bool two = grp.size() == 2;
    const HighSweep A = two ? grp[0] : HighSweep{};
    const HighSweep& B = grp.back();
    // complex64 runs use the tan butterfly when every step allows it: 4 FMAs per
    // pair instead of 8 operations, with the step's c^n and the rounding
    // feedback in the phase table and c^2m weights on the gradient terms
    bool tan = std::is_same_v<V, float2> && rs.fb_applied != nullptr;
    for (int i = 0; tan && i < K; ++i) tan = mix_coef(n, rs.gdt[i]).tan_form;
    SweepArgs a{};
    a.tab = tan ? rs.fb_applied : rs.phase_tab;
    a.n = n;
    a.psi2 = chi;
    // every pass over one tile kind (LOW, group A, group B) uses one workgroup
    // count, so each kind's partials slot of a region has a fixed size
    int gL = 0, gA = 0, gB = 0;
    reverse_grids(dim, C::TB, C::HTB, tun, &gL, &gA, &gB);
    const int offA = gL, offB = gL + (two ? gA : 0), GS = offB + gB;
    auto region = [&](int i) { return gpart + static_cast<size_t>(i & 1) * static_cast<size_t>(GS); };
    // coefficients of step i: (c, s), or (unused, t) with c^2 = 1 / (1 + t_f^2)
    // for the FP32 tangent the kernel uses
    auto coef = [&](int i, double* bc, double* bs) {
        if (tan) {
            const double t = mix_coef(n, rs.gdt[i]).bs;  // as exact_scale rounds it
            const double tf = static_cast<double>(static_cast<float>(t));
            *bc = 1.0 / (1.0 + tf * tf);
            *bs = t;
        } else {
            *bc = std::cos(rs.gdt[i]);
            *bs = std::sin(rs.gdt[i]);
        }
    };
    // qubits per group and the order the groups of step i are rotated in:
    // two groups, even i: LOW, A, B; odd i: B, A, LOW; one group: even i: H,
    // LOW; odd i: LOW, H (the group a phase pass rotates after its phase comes
    // first in the next step)
    const int nL = C::TB, nA = two ? A.a1 + A.a2 : 0, nB = B.a1 + B.a2;
    // weights of a pass doing `count` rotations of step i after `before` of them
    auto weights = [&](SweepArgs& b, int i, int before, int count) {
        if (!tan) return;
        double c2 = 0.0, t = 0.0;
        coef(i, &c2, &t);
        b.gw = std::pow(c2, before + count - 1);
        b.gic = 1.0 + static_cast<double>(static_cast<float>(t)) * static_cast<double>(static_cast<float>(t));
    };
    const double pass_bytes = 4.0 * static_cast<double>(sizeof(V)) * static_cast<double>(dim);
    auto timed = [&](const char* name, auto&& launch) -> hipError_t {
        if (tun.timer) tun.timer->begin(st);
        const hipError_t r = launch();
        if (r == hipSuccess && tun.timer) tun.timer->end(st, name, pass_bytes);
        return r;
    };
    const bool ldstab = std::is_same_v<V, float2> && tun.lds_tab13 && n <= kLdsTabMaxN;
    auto low_rev = [&](bool phase, const SweepArgs& b) -> hipError_t {
        if constexpr (std::is_same_v<V, float2>) {
            if (tan && phase && ldstab) return launch_low_rev<V, kRevPass | F_TAN | F_LDSTAB>(phi, b, gL, st);
            if (tan) return phase ? launch_low_rev<V, kRevPass | F_TAN>(phi, b, gL, st)
                                  : launch_low_rev<V, F_GRAD | F_TAN>(phi, b, gL, st);
        }
        return phase ? launch_low_rev<V, kRevPass>(phi, b, gL, st) : launch_low_rev<V, F_GRAD>(phi, b, gL, st);
    };
    // phase passes run on the group's own tile (the narrow phase-carrying
    // group included); plain passes on 2^HTB-amplitude tiles
    auto high_rev = [&](bool phase, const HighSweep& hs, const SweepArgs& b) -> hipError_t {
        const int R = hs.R, htb = group_tb(hs);
        const int gH = &hs == &B ? gB : gA;
        if (!phase && htb != C::HTB) return hipErrorInvalidValue;
        if constexpr (std::is_same_v<V, float2>) {
            if (tan && phase && ldstab && htb == kNarrowTB)
                return launch_high_rev_ppass<V, kRevPass | F_TAN | F_LDSTAB>(phi, htb, R, b, gH, st);
            if (tan) return phase ? launch_high_rev_ppass<V, kRevPass | F_TAN>(phi, htb, R, b, gH, st)
                                  : launch_high_rev<V, F_GRAD | F_TAN>(phi, R, b, gH, st);
        }
        return phase ? launch_high_rev_ppass<V>(phi, htb, R, b, gH, st) : launch_high_rev<V>(phi, R, b, gH, st);
    };
    if (two) {
        SweepArgs b = a;
        b.ntiles = dim >> C::TB;
        coef(0, &b.bc, &b.bs);
        weights(b, 0, 0, nL);
        b.partials = region(0);
        e = timed("adjoint_open", [&] { return low_rev(false, b); });
    } else {
        SweepArgs b = high_args(a, B, dim);
        coef(0, &b.bc, &b.bs);
        weights(b, 0, 0, nB);
        b.partials = region(0) + offB;
        e = timed("adjoint_open", [&] { return high_rev(false, B, b); });
    }
    if (e != hipSuccess) return e;
    // step i's gradient sum: by the first workgroup of step i+1's A pass (two
    // groups; it runs before the phase pass that rewrites region i & 1), else
    // by k_finish_sum
    int fin_step = -1;
    auto finish_alone = [&]() -> hipError_t {
        if (fin_step < 0) return hipSuccess;
        hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kThreads), 0, st, region(fin_step), GS,
                           rs.grad_slots[fin_step]);
        fin_step = -1;
        return hipGetLastError();
    };
    for (int i = 0; i < K; ++i) {
        const bool even = (i & 1) == 0;
        if (two) {
            SweepArgs b = high_args(a, A, dim);
            coef(i, &b.bc, &b.bs);
            weights(b, i, even ? nL : nB, nA);
            b.partials = region(i) + offA;
            if (fin_step >= 0 && tun.fold_finish) {
                b.fin_partials = region(fin_step);
                b.fin_count = GS;
                b.fin_out = rs.grad_slots[fin_step];
                b.fin_single = 1;
                fin_step = -1;
            }
            if ((e = finish_alone()) != hipSuccess) return e;
            if ((e = timed("adjoint_high_a", [&] { return high_rev(false, A, b); })) != hipSuccess) return e;
        }
        if ((e = finish_alone()) != hipSuccess) return e;
        // step i's table: P(-dt) times the exact c^n of its rotations, with
        // feedback; its own kernel, or (two table buffers) written by the
        // previous step's phase pass
        const bool fold = tan && rs.fb_applied2 != nullptr;
        void* tab_i = fold && (i & 1) ? rs.fb_applied2 : rs.fb_applied;
        if (tan && (!fold || i == 0)) {
            const int fg = std::max(1, std::min((rs.fb_count + kThreads - 1) / kThreads, 64));
            hipLaunchKernelGGL(k_phase_feedback<V>, dim3(fg), dim3(kThreads), 0, st,
                               reinterpret_cast<const double2*>(rs.fb_master), reinterpret_cast<double2*>(rs.fb_carry),
                               static_cast<V*>(tab_i), rs.fb_count, exact_scale<float>(mix_coef(n, rs.gdt[i]), n));
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        const bool low = two ? !even : even;
        SweepArgs b = low ? a : high_args(a, B, dim);
        if (low) b.ntiles = dim >> C::TB;
        b.ecache = low ? rs.ecache : rs.ecache_high;
        if (tan) b.tab = tab_i;
        if (fold && i + 1 < K) {
            b.fb_next = (i & 1) ? rs.fb_applied : rs.fb_applied2;
            b.fb_master = rs.fb_master;
            b.fb_carry = rs.fb_carry;
            b.fb_count = rs.fb_count;
            b.fb_scale = exact_scale<float>(mix_coef(n, rs.gdt[i + 1]), n);
        }
        coef(i, &b.bc, &b.bs);
        // the phase pass's group is the last of step i
        const int cnt = low ? nL : nB;
        weights(b, i, two ? (even ? nL + nA : nB + nA) : (even ? nB : nL), cnt);
        if (i + 1 < K) {
            coef(i + 1, &b.bc2, &b.bs2);
            SweepArgs w{};
            weights(w, i + 1, 0, cnt);  // the group opens step i + 1
            b.gw2 = w.gw;
            b.gic2 = w.gic;
        } else {
            b.bc2 = 1.0;  // identity after the last phase (t = 0 or (c, s) = (1, 0))
            b.bs2 = 0.0;
            b.gw2 = 1.0;
            b.gic2 = 1.0;
        }
        const int off = low ? 0 : offB;
        b.partials = region(i) + off;
        b.partials2 = region(i + 1) + off;
        e = low ? timed("adjoint_low_phase", [&] { return low_rev(true, b); })
                : timed("adjoint_high_phase", [&] { return high_rev(true, B, b); });
        if (e != hipSuccess) return e;
        fin_step = i;
    }
    return finish_alone();
}

template <typename V, int FL, int R = kFusedRMin>
hipError_t launch_high_batch_fl(int rbits, V* psi, const SweepArgs& a, int grid, hipStream_t st)
{
    if constexpr (R > kFusedRMax || R >= Cfg<V>::HTB) {
        return hipErrorInvalidValue;
    } else {
        if (rbits == R) return launch_high_fl<V, R, FL>(psi, a, grid, st);
        return launch_high_batch_fl<V, FL, R + 1>(rbits, psi, a, grid, st);
    }
}

template <typename V>
hipError_t batch_pass_impl(const BatchGroup& g, const BatchPass& bp, const SweepTuning& tun, double* partials,
                           hipStream_t st)
{
    using C = Cfg<V>;
    const int n = g.n;
    SweepArgs a{};
    a.tab = g.tab;
    a.tab_stride = g.tab_stride;
    a.n = n;
    a.bc_arr = g.bc;
    a.bs_arr = g.bs;
    a.s0_arr = g.s0;
    a.e0_arr = g.e0;
    a.partials = partials;
    a.dst = g.psi_out;
    if (g.tab_next && (bp.kind == BatchPass::kLowPhase || bp.kind == BatchPass::kHighPhase)) {
        a.fb_next = g.tab_next;
        a.fb_master = g.fb_master;
        a.fb_carry = g.fb_carry;
        a.fb_count = g.tab_stride;
        a.fb_sd_arr = g.fb_sd;
        a.fb_nstates = g.nstates;
    }
    V* psi = static_cast<V*>(g.psi);
    const bool low = bp.kind == BatchPass::kLowPlain || bp.kind == BatchPass::kLowPhase;
    if (low) {
        a.ecache = g.ecache;
        a.inst_shift = n - C::TB;
        a.ntiles = static_cast<uint64_t>(g.nstates) << (n - C::TB);
    } else {
        a.h = bp.hs.h1;
        a.a1 = bp.hs.a1;
        a.h2 = bp.hs.h2;
        a.lo_bits = bp.hs.h1 - bp.hs.R;
        a.mid_bits = bp.hs.h2 - (bp.hs.h1 + bp.hs.a1);
        a.ecache = bp.ecache_high;
        a.inst_shift = n - C::HTB;
        a.ntiles = static_cast<uint64_t>(g.nstates) << (n - C::HTB);
    }
    const int grid = clamp_grid(a.ntiles, tun.grid);
    constexpr int BP = F_BATCH | F_TAN | F_PHASE | F_ECACHE;
    const int fl = (bp.measure ? F_MEASURE : 0) | (bp.post ? F_POSTROT : 0);
    switch (bp.kind) {
    case BatchPass::kLowPlain: return launch_low_batch<V, F_BATCH | F_TAN>(psi, a, grid, st);
    case BatchPass::kHighPlain: return launch_high_batch<V, true>(psi, bp.hs.R, a, grid, st);
    case BatchPass::kLowPhase:
        switch (fl) {
        case 0: return launch_low_batch<V, BP>(psi, a, grid, st);
        case F_MEASURE: return launch_low_batch<V, BP | F_MEASURE>(psi, a, grid, st);
        case F_POSTROT: return launch_low_batch<V, BP | F_POSTROT>(psi, a, grid, st);
        default: return launch_low_batch<V, BP | F_MEASURE | F_POSTROT>(psi, a, grid, st);
        }
    case BatchPass::kHighPhase:
        switch (fl) {
        case 0: return launch_high_batch_fl<V, BP>(bp.hs.R, psi, a, grid, st);
        case F_MEASURE: return launch_high_batch_fl<V, BP | F_MEASURE>(bp.hs.R, psi, a, grid, st);
        case F_POSTROT: return launch_high_batch_fl<V, BP | F_POSTROT>(bp.hs.R, psi, a, grid, st);
        default: return launch_high_batch_fl<V, BP | F_MEASURE | F_POSTROT>(bp.hs.R, psi, a, grid, st);
        }
    }
    return hipErrorInvalidValue;
}

template <typename V>
hipError_t phase_from_impl(V* dst, const V* src, int n, const void* tab, double* partials, double* out,
                           hipStream_t st)
{
    const uint64_t dim = 1ull << n;
    uint64_t chunk = 4096;
    while ((dim + chunk - 1) / chunk > 4096) chunk <<= 1;
    const int grid = static_cast<int>((dim + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_phase_from<V>, dim3(grid), dim3(kThreads), 0, st, dst, src, dim, chunk, tab, n, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kThreads), 0, st, partials, grid, out);
    return hipGetLastError();
}

}  // namespace

SweepTuning default_tuning(int n, int precision)
{
    SweepTuning t{};
    t.tile_bits = precision == kC64 ? kLowTB64 : kLowTB128;
    t.row_bits = precision == kC64 ? 5 : 4;
    t.grid = 4096;
    t.pingpong = false;
    t.wide_low = true;
    // narrow phase group: N=26 complex64 B phase pass 0.283 -> 0.237 ms (step
    // 0.460 -> 0.436 ms), N=27 0.485 -> 0.465, N=28 0.918 -> 0.899; the
    // adjoint's reverse B phase pass at N=26 0.513 -> 0.460 ms
    t.narrow_phase = true;
    // LDS phase table for the 2^13-amplitude phase passes: N=26 complex64 step
    // 0.4255 -> 0.4116 ms, adjoint forward 0.531 -> 0.520 s and backward
    // 1.145 -> 1.134 s; N=27 and N=24 unchanged, N=25, 28-29 and 20 0.5-3 %
    // slower with it
    // complex128 high phase passes (one workgroup per CU either way): N=22-28
    // 0.6-1.0 % faster per step with the table in LDS, N=29 unchanged; N=30-32
    // with its first 2032 entries in LDS the B phase pass runs 1-2 % faster
    t.lds_tab13 = precision == kC64 ? (n == 26 || n == 27) : n >= 22;
    t.prefetch_low = false;
    t.prefetch_high = false;
    // page-balancing granule of the high groups' planner, in amplitudes: 2^20
    // through N=30, 2^24 from N=31.  Complex64: N=30 7.40 ms per step against
    // 7.53-7.58 ms at 2^22-2^24; N=31 15.50 -> 15.16 ms, N=32 33.0 -> 31.8 ms.
    // Complex128: 2^19 and 2^20 equal at N=26-28; N=30 15.24 -> 14.84 ms,
    // N=31 35.2 -> 31.6 ms, N=32 71.7 -> 67.7 ms.
    t.page_bits = n >= 31 ? 24 : 20;
    // nontemporal loads in the complex64 plain high-group pass of a fused run:
    // per step N=27 -1.8 %, N=28 -1.3 %, N=29 -1.1 %, N=30 unchanged, N=31 -1.1 %,
    // N=32 -0.8 %; N=26 +0.4 % (same box)
    t.nt_plain = precision == kC64 && n >= 27;
    // complex64 phase passes from n = 30: the high group's (one workgroup per
    // CU) as one persistent workgroup on each of the 256 CUs, the LOW tile's
    // as 16384 workgroups (step N=30 6.957 -> 6.886 ms, N=31 14.28 -> 14.04 ms,
    // N=32 unchanged; no gain at N=26-29, same box)
    if (precision == kC64 && n >= 30) {
        t.grid_bphase = 256;
        t.grid_lphase = 16384;
    }
    // complex64 reverse runs at n = 26: 512 workgroups for the LOW tile's
    // passes (backward 0.550 -> 0.542 s over 640 steps, same box); slower at
    // n = 24, 25, 27 and 28
    if (precision == kC64 && n == 26) t.rev_grid_low = 512;
    return t;
}

// two regions: a fused run's measured pass k writes region k & 1 while the
// next pass finishes region (k - 1) & 1
size_t partials_bytes(int grid) { return 2 * sizeof(double) * 3 * static_cast<size_t>(std::max(grid, 4096)); }

int max_grid(const SweepTuning& t) { return std::max({t.grid, t.grid_bphase, t.grid_lphase}); }

hipError_t launch_step(void* psi, void* psi_alt, int* result_in_alt, const StepSpec& spec,
                       const SweepTuning& tun, double* partials, hipStream_t stream)
{
    if (spec.n < 1 || spec.n > 32) return hipErrorInvalidValue;
    if (spec.precision == kC64)
        return step_impl(static_cast<float2*>(psi), static_cast<float2*>(psi_alt), result_in_alt, spec, tun,
                         partials, stream);
    if (spec.precision == kC128)
        return step_impl(static_cast<double2*>(psi), static_cast<double2*>(psi_alt), result_in_alt, spec, tun,
                         partials, stream);
    return hipErrorInvalidValue;
}

size_t reverse_partials_count(int n, int precision, const SweepTuning& tun)
{
    const int tb = precision == kC64 ? Cfg<float2>::TB : Cfg<double2>::TB;
    const int htb = precision == kC64 ? Cfg<float2>::HTB : Cfg<double2>::HTB;
    const uint64_t dim = 1ull << n;
    if (n < tb) return static_cast<size_t>(n) * 64;
    size_t total = static_cast<size_t>(clamp_grid(dim >> tb, tun.grid));
    for (size_t i = 0; i < plan_high_sweeps(n, tb, htb, std::max(tun.row_bits, 3), tun.page_bits).size(); ++i)
        total += static_cast<size_t>(clamp_grid(dim >> htb, tun.grid));
    // fused reverse runs: two regions of LOW + two high groups
    int gl = 0, ga = 0, gb = 0;
    reverse_grids(dim, tb, htb, tun, &gl, &ga, &gb);
    const size_t fused = 2 * (static_cast<size_t>(gl) + static_cast<size_t>(ga) + static_cast<size_t>(gb));
    return std::max(total, fused);
}

hipError_t launch_reverse_run(void* phi, void* chi, const RevRunSpec& rs, const SweepTuning& tun, double* gpart,
                              hipStream_t stream)
{
    if (rs.n < 1 || rs.n > 32 || rs.steps < 0) return hipErrorInvalidValue;
    if (rs.precision == kC64)
        return reverse_run_impl(static_cast<float2*>(phi), static_cast<float2*>(chi), rs, tun, gpart, stream);
    if (rs.precision == kC128)
        return reverse_run_impl(static_cast<double2*>(phi), static_cast<double2*>(chi), rs, tun, gpart, stream);
    return hipErrorInvalidValue;
}

hipError_t launch_reverse_step(void* phi, void* chi, const StepSpec& spec, const SweepTuning& tun, double* gpart,
                               double* grad_slot, hipStream_t stream)
{
    if (spec.n < 1 || spec.n > 32) return hipErrorInvalidValue;
    if (spec.precision == kC64)
        return reverse_step_impl(static_cast<float2*>(phi), static_cast<float2*>(chi), spec, tun, gpart, grad_slot,
                                 stream);
    if (spec.precision == kC128)
        return reverse_step_impl(static_cast<double2*>(phi), static_cast<double2*>(chi), spec, tun, gpart,
                                 grad_slot, stream);
    return hipErrorInvalidValue;
}

hipError_t launch_adj_init(const void* src, void* phi, void* chi, int n, int precision, const void* tab, uint32_t s0,
                           uint32_t e0, double wp, double wd, hipStream_t stream)
{
    if (precision == kC64)
        return adj_init_impl(static_cast<const float2*>(src), static_cast<float2*>(phi), static_cast<float2*>(chi), n,
                             tab, s0, e0, wp, wd, stream);
    if (precision == kC128)
        return adj_init_impl(static_cast<const double2*>(src), static_cast<double2*>(phi), static_cast<double2*>(chi),
                             n, tab, s0, e0, wp, wd, stream);
    return hipErrorInvalidValue;
}

hipError_t launch_phase_from(void* dst, const void* src, int n, int precision, const void* tab, double* partials,
                             double* out, hipStream_t stream)
{
    if (precision == kC64)
        return phase_from_impl(static_cast<float2*>(dst), static_cast<const float2*>(src), n, tab, partials, out,
                               stream);
    if (precision == kC128)
        return phase_from_impl(static_cast<double2*>(dst), static_cast<const double2*>(src), n, tab, partials, out,
                               stream);
    return hipErrorInvalidValue;
}

int fused_low_bits(int n, int precision, const SweepTuning& tun)
{
    if (precision == kC64) return fused_ltb<float2>(n, tun);
    return Cfg<double2>::TB;
}

std::vector<HighSweep> fused_sweeps(int n, int precision, const SweepTuning& tun, int low_bits)
{
    if (precision == kC64) return fused_groups<float2>(n, tun, low_bits);
    if (precision == kC128) return fused_groups<double2>(n, tun, low_bits);
    return {};
}

hipError_t launch_run(void* psi, const RunSpec& rs, const SweepTuning& tun, double* partials, hipStream_t stream)
{
    if (rs.n < 1 || rs.n > 32 || rs.steps < 0) return hipErrorInvalidValue;
    if (rs.precision == kC64) return run_impl(static_cast<float2*>(psi), rs, tun, partials, stream);
    if (rs.precision == kC128) return run_impl(static_cast<double2*>(psi), rs, tun, partials, stream);
    return hipErrorInvalidValue;
}

template <class C, int HTB = C::HTB, int R = kFusedRMin, typename G>
void fill_high_r(int rbits, G& go)
{
    constexpr int RMAX = HTB == C::HTB ? kFusedRMax : kNarrowRMax;
    if constexpr (R <= RMAX && R < HTB) {
        if (rbits == R) go(k_energy_fill_high<HighPlan<HTB, C::HRB, R>>);
        else fill_high_r<C, HTB, R + 1>(rbits, go);
    }
}

hipError_t launch_energy_fill_high(uint32_t* out, int n, int precision, const HighSweep& hs, hipStream_t stream)
{
    const uint64_t dim = 1ull << n;
    if (n < 3) return hipErrorInvalidValue;  // codes go in groups of eight
    const int grid = static_cast<int>(std::min<uint64_t>((dim / 8 + kThreads - 1) / kThreads, 65536));
    const int lo = hs.h1 - hs.R, mid = hs.h2 - (hs.h1 + hs.a1);
    hipError_t e = hipErrorInvalidValue;
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), 0, stream, out, n, hs.h1, hs.a1, hs.h2, lo, mid, dim);
        e = hipGetLastError();
    };
    const int htb = group_tb(hs);
    if (precision == kC64) {
        if (htb == Cfg<float2>::HTB) fill_high_r<Cfg<float2>>(hs.R, go);
        else if (htb == kNarrowTB) fill_high_r<Cfg<float2>, kNarrowTB, kNarrowRMin>(hs.R, go);
    } else if (htb == Cfg<double2>::HTB) {
        fill_high_r<Cfg<double2>>(hs.R, go);
    }
    return e;
}

MixInfo mix_info(int n, double gdt, int precision)
{
    const MixCoef mc = mix_coef(n, gdt);
    MixInfo m{};
    m.tan_form = mc.tan_form;
    m.bc = mc.bc;
    m.bs = mc.bs;
    m.sd = precision == kC64 ? exact_scale<float>(mc, n) : exact_scale<double>(mc, n);
    return m;
}

int low_tile_bits(int precision) { return precision == kC64 ? Cfg<float2>::TB : Cfg<double2>::TB; }

hipError_t launch_batch_step(const BatchGroup& g, const SweepTuning& tun, double* partials, hipStream_t stream)
{
    if (g.n < 1 || g.n > 32 || g.nstates < 1) return hipErrorInvalidValue;
    if (g.precision == kC64) return batch_step_impl<float2>(g, tun, partials, stream);
    if (g.precision == kC128) return batch_step_impl<double2>(g, tun, partials, stream);
    return hipErrorInvalidValue;
}

hipError_t launch_batch_pass(const BatchGroup& g, const BatchPass& bp, const SweepTuning& tun, double* partials,
                             hipStream_t stream)
{
    if (g.n < 1 || g.n > 32 || g.nstates < 1) return hipErrorInvalidValue;
    if (g.precision == kC64) return batch_pass_impl<float2>(g, bp, tun, partials, stream);
    if (g.precision == kC128) return batch_pass_impl<double2>(g, bp, tun, partials, stream);
    return hipErrorInvalidValue;
}

hipError_t launch_batch_feedback(const double* master, double* carry, void* applied, int count, int nstates,
                                 const double* sd, int precision, hipStream_t stream)
{
    const int64_t total = static_cast<int64_t>(count) * nstates;
    const int g = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((total + kThreads - 1) / kThreads, 2048)));
    if (precision == kC64)
        hipLaunchKernelGGL(k_phase_feedback_batch<float2>, dim3(g), dim3(kThreads), 0, stream,
                           reinterpret_cast<const double2*>(master), reinterpret_cast<double2*>(carry),
                           static_cast<float2*>(applied), count, nstates, sd);
    else
        hipLaunchKernelGGL(k_phase_feedback_batch<double2>, dim3(g), dim3(kThreads), 0, stream,
                           reinterpret_cast<const double2*>(master), reinterpret_cast<double2*>(carry),
                           static_cast<double2*>(applied), count, nstates, sd);
    return hipGetLastError();
}

hipError_t launch_batch_finish(const double* partials, int tiles_per_state, int nstates, double* out,
                               hipStream_t stream)
{
    hipLaunchKernelGGL(k_finish3_batch, dim3(nstates), dim3(kThreads), 0, stream, partials, tiles_per_state, out);
    return hipGetLastError();
}

hipError_t launch_batch_init(void* psi, int n, int nstates, int precision, const uint32_t* s0, const double* re,
                             const double* im, hipStream_t stream)
{
    const int g = std::max(1, std::min((nstates + kThreads - 1) / kThreads, 1024));
    if (precision == kC64)
        hipLaunchKernelGGL(k_set_basis_batch<float2>, dim3(g), dim3(kThreads), 0, stream, static_cast<float2*>(psi), n,
                           nstates, s0, re, im);
    else
        hipLaunchKernelGGL(k_set_basis_batch<double2>, dim3(g), dim3(kThreads), 0, stream, static_cast<double2*>(psi),
                           n, nstates, s0, re, im);
    return hipGetLastError();
}

template <typename V>
static hipError_t measure_impl(V* psi, int n, uint32_t s0, uint32_t e0, double* slot, double* partials,
                               hipStream_t st)
{
    const uint64_t dim = 1ull << n;
    uint64_t chunk = 4096;
    while ((dim + chunk - 1) / chunk > 4096) chunk <<= 1;
    const int grid = static_cast<int>((dim + chunk - 1) / chunk);
    hipLaunchKernelGGL((k_elementwise<V, F_MEASURE>), dim3(grid), dim3(kThreads), 0, st, psi, dim, chunk,
                       nullptr, 1.0, n, s0, e0, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finish3, dim3(1), dim3(kThreads), 0, st, partials, grid, slot);
    return hipGetLastError();
}

hipError_t launch_measure(const void* psi, int n, int precision, uint32_t s0, uint32_t e0,
                          double* series_slot, double* partials, int grid, hipStream_t stream)
{
    (void)grid;
    if (precision == kC64)
        return measure_impl(static_cast<float2*>(const_cast<void*>(psi)), n, s0, e0, series_slot, partials,
                            stream);
    if (precision == kC128)
        return measure_impl(static_cast<double2*>(const_cast<void*>(psi)), n, s0, e0, series_slot, partials,
                            stream);
    return hipErrorInvalidValue;
}

template <typename V>
static hipError_t phase_impl(V* psi, int n, const void* tab, uint32_t s0, uint32_t e0, double* slot,
                             double* partials, hipStream_t st)
{
    const uint64_t dim = 1ull << n;
    uint64_t chunk = 4096;
    while ((dim + chunk - 1) / chunk > 4096) chunk <<= 1;
    const int grid = static_cast<int>((dim + chunk - 1) / chunk);
    if (slot) {
        hipLaunchKernelGGL((k_elementwise<V, F_PHASE | F_MEASURE>), dim3(grid), dim3(kThreads), 0, st, psi, dim,
                           chunk, tab, 1.0, n, s0, e0, partials);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_finish3, dim3(1), dim3(kThreads), 0, st, partials, grid, slot);
    } else {
        hipLaunchKernelGGL((k_elementwise<V, F_PHASE>), dim3(grid), dim3(kThreads), 0, st, psi, dim, chunk, tab,
                           1.0, n, 0u, 0u, nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_phase(void* psi, int n, int precision, const void* phase_tab, uint32_t s0, uint32_t e0,
                        double* series_slot, double* partials, hipStream_t stream)
{
    if (precision == kC64)
        return phase_impl(static_cast<float2*>(psi), n, phase_tab, s0, e0, series_slot, partials, stream);
    if (precision == kC128)
        return phase_impl(static_cast<double2*>(psi), n, phase_tab, s0, e0, series_slot, partials, stream);
    return hipErrorInvalidValue;
}

hipError_t launch_init_basis(void* psi, int n, int precision, uint64_t s0, double re, double im,
                             hipStream_t stream)
{
    const size_t esz = precision == kC64 ? sizeof(float2) : sizeof(double2);
    hipError_t e = hipMemsetAsync(psi, 0, esz << n, stream);
    if (e != hipSuccess) return e;
    if (precision == kC64) hipLaunchKernelGGL(k_set_basis<float2>, dim3(1), dim3(1), 0, stream,
                                              static_cast<float2*>(psi), s0, re, im);
    else hipLaunchKernelGGL(k_set_basis<double2>, dim3(1), dim3(1), 0, stream, static_cast<double2*>(psi), s0,
                            re, im);
    return hipGetLastError();
}

hipError_t launch_carry_reset(double* carry, int count, hipStream_t stream)
{
    const int g = std::max(1, std::min((count + kThreads - 1) / kThreads, 64));
    hipLaunchKernelGGL(k_carry_reset, dim3(g), dim3(kThreads), 0, stream, reinterpret_cast<double2*>(carry), count);
    return hipGetLastError();
}

hipError_t launch_energy_fill(uint32_t* out, int n, int precision, hipStream_t stream, int low_bits)
{
    const uint64_t dim = 1ull << n;
    if (n < 3) return hipErrorInvalidValue;  // codes go in groups of eight
    const int grid = static_cast<int>(std::min<uint64_t>((dim / 8 + kThreads - 1) / kThreads, 65536));
    auto k64 = k_energy_fill<LowPlan<Cfg<float2>::TB, Cfg<float2>::RB>>;
    auto k64w = k_energy_fill<LowPlan<kWideLowTB, Cfg<float2>::RB>>;
    auto k128 = k_energy_fill<LowPlan<Cfg<double2>::TB, Cfg<double2>::RB>>;
    const int def = precision == kC64 ? Cfg<float2>::TB : Cfg<double2>::TB;
    if (low_bits > 0 && low_bits != def && !(precision == kC64 && low_bits == kWideLowTB)) return hipErrorInvalidValue;
    if (precision == kC64)
        hipLaunchKernelGGL(low_bits == kWideLowTB ? k64w : k64, dim3(grid), dim3(kThreads), 0, stream, out, n, dim);
    else
        hipLaunchKernelGGL(k128, dim3(grid), dim3(kThreads), 0, stream, out, n, dim);
    return hipGetLastError();
}

hipError_t launch_count_below(int n, uint32_t e0, unsigned long long* d_count, unsigned long long* partials,
                              int grid, hipStream_t stream)
{
    (void)grid;
    const uint64_t dim = 1ull << n;
    uint64_t chunk = 4096;
    while ((dim + chunk - 1) / chunk > 4096) chunk <<= 1;
    const int g = static_cast<int>((dim + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_count_below, dim3(g), dim3(kThreads), 0, stream, n, e0, dim, chunk, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finish_u64, dim3(1), dim3(kThreads), 0, stream, partials, g, d_count);
    return hipGetLastError();
}

}  // namespace bk
