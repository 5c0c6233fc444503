// bk_linalg.hip -- gfx950 kernels around the Trotter sweeps:
//  * projective sampling by inverse CDF over |psi|^2 (replaces the cuStateVec
//    sampler calls of sampler.cuh:464-471, 564-573): FP64 block sums, one
//    workgroup exclusive scan, one wave per shot for the search;
//  * matrix-free H.v (diagonalizer.cuh:60-77): y = d*x - g sum_j X_j x with
//    d = E/N, tiled like the sweeps (a LOW pass with the diagonal and qubits
//    [0,TB), HIGH passes accumulating the remaining qubits);
//  * FP64 BLAS-1 with fixed-order reductions for Lanczos/Krylov (replaces the
//    cuBLAS calls of diagonalizer.cuh:449-538).
#include "bk_kernels.hpp"
#include "bk_plan.hpp"
#include "bk_sweeps.hpp"

#include <algorithm>

namespace bk {

__device__ __forceinline__ uint32_t labs_energy_u32_l(uint32_t s, int n)
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

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename V>
__device__ __forceinline__ double prob(const V& x)
{
    const double a = static_cast<double>(x.x), b = static_cast<double>(x.y);
    return fma(a, a, b * b);
}

// ---------------------------------------------------------------------------
// Sampling
// ---------------------------------------------------------------------------
template <typename V>
__global__ __launch_bounds__(kThreads) void k_block_prob(const V* __restrict__ psi, double* __restrict__ sums)
{
    __shared__ double sc[4];
    const uint64_t beg = static_cast<uint64_t>(blockIdx.x) << kSampleBlockBits;
    double a = 0.0;
    for (int i = threadIdx.x; i < (1 << kSampleBlockBits); i += kThreads) a += prob(psi[beg + i]);
    a = wsum(a);
    if ((threadIdx.x & 63) == 0) sc[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = (sc[0] + sc[1]) + (sc[2] + sc[3]);
}

// exclusive scan of nb block sums into prefix[0..nb] (prefix[nb] = total)
__global__ __launch_bounds__(kThreads) void k_scan(const double* __restrict__ sums, int nb,
                                                  double* __restrict__ prefix)
{
    __shared__ double seg[kThreads];
    const int per = (nb + kThreads - 1) / kThreads;
    const int b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
    double s = 0.0;
    for (int i = b0; i < b1; ++i) s += sums[i];
    seg[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int i = 0; i < kThreads; ++i) {
            const double v = seg[i];
            seg[i] = run;
            run += v;
        }
        prefix[nb] = run;
    }
    __syncthreads();
    double run = seg[threadIdx.x];
    for (int i = b0; i < b1; ++i) {
        prefix[i] = run;
        run += sums[i];
    }
}

// one wave per shot: index = min{ i : cdf_i > r * total }, or dim-1 if none
template <typename V>
__global__ __launch_bounds__(kThreads) void k_search(const V* __restrict__ psi, int nb,
                                                    const double* __restrict__ prefix,
                                                    const double* __restrict__ rnd, int shots,
                                                    int64_t* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int shot = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (shot >= shots) return;
    const double total = prefix[nb];
    const double target = rnd[shot] * total;
    // smallest b with prefix[b+1] > target
    int lo = 0, hi = nb;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid + 1] > target) hi = mid;
        else lo = mid + 1;
    }
    const int64_t dim = static_cast<int64_t>(nb) << kSampleBlockBits;
    int64_t found = -1;
    for (int b = lo; b < nb && found < 0; ++b) {
        double run = prefix[b];
        const int64_t beg = static_cast<int64_t>(b) << kSampleBlockBits;
        for (int c = 0; c < (1 << kSampleBlockBits) && found < 0; c += 64) {
            double v = prob(psi[beg + c + lane]);
            // inclusive wave scan
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const double u = __shfl_up(v, off, 64);
                if (lane >= off) v += u;
            }
            const double cum = run + v;
            const unsigned long long hit = __ballot(cum > target);
            if (hit) found = beg + c + __ffsll(static_cast<long long>(hit)) - 1;
            run = __shfl(cum, 63, 64);
        }
    }
    if (found < 0) found = dim - 1;
    if (lane == 0) out[shot] = found;
}

size_t sample_scratch_doubles(int n)
{
    const int bits = std::max(n - kSampleBlockBits, 0);
    return 2 * ((size_t(1) << bits) + 1);
}

template <typename V>
__global__ void k_search_small(const V* __restrict__ psi, int64_t dim, const double* __restrict__ rnd,
                               int shots, int64_t* __restrict__ out)
{
    // tiny states (< one block): sequential FP64 CDF per shot
    const int shot = blockIdx.x * blockDim.x + threadIdx.x;
    if (shot >= shots) return;
    double total = 0.0;
    for (int64_t i = 0; i < dim; ++i) total += prob(psi[i]);
    const double target = rnd[shot] * total;
    double run = 0.0;
    int64_t idx = dim - 1;
    for (int64_t i = 0; i < dim; ++i) {
        run += prob(psi[i]);
        if (run > target) {
            idx = i;
            break;
        }
    }
    out[shot] = idx;
}

template <typename V>
static hipError_t sample_impl(const V* psi, int n, const double* rnd, int shots, int64_t* out, double* sums,
                              double* prefix, hipStream_t st)
{
    if (shots <= 0) return hipSuccess;
    if (n < kSampleBlockBits) {
        hipLaunchKernelGGL(k_search_small<V>, dim3((shots + 63) / 64), dim3(64), 0, st, psi, int64_t(1) << n,
                           rnd, shots, out);
        return hipGetLastError();
    }
    const int nb = 1 << (n - kSampleBlockBits);
    hipLaunchKernelGGL(k_block_prob<V>, dim3(nb), dim3(kThreads), 0, st, psi, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kThreads), 0, st, sums, nb, prefix);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int wpb = kThreads / 64;
    hipLaunchKernelGGL(k_search<V>, dim3((shots + wpb - 1) / wpb), dim3(kThreads), 0, st, psi, nb, prefix, rnd,
                       shots, out);
    return hipGetLastError();
}

hipError_t launch_sample(const void* psi, int n, int precision, const double* d_rand, int shots, int64_t* d_out,
                         double* d_block_sums, double* d_block_prefix, hipStream_t stream)
{
    if (precision == kC64)
        return sample_impl(static_cast<const float2*>(psi), n, d_rand, shots, d_out, d_block_sums,
                           d_block_prefix, stream);
    if (precision == kC128)
        return sample_impl(static_cast<const double2*>(psi), n, d_rand, shots, d_out, d_block_sums,
                           d_block_prefix, stream);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Overlap of two states
// ---------------------------------------------------------------------------
template <typename VA, typename VB>
__global__ __launch_bounds__(kThreads) void k_overlap(const VA* __restrict__ a, const VB* __restrict__ b, uint64_t dim,
                                                      double* __restrict__ partials)
{
    __shared__ double sc[2][kThreads / 64];
    const uint64_t chunk = (dim + gridDim.x - 1) / gridDim.x;
    const uint64_t beg = static_cast<uint64_t>(blockIdx.x) * chunk;
    const uint64_t end = min(dim, beg + chunk);
    double re = 0.0, im = 0.0;
    for (uint64_t i = beg + threadIdx.x; i < end; i += kThreads) {
        const double ar = static_cast<double>(a[i].x), ai = static_cast<double>(a[i].y);
        const double br = static_cast<double>(b[i].x), bi = static_cast<double>(b[i].y);
        re = fma(ar, br, fma(ai, bi, re));
        im = fma(ar, bi, fma(-ai, br, im));
    }
    re = wsum(re);
    im = wsum(im);
    if ((threadIdx.x & 63) == 0) {
        sc[0][threadIdx.x >> 6] = re;
        sc[1][threadIdx.x >> 6] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = (sc[0][0] + sc[0][1]) + (sc[0][2] + sc[0][3]);
        partials[2 * blockIdx.x + 1] = (sc[1][0] + sc[1][1]) + (sc[1][2] + sc[1][3]);
    }
}

template <typename VA>
static hipError_t overlap_b(const VA* a, const void* b, int prec_b, uint64_t dim, double* partials, int grid,
                            hipStream_t st)
{
    if (prec_b == kC64)
        hipLaunchKernelGGL((k_overlap<VA, float2>), dim3(grid), dim3(kThreads), 0, st, a,
                           static_cast<const float2*>(b), dim, partials);
    else if (prec_b == kC128)
        hipLaunchKernelGGL((k_overlap<VA, double2>), dim3(grid), dim3(kThreads), 0, st, a,
                           static_cast<const double2*>(b), dim, partials);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_overlap(const void* a, int prec_a, const void* b, int prec_b, int n, double* partials, int grid,
                          hipStream_t stream)
{
    static_assert(kThreads == 256, "four waves per workgroup");
    const uint64_t dim = uint64_t(1) << n;
    if (prec_a == kC64) return overlap_b(static_cast<const float2*>(a), b, prec_b, dim, partials, grid, stream);
    if (prec_a == kC128) return overlap_b(static_cast<const double2*>(a), b, prec_b, dim, partials, grid, stream);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// H.v
// ---------------------------------------------------------------------------
struct Dbl1 {
    double v;
};

template <typename T>
__device__ __forceinline__ T tz()
{
    return T{};
}
__device__ __forceinline__ void axpy_(double& y, double a, double x) { y = fma(a, x, y); }
__device__ __forceinline__ void axpy_(double2& y, double a, const double2& x)
{
    y.x = fma(a, x.x, y.x);
    y.y = fma(a, x.y, y.y);
}
__device__ __forceinline__ double scl(double a, double x) { return a * x; }
__device__ __forceinline__ double2 scl(double a, const double2& x) { return make_double2(a * x.x, a * x.y); }

// H.v tiles work on 16-byte units (two FP64 reals, or one complex128): a tile
// is 2^13 units (128 KB of LDS) held by 512 threads, 16 units each, loaded and
// stored with 16-byte accesses by consecutive lanes.  The LOW pass covers the
// element bits [0, HB) (HB = 14 real, 13 complex); each high pass covers up
// to 9 further qubits with rows of at least 16 units (256 B) in memory.
template <typename T>
struct HvCfg;
template <>
struct HvCfg<double> {
    static constexpr int HB = 14, EPU = 2;  // element bits of a tile, elements per unit
};
template <>
struct HvCfg<double2> {
    static constexpr int HB = 13, EPU = 1;
};
// high passes: 2^9 threads x 16 units; the LOW pass: 2^10 threads x 8 units
// (its accumulators and E(s) words take 16 units per thread to 256 VGPRs; at
// 8 units, 4 waves per SIMD hide its two waits on memory per tile: N=30 real
// FP64 H.v 14.65 -> 13.7 ms, complex128 N=28 7.15 -> 6.56 ms).  1024-thread
// high passes measured no faster.
constexpr int kHvUnitBits = 13, kHvThreadBits = 9, kHvRowUnitBits = 4;
constexpr int kHvThreads = 1 << kHvThreadBits, kHvUPT = 1 << (kHvUnitBits - kHvThreadBits);
constexpr int kHvLowThreadBits = 10, kHvLowThreads = 1 << kHvLowThreadBits,
              kHvLowUPT = 1 << (kHvUnitBits - kHvLowThreadBits);

// registers and LDS use a native two-double vector: arrays of it are promoted
// to registers, where arrays of the HIP vector struct stayed in scratch memory
typedef double dv2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ dv2 mkv(double a, double b)
{
    dv2 r;
    r.x = a;
    r.y = b;
    return r;
}
__device__ __forceinline__ dv2 d2_axpy(double a, const dv2& x, const dv2& y)
{
    return mkv(fma(a, x.x, y.x), fma(a, x.y, y.y));
}

// LOW pass: y = d(E) x - g sum_{j<HB} x[s ^ 2^j] + (the high groups' sums in
// acc0, acc1) on contiguous tiles, E(s) from the state-order uint16 cache
template <typename T, int NACC>
__global__ __launch_bounds__(kHvLowThreads) void k_hv_low(const T* __restrict__ x, T* y, uint64_t ntiles, double g,
                                                     double dscale, const uint16_t* __restrict__ ecache,
                                                     const T* acc0, const T* acc1)
{
    constexpr int EPU = HvCfg<T>::EPU;
    __shared__ __align__(16) dv2 lds[1 << kHvUnitBits];
    const uint32_t t = threadIdx.x;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t ubase = tile << kHvUnitBits;  // first unit of the tile
        const dv2* xu = reinterpret_cast<const dv2*>(x) + ubase;
        dv2 v[kHvLowUPT];
#pragma unroll
        for (int k = 0; k < kHvLowUPT; ++k) v[k] = xu[(k << kHvLowThreadBits) + t];
        __syncthreads();  // the previous tile's readers are done
#pragma unroll
        for (int k = 0; k < kHvLowUPT; ++k) lds[(k << kHvLowThreadBits) + t] = v[k];
        __syncthreads();
        // E(s) of the units' elements (two uint16 for real FP64) and the high
        // groups' sums (written by the high passes), which start the
        // accumulators: read once x's registers are free
        uint32_t ew[kHvLowUPT];
#pragma unroll
        for (int k = 0; k < kHvLowUPT; ++k) {
            const uint64_t gu = ubase + (k << kHvLowThreadBits) + t;
            ew[k] = EPU == 2 ? reinterpret_cast<const uint32_t*>(ecache)[gu] : static_cast<uint32_t>(ecache[gu]);
        }
        dv2 ad[kHvLowUPT];
#pragma unroll
        for (int k = 0; k < kHvLowUPT; ++k) {
            ad[k] = mkv(0.0, 0.0);
            if constexpr (NACC >= 1) {
                const dv2 w = reinterpret_cast<const dv2*>(acc0)[ubase + (k << kHvLowThreadBits) + t];
                ad[k].x += w.x;
                ad[k].y += w.y;
            }
            if constexpr (NACC >= 2) {
                const dv2 w = reinterpret_cast<const dv2*>(acc1)[ubase + (k << kHvLowThreadBits) + t];
                ad[k].x += w.x;
                ad[k].y += w.y;
            }
        }
        __builtin_amdgcn_sched_barrier(0);  // these loads stay ahead of the unit loop
        dv2* yu = reinterpret_cast<dv2*>(y) + ubase;
        // the unit and its neighbours come from LDS, so x's registers are free
        // for the accumulators.  Fully unrolled (register arrays indexed by a
        // loop variable would live in scratch memory); the scheduling barrier
        // keeps each unit's LDS reads after the previous unit's, so they are
        // not all hoisted into registers at once.
#pragma unroll
        for (int k = 0; k < kHvLowUPT; ++k) {
            __builtin_amdgcn_sched_barrier(0);
            const uint32_t u = (k << kHvLowThreadBits) + t;
            const dv2 self = lds[u];
#pragma unroll
            for (int j = 0; j < kHvUnitBits; ++j) ad[k] = d2_axpy(-g, lds[u ^ (1u << j)], ad[k]);
            dv2 acc;
            if constexpr (EPU == 2) {
                // elements e0 (x) and e0 + 1 (y): bit 0 pairs them inside the unit
                const uint32_t E0 = ew[k] & 0xFFFFu, E1 = ew[k] >> 16;
                acc.x = fma(-g, self.y, fma(static_cast<double>(E0) * dscale, self.x, ad[k].x));
                acc.y = fma(-g, self.x, fma(static_cast<double>(E1) * dscale, self.y, ad[k].y));
            } else {
                const uint32_t E = ew[k];
                const double d = static_cast<double>(E) * dscale;
                acc = mkv(fma(d, self.x, ad[k].x), fma(d, self.y, ad[k].y));
            }
            yu[u] = acc;
        }
    }
}

// high pass: y = -g sum over the group's qubits of x[s ^ 2^q] (ACC: y += ...),
// so that a pass reads one array where it can (ACC, used only past the
// second group or without scratch, reads y as it goes).  The group is
// a1 element bits at h1 and A - a1 at h2 (the sweeps' planner, which spreads
// the page-crossing qubits over the groups); a tile is 2^(13-A) contiguous
// units (a row) times the 2^A rows of the group.  The tile's x is read before
// the LDS exchange.
template <typename T, int A, bool ACC>
__global__ __launch_bounds__(kHvThreads) void k_hv_high(const T* __restrict__ x, T* __restrict__ y, uint64_t ntiles,
                                                      int h1, int a1, int h2, double g)
{
    constexpr int EPU = HvCfg<T>::EPU;
    constexpr int RU = kHvUnitBits - A;  // unit bits of a row
    constexpr uint32_t rowmask = (1u << RU) - 1u;
    __shared__ __align__(16) dv2 lds[1 << kHvUnitBits];
    const uint32_t t = threadIdx.x;
    const int e = EPU == 2 ? 1 : 0;       // element bits inside a unit
    const int hs1 = h1 - e, hs2 = h2 - e;  // the segments as unit bits
    const int lo_bits = hs1 - RU, mid_bits = hs2 - (hs1 + a1);
    const uint32_t m1 = (1u << a1) - 1u;
    const dv2* xu = reinterpret_cast<const dv2*>(x);
    dv2* yu = reinterpret_cast<dv2*>(y);
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t lo = tile & ((1ull << lo_bits) - 1ull);
        const uint64_t mid = (tile >> lo_bits) & ((1ull << mid_bits) - 1ull);
        const uint64_t hi = tile >> (lo_bits + mid_bits);
        const uint64_t ubase = (lo << RU) | (mid << (hs1 + a1)) | (hi << (hs2 + (A - a1)));
        auto gu = [&](uint32_t u) -> uint64_t {
            const uint32_t ua = u >> RU;
            return ubase + (u & rowmask) + (static_cast<uint64_t>(ua & m1) << hs1) +
                   (static_cast<uint64_t>(ua >> a1) << hs2);
        };
        dv2 v[kHvUPT];
#pragma unroll
        for (int k = 0; k < kHvUPT; ++k) v[k] = xu[gu((k << kHvThreadBits) + t)];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kHvUPT; ++k) lds[(k << kHvThreadBits) + t] = v[k];
        __syncthreads();
        // fully unrolled (register arrays indexed by a loop variable would live
        // in scratch memory); the scheduling barrier keeps each unit's LDS
        // reads after the previous unit's, so they are not all hoisted at once
#pragma unroll
        for (int k = 0; k < kHvUPT; ++k) {
            __builtin_amdgcn_sched_barrier(0);
            const uint32_t u = (k << kHvThreadBits) + t;
            dv2 sm = mkv(0.0, 0.0);
#pragma unroll
            for (int b = RU; b < kHvUnitBits; ++b) {
                const dv2 w = lds[u ^ (1u << b)];
                sm.x += w.x;
                sm.y += w.y;
            }
            const uint64_t gk = gu(u);
            yu[gk] = ACC ? d2_axpy(-g, sm, yu[gk]) : mkv(-g * sm.x, -g * sm.y);
        }
    }
}

template <typename T, int A = 1>
hipError_t launch_hv_high(int a, bool acc, const T* x, T* y, uint64_t ntiles, int h1, int a1, int h2, double g,
                          int grid, hipStream_t st)
{
    if constexpr (A > kHvUnitBits - kHvRowUnitBits) {
        return hipErrorInvalidValue;
    } else {
        if (a != A) return launch_hv_high<T, A + 1>(a, acc, x, y, ntiles, h1, a1, h2, g, grid, st);
        if (acc)
            hipLaunchKernelGGL((k_hv_high<T, A, true>), dim3(grid), dim3(kHvThreads), 0, st, x, y, ntiles, h1, a1, h2,
                               g);
        else
            hipLaunchKernelGGL((k_hv_high<T, A, false>), dim3(grid), dim3(kHvThreads), 0, st, x, y, ntiles, h1, a1,
                               h2, g);
        return hipGetLastError();
    }
}

// small n: one row per thread, as diagonalizer.cuh:60-77
template <typename T>
__global__ void k_hv_rows(const T* __restrict__ x, T* __restrict__ y, int n, double g, double dscale)
{
    const uint64_t dim = 1ull << n;
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= dim) return;
    T acc = scl(static_cast<double>(labs_energy_u32_l(static_cast<uint32_t>(i), n)) * dscale, x[i]);
    for (int j = 0; j < n; ++j) axpy_(acc, -g, x[i ^ (1ull << j)]);
    y[i] = acc;
}

template <typename T>
static hipError_t launch_hv_low(int nacc, const T* x, T* y, uint64_t ntiles, double g, double dscale,
                                const uint16_t* ecache, const T* acc0, const T* acc1, int grid, hipStream_t st)
{
    switch (nacc) {
    case 0:
        hipLaunchKernelGGL((k_hv_low<T, 0>), dim3(grid), dim3(kHvLowThreads), 0, st, x, y, ntiles, g, dscale, ecache,
                           acc0, acc1);
        break;
    case 1:
        hipLaunchKernelGGL((k_hv_low<T, 1>), dim3(grid), dim3(kHvLowThreads), 0, st, x, y, ntiles, g, dscale, ecache,
                           acc0, acc1);
        break;
    default:
        hipLaunchKernelGGL((k_hv_low<T, 2>), dim3(grid), dim3(kHvLowThreads), 0, st, x, y, ntiles, g, dscale, ecache,
                           acc0, acc1);
        break;
    }
    return hipGetLastError();
}

// the high-qubit groups of H.v on n spins (qubits [HB, n) in the fewest groups
// of at most 9, rows of at least 16 units, the page-crossing qubits of 8 MB
// pages spread over the groups)
template <typename T>
static std::vector<HighSweep> hv_groups(int n)
{
    constexpr int EB = HvCfg<T>::EPU == 2 ? 1 : 0;
    constexpr int page_bits = HvCfg<T>::EPU == 2 ? 20 : 19;
    return plan_high_sweeps(n, HvCfg<T>::HB, HvCfg<T>::HB, kHvRowUnitBits + EB, page_bits);
}

// High passes first, each writing its group's sum: the first into y, the
// second into `scratch` (one vector; null: it and any further group add into
// y), further groups adding into the scratch vector; then the LOW pass adds
// d(E) x and its own qubits' terms to them and writes y.
template <typename T>
static hipError_t hv_impl(int n, double g, double dscale, const T* x, T* y, const uint16_t* ecache, T* scratch,
                          hipStream_t st)
{
    constexpr int HB = HvCfg<T>::HB;
    const uint64_t dim = 1ull << n;
    hipError_t e;
    if (n < HB) {
        hipLaunchKernelGGL(k_hv_rows<T>, dim3(static_cast<unsigned>((dim + 255) / 256)), dim3(256), 0, st, x, y, n,
                           g, dscale);
        return hipGetLastError();
    }
    if (!ecache) return hipErrorInvalidValue;  // the tiled passes read E(s) from the cache
    const uint64_t ntiles = dim >> HB;
    const int grid = static_cast<int>(std::min<uint64_t>(ntiles, 4096));
    const std::vector<HighSweep> groups = hv_groups<T>(n);
    for (size_t i = 0; i < groups.size(); ++i) {
        const HighSweep& hs = groups[i];
        T* out = (i >= 1 && scratch) ? scratch : y;
        const bool acc = scratch ? i >= 2 : i >= 1;
        if ((e = launch_hv_high<T>(hs.a1 + hs.a2, acc, x, out, ntiles, hs.h1, hs.a1, hs.h2, g, grid, st)) !=
            hipSuccess)
            return e;
    }
    const int nacc = groups.empty() ? 0 : (scratch && groups.size() >= 2 ? 2 : 1);
    return launch_hv_low<T>(nacc, x, y, ntiles, g, dscale, ecache, y, scratch, grid, st);
}

size_t hv_scratch_doubles(int n, int is_complex)
{
    if (n < 1 || n > 32) return 0;
    const size_t groups = is_complex ? hv_groups<double2>(n).size() : hv_groups<double>(n).size();
    if (groups < 2) return 0;
    return (static_cast<size_t>(1) << n) * (is_complex ? 2 : 1);
}

hipError_t launch_hv(int n, double g, int normalize_by_n, int is_complex, const double* x, double* y,
                     hipStream_t stream, const uint16_t* ecache, double* scratch)
{
    if (n < 1 || n > 32) return hipErrorInvalidValue;
    const double dscale = normalize_by_n ? 1.0 / static_cast<double>(n) : 1.0;
    if (is_complex)
        return hv_impl(n, g, dscale, reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y), ecache,
                       reinterpret_cast<double2*>(scratch), stream);
    return hv_impl(n, g, dscale, x, y, ecache, scratch, stream);
}

__global__ __launch_bounds__(kThreads) void k_energy_natural(uint16_t* __restrict__ out, int n, uint64_t dim)
{
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < dim;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        out[i] = static_cast<uint16_t>(labs_energy_u32_l(static_cast<uint32_t>(i), n));
}

hipError_t launch_energy_natural(uint16_t* out, int n, hipStream_t stream)
{
    if (n < 1 || n > 32) return hipErrorInvalidValue;
    const uint64_t dim = 1ull << n;
    const int grid = static_cast<int>(std::min<uint64_t>((dim + kThreads - 1) / kThreads, 65536));
    hipLaunchKernelGGL(k_energy_natural, dim3(grid), dim3(kThreads), 0, stream, out, n, dim);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// BLAS-1 (FP64, fixed-order reductions)
// ---------------------------------------------------------------------------
constexpr int kBlasGrid = 1024;
size_t blas_partials_doubles() { return 2 * kBlasGrid + 8; }

template <bool CPLX>
__global__ __launch_bounds__(kThreads) void k_dot(int64_t len, const double* __restrict__ x,
                                                 const double* __restrict__ y, int64_t chunk,
                                                 double* __restrict__ partials)
{
    __shared__ double sc[8];
    double a = 0.0, b = 0.0;
    const int64_t beg = blockIdx.x * chunk;
    const int64_t end = min(len, beg + chunk);
    for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) {
        if (CPLX) {
            const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
            a = fma(xr, yr, fma(xi, yi, a));
            b = fma(xr, yi, fma(-xi, yr, b));
        } else {
            a = fma(x[i], y[i], a);
        }
    }
    a = wsum(a);
    if (CPLX) b = wsum(b);
    if ((threadIdx.x & 63) == 0) {
        sc[2 * (threadIdx.x >> 6)] = a;
        sc[2 * (threadIdx.x >> 6) + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = (sc[0] + sc[2]) + (sc[4] + sc[6]);
        partials[2 * blockIdx.x + 1] = (sc[1] + sc[3]) + (sc[5] + sc[7]);
    }
}

__global__ __launch_bounds__(kThreads) void k_dot_finish(const double* __restrict__ partials, int count,
                                                        double* __restrict__ out)
{
    __shared__ double sc[8];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < count; i += kThreads) {
        a += partials[2 * i];
        b += partials[2 * i + 1];
    }
    a = wsum(a);
    b = wsum(b);
    if ((threadIdx.x & 63) == 0) {
        sc[2 * (threadIdx.x >> 6)] = a;
        sc[2 * (threadIdx.x >> 6) + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = (sc[0] + sc[2]) + (sc[4] + sc[6]);
        out[1] = (sc[1] + sc[3]) + (sc[5] + sc[7]);
    }
}

hipError_t launch_dot(int64_t len, int is_complex, const double* x, const double* y, double* out, double* partials,
                      hipStream_t st)
{
    int64_t chunk = 4096;
    while ((len + chunk - 1) / chunk > kBlasGrid) chunk <<= 1;
    const int grid = static_cast<int>(std::max<int64_t>(1, (len + chunk - 1) / chunk));
    if (is_complex)
        hipLaunchKernelGGL(k_dot<true>, dim3(grid), dim3(kThreads), 0, st, len, x, y, chunk, partials);
    else
        hipLaunchKernelGGL(k_dot<false>, dim3(grid), dim3(kThreads), 0, st, len, x, y, chunk, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dot_finish, dim3(1), dim3(kThreads), 0, st, partials, grid, out);
    return hipGetLastError();
}

__global__ void k_axpy(int64_t len, int cplx, double ar, double ai, const double* __restrict__ x,
                       double* __restrict__ y)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < len;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (cplx) {
            const double xr = x[2 * i], xi = x[2 * i + 1];
            y[2 * i] = fma(ar, xr, fma(-ai, xi, y[2 * i]));
            y[2 * i + 1] = fma(ar, xi, fma(ai, xr, y[2 * i + 1]));
        } else {
            y[i] = fma(ar, x[i], y[i]);
        }
    }
}

hipError_t launch_axpy(int64_t len, int is_complex, double ar, double ai, const double* x, double* y,
                       hipStream_t st)
{
    const int grid = static_cast<int>(std::min<int64_t>((len + 255) / 256, 8192));
    hipLaunchKernelGGL(k_axpy, dim3(std::max(grid, 1)), dim3(256), 0, st, len, is_complex, ar, ai, x, y);
    return hipGetLastError();
}

__global__ void k_scal(int64_t len, double a, double* __restrict__ x)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < len;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        x[i] *= a;
}

hipError_t launch_scal(int64_t len, double a, double* x, hipStream_t st)
{
    const int grid = static_cast<int>(std::min<int64_t>((len + 255) / 256, 8192));
    hipLaunchKernelGGL(k_scal, dim3(std::max(grid, 1)), dim3(256), 0, st, len, a, x);
    return hipGetLastError();
}

// y = sum_j coef_j * V_j ; coef complex (re, im) pairs when cplx
__global__ void k_combine(int64_t len, int cplx, int m, const double* __restrict__ V, const double* __restrict__ coef,
                          double* __restrict__ y)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < len;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (cplx) {
            double ar = 0.0, ai = 0.0;
            for (int j = 0; j < m; ++j) {
                const double cr = coef[2 * j], ci = coef[2 * j + 1];
                const double vr = V[(2 * len) * j + 2 * i], vi = V[(2 * len) * j + 2 * i + 1];
                ar = fma(cr, vr, fma(-ci, vi, ar));
                ai = fma(cr, vi, fma(ci, vr, ai));
            }
            y[2 * i] = ar;
            y[2 * i + 1] = ai;
        } else {
            double a = 0.0;
            for (int j = 0; j < m; ++j) a = fma(coef[j], V[len * j + i], a);
            y[i] = a;
        }
    }
}

hipError_t launch_combine(int64_t len, int is_complex, int m, const double* V, const double* coef, double* y,
                          hipStream_t st)
{
    const int grid = static_cast<int>(std::min<int64_t>((len + 255) / 256, 8192));
    hipLaunchKernelGGL(k_combine, dim3(std::max(grid, 1)), dim3(256), 0, st, len, is_complex, m, V, coef, y);
    return hipGetLastError();
}

// ---- Gram-Schmidt against a block of basis vectors (complex128) ----
// One kernel, k_gs<M, MODE>, for the sweeps of classical Gram-Schmidt against
// M <= kMdMax basis vectors (a compile-time count, so the M loads of an
// element issue back to back):
//   GS_DOT       partial sums of c_i = <V_i, w>                  (reads V, w)
//   GS_SUB       w -= sum_i c_i V_i                              (reads V, w; writes w)
//   GS_SUB_DOT   w -= sum_i c_i V_i, then partial sums of <V_i, w> on the new w
//                (the second sweep's dots ride on the first sweep's subtraction)
//   GS_SUB_NORM  out = w - sum_i c_i V_i, partial sum of |out|^2 (reads V, w; writes out)
// A workgroup reduces a contiguous chunk; k_gs_finish sums the workgroups'
// partials in a fixed order, so every coefficient is deterministic.
constexpr int kMdGrid = 16384;  // most workgroups of a sweep (partials: kMdGrid x 2 kMdMax doubles)
enum GsMode { GS_DOT = 0, GS_SUB = 1, GS_SUB_DOT = 2, GS_SUB_NORM = 3 };

template <int M, int MODE>
__global__ __launch_bounds__(kThreads) void k_gs(int64_t len, const double* __restrict__ V, double* w,
                                                 double* __restrict__ out, const double* __restrict__ coef,
                                                 int64_t chunk, double* __restrict__ partials)
{
    constexpr bool SUB = MODE != GS_DOT;
    constexpr bool DOT = MODE == GS_DOT || MODE == GS_SUB_DOT;
    constexpr bool NORM = MODE == GS_SUB_NORM;
    __shared__ double sc[kThreads / 64][2 * kMdMax];
    double ar[DOT ? M : 1], ai[DOT ? M : 1];
#pragma unroll
    for (int i = 0; i < (DOT ? M : 1); ++i) ar[i] = ai[i] = 0.0;
    double nr = 0.0;
    const double2* V2 = reinterpret_cast<const double2*>(V);
    double2* w2 = reinterpret_cast<double2*>(w);
    const int64_t beg = blockIdx.x * chunk;
    const int64_t end = min(len, beg + chunk);
    for (int64_t e = beg + threadIdx.x; e < end; e += kThreads) {
        double2 wv = w2[e];
        double2 v[M];
#pragma unroll
        for (int i = 0; i < M; ++i) v[i] = V2[len * i + e];
        if constexpr (SUB) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const double cr = coef[2 * i], ci = coef[2 * i + 1];
                wv.x = fma(-cr, v[i].x, fma(ci, v[i].y, wv.x));
                wv.y = fma(-cr, v[i].y, fma(-ci, v[i].x, wv.y));
            }
            if constexpr (NORM) {
                reinterpret_cast<double2*>(out)[e] = wv;
                nr = fma(wv.x, wv.x, fma(wv.y, wv.y, nr));
            } else {
                w2[e] = wv;
            }
        }
        if constexpr (DOT) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                ar[i] = fma(v[i].x, wv.x, fma(v[i].y, wv.y, ar[i]));
                ai[i] = fma(v[i].x, wv.y, fma(-v[i].y, wv.x, ai[i]));
            }
        }
    }
    if constexpr (DOT || NORM) {
        constexpr int NS = DOT ? 2 * M : 2;  // slots this mode writes
        const int wave = threadIdx.x >> 6;
        const bool lead = (threadIdx.x & 63) == 0;
        if constexpr (DOT) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const double r = wsum(ar[i]), q = wsum(ai[i]);
                if (lead) {
                    sc[wave][2 * i] = r;
                    sc[wave][2 * i + 1] = q;
                }
            }
        } else {
            const double r = wsum(nr);
            if (lead) {
                sc[wave][0] = r;
                sc[wave][1] = 0.0;
            }
        }
        __syncthreads();
        if (threadIdx.x < NS) {
            double a = 0.0;
#pragma unroll
            for (int k = 0; k < kThreads / 64; ++k) a += sc[k][threadIdx.x];
            partials[static_cast<int64_t>(blockIdx.x) * 2 * kMdMax + threadIdx.x] = a;
        }
    }
}

// one workgroup per slot: thread t sums workgroups t, t + kThreads, ...; the
// waves then reduce in a fixed tree, so the result does not depend on timing
__global__ __launch_bounds__(kThreads) void k_gs_finish(const double* __restrict__ partials, int blocks,
                                                       double* __restrict__ coef)
{
    __shared__ double sw[kThreads / 64];
    const int slot = blockIdx.x;
    double a = 0.0;
    for (int b = threadIdx.x; b < blocks; b += kThreads) a += partials[static_cast<int64_t>(b) * 2 * kMdMax + slot];
    a = wsum(a);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) coef[slot] = (sw[0] + sw[1]) + (sw[2] + sw[3]);
}

template <int MODE, int M = 1>
hipError_t launch_gs_m(int m, int grid, int64_t len, const double* V, double* w, double* out, const double* coef,
                       int64_t chunk, double* partials, hipStream_t st)
{
    if constexpr (M > kMdMax) {
        return hipErrorInvalidValue;
    } else {
        if (m != M) return launch_gs_m<MODE, M + 1>(m, grid, len, V, w, out, coef, chunk, partials, st);
        hipLaunchKernelGGL((k_gs<M, MODE>), dim3(grid), dim3(kThreads), 0, st, len, V, w, out, coef, chunk, partials);
        return hipGetLastError();
    }
}

// chunk of elements per workgroup (at least 4096: 64 KB of each vector) and
// the grid that covers len with it; grid-strided blocks and larger chunks
// measured no faster at N=26
static void gs_grid(int64_t len, int64_t* chunk, int* grid)
{
    int64_t c = 4096;
    while ((len + c - 1) / c > kMdGrid) c <<= 1;
    *chunk = c;
    *grid = static_cast<int>(std::max<int64_t>(1, (len + c - 1) / c));
}

size_t multidot_partials_doubles() { return static_cast<size_t>(kMdGrid) * 2 * kMdMax; }

hipError_t launch_gs_block(int64_t len, int m, const double* V, double* w, double* coef, double* partials,
                           hipStream_t st)
{
    if (m < 1 || m > kMdMax) return hipErrorInvalidValue;
    int64_t chunk;
    int grid;
    gs_grid(len, &chunk, &grid);
    hipError_t e = launch_gs_m<GS_DOT>(m, grid, len, V, w, nullptr, nullptr, chunk, partials, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gs_finish, dim3(2 * m), dim3(kThreads), 0, st, partials, grid, coef);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_gs_m<GS_SUB>(m, grid, len, V, w, nullptr, coef, chunk, partials, st);
}

hipError_t launch_cgs2(int64_t len, int m, const double* V, double* w, double* next, double* coef, double* partials,
                       hipStream_t st)
{
    if (m < 1 || m > kMdMax || len < 1) return hipErrorInvalidValue;
    if (next) {
        // the last pass reads w and the basis while it writes next
        const double* vend = V + 2 * len * static_cast<int64_t>(m);
        if (next == w || (next + 2 * len > V && next < vend)) return hipErrorInvalidValue;
    }
    int64_t chunk;
    int grid;
    gs_grid(len, &chunk, &grid);
    double* c1 = coef;
    double* c2 = coef + 2 * kMdMax;
    hipError_t e = launch_gs_m<GS_DOT>(m, grid, len, V, w, nullptr, nullptr, chunk, partials, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gs_finish, dim3(2 * m), dim3(kThreads), 0, st, partials, grid, c1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_gs_m<GS_SUB_DOT>(m, grid, len, V, w, nullptr, c1, chunk, partials, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_gs_finish, dim3(2 * m), dim3(kThreads), 0, st, partials, grid, c2);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (!next) return hipSuccess;
    if ((e = launch_gs_m<GS_SUB_NORM>(m, grid, len, V, w, next, c2, chunk, partials, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_gs_finish, dim3(2), dim3(kThreads), 0, st, partials, grid, coef + 4 * kMdMax);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Subspace Hamiltonian (diagonalizer.cuh:79-114, 188-245).  The basis is held
// sorted; nbr[i * n + j] is the index of basis[i] ^ 2^j in it, or -1.  The
// table is built once per basis (one binary search per state and spin), so
// every H.v of a Lanczos run is a gather instead of n searches per row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_sub_neighbors(const uint64_t* __restrict__ basis, int64_t dim, int n,
                                                            int32_t* __restrict__ nbr)
{
    const int64_t total = dim * n;
    for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < total;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t i = k / n;
        const int j = static_cast<int>(k - i * n);
        const uint64_t target = basis[i] ^ (1ull << j);
        int64_t lo = 0, hi = dim - 1;
        int32_t found = -1;
        while (lo <= hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            const uint64_t b = basis[mid];
            if (b == target) {
                found = static_cast<int32_t>(mid);
                break;
            }
            if (b < target) lo = mid + 1;
            else hi = mid - 1;
        }
        nbr[k] = found;
    }
}

// y_i = diag_i x_i - g * x_b for every neighbour b of i, subtracted in spin order
__global__ __launch_bounds__(kThreads) void k_sub_hv(const double* __restrict__ diag, const int32_t* __restrict__ nbr,
                                                     int64_t dim, int n, double g, const double* __restrict__ x,
                                                     double* __restrict__ y)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < dim;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double yi = diag[i] * x[i];
        const int32_t* row = nbr + i * n;
        for (int j = 0; j < n; ++j) {
            const int32_t b = row[j];
            if (b >= 0) yi -= g * x[b];
        }
        y[i] = yi;
    }
}

// dense H, column-major dim x dim, from the table (H zero on entry)
__global__ __launch_bounds__(kThreads) void k_sub_dense(const double* __restrict__ diag, const int32_t* __restrict__ nbr,
                                                        int64_t dim, int n, double g, double* __restrict__ H)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < dim;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        H[i + i * dim] = diag[i];
        for (int j = 0; j < n; ++j) {
            const int32_t b = nbr[i * n + j];
            if (b >= 0) H[i + static_cast<int64_t>(b) * dim] = -g;
        }
    }
}

static int grid_for(int64_t work)
{
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((work + kThreads - 1) / kThreads, 16384)));
}

hipError_t launch_sub_neighbors(const uint64_t* basis, int64_t dim, int n, int32_t* nbr, hipStream_t st)
{
    hipLaunchKernelGGL(k_sub_neighbors, dim3(grid_for(dim * n)), dim3(kThreads), 0, st, basis, dim, n, nbr);
    return hipGetLastError();
}

hipError_t launch_sub_hv(const double* diag, const int32_t* nbr, int64_t dim, int n, double g, const double* x,
                         double* y, hipStream_t st)
{
    hipLaunchKernelGGL(k_sub_hv, dim3(grid_for(dim)), dim3(kThreads), 0, st, diag, nbr, dim, n, g, x, y);
    return hipGetLastError();
}

hipError_t launch_sub_dense(const double* diag, const int32_t* nbr, int64_t dim, int n, double g, double* H,
                            hipStream_t st)
{
    hipLaunchKernelGGL(k_sub_dense, dim3(grid_for(dim)), dim3(kThreads), 0, st, diag, nbr, dim, n, g, H);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Plain copy (HBM ceiling probe and PMC calibration: known bytes per launch)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b,
                                                    uint64_t n)
{
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        b[i] = a[i];
}

hipError_t launch_copy_probe(const void* src, void* dst, uint64_t bytes, hipStream_t st)
{
    const uint64_t n = bytes / sizeof(float4);
    hipLaunchKernelGGL(k_copy_probe, dim3(16384), dim3(256), 0, st, static_cast<const float4*>(src),
                       static_cast<float4*>(dst), n);
    return hipGetLastError();
}

// One-shot copies, one chunk of BS x U float4 per workgroup (no grid-stride
// loop).  BS = 256, U = 1 with nontemporal accesses: 4 KB workgroups, the
// fastest read+write stream of the part; consecutive 4 KB chunks go to
// consecutive workgroups and so round-robin over the XCDs.  BS = 512, U = 16:
// a workgroup holds 128 KB, the footprint of a 14-qubit complex64 sweep tile.
typedef float cp4 __attribute__((ext_vector_type(4)));
template <int BS, int U, bool NT>
__global__ __launch_bounds__(BS) void k_copy_chunk(const cp4* __restrict__ a, cp4* __restrict__ b)
{
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * (BS * U) + threadIdx.x;
    cp4 r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) r[u] = NT ? __builtin_nontemporal_load(a + base + u * BS) : a[base + u * BS];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if constexpr (NT) __builtin_nontemporal_store(r[u], b + base + u * BS);
        else b[base + u * BS] = r[u];
    }
}

uint64_t copy_pattern_chunk(int pattern)
{
    return pattern == 1 ? 256ull * 16 : pattern == 2 ? 512ull * 16 * 16 : 16;
}

hipError_t launch_copy_pattern(int pattern, const void* src, void* dst, uint64_t bytes, hipStream_t st)
{
    if (pattern == 0) return launch_copy_probe(src, dst, bytes, st);
    const uint64_t chunk = copy_pattern_chunk(pattern);
    if ((pattern != 1 && pattern != 2) || bytes == 0 || bytes % chunk != 0 || bytes / chunk > 0x7fffffffull)
        return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(bytes / chunk));
    const cp4* a = static_cast<const cp4*>(src);
    cp4* b = static_cast<cp4*>(dst);
    if (pattern == 1) hipLaunchKernelGGL((k_copy_chunk<256, 1, true>), grid, dim3(256), 0, st, a, b);
    else hipLaunchKernelGGL((k_copy_chunk<512, 16, false>), grid, dim3(512), 0, st, a, b);
    return hipGetLastError();
}

}  // namespace bk
