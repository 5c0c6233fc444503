// bk_minima.hip -- the exhaustive local-minima explorer of the reference's
// src/tools/exhaustive_minima.cpp on the GPU, and exact d_min for many states.
//
// Scan (exhaustive_minima.cpp:109-160).  Every s in [0, 2^(n-1)) is tested for
// being a one-flip minimum: no single flip lowers E(s).  Global negation maps
// minima onto minima, so the half space holds half of them; it also holds the
// canonical representative of every symmetry orbit (SpinState::canonical is
// the smallest image, and one of s, ~s has the top bit clear).  Hence the set
// of orbits at an energy, which the reference builds by inserting canonical(s)
// for every minimum into a std::set, is exactly the set of half-space minima
// with canonical(s) == s: the scan counts them with no set at all.  A first
// pass histograms minima and orbit representatives by energy (atomics on a
// ~10^4-entry table; minima are rare); E* and the trap levels follow from the
// histogram; a second pass lists the representatives at the level energies.
//
// d_min (labs_energy.hpp:175-207): the smallest Hamming radius d at which some
// state has E < target.  Every state of a sphere is tested (existence does not
// depend on the order the reference walks it in).  A launch covers one radius
// for all unresolved states: a thread unranks the first d-subset of its chunk
// in the combinatorial number system (the numeric order the reference's
// next-combination rule walks) and steps through the chunk with that rule.
//
// Integer arithmetic throughout; results are exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <random>
#include <unordered_map>
#include <vector>

#include "../../include/bk_capi.h"
#include "bk_ctx.hpp"
#include "bk_host.hpp"
#include "bk_kernels.hpp"

namespace bk {
namespace {

constexpr int kScanThreads = 256;
constexpr int kSphereChunk = 32;  // d-subsets per thread in a sphere launch

// E(s) for n <= 32 (labs_energy.hpp:121-132 in 32-bit words)
__device__ __forceinline__ int energy32(uint32_t s, int n)
{
    int e = 0;
#pragma unroll 4
    for (int k = 1; k < n; ++k) {
        const uint32_t pairs = (s ^ (s >> k)) & (0xFFFFFFFFu >> (32 - (n - k)));
        const int c = (n - k) - 2 * static_cast<int>(__popc(pairs));
        e += c * c;
    }
    return e;
}

// SpinState::canonical (labs_energy.hpp:51-78) for n <= 32
__device__ __forceinline__ uint32_t canonical32(uint32_t s, int n)
{
    const uint32_t m = 0xFFFFFFFFu >> (32 - n);
    const uint32_t alt = 0x55555555u & m;
    const uint32_t a = s & m, b = a ^ alt;
    const uint32_t ra = __brev(a) >> (32 - n), rb = __brev(b) >> (32 - n);
    uint32_t c = min(min(a, b), min(ra, rb));
    c = min(c, min(min(~a & m, ~b & m), min(~ra & m, ~rb & m)));
    return c;
}

// one-flip minimum: E(s ^ 2^p) >= E(s) for every p (exhaustive_minima.cpp:74-107)
__device__ __forceinline__ bool one_flip_minimum(uint32_t s, int n, int e)
{
    for (int p = 0; p < n; ++p)
        if (energy32(s ^ (1u << p), n) < e) return false;
    return true;
}

// pass 0: hist_min[E] += minima, hist_rep[E] += orbit representatives.
// A state's test stops at its first lower neighbour, after a few energies for
// most states, but a wave would run as long as its slowest lane.  So every lane
// computes one energy per iteration -- E(s) of a state it has just taken, or
// E(s ^ 2^p) of the next flip of the state it holds -- and a lane whose state
// is settled takes the next one of its wave's chunk.  The lanes stay busy and
// the energies computed per state are the ones the test needs.
constexpr uint64_t kScanChunk = 1u << 14;  // states per wave and chunk
__global__ __launch_bounds__(kScanThreads) void k_minima_hist(int n, uint64_t half, uint32_t* __restrict__ hist_min,
                                                              uint32_t* __restrict__ hist_rep)
{
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const uint64_t below = (uint64_t(1) << lane) - 1u;
    const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * (kScanThreads / 64);
    const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * (kScanThreads / 64) + (threadIdx.x >> 6);
    for (uint64_t c0 = wave * kScanChunk; c0 < half; c0 += nwaves * kScanChunk) {
        const uint64_t c1 = min(half, c0 + kScanChunk);
        uint64_t next = c0;  // the chunk's next untaken state (the same in every lane)
        uint32_t s = 0;
        int e = 0;
        int p = -1;  // -1: needs a state; 0..n-1: next flip to test; -2: chunk exhausted
        for (;;) {
            const bool need = p == -1;
            const uint64_t want = __ballot(need);
            const uint64_t idx = next + static_cast<uint64_t>(__popcll(want & below));
            next += static_cast<uint64_t>(__popcll(want));
            if (need) {
                if (idx < c1) s = static_cast<uint32_t>(idx);
                else p = -2;
            }
            if (__ballot(p != -2) == 0) break;
            if (p == -2) continue;
            const int E = energy32(s ^ (need ? 0u : 1u << (p & 31)), n);
            if (need) {
                e = E;
                p = 0;
            } else if (E < e) {
                p = -1;  // a lower neighbour: not a minimum
            } else if (++p == n) {
                atomicAdd(&hist_min[e], 1u);
                if (canonical32(s, n) == s) atomicAdd(&hist_rep[e], 1u);
                p = -1;
            }
        }
    }
}

// pass 1: list the representatives whose energy lies in [e_lo, e_hi] and has
// want[E] set, as (E << 32 | s); at most cap are stored, *count counts all
__global__ __launch_bounds__(kScanThreads) void k_minima_list(int n, uint64_t half, int e_lo, int e_hi,
                                                              const uint8_t* __restrict__ want,
                                                              unsigned long long* __restrict__ out,
                                                              unsigned int* __restrict__ count, unsigned int cap)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kScanThreads;
    for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * kScanThreads + threadIdx.x; s < half; s += stride) {
        const uint32_t b = static_cast<uint32_t>(s);
        const int e = energy32(b, n);
        if (e < e_lo || e > e_hi || !want[e]) continue;
        if (!one_flip_minimum(b, n, e) || canonical32(b, n) != b) continue;
        const unsigned int i = atomicAdd(count, 1u);
        if (i < cap) out[i] = (static_cast<unsigned long long>(e) << 32) | b;
    }
}

__device__ __forceinline__ uint64_t binom(int a, int b)
{
    if (b < 0 || b > a) return 0;
    if (b > a - b) b = a - b;
    uint64_t r = 1;
    for (int i = 1; i <= b; ++i) r = r * static_cast<uint64_t>(a - b + i) / static_cast<uint64_t>(i);
    return r;
}

// One radius d for 'count' states: found[i] = 1 when some state at Hamming
// distance d from base[i] has E < target[i].  Work item (i, chunk q) covers
// the d-subsets of rank [q*kSphereChunk, (q+1)*kSphereChunk) of n bits.
__global__ __launch_bounds__(kScanThreads) void k_sphere(int n, int d, int count, const uint32_t* __restrict__ base,
                                                         const int* __restrict__ target, uint64_t chunks,
                                                         uint64_t total, uint8_t* __restrict__ found)
{
    __shared__ uint64_t bin[33][33];
    for (int t = threadIdx.x; t < 33 * 33; t += kScanThreads) bin[t / 33][t % 33] = binom(t / 33, t % 33);
    __syncthreads();
    const uint64_t items = chunks * static_cast<uint64_t>(count);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kScanThreads;
    for (uint64_t it = static_cast<uint64_t>(blockIdx.x) * kScanThreads + threadIdx.x; it < items; it += stride) {
        const int i = static_cast<int>(it / chunks);
        const uint64_t q = it - static_cast<uint64_t>(i) * chunks;
        if (found[i]) continue;
        // unrank the first subset of the chunk: rank = sum_j C(c_j, j), c_d > ... > c_1
        uint64_t r = q * kSphereChunk;
        uint32_t mask = 0;
        int c = n - 1;
        for (int j = d; j >= 1; --j) {
            while (bin[c][j] > r) --c;
            mask |= 1u << c;
            r -= bin[c][j];
            --c;
        }
        const uint64_t last = min(total, (q + 1) * kSphereChunk);
        const uint32_t b = base[i];
        const int tgt = target[i];
        bool hit = false;
        for (uint64_t k = q * kSphereChunk; k < last; ++k) {
            if (energy32(b ^ mask, n) < tgt) {
                hit = true;
                break;
            }
            // next d-subset in numeric order (the reference's rule, labs_energy.hpp:195-200)
            const uint32_t low = mask & (~mask + 1u);
            const uint32_t ripple = mask + low;
            mask = ripple | (((mask ^ ripple) >> 2) / low);
        }
        if (hit) found[i] = 1;
    }
}

int grid_for(uint64_t items)
{
    const uint64_t blocks = (items + kScanThreads - 1) / kScanThreads;
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(blocks, 16384)));
}

}  // namespace
}  // namespace bk

using bkc::fail;
using bkc::ScopedBuf;

namespace {

uint64_t host_binom(int a, int b)
{
    if (b < 0 || b > a) return 0;
    if (b > a - b) b = a - b;
    uint64_t r = 1;
    for (int i = 1; i <= b; ++i) r = r * static_cast<uint64_t>(a - b + i) / static_cast<uint64_t>(i);
    return r;
}

// compute_d_min (labs_energy.hpp:175-207) for many states on the device.
// targets[i] < 0 means E(states[i]).
int d_min_device(bk_ctx* c, int n, const uint64_t* states, const int64_t* targets, int64_t count, int max_d, int* out)
{
    const uint32_t m = n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
    const int not_found = max_d > 0 ? max_d + 1 : n;
    const int limit = (max_d > 0 && max_d <= n) ? max_d : n;
    const int64_t opt = bk::host_known_optimal(n);
    std::vector<uint32_t> base;
    std::vector<int> tgt;
    std::vector<int64_t> idx;  // unresolved entries
    for (int64_t i = 0; i < count; ++i) {
        const uint32_t b = static_cast<uint32_t>(states[i]) & m;
        const int64_t t = targets && targets[i] >= 0 ? targets[i] : bk::host_energy(n, b);
        out[i] = not_found;
        if (opt >= 0 && t <= opt) continue;  // nothing lies below the optimum
        if (t <= 0) continue;
        base.push_back(b);
        tgt.push_back(static_cast<int>(std::min<int64_t>(t, 1 << 30)));
        idx.push_back(i);
    }
    ScopedBuf dbase, dtgt, dfound;
    std::vector<uint8_t> found;
    for (int d = 1; d <= limit && !idx.empty(); ++d) {
        if (d == n) {  // the complement alone
            for (size_t k = 0; k < idx.size(); ++k)
                if (bk::host_energy(n, base[k] ^ m) < tgt[k]) out[idx[k]] = n;
            break;
        }
        const int cnt = static_cast<int>(idx.size());
        HIPCALL(dbase.ensure(sizeof(uint32_t) * cnt));
        HIPCALL(dtgt.ensure(sizeof(int) * cnt));
        HIPCALL(dfound.ensure(cnt));
        HIPCALL(hipMemcpyAsync(dbase.p, base.data(), sizeof(uint32_t) * cnt, hipMemcpyHostToDevice, c->stream));
        HIPCALL(hipMemcpyAsync(dtgt.p, tgt.data(), sizeof(int) * cnt, hipMemcpyHostToDevice, c->stream));
        HIPCALL(hipMemsetAsync(dfound.p, 0, cnt, c->stream));
        const uint64_t total = host_binom(n, d);
        const uint64_t chunks = (total + bk::kSphereChunk - 1) / bk::kSphereChunk;
        hipLaunchKernelGGL(bk::k_sphere, dim3(bk::grid_for(chunks * cnt)), dim3(bk::kScanThreads), 0, c->stream, n,
                           d, cnt, dbase.as<uint32_t>(), dtgt.as<int>(), chunks, total, dfound.as<uint8_t>());
        HIPCALL(hipGetLastError());
        found.assign(cnt, 0);
        HIPCALL(hipMemcpyAsync(found.data(), dfound.p, cnt, hipMemcpyDeviceToHost, c->stream));
        HIPCALL(hipStreamSynchronize(c->stream));
        std::vector<uint32_t> nb;
        std::vector<int> nt;
        std::vector<int64_t> ni;
        for (int k = 0; k < cnt; ++k) {
            if (found[k]) {
                out[idx[k]] = d;
            } else {
                nb.push_back(base[k]);
                nt.push_back(tgt[k]);
                ni.push_back(idx[k]);
            }
        }
        base.swap(nb);
        tgt.swap(nt);
        idx.swap(ni);
    }
    return BK_OK;
}

}  // namespace

extern "C" {

static int bk_d_min_batch_impl(bk_ctx* c, int n, const uint64_t* states, const int64_t* targets, int64_t count,
                               int max_d, int* out)
{
    int rc = bkc::check_ctx(c);
    if (rc) return rc;
    if (n < 2 || n > 32 || count < 0 || (count > 0 && (!states || !out)) || count > (int64_t(1) << 24))
        return fail(BK_E_INVALID, "bk_d_min_batch: n=%d count=%lld", n, static_cast<long long>(count));
    if (count == 0) return BK_OK;
    return d_min_device(c, n, states, targets, count, max_d, out);
}

int bk_d_min_batch(bk_ctx* c, int n, const uint64_t* states, const int64_t* targets, int64_t count, int max_d,
                   int* out)
{
    return bkc::guard_abi([&] { return bk_d_min_batch_impl(c, n, states, targets, count, max_d, out); });
}

static int bk_exhaustive_minima_impl(bk_ctx* c, int n, int levels, uint64_t num_sd_starts, uint64_t sd_seed,
                                     bk_exhaustive_result* res, bk_minima_orbit* orbits, int64_t cap)
{
    int rc = bkc::check_ctx(c);
    if (rc) return rc;
    if (n < 3 || n > 32 || levels < 0 || levels > BK_MINIMA_MAX_LEVELS || !res || cap < 0 || (cap > 0 && !orbits))
        return fail(BK_E_INVALID, "bk_exhaustive_minima: n=%d levels=%d", n, levels);
    *res = bk_exhaustive_result{};
    res->n = n;
    res->e_star = -1;
    const uint64_t half = uint64_t(1) << (n - 1);
    const int emax = (n - 1) * n * (2 * n - 1) / 6;  // sum_{j<n} j^2
    ScopedBuf hist;
    HIPCALL(hist.ensure(sizeof(uint32_t) * 2 * (emax + 1)));
    HIPCALL(hipMemsetAsync(hist.p, 0, sizeof(uint32_t) * 2 * (emax + 1), c->stream));
    bkc::ScopedEvent e0, e1;
    HIPCALL(e0.create());
    HIPCALL(e1.create());
    HIPCALL(hipEventRecord(e0.e, c->stream));
    uint32_t* hmin = hist.as<uint32_t>();
    uint32_t* hrep = hmin + (emax + 1);
    const uint64_t chunks = (half + bk::kScanChunk - 1) / bk::kScanChunk;
    const int hist_grid = static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>((chunks + 3) / 4, 8192)));
    hipLaunchKernelGGL(bk::k_minima_hist, dim3(hist_grid), dim3(bk::kScanThreads), 0, c->stream, n, half, hmin,
                       hrep);
    HIPCALL(hipGetLastError());
    std::vector<uint32_t> h(2 * (emax + 1));
    HIPCALL(hipMemcpyAsync(h.data(), hist.p, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    int64_t minima = 0, orbs = 0, nen = 0;
    std::vector<int> level_e;
    for (int e = 0; e <= emax; ++e) {
        if (!h[e]) continue;
        minima += h[e];
        orbs += h[emax + 1 + e];
        if (nen == 0) res->e_star = e;
        else if (static_cast<int>(level_e.size()) < levels) level_e.push_back(e);
        ++nen;
    }
    res->total_minima_found = 2 * minima;
    res->total_orbits_found = orbs;
    res->num_energies = nen;
    res->num_levels = static_cast<int>(level_e.size());
    int64_t want_orbits = 0;
    for (size_t l = 0; l < level_e.size(); ++l) {
        res->level_energy[l] = level_e[l];
        res->level_orbits[l] = h[emax + 1 + level_e[l]];
        want_orbits += res->level_orbits[l];
    }
    // representatives at the level energies
    std::vector<unsigned long long> reps;
    if (!level_e.empty()) {
        std::vector<uint8_t> want(emax + 1, 0);
        for (int e : level_e) want[e] = 1;
        ScopedBuf dwant, dout, dcount;
        HIPCALL(dwant.ensure(want.size()));
        HIPCALL(dout.ensure(sizeof(unsigned long long) * static_cast<size_t>(want_orbits)));
        HIPCALL(dcount.ensure(sizeof(unsigned int)));
        HIPCALL(hipMemcpyAsync(dwant.p, want.data(), want.size(), hipMemcpyHostToDevice, c->stream));
        HIPCALL(hipMemsetAsync(dcount.p, 0, sizeof(unsigned int), c->stream));
        hipLaunchKernelGGL(bk::k_minima_list, dim3(bk::grid_for(half)), dim3(bk::kScanThreads), 0, c->stream, n,
                           half, level_e.front(), level_e.back(), dwant.as<uint8_t>(),
                           dout.as<unsigned long long>(), dcount.as<unsigned int>(),
                           static_cast<unsigned int>(want_orbits));
        HIPCALL(hipGetLastError());
        unsigned int got = 0;
        reps.resize(static_cast<size_t>(want_orbits));
        HIPCALL(hipMemcpyAsync(&got, dcount.p, sizeof(got), hipMemcpyDeviceToHost, c->stream));
        HIPCALL(hipMemcpyAsync(reps.data(), dout.p, sizeof(unsigned long long) * reps.size(), hipMemcpyDeviceToHost,
                               c->stream));
        HIPCALL(hipEventRecord(e1.e, c->stream));
        HIPCALL(hipStreamSynchronize(c->stream));
        if (static_cast<int64_t>(got) != want_orbits)
            return fail(BK_E_HIP, "bk_exhaustive_minima: listed %u representatives, counted %lld", got,
                        static_cast<long long>(want_orbits));
        std::sort(reps.begin(), reps.end());  // by energy, then canonical bits (the std::set order)
    } else {
        HIPCALL(hipEventRecord(e1.e, c->stream));
        HIPCALL(hipStreamSynchronize(c->stream));
    }
    float ms = 0.f;
    HIPCALL(hipEventElapsedTime(&ms, e0.e, e1.e));
    res->search_time_sec = ms * 1e-3;
    // exact d_min of every listed orbit at its level energy (exhaustive_minima.cpp:200-203)
    const int64_t nrep = static_cast<int64_t>(reps.size());
    std::vector<uint64_t> st(reps.size());
    std::vector<int64_t> tg(reps.size());
    std::vector<int> dm(reps.size());
    for (int64_t i = 0; i < nrep; ++i) {
        st[i] = reps[i] & 0xFFFFFFFFull;
        tg[i] = static_cast<int64_t>(reps[i] >> 32);
    }
    if (nrep > 0 && (rc = d_min_device(c, n, st.data(), tg.data(), nrep, -1, dm.data()))) return rc;
    // share of steepest descents from mt19937_64(sd_seed) draws that end in each
    // orbit (run_sd_benchmarks, exhaustive_minima.cpp:218-271)
    std::vector<uint64_t> hits(reps.size(), 0);
    if (num_sd_starts > 0 && nrep > 0) {
        const uint64_t mask = (uint64_t(1) << n) - 1;
        std::mt19937_64 rng(sd_seed);
        const int64_t batch = 1 << 20;
        std::vector<uint64_t> starts, fin;
        std::vector<int64_t> fe;
        std::unordered_map<uint64_t, int64_t> where;  // canonical bits -> orbit (the reference keys hits by them)
        for (int64_t j = 0; j < nrep; ++j) where.emplace(st[j], j);
        uint64_t done = 0;
        while (done < num_sd_starts) {
            const int64_t k = static_cast<int64_t>(std::min<uint64_t>(batch, num_sd_starts - done));
            starts.resize(k);
            fin.resize(k);
            fe.resize(k);
            for (auto& s : starts) s = rng() & mask;
            if ((rc = bk_descent_batch(c, n, starts.data(), k, 0, fin.data(), fe.data(), nullptr))) return rc;
            for (int64_t i = 0; i < k; ++i) {
                auto it = where.find(bk::host_canonical(n, fin[i]));
                if (it != where.end()) ++hits[it->second];
            }
            done += static_cast<uint64_t>(k);
        }
    }
    const int64_t w = std::min<int64_t>(cap, nrep);
    for (int64_t i = 0; i < w; ++i) {
        orbits[i].canonical_state = st[i];
        orbits[i].energy = tg[i];
        int lv = 0;
        while (lv < res->num_levels && res->level_energy[lv] != tg[i]) ++lv;
        orbits[i].level = lv + 1;
        orbits[i].d_min = dm[i];
        orbits[i].sd_hit_fraction =
            num_sd_starts > 0 ? static_cast<double>(hits[i]) / static_cast<double>(num_sd_starts) : 0.0;
    }
    res->orbits_written = w;
    return BK_OK;
}

int bk_exhaustive_minima(bk_ctx* c, int n, int levels, uint64_t num_sd_starts, uint64_t sd_seed,
                         bk_exhaustive_result* res, bk_minima_orbit* orbits, int64_t cap)
{
    return bkc::guard_abi(
        [&] { return bk_exhaustive_minima_impl(c, n, levels, num_sd_starts, sd_seed, res, orbits, cap); });
}

}  // extern "C"
