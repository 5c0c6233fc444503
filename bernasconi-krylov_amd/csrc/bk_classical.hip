// bk_classical.hip -- trap supply at scale: batched single-flip steepest
// descent on the GPU and the multi-start collector of distinct minima
// (classical.cuh:459-535, 861-950), for the traps the sampler starts from.
//
// One wave per start.  Lane p scores the flip of spin p from the
// autocorrelations C_k held in LDS (O(n) per lane per pass, where a thread per
// start pays O(n^2)); a wave-wide minimum over (energy, p) picks the best
// flip, lowest p first on ties, which is the reference's strict-< scan in
// increasing p.  Lane k then updates C_k.  Integer arithmetic throughout, so
// the result is bit-identical to the host descent.
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

constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ int spin(uint64_t s, int i) { return ((s >> i) & 1ull) ? -1 : 1; }

// change of C_k when spin p flips: -2 s_p (s_{p+k} + s_{p-k}) over the pairs present
__device__ __forceinline__ int delta_c(uint64_t s, int n, int p, int k)
{
    int d = 0;
    const int sp = spin(s, p);
    if (p + k < n) d -= 2 * sp * spin(s, p + k);
    if (p - k >= 0) d -= 2 * spin(s, p - k) * sp;
    return d;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(64 * kWavesPerBlock) void k_descent(int n, int64_t count, const uint64_t* __restrict__ starts,
                                                                 uint64_t max_evals, uint64_t* __restrict__ fin,
                                                                 int64_t* __restrict__ fe, int64_t* __restrict__ evc)
{
    __shared__ int32_t cks[kWavesPerBlock][64];
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int w = static_cast<int>(threadIdx.x >> 6);
    int32_t* C = cks[w];
    const uint64_t mask = n >= 64 ? ~0ull : (1ull << n) - 1ull;
    const uint64_t budget = max_evals == 0 ? ~0ull : max_evals;
    for (int64_t st = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w; st < count;
         st += static_cast<int64_t>(gridDim.x) * kWavesPerBlock) {
        uint64_t s = starts[st] & mask;
        int32_t ck = 0;
        if (lane >= 1 && lane < n)
            ck = (n - lane) - 2 * __popcll((s ^ (s >> lane)) & ((1ull << (n - lane)) - 1ull));
        C[lane] = ck;
        int64_t e = wave_sum_i64(static_cast<int64_t>(ck) * ck);
        uint64_t evals = 1;
        __builtin_amdgcn_wave_barrier();
        for (;;) {
            if (evals + static_cast<uint64_t>(n) > budget) break;
            evals += static_cast<uint64_t>(n);
            uint64_t key = ~0ull;
            if (lane < n) {
                int64_t ne = 0;
                for (int k = 1; k < n; ++k) {
                    const int64_t c = C[k] + delta_c(s, n, lane, k);
                    ne += c * c;
                }
                key = (static_cast<uint64_t>(ne) << 6) | static_cast<uint64_t>(lane);  // E <= 85344 at n = 64
            }
            key = wave_min_u64(key);
            const int64_t best_e = static_cast<int64_t>(key >> 6);
            if (best_e >= e) break;  // no strictly better neighbour
            const int p = static_cast<int>(key & 63u);
            if (lane >= 1 && lane < n) C[lane] += delta_c(s, n, p, lane);
            __builtin_amdgcn_wave_barrier();
            s ^= 1ull << p;
            e = best_e;
        }
        if (lane == 0) {
            fin[st] = s;
            fe[st] = e;
            if (evc) evc[st] = static_cast<int64_t>(evals);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace

hipError_t launch_descent(int n, int64_t count, const uint64_t* starts, uint64_t max_evals, uint64_t* fin, int64_t* fe,
                          int64_t* evc, hipStream_t st)
{
    const int64_t blocks = std::min<int64_t>((count + kWavesPerBlock - 1) / kWavesPerBlock, 8192);
    hipLaunchKernelGGL(k_descent, dim3(static_cast<unsigned>(std::max<int64_t>(blocks, 1))), dim3(64 * kWavesPerBlock),
                       0, st, n, count, starts, max_evals, fin, fe, evc);
    return hipGetLastError();
}

}  // namespace bk

using bkc::fail;
using bkc::ScopedBuf;

namespace {

// device buffers of one batch of descents
struct DescentBufs {
    ScopedBuf starts, fin, fe, evc;
};

int run_descents(bk_ctx* c, int n, const uint64_t* starts, int64_t count, uint64_t max_evals, DescentBufs* b,
                 uint64_t* fin, int64_t* fe, int64_t* evc)
{
    const size_t bytes = sizeof(uint64_t) * static_cast<size_t>(count);
    HIPCALL(b->starts.ensure(bytes));
    HIPCALL(b->fin.ensure(bytes));
    HIPCALL(b->fe.ensure(bytes));
    HIPCALL(b->evc.ensure(bytes));
    HIPCALL(hipMemcpyAsync(b->starts.p, starts, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCALL(bk::launch_descent(n, count, b->starts.as<uint64_t>(), max_evals, b->fin.as<uint64_t>(), b->fe.as<int64_t>(),
                               b->evc.as<int64_t>(), c->stream));
    HIPCALL(hipMemcpyAsync(fin, b->fin.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(hipMemcpyAsync(fe, b->fe.p, bytes, hipMemcpyDeviceToHost, c->stream));
    if (evc) HIPCALL(hipMemcpyAsync(evc, b->evc.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    return BK_OK;
}

}  // namespace

extern "C" {

uint64_t bk_canonical(int n, uint64_t s) { return bk::host_canonical(n, s); }

static int bk_descent_batch_impl(bk_ctx* c, int n, const uint64_t* starts, int64_t count, uint64_t max_evals_per_start,
                                 uint64_t* final_states, int64_t* final_energies, int64_t* eval_counts)
{
    int rc = bkc::check_ctx(c);
    if (rc) return rc;
    if (n < 2 || n > 64 || count < 0 || (count > 0 && (!starts || !final_states || !final_energies)))
        return fail(BK_E_INVALID, "bk_descent_batch: n=%d count=%lld", n, static_cast<long long>(count));
    if (count == 0) return BK_OK;
    DescentBufs b;
    return run_descents(c, n, starts, count, max_evals_per_start, &b, final_states, final_energies, eval_counts);
}

int bk_descent_batch(bk_ctx* c, int n, const uint64_t* starts, int64_t count, uint64_t max_evals_per_start,
                     uint64_t* final_states, int64_t* final_energies, int64_t* eval_counts)
{
    return bkc::guard_abi(
        [&] { return bk_descent_batch_impl(c, n, starts, count, max_evals_per_start, final_states, final_energies,
                                           eval_counts); });
}

static int bk_collect_minima_impl(bk_ctx* c, int n, uint64_t seed, int64_t target_families, int64_t batch_size,
                                  int64_t max_starts, bk_minimum_family* out, int64_t cap, int64_t* count,
                                  int64_t* total_starts, int64_t* total_evals)
{
    int rc = bkc::check_ctx(c);
    if (rc) return rc;
    if (n < 2 || n > 64 || target_families < 1 || batch_size < 1 || max_starts < 1 || !count || (cap > 0 && !out))
        return fail(BK_E_INVALID, "bk_collect_minima: bad argument");
    const uint64_t mask = n >= 64 ? ~0ull : (1ull << n) - 1ull;
    std::mt19937_64 rng(seed);  // the optimizer's generator: random_state() = rng() & mask
    struct Fam {
        uint64_t canon, sample;
        int64_t energy;
        uint64_t visits, evals_first;
    };
    std::unordered_map<uint64_t, size_t> index;
    std::vector<Fam> fams;
    std::vector<uint64_t> starts(static_cast<size_t>(batch_size)), fin(starts.size());
    std::vector<int64_t> fe(starts.size()), evc(starts.size());
    DescentBufs b;
    uint64_t evals = 0;
    int64_t started = 0;
    int stagnant = 0;
    while (static_cast<int64_t>(fams.size()) < target_families && started < max_starts) {
        for (auto& s : starts) s = rng() & mask;
        if ((rc = run_descents(c, n, starts.data(), batch_size, 0, &b, fin.data(), fe.data(), evc.data()))) return rc;
        started += batch_size;
        size_t fresh = 0;
        for (size_t i = 0; i < starts.size(); ++i) {
            evals += static_cast<uint64_t>(evc[i]);
            const uint64_t canon = bk::host_canonical(n, fin[i]);
            auto it = index.find(canon);
            if (it == index.end()) {
                index.emplace(canon, fams.size());
                fams.push_back({canon, fin[i], fe[i], 1, evals});
                ++fresh;
                if (static_cast<int64_t>(fams.size()) >= target_families) break;
            } else {
                ++fams[it->second].visits;
            }
        }
        stagnant = fresh == 0 ? stagnant + 1 : 0;
        if (stagnant >= 4) break;  // every reachable minimum found
    }
    // deepest first, then most visited; canonical bits make the order total
    std::sort(fams.begin(), fams.end(), [](const Fam& a, const Fam& b2) {
        if (a.energy != b2.energy) return a.energy < b2.energy;
        if (a.visits != b2.visits) return a.visits > b2.visits;
        return a.canon < b2.canon;
    });
    *count = static_cast<int64_t>(fams.size());
    if (total_starts) *total_starts = started;
    if (total_evals) *total_evals = static_cast<int64_t>(evals);
    for (int64_t i = 0; i < std::min<int64_t>(cap, static_cast<int64_t>(fams.size())); ++i) {
        const Fam& f = fams[static_cast<size_t>(i)];
        out[i].id = i + 1;
        out[i].canonical_state = f.canon;
        out[i].sample_state = f.sample;
        out[i].energy = f.energy;
        out[i].times_visited = f.visits;
        out[i].evals_to_first_find = f.evals_first;
    }
    return BK_OK;
}

int bk_collect_minima(bk_ctx* c, int n, uint64_t seed, int64_t target_families, int64_t batch_size, int64_t max_starts,
                      bk_minimum_family* out, int64_t cap, int64_t* count, int64_t* total_starts, int64_t* total_evals)
{
    return bkc::guard_abi(
        [&] { return bk_collect_minima_impl(c, n, seed, target_families, batch_size, max_starts, out, cap, count,
                                            total_starts, total_evals); });
}

}  // extern "C"
