// bk_ctx.hpp -- the library context and the helpers the C ABI files share
// (private to libbk; not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <exception>
#include <functional>
#include <new>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bk_capi.h"
#include "bk_kernels.hpp"

namespace bkc {

// set the thread's last-error message and return code
int fail(int code, const char* fmt, ...);

// Run an entry point's body with C++ exceptions (host allocation failures,
// container errors) turned into error codes: no exception crosses the C ABI.
template <typename F>
int guard_abi(F&& body)
{
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(BK_E_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(BK_E_INVALID, "%s", e.what());
    } catch (...) {
        return fail(BK_E_INVALID, "unexpected exception");
    }
}

#define HIPCALL(expr)                                                                           \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            return fail(_e == hipErrorOutOfMemory ? BK_E_NOMEM : BK_E_HIP, "%s failed: %s (%s:%d)", \
                        #expr, hipGetErrorString(_e), __FILE__, __LINE__);                      \
    } while (0)

inline size_t amp_bytes(int precision) { return precision == BK_C64 ? 8 : 16; }

// Device bytes the library holds in this process, and the cap a caller may set
// on them (bk_set_device_memory_limit; 0 = none): every device allocation
// goes through DevBuf, and one beyond the cap fails as hipErrorOutOfMemory.
inline std::atomic<uint64_t> g_dev_held{0};
inline std::atomic<uint64_t> g_dev_limit{0};

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            g_dev_held -= cap;
            p = nullptr;
            cap = 0;
        }
        const uint64_t limit = g_dev_limit.load();
        if (limit && g_dev_held.load() + bytes > limit) return hipErrorOutOfMemory;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) {
            cap = bytes;
            g_dev_held += bytes;
        }
        return e;
    }
    void release()
    {
        if (p) {
            (void)hipFree(p);
            g_dev_held -= cap;
        }
        p = nullptr;
        cap = 0;
    }
    template <typename T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

// A DevBuf freed when it goes out of scope (device memory local to one call,
// released on every return path).
struct ScopedBuf : DevBuf {
    ScopedBuf() = default;
    ScopedBuf(const ScopedBuf&) = delete;
    ScopedBuf& operator=(const ScopedBuf&) = delete;
    ~ScopedBuf() { release(); }
};

// A HIP event destroyed when it goes out of scope.
struct ScopedEvent {
    hipEvent_t e = nullptr;
    ScopedEvent() = default;
    ScopedEvent(const ScopedEvent&) = delete;
    ScopedEvent& operator=(const ScopedEvent&) = delete;
    hipError_t create() { return hipEventCreate(&e); }
    ~ScopedEvent()
    {
        if (e) (void)hipEventDestroy(e);
    }
};

// HIP-event timer: events are drawn from a pool, recorded around each launch,
// and resolved into per-name totals when read.
struct EventTimer : bk::KernelTimer {
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    struct Rec {
        size_t ev0, ev1;
        std::string name;
        double bytes;
    };
    std::vector<Rec> recs;
    size_t open_ev = 0;
    hipError_t err = hipSuccess;
    hipEvent_t take()
    {
        if (used == pool.size()) {
            hipEvent_t e = nullptr;
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) {
                err = r;
                return nullptr;
            }
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(hipStream_t s) override
    {
        open_ev = used;
        hipEvent_t e = take();
        if (e) (void)hipEventRecord(e, s);
    }
    void end(hipStream_t s, const char* name, double bytes) override
    {
        const size_t i1 = used;
        hipEvent_t e = take();
        if (e) (void)hipEventRecord(e, s);
        if (err == hipSuccess) recs.push_back({open_ev, i1, name, bytes});
    }
    void reset()
    {
        used = 0;
        recs.clear();
        err = hipSuccess;
    }
    ~EventTimer() override
    {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace bkc

struct bk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bkc::DevBuf psi;
    bkc::DevBuf psi_alt;  // second state buffer for out-of-place sweeps
    int n = 0;
    int precision = 0;
    bool has_state = false;
    // The stored state is exp(-i theta_pre E(s)/n) times the true state: the
    // opening phase of the next Trotter step is applied by the last sweep of
    // the previous one.  Exported states are corrected first (flush_phase).
    double theta_pre = 0.0;
    bkc::DevBuf partials;           // 3 doubles per workgroup of the sweep
    bkc::DevBuf series;             // 3 doubles per measurement slot
    bkc::DevBuf count_partials;
    bkc::DevBuf count_out;
    bkc::DevBuf samp_sums, samp_prefix, samp_rand, samp_out;
    bkc::DevBuf blas_partials, blas_out;
    std::map<std::pair<int, std::pair<int, double>>, bkc::DevBuf> phase_tabs;  // (n, (prec, dtp))
    bk::SweepTuning tuning{};
    int energy_mode = BK_ENERGY_CACHE16;  // how the LOW sweep gets E(s)
    std::map<std::string, int> tuning_overrides;  // keys set by bk_set_tuning, laid over each state's defaults
    // P_<, <d_H>, norm of the held state as the last bk_evolve measured it at
    // its end (valid while state_gen stays obs_gen), for reference s0 / threshold e0
    double obs[3] = {0.0, 0.0, 0.0};
    bool obs_valid = false;
    uint64_t obs_gen = 0;
    uint64_t obs_s0 = 0;
    uint32_t obs_e0 = 0;
    bkc::DevBuf ecache;                        // packed 12-bit codes E(s) >> 2 (LOW tile layout)
    int ecache_n = 0;
    int ecache_prec = 0;
    int ecache_lb = 0;                         // LOW tile bits of its layout
    bkc::DevBuf hv_ecache;                     // uint16 E(s) in state order, for H.v
    int hv_ecache_n = 0;
    bkc::DevBuf hv_scratch;                    // one vector for the high-qubit passes of H.v
    bkc::DevBuf ecache_hi;                     // E(s) in the layout of a fused run's phase-carrying high pass
    std::vector<int> ecache_hi_key;            // (n, precision, h1, a1, h2, a2, R) it was filled for
    bool timing = false;
    bkc::EventTimer timer;
    // bumped by every call that changes the held state; an adjoint tape is
    // valid only for the generation it was recorded at
    uint64_t state_gen = 0;
    // bumped when the held state is set anew (init / upload): the rounding
    // carries of the step-phase tables start over with each history
    uint64_t history_id = 0;
    struct FbTab {
        bkc::DevBuf master, carry, applied;
        int count = 0;
        uint64_t hist = ~0ull;
    };
    std::map<std::pair<int, std::pair<int, double>>, FbTab> fb_tabs;  // (n, (prec, dtp))
    // adjoint tape: the forward run's parameters and its checkpoint states
    struct Tape {
        bool valid = false;
        uint64_t gen = 0;
        int n = 0, precision = 0, steps = 0, every = 0, nckpt = 0;
        double dt = 0.0, theta = 0.0;
        uint64_t s0 = 0;
        uint32_t e0 = 0;
        std::vector<double> g;  // field of each step
        bkc::DevBuf ckpt;            // state after every-th step, 2*every-th, ... (< steps)
    } tape;
    bkc::DevBuf adj_phi, adj_chi, adj_part, adj_grad, adj_drift;
    // batched runs: arena buffers by role, E(s) caches by (n, precision), tuned
    // parameters by workload shape
    std::map<std::string, bkc::DevBuf> batch_bufs;
    std::map<std::pair<int, int>, bkc::DevBuf> batch_ecache;
    std::map<std::vector<int>, bkc::DevBuf> batch_ecache_hi;  // (n, precision, h1, a1, h2, a2, R)
    struct BatchTune {
        int grid = 4096, row_bits = 5;
        int64_t block_bytes = 0;
        bool out_of_place = false;
    };
    std::map<std::string, BatchTune> batch_tuned;
};


namespace bkc {

int check_ctx(bk_ctx* c);
// bound the phase-table cache; call before an entry point fetches tables
int trim_phase_tabs(bk_ctx* c);
// one bk_set_tuning key applied to t (validated; "grid" also sizes the partials)
int apply_tuning_key(bk_ctx* c, bk::SweepTuning& t, const std::string& k, int value);
bool is_fin(double x);
// exp(-i dtp E/n) for E = 4m + c_n, m = 0..count-1, in FP64 (re, im pairs)
std::vector<double> phase_master(int n, double dtp);
// the uint16 E(s) cache of the held state's n in the layout of a LOW tile of
// low_bits (0: the precision's default tile)
int ensure_ecache(bk_ctx* c, int low_bits = 0);
// plain Lanczos (diagonalizer.cuh:420-554) over a device matvec on vectors of
// length dim; returns the number of pairs written to evals (and evecs, host)
int lanczos_core(bk_ctx* c, int64_t dim, const std::function<hipError_t(const double*, double*)>& matvec,
                 int max_krylov_dim, double tol, int k_pairs, double* evals, double* evecs);


}  // namespace bkc
