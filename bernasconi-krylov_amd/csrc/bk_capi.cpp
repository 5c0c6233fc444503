// bk_capi.cpp -- C ABI of libbk.so (declared in include/bk_capi.h).
//
// Host orchestration in C++17 over the gfx950 kernels: a context owns one
// device, one HIP stream (default priority), the statevector and the scratch
// buffers.  The Trotter loop enqueues kernels only; observables land in a
// device-side series buffer that is copied back once per call, so there is
// no host synchronisation per step (the reference syncs and copies two
// doubles per measurement, sampler.cuh:369-380).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "../../include/bk_capi.h"
#include "bk_ctx.hpp"
#include "bk_host.hpp"
#include "bk_kernels.hpp"
#include "bk_sweeps.hpp"

namespace bkc {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace bkc

using namespace bkc;

namespace bkc {

int check_ctx(bk_ctx* c)
{
    if (!c) return fail(BK_E_INVALID, "null context");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return fail(BK_E_HIP, "hipSetDevice(%d): %s", c->device, hipGetErrorString(e));
    // every entry point starts here, before it fetches a phase table
    return trim_phase_tabs(c);
}

// Bound the phase-table cache.  Called at the start of an entry point, before
// it fetches any table, so no table handed out during a call is freed by a
// later fetch in the same call.  Waits for the stream: queued work may still
// read the tables.
int trim_phase_tabs(bk_ctx* c)
{
    if (c->phase_tabs.size() <= 64) return BK_OK;
    HIPCALL(hipStreamSynchronize(c->stream));
    for (auto& kv : c->phase_tabs) kv.second.release();
    c->phase_tabs.clear();
    return BK_OK;
}

// Set one tuning key on t (validated); the context keeps the keys a caller set
// and lays them over the per-state defaults (bk_set_tuning).
int apply_tuning_key(bk_ctx* c, bk::SweepTuning& t, const std::string& k, int value)
{
    if (k == "grid") {
        if (value < 1 || value > (1 << 20)) return fail(BK_E_INVALID, "grid %d", value);
        t.grid = value;
        HIPCALL(c->partials.ensure(bk::partials_bytes(bk::max_grid(t))));
    } else if (k == "grid_bphase" || k == "grid_lphase" || k == "rev_grid_low" || k == "rev_grid_bphase") {
        if (value < 0 || value > (1 << 20)) return fail(BK_E_INVALID, "%s %d", k.c_str(), value);
        (k == "grid_bphase" ? t.grid_bphase : k == "grid_lphase" ? t.grid_lphase
                            : k == "rev_grid_low" ? t.rev_grid_low : t.rev_grid_bphase) = value;
        HIPCALL(c->partials.ensure(bk::partials_bytes(bk::max_grid(t))));
    } else if (k == "row_bits") {
        if (value < 3 || value > 8) return fail(BK_E_INVALID, "row_bits %d", value);
        t.row_bits = value;
    } else if (k == "page_bits") {
        if (value < 0 || value > 32) return fail(BK_E_INVALID, "page_bits %d", value);
        t.page_bits = value;
    } else if (k == "prefetch_low") {
        t.prefetch_low = value != 0;
    } else if (k == "prefetch_high") {
        t.prefetch_high = value != 0;
    } else if (k == "pingpong") {
        t.pingpong = value != 0;
    } else if (k == "fused") {
        t.fused = value != 0;
    } else if (k == "wide_low") {
        t.wide_low = value != 0;
    } else if (k == "lds_tab13") {
        t.lds_tab13 = value != 0;
    } else if (k == "nt_plain") {
        t.nt_plain = value != 0;
    } else if (k == "fold_finish") {
        t.fold_finish = value != 0;
    } else if (k == "narrow_phase") {
        t.narrow_phase = value != 0;
    } else {
        return fail(BK_E_INVALID, "unknown tuning key '%s'", k.c_str());
    }
    return BK_OK;
}

// exp(-i dtp E / n) for E = 0..Emax in the state's precision.  The argument is
// -(dtp/n)*E as in sampler.cuh:83, evaluated and reduced mod 2*pi in FP64 by
// the C library's cos/sin, then rounded once to the table precision.
int get_phase_tab(bk_ctx* c, int n, int precision, double dtp, const void** out)
{
    auto key = std::make_pair(n, std::make_pair(precision, dtp));
    auto it = c->phase_tabs.find(key);
    if (it != c->phase_tabs.end()) {
        *out = it->second.p;
        return BK_OK;
    }
    // C_k = (n-k) mod 2, so every E(s) = 4m + c with c = #{k : n-k odd} mod 4:
    // entry m holds the phase of E = 4m + c and kernels index it by E >> 2.
    const int64_t emax = static_cast<int64_t>(n - 1) * n * (2 * n - 1) / 6;  // sum_{j<n} j^2
    int64_t codd = 0;
    for (int k = 1; k < n; ++k) codd += (n - k) & 1;
    const int64_t c4 = codd % 4;
    const size_t cnt = static_cast<size_t>(emax >> 2) + 1;
    const double w = -(dtp / static_cast<double>(n));
    DevBuf buf;
    if (precision == BK_C64) {
        std::vector<float> h(2 * cnt);
        for (size_t m = 0; m < cnt; ++m) {
            const double th = w * static_cast<double>(4 * static_cast<int64_t>(m) + c4);
            h[2 * m] = static_cast<float>(std::cos(th));
            h[2 * m + 1] = static_cast<float>(std::sin(th));
        }
        HIPCALL(buf.ensure(h.size() * sizeof(float)));
        HIPCALL(hipMemcpy(buf.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    } else {
        std::vector<double> h(2 * cnt);
        for (size_t m = 0; m < cnt; ++m) {
            const double th = w * static_cast<double>(4 * static_cast<int64_t>(m) + c4);
            h[2 * m] = std::cos(th);
            h[2 * m + 1] = std::sin(th);
        }
        HIPCALL(buf.ensure(h.size() * sizeof(double)));
        HIPCALL(hipMemcpy(buf.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    *out = buf.p;
    c->phase_tabs.emplace(key, buf);
    return BK_OK;
}

std::vector<double> phase_master(int n, double dtp)
{
    const int64_t emax = static_cast<int64_t>(n - 1) * n * (2 * n - 1) / 6;  // sum_{j<n} j^2
    int64_t codd = 0;
    for (int k = 1; k < n; ++k) codd += (n - k) & 1;
    const int64_t c4 = codd % 4;
    const size_t cnt = static_cast<size_t>(emax >> 2) + 1;
    const double w = -(dtp / static_cast<double>(n));
    std::vector<double> h(2 * cnt);
    for (size_t m = 0; m < cnt; ++m) {
        const double th = w * static_cast<double>(4 * static_cast<int64_t>(m) + c4);
        h[2 * m] = std::cos(th);
        h[2 * m + 1] = std::sin(th);
    }
    return h;
}

// Step-phase table with error feedback: exp(-i dtp E/n) in FP64 (master), the
// per-entry rounding carry, and the rounded table the sweeps read.  The carry
// restarts (all ones) the first time the table is used in a new history.
int get_fb_tab(bk_ctx* c, int n, int precision, double dtp, bk_ctx::FbTab** out)
{
    auto key = std::make_pair(n, std::make_pair(precision, dtp));
    auto it = c->fb_tabs.find(key);
    if (it == c->fb_tabs.end()) {
        if (c->fb_tabs.size() >= 16) {  // bound the cache (a cleared table restarts its carry)
            // queued passes may still read these tables; each call fetches one
            // table, so none handed out earlier in this call is freed here
            HIPCALL(hipStreamSynchronize(c->stream));
            for (auto& kv : c->fb_tabs) {
                kv.second.master.release();
                kv.second.carry.release();
                kv.second.applied.release();
            }
            c->fb_tabs.clear();
        }
        const std::vector<double> h = phase_master(n, dtp);
        const int cnt = static_cast<int>(h.size() / 2);
        bk_ctx::FbTab t;
        HIPCALL(t.master.ensure(h.size() * sizeof(double)));
        HIPCALL(t.carry.ensure(h.size() * sizeof(double)));
        // two buffers: a fused run's phase pass of step k writes step k+1's table into the other
        HIPCALL(t.applied.ensure(2 * static_cast<size_t>(cnt) * amp_bytes(precision)));
        HIPCALL(hipMemcpy(t.master.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        t.count = cnt;
        it = c->fb_tabs.emplace(key, t).first;
    }
    bk_ctx::FbTab& t = it->second;
    if (t.hist != c->history_id) {
        HIPCALL(bk::launch_carry_reset(t.carry.as<double>(), t.count, c->stream));
        t.hist = c->history_id;
    }
    *out = &t;
    return BK_OK;
}

int ensure_state(bk_ctx* c, int n, int precision)
{
    if (n < 1 || n > BK_MAX_N) return fail(BK_E_INVALID, "n=%d outside [1, %d]", n, BK_MAX_N);
    if (precision != BK_C64 && precision != BK_C128) return fail(BK_E_INVALID, "precision=%d", precision);
    const size_t bytes = amp_bytes(precision) << n;
    if (c->psi.cap < bytes) {
        c->psi.release();
        HIPCALL(c->psi.ensure(bytes));
    }
    c->n = n;
    c->precision = precision;
    c->has_state = true;
    c->theta_pre = 0.0;
    ++c->state_gen;
    ++c->history_id;
    // the defaults for this n and precision, with the keys a caller set on top
    c->tuning = bk::default_tuning(n, precision);
    for (const auto& kv : c->tuning_overrides) {
        const int rc = apply_tuning_key(c, c->tuning, kv.first, kv.second);
        if (rc) return rc;
    }
    c->tuning.timer = c->timing ? &c->timer : nullptr;
    HIPCALL(c->partials.ensure(bk::partials_bytes(bk::max_grid(c->tuning))));
    return BK_OK;
}

// remove the pre-applied opening phase so the held state is the true one
int flush_phase(bk_ctx* c)
{
    if (!c->has_state || c->theta_pre == 0.0) return BK_OK;
    const void* tab = nullptr;
    int rc = get_phase_tab(c, c->n, c->precision, -c->theta_pre, &tab);
    if (rc != BK_OK) return rc;
    HIPCALL(bk::launch_phase(c->psi.p, c->n, c->precision, tab, 0u, 0u, nullptr, nullptr, c->stream));
    c->theta_pre = 0.0;
    return BK_OK;
}

bool is_fin(double x) { return std::isfinite(x); }

}  // namespace

// the energy-code cache of the held state's n, in the layout of a LOW tile of
// low_bits (0: the precision's default tile); refilled when n, precision or
// the layout change
int bkc::ensure_ecache(bk_ctx* c, int low_bits)
{
    const int lb = low_bits > 0 ? low_bits : bk::low_tile_bits(c->precision);
    if (c->ecache_n != c->n || c->ecache_prec != c->precision || c->ecache_lb != lb) {
        HIPCALL(c->ecache.ensure(bk::energy_cache_bytes(c->n)));
        HIPCALL(bk::launch_energy_fill(c->ecache.as<uint32_t>(), c->n, c->precision, c->stream, lb));
        c->ecache_n = c->n;
        c->ecache_prec = c->precision;
        c->ecache_lb = lb;
    }
    return BK_OK;
}

// H.v buffers for n spins: E(s) in state order (null below the tiled sizes,
// whose kernel computes E itself; filled on the context stream, complete on
// return) and the scratch vector of the high-qubit passes (null if none is needed)
static int hv_buffers(bk_ctx* c, int n, int is_complex, const uint16_t** out, double** scratch)
{
    *out = nullptr;
    *scratch = nullptr;
    const size_t sd = bk::hv_scratch_doubles(n, is_complex);
    if (sd > 0) {
        HIPCALL(c->hv_scratch.ensure(sizeof(double) * sd));
        *scratch = c->hv_scratch.as<double>();
    }
    if (n < 13) return BK_OK;
    if (c->hv_ecache_n != n) {
        c->hv_ecache_n = 0;
        HIPCALL(c->hv_ecache.ensure(sizeof(uint16_t) << n));
        HIPCALL(bk::launch_energy_natural(c->hv_ecache.as<uint16_t>(), n, c->stream));
        HIPCALL(hipStreamSynchronize(c->stream));
        c->hv_ecache_n = n;
    }
    *out = c->hv_ecache.as<uint16_t>();
    return BK_OK;
}

// Plain Lanczos over a device matvec on vectors of length dim
// (run_gpu_lanczos_internal, diagonalizer.cuh:420-554): v0 from mt19937_64(42),
// uniform(-0.5, 0.5); no reorthogonalisation; breakdown at |beta| < tol; the
// k_pairs lowest Ritz values (and, with evecs, the Ritz vectors, host, pair
// after pair).  Returns the number of pairs.
int bkc::lanczos_core(bk_ctx* c, int64_t dim, const std::function<hipError_t(const double*, double*)>& matvec,
                      int max_krylov_dim, double tol, int k_pairs, double* evals, double* evecs)
{
    int rc = BK_OK;
    const int m = static_cast<int>(std::min<int64_t>(dim, max_krylov_dim));
    ScopedBuf V, w;
    HIPCALL(V.ensure(sizeof(double) * static_cast<size_t>(dim) * m));
    HIPCALL(w.ensure(sizeof(double) * static_cast<size_t>(dim)));
    HIPCALL(c->blas_partials.ensure(sizeof(double) * bk::blas_partials_doubles()));
    HIPCALL(c->blas_out.ensure(sizeof(double) * 4));
    double* dV = V.as<double>();
    double* dw = w.as<double>();
    auto dot = [&](const double* x, const double* y, double* r) -> int {
        HIPCALL(bk::launch_dot(dim, 0, x, y, c->blas_out.as<double>(), c->blas_partials.as<double>(), c->stream));
        double h[2];
        HIPCALL(hipMemcpyAsync(h, c->blas_out.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCALL(hipStreamSynchronize(c->stream));
        *r = h[0];
        return BK_OK;
    };
    // v0: mt19937_64(42), uniform(-0.5, 0.5)  (diagonalizer.cuh:441-446)
    {
        std::mt19937_64 rng(42);
        std::uniform_real_distribution<double> dist(-0.5, 0.5);
        std::vector<double> v0(static_cast<size_t>(dim));
        for (auto& x : v0) x = dist(rng);
        HIPCALL(hipMemcpy(dV, v0.data(), v0.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    double nrm2 = 0.0;
    if ((rc = dot(dV, dV, &nrm2))) return rc;
    HIPCALL(bk::launch_scal(dim, 1.0 / std::sqrt(nrm2), dV, c->stream));
    std::vector<double> alpha(m, 0.0), beta(m, 0.0);
    int m_act = m;
    for (int j = 0; j < m; ++j) {
        double* vj = dV + static_cast<size_t>(j) * dim;
        HIPCALL(matvec(vj, dw));
        if (j > 0) HIPCALL(bk::launch_axpy(dim, 0, -beta[j - 1], 0.0, vj - dim, dw, c->stream));
        if ((rc = dot(vj, dw, &alpha[j]))) return rc;
        HIPCALL(bk::launch_axpy(dim, 0, -alpha[j], 0.0, vj, dw, c->stream));
        if (j < m - 1) {
            double b2 = 0.0;
            if ((rc = dot(dw, dw, &b2))) return rc;
            beta[j] = std::sqrt(b2);
            if (std::fabs(beta[j]) < tol) {
                m_act = j + 1;
                break;
            }
            double* vn = vj + dim;
            HIPCALL(hipMemcpyAsync(vn, dw, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->stream));
            HIPCALL(bk::launch_scal(dim, 1.0 / beta[j], vn, c->stream));
        }
    }
    std::vector<double> d(alpha.begin(), alpha.begin() + m_act);
    std::vector<double> e(beta.begin(), beta.begin() + std::max(0, m_act - 1));
    std::vector<double> z;
    if (!bk::tridiag_eigen(d, e, &z)) return fail(BK_E_INVALID, "tridiagonal eigensolver did not converge");
    const int kp = std::min(k_pairs, m_act);
    for (int k = 0; k < kp; ++k) evals[k] = d[k];
    if (evecs) {
        // every pair's coefficients uploaded once, ordered on the stream before
        // the combines; each combine reads its own column
        ScopedBuf coef;
        HIPCALL(coef.ensure(sizeof(double) * m_act * static_cast<size_t>(kp)));
        HIPCALL(hipMemcpyAsync(coef.p, z.data(), sizeof(double) * m_act * static_cast<size_t>(kp),
                               hipMemcpyHostToDevice, c->stream));
        for (int k = 0; k < kp; ++k) {
            HIPCALL(bk::launch_combine(dim, 0, m_act, dV, coef.as<double>() + static_cast<size_t>(k) * m_act, dw,
                                       c->stream));
            HIPCALL(hipMemcpyAsync(evecs + static_cast<size_t>(k) * dim, dw, sizeof(double) * dim,
                                   hipMemcpyDeviceToHost, c->stream));
        }
        HIPCALL(hipStreamSynchronize(c->stream));
    }
    return kp;
}


extern "C" {

const char* bk_last_error(void) { return g_err.c_str(); }

int bk_version(void) { return 100; }

int bk_set_device_memory_limit(uint64_t bytes)
{
    bkc::g_dev_limit = bytes;
    return BK_OK;
}

uint64_t bk_device_memory_held(void) { return bkc::g_dev_held.load(); }

int bk_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

static int bk_ctx_create_impl(int device, bk_ctx** out)
{
    if (!out) return fail(BK_E_INVALID, "null out pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(BK_E_NODEVICE, "no HIP device available");
    if (device < 0 || device >= count) return fail(BK_E_INVALID, "device %d of %d", device, count);
    HIPCALL(hipSetDevice(device));
    bk_ctx* c = new bk_ctx();
    c->device = device;
    c->tuning = bk::default_tuning(30, BK_C64);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(BK_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = c;
    return BK_OK;
}

int bk_ctx_create(int device, bk_ctx** out)
{
    return bkc::guard_abi([&] { return bk_ctx_create_impl(device, out); });
}

void bk_ctx_destroy(bk_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->psi.release();
    c->psi_alt.release();
    c->partials.release();
    c->series.release();
    c->count_partials.release();
    c->count_out.release();
    c->samp_sums.release();
    c->samp_prefix.release();
    c->samp_rand.release();
    c->samp_out.release();
    c->blas_partials.release();
    c->blas_out.release();
    c->ecache.release();
    c->ecache_hi.release();
    c->hv_ecache.release();
    c->hv_scratch.release();
    c->tape.ckpt.release();
    c->adj_phi.release();
    c->adj_chi.release();
    c->adj_part.release();
    c->adj_grad.release();
    c->adj_drift.release();
    for (auto& kv : c->batch_bufs) kv.second.release();
    for (auto& kv : c->batch_ecache) kv.second.release();
    for (auto& kv : c->batch_ecache_hi) kv.second.release();
    for (auto& kv : c->phase_tabs) kv.second.release();
    for (auto& kv : c->fb_tabs) {
        kv.second.master.release();
        kv.second.carry.release();
        kv.second.applied.release();
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int64_t bk_energy(int n, uint64_t s) { return bk::host_energy(n, s); }

int bk_d_min(int n, uint64_t s0, int64_t target, int max_d) { return bk::host_d_min(n, s0, target, max_d); }

int64_t bk_known_optimal_energy(int n) { return bk::host_known_optimal(n); }

int bk_make_trap(int n, uint64_t seed, uint64_t* start, uint64_t* s0, int64_t* e0)
{
    if (n < 1 || n > 64 || !start || !s0 || !e0) return fail(BK_E_INVALID, "bk_make_trap: bad argument");
    bk::host_make_trap(n, seed, start, s0, e0);
    return BK_OK;
}

static int bk_state_init_impl(bk_ctx* c, int n, int precision, uint64_t start_bits)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = ensure_state(c, n, precision))) return rc;
    const uint64_t s = start_bits & ((1ull << n) - 1ull);
    HIPCALL(bk::launch_init_basis(c->psi.p, n, precision, s, 1.0, 0.0, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    return BK_OK;
}

int bk_state_init(bk_ctx* c, int n, int precision, uint64_t start_bits)
{
    return bkc::guard_abi([&] { return bk_state_init_impl(c, n, precision, start_bits); });
}

static int bk_state_upload_impl(bk_ctx* c, int n, int precision, const void* host)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!host) return fail(BK_E_INVALID, "null host buffer");
    if ((rc = ensure_state(c, n, precision))) return rc;
    HIPCALL(hipMemcpyAsync(c->psi.p, host, amp_bytes(precision) << n, hipMemcpyHostToDevice, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    return BK_OK;
}

int bk_state_upload(bk_ctx* c, int n, int precision, const void* host)
{
    return bkc::guard_abi([&] { return bk_state_upload_impl(c, n, precision, host); });
}

static int bk_state_download_impl(bk_ctx* c, void* host)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!host) return fail(BK_E_INVALID, "null host buffer");
    if (!c->has_state) return fail(BK_E_NOSTATE, "no statevector held");
    if ((rc = flush_phase(c))) return rc;
    HIPCALL(hipMemcpyAsync(host, c->psi.p, amp_bytes(c->precision) << c->n, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    return BK_OK;
}

int bk_state_download(bk_ctx* c, void* host)
{
    return bkc::guard_abi([&] { return bk_state_download_impl(c, host); });
}

static int bk_state_device_impl(bk_ctx* c, void** d, void** st, int* n, int* precision)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->has_state) return fail(BK_E_NOSTATE, "no statevector held");
    if ((rc = flush_phase(c))) return rc;
    if (d) *d = c->psi.p;
    if (st) *st = c->stream;
    if (n) *n = c->n;
    if (precision) *precision = c->precision;
    return BK_OK;
}

int bk_state_device(bk_ctx* c, void** d, void** st, int* n, int* precision)
{
    return bkc::guard_abi([&] { return bk_state_device_impl(c, d, st, n, precision); });
}

static int bk_count_below_impl(bk_ctx* c, int n, int64_t e0, uint64_t* count)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n < 1 || n > BK_MAX_N || !count) return fail(BK_E_INVALID, "bk_count_below: bad argument");
    if (e0 <= 0) {
        *count = 0;
        return BK_OK;
    }
    const uint32_t e = static_cast<uint32_t>(std::min<int64_t>(e0, 0xFFFFFFFFll));
    HIPCALL(c->count_partials.ensure(sizeof(unsigned long long) * 4096));
    HIPCALL(c->count_out.ensure(sizeof(unsigned long long)));
    HIPCALL(bk::launch_count_below(n, e, c->count_out.as<unsigned long long>(),
                                   c->count_partials.as<unsigned long long>(), 4096, c->stream));
    unsigned long long h = 0;
    HIPCALL(hipMemcpyAsync(&h, c->count_out.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    *count = h;
    return BK_OK;
}

int bk_count_below(bk_ctx* c, int n, int64_t e0, uint64_t* count)
{
    return bkc::guard_abi([&] { return bk_count_below_impl(c, n, e0, count); });
}

static int bk_observables_impl(bk_ctx* c, uint64_t s0, int64_t e0, double* out3)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!out3) return fail(BK_E_INVALID, "null output");
    if (!c->has_state) return fail(BK_E_NOSTATE, "no statevector held");
    HIPCALL(c->series.ensure(3 * sizeof(double)));
    const uint32_t e = static_cast<uint32_t>(std::max<int64_t>(0, std::min<int64_t>(e0, 0xFFFFFFFFll)));
    HIPCALL(bk::launch_measure(c->psi.p, c->n, c->precision, static_cast<uint32_t>(s0), e, c->series.as<double>(),
                               c->partials.as<double>(), c->tuning.grid, c->stream));
    HIPCALL(hipMemcpyAsync(out3, c->series.p, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    return BK_OK;
}

int bk_observables(bk_ctx* c, uint64_t s0, int64_t e0, double* out3)
{
    return bkc::guard_abi([&] { return bk_observables_impl(c, s0, e0, out3); });
}

static int bk_state_overlap_impl(bk_ctx* a, bk_ctx* b, double* out2)
{
    int rc = check_ctx(b);
    if (rc) return rc;
    if ((rc = check_ctx(a))) return rc;
    if (!out2) return fail(BK_E_INVALID, "null output");
    if (!a->has_state || !b->has_state) return fail(BK_E_NOSTATE, "no statevector held");
    if (a->device != b->device || a->n != b->n)
        return fail(BK_E_INVALID, "states of n=%d (device %d) and n=%d (device %d)", a->n, a->device, b->n, b->device);
    if ((rc = flush_phase(b))) return rc;
    HIPCALL(hipStreamSynchronize(b->stream));
    if ((rc = flush_phase(a))) return rc;
    const int grid = 4096;
    ScopedBuf part;
    HIPCALL(part.ensure(2 * sizeof(double) * grid));
    HIPCALL(bk::launch_overlap(a->psi.p, a->precision, b->psi.p, b->precision, a->n, part.as<double>(), grid,
                               a->stream));
    std::vector<double> h(2 * grid);
    HIPCALL(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, a->stream));
    HIPCALL(hipStreamSynchronize(a->stream));
    double re = 0.0, im = 0.0;
    for (int i = 0; i < grid; ++i) {  // fixed order: the same bits every call
        re += h[2 * i];
        im += h[2 * i + 1];
    }
    out2[0] = re;
    out2[1] = im;
    return BK_OK;
}

int bk_state_overlap(bk_ctx* a, bk_ctx* b, double* out2)
{
    return bkc::guard_abi([&] { return bk_state_overlap_impl(a, b, out2); });
}

// Make the stored state carry the opening phase of a run's first step
// (phi0 = dt for Lie, dt/2 for Strang).  fresh_start: the state is reset to
// |s0> with that phase folded into its single amplitude, and the state-0
// observables are exact (P_< = [E(s0) < e0], <d_H> = 0, norm = 1).  Otherwise
// the missing phase is applied by one elementwise pass that also reduces the
// state-0 observables into slot0 (or a read-only pass when no phase is missing).
static int prepare_opening(bk_ctx* c, int order, double dt, bool fresh_start, uint64_t s0, uint32_t e0,
                           double* slot0)
{
    const double phi0 = (order == BK_STRANG) ? 0.5 * dt : dt;
    if (fresh_start) {
        // the state was just set to |s0> (bk_state_init): overwrite its amplitude
        const double th = -(phi0 / static_cast<double>(c->n)) * static_cast<double>(bk::host_energy(c->n, s0));
        const double re = std::cos(th), im = std::sin(th);
        char* dst = static_cast<char*>(c->psi.p) + amp_bytes(c->precision) * s0;
        HIPCALL(hipStreamSynchronize(c->stream));
        if (c->precision == BK_C64) {
            const float v[2] = {static_cast<float>(re), static_cast<float>(im)};
            HIPCALL(hipMemcpy(dst, v, sizeof(v), hipMemcpyHostToDevice));
        } else {
            const double v[2] = {re, im};
            HIPCALL(hipMemcpy(dst, v, sizeof(v), hipMemcpyHostToDevice));
        }
        c->theta_pre = phi0;
        if (slot0) {
            const double h[3] = {bk::host_energy(c->n, s0) < static_cast<int64_t>(e0) ? 1.0 : 0.0, 0.0, 1.0};
            HIPCALL(hipMemcpy(slot0, h, sizeof(h), hipMemcpyHostToDevice));
        }
        return BK_OK;
    }
    const double need = phi0 - c->theta_pre;
    if (need == 0.0 && slot0 && c->obs_valid && c->obs_gen == c->state_gen && c->obs_s0 == s0 && c->obs_e0 == e0) {
        // the held state is the one the last run measured at its end (same
        // reference state and threshold): its observables are known
        HIPCALL(hipMemcpyAsync(slot0, c->obs, sizeof(c->obs), hipMemcpyHostToDevice, c->stream));
        return BK_OK;
    }
    if (need != 0.0) {
        const void* tab = nullptr;
        int rc = get_phase_tab(c, c->n, c->precision, need, &tab);
        if (rc) return rc;
        HIPCALL(bk::launch_phase(c->psi.p, c->n, c->precision, tab, static_cast<uint32_t>(s0), e0, slot0,
                                 c->partials.as<double>(), c->stream));
    } else if (slot0) {
        HIPCALL(bk::launch_measure(c->psi.p, c->n, c->precision, static_cast<uint32_t>(s0), e0, slot0,
                                   c->partials.as<double>(), c->tuning.grid, c->stream));
    }
    c->theta_pre = phi0;
    return BK_OK;
}


// Enqueue one Trotter step of the held state: the rotations, then the phase
// exp(-i dt E/n) -- for Lie the next step's opening phase, for Strang this
// step's closing half plus the next step's opening half -- so theta_pre is
// unchanged.  slot: device pointer for the stepped state's observables, or null.
static int enqueue_step(bk_ctx* c, double dt, double gdt, uint64_t s0, uint32_t e0, double* slot)
{
    bk_ctx::FbTab* fb = nullptr;
    int rc = get_fb_tab(c, c->n, c->precision, dt, &fb);
    if (rc) return rc;
    bk::StepSpec sp{};
    sp.n = c->n;
    sp.precision = c->precision;
    sp.gdt = gdt;
    sp.dtp = dt;
    sp.phase_tab = fb->applied.p;
    sp.fb_master = fb->master.as<double>();
    sp.fb_carry = fb->carry.as<double>();
    sp.fb_count = fb->count;
    sp.series_slot = slot;
    sp.s0 = static_cast<uint32_t>(s0);
    sp.e0 = e0;
    if (c->energy_mode == BK_ENERGY_CACHE16 && c->n >= 13) {
        if ((rc = ensure_ecache(c))) return rc;
        sp.ecache = c->ecache.as<uint32_t>();
    }
    ++c->state_gen;
    void* alt = nullptr;
    if (c->tuning.pingpong && c->n >= 13) {
        HIPCALL(c->psi_alt.ensure(amp_bytes(c->precision) << c->n));
        alt = c->psi_alt.p;
    }
    int in_alt = 0;
    HIPCALL(bk::launch_step(c->psi.p, alt, &in_alt, sp, c->tuning, c->partials.as<double>(), c->stream));
    if (in_alt) std::swap(c->psi, c->psi_alt);
    return BK_OK;
}

// E(s) of the held n in the layout of the phase-carrying high group hs
static int ensure_ecache_high(bk_ctx* c, const bk::HighSweep& hs)
{
    const std::vector<int> key = {c->n, c->precision, hs.h1, hs.a1, hs.h2, hs.a2, hs.R};
    if (c->ecache_hi_key != key) {
        c->ecache_hi_key.clear();
        HIPCALL(c->ecache_hi.ensure(bk::energy_cache_bytes(c->n)));
        HIPCALL(bk::launch_energy_fill_high(c->ecache_hi.as<uint32_t>(), c->n, c->precision, hs, c->stream));
        c->ecache_hi_key = key;
    }
    return BK_OK;
}

static int validate_evolve(const bk_evolve_params* p)
{
    if (!p) return fail(BK_E_INVALID, "null params");
    if (p->n < 1 || p->n > BK_MAX_N) return fail(BK_E_INVALID, "n=%d outside [1, %d]", p->n, BK_MAX_N);
    if (p->precision != BK_C64 && p->precision != BK_C128) return fail(BK_E_INVALID, "precision=%d", p->precision);
    if (p->order != BK_LIE && p->order != BK_STRANG) return fail(BK_E_INVALID, "order=%d", p->order);
    if (!is_fin(p->dt) || !is_fin(p->g) || p->steps < 0) return fail(BK_E_INVALID, "non-finite dt/g or steps < 0");
    if (p->g_steps)
        for (int k = 0; k < p->steps; ++k)
            if (!is_fin(p->g_steps[k])) return fail(BK_E_INVALID, "non-finite g_steps[%d]", k);
    return BK_OK;
}

// Enqueue gdt.size() Trotter steps of the held state (g dt of each step in gdt,
// observables of step k into slots[k] when non-null): a fused run when the plan
// allows, else step by step.  theta_pre is unchanged.
static int enqueue_run(bk_ctx* c, double dt, const std::vector<double>& gdt, const std::vector<double*>& slots,
                       uint64_t s0, uint32_t e0)
{
    const int K = static_cast<int>(gdt.size());
    if (K == 0) return BK_OK;
    const int lb = bk::fused_low_bits(c->n, c->precision, c->tuning);
    std::vector<bk::HighSweep> grp;
    if (c->energy_mode == BK_ENERGY_CACHE16 && !c->tuning.pingpong)
        grp = bk::fused_sweeps(c->n, c->precision, c->tuning, lb);
    // a run that would fall back to single steps (a (c, s) step) keeps the default LOW tile
    for (double x : gdt)
        if (!bk::mix_info(c->n, x, c->precision).tan_form) grp.clear();
    const bool fusable = !grp.empty();
    if (!fusable) {
        for (int k = 0; k < K; ++k) {
            const int rc = enqueue_step(c, dt, gdt[static_cast<size_t>(k)], s0, e0, slots[static_cast<size_t>(k)]);
            if (rc) return rc;
        }
        return BK_OK;
    }
    bk_ctx::FbTab* fb = nullptr;
    int rc = get_fb_tab(c, c->n, c->precision, dt, &fb);
    if (rc) return rc;
    if ((rc = ensure_ecache(c, lb))) return rc;
    if ((rc = ensure_ecache_high(c, grp.back()))) return rc;
    bk::RunSpec rs{};
    rs.low_bits = lb;
    rs.n = c->n;
    rs.precision = c->precision;
    rs.steps = K;
    rs.gdt = gdt.data();
    rs.slots = slots.data();
    rs.phase_tab = fb->applied.p;
    rs.phase_tab2 = static_cast<char*>(fb->applied.p) + static_cast<size_t>(fb->count) * amp_bytes(c->precision);
    rs.fb_master = fb->master.as<double>();
    rs.fb_carry = fb->carry.as<double>();
    rs.fb_count = fb->count;
    rs.s0 = static_cast<uint32_t>(s0);
    rs.e0 = e0;
    rs.ecache = c->ecache.as<uint32_t>();
    rs.ecache_high = c->ecache_hi.as<uint32_t>();
    HIPCALL(bk::launch_run(c->psi.p, rs, c->tuning, c->partials.as<double>(), c->stream));
    ++c->state_gen;
    return BK_OK;
}

static int bk_evolve_impl(bk_ctx* c, const bk_evolve_params* p, bk_series* out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = validate_evolve(p))) return rc;
    const uint64_t mask = (1ull << p->n) - 1ull;
    const uint64_t s0 = p->start_bits & mask;
    const int64_t e0 = p->e0 >= 0 ? p->e0 : bk::host_energy(p->n, s0);
    if (p->init_from_start) {
        if ((rc = bk_state_init(c, p->n, p->precision, s0))) return rc;
    } else {
        if (!c->has_state) return fail(BK_E_NOSTATE, "init_from_start=0 but no state held");
        if (c->n != p->n || c->precision != p->precision)
            return fail(BK_E_INVALID, "held state is n=%d prec=%d", c->n, c->precision);
    }
    // measured state indices: 0 and multiples of measure_every, plus the last
    const int every = p->measure_every;
    const int K = p->steps;
    std::vector<int> meas;
    if (every > 0) meas.push_back(0);
    for (int st = 1; st <= K; ++st)
        if ((every > 0 && st % every == 0) || st == K) meas.push_back(st);
    if (meas.empty()) meas.push_back(0);  // K == 0 and every == 0: the state itself
    const int64_t slots = static_cast<int64_t>(meas.size());
    if (out && out->capacity < slots && (out->p_less || out->dh || out->norm))
        return fail(BK_E_INVALID, "series capacity %lld < %lld", static_cast<long long>(out->capacity),
                    static_cast<long long>(slots));
    HIPCALL(c->series.ensure(3 * sizeof(double) * static_cast<size_t>(slots)));
    double* ds = c->series.as<double>();
    const uint32_t e0u = static_cast<uint32_t>(std::max<int64_t>(0, std::min<int64_t>(e0, 0xFFFFFFFFll)));
    int64_t slot = 0;
    double* slot0 = (meas[0] == 0) ? ds : nullptr;
    if ((rc = prepare_opening(c, p->order, p->dt, p->init_from_start != 0, s0, e0u, slot0))) return rc;
    if (slot0) slot = 1;
    std::vector<double> gdt(static_cast<size_t>(K));
    std::vector<double*> sls(static_cast<size_t>(K), nullptr);
    for (int k = 0; k < K; ++k) {
        sls[static_cast<size_t>(k)] =
            (slot < slots && meas[static_cast<size_t>(slot)] == k + 1) ? ds + 3 * (slot++) : nullptr;
        gdt[static_cast<size_t>(k)] = (p->g_steps ? p->g_steps[k] : p->g) * p->dt;
    }
    if ((rc = enqueue_run(c, p->dt, gdt, sls, s0, e0u))) return rc;
    if (out) {
        std::vector<double> h(3 * static_cast<size_t>(slot));
        HIPCALL(hipMemcpyAsync(h.data(), ds, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCALL(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < slot; ++i) {
            if (out->p_less) out->p_less[i] = h[3 * i];
            if (out->dh) out->dh[i] = h[3 * i + 1];
            if (out->norm) out->norm[i] = h[3 * i + 2];
        }
        if (slot > 0 && meas[static_cast<size_t>(slot - 1)] == K) {  // the held state's observables
            std::copy(h.end() - 3, h.end(), c->obs);
            c->obs_valid = true;
            c->obs_gen = c->state_gen;
            c->obs_s0 = s0;
            c->obs_e0 = e0u;
        }
        out->count = slot;
        out->e0 = e0;
        out->p_blind = -1.0;  // only computed for a fresh start (a continuation keeps E0)
        if (p->init_from_start) {
            uint64_t cnt = 0;
            if ((rc = bk_count_below(c, p->n, e0, &cnt))) return rc;
            out->p_blind = static_cast<double>(cnt) / static_cast<double>(1ull << p->n);
        }
    } else {
        HIPCALL(hipStreamSynchronize(c->stream));
    }
    return BK_OK;
}

int bk_evolve(bk_ctx* c, const bk_evolve_params* p, bk_series* out)
{
    return bkc::guard_abi([&] { return bk_evolve_impl(c, p, out); });
}

// ---- adjoint gradient ----
// Stored states carry P(theta) = exp(-i theta E/n), theta = the opening phase
// (phi_k = P(theta) psi_k), so one step is phi_k = P(dt) X(g_k dt) phi_{k-1}
// whatever the order.  The reverse pass holds phi'_k = P(-dt) phi_k and the
// adjoint chi'_k = P(-dt) P(theta) lambda_k (lambda_T = W psi_T, W the
// objective's diagonal weight): then phi'_{k-1} = P(-dt) X(-g_k dt) phi'_k,
// the same for chi', and dL/dg_k = 2 Re <lambda_k| dU_k/dg |psi_{k-1}>
// = -2 dt Im <chi'_k| sum_q X_q |phi'_k> (dX/dg = i dt sum_q X_q X).
static int bk_adjoint_forward_impl(bk_ctx* c, const bk_adjoint_params* p, double* obs3)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!p) return fail(BK_E_INVALID, "null params");
    const bk_evolve_params& ev = p->ev;
    if ((rc = validate_evolve(&ev))) return rc;
    if (ev.steps < 1) return fail(BK_E_INVALID, "adjoint needs steps >= 1");
    if (!ev.init_from_start) return fail(BK_E_INVALID, "adjoint runs start from |start_bits> (init_from_start = 1)");
    if (p->checkpoint_every < 1) return fail(BK_E_INVALID, "checkpoint_every=%d < 1", p->checkpoint_every);
    c->tape.valid = false;
    const uint64_t s0 = ev.start_bits & ((1ull << ev.n) - 1ull);
    const int64_t e0 = ev.e0 >= 0 ? ev.e0 : bk::host_energy(ev.n, s0);
    const uint32_t e0u = static_cast<uint32_t>(std::max<int64_t>(0, std::min<int64_t>(e0, 0xFFFFFFFFll)));
    if ((rc = bk_state_init(c, ev.n, ev.precision, s0))) return rc;
    const size_t bytes = amp_bytes(ev.precision) << ev.n;
    const int K = p->checkpoint_every;
    const int nck = (ev.steps - 1) / K;
    if (nck > 0) {
        hipError_t e = c->tape.ckpt.ensure(bytes * static_cast<size_t>(nck));
        if (e != hipSuccess)
            return fail(e == hipErrorOutOfMemory ? BK_E_NOMEM : BK_E_HIP, "adjoint tape of %d states (%.2f GB): %s", nck,
                        1e-9 * static_cast<double>(bytes) * nck, hipGetErrorString(e));
    }
    HIPCALL(c->series.ensure(3 * sizeof(double)));
    if ((rc = prepare_opening(c, ev.order, ev.dt, true, s0, e0u, nullptr))) return rc;
    char* tape = static_cast<char*>(c->tape.ckpt.p);
    // one run per checkpoint interval: a run leaves the state in the plain
    // frame, which is what the tape holds
    for (int k = 0; k < ev.steps;) {
        const int L = std::min(K, ev.steps - k);
        std::vector<double> gdt(static_cast<size_t>(L));
        std::vector<double*> slots(static_cast<size_t>(L), nullptr);
        for (int i = 0; i < L; ++i) gdt[static_cast<size_t>(i)] = (ev.g_steps ? ev.g_steps[k + i] : ev.g) * ev.dt;
        if (k + L == ev.steps) slots.back() = c->series.as<double>();
        if ((rc = enqueue_run(c, ev.dt, gdt, slots, s0, e0u))) return rc;
        k += L;
        if (k % K == 0 && k < ev.steps)
            HIPCALL(hipMemcpyAsync(tape + bytes * static_cast<size_t>(k / K - 1), c->psi.p, bytes,
                                   hipMemcpyDeviceToDevice, c->stream));
    }
    double h[3];
    HIPCALL(hipMemcpyAsync(h, c->series.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    if (obs3) std::memcpy(obs3, h, sizeof(h));
    auto& t = c->tape;
    t.n = ev.n;
    t.precision = ev.precision;
    t.steps = ev.steps;
    t.every = K;
    t.nckpt = nck;
    t.dt = ev.dt;
    t.theta = c->theta_pre;
    t.s0 = s0;
    t.e0 = e0u;
    t.g.resize(static_cast<size_t>(ev.steps));
    for (int k = 0; k < ev.steps; ++k) t.g[static_cast<size_t>(k)] = ev.g_steps ? ev.g_steps[k] : ev.g;
    t.gen = c->state_gen;
    t.valid = true;
    return BK_OK;
}

int bk_adjoint_forward(bk_ctx* c, const bk_adjoint_params* p, double* obs3)
{
    return bkc::guard_abi([&] { return bk_adjoint_forward_impl(c, p, obs3); });
}

static int bk_adjoint_backward_impl(bk_ctx* c, double w_pless, double w_dh, double* grad_steps, double* drift)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!grad_steps || !is_fin(w_pless) || !is_fin(w_dh)) return fail(BK_E_INVALID, "bk_adjoint_backward: bad argument");
    auto& t = c->tape;
    if (!t.valid || !c->has_state || t.gen != c->state_gen || c->n != t.n || c->precision != t.precision)
        return fail(BK_E_NOSTATE, "no adjoint tape for the held state (run bk_adjoint_forward first)");
    const int n = t.n, prec = t.precision;
    const size_t bytes = amp_bytes(prec) << n;
    const int low_tb = bk::default_tuning(n, prec).tile_bits;
    HIPCALL(c->adj_phi.ensure(bytes));
    HIPCALL(c->adj_chi.ensure(bytes));
    HIPCALL(c->adj_grad.ensure(sizeof(double) * static_cast<size_t>(t.steps)));
    HIPCALL(c->adj_drift.ensure(sizeof(double) * static_cast<size_t>(std::max(t.nckpt, 1))));
    HIPCALL(c->adj_part.ensure(sizeof(double) * std::max<size_t>(4096, bk::reverse_partials_count(n, prec, c->tuning))));
    bk::RevRunSpec rs{};
    rs.n = n;
    rs.precision = prec;
    if (n >= low_tb) {
        if ((rc = ensure_ecache(c))) return rc;
        rs.ecache = c->ecache.as<uint32_t>();
        const std::vector<bk::HighSweep> grp = bk::fused_sweeps(n, prec, c->tuning);
        if (!grp.empty()) {
            if ((rc = ensure_ecache_high(c, grp.back()))) return rc;
            rs.ecache_high = c->ecache_hi.as<uint32_t>();
        }
    }
    const void* tab_init = nullptr;
    const void* tab_rev = nullptr;
    // phi'_T = P(-dt) P(theta_tape - theta_now) * held state
    if ((rc = get_phase_tab(c, n, prec, t.theta - c->theta_pre - t.dt, &tab_init))) return rc;
    if ((rc = get_phase_tab(c, n, prec, -t.dt, &tab_rev))) return rc;
    rs.phase_tab = tab_rev;
    if (prec == BK_C64 && rs.ecache_high) {
        // the fused complex64 runs fold each step's c^n into this table; its
        // carry restarts with every backward so that repeated calls agree bit for bit
        bk_ctx::FbTab* fb = nullptr;
        if ((rc = get_fb_tab(c, n, prec, -t.dt, &fb))) return rc;
        HIPCALL(bk::launch_carry_reset(fb->carry.as<double>(), fb->count, c->stream));
        rs.fb_applied = fb->applied.p;
        rs.fb_applied2 = static_cast<char*>(fb->applied.p) + static_cast<size_t>(fb->count) * amp_bytes(prec);
        rs.fb_master = fb->master.as<double>();
        rs.fb_carry = fb->carry.as<double>();
        rs.fb_count = fb->count;
    }
    HIPCALL(bk::launch_adj_init(c->psi.p, c->adj_phi.p, c->adj_chi.p, n, prec, tab_init, static_cast<uint32_t>(t.s0), t.e0,
                                w_pless, w_dh, c->stream));
    const char* tape = static_cast<const char*>(t.ckpt.p);
    double* dg = c->adj_grad.as<double>();
    double* dd = c->adj_drift.as<double>();
    double* part = c->adj_part.as<double>();
    bk::SweepTuning tun = c->tuning;
    tun.timer = c->timing ? &c->timer : nullptr;
    // reverse steps k = steps..1 in runs that end at the checkpoints, where
    // phi is replaced by the taped state (and the difference recorded)
    std::vector<double> gdt;
    std::vector<double*> slots;
    for (int k = t.steps; k >= 1;) {
        if (k < t.steps && k % t.every == 0) {
            const int i = k / t.every - 1;
            HIPCALL(bk::launch_phase_from(c->adj_phi.p, tape + bytes * static_cast<size_t>(i), n, prec, tab_rev, part,
                                          dd + i, c->stream));
        }
        const int lo = ((k - 1) / t.every) * t.every;
        gdt.clear();
        slots.clear();
        for (int j = k; j > lo; --j) {
            gdt.push_back(-t.g[static_cast<size_t>(j - 1)] * t.dt);
            slots.push_back(dg + (j - 1));
        }
        rs.steps = k - lo;
        rs.gdt = gdt.data();
        rs.grad_slots = slots.data();
        HIPCALL(bk::launch_reverse_run(c->adj_phi.p, c->adj_chi.p, rs, tun, part, c->stream));
        k = lo;
    }
    std::vector<double> raw(static_cast<size_t>(t.steps));
    std::vector<double> dr(static_cast<size_t>(std::max(t.nckpt, 1)), 0.0);
    HIPCALL(hipMemcpyAsync(raw.data(), dg, raw.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (t.nckpt > 0)
        HIPCALL(hipMemcpyAsync(dr.data(), dd, static_cast<size_t>(t.nckpt) * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    for (int k = 0; k < t.steps; ++k) grad_steps[k] = -2.0 * t.dt * raw[static_cast<size_t>(k)];
    if (drift) *drift = t.nckpt > 0 ? *std::max_element(dr.begin(), dr.begin() + t.nckpt) : 0.0;
    return BK_OK;
}

int bk_adjoint_backward(bk_ctx* c, double w_pless, double w_dh, double* grad_steps, double* drift)
{
    return bkc::guard_abi([&] { return bk_adjoint_backward_impl(c, w_pless, w_dh, grad_steps, drift); });
}

static int bk_adjoint_grad_impl(bk_ctx* c, const bk_adjoint_params* p, double w_pless, double w_dh, double* obs3,
                                double* grad_steps, double* drift)
{
    int rc = bk_adjoint_forward(c, p, obs3);
    if (rc) return rc;
    return bk_adjoint_backward(c, w_pless, w_dh, grad_steps, drift);
}

int bk_adjoint_grad(bk_ctx* c, const bk_adjoint_params* p, double w_pless, double w_dh, double* obs3,
                    double* grad_steps, double* drift)
{
    return bkc::guard_abi([&] { return bk_adjoint_grad_impl(c, p, w_pless, w_dh, obs3, grad_steps, drift); });
}

static int bk_sample_impl(bk_ctx* c, const double* randnums, int64_t shots, int64_t* out_idx)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (shots < 0 || (shots > 0 && (!randnums || !out_idx))) return fail(BK_E_INVALID, "bk_sample: bad argument");
    if (!c->has_state) return fail(BK_E_NOSTATE, "no statevector held");
    if (shots == 0) return BK_OK;
    if (shots > (1ll << 30)) return fail(BK_E_INVALID, "too many shots");
    const size_t nsc = bk::sample_scratch_doubles(c->n);
    HIPCALL(c->samp_sums.ensure(nsc * sizeof(double)));
    HIPCALL(c->samp_prefix.ensure(nsc * sizeof(double)));
    HIPCALL(c->samp_rand.ensure(static_cast<size_t>(shots) * sizeof(double)));
    HIPCALL(c->samp_out.ensure(static_cast<size_t>(shots) * sizeof(int64_t)));
    HIPCALL(hipMemcpyAsync(c->samp_rand.p, randnums, static_cast<size_t>(shots) * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
    HIPCALL(bk::launch_sample(c->psi.p, c->n, c->precision, c->samp_rand.as<double>(), static_cast<int>(shots),
                              c->samp_out.as<int64_t>(), c->samp_sums.as<double>(), c->samp_prefix.as<double>(),
                              c->stream));
    HIPCALL(hipMemcpyAsync(out_idx, c->samp_out.p, static_cast<size_t>(shots) * sizeof(int64_t),
                           hipMemcpyDeviceToHost, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    return BK_OK;
}

int bk_sample(bk_ctx* c, const double* randnums, int64_t shots, int64_t* out_idx)
{
    return bkc::guard_abi([&] { return bk_sample_impl(c, randnums, shots, out_idx); });
}

static int bk_hv_impl(bk_ctx* c, int n, double g, int normalize_by_n, int is_complex, const double* d_x, double* d_y,
                      void* stream)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n < 1 || n > BK_MAX_N || !d_x || !d_y || d_x == d_y || !is_fin(g))
        return fail(BK_E_INVALID, "bk_hv: bad argument");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const uint16_t* ec = nullptr;
    double* scr = nullptr;
    if ((rc = hv_buffers(c, n, is_complex ? 1 : 0, &ec, &scr))) return rc;
    HIPCALL(bk::launch_hv(n, g, normalize_by_n, is_complex, d_x, d_y, st, ec, scr));
    HIPCALL(hipStreamSynchronize(st));
    return BK_OK;
}

int bk_hv(bk_ctx* c, int n, double g, int normalize_by_n, int is_complex, const double* d_x, double* d_y,
          void* stream)
{
    return bkc::guard_abi([&] { return bk_hv_impl(c, n, g, normalize_by_n, is_complex, d_x, d_y, stream); });
}

static int bk_lanczos_impl(bk_ctx* c, int n, double g, int normalize_by_n, int max_krylov_dim, double tol, int k_pairs,
                           double* evals, double* evecs)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n < 1 || n > 30 || max_krylov_dim < 1 || k_pairs < 1 || !evals || !is_fin(g))
        return fail(BK_E_INVALID, "bk_lanczos: bad argument");
    const uint16_t* ec = nullptr;
    double* scr = nullptr;
    if ((rc = hv_buffers(c, n, 0, &ec, &scr))) return rc;
    return lanczos_core(
        c, 1ll << n,
        [&](const double* x, double* y) { return bk::launch_hv(n, g, normalize_by_n, 0, x, y, c->stream, ec, scr); },
        max_krylov_dim, tol, k_pairs, evals, evecs);
}

int bk_lanczos(bk_ctx* c, int n, double g, int normalize_by_n, int max_krylov_dim, double tol, int k_pairs,
               double* evals, double* evecs)
{
    return bkc::guard_abi(
        [&] { return bk_lanczos_impl(c, n, g, normalize_by_n, max_krylov_dim, tol, k_pairs, evals, evecs); });
}

static int bk_krylov_propagate_impl(bk_ctx* c, double g, int normalize_by_n, double tau, double tau_restart, int m)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->has_state) return fail(BK_E_NOSTATE, "no statevector held");
    if (c->precision != BK_C128) return fail(BK_E_INVALID, "Krylov propagation needs a complex128 state");
    if (m < 2 || !is_fin(g) || !is_fin(tau) || !(tau_restart > 0.0)) return fail(BK_E_INVALID, "bad argument");
    if ((rc = flush_phase(c))) return rc;
    ++c->state_gen;
    const int n = c->n;
    const int64_t dim = 1ll << n;
    const uint16_t* ec = nullptr;
    double* scr = nullptr;
    if ((rc = hv_buffers(c, n, 1, &ec, &scr))) return rc;
    ScopedBuf V, w, coef, gsc, gsp;  // freed on every return path
    HIPCALL(V.ensure(sizeof(double) * 2 * static_cast<size_t>(dim) * m));
    HIPCALL(w.ensure(sizeof(double) * 2 * static_cast<size_t>(dim)));
    HIPCALL(coef.ensure(sizeof(double) * 2 * m));
    HIPCALL(gsc.ensure(sizeof(double) * (4 * static_cast<size_t>(bk::kMdMax) + 2)));  // both sweeps' coefficients, |w|^2
    HIPCALL(gsp.ensure(sizeof(double) * bk::multidot_partials_doubles()));
    HIPCALL(c->blas_partials.ensure(sizeof(double) * bk::blas_partials_doubles()));
    HIPCALL(c->blas_out.ensure(sizeof(double) * 4));
    double* dV = V.as<double>();
    double* dw = w.as<double>();
    double* psi = c->psi.as<double>();
    auto zdot = [&](const double* x, const double* y, double* re, double* im) -> int {
        HIPCALL(bk::launch_dot(dim, 1, x, y, c->blas_out.as<double>(), c->blas_partials.as<double>(), c->stream));
        double h[2];
        HIPCALL(hipMemcpyAsync(h, c->blas_out.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCALL(hipStreamSynchronize(c->stream));
        *re = h[0];
        if (im) *im = h[1];
        return BK_OK;
    };
    const int nsub = std::max(1, static_cast<int>(std::ceil(std::fabs(tau) / tau_restart - 1e-12)));
    const double h = tau / nsub;
    for (int sub = 0; sub < nsub; ++sub) {
        double b0sq = 0.0;
        if ((rc = zdot(psi, psi, &b0sq, nullptr))) return rc;
        const double b0 = std::sqrt(b0sq);
        if (b0 == 0.0) return BK_OK;
        HIPCALL(hipMemcpyAsync(dV, psi, sizeof(double) * 2 * dim, hipMemcpyDeviceToDevice, c->stream));
        HIPCALL(bk::launch_scal(2 * dim, 1.0 / b0, dV, c->stream));
        std::vector<double> al(m, 0.0), be(m, 0.0);
        int mm = m;
        for (int j = 0; j < m; ++j) {
            double* vj = dV + static_cast<size_t>(j) * 2 * dim;
            HIPCALL(bk::launch_hv(n, g, normalize_by_n, 1, vj, dw, c->stream, ec, scr));
            // classical Gram-Schmidt twice against v_0..v_j; the coefficient of
            // v_j in each sweep adds to alpha_j (each sweep keeps its coefficients
            // in its own region of dgs)
            double* dgs = gsc.as<double>();
            double* vn = j + 1 < m ? vj + 2 * dim : nullptr;
            const bool fused = j + 1 <= bk::kMdMax;
            double aj[2] = {0.0, 0.0};
            double bsq = 0.0;
            if (fused) {
                // one block: three passes, the second sweep's subtraction writing
                // v_{j+1} (unnormalised) and |v_{j+1}|^2
                HIPCALL(bk::launch_cgs2(dim, j + 1, dV, dw, vn, dgs, gsp.as<double>(), c->stream));
                for (int pass = 0; pass < 2; ++pass)
                    HIPCALL(hipMemcpyAsync(aj + pass, dgs + 2 * bk::kMdMax * pass + 2 * j, sizeof(double),
                                           hipMemcpyDeviceToHost, c->stream));
                if (vn)
                    HIPCALL(hipMemcpyAsync(&bsq, dgs + 4 * bk::kMdMax, sizeof(double), hipMemcpyDeviceToHost,
                                           c->stream));
                HIPCALL(hipStreamSynchronize(c->stream));
            } else {
                // blocks of at most kMdMax vectors; the last block of a sweep,
                // which holds v_j, is the one left in its region
                for (int pass = 0; pass < 2; ++pass)
                    for (int i0 = 0; i0 <= j; i0 += bk::kMdMax) {
                        const int mb = std::min(bk::kMdMax, j + 1 - i0);
                        HIPCALL(bk::launch_gs_block(dim, mb, dV + static_cast<size_t>(i0) * 2 * dim, dw,
                                                    dgs + 2 * bk::kMdMax * pass, gsp.as<double>(), c->stream));
                    }
                const int last = j % bk::kMdMax;  // v_j's position in its block
                for (int pass = 0; pass < 2; ++pass)
                    HIPCALL(hipMemcpyAsync(aj + pass, dgs + 2 * bk::kMdMax * pass + 2 * last, sizeof(double),
                                           hipMemcpyDeviceToHost, c->stream));
                if ((rc = zdot(dw, dw, &bsq, nullptr))) return rc;  // synchronises the stream (aj too)
            }
            al[j] = aj[0] + aj[1];
            if (j == m - 1) break;
            const double bj = std::sqrt(bsq);
            if (bj < 1e-13 * (std::fabs(al[j]) + 1.0)) {
                mm = j + 1;
                break;
            }
            be[j] = bj;
            if (!fused)
                HIPCALL(hipMemcpyAsync(vn, dw, sizeof(double) * 2 * dim, hipMemcpyDeviceToDevice, c->stream));
            HIPCALL(bk::launch_scal(2 * dim, 1.0 / bj, vn, c->stream));
        }
        std::vector<double> d(al.begin(), al.begin() + mm);
        std::vector<double> e(be.begin(), be.begin() + std::max(0, mm - 1));
        std::vector<double> z;
        if (!bk::tridiag_eigen(d, e, &z)) return fail(BK_E_INVALID, "tridiagonal eigensolver did not converge");
        // y = b0 * S exp(-i L h) S^T e1
        std::vector<double> y(2 * static_cast<size_t>(mm), 0.0);
        for (int k = 0; k < mm; ++k) {
            const double s0k = z[static_cast<size_t>(k) * mm];
            const double cr = std::cos(-d[k] * h) * s0k * b0, ci = std::sin(-d[k] * h) * s0k * b0;
            for (int j = 0; j < mm; ++j) {
                y[2 * j] += z[static_cast<size_t>(k) * mm + j] * cr;
                y[2 * j + 1] += z[static_cast<size_t>(k) * mm + j] * ci;
            }
        }
        HIPCALL(hipMemcpyAsync(coef.p, y.data(), sizeof(double) * y.size(), hipMemcpyHostToDevice, c->stream));
        HIPCALL(bk::launch_combine(dim, 1, mm, dV, coef.as<double>(), psi, c->stream));
        HIPCALL(hipStreamSynchronize(c->stream));
    }
    return BK_OK;
}

int bk_krylov_propagate(bk_ctx* c, double g, int normalize_by_n, double tau, double tau_restart, int m)
{
    return bkc::guard_abi([&] { return bk_krylov_propagate_impl(c, g, normalize_by_n, tau, tau_restart, m); });
}

void bk_sampler_config_defaults(bk_sampler_config* cfg)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->n = 16;
    cfg->g = 0.5;
    cfg->dt = 0.05;
    cfg->t_max = 60.0;
    cfg->num_sample_points = 20;
    cfg->shots_per_point = 100;
    cfg->seed = 42;
    cfg->num_sampling_seeds = 3;
    cfg->seed_stride = 10007;
    cfg->t_start = 0.0;
    cfg->dt_sample = 0.0;
    cfg->track_p_less = 1;
    cfg->verbose = 1;
    cfg->record_trajectory = 0;
    cfg->dt_trajectory = 0.0;
    cfg->precision = BK_C128;
    cfg->order = BK_LIE;
}

void bk_result_free(bk_sampling_result* r)
{
    if (!r) return;
    delete[] r->unique_basis;
    delete[] r->seed_stats;
    delete[] r->time_points;
    delete[] r->p_less_history;
    delete[] r->dh_history;
    delete[] r->samples;
    delete r;
}

static int bk_sampler_run_impl(bk_ctx* c, const bk_sampler_config* cfg, uint64_t start_bits, bk_sampling_result** out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!cfg || !out) return fail(BK_E_INVALID, "null argument");
    *out = nullptr;
    const int n = cfg->n;
    if (n < 1 || n > BK_MAX_N) return fail(BK_E_INVALID, "n=%d outside [1, %d]", n, BK_MAX_N);
    if (!is_fin(cfg->g) || !is_fin(cfg->dt) || !(cfg->dt > 0.0) || !is_fin(cfg->t_max) || cfg->t_max < 0.0)
        return fail(BK_E_INVALID, "non-finite or non-positive g/dt/t_max");
    const int precision = cfg->precision ? cfg->precision : BK_C128;
    const int order = cfg->order ? cfg->order : BK_LIE;
    const uint64_t mask = (1ull << n) - 1ull;
    const uint64_t start = start_bits & mask;

    std::vector<double> tp, pl, dh;
    bk_sampling_result* r = new bk_sampling_result();
    std::memset(r, 0, sizeof(*r));
    // freed on every early return; released to the caller on success
    std::unique_ptr<bk_sampling_result, void (*)(bk_sampling_result*)> guard(r, bk_result_free);
    auto cleanup_fail = [&](int code) { return code; };
    // sampler.cuh:420-423
    r->start_energy = bk::host_energy(n, start);
    if (n >= 20) {  // radii up to ~13 at n = 30: 10^8 and more energies, on the device
        int dm = 0;
        const int64_t tgt = r->start_energy;
        if ((rc = bk_d_min_batch(c, n, &start, &tgt, 1, -1, &dm))) return cleanup_fail(rc);
        r->d_min = dm;
    } else {
        r->d_min = bk::host_d_min(n, start, r->start_energy, -1);
    }
    if ((rc = bk_state_init(c, n, precision, start))) return cleanup_fail(rc);
    uint64_t cnt = 0;
    if ((rc = bk_count_below(c, n, r->start_energy, &cnt))) return cleanup_fail(rc);
    r->p_blind = static_cast<double>(cnt) / static_cast<double>(1ull << n);

    bk::ScheduleIn si{};
    si.t_max = cfg->t_max;
    si.dt = cfg->dt;
    si.t_start = cfg->t_start;
    si.dt_sample = cfg->dt_sample;
    si.dt_trajectory = cfg->dt_trajectory;
    si.num_sample_points = cfg->num_sample_points;
    si.shots_per_point = cfg->shots_per_point;
    si.track_p_less = cfg->track_p_less != 0;
    si.record_trajectory = cfg->record_trajectory != 0;
    std::vector<unsigned char> flags;
    const int total = bk::build_schedule(si, flags);

    int n_meas = 0, n_pts = 0;
    for (unsigned char f : flags) {
        n_meas += (f & 1) ? 1 : 0;
        n_pts += (f & 2) ? 1 : 0;
    }
    const int shots = std::max(1, cfg->shots_per_point);
    const int n_seeds = std::max(1, cfg->num_sampling_seeds);
    // random numbers in the reference's draw order (sampler.cuh:481-487, 567-568)
    std::vector<std::mt19937_64> rngs(n_seeds);
    for (int s = 0; s < n_seeds; ++s) rngs[s].seed(cfg->seed + static_cast<uint64_t>(s) * cfg->seed_stride);
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    const size_t per_point = static_cast<size_t>(n_seeds) * shots;
    std::vector<double> rnd(per_point * static_cast<size_t>(n_pts));
    for (int ptn = 0; ptn < n_pts; ++ptn)
        for (int s = 0; s < n_seeds; ++s)
            for (int i = 0; i < shots; ++i) rnd[ptn * per_point + s * shots + i] = dist(rngs[s]);
    std::vector<int64_t> idx(rnd.size());
    if (n_pts > 0) {
        const size_t nsc = bk::sample_scratch_doubles(n);
        HIPCALL(c->samp_sums.ensure(nsc * sizeof(double)));
        HIPCALL(c->samp_prefix.ensure(nsc * sizeof(double)));
        HIPCALL(c->samp_rand.ensure(rnd.size() * sizeof(double)));
        HIPCALL(c->samp_out.ensure(rnd.size() * sizeof(int64_t)));
        HIPCALL(hipMemcpyAsync(c->samp_rand.p, rnd.data(), rnd.size() * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
    }
    HIPCALL(c->series.ensure(3 * sizeof(double) * static_cast<size_t>(std::max(n_meas, 1))));
    double* ds = c->series.as<double>();
    const uint32_t e0u = static_cast<uint32_t>(std::max<int64_t>(0, r->start_energy));
    int mi = 0, pi = 0;
    // state 0 (|start> with the first step's opening phase folded in)
    double* slot0 = (flags[0] & 1) ? ds : nullptr;
    if ((rc = prepare_opening(c, order, cfg->dt, true, start, e0u, slot0))) return cleanup_fail(rc);
    if (slot0) mi = 1;
    // the steps between two sampling points go as one run (a run ends in the
    // plain frame, whose |psi|^2 the sampler reads)
    std::vector<double> run_gdt;
    std::vector<double*> run_slots;
    auto flush_run = [&]() -> int {
        if (run_gdt.empty()) return BK_OK;
        const int e = enqueue_run(c, cfg->dt, run_gdt, run_slots, start, e0u);
        run_gdt.clear();
        run_slots.clear();
        return e;
    };
    for (int step = 0; step <= total; ++step) {
        const unsigned char f = flags[static_cast<size_t>(step)];
        if (f & 1) tp.push_back(step * cfg->dt);
        if (f & 2) {  // sampling reads |psi|^2 of state 'step' (before this step's evolution)
            if ((rc = flush_run())) return cleanup_fail(rc);
            for (int s = 0; s < n_seeds; ++s) {
                const size_t off = pi * per_point + static_cast<size_t>(s) * shots;
                HIPCALL(bk::launch_sample(c->psi.p, n, precision, c->samp_rand.as<double>() + off, shots,
                                          c->samp_out.as<int64_t>() + off, c->samp_sums.as<double>(),
                                          c->samp_prefix.as<double>(), c->stream));
            }
            ++pi;
        }
        if (step < total) {
            // the step's last sweep reduces the observables of state step+1
            run_slots.push_back((flags[static_cast<size_t>(step) + 1] & 1) ? ds + 3 * (mi++) : nullptr);
            run_gdt.push_back(cfg->g * cfg->dt);
        }
    }
    if ((rc = flush_run())) return cleanup_fail(rc);
    std::vector<double> hs(3 * static_cast<size_t>(n_meas));
    if (n_meas > 0)
        HIPCALL(hipMemcpyAsync(hs.data(), ds, hs.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (n_pts > 0)
        HIPCALL(hipMemcpyAsync(idx.data(), c->samp_out.p, idx.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                               c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n_meas; ++i) {
        pl.push_back(hs[3 * i]);
        dh.push_back(hs[3 * i + 1]);
    }
    if (cfg->verbose)
        for (size_t i = 0; i < tp.size(); ++i) {
            const int step = static_cast<int>(std::llround(tp[i] / cfg->dt));
            if (!cfg->record_trajectory || !(step % 200 == 0 || step == total)) continue;
            std::cout << "    [Progress " << std::setw(3)
                      << static_cast<int>(total > 0 ? 100.0 * step / total : 100.0) << "%] Step " << std::setw(4)
                      << step << "/" << total << " (t = " << std::fixed << std::setprecision(1) << tp[i] << "s)"
                      << " | P_<(t) = " << std::scientific << std::setprecision(3) << pl[i]
                      << " | <dH> = " << std::fixed << std::setprecision(2) << dh[i] << std::defaultfloat
                      << std::endl;
        }

    // per-shot statistics (sampler.cuh:575-597, 622-640)
    std::set<uint64_t> uniq;
    std::vector<bk_seed_stats> ss(n_seeds);
    for (int s = 0; s < n_seeds; ++s) {
        std::memset(&ss[s], 0, sizeof(bk_seed_stats));
        ss[s].seed = cfg->seed + static_cast<uint64_t>(s) * cfg->seed_stride;
        ss[s].best_energy = r->start_energy;
    }
    uint64_t total_flips = 0;
    std::vector<double> seed_flips(n_seeds, 0.0);
    for (int ptn = 0; ptn < n_pts; ++ptn) {
        for (int s = 0; s < n_seeds; ++s) {
            uint64_t flips_this = 0;
            for (int i = 0; i < shots; ++i) {
                const uint64_t b = static_cast<uint64_t>(idx[ptn * per_point + static_cast<size_t>(s) * shots + i]);
                uniq.insert(b);
                const int d = __builtin_popcountll(start ^ b);
                flips_this += d;
                if (s == 0) total_flips += d;
                const int64_t e = bk::host_energy(n, b);
                if (e < ss[s].best_energy) ss[s].best_energy = e;
                if (e < r->start_energy) {
                    ss[s].improved_shots++;
                    if (s == 0) r->improved_shots++;
                }
            }
            ss[s].total_shots_collected += shots;
            seed_flips[s] += static_cast<double>(flips_this);
        }
        r->total_shots_collected += shots;
    }
    r->total_evals = r->total_shots_collected;
    if (r->total_shots_collected > 0) {
        r->mean_hamming_flips = static_cast<double>(total_flips) / static_cast<double>(r->total_shots_collected);
        r->success_rate = static_cast<double>(r->improved_shots) / static_cast<double>(r->total_shots_collected);
        r->per_eval_rate = r->total_evals > 0
                               ? static_cast<double>(r->improved_shots) / static_cast<double>(r->total_evals)
                               : 0.0;
    }
    for (int s = 0; s < n_seeds; ++s) {
        ss[s].mean_hamming_flips = seed_flips[s];
        if (ss[s].total_shots_collected > 0) {
            ss[s].success_rate =
                static_cast<double>(ss[s].improved_shots) / static_cast<double>(ss[s].total_shots_collected);
            ss[s].per_eval_rate = ss[s].success_rate;
            ss[s].mean_hamming_flips /= static_cast<double>(ss[s].total_shots_collected);
        }
    }
    const double avg_end = (cfg->num_sample_points > 1 && cfg->dt_sample > 0.0)
                               ? cfg->t_start + (cfg->num_sample_points - 1) * cfg->dt_sample
                               : cfg->t_max;
    const bk::Derived dm = bk::derived_metrics(tp, pl, dh, cfg->t_start, avg_end);
    r->initial_rise_rate_avg = dm.rise_avg;
    r->initial_rise_rate_max = dm.rise_max;
    r->peak_p_less = dm.peak_p;
    r->peak_time = dm.peak_t;
    r->window_avg_dh = dm.window_dh;

    r->history_len = static_cast<int64_t>(tp.size());
    r->time_points = new double[tp.size() + 1];
    r->p_less_history = new double[tp.size() + 1];
    r->dh_history = new double[tp.size() + 1];
    std::copy(tp.begin(), tp.end(), r->time_points);
    std::copy(pl.begin(), pl.end(), r->p_less_history);
    std::copy(dh.begin(), dh.end(), r->dh_history);
    r->unique_basis_len = static_cast<int64_t>(uniq.size());
    r->unique_basis = new uint64_t[uniq.size() + 1];
    std::copy(uniq.begin(), uniq.end(), r->unique_basis);
    r->seed_stats_len = n_seeds;
    r->seed_stats = new bk_seed_stats[n_seeds];
    std::copy(ss.begin(), ss.end(), r->seed_stats);
    r->samples_len = static_cast<int64_t>(idx.size());
    r->samples = new int64_t[idx.size() + 1];
    std::copy(idx.begin(), idx.end(), r->samples);
    if (cfg->verbose) {
        std::cout << "[Quantum Sampler] Successes / Trials: " << r->improved_shots << " / "
                  << r->total_shots_collected << " (" << std::fixed << std::setprecision(2)
                  << (r->success_rate * 100.0) << "%)"
                  << " | E < E_og (E_og = " << r->start_energy << ")"
                  << " | Window <dH>: " << std::fixed << std::setprecision(2) << r->window_avg_dh
                  << " | Rise Rate (avg): " << std::scientific << std::setprecision(3) << r->initial_rise_rate_avg
                  << " | Evals: " << r->total_evals << " | Per-Eval Rate: " << std::scientific
                  << std::setprecision(4) << r->per_eval_rate << std::defaultfloat << "\n";
    }
    *out = guard.release();
    return BK_OK;
}

int bk_sampler_run(bk_ctx* c, const bk_sampler_config* cfg, uint64_t start_bits, bk_sampling_result** out)
{
    return bkc::guard_abi([&] { return bk_sampler_run_impl(c, cfg, start_bits, out); });
}

static int bk_set_tuning_impl(bk_ctx* c, const char* key, int value)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!key) return fail(BK_E_INVALID, "null key");
    const std::string k = key;
    HIPCALL(hipStreamSynchronize(c->stream));
    if ((rc = apply_tuning_key(c, c->tuning, k, value))) return rc;
    if (k == "pingpong" && !value) c->psi_alt.release();
    c->tuning_overrides[k] = value;
    return BK_OK;
}

int bk_set_tuning(bk_ctx* c, const char* key, int value)
{
    return bkc::guard_abi([&] { return bk_set_tuning_impl(c, key, value); });
}

static int bk_set_energy_mode_impl(bk_ctx* c, int mode)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (mode != BK_ENERGY_ONTHEFLY && mode != BK_ENERGY_CACHE16)
        return fail(BK_E_INVALID, "energy mode %d", mode);
    c->energy_mode = mode;
    if (mode == BK_ENERGY_ONTHEFLY) {
        HIPCALL(hipStreamSynchronize(c->stream));
        c->ecache.release();
        c->ecache_n = 0;
    }
    return BK_OK;
}

int bk_set_energy_mode(bk_ctx* c, int mode)
{
    return bkc::guard_abi([&] { return bk_set_energy_mode_impl(c, mode); });
}

static int bk_sweep_plan_impl(int n, int precision, int row_bits, int page_bits, int* out, int cap)
{
    if (n < 1 || n > BK_MAX_N || (precision != BK_C64 && precision != BK_C128) || (cap > 0 && !out))
        return fail(BK_E_INVALID, "bk_sweep_plan: bad argument");
    const int tb = precision == BK_C64 ? 13 : 12, htb = precision == BK_C64 ? 14 : 13;
    if (n < tb) return 0;  // small states use per-qubit kernels
    const std::vector<bk::HighSweep> pl = bk::plan_high_sweeps(n, tb, htb, std::max(row_bits, 3), page_bits);
    for (size_t i = 0; i < pl.size() && static_cast<int>(i) < cap; ++i) {
        out[5 * i + 0] = pl[i].h1;
        out[5 * i + 1] = pl[i].a1;
        out[5 * i + 2] = pl[i].h2;
        out[5 * i + 3] = pl[i].a2;
        out[5 * i + 4] = pl[i].R;
    }
    return static_cast<int>(pl.size());
}

int bk_sweep_plan(int n, int precision, int row_bits, int page_bits, int* out, int cap)
{
    return bkc::guard_abi([&] { return bk_sweep_plan_impl(n, precision, row_bits, page_bits, out, cap); });
}

static int bk_copy_probe_impl(bk_ctx* c, uint64_t bytes, int reps, int pattern, double* gbs)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (pattern < 0 || pattern > 2) return fail(BK_E_INVALID, "bk_copy_probe: pattern %d", pattern);
    const uint64_t chunk = bk::copy_pattern_chunk(pattern);
    if (bytes < chunk || bytes % chunk || reps < 1 || !gbs)
        return fail(BK_E_INVALID, "bk_copy_probe: bad argument (bytes must be a positive multiple of %llu)",
                    static_cast<unsigned long long>(chunk));
    ScopedBuf a, b;
    HIPCALL(a.ensure(bytes));
    HIPCALL(b.ensure(bytes));
    HIPCALL(hipMemsetAsync(a.p, 0, bytes, c->stream));
    HIPCALL(bk::launch_copy_pattern(pattern, a.p, b.p, bytes, c->stream));  // warm
    ScopedEvent e0, e1;
    HIPCALL(e0.create());
    HIPCALL(e1.create());
    HIPCALL(hipEventRecord(e0.e, c->stream));
    for (int r = 0; r < reps; ++r) HIPCALL(bk::launch_copy_pattern(pattern, a.p, b.p, bytes, c->stream));
    HIPCALL(hipEventRecord(e1.e, c->stream));
    HIPCALL(hipEventSynchronize(e1.e));
    float ms = 0.0f;
    HIPCALL(hipEventElapsedTime(&ms, e0.e, e1.e));
    *gbs = 2.0 * static_cast<double>(bytes) * reps / (ms * 1e-3) / 1e9;
    return BK_OK;
}

int bk_copy_probe(bk_ctx* c, uint64_t bytes, int reps, double* gbs)
{
    return bkc::guard_abi([&] { return bk_copy_probe_impl(c, bytes, reps, 0, gbs); });
}

int bk_copy_probe_pattern(bk_ctx* c, uint64_t bytes, int reps, int pattern, double* gbs)
{
    return bkc::guard_abi([&] { return bk_copy_probe_impl(c, bytes, reps, pattern, gbs); });
}

static int bk_timing_enable_impl(bk_ctx* c, int enable)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    HIPCALL(hipStreamSynchronize(c->stream));
    c->timing = enable != 0;
    c->timer.reset();
    c->tuning.timer = c->timing ? &c->timer : nullptr;
    return BK_OK;
}

int bk_timing_enable(bk_ctx* c, int enable)
{
    return bkc::guard_abi([&] { return bk_timing_enable_impl(c, enable); });
}

static int bk_timing_read_impl(bk_ctx* c, bk_kernel_timing* out, int cap, int* count)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!count || (cap > 0 && !out)) return fail(BK_E_INVALID, "bk_timing_read: bad argument");
    HIPCALL(hipStreamSynchronize(c->stream));
    if (c->timer.err != hipSuccess) {
        hipError_t e = c->timer.err;
        c->timer.reset();
        return fail(BK_E_HIP, "timing events: %s", hipGetErrorString(e));
    }
    std::vector<bk_kernel_timing> rows;
    for (const auto& r : c->timer.recs) {
        float ms = 0.0f;
        HIPCALL(hipEventElapsedTime(&ms, c->timer.pool[r.ev0], c->timer.pool[r.ev1]));
        auto it = std::find_if(rows.begin(), rows.end(),
                               [&](const bk_kernel_timing& k) { return r.name == k.name; });
        if (it == rows.end()) {
            bk_kernel_timing k{};
            std::snprintf(k.name, sizeof(k.name), "%s", r.name.c_str());
            k.bytes_per_launch = r.bytes;
            rows.push_back(k);
            it = rows.end() - 1;
        }
        it->launches += 1;
        it->total_ms += ms;
    }
    c->timer.reset();
    *count = static_cast<int>(rows.size());
    for (int i = 0; i < cap && i < static_cast<int>(rows.size()); ++i) out[i] = rows[i];
    return BK_OK;
}

int bk_timing_read(bk_ctx* c, bk_kernel_timing* out, int cap, int* count)
{
    return bkc::guard_abi([&] { return bk_timing_read_impl(c, out, cap, count); });
}

}  // extern "C"
