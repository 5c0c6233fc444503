// bk_wire.cpp -- C ABI of the benchmark connector's output files: C records in,
// bk::wire writers (include/bk/wire.hpp) out.  Host code only.
#include <cstdio>
#include <exception>
#include <vector>

#include "../../include/bk/wire.hpp"
#include "bk_ctx.hpp"

namespace {

int to_records(const bk_trap_record* recs, int count, std::vector<bk::wire::TrapBenchmark>* out)
{
    if (count < 0 || (count > 0 && !recs)) return bkc::fail(BK_E_INVALID, "bad record list");
    out->clear();
    out->reserve(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) {
        const bk_trap_record& r = recs[i];
        if (r.n < 1 || r.n > 64 || r.trap_rank < 0) return bkc::fail(BK_E_INVALID, "record %d: n=%d rank=%lld", i, r.n,
                                                                     static_cast<long long>(r.trap_rank));
        bk::wire::TrapBenchmark b;
        b.n = r.n;
        b.g = r.g;
        b.trap_rank = static_cast<size_t>(r.trap_rank);
        b.trap_state = SpinState(r.trap_bits);
        b.trap_energy = r.trap_energy;
        b.d_min = r.d_min;
        b.quantum_time_sec = r.quantum_time_sec;
        bk::wire::set_canonical(b, r.canonical_bits, r.n);
        if (r.result) bk::wire::fill_quantum(b, bk::to_sampling_result(*r.result));
        b.window_avg_p_less = bk::wire::window_avg_p_less(b, r.window_t_start, r.window_points, r.window_dt, r.t_max);
        out->push_back(std::move(b));
    }
    return BK_OK;
}

template <typename F>
int guarded(const char* path, F f)
{
    if (!path) return bkc::fail(BK_E_INVALID, "null path");
    try {
        f();
    } catch (const std::exception& e) {
        return bkc::fail(BK_E_INVALID, "%s", e.what());
    }
    return BK_OK;
}

}  // namespace

extern "C" {

static int bk_export_results_csv_impl(const char* path, const bk_trap_record* recs, int count)
{
    std::vector<bk::wire::TrapBenchmark> v;
    const int rc = to_records(recs, count, &v);
    if (rc) return rc;
    return guarded(path, [&] { bk::wire::export_results_csv(path, v); });
}

int bk_export_results_csv(const char* path, const bk_trap_record* recs, int count)
{
    return bkc::guard_abi([&] { return bk_export_results_csv_impl(path, recs, count); });
}

static int bk_export_time_series_csv_impl(const char* path, const bk_trap_record* recs, int count)
{
    std::vector<bk::wire::TrapBenchmark> v;
    const int rc = to_records(recs, count, &v);
    if (rc) return rc;
    return guarded(path, [&] { bk::wire::export_time_series_csv(path, v); });
}

int bk_export_time_series_csv(const char* path, const bk_trap_record* recs, int count)
{
    return bkc::guard_abi([&] { return bk_export_time_series_csv_impl(path, recs, count); });
}

static int bk_export_structured_json_impl(const char* path, const bk_trap_record* recs, int count, int n_override)
{
    std::vector<bk::wire::TrapBenchmark> v;
    const int rc = to_records(recs, count, &v);
    if (rc) return rc;
    return guarded(path, [&] { bk::wire::export_structured_json(path, v, n_override); });
}

int bk_export_structured_json(const char* path, const bk_trap_record* recs, int count, int n_override)
{
    return bkc::guard_abi([&] { return bk_export_structured_json_impl(path, recs, count, n_override); });
}

static int bk_export_exhaustive_json_impl(const char* path, const bk_exhaustive_result* res,
                                          const bk_minima_orbit* orbits, int64_t count)
{
    if (!res || count < 0 || (count > 0 && !orbits) || res->num_levels < 0 ||
        res->num_levels > BK_MINIMA_MAX_LEVELS || res->n < 1 || res->n > 64)
        return bkc::fail(BK_E_INVALID, "bk_export_exhaustive_json: bad argument");
    bk::wire::ExhaustiveMinima r;
    r.n = res->n;
    r.e_star = res->e_star;
    r.levels.resize(static_cast<size_t>(res->num_levels));
    for (int l = 0; l < res->num_levels; ++l) r.levels[static_cast<size_t>(l)].energy = res->level_energy[l];
    for (int64_t i = 0; i < count; ++i) {
        const bk_minima_orbit& o = orbits[i];
        if (o.level < 1 || o.level > res->num_levels)
            return bkc::fail(BK_E_INVALID, "bk_export_exhaustive_json: orbit %lld has level %d",
                             static_cast<long long>(i), o.level);
        r.levels[static_cast<size_t>(o.level - 1)].orbits.push_back({o.canonical_state, o.d_min, o.sd_hit_fraction});
    }
    return guarded(path, [&] { bk::wire::export_exhaustive_json(path, r); });
}

int bk_export_exhaustive_json(const char* path, const bk_exhaustive_result* res, const bk_minima_orbit* orbits,
                              int64_t count)
{
    return bkc::guard_abi([&] { return bk_export_exhaustive_json_impl(path, res, orbits, count); });
}

}  // extern "C"
