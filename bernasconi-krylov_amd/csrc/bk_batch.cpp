// bk_batch.cpp -- bk_batch_run: many traps of small n in shared launches.
//
// The reference evolves one trap per ISampler::sample call (sampler.cuh:412-707);
// at n = 16..22 a single state is 0.5-32 MB and one Trotter step of it is a few
// microseconds of HBM time, so one-trap-at-a-time runs are bound by launch and
// latency (measured 16-33 us per step).  Here all instances of one n share every
// launch: a sweep's tiles run over the states stored one after another, each
// state with its own error-feedback phase table, field and references.
//
// Order of work.  Without cache blocking every step visits all states (one
// feedback update, the high-qubit sweeps and the LOW sweep per (n, butterfly
// form) group).  With a cache block of B bytes, the states are cut into blocks
// of at most B bytes and each block runs all its steps before the next starts,
// so a block that fits the 256 MB last-level cache is re-read from there.
// Out of place, a block's fused passes alternate between its states and a
// block-sized second buffer: the part's cache serves a copy-shaped pass at
// close to twice the rate of an in-place update of the same lines, so blocks
// of a few tens of MB can run above the HBM rate.  Only the observables of the
// last step leave a run, so the final pass may end in either buffer.
// The whole sequence can be captured once as a hipGraph and launched once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/bk_capi.h"
#include "bk_ctx.hpp"
#include "bk_host.hpp"
#include "bk_kernels.hpp"
#include "bk_sweeps.hpp"

using namespace bkc;

namespace {

struct Group {
    int n = 0;
    bool tan = true;
    std::vector<int> idx;  // instance indices (input order within the group)
    size_t state_off = 0;  // amplitude offset in the state arena
    size_t tab_off = 0;    // entry offset in the table arenas
    size_t arr_off = 0;    // offset in the per-state arrays
    size_t part_off = 0;   // partial-triple offset
    int tab_count = 0;     // entries per table
    size_t master_off = 0; // double offset in the master arena
    std::vector<bk::HighSweep> fused;   // high groups of a fused run (empty: step by step)
    const uint32_t* ecache_hi = nullptr; // E(s) in the phase-carrying high group's layout
};

struct Tune {
    int grid, row_bits;
    int64_t block_bytes;
    bool oop;  // blocks of fused groups run out of place
};

// device pointer of a named arena buffer of at least 'bytes'
int arena(bk_ctx* c, const char* name, size_t bytes, void** out)
{
    DevBuf& b = c->batch_bufs[name];
    if (b.cap < bytes) {
        b.release();
        HIPCALL(b.ensure(std::max<size_t>(bytes, 256)));
    }
    *out = b.p;
    return BK_OK;
}

struct Plan {
    std::vector<Group> groups;
    int precision = 0;
    size_t esz = 0;
    char* states = nullptr;
    char* applied = nullptr;
    char* applied2 = nullptr;  // second table buffer: fused runs alternate per step
    double* carry = nullptr;
    double* master = nullptr;
    double *bc = nullptr, *bs = nullptr, *sd = nullptr, *re = nullptr, *im = nullptr;
    uint32_t *s0 = nullptr, *e0 = nullptr;
    double* partials = nullptr;
    double* results = nullptr;
    char* alt = nullptr;  // out-of-place blocks: the second buffer of a block
    std::map<int, const uint32_t*> ecache;
    size_t total_states_bytes = 0;
};

// states of group g per block of t
int block_states(const Plan& P, const Group& g, const Tune& t)
{
    const int64_t sb = static_cast<int64_t>(P.esz) << g.n;
    const int ns = static_cast<int>(g.idx.size());
    return t.block_bytes <= 0 ? ns : static_cast<int>(std::min<int64_t>(ns, std::max<int64_t>(1, t.block_bytes / sb)));
}

// the second buffer of out-of-place blocks (one block of the largest fused group)
int ensure_alt(bk_ctx* c, Plan& P, const Tune& t)
{
    if (!t.oop || t.block_bytes <= 0) return BK_OK;
    size_t need = 0;
    for (const Group& g : P.groups)
        if (!g.fused.empty())
            need = std::max(need, static_cast<size_t>(block_states(P, g, t)) * (P.esz << g.n));
    if (need == 0) return BK_OK;
    void* v = nullptr;
    int rc = arena(c, "alt", need, &v);
    if (rc) return rc;
    P.alt = static_cast<char*>(v);
    return BK_OK;
}

// enqueue one run of 'steps' steps (starting states prepared by init()) with
// tuning t; counts the launches
int enqueue_run(bk_ctx* c, const Plan& P, int steps, const Tune& t, int64_t* launches)
{
    bk::SweepTuning tun = bk::default_tuning(20, P.precision);
    tun.grid = t.grid;
    tun.row_bits = t.row_bits;
    tun.timer = nullptr;
    int64_t nl = 0;
    auto group_block = [&](const Group& g, int first, int count) {
        bk::BatchGroup b{};
        b.n = g.n;
        b.precision = P.precision;
        b.nstates = count;
        b.psi = P.states + (g.state_off + (static_cast<size_t>(first) << g.n)) * P.esz;
        b.tab = P.applied + (g.tab_off + static_cast<size_t>(first) * g.tab_count) * P.esz;
        b.tab_stride = g.tab_count;
        b.ecache = P.ecache.at(g.n);
        b.bc = P.bc + g.arr_off + first;
        b.bs = P.bs + g.arr_off + first;
        b.s0 = P.s0 + g.arr_off + first;
        b.e0 = P.e0 + g.arr_off + first;
        b.tan_form = g.tan;
        return b;
    };
    const int tb = bk::low_tile_bits(P.precision);
    const int htb = tb + 1;  // high-pass tile bits (2^14 c64, 2^13 c128)
    // the tables of a block of states in table buffer 0 or 1
    auto tabs = [&](const Group& g, int first, int buf) {
        return (buf ? P.applied2 : P.applied) + (g.tab_off + static_cast<size_t>(first) * g.tab_count) * P.esz;
    };
    auto feedback = [&](const Group& g, int first, int count, int buf = 0) -> int {
        HIPCALL(bk::launch_batch_feedback(P.master + g.master_off,
                                          P.carry + 2 * (g.tab_off + static_cast<size_t>(first) * g.tab_count),
                                          tabs(g, first, buf), g.tab_count, count, P.sd + g.arr_off + first,
                                          P.precision, c->stream));
        ++nl;
        return BK_OK;
    };
    // tiles per state of the pass that measures the last step (the partials' stride)
    auto final_tiles = [&](const Group& g) -> int {
        const std::vector<bk::HighSweep>& fg = g.fused;
        if (fg.empty()) return 1 << (g.n - tb);
        const int k = steps - 1;
        const bool low = fg.size() == 2 ? (k & 1) != 0 : (k & 1) == 0;
        return 1 << (g.n - (low ? tb : htb));
    };
    // one step at a time (classic sweeps), or the fused schedule of launch_run
    auto run_block = [&](const Group& g, int first, int count, int nsteps) -> int {
        int rc = BK_OK;
        bk::BatchGroup b = group_block(g, first, count);
        const std::vector<bk::HighSweep>& fg = g.fused;
        // out of place: each pass writes the other buffer, then the two swap
        const bool oop = t.oop && t.block_bytes > 0 && !fg.empty() && P.alt;
        auto pass = [&](const bk::BatchPass& bp, double* part) -> int {
            if (oop) b.psi_out = b.psi == P.alt ? group_block(g, first, count).psi : P.alt;
            HIPCALL(bk::launch_batch_pass(b, bp, tun, part, c->stream));
            if (oop) b.psi = b.psi_out;
            ++nl;
            return BK_OK;
        };
        double* part_last = P.partials + 3 * (g.part_off + static_cast<size_t>(first) * final_tiles(g));
        if (fg.empty()) {
            for (int k = 0; k < nsteps; ++k) {
                if ((rc = feedback(g, first, count))) return rc;
                HIPCALL(bk::launch_batch_step(b, tun, k + 1 == nsteps ? part_last : nullptr, c->stream));
                nl += 1 + static_cast<int64_t>(
                              bk::plan_high_sweeps(g.n, tb, htb, std::max(tun.row_bits, 3), tun.page_bits).size());
            }
            return BK_OK;
        }
        const bool two = fg.size() == 2;
        bk::BatchPass open{};
        open.kind = two ? bk::BatchPass::kLowPlain : bk::BatchPass::kHighPlain;
        open.hs = fg.back();
        if ((rc = pass(open, nullptr))) return rc;
        for (int k = 0; k < nsteps; ++k) {
            if (two) {
                bk::BatchPass pa{};
                pa.kind = bk::BatchPass::kHighPlain;
                pa.hs = fg[0];
                if ((rc = pass(pa, nullptr))) return rc;
            }
            // step k's tables: buffer k & 1, set by step k-1's phase pass (the
            // first by its own kernel)
            if (k == 0 && (rc = feedback(g, first, count, 0))) return rc;
            const bool low = two ? (k & 1) != 0 : (k & 1) == 0;
            const bool last = k + 1 == nsteps;
            bk::BatchPass pp{};
            pp.kind = low ? bk::BatchPass::kLowPhase : bk::BatchPass::kHighPhase;
            pp.hs = fg.back();
            pp.ecache_high = g.ecache_hi;
            pp.measure = last;
            pp.post = !last;
            b.tab = tabs(g, first, k & 1);
            b.tab_next = last ? nullptr : tabs(g, first, (k + 1) & 1);
            b.fb_master = P.master + g.master_off;
            b.fb_carry = P.carry + 2 * (g.tab_off + static_cast<size_t>(first) * g.tab_count);
            b.fb_sd = P.sd + g.arr_off + first;
            if ((rc = pass(pp, last ? part_last : nullptr))) return rc;
            b.tab_next = nullptr;
        }
        return BK_OK;
    };
    int rc = BK_OK;
    if (t.block_bytes <= 0) {
        // all states advance together: one pass sequence per group, interleaved by group
        if (std::all_of(P.groups.begin(), P.groups.end(), [&](const Group& g) { return g.fused.empty(); })) {
            for (int k = 0; k < steps; ++k)
                for (const Group& g : P.groups) {
                    const bk::BatchGroup b = group_block(g, 0, static_cast<int>(g.idx.size()));
                    if ((rc = feedback(g, 0, static_cast<int>(g.idx.size())))) return rc;
                    HIPCALL(bk::launch_batch_step(
                        b, tun, k + 1 == steps ? P.partials + 3 * g.part_off : nullptr, c->stream));
                    nl += 1 + static_cast<int64_t>(
                                  bk::plan_high_sweeps(g.n, tb, htb, std::max(tun.row_bits, 3), tun.page_bits).size());
                }
        } else {
            for (const Group& g : P.groups)
                if ((rc = run_block(g, 0, static_cast<int>(g.idx.size()), steps))) return rc;
        }
    } else {
        for (const Group& g : P.groups) {
            const int per = block_states(P, g, t);
            const int ns = static_cast<int>(g.idx.size());
            for (int first = 0; first < ns; first += per)
                if ((rc = run_block(g, first, std::min(per, ns - first), steps))) return rc;
        }
    }
    for (const Group& g : P.groups) {
        HIPCALL(bk::launch_batch_finish(P.partials + 3 * g.part_off, final_tiles(g), static_cast<int>(g.idx.size()),
                                        P.results + 3 * g.arr_off, c->stream));
        ++nl;
    }
    if (launches) *launches = nl;
    return BK_OK;
}

// starting states: zero, the opening-phase amplitude at s0, carries reset
int init_states(bk_ctx* c, const Plan& P)
{
    HIPCALL(hipMemsetAsync(P.states, 0, P.total_states_bytes, c->stream));
    for (const Group& g : P.groups)
        HIPCALL(bk::launch_batch_init(P.states + g.state_off * P.esz, g.n, static_cast<int>(g.idx.size()), P.precision,
                                      P.s0 + g.arr_off, P.re + g.arr_off, P.im + g.arr_off, c->stream));
    size_t ncarry = 0;
    for (const Group& g : P.groups) ncarry += g.idx.size() * static_cast<size_t>(g.tab_count);
    HIPCALL(bk::launch_carry_reset(P.carry, static_cast<int>(ncarry), c->stream));
    return BK_OK;
}

std::string shape_key(const bk_batch_params* p, const std::vector<Group>& groups)
{
    std::string k = "p" + std::to_string(p->precision);
    for (const Group& g : groups) k += ":" + std::to_string(g.n) + (g.tan ? "t" : "c") + std::to_string(g.idx.size());
    return k;
}

}  // namespace

extern "C" {

static int bk_batch_run_impl(bk_ctx* c, const bk_instance* inst, int count, const bk_batch_params* p,
                             bk_batch_result* out, bk_batch_stats* stats)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!inst || !p || !out || count < 1) return fail(BK_E_INVALID, "bk_batch_run: bad argument");
    if (p->precision != BK_C64 && p->precision != BK_C128) return fail(BK_E_INVALID, "precision=%d", p->precision);
    if (p->order != BK_LIE && p->order != BK_STRANG) return fail(BK_E_INVALID, "order=%d", p->order);
    if (p->steps < 1 || !is_fin(p->dt)) return fail(BK_E_INVALID, "steps=%d dt=%g", p->steps, p->dt);
    if (p->block_bytes < 0) return fail(BK_E_INVALID, "block_bytes=%lld", static_cast<long long>(p->block_bytes));
    const int prec = p->precision;
    const int tb = bk::low_tile_bits(prec);
    for (int i = 0; i < count; ++i)
        if (inst[i].n < tb || inst[i].n > BK_MAX_N || !is_fin(inst[i].g))
            return fail(BK_E_INVALID, "instance %d: n=%d outside [%d, %d] or bad g", i, inst[i].n, tb, BK_MAX_N);

    // ---- groups by (n, butterfly form) ----
    std::map<std::pair<int, int>, Group> gm;
    std::vector<bk::MixInfo> mix(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) {
        mix[static_cast<size_t>(i)] = bk::mix_info(inst[i].n, inst[i].g * p->dt, prec);
        const bool tan = mix[static_cast<size_t>(i)].tan_form;
        Group& g = gm[{inst[i].n, tan ? 1 : 0}];
        g.n = inst[i].n;
        g.tan = tan;
        g.idx.push_back(i);
    }
    Plan P;
    P.precision = prec;
    P.esz = prec == BK_C64 ? 8 : 16;
    size_t st_off = 0, tab_off = 0, arr_off = 0, part_off = 0, master_off = 0;
    std::map<int, size_t> master_at;
    std::vector<double> master_h;
    for (auto& kv : gm) {
        Group g = kv.second;
        g.state_off = st_off;
        g.arr_off = arr_off;
        g.tab_off = tab_off;
        g.part_off = part_off;
        auto mit = master_at.find(g.n);
        if (mit == master_at.end()) {
            const std::vector<double> m = phase_master(g.n, p->dt);
            master_at[g.n] = master_off;
            master_h.insert(master_h.end(), m.begin(), m.end());
            master_off += m.size();
            mit = master_at.find(g.n);
        }
        g.master_off = mit->second;
        g.tab_count = static_cast<int>(phase_master(g.n, p->dt).size() / 2);
        const size_t ns = g.idx.size();
        st_off += ns << g.n;
        arr_off += ns;
        tab_off += ns * static_cast<size_t>(g.tab_count);
        part_off += ns << (g.n - tb);
        P.groups.push_back(g);
    }
    P.total_states_bytes = st_off * P.esz;

    // ---- arenas ----
    void* v = nullptr;
    if ((rc = arena(c, "states", P.total_states_bytes, &v))) return rc;
    P.states = static_cast<char*>(v);
    if ((rc = arena(c, "applied", 2 * tab_off * P.esz, &v))) return rc;
    P.applied = static_cast<char*>(v);
    P.applied2 = P.applied + tab_off * P.esz;
    if ((rc = arena(c, "carry", tab_off * 16, &v))) return rc;
    P.carry = static_cast<double*>(v);
    if ((rc = arena(c, "master", master_h.size() * 8, &v))) return rc;
    P.master = static_cast<double*>(v);
    if ((rc = arena(c, "dparams", arr_off * 5 * 8, &v))) return rc;
    P.bc = static_cast<double*>(v);
    P.bs = P.bc + arr_off;
    P.sd = P.bs + arr_off;
    P.re = P.sd + arr_off;
    P.im = P.re + arr_off;
    if ((rc = arena(c, "uparams", arr_off * 2 * 4, &v))) return rc;
    P.s0 = static_cast<uint32_t*>(v);
    P.e0 = P.s0 + arr_off;
    if ((rc = arena(c, "partials", part_off * 3 * 8, &v))) return rc;
    P.partials = static_cast<double*>(v);
    if ((rc = arena(c, "results", arr_off * 3 * 8, &v))) return rc;
    P.results = static_cast<double*>(v);
    for (Group& g : P.groups) {
        DevBuf& ec = c->batch_ecache[{g.n, prec}];
        if (!ec.p) {
            HIPCALL(ec.ensure(bk::energy_cache_bytes(g.n)));
            HIPCALL(bk::launch_energy_fill(ec.as<uint32_t>(), g.n, prec, c->stream));
        }
        P.ecache[g.n] = ec.as<uint32_t>();
        // fused schedule (default tuning: the tuner's row bits steer step-by-step sweeps only)
        std::vector<bk::HighSweep> fg;
        if (g.tan && c->tuning.fused) {
            // batched passes run every high group on 2^HTB-amplitude tiles
            bk::SweepTuning bt = bk::default_tuning(g.n, prec);
            bt.narrow_phase = false;
            fg = bk::fused_sweeps(g.n, prec, bt);
        }
        g.fused = fg;
        if (!fg.empty()) {
            const bk::HighSweep& h = fg.back();
            DevBuf& eh = c->batch_ecache_hi[{g.n, prec, h.h1, h.a1, h.h2, h.a2, h.R}];
            if (!eh.p) {
                HIPCALL(eh.ensure(bk::energy_cache_bytes(g.n)));
                HIPCALL(bk::launch_energy_fill_high(eh.as<uint32_t>(), g.n, prec, h, c->stream));
            }
            g.ecache_hi = eh.as<uint32_t>();
        }
    }

    // ---- per-state parameters ----
    const double phi0 = p->order == BK_STRANG ? 0.5 * p->dt : p->dt;
    std::vector<double> hd(arr_off * 5);
    std::vector<uint32_t> hu(arr_off * 2);
    for (const Group& g : P.groups)
        for (size_t j = 0; j < g.idx.size(); ++j) {
            const bk_instance& in = inst[g.idx[j]];
            const bk::MixInfo& m = mix[static_cast<size_t>(g.idx[j])];
            const size_t a = g.arr_off + j;
            const uint64_t s0 = in.start_bits & ((1ull << in.n) - 1ull);
            const int64_t es0 = bk::host_energy(in.n, s0);
            const int64_t e0 = in.e0 >= 0 ? in.e0 : es0;
            const double th = -(phi0 / static_cast<double>(in.n)) * static_cast<double>(es0);
            hd[a] = m.bc;
            hd[arr_off + a] = m.bs;
            hd[2 * arr_off + a] = m.sd;
            hd[3 * arr_off + a] = std::cos(th);
            hd[4 * arr_off + a] = std::sin(th);
            hu[a] = static_cast<uint32_t>(s0);
            hu[arr_off + a] = static_cast<uint32_t>(std::max<int64_t>(0, std::min<int64_t>(e0, 0xFFFFFFFFll)));
        }
    HIPCALL(hipMemcpyAsync(P.master, master_h.data(), master_h.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCALL(hipMemcpyAsync(P.bc, hd.data(), hd.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCALL(hipMemcpyAsync(P.s0, hu.data(), hu.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));  // host vectors go out of scope below

    // ---- tuning ----
    Tune t{4096, prec == BK_C64 ? 5 : 4, 0, false};
    const std::string key = shape_key(p, P.groups);
    bool tuned = false;
    auto tit = c->batch_tuned.find(key);
    if (!p->autotune) {
        t.block_bytes = p->block_bytes;
        t.oop = p->block_bytes > 0 && p->out_of_place != 0;
    } else if (tit != c->batch_tuned.end()) {
        t = {tit->second.grid, tit->second.row_bits, tit->second.block_bytes, tit->second.out_of_place};
        tuned = true;
    } else {
        ScopedEvent e0, e1;
        HIPCALL(e0.create());
        HIPCALL(e1.create());
        const int trial_steps = std::min(p->steps, 8);
        auto time_of = [&](const Tune& tt, float* ms) -> int {
            int r = ensure_alt(c, P, tt);
            if (r) return r;
            if ((r = init_states(c, P))) return r;
            HIPCALL(hipEventRecord(e0.e, c->stream));
            if ((r = enqueue_run(c, P, trial_steps, tt, nullptr))) return r;
            HIPCALL(hipEventRecord(e1.e, c->stream));
            HIPCALL(hipEventSynchronize(e1.e));
            HIPCALL(hipEventElapsedTime(ms, e0.e, e1.e));
            return BK_OK;
        };
        float best = 0.f;
        if ((rc = time_of(t, &best))) return rc;
        // coordinate descent: grid, then row bits, then cache block
        for (int gsz : {1024, 2048, 8192, 16384}) {
            Tune cand = t;
            cand.grid = gsz;
            float ms = 0.f;
            if ((rc = time_of(cand, &ms))) return rc;
            if (ms < best) best = ms, t = cand;
        }
        for (int rb : {3, 4, 5, 6}) {
            if (rb == t.row_bits) continue;
            Tune cand = t;
            cand.row_bits = rb;
            float ms = 0.f;
            if ((rc = time_of(cand, &ms))) return rc;
            if (ms < best) best = ms, t = cand;
        }
        const Tune base = t;
        for (int64_t bb : {int64_t(256) << 20, int64_t(128) << 20, int64_t(48) << 20, int64_t(8) << 20}) {
            if (bb >= static_cast<int64_t>(P.total_states_bytes)) continue;
            Tune cand = base;
            cand.block_bytes = bb;
            float ms = 0.f;
            if ((rc = time_of(cand, &ms))) return rc;
            if (ms < best) best = ms, t = cand;
        }
        // out of place: the block and its second buffer together within the cache
        for (int64_t bb : {int64_t(128) << 20, int64_t(64) << 20, int64_t(32) << 20, int64_t(16) << 20}) {
            if (bb >= static_cast<int64_t>(P.total_states_bytes)) continue;
            Tune cand = base;
            cand.block_bytes = bb;
            cand.oop = true;
            float ms = 0.f;
            if ((rc = time_of(cand, &ms))) return rc;
            if (ms < best) best = ms, t = cand;
        }
        c->batch_tuned[key] = {t.grid, t.row_bits, t.block_bytes, t.oop};
        tuned = true;
    }

    // ---- run ----
    if ((rc = ensure_alt(c, P, t))) return rc;
    if ((rc = init_states(c, P))) return rc;
    ScopedEvent ev0, ev1;
    HIPCALL(ev0.create());
    HIPCALL(ev1.create());
    int64_t launches = 0;
    if (p->use_graph) {
        struct GraphGuard {  // graph and executable destroyed on every return path, after the stream drains
            hipStream_t stream;
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            ~GraphGuard()
            {
                if (exec) {
                    (void)hipStreamSynchronize(stream);
                    (void)hipGraphExecDestroy(exec);
                }
                if (graph) (void)hipGraphDestroy(graph);
            }
        } gg{c->stream};
        HIPCALL(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        rc = enqueue_run(c, P, p->steps, t, &launches);
        const hipError_t ce = hipStreamEndCapture(c->stream, &gg.graph);
        if (rc) return rc;
        HIPCALL(ce);
        size_t nodes = 0;
        HIPCALL(hipGraphGetNodes(gg.graph, nullptr, &nodes));
        launches = static_cast<int64_t>(nodes);
        const hipError_t ie = hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            gg.exec = nullptr;
            return fail(BK_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
        }
        HIPCALL(hipEventRecord(ev0.e, c->stream));
        HIPCALL(hipGraphLaunch(gg.exec, c->stream));
        HIPCALL(hipEventRecord(ev1.e, c->stream));
        HIPCALL(hipStreamSynchronize(c->stream));
    } else {
        HIPCALL(hipEventRecord(ev0.e, c->stream));
        if ((rc = enqueue_run(c, P, p->steps, t, &launches))) return rc;
        HIPCALL(hipEventRecord(ev1.e, c->stream));
        HIPCALL(hipStreamSynchronize(c->stream));
    }
    float ms = 0.f;
    HIPCALL(hipEventElapsedTime(&ms, ev0.e, ev1.e));
    std::vector<double> res(arr_off * 3);
    HIPCALL(hipMemcpy(res.data(), P.results, res.size() * 8, hipMemcpyDeviceToHost));
    for (const Group& g : P.groups)
        for (size_t j = 0; j < g.idx.size(); ++j) {
            const size_t a = g.arr_off + j;
            out[g.idx[j]] = {res[3 * a], res[3 * a + 1], res[3 * a + 2]};
        }
    if (stats) {
        stats->device_ms = ms;
        stats->grid = t.grid;
        stats->row_bits = t.row_bits;
        stats->block_bytes = t.block_bytes;
        stats->launches = launches;
        stats->tuned = tuned ? 1 : 0;
        stats->out_of_place = t.oop ? 1 : 0;
    }
    return BK_OK;
}

int bk_batch_run(bk_ctx* c, const bk_instance* inst, int count, const bk_batch_params* p, bk_batch_result* out,
                 bk_batch_stats* stats)
{
    return bkc::guard_abi([&] { return bk_batch_run_impl(c, inst, count, p, out, stats); });
}

}  // extern "C"
