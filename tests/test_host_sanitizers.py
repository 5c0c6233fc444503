"""The engine's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only): the integer math, trap construction and tridiagonal eigensolver of
bernasconi-krylov_amd/csrc/bk_host.cpp, the schedule and derived metrics of
the sampler, the configuration parser (include/bk/config.hpp) on every
reference config kept under tests/golden/configs plus malformed input, and the
connector's and explorer's file writers (include/bk/wire.hpp), compiled with
g++ -fsanitize=address,undefined and run to completion with every report fatal.
The driver checks its own results too (energies against a direct pair sum,
eigenvalues against the trace and the determinant-free Sturm count), so the
run exercises real values, not only the code paths."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from conftest import ROOT

_DRIVER = r"""
#include "bk/config.hpp"
#include "bk/wire.hpp"
#include "bk_host.hpp"

#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#define CHECK(c) do { if (!(c)) { std::fprintf(stderr, "CHECK failed: %s (line %d)\n", #c, __LINE__); return 1; } } while (0)

// E(s) as the sum over pairs of the autocorrelations, spin by spin
static long long energy_direct(int n, unsigned long long s)
{
    long long e = 0;
    for (int k = 1; k < n; ++k) {
        long long c = 0;
        for (int i = 0; i + k < n; ++i) c += (((s >> i) & 1) ? -1 : 1) * (((s >> (i + k)) & 1) ? -1 : 1);
        e += c * c;
    }
    return e;
}

// eigenvalues of the symmetric tridiagonal matrix below x (Sturm sequence count)
static int count_below(const std::vector<double>& d, const std::vector<double>& e, double x)
{
    int cnt = 0;
    double q = 1.0;
    for (size_t i = 0; i < d.size(); ++i) {
        const double off = i ? e[i - 1] * e[i - 1] : 0.0;
        q = d[i] - x - (i ? off / (q == 0.0 ? 1e-300 : q) : 0.0);
        if (q < 0.0) ++cnt;
    }
    return cnt;
}

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    std::mt19937_64 rng(12345);
    // integer math over the whole range of n, the 64-bit edge included
    for (int n = 1; n <= 64; ++n) {
        for (int t = 0; t < 40; ++t) {
            const unsigned long long s = rng() & (n >= 64 ? ~0ull : ((1ull << n) - 1ull));
            CHECK(bk::host_energy(n, s) == energy_direct(n, s));
            const unsigned long long c = bk::host_canonical(n, s);
            CHECK(c <= s && bk::host_canonical(n, c) == c && bk::host_energy(n, c) == bk::host_energy(n, s));
        }
        if (n >= 2 && n <= 18) {
            const unsigned long long s = rng() & ((1ull << n) - 1ull);
            const int d = bk::host_d_min(n, s, -1, -1);
            CHECK(d >= 1 && d <= n);
            CHECK(bk::host_d_min(n, s, 0, 3) == 4);  // nothing lies below energy 0
        }
        if (n >= 3 && n <= 40) {
            uint64_t start = 0, s0 = 0;
            int64_t e0 = 0;
            bk::host_make_trap(n, 7 + n, &start, &s0, &e0);
            CHECK(e0 == bk::host_energy(n, s0));
            for (int p = 0; p < n; ++p) CHECK(bk::host_energy(n, s0 ^ (1ull << p)) >= e0);
        }
    }
    CHECK(bk::host_known_optimal(2) == -1 && bk::host_known_optimal(33) == -1 && bk::host_known_optimal(30) == 59);
    // tridiagonal eigensolver: ascending, the trace kept, each eigenvalue at a Sturm boundary
    for (int m : {1, 2, 3, 7, 30, 100}) {
        std::vector<double> d(m), e(m > 0 ? m - 1 : 0);
        std::uniform_real_distribution<double> u(-1.0, 1.0);
        double tr = 0.0;
        for (auto& x : d) tr += (x = u(rng));
        for (auto& x : e) x = u(rng);
        const std::vector<double> d0 = d, e0 = e;
        std::vector<double> z;
        CHECK(bk::tridiag_eigen(d, e, &z));
        CHECK(z.size() == static_cast<size_t>(m) * m);
        double sum = 0.0;
        for (int i = 0; i < m; ++i) {
            sum += d[i];
            if (i) CHECK(d[i] >= d[i - 1]);
            CHECK(count_below(d0, e0, d[i] - 1e-9) <= i && count_below(d0, e0, d[i] + 1e-9) >= i + 1);
        }
        CHECK(std::fabs(sum - tr) <= 1e-10 * m);
    }
    // schedules, including degenerate configurations
    for (double tmax : {0.0, 0.05, 1.0, 60.0}) {
        for (int pts : {0, 1, 20}) {
            bk::ScheduleIn si{};
            si.t_max = tmax;
            si.dt = 0.05;
            si.t_start = 0.5;
            si.dt_sample = pts > 1 ? 0.1 : 0.0;
            si.dt_trajectory = 0.0;
            si.num_sample_points = pts;
            si.shots_per_point = pts ? 10 : 0;
            si.track_p_less = true;
            si.record_trajectory = pts == 0;
            std::vector<unsigned char> fl;
            const int total = bk::build_schedule(si, fl);
            CHECK(total >= 0 && fl.size() == static_cast<size_t>(total) + 1);
        }
    }
    {
        const std::vector<double> none;
        bk::derived_metrics(none, none, none, 0.0, 1.0);
        const std::vector<double> t = {0.0, 0.05, 0.1}, p = {0.0, 1e-6, 3e-6}, dh = {0.0, 0.1, 0.2};
        const bk::Derived m = bk::derived_metrics(t, p, dh, 0.0, 0.1);
        CHECK(m.peak_p == 3e-6 && m.peak_t == 0.1);
    }
    // the configuration parser: the reference's files, then malformed text
    for (int i = 3; i < argc; ++i) {
        const BenchmarkConfig cfg = BenchmarkConfig::load_from_file(argv[i]);
        CHECK(!cfg.system.n_values.empty());
        const SamplerConfig sc = bk::config::to_sampler_config(cfg.quantum, cfg.system.n_values[0], 0.3, 42);
        CHECK(sc.dt > 0.0);
    }
    for (const char* text : {"", "=\n", "[\n", "[quantum\nn = \n", "n = 1,,2,\n", "g_values = x, 0.3\n",
                             "dt = 1e999\n", "t_max = -5\n", "k_sample_points = 99999999999999999999\n",
                             "unknown_key = 3\n; comment\n# comment\n", "[system]\nn_values = 16 20 24\n"}) {
        try {
            (void)BenchmarkConfig::load_from_string(text);
        } catch (const std::exception&) {
        }
    }
    // writers: empty records and a filled one
    {
        std::vector<bk::wire::TrapBenchmark> v(2);
        v[0].n = 16;
        v[0].g = 0.3;
        v[0].trap_state = SpinState(0xD2EE);
        bk::wire::set_canonical(v[0], 0xD2EE, 16);
        v[1] = v[0];
        v[1].time_points = {0.0, 0.05};
        v[1].p_less_history = {0.0, 1e-9};
        v[1].dh_history = {0.0, 0.01};
        v[1].rate_by_d[2] = 0.5;
        const std::string dir = argv[2];
        bk::wire::export_results_csv(dir + "/r.csv", v);
        bk::wire::export_time_series_csv(dir + "/t.csv", v);
        bk::wire::export_structured_json(dir + "/s.json", v, 0);
        bk::wire::ExhaustiveMinima x;
        x.n = 20;
        x.e_star = 26;
        x.levels.resize(2);
        x.levels[0].energy = 34;
        x.levels[0].orbits.push_back({0x52CC, 8, 0.0022});
        bk::wire::export_exhaustive_json(dir + "/x.json", x);
    }
    std::printf("%s ok\n", argv[1]);
    return 0;
}
"""


def test_host_code_under_asan_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "host_check.cpp"
    src.write_text(_DRIVER)
    exe = tmp_path / "host_check"
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=all", "-Wall", "-Wextra", "-Werror"]
    subprocess.run(["g++", *flags, "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "bernasconi-krylov_amd", "csrc"),
                    str(src), os.path.join(ROOT, "bernasconi-krylov_amd", "csrc", "bk_host.cpp"), "-o", str(exe)],
                   check=True)
    cfgdir = os.path.join(ROOT, "tests", "golden", "configs")
    cfgs = sorted(os.path.join(cfgdir, f) for f in os.listdir(cfgdir))
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "halt_on_error=1:detect_leaks=1:verify_asan_link_order=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    out = subprocess.run([str(exe), "host", str(tmp_path), *cfgs], capture_output=True, text=True, timeout=300,
                         env=env)
    assert out.returncode == 0, out.stdout + out.stderr[-4000:]
    assert "host ok" in out.stdout
    assert (tmp_path / "x.json").read_text().startswith('{\n  "N": 20,\n  "E_star": 26,')
