"""CPU tests of the C ABI surface: libbk.so loads, exports every function that
include/bk_capi.h declares, reports the absence of a device cleanly, and its
host-side integer math (energy, d_min, trap construction) agrees with the
oracle.  No GPU compute is called here."""
from __future__ import annotations

import ctypes as C
import os
import random
import re

import pytest

from conftest import ROOT


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "bk_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(bk_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_all_declared_symbols(bk):
    L = bk.lib()
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    # the Python mirror binds exactly the declared set
    assert sorted(bk.exported_symbols()) == declared


def test_no_device_path(bk):
    if bk.lib().bk_device_count() > 0:
        pytest.skip("a device is present")
    h = C.c_void_p()
    rc = bk.lib().bk_ctx_create(0, C.byref(h))
    assert rc == -5 and not h.value
    assert "no HIP device" in bk.lib().bk_last_error().decode()
    with pytest.raises(bk.BKError):
        bk.Engine(0)


def test_host_math_matches_oracle(bk, orc):
    rng = random.Random(17)
    for _ in range(2000):
        n = rng.randint(2, 32)
        s = rng.getrandbits(n)
        assert bk.energy(n, s) == orc.energy(n, s)
    for n in range(0, 34):
        assert bk.known_optimal_energy(n) == orc.known_optimal_energy(n)
    for _ in range(200):
        n = rng.randint(3, 18)
        s = rng.getrandbits(n)
        assert bk.d_min(n, s) == orc.d_min(n, s)
        md = rng.randint(1, n)
        assert bk.d_min(n, s, -1, md) == orc.d_min(n, s, -1, md)
    for seed in range(60):
        n = rng.choice([12, 16, 20, 24, 30, 32])
        assert bk.make_trap(n, seed) == orc.make_trap(n, seed)


def test_canonical_matches_oracle(bk, orc):
    """SpinState::canonical: the oracle's (pinned by the reference's fixtures)
    and the group's defining properties: idempotent, constant on an orbit."""
    rng = random.Random(23)
    for _ in range(3000):
        n = rng.randint(2, 64)
        s = rng.getrandbits(n)
        c = bk.canonical(n, s)
        assert c == orc.canonical(n, s)
        assert bk.canonical(n, c) == c and c <= s
        mask = (1 << n) - 1
        assert bk.canonical(n, ~s & mask) == c
        assert bk.canonical(n, int(format(s, f"0{n}b")[::-1], 2)) == c


def test_trap_with_dmin_filter(bk, orc):
    """Config-1 trap rule: iterate seeds until exact d_min >= 6 (N=24)."""
    seed, start, s0, e0, dm = bk.make_trap_min_dmin(24, 1, 6)
    assert dm >= 6 and orc.d_min(24, s0, e0) == dm
    assert orc.make_trap(24, seed) == (start, s0, e0)


def test_sampler_config_defaults(bk):
    cfg = bk._SamplerConfigC()
    bk.lib().bk_sampler_config_defaults(C.byref(cfg))
    py = bk.SamplerConfig()
    for f in ("n", "g", "dt", "t_max", "num_sample_points", "shots_per_point", "seed",
              "num_sampling_seeds", "seed_stride", "t_start", "dt_sample", "dt_trajectory"):
        assert getattr(cfg, f) == getattr(py, f), f
    assert bool(cfg.track_p_less) is py.track_p_less
    assert bool(cfg.record_trajectory) is py.record_trajectory
    # sampler.cuh:235-260 initialisers
    assert (py.n, py.g, py.dt, py.t_max, py.num_sample_points, py.shots_per_point) == \
        (16, 0.5, 0.05, 60.0, 20, 100)
    assert (py.seed, py.num_sampling_seeds, py.seed_stride) == (42, 3, 10007)


def test_factory_above_max_n_raises(bk):
    s = bk.create_sampler(33)
    with pytest.raises(RuntimeError):
        s.sample(0, bk.SamplerConfig(n=33))


@pytest.mark.parametrize("prec", [1, 2])
def test_sweep_plan_covers_every_qubit_once(bk, prec):
    """Host planner of the HIGH sweeps: every qubit in [TB, n) is rotated by
    exactly one sweep, rows stay below the active bits, every tile index stays
    inside the 2^n state."""
    tb, htb = (13, 14) if prec == 1 else (12, 13)
    for n in range(1, 33):
        for rb in (3, 4, 5, 6, 7):
            for pb in (0, 15, 17, 18, 19, 20, 24):
                plan = bk.sweep_plan(n, prec, rb, pb)
                if n <= tb:
                    assert plan == []
                    continue
                seen = []
                for h1, a1, h2, a2, R in plan:
                    A = a1 + a2
                    assert R + A == htb and R >= max(rb, 3) and a1 >= 1
                    assert R <= h1 and h1 + a1 <= h2 and h2 + a2 <= n
                    seen += list(range(h1, h1 + a1)) + list(range(h2, h2 + a2))
                    # tile index fields: [R,h1), [h1+a1,h2), [h2+a2,n) sum to n-htb bits
                    assert (h1 - R) + (h2 - h1 - a1) + (n - h2 - a2) == n - htb
                assert sorted(seen) == list(range(tb, n)), (n, rb, pb, plan)


_CXX_PROBE = r"""
#include "bk/sampler.hpp"
#include "bk/diagonalizer.hpp"
#include "bk/wire.hpp"
#include <cstdio>
int main() {
    // both Diagonalizer::solve overloads and the connector writers instantiate
    auto sub = static_cast<Eigensolution (Diagonalizer::*)(const std::vector<SpinState>&, double, int, DiagConfig)
                               const>(&Diagonalizer::solve);
    (void)sub;
    (void)&bk::wire::export_structured_json;
    SamplerConfig c;
    std::printf("cfg %d %g %g %g %d %d %llu\n", c.n, c.g, c.dt, c.t_max, c.num_sample_points,
                c.shots_per_point, (unsigned long long)c.seed);
    DiagConfig d;
    std::printf("diag %d %d %g\n", d.num_eigenpairs, d.max_krylov_dim, d.tolerance);
    auto big = create_sampler(33);
    try { big->sample(SpinState(0), c); return 1; }
    catch (const std::runtime_error& e) { std::printf("big ok\n"); }
    if (bk_device_count() == 0) {
        try { auto s = create_sampler(16); return 2; }
        catch (const std::runtime_error& e) { std::printf("nodev ok\n"); }
    }
    return 0;
}
"""


def test_cxx_dropin_headers_compile_and_link(bk, tmp_path):
    """include/bk/*.hpp compile as C++17 with -Wall -Werror, link against
    libbk.so and keep the reference's SamplerConfig / DiagConfig defaults."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "probe.cpp"
    src.write_text(_CXX_PROBE)
    exe = tmp_path / "probe"
    libdir = os.path.dirname(bk.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-Wno-unused-variable",
                    "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir, "-lbk",
                    "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    assert lines[0] == "cfg 16 0.5 0.05 60 20 100 42"
    assert lines[1] == "diag 1 100 1e-12"
    assert "big ok" in lines


def test_host_mt19937_64_and_field_schedule(bk, orc):
    """The package's mt19937_64 (used for seed-drawn fields) against the C++
    standard's fixed value (10000th output of the default seed) and the
    oracle's generator; the field schedule uses the portable (x>>11)*2^-53."""
    g = bk.MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042
    a, b = bk.MT19937_64(42), orc.MT64(42)
    for _ in range(700):  # crosses a state refill
        assert a() == b.next()
    f = bk.field_schedule(7, 64)
    c = bk.MT19937_64(7)
    ref = [0.35 + 0.05 * (2.0 * ((c() >> 11) * 2.0 ** -53) - 1.0) for _ in range(64)]
    assert f.tolist() == ref
    assert f.min() >= 0.30 and f.max() <= 0.40


def test_batch_workload_shape(bk):
    """Seed-drawn batch: four sizes chosen by the top two bits of mt19937_64
    draws, 16 evenly spaced fields, traps from consecutive seeds."""
    w = bk.batch_workload(64, seed=1)
    assert len(w) == 64 and {n for n, *_ in w} <= {16, 18, 20, 22}
    g = bk.MT19937_64(1)
    assert [n for n, *_ in w] == [(16, 18, 20, 22)[g() >> 62] for _ in range(64)]
    assert w[0][3] == pytest.approx(0.10) and w[15][3] == pytest.approx(0.60)
    assert w[16][3] == pytest.approx(0.10)
    n, s0, e0, _ = w[5]
    assert bk.make_trap(n, 1005)[1:] == (s0, e0)
