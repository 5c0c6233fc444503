"""The benchmark configuration front-end (include/bk/config.hpp and its Python
twin bernasconi-krylov_amd/config.py) against the expectations of the
reference's tests/test_config.cpp and its configuration files (the four kept
in tests/golden/configs), with the C++ header and the Python parser required to
agree field by field on every file.  Host code only."""
from __future__ import annotations

import dataclasses
import importlib
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONFIGS = os.path.join(ROOT, "tests", "golden", "configs")


@pytest.fixture(scope="module")
def cfgmod(bk):
    return importlib.import_module("bernasconi-krylov_amd.config")


def test_defaults(cfgmod):
    """test_config.cpp test_default_config."""
    c = cfgmod.BenchmarkConfig()
    assert c.system.n_values == [16, 20, 24] and c.system.seed_base == 42 and c.system.device == "auto"
    assert (c.traps.pool_minima_count, c.traps.benchmark_trap_count, c.traps.minima_cache_dir) == (100, 5, "seeds")
    assert not c.traps.force_regenerate_traps
    q = c.quantum
    assert q.g_values == [0.25, 0.5, 0.75] and (q.dt, q.t_start, q.dt_sample) == (0.05, 10.0, 1.5)
    assert (q.k_sample_points, q.m_shots_per_point) == (20, 100) and q.track_p_less and not q.record_trajectory
    assert q.dt_trajectory == 0.0 and q.run_classical
    k = c.classical
    assert (k.kick_trials_per_d, k.kick_d_values, k.restart_trials, k.max_evals_per_start, k.batch_size) == \
        (50000, [1, 2, 3, 4, 5, 6], 50000, 0, 10000)
    o = c.output
    assert (o.output_dir, o.results_csv, o.time_series_csv, o.summary_csv, o.verbose) == \
        ("data_raw", "benchmark_results_n{N}.csv", "p_less_dynamics_n{N}.csv", "benchmark_summary.csv", True)


def test_derived_quantum_helpers(cfgmod):
    q = cfgmod.QuantumConfig(t_start=10.0, dt_sample=1.5, k_sample_points=20, dt=0.05)
    assert q.t_max() == pytest.approx(38.5)
    t = q.sample_times()
    assert len(t) == 20 and t[0] == pytest.approx(10.0) and t[1] == pytest.approx(11.5) and t[19] == pytest.approx(38.5)
    assert q.trotter_steps_per_sample() == 30


def test_path_resolution(cfgmod):
    o = cfgmod.OutputConfig(output_dir="data_raw", results_csv="results_n{N}.csv",
                            time_series_csv="dynamics_n{n}.csv", summary_csv="summary.csv")
    assert o.get_results_path(24) == "data_raw/results_n24.csv"
    assert o.get_time_series_path(24) == "data_raw/dynamics_n24.csv"
    assert o.get_summary_path() == "data_raw/summary.csv"
    o.results_csv = "/tmp/results_n{N}.csv"
    assert o.get_results_path(30) == "/tmp/results_n30.csv"


INI = """
# Benchmark experiment definition
[system]
n_values = 20, 24, 28 ; System sizes
seed_base = 12345
device = gpu

[traps]
pool_minima_count = 200
benchmark_trap_count = 10
minima_cache_dir = data_raw/custom_minima
force_regenerate_traps = true

[quantum]
g_values = 0.35, 0.50
dt = 0.02
t_start = 12.0
dt_sample = 2.0
k_sample_points = 15
m_shots_per_point = 250
track_p_less = true

[classical]
kick_trials_per_d = 100000
kick_d_values = 2, 4, 6, 8
restart_trials = 80000
max_evals_per_start = 500
batch_size = 20000

[output]
output_dir = data_raw/runs
results_csv = run_n{N}.csv
time_series_csv = p_n{N}.csv
summary_csv = all_summary.csv
verbose = false
"""


def test_ini_string_parsing(cfgmod):
    """test_config.cpp test_ini_string_parsing (its own input text)."""
    c = cfgmod.load_from_string(INI)
    assert c.system.n_values == [20, 24, 28] and c.system.seed_base == 12345 and c.system.device == "gpu"
    assert (c.traps.pool_minima_count, c.traps.benchmark_trap_count) == (200, 10)
    assert c.traps.minima_cache_dir == "data_raw/custom_minima" and c.traps.force_regenerate_traps
    assert c.quantum.g_values == [0.35, 0.5] and c.quantum.dt == 0.02 and c.quantum.k_sample_points == 15
    assert c.quantum.m_shots_per_point == 250 and c.quantum.t_start == 12.0 and c.quantum.dt_sample == 2.0
    assert c.classical.kick_trials_per_d == 100000 and c.classical.kick_d_values == [2, 4, 6, 8]
    assert c.classical.restart_trials == 80000 and c.classical.max_evals_per_start == 500
    assert c.classical.batch_size == 20000
    assert c.output.get_results_path(24) == "data_raw/runs/run_n24.csv" and not c.output.verbose
    assert not c.quantum.verbose


def test_reference_config_files(cfgmod):
    """test_config.cpp test_legacy_config_backward_compatibility, plus the
    trajectory scan and an .ini of the reference's configs/ directory."""
    c = cfgmod.load_from_file(os.path.join(CONFIGS, "n24_delayed_window.cfg"))
    assert c.system.n_values == [24] and c.traps.pool_minima_count == 200 and c.traps.benchmark_trap_count == 10
    assert len(c.quantum.g_values) == 5 and (c.quantum.t_start, c.quantum.dt_sample, c.quantum.dt) == (10.0, 1.5, 0.05)
    assert (c.quantum.k_sample_points, c.quantum.m_shots_per_point) == (20, 100)
    assert len(c.classical.kick_d_values) == 6 and c.classical.kick_trials_per_d == 50000
    assert c.classical.restart_trials == 50000 and c.classical.max_evals_per_start == 2000
    c = cfgmod.load_from_file(os.path.join(CONFIGS, "default_benchmark.cfg"))
    assert c.system.n_values == [16, 20, 24] and c.traps.pool_minima_count == 100
    assert c.traps.benchmark_trap_count == 5 and c.quantum.g_values == [0.5]
    c = cfgmod.load_from_file(os.path.join(CONFIGS, "probability_scan_n30.cfg"))
    assert c.system.n_values == [30] and c.quantum.g_values == [0.3, 0.35, 0.45]
    assert c.quantum.record_trajectory and not c.quantum.run_classical and c.quantum.m_shots_per_point == 0
    assert c.quantum.t_max() == 60.0 and len(c.quantum.sample_times()) == 1201
    assert c.output.get_time_series_path(30) == "data_processed/probability_dynamics_n30.csv"
    s = cfgmod.to_sampler_config(c.quantum, 30, 0.35, c.system.seed_base)
    assert (s.n, s.g, s.dt, s.t_max, s.record_trajectory, s.shots_per_point) == (30, 0.35, 0.05, 60.0, True, 0)
    c = cfgmod.load_from_file(os.path.join(CONFIGS, "n16.ini"))
    assert c.system.n_values == [16] and c.traps.benchmark_trap_count == 0 and c.classical.tabu_kick_d_values[-1] == 12
    assert c.output.get_structured_json_path(16) == "data_processed/structured_benchmark_n16.json"


def test_trajectory_flags_and_warnings(cfgmod, capsys):
    c = cfgmod.load_from_string("[quantum]\nrecord_trajectory = true\ndt_trajectory = 0.1\nt_max = 60.0\n"
                                "run_classical = false\nno_such_key = 3\nthis line has no delimiter\n")
    assert c.quantum.record_trajectory and c.quantum.dt_trajectory == 0.1 and c.quantum.t_max() == 60.0
    assert not c.quantum.run_classical and len(c.quantum.sample_times()) == 601
    err = capsys.readouterr().err
    assert "Unknown parameter on line 6: 'no_such_key'" in err and "Malformed line 7" in err
    with pytest.raises(ValueError):
        cfgmod.load_from_string("track_p_less = maybe")
    with pytest.raises(RuntimeError):
        cfgmod.load_from_file(os.path.join(CONFIGS, "missing.cfg"))


_CXX = r"""
#include "bk/config.hpp"
#include <cstdio>
#include <iostream>
template <typename T> static void list(const char* name, const std::vector<T>& v, bool last = false) {
    std::cout << "\"" << name << "\": [";
    for (size_t i = 0; i < v.size(); ++i) std::cout << (i ? ", " : "") << v[i];
    std::cout << "]" << (last ? "" : ", ");
}
int main(int argc, char** argv) {
    if (argc != 2) return 2;
    BenchmarkConfig c = std::string(argv[1]) == "-" ? BenchmarkConfig{} : BenchmarkConfig::load_from_file(argv[1]);
    if (std::string(argv[1]) == "-") c = BenchmarkConfig::load_from_string(c.to_ini_string());
    std::cout.precision(17);
    std::cout << "{";
    list("n_values", c.system.n_values);
    std::cout << "\"seed_base\": " << c.system.seed_base << ", \"device\": \"" << c.system.device << "\", ";
    std::cout << "\"pool\": " << c.traps.pool_minima_count << ", \"traps\": " << c.traps.benchmark_trap_count << ", ";
    std::cout << "\"ranks\": " << c.traps.max_energy_ranks << ", ";
    list("g_values", c.quantum.g_values);
    std::cout << "\"dt\": " << c.quantum.dt << ", \"t_start\": " << c.quantum.t_start << ", \"dt_sample\": "
              << c.quantum.dt_sample << ", \"k\": " << c.quantum.k_sample_points << ", \"m\": "
              << c.quantum.m_shots_per_point << ", \"t_max\": " << c.quantum.t_max() << ", \"times\": "
              << c.quantum.sample_times().size() << ", \"traj\": " << c.quantum.record_trajectory << ", ";
    std::cout << "\"kick\": " << c.classical.kick_trials_per_d << ", \"restart\": " << c.classical.restart_trials
              << ", \"budget\": " << c.classical.max_evals_per_start << ", ";
    list("kick_d", c.classical.kick_d_values);
    std::cout << "\"results\": \"" << c.output.get_results_path(24) << "\", \"series\": \""
              << c.output.get_time_series_path(24) << "\", \"verbose\": " << c.output.verbose << "}\n";
    return 0;
}
"""


def _py_view(c):
    return {"n_values": c.system.n_values, "seed_base": c.system.seed_base, "device": c.system.device,
            "pool": c.traps.pool_minima_count, "traps": c.traps.benchmark_trap_count,
            "ranks": c.traps.max_energy_ranks, "g_values": c.quantum.g_values, "dt": c.quantum.dt,
            "t_start": c.quantum.t_start, "dt_sample": c.quantum.dt_sample, "k": c.quantum.k_sample_points,
            "m": c.quantum.m_shots_per_point, "t_max": c.quantum.t_max(), "times": len(c.quantum.sample_times()),
            "traj": int(c.quantum.record_trajectory), "kick": c.classical.kick_trials_per_d,
            "restart": c.classical.restart_trials, "budget": c.classical.max_evals_per_start,
            "kick_d": c.classical.kick_d_values, "results": c.output.get_results_path(24),
            "series": c.output.get_time_series_path(24), "verbose": int(c.output.verbose)}


def test_cxx_header_agrees_with_python(bk, cfgmod, tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "c.cpp"
    src.write_text(_CXX)
    exe = tmp_path / "c"
    libdir = os.path.dirname(bk.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", libdir, "-lbk", "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    for name in sorted(os.listdir(CONFIGS)):
        out = subprocess.run([str(exe), os.path.join(CONFIGS, name)], capture_output=True, text=True, timeout=60,
                             check=True)
        assert json.loads(out.stdout) == _py_view(cfgmod.load_from_file(os.path.join(CONFIGS, name))), name
    # defaults through the C++ to_ini_string and back
    out = subprocess.run([str(exe), "-"], capture_output=True, text=True, timeout=60, check=True)
    assert json.loads(out.stdout) == _py_view(cfgmod.BenchmarkConfig())
