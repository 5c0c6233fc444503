"""The C++ drop-in headers on the device, as the reference's own code would use
them: a probe compiled against include/bk/config.hpp, sampler.hpp,
diagonalizer.hpp and wire.hpp reads the reference's probability_scan_n16.cfg
(kept in tests/golden/configs/), maps it onto a sampler run with
to_sampler_config (connector.cuh:628-648) for its first field g = 0.3 and the
N=16 trap 0xD2EE, runs create_sampler(16)->sample() on the GPU
(test_sampler.cu:20-53 calls it the same way) and writes the connector's
time-series CSV.  The file is checked against the rows the reference wrote
for that trap (data_processed/probability_dynamics_n16.csv); the probe's
Diagonalizer::solve (Lanczos and ExactDense, test_lanczos.cu:22-29) against
the oracle's dense eigenvalues for n = 4..10."""
from __future__ import annotations

import csv
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

_PROBE = r"""
#include "bk/config.hpp"
#include "bk/diagonalizer.hpp"
#include "bk/sampler.hpp"
#include "bk/wire.hpp"

#include <cstdio>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    BenchmarkConfig cfg = BenchmarkConfig::load_from_file(argv[1]);
    const int n = cfg.system.n_values.at(0);
    const double g = cfg.quantum.g_values.at(0);
    SamplerConfig sc = bk::config::to_sampler_config(cfg.quantum, n, g, cfg.system.seed_base);
    sc.verbose = false;
    const SpinState trap(0xD2EEull);
    auto sampler = create_sampler(n);
    const SamplingResult res = sampler->sample(trap, sc);
    bk::wire::TrapBenchmark b;
    b.n = n;
    b.g = g;
    b.trap_rank = 1;
    b.trap_state = trap;
    b.trap_energy = res.start_energy;
    bk::wire::set_canonical(b, bk_canonical(n, trap.bits), n);
    bk::wire::fill_quantum(b, res);
    bk::wire::export_time_series_csv(argv[2], std::vector<bk::wire::TrapBenchmark>{b});
    std::printf("run n=%d g=%.2f t_max=%.1f dt=%.2f points=%zu E0=%lld d_min=%d p_blind=%.6e\n", n, g, sc.t_max,
                sc.dt, res.time_points.size(), static_cast<long long>(res.start_energy), res.d_min, res.p_blind);
    Diagonalizer diag;
    for (int m = 4; m <= 10; ++m) {
        DiagConfig lz;
        lz.num_eigenpairs = 1;
        lz.max_krylov_dim = 60;
        const Eigensolution a = diag.solve(m, 0.5, lz);
        DiagConfig ed;
        ed.method = DiagMethod::ExactDense;
        ed.num_eigenpairs = 3;
        const Eigensolution d = diag.solve(m, 0.5, ed);
        std::printf("diag %d %.15e %.15e %.15e %.15e\n", m, a.ground_energy(), d.pairs[0].energy,
                    d.pairs[1].energy, d.pairs[2].energy);
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def probe(bk, tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    d = tmp_path_factory.mktemp("cxx")
    src = d / "probe.cpp"
    src.write_text(_PROBE)
    exe = d / "probe"
    libdir = os.path.dirname(bk.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-L", libdir, "-lbk", "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    csv_path = d / "probability_dynamics_n16.csv"
    cfg = os.path.join(ROOT, "tests", "golden", "configs", "probability_scan_n16.cfg")
    print("C++ probe: probability_scan_n16.cfg, trap 0xD2EE, g = 0.3 on the GPU", flush=True)
    out = subprocess.run([str(exe), cfg, str(csv_path)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout.splitlines(), csv_path.read_text().splitlines()


def test_cxx_sampler_from_reference_config(probe, golden):
    lines, rows = probe
    run = [ln for ln in lines if ln.startswith("run ")][0]
    # the configuration's values reach the sampler: N, first g, t_max = 60, dt = 0.05,
    # a trajectory point every dt (1201), the trap's energy E* = 24 (no lower state)
    assert run == "run n=16 g=0.30 t_max=60.0 dt=0.05 points=1201 E0=24 d_min=16 p_blind=0.000000e+00"
    ref_lines = golden("wire_n16.json")["time_series_lines"]
    assert rows[0] == ref_lines[0] and len(rows) == 1 + 1201
    # the reference's first rows for this trap and field, text for text
    trap_rows = [ln for ln in ref_lines[1:] if ln.split(",")[2] == "1" and float(ln.split(",")[1]) == 0.3]
    assert len(trap_rows) >= 5
    assert rows[1:1 + len(trap_rows)] == trap_rows
    # every committed point of the reference's 1200-step trajectory
    tr = [t for t in golden("trajectories.json") if t["n"] == 16 and t["trap_bits"] == "0xD2EE" and t["g"] == 0.3][0]
    got = list(csv.reader(rows[1:]))
    for k, p, dh in zip(tr["step"], tr["p_less"], tr["dh"]):
        assert float(got[k][5]) == pytest.approx(float(p), rel=2e-6, abs=1e-15), k
        assert abs(float(got[k][6]) - float(dh)) <= 1.01e-4, k


def test_cxx_diagonalizer_against_oracle(probe, orc):
    lines, _ = probe
    diag = [ln.split() for ln in lines if ln.startswith("diag ")]
    assert [int(d[1]) for d in diag] == list(range(4, 11))
    for d in diag:
        m = int(d[1])
        lz, e0, e1, e2 = (float(x) for x in d[2:6])
        w = np.linalg.eigvalsh(orc.dense_h(m, 0.5))
        assert abs(lz - w[0]) <= 1e-6  # test_lanczos.cu:38
        assert np.allclose([e0, e1, e2], w[:3], rtol=0, atol=1e-10), (m, [e0, e1, e2], w[:3])
