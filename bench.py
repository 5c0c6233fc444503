#!/usr/bin/env python3
"""Headline benchmark: Trotter amplitude-updates/s and % of HBM peak on one MI355X.

Workload (config "large plain run"): N=30 statevector in complex64, trap drawn
from a seed (mt19937_64 -> steepest descent), second-order (Strang) steps with
dt=0.02 and g=0.35, P_< and <d_H> reduced in FP64 and recorded every step.
A step = one Trotter step of the whole 2^30-amplitude state (phase, 30 qubit
rotations, fused reductions).  value = 2^N * steps / seconds ("amp-updates/s":
one amplitude through one full Trotter step).

Prints ONE JSON line.  The roofline object is measured live with HIP events
around every sweep launch on the library's stream; the cpu_baseline object
times the FP64 CPU restatement (oracle) on a bounded sample of the same
workload on this host's cores.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


def _cpu_baseline(n: int, seed: int, threads: int, what: str = "same workload"):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # CPU restatement, used only as the timed baseline here
    start, s0, e0 = O.make_trap(n, seed)
    sec = O.trotter_bench(n, s0, 0.35, 0.02, 1, order=2, threads=threads)
    return {
        "value": float((1 << n) / sec),
        "unit": "amp-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"1 Strang step of the N={n} trap ({what}), FP64 complex128 "
                  f"statevector restatement of sampler.cuh (int energy cache, per-qubit "
                  f"passes, OpenMP {threads} threads); setup excluded; {sec:.2f} s",
    }


def _step_algorithmic_bytes(n: int, precision: str) -> float:
    """2 x b x 2^n x ceil(n / q) bytes for one Trotter step: every amplitude read
    and written once per sweep, q = 14 (complex64) or 13 (complex128) qubits per
    sweep, the largest tile that fits in 160 KB of LDS."""
    b, q = (8, 14) if precision == "c64" else (16, 13)
    return 2.0 * b * float(1 << n) * -(-n // q)


def _traffic(kernel: str, n: int, precision: str, energy: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE / WRITE_SIZE scaled by a copy-kernel calibration), for
    the workload that summary was measured on; None otherwise."""
    path = os.path.join(ROOT, "profiles", "traffic_n30_c64.json")
    if not (os.path.exists(path) and n == 30 and precision == "c64" and energy == "cache16"):
        return None
    with open(path) as f:
        d = json.load(f)["kernels"]
    entry = d.get(kernel)
    return entry["traffic_bytes_per_launch"] if entry else None


def _adjoint_bench(bk, eng, seed: int) -> dict:
    """Differentiation workload: N=26 c64 Strang, 64 segments x 20 steps of
    dt=0.02 with g_j = 0.35 +- 0.05 drawn from mt19937_64(seed), loss -P_<(T),
    checkpoints every 64 steps.  Times the plain forward (bk_evolve, observables
    at T only), the forward that records the tape, and the reverse pass."""
    import numpy as np
    import torch
    n, nseg, sps, dt, K = 26, 64, 20, 0.02, 64
    _, s0, e0 = bk.make_trap(n, seed)
    g_steps = np.repeat(bk.field_schedule(seed, nseg), sps)
    steps = nseg * sps
    # warm: allocations, energy cache, tape
    eng.adjoint_forward(n, s0, g_steps, dt, precision=bk.BK_C64, e0=e0, checkpoint_every=K)
    eng.adjoint_backward(-1.0, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd = eng.evolve(n, s0, 0.0, dt, steps, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0,
                     g_steps=g_steps, measure_every=0)
    t1 = time.perf_counter()
    obs = eng.adjoint_forward(n, s0, g_steps, dt, precision=bk.BK_C64, e0=e0, checkpoint_every=K)
    t2 = time.perf_counter()
    grad, drift = eng.adjoint_backward(-1.0, 0.0)  # loss = -P_<(T)
    t3 = time.perf_counter()
    eng.timing_enable(True)  # per-sweep device times from a second pass over the same tape
    eng.adjoint_backward(-1.0, 0.0)
    rows = eng.timing_read()
    eng.timing_enable(False)
    t_fwd, t_tape, t_bwd = t1 - t0, t2 - t1, t3 - t2
    sweeps = {r["name"]: {"avg_ms": round(r["total_ms"] / r["launches"], 4), "launches": r["launches"],
                          "gbs": round(r["bytes_per_launch"] / (r["total_ms"] / r["launches"]) * 1e-6, 1)}
              for r in rows}
    gseg = grad.reshape(nseg, sps).sum(axis=1)
    return {
        "workload": f"N={n} c64 Strang dt={dt}, {nseg} segments x {sps} steps, g=0.35+-0.05 "
                    f"(mt19937_64 seed {seed}), loss -P_<(T), checkpoints every {K}",
        "forward_s": t_fwd,
        "forward_with_tape_s": t_tape,
        "backward_s": t_bwd,
        "backward_over_forward": t_bwd / t_fwd,
        "gradient_over_forward": (t_tape + t_bwd) / t_fwd,
        "p_less_T": obs[0],
        "p_less_T_plain_forward": float(fwd["p_less"][-1]),
        "max_abs_dL_dg_segment": float(np.abs(gseg).max()),
        "reversal_drift": drift,
        "reverse_sweeps": sweeps,
    }


def _adam_bench(bk, eng, seed: int, iters: int = 50, lr: float = 0.005) -> dict:
    """The differentiation workload's optimisation: Adam on the 64 segment
    fields of the N=26 c64 adjoint workload (loss -P_<(T)), each iteration one
    taped forward and one reverse pass through the torch.autograd.Function."""
    ag = importlib.import_module("bernasconi-krylov_amd.autograd")
    n, nseg, sps, dt, K = 26, 64, 20, 0.02, 64
    _, s0, e0 = bk.make_trap(n, seed)
    g0 = bk.field_schedule(seed, nseg)
    t0 = time.perf_counter()
    res = ag.optimize_schedule(eng, n, s0, g0, dt, sps, iters=iters, lr=lr, e0=e0, checkpoint_every=K)
    sec = time.perf_counter() - t0
    loss = res["loss"]
    return {"workload": f"N={n} c64 Strang dt={dt}, {nseg} segments x {sps} steps, loss -P_<(T), "
                        f"checkpoints every {K}", "iters": iters, "lr": lr, "seconds": sec,
            "p_less_T_first": float(-loss[0]), "p_less_T_last": float(-loss[-1]),
            "g_min": float(res["g"].min()), "g_max": float(res["g"].max())}


def _linalg_bench(bk, eng, seed: int) -> dict:
    """The subspace side and sampling: the matrix-free H.v (real FP64 at N=30,
    complex128 at N=28; device tensors, five timed calls after a warm one), one
    Krylov step of the propagator at N=26 (m=30, complex128), 100 projective
    samples of the evolved N=30 complex64 state, and subspace Lanczos / dense
    ED on the unique states of 6000 samples of it.  H.v rates count the bytes its
    passes move: each high-qubit group reads x and writes its sum (2 x 8 B per
    real element), the LOW pass reads x, E(s) (2 B) and the groups' sums and
    writes y."""
    import torch
    out = {}

    def timed(f, reps=5):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    for n, cplx in ((30, False), (28, True)):
        dt = torch.complex128 if cplx else torch.float64
        x = torch.randn(1 << n, dtype=dt, device="cuda")
        y = torch.empty_like(x)
        sec = timed(lambda: eng.hv(n, 0.35, x, y))
        b = 16 if cplx else 8                                   # bytes per element
        hb = 13 if cplx else 14                                 # element bits of a tile
        ng = max(0, -(-(n - hb) // 9))                          # high groups of at most 9 qubits
        # LOW: x read, y written, E(s), up to two group sums read; each group: x read, sum written
        moved = (1 << n) * ((2 + min(ng, 2)) * b + 2 + ng * 2 * b)
        out[f"hv_n{n}_{'c128' if cplx else 'f64'}"] = {"ms": sec * 1e3, "high_groups": ng,
                                                       "bytes_moved": moved, "gbs_moved": moved / sec / 1e9}
        del x, y
        torch.cuda.empty_cache()
    n = 26
    _, s0, _ = bk.make_trap(n, seed)
    eng.state_init(n, bk.BK_C128, s0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.krylov_propagate(0.35, 0.25, tau_restart=0.25, m=30)
    torch.cuda.synchronize()
    out["krylov_n26_m30_restart_s"] = time.perf_counter() - t0
    n = 30
    _, s0, e0 = bk.make_trap(n, seed)
    eng.evolve(n, s0, 0.35, 0.02, 10, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0, measure_every=0)
    import numpy as np
    u = np.random.default_rng(seed).random(100)
    sec = timed(lambda: eng.sample(u))
    out["sample_n30_c64_100_shots"] = {"ms": sec * 1e3, "gbs_one_read": 8.0 * (1 << n) / sec / 1e9}
    # the subspace path on a sampled basis: the unique states of 6000 shots (the
    # reference's default 20 points x 100 shots x 3 seeds) of the state evolved
    # on to t = 5.2, where it has spread over many basis states
    eng.evolve(n, s0, 0.35, 0.02, 250, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0, measure_every=0,
               init=False)
    basis = np.unique(eng.sample(np.random.default_rng(seed + 1).random(6000))).astype(np.uint64)
    sl = timed(lambda: eng.subspace_lanczos(n, 0.35, basis, max_krylov_dim=100, k_pairs=2, vectors=False), reps=3)
    sd = timed(lambda: eng.subspace_dense(n, 0.35, basis, k_pairs=2, vectors=False), reps=1)
    out["subspace_n30_sampled_basis"] = {"t": 5.2, "shots": 6000, "basis_states": int(len(basis)),
                                         "lanczos_m100_ms": sl * 1e3, "dense_ms": sd * 1e3}
    return out


# quantum_time_sec of the reference's own run of this configuration and trap
# (data_processed/probability_dynamics_n30_results.csv:2, N=30 g=0.3 trap 0x38C06C95)
REF_SAMPLER_N30_S = 324.6727


def _sampler_e2e_bench(bk) -> dict:
    """ISampler::sample end to end on the reference's N=30 probability-dynamics
    configuration (configs/probability_scan_n30.cfg: Lie dt=0.05 to t=60, 1200
    steps, P_< and <d_H> recorded every step, no shots) for g=0.3 and the
    reference's trap 0x38C06C95, in complex128 (the reference's precision) and
    complex64; wall time beside the reference's recorded time for the run."""
    config = importlib.import_module("bernasconi-krylov_amd.config")
    cfg = config.load_from_file(os.path.join(ROOT, "tests", "golden", "configs", "probability_scan_n30.cfg"))
    out = {"reference_s": REF_SAMPLER_N30_S}
    for name, prec in (("c128", bk.BK_C128), ("c64", bk.BK_C64)):
        sc = config.to_sampler_config(cfg.quantum, 30, 0.3, 42)
        sc.precision = prec
        sc.verbose = False
        t0 = time.perf_counter()
        r = bk.create_sampler(30).sample(0x38C06C95, sc)
        sec = time.perf_counter() - t0
        out[name] = {"wall_s": sec, "speedup": REF_SAMPLER_N30_S / sec, "peak_p_less": r.peak_p_less,
                     "peak_time": r.peak_time, "p_blind": r.p_blind, "d_min": r.d_min,
                     "initial_rise_rate_avg": r.initial_rise_rate_avg, "points": len(r.time_points)}
    # the reference's row for this run (probability_dynamics_n30_results.csv)
    with open(os.path.join(ROOT, "tests", "golden", "results.json")) as f:
        row = [x for x in json.load(f) if x["N"] == "30" and int(x["trap_bits"], 16) == 0x38C06C95
               and float(x["g"]) == 0.3]
    if row:
        out["reference_row"] = {k: row[0][k] for k in ("peak_p_less", "peak_time", "p_blind", "d_min",
                                                       "initial_rise_rate_avg", "quantum_time_sec")}
    return out


def _batch_bench(bk, eng, seed: int) -> dict:
    """Batched workload: 512 traps, n drawn from {16, 18, 20, 22} (mt19937_64),
    16 fields evenly spaced in [0.10, 0.60], Strang dt=0.02 x 500 steps, c64,
    the whole run one hipGraph launch with autotuned sweep parameters.  The
    serial baseline runs 8 instances of each n one bk_evolve at a time (same
    kernels, observables at the end) and scales by the instance counts; each
    n's instances also run as a batch of their own (per-n ratio and HBM rate,
    counting 17.5 B per amplitude and step: state read and written, the 12-bit
    energy codes)."""
    import numpy as np
    steps, dt = 500, 0.02
    inst = bk.batch_workload(512, seed=seed)
    eng.batch_run(inst, dt, steps, use_graph=True, autotune=True)  # tune (cached) + warm
    t0 = time.perf_counter()
    res, st = eng.batch_run(inst, dt, steps, use_graph=True, autotune=True)
    t_batch = time.perf_counter() - t0
    amps = float(sum(1 << n for n, *_ in inst)) * steps
    serial_est, sample, identical = 0.0, {}, True
    for n in sorted({n for n, *_ in inst}):
        idx = [i for i, w in enumerate(inst) if w[0] == n]
        pick = idx[:8]
        t0 = time.perf_counter()
        for i in pick:
            _, s0, e0, g = inst[i]
            r = eng.evolve(n, s0, g, dt, steps, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0, measure_every=0)
            identical &= bool(r["p_less"][-1] == res[i, 0] and r["dh"][-1] == res[i, 1])
        per = (time.perf_counter() - t0) / len(pick)
        serial_est += per * len(idx)
        # the same instances of this n as a batch of their own
        sub = [inst[i] for i in idx]
        eng.batch_run(sub, dt, steps, use_graph=True, autotune=True)
        t0 = time.perf_counter()
        rs, sst = eng.batch_run(sub, dt, steps, use_graph=True, autotune=True)
        tb = time.perf_counter() - t0
        identical &= bool(np.array_equal(rs, res[idx]))
        step_bytes = len(idx) * float(1 << n) * 17.5  # state read + written (c64) + 1.5 B energy code
        sample[str(n)] = {"instances": len(idx), "serial_s_per_instance": per,
                          "serial_s_estimated": per * len(idx), "batched_s": tb,
                          "batched_over_serial": per * len(idx) / tb,
                          "batched_hbm_gbs": step_bytes * steps / tb / 1e9,
                          "block_bytes": sst["block_bytes"], "out_of_place": sst["out_of_place"]}
    return {
        "workload": f"512 traps, n in {{16,18,20,22}} (mt19937_64 seed {seed}), g = 0.10..0.60 (16 values), "
                    f"Strang dt={dt} x {steps} steps, c64, one hipGraph launch",
        "batched_s": t_batch,
        "batched_device_s": st["device_ms"] * 1e-3,
        "batched_amp_updates_per_s": amps / t_batch,
        "serial_s_estimated": serial_est,
        "serial_amp_updates_per_s": amps / serial_est,
        "batched_over_serial": serial_est / t_batch,
        "per_n": sample,
        "bit_identical_to_serial_sample": identical,
        "tuning": {"grid": st["grid"], "row_bits": st["row_bits"], "block_bytes": st["block_bytes"],
                   "out_of_place": st["out_of_place"]},
        "graph_nodes": st["launches"],
    }


def _minima_bench(bk, eng) -> dict:
    """The exhaustive local-minima explorer at N=30 (2^29 half-space states):
    scan, the two trap levels, exact d_min of every orbit and 10^5 seed-42
    descents for the hit fractions, checked against the reference's
    exhaustive_minima_n30.json (tests/golden)."""
    n = 30
    eng.exhaustive_minima(20)  # warm
    t0 = time.perf_counter()
    r = eng.exhaustive_minima(n, num_sd_starts=100000, sd_seed=42)
    wall = time.perf_counter() - t0
    with open(os.path.join(ROOT, "tests", "golden", "exhaustive_minima.json")) as f:
        rows = [x for x in json.load(f) if x["n"] == n]
    got = [(lv["E0"], o["canonical"], o["d_min"]) for lv in r["levels"] for o in lv["orbits"]]
    want = [(x["e0"], int(x["canonical_id"], 16), x["d_min"]) for x in rows]
    out = {"workload": f"N={n}: 2^{n - 1} states scanned for one-flip minima, 2 trap levels with exact d_min, "
                       f"100000 seed-42 descents",
           "scan_s": r["search_time_sec"], "total_s": wall,
           "states_per_s": float(1 << (n - 1)) / r["search_time_sec"],
           "minima": r["total_minima_found"], "orbits": r["total_orbits_found"],
           "matches_reference_file": r["e_star"] == rows[0]["e_star"] and got == want}
    return out


def _trap_supply_bench(bk, eng, seed: int):
    """Trap supply: steepest descent of 2^18 random N=30 starts on the GPU
    (bk_descent_batch).  Returns the bench object and the first starts with
    their GPU results, which the CPU-baseline leg descends again on the host."""
    import numpy as np
    n, count = 30, 1 << 18
    starts = np.random.default_rng(seed).integers(0, 1 << n, size=count, dtype=np.uint64)
    eng.descent_batch(n, starts[:4096])  # warm
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fin, fe, ev = eng.descent_batch(n, starts)
    t_gpu = time.perf_counter() - t0
    k = 2000
    out = {
        "workload": f"steepest descent of {count} random N={n} starts (numpy default_rng seed {seed})",
        "gpu_descents_per_s": count / t_gpu,
        "gpu_energy_evals_per_s": float(ev.sum()) / t_gpu,
    }
    return out, (n, starts[:k], fin[:k], fe[:k], ev[:k])


def _trap_supply_cpu(O, sample) -> dict:
    """The oracle's host descent (one thread) of the same first starts: rate,
    and bit-identity with the GPU's final states, energies and evaluations."""
    n, starts, fin, fe, ev = sample
    t0 = time.perf_counter()
    ref = [O.steepest_descent(n, int(s)) for s in starts]
    t_cpu = time.perf_counter() - t0
    same = all((int(fin[i]), int(fe[i]), int(ev[i])) == ref[i][:3] for i in range(len(starts)))
    return {"value": len(starts) / t_cpu, "unit": "descents/s", "cores": 1, "kind": "port",
            "sample": f"oracle steepest descent, 1 thread, the first {len(starts)} N={n} starts of the trap-supply leg",
            "bit_identical_to_gpu": same}


def _minima_cpu(O, eng) -> dict:
    """The reference's own run_exploration (oracle/_ref, one host thread) at N=24
    beside the GPU explorer at the same n."""
    m = 24
    t0 = time.perf_counter()
    R = O.ref_exhaustive_minima(m)
    t_ref = time.perf_counter() - t0
    g0 = time.perf_counter()
    G = eng.exhaustive_minima(m)
    t_gpu = time.perf_counter() - g0
    return {"value": float(1 << (m - 1)) / t_ref, "unit": "states/s", "cores": 1, "kind": "reference",
            "sample": f"the reference's exhaustive-minima explorer (oracle/_ref) at N={m}: {t_ref:.2f} s",
            "gpu_seconds_same_n": t_gpu,
            "same_result": (G["e_star"], G["total_minima_found"], G["total_orbits_found"]) ==
                           (R["e_star"], R["total_minima"], R["total_orbits"])}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--precision", choices=["c64", "c128"], default="c64")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--g", type=float, default=0.35)
    ap.add_argument("--dt", type=float, default=0.02)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=30)
    ap.add_argument("--energy", choices=["cache16", "onthefly"], default="cache16",
                    help="how the first sweep obtains E(s)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the N=26 adjoint/forward timing")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 512-trap batched timing")
    ap.add_argument("--no-traps", action="store_true",
                    help="skip the trap-supply (batched steepest descent) timing")
    ap.add_argument("--no-minima", action="store_true",
                    help="skip the exhaustive local-minima explorer at N=30")
    ap.add_argument("--no-linalg", action="store_true",
                    help="skip the H.v / Krylov / sampling timings")
    ap.add_argument("--no-sampler-e2e", action="store_true",
                    help="skip ISampler::sample on the reference's N=30 configuration and trap")
    ap.add_argument("--adam", action="store_true",
                    help="also run 50 Adam steps on the N=26 adjoint workload's fields")
    ap.add_argument("--no-fp64-check", action="store_true",
                    help="skip the complex128 run of the same steps that the complex64 "
                         "series are compared with")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the plain 8 GiB device copy (HBM ceiling / PMC calibration)")
    ap.add_argument("--cache-sweep", action="store_true",
                    help="only run the plain copy over working sets from 16 MB to 8 GiB "
                         "(how much faster a cache-resident working set streams than HBM)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py runs on one GPU only (asked for {max(args.gpus, world)}): the "
              f"statevector path has no multi-GPU sharding in this build.", file=sys.stderr)
        return 2

    bk = importlib.import_module("bernasconi-krylov_amd")
    if args.cache_sweep:
        eng = bk.Engine(0)
        rows = {}
        for mb in (8, 16, 32, 64, 128, 256, 1024, 8192):  # bytes copied each way; working set is twice that
            rows[f"{2 * mb} MB working set"] = eng.copy_probe(mb << 20, 50 if mb <= 256 else 5)
            rows[f"{2 * mb} MB working set, 4 KB workgroups"] = eng.copy_probe(mb << 20, 50 if mb <= 256 else 5,
                                                                             pattern=1)
        print(json.dumps({"copy_gbs_read_plus_write": rows}))
        eng.close()
        return 0
    if args.copy_probe:
        eng = bk.Engine(0)
        print(json.dumps({"copy_probe_bytes": 1 << 33,
                          "gbs_read_plus_write": {"grid_stride": eng.copy_probe(1 << 33, 5),
                                                  "4KB_workgroups": eng.copy_probe(1 << 33, 5, pattern=1),
                                                  "128KB_workgroups": eng.copy_probe(1 << 33, 5, pattern=2)}}))
        eng.close()
        return 0
    prec = bk.BK_C64 if args.precision == "c64" else bk.BK_C128
    n = args.n
    start, s0, e0 = bk.make_trap(n, args.seed)
    eng = bk.Engine(0)
    eng.set_energy_mode(bk.BK_ENERGY_CACHE16 if args.energy == "cache16"
                        else bk.BK_ENERGY_ONTHEFLY)

    import torch  # device plumbing only: synchronisation around the timed regions

    def timed_run(e, precision):
        """Warmup steps from the trap, then the timed continuation of K steps
        (P_< / <d_H> / norm every step); no per-launch events in the timed region."""
        e.evolve(n, s0, args.g, args.dt, args.warmup, order=bk.BK_STRANG, precision=precision, e0=e0,
                 init=True, measure_every=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = e.evolve(n, s0, args.g, args.dt, args.steps, order=bk.BK_STRANG, precision=precision, e0=e0,
                     init=False, measure_every=1)
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    print(f"[bench] N={n} {args.precision} warmup={args.warmup} steps={args.steps}",
          file=sys.stderr, flush=True)
    res, sec = timed_run(eng, prec)
    ms_per_step = 1e3 * sec / args.steps
    value = float((1 << n) * args.steps / sec)
    # per-kernel device times: the same continuation once more, every launch
    # bracketed by HIP events on the library's stream
    eng.timing_enable(True)
    eng.evolve(n, s0, args.g, args.dt, args.steps, order=bk.BK_STRANG, precision=prec, e0=e0,
               init=False, measure_every=1)
    rows = eng.timing_read()
    eng.timing_enable(False)

    dom = max(rows, key=lambda r: r["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    achieved = dom["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9
    step_bytes = sum(r["bytes_per_launch"] * r["launches"] for r in rows) / args.steps
    step_dev_ms = sum(r["total_ms"] for r in rows) / args.steps
    kernels = {r["name"]: {"avg_ms": round(r["total_ms"] / r["launches"], 4),
                           "launches": r["launches"],
                           "gbs": round(r["bytes_per_launch"] / (r["total_ms"] / r["launches"])
                                        * 1e-6, 1)} for r in rows}

    # the same warmup + timed steps in complex128 (the reference's precision),
    # timed the same way: the error of the complex64 series against it, step by
    # step, and of the final state, 1 - |<psi64|psi128>|
    fp64 = None
    if prec == bk.BK_C64 and not args.no_fp64_check:
        print(f"[bench] N={n} c128 of the same {args.warmup} + {args.steps} steps", file=sys.stderr, flush=True)
        # the instrumented pass moved the complex64 state on: redo the warmup and
        # the timed steps exactly as timed, so both states sit at step W + K
        for k, init in ((args.warmup, True), (args.steps, False)):
            eng.evolve(n, s0, args.g, args.dt, k, order=bk.BK_STRANG, precision=prec, e0=e0, init=init,
                       measure_every=1)
        e128 = bk.Engine(0)
        try:
            ref, sec128 = timed_run(e128, bk.BK_C128)
            ov = eng.overlap(e128)
            n64, n128 = eng.overlap(eng).real, e128.overlap(e128).real  # squared norms, FP64 sums
        finally:
            e128.close()
        rp, rd = np.asarray(ref["p_less"]), np.asarray(ref["dh"])
        fp64 = {"steps_compared": int(len(rp)),
                "p_less_max_abs_err": float(np.abs(np.asarray(res["p_less"]) - rp).max()),
                "dh_max_abs_err": float(np.abs(np.asarray(res["dh"]) - rd).max()),
                "norm_max_abs_dev": float(np.abs(np.asarray(res["norm"]) - 1.0).max()),
                # 1 - |<psi64|psi128>| / (|psi64| |psi128|), every sum in FP64
                "state_overlap_err": float(1.0 - abs(ov) / np.sqrt(n64 * n128)),
                "state_overlap_raw": [ov.real, ov.imag],
                "c128_ms_per_step": 1e3 * sec128 / args.steps,
                "c128_amp_updates_per_s": float((1 << n) * args.steps / sec128)}

    # read+write copies of 8 GiB on this device: the grid-stride copy, the part's
    # fastest stream (4 KB per workgroup) and a copy whose workgroups hold 128 KB,
    # as a complex64 sweep tile does
    copy_gbs = eng.copy_probe(1 << 33, 3)
    copy_best_gbs = eng.copy_probe(1 << 33, 3, pattern=1)
    copy_tile_gbs = eng.copy_probe(1 << 33, 3, pattern=2)

    adj = None
    if not args.no_adjoint:
        print("[bench] adjoint N=26", file=sys.stderr, flush=True)
        adj = _adjoint_bench(bk, eng, args.seed)

    linalg = None
    if not args.no_linalg:
        print("[bench] H.v / Krylov / sampling", file=sys.stderr, flush=True)
        linalg = _linalg_bench(bk, eng, args.seed)

    e2e = None
    if not args.no_sampler_e2e:
        print("[bench] sampler run, N=30 reference configuration", file=sys.stderr, flush=True)
        e2e = _sampler_e2e_bench(bk)

    adam = None
    if args.adam:
        print("[bench] 50 Adam steps N=26", file=sys.stderr, flush=True)
        adam = _adam_bench(bk, eng, args.seed)

    batch = None
    if not args.no_batch:
        print("[bench] batched 512 traps", file=sys.stderr, flush=True)
        batch = _batch_bench(bk, eng, args.seed)

    traps = trap_sample = None
    if not args.no_traps:
        print("[bench] trap supply N=30", file=sys.stderr, flush=True)
        traps, trap_sample = _trap_supply_bench(bk, eng, args.seed)

    minima = None
    if not args.no_minima:
        print("[bench] exhaustive local minima N=30", file=sys.stderr, flush=True)
        minima = _minima_bench(bk, eng)

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        print(f"[bench] cpu baseline N={args.cpu_n} threads={threads}", file=sys.stderr,
              flush=True)
        cpu = _cpu_baseline(args.cpu_n, args.seed, threads)
        # the same restatement on one core, on a 16x smaller trap of the same construction
        n1 = max(args.cpu_n - 4, 12)
        print(f"[bench] cpu baseline N={n1} threads=1", file=sys.stderr, flush=True)
        one = _cpu_baseline(n1, args.seed, 1, f"the N={args.cpu_n} trap construction at N={n1}")
        cpu["single_core"] = {"value": one["value"], "unit": one["unit"], "cores": 1, "sample": one["sample"]}
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # CPU restatement and the reference's compiled code: timed baselines and checkers only
        if trap_sample is not None:
            print("[bench] cpu baseline: trap-supply descents", file=sys.stderr, flush=True)
            cpu["trap_supply"] = _trap_supply_cpu(O, trap_sample)
        if minima is not None and os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libbk_ref.so")):
            print("[bench] cpu baseline: the reference's explorer N=24", file=sys.stderr, flush=True)
            cpu["exhaustive_minima"] = _minima_cpu(O, eng)

    out = {
        "metric": "Trotter amp-updates/s & % HBM peak, N=30 c64; adjoint/forward time ratio",
        "value": value,
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seed-drawn steepest-descent trap, no external data)",
        "config": {
            "workload": f"N={n} {args.precision} Strang dt={args.dt} g={args.g} trotter step, "
                        f"P_< and <d_H> (FP64) fused every step",
            "n": n, "energy": args.energy, "trap_seed": args.seed, "trap_bits": hex(s0), "E0": e0,
            "p_less_final": float(res["p_less"][-1]), "norm_final": float(res["norm"][-1]),
            "vs_complex128": fp64,
            "step_hbm_bytes": step_bytes,
            "step_hbm_frac": step_bytes / (step_dev_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            # the step's algorithmic bytes as a sweep-per-qubit-group schedule
            # moves them: 3 sweeps x (8 B read + 8 B written) per amplitude for
            # complex64 at N <= 32 (a 2^14-amplitude LDS tile rotates at most 14
            # qubits), against the measured time of one step
            "step_algorithmic_bytes": _step_algorithmic_bytes(n, args.precision),
            "step_algorithmic_frac": _step_algorithmic_bytes(n, args.precision) / (ms_per_step * 1e-3) / 1e9
                                     / PEAK_HBM_GBS,
            "kernels": kernels,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom["name"],
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": _traffic(dom["name"], n, args.precision, args.energy),
            "copy_ceiling_gbs": copy_gbs,
            "frac_of_copy_ceiling": achieved / copy_gbs,
            "copy_best_gbs": copy_best_gbs,
            "frac_of_copy_best": achieved / copy_best_gbs,
            "copy_tile_gbs": copy_tile_gbs,
            "frac_of_copy_tile": achieved / copy_tile_gbs,
        },
        "adjoint": adj,
        "adam": adam,
        "sampler_e2e": e2e,
        "linalg": linalg,
        "batched": batch,
        "trap_supply": traps,
        "exhaustive_minima": minima,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
