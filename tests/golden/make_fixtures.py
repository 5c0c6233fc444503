"""Extract small golden vectors from the reference's committed data.

Run in the development container (it reads the reference checkout, which
the GPU machines do not have):

    python tests/golden/make_fixtures.py [/path/to/reference]

Writes JSON fixtures next to this script.  Every value is copied from the
reference's own data files (no computation), subsampled to keep them small:

* energy_tables.json   : E(s) and C_k from data_raw/classical/exhaustive/{N}.csv
                         (full tables for N <= 12, 2000 seeded rows for N=16, 20)
* trajectories.json    : P_<(t), <d_H>(t) from data_processed/probability_dynamics_n*.csv
                         (1200 Lie steps of dt=0.05, 7/4 printed digits)
* results.json         : per-(trap, g) rows of probability_dynamics_n*_results.csv
* exhaustive_minima.json : (N, E*, E0, canonical_id, bitstring, d_min, sd_hit_fraction)
                         from exhaustive_minima_n*.json
* exhaustive_minima_files.json : the text of exhaustive_minima_n{20,24,28}.json
* seed_traps.json      : seed -> (start, s0, E0, evals, steps, d_min) from the
                         reference's CPU code compiled into oracle/_ref (if built)
"""
from __future__ import annotations

import csv
import ctypes as C
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def energy_tables():
    out = {}
    rng = random.Random(1234)
    for n in (4, 6, 8, 10, 12, 16, 20):
        path = os.path.join(REF, "data_raw/classical/exhaustive", f"{n}.csv")
        with open(path) as f:
            rows = list(csv.reader(f))
        hdr, body = rows[0], rows[1:]
        assert hdr[0] == "Sequence" and hdr[-1] == "E(s)"
        if n <= 12:
            pick = body
        else:
            pick = [body[i] for i in sorted(rng.sample(range(len(body)), 2000))]
        out[str(n)] = {
            "s": [int(r[0]) for r in pick],
            "ck": [[int(x) for x in r[1:-1]] for r in pick],
            "e": [int(r[-1]) for r in pick],
        }
    return out


def trajectories():
    """(N, trap_rank, g) -> series; step stride chosen per N."""
    plan = {
        16: ([1, 2], [0.3, 0.45], 1),
        20: ([1, 2], [0.3, 0.35, 0.45], 1),
        24: ([1], [0.35], 5),
        26: ([1], [0.35], 20),
        28: ([1], [0.35], 20),
        30: ([1], [0.35], 20),
    }
    out = []
    for n, (ranks, gs, stride) in plan.items():
        path = os.path.join(REF, "data_processed", f"probability_dynamics_n{n}.csv")
        series = {}
        meta = {}
        with open(path) as f:
            rd = csv.reader(f)
            next(rd)
            for r in rd:
                rank, g = int(r[2]), float(r[1])
                key = (rank, round(g, 6))
                if rank not in ranks or round(g, 6) not in [round(x, 6) for x in gs]:
                    continue
                series.setdefault(key, []).append((float(r[4]), r[5], r[6]))
                meta[key] = (int(r[3]), r[7])
        res_path = os.path.join(REF, "data_processed", f"probability_dynamics_n{n}_results.csv")
        bits = {}
        with open(res_path) as f:
            for row in csv.DictReader(f):
                bits[(int(row["trap_rank"]), round(float(row["g"]), 6))] = row["trap_bits"]
        for key, pts in sorted(series.items()):
            steps = list(range(0, len(pts), stride))
            out.append({
                "n": n, "trap_rank": key[0], "g": key[1], "trap_energy": meta[key][0],
                "trap_bits": bits[key], "p_blind": meta[key][1], "dt": 0.05,
                "steps_total": len(pts) - 1, "step": steps,
                "p_less": [pts[i][1] for i in steps], "dh": [pts[i][2] for i in steps],
            })
    return out


def results():
    out = []
    for n in (16, 20, 24, 26, 28, 30):
        path = os.path.join(REF, "data_processed", f"probability_dynamics_n{n}_results.csv")
        with open(path) as f:
            for row in csv.DictReader(f):
                out.append({k: row[k] for k in (
                    "N", "g", "trap_rank", "trap_energy", "d_min", "window_avg_p_less",
                    "window_avg_dh", "initial_rise_rate_avg", "initial_rise_rate_max",
                    "peak_p_less", "peak_time", "p_blind", "quantum_time_sec", "trap_bits")})
    return out


def exhaustive_minima():
    out = []
    for n in (20, 22, 24, 26, 28, 30):
        with open(os.path.join(REF, "data_processed", f"exhaustive_minima_n{n}.json")) as f:
            d = json.load(f)
        for lvl in d["levels"]:
            for o in lvl["orbits"]:
                out.append({"n": d["N"], "e_star": d["E_star"], "e0": lvl["E0"],
                            "canonical_id": o["canonical_id"], "bitstring": o["bitstring"],
                            "d_min": o["d_min"], "sd_hit_fraction": o["sd_hit_fraction"]})
    return out


def exhaustive_minima_files():
    """The explorer's output files themselves (the three smallest), as text: the
    expected bytes of the JSON writer."""
    out = {}
    for n in (20, 24, 28):
        with open(os.path.join(REF, "data_processed", f"exhaustive_minima_n{n}.json")) as f:
            out[str(n)] = f.read()
    return out


def seed_traps():
    so = os.path.join(HERE, "..", "..", "oracle", "_ref", "libbk_ref.so")
    if not os.path.exists(so):
        return None
    R = C.CDLL(so)
    R.bkref_random_state.restype = C.c_uint64
    R.bkref_random_state.argtypes = [C.c_int, C.c_uint64]
    R.bkref_steepest_descent.argtypes = [C.c_int, C.c_uint64] + [C.c_void_p] * 4
    R.bkref_d_min.restype = C.c_int
    R.bkref_d_min.argtypes = [C.c_int, C.c_uint64, C.c_int64, C.c_int]
    out = []
    for n in (12, 16, 20, 24, 26):
        for seed in range(0, 40):
            start = R.bkref_random_state(n, seed)
            fs, fe, ev, st = C.c_uint64(), C.c_int64(), C.c_uint64(), C.c_uint64()
            R.bkref_steepest_descent(n, start, C.byref(fs), C.byref(fe), C.byref(ev), C.byref(st))
            dm = R.bkref_d_min(n, fs.value, fe.value, -1) if n <= 24 else -1
            out.append({"n": n, "seed": seed, "start": start, "s0": fs.value, "e0": fe.value,
                        "evals": ev.value, "steps": st.value, "d_min": dm})
    return out


def main():
    fixtures = {
        "energy_tables.json": energy_tables(),
        "trajectories.json": trajectories(),
        "results.json": results(),
        "exhaustive_minima.json": exhaustive_minima(),
        "exhaustive_minima_files.json": exhaustive_minima_files(),
    }
    st = seed_traps()
    if st is not None:
        fixtures["seed_traps.json"] = st
    for name, data in fixtures.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print(name, os.path.getsize(os.path.join(HERE, name)), "bytes")


if __name__ == "__main__":
    main()
