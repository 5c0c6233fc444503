"""Write tests/golden/wire_n16.json from the reference checkout's N=16 time-series
CSV (data_processed/probability_dynamics_n16.csv, written by the connector's
export_time_series_csv) and results CSV (for the trap bits).

Kept: the header and the first rows of three records -- (rank 1, g=0.3),
(rank 1, g=0.35), (rank 5, g=0.3) -- as the lines the reference wrote, with the
values parsed from them, plus the rank-5 trap's rows for t <= 2 at g=0.3 (41
rows) for an end-to-end check of a sampler run written through the writers.

Run in a container that has the reference at /root/reference:
    python tests/golden/make_wire_fixture.py
"""
import csv
import json
import os

REF = "/root/reference/data_processed"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wire_n16.json")


def main():
    with open(os.path.join(REF, "probability_dynamics_n16.csv")) as f:
        lines = f.read().splitlines()
    header, rows = lines[0], lines[1:]
    parsed = [r.split(",") for r in rows]

    def block(rank, g, k):
        out = []
        for line, p in zip(rows, parsed):
            if int(p[2]) == rank and abs(float(p[1]) - g) < 1e-12:
                out.append((line, p))
                if len(out) == k:
                    break
        return out

    sel = [(1, 0.30, 6), (1, 0.35, 3), (5, 0.30, 4)]
    records, expect = [], [header]
    for rank, g, k in sel:
        b = block(rank, g, k)
        expect += [line for line, _ in b]
        p0 = b[0][1]
        records.append({"n": int(p0[0]), "g": g, "trap_rank": rank, "trap_energy": int(p0[3]),
                        "p_blind": float(p0[7]),
                        "time": [float(p[4]) for _, p in b],
                        "p_less": [float(p[5]) for _, p in b],
                        "dh": [float(p[6]) for _, p in b]})
    traps = {}
    with open(os.path.join(REF, "probability_dynamics_n16_results.csv")) as f:
        for r in csv.DictReader(f):
            traps.setdefault(r["trap_rank"], {"bits": int(r["trap_bits"], 16), "E0": int(r["trap_energy"])})
    e2e = [(line, p) for line, p in block(5, 0.30, 41)]
    json.dump({
        "source": "data_processed/probability_dynamics_n16.csv and _results.csv of the reference checkout",
        "time_series_lines": expect,
        "records": records,
        "traps": traps,
        "rank5_g030_t_le_2": {"lines": [line for line, _ in e2e],
                              "time": [float(p[4]) for _, p in e2e],
                              "p_less": [float(p[5]) for _, p in e2e],
                              "dh": [float(p[6]) for _, p in e2e],
                              "p_blind": float(e2e[0][1][7])},
    }, open(OUT, "w"), indent=1)
    print("wrote", OUT, len(expect), "lines")


if __name__ == "__main__":
    main()
