"""Golden series computed by the oracle where the oracle is too slow to run
inside a GPU test (dense exact diagonalisation at N=12 takes about 90 s of
CPU time):

    python tests/golden/make_oracle_fixtures.py

* ed_n12.json : the N=12 plain-method case -- trap from mt19937_64(1) +
  steepest descent, g = 0.35 -- P_<(t) and <d_H>(t) of the exact evolution
  exp(-iHt)|s0> (bko_ed_evolve: dense H, own symmetric eigensolver, spectral
  sum) at t = 0, 0.05, ..., 10 and at t = 40.  bko_ed_evolve itself is pinned
  against scipy's eigh / expm_multiply in test_oracle.py.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402


def ed_n12():
    n, g, seed = 12, 0.35, 1
    start, s0, e0 = O.make_trap(n, seed)
    t = np.concatenate([np.arange(201) * 0.05, [40.0]])
    pl, dh = O.ed_evolve(n, g, s0, t)
    return {"n": n, "g": g, "seed": seed, "start": start, "s0": s0, "e0": e0,
            "t": t.tolist(), "p_less": pl.tolist(), "dh": dh.tolist()}


def main():
    O.build()
    for name, fn in (("ed_n12.json", ed_n12),):
        data = fn()
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print(name, os.path.getsize(os.path.join(HERE, name)), "bytes")


if __name__ == "__main__":
    main()
