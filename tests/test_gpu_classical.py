"""Trap supply at scale on the GPU: batched steepest descent (bk_descent_batch)
bit-for-bit against the oracle's descent (pinned by the reference's seed
fixtures), including the reference's GPU-vs-CPU case (test_steepest_descent.cu
Test 3), evaluation budgets, and the multi-start collector of distinct minima
(bk_collect_minima) against the collector loop of classical.cuh:861-950 run
with the oracle, plus the reference's Test 4 properties."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [5, 12, 16, 24, 32, 45, 64])
def test_descent_batch_matches_oracle(bk, orc, engine, n):
    rng = np.random.default_rng(n)
    mask = (1 << n) - 1
    starts = [int(x) & mask for x in rng.integers(0, 2 ** 63, size=300, dtype=np.uint64)]
    starts += [0, mask]  # all-up and all-down starts
    fin, fe, ev = engine.descent_batch(n, starts)
    for i, s in enumerate(starts):
        fs, e, evals, _ = orc.steepest_descent(n, s)
        assert (int(fin[i]), int(fe[i]), int(ev[i])) == (fs, e, evals), (n, hex(s))


def test_reference_gpu_cpu_equivalence_case(bk, orc, engine):
    """test_steepest_descent.cu Test 3: n=16, ClassicalOptimizer seed 9999, 64
    starts from random_state() (mt19937_64 draws masked to n bits)."""
    n = 16
    g = orc.MT64(9999)
    starts = [g.next() & 0xFFFF for _ in range(64)]
    fin, fe, ev = engine.descent_batch(n, starts)
    for i, s in enumerate(starts):
        assert (int(fin[i]), int(fe[i]), int(ev[i])) == orc.steepest_descent(n, s)[:3]


def test_descent_budget(bk, orc, engine):
    """A pass runs only while 1 + n * passes stays within the budget; a descent
    that stops early ends above (or at) its unbudgeted minimum."""
    n, budget = 20, 61  # at most 3 passes of 20
    rng = np.random.default_rng(3)
    starts = [int(x) & ((1 << n) - 1) for x in rng.integers(0, 2 ** 62, size=200, dtype=np.uint64)]
    fin, fe, ev = engine.descent_batch(n, starts, max_evals=budget)
    for i, s in enumerate(starts):
        assert ev[i] <= budget and (ev[i] - 1) % n == 0
        assert fe[i] == bk.energy(n, int(fin[i]))
        fs, e, evals, _ = orc.steepest_descent(n, s)
        if evals <= budget:
            assert (int(fin[i]), int(fe[i])) == (fs, e)
        else:
            assert fe[i] >= e and ev[i] == 61


def _collect_reference_loop(orc, n, seed, target, batch, max_starts):
    g = orc.MT64(seed)
    mask = (1 << n) - 1
    fams, order, evals, started, stagnant = {}, [], 0, 0, 0
    while len(fams) < target and started < max_starts:
        starts = [g.next() & mask for _ in range(batch)]
        started += batch
        fresh = 0
        for s in starts:
            fs, e, ev, _ = orc.steepest_descent(n, s)
            evals += ev
            c = orc.canonical(n, fs)
            if c not in fams:
                fams[c] = {"canonical_state": c, "sample_state": fs, "energy": e, "times_visited": 1,
                           "evals_to_first_find": evals}
                fresh += 1
                if len(fams) >= target:
                    break
            else:
                fams[c]["times_visited"] += 1
        stagnant = stagnant + 1 if fresh == 0 else 0
        if stagnant >= 4:
            break
    out = sorted(fams.values(), key=lambda f: (f["energy"], -f["times_visited"], f["canonical_state"]))
    for i, f in enumerate(out):
        f["id"] = i + 1
    return out, started, evals


@pytest.mark.parametrize("n,seed,target,batch", [(14, 5, 12, 64), (20, 11, 40, 256), (10, 2026, 50, 128)])
def test_collect_minima_matches_reference_loop(bk, orc, engine, n, seed, target, batch):
    got = engine.collect_minima(n, seed, target, batch_size=batch, max_starts=20000)
    ref, started, evals = _collect_reference_loop(orc, n, seed, target, batch, 20000)
    assert got["families"] == ref
    assert (got["total_starts"], got["total_evals"]) == (started, evals)


def test_reference_distinct_minima_properties(bk, orc, engine):
    """test_steepest_descent.cu Test 4: n=12, seed 777, 5 families: unique
    canonical states that are their own canonical form, energies ascending."""
    got = engine.collect_minima(12, 777, 5, batch_size=4096, max_starts=10000)["families"]
    assert len(got) == 5
    canon = [f["canonical_state"] for f in got]
    assert len(set(canon)) == 5 and all(bk.canonical(12, c) == c for c in canon)
    assert all(got[i]["energy"] >= got[i - 1]["energy"] for i in range(1, 5))
    for f in got:
        assert bk.canonical(12, f["sample_state"]) == f["canonical_state"]
        assert f["energy"] == bk.energy(12, f["sample_state"])
