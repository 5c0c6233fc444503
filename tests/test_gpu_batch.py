"""GPU tests of the batched evolution (bk_batch_run): every instance must come
out bit-identical to a bk_evolve of that instance alone (same kernels, same
arithmetic), with and without graph capture, cache blocking or tuning."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _instances(bk, ns, gs, seed0=40):
    out = []
    for i, (n, g) in enumerate(zip(ns, gs)):
        _, s0, e0 = bk.make_trap(n, seed0 + i)
        out.append((n, s0, e0, g))
    return out


def _serial(bk, engine, inst, dt, steps, order, prec):
    rows = []
    for n, s0, e0, g in inst:
        r = engine.evolve(n, s0, g, dt, steps, order=order, precision=prec, e0=e0, measure_every=0)
        rows.append([r["p_less"][-1], r["dh"][-1], r["norm"][-1]])
    return np.array(rows)


@pytest.mark.parametrize("prec,ns", [("c64", [13, 14, 15, 16, 17, 16, 14, 13, 18]),
                                     ("c128", [12, 13, 14, 15, 13, 12]),
                                     ("c64", [19, 21, 23, 20, 22]),  # one group, upper sizes
                                     ("c64", [24, 25, 24]),    # two high-qubit groups
                                     ("c128", [22, 23])])
@pytest.mark.parametrize("order", ["strang", "lie"])
def test_batch_bit_identical_to_serial(bk, engine, orc, prec, ns, order):
    precision = bk.BK_C64 if prec == "c64" else bk.BK_C128
    o = bk.BK_STRANG if order == "strang" else bk.BK_LIE
    gs = [0.1 + 0.07 * i for i in range(len(ns))]
    inst = _instances(bk, ns, gs)
    ref = _serial(bk, engine, inst, 0.02, 37, o, precision)
    for graph in (False, True):
        got, st = engine.batch_run(inst, 0.02, 37, order=o, precision=precision, use_graph=graph)
        assert np.array_equal(got, ref), (graph, np.abs(got - ref).max())
        assert st["launches"] > 0 and st["device_ms"] > 0
    # and one instance of the mix against the FP64 oracle
    i = int(np.argmin(ns))
    n, s0, e0, g = inst[i]
    fp = orc.trotter(n, s0, g, 0.02, 37, order=2 if o == bk.BK_STRANG else 1, e0=e0)
    tol = (1e-12, 1e-11) if prec == "c128" else (1e-5, 1e-4)
    assert abs(got[i, 0] - fp["p_less"][-1]) <= tol[0]
    assert abs(got[i, 1] - fp["dh"][-1]) <= tol[1]


def test_batch_mixed_butterfly_forms(bk, engine):
    """Fields large enough for |cos(g dt)| < 0.5 take the (c, s) butterfly; they
    form their own launch group and still match serial runs exactly."""
    ns = [14, 14, 15, 15, 16]
    gs = [0.35, 70.0, 0.4, 80.0, 65.0]  # g dt = 1.4, 1.6, 1.3 -> (c, s) form
    inst = _instances(bk, ns, gs, seed0=7)
    ref = _serial(bk, engine, inst, 0.02, 25, bk.BK_STRANG, bk.BK_C64)
    got, _ = engine.batch_run(inst, 0.02, 25, order=bk.BK_STRANG, precision=bk.BK_C64)
    assert np.array_equal(got, ref)


def test_batch_tuned_blocks_identical(bk, engine):
    """The autotuner may pick cache blocks (a block runs all its steps before the
    next): the per-instance arithmetic, and so the results, do not change."""
    ns = [16, 18, 16, 18, 17, 17, 16, 18] * 2
    gs = [0.10 + 0.50 * (i % 16) / 15 for i in range(len(ns))]
    inst = _instances(bk, ns, gs, seed0=90)
    ref, _ = engine.batch_run(inst, 0.02, 30, precision=bk.BK_C64, use_graph=False)
    tuned, st = engine.batch_run(inst, 0.02, 30, precision=bk.BK_C64, use_graph=True, autotune=True)
    assert st["tuned"]
    assert np.array_equal(tuned, ref)
    again, st2 = engine.batch_run(inst, 0.02, 30, precision=bk.BK_C64, use_graph=True, autotune=True)
    assert st2["tuned"] and (st2["grid"], st2["row_bits"], st2["block_bytes"], st2["out_of_place"]) == \
        (st["grid"], st["row_bits"], st["block_bytes"], st["out_of_place"])
    assert np.array_equal(again, ref)


@pytest.mark.parametrize("prec,ns", [("c64", [13, 16, 18, 16, 17, 18, 16, 24]),
                                     ("c128", [12, 15, 16, 15, 22])])
@pytest.mark.parametrize("oop", [False, True])
def test_batch_forced_blocks_bit_identical(bk, engine, prec, ns, oop):
    """Cache blocks chosen by the caller, in place or out of place (a block's
    passes alternate between its states and a second buffer), with blocks of
    one and of several states, a partial last block, and both parities of the
    pass count (so the last pass ends in either buffer): every instance equals
    its serial run bit for bit.  Step-by-step groups (c64 n = 13) stay in place."""
    precision = bk.BK_C64 if prec == "c64" else bk.BK_C128
    gs = [0.1 + 0.06 * i for i in range(len(ns))]
    inst = _instances(bk, ns, gs, seed0=120)
    for steps in (20, 21):
        ref = _serial(bk, engine, inst, 0.02, steps, bk.BK_STRANG, precision)
        for bb in (1 << 20, 3 << 20, 64 << 20):
            for graph in (False, True):
                got, st = engine.batch_run(inst, 0.02, steps, precision=precision, use_graph=graph,
                                           block_bytes=bb, out_of_place=oop)
                assert st["block_bytes"] == bb and st["out_of_place"] == oop and not st["tuned"]
                assert np.array_equal(got, ref), (steps, bb, graph, np.abs(got - ref).max())


def test_batch_rejects_small_n(bk, engine):
    with pytest.raises(bk.BKError) as ei:
        engine.batch_run([(10, 0, -1, 0.3)], 0.02, 5)
    assert ei.value.code == -1



@pytest.mark.slow
def test_batched_workload_512_traps(bk, engine, orc):
    """The batched workload at its stated size: 512 traps, n drawn from
    {16, 18, 20, 22} by mt19937_64(1), g_i = 0.10 + 0.50 (i mod 16) / 15, Strang
    dt = 0.02 x 500 steps, complex64, one autotuned hipGraph launch.  Checks:
    4 instances of every n bit-identical to serial bk_evolve runs; 3 instances
    each of n = 16 and 18 against the FP64 oracle (P_< to 1e-5, <d_H> to 1e-4);
    every n = 20 / 22 instance against a complex128 batched run of the same
    instances (P_< to 1e-5, <d_H> to 1e-4)."""
    dt, steps = 0.02, 500
    work = bk.batch_workload(512, seed=1)
    print("batched workload: 512 traps, 500 steps, graph + autotune", flush=True)
    got, st = engine.batch_run(work, dt, steps, order=bk.BK_STRANG, precision=bk.BK_C64, use_graph=True,
                               autotune=True)
    assert st["tuned"] and got.shape == (512, 3)
    assert np.all(np.abs(got[:, 2] - 1.0) <= 1e-5)
    by_n = {}
    for i, (n, _, _, _) in enumerate(work):
        by_n.setdefault(n, []).append(i)
    assert sorted(by_n) == [16, 18, 20, 22]
    for n, idx in sorted(by_n.items()):
        print(f"  n={n}: {len(idx)} traps; serial check of 4", flush=True)
        for i in idx[:4]:
            _, s0, e0, g = work[i]
            r = engine.evolve(n, s0, g, dt, steps, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0, measure_every=0)
            assert np.array_equal(got[i], [r["p_less"][-1], r["dh"][-1], r["norm"][-1]]), (n, i)
    for n in (16, 18):
        for i in by_n[n][4:7]:
            _, s0, e0, g = work[i]
            print(f"  n={n} instance {i}: FP64 oracle", flush=True)
            fp = orc.trotter(n, s0, g, dt, steps, order=2, e0=e0)
            assert abs(got[i, 0] - fp["p_less"][-1]) <= 1e-5, (n, i)
            assert abs(got[i, 1] - fp["dh"][-1]) <= 1e-4, (n, i)
    big = by_n[20] + by_n[22]
    print(f"  {len(big)} n = 20 / 22 traps in complex128", flush=True)
    ref, _ = engine.batch_run([work[i] for i in big], dt, steps, order=bk.BK_STRANG, precision=bk.BK_C128,
                              use_graph=True)
    assert np.abs(got[big, 0] - ref[:, 0]).max() <= 1e-5
    assert np.abs(got[big, 1] - ref[:, 1]).max() <= 1e-4
    print(f"  max|dP_<| {np.abs(got[big, 0] - ref[:, 0]).max():.2e}, "
          f"max|d<d_H>| {np.abs(got[big, 1] - ref[:, 1]).max():.2e}", flush=True)
