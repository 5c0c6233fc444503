"""GPU parity of the adjoint gradient (bk_adjoint_forward / bk_adjoint_backward)
against the oracle's FP64 reverse mode (bko_adjoint_grad, itself pinned by
central differences and a closed form in test_oracle.py), against central
differences of the GPU complex128 forward, and across checkpoint spacings."""
from __future__ import annotations

import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(n_seg, seed, spread=0.05, base=0.35):
    rng = np.random.default_rng(seed)
    return base + spread * rng.uniform(-1.0, 1.0, n_seg)


# n: 8 runs the per-qubit path, 12/13+ the tiled sweeps (c128 LOW tile = 12 bits,
# c64 = 13), 15-16 add the high-qubit sweeps
@pytest.mark.parametrize("n,order,w", [
    (8, 2, (1.0, 0.0)), (8, 1, (0.0, 1.0)),
    (12, 2, (1.0, 0.0)), (13, 1, (0.5, 0.25)),
    (15, 2, (1.0, 0.0)), (16, 1, (0.0, 1.0)), (16, 2, (0.7, -0.2)),
])
def test_adjoint_c128_matches_oracle(bk, orc, engine, n, order, w):
    _, s0, e0 = bk.make_trap(n, 3)
    gs = np.repeat(_fields(6, n), 8)  # 6 segments x 8 steps
    dt = 0.05
    ref, (pl, dh) = orc.adjoint_grad(n, s0, gs, dt, order=order, e0=e0, wp=w[0], wd=w[1])
    obs = engine.adjoint_forward(n, s0, gs, dt, order=order, precision=bk.BK_C128, e0=e0,
                                 checkpoint_every=10)
    assert obs[0] == pytest.approx(pl, abs=1e-12)
    assert obs[1] == pytest.approx(dh, abs=1e-11)
    assert obs[2] == pytest.approx(1.0, abs=1e-12)
    grad, drift = engine.adjoint_backward(*w)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(grad - ref).max() <= 1e-10 * scale, (np.abs(grad - ref).max(), scale)
    assert drift < 1e-20


@pytest.mark.parametrize("n,order", [(14, 2), (16, 1), (18, 2)])
def test_adjoint_c64_matches_oracle(bk, orc, engine, n, order):
    """complex64 state, FP32 butterflies, FP64 reductions: within 5e-5 of the
    largest gradient component (FP32 rounding accumulates along the reverse
    pass; measured 2e-5 at n = 18 over 80 steps)."""
    _, s0, e0 = bk.make_trap(n, 5)
    gs = np.repeat(_fields(8, 100 + n), 10)
    dt = 0.02
    ref, (pl, _) = orc.adjoint_grad(n, s0, gs, dt, order=order, e0=e0, wp=1.0, wd=0.0)
    obs = engine.adjoint_forward(n, s0, gs, dt, order=order, precision=bk.BK_C64, e0=e0,
                                 checkpoint_every=16)
    assert obs[0] == pytest.approx(pl, abs=1e-6)
    grad, drift = engine.adjoint_backward(1.0, 0.0)
    scale = np.abs(ref).max()
    assert np.abs(grad - ref).max() <= 5e-5 * scale + 1e-9, (np.abs(grad - ref).max(), scale)
    assert drift < 1e-9


def test_adjoint_checkpoint_spacing_invariance(bk, engine):
    """The gradient does not depend on where the tape resets the reversed state
    (complex128: reversal is exact to rounding), and without checkpoints the
    result is the same too."""
    n = 14
    _, s0, e0 = bk.make_trap(n, 9)
    gs = np.repeat(_fields(5, 4), 12)
    out = []
    for K in (1, 7, 25, 60, 1000):
        engine.adjoint_forward(n, s0, gs, 0.04, order=bk.BK_STRANG, precision=bk.BK_C128, e0=e0,
                               checkpoint_every=K)
        g, drift = engine.adjoint_backward(1.0, 0.3)
        assert drift < 1e-20
        out.append(g)
    for g in out[1:]:
        assert np.abs(g - out[0]).max() <= 1e-12 * max(1.0, np.abs(out[0]).max())


def test_adjoint_matches_fd_of_gpu_forward(bk, engine):
    """c64 adjoint against central differences (h = 1e-4) of complex128 GPU
    forward runs, per segment, to 1e-4."""
    n, sps, dt = 16, 20, 0.02
    _, s0, e0 = bk.make_trap(n, 2)
    g = _fields(8, 21)
    engine.adjoint_forward(n, s0, np.repeat(g, sps), dt, order=bk.BK_STRANG, precision=bk.BK_C64,
                           e0=e0, checkpoint_every=32)
    grad, _ = engine.adjoint_backward(1.0, 0.0)
    gseg = grad.reshape(len(g), sps).sum(axis=1)
    h = 1e-4
    for j in (0, 3, 7):
        vals = []
        for sgn in (1, -1):
            gg = g.copy()
            gg[j] += sgn * h
            r = engine.evolve(n, s0, 0.0, dt, len(g) * sps, order=bk.BK_STRANG, precision=bk.BK_C128,
                              e0=e0, g_steps=np.repeat(gg, sps), measure_every=0)
            vals.append(r["p_less"][-1])
        fd = (vals[0] - vals[1]) / (2 * h)
        assert gseg[j] == pytest.approx(fd, abs=1e-4), (j, gseg[j], fd)


@pytest.mark.slow
def test_differentiation_workload_n26_against_fd(bk, engine):
    """The differentiation workload at full size: N=26 complex64, 64 segments x
    20 steps of Strang dt=0.02, g_j = 0.35 +- 0.05 from mt19937_64(1), loss
    -P_<(T), checkpoints every 64 steps.  The adjoint's segment gradient matches
    central differences (h = 1e-4) of complex128 forward runs to 1e-6 on the
    first, a middle and the last segment (the largest component is about 0.03;
    the FD truncation error, h^2 times the third derivative, is near 1e-8)."""
    n, nseg, sps, dt, K = 26, 64, 20, 0.02, 64
    _, s0, e0 = bk.make_trap(n, 1)
    g = bk.field_schedule(1, nseg)
    engine.adjoint_forward(n, s0, np.repeat(g, sps), dt, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0,
                           checkpoint_every=K)
    grad, drift = engine.adjoint_backward(-1.0, 0.0)
    gseg = grad.reshape(nseg, sps).sum(axis=1)
    assert drift < 1e-9
    h = 1e-4
    for j in (0, 31, 63):
        vals = []
        for sgn in (1, -1):
            gg = g.copy()
            gg[j] += sgn * h
            r = engine.evolve(n, s0, 0.0, dt, nseg * sps, order=bk.BK_STRANG, precision=bk.BK_C128, e0=e0,
                              g_steps=np.repeat(gg, sps), measure_every=0)
            vals.append(-r["p_less"][-1])
        fd = (vals[0] - vals[1]) / (2 * h)
        print(f"\n[N=26 segment {j}] adjoint {gseg[j]:+.8e} fd {fd:+.8e}")
        assert gseg[j] == pytest.approx(fd, abs=1e-6)


@pytest.mark.slow
def test_differentiation_workload_n26_c64_against_c128_adjoint(bk, engine):
    """The same workload's gradient on every one of the 64 segments: the
    complex64 adjoint (tan butterflies, weighted gradient terms) against the
    complex128 adjoint (FP64 throughout), and the forward observables.  The
    complex64 trajectory itself carries a relative error near 1e-6 (P_<(T) is
    off by 1.2e-8 of 0.0144), and the bilinear gradient inherits it: measured
    2.2e-6 of the largest component (6.5e-8 absolute), checked at 5e-6."""
    n, nseg, sps, dt, K = 26, 64, 20, 0.02, 64
    _, s0, e0 = bk.make_trap(n, 1)
    g = np.repeat(bk.field_schedule(1, nseg), sps)
    out = {}
    for prec in (bk.BK_C64, bk.BK_C128):
        obs = engine.adjoint_forward(n, s0, g, dt, order=bk.BK_STRANG, precision=prec, e0=e0, checkpoint_every=K)
        grad, drift = engine.adjoint_backward(-1.0, 0.0)
        out[prec] = (np.asarray(obs), grad.reshape(nseg, sps).sum(axis=1), drift)
    (o64, g64, d64), (o128, g128, d128) = out[bk.BK_C64], out[bk.BK_C128]
    scale = np.abs(g128).max()
    err = np.abs(g64 - g128).max()
    print(f"\n[N=26 adjoint c64 vs c128] max|dg| {err:.3e} of {scale:.3e} ({err / scale:.2e}); "
          f"P_< {o64[0]:.10f} / {o128[0]:.10f}; drift {d64:.1e} / {d128:.1e}")
    assert err <= 5e-6 * scale
    assert abs(o64[0] - o128[0]) <= 1e-6
    assert d64 < 1e-9 and d128 < 1e-20


@pytest.mark.slow
def test_differentiation_workload_n26_adam(bk, engine):
    """The optimisation leg of the workload: 50 Adam steps (lr 0.005) on the 64
    segment fields of N=26, -P_<(T) as the loss, complex64 with checkpoints every
    64 steps.  P_<(T) rises and every field stays finite."""
    torch = pytest.importorskip("torch")  # noqa: F841
    ag = importlib.import_module("bernasconi-krylov_amd.autograd")
    n, nseg, sps, dt = 26, 64, 20, 0.02
    _, s0, e0 = bk.make_trap(n, 1)
    g0 = bk.field_schedule(1, nseg)
    print("N=26 Adam: 50 iterations", flush=True)
    res = ag.optimize_schedule(engine, n, s0, g0, dt, sps, iters=50, lr=0.005, e0=e0, checkpoint_every=64)
    loss = np.asarray(res["loss"])
    print(f"\n[N=26 Adam] P_<(T) {-loss[0]:.5f} -> {-loss[-1]:.5f}; fields in "
          f"[{np.min(res['g']):.3f}, {np.max(res['g']):.3f}]")
    assert len(loss) == 50
    assert np.all(np.isfinite(loss)) and np.all(np.isfinite(res["g"]))
    assert loss[-1] < loss[0] - 0.01


@pytest.mark.parametrize("n,prec", [(17, "c128"), (22, "c128"), (23, "c64"), (24, "c64"), (25, "c64")])
def test_fused_reverse_matches_step_by_step(bk, n, prec):
    """Fused reverse runs (the forward run's schedule on both states; the
    gradient terms of the rotations after a phase belong to the next step;
    runs end at the checkpoints) against the step-by-step reverse pass, which
    the oracle pins above.  One high group: n = 17 (c128), 23 (c64); two: the
    rest.  50 steps with checkpoints every 16 give runs of 2, 16, 16, 16 steps.
    complex128: 1e-10 of the largest component; complex64: 1e-4 (the fused
    runs use the tan butterfly with c^2m-weighted gradient terms, the
    step-by-step reverse (c, s); each is within 5e-5 of FP64).  A second
    backward over the same tape must repeat the gradient bit for bit."""
    precision = bk.BK_C128 if prec == "c128" else bk.BK_C64
    _, s0, e0 = bk.make_trap(n, 7)
    gs = np.repeat(_fields(5, 40 + n), 10)
    out = {}
    for fused in (0, 1):
        eng = bk.Engine(0)
        try:
            eng.set_tuning("fused", fused)
            obs = eng.adjoint_forward(n, s0, gs, 0.02, order=bk.BK_STRANG, precision=precision, e0=e0,
                                      checkpoint_every=16)
            grad, drift = eng.adjoint_backward(1.0, 0.25)
            # the reverse phase table's rounding carry restarts with each backward
            again, _ = eng.adjoint_backward(1.0, 0.25)
            assert np.array_equal(grad, again)
        finally:
            eng.close()
        out[fused] = (np.asarray(obs), grad, drift)
    (o0, g0, d0), (o1, g1, d1) = out[0], out[1]
    scale = np.abs(g0).max()
    tol = 1e-10 if prec == "c128" else 1e-4
    print(f"\n[fused reverse n={n} {prec}] max|dg| {np.abs(g1 - g0).max():.3e} of {scale:.3e}, "
          f"drift {d0:.2e} / {d1:.2e}")
    # P_<, <d_H> (O(n)), norm of the forward runs
    otol = np.array([1e-12, 1e-11, 1e-12]) if prec == "c128" else np.array([1e-6, 1e-5, 1e-6])
    assert np.all(np.abs(o1 - o0) <= otol), np.abs(o1 - o0)
    assert np.abs(g1 - g0).max() <= tol * scale + 1e-12
    assert d1 < (1e-20 if prec == "c128" else 1e-9)


def test_narrow_phase_group_reverse(bk):
    """n = 26 complex64 (the differentiation workload's size): the adjoint with
    the phase-carrying group on 2^13-amplitude tiles (the default) against
    2^14-amplitude tiles.  The forward observables agree to FP64 summation
    order, the states bit for bit (so the drift is the same), and the
    gradients to 1e-6 of the largest component (per-thread FP32 partials of
    the pair terms group differently on the two tilings)."""
    n = 26
    _, s0, e0 = bk.make_trap(n, 7)
    gs = np.repeat(_fields(4, 91), 10)
    out = {}
    for narrow in (0, 1):
        eng = bk.Engine(0)
        try:
            eng.set_tuning("narrow_phase", narrow)
            obs = eng.adjoint_forward(n, s0, gs, 0.02, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0,
                                      checkpoint_every=16)
            grad, drift = eng.adjoint_backward(1.0, 0.25)
        finally:
            eng.close()
        out[narrow] = (np.asarray(obs), grad, drift)
    (o0, g0, d0), (o1, g1, d1) = out[0], out[1]
    assert np.abs(o1 - o0).max() <= 1e-10
    assert d0 == d1
    assert np.abs(g1 - g0).max() <= 1e-6 * np.abs(g0).max()


def test_reverse_low_grid(bk):
    """The reverse run's LOW-tile passes on 512 workgroups (rev_grid_low, the
    n = 26 default) against the default grid at n = 24: every tile is
    transformed the same way, so the drift is the same; the gradient partials
    of the LOW passes sum in another grouping (per-workgroup FP64 partials),
    so the gradients agree to rounding."""
    n = 24
    _, s0, e0 = bk.make_trap(n, 5)
    gs = np.repeat(_fields(4, 93), 10)
    out = {}
    for g in (0, 512):
        eng = bk.Engine(0)
        try:
            eng.set_tuning("rev_grid_low", g)
            eng.adjoint_forward(n, s0, gs, 0.02, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0,
                                checkpoint_every=16)
            out[g] = eng.adjoint_backward(1.0, 0.25)
        finally:
            eng.close()
    (g0, d0), (g1, d1) = out[0], out[512]
    assert d0 == d1
    assert np.abs(g1 - g0).max() <= 1e-10 * np.abs(g0).max()


def test_backward_leaves_state_and_detects_stale_tape(bk, engine):
    n = 13
    _, s0, e0 = bk.make_trap(n, 1)
    gs = np.full(30, 0.4)
    engine.adjoint_forward(n, s0, gs, 0.05, precision=bk.BK_C128, e0=e0, checkpoint_every=8)
    before = engine.download()
    g1, _ = engine.adjoint_backward(1.0, 0.0)
    after = engine.download()
    assert np.array_equal(before, after)
    g2, _ = engine.adjoint_backward(1.0, 0.0)  # the tape can be reused
    assert np.array_equal(g1, g2)
    engine.evolve(n, s0, 0.4, 0.05, 3, precision=bk.BK_C128)
    with pytest.raises(bk.BKError) as ei:
        engine.adjoint_backward(1.0, 0.0)
    assert ei.value.code == -4
    with pytest.raises(bk.BKError) as ei:
        engine.adjoint_forward(n, s0, gs, 0.05, precision=bk.BK_C128, checkpoint_every=0)
    assert ei.value.code == -1


def test_autograd_function_and_adam(bk, engine):
    """torch.autograd.Function: the segment gradient equals the per-step
    gradient summed per segment, the (w_p, w_d) weights come from the incoming
    gradient, and 50 Adam steps on -P_<(T) raise P_<(T)."""
    torch = pytest.importorskip("torch")
    ag = importlib.import_module("bernasconi-krylov_amd.autograd")
    n, sps, dt = 14, 10, 0.02
    _, s0, e0 = bk.make_trap(n, 4)
    g0 = _fields(8, 77)
    g = torch.tensor(g0, requires_grad=True)
    obs = ag.trotter_observables(g, engine, n, s0, dt, sps, precision=bk.BK_C64, e0=e0,
                                 checkpoint_every=16)
    (2.0 * obs[0] - 0.5 * obs[1]).backward()
    engine.adjoint_forward(n, s0, np.repeat(g0, sps), dt, precision=bk.BK_C64, e0=e0, checkpoint_every=16)
    ref, _ = engine.adjoint_backward(2.0, -0.5)
    assert np.allclose(g.grad.numpy(), ref.reshape(8, sps).sum(axis=1), rtol=0, atol=1e-12)
    res = ag.optimize_schedule(engine, n, s0, g0, dt, sps, iters=50, lr=0.02, e0=e0,
                               checkpoint_every=16)
    assert res["loss"][-1] < res["loss"][0]
    assert np.all(np.isfinite(res["g"]))


def test_backward_with_a_full_phase_table_cache(bk):
    """A context keeps at most 64 phase tables and trims them only when an entry
    point starts.  A download between forward and backward makes the backward's
    two table lookups differ, so with the cache filled up to the bound the
    second lookup would overflow it inside the call: the gradient must still
    equal a fresh context's, bit for bit."""
    n, dt = 12, 0.05
    _, s0, e0 = bk.make_trap(n, 3)
    gs = np.full(24, 0.4)

    def grad_with(fill):
        eng = bk.Engine(0)
        try:
            for i in range(fill):  # each download flushes the opening phase of its own dt: one table each
                eng.evolve(n, s0, 0.3, 0.01 + 1e-3 * i, 1, order=bk.BK_STRANG, precision=bk.BK_C128)
                eng.download()
            eng.adjoint_forward(n, s0, gs, dt, precision=bk.BK_C128, e0=e0, checkpoint_every=8)
            eng.download()
            return eng.adjoint_backward(1.0, 0.5)[0]
        finally:
            eng.close()

    ref = grad_with(0)
    assert np.abs(ref).max() > 0
    for fill in range(60, 68):
        print(f"phase-table cache filled with {fill} tables", flush=True)
        assert np.array_equal(grad_with(fill), ref), fill
