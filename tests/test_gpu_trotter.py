"""GPU parity of the Trotter hot path (LOW/HIGH sweeps, fused phase and
reductions) against the FP64 oracle and the reference's golden trajectories.

Tolerances: complex128 runs match the FP64 oracle to 1e-12 (P_<, norm) /
1e-10 (<d_H>, which is O(N)); complex64 runs match to 1e-5 in P_< (the stated
target), 1e-4 in <d_H>, and 1 - |<psi64|psi128>| <= 1e-5."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C64, C128, LIE, STRANG = 1, 2, 1, 2

CASES = [
    # n, precision, order, g, dt, steps, seed
    (4, C128, LIE, 0.5, 0.05, 40, 1),
    (10, C128, LIE, 0.35, 0.05, 100, 2),
    (12, C128, STRANG, 0.35, 0.02, 100, 3),   # c128 LOW sweep only
    (13, C64, LIE, 0.35, 0.05, 100, 4),       # c64 LOW sweep only
    (16, C128, LIE, 0.3, 0.05, 80, 5),        # LOW + 1 HIGH
    (16, C64, STRANG, 0.3, 0.02, 80, 6),
    (20, C128, STRANG, 0.35, 0.02, 40, 7),
    (21, C64, LIE, 0.45, 0.05, 40, 8),
    (22, C128, LIE, 0.35, 0.05, 20, 9),       # LOW + 2 HIGH (c128)
    (23, C64, STRANG, 0.35, 0.02, 20, 10),    # LOW + 2 HIGH (c64)
]


def _tol(prec):
    return (1e-12, 1e-10) if prec == C128 else (1e-5, 1e-4)


@pytest.mark.parametrize("n,prec,order,g,dt,steps,seed", CASES)
def test_trotter_parity_from_trap(engine, orc, bk, n, prec, order, g, dt, steps, seed):
    print(f"case n={n} prec={prec} order={order} steps={steps}", flush=True)
    _, s0, e0 = orc.make_trap(n, seed)
    ref = orc.trotter(n, s0, g, dt, steps, order=order, want_state=True)
    got = engine.evolve(n, s0, g, dt, steps, order=order, precision=prec)
    tp, td = _tol(prec)
    assert got["e0"] == e0
    assert np.max(np.abs(got["p_less"] - ref["p_less"])) <= tp
    assert np.max(np.abs(got["dh"] - ref["dh"])) <= td
    assert np.max(np.abs(got["norm"] - ref["norm"])) <= (1e-12 if prec == C128 else 1e-5)
    psi = engine.download().astype(np.complex128)
    if prec == C128:
        assert np.max(np.abs(psi - ref["psi"])) <= 1e-11
    else:
        ov = abs(np.vdot(ref["psi"], psi))
        assert 1.0 - ov <= 1e-5
    assert got["p_blind"] == pytest.approx(orc.count_below(n, e0) / 2.0 ** n, rel=0, abs=0)


@pytest.mark.parametrize("n,prec", [(11, C128), (14, C128), (15, C64), (19, C64), (24, C64)])
def test_trotter_parity_generic_state(engine, orc, n, prec):
    """Random input state: exercises every phase/mixer entry, not just a basis state."""
    print(f"generic n={n} prec={prec}", flush=True)
    rng = np.random.default_rng(n)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    if prec == C64:
        psi = psi.astype(np.complex64).astype(np.complex128)
    s0, e0, g, dt, steps = 5, orc.energy(n, 5), 0.41, 0.03, 12
    engine.upload(psi.astype(np.complex64 if prec == C64 else np.complex128), prec)
    got = engine.evolve(n, s0, g, dt, steps, order=STRANG, precision=prec, init=False)
    ref = orc.trotter(n, s0, g, dt, steps, order=2, psi=psi, want_state=True)
    tp, td = _tol(prec)
    assert np.max(np.abs(got["p_less"] - ref["p_less"])) <= tp
    assert np.max(np.abs(got["dh"] - ref["dh"])) <= td
    out = engine.download().astype(np.complex128)
    tol = 1e-11 if prec == C128 else 3e-6
    assert np.max(np.abs(out - ref["psi"])) <= tol


def test_field_schedule_and_large_angle(engine, orc):
    """Piecewise g(t) and |cos(g dt)| < 0.5 (full cos/sin butterfly form)."""
    n, steps = 17, 24
    _, s0, e0 = orc.make_trap(n, 12)
    g_steps = np.repeat([0.3, 0.5, 1.4, 0.2], steps // 4)
    for prec in (C128, C64):
        got = engine.evolve(n, s0, 0.0, 1.0, steps, order=LIE, precision=prec, g_steps=g_steps)
        ref = orc.trotter(n, s0, 0.0, 1.0, steps, order=1, g_steps=g_steps)
        tp, td = _tol(prec)
        assert np.max(np.abs(got["p_less"] - ref["p_less"])) <= tp * 10
        assert np.max(np.abs(got["dh"] - ref["dh"])) <= td * 10


def test_measure_every_and_continue(engine, orc):
    n, g, dt = 18, 0.35, 0.02
    _, s0, e0 = orc.make_trap(n, 13)
    ref = orc.trotter(n, s0, g, dt, 30, order=2)
    a = engine.evolve(n, s0, g, dt, 10, order=STRANG, precision=C128, measure_every=5)
    assert len(a["p_less"]) == 3
    assert np.allclose(a["p_less"], ref["p_less"][[0, 5, 10]], atol=1e-13, rtol=0)
    b = engine.evolve(n, s0, g, dt, 20, order=STRANG, precision=C128, init=False,
                      measure_every=0)
    assert len(b["p_less"]) == 1
    assert abs(b["p_less"][0] - ref["p_less"][30]) <= 1e-12
    assert abs(b["dh"][0] - ref["dh"][30]) <= 1e-10


def _golden(golden, n):
    return [t for t in golden("trajectories.json") if t["n"] == n]


@pytest.mark.parametrize("n", [16, 20, 24, 26, 28, 30])
def test_reference_golden_trajectories_c128(engine, golden, n):
    """data_processed/probability_dynamics_n{N}.csv (reference GPU runs, c128 Lie,
    dt=0.05, 1200 steps) reproduced to their printed precision."""
    for tr in _golden(golden, n):
        print(f"golden n={n} rank={tr['trap_rank']} g={tr['g']}", flush=True)
        s0 = int(tr["trap_bits"], 16)
        got = engine.evolve(n, s0, tr["g"], tr["dt"], tr["steps_total"], order=LIE,
                            precision=C128)
        idx = np.array(tr["step"])
        pl = np.array([float(x) for x in tr["p_less"]])
        dh = np.array([float(x) for x in tr["dh"]])
        big = pl > 1e-25
        assert np.all(np.abs(got["p_less"][idx][big] - pl[big]) <= 6e-7 * pl[big] + 1e-30)
        assert np.max(np.abs(got["dh"][idx] - dh)) <= 5.1e-5
        assert abs(got["p_blind"] - float(tr["p_blind"])) <= 5e-7 * float(tr["p_blind"])
        assert got["e0"] == tr["trap_energy"]


def test_c64_against_c128_n24(engine, bk):
    """Config-1 shape: N=24 trap with d_min >= 6, Strang dt=0.02; c64 vs c128."""
    seed, start, s0, e0, dm = bk.make_trap_min_dmin(24, 1, 6)
    steps = 200
    a = engine.evolve(24, s0, 0.35, 0.02, steps, order=STRANG, precision=C128, measure_every=10)
    psi128 = engine.download()
    b = engine.evolve(24, s0, 0.35, 0.02, steps, order=STRANG, precision=C64, measure_every=10)
    psi64 = engine.download().astype(np.complex128)
    assert np.max(np.abs(a["p_less"] - b["p_less"])) <= 1e-5
    assert np.max(np.abs(a["dh"] - b["dh"])) <= 1e-4
    assert 1.0 - abs(np.vdot(psi128, psi64)) <= 1e-5


# n mod 8 covers every residue c_n = (n / 2) mod 4 of E(s) mod 4
@pytest.mark.parametrize("n,prec", [(14, C128), (16, C64), (18, C128), (20, C64), (22, C128)])
@pytest.mark.parametrize("shift", [0, 1, 2])
def test_energy_modes_bit_identical(engine, orc, bk, n, prec, shift):
    """The energy-code cache (12-bit E >> 2) and on-the-fly popcount energies
    feed the same phases and the same P_< test to the same kernels: the series
    must agree bit for bit.  On-the-fly energies run the three-pass step
    schedule, so the cached run takes it too.  shift moves the P_< threshold
    off the lattice E = 4m + c_n that every energy lies on (e0 + shift), where
    the cached test m < ceil((e0 - c_n) / 4) must still equal E < e0."""
    _, s0, e0 = orc.make_trap(n, 31)
    e0 += shift
    engine.set_tuning("fused", 0)
    try:
        engine.set_energy_mode(bk.BK_ENERGY_ONTHEFLY)
        a = engine.evolve(n, s0, 0.35, 0.02, 15, order=STRANG, precision=prec, e0=e0)
        psi_a = engine.download()
        engine.set_energy_mode(bk.BK_ENERGY_CACHE16)
        b = engine.evolve(n, s0, 0.35, 0.02, 15, order=STRANG, precision=prec, e0=e0)
        psi_b = engine.download()
    finally:
        engine.set_energy_mode(bk.BK_ENERGY_CACHE16)
        engine.set_tuning("fused", 1)
    assert np.array_equal(a["p_less"], b["p_less"]) and np.array_equal(a["dh"], b["dh"])
    assert np.array_equal(psi_a, psi_b)
    assert a["p_less"][-1] > 0.0 or shift == 0


@pytest.mark.parametrize("n", [24, 26, 28])
@pytest.mark.parametrize("shift", [1, 3])
def test_fused_threshold_off_lattice(engine, orc, bk, n, shift):
    """The fused schedule reads both energy-code caches (the LOW tile's and the
    phase-carrying high group's); with a P_< threshold off the energy lattice
    its series matches the step-by-step schedule on on-the-fly energies in
    complex128 (two schedules: to 1e-12, as for on-lattice thresholds)."""
    _, s0, e0 = orc.make_trap(n, 5)
    e0 += shift
    a = engine.evolve(n, s0, 0.35, 0.02, 12, order=STRANG, precision=C128, e0=e0)
    engine.set_tuning("fused", 0)
    try:
        engine.set_energy_mode(bk.BK_ENERGY_ONTHEFLY)
        b = engine.evolve(n, s0, 0.35, 0.02, 12, order=STRANG, precision=C128, e0=e0)
    finally:
        engine.set_energy_mode(bk.BK_ENERGY_CACHE16)
        engine.set_tuning("fused", 1)
    assert np.max(np.abs(a["p_less"] - b["p_less"])) <= 1e-12
    assert np.max(np.abs(a["dh"] - b["dh"])) <= 1e-11
    assert a["p_less"][-1] > 0.0


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("n,prec", [(17, C128), (22, C64), (25, C64)])
def test_pingpong_and_grid_bit_identical(engine, orc, bk, n, prec, fused):
    """Out-of-place sweeps and other persistent grid sizes change only where
    data lives, not the arithmetic: results must match the in-place run exactly.
    Out-of-place sweeps exist in the step schedule only."""
    _, s0, e0 = orc.make_trap(n, 41)
    engine.set_tuning("fused", fused)
    try:
        base = engine.evolve(n, s0, 0.35, 0.02, 9, order=STRANG, precision=prec)
        psi0 = engine.download()
        knobs = (("grid", 333), ("grid", 100000)) if fused else (("pingpong", 1), ("grid", 333), ("grid", 100000))
        for key, val in knobs:
            engine.set_tuning(key, val)
            got = engine.evolve(n, s0, 0.35, 0.02, 9, order=STRANG, precision=prec)
            psi = engine.download()
            assert np.array_equal(psi, psi0), (key, val)
            # the FP64 reduction order follows the grid; values agree to rounding
            assert np.max(np.abs(got["p_less"] - base["p_less"])) <= 1e-15
            assert np.max(np.abs(got["dh"] - base["dh"])) <= 1e-13
    finally:
        engine.set_tuning("pingpong", 0)
        engine.set_tuning("grid", 2048)
        engine.set_tuning("fused", 1)


@pytest.mark.parametrize("n", [14, 17, 19, 21, 22, 23, 25, 27, 28, 29, 30, 31])
def test_fused_run_matches_step_by_step(bk, n):
    """Fused runs carry each step's phase inside a pass that also opens the next
    step on its qubits: one high group takes K+1 passes, two take 2K+1, against
    3K (or 2K) for the step schedule.  Complex128: equal to the step-by-step run
    to 1e-12 in every recorded P_<, <d_H>, norm and to 1e-13 in the final
    state; complex64: within the c64 error of complex128."""
    _, s0, e0 = bk.make_trap(n, 5)
    steps, dt, g = 9, 0.02, 0.35
    out = {}
    for prec in (bk.BK_C128, bk.BK_C64):
        for fused in (0, 1):
            eng = bk.Engine(0)
            try:
                eng.set_tuning("fused", fused)
                r = eng.evolve(n, s0, g, dt, steps, order=bk.BK_STRANG, precision=prec, e0=e0, measure_every=2)
                psi = eng.download() if prec == bk.BK_C128 and n <= 28 else None
            finally:
                eng.close()
            out[(prec, fused)] = (r, psi)
    a, b = out[(bk.BK_C128, 0)][0], out[(bk.BK_C128, 1)][0]
    for key in ("p_less", "dh", "norm"):
        assert a[key].shape == b[key].shape
        assert np.abs(a[key] - b[key]).max() <= 1e-12, key
    if n <= 28:
        pa, pb = out[(bk.BK_C128, 0)][1], out[(bk.BK_C128, 1)][1]
        assert np.abs(pa - pb).max() <= 1e-13
    c = out[(bk.BK_C64, 1)][0]
    assert np.abs(c["p_less"] - a["p_less"]).max() <= 1e-6
    assert np.abs(c["dh"] - a["dh"]).max() <= 1e-5


@pytest.mark.parametrize("n", [26, 27, 28])
def test_narrow_phase_group_bit_identical(bk, n):
    """The phase-carrying high group of a complex64 run on 2^13-amplitude tiles
    (the default at n = 26..28) against 2^14-amplitude tiles: every butterfly,
    phase product and store is the same operation on the same operands, so
    the state after the run is equal bit for bit; the recorded observables
    differ only by the order of their FP64 partial sums."""
    _, s0, e0 = bk.make_trap(n, 3)
    out = {}
    for narrow in (0, 1):
        eng = bk.Engine(0)
        try:
            eng.set_tuning("narrow_phase", narrow)
            r = eng.evolve(n, s0, 0.35, 0.02, 12, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0,
                           measure_every=1)
            out[narrow] = (r, eng.download())
        finally:
            eng.close()
    (ra, pa), (rb, pb) = out[0], out[1]
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    for key, tol in (("p_less", 1e-12), ("dh", 1e-10), ("norm", 1e-12)):
        assert np.abs(ra[key] - rb[key]).max() <= tol, key


@pytest.mark.parametrize("n,prec", [(22, "c64"), (26, "c64"), (28, "c64"), (24, "c128"), (30, "c128")])
def test_lds_phase_table_bit_identical(bk, n, prec):
    """Phase passes on 2^13-amplitude tiles with the phase table in LDS
    (complex64 LOW and narrow passes, default at n = 26..27; complex128 high
    groups, default from n = 22, only the first 2032 entries beyond n = 29)
    against the same passes reading it from
    global memory: the same entries multiply the same amplitudes, so the
    state and every observable agree bit for bit."""
    _, s0, e0 = bk.make_trap(n, 13)
    out = {}
    for lds in (1, 0):
        eng = bk.Engine(0)
        try:
            eng.set_tuning("lds_tab13", lds)
            r = eng.evolve(n, s0, 0.35, 0.02, 9 if n < 30 else 4, order=bk.BK_STRANG,
                           precision=bk.BK_C64 if prec == "c64" else bk.BK_C128, e0=e0, measure_every=1)
            # beyond n = 28 the states are compared through their observables
            out[lds] = (r, eng.download() if n <= 28 else None)
        finally:
            eng.close()
    (ra, pa), (rb, pb) = out[0], out[1]
    if pa is not None:
        assert np.array_equal(pa.view(np.uint64 if prec == "c128" else np.uint32),
                              pb.view(np.uint64 if prec == "c128" else np.uint32))
    for key in ("p_less", "dh", "norm"):
        assert np.array_equal(ra[key], rb[key]), key


@pytest.mark.parametrize("n,prec", [(22, "c64"), (26, "c64"), (24, "c128")])
def test_folded_finish_bit_identical(bk, n, prec):
    """A fused run finishes each measured pass's partials in the first
    workgroup of the run's next pass (and k_finish3 only after its last pass);
    the same arithmetic as a k_finish3 kernel after every measured pass, so
    every recorded observable agrees bit for bit (one group at n = 22, two at
    26 and 24)."""
    precision = bk.BK_C64 if prec == "c64" else bk.BK_C128
    _, s0, e0 = bk.make_trap(n, 11)
    out = {}
    for fold in (1, 0):
        eng = bk.Engine(0)
        try:
            eng.set_tuning("fold_finish", fold)
            out[fold] = eng.evolve(n, s0, 0.3, 0.02, 15, order=bk.BK_STRANG, precision=precision, e0=e0,
                                   measure_every=1)
        finally:
            eng.close()
    for key in ("p_less", "dh", "norm"):
        assert np.array_equal(out[0][key], out[1][key]), key


@pytest.mark.parametrize("n", [24, 27])
def test_nontemporal_plain_pass_bit_identical(bk, n):
    """The complex64 plain high-group pass with nontemporal amplitude loads
    (default from n = 27) against cached loads: only the cache policy differs,
    so the state and every observable agree bit for bit."""
    _, s0, e0 = bk.make_trap(n, 17)
    out = {}
    for nt in (1, 0):
        eng = bk.Engine(0)
        try:
            eng.set_tuning("nt_plain", nt)
            r = eng.evolve(n, s0, 0.35, 0.02, 8, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0,
                           measure_every=1)
            out[nt] = (r, eng.download())
        finally:
            eng.close()
    (ra, pa), (rb, pb) = out[0], out[1]
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    for key in ("p_less", "dh", "norm"):
        assert np.array_equal(ra[key], rb[key]), key


def test_phase_pass_grids(bk):
    """Per-pass workgroup counts of the phase passes (grid_bphase /
    grid_lphase, the defaults from n = 30) set at n = 24, beyond the default
    grid for the LOW pass: every tile is transformed the same way, so the state
    is bit-identical; only the order in which the workgroups' FP64 partials
    are summed changes, so the observables agree to rounding."""
    n = 24
    _, s0, e0 = bk.make_trap(n, 19)
    out = {}
    for grids in ((0, 0), (256, 16384)):
        eng = bk.Engine(0)
        try:
            eng.set_tuning("grid_bphase", grids[0])
            eng.set_tuning("grid_lphase", grids[1])
            r = eng.evolve(n, s0, 0.35, 0.02, 12, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0,
                           measure_every=1)
            out[grids] = (r, eng.download())
        finally:
            eng.close()
    (ra, pa), (rb, pb) = out[(0, 0)], out[(256, 16384)]
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    for key in ("p_less", "dh", "norm"):
        np.testing.assert_allclose(ra[key], rb[key], rtol=1e-12, atol=1e-15, err_msg=key)


def test_tuning_key_keeps_other_defaults(bk):
    """bk_set_tuning pins only the key it sets: a later state of another n or
    precision takes that state's defaults for every other key (row bits,
    planner page granule), so setting 'grid' to its default value changes
    nothing, bit for bit, against a context that never set a key."""
    n = 24
    _, s0, e0 = bk.make_trap(n, 2)
    out = []
    for pin in (False, True):
        eng = bk.Engine(0)
        try:
            if pin:
                eng.set_tuning("grid", 4096)
                eng.evolve(16, 1, 0.3, 0.02, 3, precision=bk.BK_C64)  # another state first
            r = eng.evolve(n, s0, 0.35, 0.02, 30, order=bk.BK_STRANG, precision=bk.BK_C128, e0=e0, measure_every=5)
            out.append(np.concatenate([r["p_less"], r["dh"], r["norm"]]))
        finally:
            eng.close()
    assert np.array_equal(out[0], out[1])
