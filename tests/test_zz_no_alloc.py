"""GPU test of bk_sample_noalloc, the allocation-free variant of bk_sample: on
the same state and random numbers it returns exactly bk_sample's indices, on
two calls in a row (the second with a smaller state and fewer shots, so it
reuses the static scratch that the first filled), and rejects sizes beyond its
static scratch."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C64, C128 = 1, 2


def _noalloc(bk, n, prec, d_amps, r, stream):
    import torch
    dr = torch.tensor(r, dtype=torch.float64, device="cuda")
    dout = torch.full((len(r),), -7, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    rc = bk.lib().bk_sample_noalloc(n, prec, C.c_void_p(d_amps), C.c_void_p(dr.data_ptr()), len(r),
                                    C.c_void_p(dout.data_ptr()), C.c_void_p(stream))
    assert rc == 0, bk.lib().bk_last_error().decode()
    torch.cuda.synchronize()
    return dout.cpu().numpy()


@pytest.mark.parametrize("first,second", [((20, C64, 4000), (16, C128, 700)),
                                          ((24, C128, 3000), (13, C64, 50))])
def test_sample_noalloc_matches_bk_sample(bk, engine, first, second):
    for n, prec, shots in (first, second):
        print(f"noalloc sample n={n} prec={prec} shots={shots}", flush=True)
        _, s0, _ = bk.make_trap(n, 11 + n)
        engine.evolve(n, s0, 0.35, 0.05, 40, order=bk.BK_STRANG, precision=prec, measure_every=0)
        rng = np.random.default_rng(n * 7 + prec)
        r = rng.random(shots)
        r[:3] = [0.0, 0.5, 1.0 - 2 ** -53]
        # device_state() first: it applies any pending phase of the stored frame,
        # which changes |psi_s|^2 by rounding; both samplers then read one state
        d_amps, stream, n_held, prec_held = engine.device_state()
        assert (n_held, prec_held) == (n, prec)
        want = engine.sample(r)
        got = _noalloc(bk, n, prec, d_amps, r, stream)
        assert np.array_equal(got, want), np.nonzero(got != want)[0][:10]
        assert got.min() >= 0 and got.max() < (1 << n)


def test_sample_noalloc_small_state_and_limits(bk, engine):
    n, prec = 9, C128
    print(f"noalloc sample n={n} (no scratch) and limits", flush=True)
    rng = np.random.default_rng(3)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    engine.upload(psi, prec)
    r = rng.random(300)
    d_amps, stream, _, _ = engine.device_state()
    want = engine.sample(r)
    assert np.array_equal(_noalloc(bk, n, prec, d_amps, r, stream), want)
    L = bk.lib()
    dummy = C.c_void_p(d_amps)
    # beyond the static scratch (n > 32), too many shots, null pointers: rejected before any launch
    assert L.bk_sample_noalloc(33, prec, dummy, dummy, 1, dummy, None) == -1
    assert "outside" in L.bk_last_error().decode()
    assert L.bk_sample_noalloc(n, prec, dummy, dummy, (1 << 30) + 1, dummy, None) == -1
    assert L.bk_sample_noalloc(n, prec, None, dummy, 1, dummy, None) == -1
    assert L.bk_sample_noalloc(n, 3, dummy, dummy, 1, dummy, None) == -1
    assert L.bk_sample_noalloc(n, prec, dummy, None, 0, None, None) == 0
