"""Failure handling on the device, forced from the host: a device-memory cap
makes allocations fail as on a full device.  Each failing call must return
BK_E_NOMEM (no exception, no abort), and the context must stay usable: the
same run after the cap is lifted gives the bits a fresh context gives.
Every kernel still runs normally; nothing here faults or hangs the GPU."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_allocation_failure_under_a_device_memory_cap(bk):
    L = bk.lib()
    n = 26
    _, s0, e0 = bk.make_trap(n, 1)
    ref = bk.Engine(0)
    try:
        want = ref.evolve(n, s0, 0.35, 0.02, 6, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0, measure_every=2)
    finally:
        ref.close()
    eng = bk.Engine(0)
    try:
        held = L.bk_device_memory_held()
        # room for the 512 MiB state of N=26 complex64, not for its energy caches
        assert L.bk_set_device_memory_limit(held + (600 << 20)) == 0
        print("cap: N=26 state fits, its energy caches do not", flush=True)
        with pytest.raises(bk.BKError) as ei:
            eng.evolve(n, s0, 0.35, 0.02, 6, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0, measure_every=2)
        assert ei.value.code == -3
        print("cap: N=32 complex128 (68.7 GB) refused before any launch", flush=True)
        with pytest.raises(bk.BKError) as ei:
            eng.evolve(32, 1, 0.35, 0.02, 1, precision=bk.BK_C128)
        assert ei.value.code == -3
        assert L.bk_device_memory_held() <= held + (600 << 20)
        # small work still runs under the cap
        small = eng.evolve(12, 5, 0.3, 0.05, 20, precision=bk.BK_C128)
        assert abs(small["norm"][-1] - 1.0) < 1e-12
        assert L.bk_set_device_memory_limit(0) == 0
        got = eng.evolve(n, s0, 0.35, 0.02, 6, order=bk.BK_STRANG, precision=bk.BK_C64, e0=e0, measure_every=2)
        for k in ("p_less", "dh", "norm"):
            assert np.array_equal(got[k], want[k]), k
    finally:
        L.bk_set_device_memory_limit(0)
        eng.close()
