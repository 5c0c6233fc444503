"""GPU parity of sampling, H.v, Lanczos and the Krylov propagator."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C64, C128 = 1, 2


def _boundary_ok(psi, r, got, ref):
    """Mismatch allowed only when r*total sits within 1e-12 of a CDF boundary."""
    cdf = np.cumsum(np.abs(psi.astype(np.complex128)) ** 2)
    t = r * cdf[-1]
    lo = min(got, ref)
    return abs(cdf[lo] - t) <= 1e-12 * cdf[-1]


@pytest.mark.parametrize("n,prec", [(5, C128), (11, C128), (14, C128), (17, C64), (20, C128)])
def test_sampling_inverse_cdf(engine, orc, n, prec):
    print(f"sample n={n} prec={prec}", flush=True)
    rng = np.random.default_rng(100 + n)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi *= rng.random(1 << n) ** 4  # uneven weights
    psi /= np.linalg.norm(psi)
    dt = np.complex64 if prec == C64 else np.complex128
    psi = psi.astype(dt)
    engine.upload(psi, prec)
    r = rng.random(3000)
    r[:3] = [0.0, 0.5, 1.0 - 2 ** -53]
    got = engine.sample(r)
    ref = orc.sample_inverse_cdf(n, psi.astype(np.complex128), r)
    bad = np.nonzero(got != ref)[0]
    for i in bad:
        assert _boundary_ok(psi, r[i], got[i], ref[i]), (i, got[i], ref[i])
    assert len(bad) <= 2


@pytest.mark.parametrize("n", [4, 9, 12, 13, 14, 17, 22, 24])
def test_hv_parity(engine, orc, n):
    """Against the oracle's H.v.  The tiled LOW pass adds 0, 1 or 2 high-group
    sums: n = 14 (real FP64: none; complex128: one), 17 and 22 (one), 24
    (two in both precisions)."""
    import torch
    print(f"hv n={n}", flush=True)
    rng = np.random.default_rng(n)
    g = 0.37
    x = rng.standard_normal(1 << n)
    dx = torch.tensor(x, device="cuda")
    dy = torch.empty_like(dx)
    engine.hv(n, g, dx, dy)
    ref = orc.hv_real(n, g, x)
    assert np.max(np.abs(dy.cpu().numpy() - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    z = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    dz = torch.tensor(z, device="cuda")
    dw = torch.empty_like(dz)
    engine.hv(n, g, dz, dw, normalize_by_n=False)
    refz = orc.hv_complex(n, g, z, normalize_by_n=False)
    assert np.max(np.abs(dw.cpu().numpy() - refz)) <= 1e-12 * max(1.0, np.max(np.abs(refz)))


def test_hv_symmetry(engine):
    import torch
    n = 15
    g = 0.3
    rng = np.random.default_rng(5)
    x = torch.tensor(rng.standard_normal(1 << n), device="cuda")
    y = torch.tensor(rng.standard_normal(1 << n), device="cuda")
    hx, hy = torch.empty_like(x), torch.empty_like(y)
    engine.hv(n, g, x, hx)
    engine.hv(n, g, y, hy)
    a = float(torch.dot(x, hy))
    b = float(torch.dot(hx, y))
    assert abs(a - b) <= 1e-10 * abs(a)


def test_lanczos_vs_dense_ed(engine, orc):
    """test_lanczos.cu:11-63 (N=4, g=0.5, m=30): |E_lanczos - E_ED| < 1e-6 and
    ||Hv - Ev|| < 1e-5."""
    n, g = 4, 0.5
    w, _ = orc.syev(orc.dense_h(n, g))
    ev, vec = engine.lanczos(n, g, max_krylov_dim=30)
    assert abs(ev[0] - w[0]) < 1e-6
    r = orc.hv_real(n, g, vec[0]) - ev[0] * vec[0]
    assert np.linalg.norm(r) < 1e-5


def test_lanczos_matches_oracle(engine, orc):
    for n, g, m in ((10, 0.35, 60), (16, 0.3, 40)):
        ev, vec = engine.lanczos(n, g, max_krylov_dim=m, k_pairs=2)
        ev2, vec2 = orc.lanczos(n, g, max_krylov_dim=m, k_pairs=2)
        # same v0 and recurrence: the converged ground Ritz value agrees to rounding;
        # the unconverged second one (no reorthogonalisation) is rounding-sensitive
        assert abs(ev[0] - ev2[0]) < 1e-10
        assert abs(ev[1] - ev2[1]) < 1e-4
        assert abs(abs(np.dot(vec[0], vec2[0])) - np.linalg.norm(vec[0]) *
                   np.linalg.norm(vec2[0])) < 1e-6


@pytest.mark.parametrize("n,m", [(8, 30), (14, 30), (14, 32), (14, 33), (12, 45), (12, 2), (4, 30)])
def test_krylov_propagator(engine, orc, n, m):
    """Against the oracle's Krylov propagator (same recurrence, full
    reorthogonalisation).  m <= 32: the one-block fused Gram-Schmidt passes;
    m = 33, 45: the later steps orthogonalise in blocks of 32; n = 4 (16
    states) with m = 30: the recurrence breaks down early."""
    rng = np.random.default_rng(n + 100 * m)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    engine.upload(psi, C128)
    engine.krylov_propagate(0.35, 1.0, tau_restart=0.5, m=m)
    got = engine.download()
    ref = orc.krylov_propagate(n, 0.35, psi, 0.5, m=m)
    ref = orc.krylov_propagate(n, 0.35, ref, 0.5, m=m)
    assert np.max(np.abs(got - ref)) < 1e-10


def test_count_below(engine, orc):
    for n, e0 in ((10, 20), (16, 40), (20, 34), (22, 60)):
        assert engine.count_below(n, e0) == orc.count_below(n, e0)


def test_errors_are_reported(engine, bk):
    with pytest.raises(bk.BKError):
        engine.evolve(33, 0, 0.3, 0.05, 1)
    with pytest.raises(bk.BKError):
        engine.evolve(10, 0, float("nan"), 0.05, 1)
    with pytest.raises(bk.BKError):
        engine.evolve(10, 0, 0.3, 0.05, 1, order=3)


def test_copy_probe_patterns(engine, bk):
    """The HBM copy probes: every pattern copies the same bytes and reports a
    finite positive rate; sizes off the pattern's chunk are rejected."""
    nbytes = 1 << 27  # a multiple of every pattern's chunk (16 B, 4 KB, 128 KB)
    for pattern in (0, 1, 2):
        gbs = engine.copy_probe(nbytes, 2, pattern=pattern)
        assert np.isfinite(gbs) and gbs > 0.0, pattern
    with pytest.raises(bk.BKError):
        engine.copy_probe(4096 + 16, 1, pattern=1)
    with pytest.raises(bk.BKError):
        engine.copy_probe((1 << 20) + 4096, 1, pattern=2)
    with pytest.raises(bk.BKError):
        engine.copy_probe(1 << 20, 1, pattern=3)
