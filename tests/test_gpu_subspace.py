"""GPU parity of the subspace ("sampled-Krylov") solvers and the dense ED
(bk_subspace_lanczos, bk_subspace_dense, bk_fullspace_dense) against numpy's
eigh of the restricted Hamiltonian built here from its definition (oracle
energies E(s)/n on the diagonal, -g between basis states one spin flip apart),
including the reference's own cases (test_lanczos.cu:66-102,
test_exact_diag.cpp:8-80)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dense_sub(orc, n, g, basis, normalize=True):
    idx = {int(b): i for i, b in enumerate(basis)}
    d = len(basis)
    H = np.zeros((d, d))
    for i, b in enumerate(basis):
        H[i, i] = orc.energy(n, int(b)) / (n if normalize else 1)
        for j in range(n):
            k = idx.get(int(b) ^ (1 << j))
            if k is not None:
                H[i, k] = -g
    return H


def test_reference_subspace_lanczos_case(bk, orc, engine):
    """test_lanczos.cu:66-102: n=6, g=0.35, basis {3i+1 : i < 16}, Krylov
    dimension 16: Lanczos and dense agree to 1e-6 (and both with eigh)."""
    n, g = 6, 0.35
    basis = [3 * i + 1 for i in range(16)]
    lz, _ = engine.subspace_lanczos(n, g, basis, max_krylov_dim=16, vectors=False)
    de, _ = engine.subspace_dense(n, g, basis, vectors=False)
    ref = np.linalg.eigvalsh(_dense_sub(orc, n, g, basis))
    assert abs(lz[0] - de[0]) < 1e-6
    assert de[0] == pytest.approx(ref[0], abs=1e-12)


def test_reference_subspace_dense_case(bk, orc, engine):
    """test_exact_diag.cpp Test 2: n=4, g=0.5, basis {0, 1, 3, 7}, 2 pairs,
    ascending, unit eigenvectors in the order of the basis."""
    n, g = 4, 0.5
    basis = [0, 1, 3, 7]
    ev, vec = engine.subspace_dense(n, g, basis, k_pairs=2)
    H = _dense_sub(orc, n, g, basis)
    w = np.linalg.eigvalsh(H)
    assert ev == pytest.approx(w[:2], abs=1e-12) and ev[0] <= ev[1]
    for k in range(2):
        assert np.linalg.norm(vec[k]) == pytest.approx(1.0, abs=1e-12)
        assert np.abs(H @ vec[k] - ev[k] * vec[k]).max() < 1e-10


@pytest.mark.parametrize("n,g,k", [(4, 0.5, 2), (10, 0.35, 3), (12, 0.35, 2)])
def test_fullspace_dense_against_eigh(bk, orc, engine, n, g, k):
    ev, vec = engine.fullspace_dense(n, g, k_pairs=k)
    H = _dense_sub(orc, n, g, list(range(1 << n)))
    w = np.linalg.eigvalsh(H)
    assert ev == pytest.approx(w[:k], abs=1e-10)
    for i in range(k):
        assert np.abs(H @ vec[i] - ev[i] * vec[i]).max() < 1e-9
    if n == 12:
        lz, _ = engine.lanczos(n, g, max_krylov_dim=100, k_pairs=1, vectors=False)
        assert lz[0] == pytest.approx(ev[0], abs=1e-8)


def test_sampled_basis_lanczos_dense_eigh(bk, orc, engine):
    """The path's own use: the unique basis of a sampler run (n=14), in the
    sampler's order, then both subspace solvers; eigenvectors come back in the
    order of the basis given."""
    n, g = 14, 0.35
    _, s0, _ = bk.make_trap(n, 3)
    cfg = bk.SamplerConfig(n=n, g=g, dt=0.05, t_max=6.0, num_sample_points=10, shots_per_point=100,
                           verbose=False)
    res = bk.create_sampler(n).sample(s0, cfg)
    basis = np.array(res.unique_basis, dtype=np.uint64)
    rng = np.random.default_rng(5)
    basis = basis[rng.permutation(basis.shape[0])]  # any order
    assert 50 < basis.shape[0] <= 3000
    H = _dense_sub(orc, n, g, basis)
    w, V = np.linalg.eigh(H)
    de, dvec = engine.subspace_dense(n, g, basis, k_pairs=3)
    assert de == pytest.approx(w[:3], abs=1e-10)
    assert abs(abs(dvec[0] @ V[:, 0]) - 1.0) < 1e-9
    lz, lvec = engine.subspace_lanczos(n, g, basis, max_krylov_dim=200, k_pairs=1)
    assert lz[0] == pytest.approx(w[0], abs=1e-8)
    assert np.abs(H @ lvec[0] - lz[0] * lvec[0]).max() < 1e-6
    # unnormalised diagonal (E(s) instead of E(s)/n)
    de2, _ = engine.subspace_dense(n, g, basis, normalize_by_n=False, vectors=False)
    assert de2[0] == pytest.approx(np.linalg.eigvalsh(_dense_sub(orc, n, g, basis, False))[0], abs=1e-9)


def test_subspace_errors(bk, engine):
    with pytest.raises(bk.BKError):
        engine.subspace_dense(6, 0.3, [1, 2, 1])          # a state twice
    with pytest.raises(bk.BKError):
        engine.subspace_lanczos(4, 0.3, [1, 2, 1 << 4])    # bits above n
    with pytest.raises(bk.BKError):
        engine.fullspace_dense(15, 0.3)                    # n <= 14 as in the reference
