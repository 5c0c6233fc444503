"""GPU tests of the ISampler drop-in (sampler.cuh:412-707), mirroring the
reference's tests/test_sampler.cu and adding value-level parity with the oracle."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_quantum_sampling_counts(bk):
    """test_sampler.cu:10-53."""
    n = 4
    s = bk.create_sampler(n)
    cfg = bk.SamplerConfig(n=n, g=0.5, dt=0.05, t_max=2.0, num_sample_points=5,
                           shots_per_point=50, track_p_less=True, verbose=False)
    r = s.sample(0, cfg)
    assert r.total_shots_collected == 5 * 50
    assert r.total_evals == r.total_shots_collected
    assert r.improved_shots <= r.total_shots_collected
    assert 0.0 <= r.success_rate <= 1.0 and 0.0 <= r.per_eval_rate <= 1.0
    assert 0 < len(r.unique_basis) <= 1 << n
    assert 0.0 <= r.mean_hamming_flips <= n
    assert len(r.seed_stats) == 3
    assert [st.seed for st in r.seed_stats] == [42, 42 + 10007, 42 + 2 * 10007]


def test_tunneling_probability_tracking(bk):
    """test_sampler.cu:56-106."""
    n = 4
    s = bk.create_sampler(n)
    cfg = bk.SamplerConfig(n=n, g=0.75, dt=0.05, t_max=3.0, num_sample_points=10,
                           shots_per_point=20, track_p_less=True, verbose=False)
    r = s.sample(7, cfg)
    assert len(r.p_less_history) > 0 and len(r.time_points) == len(r.p_less_history)
    assert np.all(r.p_less_history >= -1e-9) and np.all(r.p_less_history <= 1 + 1e-9)
    cfg.track_p_less = False
    r2 = s.sample(7, cfg)
    assert len(r2.p_less_history) == 0 and len(r2.time_points) == 0


def test_delayed_window(bk):
    """test_sampler.cu:109-150."""
    s = bk.create_sampler(4)
    cfg = bk.SamplerConfig(n=4, g=0.5, dt=0.05, t_start=1.0, t_max=2.5, dt_sample=0.5,
                           shots_per_point=25, verbose=False)
    r = s.sample(0, cfg)
    assert np.all(r.time_points >= 1.0 - 1e-9)
    assert np.allclose(r.time_points, [1.0, 1.5, 2.0, 2.5])
    assert r.total_shots_collected == 4 * 25


def test_trajectory_parity_and_metrics(bk, orc):
    """record_trajectory run at N=12 (config-0 shape) vs the oracle: histories,
    d_min, p_blind, E0 and the derived metrics of sampler.cuh:642-683."""
    n, g = 12, 0.35
    _, s0, e0 = orc.make_trap(n, 21)
    cfg = bk.SamplerConfig(n=n, g=g, dt=0.05, t_max=10.0, record_trajectory=True,
                           shots_per_point=0, verbose=False)
    r = bk.create_sampler(n).sample(s0, cfg)
    ref = orc.trotter(n, s0, g, 0.05, 200, order=1)
    assert len(r.time_points) == 201
    assert np.allclose(r.time_points, np.arange(201) * 0.05)
    assert np.max(np.abs(r.p_less_history - ref["p_less"])) <= 1e-12
    assert np.max(np.abs(r.dh_history - ref["dh"])) <= 1e-10
    assert r.start_energy == e0 == orc.energy(n, s0)
    assert r.d_min == orc.d_min(n, s0, e0)
    assert r.p_blind == orc.count_below(n, e0) / 2.0 ** n
    m = orc.derived_metrics(r.time_points, ref["p_less"], ref["dh"], 0.0, 10.0)
    assert r.initial_rise_rate_avg == pytest.approx(m["rise_avg"], rel=1e-9, abs=1e-15)
    assert r.initial_rise_rate_max == pytest.approx(m["rise_max"], rel=1e-9, abs=1e-15)
    assert r.peak_p_less == pytest.approx(m["peak_p"], rel=1e-10)
    assert r.peak_time == pytest.approx(m["peak_t"])
    assert r.window_avg_dh == pytest.approx(m["window_dh"], rel=1e-10)


def test_projective_samples_match_oracle(bk, orc):
    """Shots drawn with the reference's RNG stream (mt19937_64(seed + s*stride),
    uniform(0,1)) give the inverse-CDF indices of the evolved state."""
    n, g, dt = 10, 0.6, 0.05
    _, s0, e0 = orc.make_trap(n, 5)
    cfg = bk.SamplerConfig(n=n, g=g, dt=dt, t_max=1.0, num_sample_points=4, shots_per_point=64,
                           seed=123, num_sampling_seeds=2, seed_stride=7, verbose=False)
    r = bk.create_sampler(n).sample(s0, cfg)
    flags = orc.schedule(1.0, dt, num_sample_points=4, shots_per_point=64)
    pts = np.nonzero(flags & 2)[0]
    assert len(pts) == 4
    rngs = [orc.MT64(123 + s * 7) for s in range(2)]
    expect, cdfs, draws = [], [], []
    for k in pts:
        st = orc.trotter(n, s0, g, dt, int(k), order=1, want_state=True)["psi"]
        cdf = np.cumsum(np.abs(st) ** 2)
        for s in range(2):
            rn = np.array([rngs[s].uniform(0.0, 1.0) for _ in range(64)])
            expect.append(orc.sample_inverse_cdf(n, st, rn))
            cdfs += [cdf] * 64
            draws.append(rn)
    expect = np.concatenate(expect)
    draws = np.concatenate(draws)
    assert r.samples.shape == expect.shape
    # index-exact, except where a draw's target r * total lies on a CDF
    # boundary to within the rounding of two FP64 evolutions
    for i in np.nonzero(r.samples != expect)[0]:
        cdf = cdfs[i]
        lo = min(int(r.samples[i]), int(expect[i]))
        assert abs(cdf[lo] - draws[i] * cdf[-1]) <= 1e-12 * cdf[-1], (i, r.samples[i], expect[i])
    # statistics recomputed from the draws (sampler.cuh:575-640)
    first = r.samples.reshape(4, 2, 64)[:, 0, :].ravel()
    assert r.mean_hamming_flips == pytest.approx(
        np.mean([bin(int(x) ^ s0).count("1") for x in first]))
    improved = sum(orc.energy(n, int(x)) < e0 for x in first)
    assert r.improved_shots == improved
    assert sorted(r.unique_basis) == sorted(set(int(x) for x in r.samples))


def test_projective_samples_follow_born_rule(bk, orc, engine):
    """Statistical pin of bk_sample (the reference's sampler is closed, so its
    draws cannot be matched): 400k shots of an evolved N=10 state against
    |psi_s|^2 of the oracle's FP64 state -- chi-square goodness of fit (cells
    with an expected count below 5 pooled) and the Kolmogorov distance of the
    index CDF at the 0.1 % level.  The shots' uniforms come from a seeded
    generator, so the outcome is fixed."""
    from scipy import stats
    n, g, dt, steps = 10, 0.6, 0.05, 30
    _, s0, e0 = orc.make_trap(n, 3)
    engine.evolve(n, s0, g, dt, steps, order=bk.BK_LIE, precision=bk.BK_C128, e0=e0, measure_every=0)
    p = np.abs(orc.trotter(n, s0, g, dt, steps, order=1, want_state=True)["psi"]) ** 2
    p = p / p.sum()
    shots = 400_000
    idx = engine.sample(np.random.default_rng(2026).random(shots))
    assert idx.min() >= 0 and idx.max() < (1 << n)
    counts = np.bincount(idx, minlength=1 << n)
    expected = p * shots
    big = expected >= 5.0
    obs = np.append(counts[big], counts[~big].sum())
    exp = np.append(expected[big], expected[~big].sum())
    if exp[-1] == 0.0:
        obs, exp = obs[:-1], exp[:-1]
    chi2, pval = stats.chisquare(obs, exp * (obs.sum() / exp.sum()))
    print(f"\n[born rule n={n}] chi2={chi2:.1f} over {len(obs) - 1} dof, p={pval:.3f}")
    assert pval > 1e-3
    ks = np.abs(np.cumsum(counts) / shots - np.cumsum(p)).max()
    assert ks < 1.95 / np.sqrt(shots)

