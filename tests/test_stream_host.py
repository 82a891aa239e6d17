"""Host side of the segment stream: the periodogram estimate of the noise variance from periodograms a
caller holds, and the credible levels (no device)."""
import numpy as np
import pytest

from exmarg_amd.search import stream


def test_psd_estimate_factor_and_distribution():
    """the mean of N periodograms is sig2 * Gamma(N, 1/N); N/(N-2) makes E[g^-2]/E[g^-1] one on average"""
    rng = np.random.default_rng(0)
    sig2 = np.full(200000, 2.5)
    N = 32
    P = stream.noise_periodograms(sig2, N, rng)
    raw = stream.psd_estimate(P, scale="none")
    g = raw / sig2
    assert abs(g.mean() - 1.0) < 3e-3 and abs(g.var() - 1.0 / N) < 1e-3          # Gamma(N, 1/N)
    est = stream.psd_estimate(P)
    np.testing.assert_allclose(est, raw * N / (N - 2.0), rtol=1e-15)
    ge = est / sig2
    # the ratio r = E[g^-2] / E[g^-1] of notebooks/07_estimated_psd.ipynb cell 2 is one with N/(N-2)
    assert abs(np.mean(ge ** -2.0) / np.mean(ge ** -1.0) - 1.0) < 5e-3
    gr = raw / sig2
    assert abs(np.mean(gr ** -2.0) / np.mean(gr ** -1.0) - N / (N - 2.0)) < 5e-3
    with pytest.raises(ValueError):
        stream.psd_estimate(P, scale="median")
    with pytest.raises(ValueError):
        stream.psd_estimate(P[:2])


def test_credible_level_and_population():
    grid = np.linspace(0.0, 1.0, 101)
    marg = np.ones(101)
    assert abs(stream.credible_level(grid, marg, 0.5) - 0.5) < 1e-12
    rng = np.random.default_rng(2)
    m1 = np.geomspace(2.0, 200.0, 60)
    z = np.linspace(0.02, 10.0, 50)
    prior = stream.population_prior(m1, z, (10.0, 500.0))
    assert prior.min() >= 0 and prior.sum() > 0
    Md = 2.0 * m1[:, None] * (1.0 + z[None, :])
    assert np.all(prior[(Md < 10.0) | (Md > 500.0)] == 0.0)
    for _ in range(20):
        a, b = stream.draw_population(rng, (10.0, 500.0))
        assert 10.0 <= 2 * a * (1 + b) <= 500.0
