"""Template factors computed on the device (exm_bank_taylorf2) against the host TaylorF2 model
(exmarg_amd.synth.template_factors), both float64 up to the complex64 rounding of W."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("T,fmin,fmax", [(4.0, 20.0, 512.0), (16.0, 20.0, 1024.0)])
def test_device_taylorf2_matches_host(hip, T, fmin, fmax):
    from exmarg_amd.search import mass_bank
    say("taylorf2: T=%g band %g-%g Hz" % (T, fmin, fmax))
    seg, _, _ = synth.simulate(1, 5, lambda r: (30.0, 0.8, 0.1), lambda r: 0.0, T, fmin, fmax)
    rng = np.random.default_rng(2)
    # masses from light (ISCO above the band) to heavy (ISCO inside the band: W is zero above it)
    rows = np.stack([np.geomspace(2.5, 300.0, 24), rng.uniform(0.1, 1.0, 24), rng.uniform(-0.9, 0.9, 24)], axis=1)
    bank = mass_bank(rows, seg, "hip")
    W_host, u_host = synth.template_factors(rows[:, 0], rows[:, 1], rows[:, 2], dict(seg, sig2_H=seg["sig2"]))
    W_dev = bank.W
    assert W_dev.shape == W_host.shape
    np.testing.assert_array_equal(W_dev == 0, W_host == 0)      # same bins cut at the ISCO frequency
    live = np.abs(W_host).max(axis=1) > 0
    assert live.sum() >= 12 and (~live).sum() >= 2               # both kinds of template are covered
    err = np.abs(W_dev[live] - W_host[live]).max(axis=1) / np.abs(W_host[live]).max(axis=1)
    say("max relative W error", err.max())
    assert err.max() < 2e-6
    np.testing.assert_allclose(bank.u, u_host, rtol=1e-12, atol=0)


def test_device_taylorf2_rejects_bad_templates(hip):
    from exmarg_amd.search import mass_bank
    seg, _, _ = synth.simulate(1, 5, lambda r: (30.0, 0.8, 0.1), lambda r: 0.0, 4.0, 20.0, 256.0)
    with pytest.raises(ValueError, match="outside"):
        mass_bank(np.array([[30.0, 1.5, 0.0]]), seg, "hip")
