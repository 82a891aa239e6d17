"""The time-domain front end on the device (exm_strain_segments, exm_strain_psd) against the oracle's numpy
transform, through size-independent round trips, and the stream fed from strain.

Tolerances: both sides transform in float64 with different FFT algorithms, so a bin differs by a few
rounding errors of the whole spectrum's scale, not of the bin's own: |d - d_oracle| <= 1e-12 max|d|, and
the periodogram mean within 1e-11 of its largest bin."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu

D_TOL = 1e-12
PSD_TOL = 1e-11
MIX_TOL = 1e-7


def say(*a):
    print(*a, flush=True)


def colored(rng, rows, n, fs, fmin, fmax):
    """band-limited Gaussian noise with the aLIGO fit as one-sided density: (rows, n) strain and its
    band spectrum D (rows, n_freq) = dt rfft(x) on the band bins (exact up to the inverse transform's
    rounding)"""
    T = n / fs
    nu = np.arange(int(np.ceil(fmin * T)), int(np.floor(fmax * T)) + 1)
    sig2 = T * synth.psd_ajith(nu / T) / 2.0
    D = np.sqrt(sig2 / 2.0) * (rng.standard_normal((rows, nu.size)) + 1j * rng.standard_normal((rows, nu.size)))
    Y = np.zeros((rows, n // 2 + 1), complex)
    Y[:, nu] = D * fs
    return np.fft.irfft(Y, n, axis=1), D, nu, sig2


@pytest.mark.parametrize("n,fs,band", [(32768, 4096.0, (20.0, 512.0)), (6001, 750.0, (10.0, 375.0)),
                                       (8192, 1024.0, (0.0, 512.0))])
def test_segments_match_the_oracle(hip, oracle, n, fs, band):
    from exmarg_amd.search import stream
    rng = np.random.default_rng(n)
    x = rng.standard_normal((5, n)) * np.linspace(1.0, 3.0, n)[None, :]
    say("strain segments: 5 rows of %d samples at %g Hz, band %s" % (n, fs, band))
    d, nu, T = stream.segment_from_strain(x, fs, *band)
    assert T == n / fs and nu[0] == int(np.ceil(band[0] * T)) and nu[-1] == int(np.floor(band[1] * T))
    ref = oracle.segment_from_strain(x, fs, int(nu[0]), nu.size)
    err = np.abs(d - ref).max() / np.abs(ref).max()
    say("max |d - d_oracle| / max |d| = %.2e" % err)
    assert err <= D_TOL


def test_segments_round_trip_in_chunks(hip):
    """strain made from a known band spectrum gives it back; 20 rows of 2^21 samples take several
    transform chunks"""
    from exmarg_amd.search import stream
    rng = np.random.default_rng(3)
    n, fs = 1 << 21, 65536.0
    x, D, nu, _ = colored(rng, 20, n, fs, 20.0, 2048.0)
    say("round trip: 20 rows of %d samples (%d band bins)" % (n, nu.size))
    d, nu2, _ = stream.segment_from_strain(x, fs, 20.0, 2048.0)
    np.testing.assert_array_equal(nu2, nu)
    err = np.abs(d - D).max() / np.abs(D).max()
    say("max |d - D| / max |D| = %.2e" % err)
    assert err <= D_TOL


@pytest.mark.parametrize("N,n", [(32, 32768), (20, 1 << 21)])
def test_psd_matches_the_oracle(hip, oracle, N, n):
    from exmarg_amd.search import stream
    rng = np.random.default_rng(N)
    fs = n / 8.0
    x, _, nu, sig2 = colored(rng, N, n, fs, 20.0, 512.0)
    say("psd: %d blocks of %d samples" % (N, n))
    p = stream.psd_from_strain(x.ravel(), fs, 8.0, 20.0, 512.0)
    ref = oracle.psd_from_strain(x, fs, int(nu[0]), nu.size, N / (N - 2.0))
    err = np.abs(p - ref).max() / ref.max()
    say("max |psd - psd_oracle| / max psd = %.2e; mean estimate / sig2 = %.4f" % (err, np.mean(p / sig2)))
    assert err <= PSD_TOL
    # the estimate is sig2 Gamma(N, 1/N) N/(N-2): its mean over the bins near N/(N-2)
    assert abs(np.mean(p / sig2) - N / (N - 2.0)) < 5.0 / np.sqrt(N * nu.size)


def test_front_end_rejects_bad_bands(hip):
    from exmarg_amd import _hip
    from exmarg_amd.search import stream
    x = np.zeros((2, 1024))
    with pytest.raises(ValueError, match="not inside"):
        stream.segment_from_strain(x, 1024.0, 20.0, 600.0)
    with pytest.raises(ValueError, match="at least 3 blocks"):
        stream.psd_from_strain(np.zeros(2048), 1024.0, 1.0, 20.0, 300.0)
    d = np.empty((2, 10), complex)
    # the library's own check of the bins against the Nyquist bin
    with pytest.raises(ValueError, match="Nyquist"):
        _hip.call("exm_strain_segments", _hip.ptr(x), 2, 1024, 1024.0, 510, 10, _hip.ptr(d))


def test_stream_from_strain_equals_frequency_domain(hip):
    """segments_from_strain on strain built from a simulated frequency-domain group gives the group's
    data back, and pe_redshift_stream on it the results of pe_redshift on the same segments"""
    from exmarg_amd.search import mass_bank, pe_redshift, pe_redshift_stream, stream, store as stm
    rng = np.random.default_rng(11)
    T, fs = 4.0, 1024.0
    n = int(T * fs)
    seg, _, _ = synth.simulate(3, 61, lambda r: (35.0, 1.0, 0.0), lambda r: 0.6, T, 20.0, 256.0)
    strain, noise = {}, {}
    for det in "HL":
        Y = np.zeros((3, n // 2 + 1), complex)
        Y[:, seg["nu"]] = seg["d_" + det] * fs
        strain[det] = np.fft.irfft(Y, n, axis=1).ravel()
        noise[det], _, _, _ = colored(rng, 32, n, fs, 20.0, 256.0)
    got = stream.segments_from_strain(strain, noise, fs, T, 20.0, 256.0)
    np.testing.assert_array_equal(got["nu"], seg["nu"])
    for det in "HL":
        err = np.abs(got["d_" + det] - seg["d_" + det]).max() / np.abs(seg["d_" + det]).max()
        assert err <= D_TOL, err
        assert got["sig2_" + det].shape == seg["sig2"].shape and np.all(got["sig2_" + det] > 0)
    say("stream from strain: 3 segments of 4 s, noise from 32 blocks per detector")
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[::4])
    masses = np.geomspace(20.0, 60.0, 10)
    kw = dict(zmax=3.0, nz=90, m1_range=(5.0, 40.0), nm1=80, mdet_range=(20.0, 60.0), seed=5, ymin=0.1267)
    res = list(pe_redshift_stream([got], masses, st, batch_segments=2, **kw))[0]
    ref = pe_redshift(got, mass_bank(masses, got, "hip"), st, resident=True, **kw)
    np.testing.assert_allclose(res["ln_b"], ref["ln_b"], rtol=0, atol=MIX_TOL)
    stm.free(st)
