"""The segment stream (pe_redshift over groups sharing an estimated noise spectrum, gathered in two
launch slots) gives per segment what pe_redshift gives on the same inputs."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu

MIX_TOL = 1e-7          # a segment batched beside others vs alone (float32 exps grouped differently)


def say(*a):
    print(*a, flush=True)


def groups():
    out = []
    for g, (n, seed) in enumerate(((3, 61), (2, 62))):
        seg, _, _ = synth.simulate(n, seed, lambda r: (35.0, 1.0, 0.0), lambda r: 0.6 if g == 0 else 0.0, 4.0, 20.0,
                                   256.0)
        from exmarg_amd.search import stream
        P = stream.noise_periodograms(seg["sig2"], 32, np.random.default_rng(seed))
        seg["sig2"] = stream.psd_estimate(P)
        out.append(seg)
    return out


def test_stream_equals_pe_redshift(hip):
    from exmarg_amd.search import mass_bank, pe_redshift, pe_redshift_stream, store as stm
    say("stream: two groups (3 + 2 segments of 4 s), slots of 2 segments")
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[::4])
    masses = np.geomspace(20.0, 60.0, 10)
    kw = dict(zmax=3.0, nz=90, m1_range=(5.0, 40.0), nm1=80, mdet_range=(20.0, 60.0), seed=5, ymin=0.1267)
    gs = groups()
    got = list(pe_redshift_stream(gs, masses, st, batch_segments=2, keep=lambda g, row: True, **kw))
    assert len(got) == 2
    for seg, res in zip(gs, got):
        bank = mass_bank(masses, seg, "hip")
        ref = pe_redshift(seg, bank, st, resident=True, **kw)
        np.testing.assert_allclose(res["ln_b"], ref["ln_b"], rtol=0, atol=MIX_TOL)
        for j, row in enumerate(ref["rows"]):
            L = res["kept"][int(row)]
            np.testing.assert_array_equal(np.isfinite(L), np.isfinite(ref["ln_lambda"][j]))
            fin = np.isfinite(L)
            np.testing.assert_allclose(L[fin], ref["ln_lambda"][j][fin], rtol=0, atol=MIX_TOL)
    assert np.all(np.isfinite(got[0]["ln_b"]))
    stm.free(st)
