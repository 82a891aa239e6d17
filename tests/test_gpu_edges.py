"""Edge cases of pe_intrinsic on the device against the FP64 oracle: no rows, one template, a segment so
short that its filter length takes the hipFFT path, the largest n_loud a segment allows, a segment of
zeros (every lookup in the clamped first beta row), and a walk longer than one launch's slice."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu

LNB_TOL = 1e-5
LNL_TOL = 1e-5


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def store(hip):
    from exmarg_amd.search import store as stm
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[::4])
    yield st
    stm.free(st)


def _check(oracle, seg, bank, st, res, **kw):
    from exmarg_amd.search import read_segments
    b = dict(W=bank.W, u=bank.u, M=bank.M, q=bank.q, chi=bank.chi, lo=bank.lo, n_time=bank.n_time)
    store = dict(s=st.s, alpha_lo=st.alpha_lo, alpha_hi=st.alpha_hi, beta=st.beta[0], delta=st.delta[0],
                 tau_idx=st.tau_idx)
    tabs = oracle.Tables(synth=dict(seed=0, kind="smooth", nodes=np.arange(st.s.size)), shape=st.table_shape(),
                         tau_idx=st.tau_idx)
    ref = oracle.pe_intrinsic(read_segments(seg), b, store, tabs, **kw)
    dl = np.max(np.abs(res["ln_lambda"] - ref["ln_lambda"]))
    db = np.max(np.abs(res["ln_b"] - ref["ln_b"]))
    say("max |d ln Lambda| %.2e, max |d ln B| %.2e" % (dl, db))
    assert np.isfinite(res["ln_lambda"]).all()
    assert dl <= LNL_TOL and db <= LNB_TOL
    return ref


def test_no_rows_gives_empty_results(hip, store):
    from exmarg_amd.search import mass_bank, pe_intrinsic
    seg, _, _ = synth.simulate_at_snr(1, 3, (30.0, 0.8, 0.1), 12.0, 2.0, 20.0, 256.0)
    bank = mass_bank(synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 4), seg, "hip")
    res = pe_intrinsic(seg, bank, store, rows=[], resident=True, n_loud=64, n_random=64)
    assert res["ln_lambda"].shape == (0, 4) and res["ln_b"].shape == (0,) and res["rows"].size == 0


def test_one_template(hip, oracle, store):
    from exmarg_amd.search import mass_bank, pe_intrinsic
    seg, _, _ = synth.simulate_at_snr(1, 4, (30.0, 0.8, 0.1), 12.0, 2.0, 20.0, 256.0)
    bank = mass_bank(np.array([[30.0, 0.8, 0.1]]), seg, "hip")
    say("one template, 2 s")
    kw = dict(n_loud=128, n_random=128, seed=1)
    res = pe_intrinsic(seg, bank, store, resident=True, **kw)
    _check(oracle, seg, bank, store, res, **kw)


@pytest.mark.parametrize("n_loud", [0, 100])
def test_short_segment_takes_the_hipfft_filter(hip, oracle, store, n_loud):
    """T = 1 s at 20-64 Hz: n_time = 8 * 64 = 512, a length the band transform does not take"""
    from exmarg_amd.search import mass_bank, pe_intrinsic
    seg, _, _ = synth.simulate_at_snr(1, 5, (60.0, 0.8, 0.1), 10.0, 1.0, 20.0, 64.0)
    bank = mass_bank(synth.chirp_mass_grid_bank(60.0, 0.8, 0.1, 5), seg, "hip")
    assert bank.n_time == 512
    say("1 s at 20-64 Hz, n_loud %d" % n_loud)
    kw = dict(n_loud=n_loud, n_random=100, seed=2)
    res = pe_intrinsic(seg, bank, store, resident=True, **kw)
    _check(oracle, seg, bank, store, res, **kw)


def test_largest_n_loud(hip, oracle, store):
    """n_loud = n_time - 1: one time is left for the random draw, which every draw picks"""
    from exmarg_amd.search import mass_bank, pe_intrinsic
    seg, _, _ = synth.simulate_at_snr(1, 6, (60.0, 0.8, 0.1), 10.0, 1.0, 20.0, 64.0)
    bank = mass_bank(synth.chirp_mass_grid_bank(60.0, 0.8, 0.1, 3), seg, "hip")
    say("n_loud = n_time - 1 = %d" % (bank.n_time - 1))
    kw = dict(n_loud=bank.n_time - 1, n_random=7, seed=3)
    res = pe_intrinsic(seg, bank, store, resident=True, **kw)
    _check(oracle, seg, bank, store, res, **kw)
    with pytest.raises(ValueError):
        pe_intrinsic(seg, bank, store, resident=True, n_loud=bank.n_time, n_random=7)


def test_segment_of_zeros(hip, oracle, store):
    """zero data: |z|^2 is clamped to 1e-38, every walked time sits in the clamped first beta row and every
    lookup goes through the edge rows' exact path"""
    from exmarg_amd.search import mass_bank, pe_intrinsic
    seg, _, _ = synth.simulate_at_snr(1, 7, (30.0, 0.8, 0.1), 12.0, 2.0, 20.0, 256.0)
    seg = dict(seg, d_H=np.zeros_like(seg["d_H"]), d_L=np.zeros_like(seg["d_L"]))
    bank = mass_bank(synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 4), seg, "hip")
    say("segment of zeros")
    kw = dict(n_loud=64, n_random=64, seed=4)
    res = pe_intrinsic(seg, bank, store, resident=True, **kw)
    _check(oracle, seg, bank, store, res, **kw)


def test_walk_longer_than_a_slice(hip, oracle, store):
    """n_loud + n_random = 12000 walked times: two slices of the band gather merged per read and lag"""
    from exmarg_amd.search import mass_bank, pe_intrinsic
    seg, _, _ = synth.simulate_at_snr(1, 8, (30.0, 0.8, 0.1), 12.0, 4.0, 20.0, 512.0)
    bank = mass_bank(synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 3), seg, "hip")
    say("walk of 12000 times, n_time %d" % bank.n_time)
    kw = dict(n_loud=6000, n_random=6000, seed=5)
    res = pe_intrinsic(seg, bank, store, resident=True, **kw)
    _check(oracle, seg, bank, store, res, **kw)
