"""The reference's CPU-runnable configuration end to end against the FP64 oracle: one 16 s segment at
20-1024 Hz (n_time 131072, 16065 bins) with one TaylorF2 injection (M_det 30, q 0.8, chi 0.1,
network SNR 12), a 64-template chirp-mass grid bank, one table node that every template reaches,
pe_intrinsic with the 'volume' prior; subsampled (1024 + 1024) and every-time walks."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu

LNB_TOL = 1e-5


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def config1(hip):
    from exmarg_amd.search import mass_bank, store as stm
    seg, _, snr = synth.simulate_at_snr(1, 20261203, (30.0, 0.8, 0.1), 12.0, 16.0, 20.0, 1024.0)
    bank = mass_bank(synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 64), seg, "hip")
    s_node = np.array([4.0 * bank.u.max()])                 # a = s / u >= 4 for every template
    st = stm.synthetic_store(seed=0, s_nodes=s_node)
    yield seg, bank, st
    stm.free(st)


@pytest.mark.parametrize("n_loud,n_random", [(1024, 1024), (0, 1024)])
def test_config1_matches_oracle(config1, oracle, n_loud, n_random):
    from exmarg_amd.search import pe_intrinsic, read_segments
    seg, bank, st = config1
    say("config 1: 64 templates x 131072 times, one node, n_loud=%d" % n_loud)
    res = pe_intrinsic(seg, bank, st, n_loud=n_loud, n_random=n_random, seed=0, resident=True)
    b = dict(W=bank.W, u=bank.u, M=bank.M, q=bank.q, chi=bank.chi, lo=bank.lo, n_time=bank.n_time)
    store = dict(s=st.s, alpha_lo=st.alpha_lo, alpha_hi=st.alpha_hi, beta=st.beta[0], delta=st.delta[0],
                 tau_idx=st.tau_idx)
    tabs = oracle.Tables(synth=dict(seed=0, kind="smooth", nodes=np.arange(st.s.size)), shape=st.table_shape(),
                         tau_idx=st.tau_idx)
    ref = oracle.pe_intrinsic(read_segments(seg), b, store, tabs, n_loud=n_loud, n_random=n_random, seed=0)
    err = np.max(np.abs(res["ln_lambda"] - ref["ln_lambda"]))
    say("ln B %.8f oracle %.8f; max |d ln Lambda| %.2e" % (res["ln_b"][0], ref["ln_b"][0], err))
    assert np.all(np.isfinite(res["ln_lambda"]))
    assert err <= LNB_TOL
    assert abs(res["ln_b"][0] - ref["ln_b"][0]) <= LNB_TOL


def test_config1_matches_reference_semantics(config1, oracle):
    """the HIP path against the oracle in the reference's own CPU arithmetic (complex64 filter, float32
    score ranked by argpartition, float64 gather), the agreement the reference's CPU/GPU check asks for"""
    from exmarg_amd.search import pe_intrinsic, read_segments
    seg, bank, st = config1
    say("config 1: HIP against the reference-semantics oracle")
    res = pe_intrinsic(seg, bank, st, n_loud=1024, n_random=1024, seed=0, resident=True)
    b = dict(W=bank.W, u=bank.u, M=bank.M, q=bank.q, chi=bank.chi, lo=bank.lo, n_time=bank.n_time)
    store = dict(s=st.s, alpha_lo=st.alpha_lo, alpha_hi=st.alpha_hi, beta=st.beta[0], delta=st.delta[0],
                 tau_idx=st.tau_idx)
    tabs = oracle.Tables(synth=dict(seed=0, kind="smooth", nodes=np.arange(st.s.size)), shape=st.table_shape(),
                         tau_idx=st.tau_idx)
    ref = oracle.pe_intrinsic(read_segments(seg), b, store, tabs, n_loud=1024, n_random=1024, seed=0,
                              semantics="ref")
    err = np.max(np.abs(res["ln_lambda"] - ref["ln_lambda"]))
    say("ln B %.8f reference semantics %.8f; max |d ln Lambda| %.2e" % (res["ln_b"][0], ref["ln_b"][0], err))
    assert abs(res["ln_b"][0] - ref["ln_b"][0]) <= LNB_TOL
