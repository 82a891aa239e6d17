"""End-to-end parity of pe_intrinsic / pe_redshift (device='hip') with the FP64 oracle on the same
segments, banks and tables."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu
# repeated runs of one analysis: the band kernel's default form adds each read's pieces by float64
# atomic adds in LDS, whose order varies from run to run, so repeats agree to float64 rounding
REPEAT_RTOL = 1e-13

LNB_TOL = 1e-5          # |ln B - ln B_oracle| bar of the analysis
LNL_TOL = 1e-5          # per-template ln Lambda
# the same segment alone or beside others in a launch: the band kernel groups a read's lookups with
# float32 exps of non-positive differences, so the grouping (which depends on the neighbours in a
# workgroup) moves ln Lambda by at most this much
MIX_TOL = 1e-7


def say(*a):
    print(*a, flush=True)


def small_problem(n_templates=12, T=4.0, fmax=256.0, nodes=slice(None, None, 4), seed=11):
    seg1, _, snr = synth.simulate_at_snr(1, seed, (30.0, 0.8, 0.1), 12.0, T, 20.0, fmax)
    seg0, _, _ = synth.simulate(1, seed + 50, lambda r: (30.0, 0.8, 0.1), lambda r: 0.0, T, 20.0, fmax)
    seg = dict(seg1)
    seg["d_H"] = np.concatenate([seg1["d_H"], seg0["d_H"]])
    seg["d_L"] = np.concatenate([seg1["d_L"], seg0["d_L"]])
    bank = synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, n_templates)
    return seg, bank, synth.S_NODES[nodes]


def oracle_inputs(bank, st, tabs_seed=0):
    b = dict(W=bank.W, u=bank.u, M=bank.M, q=bank.q, chi=bank.chi, lo=bank.lo, n_time=bank.n_time)
    store = dict(s=st.s, alpha_lo=st.alpha_lo, alpha_hi=st.alpha_hi, beta=st.beta[0], delta=st.delta[0],
                 tau_idx=st.tau_idx)
    return b, store


@pytest.fixture(scope="module")
def problem(hip):
    from exmarg_amd.search import mass_bank, read_segments, store as stm
    seg, bankrows, s_nodes = small_problem()
    st = stm.synthetic_store(seed=0, s_nodes=s_nodes)
    bank = mass_bank(bankrows, seg, "hip")
    yield seg, bank, st
    stm.free(st)


@pytest.mark.parametrize("prior,n_loud", [("volume", 1024), ("volume", 0), ("madau-dickinson", 256)])
def test_pe_intrinsic_matches_oracle(problem, oracle, prior, n_loud):
    from exmarg_amd.search import pe_intrinsic, read_segments
    seg, bank, st = problem
    say("pe_intrinsic prior=%s n_loud=%d n_tmpl=%d nodes=%d" % (prior, n_loud, bank.u.size, st.s.size))
    ymin = 0.1267 if prior == "volume" else 0.0982
    res = pe_intrinsic(seg, bank, st, prior=prior, ymin=ymin, n_loud=n_loud, n_random=n_loud or 1024, seed=3,
                       resident=True)
    b, store = oracle_inputs(bank, st)
    tabs = oracle.Tables(synth=dict(seed=0, kind="smooth", nodes=np.arange(st.s.size)), shape=st.table_shape(),
                         tau_idx=st.tau_idx)
    ref = oracle.pe_intrinsic(read_segments(seg), b, store, tabs, prior=prior, ymin=ymin, n_loud=n_loud,
                              n_random=n_loud or 1024, seed=3)
    say("ln_b hip", res["ln_b"], "oracle", ref["ln_b"])
    np.testing.assert_allclose(res["ln_lambda"], ref["ln_lambda"], rtol=0, atol=LNL_TOL)
    np.testing.assert_allclose(res["ln_b"], ref["ln_b"], rtol=0, atol=LNB_TOL)
    assert res["ln_b"][0] > res["ln_b"][1]          # the injection outweighs noise


def test_pe_redshift_matches_oracle(hip, oracle):
    from exmarg_amd.search import mass_bank, pe_redshift, read_segments, store as stm
    say("pe_redshift 10 templates, 80x90 grid")
    seg, _, s_nodes = small_problem()
    masses = np.geomspace(20.0, 60.0, 10)
    bank = mass_bank(masses, seg, "hip")
    st = stm.synthetic_store(seed=0, s_nodes=s_nodes)
    kw = dict(zmax=3.0, nz=90, m1_range=(5.0, 40.0), nm1=80, mdet_range=(20.0, 60.0), seed=5)
    res = pe_redshift(seg, bank, st, resident=False, **kw)
    b, store = oracle_inputs(bank, st)
    tabs = oracle.Tables(synth=dict(seed=0, kind="smooth", nodes=np.arange(st.s.size)), shape=st.table_shape(),
                         tau_idx=st.tau_idx)
    rseg = read_segments(seg)
    geometry = oracle.node_geometry(store["s"], store["alpha_lo"], store["alpha_hi"], bank.u, 0.1267)
    z = np.linspace(0.02, 3.0, 90)
    m1 = np.geomspace(5.0, 40.0, 80)
    yq = oracle.D0 / oracle.dl_of_z(z)
    w = oracle.pe_redshift_weights(m1, z, (20.0, 60.0))
    for e in range(2):
        row = dict(d_H=rseg["d_H"][e], d_L=rseg["d_L"][e], sig2_H=rseg["sig2_H"], sig2_L=rseg["sig2_L"], row=e)
        _, _, lamN = oracle.analyze_segment(row, b, store, tabs, geometry, 1024, 1024, 5)
        L = oracle.to_source_frame(lamN, geometry[2], bank.M, yq, m1, z)
        fin = np.isfinite(L)
        np.testing.assert_array_equal(np.isfinite(res["ln_lambda"][e]), fin)
        np.testing.assert_allclose(res["ln_lambda"][e][fin], L[fin], rtol=0, atol=LNL_TOL)
        assert abs(res["ln_b"][e] - oracle.log_mean(L, w)) < LNB_TOL
    # a store generated on the device has no host copy to upload again: resident=False keeps it
    assert st.device is not None
    stm.free(st)


def test_segments_sharing_a_launch_match_one_launch_each(hip, monkeypatch):
    """with chunk > 1 the segments of a bank with few reads per table share gather launches (a SegmentBatch
    of up to chunk segments); every segment's pe_redshift grid and ln B, and pe_intrinsic ln Lambda and ln B,
    equal those of one launch per segment (chunk=1) to float64 rounding"""
    import exmarg_amd.search.batch as bmod
    from exmarg_amd.search import mass_bank, pe_intrinsic, pe_redshift, store as stm
    say("shared launches: 3 segments, chunk 2 (pe_redshift) and 3 (pe_intrinsic) against chunk 1")
    seg, bankrows, s_nodes = small_problem()
    seg = dict(seg, d_H=np.concatenate([seg["d_H"], seg["d_H"][:1] * 0.5]),
               d_L=np.concatenate([seg["d_L"], seg["d_L"][:1] * 0.5]))
    st = stm.synthetic_store(seed=0, s_nodes=s_nodes)
    made = []

    class Spy(bmod.SegmentBatch):
        def __init__(self, *a, **k):
            super().__init__(*a, **k)
            made.append(self.cap)

    monkeypatch.setattr(bmod, "SegmentBatch", Spy)
    masses = np.geomspace(20.0, 60.0, 10)
    rbank = mass_bank(masses, seg, "hip")
    kw = dict(zmax=3.0, nz=90, m1_range=(5.0, 40.0), nm1=80, mdet_range=(20.0, 60.0), seed=5, resident=True)
    one = pe_redshift(seg, rbank, st, chunk=1, **kw)
    assert made == []
    shared = pe_redshift(seg, rbank, st, chunk=2, **kw)
    assert made == [2 * masses.size]          # segments 0 and 1 in one launch, segment 2 in the next
    fin = np.isfinite(one["ln_lambda"])
    np.testing.assert_array_equal(np.isfinite(shared["ln_lambda"]), fin)
    np.testing.assert_allclose(shared["ln_lambda"][fin], one["ln_lambda"][fin], rtol=0, atol=1e-12)
    np.testing.assert_allclose(shared["ln_b"], one["ln_b"], rtol=0, atol=1e-12)
    ibank = mass_bank(bankrows, seg, "hip")
    made.clear()
    one = pe_intrinsic(seg, ibank, st, chunk=1, seed=3, resident=True)
    shared = pe_intrinsic(seg, ibank, st, chunk=3, seed=3, resident=True)
    assert made == [3 * bankrows.shape[0]]
    np.testing.assert_allclose(shared["ln_lambda"], one["ln_lambda"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(shared["ln_b"], one["ln_b"], rtol=0, atol=1e-12)
    stm.free(st)


def test_batch_add_many_equals_one_add_each(hip):
    """SegmentBatch.add_many (exm_batch_add_seeded_many: the segments of one bank filtered one by one, selected
    and walked in one launch) gives the node values of add() per segment, to float64 rounding"""
    from exmarg_amd.search import mass_bank, read_segments, store as stm
    from exmarg_amd.search.batch import SegmentBatch
    from exmarg_amd.search.pe import _Analysis
    seg, bankrows, s_nodes = small_problem()
    seg = dict(seg, d_H=np.concatenate([seg["d_H"], seg["d_H"][:1] * 0.5]),
               d_L=np.concatenate([seg["d_L"], seg["d_L"][:1] * 0.5]))
    say("batch: 3 segments admitted one by one and at once")
    st = stm.synthetic_store(seed=0, s_nodes=s_nodes)
    bank = mass_bank(bankrows, seg, "hip")
    an = _Analysis(bank, st, 0.1267, True).start()
    rseg = read_segments(seg)
    sig2 = {I: rseg["sig2_" + I] for I in "HL"}
    b = SegmentBatch(st, 128, 128, max_templates=4 * bank.u.size)
    try:
        b.add(an, {I: rseg["d_" + I][0] for I in "HL"}, sig2, 5)      # a segment ahead of the group
        for j in range(3):
            b.add(an, {I: rseg["d_" + I][j] for I in "HL"}, sig2, 7 + j)
        one = b.run()
        b.reset()
        b.add(an, {I: rseg["d_" + I][0] for I in "HL"}, sig2, 5)
        assert b.add_many(an, rseg["d_H"], rseg["d_L"], sig2, [7, 8, 9]) == 1
        many = b.run()
    finally:
        b.free()
    assert len(one) == len(many) == 4
    for x, y in zip(one, many):
        fin = np.isfinite(x)
        np.testing.assert_array_equal(np.isfinite(y), fin)
        np.testing.assert_allclose(y[fin], x[fin], rtol=REPEAT_RTOL, atol=0)
    stm.free(st)


def test_pe_intrinsic_rejects_templates_outside_the_store(hip):
    from exmarg_amd.search import mass_bank, pe_intrinsic, store as stm
    seg, bankrows, _ = small_problem(n_templates=3)
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[:2])
    bank = mass_bank(bankrows, seg, "hip")
    with pytest.raises(ValueError, match="fewer than two"):
        pe_intrinsic(seg, bank, st)
    stm.free(st)


def test_batch_of_uneven_segments_matches_single_calls(hip, oracle):
    """segments of 2 s and 4 s with their own banks in one batch (capacity forcing two gather launches)
    give the same ln Lambda as pe_intrinsic one set at a time (to MIX_TOL), and match the oracle"""
    from exmarg_amd.search import mass_bank, pe_intrinsic, pe_intrinsic_batch, read_segments, store as stm
    say("batch: 2 s and 4 s segments, two banks, capacity of 3 segments")
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[::4])
    items = []
    for T, seed in ((2.0, 31), (4.0, 32)):
        seg1, _, _ = synth.simulate_at_snr(2, seed, (30.0, 0.8, 0.1), 14.0, T, 20.0, 256.0)
        seg0, _, _ = synth.simulate(1, seed + 9, lambda r: (30.0, 0.8, 0.1), lambda r: 0.0, T, 20.0, 256.0)
        seg = dict(seg1, d_H=np.concatenate([seg1["d_H"], seg0["d_H"]]), d_L=np.concatenate([seg1["d_L"], seg0["d_L"]]))
        bank = mass_bank(synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 8), seg, "hip")
        items.append((seg, bank))
    res = pe_intrinsic_batch(items, st, n_loud=128, n_random=128, seed=4, max_templates=3 * 8)
    for (seg, bank), r in zip(items, res):
        single = pe_intrinsic(seg, bank, st, n_loud=128, n_random=128, seed=4, resident=True)
        np.testing.assert_allclose(r["ln_lambda"], single["ln_lambda"], rtol=0, atol=MIX_TOL)
        b, store = oracle_inputs(bank, st)
        tabs = oracle.Tables(synth=dict(seed=0, kind="smooth", nodes=np.arange(st.s.size)), shape=st.table_shape(),
                             tau_idx=st.tau_idx)
        ref = oracle.pe_intrinsic(read_segments(seg), b, store, tabs, n_loud=128, n_random=128, seed=4)
        np.testing.assert_allclose(r["ln_lambda"], ref["ln_lambda"], rtol=0, atol=LNL_TOL)
        np.testing.assert_allclose(r["ln_b"], ref["ln_b"], rtol=0, atol=LNB_TOL)
    # the batch buffer kept with the store between calls: the same results, one buffer reused
    from exmarg_amd.search.batch import release_kept
    kw = dict(n_loud=128, n_random=128, seed=4, max_templates=3 * 8, keep_buffer=True)
    k1 = pe_intrinsic_batch(items, st, **kw)
    buf = st.__dict__["_batch"][1]
    k2 = pe_intrinsic_batch(items, st, **kw)
    assert st.__dict__["_batch"][1] is buf
    for r, a, b in zip(res, k1, k2):
        np.testing.assert_allclose(a["ln_lambda"], r["ln_lambda"], rtol=REPEAT_RTOL, atol=0)
        np.testing.assert_allclose(b["ln_lambda"], r["ln_lambda"], rtol=REPEAT_RTOL, atol=0)
    release_kept(st)
    assert "_batch" not in st.__dict__
    pe_intrinsic_batch(items[:1], st, **kw)
    stm.free(st)                                          # frees the kept buffer with the store
    assert "_batch" not in st.__dict__
