"""The serving loop: each request's result is the same alone and in a mix of uneven requests
admitted while other gathers run, equals pe_intrinsic on the same inputs, and matches the oracle."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu

LNB_TOL = 1e-5
# alone vs in the mix: a read's lookups are grouped with float32 exps of non-positive differences, and
# the grouping depends on the neighbouring reads of its workgroup
MIX_TOL = 1e-7


def say(*a):
    print(*a, flush=True)


def requests():
    """four segments of 2 s and 4 s with banks of 8 to 24 templates"""
    out = []
    for T, seed, n, snr in ((4.0, 41, 12, 15.0), (2.0, 42, 8, 0.0), (4.0, 43, 24, 9.0), (2.0, 44, 16, 20.0)):
        if snr > 0:
            seg, _, _ = synth.simulate_at_snr(1, seed, (30.0, 0.8, 0.1), snr, T, 20.0, 256.0)
        else:
            seg, _, _ = synth.simulate(1, seed, lambda r: (30.0, 0.8, 0.1), lambda r: 0.0, T, 20.0, 256.0)
        out.append((seg, synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, n), seed))
    return out


@pytest.fixture(scope="module")
def store(hip):
    from exmarg_amd.search import store as stm
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[::4])
    yield st
    stm.free(st)


def serve(store, reqs, slot_templates):
    from exmarg_amd.search import Server
    srv = Server(store, n_loud=128, n_random=128, slot_templates=slot_templates, max_request=(24, 32768, 1000))
    handles = [srv.submit(seg, rows, seed=seed) for seg, rows, seed in reqs]
    srv.run()
    stats = dict(srv.stats)
    srv.close()
    for r in handles:
        assert r.error is None, r.error
    return handles, stats


def test_alone_and_in_the_mix(store):
    reqs = requests()
    say("serve: 4 requests alone, then together in slots of 30 templates")
    alone = [serve(store, [r], 30)[0][0] for r in reqs]
    mixed, stats = serve(store, reqs, 30)
    say("stats", stats)
    assert stats["batches"] >= 2 and stats["admitted_during_gather"] >= 1
    for a, m in zip(alone, mixed):
        np.testing.assert_allclose(m.ln_lambda, a.ln_lambda, rtol=0, atol=MIX_TOL)
        assert abs(m.ln_b - a.ln_b) <= MIX_TOL
        assert m.latency is not None and m.latency >= 0


def test_served_equals_pe_intrinsic_and_oracle(store, oracle):
    from exmarg_amd.search import mass_bank, pe_intrinsic, read_segments
    reqs = requests()
    mixed, _ = serve(store, reqs, 60)
    for (seg, rows, seed), r in zip(reqs, mixed):
        bank = mass_bank(rows, seg, "hip")
        ref = pe_intrinsic(seg, bank, store, n_loud=128, n_random=128, seed=seed, resident=True)
        np.testing.assert_allclose(r.ln_lambda, ref["ln_lambda"][0], rtol=0, atol=MIX_TOL)
    seg, rows, seed = reqs[0]
    bank = mass_bank(rows, seg, "hip")
    b = dict(W=bank.W, u=bank.u, M=bank.M, q=bank.q, chi=bank.chi, lo=bank.lo, n_time=bank.n_time)
    st = dict(s=store.s, alpha_lo=store.alpha_lo, alpha_hi=store.alpha_hi, beta=store.beta[0], delta=store.delta[0],
              tau_idx=store.tau_idx)
    tabs = oracle.Tables(synth=dict(seed=0, kind="smooth", nodes=np.arange(store.s.size)), shape=store.table_shape(),
                         tau_idx=store.tau_idx)
    ref = oracle.pe_intrinsic(read_segments(seg), b, st, tabs, n_loud=128, n_random=128, seed=seed)
    np.testing.assert_allclose(mixed[0].ln_lambda, ref["ln_lambda"][0], rtol=0, atol=LNB_TOL)
    assert abs(mixed[0].ln_b - ref["ln_b"][0]) <= LNB_TOL


def test_request_larger_than_a_slot_is_refused(store):
    reqs = requests()
    from exmarg_amd.search import Server
    srv = Server(store, n_loud=128, n_random=128, slot_templates=10)
    big = srv.submit(*reqs[2][:2], seed=1)
    small = srv.submit(*reqs[1][:2], seed=1)
    srv.run()
    srv.close()
    assert isinstance(big.error, MemoryError)
    assert small.error is None and np.isfinite(small.ln_b)


def test_repeated_banks_are_cached(store):
    """requests that repeat a bank (same rows, segment length, band and noise curve) share one device
    bank: built once, results as with a bank of their own; with no cache budget a bank is shared only
    while a request still uses it, and none is held once every request has ended"""
    from exmarg_amd.search import Server
    reqs = requests()
    seg, rows, _ = reqs[0]
    work = [(seg, rows, 5), reqs[1], (seg, rows, 6), (seg, rows, 5)]
    say("serve: a bank repeated three times among four requests")
    for budget, builds in ((8 << 30, 2), (0, None)):
        srv = Server(store, n_loud=128, n_random=128, slot_templates=30, max_request=(24, 32768, 1000),
                     bank_cache_bytes=budget)
        handles = [srv.submit(s, r, seed=sd) for s, r, sd in work]
        srv.run()
        stats = dict(srv.stats)
        held = len(srv.banks)
        srv.close()
        say("budget", budget, "stats", stats, "banks held", held)
        assert stats["bank_builds"] + stats["bank_hits"] == 4
        if budget:
            assert stats["bank_builds"] == builds and held == 2
        else:
            assert stats["bank_builds"] >= 2 and held == 0
        for h in handles:
            assert h.error is None and np.isfinite(h.ln_b)
        alone = serve(store, [(seg, rows, 5)], 30)[0][0]
        for h in (handles[0], handles[3]):                  # same inputs as the request served alone
            np.testing.assert_allclose(h.ln_lambda, alone.ln_lambda, rtol=0, atol=MIX_TOL)
        assert abs(handles[2].ln_b - handles[0].ln_b) < 1.0       # another draw, the same bank


def test_bad_requests_do_not_hold_up_others(store):
    """a segment too short for the walk (n_time <= n_loud) and a bank row the waveform model rejects
    (q > 1), queued ahead of a valid request, are refused with their own errors; the valid request is
    served, and the refused ones hold no bank"""
    from exmarg_amd.search import Server
    reqs = requests()
    short, _, _ = synth.simulate(1, 45, lambda r: (30.0, 0.8, 0.1), lambda r: 0.0, 0.25, 20.0, 64.0)
    bad_rows = np.array([[30.0, 1.5, 0.1], [31.0, 0.8, 0.1]])
    srv = Server(store, n_loud=128, n_random=128, slot_templates=30, max_request=(24, 32768, 1000),
                 bank_cache_bytes=0)
    try:
        h_short = srv.submit(short, reqs[1][1], seed=1)
        h_bad = srv.submit(reqs[1][0], bad_rows, seed=1)
        h_good = srv.submit(*reqs[1][:2], seed=1)
        srv.run()
        held = len(srv.banks)
        stats = dict(srv.stats)
    finally:
        srv.close()
    say("short:", h_short.error, "| bad row:", h_bad.error, "| stats", stats)
    assert isinstance(h_short.error, ValueError) and "n_loud" in str(h_short.error)
    assert isinstance(h_bad.error, ValueError)
    assert h_good.error is None and np.isfinite(h_good.ln_b)
    assert stats["refused"] == 2 and stats["finished"] == 1 and held == 0
