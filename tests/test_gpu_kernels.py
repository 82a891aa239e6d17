"""Parity of each HIP stage with the oracle on the same inputs (MI355X)."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


# ---------------------------------------------------------------- store
def test_synthetic_tables_match_host_recipe(hip, oracle):
    from exmarg_amd.search import store as st
    say("store: synthetic smooth+rough, reduced geometry")
    for kind in ("smooth", "rough"):
        s = st.synthetic_store(n_nodes=2, seed=5, kind=kind, shape=(7, 49, 12, 15), host=True)
        for k in range(2):
            np.testing.assert_array_equal(s.device.read_table(k), s.tables[k])
        st.free(s)


def test_synthetic_full_geometry_sample(hip):
    from exmarg_amd.search import store as st
    say("store: one node at full geometry (87, 49, 200, 219)")
    s = st.synthetic_store(n_nodes=1, seed=0)
    T = s.device.read_table(0)
    idx = np.random.default_rng(0).integers(0, T.size, 20000)
    ref = synth.table_values(0, 0, T.shape, s.tau_idx, "smooth", flat=idx)
    np.testing.assert_array_equal(T.ravel()[idx], ref)
    st.free(s)


def test_set_table_roundtrip(hip):
    from exmarg_amd.search import store as st
    say("store: upload and read back")
    arr = synth.synthetic_store_arrays(n_nodes=3, seed=2, kind="rough", shape=(5, 49, 6, 7))
    rs = st.RedshiftStore(s=arr["s"], alpha_lo=np.full(3, 0.16), alpha_hi=np.full(3, 4e4), beta=[arr["beta"]] * 3,
                          delta=[arr["delta"]] * 3, tau_idx=arr["tau_idx"], tables=arr["tables"], n_max=120)
    dev = st.upload(rs)
    for k in range(3):
        np.testing.assert_array_equal(dev.read_table(k), arr["tables"][k])
    st.free(rs)


# ---------------------------------------------------------------- filter
def _segment(T=4.0, fmin=20.0, fmax=256.0, snr=10.0, seed=3):
    seg, truth, _ = synth.simulate_at_snr(1, seed, (30.0, 0.8, 0.1), snr, T, fmin, fmax)
    seg = dict(seg, sig2_H=seg["sig2"], sig2_L=seg["sig2"])
    return seg


def test_filter_matches_fp64_oracle(hip, oracle):
    from exmarg_amd.search.bank import MassBank
    say("filter: 24 templates, T=4 s, 20-256 Hz")
    seg = _segment()
    bank_rows = synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 24)
    bank = MassBank(bank_rows[:, 0], bank_rows[:, 1], bank_rows[:, 2], seg)
    d = {I: seg["d_" + I][0] for I in "HL"}
    s2 = {I: seg["sig2"] for I in "HL"}
    got = bank.trajectories(d, s2).to_host()
    ref = oracle.filter_fp64(bank.W, bank.lo, bank.n_time, d, s2)
    for k in ("lqH", "lqL"):
        diff = np.abs(got[k].astype(np.float64) - ref[k])
        ulp = np.spacing(np.abs(ref[k])).astype(np.float64)
        assert (diff <= ulp).all(), k
        assert np.mean(got[k] == ref[k]) > 0.999, k
    for k in ("phH", "phL"):
        dph = np.angle(np.exp(1j * (got[k].astype(np.float64) - ref[k])))
        assert np.abs(dph).max() < 1e-6, k


@pytest.mark.parametrize("T,fmin,fmax", [(1.0, 20.0, 1024.0), (2.0, 20.0, 1024.0), (16.0, 20.0, 1024.0),
                                          (64.0, 20.0, 1024.0), (2.0, 100.0, 1024.0)])
def test_band_transform_matches_fp64_oracle(hip, oracle, T, fmin, fmax):
    """the default float64 filter (the band transform: two passes reading only the band) against the
    FP64 oracle (numpy's transform of the zero-padded band) and against the hipFFT path, from the
    shortest length it takes (n_time / 512 = 16) through one with a radix-2 stage to the longest segment"""
    from exmarg_amd.search.bank import MassBank
    seg = _segment(T=T, fmin=fmin, fmax=fmax, seed=5)
    bank_rows = synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 6)
    bank = MassBank(bank_rows[:, 0], bank_rows[:, 1], bank_rows[:, 2], seg)
    n1 = bank.n_time // 512
    say("filter: band transform, T=%g s, %g-%g Hz, n_time %d, band %d bins" % (T, fmin, fmax, bank.n_time,
                                                                             bank.W.shape[1]))
    assert bank.n_time % 512 == 0 and 16 <= n1 <= 2048 and n1 & (n1 - 1) == 0
    d = {I: seg["d_" + I][0] for I in "HL"}
    s2 = {I: seg["sig2"] for I in "HL"}
    got = bank.trajectories(d, s2).to_host()
    hip.call("exm_set_filter_transform", 1)
    try:
        fft = bank.trajectories(d, s2).to_host()
    finally:
        hip.call("exm_set_filter_transform", 0)
    ref = oracle.filter_fp64(bank.W, bank.lo, bank.n_time, d, s2)
    for k in ("lqH", "lqL"):
        ulp = np.spacing(np.abs(ref[k])).astype(np.float64)
        for g in (got, fft):
            assert (np.abs(g[k].astype(np.float64) - ref[k]) <= ulp).all(), k
            assert np.mean(g[k] == ref[k]) > 0.999, k
        assert (np.abs(got[k].astype(np.float64) - fft[k]) <= 2 * ulp).all(), k
    for k in ("phH", "phL"):
        dph = np.angle(np.exp(1j * (got[k].astype(np.float64) - ref[k])))
        assert np.abs(dph).max() < 1e-6, k


@pytest.mark.parametrize("T,fmin", [(1.0, 20.0), (16.0, 20.0), (2.0, 100.0)])
def test_band_transform_float32(hip, oracle, T, fmin):
    """the float32 band transform (float32 form of the filter) against the FP64 oracle with the float32
    transform's bar (ln|z| and the phase within 2e-5 where |z| is above a hundredth of its maximum), and
    against the float32 hipFFT path on the same bar"""
    from exmarg_amd.search.bank import MassBank
    seg = _segment(T=T, fmin=fmin, fmax=1024.0, seed=6)
    bank_rows = synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 6)
    bank = MassBank(bank_rows[:, 0], bank_rows[:, 1], bank_rows[:, 2], seg)
    say("filter: float32 band transform, T=%g s, n_time %d" % (T, bank.n_time))
    d = {I: seg["d_" + I][0] for I in "HL"}
    s2 = {I: seg["sig2"] for I in "HL"}
    hip.call("exm_set_filter_precision", 1)
    try:
        got = bank.trajectories(d, s2).to_host()
        hip.call("exm_set_filter_transform", 1)
        fft = bank.trajectories(d, s2).to_host()
    finally:
        hip.call("exm_set_filter_transform", 0)
        hip.call("exm_set_filter_precision", 0)
    ref = oracle.filter_fp64(bank.W, bank.lo, bank.n_time, d, s2)
    for I in "HL":
        lq, ph = ref["lq" + I], ref["ph" + I]
        big = lq > lq.max(axis=1, keepdims=True) - np.log(1e2)
        for g in (got, fft):
            assert np.abs(g["lq" + I].astype(np.float64) - lq)[big].max() < 2e-5
            dph = np.angle(np.exp(1j * (g["ph" + I].astype(np.float64) - ph)))
            assert np.abs(dph[big]).max() < 2e-5
        assert np.abs(got["lq" + I].astype(np.float64) - lq).max() > 0.0      # a different computation


def test_filter_transform_choice_is_checked(hip):
    with pytest.raises(ValueError, match="filter transform"):
        hip.call("exm_set_filter_transform", 2)


def test_filter_float32_transform(hip, oracle):
    """the float32 transform (precision form) against the FP64 oracle: ln|z| and the phase within 2e-5
    wherever |z| is above a hundredth of its maximum (the transform's float32 error is relative to the
    largest values, so it grows as 1 / |z| below)"""
    from exmarg_amd.search.bank import MassBank
    say("filter: float32 transform, 24 templates, T=4 s, 20-256 Hz")
    seg = _segment()
    bank_rows = synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 24)
    bank = MassBank(bank_rows[:, 0], bank_rows[:, 1], bank_rows[:, 2], seg)
    d = {I: seg["d_" + I][0] for I in "HL"}
    s2 = {I: seg["sig2"] for I in "HL"}
    hip.call("exm_set_filter_precision", 1)
    try:
        got = bank.trajectories(d, s2).to_host()
    finally:
        hip.call("exm_set_filter_precision", 0)
    ref = oracle.filter_fp64(bank.W, bank.lo, bank.n_time, d, s2)
    for I in "HL":
        lq, ph = ref["lq" + I], ref["ph" + I]
        big = lq > lq.max(axis=1, keepdims=True) - np.log(1e2)
        assert np.abs(got["lq" + I].astype(np.float64) - lq)[big].max() < 2e-5
        dph = np.angle(np.exp(1j * (got["ph" + I].astype(np.float64) - ph)))
        assert np.abs(dph[big]).max() < 2e-5
        assert np.abs(got["lq" + I].astype(np.float64) - lq).max() > 0.0      # a different computation


def test_filter_padding_follows_the_band(hip, oracle):
    """the transform input keeps its zero padding between calls; banks whose frequency bands differ
    (same transform length), and the float32 transform in between, must each still match the oracle"""
    from exmarg_amd.search.bank import MassBank
    say("filter: alternating frequency bands on one transform length")
    bank_rows = synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 8)
    cases = [(20.0, 256.0), (60.0, 200.0), (20.0, 256.0), (30.0, 128.0)]
    for i, (fmin, fmax) in enumerate(cases):
        seg = _segment(fmin=fmin, fmax=fmax, seed=11 + i)
        bank = MassBank(bank_rows[:, 0], bank_rows[:, 1], bank_rows[:, 2], seg)
        d = {I: seg["d_" + I][0] for I in "HL"}
        s2 = {I: seg["sig2"] for I in "HL"}
        if i == 2:                                        # a float32 transform between two float64 ones
            hip.call("exm_set_filter_precision", 1)
            try:
                bank.trajectories(d, s2).free()
            finally:
                hip.call("exm_set_filter_precision", 0)
        got = bank.trajectories(d, s2).to_host()
        ref = oracle.filter_fp64(bank.W, bank.lo, bank.n_time, d, s2)
        for k in ("lqH", "lqL"):
            diff = np.abs(got[k].astype(np.float64) - ref[k])
            assert (diff <= np.spacing(np.abs(ref[k]))).all(), (fmin, fmax, k)


def test_filter_single_bin_closed_form(hip):
    from exmarg_amd.search.bank import MassBank
    say("filter: single-bin closed form")
    nu = np.arange(40, 72)
    seg = dict(f=nu / 2.0, nu=nu, T=2.0, fmin=20.0, fmax=36.0, sig2_H=np.full(nu.size, 4.0))
    W = np.zeros((1, nu.size), np.complex64)
    W[0, 5] = 0.5 - 0.25j
    bank = MassBank([30.0], [1.0], [0.0], seg, factors=(W, np.ones(1)))
    d = np.full(nu.size, 2.0 - 1.0j)
    got = bank.trajectories({"H": d, "L": d}, {"H": seg["sig2_H"], "L": seg["sig2_H"]}).to_host()
    w = complex(W[0, 5]) * np.conj(2.0 - 1.0j) / 4.0
    np.testing.assert_allclose(got["lqH"][0], np.log(abs(w)), atol=1e-6)
    t = np.arange(bank.n_time)
    ph = np.angle(w) - 2 * np.pi * 45 * t / bank.n_time
    assert np.abs(np.angle(np.exp(1j * (got["phH"][0] - ph)))).max() < 2e-6


# ---------------------------------------------------------------- select_times
def _rand_traj(nm, nt, seed):
    rng = np.random.default_rng(seed)
    return {"lqH": rng.standard_normal((nm, nt)).astype(np.float32),
            "phH": rng.uniform(-np.pi, np.pi, (nm, nt)).astype(np.float32),
            "lqL": rng.standard_normal((nm, nt)).astype(np.float32),
            "phL": rng.uniform(-np.pi, np.pi, (nm, nt)).astype(np.float32)}


@pytest.mark.parametrize("nm,nt,K,S,lag", [(5, 4096, 64, 100, 43), (3, 20000, 1024, 1024, 43), (4, 300, 299, 5, 3),
                                           (2, 131072, 1024, 1024, 43), (2, 9000, 4000, 512, 1),
                                           (4, 524288, 1024, 1024, 43)])
def test_select_matches_oracle(hip, oracle, nm, nt, K, S, lag):
    from exmarg_amd.search import gather as g
    say("select: nm=%d nt=%d K=%d S=%d" % (nm, nt, K, S))
    traj = _rand_traj(nm, nt, seed=nt)
    tau = synth.lag_indices(lag)
    sel, lw = g.select_times(traj, tau, K, S, seed=7).to_host()
    rsel, rlw = oracle.select_times(traj, tau, K, S, seed=7)
    np.testing.assert_array_equal(sel, rsel)
    np.testing.assert_array_equal(lw, rlw)


def _sample_layout(nt, K, high, seed):
    """template scores e^{2 lqL} (lqH far below) whose samples, the keys at every stride-th time (stride
    the least power of two with 2048 stride >= nt), take the values v_i = 10 - 0.05 i (i < 64) and -5
    elsewhere; for each (f, n) in high, n keys at times off the samples take the value v_rank(f), where
    rank(f) is the sample rank at an expected count of f K as the candidate pass ranks them (2.2 K its
    threshold, 1.3, 1.7, 2.8, 3.5 K the alternates)"""
    stride = 1
    while 2048 * stride < nt:
        stride *= 2
    rank = lambda f: int(f * K / stride + 0.5) - 1
    lqL = np.full(nt, -5.0, np.float32)
    v = (10.0 - 0.05 * np.arange(64)).astype(np.float32)
    lqL[np.arange(64) * stride] = v
    off = np.flatnonzero(np.arange(nt) % stride != 0)
    np.random.default_rng(seed).shuffle(off)
    o = 0
    for f, n in high:
        lqL[off[o:o + n]] = v[rank(f)]
        o += n
    return lqL, stride, rank


@pytest.mark.parametrize("layout", ["off_sample", "one_template_flat", "retry_high", "retry_low"])
def test_select_threshold_fallbacks(hip, oracle, layout):
    """scores arranged so that the sample threshold catches too many or too few keys: the exact
    selection must give the oracle's times all the same.  retry_high: 3000 keys off the samples just above
    the threshold push its count past the 4096 candidates while the highest alternate holds 1500 + 10;
    retry_low: the threshold catches 18 keys and the third alternate 1500 + 22; both take the second pass
    at an alternate"""
    from exmarg_amd.search import gather as g
    say("select: %s" % layout)
    nm, nt = 2, 65536
    if layout.startswith("retry"):
        nt = 262144                               # samples sparse enough for the alternates (stride 128)
    traj = _rand_traj(nm, nt, seed=4)
    if layout == "off_sample":
        # loud times only where no sample is taken (every time but multiples of 32)
        traj["lqL"][:, np.arange(nt) % 32 != 0] += 6.0
    elif layout == "one_template_flat":
        traj["lqL"][0] = 0.25                     # template 0: many equal scores around the threshold
        traj["lqH"][0] = -3.0
    else:
        K = 1024
        if layout == "retry_high":
            # 1500 keys at the value of sample rank(1.3) - 2, 3000 between rank(1.3) and the threshold's
            spec = [(1.0, 1500), (1.7, 3000)]
        else:
            spec = [(2.75, 1500)]                 # between the threshold's sample and the third alternate's
        lqL, stride, rank = _sample_layout(nt, K, spec, seed=11)
        assert stride == 128
        if layout == "retry_high":
            assert rank(1.0) < rank(1.3) < rank(1.7) < rank(2.2)
        else:
            assert rank(2.2) < rank(2.75) <= rank(2.8)
        traj["lqL"][0] = lqL
        traj["lqH"][0] = -20.0
    tau = synth.lag_indices(43)
    sel, lw = g.select_times(traj, tau, 1024, 512, seed=3).to_host()
    rsel, rlw = oracle.select_times(traj, tau, 1024, 512, seed=3)
    np.testing.assert_array_equal(sel, rsel)
    np.testing.assert_array_equal(lw, rlw)


def test_select_ties_take_earliest(hip, oracle):
    from exmarg_amd.search import gather as g
    say("select: all-equal scores")
    traj = {k: np.zeros((2, 1000), np.float32) for k in ("lqH", "phH", "lqL", "phL")}
    sel, lw = g.select_times(traj, synth.lag_indices(5), 100, 50, seed=1).to_host()
    np.testing.assert_array_equal(sel[:, :100][lw[:, :100] == 0].reshape(2, -1)[:, :10], np.tile(np.arange(10), (2, 1)))
    rsel, rlw = oracle.select_times(traj, synth.lag_indices(5), 100, 50, seed=1)
    np.testing.assert_array_equal(sel, rsel)
    # all equal over more times than the candidate sort holds
    traj = {k: np.zeros((1, 20000), np.float32) for k in ("lqH", "phH", "lqL", "phL")}
    sel, lw = g.select_times(traj, synth.lag_indices(5), 100, 50, seed=1).to_host()
    rsel, rlw = oracle.select_times(traj, synth.lag_indices(5), 100, 50, seed=1)
    np.testing.assert_array_equal(sel, rsel)


def test_select_rejects_bad_arguments(hip):
    from exmarg_amd.search import gather as g
    traj = _rand_traj(1, 100, 0)
    with pytest.raises(ValueError):
        g.select_times(traj, synth.lag_indices(3), 100, 5, seed=0)
    with pytest.raises(ValueError):
        g.select_times(traj, np.array([0, 1, 3]), 10, 5, seed=0)


# ---------------------------------------------------------------- gather
def _tiny_store(kind, shape, n_nodes=2, seed=1):
    from exmarg_amd.search import store as st
    return st.synthetic_store(n_nodes=n_nodes, seed=seed, kind=kind, shape=shape, host=True)


KERNELS = {1: 1e-12, 2: 1e-6, 3: 1e-6, 4: 1e-6}     # reference-order kernel: same operation order as the oracle; band kernel


@pytest.fixture
def kernel(hip, request):
    hip.call("exm_set_gather_kernel", request.param)
    yield request.param
    hip.call("exm_set_gather_kernel", 0)


@pytest.mark.parametrize("kernel", [1, 2, 3, 4], indirect=True)
@pytest.mark.parametrize("kind", ["smooth", "rough"])
@pytest.mark.parametrize("subsampled", [False, True])
def test_gather_matches_oracle(hip, oracle, kernel, kind, subsampled):
    from exmarg_amd.search import gather as g
    say("gather: kernel=%d kind=%s subsampled=%s" % (kernel, kind, subsampled))
    shape = (9, 49, 40, 31)
    s = _tiny_store(kind, shape)
    nm, nt = 6, 2048
    traj = _rand_traj(nm, nt, seed=21)
    traj["lqH"] *= 2.0
    traj["lqL"] *= 2.0
    tau = s.tau_idx
    times_dev = g.select_times(traj, tau, 64, 96, seed=3) if subsampled else None
    times = times_dev.to_host() if subsampled else None
    ids = np.array([0, 2, 3, 5])
    lna = np.array([[0.5, -2.0], [1.0, 3.0], [-4.0, 0.1], [2.5, 2.5]])
    tab = np.array([[0, 1], [1, 0], [0, 0], [1, 1]])
    beta, delta = s.beta[0], s.delta[0]
    got = g.gather(traj, s, ids, lna, tab, times=times_dev)
    tabs = oracle.Tables(tables=s.tables, tau_idx=tau)
    ref = oracle.gather(traj, tabs, beta, delta, ids, lna, tab, times)
    np.testing.assert_allclose(got, ref, rtol=0, atol=KERNELS[kernel])


@pytest.fixture
def piece(hip, request):
    hip.call("exm_set_gather_piece", request.param)
    yield request.param
    hip.call("exm_set_gather_piece", 0)


@pytest.mark.parametrize("kernel,piece", [(2, 4), (2, 8), (2, 16), (3, 4), (3, 6), (3, 8), (4, 4), (4, 6), (4, 8)],
                         indirect=True)
def test_band_gather_clamps_and_many_reads(hip, oracle, kernel, piece):
    # beta clamped at both table edges, Delta clamped, more reads of one table than one tile holds
    from exmarg_amd.search import gather as g
    say("gather: band kernel %d, piece %d, clamped rows, 1500 reads of one table" % (kernel, piece))
    s = _tiny_store("rough", (5, 49, 24, 17), n_nodes=1, seed=4)
    nm, nt = 3, 4096
    traj = _rand_traj(nm, nt, seed=8)
    traj["lqH"] *= 4.0
    traj["lqL"] *= 4.0
    times_dev = g.select_times(traj, s.tau_idx, 300, 700, seed=9)
    times = times_dev.to_host()
    rng = np.random.default_rng(1)
    ids = rng.integers(0, nm, 1500)
    lna = rng.uniform(-12.0, 12.0, (1500, 1))
    tab = np.zeros((1500, 1), np.int64)
    got = g.gather(traj, s, ids, lna, tab, times=times_dev)
    ref = oracle.gather(traj, oracle.Tables(tables=s.tables, tau_idx=s.tau_idx), s.beta[0], s.delta[0], ids, lna,
                        tab, times)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)


@pytest.fixture
def tile(hip, request):
    hip.call("exm_set_gather_tile", request.param)
    yield request.param
    hip.call("exm_set_gather_tile", 0)


@pytest.mark.parametrize("kernel", [2, 3, 4], indirect=True)
@pytest.mark.parametrize("tile", [256, 512, 2048], indirect=True)
def test_band_gather_tiles_over_xcd_queues(hip, oracle, kernel, tile):
    # two tables cut into 9 and 3 tiles (256 reads per tile), 5 and 2 (512) or 2 and 1 (2048, each cut in
    # two for the wave and piece forms): the tile list is built from per-XCD queues, piece q of every
    # table in one queue (>= 8 pieces) or shared by several (< 8); every read must be gathered exactly once
    from exmarg_amd.search import gather as g
    say("gather: band kernel %d, tile %d, 2300 + 700 reads of two tables" % (kernel, tile))
    s = _tiny_store("rough", (5, 49, 24, 17), n_nodes=2, seed=6)
    nm, nt = 3, 4096
    traj = _rand_traj(nm, nt, seed=12)
    traj["lqH"] *= 3.0
    traj["lqL"] *= 3.0
    times_dev = g.select_times(traj, s.tau_idx, 300, 700, seed=2)
    times = times_dev.to_host()
    rng = np.random.default_rng(7)
    n = 3000
    ids = rng.integers(0, nm, n)
    lna = rng.uniform(-10.0, 10.0, (n, 1))
    tab = np.zeros((n, 1), np.int64)
    tab[rng.permutation(n)[:700], 0] = 1
    got = g.gather(traj, s, ids, lna, tab, times=times_dev)
    ref = oracle.gather(traj, oracle.Tables(tables=s.tables, tau_idx=s.tau_idx), s.beta[0], s.delta[0], ids, lna,
                        tab, times)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)


@pytest.mark.parametrize("form", [2, 3, 4])
@pytest.mark.parametrize("kind", ["smooth", "rough"])
def test_band_gather_float32_form(hip, oracle, kind, form):
    """float32 table interpolation (index arithmetic and log-sum-exp in float64) stays within 2e-6
    of the float64 oracle; the float64 form of the same reads (float32 exp inside the log-sum-exp)
    stays within 1e-7 of it"""
    from exmarg_amd.search import gather as g
    say("gather: band kernel %d, float32 interpolation, %s table" % (form, kind))
    s = _tiny_store(kind, (7, 49, 30, 23), n_nodes=2, seed=6)
    traj = _rand_traj(4, 4096, seed=12)
    traj["lqH"] *= 3.0
    traj["lqL"] *= 3.0
    times_dev = g.select_times(traj, s.tau_idx, 200, 300, seed=2)
    times = times_dev.to_host()
    rng = np.random.default_rng(3)
    ids = rng.integers(0, 4, 300)
    lna = rng.uniform(-6.0, 6.0, (300, 2))
    tab = rng.integers(0, 2, (300, 2))
    ref = oracle.gather(traj, oracle.Tables(tables=s.tables, tau_idx=s.tau_idx), s.beta[0], s.delta[0], ids, lna,
                        tab, times)
    hip.call("exm_set_gather_kernel", form)
    try:
        hip.call("exm_set_gather_precision", 1)
        got32 = g.gather(traj, s, ids, lna, tab, times=times_dev)
        hip.call("exm_set_gather_precision", 0)
        got64 = g.gather(traj, s, ids, lna, tab, times=times_dev)
    finally:
        hip.call("exm_set_gather_precision", 0)
        hip.call("exm_set_gather_kernel", 0)
    np.testing.assert_allclose(got64, ref, rtol=0, atol=1e-7)
    np.testing.assert_allclose(got32, ref, rtol=0, atol=2e-6)
    assert np.abs(got32 - ref).max() > 0.0          # the float32 form is a different computation


def test_gather_full_geometry_one_node(hip, oracle):
    from exmarg_amd.search import gather as g
    from exmarg_amd.search import store as st
    say("gather: full table geometry, synthetic node, 4 templates x 8192 times (every time, band kernel)")
    s = st.synthetic_store(n_nodes=1, seed=0)
    traj = _rand_traj(4, 8192, seed=2)
    ids = np.arange(4)
    lna = np.array([[-2.0], [0.0], [3.0], [6.0]])
    tab = np.zeros((4, 1), np.int64)
    got = g.gather(traj, s, ids, lna, tab)
    tabs = oracle.Tables(synth=dict(seed=0, kind="smooth", nodes=[0]), shape=s.table_shape(), tau_idx=s.tau_idx)
    ref = oracle.gather(traj, tabs, s.beta[0], s.delta[0], ids, lna, tab)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)
    st.free(s)


@pytest.mark.parametrize("kernel", [2, 3, 4], indirect=True)
@pytest.mark.parametrize("walk", ["every_time", "long_selection"])
def test_band_gather_in_slices(hip, oracle, kernel, walk):
    """walks longer than one band launch takes (every arrival time of 20000, or a selection of 10000
    times) run in slices of 8192 walked times whose (max, sum) per read and lag are merged"""
    from exmarg_amd.search import gather as g
    say("gather: band kernel %d, %s walk in slices" % (kernel, walk))
    s = _tiny_store("rough", (7, 49, 30, 25), n_nodes=2, seed=17)
    nm, nt = 3, 20000
    traj = _rand_traj(nm, nt, seed=19)
    traj["lqH"] *= 2.5
    traj["lqL"] *= 2.5
    times_dev = g.select_times(traj, s.tau_idx, 5000, 5000, seed=2) if walk == "long_selection" else None
    times = times_dev.to_host() if times_dev is not None else None
    rng = np.random.default_rng(23)
    nr = 40
    ids = rng.integers(0, nm, nr)
    lna = rng.uniform(-6.0, 6.0, (nr, 1))
    tab = rng.integers(0, 2, (nr, 1))
    got = g.gather(traj, s, ids, lna, tab, times=times_dev)
    ref = oracle.gather(traj, oracle.Tables(tables=s.tables, tau_idx=s.tau_idx), s.beta[0], s.delta[0], ids, lna,
                        tab, times)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)


@pytest.mark.parametrize("kernel,piece", [(2, 8), (3, 4), (3, 6), (3, 8), (4, 4), (4, 6), (4, 8)], indirect=True)
@pytest.mark.parametrize("n_walk", [8192, 24])
def test_band_gather_skewed_rows(hip, oracle, kernel, piece, n_walk):
    """reads whose whole walk falls in one clamped row (one read's pieces fill whole rounds of lanes)
    beside reads spread over many rows and reads with a single piece per row, at the longest walk and a
    short one"""
    from exmarg_amd.search import gather as g
    say("gather: band kernel %d, piece %d, skewed rows, walk %d" % (kernel, piece, n_walk))
    s = _tiny_store("rough", (5, 49, 24, 17), n_nodes=1, seed=11)
    nm = 4
    traj = _rand_traj(nm, 16384, seed=13)
    traj["lqH"] *= 3.0
    traj["lqL"] *= 3.0
    times_dev = g.select_times(traj, s.tau_idx, n_walk // 2, n_walk // 2, seed=5)
    times = times_dev.to_host()
    rng = np.random.default_rng(7)
    nr = 150
    ids = rng.integers(0, nm, nr)
    lna = rng.uniform(-5.0, 5.0, (nr, 1))
    lna[::3] = -40.0                                  # every walked time in the lowest row
    lna[1::7] = 40.0                                  # every walked time in the highest row
    tab = np.zeros((nr, 1), np.int64)
    got = g.gather(traj, s, ids, lna, tab, times=times_dev)
    ref = oracle.gather(traj, oracle.Tables(tables=s.tables, tau_idx=s.tau_idx), s.beta[0], s.delta[0], ids, lna,
                        tab, times)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)


@pytest.mark.parametrize("kernel,piece", [(2, 8), (3, 4), (3, 6), (3, 8), (4, 4), (4, 6), (4, 8)], indirect=True)
def test_band_gather_negative_log_weight(hip, oracle, kernel, piece):
    """more random draws than times outside the top K: the drawn times' log weight ln((nt - K) / S)
    (gather.py:433-435) is negative, so the drawn terms weigh less than the loud ones"""
    from exmarg_amd.search import gather as g
    say("gather: band kernel %d, piece %d, negative log weight" % (kernel, piece))
    s = _tiny_store("rough", (5, 49, 24, 17), n_nodes=1, seed=12)
    nm, nt, K, S = 3, 2048, 1024, 4096
    traj = _rand_traj(nm, nt, seed=14)
    traj["lqH"] *= 3.0
    traj["lqL"] *= 3.0
    times_dev = g.select_times(traj, s.tau_idx, K, S, seed=6)
    times = times_dev.to_host()
    assert np.all(times[1][:, :] <= 0.0) and np.any(times[1] < -1.0)
    rng = np.random.default_rng(9)
    nr = 200
    ids = rng.integers(0, nm, nr)
    lna = rng.uniform(-5.0, 5.0, (nr, 1))
    tab = np.zeros((nr, 1), np.int64)
    got = g.gather(traj, s, ids, lna, tab, times=times_dev)
    ref = oracle.gather(traj, oracle.Tables(tables=s.tables, tau_idx=s.tau_idx), s.beta[0], s.delta[0], ids, lna,
                        tab, times)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)


@pytest.mark.parametrize("kernel", [3, 4], indirect=True)
@pytest.mark.parametrize("loud", ["low_rows", "high_rows"])
def test_band_gather_values_spanning_beyond_float64(hip, oracle, kernel, loud):
    """table values 1500 above the rest in the lower (or the upper) half of the beta rows: a read's
    terms then span e^1500, beyond float64.  The scheduled form sums each read against the value of its
    loudest walked time; with the lower rows loud that reference lies ~1500 below the largest terms, the
    sum overflows and the (read, lag) is flagged and recomputed exactly, and with the upper rows loud the
    lower rows' terms vanish against it.  Both must match the oracle's online log-sum-exp."""
    from exmarg_amd.search import gather as g
    from exmarg_amd.search import store as st
    say("gather: band kernel %d, values spanning 1500 (%s)" % (kernel, loud))
    arr = synth.synthetic_store_arrays(n_nodes=1, seed=4, kind="rough", shape=(5, 49, 24, 17))
    T = arr["tables"][0]
    nb = T.shape[2]
    rows = np.arange(nb) < nb // 2 if loud == "low_rows" else np.arange(nb) >= nb // 2
    T = (T + np.where(rows, 1500.0, 0.0)[None, None, :, None]).astype(np.float32)
    rs = st.RedshiftStore(s=arr["s"], alpha_lo=np.full(1, 0.16), alpha_hi=np.full(1, 4e4), beta=[arr["beta"]],
                          delta=[arr["delta"]], tau_idx=arr["tau_idx"], tables=[T], n_max=120)
    st.upload(rs)
    try:
        nm = 3
        traj = _rand_traj(nm, 4096, seed=31)
        traj["lqH"] *= 3.0
        traj["lqL"] *= 3.0
        times_dev = g.select_times(traj, rs.tau_idx, 300, 500, seed=4)
        times = times_dev.to_host()
        rng = np.random.default_rng(5)
        nr = 60
        ids = rng.integers(0, nm, nr)
        lna = rng.uniform(-2.0, 2.0, (nr, 1))
        tab = np.zeros((nr, 1), np.int64)
        got = g.gather(traj, rs, ids, lna, tab, times=times_dev)
        ref = oracle.gather(traj, oracle.Tables(tables=[T], tau_idx=rs.tau_idx), rs.beta[0], rs.delta[0], ids, lna,
                            tab, times)
        assert np.all(ref > 1000.0)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)
    finally:
        st.free(rs)


def test_band_gather_repeat_agreement(hip):
    """the wave form (kernel 3) adds in a fixed order and repeats bit for bit; the scheduled form (kernel 4)
    adds with float64 atomics in the order waves reach them, so its repeats differ only in the last bits.
    The two forms round differently (float32 exponentials of terms against different references), and
    agree with each other well inside the oracle tolerance of 1e-6"""
    from exmarg_amd.search import gather as g
    say("gather: repeats of the wave and scheduled forms")
    s = _tiny_store("rough", (9, 49, 40, 31), n_nodes=2, seed=6)
    traj = _rand_traj(5, 8192, seed=40)
    traj["lqH"] *= 2.0
    traj["lqL"] *= 2.0
    times = g.select_times(traj, s.tau_idx, 512, 700, seed=8)
    rng = np.random.default_rng(9)
    nr = 300
    ids = rng.integers(0, 5, nr)
    lna = rng.uniform(-3.0, 3.0, (nr, 2))
    tab = rng.integers(0, 2, (nr, 2))
    out = {}
    try:
        for k in (3, 4):
            hip.call("exm_set_gather_kernel", k)
            out[k] = [g.gather(traj, s, ids, lna, tab, times=times) for _ in range(3)]
    finally:
        hip.call("exm_set_gather_kernel", 0)
    for r in out[3][1:]:
        np.testing.assert_array_equal(r, out[3][0])
    for r in out[4][1:]:
        np.testing.assert_allclose(r, out[4][0], rtol=1e-13, atol=0)
    np.testing.assert_allclose(out[4][0], out[3][0], rtol=0, atol=1e-7)


def test_wave_form_rejects_pieces_of_16(hip):
    from exmarg_amd.search import gather as g
    s = _tiny_store("smooth", (5, 49, 8, 9), n_nodes=1)
    traj = _rand_traj(2, 256, 0)
    times = g.select_times(traj, s.tau_idx, 16, 16, seed=1)
    hip.call("exm_set_gather_kernel", 3)
    hip.call("exm_set_gather_piece", 16)
    try:
        with pytest.raises(ValueError, match="4, 6 or 8"):
            g.gather(traj, s, np.array([0]), np.array([[0.0]]), np.array([[0]]), times=times)
    finally:
        hip.call("exm_set_gather_piece", 0)
        hip.call("exm_set_gather_kernel", 0)


def test_gather_rejects_bad_ids(hip):
    from exmarg_amd.search import gather as g
    s = _tiny_store("smooth", (5, 49, 8, 9), n_nodes=1)
    traj = _rand_traj(2, 64, 0)
    with pytest.raises(ValueError):
        g.gather(traj, s, np.array([5]), np.zeros((1, 1)), np.zeros((1, 1), np.int64))
    with pytest.raises(ValueError):
        g.gather(traj, s, np.array([0]), np.zeros((1, 1)), np.ones((1, 1), np.int64))


@pytest.mark.parametrize("seed,n,high", [(5, 2_048_000, 130048), (0, 100_000, 3 << 30), (123, 9999, 7),
                                         (9, 30_000, (1 << 32) - 5), (77, 1, 2), (4, 50_000, 1)])
def test_device_draw_equals_numpy(hip, seed, n, high):
    """the device draw is numpy's default_rng(seed).integers(0, high, n) bit for bit, at rejection rates
    from none to a quarter of the words"""
    from exmarg_amd.search.gather import rng_state
    say("draw: seed %d, %d values below %d" % (seed, n, high))
    out = np.empty(n, np.int64)
    hip.call("exm_draw_integers", hip.ptr(rng_state(seed)), n, high, hip.ptr(out))
    np.testing.assert_array_equal(out, np.random.default_rng(seed).integers(0, high, size=n))
