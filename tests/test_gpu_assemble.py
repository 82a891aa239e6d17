"""The float64 assembly on the device: the distance-prior weights of every template (exm_distance_weights)
against the oracle and the closed form, and ln Lambda / ln B assembled from the read values
(exm_plan_gather_lse, exm_batch_collect_lse) against the oracle's assembly of the same read values.

Tolerances: the weights of a bank within 5e-11 of each template's total weight (the float64 closed form
cancels on short intervals; the device sums the trapezoid in order, numpy pairwise), single sets within
1e-12 relative;
the assembled ln Lambda and ln B within 1e-12 absolute of the host assembly of the same values (exp and
log of two libraries, sums in the same order)."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu

W_TOL = 1e-12
# the reference's closed form of the volume weights evaluated in float64 is itself off by up to 6.4e-12 of
# a template's total weight on this geometry (against an 80-bit evaluation: intervals down to h = 4.8e-5
# in ln y), so two float64 evaluations with different exp may differ by a few times that
W_BANK_TOL = 5e-11
A_TOL = 1e-12


def say(*a):
    print(*a, flush=True)


class _Bank:
    def __init__(self, u):
        self.u = np.asarray(u, float)
        self.M = np.geomspace(10.0, 200.0, self.u.size)
        self.q = np.ones(self.u.size)
        self.chi = np.zeros(self.u.size)


def _store(s):
    from exmarg_amd.search import store as stm
    return stm.RedshiftStore(s=np.asarray(s), alpha_lo=np.full(len(s), 0.16), alpha_hi=np.full(len(s), 4e4),
                             beta=[], delta=[], tau_idx=np.arange(3), tables=None, n_max=120)


@pytest.mark.parametrize("prior,ymin", [("volume", 0.1267), ("madau-dickinson", 0.0982)])
def test_bank_weights_match_the_oracle(hip, oracle, prior, ymin):
    from exmarg_amd.search import pe
    u = np.geomspace(0.3, 80.0, 2000)
    yN = pe._node_geometry(_Bank(u), _store(synth.S_NODES), ymin)[2]
    say("distance weights: %s prior, %d templates over %d rows" % (prior, u.size, yN.shape[0]))
    got = pe._bank_weights(yN, prior, ymin, _Bank(u))
    ref = oracle.bank_weights(yN, prior, ymin)
    np.testing.assert_array_equal(np.isfinite(got), np.isfinite(ref))
    # the closed form of an interval's weights cancels to O(h^2) terms (h its length in ln y), so on the
    # rare very short interval (a node just above the boundary point) the weight's own relative error is
    # eps / h in both evaluations; what ln Lambda feels is the error against the template's total weight
    wg, wr = np.exp(got), np.exp(ref)
    err = np.max(np.abs(wg - wr).max(axis=0) / wr.sum(axis=0))
    say("max |w - w_oracle| / sum w = %.2e" % err)
    assert err <= W_BANK_TOL


def test_volume_weights_closed_form(hip):
    """for Lambda = 1 the weights integrate the prior: sum w = ymin^3 (y_0^-3 - y_last^-3); for Lambda = ln y
    the exact integral of ln y against 3 ymin^3 y^-4"""
    from exmarg_amd.search import pe
    ymin = 0.1267
    y = np.geomspace(0.15, 7.0, 40)
    w = np.exp(pe.distance_weights(y, "volume", ymin))
    want = ymin ** 3 * (y[0] ** -3.0 - y[-1] ** -3.0)
    assert abs(w.sum() / want - 1.0) < 1e-13

    def F(t):   # antiderivative of 3 ymin^3 y^-4 ln y
        return -ymin ** 3 * t ** -3.0 * (np.log(t) + 1.0 / 3.0)
    assert abs(np.sum(w * np.log(y)) / (F(y[-1]) - F(y[0])) - 1.0) < 1e-12
    assert np.all(pe.distance_weights(y[:1], "volume", ymin) == -np.inf)
    with pytest.raises(ValueError, match="ascending"):
        pe.distance_weights(y[::-1], "volume", ymin)


def test_distance_weights_match_the_oracle(hip, oracle):
    from exmarg_amd.search import pe
    y = np.sort(np.random.default_rng(1).uniform(0.1, 4.0, 11))
    for prior in ("volume", "madau-dickinson"):
        np.testing.assert_allclose(pe.distance_weights(y, prior, 0.1), oracle.distance_weights(y, prior, 0.1),
                                   rtol=W_TOL)


def _host_assembly(oracle, an, vals, lnW, cell):
    """the oracle's node log-sum-exp and bank average of read values vals (plan order)"""
    nn = an.geometry[0].shape[0]
    lamN = np.full((nn + 1, an.bank.u.size), -np.inf)
    lamN[an.row, an.plan.job_m] = vals
    v = lnW + lamN
    mx = v.max(axis=0)
    L = mx + np.log(np.exp(v - mx[None, :]).sum(axis=0))
    return L, oracle.log_mean(L, cell)


@pytest.mark.parametrize("prior,ymin", [("volume", 0.1267), ("madau-dickinson", 0.0982)])
def test_plan_assembly_matches_host_assembly(hip, oracle, prior, ymin):
    """exm_plan_gather_lse against exm_plan_gather of the same segment assembled on the host by the
    oracle (weights, node LSE and bank average)"""
    from exmarg_amd.search import mass_bank, pe, read_segments, store as stm
    seg1, _, _ = synth.simulate_at_snr(1, 5, (30.0, 0.8, 0.1), 12.0, 4.0, 20.0, 256.0)
    rows = synth.grid_bank(M_range=(25.0, 40.0), shape=(6, 3, 2))
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[::3])
    bank = mass_bank(rows, seg1, "hip")
    seg = read_segments(seg1)
    say("plan assembly: %s prior, %d templates, %d nodes" % (prior, bank.u.size, st.s.size))
    an = pe._Analysis(bank, st, ymin, True)
    an.start().use_weights(prior)
    L, lnb = an.ln_lambda(seg, 0, 256, 256, 4)
    traj, times = an._filtered(seg, 0, 256, 256, 4)
    vals = an.plan.run(traj, an.dev, times)
    times.free()
    traj.free()
    lnW = oracle.bank_weights(an.geometry[2], prior, ymin)
    Lh, lnbh = _host_assembly(oracle, an, vals, lnW, oracle.cell_weights(bank.M, bank.q, bank.chi))
    dl, db = np.max(np.abs(L - Lh)), abs(lnb - lnbh)
    say("max |d ln Lambda| %.2e, |d ln B| %.2e" % (dl, db))
    assert dl <= A_TOL and db <= A_TOL
    stm.free(st)


def test_batch_assembly_matches_values(hip, oracle):
    """exm_batch_collect_lse against exm_batch_collect of the same batch (two segments, two banks)
    assembled on the host by the oracle"""
    from exmarg_amd.search import SegmentBatch, mass_bank, pe, read_segments, store as stm
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[::4])
    sb = SegmentBatch(st, 128, 128, max_templates=64)
    ans, segs = [], []
    for T, seed in ((2.0, 21), (4.0, 22)):
        seg1, _, _ = synth.simulate_at_snr(1, seed, (30.0, 0.8, 0.1), 14.0, T, 20.0, 256.0)
        bank = mass_bank(synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 9), seg1, "hip")
        an = pe._Analysis(bank, st, 0.1267, True)
        an.start().use_weights("volume")
        seg = read_segments(seg1)
        ans.append(an)
        segs.append(seg)
    say("batch assembly: 2 segments of 2 s and 4 s, 9 templates each")

    def fill():
        sb.reset()
        for an, seg in zip(ans, segs):
            sb.add(an, {I: seg["d_" + I][0] for I in "HL"}, {I: seg["sig2_" + I] for I in "HL"}, 7)
        sb.launch()

    fill()
    got = sb.collect_lse()
    fill()
    lamNs = sb.collect()
    for an, (L, lnb), lamN in zip(ans, got, lamNs):
        lnW = oracle.bank_weights(an.geometry[2], "volume", 0.1267)
        v = lnW + lamN
        mx = v.max(axis=0)
        Lh = mx + np.log(np.exp(v - mx[None, :]).sum(axis=0))
        lnbh = oracle.log_mean(Lh, oracle.cell_weights(an.bank.M, an.bank.q, an.bank.chi))
        assert np.max(np.abs(L - Lh)) <= A_TOL and abs(lnb - lnbh) <= A_TOL
    # a batch holding a segment whose plan has no weights refuses the device assembly
    sb.reset()
    an = ans[0]
    bank2 = mass_bank(synth.chirp_mass_grid_bank(31.0, 0.8, 0.1, 5), segs[0], "hip")
    an2 = pe._Analysis(bank2, st, 0.1267, True).start()
    sb.add(an2, {I: segs[0]["d_" + I][0] for I in "HL"}, {I: segs[0]["sig2_" + I] for I in "HL"}, 7)
    sb.launch()
    with pytest.raises(RuntimeError, match="node weights"):
        sb.collect_lse()
    sb.collect()
    sb.free()
    stm.free(st)
