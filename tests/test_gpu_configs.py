"""Every benchmark configuration at its own full-size workload on the device, against the FP64 oracle.

The oracle's results for these workloads take minutes of host time each (about 1.6e10 table lookups per
full bank), so they are computed once by tests/golden/make_config_fixtures.py from the same seeded
generators (bench.py) and committed as tests/golden/config*.npz.  The device banks here are built from
the same host TaylorF2 factors the oracle used (mass_bank(..., factors=(W, u))); the device TaylorF2 is
checked against the host model in test_gpu_bank.py.

  config 2  one 16 s segment, 2000 templates, 53 nodes: ln Lambda of every template and ln B (both forms)
  config 3  256 uneven segments in the batched float32 form: ln B of one segment per length against the
            oracle, every segment against the batched float64 form
  config 4  16 requests of the serving workload (500-7760 templates, 4-64 s) served together: two of them
            against the oracle, and alone against served
  config 5  8 s segments at 20-512 Hz, the 3517 masses of the reference's two-parameter bank, pe_redshift
            on the 1600 x 1600 grid: ln B and the grid
Beside them, the reference's own segment file (two rows of data/segments/two_parameters.npz) through
pe_redshift with the quickstart's bank, against the oracle computed here.
"""
import os

import numpy as np
import pytest

import bench
from exmarg_amd import synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LNB_TOL = 1e-5          # |ln B - ln B_FP64|: the bar of the analysis
LNL_TOL = 1e-5          # per-template ln Lambda of the float64 forms
# the float32 form rounds the filter output to float32 before the arrival-time ranking, so a time whose
# score lies within a float32 rounding of the n_loud-th largest may be exchanged for its neighbour in the
# ranking: a different (equally valid) walk of that template, which moves its ln Lambda by up to ~1e-3;
# ln B stays within LNB_TOL
LNL_TOL_F32 = 2e-3
MIX_TOL = 1e-7          # one request alone or served beside others


def say(*a):
    print(*a, flush=True)


def golden(name):
    path = os.path.join(GOLDEN, name + ".npz")
    if not os.path.exists(path):
        pytest.fail("fixture %s missing: python tests/golden/make_config_fixtures.py %s" % (path, name))
    return np.load(path)


def host_factors(rows, seg):
    from exmarg_amd.search import read_segments
    rows = np.asarray(rows, float)
    if rows.ndim == 1:
        rows = np.stack([rows, np.ones_like(rows), np.zeros_like(rows)], axis=1)
    return synth.template_factors(rows[:, 0], rows[:, 1], rows[:, 2], read_segments(seg, rows=[0]))


def host_bank(rows, seg):
    from exmarg_amd.search import mass_bank
    return mass_bank(rows, seg, "hip", factors=host_factors(rows, seg))


def set_precision(hip, f32):
    hip.call("exm_set_gather_precision", 1 if f32 else 0)
    hip.call("exm_set_filter_precision", 1 if f32 else 0)


@pytest.fixture(scope="module")
def store53(hip):
    from exmarg_amd.search import store as stm
    say("53 synthetic table nodes on the device (39.6 GB)")
    st = stm.synthetic_store(n_nodes=53, seed=0)
    yield st
    stm.free(st)


class _Grid:
    grid = "20,10,10"


@pytest.mark.parametrize("f32", [False, True])
def test_config2_full_bank(hip, store53, f32):
    from exmarg_amd.search import pe_intrinsic
    ref = golden("config2")
    seg, rows, _ = bench.workload(_Grid())
    bank = host_bank(rows, seg)
    say("config 2: 2000 templates, 16 s, 53 nodes, %s form" % ("float32" if f32 else "float64"))
    set_precision(hip, f32)
    try:
        res = pe_intrinsic(seg, bank, store53, resident=True)
    finally:
        set_precision(hip, False)
    dl = np.abs(res["ln_lambda"][0] - ref["ln_lambda"])
    db = abs(res["ln_b"][0] - float(ref["ln_b"]))
    say("ln B %.8f oracle %.8f |d ln B| %.2e; max |d ln Lambda| %.2e (template %d)"
        % (res["ln_b"][0], float(ref["ln_b"]), db, dl.max(), int(dl.argmax())))
    assert np.isfinite(res["ln_lambda"]).all()
    assert db <= LNB_TOL
    assert dl.max() <= (LNL_TOL_F32 if f32 else LNL_TOL)


def test_config3_batched_float32(hip, store53):
    from exmarg_amd.search import pe_intrinsic_batch
    ref = golden("config3")
    rows = synth.grid_bank(shape=(20, 10, 10))
    raw = bench.batch_items(256, 7)
    items = [(seg, host_bank(rows, seg)) for seg, _ in raw]
    say("config 3: %d segments of 4-64 s, 2000 templates per length, batched" % sum(s["d_H"].shape[0] for s, _ in raw))
    set_precision(hip, True)
    try:
        r32 = pe_intrinsic_batch(items, store53, seed=0)
    finally:
        set_precision(hip, False)
    r64 = pe_intrinsic_batch(items, store53, seed=0)
    worst = 0.0
    for (seg, T), a, b in zip(raw, r32, r64):
        t = int(T)
        db = abs(a["ln_b"][0] - float(ref["ln_b_%d" % t]))
        dl = np.abs(a["ln_lambda"][0] - ref["ln_lambda_%d" % t]).max()
        db64 = abs(b["ln_b"][0] - float(ref["ln_b_%d" % t]))
        dl64 = np.abs(b["ln_lambda"][0] - ref["ln_lambda_%d" % t]).max()
        say("T %2d s row 0: float32 |d ln B| %.2e max |d ln Lambda| %.2e; float64 %.2e / %.2e"
            % (t, db, dl, db64, dl64))
        assert db <= LNB_TOL and dl <= LNL_TOL_F32
        assert db64 <= LNB_TOL and dl64 <= LNL_TOL
        worst = max(worst, float(np.max(np.abs(a["ln_b"] - b["ln_b"]))))
    say("every segment: max |ln B(float32) - ln B(float64)| %.2e" % worst)
    assert worst <= LNB_TOL


def test_config4_serving(hip, store53):
    from exmarg_amd.search import Server, pe_intrinsic
    ref = golden("config4")
    reqs = bench.serve_requests(16, 7, 50.0)
    say("config 4: 16 requests (%d-%d templates, 4-64 s) served together"
        % (min(r[2].shape[0] for r in reqs), max(r[2].shape[0] for r in reqs)))
    srv = Server(store53, max_request=(8000, 8 * 65536, 64257))
    try:
        handles = [srv.submit(seg, rows, seed=i, factors=host_factors(rows, seg)) for i, (_, seg, rows) in
                   enumerate(reqs)]
        srv.run()
        assert all(h.error is None for h in handles), [str(h.error) for h in handles if h.error]
        assert srv.stats["batches"] >= 2
    finally:
        srv.close()
    for i in (4, 5):
        h = handles[i]
        db = abs(h.ln_b - float(ref["ln_b_%d" % i]))
        dl = np.abs(h.ln_lambda - ref["ln_lambda_%d" % i]).max()
        _, seg, rows = reqs[i]
        alone = pe_intrinsic(seg, host_bank(rows, seg), store53, seed=i, resident=True)
        dm = np.abs(alone["ln_lambda"][0] - h.ln_lambda).max()
        say("request %d (%d templates, %g s): |d ln B| %.2e, max |d ln Lambda| %.2e, alone vs served %.2e"
            % (i, rows.shape[0], seg["T"], db, dl, dm))
        assert db <= LNB_TOL and dl <= LNL_TOL and dm <= MIX_TOL


def test_config5_redshift_stream_segments(hip, store53):
    from exmarg_amd.priors import madau_dickinson_ymin
    from exmarg_amd.search import pe_redshift, read_segments
    ref = golden("config5")
    groups, _ = bench.stream_groups(64, 32, 7)
    seg = read_segments(groups[0])
    masses = np.load(os.path.join(os.path.dirname(GOLDEN), "..", "exmarg_amd", "data", "bank_2p.npy"))
    bank = host_bank(masses, seg)
    say("config 5: 8 s, 20-512 Hz, %d masses, 1600 x 1600 grid, rows 0 and 6" % masses.size)
    res = pe_redshift(seg, bank, store53, rows=[0, 6], zmax=10.0, nz=1600, m1_range=(2.0, 200.0), nm1=1600,
                      mdet_range=(10.0, 500.0), ymin=madau_dickinson_ymin(), seed=0, resident=True)
    step = 40
    for k, j in enumerate((0, 6)):
        db = abs(res["ln_b"][k] - float(ref["ln_b_%d" % j]))
        g, gr = res["ln_lambda"][k][::step, ::step], ref["grid_%d" % j]
        fin = np.isfinite(gr)
        np.testing.assert_array_equal(np.isfinite(g), fin)
        dg = np.abs(g[fin] - gr[fin]).max()
        say("row %d: ln B %.6f |d ln B| %.2e; grid sample max |d ln Lambda| %.2e" % (j, res["ln_b"][k], db, dg))
        assert db <= LNB_TOL and dg <= LNL_TOL


@pytest.mark.parametrize("bank_name", ["quickstart", "two_parameters"])
def test_reference_segments_through_pe_redshift(hip, store53, oracle, bank_name):
    """the two rows of the reference's data/segments/two_parameters.npz (8 s, 20-512 Hz) with the bank of
    its quickstart (120 masses geomspace(20, 200), M_det prior on (20, 200)) or every 8th mass of its
    two-parameter bank (M_det prior on (10, 500)), against the oracle run here"""
    from exmarg_amd.search import pe_redshift, read_segments
    path = os.path.join(GOLDEN, "segments_2p.npz")
    if bank_name == "quickstart":
        masses, mdet = np.geomspace(20.0, 200.0, 120), (20.0, 200.0)
    else:
        masses = np.load(os.path.join(os.path.dirname(GOLDEN), "..", "exmarg_amd", "data", "bank_2p.npy"))[::8]
        mdet = (10.0, 500.0)
    bank = host_bank(masses, path)
    say("reference segments: %d masses, prior on M_det %s" % (masses.size, mdet))
    kw = dict(zmax=8.0, nz=400, m1_range=(3.0, 150.0), nm1=400, mdet_range=mdet)
    res = pe_redshift(path, bank, store53, seed=0, resident=True, **kw)
    W, u = host_factors(masses, path)
    seg = read_segments(path)
    b = dict(W=W, u=u, M=bank.M, q=bank.q, chi=bank.chi, lo=bank.lo, n_time=bank.n_time)
    st = dict(s=store53.s, alpha_lo=store53.alpha_lo, alpha_hi=store53.alpha_hi, beta=store53.beta[0],
              delta=store53.delta[0], tau_idx=store53.tau_idx)
    tabs = oracle.Tables(synth=dict(seed=0, kind="smooth", nodes=np.arange(53)), shape=store53.table_shape(),
                         tau_idx=store53.tau_idx, materialize=True)
    ref = oracle.pe_redshift(seg, b, st, tabs, seed=0, **kw)
    for e in range(2):
        fin = np.isfinite(ref["ln_lambda"][e])
        np.testing.assert_array_equal(np.isfinite(res["ln_lambda"][e]), fin)
        dl = np.abs(res["ln_lambda"][e][fin] - ref["ln_lambda"][e][fin]).max()
        say("row %d: ln B %.6f oracle %.6f; max |d ln Lambda| %.2e" % (e, res["ln_b"][e], ref["ln_b"][e], dl))
        assert abs(res["ln_b"][e] - ref["ln_b"][e]) <= LNB_TOL and dl <= LNL_TOL
