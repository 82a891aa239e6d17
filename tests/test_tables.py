"""Stage 3 of the tables (the sum over the orders, scripts/build_tables/build_bd.py): the oracle pinned
by closed forms and brute force on the CPU, the device build against the oracle on the GPU."""
import math
import os

import numpy as np
import pytest

say = print
QMAX, QSPAN, NQ = 100.0, 2000.0, 110            # the stage-2 q grid (scripts/build_tables/build_k.py:30-31,87)


def _dyadic_lnq(nq, step=0.25, start=-1.0):
    """a ln q grid of exactly representable values (so that 2 n ln q is exact in float32)"""
    return start + step * np.arange(nq)


def _constant_g(nord, ntau, nq, nphif, lnq, c):
    """col and row whose g_n is c[ti, pi] for every order and Delta (exact in float32)"""
    n = np.arange(nord, dtype=np.float64)[:, None, None, None]
    col = (2.0 * n * lnq[-1] + c[None, :, None, :]).astype(np.float32) + np.zeros((1, 1, nq, 1), np.float32)
    row = (2.0 * n * lnq[None, None, :, None] + c[None, :, None, :]).astype(np.float32)
    return np.ascontiguousarray(col), np.ascontiguousarray(row)


def _random_stage2(nord, ntau, nq, nphif, seed):
    """col and row with g_n decaying like -ln n! (a converging sum) plus random detail"""
    rng = np.random.default_rng(seed)
    lnq = np.log(np.geomspace(0.5, 8.0, nq))
    n = np.arange(nord, dtype=np.float64)[:, None, None, None]
    lgam = np.array([math.lgamma(k + 1.0) for k in range(nord)])[:, None, None, None]
    col = 2.0 * n * lnq[-1] - lgam + rng.normal(0.0, 1.0, (nord, ntau, nq, nphif))
    row = 2.0 * n * lnq[None, None, :, None] - lgam + rng.normal(0.0, 1.0, (nord, ntau, nq, nphif))
    return col.astype(np.float32), row.astype(np.float32), lnq


def _ulps(a, b):
    """distance in float32 units in the last place"""
    a = a.astype(np.float32).view(np.int32).astype(np.int64)
    b = b.astype(np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


# ---------------------------------------------------------------- the oracle on the CPU
def test_grids_are_the_store_geometry(oracle):
    from exmarg_amd import synth
    lnq = np.log(np.geomspace(QMAX / QSPAN, QMAX, NQ))
    beta, delta = oracle.bd_grids(lnq)
    assert beta.size == 200 and delta.size == 219
    b, d, _ = synth.store_grids()
    np.testing.assert_allclose(beta, b, rtol=0, atol=1e-12)
    np.testing.assert_allclose(delta, d, rtol=0, atol=1e-12)
    assert np.all(np.diff(delta) > 0) and delta[NQ - 1] == 0.0


def test_constant_orders_give_the_geometric_series(oracle):
    """g_n = c for every n: T = c + ln sum_{n <= n_max} e^{n beta} = c + ln((e^{(n_max+1) beta} - 1)/(e^beta - 1))"""
    nord, ntau, nq, nphif = 9, 2, 5, 3
    lnq = _dyadic_lnq(nq)
    c = np.array([[0.5, -1.25, 2.0], [0.0, 3.5, -0.75]])
    col, row = _constant_g(nord, ntau, nq, nphif, lnq, c)
    out = oracle.bd_sum_orders(col, row, lnq, [0, 4, 8])
    beta, delta = oracle.bd_grids(lnq)
    for nm, T in out.items():
        series = np.array([math.log(sum(math.exp(k * b) for k in range(nm + 1))) for b in beta])
        ref = c[:, :, None, None] + series[None, None, :, None] + np.zeros(delta.size)
        assert _ulps(T, ref).max() <= 1, nm


def test_order_zero_is_g0(oracle):
    """a single order: T = g_0(Delta), the q_max column for Delta <= 0 and the reversed row after"""
    col, row, lnq = _random_stage2(4, 2, 6, 3, seed=1)
    T = oracle.bd_sum_orders(col, row, lnq, [0])[0]
    g0 = np.concatenate([col[0], row[0][:, ::-1, :][:, 1:, :]], axis=1)          # (ntau, nd, nphif)
    np.testing.assert_array_equal(T, np.broadcast_to(g0.transpose(0, 2, 1)[:, :, None, :], T.shape))


def test_matches_brute_force(oracle):
    nord, ntau, nq, nphif = 6, 2, 4, 3
    col, row, lnq = _random_stage2(nord, ntau, nq, nphif, seed=2)
    out = oracle.bd_sum_orders(col, row, lnq, [2, 5])
    beta, _ = oracle.bd_grids(lnq)
    nd = 2 * nq - 1
    for nm, T in out.items():
        for ti in range(ntau):
            for pi in range(nphif):
                for b in range(0, beta.size, 7):
                    for d in range(nd):
                        E = []
                        for k in range(nm + 1):
                            if d < nq:
                                g = float(col[k, ti, d, pi]) - 2.0 * k * lnq[-1]
                            else:
                                q2 = 2 * nq - 2 - d
                                g = float(row[k, ti, q2, pi]) - 2.0 * k * lnq[q2]
                            E.append(k * beta[b] + g)
                        m = max(E)
                        ref = m + math.log(sum(math.exp(e - m) for e in E))
                        assert _ulps(np.float32(T[ti, pi, b, d]), np.float32(ref)) <= 1


# ---------------------------------------------------------------- the device build
@pytest.mark.gpu
@pytest.mark.parametrize("shape,ladder", [((6, 2, 4, 3), [0, 2, 5]), ((121, 3, 9, 5), [1, 20, 120])])
def test_device_sum_matches_oracle(hip, oracle, shape, ladder):
    from exmarg_amd import build_bd
    say("tables: stage 3, shape %s, ladder %s" % (shape, ladder), flush=True)
    col, row, lnq = _random_stage2(*shape, seed=3)
    got = build_bd.sum_over_orders(col, row, lnq, ladder)
    ref = oracle.bd_sum_orders(col, row, lnq, ladder)
    for nm in ladder:
        assert got[nm].shape == ref[nm].shape
        assert _ulps(got[nm], ref[nm]).max() <= 1, nm
        assert np.mean(got[nm] == ref[nm]) > 0.99, nm


@pytest.mark.gpu
def test_device_sum_closed_form(hip, oracle):
    from exmarg_amd import build_bd
    say("tables: stage 3, constant orders (geometric series)", flush=True)
    nord, ntau, nq, nphif = 17, 2, 5, 2
    lnq = _dyadic_lnq(nq)
    c = np.array([[1.5, -2.0], [0.25, 0.0]])
    col, row = _constant_g(nord, ntau, nq, nphif, lnq, c)
    got = build_bd.sum_over_orders(col, row, lnq, [16])[16]
    beta, delta = build_bd.grids(lnq)
    series = np.array([math.log(sum(math.exp(k * b) for k in range(17))) for b in beta])
    ref = c[:, :, None, None] + series[None, None, :, None] + np.zeros(delta.size)
    assert _ulps(got, ref).max() <= 1


@pytest.mark.gpu
def test_build_node_writes_store_files(hip, oracle, tmp_path):
    """stage-2 files in, bd_*.npz files out (with the truncated orders of a ladder beside them), readable
    by load_redshift_store and equal to the oracle"""
    from exmarg_amd import build_bd
    from exmarg_amd.search import store as stm
    say("tables: build_node from stage-2 files", flush=True)
    nord, ntau, nq, nphi = 9, 3, 6, 96
    rng = np.random.default_rng(4)
    lnq = np.log(np.geomspace(0.5, 8.0, nq))
    n = np.arange(nord, dtype=np.float64)
    lgam = np.array([math.lgamma(k + 1.0) for k in range(nord)])
    lnK = (2.0 * n[:, None, None, None, None] * lnq[None, None, None, :, None] - lgam[:, None, None, None, None]
           + rng.normal(0.0, 0.5, (nord, ntau, nq, nq, nphi))).astype(np.float32)
    kdir = tmp_path / "k"
    kdir.mkdir()
    tau_idx = np.array([0, 1, -1])
    for s in (3.5, 12.25):
        stem = str(kdir / ("kstore_n%03d_s%09.3f" % (nord - 1, s)))
        np.save(stem + ".lnK.npy", lnK)
        np.savez(stem + ".meta.npz", s_node=s, u_node=s, lnq=lnq, tau_idx=tau_idx)
    out = str(tmp_path / "bd")
    stems = sorted(str(p)[:-len(".meta.npz")] for p in kdir.glob("*.meta.npz"))
    written = [build_bd.build_node(st, out, ladder=[3, nord - 1]) for st in stems]
    assert all(len(w) == 2 for w in written)
    assert build_bd.build_node(stems[0], out, ladder=[3, nord - 1]) == []          # already there: skipped
    st = stm.load_redshift_store(out)
    assert list(st.s) == [3.5, 12.25] and st.n_max == nord - 1
    col = np.ascontiguousarray(lnK[:, :, :, nq - 1, :49])
    row = np.ascontiguousarray(lnK[:, :, nq - 1, :, :49])
    ref = oracle.bd_sum_orders(col, row, lnq, [3, nord - 1])
    for k in range(2):
        assert st.tables[k].shape == ref[nord - 1].shape
        assert _ulps(st.tables[k], ref[nord - 1]).max() <= 1
    trunc = stm.load_redshift_store(out + "_n003")
    assert trunc.n_max == 3 and _ulps(trunc.tables[0], ref[3]).max() <= 1
    with np.load(os.path.join(out, "bd_n%03d_s%09.3f.npz" % (nord - 1, 3.5))) as h:
        assert set(h.files) >= {"T", "beta", "delta", "tau_idx", "lnq", "alpha", "s_node", "u_node", "n_max",
                                "nphif", "norm"}
        assert int(h["nphif"]) == 49 and str(h["norm"]) == "taylor_off1"


@pytest.mark.gpu
def test_device_sum_rejects_bad_ladders(hip):
    from exmarg_amd import build_bd
    col, row, lnq = _random_stage2(5, 1, 3, 2, seed=5)
    with pytest.raises(ValueError, match="outside"):
        build_bd.sum_over_orders(col, row, lnq, [5])
    with pytest.raises(ValueError, match="ascend"):
        build_bd.sum_over_orders(col, row, lnq, [3, 1])
    with pytest.raises(ValueError, match="lnq"):
        build_bd.sum_over_orders(col, row, lnq[:2], [1])


def test_cli_indices_name_grid_nodes(tmp_path):
    """a node index means the node at that position of the 53-node s grid (the reference's
    build_bd.py:88-92), whichever nodes the stage-2 directory holds"""
    from exmarg_amd import build_bd
    s_grid = np.load(os.path.join(os.path.dirname(__file__), "golden", "s_grid.npz"))["s_nodes"]
    kdir = tmp_path / "k"
    kdir.mkdir()
    for i in (7, 3, 40):                        # a partial directory
        stem = kdir / ("kstore_n120_s%09.3f" % s_grid[i])
        np.savez(str(stem) + ".meta.npz", s_node=s_grid[i], lnq=np.zeros(3), tau_idx=np.zeros(1))
    stems = build_bd.node_stems(str(kdir), [40, 3])
    assert [os.path.basename(p) for p in stems] == ["kstore_n120_s%09.3f" % s_grid[i] for i in (40, 3)]
    assert [os.path.basename(p) for p in build_bd.node_stems(str(kdir))] == \
        ["kstore_n120_s%09.3f" % s_grid[i] for i in (3, 7, 40)]
    with pytest.raises(FileNotFoundError, match="node 5"):
        build_bd.node_stems(str(kdir), [5])
    with pytest.raises(IndexError):
        build_bd.node_stems(str(kdir), [53])
    node = build_bd.Stage2Node.open(stems[1])
    assert node.n_full == 120 and node.s == pytest.approx(s_grid[3])
    assert build_bd.store_path("out", 7, 120, node.s).startswith(os.path.join("out_n007", "bd_n007_s"))
    assert build_bd.truncation_orders(120, [5, 1, 5]) == [1, 5]
    with pytest.raises(ValueError, match="above"):
        build_bd.truncation_orders(10, [11])
