"""Pins of the CPU oracle (oracle/) against what the mathematics fixes: closed forms, library routines,
brute force on tiny inputs and the golden vectors held in the reference's data files.  Each test is built
so that a plausible mistake in the oracle (a dropped term, a wrong sign or index, a transposed operand)
fails it."""
import math
import os

import numpy as np
import pytest
from scipy import integrate

from exmarg_amd import synth

NT, NM = 64, 3
SHAPE = (5, 49, 8, 9)          # lags -2..2, 49 phases, 8 beta rows, 9 Delta columns


def tiny_grids(nb=8, nd=9):
    beta, delta, _ = synth.store_grids()
    db, dd = beta[1] - beta[0], delta[1] - delta[0]
    return beta[0] + db * np.arange(nb), dd * (np.arange(nd) - nd // 2)


def tiny_traj(seed=1, nm=NM, nt=NT, spread=1.0):
    rng = np.random.default_rng(seed)
    return {"lqH": (spread * rng.standard_normal((nm, nt))).astype(np.float32),
            "phH": rng.uniform(-np.pi, np.pi, (nm, nt)).astype(np.float32),
            "lqL": (spread * rng.standard_normal((nm, nt))).astype(np.float32),
            "phL": rng.uniform(-np.pi, np.pi, (nm, nt)).astype(np.float32)}


def table_from(fn, shape=SHAPE):
    r, c, b, d = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return fn(r, c, b, d).astype(np.float32)


def run(oracle, traj, T, lna, times=None, ntab_ids=None):
    beta, delta = tiny_grids(T.shape[2], T.shape[3])
    tau = synth.lag_indices((T.shape[0] - 1) // 2)
    tabs = oracle.Tables(tables=[T], tau_idx=tau)
    ids = np.arange(NM)
    val = oracle.gather(traj, tabs, beta, delta, ids, np.asarray(lna, float).reshape(NM, 1),
                        np.zeros((NM, 1), np.int64), times)
    return val[0], beta, delta, tau


def index_coords(traj, lna, beta, delta, tau):
    """xb (nm, nt), xd (nm, ntau, nt), x3 (nm, ntau, nt) of every lookup, from their definitions"""
    nb, nd = beta.size, delta.size
    db, dd = beta[1] - beta[0], delta[1] - delta[0]
    aL = traj["lqL"].astype(np.float64)
    xbr = aL * (2.0 / db) + ((np.asarray(lna)[:, None] - beta[0]) / db)
    xb = np.clip(xbr, 0.0, nb - 1 - 1e-6)
    aLc = aL + (xb - xbr) / (2.0 / db)
    t = np.arange(NT)
    xd = np.empty((NM, tau.size, NT))
    x3 = np.empty((NM, tau.size, NT))
    for r, tv in enumerate(tau):
        i1 = (t + tv) % NT
        xd[:, r] = np.clip((traj["lqH"][:, i1] - aLc) * (1.0 / dd) + (-delta[0] / dd), 0.0, nd - 1 - 1e-6)
        dph = (traj["phH"][:, i1] - traj["phL"]).astype(np.float32).astype(np.float64)
        x3[:, r] = np.mod(dph * (96 / (2 * np.pi)), 96.0)
    return xb, xd, x3


def lse(v, axis=None):
    mx = np.max(v, axis=axis, keepdims=True)
    return np.squeeze(mx, axis=axis) + np.log(np.sum(np.exp(v - mx), axis=axis))


LNA = [0.3, -1.2, 2.0]


def test_constant_table(oracle):
    traj = tiny_traj()
    val, *_ = run(oracle, traj, np.full(SHAPE, 1.75, np.float32), LNA)
    np.testing.assert_allclose(val, 1.75 + np.log(SHAPE[0]), rtol=0, atol=1e-12)


def test_constant_table_subsampled_weights_restore_the_sum(oracle):
    traj = tiny_traj()
    times = oracle.select_times(traj, synth.lag_indices(2), 10, 12, seed=3)
    val, *_ = run(oracle, traj, np.full(SHAPE, -0.5, np.float32), LNA, times)
    np.testing.assert_allclose(val, -0.5 + np.log(SHAPE[0]), atol=2e-7)


def test_trilinear_table_is_reproduced_exactly(oracle):
    # T = alpha b d c' + d/128 - b/32 (dyadic, so exact in float32): bilinear interpolation in
    # (beta, Delta) and linear interpolation between the two (mirrored) phase bins reproduce it exactly
    traj = tiny_traj(spread=1.4)
    alpha = 0.0625
    T = table_from(lambda r, c, b, d: alpha * b * d * c + d / 128.0 - b / 32.0)
    val, beta, delta, tau = run(oracle, traj, T, LNA)
    xb, xd, x3 = index_coords(traj, LNA, beta, delta, tau)
    tri = np.minimum(x3, 96.0 - x3)
    acc = alpha * xb[:, None, :] * xd * tri + xd / 128.0 - xb[:, None, :] / 32.0
    expect = lse(acc.reshape(NM, -1), axis=1) - np.log(NT)
    np.testing.assert_allclose(val, expect, rtol=0, atol=1e-9)
    # clamps were exercised on both axes
    assert (xb == 0).any() or (xb >= 6.999).any()


def test_lag_slab_goes_with_its_lag(oracle):
    # slab r carries an offset delta_r and a Delta slope: a slab read at the wrong lag changes the sum
    traj = tiny_traj(seed=5)
    offs = np.array([0.0, 0.875, -0.375, 1.25, 0.25])
    T = table_from(lambda r, c, b, d: offs[r] + 0.375 * d)
    val, beta, delta, tau = run(oracle, traj, T, LNA)
    _, xd, _ = index_coords(traj, LNA, beta, delta, tau)
    acc = offs[None, :, None] + 0.375 * xd
    expect = lse(acc.reshape(NM, -1), axis=1) - np.log(NT)
    np.testing.assert_allclose(val, expect, rtol=0, atol=1e-9)


def test_lag_wraps_around_the_segment(oracle):
    # shifting every trajectory in time by a full period of lags leaves ln Lambda unchanged
    traj = tiny_traj(seed=7)
    T = table_from(lambda r, c, b, d: np.sin(r + 0.3 * c) + 0.2 * b - 0.1 * d)
    v0, *_ = run(oracle, traj, T, LNA)
    rolled = {k: np.roll(v, 17, axis=1) for k, v in traj.items()}
    v1, *_ = run(oracle, rolled, T, LNA)
    np.testing.assert_allclose(v0, v1, rtol=0, atol=1e-12)


def test_phase_is_read_evenly(oracle):
    # swapping the phases of the two detectors flips dphi; the tables are even in dphi
    traj = tiny_traj(seed=9)
    T = table_from(lambda r, c, b, d: np.cos(0.2 * c) * (1 + 0.1 * b) + 0.05 * d)
    swapped = dict(traj, phH=(-traj["phH"]).astype(np.float32), phL=(-traj["phL"]).astype(np.float32))
    v0, *_ = run(oracle, traj, T, LNA)
    v1, *_ = run(oracle, swapped, T, LNA)
    np.testing.assert_allclose(v0, v1, rtol=0, atol=1e-5)


def test_pexp_accuracy(oracle):
    xs = np.linspace(-100.0, 60.0, 20001)
    got = np.array([oracle.lib().oracle_pexp(float(x)) for x in xs[::50]])
    ref = np.array([math.exp(x) for x in xs[::50]])
    assert np.max(np.abs(got / ref - 1)) < 5e-15
    # the numpy evaluation of the same operation sequence is bit-identical
    np.testing.assert_array_equal(oracle.pexp_np(xs[::50]), got)


def test_select_c_matches_numpy_lexsort(oracle):
    traj = tiny_traj(seed=11, nt=256)
    traj = {k: v[:, :256] for k, v in traj.items()}
    tau = synth.lag_indices(4)
    a = oracle.select_times(traj, tau, 20, 30, seed=4)
    b = oracle.select_times_np(traj, tau, 20, 30, seed=4)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_select_top_times_outrank_the_rest(oracle):
    nt, K, S = 300, 25, 40
    rng = np.random.default_rng(2)
    traj = {"lqH": rng.standard_normal((2, nt)).astype(np.float32),
            "lqL": rng.standard_normal((2, nt)).astype(np.float32)}
    sel, lw = oracle.select_times(traj, synth.lag_indices(3), K, S, seed=8)
    assert (np.diff(sel, axis=1) >= 0).all()
    for m in range(2):
        mx = np.array([max(traj["lqH"][m, (t + k) % nt] for k in range(-3, 4)) for t in range(nt)])
        score = np.exp(2.0 * traj["lqL"][m].astype(float)) + np.exp(2.0 * mx.astype(float))
        top = sel[m][lw[m] == 0]
        others = np.setdiff1d(np.arange(nt), top)
        assert top.size == K and score[top].min() > score[others].max()
        drawn = sel[m][lw[m] != 0]
        pick = np.random.default_rng(8).integers(0, nt - K, size=(2, S))[m]
        np.testing.assert_array_equal(np.sort(others[pick]), drawn)
    np.testing.assert_allclose(lw[lw != 0], np.float32(np.log((nt - K) / S)))


def test_select_ties_go_to_earlier_times(oracle):
    traj = {"lqH": np.zeros((1, 128), np.float32), "lqL": np.zeros((1, 128), np.float32)}
    sel, lw = oracle.select_times(traj, synth.lag_indices(2), 16, 8, seed=0)
    np.testing.assert_array_equal(sel[0][lw[0] == 0], np.arange(16))


def test_filter_single_bin_closed_form(oracle):
    # one nonzero bin nu0: z(t) = W w e^{-2 pi i nu0 t / N}, constant modulus, linear phase
    n_time, lo, nf, nu0 = 256, 10, 20, 13
    W = np.zeros((1, nf), np.complex64)
    W[0, nu0 - lo] = 0.5 - 0.25j
    d = np.full(nf, 2.0 - 1.0j)
    s2 = np.full(nf, 4.0)
    traj = oracle.filter_fp64(W, lo, n_time, {"H": d, "L": d}, {"H": s2, "L": s2})
    w = complex(W[0, nu0 - lo]) * np.conj(2.0 - 1.0j) / 4.0
    t = np.arange(n_time)
    np.testing.assert_allclose(traj["lqH"][0], np.float32(np.log(abs(w))), rtol=0, atol=1e-6)
    ph = np.angle(w) - 2 * np.pi * nu0 * t / n_time
    dphi = np.angle(np.exp(1j * (traj["phL"][0].astype(float) - ph)))
    assert np.abs(dphi).max() < 1e-6


def test_filter_matches_scipy_fft(oracle):
    from scipy import fft
    rng = np.random.default_rng(0)
    n_time, lo, nf = 512, 30, 33
    W = (rng.standard_normal((4, nf)) + 1j * rng.standard_normal((4, nf))).astype(np.complex64)
    d = {I: rng.standard_normal(nf) + 1j * rng.standard_normal(nf) for I in "HL"}
    s2 = {I: rng.uniform(0.5, 2, nf) for I in "HL"}
    traj = oracle.filter_fp64(W, lo, n_time, d, s2)
    pad = np.zeros((4, n_time), complex)
    pad[:, lo:lo + nf] = W.astype(complex) * (np.conj(d["L"]) / s2["L"])
    z = fft.fft(pad, axis=1)
    np.testing.assert_allclose(traj["lqL"], np.log(np.abs(z)), rtol=0, atol=2e-6)


def test_volume_weights_closed_form(oracle):
    ymin = 0.1267
    y = np.sort(np.random.default_rng(3).uniform(0.13, 3.0, 9))
    w = np.exp(oracle.distance_weights(y, "volume", ymin))
    np.testing.assert_allclose(w.sum(), ymin ** 3 * (y[0] ** -3 - y[-1] ** -3), rtol=1e-13)
    lam = 0.4 + 1.7 * np.log(y)
    ref = integrate.quad(lambda v: 3 * ymin ** 3 * v ** -4 * (0.4 + 1.7 * np.log(v)), y[0], y[-1],
                         epsabs=0, epsrel=1e-13, limit=200)[0]
    np.testing.assert_allclose(np.sum(w * lam), ref, rtol=1e-11)


def test_madau_dickinson_weights_against_quad(oracle):
    lny, p = oracle.distance_density("madau-dickinson")
    y = np.exp(np.linspace(-2.0, 1.5, 7))
    w = np.exp(oracle.distance_weights(y, "madau-dickinson", 0.0982))
    lam = 2.0 - 0.8 * np.log(y)
    f = lambda x: np.interp(x, lny, p, left=0.0, right=0.0) * (2.0 - 0.8 * x)
    ref = integrate.quad(f, np.log(y[0]), np.log(y[-1]), points=list(np.log(y[1:-1])), limit=500)[0]
    np.testing.assert_allclose(np.sum(w * lam), ref, rtol=2e-5)


def test_log_mean_closed_form(oracle):
    vals = np.array([0.5, 2.0, 3.5, 1.0])
    assert abs(oracle.log_mean(np.log(vals), np.full(4, 0.25)) - np.log(vals.mean())) < 1e-14
    assert oracle.log_mean(np.array([-np.inf, np.log(2.0)]), np.array([0.5, 0.5])) == pytest.approx(np.log(1.0))
    assert oracle.log_mean(np.array([-np.inf]), np.array([1.0])) == -np.inf


def test_node_geometry_inequalities(oracle):
    s = synth.S_NODES
    u = np.array([0.9, 5.0, 60.0])
    ymin = 0.1267
    valid, lna, y, bnode, blna = oracle.node_geometry(s, np.full(53, 0.16), np.full(53, 4e4), u, ymin)
    a = s[:, None] / u[None, :]
    np.testing.assert_array_equal(valid, (a >= ymin ** 2) & (a > ymin ** 2) & (a < 4e4))
    for m in range(3):
        k = int(np.nonzero(valid[:, m])[0][0])
        assert bnode[m] == k and s[k - 1] / u[m] <= ymin ** 2
        assert blna[m] == pytest.approx(np.log(ymin ** 2 * (1 + 1e-9)), abs=1e-15)


def test_dl_matches_reference_planck18_pairs(oracle):
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "planck18_pairs.npz"))
    np.testing.assert_allclose(oracle.dl_of_z(z["z"]), z["dl"], rtol=1e-7)


def test_to_source_frame_brute_force(oracle):
    rng = np.random.default_rng(4)
    nn, nm = 6, 5
    yN = np.sort(rng.uniform(0.2, 2.0, (nn, nm)), axis=0)
    lamN = rng.standard_normal((nn, nm))
    lamN[2, 1] = -np.inf
    Mdet = np.linspace(20, 60, nm)
    yq, m1, z = np.array([0.3, 0.9, 1.5]), np.array([9.0, 12.0, 20.0]), np.array([0.1, 0.5, 0.8])
    out = oracle.to_source_frame(lamN, yN, Mdet, yq, m1, z)
    for iz in range(3):
        col = []
        for m in range(nm):
            ok = np.isfinite(lamN[:, m])
            ys, ls = yN[ok, m], lamN[ok, m]
            col.append(np.interp(yq[iz], ys, ls) if ys[0] <= yq[iz] <= ys[-1] else -np.inf)
        col = np.array(col)
        g = np.isfinite(col)
        for i in range(3):
            Md = 2 * m1[i] * (1 + z[iz])
            exp_ = np.interp(Md, Mdet[g], col[g]) if g.sum() >= 2 and Mdet[g][0] <= Md <= Mdet[g][-1] else -np.inf
            assert out[i, iz] == pytest.approx(exp_) if np.isfinite(exp_) else out[i, iz] == -np.inf


# ------------------------------------------------------------------ bank assembly pins
def test_cell_weights_by_hand(oracle):
    # np.gradient of the distinct values of an axis: end cells one-sided, inner cells centred
    M = np.array([10.0, 20.0, 40.0])            # spacings 10, 15, 20
    q = np.array([0.5, 1.0])                    # 0.5, 0.5
    chi = np.array([-0.5, 0.0, 0.25])           # 0.5, 0.375, 0.25
    grid = np.array(np.meshgrid(M, q, chi, indexing="ij")).reshape(3, -1).T
    grid = grid[np.random.default_rng(6).permutation(grid.shape[0])]       # any row order
    got = oracle.cell_weights(grid[:, 0], grid[:, 1], grid[:, 2])
    gm = {10.0: 10.0, 20.0: 15.0, 40.0: 20.0}
    gc = {-0.5: 0.5, 0.0: 0.375, 0.25: 0.25}
    raw = np.array([gm[r[0]] * 0.5 * gc[r[2]] for r in grid])
    np.testing.assert_allclose(got, raw / raw.sum(), rtol=1e-14)
    # an axis that does not vary contributes nothing; a uniform grid weighs every cell alike
    flat = oracle.cell_weights(grid[:, 0], grid[:, 1], np.zeros(grid.shape[0]))
    np.testing.assert_allclose(flat, np.array([gm[r[0]] for r in grid]) / (2 * 3 * 45.0), rtol=1e-14)
    uni = np.array(np.meshgrid([1.0, 2.0, 3.0], [0.2, 0.4], indexing="ij")).reshape(2, -1).T
    np.testing.assert_allclose(oracle.cell_weights(uni[:, 0], uni[:, 1], np.zeros(6)), np.full(6, 1 / 6), rtol=1e-14)


def _dl_derivative(oracle, z, h=1e-3):
    """dD_L/dz by the five-point central stencil on the (pinned) dl_of_z"""
    f = lambda zz: oracle.dl_of_z(zz)
    return (f(z - 2 * h) - 8 * f(z - h) + 8 * f(z + h) - f(z + 2 * h)) / (12 * h)


def test_redshift_prior_weights_against_quad(oracle):
    # p(m1, z) ~ (1 + z) D_L^2 dD_L/dz, uniform in M_det = 2 m1 (1 + z) on mdet_range: the weight of a
    # z column is the integral of that density over the column's cell, the weight of a m1 row its
    # cell width; cells outside the M_det range weigh nothing
    z = np.linspace(0.02, 8.0, 1600)
    m1 = np.geomspace(3.0, 150.0, 60)
    mdet = (10.0, 500.0)
    w = oracle.pe_redshift_weights(m1, z, mdet)
    p = lambda zz: (1.0 + zz) * oracle.dl_of_z(zz) ** 2 * _dl_derivative(oracle, np.atleast_1d(zz))
    dz = np.gradient(z)
    cells = np.array([integrate.quad(lambda x: float(p(x)[0]), zj - hj / 2, zj + hj / 2, epsrel=1e-10)[0]
                      for zj, hj in zip(z[::97], dz[::97])])
    Md = 2.0 * m1[:, None] * (1.0 + z[None, :])
    inside = (Md >= mdet[0]) & (Md <= mdet[1])
    assert (w[~inside] == 0).all() and (w[inside] > 0).all()
    # the column profile: w[i, j] / (dm1_i) is the same for every row inside the range
    prof = np.where(inside, w / np.gradient(m1)[:, None], np.nan)
    col = np.nanmax(prof, axis=0)
    np.testing.assert_allclose(np.where(inside, prof, col[None, :]), np.broadcast_to(col, prof.shape), rtol=1e-12)
    # and proportional to the cell integral of the density (the ratio is one constant).  Interior
    # columns only: the end columns take np.gradient's one-sided differences (first order), the
    # convention of the reference's prior (src/exmarg/search/pe.py:250-255)
    ratio = col[97::97] / cells[1:]
    np.testing.assert_allclose(ratio / ratio[0], 1.0, rtol=1e-4)        # O(dz^2) cell rule
    assert w.sum() == pytest.approx(1.0, abs=1e-12)


def _assembly_case(oracle, prior, ymin):
    """pe_intrinsic of the oracle on tables constant per node, e^{c_k} = A + B ln s_k, with every arrival
    time walked: node k of template m then has Lambda = 87 (A + B ln s_k) at y = sqrt(s_k / u_m), which is
    linear in ln y along each template (ln s_k = 2 ln y + ln u_m)"""
    A, B = 10.0, 1.0
    s = synth.S_NODES
    nn = s.size
    shape = (87, 49, 4, 5)
    c = np.log(A + B * np.log(s)).astype(np.float32)
    tabs = oracle.Tables(tables=[np.full(shape, ck, np.float32) for ck in c], tau_idx=synth.lag_indices(43))
    beta, delta = tiny_grids(4, 5)
    store = dict(s=s, alpha_lo=np.full(nn, 0.16), alpha_hi=np.full(nn, 4e4), beta=beta, delta=delta,
                 tau_idx=synth.lag_indices(43))
    rng = np.random.default_rng(12)
    nf, lo, nt = 40, 10, 256
    u = np.array([0.9, 5.0, 60.0, 800.0])
    nm = u.size
    bank = dict(W=(rng.standard_normal((nm, nf)) + 1j * rng.standard_normal((nm, nf))).astype(np.complex64),
                u=u, M=np.array([10.0, 20.0, 40.0, 80.0]), q=np.ones(nm), chi=np.zeros(nm), lo=lo, n_time=nt)
    seg = dict(d_H=rng.standard_normal((1, nf)) + 1j * rng.standard_normal((1, nf)),
               d_L=rng.standard_normal((1, nf)) + 1j * rng.standard_normal((1, nf)),
               sig2_H=np.ones(nf), sig2_L=np.ones(nf), rows=np.array([0]))
    res = oracle.pe_intrinsic(seg, bank, store, tabs, prior=prior, ymin=ymin, n_loud=0)
    return res, s, u, A, B, c


@pytest.mark.parametrize("prior,ymin", [("volume", 0.1267), ("madau-dickinson", 0.0982)])
def test_pe_assembly_closed_form(oracle, prior, ymin):
    res, s, u, A, B, c = _assembly_case(oracle, prior, ymin)
    density = oracle.distance_density(prior)
    for m in range(u.size):
        a = s / u[m]
        alo = min(0.16, ymin * ymin)
        valid = (np.sqrt(a) >= ymin) & (a > alo) & (a < 4e4)
        ks = np.nonzero(valid)[0]
        y_first, y_last = np.sqrt(a[ks[0]]), np.sqrt(a[ks[-1]])
        lam_first = 87.0 * np.exp(np.float64(c[ks[0]]))
        has_b = ks[0] == 0 or s[ks[0] - 1] / u[m] <= alo
        y_lo = np.sqrt(alo * (1 + 1e-9)) if has_b else y_first
        lam = lambda y: 87.0 * (A + B * (2.0 * np.log(y) + np.log(u[m])))
        if density is None:
            p = lambda y: 3.0 * ymin ** 3 * y ** -4
        else:
            p = lambda y: np.interp(np.log(y), density[0], density[1], left=0.0, right=0.0) / y
        pts = list(np.sqrt(a[ks[1:-1]]))
        total = integrate.quad(lambda y: p(y) * lam(y), y_first, y_last, points=pts or None, epsrel=1e-12,
                               limit=500)[0]
        if y_lo < y_first:
            total += lam_first * integrate.quad(p, y_lo, y_first, epsrel=1e-12)[0]
        tol = 1e-6 if density is None else 2e-5
        assert res["ln_lambda"][0, m] == pytest.approx(np.log(total), abs=tol)
    w = oracle.cell_weights(np.array([10.0, 20.0, 40.0, 80.0]), np.ones(4), np.zeros(4))
    assert res["ln_b"][0] == pytest.approx(np.log(np.sum(w * np.exp(res["ln_lambda"][0]))), abs=1e-12)


# ------------------------------------------------------------------ reference-semantics mode
def test_filter_ref_single_bin_and_against_fp64(oracle):
    # the complex64 filter reproduces the single-bin closed form to float32 precision and the complex128
    # filter to the precision of a complex64 FFT
    n_time, lo, nf, nu0 = 256, 10, 20, 13
    W = np.zeros((1, nf), np.complex64)
    W[0, nu0 - lo] = 0.5 - 0.25j
    d = np.full(nf, 2.0 - 1.0j)
    s2 = np.full(nf, 4.0)
    traj = oracle.filter_ref(W, lo, n_time, {"H": d, "L": d}, {"H": s2, "L": s2})
    assert all(v.dtype == np.float32 for v in traj.values())
    w = complex(W[0, nu0 - lo]) * np.conj(2.0 - 1.0j) / 4.0
    np.testing.assert_allclose(traj["lqH"][0], np.log(abs(w)), rtol=0, atol=2e-6)
    rng = np.random.default_rng(1)
    n_time, lo, nf = 2048, 40, 200
    W = (rng.standard_normal((3, nf)) + 1j * rng.standard_normal((3, nf))).astype(np.complex64)
    dd = {I: rng.standard_normal(nf) + 1j * rng.standard_normal(nf) for I in "HL"}
    ss = {I: rng.uniform(0.5, 2, nf) for I in "HL"}
    a = oracle.filter_ref(W, lo, n_time, dd, ss)
    b = oracle.filter_fp64(W, lo, n_time, dd, ss)
    loud = b["lqL"] > np.log(3.0)                   # away from the zeros of z
    assert np.abs(a["lqL"] - b["lqL"])[loud].max() < 1e-5
    dph = np.angle(np.exp(1j * (a["phL"].astype(float) - b["phL"])))[loud]
    assert np.abs(dph).max() < 1e-5


def test_window_max_brute_force(oracle):
    rng = np.random.default_rng(3)
    lq = rng.standard_normal((2, 50)).astype(np.float32)
    got = oracle.window_max(lq, -4, 6)
    ref = np.array([[max(lq[m, (t + k) % 50] for k in range(-4, 7)) for t in range(50)] for m in range(2)])
    np.testing.assert_array_equal(got, ref)


def test_select_ref_ranks_the_float32_score(oracle):
    nt, K, S = 400, 30, 50
    rng = np.random.default_rng(9)
    traj = {"lqH": rng.standard_normal((3, nt)).astype(np.float32),
            "lqL": rng.standard_normal((3, nt)).astype(np.float32)}
    tau = synth.lag_indices(3)
    sel, lw = oracle.select_times_ref(traj, tau, K, S, seed=5)
    pick = np.random.default_rng(5).integers(0, nt - K, size=(3, S))
    for m in range(3):
        mx = np.array([max(traj["lqH"][m, (t + k) % nt] for k in range(-3, 4)) for t in range(nt)], np.float32)
        score = np.exp(np.float32(2) * traj["lqL"][m]) + np.exp(np.float32(2) * mx)
        assert score.dtype == np.float32
        top = sel[m][lw[m] == 0]
        others = np.setdiff1d(np.arange(nt), top)
        assert top.size == K and score[top].min() >= score[others].max()
        np.testing.assert_array_equal(np.sort(others[pick[m]]), sel[m][lw[m] != 0])
    # well separated scores: the float32 and the float64 rankings pick the same times
    a = oracle.select_times(traj, tau, K, S, seed=5)
    if (oracle.top_set_differences(traj, tau, K, S, seed=5) == 0).all():
        np.testing.assert_array_equal(a[0], sel)


def test_top_set_flips_on_config1_trajectories(oracle):
    """how often the reference's float32 ranking leaves out a time of the float64 top set, on the
    trajectories of the 16 s, 20-1024 Hz segment with its 64-template bank (n_loud 1024): a flip is a
    near-tie at the K-th score, so it is rare"""
    from exmarg_amd.search.segments import read_segments
    seg = read_segments(synth.simulate_at_snr(1, 20261203, (30.0, 0.8, 0.1), 12.0, 16.0, 20.0, 1024.0)[0])
    rows = synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 64)
    W, u = synth.template_factors(rows[:, 0], rows[:, 1], rows[:, 2], seg)
    lo, n_time = int(seg["nu"][0]), 8 * int(seg["nu"][-1])
    traj = oracle.filter_fp64(W, lo, n_time, {I: seg["d_" + I][0] for I in "HL"},
                              {I: seg["sig2_" + I] for I in "HL"})
    flips = oracle.top_set_differences(traj, synth.lag_indices(43), 1024, 1024, seed=0)
    print("top-1024 times left out by the float32 ranking, per template: total %d, max %d, templates %d of 64"
          % (flips.sum(), flips.max(), (flips > 0).sum()))
    assert flips.max() <= 4 and flips.sum() <= 32


def test_gather_orders_agree_bit_for_bit(oracle):
    # the checker's slab order and the reference's (template, lag) job order compute the same sums
    traj = tiny_traj(seed=21)
    T = table_from(lambda r, c, b, d: np.sin(0.7 * r + 0.3 * c) + 0.2 * b - 0.1 * d)
    beta, delta = tiny_grids(T.shape[2], T.shape[3])
    tau = synth.lag_indices(2)
    tabs = oracle.Tables(tables=[T, T + 0.5], tau_idx=tau)
    times = oracle.select_times(traj, tau, 8, 8, seed=2)
    ids = np.arange(NM)
    lna = np.array([[0.3, 0.1], [-1.2, 0.0], [2.0, 1.5]])
    tab = np.array([[0, 1], [1, 0], [0, 1]])
    a = oracle.gather_pairs(traj, tabs, beta, delta, ids, lna, tab, times, order="slab")
    b = oracle.gather_pairs(traj, tabs, beta, delta, ids, lna, tab, times, order="reference")
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_oracle_under_sanitizers():
    """every public function of the C oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
    and run on inputs that reach its edges (oracle/sanitize_check.c), with closed-form value checks"""
    import shutil
    import subprocess
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle")
    if shutil.which("make") is None or shutil.which("gcc") is None:
        pytest.skip("make and gcc are needed to build the sanitized oracle")
    r = subprocess.run(["make", "-s", "-C", here, "sanitize"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "all checks passed" in r.stdout


def test_strain_segment_closed_form(oracle):
    """a tone A cos(2 pi k t / T + phi) on bin k gives d_k = A T / 2 e^{i phi} and nothing on the other
    bins (for tones on distinct bins of the band, odd and even lengths)"""
    for n, fs in ((8192, 1024.0), (6001, 750.0)):
        T = n / fs
        t = np.arange(n) / fs
        tones = {37: (1.5, 0.3), 38: (0.25, -2.0), 900: (3.0, 1.1)}
        x = sum(A * np.cos(2 * np.pi * k * t / T + ph) for k, (A, ph) in tones.items())
        d = oracle.segment_from_strain(x, fs, 30, 1000)[0]
        want = np.zeros(1000, complex)
        for k, (A, ph) in tones.items():
            want[k - 30] = A * T / 2 * np.exp(1j * ph)
        np.testing.assert_allclose(d, want, rtol=0, atol=1e-11 * T)


def test_strain_white_noise_normalization(oracle):
    """white noise of one-sided density S gives E|d|^2 = T S / 2 on every bin of the band"""
    rng = np.random.default_rng(1)
    fs, T, S = 1024.0, 4.0, 3e-3
    x = rng.normal(0.0, np.sqrt(S * fs / 2.0), (400, int(fs * T)))
    d = oracle.segment_from_strain(x, fs, 80, 1121)
    np.testing.assert_allclose(np.mean(np.abs(d) ** 2), T * S / 2.0, rtol=1e-2)


def test_strain_psd_closed_form(oracle):
    """blocks carrying tones of amplitude A_b on bin k: the estimate on k is scale / N sum_b (A_b T / 2)^2,
    and the tone-free bins stay at zero"""
    n, fs, N = 4096, 512.0, 5
    T = n / fs
    t = np.arange(n) / fs
    amps = np.array([1.0, 2.0, 0.5, 3.0, 1.5])
    blocks = np.stack([A * np.cos(2 * np.pi * 100 * t / T + 0.7 * b) for b, A in enumerate(amps)])
    scale = N / (N - 2.0)
    p = oracle.psd_from_strain(blocks, fs, 90, 20, scale)
    want = np.zeros(20)
    want[10] = scale / N * np.sum((amps * T / 2) ** 2)
    np.testing.assert_allclose(p, want, rtol=1e-12, atol=1e-20 * T * T)
