"""
CPU oracle of the exmarg segment analysis (exmarg.search), restated from the
reference in numpy plus the C gather/selection of oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, as the checker.  The analysis package
(exmarg_amd) never imports it.

Precision ("the FP64 oracle"): the matched filter runs in complex128 and its
outputs are rounded once to float32 trajectories (the storage type of the
reference, src/exmarg/search/bank.py:62-63); every later step is float64 on
those float32 trajectories and float32 tables, as the reference CPU path
computes (src/exmarg/search/gather.py:32-100 under numba).  Arrival-time
ranking uses the float64 score of oracle.c (oracle_pexp) with ties to the
smaller time.

Reference-semantics mode (the reference's own device='cpu' arithmetic, for comparison with the
reference's behaviour and as the CPU baseline): filter_ref keeps the matched filter in complex64 with
numpy's single-precision FFT, select_times_ref ranks a float32 score with numpy's float32 exp and
np.argpartition; the gather is the same float64 C gather (numba's promotion in the reference).
pe_intrinsic(..., semantics="ref") runs the whole chain that way.

Functions and the reference lines they follow:
  filter_fp64          src/exmarg/search/bank.py:56-78  (MassBank.trajectories)
  filter_ref           the same in complex64 / float32 as numpy >= 2 evaluates it
  select_times         src/exmarg/search/gather.py:374-420
  select_times_np      the same in numpy (lexsort), for small inputs
  select_times_ref     src/exmarg/search/gather.py:396-420 with its float32 score and argpartition
  gather               src/exmarg/search/gather.py:114-140 + _lse_lags :103-107
  node_geometry        src/exmarg/search/pe.py:58-84
  distance_weights     src/exmarg/search/pe.py:260-286 (+ priors.py:31-38)
  node_values          src/exmarg/search/pe.py:168-200
  pe_intrinsic         src/exmarg/search/pe.py:289-346
  log_mean             src/exmarg/search/pe.py:135-142
  to_source_frame      src/exmarg/search/pe.py:87-132
  pe_redshift          src/exmarg/search/pe.py:203-257
  dl_of_z              src/exmarg/search/pe.py:26-28 (astropy Planck18 restated: flat LCDM with
                       photons and one massive neutrino species, the Komatsu et al. 2011
                       Eq. 26 fitting function for the neutrino density)
  segment_from_strain  src/exmarg/pp.py:38-72 (d = dt rfft(x) on the band bins)
  psd_from_strain      notebooks/07_estimated_psd.ipynb cell 5 (mean periodogram times N/(N-2))
  bd_grids             scripts/build_tables/build_bd.py:52-60 (beta and Delta grids of stage 3)
  bd_sum_orders        scripts/build_tables/build_bd.py:61-76 (stage 3: the sum over the orders)
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
NPHI = 96
D0 = 10410.0

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        i64, i32, f64, f32, u64 = ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_uint64
        L.oracle_pexp.restype = f64
        L.oracle_pexp.argtypes = [f64]
        L.oracle_synth_fill.argtypes = [u64, i64, i32, i32, i32, i32, i32, P, P]
        L.oracle_gather.argtypes = [P, P, P, P, i64, P, P, i64, i32, P, i32, P, i32, i32, i32,
                                    P, f64, f64, f64, P, P, i64, i32, P, P, i32]
        L.oracle_gather_synth.argtypes = [P, P, P, P, i64, P, P, i64, i32, P, i32, P, u64, i32, i32, i32, i32,
                                          P, f64, f64, f64, P, P, i64, i32, P, P, i32]
        L.oracle_select.argtypes = [P, P, i64, i64, i32, i32, i32, i32, P, f32, P, P]
        L.oracle_scores.argtypes = [P, P, i64, i32, i32, P]
        L.oracle_set_threads.argtypes = [i32]
        _lib = L
    return _lib


def set_threads(n):
    """threads of the oracle's OpenMP loops"""
    lib().oracle_set_threads(int(n))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ filter
def filter_fp64(W, lo, n_time, d, sig2, threads=None):
    """trajectories of one segment: dict lqH, phH, lqL, phL of float32 (n_template, n_time).

    W (n_template, n_freq) complex64; d, sig2 dicts {'H','L'} of (n_freq,) arrays.
    z = FFT_N[pad(W conj(d)/sig2)] in complex128 (forward, unnormalized, scipy.fft's pocketfft);
    lq = 1/2 ln max(|z|^2, 1e-38), ph = atan2(Im z, Re z), each rounded once to float32.
    Blocks of 64 templates run on `threads` host threads (default OMP_NUM_THREADS, else at most 16)."""
    from concurrent.futures import ThreadPoolExecutor
    from scipy import fft as sfft
    W = np.asarray(W)
    nm, nf = W.shape
    out = {}
    nthr = max(1, int(threads or os.environ.get("OMP_NUM_THREADS") or min(16, os.cpu_count() or 1)))
    for I in "HL":
        w = np.conj(np.asarray(d[I], np.complex128)) / np.asarray(sig2[I], np.float64)
        lq = np.empty((nm, n_time), np.float32)
        ph = np.empty((nm, n_time), np.float32)

        def block(a, w=w, lq=lq, ph=ph):
            b = min(a + 64, nm)
            pad = np.zeros((b - a, n_time), np.complex128)
            pad[:, lo:lo + nf] = W[a:b].astype(np.complex128) * w[None, :]
            z = sfft.fft(pad, axis=-1, overwrite_x=True)
            lq[a:b] = (0.5 * np.log(np.maximum(z.real ** 2 + z.imag ** 2, 1e-38))).astype(np.float32)
            ph[a:b] = np.arctan2(z.imag, z.real).astype(np.float32)

        with ThreadPoolExecutor(max_workers=nthr) as pool:
            list(pool.map(block, range(0, nm, 64)))
        out["lq" + I], out["ph" + I] = lq, ph
    return out


def filter_ref(W, lo, n_time, d, sig2, chunk=1024):
    """trajectories in the reference's CPU arithmetic (src/exmarg/search/bank.py:66-75): the weighted
    data rounded to complex64, the product with W and the FFT in complex64 (numpy >= 2 keeps single
    precision), |z|^2, the 1e-38 floor, ln and atan2 in float32"""
    W = np.asarray(W, np.complex64)
    nm, nf = W.shape
    out = {}
    for I in "HL":
        w = (np.conj(np.asarray(d[I], np.complex128)) / np.asarray(sig2[I], np.float64)).astype(np.complex64)
        lq = np.empty((nm, n_time), np.float32)
        ph = np.empty((nm, n_time), np.float32)
        for a in range(0, nm, chunk):
            b = min(a + chunk, nm)
            pad = np.zeros((b - a, n_time), np.complex64)
            pad[:, lo:lo + nf] = W[a:b] * w[None, :]
            z = np.fft.fft(pad, axis=-1)
            if z.dtype != np.complex64:
                raise RuntimeError("numpy.fft returned %s: the reference semantics need numpy >= 2" % z.dtype)
            re, im = z.real, z.imag
            lq[a:b] = np.float32(0.5) * np.log(np.maximum(re * re + im * im, np.float32(1e-38)))
            ph[a:b] = np.arctan2(im, re)
        out["lq" + I], out["ph" + I] = lq, ph
    return out


# ------------------------------------------------------------------ selection
def lag_range(tau_idx):
    lags = np.asarray(tau_idx).ravel().astype(int)
    lo, hi = int(lags.min()), int(lags.max())
    if not np.array_equal(np.sort(lags), np.arange(lo, hi + 1)):
        raise ValueError("lags are not contiguous")
    return lo, hi


def draw_picks(n_template, n_time, K, S, seed):
    """the host random draw of select_times (src/exmarg/search/gather.py:397)"""
    return np.random.default_rng(seed).integers(0, n_time - K, size=(n_template, S))


def random_log_weight(n_time, K, S):
    return np.float32(np.log((n_time - K) / S))


def select_times(traj, tau_idx, n_loud, n_random, seed, pick=None):
    """(sel (n_template, K+S) int32 sorted, lw float32) by the C oracle; pick overrides the draw
    (rows of the draw of a larger bank)"""
    K, S = int(n_loud), int(n_random)
    lqH = np.ascontiguousarray(traj["lqH"], np.float32)
    lqL = np.ascontiguousarray(traj["lqL"], np.float32)
    nm, nt = lqH.shape
    lo, hi = lag_range(tau_idx)
    if pick is None:
        pick = draw_picks(nm, nt, K, S, seed)
    pick = np.ascontiguousarray(pick, np.int64)
    sel = np.empty((nm, K + S), np.int32)
    lw = np.empty((nm, K + S), np.float32)
    lib().oracle_select(_p(lqH), _p(lqL), nm, nt, lo, hi, K, S, _p(pick), float(random_log_weight(nt, K, S)),
                        _p(sel), _p(lw))
    return sel, lw


def pexp_np(x):
    """numpy evaluation of the oracle_pexp operation sequence (no fused operations)"""
    import math
    x = np.asarray(x, np.float64)
    k = np.rint(x * 1.4426950408889634)
    r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10
    p = np.full_like(x, 1.0 / math.factorial(13))
    for i in range(12, -1, -1):
        p = p * r + 1.0 / math.factorial(i)
    return np.ldexp(p, k.astype(np.int64))


def select_times_np(traj, tau_idx, n_loud, n_random, seed):
    """select_times in numpy with np.lexsort for the ranking (small inputs)"""
    K, S = int(n_loud), int(n_random)
    lqH = np.asarray(traj["lqH"], np.float32)
    lqL = np.asarray(traj["lqL"], np.float32)
    nm, nt = lqH.shape
    lo, hi = lag_range(tau_idx)
    pick = draw_picks(nm, nt, K, S, seed)
    sel = np.empty((nm, K + S), np.int32)
    lw = np.zeros((nm, K + S), np.float32)
    t = np.arange(nt)
    for m in range(nm):
        mx = np.max(np.stack([np.roll(lqH[m], -tau) for tau in range(lo, hi + 1)]), axis=0)
        score = pexp_np(2.0 * lqL[m].astype(np.float64)) + pexp_np(2.0 * mx.astype(np.float64))
        order = np.lexsort((t, -score))
        top = order[:K]
        rest = np.setdiff1d(t, top)
        both = np.concatenate([top, rest[pick[m]]])
        w = np.concatenate([np.zeros(K, np.float32), np.full(S, random_log_weight(nt, K, S), np.float32)])
        o = np.argsort(both, kind="stable")
        sel[m], lw[m] = both[o], w[o]
    return sel, lw


def window_max(lq, lo, hi):
    """max over tau in [lo, hi] of lq[:, (t + tau) mod n_time], exact (any evaluation order gives the
    same float32 maximum)"""
    lq = np.asarray(lq, np.float32)
    nt = lq.shape[1]
    ext = np.concatenate([lq, lq[:, :hi - lo]], axis=1)
    win = np.lib.stride_tricks.sliding_window_view(ext, hi - lo + 1, axis=1)[:, :nt].max(axis=-1)
    return np.roll(win, -lo, axis=1)


def select_times_ref(traj, tau_idx, n_loud, n_random, seed, block=256):
    """select_times with the reference's own ranking (src/exmarg/search/gather.py:396-420): the score
    e^{2 lqL} + e^{2 max lqH} in float32 (numpy's float32 exp), the n_loud largest by np.argpartition
    (whichever of a tie it returns), the drawn times indexing the rest in time order.
    returns (sel, lw) like select_times."""
    K, S = int(n_loud), int(n_random)
    lqH = np.asarray(traj["lqH"], np.float32)
    lqL = np.asarray(traj["lqL"], np.float32)
    nm, nt = lqH.shape
    lo, hi = lag_range(tau_idx)
    pick = draw_picks(nm, nt, K, S, seed)
    two = np.float32(2.0)
    sel = np.empty((nm, K + S), np.int32)
    for a in range(0, nm, block):
        b = min(a + block, nm)
        score = np.exp(two * lqL[a:b]) + np.exp(two * window_max(lqH[a:b], lo, hi))
        top = np.argpartition(-score, K, axis=1)[:, :K]
        keep = np.ones((b - a, nt), bool)
        np.put_along_axis(keep, top, False, axis=1)
        rest = np.nonzero(keep)[1].reshape(b - a, nt - K)
        sel[a:b] = np.concatenate([top, np.take_along_axis(rest, pick[a:b], axis=1)], axis=1)
    lw = np.zeros((nm, K + S), np.float32)
    lw[:, K:] = random_log_weight(nt, K, S)
    order = np.argsort(sel, axis=1, kind="stable")
    return (np.ascontiguousarray(np.take_along_axis(sel, order, axis=1)),
            np.ascontiguousarray(np.take_along_axis(lw, order, axis=1)))


def top_set_differences(traj, tau_idx, n_loud, n_random, seed):
    """per template, the number of times in the top-n_loud set of the float64 ranking (select_times)
    that the reference's float32 ranking (select_times_ref) leaves out"""
    a = select_times(traj, tau_idx, n_loud, n_random, seed)
    b = select_times_ref(traj, tau_idx, n_loud, n_random, seed)
    out = np.empty(a[0].shape[0], np.int64)
    for m in range(out.size):
        out[m] = np.setdiff1d(a[0][m][a[1][m] == 0], b[0][m][b[1][m] == 0]).size
    return out


# ------------------------------------------------------------------ gather
def synth_table(seed, node, shape, tau_idx, kind="smooth"):
    """one synthetic table in host memory (reference layout), by the C recipe"""
    out = np.empty(shape, np.float32)
    tau = np.ascontiguousarray(tau_idx, np.int64)
    lib().oracle_synth_fill(int(seed), int(node), 1 if kind == "rough" else 0, *[int(v) for v in shape],
                            _p(tau), _p(out))
    return out


class Tables:
    """the tables a gather reads: host arrays (reference layout) or a synthetic recipe.  With a recipe and
    materialize=True, each gather fills the tables it reads into host memory first (one pass of
    oracle_synth_fill per table, OpenMP) and reads them from there: the same values, read at memory speed
    instead of recomputed per lookup."""

    def __init__(self, tables=None, synth=None, shape=None, tau_idx=None, materialize=False):
        self.tables = tables
        self.synth = synth            # dict(seed, kind, nodes) with nodes the node id of each table
        self.materialize = bool(materialize)
        if tables is not None:
            self.arrs = [np.ascontiguousarray(t, np.float32) for t in tables]
            self.shape = self.arrs[0].shape
        else:
            self.shape = tuple(shape)
        self.tau_idx = np.ascontiguousarray(tau_idx, np.int64)

    def filled(self, tab_ids):
        """Tables in host memory holding (at their own indices) the tables of tab_ids"""
        used = np.unique(np.asarray(tab_ids))
        kind = self.synth.get("kind", "smooth")
        arrs = {int(k): synth_table(self.synth["seed"], int(self.synth["nodes"][k]), self.shape, self.tau_idx, kind)
                for k in used}
        first = arrs[int(used[0])]
        ntab = len(self.synth["nodes"])
        return Tables(tables=[arrs.get(k, first) for k in range(ntab)], tau_idx=self.tau_idx)


def lse_lags(om, os_, n_time):
    mx = om.max(axis=1)
    tot = (os_ * np.exp(om - mx[:, None, :])).sum(axis=1)
    return mx + np.log(tot) - np.log(n_time)


def gather_pairs(traj, tabs, beta, delta, mass_ids, lna, tab_ids, times=None, order="slab"):
    """(om, os) (NS, ntau, nsel) running max and sum of every (slot, lag, template).  order 'slab' runs
    the threads over the templates of one (slot, lag) at a time (cache-friendly, the checker's order);
    'reference' hands (template, lag) jobs to the threads as the reference's prange does (the CPU
    baseline's order).  Both give the same values."""
    if order not in ("slab", "reference"):
        raise ValueError("order is 'slab' or 'reference', not %r" % (order,))
    slab = 1 if order == "slab" else 0
    lqH, phH, lqL, phL = (np.ascontiguousarray(traj[k], np.float32) for k in ("lqH", "phH", "lqL", "phL"))
    n_time = lqH.shape[1]
    ids = np.ascontiguousarray(mass_ids, np.int64)
    tab = np.ascontiguousarray(tab_ids, np.int64)
    nsel, ns = tab.shape
    ntau, nphif, nb, nd = tabs.shape
    offb = np.ascontiguousarray((np.asarray(lna, np.float64) - float(beta[0])) / float(beta[1] - beta[0]))
    rdb = 2.0 / float(beta[1] - beta[0])
    dd = float(delta[1] - delta[0])
    offd, rdd = -float(delta[0]) / dd, 1.0 / dd
    if times is None:
        sel, lw, nw, use = np.zeros((1, 1), np.int32), np.zeros((1, 1), np.float32), 0, 0
    else:
        sel = np.ascontiguousarray(times[0], np.int32)
        lw = np.ascontiguousarray(times[1], np.float32)
        nw, use = sel.shape[1], 1
    om = np.empty((ns, ntau, nsel))
    os_ = np.empty((ns, ntau, nsel))
    if tabs.tables is None and tabs.materialize:
        tabs = tabs.filled(tab)
    L = lib()
    if tabs.tables is not None:
        ptrs = (ctypes.c_void_p * len(tabs.arrs))(*[a.ctypes.data for a in tabs.arrs])
        L.oracle_gather(_p(lqH), _p(phH), _p(lqL), _p(phL), n_time, _p(ids), _p(tab), nsel, ns,
                        _p(tabs.tau_idx), ntau, ptrs, nphif, nb, nd, _p(offb), rdb, offd, rdd,
                        _p(sel), _p(lw), nw, use, _p(om), _p(os_), slab)
    else:
        nodes = np.ascontiguousarray(tabs.synth["nodes"], np.int64)
        L.oracle_gather_synth(_p(lqH), _p(phH), _p(lqL), _p(phL), n_time, _p(ids), _p(tab), nsel, ns,
                              _p(tabs.tau_idx), ntau, _p(nodes), int(tabs.synth["seed"]),
                              1 if tabs.synth.get("kind", "smooth") == "rough" else 0,
                              nphif, nb, nd, _p(offb), rdb, offd, rdd, _p(sel), _p(lw), nw, use, _p(om), _p(os_),
                              slab)
    return om, os_


def gather(traj, tabs, beta, delta, mass_ids, lna, tab_ids, times=None, order="slab"):
    """ln Lambda (NS, nsel) float64 (src/exmarg/search/gather.py:114-140)"""
    om, os_ = gather_pairs(traj, tabs, beta, delta, mass_ids, lna, tab_ids, times, order)
    return lse_lags(om, os_, traj["lqH"].shape[1])


# ------------------------------------------------------------------ assembly
def node_geometry(s, alpha_lo, alpha_hi, u, ymin):
    """valid (nn, nm), lna (nn, nm), y (nn+1, nm), bnode (nm,), blna (nm,) (src/exmarg/search/pe.py:58-84)"""
    nn, nm = len(s), len(u)
    valid = np.zeros((nn, nm), bool)
    lna = np.zeros((nn, nm))
    y = np.zeros((nn + 1, nm))
    bnode = np.full(nm, -1)
    blna = np.zeros(nm)
    for m in range(nm):
        for k in range(nn):
            alo = min(alpha_lo[k], ymin * ymin)
            a = s[k] / u[m]
            ok = (np.sqrt(a) >= ymin) and (a > alo) and (a < alpha_hi[k])
            if not ok:
                continue
            valid[k, m] = True
            lna[k, m] = np.log(a)
            y[k, m] = np.sqrt(a)
            s_prev = s[k - 1] if k > 0 else 0.0
            if s_prev / u[m] <= alo:
                a_b = float(alo) * (1.0 + 1e-9)
                bnode[m] = k
                blna[m] = np.log(a_b)
                y[nn, m] = np.sqrt(a_b)
    return valid, lna, y, bnode, blna


_PRIOR = None


def distance_density(name):
    """None for 'volume'; (lny, p_lny) for 'madau-dickinson' (src/exmarg/priors.py:31-38)"""
    global _PRIOR
    if name == "volume":
        return None
    if name == "madau-dickinson":
        if _PRIOR is None:
            z = np.load(os.path.join(HERE, "..", "exmarg_amd", "data", "distance_prior.npz"))
            _PRIOR = (z["lny"], z["p_lny"])
        return _PRIOR
    raise ValueError("the distance prior is 'volume' or 'madau-dickinson', not %r" % name)


def distance_weights(y, prior, ymin):
    """ln w_j with sum_j w_j Lambda(y_j) = int p(y) Lambda(y) dy, Lambda linear in ln y
    (src/exmarg/search/pe.py:260-286)"""
    density = distance_density(prior)
    x = np.log(np.asarray(y, float))
    w = np.zeros_like(x)
    for i in range(x.size - 1):
        a, b = x[i], x[i + 1]
        h = b - a
        if density is None:
            c, k = 3.0 * ymin ** 3, 3.0
            ea, eb = np.exp(-k * a), np.exp(-k * b)
            I0 = (ea - eb) / k
            I1 = (a * ea - b * eb) / k + (ea - eb) / k ** 2
            w[i] += c * (b * I0 - I1) / h
            w[i + 1] += c * (I1 - a * I0) / h
        else:
            t = np.linspace(a, b, 401)
            f = np.interp(t, density[0], density[1], left=0.0, right=0.0)
            w[i] += np.trapezoid(f * (b - t) / h, t)
            w[i + 1] += np.trapezoid(f * (t - a) / h, t)
    with np.errstate(divide="ignore"):
        return np.log(w)


def bank_weights(yN, prior, ymin, M=None, u=None):
    """lnW (nn+1, nm) of pe_intrinsic (src/exmarg/search/pe.py:321-327)"""
    lnW = np.full(yN.shape, -np.inf)
    for m in range(yN.shape[1]):
        pts = np.nonzero(yN[:, m] > 0)[0]
        if pts.size < 2:
            raise ValueError("template %d reaches fewer than two store nodes" % m)
        order = pts[np.argsort(yN[pts, m])]
        lnW[order, m] = distance_weights(yN[order, m], prior, ymin)
    return lnW


def node_values(traj, times, store, tabs, geometry, groups):
    """lamN (nn+1, nm) of one segment (src/exmarg/search/pe.py:168-200) with one gather per node group
    and one per group for the boundary points"""
    valid, lna, _, bnode, blna = geometry
    nn, nm = valid.shape
    lamN = np.full((nn + 1, nm), -np.inf)
    for ks in groups:
        k0 = ks[0]
        beta, delta = store["beta"], store["delta"]
        ok = valid[ks]
        ids = np.nonzero(ok.any(axis=0))[0]
        lnaM = np.where(ok[:, ids], lna[ks][:, ids], 0.0).T
        tid = np.tile(np.asarray(ks), (ids.size, 1))
        bids = np.nonzero((bnode >= k0) & (bnode <= ks[-1]))[0]
        if ids.size:
            val = gather(traj, tabs, beta, delta, ids, lnaM, tid, times)
            for i, k in enumerate(ks):
                keep = ok[i][ids]
                lamN[k, ids[keep]] = val[i][keep]
        if bids.size:
            val = gather(traj, tabs, beta, delta, bids, blna[bids][:, None], bnode[bids][:, None], times)
            lamN[nn, bids] = val[0]
    return lamN


def log_mean(lam, w):
    """ln sum_i w_i e^{lam_i} over finite lam and w > 0 (src/exmarg/search/pe.py:135-142)"""
    lam, w = np.asarray(lam), np.asarray(w)
    fin = np.isfinite(lam) & (w > 0)
    if not fin.any():
        return -np.inf
    mx = np.where(fin, lam, -np.inf).max()
    return float(mx + np.log(np.sum(np.where(fin, w * np.exp(np.where(fin, lam, mx) - mx), 0.0))))


def cell_weights(M, q, chi):
    cell = np.ones(np.asarray(M).size)
    for ax in (M, q, chi):
        vals = np.unique(ax)
        if vals.size > 1:
            cell *= np.gradient(vals)[np.searchsorted(vals, ax)]
    return cell / cell.sum()


def analyze_segment(seg_row, bank, store, tabs, geometry, n_loud, n_random, seed, nodes_per_launch=6,
                    semantics="fp64"):
    """trajectories, times and lamN of one segment row.
    seg_row: dict d_H, d_L (n_freq,), sig2_H, sig2_L, row (int); bank: dict W, lo, n_time.
    semantics 'fp64' (the FP64 oracle) or 'ref' (the reference's CPU arithmetic: filter_ref and
    select_times_ref)."""
    if semantics not in ("fp64", "ref"):
        raise ValueError("semantics is 'fp64' or 'ref', not %r" % (semantics,))
    filt = filter_ref if semantics == "ref" else filter_fp64
    select = select_times_ref if semantics == "ref" else select_times
    traj = filt(bank["W"], bank["lo"], bank["n_time"], {"H": seg_row["d_H"], "L": seg_row["d_L"]},
                {"H": seg_row["sig2_H"], "L": seg_row["sig2_L"]})
    times = (select(traj, store["tau_idx"], n_loud, n_random, seed + int(seg_row["row"]))
             if n_loud > 0 else None)
    nn = len(store["s"])
    groups = [list(range(g, min(g + nodes_per_launch, nn))) for g in range(0, nn, nodes_per_launch)]
    return traj, times, node_values(traj, times, store, tabs, geometry, groups)


def pe_intrinsic(seg, bank, store, tabs, prior="volume", ymin=0.1267, n_loud=1024, n_random=1024, seed=0,
                 rows=None, semantics="fp64"):
    """dict ln_lambda (n_seg, nm), ln_b (n_seg,), lamN (n_seg, nn+1, nm) (src/exmarg/search/pe.py:289-346).
    seg: the read_segments dict; bank: dict W, u, M, q, chi, lo, n_time; store: dict s, alpha_lo, alpha_hi,
    beta, delta, tau_idx."""
    geometry = node_geometry(store["s"], store["alpha_lo"], store["alpha_hi"], bank["u"], ymin)
    lnW = bank_weights(geometry[2], prior, ymin)
    rows = np.arange(np.asarray(seg["d_H"]).shape[0]) if rows is None else np.asarray(rows)
    out = np.full((rows.size, bank["u"].size), -np.inf)
    lamNs = []
    for j, e in enumerate(rows):
        row = dict(d_H=seg["d_H"][e], d_L=seg["d_L"][e], sig2_H=seg["sig2_H"], sig2_L=seg["sig2_L"],
                   row=seg["rows"][e] if "rows" in seg else e)
        _, _, lamN = analyze_segment(row, bank, store, tabs, geometry, n_loud, n_random, seed, semantics=semantics)
        v = lnW + lamN
        mx = v.max(axis=0)
        out[j] = mx + np.log(np.exp(v - mx[None, :]).sum(axis=0))
        lamNs.append(lamN)
    w = cell_weights(bank["M"], bank["q"], bank["chi"])
    ln_b = np.array([log_mean(L, w) for L in out])
    return dict(ln_lambda=out, ln_b=ln_b, lamN=np.array(lamNs), lnW=lnW, geometry=geometry)


# ------------------------------------------------------------------ redshift
# Planck 2018 (TT,TE,EE+lowE+lensing+BAO, Planck Collaboration 2020 Table 2 col. 6) as astropy's Planck18
_H0, _OM0, _TCMB0, _NEFF, _MNU = 67.66, 0.30966, 2.7255, 3.046, (0.0, 0.0, 0.06)
_C_KMS = 299792.458


def _planck18_E(z):
    """H(z)/H0 of flat LCDM with photons and neutrinos (one massive species)"""
    sigma_sb, c, G = 5.670374419e-8, 299792458.0, 6.67430e-11
    h0_si = _H0 * 1e3 / 3.0856775814913673e22
    rho_c = 3.0 * h0_si ** 2 / (8.0 * np.pi * G)
    og0 = 4.0 * sigma_sb * _TCMB0 ** 4 / c ** 3 / rho_c
    tnu0 = (4.0 / 11.0) ** (1.0 / 3.0) * _TCMB0
    kb_ev = 8.617333262e-5
    p, k = 1.83, 0.3173
    prefac = 7.0 / 8.0 * (4.0 / 11.0) ** (4.0 / 3.0)
    nmassless = sum(1 for m in _MNU if m == 0)
    ymass = np.array([m for m in _MNU if m > 0]) / (kb_ev * tnu0)

    def nu_rel(zz):
        zz = np.asarray(zz, float)
        per = (1.0 + (k * ymass[:, None] / (1.0 + zz[None, :])) ** p) ** (1.0 / p)
        return prefac * (_NEFF / 3.0) * (per.sum(axis=0) + nmassless)

    onu0 = og0 * float(nu_rel(np.array([0.0]))[0])
    ode0 = 1.0 - _OM0 - og0 - onu0
    z = np.asarray(z, float)
    zp1 = 1.0 + z
    return np.sqrt(_OM0 * zp1 ** 3 + og0 * zp1 ** 4 * (1.0 + nu_rel(z.ravel()).reshape(z.shape)) + ode0)


def dl_of_z(z):
    """luminosity distance (Mpc) by Gauss-Legendre quadrature of 1/E on [0, z] (src/exmarg/search/pe.py:26-28)"""
    z = np.atleast_1d(np.asarray(z, float))
    x, wq = np.polynomial.legendre.leggauss(64)
    out = np.empty(z.size)
    for i, zi in enumerate(z):
        zz = 0.5 * zi * (x + 1.0)
        out[i] = 0.5 * zi * np.sum(wq / _planck18_E(zz))
    return (1.0 + z) * _C_KMS / _H0 * out


def to_source_frame(lamN, yN, Mdet, yq, m1, z):
    """ln Lambda (nm1, nz) from node values, column by column with np.interp (src/exmarg/search/pe.py:87-132)"""
    nn1, nm = lamN.shape
    surf = np.full((yq.size, nm), -np.inf)
    for m in range(nm):
        ok = np.isfinite(lamN[:, m]) & (yN[:, m] > 0)
        if ok.sum() < 2:
            continue
        ys, ls = yN[ok, m], lamN[ok, m]
        o = np.argsort(ys)
        ys, ls = ys[o], ls[o]
        inside = (yq >= ys[0]) & (yq <= ys[-1])
        surf[inside, m] = np.interp(yq[inside], ys, ls)
    out = np.full((m1.size, z.size), -np.inf)
    for iz in range(z.size):
        good = np.isfinite(surf[iz])
        if good.sum() < 2:
            continue
        Mg, Lg = Mdet[good], surf[iz, good]
        Md = 2.0 * m1 * (1.0 + z[iz])
        inb = (Md >= Mg[0]) & (Md <= Mg[-1])
        out[inb, iz] = np.interp(Md[inb], Mg, Lg)
    return out


def pe_redshift_weights(m1, z, mdet_range):
    dl = dl_of_z(z)
    lnp = np.log(np.maximum((1.0 + z) * dl ** 2 * np.gradient(dl, z), 1e-300))
    Md = 2.0 * m1[:, None] * (1.0 + z[None, :])
    inside = (Md >= mdet_range[0]) & (Md <= mdet_range[1])
    w = np.where(inside, np.exp(lnp)[None, :], 0.0) * np.gradient(m1)[:, None] * np.gradient(z)[None, :]
    return w / w.sum()


def pe_redshift(seg, bank, store, tabs, zmax=8.0, nz=1600, m1_range=(3.0, 150.0), nm1=1600, ymin=0.1267,
                mdet_range=(10.0, 500.0), n_loud=1024, n_random=1024, seed=0, rows=None, semantics="fp64"):
    """dict ln_lambda (n_seg, nm1, nz), ln_b (n_seg,) (src/exmarg/search/pe.py:203-257): node values per
    segment, the source-frame grid by to_source_frame and ln B by log_mean with pe_redshift_weights.
    bank: dict W, u, M, lo, n_time; seg and store as pe_intrinsic."""
    z = np.linspace(0.02, zmax, int(nz))
    m1 = np.geomspace(m1_range[0], m1_range[1], int(nm1))
    yq = D0 / dl_of_z(z)
    w = pe_redshift_weights(m1, z, mdet_range)
    geometry = node_geometry(store["s"], store["alpha_lo"], store["alpha_hi"], bank["u"], ymin)
    rows = np.arange(np.asarray(seg["d_H"]).shape[0]) if rows is None else np.asarray(rows)
    out = np.full((rows.size, m1.size, z.size), -np.inf)
    ln_b = np.empty(rows.size)
    for j, e in enumerate(rows):
        row = dict(d_H=seg["d_H"][e], d_L=seg["d_L"][e], sig2_H=seg["sig2_H"], sig2_L=seg["sig2_L"],
                   row=seg["rows"][e] if "rows" in seg else e)
        _, _, lamN = analyze_segment(row, bank, store, tabs, geometry, n_loud, n_random, seed, semantics=semantics)
        out[j] = to_source_frame(lamN, geometry[2], np.asarray(bank["M"], float), yq, m1, z)
        ln_b[j] = log_mean(out[j], w)
    return dict(m1=m1, z=z, ln_lambda=out, ln_b=ln_b)


# ---------------------------------------------------------------- stream front end
def segment_from_strain(x, fs, nu0, n_freq):
    """d = rfft(x)[nu0 : nu0 + n_freq] / fs of each row of x, complex128 (numpy's pocketfft): the
    segment convention of src/exmarg/pp.py:38-72 (E|d|^2 = T S / 2 for white noise of one-sided density S)"""
    x = np.atleast_2d(np.asarray(x, np.float64))
    return np.fft.rfft(x, axis=1)[:, nu0:nu0 + n_freq] / fs


def psd_from_strain(blocks, fs, nu0, n_freq, scale):
    """scale / N * sum over the N blocks (rows), in order, of |d_b|^2: the mean periodogram estimate of
    notebooks/07_estimated_psd.ipynb cell 5 (there from simulated periodograms), scale N/(N-2)"""
    d = segment_from_strain(blocks, fs, nu0, n_freq)
    acc = np.zeros(n_freq)
    for b in range(d.shape[0]):
        acc += d[b].real * d[b].real + d[b].imag * d[b].imag
    return acc * (scale / d.shape[0])


# ---------------------------------------------------------------- table build, stage 3
BD_ALPHA = np.geomspace(0.16, 4.0e4, 30)        # scripts/build_tables/build_bd.py:26


def bd_grids(lnq):
    """(beta, delta) grids of one node from its ln q grid (scripts/build_tables/build_bd.py:52-60):
    Delta = ln q1 - ln q2 on the 2 nq - 1 differences of the grid, sorted; beta from ln alpha_min + 2 ln q_min
    to ln alpha_max + 2 ln q_max in steps of 2 d(ln q)"""
    lnq = np.asarray(lnq, np.float64)
    last = lnq.size - 1
    dlq = float(lnq[1] - lnq[0])
    delta = np.concatenate([lnq - lnq[last], (lnq[last] - lnq)[::-1][1:]])
    dbeta = 2.0 * dlq
    b_lo = np.log(BD_ALPHA[0]) + 2.0 * lnq[0]
    b_hi = np.log(BD_ALPHA[-1]) + 2.0 * lnq[-1]
    nb = int(np.ceil((b_hi - b_lo) / dbeta)) + 1
    return b_lo + dbeta * np.arange(nb), delta


def bd_sum_orders(col, row, lnq, ladder):
    """stage 3 of the tables for one node (scripts/build_tables/build_bd.py:61-76), restated in numpy.

    col (nord, ntau, nq, nphif): lnK_n at q2 = q_max (the last column of each (q1, q2) block),
    row (nord, ntau, nq, nphif): lnK_n at q1 = q_max (the last row); float32 as stage 2 stores them.
    g_n(Delta) = col - 2 n ln q_max for Delta <= 0, row - 2 n ln q2 for Delta >= 0 (sorted in Delta);
    E_n(beta, Delta) = n beta + g_n(Delta); for every n_max of the ladder
    T = max_{n <= n_max} E_n + ln sum_{n <= n_max} exp(E_n - max), in float64, rounded to float32.
    Returns {n_max: T (ntau, nphif, nb, nd) float32}."""
    col = np.asarray(col)
    row = np.asarray(row)
    lnq = np.asarray(lnq, np.float64)
    nord, ntau, nq, nphif = col.shape
    last = nq - 1
    beta, delta = bd_grids(lnq)
    n = np.arange(nord, dtype=np.float64)
    E_nb = n[:, None] * beta[None, :]
    out = {nm: np.empty((ntau, nphif, beta.size, delta.size), np.float32) for nm in ladder}
    for ti in range(ntau):
        c = col[:, ti].astype(np.float64)
        r = row[:, ti].astype(np.float64)
        g_neg = c - 2.0 * n[:, None, None] * lnq[last]
        g_pos = r - 2.0 * n[:, None, None] * lnq[None, :, None]
        g = np.concatenate([g_neg, g_pos[:, ::-1, :][:, 1:, :]], axis=1)
        for pi in range(nphif):
            E = E_nb[:, :, None] + g[:, None, :, pi]
            mx_run = np.maximum.accumulate(E, axis=0)
            for nm in ladder:
                mxn = mx_run[nm]
                tot = np.zeros_like(mxn)
                for k in range(nm + 1):                  # in order of n
                    tot += np.exp(E[k] - mxn)
                out[nm][ti, pi] = (mxn + np.log(tot)).astype(np.float32)
    return out
