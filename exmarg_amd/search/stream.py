"""
A stream of short segments analysed with pe_redshift, the noise spectrum estimated from neighbouring
data.

    for res in pe_redshift_stream(groups, masses, store, keep=lambda g, row: row in rows_of_interest):
        res["ln_b"], res["kept"]

The reference analyses such a stream in batches of segments (src/exmarg/pp.py:112-161, run_redshift)
with the true spectrum, and studies the estimated spectrum separately
(notebooks/07_estimated_psd.ipynb): the estimate is the mean of N periodograms of independent noise
blocks, multiplied by N/(N-2) so that the noise mean of the likelihood ratio is unchanged.

Here a group of segments shares one estimate; its bank is built on the device with that estimate;
the segments of consecutive groups are gathered in batches of two slots (one gathering while the
next fills), and every segment's node values go through the device source-frame step (ln Lambda on
the (m1, z) grid and ln B).  The population posterior of pp.py:143-161 (credible levels of m1 and z)
is computed for the rows a caller marks.

Time-domain front end on the device (the segment convention of src/exmarg/pp.py:38-72: d = dt rfft(x) on
the bins ceil(fmin T) .. floor(fmax T), E|d|^2 = sig2 = T S / 2): segment_from_strain, psd_from_strain and
segments_from_strain (a segment set with its estimated noise variance from strain series).
"""

import numpy as np

from .. import _hip, cosmology
from ..priors import madau_dickinson_ymin, population_m1_z
from . import pe as _pe, store as _store
from .bank import mass_bank
from .batch import SegmentBatch
from .segments import D0, read_segments


# ---------------------------------------------------------------- spectra
def noise_periodograms(sig2, n_blocks, rng):
    """(n_blocks, n_freq) periodograms |d|^2 of independent complex Gaussian noise blocks with variance
    sig2 per bin: sig2 / 2 (g0^2 + g1^2) (notebooks/07_estimated_psd.ipynb cell 5)"""
    sig2 = np.asarray(sig2, float)
    g = rng.standard_normal((2, int(n_blocks), sig2.size))
    return sig2 / 2.0 * (g[0] ** 2 + g[1] ** 2)


_SCALES = {"n-2": lambda N: N / (N - 2.0), "n-1": lambda N: N / (N - 1.0), "none": lambda N: 1.0}


def _scale(scale, N):
    if scale not in _SCALES:
        raise ValueError("scale must be 'n-2', 'n-1' or 'none'")
    return _SCALES[scale](N)


def psd_estimate(periodograms, scale="n-2"):
    """the noise variance per bin from N periodograms a caller already holds: their mean times N/(N-2)
    (scale 'n-2', which keeps the noise mean of the likelihood ratio), N/(N-1) ('n-1', unbiased 1/S) or 1
    ('none') (notebooks/07_estimated_psd.ipynb cell 2: the mean is S g with g ~ Gamma(N, 1/N)).  From
    strain data, psd_from_strain forms the periodograms and their mean on the device."""
    P = np.asarray(periodograms, float)
    if P.ndim != 2 or P.shape[0] < 3:
        raise ValueError("need (N >= 3, n_freq) periodograms")
    return P.mean(axis=0) * _scale(scale, P.shape[0])


def frequency_bins(T, fmin, fmax):
    """the bins nu = ceil(fmin T) .. floor(fmax T) of a segment of T seconds (src/exmarg/pp.py:44)"""
    return np.arange(int(np.ceil(fmin * T)), int(np.floor(fmax * T)) + 1)


def _band(n, fs, fmin, fmax):
    T = n / float(fs)
    nu = frequency_bins(T, fmin, fmax)
    if nu.size == 0:
        raise ValueError("no frequency bin of a %g s block lies in [%g, %g] Hz" % (T, fmin, fmax))
    if nu[0] < 0 or nu[-1] > n // 2:
        raise ValueError("the band [%g, %g] Hz is not inside [0, fs / 2] = [0, %g] Hz" % (fmin, fmax, fs / 2.0))
    return nu, T


def segment_from_strain(x, fs, fmin, fmax):
    """frequency-domain segment of each row of the time series x (n_row, n) sampled at fs, on the device:
    d = dt rfft(x) on the bins of the band, so that white noise of one-sided density S gives
    E|d|^2 = T S / 2 (the convention of src/exmarg/pp.py:38-72).  returns d (n_row, n_freq) complex128,
    nu and T."""
    x = _hip.contiguous(np.atleast_2d(x), np.float64)
    if x.ndim != 2:
        raise ValueError("x must be one series or rows of equally long series")
    n = x.shape[1]
    nu, T = _band(n, fs, fmin, fmax)
    d = np.empty((x.shape[0], nu.size), np.complex128)
    _hip.call("exm_strain_segments", _hip.ptr(x), x.shape[0], n, float(fs), int(nu[0]), nu.size, _hip.ptr(d))
    return d, nu, T


def psd_from_strain(x, fs, T, fmin, fmax, scale="n-2"):
    """the noise variance per bin of the band estimated from the consecutive blocks of T seconds of the
    series x (1-d; a partial last block is dropped), on the device: the mean of their periodograms
    |d|^2 (d as segment_from_strain) times the factor of psd_estimate (notebooks/07_estimated_psd.ipynb
    cell 5, there with simulated periodograms)"""
    x = np.asarray(x, np.float64).ravel()
    n = int(round(T * fs))
    N = x.size // n if n > 0 else 0
    if N < 3:
        raise ValueError("need at least 3 blocks of %g s, the series holds %d" % (T, N))
    nu, _ = _band(n, fs, fmin, fmax)
    blocks = _hip.contiguous(x[:N * n], np.float64)
    out = np.empty(nu.size, np.float64)
    _hip.call("exm_strain_psd", _hip.ptr(blocks), N, n, float(fs), int(nu[0]), nu.size, float(_scale(scale, N)),
              _hip.ptr(out))
    return out


def segments_from_strain(strain, noise, fs, T, fmin, fmax, scale="n-2"):
    """a segment set (the dict read_segments takes) from strain data: strain[I] (1-d) cut into consecutive
    blocks of T seconds, one segment each, and the noise variance sig2_I estimated from the blocks of
    noise[I] (psd_from_strain), for I in 'H', 'L'"""
    out = {}
    for det in ("H", "L"):
        x = np.asarray(strain[det], np.float64).ravel()
        n = int(round(T * fs))
        if x.size < n:
            raise ValueError("the %s series is shorter than one segment" % det)
        k = x.size // n
        d, nu, T_ = segment_from_strain(x[:k * n].reshape(k, n), fs, fmin, fmax)
        out["d_" + det] = d
        out["sig2_" + det] = psd_from_strain(noise[det], fs, T, fmin, fmax, scale)
    out.update(f=nu / T_, nu=nu, T=T_, fmin=float(fmin), fmax=float(fmax))
    return out


# ---------------------------------------------------------------- population
def credible_level(grid, marginal, truth):
    """fraction of the posterior below the true value, from a marginal on a grid (src/exmarg/pp.py:75-78)"""
    cdf = (np.cumsum(marginal) - 0.5 * marginal) / marginal.sum()
    return float(np.interp(truth, grid, cdf))


def population_prior(m1, z, mdet_range):
    """the population on the (m1, z) cells inside mdet_range (src/exmarg/pp.py:143-146)"""
    (m1p, pm1), (zp, pz) = population_m1_z()
    Md = 2.0 * m1[:, None] * (1.0 + z[None, :])
    inside = (Md >= mdet_range[0]) & (Md <= mdet_range[1])
    return (np.where(inside, np.interp(m1, m1p, pm1, left=0.0, right=0.0)[:, None]
                     * np.interp(z, zp, pz, left=0.0, right=0.0)[None, :], 0.0)
            * np.gradient(m1)[:, None] * np.gradient(z)[None, :])


def draw_population(rng, mdet_range):
    """one (m1, z) of the population with M_det = 2 m1 (1 + z) on mdet_range, by inverse transform of the
    tabulated densities (src/exmarg/pp.py:125-135)"""
    (m1p, pm1), (zp, pz) = population_m1_z()

    def inv_cdf(x, p, r):
        c = np.concatenate([[0.0], np.cumsum(0.5 * (p[1:] + p[:-1]) * np.diff(x))])
        return np.interp(r, c / c[-1], x)

    while True:
        m1_, z_ = inv_cdf(m1p, pm1, rng.uniform()), inv_cdf(zp, pz, rng.uniform())
        if mdet_range[0] <= 2 * m1_ * (1 + z_) <= mdet_range[1]:
            return float(m1_), float(z_)


def posterior_levels(L, prior_mz, m1, z, m1_true, z_true):
    """credible levels of the true m1 and z under the population posterior of one segment
    (src/exmarg/pp.py:152-160)"""
    ok = np.isfinite(L) & (prior_mz > 0)
    post = np.zeros_like(L)
    if ok.any():
        post[ok] = prior_mz[ok] * np.exp(L[ok] - L[ok].max())
    if post.sum() <= 0:
        return np.nan, np.nan
    return credible_level(m1, post.sum(axis=1), m1_true), credible_level(z, post.sum(axis=0), z_true)


# ---------------------------------------------------------------- the stream
def pe_redshift_stream(groups, masses, store, zmax=10.0, nz=1600, m1_range=(2.0, 200.0), nm1=1600, ymin=None,
                       mdet_range=(10.0, 500.0), n_loud=1024, n_random=1024, seed=0, batch_segments=8,
                       keep=None):
    """pe_redshift over a stream of segment groups (src/exmarg/search/pe.py:203-257 per segment).

    groups          iterable of segment dicts (read_segments format); the sig2 of a group is the noise
                    estimate its analysis uses (its bank is built with it)
    masses          ascending detector-frame total masses of the one-column bank
    ymin            smallest y of the store; default the Madau-Dickinson value (src/exmarg/priors.py:47-49)
    batch_segments  segments per gather launch (two launch slots)
    keep            keep(group, row) -> truthy: return that segment's ln Lambda grid (nm1, nz)

    yields one dict per group, in order: ln_b (n_rows,), rows, kept {row: ln Lambda}, and the grids
    m1, z."""
    ymin = madau_dickinson_ymin() if ymin is None else float(ymin)
    masses = np.asarray(masses, float)
    if masses.ndim != 1 or np.any(np.diff(masses) <= 0):
        raise ValueError("masses must be one ascending column of detector-frame total masses")
    z = np.linspace(0.02, zmax, int(nz))
    m1 = np.geomspace(m1_range[0], m1_range[1], int(nm1))
    dl = cosmology.luminosity_distance(z)
    yq = D0 / dl
    w = _pe.redshift_prior_weights(m1, z, dl, mdet_range)
    _store.upload(store)
    slots = [SegmentBatch(store, n_loud, n_random, max_templates=batch_segments * masses.size) for _ in range(2)]
    members = {0: [], 1: []}
    state = dict(cur=0, flying=None)
    results = {}            # group -> dict being filled
    remaining = {}          # group -> segments not yet finished
    analyses = {}
    done_order = []

    grids = {}              # group -> its source-frame constants on the device

    def finish(k):
        lamNs = slots[k].collect()
        for (g, j, an), lamN in zip(members[k], lamNs):
            res = results[g]
            want = keep is not None and keep(g, int(res["rows"][j]))
            out = np.empty((m1.size, z.size)) if want else None
            if g not in grids:
                grids[g] = _pe.SourceGrid(an.geometry[2], an.bank.M, yq, m1, z, w)
            res["ln_b"][j] = grids[g].run(lamN, out=out)
            if want:
                res["kept"][int(res["rows"][j])] = out
            remaining[g] -= 1
        members[k] = []
        slots[k].reset()

    def launch():
        """launch the filling slot, then finish the one launched before it (its source-frame steps and
        the next admissions run while the new gather does)"""
        k = state["cur"]
        slots[k].launch()
        prev, state["flying"], state["cur"] = state["flying"], k, 1 - k
        if prev is not None:
            finish(prev)

    def ready():
        """groups finished in order: release their banks and yield them"""
        out = []
        while done_order and remaining[done_order[0]] == 0:
            g = done_order.pop(0)
            an = analyses.pop(g)
            grid = grids.pop(g, None)
            if grid is not None:
                grid.free()
            an.plan.handle.free()
            an.bank.handle.free()
            out.append(results.pop(g))
        return out

    try:
        for g, group in enumerate(groups):
            seg = read_segments(group)
            bank = mass_bank(masses, seg, "hip")
            an = _pe._Analysis(bank, store, ymin, True).start()
            analyses[g] = an
            n = seg["rows"].size
            results[g] = dict(ln_b=np.full(n, -np.inf), rows=seg["rows"], kept={}, m1=m1, z=z)
            remaining[g] = n
            done_order.append(g)
            for j in range(n):
                if not slots[state["cur"]].fits(bank):
                    launch()
                slots[state["cur"]].add(an, {I: seg["d_" + I][j] for I in "HL"}, {I: seg["sig2_" + I] for I in "HL"},
                                        seed + int(seg["rows"][j]))
                members[state["cur"]].append((g, j, an))
            for r in ready():
                yield r
        if members[state["cur"]]:
            launch()
        if state["flying"] is not None:
            finish(state["flying"])
            state["flying"] = None
        for r in ready():
            yield r
    finally:
        for s in slots:
            s.free()
        for grid in grids.values():
            grid.free()
