"""
Parameter estimation of single segments with the stored tables on the GPU
(src/exmarg/search/pe.py:1-346).

pe_intrinsic  (M_det, q, chi) on a template bank, the distance integrated inside the tables
pe_redshift   (m1_source, z) at equal mass and zero spin, the distance a parameter through y = D0 / D_L(z)

Per segment the device runs the matched filter, the arrival-time selection, one gather of every
(template, node) pair plus the boundary points, and the float64 assembly: the distance-prior weights
(once per bank), the log-sum-exp over each template's points and the average over the bank.
"""

import functools
import itertools

import numpy as np

from .. import _hip, cosmology
from ..priors import distance_density
from . import gather as _gather, store as _store
from .bank import check_device
from .segments import D0, read_segments


_store_tokens = itertools.count()


def _store_key(store):
    """a key that identifies the store object for the bank's caches (never reused, unlike id())"""
    return store.__dict__.setdefault("_cache_token", next(_store_tokens))


def _dl_of_z(z):
    return cosmology.luminosity_distance(z)


def _node_geometry(bank, st, ymin):
    """which (node, template) pairs the store covers, their ln a and y, and the boundary points
    (src/exmarg/search/pe.py:58-84): a = s/u is valid for sqrt(a) >= ymin and
    min(alpha_lo, ymin^2) < a < alpha_hi; a template whose a crosses that lower limit between two
    nodes also gets a point at a = alpha_lo' (1 + 1e-9) read from the upper node's table (row nn)."""
    s = np.asarray(st.s, float)
    u = np.asarray(bank.u, float)
    nn, nm = s.size, u.size
    alo = np.minimum(np.asarray(st.alpha_lo, float), ymin * ymin)
    with np.errstate(divide="ignore", invalid="ignore"):   # u = 0 (no power in the band): never valid
        a = s[:, None] / u[None, :]
        s_prev_a = np.concatenate([[0.0], s[:-1]])[:, None] / u[None, :]
    valid = (np.sqrt(a) >= ymin) & (a > alo[:, None]) & (a < np.asarray(st.alpha_hi, float)[:, None])
    lna = np.where(valid, np.log(np.where(valid, a, 1.0)), 0.0)
    y = np.zeros((nn + 1, nm))
    y[:nn] = np.where(valid, np.sqrt(a), 0.0)
    low = valid & (s_prev_a <= alo[:, None])
    bnode = np.where(low, np.arange(nn)[:, None], -1).max(axis=0)
    blna = np.zeros(nm)
    has = bnode >= 0
    a_b = alo[bnode[has]] * (1.0 + 1e-9)
    blna[has] = np.log(a_b)
    y[nn, has] = np.sqrt(a_b)
    return valid, lna, y, bnode, blna


_PRIOR_KIND = {"volume": 0, "madau-dickinson": 1}


def _device_weights(yN, prior, ymin):
    """lnW (n_rows, nm) of exm_distance_weights for the points yN (y <= 0: not reached)"""
    if prior not in _PRIOR_KIND:
        raise ValueError("prior must be 'volume' or 'madau-dickinson', got %r" % (prior,))
    density = distance_density(prior)
    yN = _hip.contiguous(yN, np.float64)
    lnW = np.empty(yN.shape, np.float64)
    if density is None:
        _hip.call("exm_distance_weights", _hip.ptr(yN), yN.shape[0], yN.shape[1], 0, float(ymin), None, None, 0,
                  _hip.ptr(lnW))
    else:
        lnx, px = (_hip.contiguous(a, np.float64) for a in density)
        _hip.call("exm_distance_weights", _hip.ptr(yN), yN.shape[0], yN.shape[1], 1, float(ymin), _hip.ptr(lnx),
                  _hip.ptr(px), lnx.size, _hip.ptr(lnW))
    return lnW


def distance_weights(y, prior, ymin):
    """ln of the weights that integrate the distance prior against a likelihood ratio linear in x = ln y
    between the points y (ascending): sum_j w_j Lambda(y_j) = integral dy p(y) Lambda(y) over
    [y_0, y_last] (src/exmarg/search/pe.py:260-286), computed on the device in float64.

    prior  'volume': p(y) = 3 ymin^3 y^-4, integrated exactly on every interval;
           'madau-dickinson': the density in ln y of exmarg_amd.priors, trapezoid rule on 401 points
           inside every interval"""
    y = np.asarray(y, float)
    if prior not in _PRIOR_KIND:
        raise ValueError("prior must be 'volume' or 'madau-dickinson', got %r" % (prior,))
    if y.ndim != 1 or np.any(y <= 0) or np.any(np.diff(y) <= 0):
        raise ValueError("y must hold positive ascending points")
    if y.size < 2:                         # no interval: every weight zero
        return np.full(y.size, -np.inf)
    return _device_weights(y[:, None], prior, ymin)[:, 0]


def _bank_weights(yN, prior, ymin, bank):
    """lnW (nn+1, nm): distance weights of every template over the points it reaches
    (src/exmarg/search/pe.py:321-327), all templates at once on the device"""
    count = (yN > 0).sum(axis=0)
    bad = np.nonzero(count < 2)[0]
    if bad.size:
        m = int(bad[0])
        raise ValueError("template %d (M %.2f, u %.3f) reaches fewer than two store nodes" % (m, bank.M[m], bank.u[m]))
    return _device_weights(yN, prior, ymin)


def bank_weights(an, prior):
    """lnW of the analysis' bank over its points for a prior, computed once per (store, ymin, prior) and
    kept with the bank (the distance weights depend on the bank and the store, not on the segment)"""
    cache = an.bank.__dict__.setdefault("_lnW", {})
    key = (_store_key(an.store), an.ymin, prior)
    if key not in cache:
        cache[key] = _bank_weights(an.geometry[2], prior, an.ymin, an.bank)
    return cache[key]


def _cell_weights(bank):
    """product of the spacings of the distinct values of every varying axis (src/exmarg/search/pe.py:338-343)"""
    cell = np.ones(bank.u.size)
    for ax in (bank.M, bank.q, bank.chi):
        vals = np.unique(ax)
        if vals.size > 1:
            cell *= np.gradient(vals)[np.searchsorted(vals, ax)]
    return cell / cell.sum()


def _jobs(geometry):
    """the flat read list: every valid (node, template) pair, then every boundary point
    (src/exmarg/search/pe.py:178-194); returns (job_m, job_tab, job_lna, row) with row the lamN row"""
    valid, lna, _, bnode, blna = geometry
    nn = valid.shape[0]
    k, m = np.nonzero(valid)
    bm = np.nonzero(bnode >= 0)[0]
    job_m = np.concatenate([m, bm]).astype(np.int32)
    job_tab = np.concatenate([k, bnode[bm]]).astype(np.int32)
    job_lna = np.concatenate([lna[k, m], blna[bm]])
    row = np.concatenate([k, np.full(bm.size, nn)])
    return job_m, job_tab, job_lna, row


class _Analysis:
    """device state of one (bank, store, ymin): the uploaded tables and the read plan"""

    def __init__(self, bank, store, ymin, resident):
        check_device(bank.device)
        self.bank, self.store, self.resident, self.ymin = bank, store, resident, float(ymin)
        # per (store, ymin): the geometry of the bank's templates over the nodes, kept with the bank
        cache = bank.__dict__.setdefault("_geometry", {})
        key = (_store_key(store), self.ymin)
        if key not in cache:
            cache[key] = _node_geometry(bank, store, ymin)
        self.geometry = cache[key]
        self.was_resident = store.device is not None
        self.dev = None

    def start(self):
        """upload the tables and build the read plan (after the host-side checks have passed)"""
        bank, store = self.bank, self.store
        self.dev = _store.upload(store)
        key = (_store_key(store), self.ymin)
        cache = bank.__dict__.setdefault("_plans", {})
        if key not in cache:
            job_m, job_tab, job_lna, row = _jobs(self.geometry)
            if job_m.size == 0:
                raise ValueError("no template of the bank reaches a node of the store")
            cache[key] = (_gather.Plan(self.dev, bank.u.size, job_m, job_tab, job_lna), row)
        self.plan, self.row = cache[key]
        return self

    def use_weights(self, prior):
        """give the plan's reads the distance weights of prior and the bank's cell weights, for the
        device assembly of ln Lambda and ln B"""
        lnW = bank_weights(self, prior)
        if self.plan.weights_key != prior:
            self.plan.set_weights(lnW[self.row, self.plan.job_m], _cell_weights(self.bank))
            self.plan.weights_key = prior
        return self

    def _filtered(self, seg, j, n_loud, n_random, seed):
        """trajectories and arrival times of segment j of seg on the device"""
        traj = self.bank.trajectories({I: seg["d_" + I][j] for I in "HL"}, {I: seg["sig2_" + I] for I in "HL"})
        times = None
        if n_loud > 0:
            try:
                times = _gather.select_times(traj, self.store.tau_idx, n_loud, n_random,
                                             seed + int(seg["rows"][j]))
            except BaseException:
                traj.free()
                raise
        return traj, times

    def node_values(self, seg, j, n_loud, n_random, seed):
        """lamN (nn+1, nm) of segment j of seg (src/exmarg/search/pe.py:168-200)"""
        traj, times = self._filtered(seg, j, n_loud, n_random, seed)
        try:
            vals = self.plan.run(traj, self.dev, times)
        finally:
            if times is not None:
                times.free()
            traj.free()
        nn = self.geometry[0].shape[0]
        lamN = np.full((nn + 1, self.bank.u.size), -np.inf)
        lamN[self.row, self.plan.job_m] = vals
        return lamN

    def ln_lambda(self, seg, j, n_loud, n_random, seed):
        """(ln Lambda (nm), ln B) of segment j of seg, assembled on the device with the weights of
        use_weights (src/exmarg/search/pe.py:320-346)"""
        traj, times = self._filtered(seg, j, n_loud, n_random, seed)
        try:
            return self.plan.run_lse(traj, self.dev, times)
        finally:
            if times is not None:
                times.free()
            traj.free()

    def close(self):
        if self.dev is not None and not self.resident and not self.was_resident:
            _store.free(self.store)


# Reads per table that one gather launch should reach before segments share it: below it a launch's
# workgroups have few lookups per staged table row, and staging the rows, not the lookups, sets its time.
READS_PER_TABLE = 4096
# device bytes the batch of one shared launch may hold (walk records and launch work: 2.2 MB per
# template-segment at 87 lags x 2048 walked times, so 8,000 template-segments)
SHARED_BYTES = 16 << 30


def _segments_per_launch(an, n_seg, chunk, n_loud, n_random):
    """how many segments share one gather launch: 1 unless the bank's reads per table are few, then up to
    chunk (the reference's segments per launch, src/exmarg/search/pe.py:173-194) and as many as bring the
    reads per table to READS_PER_TABLE, within SHARED_BYTES of batch memory"""
    if n_loud <= 0 or n_seg < 2 or int(chunk) < 2:
        return 1
    from .batch import template_bytes
    n_tab = max(int(np.asarray(an.store.s).size), 1)
    per_tab = max(an.plan.job_m.size / n_tab, 1.0)
    want = int(np.ceil(READS_PER_TABLE / per_tab))
    per = template_bytes(int(np.asarray(an.store.tau_idx).size), int(n_loud) + int(n_random), n_tab + 1,
                         an.store.table_shape()[2]) * max(an.bank.u.size, 1)
    return max(1, min(int(chunk), n_seg, want, SHARED_BYTES // per))


def _segment_groups(an, seg, chunk, n_loud, n_random, seed, lse):
    """(js, results) for consecutive groups of the segments of seg, in order: per segment j of the range js,
    (ln Lambda, ln B) assembled on the device when lse, else the node values lamN (n_node + 1, n_template).
    Segments whose bank has few reads per table share gather launches (a SegmentBatch of
    _segments_per_launch segments, one group), so that the table rows one launch stages serve the reads of
    all of them; otherwise, or when the device memory for a shared launch is not free, each segment has a
    launch (and a group) of its own.  (Two slots, one gathering while the other's results are used,
    measured no faster: the gather holds every CU, so the source-frame kernels wait for it.)"""
    n = seg["rows"].size
    g = _segments_per_launch(an, n, chunk, n_loud, n_random)
    batch = None
    if g > 1:
        from .batch import SegmentBatch
        try:
            batch = SegmentBatch(an.store, n_loud, n_random, max_templates=g * an.bank.u.size)
        except MemoryError:
            batch = None
    if batch is None:
        for j in range(n):
            yield range(j, j + 1), [an.ln_lambda(seg, j, n_loud, n_random, seed) if lse
                                    else an.node_values(seg, j, n_loud, n_random, seed)]
        return
    try:
        sig2 = {I: seg["sig2_" + I] for I in "HL"}
        for j0 in range(0, n, g):
            js = range(j0, min(n, j0 + g))
            batch.add_many(an, seg["d_H"][js.start:js.stop], seg["d_L"][js.start:js.stop], sig2,
                           [seed + int(seg["rows"][j]) for j in js])
            batch.launch()
            res = batch.collect_lse() if lse else batch.collect()
            batch.reset()
            yield js, res
    finally:
        batch.free()


def pe_intrinsic(segments, bank, store, rows=None, prior="volume", ymin=0.1267, chunk=4, nodes_per_launch=6,
                 block=256, resident=False, n_loud=1024, n_random=1024, seed=0):
    """ln Lambda of each template of the bank, the distance integrated over its prior
    (src/exmarg/search/pe.py:289-346).

    segments          anything read_segments takes
    bank              mass_bank built on the same segments, rows (M_det, q, chi), device 'hip'
    store             load_redshift_store (or store.synthetic_store)
    rows              segment indices (default all)
    prior             'volume' or 'madau-dickinson'
    ymin              smallest y = D0 / D_L of the prior
    resident          keep the tables on the device after the call
    n_loud, n_random  arrival times visited per template (select_times); n_loud = 0 visits all of them
    seed              seed of the random draw (plus the segment index)
    chunk             segments per gather launch at most (see _segment_groups); nodes_per_launch and
                      block are accepted for the reference signature: every node of a segment is read in
                      one launch.

    returns dict: M, q, chi, u (n_template,), ln_lambda (n_segment, n_template), ln_b (n_segment,), rows."""
    seg = read_segments(segments, rows)
    an = _Analysis(bank, store, ymin, resident)
    bank_weights(an, prior)                 # the host-side checks before the tables go up
    an.start().use_weights(prior)
    try:
        out = np.full((seg["rows"].size, bank.u.size), -np.inf)
        ln_b = np.full(seg["rows"].size, -np.inf)
        for js, res in _segment_groups(an, seg, chunk, n_loud, n_random, seed, lse=True):
            for j, (L, lnb) in zip(js, res):
                out[j], ln_b[j] = L, lnb
    finally:
        an.close()
    return dict(M=bank.M, q=bank.q, chi=bank.chi, u=bank.u, ln_lambda=out, ln_b=ln_b, rows=seg["rows"])


def pe_redshift(segments, bank, store, rows=None, chunk=8, zmax=8.0, nz=1600, m1_range=(3.0, 150.0), nm1=1600,
                ymin=0.1267, nodes_per_launch=6, block=256, mdet_range=(10.0, 500.0), resident=False, n_loud=1024,
                n_random=1024, seed=0):
    """ln Lambda(m1_source, z) of each segment (src/exmarg/search/pe.py:203-257).

    bank is a mass_bank of a one-column bank of detector-frame total masses (ascending, equal mass, zero
    spin) on the same segments.  ln B averages Lambda over the prior uniform in M_det = 2 m1 (1+z) on
    mdet_range and proportional to D_L^2 in luminosity distance, p(m1, z) ~ (1+z) D_L^2 dD_L/dz.

    chunk is the number of segments one gather launch reads at most (see _segment_groups).

    returns dict: m1 (nm1,), z (nz,), ln_lambda (n_segment, nm1, nz), ln_b (n_segment,), rows."""
    if np.any(np.diff(bank.M) <= 0) or np.any(bank.q != 1.0) or np.any(bank.chi != 0.0):
        raise ValueError("pe_redshift needs an ascending equal-mass zero-spin bank")
    seg = read_segments(segments, rows)
    z, m1, yq, w = _redshift_grid(float(zmax), int(nz), float(m1_range[0]), float(m1_range[1]), int(nm1),
                                  float(mdet_range[0]), float(mdet_range[1]))
    an = _Analysis(bank, store, ymin, resident).start()
    # every cell of every segment's grid is written by its source-frame step
    out = np.empty((seg["rows"].size, m1.size, z.size))
    ln_b = np.empty(seg["rows"].size)
    try:
        grid = SourceGrid(an.geometry[2], bank.M, yq, m1, z, w)
        try:
            for js, lamNs in _segment_groups(an, seg, chunk, n_loud, n_random, seed, lse=False):
                ln_b[js.start:js.stop] = grid.run_many(np.stack(lamNs), out=out[js.start:js.stop])
        finally:
            grid.free()
    finally:
        an.close()
    return dict(m1=m1.copy(), z=z.copy(), ln_lambda=out, ln_b=ln_b, rows=seg["rows"])


@functools.lru_cache(maxsize=4)
def _redshift_grid(zmax, nz, m1_lo, m1_hi, nm1, mdet_lo, mdet_hi):
    """(z, m1, yq = D0 / D_L(z), prior weights w) of a pe_redshift grid (src/exmarg/search/pe.py:236-255),
    kept for the next call with the same grid (read-only arrays; w is nm1 x nz float64)"""
    z = np.linspace(0.02, zmax, nz)
    m1 = np.geomspace(m1_lo, m1_hi, nm1)
    dl = _dl_of_z(z)
    out = (z, m1, D0 / dl, redshift_prior_weights(m1, z, dl, (mdet_lo, mdet_hi)))
    for a in out:
        a.setflags(write=False)
    return out


class SourceGrid:
    """the source-frame step of many segments with one bank geometry, grid and prior (one pe_redshift call):
    the constants stay on the device (exm_source_grid_create), each segment sends its node values only
    (exm_source_grid_run); the same arithmetic as source_frame"""

    def __init__(self, yN, Mdet, yq, m1, z, w=None):
        yN_, Md, yq_, m1_, z_ = [np.ascontiguousarray(a, np.float64) for a in (yN, Mdet, yq, m1, z)]
        if yN_.ndim != 2 or Md.shape != (yN_.shape[1],) or yq_.shape != z_.shape:
            raise ValueError("yN must be (nrow, nm) with one mass per template, yq one value per z")
        wv = None if w is None else np.ascontiguousarray(w, np.float64)
        if wv is not None and wv.shape != (m1_.size, z_.size):
            raise ValueError("w has shape %s, expected %s" % (wv.shape, (m1_.size, z_.size)))
        self.shape = (m1_.size, z_.size)
        self.nodes = yN_.shape
        self.has_w = wv is not None
        self.handle = _hip.new_handle("exm_source_grid_create", _hip.ptr(yN_), yN_.shape[0], yN_.shape[1],
                                      _hip.ptr(Md), _hip.ptr(yq_), _hip.ptr(m1_), m1_.size, _hip.ptr(z_), z_.size,
                                      _hip.ptr(wv) if wv is not None else None, kind="grid")

    def run(self, lamN, out=None):
        """ln Lambda of one segment into out (nm1, nz) when given; returns ln B when the grid has w"""
        lamN = np.ascontiguousarray(lamN, np.float64)
        if lamN.shape != self.nodes:
            raise ValueError("lamN has shape %s, expected %s" % (lamN.shape, self.nodes))
        if out is not None and (out.shape != self.shape or out.dtype != np.float64 or not out.flags.c_contiguous):
            raise ValueError("out must be a C-contiguous float64 (nm1, nz) array")
        ln_b = np.zeros(1)
        _hip.call("exm_source_grid_run", int(self.handle), _hip.ptr(lamN), _hip.ptr(out) if out is not None else None,
                  _hip.ptr(ln_b))
        return float(ln_b[0]) if self.has_w else None

    def run_many(self, lamNs, out=None):
        """run over n segments in one call: lamNs (n, nrow, nm); out (n, nm1, nz) when given; returns ln B (n,)
        when the grid has w.  Each grid's copy to the host overlaps the next segment's work on the device."""
        lamNs = np.ascontiguousarray(lamNs, np.float64)
        if lamNs.ndim != 3 or lamNs.shape[1:] != self.nodes:
            raise ValueError("lamNs has shape %s, expected (n,) + %s" % (lamNs.shape, self.nodes))
        n = lamNs.shape[0]
        if out is not None and (out.shape != (n,) + self.shape or out.dtype != np.float64
                                or not out.flags.c_contiguous):
            raise ValueError("out must be a C-contiguous float64 (%d, nm1, nz) array" % n)
        ln_b = np.zeros(max(n, 1))
        _hip.call("exm_source_grid_run_many", int(self.handle), n, _hip.ptr(lamNs),
                  _hip.ptr(out) if out is not None else None, _hip.ptr(ln_b))
        return ln_b[:n] if self.has_w else None

    def free(self):
        self.handle.free()


def redshift_prior_weights(m1, z, dl, mdet_range):
    """normalized weights of the pe_redshift prior on the (m1, z) cells: uniform in M_det = 2 m1 (1 + z)
    on mdet_range and p(z) ~ (1 + z) D_L^2 dD_L/dz (src/exmarg/search/pe.py:250-255)"""
    lnp = np.log(np.maximum((1.0 + z) * dl ** 2 * np.gradient(dl, z), 1e-300))
    Md = 2.0 * m1[:, None] * (1.0 + z[None, :])
    inside = (Md >= mdet_range[0]) & (Md <= mdet_range[1])
    w = np.where(inside, np.exp(lnp)[None, :], 0.0) * np.gradient(m1)[:, None] * np.gradient(z)[None, :]
    return w / w.sum()


def source_frame(lamN, yN, Mdet, yq, m1, z, w=None, out=None):
    """ln Lambda on the (m1, z) grid from the node values of one segment, on the device
    (exm_source_frame: src/exmarg/search/pe.py:87-142).  Writes the grid into out (nm1, nz) when given
    and returns ln B = ln sum w e^Lambda when w is given (else None)."""
    lamN = np.ascontiguousarray(lamN, np.float64)
    yN = np.ascontiguousarray(yN, np.float64)
    nrow, nm = lamN.shape
    if yN.shape != (nrow, nm):
        raise ValueError("yN has shape %s, expected %s" % (yN.shape, (nrow, nm)))
    arrs = [np.ascontiguousarray(a, np.float64) for a in (Mdet, yq, m1, z)]
    Md, yq_, m1_, z_ = arrs
    if Md.shape != (nm,) or yq_.shape != z_.shape:
        raise ValueError("Mdet must have one mass per template and yq one value per z")
    wv = None if w is None else np.ascontiguousarray(w, np.float64)
    if wv is not None and wv.shape != (m1_.size, z_.size):
        raise ValueError("w has shape %s, expected %s" % (wv.shape, (m1_.size, z_.size)))
    if out is not None and (out.shape != (m1_.size, z_.size) or out.dtype != np.float64 or not out.flags.c_contiguous):
        raise ValueError("out must be a C-contiguous float64 (nm1, nz) array")
    ln_b = np.zeros(1)
    _hip.call("exm_source_frame", _hip.ptr(lamN), _hip.ptr(yN), nrow, nm, _hip.ptr(Md), _hip.ptr(yq_), _hip.ptr(m1_),
              m1_.size, _hip.ptr(z_), z_.size, _hip.ptr(wv) if wv is not None else None,
              _hip.ptr(out) if out is not None else None, _hip.ptr(ln_b))
    return float(ln_b[0]) if wv is not None else None
