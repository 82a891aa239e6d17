"""
The table read on the GPU: for one template and one table, Eq. (C7) of the paper
(src/exmarg/search/gather.py:1-24),

  ln Lambda = ln( sum_tau sum_t exp lnS(tau; beta, Delta, dphi) ) - ln n_time,

and the arrival-time selection (select_times).  Both run in the HIP library;
this module marshals arguments and maps library errors to Python exceptions.
"""

import numpy as np

from .. import _hip
from .bank import Trajectories

NPHI = 96


class Times:
    """arrival times walked per template on the device: sel (n_template, K+S) int32 sorted, lw float32"""

    def __init__(self, handle, n_template, n_walk):
        self.handle = handle
        self.shape = (n_template, n_walk)

    def to_host(self):
        sel = np.empty(self.shape, np.int32)
        lw = np.empty(self.shape, np.float32)
        _hip.call("exm_times_read", int(self.handle), _hip.ptr(sel), _hip.ptr(lw))
        return sel, lw

    def __getitem__(self, i):
        return self.to_host()[i]

    def free(self):
        self.handle.free()


def lag_range(tau_idx):
    """(lo, hi) of contiguous integer lags (src/exmarg/search/gather.py:392-395)"""
    lags = np.asarray(tau_idx).ravel().astype(int)
    lo, hi = int(lags.min()), int(lags.max())
    if not np.array_equal(np.sort(lags), np.arange(lo, hi + 1)):
        raise ValueError("lags are not contiguous")
    return lo, hi


def random_draw(n_template, n_time, n_loud, n_random, seed):
    """the host draw of the random times: indices into the non-loud times
    (src/exmarg/search/gather.py:397, the same numpy stream as the reference)"""
    return np.random.default_rng(seed).integers(0, n_time - n_loud, size=(n_template, n_random))


def rng_state(seed):
    """the PCG64 state of numpy.random.default_rng(seed) as four uint64 (state hi, lo, inc hi, lo): the
    device makes the draw of random_draw from it (exm_select_times_seeded), bit for bit"""
    st = np.random.default_rng(seed).bit_generator.state["state"]
    s, inc = int(st["state"]), int(st["inc"])
    m = (1 << 64) - 1
    return np.array([s >> 64, s & m, inc >> 64, inc & m], dtype=np.uint64)


def select_times(traj, tau_idx, n_loud, n_random, seed, xp=None, block=256):
    """the arrival times visited for every template (src/exmarg/search/gather.py:374-420).

    The n_loud times with the largest |z_L(t)|^2 + max_tau |z_H(t + tau)|^2 are kept exactly (float64
    score, ties to the earlier time); n_random more are drawn uniformly with replacement from the rest
    (numpy's default_rng(seed) stream, drawn on the device) and carry the log weight
    ln((n_time - n_loud) / n_random).  xp and block are accepted for the reference signature.
    returns Times on the device."""
    if not isinstance(traj, Trajectories):
        traj = Trajectories.from_host(traj)
    K, S = int(n_loud), int(n_random)
    nm, nt = traj.shape
    lo, hi = lag_range(tau_idx)
    if not 0 < K < nt:
        raise ValueError("n_loud must lie in (0, n_time), got %d" % K)
    if S <= 0:
        raise ValueError("n_random must be positive, got %d" % S)
    lw = np.float32(np.log((nt - K) / S))
    state = rng_state(seed)
    h = _hip.new_handle("exm_select_times_seeded", int(traj.handle), lo, hi, K, S, _hip.ptr(state), float(lw),
                        kind="times")
    return Times(h, nm, K + S)


def _store_handle(tables):
    dev = tables.device if hasattr(tables, "device") and tables.device is not None else tables
    return dev


def gather(traj, tables, mass_ids, lna, tab_ids, block=256, times=None):
    """ln Lambda of (template, table) pairs on the GPU (src/exmarg/search/gather.py:340-371).

    traj      Trajectories (or a host dict of lqH, phH, lqL, phL)
    tables    DeviceStore (exmarg_amd.search.store.upload) or a RedshiftStore uploaded to the device
    mass_ids  (nsel,) template rows; lna, tab_ids (nsel, NS) ln a and the table read by each slot
    times     None (every arrival time) or Times from select_times
    block     accepted for the reference signature

    returns (NS, nsel) float64 on the host."""
    if not isinstance(traj, Trajectories):
        traj = Trajectories.from_host(traj)
    dev = _store_handle(tables)
    ids = _hip.contiguous(mass_ids, np.int64).ravel()
    lna = _hip.contiguous(np.atleast_2d(np.asarray(lna, float).reshape(ids.size, -1)), np.float64)
    tab = _hip.contiguous(np.asarray(tab_ids).reshape(ids.size, -1), np.int64)
    if lna.shape != tab.shape:
        raise ValueError("lna %s and tab_ids %s differ in shape" % (lna.shape, tab.shape))
    ns = tab.shape[1]
    out = np.empty((ns, ids.size), np.float64)
    _hip.call("exm_gather", int(traj.handle), int(dev.handle), _hip.ptr(ids), ids.size, _hip.ptr(lna), _hip.ptr(tab),
              ns, int(times.handle) if times is not None else 0, _hip.ptr(out))
    return out


class Plan:
    """a fixed list of (template, table, ln a) reads gathered in one launch per segment"""

    def __init__(self, dev, n_template, job_m, job_tab, job_lna):
        self.job_m = _hip.contiguous(job_m, np.int32)
        self.job_tab = _hip.contiguous(job_tab, np.int32)
        self.job_lna = _hip.contiguous(job_lna, np.float64)
        self.n_template = int(n_template)
        self.handle = _hip.new_handle("exm_plan_create", int(dev.handle), self.n_template, self.job_m.size,
                                      _hip.ptr(self.job_m), _hip.ptr(self.job_tab), _hip.ptr(self.job_lna),
                                      kind="plan")

    weights_key = None

    def run(self, traj, dev, times=None):
        out = np.empty(self.job_m.size, np.float64)
        _hip.call("exm_plan_gather", int(self.handle), int(traj.handle), int(dev.handle),
                  int(times.handle) if times is not None else 0, _hip.ptr(out))
        return out

    def set_weights(self, lnw, cellw):
        """ln w of every read (the distance weight of its point) and the template average weights, for
        run_lse and for batches admitting this plan"""
        lnw = _hip.contiguous(lnw, np.float64)
        cellw = _hip.contiguous(cellw, np.float64)
        if lnw.shape != self.job_m.shape or cellw.shape != (self.n_template,):
            raise ValueError("weights of shape %s / %s for %d reads and %d templates"
                             % (lnw.shape, cellw.shape, self.job_m.size, self.n_template))
        _hip.call("exm_plan_set_weights", int(self.handle), _hip.ptr(lnw), _hip.ptr(cellw))

    def run_lse(self, traj, dev, times=None):
        """(ln Lambda (n_template), ln B) of one segment, assembled on the device"""
        out = np.empty(self.n_template, np.float64)
        ln_b = np.empty(1, np.float64)
        _hip.call("exm_plan_gather_lse", int(self.handle), int(traj.handle), int(dev.handle),
                  int(times.handle) if times is not None else 0, _hip.ptr(out), _hip.ptr(ln_b))
        return out, float(ln_b[0])
