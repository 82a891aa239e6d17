"""
Serving: a request loop with continuous batching over the HBM-resident tables.

A request is one segment and its own template bank (any length, band and size).  The server keeps
two batch slots.  While the gather of one slot runs on its own stream, arriving requests are
admitted into the other: their bank is built on the device (TaylorF2), their segment filtered, its
arrival times selected and walked, and the walk records (8 bytes per template, lag and walked time)
stay in the slot's HBM buffer.  When the running gather ends, its requests are finished on the host
(node log-sum-exp with the distance-prior weights, ln B) and the other slot is launched at once, so
admission and gather overlap on the device and a request never waits for a full batch.

Admission is by HBM budget: a slot holds at most `slot_templates` templates of walk records
(template_bytes each); a request larger than a slot is refused.  The reference has no serving
loop; it analyses one segment group at a time (src/exmarg/search/pe.py:289-346, pp.py:96-97).

Banks are cached: requests with the same bank rows, segment length, band and noise curve share one
device bank (and the node geometry and prior weights cached on it), least recently used banks beyond
`bank_cache_bytes` of template factors being freed once no admitted request uses them.

    srv = Server(store)
    r = srv.submit(segment, bank_rows)          # bank rows (M_det, q, chi)
    while not r.done: srv.step()
    r.ln_b, r.ln_lambda, r.latency
"""

import collections
import hashlib
import time

import numpy as np

from .. import _hip
from . import pe as _pe, store as _store
from .bank import mass_bank
from .batch import SegmentBatch, template_bytes
from .segments import read_segments


class Request:
    """one segment (a single row) analysed with its own bank"""

    def __init__(self, rid, seg, rows, seed, t_submit, factors=None):
        self.id, self.seg, self.rows, self.seed = rid, seg, rows, int(seed)
        self.factors = factors
        self.t_submit = t_submit
        self.t_admit = self.t_done = None
        self.ln_lambda = self.ln_b = None
        self.bank = self.an = None
        self.error = None

    @property
    def done(self):
        return self.t_done is not None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit


class Server:
    """continuous-batching request loop (two batch slots on their own streams)

    store           RedshiftStore (uploaded on construction)
    n_loud/n_random arrival times walked per template (select_times)
    slot_templates  templates of walk records per slot (default: the HBM left after reserve_bytes,
                    split over the two slots, at most SLOT_TEMPLATES)
    max_request     (n_tmpl, n_time, n_freq) of the largest request expected: the work memory for
                    admitting it is reserved up front so that no admission allocates while a gather
                    runs
    prior, ymin     distance prior of the node weights (pe_intrinsic)
    bank_cache_bytes  device bytes of template factors kept for banks no request uses (0: a bank is
                    freed as soon as its last request finishes)"""

    SLOT_TEMPLATES = 32768          # default cap of a slot: larger batches only lengthen the wait of others

    def __init__(self, store, n_loud=1024, n_random=1024, slot_templates=None, reserve_bytes=16 << 30,
                 max_request=None, prior="volume", ymin=0.1267, clock=time.perf_counter, bank_cache_bytes=8 << 30):
        self.store, self.prior, self.ymin, self.clock = store, prior, float(ymin), clock
        _store.upload(store)
        self.K, self.S = int(n_loud), int(n_random)
        if max_request is not None:
            nm, nt, nf = (int(v) for v in max_request)
            _hip.call("exm_reserve", nm, nt, nf, self.K + self.S)
        per = template_bytes(np.asarray(store.tau_idx).size, self.K + self.S, np.asarray(store.s).size + 1,
                             store.table_shape()[2])
        if slot_templates is None:
            free, _ = _hip.mem_info()
            slot_templates = int(min((free - reserve_bytes) // (2 * per), self.SLOT_TEMPLATES))
        if slot_templates <= 0:
            raise MemoryError("no device memory for a batch slot (%d bytes per template)" % per)
        self.slots = [SegmentBatch(store, self.K, self.S, max_templates=slot_templates) for _ in range(2)]
        self.filling, self.flying = 0, None           # slot indices
        self.members = {0: [], 1: []}                   # admitted requests per slot
        self.queue = collections.deque()
        self.release = []                               # device objects freed once nothing is in flight
        self.banks = collections.OrderedDict()          # bank key -> [MassBank, requests using it]
        self.bank_cache_bytes = int(bank_cache_bytes)
        self.next_id = 0
        self.stats = dict(batches=0, admitted=0, admitted_during_gather=0, finished=0, refused=0, bank_builds=0,
                          bank_hits=0)

    # ------------------------------------------------------------ client side
    def submit(self, segments, bank_rows, seed=0, t_submit=None, factors=None):
        """queue one segment (the first row of `segments`) with its bank rows (M_det, q, chi); factors
        (W, u) computed elsewhere replace the device TaylorF2 of the bank (mass_bank's factors)"""
        seg = read_segments(segments, rows=[0])
        rows = np.asarray(bank_rows, float)
        req = Request(self.next_id, seg, rows, seed, self.clock() if t_submit is None else t_submit, factors)
        self.next_id += 1
        self.queue.append(req)
        return req

    @property
    def idle(self):
        return not self.queue and self.flying is None and not self.members[self.filling]

    # ------------------------------------------------------------ server loop
    def step(self):
        """one pass of the loop: finish a completed gather, admit every queued request that fits the
        filling slot (a gather that ends meanwhile is finished at once), and launch the filling slot
        when the device has no gather and the next queued request would not join it; returns whether
        anything happened"""
        progress = False
        while True:
            moved = self._finish_done()
            if self.flying is None and self.members[self.filling] and not (self.queue and self._fits(self.queue[0])):
                self.slots[self.filling].launch()
                self.flying, self.filling = self.filling, 1 - self.filling
                self.stats["batches"] += 1
                moved = True
            if self.queue and self._admit(self.queue[0]):
                self.queue.popleft()
                moved = True
            if not moved:
                return progress
            progress = True

    def _finish_done(self):
        if self.flying is not None and self.slots[self.flying].done():
            self._finish(self.flying)
            return True
        return False

    def run(self, until=None, poll=2e-4):
        """step until every submitted request has finished (or until() is true)"""
        while not self.idle and not (until and until()):
            if not self.step():
                time.sleep(poll)

    def close(self):
        if self.flying is not None:
            self._finish(self.flying)
        for s in self.slots:
            s.free()
        for bank, _ in self.banks.values():
            self._free_bank(bank)
        self.banks.clear()
        self._release()

    # ------------------------------------------------------------ internals
    @staticmethod
    def _size(req):
        return req.rows.shape[0] if req.rows.ndim == 2 else req.rows.size

    def _fits(self, req):
        """whether the filling slot takes the request now (a request larger than any slot is taken, to
        be refused)"""
        batch = self.slots[self.filling]
        n = self._size(req)
        return n > batch.cap or batch.used + n <= batch.cap

    def _refuse(self, req, err):
        req.bank = req.an = None
        req.error = err
        req.t_done = self.clock()
        self.stats["refused"] += 1
        return True

    def _admit(self, req):
        """admit the request into the filling slot if it fits; False leaves it queued.  A request that
        cannot be analysed (too large for a slot, a segment too short for the walk, a bank row the
        waveform model rejects, a bank that reaches no node, a device error while it is filtered) is
        finished at once with req.error set, so that it never holds up the requests behind it."""
        batch = self.slots[self.filling]
        n = self._size(req)
        if n > batch.cap:
            return self._refuse(req, MemoryError("request of %d templates exceeds a batch slot of %d" % (n, batch.cap)))
        if batch.used + n > batch.cap:
            return False
        seg = req.seg
        n_time = 8 * int(np.max(seg["nu"]))
        if not 0 < self.K < n_time:
            return self._refuse(req, ValueError("n_loud %d does not lie in (0, n_time = %d)" % (self.K, n_time)))
        bank = None
        try:
            bank = self._bank(req.rows, seg, req.factors)
            an = _pe._Analysis(bank, self.store, self.ymin, True)
            _pe.bank_weights(an, self.prior)
            an.start().use_weights(self.prior)
            batch.add(an, {I: seg["d_" + I][0] for I in "HL"}, {I: seg["sig2_" + I] for I in "HL"},
                      req.seed + int(seg["rows"][0]))
        except Exception as e:          # noqa: BLE001 - the request carries its own error
            if bank is not None:
                self._unuse(bank)
            return self._refuse(req, e)
        req.bank, req.an = bank, an
        req.t_admit = self.clock()
        self.members[self.filling].append(req)
        self.stats["admitted"] += 1
        if self.flying is not None:
            self.stats["admitted_during_gather"] += 1
        return True

    def _finish(self, k):
        batch = self.slots[k]
        results = batch.collect_lse()
        t = self.clock()
        for req, (L, lnb) in zip(self.members[k], results):
            req.ln_lambda = L
            req.ln_b = lnb
            req.t_done = t
            self._unuse(req.bank)
            req.an = None
            self.stats["finished"] += 1
        self.members[k] = []
        batch.reset()
        self.flying = None
        self._release()

    @staticmethod
    def _bank_key(rows, seg, factors=None):
        """what a bank depends on: its rows, the segment length, band and the noise curve of its norm (and
        the factors when given)"""
        h = hashlib.sha1(np.ascontiguousarray(rows, np.float64).tobytes())
        h.update(np.ascontiguousarray(seg["sig2_H"], np.float64).tobytes())
        if factors is not None:
            h.update(np.ascontiguousarray(factors[0], np.complex64).tobytes())
            h.update(np.ascontiguousarray(factors[1], np.float64).tobytes())
        nu = np.asarray(seg["nu"])
        return (h.hexdigest(), float(seg["T"]), float(seg["fmin"]), float(seg["fmax"]), int(nu[0]), int(nu.size))

    def _bank(self, rows, seg, factors=None):
        """the cached bank of these rows and segment geometry, built on the device when absent"""
        key = self._bank_key(rows, seg, factors)
        ent = self.banks.get(key)
        if ent is None:
            ent = [mass_bank(rows, seg, "hip", factors=factors), 0]
            ent[0]._cache_key = key
            self.banks[key] = ent
            self.stats["bank_builds"] += 1
        else:
            self.banks.move_to_end(key)
            self.stats["bank_hits"] += 1
        ent[1] += 1
        return ent[0]

    def _unuse(self, bank):
        """one request less uses the bank; banks no request uses are freed, least recently used first,
        while the unused ones hold more than bank_cache_bytes"""
        self.banks[bank._cache_key][1] -= 1
        idle = [(k, b) for k, (b, n) in self.banks.items() if n == 0]
        held = sum(b.M.size * b.nu.size * 8 for _, b in idle)
        for k, b in idle:
            if held <= self.bank_cache_bytes:
                break
            held -= b.M.size * b.nu.size * 8
            self._free_bank(b)
            del self.banks[k]

    def _free_bank(self, bank):
        """free a bank and the read plans cached with it (pe._Analysis keeps one per store and ymin)"""
        for plan, _ in bank.__dict__.get("_plans", {}).values():
            self.release.append(plan.handle)
        self.release.append(bank.handle)

    def _release(self):
        """free banks and read plans of finished requests (a free may wait for the device, so it is
        done right after a gather has been collected, before the next launch)"""
        for h in self.release:
            h.free()
        self.release = []
