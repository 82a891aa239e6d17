"""
Batched analysis: many segments, possibly of different lengths and banks, read in one gather launch.

The reference analyses segments chunk by chunk with one gather launch and one device-to-host copy per
node group and segment (src/exmarg/search/pe.py:178-194).  Here every admitted segment is filtered and
its arrival times selected and sorted on the device; its walk records (8 bytes per template, lag and
walked time) stay in a batch buffer sized by the free HBM, and exm_batch_run gathers every
(segment, template, node) of the batch at once.

    res = pe_intrinsic_batch([(segments_16s, bank_16s), (segments_64s, bank_64s)], store)
    res[0]["ln_b"], res[1]["ln_lambda"]
"""

import numpy as np

from .. import _hip
from . import pe as _pe
from .gather import lag_range, rng_state
from .segments import read_segments

RECORD_BYTES = 12         # bytes per template, lag and walked time: the walk record and ln|z_H|


def template_bytes(ntau, n_walk, reads=54, nb=200):
    """device bytes one template of one segment needs in a batch: its walk records, and the launch's
    work for its reads (at most one per node plus a boundary point): (max, sum) per lag, row bounds
    and the read list"""
    return ntau * n_walk * RECORD_BYTES + n_walk * 4 + 12 + reads * (16 * ntau + 2 * (nb + 1) + 32)


class SegmentBatch:
    """walk records of up to max_templates templates over all admitted segments"""

    def __init__(self, store, n_loud=1024, n_random=1024, max_templates=None, reserve_bytes=8 << 30,
                 max_bytes=None):
        self.store = store
        self.dev = store.device
        if self.dev is None:
            raise ValueError("upload the store first (exmarg_amd.search.store.upload)")
        self.K, self.S = int(n_loud), int(n_random)
        if self.K <= 0 or self.S <= 0:
            raise ValueError("batched analysis walks a selection: n_loud and n_random must be positive")
        self.nw = self.K + self.S
        self.ntau = int(np.asarray(store.tau_idx).size)
        self.lag_lo, self.lag_hi = lag_range(store.tau_idx)
        per = template_bytes(self.ntau, self.nw, np.asarray(store.s).size + 1, store.table_shape()[2])
        if max_templates is None:
            free, _ = _hip.mem_info()
            budget = free - reserve_bytes
            if max_bytes is not None:
                budget = min(budget, max_bytes)
            max_templates = int(budget // per)
        if max_templates <= 0:
            raise MemoryError("no device memory for a batch (%d bytes per template)" % per)
        self.cap = int(max_templates)
        self.handle = _hip.new_handle("exm_batch_create", int(self.dev.handle), self.cap, self.nw, kind="batch")
        self.used = 0
        self.entries = []          # (analysis, reads) per admitted segment

    def fits(self, bank):
        return self.used + bank.u.size <= self.cap

    def add(self, analysis, d, sig2, seed):
        """admit one segment (d, sig2: dicts 'H', 'L' of (n_freq,) arrays) analysed with analysis.bank"""
        bank = analysis.bank
        nm, nt = bank.u.size, bank.n_time
        if not 0 < self.K < nt:
            raise ValueError("n_loud must lie in (0, n_time), got %d" % self.K)
        state = rng_state(seed)
        lw = np.float32(np.log((nt - self.K) / self.S))
        arrs = [_hip.contiguous(d["H"], np.complex128), _hip.contiguous(d["L"], np.complex128),
                _hip.contiguous(sig2["H"], np.float64), _hip.contiguous(sig2["L"], np.float64)]
        slot = _hip.ctypes.c_int()
        _hip.call("exm_batch_add_seeded", int(self.handle), int(bank.handle), int(analysis.plan.handle),
                  *[_hip.ptr(a) for a in arrs], self.lag_lo, self.lag_hi, self.K, self.S, _hip.ptr(state), float(lw),
                  _hip.ctypes.byref(slot))
        self.used += nm
        self.entries.append(analysis)
        return slot.value

    def add_many(self, analysis, dH, dL, sig2, seeds):
        """admit len(seeds) segments of one bank at once: dH, dL (n, n_freq) data, sig2 (dict 'H', 'L') the
        noise variance they share, seeds one per segment; the same as add() of each in turn, with the
        selection and the walk launched once for all of them.  Returns the first segment's slot."""
        bank = analysis.bank
        nm, nt = bank.u.size, bank.n_time
        n = len(seeds)
        if n == 0:
            return None
        if not 0 < self.K < nt:
            raise ValueError("n_loud must lie in (0, n_time), got %d" % self.K)
        states = np.stack([rng_state(s) for s in seeds])
        lw = np.float32(np.log((nt - self.K) / self.S))
        arrs = [_hip.contiguous(dH, np.complex128), _hip.contiguous(dL, np.complex128),
                _hip.contiguous(sig2["H"], np.float64), _hip.contiguous(sig2["L"], np.float64)]
        if arrs[0].shape != (n, bank.nu.size) or arrs[1].shape != arrs[0].shape:
            raise ValueError("dH and dL must be (%d, %d) arrays" % (n, bank.nu.size))
        slot = _hip.ctypes.c_int()
        _hip.call("exm_batch_add_seeded_many", int(self.handle), int(bank.handle), int(analysis.plan.handle), n,
                  *[_hip.ptr(a) for a in arrs], self.lag_lo, self.lag_hi, self.K, self.S,
                  _hip.ptr(_hip.contiguous(states, np.uint64)), float(lw), _hip.ctypes.byref(slot))
        self.used += n * nm
        self.entries.extend([analysis] * n)
        return slot.value

    def run(self):
        """lamN (n_node + 1, n_template) of every admitted segment, in admission order"""
        self.launch()
        return self.collect()

    def launch(self):
        """enqueue the gather of every admitted segment on the batch's own stream and return"""
        _hip.call("exm_batch_launch", int(self.handle))

    def done(self):
        """whether the launched gather has finished"""
        flag = _hip.ctypes.c_int()
        _hip.call("exm_batch_poll", int(self.handle), _hip.ctypes.byref(flag))
        return bool(flag.value)

    def collect(self):
        """wait for the launched gather; lamN of every admitted segment, in admission order"""
        nreads = sum(a.plan.job_m.size for a in self.entries)
        out = np.empty(max(nreads, 1), np.float64)
        _hip.call("exm_batch_collect", int(self.handle), _hip.ptr(out))
        res, o = [], 0
        for a in self.entries:
            n = a.plan.job_m.size
            nn = a.geometry[0].shape[0]
            lamN = np.full((nn + 1, a.bank.u.size), -np.inf)
            lamN[a.row, a.plan.job_m] = out[o:o + n]
            res.append(lamN)
            o += n
        return res

    def collect_lse(self):
        """wait for the launched gather; (ln Lambda (n_template), ln B) of every admitted segment, in admission
        order, assembled on the device (every segment's plan carries node weights: _Analysis.use_weights)"""
        L = np.empty(max(self.used, 1), np.float64)
        ln_b = np.empty(max(len(self.entries), 1), np.float64)
        _hip.call("exm_batch_collect_lse", int(self.handle), _hip.ptr(L), _hip.ptr(ln_b))
        res, o = [], 0
        for k, a in enumerate(self.entries):
            n = a.bank.u.size
            res.append((L[o:o + n].copy(), float(ln_b[k])))
            o += n
        return res

    def reset(self):
        _hip.call("exm_batch_reset", int(self.handle))
        self.used = 0
        self.entries = []

    def free(self):
        self.handle.free()


class _Slots:
    """the two batch slots of pe_intrinsic_batch (kept with the store between calls on request)"""

    def __init__(self, slots):
        self.slots = slots
        self.dev = slots[0].dev

    def reset(self):
        for b in self.slots:
            b.reset()

    def free(self):
        for b in self.slots:
            b.free()


def release_kept(store):
    """free the batch buffers pe_intrinsic_batch(..., keep_buffer=True) keeps with the store"""
    kept = store.__dict__.pop("_batch", None)
    if kept is not None:
        kept[1].free()


def pe_intrinsic_batch(items, store, prior="volume", ymin=0.1267, n_loud=1024, n_random=1024, seed=0,
                       max_templates=None, max_bytes=None, keep_buffer=False):
    """pe_intrinsic over many segment sets in batches that share one gather launch per batch.

    items  list of (segments, bank): anything read_segments takes and a mass_bank built on those
           segments (device 'hip'); banks may differ in length, band and templates
    other arguments as pe_intrinsic (src/exmarg/search/pe.py:289-346).

    returns a list with one pe_intrinsic result dict per item.

    The device memory (sized from the free HBM unless max_templates / max_bytes bound it) is split into
    two batch slots: while the gather of one slot runs on its own stream, the next segments are filtered,
    their times selected and walked into the other, so that producing walk records and consuming them
    overlap on the device.

    keep_buffer  keep the slots with the store for the next call with the same arguments, for callers that
           repeat batches: allocating them anew after freeing them costs seconds.  They hold that memory
           until release_kept(store) or exmarg_amd.search.store.free(store)."""
    from . import store as _store
    _store.upload(store)
    prepared = []
    for segments, bank in items:
        seg = read_segments(segments)
        an = _pe._Analysis(bank, store, ymin, True)
        _pe.bank_weights(an, prior)
        prepared.append((seg, bank, an.start().use_weights(prior)))
    # work memory for admitting the largest segment first, so that the slots are sized on what is left
    big = max(prepared, key=lambda p: p[1].u.size * p[1].n_time)[1]
    _hip.call("exm_reserve", big.u.size, big.n_time, big.nu.size, int(n_loud) + int(n_random))
    key = (int(n_loud), int(n_random), max_templates, max_bytes)
    kept = store.__dict__.get("_batch")
    if keep_buffer and kept is not None and kept[0] == key and kept[1].dev is store.device:
        slots = kept[1]
        slots.reset()
    else:
        if keep_buffer:
            release_kept(store)
        half_t = None if max_templates is None else max(1, int(max_templates) // 2)
        if max_bytes is None:
            free, _ = _hip.mem_info()
            half_b = (free - (8 << 30)) // 2
        else:
            half_b = int(max_bytes) // 2
        first = SegmentBatch(store, n_loud, n_random, max_templates=half_t, max_bytes=half_b)
        try:
            slots = _Slots([first, SegmentBatch(store, n_loud, n_random, max_templates=first.cap)])
        except BaseException:
            first.free()
            raise
        if keep_buffer:
            store.__dict__["_batch"] = (key, slots)
    outs = [np.full((seg["rows"].size, bank.u.size), -np.inf) for seg, bank, _ in prepared]
    lnbs = [np.full(seg["rows"].size, -np.inf) for seg, _, _ in prepared]
    members = {0: [], 1: []}
    state = dict(cur=0, flying=None)

    def finish(k):
        for (i, j), (L, lnb) in zip(members[k], slots.slots[k].collect_lse()):
            outs[i][j] = L
            lnbs[i][j] = lnb
        members[k] = []
        slots.slots[k].reset()

    def launch():
        """launch the filling slot, then finish the one launched before it (it is filled next)"""
        k = state["cur"]
        slots.slots[k].launch()
        prev, state["flying"], state["cur"] = state["flying"], k, 1 - k
        if prev is not None:
            finish(prev)

    try:
        for i, (seg, bank, an) in enumerate(prepared):
            if bank.u.size > slots.slots[0].cap:
                raise MemoryError("a bank of %d templates exceeds the batch slot capacity %d"
                                  % (bank.u.size, slots.slots[0].cap))
            for j in range(seg["rows"].size):
                if not slots.slots[state["cur"]].fits(bank):
                    launch()
                slots.slots[state["cur"]].add(an, {I: seg["d_" + I][j] for I in "HL"},
                                              {I: seg["sig2_" + I] for I in "HL"}, seed + int(seg["rows"][j]))
                members[state["cur"]].append((i, j))
        if members[state["cur"]]:
            launch()
        if state["flying"] is not None:
            finish(state["flying"])
            state["flying"] = None
    finally:
        if state["flying"] is not None:          # an error left a launch in flight: wait for it
            try:
                slots.slots[state["flying"]].collect()
            except Exception:                    # noqa: BLE001 - the original error is the one raised
                pass
        if keep_buffer:
            slots.reset()
        else:
            slots.free()
    return [dict(M=bank.M, q=bank.q, chi=bank.chi, u=bank.u, ln_lambda=out, ln_b=lnb, rows=seg["rows"])
            for (seg, bank, _), out, lnb in zip(prepared, outs, lnbs)]
