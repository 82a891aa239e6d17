"""
Autotuner of the band gather kernel.

The band kernel has a form and two shape knobs: the form (4: pieces listed once for every lag and
dealt to the waves in chunks; 3: each wave deals its own reads' pieces; 2: pieces dealt over the whole
workgroup), the piece length (walked times one lane reads in a row: 4, 6 or 8 in the scheduled and wave
forms, 4, 8 or 16 in the piece form) and the tile (reads per workgroup, 256, 512, 1024 or 2048; the wave
and piece forms cut tiles of 2048 in two).  The best choice depends on the walk, the table geometry and how the reads spread over the
beta rows, so it is measured, not modelled: tune_batch times every candidate on the caller's own
filled batch (the gather is rerun, nothing is consumed) with the kernel's own HIP events, keeps the
fastest and records it in a JSON cache keyed by device, table geometry, walk length and precision;
apply() restores a recorded choice.

    batch = SegmentBatch(store, ...); batch.add(...)     # a sample of the workload
    best = tune.tune_batch(batch)                        # sets the knobs
    tune.apply(store, n_walk)                            # later runs, from the cache

No reference counterpart (the reference's CUDA kernel has one fixed block size,
src/exmarg/search/gather.py:313-337).
"""

import ctypes
import json
import os

from . import _hip

FORMS = (4, 3, 2)
PIECES = (4, 6, 8, 16)
TILES = (256, 512, 1024, 2048)


def cache_path():
    return os.environ.get("EXMARG_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "exmarg_amd",
                                                               "tuning.json")


def current():
    """(form, piece, tile, precision) in force (0 for the defaults)"""
    v = [ctypes.c_int() for _ in range(4)]
    _hip.call("exm_get_gather_tuning", *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def set_knobs(piece=0, tile=0, precision=None, form=None):
    _hip.call("exm_set_gather_piece", int(piece))
    _hip.call("exm_set_gather_tile", int(tile))
    if precision is not None:
        _hip.call("exm_set_gather_precision", int(precision))
    if form is not None:
        _hip.call("exm_set_gather_kernel", int(form))


def key(store, n_walk, precision):
    name = _hip.device_info()["name"]
    shp = store.table_shape()
    return "%s|%s|walk%d|%s" % (name, "x".join(str(int(v)) for v in shp), int(n_walk), "f32" if precision else "f64")


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(table):
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune_batch(batch, repeats=2, forms=FORMS, pieces=PIECES, tiles=TILES):
    """time the gather of the filled batch for every (form, piece, tile), keep the fastest (knobs set)
    and record it; returns dict(form, piece, tile, ms, table={"form,piece,tile": ms}).  Uses (and
    resets) the library's stage timers."""
    if not batch.entries:
        raise ValueError("tune a batch that holds segments")
    precision = current()[3]
    _hip.profile(True)
    times = {}
    try:
        for f in forms:
            for p in pieces:
                if (f in (3, 4) and p == 16) or (f == 2 and p == 6):
                    continue                          # scheduled, wave form: 4, 6 or 8; piece form: 4, 8 or 16
                for t in tiles:
                    set_knobs(p, t, form=f)
                    batch.run()                      # warm (code objects, work memory)
                    ms0, n0 = _hip.profile_get("gather")
                    for _ in range(repeats):
                        batch.run()
                    ms1, n1 = _hip.profile_get("gather")
                    times[(f, p, t)] = (ms1 - ms0) / max(n1 - n0, 1)
    finally:
        _hip.profile(False)
    (f, p, t), ms = min(times.items(), key=lambda kv: kv[1])
    set_knobs(p, t, form=f)
    table = _load()
    table[key(batch.store, batch.nw, precision)] = dict(form=f, piece=p, tile=t, ms=ms)
    _save(table)
    return dict(form=f, piece=p, tile=t, ms=ms, table={"%d,%d,%d" % k: v for k, v in times.items()})


def apply(store, n_walk):
    """set the recorded knobs for this store geometry and walk (precision in force); returns them or None"""
    precision = current()[3]
    rec = _load().get(key(store, n_walk, precision))
    if rec is None:
        return None
    set_knobs(rec["piece"], rec["tile"], form=rec.get("form", 0))
    return rec
