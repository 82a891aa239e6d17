"""
The (beta, Delta) tables of the redshift store (src/exmarg/search/store.py:1-127).

Host side: bd_*.npz files read into host memory (optionally pinned) or mapped
from disk, sorted by signal power s.  Device side: one HBM-resident stack of
every node, uploaded once and kept for the life of the store object (288 GB of
HBM holds the 40 GB store whole).
"""

import ctypes
import glob
import os
import re
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .. import _hip, synth
from . import npzmember

NPHI = 96                 # phase bins of the full circle; the tables hold NPHI/2 + 1 (store.py:20)


def read_npz_member(path, name):
    """one uncompressed .npy member of an npz, read sequentially (src/exmarg/search/store.py:23-36)"""
    return npzmember.locate(path, name).read()


def map_npz_member(path, name):
    """read-only memory map of one uncompressed .npy member (src/exmarg/search/store.py:39-52)"""
    return np.asarray(npzmember.locate(path, name).mapped())


@dataclass
class RedshiftStore:
    """signal power nodes s (sorted), their amplitude limits, grids, host tables and the truncation order
    (src/exmarg/search/store.py:76-87).  tables may be None for a store that lives on the device only;
    device holds the DeviceStore once uploaded."""
    s: np.ndarray
    alpha_lo: np.ndarray
    alpha_hi: np.ndarray
    beta: list
    delta: list
    tau_idx: np.ndarray
    tables: list
    n_max: int
    device: object = field(default=None, repr=False)
    shape: tuple = None
    synthetic: dict = None
    files: list = None          # (path, byte offset of the float32 T member) per node, for direct ingest

    def table_shape(self):
        if self.tables is not None:
            return tuple(self.tables[0].shape)
        return tuple(self.shape)


_FILE_NAME = re.compile(r"bd_n(\d+)_s[0-9.]+\.npz$")


@dataclass
class _NodeFile:
    """the grids and node values of one bd_*.npz file and where its table lies in the file"""
    s: float
    alpha_lo: float
    alpha_hi: float
    beta: np.ndarray
    delta: np.ndarray
    tau_idx: np.ndarray
    n_max: int
    table: npzmember.Member

    @classmethod
    def scan(cls, path):
        m = _FILE_NAME.search(os.path.basename(path))
        if m is None:
            raise ValueError("%s is not named bd_n<order>_s<s>.npz" % path)
        with np.load(path) as z:
            alpha = np.asarray(z["alpha"], np.float64)
            vals = dict(s=float(z["s_node"]), alpha_lo=float(alpha[0]), alpha_hi=float(alpha[-1]),
                        beta=np.asarray(z["beta"]), delta=np.asarray(z["delta"]), tau_idx=np.asarray(z["tau_idx"]))
        table = npzmember.locate(path, "T.npy")
        if table.fortran or table.dtype != np.float32 or len(table.shape) != 4:
            raise ValueError("%s: T must be a C-order float32 (ntau, nphif, nb, nd) array" % path)
        return cls(n_max=int(m.group(1)), table=table, **vals)


def load_redshift_store(store_dir, pinned=False, readers=6, mmap=False, host=True):
    """every node of the redshift store (src/exmarg/search/store.py:90-107).

    host=True (the reference's behaviour): the tables are read into host memory (mmap=True maps them
    from disk instead).  host=False: only the node metadata and the files' member offsets are read;
    upload() then streams the tables from the files into HBM with `readers` threads through pinned
    buffers (exm_store_load_files), never holding them in host memory.  pinned is accepted for the
    reference signature: uploads stage through the library's own pinned buffers."""
    nodes = sorted((_NodeFile.scan(p) for p in glob.glob(os.path.join(store_dir, "bd_*.npz"))),
                   key=lambda nf: nf.s)
    if not nodes:
        raise FileNotFoundError("no bd_*.npz in %s" % store_dir)
    shapes = {nf.table.shape for nf in nodes}
    if len(shapes) != 1:
        raise ValueError("the tables of a store must share their shape, found %s" % sorted(shapes))
    if not host:
        tables = None
    elif mmap:
        tables = [np.asarray(nf.table.mapped()) for nf in nodes]
    else:
        with ThreadPoolExecutor(max_workers=max(1, int(readers))) as pool:
            tables = list(pool.map(lambda nf: nf.table.read(), nodes))
    st = RedshiftStore(s=np.array([nf.s for nf in nodes]), alpha_lo=np.array([nf.alpha_lo for nf in nodes]),
                       alpha_hi=np.array([nf.alpha_hi for nf in nodes]), beta=[nf.beta for nf in nodes],
                       delta=[nf.delta for nf in nodes], tau_idx=nodes[0].tau_idx, tables=tables,
                       n_max=nodes[0].n_max, shape=nodes[0].table.shape,
                       files=[(nf.table.path, nf.table.offset) for nf in nodes])
    st.readers = int(readers)
    return st


class DeviceStore:
    """the tables of every node stacked on the device (src/exmarg/search/store.py:110-120)"""

    def __init__(self, handle, shape, n_tables, beta, delta, tau_idx):
        self.handle = handle
        self.shape = tuple(shape)
        self.n_tables = n_tables
        self.beta = np.asarray(beta, float)
        self.delta = np.asarray(delta, float)
        self.tau_idx = np.asarray(tau_idx, np.int64)

    def read_table(self, k):
        out = np.empty(self.shape, np.float32)
        _hip.call("exm_store_read_table", int(self.handle), int(k), _hip.ptr(out))
        return out

    def free(self):
        self.handle.free()

    @property
    def nbytes(self):
        return 4 * int(np.prod(self.shape)) * self.n_tables


def _check_grids(st):
    b0, d0 = np.asarray(st.beta[0]), np.asarray(st.delta[0])
    for b, d in zip(st.beta, st.delta):
        if not (np.array_equal(np.asarray(b), b0) and np.array_equal(np.asarray(d), d0)):
            raise ValueError("every node of a store must share its beta and Delta grids")


def _create(shape, n_tables, beta, delta, tau_idx):
    _hip.init()
    ntau, nphif, nb, nd = shape
    if len(tau_idx) != ntau:
        raise ValueError("tau_idx has %d lags, tables %d" % (len(tau_idx), ntau))
    beta, delta = np.asarray(beta, float), np.asarray(delta, float)
    tau = _hip.contiguous(tau_idx, np.int64)
    h = _hip.new_handle("exm_store_create", int(n_tables), int(ntau), int(nphif), int(nb), int(nd), _hip.ptr(tau),
                        float(beta[0]), float(beta[1] - beta[0]), float(delta[0]), float(delta[1] - delta[0]),
                        kind="store")
    return DeviceStore(h, shape, n_tables, beta, delta, tau)


def upload(store):
    """host tables -> one device stack, cached on store.device (src/exmarg/search/pe.py:146-154 with
    resident=True)"""
    if store.device is not None:
        return store.device
    _check_grids(store)
    in_ram = store.tables is not None and not any(isinstance(T, np.memmap) or
                                                  (getattr(T, "base", None) is not None and
                                                   isinstance(T.base, np.memmap)) for T in store.tables)
    if store.files is not None and not in_ram:
        # straight from the files (mapped or not read): reader threads, pinned buffers, async copies
        dev = _create(tuple(store.shape), len(store.files), store.beta[0], store.delta[0], store.tau_idx)
        paths = [os.fsencode(p) for p, _ in store.files]
        cpaths = (ctypes.c_char_p * len(paths))(*paths)
        offs = np.array([o for _, o in store.files], np.int64)
        idx = np.arange(len(paths), dtype=np.int32)
        secs = ctypes.c_double()
        _hip.call("exm_store_load_files", int(dev.handle), len(paths), ctypes.cast(cpaths, ctypes.c_void_p),
                  _hip.ptr(offs), _hip.ptr(idx),
                  int(getattr(store, "readers", 6)), ctypes.byref(secs))
        store.ingest_seconds = secs.value
        store.device = dev
        return dev
    if store.tables is None:
        raise ValueError("store has no host tables to upload")
    shape = tuple(store.tables[0].shape)
    dev = _create(shape, len(store.tables), store.beta[0], store.delta[0], store.tau_idx)
    for k, T in enumerate(store.tables):
        if tuple(T.shape) != shape:
            raise ValueError("table %d has shape %s, expected %s" % (k, T.shape, shape))
        a = _hip.contiguous(T, np.float32)
        _hip.call("exm_store_set_table", int(dev.handle), k, _hip.ptr(a))
    store.device = dev
    return dev


def free(store):
    """release the device copy of a store (src/exmarg/search/store.py:123-127), and the batch buffer
    pe_intrinsic_batch keeps with it"""
    if store is None:
        return
    kept = store.__dict__.pop("_batch", None)
    if kept is not None:
        kept[1].free()
    if store.device is not None:
        store.device.free()
        store.device = None


def synthetic_store(n_nodes=1, seed=0, kind="smooth", s_nodes=None, shape=None, max_lag=43, host=False):
    """a store of synthetic tables generated on the device (and on the host when host=True), with the
    grids and lags of the reference store files (exmarg_amd.synth.synthetic_store_arrays gives the same
    values on the host)"""
    beta, delta, _ = synth.store_grids()
    tau = synth.lag_indices(max_lag)
    if shape is not None:
        ntau, nphif, nb, nd = shape
        tau = synth.lag_indices((ntau - 1) // 2)
        beta = beta[0] + (beta[1] - beta[0]) * np.arange(nb)
        delta = (delta[1] - delta[0]) * (np.arange(nd) - nd // 2)
    shp = (tau.size, synth.NPHIF, beta.size, delta.size)
    if s_nodes is None:
        s_nodes = synth.S_NODES[-n_nodes:] if n_nodes < synth.S_NODES.size else synth.S_NODES
    s_nodes = np.asarray(s_nodes, float)
    nn = s_nodes.size
    tables = None
    if host:
        tables = [synth.table_values(seed, k, shp, tau, kind).reshape(shp) for k in range(nn)]
    st = RedshiftStore(s=s_nodes, alpha_lo=np.full(nn, synth.ALPHA[0]), alpha_hi=np.full(nn, synth.ALPHA[-1]),
                       beta=[beta] * nn, delta=[delta] * nn, tau_idx=tau, tables=tables, n_max=120, shape=shp,
                       synthetic=dict(seed=seed, kind=kind, nodes=np.arange(nn)))
    dev = _create(shp, nn, beta, delta, tau)
    for k in range(nn):
        _hip.call("exm_store_fill_synthetic", int(dev.handle), k, ctypes.c_uint64(seed), k,
                  1 if kind == "rough" else 0)
    st.device = dev
    return st
