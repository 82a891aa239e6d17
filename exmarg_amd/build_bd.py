"""
Stage 3 of the tables on the device: the sum over the orders (scripts/build_tables/build_bd.py).

Every order of stage 2 is homogeneous, lnK_n(q1, q2) = 2n ln q2 + g_n(ln q1 - ln q2), so the table read by
the analysis,

    T(tau, phi, beta, Delta) = logsumexp_n [ n beta + g_n(Delta) ],   beta = ln alpha + 2 ln q2,

needs only the last row and the last column of each (q1, q2) block of stage 2.  This module reads them
from a node's stage-2 files, forms T for every truncation order of the ladder on the GPU
(exm_tables_sum_orders: float64 arithmetic, float32 tables) and writes the node's bd_*.npz files with
the keys and names of the reference (build_bd.py:43,83-84), so that load_redshift_store reads them.

    python -m exmarg_amd.build_bd <k_dir> <out_dir> [node indices, default all]
    -> <out_dir>/bd_n120_s<s>.npz

A node index is the position of the node's s in the 53-node grid (data/tables/s_grid.npz of the
reference, build_bd.py:88-92), whichever nodes k_dir holds.  BD_NMAX_LADDER=1,2,5,... also writes the
tables truncated at those orders, from the same pass, into <out_dir>_n001, ...; BD_INMEM=1 reads a
node's stage-2 array into memory first (build_bd.py:14-16).
"""

import ctypes
import glob
import os
import re
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import _hip, synth

ALPHA = np.geomspace(0.16, 4.0e4, 30)               # sets the range of beta (build_bd.py:26)
_STEM = re.compile(r"kstore_n(\d+)_s([0-9.]+)$")


def grids(lnq):
    """(beta, delta) of a node with ln q grid lnq (build_bd.py:52-60): Delta runs over ln q1 - ln q2
    with one of them at q_max (2 nq - 1 sorted values), beta from ln alpha_min + 2 ln q_min in steps
    of 2 dln q until ln alpha_max + 2 ln q_max is covered"""
    lnq = np.asarray(lnq, np.float64)
    top = lnq[-1]
    delta = np.concatenate([lnq - top, top - lnq[-2::-1]])
    step = 2.0 * float(lnq[1] - lnq[0])
    first = np.log(ALPHA[0]) + 2.0 * lnq[0]
    span = np.log(ALPHA[-1]) + 2.0 * top - first
    return first + step * np.arange(int(np.ceil(span / step)) + 1), delta


def sum_over_orders(col, row, lnq, ladder=None):
    """{n_max: T (ntau, nphif, nb, 2 nq - 1) float32} for every truncation order of the ladder (default the
    full order, nord - 1).  col and row (nord, ntau, nq, nphif): lnK_n at q2 = q_max and at q1 = q_max."""
    col = _hip.contiguous(col, np.float32)
    row = _hip.contiguous(row, np.float32)
    if col.ndim != 4 or row.shape != col.shape:
        raise ValueError("col and row must share a (nord, ntau, nq, nphif) shape, got %s and %s"
                         % (col.shape, row.shape))
    nord, ntau, nq, nphif = col.shape
    lnq = _hip.contiguous(lnq, np.float64)
    if lnq.shape != (nq,):
        raise ValueError("lnq has %s points, the tables %d" % (lnq.shape, nq))
    ladder = [nord - 1] if ladder is None else [int(v) for v in ladder]
    beta, _ = grids(lnq)
    beta = _hip.contiguous(beta, np.float64)
    lad = np.asarray(ladder, np.int32)
    outs = [np.empty((ntau, nphif, beta.size, 2 * nq - 1), np.float32) for _ in ladder]
    ptrs = (ctypes.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
    _hip.call("exm_tables_sum_orders", _hip.ptr(col), _hip.ptr(row), nord, ntau, nq, nphif, _hip.ptr(lnq),
              _hip.ptr(beta), int(beta.size), _hip.ptr(lad), len(ladder), ctypes.cast(ptrs, ctypes.c_void_p))
    return dict(zip(ladder, outs))


def store_path(base, n_max, n_full, s):
    """where the table of node s truncated at order n_max goes: the full order's store is base, each
    truncation a sibling directory base_nNNN (build_bd.py:30-32,43)"""
    folder = base if n_max == n_full else "%s_n%03d" % (base, n_max)
    return os.path.join(folder, "bd_n%03d_s%09.3f.npz" % (n_max, s))


@dataclass
class Stage2Node:
    """one node's stage-2 output: <stem>.meta.npz (s, ln q grid, lags) and <stem>.lnK.npy"""
    stem: str
    s: float
    lnq: np.ndarray
    tau_idx: np.ndarray
    n_full: int

    @classmethod
    def open(cls, stem):
        m = _STEM.search(os.path.basename(stem))
        if m is None:
            raise ValueError("%s is not a kstore_n<order>_s<s> stem" % stem)
        with np.load(stem + ".meta.npz") as meta:
            return cls(stem=stem, s=float(meta["s_node"]), lnq=np.asarray(meta["lnq"], np.float64),
                       tau_idx=np.asarray(meta["tau_idx"]), n_full=int(m.group(1)))

    def edges(self, in_memory=False):
        """(col, row) (nord, ntau, nq, nphif): lnK_n on the q2 = q_max column and the q1 = q_max row
        of each block, the stored half of the phases (build_bd.py:61-62)"""
        path = self.stem + ".lnK.npy"
        A = np.load(path) if in_memory else np.load(path, mmap_mode="r")
        nphif = A.shape[4] // 2 + 1
        return (np.ascontiguousarray(A[:, :, :, -1, :nphif], np.float32),
                np.ascontiguousarray(A[:, :, -1, :, :nphif], np.float32))


def truncation_orders(n_full, ladder=None):
    """the orders to write: the ladder given, else BD_NMAX_LADDER, else the full order; ascending"""
    if ladder is None:
        spec = os.environ.get("BD_NMAX_LADDER", "").strip()
        ladder = spec.split(",") if spec else [n_full]
    orders = sorted({int(v) for v in ladder})
    if orders[-1] > n_full:
        raise ValueError("truncation order %d above the stored order %d" % (orders[-1], n_full))
    return orders


def _write(path, T, node, n_max):
    if not np.isfinite(T).all():
        raise FloatingPointError("table has non-finite entries at n_max=%d" % n_max)
    beta, delta = grids(node.lnq)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, T=T, beta=beta, delta=delta, tau_idx=node.tau_idx, lnq=node.lnq, alpha=ALPHA, s_node=node.s,
             u_node=node.s, n_max=n_max, nphif=T.shape[1], norm="taylor_off1")


def build_node(stem, base, ladder=None):
    """the bd tables of one node from its stage-2 files (build_bd.py:35-84).  Orders whose file already
    exists are skipped; returns the paths written."""
    node = Stage2Node.open(stem)
    targets = {n: store_path(base, n, node.n_full, node.s) for n in truncation_orders(node.n_full, ladder)}
    pending = [n for n, p in targets.items() if not os.path.exists(p)]
    if not pending:
        print("node s=%g: every order present, nothing to do" % node.s, flush=True)
        return []
    t0 = time.time()
    col, row = node.edges(in_memory=os.environ.get("BD_INMEM", "0") not in ("", "0"))
    tables = sum_over_orders(col, row, node.lnq, pending)
    for n in pending:
        _write(targets[n], tables[n], node, n)
    print("node s=%g: orders %s written in %.1f s" % (node.s, pending, time.time() - t0), flush=True)
    return [targets[n] for n in pending]


def node_stems(k_dir, indices=None):
    """the stage-2 stems of the nodes at the given positions of the 53-node s grid (default: every node
    k_dir holds, in s order); a position whose node k_dir lacks raises FileNotFoundError"""
    found = {}
    for meta in glob.glob(os.path.join(k_dir, "kstore_n*_s*.meta.npz")):
        stem = meta[:-len(".meta.npz")]
        m = _STEM.search(os.path.basename(stem))
        if m is not None:
            found[m.group(2)] = stem
    if not found:
        raise FileNotFoundError("no stage-2 node (kstore_n*_s*.meta.npz) in %s" % k_dir)
    if indices is None:
        return [found[k] for k in sorted(found, key=float)]
    out = []
    for i in indices:
        if not 0 <= int(i) < synth.S_NODES.size:
            raise IndexError("node index %d outside the %d-node grid" % (int(i), synth.S_NODES.size))
        key = "%09.3f" % synth.S_NODES[int(i)]
        if key not in found:
            raise FileNotFoundError("node %d (s=%s) has no stage-2 files in %s" % (int(i), key, k_dir))
        out.append(found[key])
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 2:
        sys.exit("usage: python -m exmarg_amd.build_bd <k_dir> <out_dir> [node indices]")
    stems = node_stems(argv[0], [int(v) for v in argv[2:]] or None)
    _hip.init(0)
    for stem in stems:
        build_node(stem, argv[1])


if __name__ == "__main__":
    main()
