"""
Stage 3 of the tables on the device: the sum over the orders (scripts/build_tables/build_bd.py).

Every order of stage 2 is homogeneous, lnK_n(q1, q2) = 2n ln q2 + g_n(ln q1 - ln q2), so the table read by
the analysis,

    T(tau, phi, beta, Delta) = logsumexp_n [ n beta + g_n(Delta) ],   beta = ln alpha + 2 ln q2,

needs only the last row and the last column of each (q1, q2) block of stage 2.  This module reads them
from a node's stage-2 files, forms T for every truncation order of the ladder on the GPU
(exm_tables_sum_orders: float64 arithmetic, float32 tables) and writes the node's bd_*.npz files with
the keys and names of the reference (build_bd.py:46-47,78-79), so that load_redshift_store reads them.

    python -m exmarg_amd.build_bd <k_dir> <out_dir> [node indices, default all]
    -> <out_dir>/bd_n120_s<s>.npz

BD_NMAX_LADDER=1,2,5,... also writes the tables truncated at those orders, from the same pass, into
<out_dir>_n001, ...; BD_INMEM=1 reads a node's stage-2 array into memory first (build_bd.py:14-16).
"""

import ctypes
import glob
import os
import sys
import time

import numpy as np

from . import _hip

ALPHA = np.geomspace(0.16, 4.0e4, 30)               # sets the range of beta (build_bd.py:26)


def grids(lnq):
    """(beta, delta) of a node with ln q grid lnq (build_bd.py:52-60)"""
    lnq = np.asarray(lnq, np.float64)
    last = lnq.size - 1
    dlq = float(lnq[1] - lnq[0])
    delta = np.concatenate([lnq - lnq[last], (lnq[last] - lnq)[::-1][1:]])
    dbeta = 2.0 * dlq
    b_lo = np.log(ALPHA[0]) + 2.0 * lnq[0]
    b_hi = np.log(ALPHA[-1]) + 2.0 * lnq[-1]
    nb = int(np.ceil((b_hi - b_lo) / dbeta)) + 1
    return b_lo + dbeta * np.arange(nb), delta


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


def out_dir(base, nmax, nmax_full):
    """the full order in base, the truncated tables next to it (build_bd.py:30-32)"""
    return base if nmax == nmax_full else '%s_n%03d' % (base, nmax)


def read_stage2(stem, inmem=False):
    """(col, row, meta) of a node's stage-2 files: <stem>.lnK.npy (nord, ntau, nq, nq, nphi) and
    <stem>.meta.npz (build_bd.py:36-49,61-62)"""
    meta = np.load(stem + '.meta.npz')
    A = np.load(stem + '.lnK.npy') if inmem else np.load(stem + '.lnK.npy', mmap_mode='r')
    nord, ntau, nq, _, nphi = A.shape
    nphif = nphi // 2 + 1
    last = nq - 1
    col = np.ascontiguousarray(A[:, :, :, last, :nphif], np.float32)
    row = np.ascontiguousarray(A[:, :, last, :, :nphif], np.float32)
    return col, row, meta


def build_node(stem, base, ladder=None):
    """the bd tables of one node from its stage-2 files (build_bd.py:35-79); returns the paths written"""
    meta = np.load(stem + '.meta.npz')
    s = float(meta['s_node'])
    lnq = meta['lnq'].astype(np.float64)
    tau_idx = meta['tau_idx']
    nmax_full = int(os.path.basename(stem).split('_n')[1].split('_')[0])
    if ladder is None:
        env = os.environ.get('BD_NMAX_LADDER', '').strip()
        ladder = [int(v) for v in env.split(',')] if env else [nmax_full]
    ladder = sorted(int(v) for v in ladder)
    if max(ladder) > nmax_full:
        raise ValueError("truncation order %d above the stored order %d" % (max(ladder), nmax_full))
    outs = {nm: os.path.join(out_dir(base, nm, nmax_full), 'bd_n%03d_s%09.3f.npz' % (nm, s)) for nm in ladder}
    for d in {os.path.dirname(v) for v in outs.values()}:
        os.makedirs(d, exist_ok=True)
    todo = [nm for nm in ladder if not os.path.exists(outs[nm])]
    if not todo:
        print('s=%g exists for every order, skipped' % s, flush=True)
        return []
    t0 = time.time()
    col, row, _ = read_stage2(stem, inmem=bool(int(os.environ.get('BD_INMEM', '0'))))
    nphif = col.shape[3]
    tables = sum_over_orders(col, row, lnq, todo)
    beta, delta = grids(lnq)
    for nm in todo:
        if not np.isfinite(tables[nm]).all():
            raise FloatingPointError('table has non-finite entries at n_max=%d' % nm)
        np.savez(outs[nm], T=tables[nm], beta=beta, delta=delta, tau_idx=tau_idx, lnq=lnq, alpha=ALPHA, s_node=s,
                 u_node=s, n_max=nm, nphif=nphif, norm='taylor_off1')
    print('s=%g done: %d orders %s in %.1f s' % (s, len(todo), todo, time.time() - t0), flush=True)
    return [outs[nm] for nm in todo]


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 2:
        sys.exit("usage: python -m exmarg_amd.build_bd <k_dir> <out_dir> [node indices]")
    k_dir, base = argv[0], argv[1]
    stems = sorted(f[:-len('.meta.npz')] for f in glob.glob(os.path.join(k_dir, 'kstore_n*_s*.meta.npz')))
    if not stems:
        raise FileNotFoundError("no stage-2 node (kstore_n*_s*.meta.npz) in %s" % k_dir)
    pick = [int(v) for v in argv[2:]] or list(range(len(stems)))
    _hip.init(0)
    for i in pick:
        build_node(stems[i], base)


if __name__ == '__main__':
    main()
