"""FP64-oracle results of each benchmark configuration at its full size, for the GPU parity tests
(tests/test_gpu_configs.py).  The inputs are the seeded synthetic workloads of bench.py (the same
generators the tests call), the template factors W are the host TaylorF2 of exmarg_amd.synth (the tests
hand the same W to the device bank), the tables the synthetic recipe of every node.  Each file holds
only results (ln Lambda per template, ln B, a sample of the source-frame grid):

  config2.npz   one 16 s segment, 20-1024 Hz, 2000-template 20x10x10 bank, 53 nodes, pe_intrinsic 'volume'
  config3.npz   row 0 of each segment length (4, 8, 16, 32, 64 s) of the 256-segment batched workload
  config4.npz   requests 4 and 5 of the serving workload (seed 7): 1000 and 500 templates
  config5.npz   rows 0 and 6 of the first group of the stream workload (seed 7): 8 s, 20-512 Hz, the
                3517 masses of the reference's two-parameter bank, pe_redshift at 1600 x 1600

    python tests/golden/make_config_fixtures.py [config2 config3 ...]     (files that exist are kept)

A full-bank oracle run takes minutes per segment on a few cores (about 1.6e10 table lookups at the
default 1024 + 1024 walked times).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import bench                                   # noqa: E402  (the workload generators)
import exm_oracle as O                         # noqa: E402
from exmarg_amd import synth                   # noqa: E402
from exmarg_amd.search.segments import read_segments    # noqa: E402

K = S = 1024
GRID_STEP = 40                                  # the source-frame grid sample: every 40th m1 and z


def store_dict(ymin_nodes=53):
    beta, delta, _ = synth.store_grids()
    s = synth.S_NODES
    return dict(s=s, alpha_lo=np.full(s.size, synth.ALPHA[0]), alpha_hi=np.full(s.size, synth.ALPHA[-1]),
                beta=beta, delta=delta, tau_idx=synth.lag_indices(43))


def tables():
    return O.Tables(synth=dict(seed=0, kind="smooth", nodes=np.arange(53)), shape=(87, synth.NPHIF, 200, 219),
                    tau_idx=synth.lag_indices(43), materialize=True)


def bank_dict(rows, seg):
    """the oracle's bank: host TaylorF2 factors of the rows on the segment's grid"""
    rows = np.asarray(rows, float)
    if rows.ndim == 1:
        rows = np.stack([rows, np.ones_like(rows), np.zeros_like(rows)], axis=1)
    W, u = synth.template_factors(rows[:, 0], rows[:, 1], rows[:, 2], seg)
    nu = np.asarray(seg["nu"])
    return dict(W=W, u=u, M=rows[:, 0], q=rows[:, 1], chi=rows[:, 2], lo=int(nu[0]), n_time=8 * int(nu[-1]))


def one_row(seg, j):
    r = read_segments(seg, rows=[j])
    r["rows"] = np.array([0])
    return r


def intrinsic(seg, rows, seed):
    t0 = time.time()
    res = O.pe_intrinsic(seg, bank_dict(rows, seg), store_dict(), tables(), n_loud=K, n_random=S, seed=seed)
    print("  %d templates, T %g s: ln B %.8f (%.0f s)" % (np.asarray(rows).shape[0], seg["T"], res["ln_b"][0],
                                                          time.time() - t0), flush=True)
    return res["ln_lambda"][0], res["ln_b"][0]


def config2():
    seg, rows, _ = bench.workload(type("A", (), {"grid": "20,10,10"})())
    lnl, lnb = intrinsic(read_segments(seg), rows, 0)
    return dict(ln_lambda=lnl, ln_b=lnb)


def config3():
    # one segment length at a time, each kept in config3.partial.npz until the file is complete
    part = os.path.join(HERE, "config3.partial.npz")
    out = dict(np.load(part)) if os.path.exists(part) else {}
    rows = synth.grid_bank(shape=(20, 10, 10))
    for seg, T in bench.batch_items(256, 7):
        if "ln_b_%d" % int(T) in out:
            continue
        lnl, lnb = intrinsic(one_row(seg, 0), rows, 0)
        out["ln_lambda_%d" % int(T)], out["ln_b_%d" % int(T)] = lnl, lnb
        np.savez(part + ".tmp.npz", **out)
        os.replace(part + ".tmp.npz", part)
    return out


def config4():
    out = {}
    reqs = bench.serve_requests(16, 7, 50.0)
    for i in (4, 5):
        _, seg, rows = reqs[i]
        lnl, lnb = intrinsic(read_segments(seg), rows, i)
        out["ln_lambda_%d" % i], out["ln_b_%d" % i] = lnl, lnb
    return out


def config5():
    from exmarg_amd.priors import madau_dickinson_ymin
    groups, truths = bench.stream_groups(64, 32, 7)
    seg = read_segments(groups[0])
    masses = np.load(os.path.join(ROOT, "exmarg_amd", "data", "bank_2p.npy"))
    res = {}
    for j in (0, 6):
        t0 = time.time()
        r = O.pe_redshift(one_row(seg, j), bank_dict(masses, seg), store_dict(), tables(), zmax=10.0, nz=1600,
                          m1_range=(2.0, 200.0), nm1=1600, ymin=madau_dickinson_ymin(), mdet_range=(10.0, 500.0),
                          n_loud=K, n_random=S, seed=j)
        print("  row %d: ln B %.8f (%.0f s)" % (j, r["ln_b"][0], time.time() - t0), flush=True)
        res["ln_b_%d" % j] = r["ln_b"][0]
        res["grid_%d" % j] = r["ln_lambda"][0][::GRID_STEP, ::GRID_STEP]
    return res


def main(names):
    for name in names:
        path = os.path.join(HERE, name + ".npz")
        if os.path.exists(path):
            print(name, "exists", flush=True)
            continue
        print(name, flush=True)
        data = globals()[name]()
        np.savez(path + ".tmp.npz", **data)
        os.replace(path + ".tmp.npz", path)
        part = os.path.join(HERE, name + ".partial.npz")
        if os.path.exists(part):
            os.remove(part)
        print(name, "written", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["config2", "config4", "config5", "config3"])
