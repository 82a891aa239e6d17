"""
The segments an analysis reads: frequency-domain data of the two detectors,
their noise variance, the frequency grid and the band
(src/exmarg/search/segments.py:1-54).
"""

import os

import numpy as np

from . import npzmember

D0 = 10410.0              # reference distance of the stores (Mpc): y = D0 / D_L (segments.py:8)

DETECTORS = ("H", "L")


class _Source:
    """the arrays of a segment set, from an .npz file or from a mapping of arrays.  From a file, the data
    rows come from a memory map of the (uncompressed) member, so only the rows asked for are read."""

    def __init__(self, segments):
        self.path = os.fspath(segments) if isinstance(segments, (str, bytes, os.PathLike)) else None
        if self.path is None:
            self.arrays = segments
            self.names = set(segments.keys())
        else:
            with np.load(self.path) as z:
                self.names = {n for n in z.files}
                self.arrays = {n: z[n] for n in self.names if not n.startswith("d_")}

    def data(self, det):
        name = "d_" + det
        if self.path is None:
            return np.asarray(self.arrays[name])
        member = npzmember.locate(self.path, name + ".npy", compressed_ok=True)
        if member is None:
            with np.load(self.path) as z:
                return z[name]
        return member.mapped()

    def noise(self, det):
        """sig2 of one detector: its own sig2_<det> when present, else the common sig2"""
        own = "sig2_" + det
        return np.asarray(self.arrays[own if own in self.names else "sig2"], np.float64)

    def scalar(self, name):
        return float(np.asarray(self.arrays[name]))


def read_segments(segments, rows=None):
    """the arrays the analysis reads from a set of segments (src/exmarg/search/segments.py:11-37).

    segments  npz path, loaded npz, or dict with d_H, d_L (n_segment, n_freq) complex data, f (n_freq,),
              nu (n_freq,) integer bins f T (contiguous), sig2 (n_freq,) or sig2_H, sig2_L, T, fmin, fmax
    rows      segment indices to keep (default all)

    returns a dict with the same keys, d_H, d_L as complex128 and sig2_H, sig2_L as float64."""
    src = _Source(segments)
    d = {det: src.data(det) for det in DETECTORS}
    keep = np.arange(d["H"].shape[0]) if rows is None else np.asarray(rows, int)
    out = dict(rows=keep, f=np.asarray(src.arrays["f"], np.float64), nu=np.asarray(src.arrays["nu"], np.int64),
               T=src.scalar("T"), fmin=src.scalar("fmin"), fmax=src.scalar("fmax"))
    for det in DETECTORS:
        out["d_" + det] = np.asarray(d[det][keep], dtype=np.complex128)
        out["sig2_" + det] = src.noise(det)
    return out
