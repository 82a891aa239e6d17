"""
The template bank and the matched filter on the GPU (src/exmarg/search/bank.py:1-94).

    z_I(t) = FFT_t[ conj(d_I) / sig2_I * W ]   (zero padded to n_time = 8 nu_max)

W = (a / D0) e^{i psi0} depends on the template only and stays on the device;
a segment gives four (n_template, n_time) float32 device arrays, ln|z_I| and
arg z_I at every arrival time (the trajectories).
"""

import numpy as np

from .. import _hip, synth
from .segments import read_segments

DEVICES = ("hip",)


def check_device(device):
    """device 'hip' is this package; 'cpu' and 'gpu' are the reference's own paths
    (src/exmarg/search/bank.py:17-24 raises ValueError for an unknown device)"""
    if device != "hip":
        raise ValueError("device must be 'hip' (the reference keeps 'cpu' and 'gpu'), got %r" % (device,))


class Trajectories:
    """matched-filter outputs of one segment on the device: lqH, phH, lqL, phL (n_template, n_time) float32.
    Indexing by name copies that array to the host."""

    KEYS = ("lqH", "phH", "lqL", "phL")

    def __init__(self, handle, n_template, n_time):
        self.handle = handle
        self.shape = (n_template, n_time)

    @classmethod
    def from_host(cls, traj):
        _hip.init()
        arrs = [_hip.contiguous(traj[k], np.float32) for k in cls.KEYS]
        nm, nt = arrs[0].shape
        h = _hip.new_handle("exm_traj_create", nm, nt, *[_hip.ptr(a) for a in arrs], kind="traj")
        return cls(h, nm, nt)

    def to_host(self):
        out = {k: np.empty(self.shape, np.float32) for k in self.KEYS}
        _hip.call("exm_traj_read", int(self.handle), *[_hip.ptr(out[k]) for k in self.KEYS])
        return out

    def __getitem__(self, key):
        if key not in self.KEYS:
            raise KeyError(key)
        return self.to_host()[key]

    def keys(self):
        return self.KEYS

    def free(self):
        self.handle.free()


class MassBank:
    """templates for one frequency grid and noise curve (src/exmarg/search/bank.py:27-54).

    M, q, chi  (n_template,) detector-frame total mass, mass ratio m2/m1 <= 1 and common aligned spin
    u          (n_template,) template power sum a^2 / sig2_H at D0
    W          (n_template, n_freq) complex64 (a / D0) e^{i psi0}, host copy; the device copy is
               behind .handle

    W comes from `factors` when given (any waveform model, uploaded from the host), else from
    TaylorF2 computed on the device (exm_bank_taylorf2, the model of exmarg_amd.synth.taylorf2);
    the host copy of a device-computed W is read back on first access."""

    def __init__(self, M, q, chi, seg, device="hip", factors=None):
        check_device(device)
        self.M, self.q, self.chi = (np.ascontiguousarray(np.atleast_1d(v), float) for v in (M, q, chi))
        self.device = device
        self.T, self.fmin, self.fmax = seg["T"], seg["fmin"], seg["fmax"]
        self.nu = np.asarray(seg["nu"])
        if not np.all(np.diff(self.nu) == 1):
            raise ValueError("frequency bins must be contiguous")
        self.lo, self.hi = int(self.nu[0]), int(self.nu[-1]) + 1
        self.n_time = 8 * int(self.nu.max())
        _hip.init()
        if factors is None:
            sig2 = _hip.contiguous(seg["sig2_H"], np.float64)
            if sig2.shape != (self.nu.size,):
                raise ValueError("sig2_H must have %d bins, got %s" % (self.nu.size, sig2.shape))
            self.u = np.empty(self.M.size)
            self._W = None
            self.handle = _hip.new_handle("exm_bank_taylorf2", _hip.ptr(self.M), _hip.ptr(self.q), _hip.ptr(self.chi),
                                          self.M.size, float(self.T), self.lo, self.nu.size, self.n_time,
                                          _hip.ptr(sig2), float(synth.D0), _hip.ptr(self.u), kind="bank")
        else:
            W, u = factors
            self._W = np.ascontiguousarray(W, np.complex64)
            self.u = np.ascontiguousarray(u, float)
            if self._W.shape != (self.M.size, self.nu.size):
                raise ValueError("W has shape %s, expected %s" % (self._W.shape, (self.M.size, self.nu.size)))
            self.handle = _hip.new_handle("exm_bank_create", _hip.ptr(self._W.view(np.float32)), self.M.size,
                                          self.nu.size, self.lo, self.n_time, kind="bank")

    @property
    def W(self):
        """(n_template, n_freq) complex64 host copy of the template factors"""
        if self._W is None:
            W = np.empty((self.M.size, self.nu.size), np.complex64)
            _hip.call("exm_bank_read", int(self.handle), _hip.ptr(W.view(np.float32)))
            self._W = W
        return self._W

    def trajectories(self, d, sig2, chunk=1024):
        """matched filter of one segment against every template (src/exmarg/search/bank.py:56-78).

        d, sig2  dicts {'H': (n_freq,), 'L': (n_freq,)} of data and noise variance
        chunk    accepted for the reference signature; the library sizes its own batches
        returns Trajectories on the device."""
        arrs = [_hip.contiguous(d["H"], np.complex128), _hip.contiguous(d["L"], np.complex128),
                _hip.contiguous(sig2["H"], np.float64), _hip.contiguous(sig2["L"], np.float64)]
        for a in arrs:
            if a.shape != (self.nu.size,):
                raise ValueError("segment arrays must have %d bins, got %s" % (self.nu.size, a.shape))
        h = _hip.new_handle("exm_filter", int(self.handle), *[_hip.ptr(a) for a in arrs], kind="traj")
        return Trajectories(h, self.M.size, self.n_time)


def mass_bank(bank, segments, device="hip", factors=None):
    """the template bank (src/exmarg/search/bank.py:81-94).

    bank      .npy path or array: (n,) detector-frame total masses (equal mass, zero spin) or (n, 3) rows
              (M_det, q, chi)
    segments  anything read_segments takes; its frequency grid, T, band and sig2_H define the templates
    device    'hip'
    factors   optional (W, u) computed elsewhere (e.g. from another waveform model)"""
    B = np.load(bank) if isinstance(bank, (str, bytes)) or hasattr(bank, "__fspath__") else np.asarray(bank)
    B = np.asarray(B, float)
    if B.ndim == 1:
        B = np.stack([B, np.ones_like(B), np.zeros_like(B)], axis=1)
    seg = read_segments(segments, rows=[0])
    return MassBank(B[:, 0], B[:, 1], B[:, 2], seg, device, factors)
