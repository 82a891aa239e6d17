"""
Synthetic inputs for the segment analysis: noise curve, frequency grid,
TaylorF2 templates, simulated two-detector segments, template banks and
stores of tables in the on-disk layout of the reference.

Nothing here is on the analysis path; it produces the segments, banks and
tables the analysis reads (the reference simulates segments the same way in
src/exmarg/pp.py:38-72 and builds W from a waveform in
src/exmarg/search/bank.py:46-53).

Table values are a deterministic function of (seed, node, flat index) built
from integer arithmetic only, so the host writer here and the device
generator of the HIP library produce bit-identical float32 tables.
"""

import os

import numpy as np

D0 = 10410.0                        # Mpc, y = D0 / D_L (src/exmarg/search/segments.py:8)
C_SI = 299792458.0
MSUN_S = 4.925490947641267e-06      # G Msun / c^3 in seconds (IAU nominal solar mass parameter)
MPC_S = 3.0856775814913674e22 / C_SI  # one megaparsec in light-seconds
EULER_GAMMA = 0.5772156649015329

# ---------------------------------------------------------------- detectors
# Vertex positions and arm unit vectors of LIGO Hanford and Livingston
# (the published LIGO site coordinates, as src/exmarg/detectors.py:7-13 uses).
DET_X = {"H": np.array([-2.16141492636e6, -3.83469517889e6, 4.60035022664e6]),
         "L": np.array([-7.42760447238e4, -5.49628371971e6, 3.22425701744e6])}
DET_U = {"H": np.array([-0.22389266154, 0.79983062746, 0.55690487831]),
         "L": np.array([-0.95457412153, -0.14158077340, -0.26218911324])}
DET_V = {"H": np.array([-0.91397818574, 0.02609403989, -0.40492342125]),
         "L": np.array([0.29774156894, -0.48791033647, -0.82054461286])}


def _det_tensor(I):
    return 0.5 * (np.outer(DET_U[I], DET_U[I]) - np.outer(DET_V[I], DET_V[I]))


def antenna(I, theta, phi, psi):
    """F+, Fx of detector I for a source at polar angle theta, azimuth phi and polarization psi"""
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    th = np.array([ct * cp, ct * sp, -st])
    ph = np.array([-sp, cp, 0.0])
    ep = np.outer(th, th) - np.outer(ph, ph)
    ex = np.outer(th, ph) + np.outer(ph, th)
    d = _det_tensor(I)
    c2, s2 = np.cos(2 * psi), np.sin(2 * psi)
    return float(np.sum(d * (ep * c2 - ex * s2))), float(np.sum(d * (ep * s2 + ex * c2)))


def sky_direction(theta, phi):
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


# ---------------------------------------------------------------- noise
def psd_ajith(f, f_low=10.0):
    """one-sided aLIGO design PSD of the analytic fit (Ajith 2011, Eq. 4.7), infinite below f_low
    (the 'ajith' curve of src/exmarg/detectors.py:89-94)"""
    f = np.asarray(f, dtype=float)
    x = np.where(f > 0, f, 1.0) / 215.0
    s = 1e-49 * (x ** -4.14 - 5.0 * x ** -2 + 111.0 * (1.0 - x ** 2 + 0.5 * x ** 4) / (1.0 + 0.5 * x ** 2))
    return np.where(f >= f_low, s, np.inf)


def freq_bins(T, fmin, fmax):
    """integer bins nu = f T with fmin <= f <= fmax (src/exmarg/detectors.py:104-109)"""
    return np.arange(int(np.ceil(fmin * T)), int(np.floor(fmax * T)) + 1)


# ---------------------------------------------------------------- TaylorF2
def taylorf2(M, q, chi, f, distance_mpc=1.0):
    """face-on h+ of a TaylorF2 waveform: amplitude a(f) and phase psi0(f) with h+ = a e^{i psi0}.

    M detector-frame total mass (Msun), q = m2/m1 <= 1, chi the common aligned spin.
    Leading-order SPA amplitude, 3.5PN point-particle phase with the 1.5PN, 2.5PN and 3.5PN
    spin-orbit and 2PN spin-spin terms (Buonanno, Iyer, Ochsner, Pan & Sathyaprakash 2009,
    PRD 80 084043, Eq. 3.18 and Table I; spin terms of Arun, Buonanno, Faye & Ochsner 2009,
    PRD 79 104023 and Bohe, Marsat & Blanchet 2013).  The amplitude is zero above the
    Schwarzschild ISCO frequency of M.  Linear-in-f and constant phase terms are dropped
    (absorbed by the arrival-time and phase marginalizations)."""
    f = np.asarray(f, float)
    eta = q / (1.0 + q) ** 2
    delta = (1.0 - q) / (1.0 + q)
    Ms = M * MSUN_S
    Mc = Ms * eta ** 0.6
    chis, chia = chi, 0.0
    v = (np.pi * Ms * f) ** (1.0 / 3.0)
    f_isco = 1.0 / (6.0 ** 1.5 * np.pi * Ms)
    pi = np.pi
    p0 = 1.0
    p2 = 3715.0 / 756.0 + 55.0 * eta / 9.0
    p3 = -16.0 * pi + (113.0 / 3.0) * delta * chia + (113.0 / 3.0 - 76.0 * eta / 3.0) * chis
    p4 = (15293365.0 / 508032.0 + 27145.0 * eta / 504.0 + 3085.0 * eta ** 2 / 72.0
          + (-405.0 / 8.0 + 200.0 * eta) * chia ** 2 - 405.0 / 4.0 * delta * chia * chis
          + (-405.0 / 8.0 + 5.0 * eta / 2.0) * chis ** 2)
    p5c = (pi * (38645.0 / 756.0 - 65.0 * eta / 9.0)
           - (732985.0 / 2268.0 - 24260.0 * eta / 81.0 - 340.0 * eta ** 2 / 9.0) * chis
           - (732985.0 / 2268.0 + 140.0 * eta / 9.0) * delta * chia)
    p6 = (11583231236531.0 / 4694215680.0 - 640.0 * pi ** 2 / 3.0 - 6848.0 * EULER_GAMMA / 21.0
          + (-15737765635.0 / 3048192.0 + 2255.0 * pi ** 2 / 12.0) * eta
          + 76055.0 * eta ** 2 / 1728.0 - 127825.0 * eta ** 3 / 1296.0
          + pi * (2270.0 / 3.0 * delta * chia + (2270.0 / 3.0 - 520.0 * eta) * chis))
    p6l = -6848.0 / 21.0
    p7 = (pi * (77096675.0 / 254016.0 + 378515.0 * eta / 1512.0 - 74045.0 * eta ** 2 / 756.0)
          + (-25150083775.0 / 3048192.0 + 10566655595.0 * eta / 762048.0
             - 1042165.0 * eta ** 2 / 3024.0 + 5345.0 * eta ** 3 / 36.0) * chis
          + delta * (-25150083775.0 / 3048192.0 + 26804935.0 * eta / 6048.0
                     - 1985.0 * eta ** 2 / 48.0) * chia)
    with np.errstate(divide="ignore", invalid="ignore"):
        lv = np.log(v)
        series = (p0 + p2 * v ** 2 + p3 * v ** 3 + p4 * v ** 4 + p5c * (1.0 + 3.0 * lv) * v ** 5
                  + (p6 + p6l * np.log(4.0 * v)) * v ** 6 + p7 * v ** 7)
        psi = 3.0 / (128.0 * eta * v ** 5) * series
        amp = (np.sqrt(5.0 / 24.0) * pi ** (-2.0 / 3.0) * Mc ** (5.0 / 6.0)
               * f ** (-7.0 / 6.0) / (distance_mpc * MPC_S))
    keep = (f > 0) & (f <= f_isco)
    return np.where(keep, amp, 0.0), np.where(keep, -psi, 0.0)


def template_factors(M, q, chi, seg):
    """W (n_template, n_freq) complex64 = (a / D0) e^{i psi0} on the bins of seg and u = sum a^2 / sig2_H
    (src/exmarg/search/bank.py:44-53, with TaylorF2 in place of IMRPhenomXAS)"""
    M, q, chi = (np.atleast_1d(np.asarray(v, float)) for v in (M, q, chi))
    f = np.asarray(seg["f"], float)
    W = np.empty((M.size, f.size), np.complex64)
    u = np.empty(M.size)
    for i in range(M.size):
        a, p = taylorf2(M[i], q[i], chi[i], f)
        a = a / D0
        W[i] = a * np.exp(1j * p)
        u[i] = float((a ** 2 / seg["sig2_H"]).sum())
    return W, u


# ---------------------------------------------------------------- segments
def simulate(n, seed, draw, draw_y, T, fmin, fmax, psd=psd_ajith):
    """n segments, each with one TaylorF2 injection in Gaussian noise colored by psd (the same curve in both
    detectors), in the convention of src/exmarg/pp.py:38-72.  draw(rng) -> (M_det, q, chi) and
    draw_y(rng) -> y = D0 / D_L; y = 0 gives noise only.  returns (segments dict, truth (n, 4), network SNR (n,))."""
    nu = freq_bins(T, fmin, fmax)
    f = nu / T
    sig2 = T * psd(f) / 2.0
    d = {I: np.empty((n, nu.size), complex) for I in "HL"}
    truth = np.empty((n, 4))
    snr = np.empty(n)
    for k in range(n):
        rng = np.random.default_rng(seed + k)
        M, q, chi = draw(rng)
        theta, phi = np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)
        psi, cos_iota = rng.uniform(0, np.pi), rng.uniform(-1, 1)
        phase, t_arr = rng.uniform(0, 2 * np.pi), rng.uniform(0, T)
        y = draw_y(rng)
        a, p = taylorf2(M, q, chi, f)
        h = a / D0 * np.exp(1j * p)
        n_hat = sky_direction(theta, phi)
        rho2 = 0.0
        for I in "HL":
            Fp, Fx = antenna(I, theta, phi, psi)
            R = 0.5 * (1 + cos_iota ** 2) * Fp + 1j * cos_iota * Fx
            tau = n_hat @ (DET_X[I] - DET_X["L"]) / C_SI
            signal = y * R * np.exp(-1j * phase) * h * np.exp(-2j * np.pi * f * (t_arr + tau))
            g = rng.standard_normal((2, nu.size))
            d[I][k] = np.sqrt(sig2 / 2.0) * (g[0] + 1j * g[1]) + signal
            rho2 += 2.0 * np.sum(np.abs(signal) ** 2 / sig2)
        truth[k] = M, q, chi, y
        snr[k] = np.sqrt(rho2)
    seg = dict(d_H=d["H"], d_L=d["L"], f=f, nu=nu, sig2=sig2, T=float(T), fmin=float(fmin), fmax=float(fmax))
    return seg, truth, snr


def simulate_at_snr(n, seed, params, snr_target, T, fmin, fmax):
    """segments with one injection of fixed (M_det, q, chi) each, rescaled to the network SNR snr_target[k]
    (0 gives noise only), as 00_quickstart.ipynb rescales y to a target SNR"""
    snr_target = np.broadcast_to(np.asarray(snr_target, float), (n,))
    _, _, snr1 = simulate(n, seed, lambda rng: params, lambda rng: 1.0, T, fmin, fmax)
    it = iter(snr_target / snr1)
    return simulate(n, seed, lambda rng: params, lambda rng: float(next(it)), T, fmin, fmax)


# ---------------------------------------------------------------- banks
def chirp_mass_grid_bank(M_det, q, chi, n, span=0.25):
    """(n, 3) bank of (M_det, q, chi) with q and chi fixed and the chirp mass on a uniform grid of relative
    half-width span around that of M_det"""
    eta = q / (1.0 + q) ** 2
    mc = M_det * eta ** 0.6
    mcs = np.linspace(mc * (1 - span), mc * (1 + span), n)
    return np.stack([mcs / eta ** 0.6, np.full(n, q), np.full(n, chi)], axis=1)


def grid_bank(M_range=(10.0, 80.0), q_range=(0.2, 1.0), chi_range=(-0.5, 0.5), shape=(20, 10, 10)):
    """(prod(shape), 3) regular grid bank over (M_det, q, chi)"""
    axes = [np.linspace(lo, hi, n) for (lo, hi), n in zip((M_range, q_range, chi_range), shape)]
    g = np.meshgrid(*axes, indexing="ij")
    return np.stack([a.ravel() for a in g], axis=1)


# ---------------------------------------------------------------- stores
S_NODES = np.geomspace(0.004264528739608287, 4000.0, 53)   # the 53 signal-power nodes (data/tables/s_grid.npz)
ALPHA = np.geomspace(0.16, 4.0e4, 30)                      # amplitude axis of the store files
NPHI = 96
NPHIF = NPHI // 2 + 1


def store_grids(nq=110, qmin=0.05, qmax=100.0):
    """the beta and Delta grids and lags of the store files, from the build constants
    (scripts/build_tables/build_k.py:30-35 and build_bd.py:26,56-62)"""
    lnq = np.log(np.geomspace(qmin, qmax, nq))
    last = nq - 1
    delta = np.concatenate([lnq - lnq[last], (lnq[last] - lnq)[::-1][1:]])
    dbeta = 2.0 * float(lnq[1] - lnq[0])
    b_lo = np.log(ALPHA[0]) + 2.0 * lnq[0]
    b_hi = np.log(ALPHA[-1]) + 2.0 * lnq[-1]
    nb = int(np.ceil((b_hi - b_lo) / dbeta)) + 1
    beta = b_lo + dbeta * np.arange(nb)
    return beta, delta, lnq


def lag_indices(max_lag=43):
    """lags in samples, fftfreq order [0, 1, .., L, -L, .., -1]"""
    return np.concatenate([np.arange(0, max_lag + 1), np.arange(-max_lag, 0)]).astype(np.int64)


_SM1 = np.uint64(0x9E3779B97F4A7C15)
_SM2 = np.uint64(0xBF58476D1CE4E5B9)
_SM3 = np.uint64(0x94D049BB133111EB)


def _splitmix(x):
    """splitmix64 finalizer on uint64 arrays (wrapping arithmetic)"""
    with np.errstate(over="ignore"):
        z = x + _SM1
        z = (z ^ (z >> np.uint64(30))) * _SM2
        z = (z ^ (z >> np.uint64(27))) * _SM3
        return z ^ (z >> np.uint64(31))


def table_values(seed, node, shape, tau_idx, kind="smooth", flat=None):
    """float32 values of synthetic table `node` (reference layout (ntau, nphif, nb, nd)).

    kind 'smooth': T = (8 max(0, b - 56)(64 - c) - 64 |d - nd//2| - 32 |tau| + h) / 4096, h a 9-bit hash,
                   rising with beta like the real tables, exact in float32;
    kind 'rough':  T = (h - 2^22) / 2^22, h a 23-bit hash, uniform on [-1, 1).
    The same integer recipe runs on the device (exm_store_fill_synthetic).
    flat: optional flat indices to evaluate (default every entry)."""
    ntau, nphif, nb, nd = shape
    n = ntau * nphif * nb * nd
    idx = np.arange(n, dtype=np.uint64) if flat is None else np.asarray(flat, np.uint64)
    key = (np.uint64(seed) * np.uint64(0x100000001B3)) ^ (np.uint64(node) << np.uint64(40)) ^ idx
    h = _splitmix(key)
    if kind == "rough":
        hv = (h >> np.uint64(41)).astype(np.int64)
        return ((hv - (1 << 22)).astype(np.float64) / float(1 << 22)).astype(np.float32)
    hv = (h >> np.uint64(55)).astype(np.int64)
    i = idx.astype(np.int64)
    d = i % nd
    b = (i // nd) % nb
    c = (i // (nd * nb)) % nphif
    r = i // (nd * nb * nphif)
    tau = np.abs(np.asarray(tau_idx, np.int64)[r])
    num = 8 * np.maximum(0, b - 56) * (64 - c) - 64 * np.abs(d - nd // 2) - 32 * tau + hv
    return (num.astype(np.float64) / 4096.0).astype(np.float32)


def synthetic_store_arrays(n_nodes=1, seed=0, s_nodes=None, kind="smooth", shape=None, max_lag=43):
    """metadata and tables of a synthetic store: dict(s, beta, delta, tau_idx, alpha, tables) with
    tables a list of (ntau, nphif, nb, nd) float32 arrays.  shape overrides the table geometry
    (tiny stores for tests keep the beta/Delta steps of the real grids)."""
    beta, delta, _ = store_grids()
    tau = lag_indices(max_lag)
    if shape is not None:
        ntau, nphif, nb, nd = shape
        assert nphif == NPHIF
        tau = lag_indices((ntau - 1) // 2)
        beta = beta[0] + (beta[1] - beta[0]) * np.arange(nb)
        dd = delta[1] - delta[0]
        delta = dd * (np.arange(nd) - nd // 2)
    shp = (tau.size, NPHIF, beta.size, delta.size)
    if s_nodes is None:
        s_nodes = S_NODES[-n_nodes:] if n_nodes < S_NODES.size else S_NODES
    s_nodes = np.asarray(s_nodes, float)
    tables = [table_values(seed, k, shp, tau, kind).reshape(shp) for k in range(s_nodes.size)]
    return dict(s=s_nodes, beta=beta, delta=delta, tau_idx=tau, alpha=ALPHA.copy(), tables=tables)


def write_store(store_dir, arrays, n_max=120):
    """write a store as uncompressed bd_n{n_max:03d}_s{s:09.3f}.npz files, one per node, with the members of
    scripts/build_tables/build_bd.py:83-84"""
    os.makedirs(store_dir, exist_ok=True)
    paths = []
    for s, T in zip(arrays["s"], arrays["tables"]):
        p = os.path.join(store_dir, "bd_n%03d_s%09.3f.npz" % (n_max, s))
        np.savez(p, T=T, beta=arrays["beta"], delta=arrays["delta"], tau_idx=arrays["tau_idx"],
                 lnq=np.zeros(1), alpha=arrays["alpha"], s_node=float(s), u_node=float(s), n_max=n_max,
                 nphif=T.shape[1], norm="taylor_off1")
        paths.append(p)
    return paths
