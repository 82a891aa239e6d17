"""
Luminosity distance of the Planck 2018 cosmology, the D_L(z) that pe_redshift
converts with (src/exmarg/search/pe.py:26-28 calls astropy's Planck18).

Flat LCDM with H0 = 67.66 km/s/Mpc, Om0 = 0.30966, Tcmb0 = 2.7255 K, Neff = 3.046 and one
massive neutrino of 0.06 eV (Planck Collaboration 2020, A&A 641 A6, Table 2, last column).
Photons from the Stefan-Boltzmann law; neutrinos by the fitting function of Komatsu et al.
2011 (ApJS 192 18, Eq. 26), rho_nu / rho_gamma = 7/8 (4/11)^{4/3} (Neff/3)
sum_i [1 + (0.3173 m_i / (k T_nu (1+z)))^1.83]^{1/1.83}.  The comoving distance integral
is Gauss-Legendre quadrature on a tabulated cumulative grid.
"""

import numpy as np

H0, OM0, TCMB0, NEFF, MNU = 67.66, 0.30966, 2.7255, 3.046, (0.0, 0.0, 0.06)
C_KMS = 299792.458


def _densities():
    sigma_sb, c, G = 5.670374419e-8, 299792458.0, 6.67430e-11
    h0 = H0 * 1e3 / 3.0856775814913673e22
    og0 = 4.0 * sigma_sb * TCMB0 ** 4 / c ** 3 / (3.0 * h0 ** 2 / (8.0 * np.pi * G))
    tnu0 = (4.0 / 11.0) ** (1.0 / 3.0) * TCMB0
    ymass = np.array([m for m in MNU if m > 0]) / (8.617333262e-5 * tnu0)
    nmassless = sum(1 for m in MNU if m == 0)
    return og0, ymass, nmassless


_OG0, _YMASS, _NMASSLESS = _densities()
_PREFAC = 7.0 / 8.0 * (4.0 / 11.0) ** (4.0 / 3.0) * NEFF / 3.0


def _nu_rel(z):
    per = (1.0 + (0.3173 * _YMASS[None, :] / (1.0 + z[:, None])) ** 1.83) ** (1.0 / 1.83)
    return _PREFAC * (per.sum(axis=1) + _NMASSLESS)


_ONU0 = _OG0 * float(_nu_rel(np.zeros(1))[0])
_ODE0 = 1.0 - OM0 - _OG0 - _ONU0


def efunc(z):
    """H(z) / H0"""
    z = np.atleast_1d(np.asarray(z, float))
    zp1 = 1.0 + z
    return np.sqrt(OM0 * zp1 ** 3 + _OG0 * zp1 ** 4 * (1.0 + _nu_rel(z.ravel()).reshape(z.shape)) + _ODE0)


def luminosity_distance(z):
    """D_L(z) in Mpc; the comoving integral by 48-point Gauss-Legendre quadrature on each interval between
    consecutive sorted redshifts"""
    z = np.atleast_1d(np.asarray(z, float))
    order = np.argsort(z)
    zs = np.concatenate([[0.0], z[order]])
    x, w = np.polynomial.legendre.leggauss(48)
    a, b = zs[:-1], zs[1:]
    mid, half = 0.5 * (a + b), 0.5 * (b - a)
    nodes = mid[:, None] + half[:, None] * x[None, :]
    seg = half * (w[None, :] / efunc(nodes.ravel()).reshape(nodes.shape)).sum(axis=1)
    dc = np.empty_like(z)
    dc[order] = np.cumsum(seg)
    return (1.0 + z) * C_KMS / H0 * dc
