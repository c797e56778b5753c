"""Synthetic event catalogues for the path's workloads.

Draws follow the event definitions of the benchmark workloads (binary black
holes: source-frame m1, m2 ~ U[5, 80] Msun, z ~ U[0.1, 2]; binary neutron
stars: m ~ U[1.1, 2.5] Msun, z ~ U[0.02, 0.5], Lambda ~ U[0, 2000]), with
isotropic sky position and inclination as in the reference tutorial
(notebooks/gwfast_tutorial.ipynb:569-589) and parameter names as gwfast's
``events`` dictionaries (docs/source/create_data.rst:29-80).

Luminosity distance uses flat LambdaCDM (H0 = 67.7 km/s/Mpc, Om = 0.31) by
Gauss-Legendre quadrature of 1/E(z) (astropy, which the reference's drivers
use for this, is not part of this package).
"""
import numpy as np

H0 = 67.7
OM0 = 0.31
C_KMS = 2.99792458e5


def dL_Gpc(z, H0=H0, Om0=OM0, n_nodes=64):
    """Luminosity distance [Gpc] in flat LambdaCDM:
    dL = (1 + z) (c/H0) int_0^z dz' / sqrt(Om (1+z')^3 + 1 - Om)."""
    z = np.asarray(z, dtype=float)
    x, w = np.polynomial.legendre.leggauss(n_nodes)
    zz = 0.5 * z[..., None] * (x + 1.0)
    Ez = np.sqrt(Om0 * (1.0 + zz) ** 3 + (1.0 - Om0))
    chi = 0.5 * z * np.sum(w / Ez, axis=-1)
    return (1.0 + z) * (C_KMS / H0) * chi / 1000.0


def Mc_eta_from_m1m2(m1, m2):
    """gwfastUtils.py:466-479."""
    Mc = ((m1 * m2) ** (3. / 5.)) / ((m1 + m2) ** (1. / 5.))
    eta = (m1 * m2) / ((m1 + m2) * (m1 + m2))
    return Mc, eta


def draw_bbh(n, seed=0, m_range=(5.0, 80.0), z_range=(0.1, 2.0), chi_max=0.8):
    """BBH catalogue: returns a dict of (n,) float64 arrays keyed like gwfast
    (Mc, eta in the detector frame, dL in Gpc, angles in rad, tcoal in GMST
    days).  Draw order is fixed: m1, m2, z, phi, cos(theta), cos(iota), psi,
    tcoal, Phicoal, chi1z, chi2z."""
    rng = np.random.default_rng(seed)
    m1 = rng.uniform(*m_range, n)
    m2 = rng.uniform(*m_range, n)
    m1, m2 = np.maximum(m1, m2), np.minimum(m1, m2)
    z = rng.uniform(*z_range, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    theta = np.arccos(rng.uniform(-1.0, 1.0, n))
    iota = np.arccos(rng.uniform(-1.0, 1.0, n))
    psi = rng.uniform(0.0, np.pi, n)
    tcoal = rng.uniform(0.0, 1.0, n)
    Phicoal = rng.uniform(0.0, 2.0 * np.pi, n)
    chi1z = rng.uniform(-chi_max, chi_max, n)
    chi2z = rng.uniform(-chi_max, chi_max, n)
    Mc, eta = Mc_eta_from_m1m2(m1 * (1.0 + z), m2 * (1.0 + z))
    return {'Mc': Mc, 'eta': eta, 'dL': dL_Gpc(z), 'theta': theta, 'phi': phi, 'iota': iota, 'psi': psi,
            'tcoal': tcoal, 'Phicoal': Phicoal, 'chi1z': chi1z, 'chi2z': chi2z, 'z': z}


def draw_bns(n, seed=0, m_range=(1.1, 2.5), z_range=(0.02, 0.5), chi_max=0.05, lam_max=2000.0):
    """BNS catalogue with tidal deformabilities Lambda1, Lambda2."""
    rng = np.random.default_rng(seed)
    ev = draw_bbh(n, seed=seed, m_range=m_range, z_range=z_range, chi_max=chi_max)
    rng = np.random.default_rng(seed + 7919)
    ev['Lambda1'] = rng.uniform(0.0, lam_max, n)
    ev['Lambda2'] = rng.uniform(0.0, lam_max, n)
    return ev
