"""Parameter conversions and input checks used on the path (host side).

Mirrors gwfast/gwfastUtils.py:151-175 (sky angles), 398-444 (tidal
combinations), 450-479 (masses) and 965-1019 (check_evparams).  These act on
(N_ev,) parameter vectors before they are copied to the device.
"""
import numpy as np


def ra_dec_from_th_phi_rad(theta, phi):
    """gwfastUtils.py:151-164: ra = phi, dec = pi/2 - theta."""
    return phi, 0.5 * np.pi - theta


def th_phi_from_ra_dec_rad(ra, dec):
    return 0.5 * np.pi - dec, ra


def _seta(eta):
    eta = np.asarray(eta, dtype=np.float64)
    return np.sqrt(np.where(eta < 0.25, 1.0 - 4.0 * eta, 0.))


def Lamt_delLam_from_Lam12(Lambda1, Lambda2, eta):
    """gwfastUtils.py:398-417 (arXiv:1402.5156 eq. 5-6)."""
    eta2 = eta * eta
    Seta = _seta(eta)
    Lamt = (8. / 13.) * ((1. + 7. * eta - 31. * eta2) * (Lambda1 + Lambda2)
                         + Seta * (1. + 9. * eta - 11. * eta2) * (Lambda1 - Lambda2))
    delLam = 0.5 * (Seta * (1. - 13272. / 1319. * eta + 8944. / 1319. * eta2) * (Lambda1 + Lambda2)
                    + (1. - 15910. / 1319. * eta + 32850. / 1319. * eta2 + 3380. / 1319. * eta2 * eta) * (Lambda1 - Lambda2))
    return Lamt, delLam


def Lam12_from_Lamt_delLam(Lamt, delLam, eta):
    """gwfastUtils.py:419-444 (inverse of the above)."""
    eta2 = eta * eta
    Seta = _seta(eta)
    mLp = (8. / 13.) * (1. + 7. * eta - 31. * eta2)
    mLm = (8. / 13.) * Seta * (1. + 9. * eta - 11. * eta2)
    mdp = Seta * (1. - (13272. / 1319.) * eta + (8944. / 1319.) * eta2) * 0.5
    mdm = (1. - (15910. / 1319.) * eta + (32850. / 1319.) * eta2 + (3380. / 1319.) * (eta2 * eta)) * 0.5
    det = (306656. / 1319.) * (eta ** 5) - (5936. / 1319.) * (eta ** 4)
    return (((mdp - mdm) * Lamt + (mLm - mLp) * delLam) / det,
            ((-mdm - mdp) * Lamt + (mLm + mLp) * delLam) / det)


def m1m2_from_Mceta(Mc, eta):
    """gwfastUtils.py:450-464."""
    Seta = _seta(eta)
    m1 = 0.5 * (Mc / (eta ** (3. / 5.))) * (1. + Seta)
    m2 = 0.5 * (Mc / (eta ** (3. / 5.))) * (1. - Seta)
    return m1, m2


def Mceta_from_m1m2(m1, m2):
    """gwfastUtils.py:466-479."""
    Mc = ((m1 * m2) ** (3. / 5.)) / ((m1 + m2) ** (1. / 5.))
    eta = (m1 * m2) / ((m1 + m2) * (m1 + m2))
    return Mc, eta


def check_evparams(evParams):
    """gwfastUtils.py:965-1019: completes the dictionary in place (iota from
    thetaJN, Mc/eta from m1/m2, theta/phi from ra/dec) and raises ValueError
    with the reference's messages when a required pair is missing.  GPS-time
    input (tGPS) needs astropy in the reference and is not supported here."""
    if 'tcoal' not in evParams:
        if 'tGPS' in evParams:
            raise ValueError('tGPS input is not supported: provide tcoal as GMST days.')
        raise ValueError('One among tGPS and tcoal has to be provided.')
    if 'iota' not in evParams:
        try:
            evParams['iota'] = evParams['thetaJN']
        except KeyError:
            raise ValueError('One among iota and thetaJN has to be provided.')
    if 'Mc' not in evParams:
        try:
            evParams['Mc'], evParams['eta'] = Mceta_from_m1m2(evParams['m1'], evParams['m2'])
        except KeyError:
            raise ValueError('Two among (Mc, eta) and (m1, m2) have to be provided.')
    if 'theta' not in evParams:
        try:
            evParams['theta'] = np.pi / 2 - evParams['dec']
            evParams['phi'] = evParams['ra']
        except KeyError:
            raise ValueError('Two among (theta, phi) and (ra, dec) have to be provided.')
    return evParams
