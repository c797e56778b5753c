"""Oracle restatement of the two waveform models on the path (test infrastructure).

Follows gwfast/waveforms.py:
  * TaylorF2_RestrictedPN: Phi 743-862, Ampl 864-876, tau_star 878-901,
    fcut 903-953 (constructor 726-741);
  * IMRPhenomD: constructor 973-990, Phi 992-1153, Ampl 1155-1254,
    _finalspin 1256-1276, _radiatednrg 1278-1297, tau_star 1299-1322,
    fcut 1324-1333.
Expressions keep the reference's association (same operand order) so that
the FP64 rounding matches it as closely as NumPy allows.  Inputs may be
ndarrays, complex ndarrays (complex-step checks) or ``Dual`` numbers.
"""
import os

import numpy as np

from . import dual as D
from gwfast_amd import gwfastGlobals as glob

EULER_GAMMA = np.euler_gamma

# IMRPhenomD phenomenological fits.  Each entry holds the 11 numbers
# (a0, a1, b0, b1, b2, c0, c1, c2, d0, d1, d2) of
#   a0 + a1*eta + (b0 + b1*eta + b2*eta2 + (c0 + c1*eta + c2*eta2)*xi
#                  + (d0 + d1*eta + d2*eta2)*xi*xi)*xi
# (waveforms.py:1039-1056, 1189-1191, 1195-1197, 1227).
PHENOMD_FITS = {
    'sigma1': (2096.551999295543, 1463.7493168261553, 1312.5493286098522, 18307.330017082117, -43534.1440746107, -833.2889543511114, 32047.31997183187, -108609.45037520859, 452.25136398112204, 8353.439546391714, -44531.3250037322),
    'sigma2': (-10114.056472621156, -44631.01109458185, -6541.308761668722, -266959.23419307504, 686328.3229317984, 3405.6372187679685, -437507.7208209015, 1631817.1307344697, -7462.648563007646, -114585.25177153319, 674402.4689098676),
    'sigma3': (22933.658273436497, 230960.00814979506, 14961.083974183695, 1194018.1342318142, -3104223.9693052764, -3038.166617199259, 1872032.2849093592, -7309145.012085539, 42738.22871475411, 467502.018616601, -3064853.498512499),
    'sigma4': (-14621.71522218357, -377812.8579387104, -9608.682631509726, -1710892.5257214056, 4332924.601416521, -22366.683262266528, -2501971.6386377467, 10274495.902259542, -85360.30079034246, -570025.3441737515, 4396844.346849777),
    'beta1': (97.89747327985583, -42.659730877489224, 153.48421037904913, -1417.0620760768954, 2752.8614143665027, 138.7406469558649, -1433.6585075135881, 2857.7418952430758, 41.025109467376126, -423.680737974639, 850.3594335657173),
    'beta2': (-3.282701958759534, -9.051384468245866, -12.415449742258042, 55.4716447709787, -106.05109938966335, -11.953044553690658, 76.80704618365418, -155.33172948098394, -3.4129261592393263, 25.572377569952536, -54.408036707740465),
    'beta3': (-2.5156429818799565e-05, 1.9750256942201327e-05, -1.8370671469295915e-05, 2.1886317041311973e-05, 8.250240316860033e-05, 7.157371250566708e-06, -5.5780000112270685e-05, 0.00019142082884072178, 5.447166261464217e-06, -3.220610095021982e-05, 7.974016714984341e-05),
    'alpha1': (43.31514709695348, 638.6332679188081, -32.85768747216059, 2415.8938269370315, -5766.875169379177, -61.85459307173841, 2953.967762459948, -8986.29057591497, -21.571435779762044, 981.2158224673428, -3239.5664895930286),
    'alpha2': (-0.07020209449091723, -0.16269798450687084, -0.1872514685185499, 1.138313650449945, -2.8334196304430046, -0.17137955686840617, 1.7197549338119527, -4.539717148261272, -0.049983437357548705, 0.6062072055948309, -1.682769616644546),
    'alpha3': (9.5988072383479, -397.05438595557433, 16.202126189517813, -1574.8286986717037, 3600.3410843831093, 27.092429659075467, -1786.482357315139, 5152.919378666511, 11.175710130033895, -577.7999423177481, 1808.730762932043),
    'alpha4': (-0.02989487384493607, 1.4022106448583738, -0.07356049468633846, 0.8337006542278661, 0.2240008282397391, -0.055202870001177226, 0.5667186343606578, 0.7186931973380503, -0.015507437354325743, 0.15750322779277187, 0.21076815715176228),
    'alpha5': (0.9974408278363099, -0.007884449714907203, -0.059046901195591035, 1.3958712396764088, -4.516631601676276, -0.05585343136869692, 1.7516580039343603, -5.990208965347804, -0.017945336522161195, 0.5965097794825992, -2.0608879367971804),
    'gamma1': (0.006927402739328343, 0.03020474290328911, 0.006308024337706171, -0.12074130661131138, 0.26271598905781324, 0.0034151773647198794, -0.10779338611188374, 0.27098966966891747, 0.0007374185938559283, -0.02749621038376281, 0.0733150789135702),
    'gamma2': (1.010344404799477, 0.0008993122007234548, 0.283949116804459, -4.049752962958005, 13.207828172665366, 0.10396278486805426, -7.025059158961947, 24.784892370130475, 0.03093202475605892, -2.6924023896851663, 9.609374464684983),
    'gamma3': (1.3081615607036106, -0.005537729694807678, -0.06782917938621007, -0.6689834970767117, 3.403147966134083, -0.05296577374411866, -0.9923793203111362, 4.820681208409587, -0.006134139870393713, -0.38429253308696365, 1.7561754421985984),
    'rho1': (3931.8979897196696, -17395.758706812805, 3132.375545898835, 343965.86092361377, -1216256.5819981997, -70698.00600428853, 1383907.177859705, -3966276.1890979446, -60017.52423652596, 803515.1181825735, -2091710.365941658),
    'rho2': (-40105.47653771657, 112253.0169706701, 23561.696065836168, -3476180.699403351, 11375936.70849482, 754313.1127166454, -13084760.44625268, 36444584.853928134, 596226.612472288, -7427790.1143564405, 18928977.514040343),
    'rho3': (83208.35471266537, -191237.7264145924, -210916.2454782992, 8717975.08352568, -26914942.420669552, -1988980.6527362722, 30888029.960154563, -83908702.79256162, -1453503.1953446497, 17063528.990822166, -42748659.731120914),
    'v2': (0.8149838730507785, 2.5747553517454658, 1.1610198035496786, -2.3627771785551537, 6.771038707057573, 0.7570782938606834, -2.7256896890432474, 7.1140380397149965, 0.1766934149293479, -0.7978690983168183, 2.1162391502005153),
}


def _fit(name, eta, eta2, xi):
    a0, a1, b0, b1, b2, c0, c1, c2, d0, d1, d2 = PHENOMD_FITS[name]
    return a0 + a1 * eta + (b0 + b1 * eta + b2 * eta2 + (c0 + c1 * eta + c2 * eta2) * xi
                            + (d0 + d1 * eta + d2 * eta2) * xi * xi) * xi


def _seta(eta):
    # sqrt(where(eta < 0.25, 1 - 4 eta, 0)): waveforms.py:759, 1013, 1172
    return D.sqrt(D.where(np.real(D.value(eta)) < 0.25, 1.0 - 4.0 * eta, 0.))


def lamt_dellam_from_lam12(L1, L2, eta):
    """gwfastUtils.py:398-417."""
    eta2 = eta * eta
    Seta = _seta(eta)
    Lamt = (8. / 13.) * ((1. + 7. * eta - 31. * eta2) * (L1 + L2) + Seta * (1. + 9. * eta - 11. * eta2) * (L1 - L2))
    delLam = 0.5 * (Seta * (1. - 13272. / 1319. * eta + 8944. / 1319. * eta2) * (L1 + L2)
                    + (1. - 15910. / 1319. * eta + 32850. / 1319. * eta2 + 3380. / 1319. * eta2 * eta) * (L1 - L2))
    return Lamt, delLam


def lam12_from_lamt_dellam(Lamt, delLam, eta):
    """gwfastUtils.py:419-444."""
    eta2 = eta * eta
    Seta = _seta(eta)
    mLp = (8. / 13.) * (1. + 7. * eta - 31. * eta2)
    mLm = (8. / 13.) * Seta * (1. + 9. * eta - 11. * eta2)
    mdp = Seta * (1. - (13272. / 1319.) * eta + (8944. / 1319.) * eta2) * 0.5
    mdm = (1. - (15910. / 1319.) * eta + (32850. / 1319.) * eta2 + (3380. / 1319.) * (eta2 * eta)) * 0.5
    det = (306656. / 1319.) * (eta ** 5) - (5936. / 1319.) * (eta ** 4)
    L1 = ((mdp - mdm) * Lamt + (mLm - mLp) * delLam) / det
    L2 = ((-mdm - mdp) * Lamt + (mLm + mLp) * delLam) / det
    return L1, L2


def tau_star_pn(f, ev):
    """3.5PN time to coalescence; identical in TaylorF2 (waveforms.py:878-901)
    and IMRPhenomD (1299-1322)."""
    Mtot_sec = ev['Mc'] * glob.GMsun_over_c3 / (ev['eta'] ** (3. / 5.))
    v = (np.pi * Mtot_sec * f) ** (1. / 3.)
    eta = ev['eta']
    eta2 = eta * eta
    OverallFac = 5. / 256 * Mtot_sec / (eta * (v ** 8.))
    t05 = (1. + (743. / 252. + 11. / 3. * eta) * (v * v) - 32. / 5. * np.pi * (v * v * v)
           + (3058673. / 508032. + 5429. / 504. * eta + 617. / 72. * eta2) * (v ** 4)
           - (7729. / 252. - 13. / 3. * eta) * np.pi * (v ** 5))
    t6 = (-10052469856691. / 23471078400. + 128. / 3. * np.pi * np.pi + 6848. / 105. * EULER_GAMMA
          + (3147553127. / 3048192. - 451. / 12. * np.pi * np.pi) * eta - 15211. / 1728. * eta2
          + 25565. / 1296. * eta2 * eta + 3424. / 105. * D.log(16. * v * v)) * (v ** 6)
    t7 = (-15419335. / 127008. - 75703. / 756. * eta + 14809. / 378. * eta2) * np.pi * (v ** 7)
    return OverallFac * (t05 + t6 + t7)


def _tf2_coeffs(eta, Seta, chi1, chi2, QuadMon1, QuadMon2, use_3p5PN_SpinHO):
    """TaylorF2 PN coefficients: waveforms.py:784-812 (IMRPhenomD repeats them
    at 1062-1077 with QuadMon = 1)."""
    eta2 = eta * eta
    m1ByM = 0.5 * (1.0 + Seta)
    m2ByM = 0.5 * (1.0 - Seta)
    chi12, chi22 = chi1 * chi1, chi2 * chi2
    chi1dotchi2 = chi1 * chi2
    chi_s, chi_a = 0.5 * (chi1 + chi2), 0.5 * (chi1 - chi2)
    chi_s2, chi_a2 = chi_s * chi_s, chi_a * chi_a
    c = {}
    c['two'] = 3715. / 756. + (55. * eta) / 9.
    c['three'] = -16. * np.pi + (113. * Seta * chi_a) / 3. + (113. / 3. - (76. * eta) / 3.) * chi_s
    c['four'] = (5. * (3058.673 / 7.056 + 5429. / 7. * eta + 617. * eta2) / 72. + 247. / 4.8 * eta * chi1dotchi2
                 - 721. / 4.8 * eta * chi1dotchi2 + (-720. / 9.6 * QuadMon1 + 1. / 9.6) * m1ByM * m1ByM * chi12
                 + (-720. / 9.6 * QuadMon2 + 1. / 9.6) * m2ByM * m2ByM * chi22
                 + (240. / 9.6 * QuadMon1 - 7. / 9.6) * m1ByM * m1ByM * chi12
                 + (240. / 9.6 * QuadMon2 - 7. / 9.6) * m2ByM * m2ByM * chi22)
    c['five_tmp'] = ((732985. / 2268. - 24260. * eta / 81. - 340. * eta2 / 9.) * chi_s
                     + (732985. / 2268. + 140. * eta / 9.) * Seta * chi_a)
    c['six'] = (11583.231236531 / 4.694215680 - 640. / 3. * np.pi * np.pi - 684.8 / 2.1 * EULER_GAMMA
                + eta * (-15737.765635 / 3.048192 + 225.5 / 1.2 * np.pi * np.pi) + eta2 * 76.055 / 1.728
                - eta2 * eta * 127.825 / 1.296 - np.log(4.) * 684.8 / 2.1
                + np.pi * chi1 * m1ByM * (1490. / 3. + m1ByM * 260.) + np.pi * chi2 * m2ByM * (1490. / 3. + m2ByM * 260.)
                + (326.75 / 1.12 + 557.5 / 1.8 * eta) * eta * chi1dotchi2
                + (4703.5 / 8.4 + 2935. / 6. * m1ByM - 120. * m1ByM * m1ByM) * m1ByM * m1ByM * QuadMon1 * chi12
                + (-4108.25 / 6.72 - 108.5 / 1.2 * m1ByM + 125.5 / 3.6 * m1ByM * m1ByM) * m1ByM * m1ByM * chi12
                + (4703.5 / 8.4 + 2935. / 6. * m2ByM - 120. * m2ByM * m2ByM) * m2ByM * m2ByM * QuadMon2 * chi22
                + (-4108.25 / 6.72 - 108.5 / 1.2 * m2ByM + 125.5 / 3.6 * m2ByM * m2ByM) * m2ByM * m2ByM * chi22)
    c['six_log'] = -(6848. / 21.)
    if use_3p5PN_SpinHO:
        c['seven'] = (77096675. * np.pi / 254016. + 378515. * np.pi * eta / 1512. - 74045. * np.pi * eta2 / 756.
                      + (-25150083775. / 3048192. + 10566655595. * eta / 762048. - 1042165. * eta2 / 3024.
                         + 5345. * eta2 * eta / 36. + (14585. / 8. - 7270. * eta + 80. * eta2) * chi_a2) * chi_s
                      + (14585. / 24. - 475. * eta / 6. + 100. * eta2 / 3.) * chi_s2 * chi_s
                      + Seta * ((-25150083775. / 3048192. + 26804935. * eta / 6048. - 1985. * eta2 / 48.) * chi_a
                                + (14585. / 24. - 2380. * eta) * chi_a2 * chi_a + (14585. / 8. - 215. * eta / 2.) * chi_a * chi_s2))
    else:
        c['seven'] = (77096675. * np.pi / 254016. + 378515. * np.pi * eta / 1512. - 74045. * np.pi * eta2 / 756.
                      + (-25150083775. / 3048192. + 10566655595. * eta / 762048. - 1042165. * eta2 / 3024.
                         + 5345. * eta2 * eta / 36.) * chi_s
                      + Seta * ((-25150083775. / 3048192. + 26804935. * eta / 6048. - 1985. * eta2 / 48.) * chi_a))
    c['m1ByM'], c['m2ByM'] = m1ByM, m2ByM
    c['chi12'], c['chi22'], c['chi1dotchi2'] = chi12, chi22, chi1dotchi2
    return c


class TaylorF2:
    """Configuration of TaylorF2_RestrictedPN (waveforms.py:726-741)."""

    name = 'TaylorF2_RestrictedPN'

    def __init__(self, fHigh=None, is_tidal=False, use_3p5PN_SpinHO=False, phiref_vlso=False,
                 which_ISCO='Schw', use_QuadMonTid=False, is_eccentric=False, fRef_ecc=None):
        if fHigh is None:
            fHigh = 1. / (6. * np.pi * np.sqrt(6.) * glob.GMsun_over_c3)
        self.fcutPar = fHigh
        self.is_tidal = is_tidal
        self.is_eccentric = is_eccentric
        self.fRef_ecc = fRef_ecc
        self.use_3p5PN_SpinHO = use_3p5PN_SpinHO
        self.phiref_vlso = phiref_vlso
        self.which_ISCO = which_ISCO
        self.use_QuadMonTid = use_QuadMonTid
        self.is_holomorphic = True
        self.objType = 'BNS' if is_tidal else 'BBH'
        self.nParams = 11 + (2 if is_tidal else 0) + (1 if is_eccentric else 0)

    def phi_ecc(self, v, Mtot_sec, eta, ecc):
        """waveforms.py:814-845: the low-eccentricity phase to 3PN
        (arXiv:1605.00304), added to the bracket of the TaylorF2 phase; v0 at
        fRef_ecc or the grid minimum of v."""
        v0 = D.amin0(v) if self.fRef_ecc is None else (np.pi * Mtot_sec * self.fRef_ecc) ** (1. / 3.)
        eta2 = eta * eta
        eg = np.euler_gamma
        k = {}
        k['twoV'] = 29.9076223 / 8.1976608 + 18.766963 / 2.927736 * eta
        k['twoV0'] = 2.833 / 1.008 - 19.7 / 3.6 * eta
        k['threeV'] = -28.19123 / 2.82600 * np.pi
        k['threeV0'] = 37.7 / 7.2 * np.pi
        k['fourV4'] = 16.237683263 / 3.330429696 + 241.33060753 / 9.71375328 * eta + 156.2608261 / 6.9383952 * eta2
        k['fourV2V02'] = 84.7282939759 / 8.2632420864 - 7.18901219 / 3.68894736 * eta - 36.97091711 / 1.05398496 * eta2
        k['fourV04'] = -1.193251 / 3.048192 - 66.317 / 9.072 * eta + 18.155 / 1.296 * eta2
        k['fiveV5'] = -28.31492681 / 1.18395270 * np.pi - 115.52066831 / 2.70617760 * np.pi * eta
        k['fiveV3V02'] = -79.86575459 / 2.84860800 * np.pi + 55.5367231 / 1.0173600 * np.pi * eta
        k['fiveV2V03'] = 112.751736071 / 5.902315776 * np.pi + 70.75145051 / 2.10796992 * np.pi * eta
        k['fiveV05'] = 76.4881 / 9.0720 * np.pi - 94.9457 / 2.2680 * np.pi * eta
        k['sixV6'] = (-436.03153867072577087 / 1.32658535116800000 + 53.6803271 / 1.9782000 * eg
                      + 157.22503703 / 3.25555200 * np.pi * np.pi
                      + (2991.72861614477 / 6.89135247360 - 15.075413 / 1.446912 * np.pi * np.pi) * eta
                      + 345.5209264991 / 4.1019955200 * eta2 + 506.12671711 / 8.78999040 * eta2 * eta
                      + 384.3505163 / 5.9346000 * np.log(2.) - 112.1397129 / 1.7584000 * np.log(3.))
        k['sixV4V02'] = (46.001356684079 / 3.357073133568 + 253.471410141755 / 5.874877983744 * eta
                         - 169.3852244423 / 2.3313007872 * eta2 - 307.833827417 / 2.497822272 * eta2 * eta)
        k['sixV3V03'] = -106.2809371 / 2.0347200 * np.pi * np.pi
        k['sixV2V04'] = (-3.56873002170973 / 2.49880440692736 - 260.399751935005 / 8.924301453312 * eta
                         + 15.0484695827 / 3.5413894656 * eta2 + 340.714213265 / 3.794345856 * eta2 * eta)
        k['sixV06'] = (265.31900578691 / 1.68991764480 - 33.17 / 1.26 * eg + 12.2833 / 1.0368 * np.pi * np.pi
                       + (91.55185261 / 5.48674560 - 3.977 / 1.152 * np.pi * np.pi) * eta - 5.732473 / 1.306368 * eta2
                       - 30.90307 / 1.39968 * eta2 * eta + 87.419 / 1.890 * np.log(2.) - 260.01 / 5.60 * np.log(3.))
        amp = -2.355 / 1.462 * ecc * ecc * ((v0 / v) ** (19. / 3.))
        return amp * (1. + (k['twoV'] * v * v + k['twoV0'] * v0 * v0)
                      + (k['threeV'] * v ** 3 + k['threeV0'] * v0 ** 3)
                      + (k['fourV4'] * v ** 4 + k['fourV2V02'] * v * v * v0 * v0 + k['fourV04'] * v0 ** 4)
                      + (k['fiveV5'] * v ** 5 + k['fiveV3V02'] * v ** 3 * v0 * v0 + k['fiveV2V03'] * v * v * v0 ** 3
                         + k['fiveV05'] * v0 ** 5)
                      + ((k['sixV6'] + 53.6803271 / 3.9564000 * D.log(16. * v * v)) * v ** 6
                         + k['sixV4V02'] * v ** 4 * v0 * v0 + k['sixV3V03'] * v ** 3 * v0 ** 3
                         + k['sixV2V04'] * v * v * v0 ** 4
                         + (k['sixV06'] - 33.17 / 2.52 * D.log(16. * v0 * v0)) * v0 ** 6))

    def Phi(self, f, ev):
        """waveforms.py:743-862."""
        Mtot_sec = ev['Mc'] * glob.GMsun_over_c3 / (ev['eta'] ** (3. / 5.))
        v = (np.pi * Mtot_sec * f) ** (1. / 3.)
        eta = ev['eta']
        Seta = _seta(eta)
        chi1, chi2 = ev['chi1z'], ev['chi2z']
        vlso = 1. / np.sqrt(6.)
        if self.is_tidal and self.use_QuadMonTid:
            QuadMon1 = _quadmon(ev['Lambda1'])
            QuadMon2 = _quadmon(ev['Lambda2'])
        else:
            QuadMon1, QuadMon2 = 1.0, 1.0
        c = _tf2_coeffs(eta, Seta, chi1, chi2, QuadMon1, QuadMon2, self.use_3p5PN_SpinHO)
        TF2OverallAmpl = 3. / (128. * eta)
        base5 = (38645. * np.pi / 756. - 65. * np.pi * eta / 9. - c['five_tmp'])
        if self.phiref_vlso:
            five = base5 * (1. - 3. * np.log(vlso))
            phiR = 0.
        else:
            five = base5
            phiR = np.pi
        five_log = base5 * 3.
        if self.is_tidal:
            Lam_t, delLam = lamt_dellam_from_lam12(ev['Lambda1'], ev['Lambda2'], eta)
            phi_Tidal = (-0.5 * 39. * Lam_t) * (v ** 10.) + (-3115. / 64. * Lam_t + 6595. / 364. * Seta * delLam) * (v ** 12.)
        else:
            phi_Tidal = 0.
        phi_Ecc = self.phi_ecc(v, Mtot_sec, eta, ev['ecc']) if self.is_eccentric else 0.
        phase = TF2OverallAmpl * (1. + 0. * v + c['two'] * v * v + c['three'] * v ** 3 + c['four'] * v ** 4
                                  + (five + five_log * D.log(v)) * v ** 5
                                  + (c['six'] + c['six_log'] * D.log(v)) * v ** 6
                                  + c['seven'] * v ** 7 + phi_Tidal + phi_Ecc) / (v ** 5.)
        return phase + phiR - np.pi * 0.25

    def Ampl(self, f, ev):
        """waveforms.py:864-876."""
        return (np.sqrt(5. / 24.) * (np.pi ** (-2. / 3.)) * glob.clightGpc / ev['dL']
                * (glob.GMsun_over_c3 * ev['Mc']) ** (5. / 6.) * (f ** (-7. / 6.)))

    def tau_star(self, f, ev):
        return tau_star_pn(f, ev)

    def fcut(self, ev):
        """waveforms.py:903-953."""
        if self.which_ISCO == 'Schw':
            return self.fcutPar / (ev['Mc'] / (ev['eta'] ** (3. / 5.)))
        eta = ev['eta']
        eta2 = eta * eta
        Mtot = ev['Mc'] / (eta ** (3. / 5.))
        chi1, chi2 = ev['chi1z'], ev['chi2z']
        Seta = _seta(eta)
        m1 = 0.5 * (1.0 + Seta)
        m2 = 0.5 * (1.0 - Seta)
        s = (m1 * m1 * chi1 + m2 * m2 * chi2) / (m1 * m1 + m2 * m2)
        atot = (chi1 + chi2 * (m2 / m1) * (m2 / m1)) / ((1. + m2 / m1) * (1. + m2 / m1))
        aeff = atot + 0.41616 * eta * (chi1 + chi2)

        def r_isco(chi):
            Z1 = 1.0 + ((1.0 - chi * chi) ** (1. / 3.)) * ((1.0 + chi) ** (1. / 3.) + (1.0 - chi) ** (1. / 3.))
            Z2 = np.sqrt(3.0 * chi * chi + Z1 * Z1)
            return np.where(chi > 0., 3.0 + Z2 - np.sqrt((3.0 - Z1) * (3.0 + Z1 + 2.0 * Z2)),
                            3.0 + Z2 + np.sqrt((3.0 - Z1) * (3.0 + Z1 + 2.0 * Z2)))

        r_ISCO = r_isco(aeff)
        EradNS = eta * (0.055974469826360077 + 0.5809510763115132 * eta - 0.9606726679372312 * eta2 + 3.352411249771192 * eta2 * eta)
        EradTot = ((EradNS * (1. + (-0.0030302335878845507 - 2.0066110851351073 * eta + 7.7050567802399215 * eta2) * s))
                   / (1. + (-0.6714403054720589 - 1.4756929437702908 * eta + 7.304676214885011 * eta2) * s))
        Mfin = Mtot * (1. - EradTot)
        L_ISCO = 2. / (3. * np.sqrt(3.)) * (1. + 2. * np.sqrt(3. * r_ISCO - 2.))
        E_ISCO = np.sqrt(1. - 2. / (3. * r_ISCO))
        chif = (atot + eta * (L_ISCO - 2. * atot * (E_ISCO - 1.)) + (-3.821158961 - 1.2019 * aeff - 1.20764 * aeff * aeff) * eta2
                + (3.79245 + 1.18385 * aeff + 4.90494 * aeff * aeff) * eta2 * eta)
        Om_ISCO = 1. / (((r_isco(chif)) ** (3. / 2.)) + chif)
        return Om_ISCO / (np.pi * Mfin * glob.GMsun_over_c3)


def _quadmon(L):
    """Spin-induced quadrupole from tidal deformability, waveforms.py:771-772."""
    lL = D.log(D.where(D.value(L) < 1., 1., L))
    return D.where(D.value(L) < 1., 1. + L * (0.427688866723244 + L * (-0.324336526985068 + L * 0.1107439432180572)),
                   D.exp(0.1940 + 0.09163 * lL + 0.04812 * lL * lL - 4.283e-3 * lL * lL * lL + 1.245e-4 * lL * lL * lL * lL))


def load_qnm_tables(path=None):
    path = path or glob.WFfilesPath
    return (np.loadtxt(os.path.join(path, 'QNMData_a.txt')),
            np.loadtxt(os.path.join(path, 'QNMData_fring.txt')),
            np.loadtxt(os.path.join(path, 'QNMData_fdamp.txt')))


class IMRPhenomD:
    """Configuration of IMRPhenomD (waveforms.py:973-990)."""

    name = 'IMRPhenomD'

    def __init__(self, fRef=None, apply_fcut=True):
        self.AMP_fJoin_INS = 0.014
        self.PHI_fJoin_INS = 0.018
        self.fcutPar = 0.2
        self.fRef = fRef
        self.apply_fcut = apply_fcut
        self.is_tidal = False
        self.is_holomorphic = False
        self.objType = 'BBH'
        self.nParams = 11
        self.QNMgrid_a, self.QNMgrid_fring, self.QNMgrid_fdamp = load_qnm_tables()

    # -- remnant fits --------------------------------------------------------
    def _finalspin(self, eta, chi1, chi2):
        """waveforms.py:1256-1276."""
        Seta = _seta(eta)
        m1 = 0.5 * (1.0 + Seta)
        m2 = 0.5 * (1.0 - Seta)
        s = (m1 * m1 * chi1 + m2 * m2 * chi2)
        af1 = eta * (3.4641016151377544 - 4.399247300629289 * eta + 9.397292189321194 * eta * eta - 13.180949901606242 * eta * eta * eta)
        af2 = eta * (s * ((1.0 / eta - 0.0850917821418767 - 5.837029316602263 * eta) + (0.1014665242971878 - 2.0967746996832157 * eta) * s))
        af3 = eta * (s * ((-1.3546806617824356 + 4.108962025369336 * eta) * s * s + (-0.8676969352555539 + 2.064046835273906 * eta) * s * s * s))
        return af1 + af2 + af3

    def _radiatednrg(self, eta, chi1, chi2):
        """waveforms.py:1278-1297."""
        Seta = _seta(eta)
        m1 = 0.5 * (1.0 + Seta)
        m2 = 0.5 * (1.0 - Seta)
        s = (m1 * m1 * chi1 + m2 * m2 * chi2) / (m1 * m1 + m2 * m2)
        EradNS = eta * (0.055974469826360077 + 0.5809510763115132 * eta - 0.9606726679372312 * eta * eta + 3.352411249771192 * eta * eta * eta)
        return ((EradNS * (1. + (-0.0030302335878845507 - 2.0066110851351073 * eta + 7.7050567802399215 * eta * eta) * s))
                / (1. + (-0.6714403054720589 - 1.4756929437702908 * eta + 7.304676214885011 * eta * eta) * s))

    def _ringdown(self, eta, chi1, chi2):
        aeff = self._finalspin(eta, chi1, chi2)
        Erad = self._radiatednrg(eta, chi1, chi2)
        fring = D.interp(D.real(aeff), self.QNMgrid_a, self.QNMgrid_fring) / (1.0 - Erad)
        fdamp = D.interp(D.real(aeff), self.QNMgrid_a, self.QNMgrid_fdamp) / (1.0 - Erad)
        return fring, fdamp

    def Phi(self, f, ev):
        """waveforms.py:992-1153."""
        co = self.phase_coeffs(f, ev)
        phis = self._phi_piecewise(co['fgrid'], co)
        extra = self._phase_extra(co['fgrid'], ev, co['eta'], co['Seta'], co['QuadMon1'], co['QuadMon2'],
                                  co['TF2OverallAmpl'])
        t0, fRef, phiRef = co['t0'], co['fRef'], co['phiRef']
        if self.apply_fcut:
            return phis + D.where(D.value(co['fgrid']) < self.fcutPar, -t0 * (co['fgrid'] - fRef) - phiRef + extra, 0.)
        return phis - t0 * (co['fgrid'] - fRef) - phiRef + extra

    def phase_coeffs(self, f, ev):
        """The per-event constants of the phase (waveforms.py:992-1143): the
        inspiral / intermediate / merger-ringdown coefficients, the C1 join
        constants, fpeak, the time shift t0, fRef (in Mf) and phiRef, plus
        fgrid = M G f."""
        M = ev['Mc'] / (ev['eta'] ** (3. / 5.))
        eta = ev['eta']
        eta2 = eta * eta
        chi1, chi2 = ev['chi1z'], ev['chi2z']
        Seta = _seta(eta)
        chi_s = 0.5 * (chi1 + chi2)
        chi_a = 0.5 * (chi1 - chi2)
        fgrid = M * glob.GMsun_over_c3 * f
        chiPN = (chi_s * (1.0 - eta * 76.0 / 113.0) + Seta * chi_a)
        xi = -1.0 + chiPN
        fring, fdamp = self._ringdown(eta, chi1, chi2)
        sigma1 = _fit('sigma1', eta, eta2, xi)
        sigma2 = _fit('sigma2', eta, eta2, xi)
        sigma3 = _fit('sigma3', eta, eta2, xi)
        sigma4 = _fit('sigma4', eta, eta2, xi)
        beta1 = _fit('beta1', eta, eta2, xi)
        beta2 = _fit('beta2', eta, eta2, xi)
        beta3 = _fit('beta3', eta, eta2, xi)
        alpha1 = _fit('alpha1', eta, eta2, xi)
        alpha2 = _fit('alpha2', eta, eta2, xi)
        alpha3 = _fit('alpha3', eta, eta2, xi)
        alpha4 = _fit('alpha4', eta, eta2, xi)
        alpha5 = _fit('alpha5', eta, eta2, xi)

        QuadMon1, QuadMon2 = self._quadmons(ev)
        c = _tf2_coeffs(eta, Seta, chi1, chi2, QuadMon1, QuadMon2, False)
        TF2OverallAmpl = 3. / (128. * eta)
        five = (38645. * np.pi / 756. - 65. * np.pi * eta / 9. - c['five_tmp'])
        five_log = five * 3.
        m1ByM, m2ByM = c['m1ByM'], c['m2ByM']
        chi12, chi22, chi1dotchi2 = c['chi12'], c['chi22'], c['chi1dotchi2']
        # 3PN spin-spin part removed (waveforms.py:1079)
        six = c['six'] - ((326.75 / 1.12 + 557.5 / 1.8 * eta) * eta * chi1dotchi2
                          + ((4703.5 / 8.4 + 2935. / 6. * m1ByM - 120. * m1ByM * m1ByM)
                             + (-4108.25 / 6.72 - 108.5 / 1.2 * m1ByM + 125.5 / 3.6 * m1ByM * m1ByM)) * m1ByM * m1ByM * chi12
                          + ((4703.5 / 8.4 + 2935. / 6. * m2ByM - 120. * m2ByM * m2ByM)
                             + (-4108.25 / 6.72 - 108.5 / 1.2 * m2ByM + 125.5 / 3.6 * m2ByM * m2ByM)) * m2ByM * m2ByM * chi22)
        six_log, seven, four, three, two = c['six_log'], c['seven'], c['four'], c['three'], c['two']
        one, zero = 0., 1.

        P = {}
        P['initial_phasing'] = five * TF2OverallAmpl
        P['two_thirds'] = seven * TF2OverallAmpl * (np.pi ** (2. / 3.))
        P['third'] = six * TF2OverallAmpl * (np.pi ** (1. / 3.))
        P['third_log'] = six_log * TF2OverallAmpl * (np.pi ** (1. / 3.))
        P['log'] = five_log * TF2OverallAmpl
        P['min_third'] = four * TF2OverallAmpl * (np.pi ** (-1. / 3.))
        P['min_two_thirds'] = three * TF2OverallAmpl * (np.pi ** (-2. / 3.))
        P['min_one'] = two * TF2OverallAmpl / np.pi
        P['min_four_thirds'] = one * TF2OverallAmpl * (np.pi ** (-4. / 3.))
        P['min_five_thirds'] = zero * TF2OverallAmpl * (np.pi ** (-5. / 3.))
        P['one'] = sigma1
        P['four_thirds'] = sigma2 * 0.75
        P['five_thirds'] = sigma3 * 0.6
        P['two'] = sigma4 * 0.5

        fInsJoin = self.PHI_fJoin_INS
        fMRDJoin = 0.5 * fring
        pf = np.pi * fInsJoin
        DPhiIns = ((2.0 * seven * TF2OverallAmpl * (pf ** (7. / 3.))
                    + (six * TF2OverallAmpl + six_log * TF2OverallAmpl * (1.0 + np.log(pf) / 3.)) * (pf ** (2.))
                    + five_log * TF2OverallAmpl * (pf ** (5. / 3.)) - four * TF2OverallAmpl * (pf ** (4. / 3.))
                    - 2. * three * TF2OverallAmpl * pf - 3. * two * TF2OverallAmpl * (pf ** (2. / 3.))
                    - 4. * one * TF2OverallAmpl * (pf ** (1. / 3.)) - 5. * zero * TF2OverallAmpl)
                   * np.pi / (3. * (pf ** (8. / 3.))))
        DPhiIns = DPhiIns + (sigma1 + sigma2 * (fInsJoin ** (1. / 3.)) + sigma3 * (fInsJoin ** (2. / 3.)) + sigma4 * fInsJoin) / eta
        DPhiInt = (beta1 + beta3 / (fInsJoin ** 4) + beta2 / fInsJoin) / eta
        C2Int = DPhiIns - DPhiInt
        PhiInsJoin = _phi_ins(P, fInsJoin, eta)
        PhiIntJoin = beta1 * fInsJoin - beta3 / (3. * fInsJoin * fInsJoin * fInsJoin) + beta2 * np.log(fInsJoin)
        C1Int = PhiInsJoin - PhiIntJoin / eta - C2Int * fInsJoin

        PhiIntTempVal = (beta1 * fMRDJoin - beta3 / (3. * fMRDJoin * fMRDJoin * fMRDJoin) + beta2 * D.log(fMRDJoin)) / eta + C1Int + C2Int * fMRDJoin
        DPhiIntTempVal = C2Int + (beta1 + beta3 / (fMRDJoin ** 4) + beta2 / fMRDJoin) / eta
        DPhiMRDVal = (alpha1 + alpha2 / (fMRDJoin * fMRDJoin) + alpha3 / (fMRDJoin ** (1. / 4.))
                      + alpha4 / (fdamp * (1. + (fMRDJoin - alpha5 * fring) * (fMRDJoin - alpha5 * fring) / (fdamp * fdamp)))) / eta
        PhiMRJoinTemp = (-(alpha2 / fMRDJoin) + (4.0 / 3.0) * (alpha3 * (fMRDJoin ** (3. / 4.))) + alpha1 * fMRDJoin
                         + alpha4 * D.arctan((fMRDJoin - alpha5 * fring) / fdamp))
        C2MRD = DPhiIntTempVal - DPhiMRDVal
        C1MRD = PhiIntTempVal - PhiMRJoinTemp / eta - C2MRD * fMRDJoin

        gamma2 = _fit('gamma2', eta, eta2, xi)
        gamma3 = _fit('gamma3', eta, eta2, xi)
        fpeak = D.where(D.value(gamma2) >= 1.0, D.fabs(fring - (fdamp * gamma3) / gamma2),
                        D.fabs(fring + (fdamp * (-1.0 + D.sqrt(1.0 - gamma2 * gamma2)) * gamma3) / gamma2))
        t0 = (alpha1 + alpha2 / (fpeak * fpeak) + alpha3 / (fpeak ** (1. / 4.))
              + alpha4 / (fdamp * (1. + (fpeak - alpha5 * fring) * (fpeak - alpha5 * fring) / (fdamp * fdamp)))) / eta

        fRef = D.amin0(fgrid)
        if self.fRef is not None:
            fRef = M * glob.GMsun_over_c3 * self.fRef
        co = dict(P=P, eta=eta, beta1=beta1, beta2=beta2, beta3=beta3, C1Int=C1Int, C2Int=C2Int, alpha1=alpha1,
                  alpha2=alpha2, alpha3=alpha3, alpha4=alpha4, alpha5=alpha5, fring=fring, fdamp=fdamp,
                  C1MRD=C1MRD, C2MRD=C2MRD, fMRDJoin=fMRDJoin, gamma2=gamma2, gamma3=gamma3, fpeak=fpeak, t0=t0,
                  fRef=fRef, fgrid=fgrid, Seta=Seta, QuadMon1=QuadMon1, QuadMon2=QuadMon2,
                  TF2OverallAmpl=TF2OverallAmpl)
        co['phiRef'] = self._phi_piecewise(fRef, co)
        return co

    def _quadmons(self, ev):
        return 1.0, 1.0

    def _phase_extra(self, fgrid, ev, eta, Seta, QuadMon1, QuadMon2, TF2OverallAmpl):
        return 0.0

    def _phi_piecewise(self, x, co):
        """The 3-branch phase at waveforms.py:1143-1147 (fcut applied when
        apply_fcut)."""
        eta = co['eta']
        ins = _phi_ins(co['P'], x, eta)
        inter = (co['beta1'] * x - co['beta3'] / (3. * x * x * x) + co['beta2'] * D.log(x)) / eta + co['C1Int'] + co['C2Int'] * x
        mrd = ((-(co['alpha2'] / x) + (4.0 / 3.0) * (co['alpha3'] * (x ** (3. / 4.))) + co['alpha1'] * x
                + co['alpha4'] * D.arctan((x - co['alpha5'] * co['fring']) / co['fdamp'])) / eta + co['C1MRD'] + co['C2MRD'] * x)
        xv = D.value(x)
        if self.apply_fcut:
            mrd = D.where(xv < self.fcutPar, mrd, 0.)
        return D.where(xv < self.PHI_fJoin_INS, ins, D.where(xv < D.value(co['fMRDJoin']), inter, mrd))

    def Ampl(self, f, ev):
        """waveforms.py:1155-1254."""
        M = ev['Mc'] / (ev['eta'] ** (3. / 5.))
        eta = ev['eta']
        eta2 = eta * eta
        chi1, chi2 = ev['chi1z'], ev['chi2z']
        chi12, chi22 = chi1 * chi1, chi2 * chi2
        Seta = _seta(eta)
        SetaPlus1 = 1.0 + Seta
        chi_s = 0.5 * (chi1 + chi2)
        chi_a = 0.5 * (chi1 - chi2)
        fgrid = M * glob.GMsun_over_c3 * f
        chiPN = (chi_s * (1.0 - eta * 76.0 / 113.0) + Seta * chi_a)
        xi = -1.0 + chiPN
        fring, fdamp = self._ringdown(eta, chi1, chi2)
        gamma1 = _fit('gamma1', eta, eta2, xi)
        gamma2 = _fit('gamma2', eta, eta2, xi)
        gamma3 = _fit('gamma3', eta, eta2, xi)
        fpeak = D.where(D.value(gamma2) >= 1.0, D.fabs(fring - (fdamp * gamma3) / gamma2),
                        fring + (fdamp * (-1.0 + D.sqrt(1.0 - gamma2 * gamma2)) * gamma3) / gamma2)
        rho1 = _fit('rho1', eta, eta2, xi)
        rho2 = _fit('rho2', eta, eta2, xi)
        rho3 = _fit('rho3', eta, eta2, xi)
        f1Interm = self.AMP_fJoin_INS
        f3Interm = fpeak
        dfInterm = 0.5 * (f3Interm - f1Interm)
        f2Interm = f1Interm + dfInterm
        amp0 = D.sqrt(2.0 * eta / 3.0) * (np.pi ** (-1. / 6.))
        pi = np.pi
        A = {}
        A['two_thirds'] = ((-969. + 1804. * eta) * (pi ** (2. / 3.))) / 672.
        A['one'] = ((chi1 * (81. * SetaPlus1 - 44. * eta) + chi2 * (81. - 81. * Seta - 44. * eta)) * pi) / 48.
        A['four_thirds'] = ((-27312085.0 - 10287648. * chi22 - 10287648. * chi12 * SetaPlus1 + 10287648. * chi22 * Seta
                             + 24. * (-1975055. + 857304. * chi12 - 994896. * chi1 * chi2 + 857304. * chi22) * eta
                             + 35371056 * eta2) * (pi ** (4. / 3.))) / 8.128512e6
        A['five_thirds'] = ((pi ** (5. / 3.)) * (chi2 * (-285197. * (-1. + Seta) + 4. * (-91902. + 1579. * Seta) * eta - 35632. * eta2)
                                                 + chi1 * (285197. * SetaPlus1 - 4. * (91902. + 1579. * Seta) * eta - 35632. * eta2)
                                                 + 42840. * (-1.0 + 4. * eta) * pi)) / 32256.
        A['two'] = -((pi ** 2.) * (-336. * (-3248849057.0 + 2943675504. * chi12 - 3339284256. * chi1 * chi2 + 2943675504. * chi22) * eta2
                                   - 324322727232. * eta2 * eta
                                   - 7. * (-177520268561. + 107414046432. * chi22 + 107414046432. * chi12 * SetaPlus1
                                           - 107414046432. * chi22 * Seta + 11087290368. * (chi1 + chi2 + chi1 * Seta - chi2 * Seta) * pi)
                                   + 12. * eta * (-545384828789. - 176491177632. * chi1 * chi2 + 202603761360. * chi22
                                                  + 77616. * chi12 * (2610335. + 995766. * Seta) - 77287373856. * chi22 * Seta
                                                  + 5841690624. * (chi1 + chi2) * pi + 21384760320. * pi * pi))) / 6.0085960704e10
        A['seven_thirds'] = rho1
        A['eight_thirds'] = rho2
        A['three'] = rho3
        f1 = f1Interm
        v1 = (1. + (f1 ** (2. / 3.)) * A['two_thirds'] + (f1 ** (4. / 3.)) * A['four_thirds'] + (f1 ** (5. / 3.)) * A['five_thirds']
              + (f1 ** (7. / 3.)) * A['seven_thirds'] + (f1 ** (8. / 3.)) * A['eight_thirds']
              + f1 * (A['one'] + f1 * A['two'] + f1 * f1 * A['three']))
        d1 = (((-969. + 1804. * eta) * (pi ** (2. / 3.))) / (1008. * (f1 ** (1. / 3.)))
              + ((chi1 * (81. * SetaPlus1 - 44. * eta) + chi2 * (81. - 81. * Seta - 44. * eta)) * pi) / 48.
              + ((-27312085. - 10287648. * chi22 - 10287648. * chi12 * SetaPlus1 + 10287648. * chi22 * Seta
                  + 24. * (-1975055. + 857304. * chi12 - 994896. * chi1 * chi2 + 857304. * chi22) * eta + 35371056. * eta2)
                 * (f1 ** (1. / 3.)) * (pi ** (4. / 3.))) / 6.096384e6
              + (5. * (f1 ** (2. / 3.)) * (pi ** (5. / 3.)) * (chi2 * (-285197. * (-1 + Seta) + 4. * (-91902. + 1579. * Seta) * eta - 35632. * eta2)
                                                               + chi1 * (285197. * SetaPlus1 - 4. * (91902. + 1579. * Seta) * eta - 35632. * eta2)
                                                               + 42840. * (-1 + 4 * eta) * pi)) / 96768.
              - (f1 * pi * pi * (-336. * (-3248849057.0 + 2943675504. * chi12 - 3339284256. * chi1 * chi2 + 2943675504. * chi22) * eta2
                                 - 324322727232. * eta2 * eta
                                 - 7. * (-177520268561. + 107414046432. * chi22 + 107414046432. * chi12 * SetaPlus1
                                         - 107414046432. * chi22 * Seta + 11087290368 * (chi1 + chi2 + chi1 * Seta - chi2 * Seta) * pi)
                                 + 12. * eta * (-545384828789.0 - 176491177632. * chi1 * chi2 + 202603761360. * chi22
                                                + 77616. * chi12 * (2610335. + 995766. * Seta) - 77287373856. * chi22 * Seta
                                                + 5841690624. * (chi1 + chi2) * pi + 21384760320 * pi * pi))) / 3.0042980352e10
              + (7.0 / 3.0) * (f1 ** (4. / 3.)) * rho1 + (8.0 / 3.0) * (f1 ** (5. / 3.)) * rho2 + 3. * (f1 * f1) * rho3)
        f3 = f3Interm
        v3 = D.exp(-(f3 - fring) * gamma2 / (fdamp * gamma3)) * (fdamp * gamma3 * gamma1) / ((f3 - fring) * (f3 - fring) + fdamp * gamma3 * fdamp * gamma3)
        d2 = (((-2. * fdamp * (f3 - fring) * gamma3 * gamma1) / ((f3 - fring) * (f3 - fring) + fdamp * gamma3 * fdamp * gamma3) - (gamma2 * gamma1))
              / (D.exp((f3 - fring) * gamma2 / (fdamp * gamma3)) * ((f3 - fring) * (f3 - fring) + fdamp * gamma3 * fdamp * gamma3)))
        v2 = _fit('v2', eta, eta2, xi)
        f2 = f2Interm
        delta0, delta1, delta2, delta3, delta4 = _phenomd_deltas(f1, f2, f3, v1, v2, v3, d1, d2)
        Overallamp = 2. * np.sqrt(5. / (64. * np.pi)) * M * glob.GMsun_over_c2_Gpc * M * glob.GMsun_over_c3 / ev['dL']
        ins = (1. + (fgrid ** (2. / 3.)) * A['two_thirds'] + (fgrid ** (4. / 3.)) * A['four_thirds']
               + (fgrid ** (5. / 3.)) * A['five_thirds'] + (fgrid ** (7. / 3.)) * A['seven_thirds']
               + (fgrid ** (8. / 3.)) * A['eight_thirds'] + fgrid * (A['one'] + fgrid * A['two'] + fgrid * fgrid * A['three']))
        inter = delta0 + fgrid * delta1 + fgrid * fgrid * (delta2 + fgrid * delta3 + fgrid * fgrid * delta4)
        mrd = D.exp(-(fgrid - fring) * gamma2 / (fdamp * gamma3)) * (fdamp * gamma3 * gamma1) / ((fgrid - fring) * (fgrid - fring) + fdamp * gamma3 * fdamp * gamma3)
        xv = D.value(fgrid)
        if self.apply_fcut:
            mrd = D.where(xv < self.fcutPar, mrd, 0.)
        amplitudeIMR = D.where(xv < self.AMP_fJoin_INS, ins, D.where(xv < D.value(fpeak), inter, mrd))
        return self._amp_final(Overallamp, amp0, fgrid, amplitudeIMR, ev, eta, Seta)

    def _amp_final(self, Overallamp, amp0, fgrid, amplitudeIMR, ev, eta, Seta):
        return Overallamp * amp0 * (fgrid ** (-7. / 6.)) * amplitudeIMR

    def tau_star(self, f, ev):
        return tau_star_pn(f, ev)

    def fcut(self, ev):
        """waveforms.py:1324-1333."""
        return self.fcutPar / (ev['Mc'] * glob.GMsun_over_c3 / (ev['eta'] ** (3. / 5.)))


def _phi_ins(P, x, eta):
    """Inspiral phase expression shared by waveforms.py:1117, 1142, 1145."""
    return (P['initial_phasing'] + P['two_thirds'] * (x ** (2. / 3.)) + P['third'] * (x ** (1. / 3.))
            + P['third_log'] * (x ** (1. / 3.)) * D.log(np.pi * x) / 3. + P['log'] * D.log(np.pi * x) / 3.
            + P['min_third'] * (x ** (-1. / 3.)) + P['min_two_thirds'] * (x ** (-2. / 3.)) + P['min_one'] / x
            + P['min_four_thirds'] * (x ** (-4. / 3.)) + P['min_five_thirds'] * (x ** (-5. / 3.))
            + (P['one'] * x + P['four_thirds'] * (x ** (4. / 3.)) + P['five_thirds'] * (x ** (5. / 3.)) + P['two'] * x * x) / eta)


def _phenomd_deltas(f1, f2, f3, v1, v2, v3, d1, d2):
    """Intermediate-amplitude collocation solution, waveforms.py:1230-1245.
    The reference writes it as explicit polynomials; they are restated here
    term by term."""
    f12 = f1 * f1
    f13 = f1 * f12
    f14 = f1 * f13
    f15 = f1 * f14
    f22 = f2 * f2
    f23 = f2 * f22
    f24 = f2 * f23
    f32 = f3 * f3
    f33 = f3 * f32
    f34 = f3 * f33
    f35 = f3 * f34
    den = ((f1 - f2) * (f1 - f2) * (f1 - f3) * (f1 - f3) * (f1 - f3) * (f3 - f2) * (f3 - f2))
    delta0 = -((d2 * f15 * f22 * f3 - 2 * d2 * f14 * f23 * f3 + d2 * f13 * f24 * f3 - d2 * f15 * f2 * f32 + d2 * f14 * f22 * f32
                - d1 * f13 * f23 * f32 + d2 * f13 * f23 * f32 + d1 * f12 * f24 * f32 - d2 * f12 * f24 * f32 + d2 * f14 * f2 * f33
                + 2 * d1 * f13 * f22 * f33 - 2 * d2 * f13 * f22 * f33 - d1 * f12 * f23 * f33 + d2 * f12 * f23 * f33 - d1 * f1 * f24 * f33
                - d1 * f13 * f2 * f34 - d1 * f12 * f22 * f34 + 2 * d1 * f1 * f23 * f34 + d1 * f12 * f2 * f35 - d1 * f1 * f22 * f35
                + 4 * f12 * f23 * f32 * v1 - 3 * f1 * f24 * f32 * v1 - 8 * f12 * f22 * f33 * v1 + 4 * f1 * f23 * f33 * v1 + f24 * f33 * v1
                + 4 * f12 * f2 * f34 * v1 + f1 * f22 * f34 * v1 - 2 * f23 * f34 * v1 - 2 * f1 * f2 * f35 * v1 + f22 * f35 * v1
                - f15 * f32 * v2 + 3 * f14 * f33 * v2 - 3 * f13 * f34 * v2 + f12 * f35 * v2 - f15 * f22 * v3 + 2 * f14 * f23 * v3
                - f13 * f24 * v3 + 2 * f15 * f2 * f3 * v3 - f14 * f22 * f3 * v3 - 4 * f13 * f23 * f3 * v3 + 3 * f12 * f24 * f3 * v3
                - 4 * f14 * f2 * f32 * v3 + 8 * f13 * f22 * f32 * v3 - 4 * f12 * f23 * f32 * v3) / den)
    delta1 = -((-(d2 * f15 * f22) + 2. * d2 * f14 * f23 - d2 * f13 * f24 - d2 * f14 * f22 * f3 + 2. * d1 * f13 * f23 * f3
                + 2. * d2 * f13 * f23 * f3 - 2 * d1 * f12 * f24 * f3 - d2 * f12 * f24 * f3 + d2 * f15 * f32 - 3 * d1 * f13 * f22 * f32
                - d2 * f13 * f22 * f32 + 2 * d1 * f12 * f23 * f32 - 2 * d2 * f12 * f23 * f32 + d1 * f1 * f24 * f32 + 2 * d2 * f1 * f24 * f32
                - d2 * f14 * f33 + d1 * f12 * f22 * f33 + 3 * d2 * f12 * f22 * f33 - 2 * d1 * f1 * f23 * f33 - 2 * d2 * f1 * f23 * f33
                + d1 * f24 * f33 + d1 * f13 * f34 + d1 * f1 * f22 * f34 - 2 * d1 * f23 * f34 - d1 * f12 * f35 + d1 * f22 * f35
                - 8 * f12 * f23 * f3 * v1 + 6 * f1 * f24 * f3 * v1 + 12 * f12 * f22 * f32 * v1 - 8 * f1 * f23 * f32 * v1
                - 4 * f12 * f34 * v1 + 2 * f1 * f35 * v1 + 2 * f15 * f3 * v2 - 4 * f14 * f32 * v2 + 4 * f12 * f34 * v2 - 2 * f1 * f35 * v2
                - 2 * f15 * f3 * v3 + 8 * f12 * f23 * f3 * v3 - 6 * f1 * f24 * f3 * v3 + 4 * f14 * f32 * v3 - 12 * f12 * f22 * f32 * v3
                + 8 * f1 * f23 * f32 * v3) / den)
    delta2 = -((d2 * f15 * f2 - d1 * f13 * f23 - 3 * d2 * f13 * f23 + d1 * f12 * f24 + 2. * d2 * f12 * f24 - d2 * f15 * f3
                + d2 * f14 * f2 * f3 - d1 * f12 * f23 * f3 + d2 * f12 * f23 * f3 + d1 * f1 * f24 * f3 - d2 * f1 * f24 * f3 - d2 * f14 * f32
                + 3 * d1 * f13 * f2 * f32 + d2 * f13 * f2 * f32 - d1 * f1 * f23 * f32 + d2 * f1 * f23 * f32 - 2 * d1 * f24 * f32
                - d2 * f24 * f32 - 2 * d1 * f13 * f33 + 2 * d2 * f13 * f33 - d1 * f12 * f2 * f33 - 3 * d2 * f12 * f2 * f33
                + 3 * d1 * f23 * f33 + d2 * f23 * f33 + d1 * f12 * f34 - d1 * f1 * f2 * f34 + d1 * f1 * f35 - d1 * f2 * f35
                + 4 * f12 * f23 * v1 - 3 * f1 * f24 * v1 + 4 * f1 * f23 * f3 * v1 - 3 * f24 * f3 * v1 - 12 * f12 * f2 * f32 * v1
                + 4 * f23 * f32 * v1 + 8 * f12 * f33 * v1 - f1 * f34 * v1 - f35 * v1 - f15 * v2 - f14 * f3 * v2 + 8 * f13 * f32 * v2
                - 8 * f12 * f33 * v2 + f1 * f34 * v2 + f35 * v2 + f15 * v3 - 4 * f12 * f23 * v3 + 3 * f1 * f24 * v3 + f14 * f3 * v3
                - 4 * f1 * f23 * f3 * v3 + 3 * f24 * f3 * v3 - 8 * f13 * f32 * v3 + 12 * f12 * f2 * f32 * v3 - 4 * f23 * f32 * v3) / den)
    delta3 = -((-2. * d2 * f14 * f2 + d1 * f13 * f22 + 3 * d2 * f13 * f22 - d1 * f1 * f24 - d2 * f1 * f24 + 2 * d2 * f14 * f3
                - 2. * d1 * f13 * f2 * f3 - 2 * d2 * f13 * f2 * f3 + d1 * f12 * f22 * f3 - d2 * f12 * f22 * f3 + d1 * f24 * f3
                + d2 * f24 * f3 + d1 * f13 * f32 - d2 * f13 * f32 - 2 * d1 * f12 * f2 * f32 + 2 * d2 * f12 * f2 * f32
                + d1 * f1 * f22 * f32 - d2 * f1 * f22 * f32 + d1 * f12 * f33 - d2 * f12 * f33 + 2 * d1 * f1 * f2 * f33
                + 2 * d2 * f1 * f2 * f33 - 3 * d1 * f22 * f33 - d2 * f22 * f33 - 2 * d1 * f1 * f34 + 2 * d1 * f2 * f34
                - 4 * f12 * f22 * v1 + 2 * f24 * v1 + 8 * f12 * f2 * f3 * v1 - 4 * f1 * f22 * f3 * v1 - 4 * f12 * f32 * v1
                + 8 * f1 * f2 * f32 * v1 - 4 * f22 * f32 * v1 - 4 * f1 * f33 * v1 + 2 * f34 * v1 + 2 * f14 * v2 - 4 * f13 * f3 * v2
                + 4 * f1 * f33 * v2 - 2 * f34 * v2 - 2 * f14 * v3 + 4 * f12 * f22 * v3 - 2 * f24 * v3 + 4 * f13 * f3 * v3
                - 8 * f12 * f2 * f3 * v3 + 4 * f1 * f22 * f3 * v3 + 4 * f12 * f32 * v3 - 8 * f1 * f2 * f32 * v3 + 4 * f22 * f32 * v3) / den)
    delta4 = -((d2 * f13 * f2 - d1 * f12 * f22 - 2 * d2 * f12 * f22 + d1 * f1 * f23 + d2 * f1 * f23 - d2 * f13 * f3
                + 2. * d1 * f12 * f2 * f3 + d2 * f12 * f2 * f3 - d1 * f1 * f22 * f3 + d2 * f1 * f22 * f3 - d1 * f23 * f3 - d2 * f23 * f3
                - d1 * f12 * f32 + d2 * f12 * f32 - d1 * f1 * f2 * f32 - 2 * d2 * f1 * f2 * f32 + 2 * d1 * f22 * f32 + d2 * f22 * f32
                + d1 * f1 * f33 - d1 * f2 * f33 + 3 * f1 * f22 * v1 - 2 * f23 * v1 - 6 * f1 * f2 * f3 * v1 + 3 * f22 * f3 * v1
                + 3 * f1 * f32 * v1 - f33 * v1 - f13 * v2 + 3 * f12 * f3 * v2 - 3 * f1 * f32 * v2 + f33 * v2 + f13 * v3
                - 3 * f1 * f22 * v3 + 2 * f23 * v3 - 3 * f12 * f3 * v3 + 6 * f1 * f2 * f3 * v3 - 3 * f22 * f3 * v3) / den)
    return delta0, delta1, delta2, delta3, delta4


# ---------------------------------------------------------------- NRTidalv2
def _kappa2T(Seta, L1, L2):
    """Tidal coupling constant, waveforms.py:1594 / 1749 (arXiv:1905.06011 eq. 8)."""
    Xa = 0.5 * (1.0 + Seta)
    Xb = 0.5 * (1.0 - Seta)
    return (3.0 / 13.0) * ((1.0 + 12.0 * Xb / Xa) * (Xa ** 5) * L1 + (1.0 + 12.0 * Xa / Xb) * (Xb ** 5) * L2)


def _f_merger(eta, Seta, kappa2T):
    """Dimensionless merger frequency of the Planck taper, waveforms.py:1760-1773."""
    q = 0.5 * (1.0 + Seta - 2.0 * eta) / eta
    k2 = kappa2T * kappa2T
    numPT = 1.0 + 3.35411203e-2 * kappa2T + 4.31460284e-5 * k2
    denPT = 1.0 + 7.54224145e-2 * kappa2T + 2.23626859e-4 * k2
    return 0.3586 / D.sqrt(q) * (numPT / denPT) / (2. * np.pi)


def planck_taper(x, y, end_inclusive=False):
    """waveforms.py:1775-1797: the taper value with the reference's custom
    JVP -- a tangent through y (the merger frequency) only; x's tangent is
    dropped (defjvps(None, ...)); NaN -> 0 (nan_to_num) in value and tangent.
    At exactly x == 1.2 y the reference's expression gives 1 (exp(+inf));
    end_inclusive=True treats that point as cut (value 0), as the HIP
    kernels do."""
    xv, yv = np.real(D.value(x)), np.real(D.value(y))
    a = 1.2
    ypv = a * yv
    with np.errstate(divide='ignore', invalid='ignore', over='ignore'):
        z = (ypv - yv) / (xv - yv) + (ypv - yv) / (xv - ypv)
        ez = np.exp(z)
        cut = (xv >= ypv) if end_inclusive else (xv > ypv)
        val = np.where(xv < yv, 1., np.where(cut, 0., 1. - 1. / (ez + 1.)))
        der = np.where(xv < yv, 0., np.where(cut, 0., ez * ((-1. + a) / (xv - yv) + (-1. + a) / (xv - ypv)
                                                                 + (-yv + ypv) / ((xv - yv) ** 2)
                                                                 + 1.2 * (-yv + ypv) / ((xv - ypv) ** 2))
                                 / ((ez + 1.) ** 2)))
    val = np.nan_to_num(val)
    der = np.nan_to_num(der)
    if isinstance(y, D.Dual):
        return D.Dual(val, D._expand(y.d, y.v.ndim, val.ndim) * der[None])
    return val


class IMRPhenomD_NRTidalv2(IMRPhenomD):
    """IMRPhenomD_NRTidalv2 (waveforms.py:1339-1837): IMRPhenomD with the
    spin-induced quadrupole in the 2PN and 3PN coefficients, the NRTidalv2
    tidal phase and amplitude, the 3.5PN spin-squared / spin-cubed terms and
    a Planck taper ending at 1.2 f_merger."""

    name = 'IMRPhenomD_NRTidalv2'

    def __init__(self, fRef=None, apply_fcut=True, taper_end_inclusive=False):
        super().__init__(fRef=fRef, apply_fcut=apply_fcut)
        self.taper_end_inclusive = taper_end_inclusive
        self.is_tidal = True
        self.objType = 'BNS'
        self.nParams = 13

    def _lambdas(self, ev):
        if 'Lambda1' in ev:
            return ev['Lambda1'], ev['Lambda2']
        return 0.0 * ev['eta'], 0.0 * ev['eta']

    def _quadmons(self, ev):
        """waveforms.py:1388-1389 (arXiv:1608.02582 eq. 15)."""
        if 'Lambda1' in ev:
            return _quadmon(ev['Lambda1']), _quadmon(ev['Lambda2'])
        return 1.0, 1.0

    def _phase_extra(self, fgrid, ev, eta, Seta, QuadMon1, QuadMon2, TF2OverallAmpl):
        """waveforms.py:1582-1616: tidal phase (arXiv:1905.06011 eq. 7),
        3.5PN spin-squared and spin-cubed terms (eq. 27)."""
        L1, L2 = self._lambdas(ev)
        chi1, chi2 = ev['chi1z'], ev['chi2z']
        chi12, chi22 = chi1 * chi1, chi2 * chi2
        m1ByM = 0.5 * (1.0 + Seta)
        m2ByM = 0.5 * (1.0 - Seta)
        kappa2T = (3.0 / 13.0) * ((1.0 + 12.0 * m2ByM / m1ByM) * (m1ByM ** 5) * L1
                                  + (1.0 + 12.0 * m1ByM / m2ByM) * (m2ByM ** 5) * L2)
        px = np.pi * fgrid
        px23 = px ** (2. / 3.)
        numTidal = (1.0 + (-12.615214237993088 * px23) + (19.0537346970349 * px) + (-21.166863146081035 * (px ** (4. / 3.)))
                    + (90.55082156324926 * (px ** (5. / 3.))) + (-60.25357801943598 * px * px))
        denTidal = 1.0 + (-15.11120782773667 * px23) + (22.195327350624694 * px) + (8.064109635305156 * (px ** (4. / 3.)))
        tidal_phase = -kappa2T * 2.4375 / (m1ByM * m2ByM) * (px ** (5. / 3.)) * numTidal / denTidal

        def octmon(Q):
            lq = D.log(Q)
            return -1. + D.exp(0.003131 + 2.071 * lq - 0.7152 * lq * lq + 0.2458 * lq * lq * lq - 0.03309 * lq * lq * lq * lq)
        Oct1, Oct2 = octmon(QuadMon1), octmon(QuadMon2)
        SS = -400. * np.pi * (QuadMon1 - 1.) * chi12 * m1ByM * m1ByM - 400. * np.pi * (QuadMon2 - 1.) * chi22 * m2ByM * m2ByM
        SSS = (10. * ((m1ByM * m1ByM + 308. / 3. * m1ByM) * chi1 + (m2ByM * m2ByM - 89. / 3. * m2ByM) * chi2) * (QuadMon1 - 1.) * m1ByM * m1ByM * chi12
               + 10. * ((m2ByM * m2ByM + 308. / 3. * m2ByM) * chi2 + (m1ByM * m1ByM - 89. / 3. * m1ByM) * chi1) * (QuadMon2 - 1.) * m2ByM * m2ByM * chi22
               - 440. * Oct1 * m1ByM * m1ByM * m1ByM * chi12 * chi1 - 440. * Oct2 * m2ByM * m2ByM * m2ByM * chi22 * chi2)
        return tidal_phase + (SS + SSS) * TF2OverallAmpl * px23

    def _amp_final(self, Overallamp, amp0, fgrid, amplitudeIMR, ev, eta, Seta):
        """waveforms.py:1744-1799: + tidal amplitude (arXiv:1905.06011 eq. 24),
        times the Planck taper."""
        L1, L2 = self._lambdas(ev)
        kappa2T = _kappa2T(Seta, L1, L2)
        xT = (np.pi * fgrid) ** (2. / 3.)
        polyTidal = (1.0 + 4.157407407407407 * xT + 2519.111111111111 * (xT ** 2.89)) / (1. + 13477.8073677 * (xT ** 4))
        ampTidal = -9.0 * kappa2T * (xT ** 3.25) * polyTidal
        taper = planck_taper(fgrid, _f_merger(eta, Seta, kappa2T), self.taper_end_inclusive)
        return Overallamp * (amp0 * (fgrid ** (-7. / 6.)) * amplitudeIMR + 2 * np.sqrt(np.pi / 5.) * ampTidal) * taper

    def fcut(self, ev):
        """waveforms.py:1803-1835: 1.2 f_merger / (M G)."""
        eta = np.asarray(ev['eta'], dtype=float)
        M = ev['Mc'] / (eta ** (3. / 5.))
        Seta = np.sqrt(np.where(eta < 0.25, 1.0 - 4.0 * eta, 0.))
        L1, L2 = self._lambdas(ev)
        fm = _f_merger(eta, Seta, _kappa2T(Seta, D.value(L1), D.value(L2)))
        return 1.2 * fm / (M * glob.GMsun_over_c3)
