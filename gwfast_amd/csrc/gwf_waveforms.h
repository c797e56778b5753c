// Waveform models of the event-batched path, split into
//   * a per-event setup that evaluates every frequency-independent coefficient
//     (value + tangents w.r.t. Mc, eta, chi1z, chi2z[, LambdaTilde,
//     deltaLambda]) once and stores it in an event record, and
//   * a per-frequency evaluation that reads the record and returns the
//     amplitude A(f) (without the 1/dL factor) and the GW phase Phi(f), each
//     as a dual number (the 3.5PN time to coalescence tau(f), the same for
//     both models, waveforms.py:878-901 and 1299-1322, is evaluated from
//     per-event coefficients of the event record: gwf_paths.h).
// Formulas follow gwfast/waveforms.py:
//   TaylorF2_RestrictedPN  Phi 743-862, Ampl 864-876, fcut 903-953;
//   IMRPhenomD             Phi 992-1153, Ampl 1155-1254, _finalspin
//                          1256-1276, _radiatednrg 1278-1297, fcut
//                          1324-1333;
//   tidal conversions      gwfast/gwfastUtils.py:398-444.
#pragma once
#include "gwf_dual.h"
#include "gwfast_hip.h"

namespace gwf {

constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double GMSUN_C3 = 4.925491025543575903411922162094833998e-6;  // s
constexpr double GMSUN_C2 = 1.476625061404649406193430731479084713e3;   // m
constexpr double UGPC = 3.085677581491367278913937957796471611e25;      // m
constexpr double GMSUN_C2_GPC = GMSUN_C2 / UGPC;
constexpr double CLIGHT = 2.99792458e5;                          // km/s
constexpr double CLIGHT_GPC = CLIGHT / 3.0856778570831e+22;     // Gpc/s
constexpr double REARTH = 6371.00;                               // km
constexpr double DAY = 3600. * 24.;
constexpr double EULER_GAMMA = 0.5772156649015329;

// --------------------------------------------------------------- records
// IMRPhenomD event record: duals of NT_PD = 4 tangents.
enum PdSlot {
  PD_MG = 0, PD_ETA,
  PD_PIP, PD_P23, PD_P13, PD_P13L, PD_PLOG, PD_PM13, PD_PM23, PD_PM1, PD_PM53,
  PD_S1, PD_S43, PD_S53, PD_S2,
  PD_B1, PD_B2, PD_B3, PD_C1I, PD_C2I,
  PD_A1, PD_A2, PD_A3, PD_A4, PD_A5, PD_C1M, PD_C2M,
  PD_FRING, PD_FDAMP, PD_FMRDJ, PD_T0, PD_FREF, PD_PHIREF,
  PD_AM23, PD_AM43, PD_AM53, PD_AM73, PD_AM83, PD_AM1, PD_AM2, PD_AM3,
  PD_DL0, PD_DL1, PD_DL2, PD_DL3, PD_DL4,
  PD_G1, PD_G2, PD_G3, PD_FPEAKA, PD_OAMP, PD_FCUT, PD_IETA,
  // IMRPhenomD_NRTidalv2 only
  PD_TPH, PD_SSS, PD_TAMP, PD_FMERG,
  PD_COUNT
};
// TaylorF2 event record: duals of 4 (6 when tidal) tangents.
enum TfSlot {
  TF_MG = 0, TF_ETA,
  TF_C0, TF_C2, TF_C3, TF_C4, TF_C5, TF_C5L, TF_C6, TF_C6L, TF_C7, TF_C10, TF_C12,
  TF_PHIR, TF_AMP, TF_FCUT,
  // eccentric TaylorF2 (waveforms.py:814-845): coefficients of v^(-34/3) x
  // (1, v^2, v^3, v^4, v^5, v^6, v^6 ln v), each times 3/(128 eta) (-2355/1462)
  // e0^2 v0^(19/3)
  TF_E0, TF_E2, TF_E3, TF_E4, TF_E5, TF_E6, TF_E6L,
  TF_COUNT
};
constexpr int REC_SLOTS = PD_COUNT;  // >= TF_COUNT
constexpr int NT_MAX = 7;
constexpr int REC_STRIDE = REC_SLOTS * (NT_MAX + 1);  // doubles per event record

// IMRPhenomD phenomenological fits: (a0, a1, b0, b1, b2, c0, c1, c2, d0, d1, d2)
// of a0 + a1*eta + (b0 + b1*eta + b2*eta2 + (c0 + c1*eta + c2*eta2)*xi
//                  + (d0 + d1*eta + d2*eta2)*xi*xi)*xi
// (waveforms.py:1039-1056, 1189-1191, 1195-1197, 1227).
enum FitId {
  FIT_SIGMA1 = 0, FIT_SIGMA2, FIT_SIGMA3, FIT_SIGMA4, FIT_BETA1, FIT_BETA2, FIT_BETA3,
  FIT_ALPHA1, FIT_ALPHA2, FIT_ALPHA3, FIT_ALPHA4, FIT_ALPHA5,
  FIT_GAMMA1, FIT_GAMMA2, FIT_GAMMA3, FIT_RHO1, FIT_RHO2, FIT_RHO3, FIT_V2, FIT_COUNT
};

#if defined(__HIPCC__)
__device__ __constant__
#endif
static const double kPhenomDFits[FIT_COUNT][11] = {
    /* sigma1 */ {2096.551999295543, 1463.7493168261553, 1312.5493286098522, 18307.330017082117, -43534.1440746107, -833.2889543511114, 32047.31997183187, -108609.45037520859, 452.25136398112204, 8353.439546391714, -44531.3250037322},
    /* sigma2 */ {-10114.056472621156, -44631.01109458185, -6541.308761668722, -266959.23419307504, 686328.3229317984, 3405.6372187679685, -437507.7208209015, 1631817.1307344697, -7462.648563007646, -114585.25177153319, 674402.4689098676},
    /* sigma3 */ {22933.658273436497, 230960.00814979506, 14961.083974183695, 1194018.1342318142, -3104223.9693052764, -3038.166617199259, 1872032.2849093592, -7309145.012085539, 42738.22871475411, 467502.018616601, -3064853.498512499},
    /* sigma4 */ {-14621.71522218357, -377812.8579387104, -9608.682631509726, -1710892.5257214056, 4332924.601416521, -22366.683262266528, -2501971.6386377467, 10274495.902259542, -85360.30079034246, -570025.3441737515, 4396844.346849777},
    /* beta1 */ {97.89747327985583, -42.659730877489224, 153.48421037904913, -1417.0620760768954, 2752.8614143665027, 138.7406469558649, -1433.6585075135881, 2857.7418952430758, 41.025109467376126, -423.680737974639, 850.3594335657173},
    /* beta2 */ {-3.282701958759534, -9.051384468245866, -12.415449742258042, 55.4716447709787, -106.05109938966335, -11.953044553690658, 76.80704618365418, -155.33172948098394, -3.4129261592393263, 25.572377569952536, -54.408036707740465},
    /* beta3 */ {-2.5156429818799565e-05, 1.9750256942201327e-05, -1.8370671469295915e-05, 2.1886317041311973e-05, 8.250240316860033e-05, 7.157371250566708e-06, -5.5780000112270685e-05, 0.00019142082884072178, 5.447166261464217e-06, -3.220610095021982e-05, 7.974016714984341e-05},
    /* alpha1 */ {43.31514709695348, 638.6332679188081, -32.85768747216059, 2415.8938269370315, -5766.875169379177, -61.85459307173841, 2953.967762459948, -8986.29057591497, -21.571435779762044, 981.2158224673428, -3239.5664895930286},
    /* alpha2 */ {-0.07020209449091723, -0.16269798450687084, -0.1872514685185499, 1.138313650449945, -2.8334196304430046, -0.17137955686840617, 1.7197549338119527, -4.539717148261272, -0.049983437357548705, 0.6062072055948309, -1.682769616644546},
    /* alpha3 */ {9.5988072383479, -397.05438595557433, 16.202126189517813, -1574.8286986717037, 3600.3410843831093, 27.092429659075467, -1786.482357315139, 5152.919378666511, 11.175710130033895, -577.7999423177481, 1808.730762932043},
    /* alpha4 */ {-0.02989487384493607, 1.4022106448583738, -0.07356049468633846, 0.8337006542278661, 0.2240008282397391, -0.055202870001177226, 0.5667186343606578, 0.7186931973380503, -0.015507437354325743, 0.15750322779277187, 0.21076815715176228},
    /* alpha5 */ {0.9974408278363099, -0.007884449714907203, -0.059046901195591035, 1.3958712396764088, -4.516631601676276, -0.05585343136869692, 1.7516580039343603, -5.990208965347804, -0.017945336522161195, 0.5965097794825992, -2.0608879367971804},
    /* gamma1 */ {0.006927402739328343, 0.03020474290328911, 0.006308024337706171, -0.12074130661131138, 0.26271598905781324, 0.0034151773647198794, -0.10779338611188374, 0.27098966966891747, 0.0007374185938559283, -0.02749621038376281, 0.0733150789135702},
    /* gamma2 */ {1.010344404799477, 0.0008993122007234548, 0.283949116804459, -4.049752962958005, 13.207828172665366, 0.10396278486805426, -7.025059158961947, 24.784892370130475, 0.03093202475605892, -2.6924023896851663, 9.609374464684983},
    /* gamma3 */ {1.3081615607036106, -0.005537729694807678, -0.06782917938621007, -0.6689834970767117, 3.403147966134083, -0.05296577374411866, -0.9923793203111362, 4.820681208409587, -0.006134139870393713, -0.38429253308696365, 1.7561754421985984},
    /* rho1 */ {3931.8979897196696, -17395.758706812805, 3132.375545898835, 343965.86092361377, -1216256.5819981997, -70698.00600428853, 1383907.177859705, -3966276.1890979446, -60017.52423652596, 803515.1181825735, -2091710.365941658},
    /* rho2 */ {-40105.47653771657, 112253.0169706701, 23561.696065836168, -3476180.699403351, 11375936.70849482, 754313.1127166454, -13084760.44625268, 36444584.853928134, 596226.612472288, -7427790.1143564405, 18928977.514040343},
    /* rho3 */ {83208.35471266537, -191237.7264145924, -210916.2454782992, 8717975.08352568, -26914942.420669552, -1988980.6527362722, 30888029.960154563, -83908702.79256162, -1453503.1953446497, 17063528.990822166, -42748659.731120914},
    /* v2 */ {0.8149838730507785, 2.5747553517454658, 1.1610198035496786, -2.3627771785551537, 6.771038707057573, 0.7570782938606834, -2.7256896890432474, 7.1140380397149965, 0.1766934149293479, -0.7978690983168183, 2.1162391502005153},
};

template <class T>
GWF_HD T fit_eval(int id, const T& eta, const T& eta2, const T& xi) {
  const double* c = kPhenomDFits[id];
  return c[0] + c[1] * eta +
         (c[2] + c[3] * eta + c[4] * eta2 + (c[5] + c[6] * eta + c[7] * eta2) * xi +
          (c[8] + c[9] * eta + c[10] * eta2) * xi * xi) *
             xi;
}

// sqrt(where(eta < 0.25, 1 - 4 eta, 0)) as in waveforms.py:759.
template <class T>
GWF_HD T seta_of(const T& eta) {
  T arg = 1.0 - 4.0 * eta;
  if (!(val(eta) < 0.25)) arg = 0.0 * eta;
  return dsqrt(arg);
}

// jnp.interp on a strictly increasing table, derivative = segment slope
// (segment chosen with searchsorted(side='right') clipped to [1, n-1]).
template <class T>
GWF_HD T interp_tab(const T& x, const double* xp, const double* fp, int n) {
  const double xv = val(x);
  int lo = 0, hi = n;  // upper_bound
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] <= xv) lo = mid + 1; else hi = mid;
  }
  int i = lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
  const double df = fp[i] - fp[i - 1];
  const double dx = xp[i] - xp[i - 1];
  T r = fp[i - 1] + ((x - xp[i - 1]) / dx) * df;
  if (xv < xp[0]) r = 0.0 * x + fp[0];
  if (xv > xp[n - 1]) r = 0.0 * x + fp[n - 1];
  return r;
}

// ------------------------------------------------------------ IMRPhenomD
template <class T>
struct PdRemnant {
  T fring, fdamp;
};

template <class T>
GWF_HD T pd_finalspin(const T& eta, const T& chi1, const T& chi2) {
  const T Seta = seta_of(eta);
  const T m1 = 0.5 * (1.0 + Seta);
  const T m2 = 0.5 * (1.0 - Seta);
  const T s = (m1 * m1 * chi1 + m2 * m2 * chi2);
  const T af1 = eta * (3.4641016151377544 - 4.399247300629289 * eta + 9.397292189321194 * eta * eta -
                       13.180949901606242 * eta * eta * eta);
  const T af2 = eta * (s * ((1.0 / eta - 0.0850917821418767 - 5.837029316602263 * eta) +
                            (0.1014665242971878 - 2.0967746996832157 * eta) * s));
  const T af3 = eta * (s * ((-1.3546806617824356 + 4.108962025369336 * eta) * s * s +
                            (-0.8676969352555539 + 2.064046835273906 * eta) * s * s * s));
  return af1 + af2 + af3;
}

template <class T>
GWF_HD T pd_radiatednrg(const T& eta, const T& chi1, const T& chi2) {
  const T Seta = seta_of(eta);
  const T m1 = 0.5 * (1.0 + Seta);
  const T m2 = 0.5 * (1.0 - Seta);
  const T s = (m1 * m1 * chi1 + m2 * m2 * chi2) / (m1 * m1 + m2 * m2);
  const T EradNS = eta * (0.055974469826360077 + 0.5809510763115132 * eta - 0.9606726679372312 * eta * eta +
                          3.352411249771192 * eta * eta * eta);
  return (EradNS * (1. + (-0.0030302335878845507 - 2.0066110851351073 * eta + 7.7050567802399215 * eta * eta) * s)) /
         (1. + (-0.6714403054720589 - 1.4756929437702908 * eta + 7.304676214885011 * eta * eta) * s);
}

// TaylorF2 PN phase coefficients (waveforms.py:784-812; IMRPhenomD repeats
// them at 1062-1077 with QuadMon = 1).
template <class T>
struct Tf2Coeffs {
  T two, three, four, five_tmp, six, seven;
};

template <class T>
GWF_HD Tf2Coeffs<T> tf2_coeffs(const T& eta, const T& Seta, const T& chi1, const T& chi2, const T& QM1,
                                const T& QM2, bool spin_ho) {
  Tf2Coeffs<T> c;
  const T eta2 = eta * eta;
  const T m1ByM = 0.5 * (1.0 + Seta);
  const T m2ByM = 0.5 * (1.0 - Seta);
  const T chi12 = chi1 * chi1, chi22 = chi2 * chi2;
  const T chi1dotchi2 = chi1 * chi2;
  const T chi_s = 0.5 * (chi1 + chi2), chi_a = 0.5 * (chi1 - chi2);
  const T chi_s2 = chi_s * chi_s, chi_a2 = chi_a * chi_a;
  c.two = 3715. / 756. + (55. * eta) / 9.;
  c.three = -16. * PI + (113. * Seta * chi_a) / 3. + (113. / 3. - (76. * eta) / 3.) * chi_s;
  c.four = 5. * (3058.673 / 7.056 + 5429. / 7. * eta + 617. * eta2) / 72. + 247. / 4.8 * eta * chi1dotchi2 -
           721. / 4.8 * eta * chi1dotchi2 + (-720. / 9.6 * QM1 + 1. / 9.6) * m1ByM * m1ByM * chi12 +
           (-720. / 9.6 * QM2 + 1. / 9.6) * m2ByM * m2ByM * chi22 + (240. / 9.6 * QM1 - 7. / 9.6) * m1ByM * m1ByM * chi12 +
           (240. / 9.6 * QM2 - 7. / 9.6) * m2ByM * m2ByM * chi22;
  c.five_tmp = (732985. / 2268. - 24260. * eta / 81. - 340. * eta2 / 9.) * chi_s +
               (732985. / 2268. + 140. * eta / 9.) * Seta * chi_a;
  c.six = 11583.231236531 / 4.694215680 - 640. / 3. * PI * PI - 684.8 / 2.1 * EULER_GAMMA +
          eta * (-15737.765635 / 3.048192 + 225.5 / 1.2 * PI * PI) + eta2 * 76.055 / 1.728 -
          eta2 * eta * 127.825 / 1.296 - log(4.) * 684.8 / 2.1 + PI * chi1 * m1ByM * (1490. / 3. + m1ByM * 260.) +
          PI * chi2 * m2ByM * (1490. / 3. + m2ByM * 260.) + (326.75 / 1.12 + 557.5 / 1.8 * eta) * eta * chi1dotchi2 +
          (4703.5 / 8.4 + 2935. / 6. * m1ByM - 120. * m1ByM * m1ByM) * m1ByM * m1ByM * QM1 * chi12 +
          (-4108.25 / 6.72 - 108.5 / 1.2 * m1ByM + 125.5 / 3.6 * m1ByM * m1ByM) * m1ByM * m1ByM * chi12 +
          (4703.5 / 8.4 + 2935. / 6. * m2ByM - 120. * m2ByM * m2ByM) * m2ByM * m2ByM * QM2 * chi22 +
          (-4108.25 / 6.72 - 108.5 / 1.2 * m2ByM + 125.5 / 3.6 * m2ByM * m2ByM) * m2ByM * m2ByM * chi22;
  if (spin_ho) {
    c.seven = 77096675. * PI / 254016. + 378515. * PI * eta / 1512. - 74045. * PI * eta2 / 756. +
              (-25150083775. / 3048192. + 10566655595. * eta / 762048. - 1042165. * eta2 / 3024. +
               5345. * eta2 * eta / 36. + (14585. / 8. - 7270. * eta + 80. * eta2) * chi_a2) *
                  chi_s +
              (14585. / 24. - 475. * eta / 6. + 100. * eta2 / 3.) * chi_s2 * chi_s +
              Seta * ((-25150083775. / 3048192. + 26804935. * eta / 6048. - 1985. * eta2 / 48.) * chi_a +
                      (14585. / 24. - 2380. * eta) * chi_a2 * chi_a + (14585. / 8. - 215. * eta / 2.) * chi_a * chi_s2);
  } else {
    c.seven = 77096675. * PI / 254016. + 378515. * PI * eta / 1512. - 74045. * PI * eta2 / 756. +
              (-25150083775. / 3048192. + 10566655595. * eta / 762048. - 1042165. * eta2 / 3024. +
               5345. * eta2 * eta / 36.) *
                  chi_s +
              Seta * ((-25150083775. / 3048192. + 26804935. * eta / 6048. - 1985. * eta2 / 48.) * chi_a);
  }
  return c;
}

// Phase-coefficient bundle used for both phiRef (setup) and per-frequency.
template <class T>
struct PdPhase {
  T pip, p23, p13, p13l, plog, pm13, pm23, pm1, pm53, s1, s43, s53, s2;
  T b1, b2, b3, c1i, c2i, a1, a2, a3, a4, a5, c1m, c2m, fring, fdamp, fmrdj, eta;
};

template <class T, class X>
GWF_HD X pd_phi_ins(const PdPhase<T>& c, const X& x) {
  const X lpx = dlog(PI * x);
  const X x13 = dpow(x, 1. / 3.);
  return c.pip + c.p23 * dpow(x, 2. / 3.) + c.p13 * x13 + c.p13l * x13 * lpx / 3. + c.plog * lpx / 3. +
         c.pm13 * dpow(x, -1. / 3.) + c.pm23 * dpow(x, -2. / 3.) + c.pm1 / x + c.pm53 * dpow(x, -5. / 3.) +
         (c.s1 * x + c.s43 * dpow(x, 4. / 3.) + c.s53 * dpow(x, 5. / 3.) + c.s2 * x * x) / c.eta;
}

template <class T, class X>
GWF_HD X pd_phi_int(const PdPhase<T>& c, const X& x) {
  return (c.b1 * x - c.b3 / (3. * x * x * x) + c.b2 * dlog(x)) / c.eta + c.c1i + c.c2i * x;
}

template <class T, class X>
GWF_HD X pd_phi_mrd(const PdPhase<T>& c, const X& x) {
  return (-(c.a2 / x) + (4.0 / 3.0) * (c.a3 * dpow(x, 3. / 4.)) + c.a1 * x + c.a4 * datan((x - c.a5 * c.fring) / c.fdamp)) /
             c.eta +
         c.c1m + c.c2m * x;
}

// Three-branch phase of waveforms.py:1143-1147 (MRD branch zeroed beyond the
// cut when apply_fcut).
template <class T, class X>
GWF_HD X pd_phi_piecewise(const PdPhase<T>& c, const X& x, bool apply_fcut) {
  const double xv = val(x);
  if (xv < 0.018) return pd_phi_ins(c, x);
  if (xv < val(c.fmrdj)) return pd_phi_int(c, x);
  if (apply_fcut && !(xv < 0.2)) return 0.0 * x;
  return pd_phi_mrd(c, x);
}

// Per-event IMRPhenomD setup. Mc, eta, chi1, chi2 are seeded duals (or
// doubles); f0 is the frequency whose M*G*f0 is fRef (grid minimum or the
// model's fixed fRef); rec receives PD_COUNT entries of stride (N+1).
template <int N>
GWF_HD void pd_setup(const D<N>& Mc, const D<N>& eta, const D<N>& chi1, const D<N>& chi2, double f0,
                     bool apply_fcut, const double* qa, const double* qfr, const double* qfd, int nq,
                     double* rec, int stride, const D<N>& QM1, const D<N>& QM2, int tk = -1) {
  typedef D<N> T;
  const T M = Mc / dpow(eta, 3. / 5.);
  const T MG = M * GMSUN_C3;
  const T eta2 = eta * eta;
  const T Seta = seta_of(eta);
  const T SetaPlus1 = 1.0 + Seta;
  const T chi_s = 0.5 * (chi1 + chi2);
  const T chi_a = 0.5 * (chi1 - chi2);
  const T chi12 = chi1 * chi1, chi22 = chi2 * chi2;
  const T chiPN = (chi_s * (1.0 - eta * 76.0 / 113.0) + Seta * chi_a);
  const T xi = -1.0 + chiPN;
  const T aeff = pd_finalspin(eta, chi1, chi2);
  const T Erad = pd_radiatednrg(eta, chi1, chi2);
  const T fring = interp_tab(aeff, qa, qfr, nq) / (1.0 - Erad);
  const T fdamp = interp_tab(aeff, qa, qfd, nq) / (1.0 - Erad);

  const T sigma1 = fit_eval(FIT_SIGMA1, eta, eta2, xi);
  const T sigma2 = fit_eval(FIT_SIGMA2, eta, eta2, xi);
  const T sigma3 = fit_eval(FIT_SIGMA3, eta, eta2, xi);
  const T sigma4 = fit_eval(FIT_SIGMA4, eta, eta2, xi);
  const T beta1 = fit_eval(FIT_BETA1, eta, eta2, xi);
  const T beta2 = fit_eval(FIT_BETA2, eta, eta2, xi);
  const T beta3 = fit_eval(FIT_BETA3, eta, eta2, xi);
  const T alpha1 = fit_eval(FIT_ALPHA1, eta, eta2, xi);
  const T alpha2 = fit_eval(FIT_ALPHA2, eta, eta2, xi);
  const T alpha3 = fit_eval(FIT_ALPHA3, eta, eta2, xi);
  const T alpha4 = fit_eval(FIT_ALPHA4, eta, eta2, xi);
  const T alpha5 = fit_eval(FIT_ALPHA5, eta, eta2, xi);

  // spin-induced quadrupole moments: 1 for IMRPhenomD; from Lambda for
  // NRTidalv2 (waveforms.py:1388-1389, 1452-1453)
  Tf2Coeffs<T> tc = tf2_coeffs(eta, Seta, chi1, chi2, QM1, QM2, false);
  const T TF2OA = 3. / (128. * eta);
  const T five = (38645. * PI / 756. - 65. * PI * eta / 9. - tc.five_tmp);
  const T five_log = five * 3.;
  const T m1ByM = 0.5 * (1.0 + Seta);
  const T m2ByM = 0.5 * (1.0 - Seta);
  const T chi1dotchi2 = chi1 * chi2;
  // 3PN spin-spin part removed (waveforms.py:1079)
  const T six = tc.six - ((326.75 / 1.12 + 557.5 / 1.8 * eta) * eta * chi1dotchi2 +
                          ((4703.5 / 8.4 + 2935. / 6. * m1ByM - 120. * m1ByM * m1ByM) +
                           (-4108.25 / 6.72 - 108.5 / 1.2 * m1ByM + 125.5 / 3.6 * m1ByM * m1ByM)) *
                              m1ByM * m1ByM * chi12 +
                          ((4703.5 / 8.4 + 2935. / 6. * m2ByM - 120. * m2ByM * m2ByM) +
                           (-4108.25 / 6.72 - 108.5 / 1.2 * m2ByM + 125.5 / 3.6 * m2ByM * m2ByM)) *
                              m2ByM * m2ByM * chi22);
  const double six_log = -6848. / 21.;

  PdPhase<T> c;
  c.pip = five * TF2OA;
  c.p23 = tc.seven * TF2OA * pow(PI, 2. / 3.);
  c.p13 = six * TF2OA * pow(PI, 1. / 3.);
  c.p13l = six_log * TF2OA * pow(PI, 1. / 3.);
  c.plog = five_log * TF2OA;
  c.pm13 = tc.four * TF2OA * pow(PI, -1. / 3.);
  c.pm23 = tc.three * TF2OA * pow(PI, -2. / 3.);
  c.pm1 = tc.two * TF2OA / PI;
  c.pm53 = 1.0 * TF2OA * pow(PI, -5. / 3.);
  c.s1 = sigma1;
  c.s43 = sigma2 * 0.75;
  c.s53 = sigma3 * 0.6;
  c.s2 = sigma4 * 0.5;
  c.b1 = beta1;
  c.b2 = beta2;
  c.b3 = beta3;
  c.a1 = alpha1;
  c.a2 = alpha2;
  c.a3 = alpha3;
  c.a4 = alpha4;
  c.a5 = alpha5;
  c.fring = fring;
  c.fdamp = fdamp;
  c.eta = eta;

  const double fInsJoin = 0.018;
  const T fMRDJoin = 0.5 * fring;
  c.fmrdj = fMRDJoin;
  const double pf = PI * fInsJoin;
  T DPhiIns = (2.0 * tc.seven * TF2OA * pow(pf, 7. / 3.) +
               (six * TF2OA + six_log * TF2OA * (1.0 + log(pf) / 3.)) * pow(pf, 2.) + five_log * TF2OA * pow(pf, 5. / 3.) -
               tc.four * TF2OA * pow(pf, 4. / 3.) - 2. * tc.three * TF2OA * pf - 3. * tc.two * TF2OA * pow(pf, 2. / 3.) -
               4. * 0.0 * TF2OA * pow(pf, 1. / 3.) - 5. * 1.0 * TF2OA) *
              PI / (3. * pow(pf, 8. / 3.));
  DPhiIns = DPhiIns + (sigma1 + sigma2 * pow(fInsJoin, 1. / 3.) + sigma3 * pow(fInsJoin, 2. / 3.) + sigma4 * fInsJoin) / eta;
  const T DPhiInt = (beta1 + beta3 / pow(fInsJoin, 4) + beta2 / fInsJoin) / eta;
  const T C2Int = DPhiIns - DPhiInt;
  const D<N> xj = mkc<N>(fInsJoin);
  const T PhiInsJoin = pd_phi_ins(c, xj);
  const T PhiIntJoin = beta1 * fInsJoin - beta3 / (3. * fInsJoin * fInsJoin * fInsJoin) + beta2 * log(fInsJoin);
  const T C1Int = PhiInsJoin - PhiIntJoin / eta - C2Int * fInsJoin;
  c.c1i = C1Int;
  c.c2i = C2Int;

  const T PhiIntTempVal = (beta1 * fMRDJoin - beta3 / (3. * fMRDJoin * fMRDJoin * fMRDJoin) + beta2 * dlog(fMRDJoin)) / eta +
                          C1Int + C2Int * fMRDJoin;
  const T DPhiIntTempVal = C2Int + (beta1 + beta3 / dpow(fMRDJoin, 4.) + beta2 / fMRDJoin) / eta;
  const T DPhiMRDVal = (alpha1 + alpha2 / (fMRDJoin * fMRDJoin) + alpha3 / dpow(fMRDJoin, 1. / 4.) +
                        alpha4 / (fdamp * (1. + (fMRDJoin - alpha5 * fring) * (fMRDJoin - alpha5 * fring) / (fdamp * fdamp)))) /
                       eta;
  const T PhiMRJoinTemp = -(alpha2 / fMRDJoin) + (4.0 / 3.0) * (alpha3 * dpow(fMRDJoin, 3. / 4.)) + alpha1 * fMRDJoin +
                          alpha4 * datan((fMRDJoin - alpha5 * fring) / fdamp);
  const T C2MRD = DPhiIntTempVal - DPhiMRDVal;
  const T C1MRD = PhiIntTempVal - PhiMRJoinTemp / eta - C2MRD * fMRDJoin;
  c.c1m = C1MRD;
  c.c2m = C2MRD;

  const T gamma1 = fit_eval(FIT_GAMMA1, eta, eta2, xi);
  const T gamma2 = fit_eval(FIT_GAMMA2, eta, eta2, xi);
  const T gamma3 = fit_eval(FIT_GAMMA3, eta, eta2, xi);
  // Phase fpeak: fabs on both branches (waveforms.py:1134)
  const T fpeakP = (val(gamma2) >= 1.0) ? dfabs(fring - (fdamp * gamma3) / gamma2)
                                        : dfabs(fring + (fdamp * (-1.0 + dsqrt(1.0 - gamma2 * gamma2)) * gamma3) / gamma2);
  const T t0 = (alpha1 + alpha2 / (fpeakP * fpeakP) + alpha3 / dpow(fpeakP, 1. / 4.) +
                alpha4 / (fdamp * (1. + (fpeakP - alpha5 * fring) * (fpeakP - alpha5 * fring) / (fdamp * fdamp)))) /
               eta;
  const T fRef = MG * f0;
  const T phiRef = pd_phi_piecewise(c, fRef, apply_fcut);

  // The phase coefficients go to the record before the amplitude section
  // starts: none of them is live past this point but fring and fdamp, which
  // shortens the live ranges through the collocation solve below.
  double* r = rec;
  auto put = [&](int slot, const T& x) { std_tk(r + slot * stride, x, tk); };
  put(PD_MG, MG);
  put(PD_ETA, eta);
  put(PD_PIP, c.pip); put(PD_P23, c.p23); put(PD_P13, c.p13); put(PD_P13L, c.p13l); put(PD_PLOG, c.plog);
  put(PD_PM13, c.pm13); put(PD_PM23, c.pm23); put(PD_PM1, c.pm1); put(PD_PM53, c.pm53);
  put(PD_S1, c.s1); put(PD_S43, c.s43); put(PD_S53, c.s53); put(PD_S2, c.s2);
  put(PD_B1, beta1); put(PD_B2, beta2); put(PD_B3, beta3); put(PD_C1I, C1Int); put(PD_C2I, C2Int);
  put(PD_A1, alpha1); put(PD_A2, alpha2); put(PD_A3, alpha3); put(PD_A4, alpha4); put(PD_A5, alpha5);
  put(PD_C1M, C1MRD); put(PD_C2M, C2MRD);
  put(PD_FRING, fring); put(PD_FDAMP, fdamp); put(PD_FMRDJ, fMRDJoin); put(PD_T0, t0); put(PD_FREF, fRef);
  put(PD_PHIREF, phiRef);

  // Amplitude (waveforms.py:1188-1248); fpeak with fabs on the first branch only (1193)
  const T fpeakA = (val(gamma2) >= 1.0) ? dfabs(fring - (fdamp * gamma3) / gamma2)
                                        : fring + (fdamp * (-1.0 + dsqrt(1.0 - gamma2 * gamma2)) * gamma3) / gamma2;
  const T rho1 = fit_eval(FIT_RHO1, eta, eta2, xi);
  const T rho2 = fit_eval(FIT_RHO2, eta, eta2, xi);
  const T rho3 = fit_eval(FIT_RHO3, eta, eta2, xi);
  const double f1 = 0.014;
  const T f3 = fpeakA;
  const T dfInterm = 0.5 * (f3 - f1);
  const T f2 = f1 + dfInterm;
  const T amp0 = dsqrt(2.0 * eta / 3.0) * pow(PI, -1. / 6.);
  const T A23 = ((-969. + 1804. * eta) * pow(PI, 2. / 3.)) / 672.;
  const T A1 = ((chi1 * (81. * SetaPlus1 - 44. * eta) + chi2 * (81. - 81. * Seta - 44. * eta)) * PI) / 48.;
  const T A43 = ((-27312085.0 - 10287648. * chi22 - 10287648. * chi12 * SetaPlus1 + 10287648. * chi22 * Seta +
                  24. * (-1975055. + 857304. * chi12 - 994896. * chi1 * chi2 + 857304. * chi22) * eta + 35371056 * eta2) *
                 pow(PI, 4. / 3.)) /
                8.128512e6;
  const T A53 = (pow(PI, 5. / 3.) * (chi2 * (-285197. * (-1. + Seta) + 4. * (-91902. + 1579. * Seta) * eta - 35632. * eta2) +
                                      chi1 * (285197. * SetaPlus1 - 4. * (91902. + 1579. * Seta) * eta - 35632. * eta2) +
                                      42840. * (-1.0 + 4. * eta) * PI)) /
                32256.;
  const T A2 = -(pow(PI, 2.) *
                 (-336. * (-3248849057.0 + 2943675504. * chi12 - 3339284256. * chi1 * chi2 + 2943675504. * chi22) * eta2 -
                  324322727232. * eta2 * eta -
                  7. * (-177520268561. + 107414046432. * chi22 + 107414046432. * chi12 * SetaPlus1 -
                        107414046432. * chi22 * Seta + 11087290368. * (chi1 + chi2 + chi1 * Seta - chi2 * Seta) * PI) +
                  12. * eta *
                      (-545384828789. - 176491177632. * chi1 * chi2 + 202603761360. * chi22 +
                       77616. * chi12 * (2610335. + 995766. * Seta) - 77287373856. * chi22 * Seta +
                       5841690624. * (chi1 + chi2) * PI + 21384760320. * PI * PI))) /
               6.0085960704e10;
  const T v1 = 1. + pow(f1, 2. / 3.) * A23 + pow(f1, 4. / 3.) * A43 + pow(f1, 5. / 3.) * A53 + pow(f1, 7. / 3.) * rho1 +
               pow(f1, 8. / 3.) * rho2 + f1 * (A1 + f1 * A2 + f1 * f1 * rho3);
  const T d1 =
      ((-969. + 1804. * eta) * pow(PI, 2. / 3.)) / (1008. * pow(f1, 1. / 3.)) +
      ((chi1 * (81. * SetaPlus1 - 44. * eta) + chi2 * (81. - 81. * Seta - 44. * eta)) * PI) / 48. +
      ((-27312085. - 10287648. * chi22 - 10287648. * chi12 * SetaPlus1 + 10287648. * chi22 * Seta +
        24. * (-1975055. + 857304. * chi12 - 994896. * chi1 * chi2 + 857304. * chi22) * eta + 35371056. * eta2) *
       pow(f1, 1. / 3.) * pow(PI, 4. / 3.)) /
          6.096384e6 +
      (5. * pow(f1, 2. / 3.) * pow(PI, 5. / 3.) *
       (chi2 * (-285197. * (-1 + Seta) + 4. * (-91902. + 1579. * Seta) * eta - 35632. * eta2) +
        chi1 * (285197. * SetaPlus1 - 4. * (91902. + 1579. * Seta) * eta - 35632. * eta2) + 42840. * (-1 + 4 * eta) * PI)) /
          96768. -
      (f1 * PI * PI *
       (-336. * (-3248849057.0 + 2943675504. * chi12 - 3339284256. * chi1 * chi2 + 2943675504. * chi22) * eta2 -
        324322727232. * eta2 * eta -
        7. * (-177520268561. + 107414046432. * chi22 + 107414046432. * chi12 * SetaPlus1 - 107414046432. * chi22 * Seta +
              11087290368. * (chi1 + chi2 + chi1 * Seta - chi2 * Seta) * PI) +
        12. * eta *
            (-545384828789.0 - 176491177632. * chi1 * chi2 + 202603761360. * chi22 +
             77616. * chi12 * (2610335. + 995766. * Seta) - 77287373856. * chi22 * Seta + 5841690624. * (chi1 + chi2) * PI +
             21384760320. * PI * PI))) /
          3.0042980352e10 +
      (7.0 / 3.0) * pow(f1, 4. / 3.) * rho1 + (8.0 / 3.0) * pow(f1, 5. / 3.) * rho2 + 3. * (f1 * f1) * rho3;
  const T fdg3 = fdamp * gamma3;
  const T v3 = dexp(-(f3 - fring) * gamma2 / (fdamp * gamma3)) * (fdamp * gamma3 * gamma1) /
               ((f3 - fring) * (f3 - fring) + fdg3 * fdamp * gamma3);
  const T d2 = ((-2. * fdamp * (f3 - fring) * gamma3 * gamma1) / ((f3 - fring) * (f3 - fring) + fdg3 * fdamp * gamma3) -
                (gamma2 * gamma1)) /
               (dexp((f3 - fring) * gamma2 / (fdamp * gamma3)) * ((f3 - fring) * (f3 - fring) + fdg3 * fdamp * gamma3));
  const T v2 = fit_eval(FIT_V2, eta, eta2, xi);

  // The amplitude coefficients, too, go out before the collocation solve,
  // which needs only f1, f2, f3, v1, v2, v3, d1 and d2.
  put(PD_AM23, A23); put(PD_AM43, A43); put(PD_AM53, A53); put(PD_AM73, rho1); put(PD_AM83, rho2);
  put(PD_AM1, A1); put(PD_AM2, A2); put(PD_AM3, rho3);
  put(PD_G1, gamma1); put(PD_G2, gamma2); put(PD_G3, gamma3); put(PD_FPEAKA, fpeakA);
  // Overallamp * amp0 without the 1/dL factor (waveforms.py:1248, 1254)
  put(PD_OAMP, 2. * sqrt(5. / (64. * PI)) * M * GMSUN_C2_GPC * M * GMSUN_C3 * amp0);

  // Collocation solution for the intermediate amplitude polynomial
  // (waveforms.py:1229-1245): the unique quartic with p(f1)=v1, p'(f1)=d1,
  // p(f2)=v2, p(f3)=v3, p'(f3)=d2, written out as in the reference.
  const T f12 = f1 * f1 + 0.0 * eta;
  const T f13 = f1 * f12, f14 = f1 * f13, f15 = f1 * f14;
  const T f22 = f2 * f2, f23 = f2 * f22, f24 = f2 * f23;
  const T f32 = f3 * f3, f33 = f3 * f32, f34 = f3 * f33, f35 = f3 * f34;
  const T F1 = f1 + 0.0 * eta;
  const T den = (F1 - f2) * (F1 - f2) * (F1 - f3) * (F1 - f3) * (F1 - f3) * (f3 - f2) * (f3 - f2);
  const T delta0 =
      -((d2 * f15 * f22 * f3 - 2 * d2 * f14 * f23 * f3 + d2 * f13 * f24 * f3 - d2 * f15 * f2 * f32 + d2 * f14 * f22 * f32 -
         d1 * f13 * f23 * f32 + d2 * f13 * f23 * f32 + d1 * f12 * f24 * f32 - d2 * f12 * f24 * f32 + d2 * f14 * f2 * f33 +
         2 * d1 * f13 * f22 * f33 - 2 * d2 * f13 * f22 * f33 - d1 * f12 * f23 * f33 + d2 * f12 * f23 * f33 - d1 * F1 * f24 * f33 -
         d1 * f13 * f2 * f34 - d1 * f12 * f22 * f34 + 2 * d1 * F1 * f23 * f34 + d1 * f12 * f2 * f35 - d1 * F1 * f22 * f35 +
         4 * f12 * f23 * f32 * v1 - 3 * F1 * f24 * f32 * v1 - 8 * f12 * f22 * f33 * v1 + 4 * F1 * f23 * f33 * v1 +
         f24 * f33 * v1 + 4 * f12 * f2 * f34 * v1 + F1 * f22 * f34 * v1 - 2 * f23 * f34 * v1 - 2 * F1 * f2 * f35 * v1 +
         f22 * f35 * v1 - f15 * f32 * v2 + 3 * f14 * f33 * v2 - 3 * f13 * f34 * v2 + f12 * f35 * v2 - f15 * f22 * v3 +
         2 * f14 * f23 * v3 - f13 * f24 * v3 + 2 * f15 * f2 * f3 * v3 - f14 * f22 * f3 * v3 - 4 * f13 * f23 * f3 * v3 +
         3 * f12 * f24 * f3 * v3 - 4 * f14 * f2 * f32 * v3 + 8 * f13 * f22 * f32 * v3 - 4 * f12 * f23 * f32 * v3) /
        den);
  put(PD_DL0, delta0);
  const T delta1 =
      -((-(d2 * f15 * f22) + 2. * d2 * f14 * f23 - d2 * f13 * f24 - d2 * f14 * f22 * f3 + 2. * d1 * f13 * f23 * f3 +
         2. * d2 * f13 * f23 * f3 - 2 * d1 * f12 * f24 * f3 - d2 * f12 * f24 * f3 + d2 * f15 * f32 - 3 * d1 * f13 * f22 * f32 -
         d2 * f13 * f22 * f32 + 2 * d1 * f12 * f23 * f32 - 2 * d2 * f12 * f23 * f32 + d1 * F1 * f24 * f32 +
         2 * d2 * F1 * f24 * f32 - d2 * f14 * f33 + d1 * f12 * f22 * f33 + 3 * d2 * f12 * f22 * f33 -
         2 * d1 * F1 * f23 * f33 - 2 * d2 * F1 * f23 * f33 + d1 * f24 * f33 + d1 * f13 * f34 + d1 * F1 * f22 * f34 -
         2 * d1 * f23 * f34 - d1 * f12 * f35 + d1 * f22 * f35 - 8 * f12 * f23 * f3 * v1 + 6 * F1 * f24 * f3 * v1 +
         12 * f12 * f22 * f32 * v1 - 8 * F1 * f23 * f32 * v1 - 4 * f12 * f34 * v1 + 2 * F1 * f35 * v1 + 2 * f15 * f3 * v2 -
         4 * f14 * f32 * v2 + 4 * f12 * f34 * v2 - 2 * F1 * f35 * v2 - 2 * f15 * f3 * v3 + 8 * f12 * f23 * f3 * v3 -
         6 * F1 * f24 * f3 * v3 + 4 * f14 * f32 * v3 - 12 * f12 * f22 * f32 * v3 + 8 * F1 * f23 * f32 * v3) /
        den);
  put(PD_DL1, delta1);
  const T delta2 =
      -((d2 * f15 * f2 - d1 * f13 * f23 - 3 * d2 * f13 * f23 + d1 * f12 * f24 + 2. * d2 * f12 * f24 - d2 * f15 * f3 +
         d2 * f14 * f2 * f3 - d1 * f12 * f23 * f3 + d2 * f12 * f23 * f3 + d1 * F1 * f24 * f3 - d2 * F1 * f24 * f3 -
         d2 * f14 * f32 + 3 * d1 * f13 * f2 * f32 + d2 * f13 * f2 * f32 - d1 * F1 * f23 * f32 + d2 * F1 * f23 * f32 -
         2 * d1 * f24 * f32 - d2 * f24 * f32 - 2 * d1 * f13 * f33 + 2 * d2 * f13 * f33 - d1 * f12 * f2 * f33 -
         3 * d2 * f12 * f2 * f33 + 3 * d1 * f23 * f33 + d2 * f23 * f33 + d1 * f12 * f34 - d1 * F1 * f2 * f34 +
         d1 * F1 * f35 - d1 * f2 * f35 + 4 * f12 * f23 * v1 - 3 * F1 * f24 * v1 + 4 * F1 * f23 * f3 * v1 -
         3 * f24 * f3 * v1 - 12 * f12 * f2 * f32 * v1 + 4 * f23 * f32 * v1 + 8 * f12 * f33 * v1 - F1 * f34 * v1 -
         f35 * v1 - f15 * v2 - f14 * f3 * v2 + 8 * f13 * f32 * v2 - 8 * f12 * f33 * v2 + F1 * f34 * v2 + f35 * v2 +
         f15 * v3 - 4 * f12 * f23 * v3 + 3 * F1 * f24 * v3 + f14 * f3 * v3 - 4 * F1 * f23 * f3 * v3 + 3 * f24 * f3 * v3 -
         8 * f13 * f32 * v3 + 12 * f12 * f2 * f32 * v3 - 4 * f23 * f32 * v3) /
        den);
  put(PD_DL2, delta2);
  const T delta3 =
      -((-2. * d2 * f14 * f2 + d1 * f13 * f22 + 3 * d2 * f13 * f22 - d1 * F1 * f24 - d2 * F1 * f24 + 2 * d2 * f14 * f3 -
         2. * d1 * f13 * f2 * f3 - 2 * d2 * f13 * f2 * f3 + d1 * f12 * f22 * f3 - d2 * f12 * f22 * f3 + d1 * f24 * f3 +
         d2 * f24 * f3 + d1 * f13 * f32 - d2 * f13 * f32 - 2 * d1 * f12 * f2 * f32 + 2 * d2 * f12 * f2 * f32 +
         d1 * F1 * f22 * f32 - d2 * F1 * f22 * f32 + d1 * f12 * f33 - d2 * f12 * f33 + 2 * d1 * F1 * f2 * f33 +
         2 * d2 * F1 * f2 * f33 - 3 * d1 * f22 * f33 - d2 * f22 * f33 - 2 * d1 * F1 * f34 + 2 * d1 * f2 * f34 -
         4 * f12 * f22 * v1 + 2 * f24 * v1 + 8 * f12 * f2 * f3 * v1 - 4 * F1 * f22 * f3 * v1 - 4 * f12 * f32 * v1 +
         8 * F1 * f2 * f32 * v1 - 4 * f22 * f32 * v1 - 4 * F1 * f33 * v1 + 2 * f34 * v1 + 2 * f14 * v2 -
         4 * f13 * f3 * v2 + 4 * F1 * f33 * v2 - 2 * f34 * v2 - 2 * f14 * v3 + 4 * f12 * f22 * v3 - 2 * f24 * v3 +
         4 * f13 * f3 * v3 - 8 * f12 * f2 * f3 * v3 + 4 * F1 * f22 * f3 * v3 + 4 * f12 * f32 * v3 -
         8 * F1 * f2 * f32 * v3 + 4 * f22 * f32 * v3) /
        den);
  put(PD_DL3, delta3);
  const T delta4 =
      -((d2 * f13 * f2 - d1 * f12 * f22 - 2 * d2 * f12 * f22 + d1 * F1 * f23 + d2 * F1 * f23 - d2 * f13 * f3 +
         2. * d1 * f12 * f2 * f3 + d2 * f12 * f2 * f3 - d1 * F1 * f22 * f3 + d2 * F1 * f22 * f3 - d1 * f23 * f3 -
         d2 * f23 * f3 - d1 * f12 * f32 + d2 * f12 * f32 - d1 * F1 * f2 * f32 - 2 * d2 * F1 * f2 * f32 +
         2 * d1 * f22 * f32 + d2 * f22 * f32 + d1 * F1 * f33 - d1 * f2 * f33 + 3 * F1 * f22 * v1 - 2 * f23 * v1 -
         6 * F1 * f2 * f3 * v1 + 3 * f22 * f3 * v1 + 3 * F1 * f32 * v1 - f33 * v1 - f13 * v2 + 3 * f12 * f3 * v2 -
         3 * F1 * f32 * v2 + f33 * v2 + f13 * v3 - 3 * F1 * f22 * v3 + 2 * f23 * v3 - 3 * f12 * f3 * v3 +
         6 * F1 * f2 * f3 * v3 - 3 * f22 * f3 * v3) /
        den);
  put(PD_DL4, delta4);

  // fcut = 0.2 / (Mc G / eta^(3/5)) (waveforms.py:1333)
  put(PD_FCUT, mkc<N>(0.2 / (Mc.v * GMSUN_C3 / pow(eta.v, 3. / 5.))));
  put(PD_IETA, 1.0 / eta);
}

template <int N>
GWF_HD PdPhase<D<N>> pd_load_phase(const double* rec, int stride) {
  PdPhase<D<N>> c;
  c.pip = ldd<N>(rec + PD_PIP * stride); c.p23 = ldd<N>(rec + PD_P23 * stride); c.p13 = ldd<N>(rec + PD_P13 * stride);
  c.p13l = ldd<N>(rec + PD_P13L * stride); c.plog = ldd<N>(rec + PD_PLOG * stride);
  c.pm13 = ldd<N>(rec + PD_PM13 * stride); c.pm23 = ldd<N>(rec + PD_PM23 * stride); c.pm1 = ldd<N>(rec + PD_PM1 * stride);
  c.pm53 = ldd<N>(rec + PD_PM53 * stride);
  c.s1 = ldd<N>(rec + PD_S1 * stride); c.s43 = ldd<N>(rec + PD_S43 * stride); c.s53 = ldd<N>(rec + PD_S53 * stride);
  c.s2 = ldd<N>(rec + PD_S2 * stride);
  c.b1 = ldd<N>(rec + PD_B1 * stride); c.b2 = ldd<N>(rec + PD_B2 * stride); c.b3 = ldd<N>(rec + PD_B3 * stride);
  c.c1i = ldd<N>(rec + PD_C1I * stride); c.c2i = ldd<N>(rec + PD_C2I * stride);
  c.a1 = ldd<N>(rec + PD_A1 * stride); c.a2 = ldd<N>(rec + PD_A2 * stride); c.a3 = ldd<N>(rec + PD_A3 * stride);
  c.a4 = ldd<N>(rec + PD_A4 * stride); c.a5 = ldd<N>(rec + PD_A5 * stride);
  c.c1m = ldd<N>(rec + PD_C1M * stride); c.c2m = ldd<N>(rec + PD_C2M * stride);
  c.fring = ldd<N>(rec + PD_FRING * stride); c.fdamp = ldd<N>(rec + PD_FDAMP * stride);
  c.fmrdj = ldd<N>(rec + PD_FMRDJ * stride); c.eta = ldd<N>(rec + PD_ETA * stride);
  return c;
}

// ---- IMRPhenomD_NRTidalv2 per-point terms (waveforms.py:1582-1616, 1744-1799)
// (pi x)^(2/3) from x^(1/3)
template <int N>
GWF_HD D<N> nrt_px23(const D<N>& x13) {
  const D<N> p13 = x13 * 1.4645918875615231;  // pi^(1/3)
  return p13 * p13;
}

// tidal phase (arXiv:1905.06011 eq. 7) + 3.5PN spin-squared / spin-cubed terms
template <int N, class X>
GWF_HD D<N> nrt_phase_extra(const double* rec, int stride, const X& x, const X& x13) {
  const auto px23 = nrt_px23(x13);
  const auto px = PI * x;
  const auto px43 = px * (x13 * 1.4645918875615231);
  const auto px53 = px * px23;
  const auto num = 1.0 + (-12.615214237993088 * px23) + (19.0537346970349 * px) + (-21.166863146081035 * px43) +
                   (90.55082156324926 * px53) + (-60.25357801943598 * px * px);
  const auto den = 1.0 + (-15.11120782773667 * px23) + (22.195327350624694 * px) + (8.064109635305156 * px43);
  return ldd<N>(rec + PD_TPH * stride) * px53 * num / den + ldd<N>(rec + PD_SSS * stride) * px23;
}

// NaN -> 0, +-inf -> +-DBL_MAX (numpy.nan_to_num)
GWF_HD double nan_to_num(double v) {
  if (v != v) return 0.0;
  if (v > 1.79769313486231570815e308) return 1.79769313486231570815e308;
  if (v < -1.79769313486231570815e308) return -1.79769313486231570815e308;
  return v;
}

// Planck taper ending at 1.2 y (waveforms.py:1777-1797): value and the
// derivative with respect to y; the reference's custom JVP propagates no
// tangent through x.  At exactly x == 1.2 y the reference's expression
// evaluates to 1 (exp(+inf)) instead of its limit 0; the grid's last point
// M G fcut lands there or a few ulp away depending on the rounding of log10
// and pow, so the end point is taken as part of the cut here (value 0).
GWF_HD void planck_taper(double x, double y, double& tv, double& dv) {
  const double a = 1.2;
  const double yp = a * y;
  if (x < y) {
    tv = 1.0;
    dv = 0.0;
    return;
  }
  if (!(x < yp)) {
    tv = 0.0;
    dv = 0.0;
    return;
  }
  const double z = (yp - y) / (x - y) + (yp - y) / (x - yp);
  const double ez = exp(z);
  tv = nan_to_num(1. - 1. / (ez + 1.));
  dv = nan_to_num(ez * ((-1. + a) / (x - y) + (-1. + a) / (x - yp) + (-y + yp) / ((x - y) * (x - y)) +
                        1.2 * (-y + yp) / ((x - yp) * (x - yp))) /
                  ((ez + 1.) * (ez + 1.)));
}

// tidal amplitude (arXiv:1905.06011 eq. 24) without 1/dL
template <int N, class X>
GWF_HD D<N> nrt_amp_tidal(const double* rec, int stride, const X& x13) {
  const auto xT = nrt_px23(x13);
  const auto xT2 = xT * xT;
  const auto xT325 = xT2 * xT * dsqrt(dsqrt(xT));
  const auto xT289 = dexp(2.89 * dlog(xT));
  const auto poly = (1.0 + 4.157407407407407 * xT + 2519.111111111111 * xT289) / (1. + 13477.8073677 * (xT2 * xT2));
  return ldd<N>(rec + PD_TAMP * stride) * xT325 * poly;
}

GWF_HD double nrt_amp_tidal_val(const double* rec, int stride, double x13) {
  const double p13 = x13 * 1.4645918875615231;
  const double xT = p13 * p13;
  const double xT2 = xT * xT;
  const double xT325 = xT2 * xT * sqrt(sqrt(xT));
  const double xT289 = exp(2.89 * log(xT));
  const double poly = (1.0 + 4.157407407407407 * xT + 2519.111111111111 * xT289) / (1. + 13477.8073677 * (xT2 * xT2));
  return rec[PD_TAMP * stride] * xT325 * poly;
}

// Per-frequency IMRPhenomD phase Phi(f) as a dual (waveforms.py:1139-1147).
// Only the coefficients of the branch that x falls in are read.  NRT adds
// the NRTidalv2 terms inside the Mf < 0.2 window (waveforms.py:1613-1616).
// x, x13, lx: duals over the first tangents only (D<2>, functions of Mc and
// eta) or D<N>; the coefficients are D<N>.
template <int N, bool NRT = false, class X = D<N>>
GWF_HD D<N> pd_phi_at(const double* rec, int stride, const X& x, const X& x13, const X& lx, bool apply_fcut) {
  typedef D<N> T;
  const double xv = x.v;
#define GWF_L(slot) ldd<N>(rec + (slot) * stride)
  const T ieta = GWF_L(PD_IETA);
  T phis;
  if (xv < 0.018) {
    // the inspiral phase as polynomials in x^(1/3) and x^(-1/3) (Horner):
    // PIP + P13 x^1/3 + P23 x^2/3 + (S1 x + S43 x^4/3 + S53 x^5/3 + S2 x^2) / eta
    // + PM13 x^-1/3 + PM23 x^-2/3 + PM1 x^-1 + PM53 x^-5/3 + (P13L x^1/3 + PLOG) ln(pi x) / 3
    const auto lpx3 = (lx + log(PI)) * (1.0 / 3.0);
    const auto ix13 = 1.0 / x13;
    T pos = GWF_L(PD_S2);
    dhorner(pos, x13, GWF_L(PD_S53));
    dhorner(pos, x13, GWF_L(PD_S43));
    dhorner(pos, x13, GWF_L(PD_S1));
    pos = pos * ieta;
    dhorner(pos, x13, GWF_L(PD_P23));
    dhorner(pos, x13, GWF_L(PD_P13));
    dhorner(pos, x13, GWF_L(PD_PIP));
    T neg = GWF_L(PD_PM53) * ix13;  // no x^-4/3 term
    dhorner(neg, ix13, GWF_L(PD_PM1));
    dhorner(neg, ix13, GWF_L(PD_PM23));
    dhorner(neg, ix13, GWF_L(PD_PM13));
    T lg = GWF_L(PD_P13L);
    dhorner(lg, x13, GWF_L(PD_PLOG));
    phis = pos + neg * ix13 + lg * lpx3;
  } else if (xv < rec[PD_FMRDJ * stride]) {
    const auto ix = 1.0 / x;
    phis = (GWF_L(PD_B1) * x - GWF_L(PD_B3) * (ix * ix * ix) * (1.0 / 3.0) + GWF_L(PD_B2) * lx) * ieta +
           GWF_L(PD_C1I) + GWF_L(PD_C2I) * x;
  } else if (apply_fcut && !(xv < 0.2)) {
    phis = mkc<N>(0.0);
  } else {
    const auto x14 = dsqrt(dsqrt(x));
    phis = (-(GWF_L(PD_A2) * (1.0 / x)) + (4.0 / 3.0) * (GWF_L(PD_A3) * (x14 * x14 * x14)) + GWF_L(PD_A1) * x +
            GWF_L(PD_A4) * datan((x - GWF_L(PD_A5) * GWF_L(PD_FRING)) / GWF_L(PD_FDAMP))) *
               ieta +
           GWF_L(PD_C1M) + GWF_L(PD_C2M) * x;
  }
  if (apply_fcut) {
    if (xv < 0.2) {
      if (NRT)
        return phis + (-GWF_L(PD_T0) * (x - GWF_L(PD_FREF)) - GWF_L(PD_PHIREF) + nrt_phase_extra<N>(rec, stride, x, x13));
      return phis + (-GWF_L(PD_T0) * (x - GWF_L(PD_FREF)) - GWF_L(PD_PHIREF));
    }
    return phis + 0.0;
  }
  if (NRT) return phis - GWF_L(PD_T0) * (x - GWF_L(PD_FREF)) - GWF_L(PD_PHIREF) + nrt_phase_extra<N>(rec, stride, x, x13);
  return phis - GWF_L(PD_T0) * (x - GWF_L(PD_FREF)) - GWF_L(PD_PHIREF);
#undef GWF_L
}

template <int N, bool NRT = false, class X = D<N>>
GWF_HD D<N> pd_amp_at(const double* rec, int stride, const X& x, const X& x13, bool apply_fcut) {
  typedef D<N> T;
  const double xv = x.v;
  T amp;
  if (xv < 0.014) {
    // 1 + x^2/3 (AM23 + AM1 x^1/3 + AM43 x^2/3 + AM53 x + AM2 x^4/3 + AM73 x^5/3
    // + AM83 x^2 + AM3 x^7/3), the polynomial in x^(1/3) by Horner's rule
    T q = ldd<N>(rec + PD_AM3 * stride);
    dhorner(q, x13, ldd<N>(rec + PD_AM83 * stride));
    dhorner(q, x13, ldd<N>(rec + PD_AM73 * stride));
    dhorner(q, x13, ldd<N>(rec + PD_AM2 * stride));
    dhorner(q, x13, ldd<N>(rec + PD_AM53 * stride));
    dhorner(q, x13, ldd<N>(rec + PD_AM43 * stride));
    dhorner(q, x13, ldd<N>(rec + PD_AM1 * stride));
    dhorner(q, x13, ldd<N>(rec + PD_AM23 * stride));
    amp = 1. + (x13 * x13) * q;
  } else if (xv < ldd<N>(rec + PD_FPEAKA * stride).v) {
    amp = ldd<N>(rec + PD_DL0 * stride) + x * ldd<N>(rec + PD_DL1 * stride) +
          x * x * (ldd<N>(rec + PD_DL2 * stride) + x * ldd<N>(rec + PD_DL3 * stride) + x * x * ldd<N>(rec + PD_DL4 * stride));
  } else if (apply_fcut && !(xv < 0.2)) {
    amp = mkc<N>(0.0);
  } else {
    const T fring = ldd<N>(rec + PD_FRING * stride);
    const T fdamp = ldd<N>(rec + PD_FDAMP * stride);
    const T g1 = ldd<N>(rec + PD_G1 * stride);
    const T g2 = ldd<N>(rec + PD_G2 * stride);
    const T g3 = ldd<N>(rec + PD_G3 * stride);
    const T fdg3 = fdamp * g3;
    amp = dexp(-(x - fring) * g2 / fdg3) * (fdg3 * g1) / ((x - fring) * (x - fring) + fdg3 * fdg3);
  }
  // x^(-7/6) = 1 / (x sqrt(x^(1/3)))
  const auto xm76 = 1.0 / (x * dsqrt(x13));
  if (!NRT) return ldd<N>(rec + PD_OAMP * stride) * xm76 * amp;
  // NRTidalv2: (Overallamp amp0 x^-7/6 A_IMR + 2 sqrt(pi/5) Overallamp A_tidal) * taper
  const T A = ldd<N>(rec + PD_OAMP * stride) * xm76 * amp + nrt_amp_tidal<N>(rec, stride, x13);
  const T fm = ldd<N>(rec + PD_FMERG * stride);
  double tv, dv;
  planck_taper(xv, fm.v, tv, dv);
  T r;
  r.v = A.v * tv;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = A.d[k] * tv + A.v * dv * fm.d[k];
  return r;
}

// Value-only IMRPhenomD amplitude (the SNR path): reads only .v of the record.
// x13 = cbrt(x), computed once by the caller (the SNR path reuses it for tau)
GWF_HD double pd_amp_val(const double* rec, int stride, double x, double x13, bool apply_fcut, bool nrt = false) {
  double amp;
  if (x < 0.014) {  // Horner in x^(1/3), as pd_amp_at
    double q = rec[PD_AM3 * stride];
    q = fma(x13, q, rec[PD_AM83 * stride]);
    q = fma(x13, q, rec[PD_AM73 * stride]);
    q = fma(x13, q, rec[PD_AM2 * stride]);
    q = fma(x13, q, rec[PD_AM53 * stride]);
    q = fma(x13, q, rec[PD_AM43 * stride]);
    q = fma(x13, q, rec[PD_AM1 * stride]);
    q = fma(x13, q, rec[PD_AM23 * stride]);
    amp = fma(x13 * x13, q, 1.0);
  } else if (x < rec[PD_FPEAKA * stride]) {
    amp = rec[PD_DL0 * stride] + x * rec[PD_DL1 * stride] +
          x * x * (rec[PD_DL2 * stride] + x * rec[PD_DL3 * stride] + x * x * rec[PD_DL4 * stride]);
  } else if (apply_fcut && !(x < 0.2)) {
    amp = 0.0;
  } else {
    const double fring = rec[PD_FRING * stride], fdamp = rec[PD_FDAMP * stride];
    const double g1 = rec[PD_G1 * stride], g2 = rec[PD_G2 * stride], g3 = rec[PD_G3 * stride];
    const double fdg3 = fdamp * g3;
    amp = exp(-(x - fring) * g2 / fdg3) * (fdg3 * g1) / ((x - fring) * (x - fring) + fdg3 * fdg3);
  }
  // x^(-7/6) = 1 / (x * x^(1/6)), x^(1/6) = sqrt(x^(1/3))
  if (!nrt) return rec[PD_OAMP * stride] * (1.0 / (x * sqrt(x13))) * amp;
  const double A = rec[PD_OAMP * stride] * (1.0 / (x * sqrt(x13))) * amp + nrt_amp_tidal_val(rec, stride, x13);
  double tv, dv;
  planck_taper(x, rec[PD_FMERG * stride], tv, dv);
  return A * tv;
}

// ------------------------------------------------------------- TaylorF2
template <class T>
GWF_HD T quadmon(const T& L) {
  // waveforms.py:771-772
  if (val(L) < 1.) return 1. + L * (0.427688866723244 + L * (-0.324336526985068 + L * 0.1107439432180572));
  const T lL = dlog(L);
  return dexp(0.1940 + 0.09163 * lL + 0.04812 * lL * lL - 4.283e-3 * lL * lL * lL + 1.245e-4 * lL * lL * lL * lL);
}

// Lambda1, Lambda2 from (LambdaTilde, deltaLambda) (gwfastUtils.py:419-444).
template <class T>
GWF_HD void lam12_from_lamt(const T& Lamt, const T& delLam, const T& eta, T& L1, T& L2) {
  const T eta2 = eta * eta;
  const T Seta = seta_of(eta);
  const T mLp = (8. / 13.) * (1. + 7. * eta - 31. * eta2);
  const T mLm = (8. / 13.) * Seta * (1. + 9. * eta - 11. * eta2);
  const T mdp = Seta * (1. - (13272. / 1319.) * eta + (8944. / 1319.) * eta2) * 0.5;
  const T mdm = (1. - (15910. / 1319.) * eta + (32850. / 1319.) * eta2 + (3380. / 1319.) * (eta2 * eta)) * 0.5;
  const T det = (306656. / 1319.) * dpow(eta, 5.) - (5936. / 1319.) * dpow(eta, 4.);
  L1 = ((mdp - mdm) * Lamt + (mLm - mLp) * delLam) / det;
  L2 = ((-mdm - mdp) * Lamt + (mLm + mLp) * delLam) / det;
}

template <class T>
GWF_HD void lamt_from_lam12(const T& L1, const T& L2, const T& eta, T& Lamt, T& delLam) {
  // gwfastUtils.py:398-417
  const T eta2 = eta * eta;
  const T Seta = seta_of(eta);
  Lamt = (8. / 13.) * ((1. + 7. * eta - 31. * eta2) * (L1 + L2) + Seta * (1. + 9. * eta - 11. * eta2) * (L1 - L2));
  delLam = 0.5 * (Seta * (1. - 13272. / 1319. * eta + 8944. / 1319. * eta2) * (L1 + L2) +
                  (1. - 15910. / 1319. * eta + 32850. / 1319. * eta2 + 3380. / 1319. * eta2 * eta) * (L1 - L2));
}

// TaylorF2 fcut (waveforms.py:903-953), value only.
GWF_HD double tf2_fcut(const gwf_waveform& w, double Mc, double eta, double chi1, double chi2) {
  if (!w.kerr_isco) return w.fcut_par / (Mc / pow(eta, 3. / 5.));
  const double eta2 = eta * eta;
  const double Mtot = Mc / pow(eta, 3. / 5.);
  const double Seta = sqrt(eta < 0.25 ? 1.0 - 4.0 * eta : 0.0);
  const double m1 = 0.5 * (1.0 + Seta), m2 = 0.5 * (1.0 - Seta);
  const double s = (m1 * m1 * chi1 + m2 * m2 * chi2) / (m1 * m1 + m2 * m2);
  const double atot = (chi1 + chi2 * (m2 / m1) * (m2 / m1)) / ((1. + m2 / m1) * (1. + m2 / m1));
  const double aeff = atot + 0.41616 * eta * (chi1 + chi2);
  auto r_isco = [](double chi) {
    const double Z1 = 1.0 + pow(1.0 - chi * chi, 1. / 3.) * (pow(1.0 + chi, 1. / 3.) + pow(1.0 - chi, 1. / 3.));
    const double Z2 = sqrt(3.0 * chi * chi + Z1 * Z1);
    return chi > 0. ? 3.0 + Z2 - sqrt((3.0 - Z1) * (3.0 + Z1 + 2.0 * Z2)) : 3.0 + Z2 + sqrt((3.0 - Z1) * (3.0 + Z1 + 2.0 * Z2));
  };
  const double rI = r_isco(aeff);
  const double EradNS = eta * (0.055974469826360077 + 0.5809510763115132 * eta - 0.9606726679372312 * eta2 +
                               3.352411249771192 * eta2 * eta);
  const double EradTot = (EradNS * (1. + (-0.0030302335878845507 - 2.0066110851351073 * eta + 7.7050567802399215 * eta2) * s)) /
                         (1. + (-0.6714403054720589 - 1.4756929437702908 * eta + 7.304676214885011 * eta2) * s);
  const double Mfin = Mtot * (1. - EradTot);
  const double L_ISCO = 2. / (3. * sqrt(3.)) * (1. + 2. * sqrt(3. * rI - 2.));
  const double E_ISCO = sqrt(1. - 2. / (3. * rI));
  const double chif = atot + eta * (L_ISCO - 2. * atot * (E_ISCO - 1.)) +
                      (-3.821158961 - 1.2019 * aeff - 1.20764 * aeff * aeff) * eta2 +
                      (3.79245 + 1.18385 * aeff + 4.90494 * aeff * aeff) * eta2 * eta;
  const double Om = 1. / (pow(r_isco(chif), 3. / 2.) + chif);
  return Om / (PI * Mfin * GMSUN_C3);
}

// Per-event TaylorF2 setup.  N = 4 (Mc, eta, chi1, chi2) or 6 (+ LambdaTilde,
// deltaLambda; tidal).  The phase is stored as coefficients of
// v^-5, v^-3, v^-2, v^-1, 1, ln v, v, v ln v, v^2, v^5, v^7 (all times
// 3/(128 eta)), so Phi(f) = sum c_k b_k(v) + phiR - pi/4.
// Eccentric TaylorF2 is the TaylorF2 instantiation with N = 5 tangents
// (Mc, eta, chi1z, chi2z, ecc); the tidal one has N = 6.
constexpr bool tf2_eccentric(int N) { return N == 5 || N == 7; }  // 7: tidal and eccentric

template <int N>
GWF_HD void tf2_setup(const gwf_waveform& w, const D<N>& Mc, const D<N>& eta, const D<N>& chi1, const D<N>& chi2,
                      const D<N>& Lamt, const D<N>& delLam, double* rec, int stride, const D<N>& ecc = D<N>(),
                      double f0 = 1.0) {
  typedef D<N> T;
  const T MG = Mc * GMSUN_C3 / dpow(eta, 3. / 5.);
  const T Seta = seta_of(eta);
  T QM1 = mkc<N>(1.0), QM2 = mkc<N>(1.0);
  T L1 = mkc<N>(0.0), L2 = mkc<N>(0.0);
  if (w.is_tidal) {
    lam12_from_lamt(Lamt, delLam, eta, L1, L2);
    if (w.use_quadmon) {
      QM1 = quadmon(L1);
      QM2 = quadmon(L2);
    }
  }
  const Tf2Coeffs<T> c = tf2_coeffs(eta, Seta, chi1, chi2, QM1, QM2, w.use_3p5PN_SpinHO != 0);
  const T TF2OA = 3. / (128. * eta);
  const T base5 = (38645. * PI / 756. - 65. * PI * eta / 9. - c.five_tmp);
  const double vlso = 1. / sqrt(6.);
  T five;
  double phiR;
  if (w.phiref_vlso) {
    five = base5 * (1. - 3. * log(vlso));
    phiR = 0.;
  } else {
    five = base5;
    phiR = PI;
  }
  const T five_log = base5 * 3.;
  T t10 = mkc<N>(0.0), t12 = mkc<N>(0.0);
  if (w.is_tidal) {
    // Phi converts back to (LambdaTilde, deltaLambda) (waveforms.py:852-856)
    T Lt, dLm;
    lamt_from_lam12(L1, L2, eta, Lt, dLm);
    t10 = (-0.5 * 39. * Lt);
    t12 = (-3115. / 64. * Lt + 6595. / 364. * Seta * dLm);
  }
  double* r = rec;
  auto put = [&](int slot, const T& x) { std_(r + slot * stride, x); };
  put(TF_MG, MG);
  put(TF_ETA, eta);
  put(TF_C0, TF2OA);
  put(TF_C2, TF2OA * c.two);
  put(TF_C3, TF2OA * c.three);
  put(TF_C4, TF2OA * c.four);
  put(TF_C5, TF2OA * five);
  put(TF_C5L, TF2OA * five_log);
  put(TF_C6, TF2OA * c.six);
  put(TF_C6L, TF2OA * (-(6848. / 21.)));
  put(TF_C7, TF2OA * c.seven);
  put(TF_C10, TF2OA * t10);
  put(TF_C12, TF2OA * t12);
  put(TF_PHIR, mkc<N>(phiR - PI * 0.25));
  // sqrt(5/24) pi^(-2/3) clightGpc (G Mc)^(5/6)  (waveforms.py:875, without 1/dL)
  put(TF_AMP, sqrt(5. / 24.) * pow(PI, -2. / 3.) * CLIGHT_GPC * dpow(GMSUN_C3 * Mc, 5. / 6.));
  put(TF_FCUT, mkc<N>(tf2_fcut(w, Mc.v, eta.v, chi1.v, chi2.v)));
  if (tf2_eccentric(N)) {
    // low-eccentricity 3PN phase of arXiv:1605.00304 as written in waveforms.py:814-845;
    // v0 at fRef_ecc, else at the grid minimum (amin of v over the grid)
    const double fe = (w.fref_ecc > 0.0) ? w.fref_ecc : f0;
    const T v0 = dcbrt(PI * MG * fe);
    const T v02 = v0 * v0, v03 = v02 * v0, v04 = v02 * v02, v05 = v04 * v0, v06 = v03 * v03;
    const T eta2 = eta * eta, eta3 = eta2 * eta;
    const double eg = 0.57721566490153286061, l2 = log(2.), l3 = log(3.);
    const T twoV = 29.9076223 / 8.1976608 + 18.766963 / 2.927736 * eta;
    const T twoV0 = 2.833 / 1.008 - 19.7 / 3.6 * eta;
    const double threeV = -28.19123 / 2.82600 * PI, threeV0 = 37.7 / 7.2 * PI;
    const T fourV4 = 16.237683263 / 3.330429696 + 241.33060753 / 9.71375328 * eta + 156.2608261 / 6.9383952 * eta2;
    const T fourV2V02 = 84.7282939759 / 8.2632420864 - 7.18901219 / 3.68894736 * eta - 36.97091711 / 1.05398496 * eta2;
    const T fourV04 = -1.193251 / 3.048192 - 66.317 / 9.072 * eta + 18.155 / 1.296 * eta2;
    const T fiveV5 = -28.31492681 / 1.18395270 * PI - 115.52066831 / 2.70617760 * PI * eta;
    const T fiveV3V02 = -79.86575459 / 2.84860800 * PI + 55.5367231 / 1.0173600 * PI * eta;
    const T fiveV2V03 = 112.751736071 / 5.902315776 * PI + 70.75145051 / 2.10796992 * PI * eta;
    const T fiveV05 = 76.4881 / 9.0720 * PI - 94.9457 / 2.2680 * PI * eta;
    const T sixV6 = -436.03153867072577087 / 1.32658535116800000 + 53.6803271 / 1.9782000 * eg +
                    157.22503703 / 3.25555200 * PI * PI +
                    (2991.72861614477 / 6.89135247360 - 15.075413 / 1.446912 * PI * PI) * eta +
                    345.5209264991 / 4.1019955200 * eta2 + 506.12671711 / 8.78999040 * eta3 +
                    384.3505163 / 5.9346000 * l2 - 112.1397129 / 1.7584000 * l3;
    const T sixV4V02 = 46.001356684079 / 3.357073133568 + 253.471410141755 / 5.874877983744 * eta -
                       169.3852244423 / 2.3313007872 * eta2 - 307.833827417 / 2.497822272 * eta3;
    const double sixV3V03 = -106.2809371 / 2.0347200 * PI * PI;
    const T sixV2V04 = -3.56873002170973 / 2.49880440692736 - 260.399751935005 / 8.924301453312 * eta +
                       15.0484695827 / 3.5413894656 * eta2 + 340.714213265 / 3.794345856 * eta3;
    const T sixV06 = 265.31900578691 / 1.68991764480 - 33.17 / 1.26 * eg + 12.2833 / 1.0368 * PI * PI +
                     (91.55185261 / 5.48674560 - 3.977 / 1.152 * PI * PI) * eta - 5.732473 / 1.306368 * eta2 -
                     30.90307 / 1.39968 * eta3 + 87.419 / 1.890 * l2 - 260.01 / 5.60 * l3;
    const double k6 = 53.6803271 / 3.9564000, k0 = 33.17 / 2.52;
    // overall -2.355/1.462 e^2 (v0/v)^(19/3), over v^5 with the rest of the phase
    const T K = TF2OA * (-2.355 / 1.462) * ecc * ecc * dpow(v0, 19. / 3.);
    put(TF_E0, K * (1. + twoV0 * v02 + threeV0 * v03 + fourV04 * v04 + fiveV05 * v05 +
                    (sixV06 - k0 * dlog(16. * v02)) * v06));
    put(TF_E2, K * (twoV + fourV2V02 * v02 + fiveV2V03 * v03 + sixV2V04 * v04));
    put(TF_E3, K * (threeV + fiveV3V02 * v02 + sixV3V03 * v03));
    put(TF_E4, K * (fourV4 + sixV4V02 * v02));
    put(TF_E5, K * fiveV5);
    put(TF_E6, K * (sixV6 + k6 * log(16.)));
    put(TF_E6L, K * (2. * k6));  // k6 ln(16 v^2) = k6 ln 16 + 2 k6 ln v
  }
}

// Per-frequency TaylorF2 phase.  v is (pi MG f)^(1/3) as a dual.
// v, lv: duals over the first tangents only (D<2>) or D<N>.
template <int N, class V = D<N>>
GWF_HD D<N> tf2_phi_at(const double* rec, int stride, const V& v, const V& lv) {
  typedef D<N> T;
  const auto iv = 1.0 / v;
  const auto iv2 = iv * iv;
  const auto iv3 = iv2 * iv;
  const auto iv5 = iv3 * iv2;
  const auto v2 = v * v;
  const auto v5 = v2 * v2 * v;
  const auto v7 = v5 * v2;
  T ph = ldd<N>(rec + TF_C0 * stride) * iv5 + ldd<N>(rec + TF_C2 * stride) * iv3 + ldd<N>(rec + TF_C3 * stride) * iv2 +
         ldd<N>(rec + TF_C4 * stride) * iv + ldd<N>(rec + TF_C5 * stride) + ldd<N>(rec + TF_C5L * stride) * lv +
         ldd<N>(rec + TF_C6 * stride) * v + ldd<N>(rec + TF_C6L * stride) * (v * lv) + ldd<N>(rec + TF_C7 * stride) * v2 +
         ldd<N>(rec + TF_C10 * stride) * v5 + ldd<N>(rec + TF_C12 * stride) * v7 + rec[TF_PHIR * stride];
  if (tf2_eccentric(N)) {
    // v^(-34/3) (E0 + E2 v^2 + E3 v^3 + E4 v^4 + E5 v^5 + (E6 + E6L ln v) v^6)
    const double p = pow(v.v, -34. / 3.);
    const auto w = chain(v, p, (-34. / 3.) * p / v.v);
    const auto v3 = v2 * v, v4 = v2 * v2, v6 = v3 * v3;
    ph = ph + w * (ldd<N>(rec + TF_E0 * stride) + ldd<N>(rec + TF_E2 * stride) * v2 + ldd<N>(rec + TF_E3 * stride) * v3 +
                   ldd<N>(rec + TF_E4 * stride) * v4 + ldd<N>(rec + TF_E5 * stride) * v5 +
                   (ldd<N>(rec + TF_E6 * stride) + ldd<N>(rec + TF_E6L * stride) * lv) * v6);
  }
  return ph;
}

// --------------------------------------------------- IMRPhenomD_NRTidalv2
// f_merger of the Planck taper (waveforms.py:1760-1773, 1818-1830).
template <class T>
GWF_HD T nrt_fmerger(const T& eta, const T& Seta, const T& kappa2T) {
  const T q = 0.5 * (1.0 + Seta - 2.0 * eta) / eta;
  const T k2 = kappa2T * kappa2T;
  const T numPT = 1.0 + 3.35411203e-2 * kappa2T + 4.31460284e-5 * k2;
  const T denPT = 1.0 + 7.54224145e-2 * kappa2T + 2.23626859e-4 * k2;
  const T Q0 = 0.3586 / dsqrt(q);
  return Q0 * (numPT / denPT) / (2. * PI);
}

// kappa2T (arXiv:1905.06011 eq. 8; waveforms.py:1594, 1749, 1816)
template <class T>
GWF_HD T nrt_kappa2T(const T& Seta, const T& L1, const T& L2) {
  const T Xa = 0.5 * (1.0 + Seta);
  const T Xb = 0.5 * (1.0 - Seta);
  // X**5 as pow (the reference's expression): the taper's end 1.2 f_merger and
  // the grid's last point M G fcut are compared at ulp level
  return (3.0 / 13.0) * ((1.0 + 12.0 * Xb / Xa) * dpow(Xa, 5.) * L1 + (1.0 + 12.0 * Xa / Xb) * dpow(Xb, 5.) * L2);
}

// fcut = 1.2 f_merger / (M G) (waveforms.py:1803-1835), value only, from
// (LambdaTilde, deltaLambda).
GWF_HD double nrt_fcut(double Mc, double eta, double Lt, double dLm) {
  const double Seta = sqrt(eta < 0.25 ? 1.0 - 4.0 * eta : 0.0);
  double L1, L2;
  lam12_from_lamt(Lt, dLm, eta, L1, L2);
  const double fm = nrt_fmerger(eta, Seta, nrt_kappa2T(Seta, L1, L2));
  const double M = Mc / pow(eta, 3. / 5.);
  return (1.2 * fm) / (M * GMSUN_C3);
}

// The NRTidalv2 slots of an event record (after pd_setup with the same
// quadrupole moments): tidal-phase factor, 3.5PN SS + SSS coefficient,
// tidal-amplitude factor, merger frequency, and the model's fcut.
template <int N>
GWF_HD void nrt_setup(const D<N>& Mc, const D<N>& eta, const D<N>& chi1, const D<N>& chi2, const D<N>& L1,
                      const D<N>& L2, const D<N>& QM1, const D<N>& QM2, double* rec, int stride, int tk = -1) {
  typedef D<N> T;
  const T Seta = seta_of(eta);
  const T m1ByM = 0.5 * (1.0 + Seta);
  const T m2ByM = 0.5 * (1.0 - Seta);
  const T chi12 = chi1 * chi1, chi22 = chi2 * chi2;
  const T kappa2T = nrt_kappa2T(Seta, L1, L2);
  const T TF2OA = 3. / (128. * eta);
  // octupole moments minus the BBH baseline (waveforms.py:1607-1608)
  auto octmon = [](const T& Q) {
    const T lq = dlog(Q);
    return -1. + dexp(0.003131 + 2.071 * lq - 0.7152 * lq * lq + 0.2458 * lq * lq * lq - 0.03309 * lq * lq * lq * lq);
  };
  const T Oct1 = octmon(QM1), Oct2 = octmon(QM2);
  const T SS = -400. * PI * (QM1 - 1.) * chi12 * m1ByM * m1ByM - 400. * PI * (QM2 - 1.) * chi22 * m2ByM * m2ByM;
  const T SSS = 10. * ((m1ByM * m1ByM + 308. / 3. * m1ByM) * chi1 + (m2ByM * m2ByM - 89. / 3. * m2ByM) * chi2) *
                    (QM1 - 1.) * m1ByM * m1ByM * chi12 +
                10. * ((m2ByM * m2ByM + 308. / 3. * m2ByM) * chi2 + (m1ByM * m1ByM - 89. / 3. * m1ByM) * chi1) *
                    (QM2 - 1.) * m2ByM * m2ByM * chi22 -
                440. * Oct1 * m1ByM * m1ByM * m1ByM * chi12 * chi1 - 440. * Oct2 * m2ByM * m2ByM * m2ByM * chi22 * chi2;
  const T M = Mc / dpow(eta, 3. / 5.);
  // Overallamp without 1/dL (waveforms.py:1248) times 2 sqrt(pi/5) and -9 kappa2T
  const T OA = 2. * sqrt(5. / (64. * PI)) * M * GMSUN_C2_GPC * M * GMSUN_C3;
  const T fm = nrt_fmerger(eta, Seta, kappa2T);
  auto put = [&](int slot, const T& x) { std_tk(rec + slot * stride, x, tk); };
  put(PD_TPH, -kappa2T * 2.4375 / (m1ByM * m2ByM));
  put(PD_SSS, (SS + SSS) * TF2OA);
  put(PD_TAMP, OA * (2. * sqrt(PI / 5.)) * (-9.0 * kappa2T));
  put(PD_FMERG, fm);
  put(PD_FCUT, mkc<N>((1.2 * fm.v) / (M.v * GMSUN_C3)));
}

}  // namespace gwf
