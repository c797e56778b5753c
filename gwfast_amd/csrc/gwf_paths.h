// Detector projection and the per-event bodies of the path's kernels.
//
// Each body processes the frequency points lane, lane + nlanes, ... of one
// event; the HIP kernels run it with one 64-wide wavefront per event and
// reduce across lanes.
//
// Projection follows gwfast/signal.py: _PatternFunction 340-387 (JKS a, b),
// _DeltLoc 401-423, GWAmplitudes 425-466, GWPhase 468-483, GWstrain 485-655.
// The Fisher body differentiates the strain in closed form per point: with
//   h = (Ap + i Ac) exp(i Psi),
// every derivative is dX h = [(dX Ap - Ac dX Psi) + i (dX Ac + Ap dX Psi)]
// exp(i Psi), and in Re(conj(da h) db h) the common factor exp(i Psi) cancels
// exactly, so the Gram never needs the (up to ~1e9 rad) phase itself.
#pragma once
#include "gwf_waveforms.h"

namespace gwf {

constexpr int SLOT = NT_MAX + 1;  // doubles per dual slot in a record
constexpr double INV_DAY = 1.0 / (3600. * 24.);

// -------------------------------------------------------------- event record
// [waveform slots: REC_SLOTS * SLOT][event constants][site constants]
enum EvConst {
  EV_IDL = 0, EV_DL, EV_CP, EV_CC, EV_DCP, EV_DCC, EV_TC, EV_RA, EV_FCUT, EV_PHIC, EV_CRA, EV_SRA,
  EV_LMG, EV_LMG_D0, EV_LMG_D1,  // ln(M G) and its (Mc, eta) tangents
  // tau_star = K v^-8 (1 + c2 v^2 + c3 v^3 + ... + (c6 + c6L ln v) v^6 + c7 v^7):
  // K = (5/256) M G / eta over (Mc, eta), the eta-dependent c_k and d c_k / d eta
  EV_TK, EV_TK_D0, EV_TK_D1, EV_TC2, EV_TC4, EV_TC4_D1, EV_TC5, EV_TC6, EV_TC6_D1, EV_TC7, EV_TC7_D1,
  EV_COUNT
};
enum SiteConst {
  SC_U1 = 0, SC_U2, SC_U3, SC_U1T, SC_U2T, SC_U3T,
  SC_E1, SC_E2, SC_E3, SC_E4, SC_E5, SC_G1, SC_G2, SC_G3, SC_G4,
  SC_E1T, SC_E2T, SC_E3T, SC_E4T, SC_E5T, SC_G1T, SC_G2T, SC_G3T, SC_G4T,
  SC_P1, SC_P2,
  SC_CL0, SC_SL0,  // cos/sin(lon + 2 pi t0), t0 = tcoal (fixed or rotating Earth) or 0 (noMotion)
  SC_COUNT
};
constexpr int REC_EV = REC_SLOTS * SLOT;
constexpr int REC_SITE = REC_EV + EV_COUNT;
constexpr int REC_DOUBLES = REC_SITE + GWF_MAX_DETECTORS * SC_COUNT;

// ------------------------------------------------------------ network (args)
enum Motion { MOTION_FIXED = 0, MOTION_EARTH = 1, MOTION_NONE = 2 };

struct PsdTab {
  const double* f;
  const double* S;
  int n;
  double lx0;  // log(f[0])
  double idl;  // (n-1) / (log f[n-1] - log f[0])
  double f_first, f_last;  // f[0], f[n-1]
};

struct SiteK {
  double clat, slat, s2lat, c2lat, lon, sinzeta;
  int motion;
  int psd;
  int arm0, narm;  // this site's physical arms: arm[arm0 .. arm0 + narm)
  int sum0, nsum;  // and its network-sum arms: sum_arm[sum0 .. sum0 + nsum)
};

struct ArmK {
  int site;
  double s2x, c2x;  // sin/cos(2 (xax + rot)); a derived (null-stream) arm holds -(sum of two)
  double wgt;       // Gram weight (1 for a physical arm)
  double swgt;      // sqrt(max(wgt, 0))
};

// Two arm tables: the physical arms (per-arm outputs, SNR, strain) and the
// arms the network Fisher sum runs over.  Every strain quantity of an arm is
// linear in its (s2x, c2x), so the sum over the arms k of one site of
// G(D_k) = G(s_k U + c_k V) equals lambda_1 G(D(q_1)) + lambda_2 G(D(q_2))
// with (lambda_j, q_j) the eigenpairs of sum_k (s_k, c_k)^T (s_k, c_k): a
// triangle's three arms reduce to two weighted arms.
struct NetK {
  int n_sites, n_arms, res, any_earth;
  int n_sum_arms, n_psd;  // n_psd: distinct PSD tables
  int lin;                // Fisher grids: linspace instead of geomspace
  int mixed_motion;       // rotating-Earth and fixed-Earth sites together
  unsigned spin_limit;    // polls per wait of the producer / consumer Fisher kernel
  double fmin, fmax;
  SiteK site[GWF_MAX_DETECTORS];
  ArmK arm[GWF_MAX_ARMS];
  ArmK sum_arm[2 * GWF_MAX_DETECTORS];
  PsdTab psd[GWF_MAX_DETECTORS];
};

// --------------------------------------------------------------------- grid
// jnp.geomspace(fmin, fcut, res) = 10**linspace(log10 fmin, log10 fcut)
// (JAX 0.4.30): lin_i = a (1 - i/div) + b (i/div), last point 10**b.
GWF_HD double grid_lin(double a, double b, int i, int div) {
  if (i >= div) return b;
  const double s = (double)i / (double)div;
  return a * (1 - s) + b * s;
}
GWF_HD double exp10d(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return exp10(x);
#else
  return pow(10.0, x);
#endif
}
// 10^x for |x| <= 300 (no overflow, underflow or special operands), within
// 2 ulp: x log2(10) = n + r with n an integer and |r| <= 1/2, the product in
// double-double (FMA), 2^r = e^u, u = r ln 2, by the Taylor series of e^u to
// the u^13 term (|u| <= 0.347: truncation below 4e-18), then 2^n exactly.
// The library exp10 spends most of its instructions on the special operands.
GWF_HD double exp10_kernel(double x) {
  constexpr double L2T_HI = 3.321928094887362, L2T_LO = 1.661617516973592e-16;  // log2(10) = HI + LO
  const double yh = x * L2T_HI;
  const double yl = fma(x, L2T_HI, -yh) + x * L2T_LO;
  const double n = rint(yh);
  const double u = ((yh - n) + yl) * 0.6931471805599453;
  double p = 1.0 / 6227020800.0;  // 1/13!
  p = fma(p, u, 1.0 / 479001600.0);
  p = fma(p, u, 1.0 / 39916800.0);
  p = fma(p, u, 1.0 / 3628800.0);
  p = fma(p, u, 1.0 / 362880.0);
  p = fma(p, u, 1.0 / 40320.0);
  p = fma(p, u, 1.0 / 5040.0);
  p = fma(p, u, 1.0 / 720.0);
  p = fma(p, u, 1.0 / 120.0);
  p = fma(p, u, 1.0 / 24.0);
  p = fma(p, u, 1.0 / 6.0);
  p = fma(p, u, 0.5);
  p = fma(p, u, 1.0);
  p = fma(p, u, 1.0);
  return ldexp(p, (int)n);
}

// 10^lin_i of grid point i: interior points by exp10_kernel, the end points
// (i = 0, i >= div) by the library exp10, the value numpy's 10**x gives there
// to the ulp in practice: a cut at the last point (IMRPhenomD's Mf = 0.2)
// compares that frequency with the cut.
GWF_HD double grid_pow10(double x, int i, int div) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (i == 0 || i >= div) ? exp10d(x) : exp10_kernel(x);
#else
  return exp10d(x);
#endif
}
GWF_HD double grid_f(double a, double b, int i, int div) { return grid_pow10(grid_lin(a, b, i, div), i, div); }

// End points of an event's grid: (log10 fmin, log10 fcut) for geomspace,
// (fmin, fcut) for linspace (jnp.linspace: f_i = fmin (1 - s) + fcut s, last
// point fcut), so that grid_lin gives log10 f or f itself.
GWF_HD void grid_ends(bool lin, double fmin, double fcut, double* a, double* b) {
  *a = lin ? fmin : log10(fmin);
  *b = lin ? fcut : log10(fcut);
}

// Trapezoid weight of point i (np.trapz / jnp.trapezoid as a weighted sum).
GWF_HD double trap_w(double a, double b, int i, int div, double fi) {
  if (div <= 0) return 0.0;
  const double fl = (i > 0) ? grid_f(a, b, i - 1, div) : fi;
  const double fr = (i < div) ? grid_f(a, b, i + 1, div) : fi;
  return 0.5 * (fr - fl);
}

// Walks a geometric grid in chunks of nlanes consecutive points (point
// base + lane in each lane) and yields each point's frequency and trapezoid
// weight.  On the device the neighbours come from the adjacent lanes
// (one exp10 per point instead of three); lane 0 takes the previous chunk's
// last point, lane 63 the next chunk's first point.
// LIN: linspace grid (a, b are fmin, fcut) instead of geomspace (log10 of them).
template <bool LIN = false>
struct GridWalk {
  double a, b;
  int div, lane, nlanes;
  double f_cur, f_next, lin_cur, lin_next, prev_last;
  GWF_HD void init(double a_, double b_, int res, int lane_, int nlanes_) {
    a = a_;
    b = b_;
    div = res - 1;
    lane = lane_;
    nlanes = nlanes_;
    const int i0 = lane < res ? lane : 0;
    lin_cur = grid_lin(a, b, i0, div);
    f_cur = LIN ? lin_cur : grid_pow10(lin_cur, i0, div);
    prev_last = 0.0;
  }
  // Call once per chunk with base = 0, nlanes, 2 nlanes, ... (all lanes);
  // returns f, its trapezoid weight and ln f of point base + lane.
  GWF_HD void step(int base, int res, double* f, double* w, double* lnf) {
    const int i = base + lane;
    const int inext = i + nlanes;
    const int in = inext < res ? inext : 0;
    lin_next = grid_lin(a, b, in, div);
    f_next = LIN ? lin_next : grid_pow10(lin_next, in, div);
#if defined(__HIP_DEVICE_COMPILE__)
    double fl = __shfl_up(f_cur, 1, 64);
    double fr = __shfl_down(f_cur, 1, 64);
    const double first_next = __shfl(f_next, 0, 64);
    if (lane == 0) fl = prev_last;
    if (lane == nlanes - 1) fr = first_next;
    prev_last = __shfl(f_cur, nlanes - 1, 64);
#else
    double fl = (i > 0) ? grid_lin(a, b, i - 1, div) : 0.0;
    double fr = (i < div) ? grid_lin(a, b, i + 1, div) : 0.0;
    if (!LIN) {
      fl = (i > 0) ? exp10d(fl) : 0.0;
      fr = (i < div) ? exp10d(fr) : 0.0;
    }
#endif
    const double fi = f_cur;
    *f = fi;
    *w = (div <= 0) ? 0.0 : 0.5 * (((i < div) ? fr : fi) - ((i > 0) ? fl : fi));
    *lnf = LIN ? log(fi) : lin_cur * 2.302585092994045684;
    f_cur = f_next;
    lin_cur = lin_next;
  }
};

// jnp.interp(f, xp, fp, left=1, right=1) on the PSD table; lnf = log(f).
GWF_HD double psd_at(const PsdTab& t, double f, double lnf) {
  const double* x = t.f;
  const int n = t.n;
  if (f < x[0] || f > x[n - 1]) return 1.0;
  // segment = largest j <= n-2 with x[j] <= f (searchsorted right, clipped):
  // a guess from log-uniform spacing, a short walk, then bisection for
  // tables that are not log-spaced (e.g. 0.5 Hz linear ASD tables)
  int j = (int)((lnf - t.lx0) * t.idl);
  j = j < 0 ? 0 : (j > n - 2 ? n - 2 : j);
  for (int s = 0; s < 4 && j < n - 2 && x[j + 1] <= f; ++s) ++j;
  for (int s = 0; s < 4 && j > 0 && x[j] > f; ++s) --j;
  if (x[j] > f || (j < n - 2 && x[j + 1] <= f)) {
    int lo = 0, hi = n - 2;  // x[0] <= f holds here
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (x[mid] <= f) lo = mid;
      else hi = mid - 1;
    }
    j = lo;
  }
  const int i = j + 1;
  return t.S[i - 1] + ((f - x[i - 1]) / (x[i] - x[i - 1])) * (t.S[i] - t.S[i - 1]);
}

// The same interpolation split in two: psd_prefetch issues the loads of the
// segment a log-spaced table predicts (independent of each other, so their
// latency overlaps the waveform work that follows), psd_finish checks the
// prediction and interpolates, falling back to psd_at when it missed.
// Results are identical to psd_at.
struct PsdPre {
  int j;
  double x0, x1, s0, s1;
};

GWF_HD void psd_prefetch(const PsdTab& t, double f, double lnf, PsdPre& p) {
  int j = (int)((lnf - t.lx0) * t.idl);
  j = j < 0 ? 0 : (j > t.n - 2 ? t.n - 2 : j);
  p.j = j;
  p.x0 = t.f[j];
  p.x1 = t.f[j + 1];
  p.s0 = t.S[j];
  p.s1 = t.S[j + 1];
}

GWF_HD double psd_finish(const PsdTab& t, double f, double lnf, const PsdPre& p) {
  if (f < t.f_first || f > t.f_last) return 1.0;
  if (p.x0 <= f && (p.j == t.n - 2 || f < p.x1)) return p.s0 + ((f - p.x0) / (p.x1 - p.x0)) * (p.s1 - p.s0);
  return psd_at(t, f, lnf);
}

GWF_HD double fcut_eff(double fcut, double fmax) { return (fmax > 0.0 && fcut > fmax) ? fmax : fcut; }

// IMRPhenomD and IMRPhenomD_NRTidalv2 share the IMRPhenomD record layout.
GWF_HD constexpr bool is_pd(int model) {
  return model == GWF_MODEL_IMRPHENOMD || model == GWF_MODEL_IMRPHENOMD_NRTIDALV2;
}

// -------------------------------------------------------------- setup body
// The IMRPhenomD models' waveform coefficients, split over NT lanes: lane tk
// evaluates every coefficient as a D<1> carrying tangent tk only (its seed
// the only nonzero one) and stores the value (tk == 0) and that tangent.
// The same arithmetic as the NT-tangent evaluation, in a fifth of the
// registers, so the dependent chain of the collocation solve no longer
// spills.
template <int MODEL>
GWF_HD void setup_waveform_split(const gwf_waveform& w, const double* par, int64_t ld, int64_t e, double f0,
                                 double* rec, int tk) {
  static_assert(is_pd(MODEL), "split setup: IMRPhenomD models");
  const double Mc = par[0 * ld + e], eta = par[1 * ld + e];
  const double chi1 = par[9 * ld + e], chi2 = par[10 * ld + e];
  auto seed = [&](double v, int slot) { return mkseed<1>(v, slot == tk ? 0 : -1); };
  const double f0u = (w.fref > 0.0) ? w.fref : f0;
  const D<1> Mcs = seed(Mc, 0), etas = seed(eta, 1), c1s = seed(chi1, 2), c2s = seed(chi2, 3);
  if (MODEL == GWF_MODEL_IMRPHENOMD) {
    pd_setup<1>(Mcs, etas, c1s, c2s, f0u, w.apply_fcut != 0, w.qnm_a, w.qnm_fring, w.qnm_fdamp, w.n_qnm, rec, SLOT,
                mkc<1>(1.0), mkc<1>(1.0), tk);
  } else {
    D<1> L1, L2;
    lam12_from_lamt(seed(par[11 * ld + e], 4), seed(par[12 * ld + e], 5), etas, L1, L2);
    const D<1> QM1 = quadmon(L1), QM2 = quadmon(L2);
    pd_setup<1>(Mcs, etas, c1s, c2s, f0u, w.apply_fcut != 0, w.qnm_a, w.qnm_fring, w.qnm_fdamp, w.n_qnm, rec, SLOT,
                QM1, QM2, tk);
    nrt_setup<1>(Mcs, etas, c1s, c2s, L1, L2, QM1, QM2, rec, SLOT, tk);
  }
}

template <int MODEL>
GWF_HD void setup_event_consts(const NetK& net, const double* par, int64_t ld, int64_t e, double* rec);

// Fills the event record: waveform coefficients (duals) + projection constants.
template <int MODEL, int NT>
GWF_HD void setup_event(const gwf_waveform& w, const NetK& net, const double* par, int64_t ld, int64_t e, double f0,
                        double* rec) {
  const double Mc = par[0 * ld + e], eta = par[1 * ld + e];
  const double chi1 = par[9 * ld + e], chi2 = par[10 * ld + e];
  if (MODEL == GWF_MODEL_IMRPHENOMD) {
    const double f0u = (w.fref > 0.0) ? w.fref : f0;
    pd_setup<NT>(mkseed<NT>(Mc, 0), mkseed<NT>(eta, 1), mkseed<NT>(chi1, 2), mkseed<NT>(chi2, 3), f0u,
                 w.apply_fcut != 0, w.qnm_a, w.qnm_fring, w.qnm_fdamp, w.n_qnm, rec, SLOT, mkc<NT>(1.0), mkc<NT>(1.0));
  } else if (MODEL == GWF_MODEL_IMRPHENOMD_NRTIDALV2) {
    // Lambda1, Lambda2 from the (LambdaTilde, deltaLambda) rows (signal.py:552-559)
    const double f0u = (w.fref > 0.0) ? w.fref : f0;
    const D<NT> Mcs = mkseed<NT>(Mc, 0), etas = mkseed<NT>(eta, 1), c1s = mkseed<NT>(chi1, 2), c2s = mkseed<NT>(chi2, 3);
    D<NT> L1, L2;
    lam12_from_lamt(mkseed<NT>(par[11 * ld + e], 4), mkseed<NT>(par[12 * ld + e], 5), etas, L1, L2);
    const D<NT> QM1 = quadmon(L1), QM2 = quadmon(L2);
    pd_setup<NT>(Mcs, etas, c1s, c2s, f0u, w.apply_fcut != 0, w.qnm_a, w.qnm_fring, w.qnm_fdamp, w.n_qnm, rec, SLOT,
                 QM1, QM2);
    nrt_setup<NT>(Mcs, etas, c1s, c2s, L1, L2, QM1, QM2, rec, SLOT);
  } else if (tf2_eccentric(NT)) {  // eccentric TaylorF2: ParNums['ecc'] = 11, or 13 after the tidal rows
    if (NT == 7)
      tf2_setup<NT>(w, mkseed<NT>(Mc, 0), mkseed<NT>(eta, 1), mkseed<NT>(chi1, 2), mkseed<NT>(chi2, 3),
                    mkseed<NT>(par[11 * ld + e], 4), mkseed<NT>(par[12 * ld + e], 5), rec, SLOT,
                    mkseed<NT>(par[13 * ld + e], 6), f0);
    else
      tf2_setup<NT>(w, mkseed<NT>(Mc, 0), mkseed<NT>(eta, 1), mkseed<NT>(chi1, 2), mkseed<NT>(chi2, 3), mkc<NT>(0.0),
                    mkc<NT>(0.0), rec, SLOT, mkseed<NT>(par[11 * ld + e], 4), f0);
  } else {
    const double Lt = w.is_tidal ? par[11 * ld + e] : 0.0;
    const double dLm = w.is_tidal ? par[12 * ld + e] : 0.0;
    tf2_setup<NT>(w, mkseed<NT>(Mc, 0), mkseed<NT>(eta, 1), mkseed<NT>(chi1, 2), mkseed<NT>(chi2, 3),
                  mkseed<NT>(Lt, 4), mkseed<NT>(dLm, 5), rec, SLOT);
  }
  setup_event_consts<MODEL>(net, par, ld, e, rec);
}

// tau_star (waveforms.py:878-901, 1299-1322) as K v^-8 P(v), P the polynomial of the
// 0-3.5PN terms in v = (pi M G f)^(1/3): the eta-dependent coefficients and
// K = (5/256) M G / eta once per event (mg: the record's M G dual, whose
// first two tangents are d/dMc and d/deta), so a grid point evaluates P by
// Horner's rule instead of re-forming the eta polynomials.
constexpr double TAU_C3 = -32. / 5. * PI;
constexpr double TAU_C2_D1 = 11. / 3.;
constexpr double TAU_C5_D1 = 13. / 3. * PI;
constexpr double TAU_C6L = 3424. / 105. * 2.0;  // of ln v

GWF_HD void tau_event_coeffs(const double* mg, double eta, double* ev) {
  const double eta2 = eta * eta;
  const double k = 5. / 256.;
  ev[EV_TK] = k * mg[0] / eta;
  ev[EV_TK_D0] = k * mg[1] / eta;
  ev[EV_TK_D1] = k * (mg[2] / eta - mg[0] / eta2);
  ev[EV_TC2] = 743. / 252. + 11. / 3. * eta;
  ev[EV_TC4] = 3058673. / 508032. + 5429. / 504. * eta + 617. / 72. * eta2;
  ev[EV_TC4_D1] = 5429. / 504. + 2. * 617. / 72. * eta;
  ev[EV_TC5] = -(7729. / 252. - 13. / 3. * eta) * PI;
  const double c6e = 3147553127. / 3048192. - 451. / 12. * PI * PI;
  ev[EV_TC6] = -10052469856691. / 23471078400. + 128. / 3. * PI * PI + 6848. / 105. * EULER_GAMMA + c6e * eta -
               15211. / 1728. * eta2 + 25565. / 1296. * eta2 * eta + 3424. / 105. * log(16.);
  ev[EV_TC6_D1] = c6e - 2. * 15211. / 1728. * eta + 3. * 25565. / 1296. * eta2;
  ev[EV_TC7] = (-15419335. / 127008. - 75703. / 756. * eta + 14809. / 378. * eta2) * PI;
  ev[EV_TC7_D1] = (-75703. / 756. + 2. * 14809. / 378. * eta) * PI;
}

// q <- c + u q over (Mc, eta), two FMAs per tangent
GWF_HD void tau_horner(D<2>& q, const D<2>& u, double c, double c_d1) {
  const double qv = q.v;
  q.v = fma(u.v, qv, c);
  q.d[0] = fma(u.d[0], qv, u.v * q.d[0]);
  q.d[1] = fma(u.d[1], qv, fma(u.v, q.d[1], c_d1));
}

GWF_HD D<2> tau_star_ev(const double* ev, const D<2>& v, const D<2>& lv) {
  D<2> q;
  q.v = ev[EV_TC7];
  q.d[0] = 0.0;
  q.d[1] = ev[EV_TC7_D1];
  // c6 + c6L ln v carries tangents through ln v
  const double c6 = fma(TAU_C6L, lv.v, ev[EV_TC6]);
  {
    const double qv = q.v;
    q.v = fma(v.v, qv, c6);
    q.d[0] = fma(v.d[0], qv, fma(v.v, q.d[0], TAU_C6L * lv.d[0]));
    q.d[1] = fma(v.d[1], qv, fma(v.v, q.d[1], fma(TAU_C6L, lv.d[1], ev[EV_TC6_D1])));
  }
  tau_horner(q, v, ev[EV_TC5], TAU_C5_D1);
  tau_horner(q, v, ev[EV_TC4], ev[EV_TC4_D1]);
  tau_horner(q, v, TAU_C3, 0.0);
  tau_horner(q, v, ev[EV_TC2], TAU_C2_D1);
  const D<2> v2 = v * v;
  D<2> p = v2 * q;
  p.v += 1.0;
  const D<2> v4 = v2 * v2;
  D<2> K;
  K.v = ev[EV_TK];
  K.d[0] = ev[EV_TK_D0];
  K.d[1] = ev[EV_TK_D1];
  return K * p / (v4 * v4);
}

// The same, value only (the SNR path)
GWF_HD double tau_star_ev_val(const double* ev, double v, double lv) {
  double q = ev[EV_TC7];
  q = fma(v, q, fma(TAU_C6L, lv, ev[EV_TC6]));
  q = fma(v, q, ev[EV_TC5]);
  q = fma(v, q, ev[EV_TC4]);
  q = fma(v, q, TAU_C3);
  q = fma(v, q, ev[EV_TC2]);
  const double v2 = v * v, v4 = v2 * v2;
  return ev[EV_TK] * fma(v2, q, 1.0) / (v4 * v4);
}

// The event constants of the record (after its waveform coefficients, whose
// fcut it reads): inclination factors, 1/dL, tcoal, sky angles, ln(M G) and
// its (Mc, eta) tangents, and the projection constants of every site.
template <int MODEL>
GWF_HD void setup_event_consts(const NetK& net, const double* par, int64_t ld, int64_t e, double* rec) {
  const double Mc = par[0 * ld + e], eta = par[1 * ld + e], dL = par[2 * ld + e];
  const double theta = par[3 * ld + e], phi = par[4 * ld + e], iota = par[5 * ld + e];
  const double psi = par[6 * ld + e], tc = par[7 * ld + e];
  double* ev = rec + REC_EV;
  const double ci = cos(iota), si = sin(iota);
  ev[EV_DL] = dL;
  ev[EV_IDL] = 1.0 / dL;
  ev[EV_CP] = 0.5 * (1. + ci * ci);
  ev[EV_CC] = ci;
  ev[EV_DCP] = -(ci * si);
  ev[EV_DCC] = -si;
  ev[EV_TC] = tc;
  ev[EV_PHIC] = par[8 * ld + e];
  ev[EV_RA] = phi;
  ev[EV_CRA] = cos(phi);
  ev[EV_SRA] = sin(phi);
  ev[EV_FCUT] = rec[(is_pd(MODEL) ? (int)PD_FCUT : (int)TF_FCUT) * SLOT];
  {  // ln(M G f) = ln(M G) + ln f on the grids, whose ln f comes with f;
     // ln(M G) = ln Mc - (3/5) ln eta + ln G: tangents 1/Mc and -(3/5)/eta
    const double* mg = rec + (is_pd(MODEL) ? (int)PD_MG : (int)TF_MG) * SLOT;
    ev[EV_LMG] = log(mg[0]);
    ev[EV_LMG_D0] = 1.0 / Mc;
    ev[EV_LMG_D1] = -0.6 / eta;
    tau_event_coeffs(mg, eta, ev);
  }
  // sky position and polarisation (ra = phi, dec = pi/2 - theta; gwfastUtils.py:151-164)
  const double ra = phi, dec = 0.5 * PI - theta;
  const double cd = cos(dec), sd = sin(dec), cr = cos(ra), sr = sin(ra);
  const double c2d = cos(2. * dec), s2d = sin(2. * dec);
  const double c2p = cos(2. * psi), s2p = sin(2 * psi);
  const double k = -REARTH / CLIGHT;
  for (int s = 0; s < net.n_sites; ++s) {
    const SiteK& S = net.site[s];
    double* sc = rec + REC_SITE + s * SC_COUNT;
    // Delta t = -R/c (cd cr clat cos L + cd sr clat sin L + sd slat), L = lon + 2 pi t
    sc[SC_U1] = k * (cd * cr * S.clat);
    sc[SC_U2] = k * (cd * sr * S.clat);
    sc[SC_U3] = k * (sd * S.slat);
    sc[SC_U1T] = k * (sd * cr * S.clat);  // explicit d/dtheta (d dec/d theta = -1)
    sc[SC_U2T] = k * (sd * sr * S.clat);
    sc[SC_U3T] = k * (-cd * S.slat);
    // a = s2x (E1 cos2al + E3 cos al + E5) - c2x (E2 sin2al + E4 sin al)
    // b = c2x (G1 cos2al + G3 cos al) + s2x (G2 sin2al + G4 sin al)
    sc[SC_E1] = 0.0625 * (3. - S.c2lat) * (3. - c2d);
    sc[SC_E2] = 0.25 * S.slat * (3. - c2d);
    sc[SC_E3] = 0.25 * S.s2lat * s2d;
    sc[SC_E4] = 0.5 * S.clat * s2d;
    sc[SC_E5] = 3. * 0.25 * (S.clat * cd) * (S.clat * cd);
    sc[SC_G1] = S.slat * sd;
    sc[SC_G2] = 0.25 * (3. - S.c2lat) * sd;
    sc[SC_G3] = S.clat * cd;
    sc[SC_G4] = 0.5 * S.s2lat * cd;
    sc[SC_E1T] = 0.0625 * (3. - S.c2lat) * (-2. * s2d);
    sc[SC_E2T] = 0.25 * S.slat * (-2. * s2d);
    sc[SC_E3T] = 0.25 * S.s2lat * (-2. * c2d);
    sc[SC_E4T] = 0.5 * S.clat * (-2. * c2d);
    sc[SC_E5T] = 2. * 3. * 0.25 * (S.clat * S.clat) * cd * sd;
    sc[SC_G1T] = -S.slat * cd;
    sc[SC_G2T] = -0.25 * (3. - S.c2lat) * cd;
    sc[SC_G3T] = S.clat * sd;
    sc[SC_G4T] = 0.5 * S.s2lat * sd;
    sc[SC_P1] = S.sinzeta * c2p;
    sc[SC_P2] = S.sinzeta * s2p;
    const double t0 = (S.motion == MOTION_NONE) ? 0.0 : tc;
    sc[SC_CL0] = cos(S.lon + 2. * PI * t0);
    sc[SC_SL0] = sin(S.lon + 2. * PI * t0);
  }
}

// ------------------------------------------------------- per-point geometry
struct Geo {
  double dt, dt_t, dt_th, dt_ph;       // Delta t [s] and partials (per day of t, per rad)
  double tn;                           // time without the location delay [GMST days]
  double Sa, Ca, Cb, Sb;               // pattern blocks at this time
  double Sa_al, Ca_al, Cb_al, Sb_al;   // d/d alpha
  double Sa_th, Ca_th, Cb_th, Sb_th;   // explicit d/d theta
};

// sin and cos of x for |x| < 2^30, within 2 ulp away from the zeros: x = k pi/2 + r
// with pi/2 in three parts (Cody and Waite's reduction, with FMAs) and
// |r| <= pi/4, the Taylor series of sin r to r^17 and of cos r to r^18
// (truncation below 1e-19), then the quadrant k mod 4.
GWF_HD void sincos_kernel(double x, double* s, double* c) {
  constexpr double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17, P3 = -1.4973849048591698e-33;
  const double k = rint(x * 0.6366197723675814);
  double r = fma(-k, P1, x);
  r = fma(-k, P2, r);
  r = fma(-k, P3, r);
  const double r2 = r * r;
  double ps = -1.0 / 355687428096000.0;  // -1/17!
  ps = fma(ps, r2, 1.0 / 1307674368000.0);
  ps = fma(ps, r2, -1.0 / 6227020800.0);
  ps = fma(ps, r2, 1.0 / 39916800.0);
  ps = fma(ps, r2, -1.0 / 362880.0);
  ps = fma(ps, r2, 1.0 / 5040.0);
  ps = fma(ps, r2, -1.0 / 120.0);
  ps = fma(ps, r2, 1.0 / 6.0);
  const double sr = fma(-r * r2, ps, r);  // r - r^3 (1/6 - r^2/120 + ...)
  double pc = 1.0 / 6402373705728000.0;  // 1/18!
  pc = fma(pc, r2, -1.0 / 20922789888000.0);
  pc = fma(pc, r2, 1.0 / 87178291200.0);
  pc = fma(pc, r2, -1.0 / 479001600.0);
  pc = fma(pc, r2, 1.0 / 3628800.0);
  pc = fma(pc, r2, -1.0 / 40320.0);
  pc = fma(pc, r2, 1.0 / 720.0);
  pc = fma(pc, r2, -1.0 / 24.0);
  pc = fma(pc, r2, 0.5);
  const double cr = fma(-r2, pc, 1.0);  // 1 - r^2 (1/2 - r^2/24 + ...)
  const int q = (int)(long long)k & 3;
  const double a = (q & 1) ? cr : sr, b = (q & 1) ? sr : cr;
  *s = (q & 2) ? -a : a;
  *c = ((q + 1) & 2) ? -b : b;
}

GWF_HD void sincos_d(double x, double* s, double* c) {
#if defined(__HIP_DEVICE_COMPILE__)
  sincos(x, s, c);
#else
  *s = sin(x);
  *c = cos(x);
#endif
}

// Site geometry at one frequency.  The Earth-rotation angle
// L = lon + 2 pi t_noloc is the per-event angle L0 = lon + 2 pi t0 rotated by
// -delta, delta = 2 pi tau / 86400 (the same for every site, so one sincos
// per frequency serves all sites: cd, sd = cos, sin delta); the pattern
// angle alpha = ra - L - eps, eps = 2 pi Delta t / 86400 (|eps| < 1.6e-6), is
// (ra - L) rotated by -eps with cos eps = 1 - eps^2/2, sin eps = eps - eps^3/6
// (truncation < 1e-24).  Angle-addition rounding is at the 1e-16 level.
template <bool WITH_DERIV>
GWF_HD void site_geo(const SiteK& S, const double* sc, const double* ev, double tn, double cdel, double sdel, Geo& g) {
  g.tn = tn;
  double cL = sc[SC_CL0], sL = sc[SC_SL0];
  if (S.motion == MOTION_EARTH) {
    const double c0 = cL, s0 = sL;
    cL = c0 * cdel + s0 * sdel;
    sL = s0 * cdel - c0 * sdel;
  }
  g.dt = sc[SC_U1] * cL + sc[SC_U2] * sL + sc[SC_U3];
  // beta = ra - L
  const double cra = ev[EV_CRA], sra = ev[EV_SRA];
  const double cb = cra * cL + sra * sL;
  const double sb = sra * cL - cra * sL;
  const double eps = (2. * PI * INV_DAY) * g.dt;
  const double ce = 1.0 - 0.5 * eps * eps;
  const double se = eps - eps * eps * eps * (1.0 / 6.0);
  const double ca = cb * ce + sb * se;
  const double sa = sb * ce - cb * se;
  const double c2a = ca * ca - sa * sa;
  const double s2a = 2. * sa * ca;
  g.Sa = sc[SC_E1] * c2a + sc[SC_E3] * ca + sc[SC_E5];
  g.Ca = sc[SC_E2] * s2a + sc[SC_E4] * sa;
  g.Cb = sc[SC_G1] * c2a + sc[SC_G3] * ca;
  g.Sb = sc[SC_G2] * s2a + sc[SC_G4] * sa;
  if (WITH_DERIV) {
    g.dt_t = 2. * PI * (sc[SC_U2] * cL - sc[SC_U1] * sL);
    g.dt_th = sc[SC_U1T] * cL + sc[SC_U2T] * sL + sc[SC_U3T];
    g.dt_ph = sc[SC_U1] * sL - sc[SC_U2] * cL;
    g.Sa_al = -2. * sc[SC_E1] * s2a - sc[SC_E3] * sa;
    g.Ca_al = 2. * sc[SC_E2] * c2a + sc[SC_E4] * ca;
    g.Cb_al = -2. * sc[SC_G1] * s2a - sc[SC_G3] * sa;
    g.Sb_al = 2. * sc[SC_G2] * c2a + sc[SC_G4] * ca;
    g.Sa_th = sc[SC_E1T] * c2a + sc[SC_E3T] * ca + sc[SC_E5T];
    g.Ca_th = sc[SC_E2T] * s2a + sc[SC_E4T] * sa;
    g.Cb_th = sc[SC_G1T] * c2a + sc[SC_G3T] * ca;
    g.Sb_th = sc[SC_G2T] * s2a + sc[SC_G4T] * sa;
  }
}

// sincos of the site-independent Earth-rotation increment 2 pi tau / 86400.
GWF_HD void earth_delta(bool any_earth, double tau, double* c, double* s) {
  if (any_earth) {
    sincos_kernel((2. * PI * INV_DAY) * tau, s, c);
  } else {
    *c = 1.0;
    *s = 0.0;
  }
}

// t without location delay for a site (signal.py:444-453).
GWF_HD double site_tn(int motion, double tc, double tau) {
  if (motion == MOTION_EARTH) return tc - tau * INV_DAY;
  if (motion == MOTION_FIXED) return tc;
  return 0.0;
}

// ------------------------------------------------------- waveform at a point
template <int MODEL, int NT>
struct WfPoint {
  D<NT> A;    // amplitude without 1/dL
  D<NT> Phi;  // GW phase
  D<2> tau;   // tau_star over (Mc, eta)
};

// Waveform quantities at one frequency.  The powers and logarithm of the
// scaled frequency are computed once and shared by phase, amplitude and tau:
// IMRPhenomD: x = M G f, tau's v = (pi x)^(1/3) = pi^(1/3) x^(1/3);
// TaylorF2:   v = (pi M G f)^(1/3).
// M G (and so x, v and their powers and logarithms) depends on Mc and eta
// only: it is carried as a D<2> over the first two tangents, and the mixed
// dual arithmetic leaves the other tangents' zero products out.
// HAVE_LNF: the caller (a grid walk) passes lnf = ln f.
template <int MODEL, int NT, bool HAVE_LNF = false>
GWF_HD void wf_point(const gwf_waveform& w, const double* rec, double f, bool need_tau, WfPoint<MODEL, NT>& o,
                     double lnf = 0.0) {
  const int smg = is_pd(MODEL) ? (int)PD_MG : (int)TF_MG;
  const D<2> MG = ldd<2>(rec + smg * SLOT);
  // ln(M G f) from the event's ln(M G) and ln f (its tangents are those of ln M G)
  constexpr bool have_lnf = HAVE_LNF;
  D<2> lmgf;
  if (have_lnf) {
    const double* ev = rec + REC_EV;
    lmgf.v = ev[EV_LMG] + lnf;
    lmgf.d[0] = ev[EV_LMG_D0];
    lmgf.d[1] = ev[EV_LMG_D1];
  }
  D<2> v, lv;
  if (is_pd(MODEL)) {
    constexpr bool NRT = MODEL == GWF_MODEL_IMRPHENOMD_NRTIDALV2;
    const D<2> x = MG * f;
    const D<2> x13 = dcbrt(x);
    const D<2> lx = have_lnf ? lmgf : dlog(x);
    o.A = pd_amp_at<NT, NRT>(rec, SLOT, x, x13, w.apply_fcut != 0 || NRT);
    o.Phi = pd_phi_at<NT, NRT>(rec, SLOT, x, x13, lx, w.apply_fcut != 0);
    if (need_tau) {
      v = x13 * cbrt(PI);
      lv = (lx + log(PI)) * (1.0 / 3.0);
    }
  } else {
    v = dcbrt(PI * MG * f);
    lv = have_lnf ? (lmgf + log(PI)) * (1.0 / 3.0) : dlog(v);
    o.Phi = tf2_phi_at<NT>(rec, SLOT, v, lv);
    const double f13 = cbrt(f);
    o.A = ldd<NT>(rec + TF_AMP * SLOT) * (1.0 / (f * sqrt(f13)));
  }
  if (need_tau) {
    o.tau = tau_star_ev(rec + REC_EV, v, lv);
  } else {
    o.tau = mkc<2>(0.0);
  }
}

// Value-only amplitude without 1/dL and, when need_tau, tau_star (SNR path;
// lnf = ln f from the grid walk, so ln v = (ln(M G) + ln f + ln pi) / 3).
// IMRPhenomD shares x^(1/3) = (M G f)^(1/3) between them: tau's
// v = (pi M G f)^(1/3) = pi^(1/3) x^(1/3), as in wf_point.
template <int MODEL>
GWF_HD double wf_amp_tau_val(const gwf_waveform& w, const double* rec, double f, double lnf, bool need_tau,
                             double* tau) {
  const double* ev = rec + REC_EV;
  if (is_pd(MODEL)) {
    constexpr bool NRT = MODEL == GWF_MODEL_IMRPHENOMD_NRTIDALV2;
    const double MG = rec[PD_MG * SLOT];
    const double x = MG * f;
    const double x13 = cbrt(x);
    if (need_tau) {
      const double v = x13 * cbrt(PI);
      *tau = tau_star_ev_val(ev, v, (ev[EV_LMG] + lnf + log(PI)) * (1.0 / 3.0));
    } else {
      *tau = 0.0;
    }
    return pd_amp_val(rec, SLOT, x, x13, w.apply_fcut != 0 || NRT, NRT);
  }
  if (need_tau) {
    const double v = cbrt(PI * rec[TF_MG * SLOT] * f);
    *tau = tau_star_ev_val(ev, v, (ev[EV_LMG] + lnf + log(PI)) * (1.0 / 3.0));
  } else {
    *tau = 0.0;
  }
  const double f13 = cbrt(f);
  return rec[TF_AMP * SLOT] * (1.0 / (f * sqrt(f13)));
}

// ------------------------------------------------------ strain derivatives
// Reduced derivative vectors of one arm at one point:
//   re[X] + i im[X] = dX h * exp(-i Psi), tcoal row already divided by 86400
// (signal.py:919-920); X in ParNums order.  With h = A (F+ c+ + i Fx cx)
// exp(i Psi) / dL (c+ = (1 + cos^2 iota)/2, cx = cos iota), Ap = A F+ c+ / dL,
// Ac = A Fx cx / dL, every row is (dX Ap - Ac dX Psi) + i (dX Ac + Ap dX Psi);
// the products A c+ / dL, A cx / dL, F+ c+, Fx cx and the pattern-angle
// terms A F'_alpha c / dL are formed once and shared by the rows.  Every
// row is linear in the amplitude and its tangents, so a scale q of all rows
// (and of Ap, Ac) costs one multiply, folded into 1/dL: the Fisher sums use
// q = sqrt(Gram weight) and then need no per-row weighting.
template <int MODEL, int NT, int P>
GWF_HD void arm_derivs(const WfPoint<MODEL, NT>& wp, const double* ev, const double* sc, const Geo& g, const ArmK& arm,
                       int motion, double f, double re[P], double im[P], double* Ap_out, double* Ac_out,
                       double q = 1.0, double* v_out = nullptr) {
  const double idL0 = ev[EV_IDL], idL = idL0 * q;
  const double cp = ev[EV_CP], cc = ev[EV_CC];
  const double P1 = sc[SC_P1], P2 = sc[SC_P2];
  const double s2x = arm.s2x, c2x = arm.c2x;
  const double a = s2x * g.Sa - c2x * g.Ca;
  const double b = c2x * g.Cb + s2x * g.Sb;
  const double a_al = s2x * g.Sa_al - c2x * g.Ca_al;
  const double b_al = c2x * g.Cb_al + s2x * g.Sb_al;
  const double a_th = s2x * g.Sa_th - c2x * g.Ca_th;
  const double b_th = c2x * g.Cb_th + s2x * g.Sb_th;
  const double Fp = a * P1 + b * P2, Fc = b * P1 - a * P2;
  const double Fp_al = a_al * P1 + b_al * P2, Fc_al = b_al * P1 - a_al * P2;
  const double Fp_th = a_th * P1 + b_th * P2, Fc_th = b_th * P1 - a_th * P2;

  // d t_noloc / d(Mc, eta, tc) [days per unit]
  double tn_mc = 0., tn_eta = 0., tn_tc = 0.;
  if (motion == MOTION_EARTH) {
    tn_mc = -wp.tau.d[0] * INV_DAY;
    tn_eta = -wp.tau.d[1] * INV_DAY;
    tn_tc = 1.0;
  } else if (motion == MOTION_FIXED) {
    tn_tc = 1.0;
  }
  const double tfac = 1.0 + g.dt_t * INV_DAY;  // d t / d t_noloc
  const double tw = 2. * PI;
  // d alpha / dX with alpha = ra - lon - 2 pi t
  const double al_mc = -tw * tn_mc * tfac, al_eta = -tw * tn_eta * tfac, al_tc = -tw * tn_tc * tfac;
  const double al_th = -tw * g.dt_th * INV_DAY, al_ph = 1.0 - tw * g.dt_ph * INV_DAY;

  const double Av = wp.A.v * idL;
  const double Avp = Av * cp, Avc = Av * cc;      // A c+ / dL, A cx / dL
  const double Fpp = Fp * cp, Fcc = Fc * cc;      // F+ c+, Fx cx
  const double Ap = Fp * Avp, Ac = Fc * Avc;
  const double Xp = Fp_al * Avp, Xc = Fc_al * Avc;  // d(Ap, Ac) / d alpha
  *Ap_out = Ap;
  *Ac_out = Ac;
  const double twopif = 2. * PI * f;
  const double tdt = twopif * g.dt_t;             // 2 pi f d(Delta t)/dt
  if (v_out) {
    // v = -2 pi tfac dU/dalpha + i tdt U (U = Ap + i Ac): the Mc, eta and
    // tcoal rows are U c_j + v tn_j (tn_j: their d t_noloc / dX, zero for
    // every row when the site does not move)
    const double k = -tw * tfac;
    v_out[0] = motion == MOTION_NONE ? 0.0 : fma(k, Xp, -tdt * Ac);
    v_out[1] = motion == MOTION_NONE ? 0.0 : fma(k, Xc, tdt * Ap);
  }

#define GWF_SET(X, dAp, dAc, dPsi)        \
  {                                       \
    const double dps_ = (dPsi);           \
    re[X] = (dAp) - Ac * dps_;            \
    im[X] = (dAc) + Ap * dps_;            \
  }
  // Mc, eta: waveform amplitude/phase + Earth-rotation time dependence
  {
    const double dA = wp.A.d[0] * idL;
    GWF_SET(0, dA * Fpp + Xp * al_mc, dA * Fcc + Xc * al_mc, tdt * tn_mc - wp.Phi.d[0]);
  }
  {
    const double dA = wp.A.d[1] * idL;
    GWF_SET(1, dA * Fpp + Xp * al_eta, dA * Fcc + Xc * al_eta, tdt * tn_eta - wp.Phi.d[1]);
  }
  // dL
  re[2] = -Ap * idL0;
  im[2] = -Ac * idL0;
  // theta, phi: pattern functions (direct + via the location delay) and phiL
  GWF_SET(3, Avp * (Fp_th + Fp_al * al_th), Avc * (Fc_th + Fc_al * al_th), twopif * g.dt_th);
  GWF_SET(4, Xp * al_ph, Xc * al_ph, twopif * g.dt_ph);
  // iota
  re[5] = Av * (Fp * ev[EV_DCP]);
  im[5] = Av * (Fc * ev[EV_DCC]);
  // psi: dFp/dpsi = 2 Fc, dFc/dpsi = -2 Fp
  re[6] = (Fc + Fc) * Avp;
  im[6] = -(Fp + Fp) * Avc;
  // tcoal, per second
  {
    const double al = al_tc * INV_DAY;
    GWF_SET(7, Xp * al, Xc * al, twopif + tdt * (tn_tc * INV_DAY));
  }
  // Phicoal: d Psi = -1
  re[8] = Ac;
  im[8] = -Ap;
  // spins (and tidal parameters): waveform amplitude and phase only
  {
    const double dA = wp.A.d[2] * idL;
    GWF_SET(9, dA * Fpp, dA * Fcc, -wp.Phi.d[2]);
  }
  {
    const double dA = wp.A.d[3] * idL;
    GWF_SET(10, dA * Fpp, dA * Fcc, -wp.Phi.d[3]);
  }
  if (P > 11) {
    // tidal parameters (phase: TaylorF2, NRTidalv2; amplitude: NRTidalv2
    // tidal amplitude and Planck taper, zero tangent for TaylorF2), or the
    // eccentricity of eccentric TaylorF2 (phase only)
    const double dA4 = wp.A.d[NT > 4 ? 4 : 0] * idL;
    GWF_SET(11, dA4 * Fpp, dA4 * Fcc, -wp.Phi.d[NT > 4 ? 4 : 0]);
  }
  if (P > 12) {
    const double dA5 = wp.A.d[NT > 5 ? 5 : 0] * idL;
    GWF_SET(12, dA5 * Fpp, dA5 * Fcc, -wp.Phi.d[NT > 5 ? 5 : 0]);
  }
  if (P > 13) {  // eccentricity after the tidal rows
    const double dA6 = wp.A.d[NT > 6 ? 6 : 0] * idL;
    GWF_SET(13, dA6 * Fpp, dA6 * Fcc, -wp.Phi.d[NT > 6 ? 6 : 0]);
  }
#undef GWF_SET
}

// Symmetric Gram update G += w (re re^T + im im^T), packed upper triangle.
// Two explicit FMAs per entry: written as G += (wr re + wi im) the compiler
// may only contract the inner sum (a multiply, an FMA and an add).
template <int P>
GWF_HD void gram_update(double* G, const double* re, const double* im, double w) {
  int k = 0;
#pragma unroll
  for (int a = 0; a < P; ++a) {
    const double wr = w * re[a], wi = w * im[a];
#pragma unroll
    for (int b = a; b < P; ++b) {
      G[k] = fma(wi, im[b], fma(wr, re[b], G[k]));
      ++k;
    }
  }
}

// G += re re^T + im im^T for rows already scaled by sqrt(weight).
template <int P>
GWF_HD void gram_update_unit(double* G, const double* re, const double* im) {
  int k = 0;
#pragma unroll
  for (int a = 0; a < P; ++a) {
#pragma unroll
    for (int b = a; b < P; ++b) {
      G[k] = fma(im[a], im[b], fma(re[a], re[b], G[k]));
      ++k;
    }
  }
}

// ------------------------------------------------------------ Fisher body
// Accumulates the packed Fisher of event e over points lane, lane+nlanes, ...
// arm_sel < 0: the network sum (over net.sum_arm); else only that physical arm.
// MASKED: the network sum over the physical arms whose bit is set in mask
// (bit a = physical arm a); a site with all its arms set uses its sum arms.
// LIN: linearly spaced grid (FisherMatr spacing='lin').
template <int MODEL, int NT, int P, bool MASKED = false, bool LIN = false>
GWF_HD void fisher_event(const gwf_waveform& w, const NetK& net, const double* rec, int res, int lane, int nlanes,
                         int arm_sel, double* G, uint32_t mask = 0u) {
  const double* ev = rec + REC_EV;
  double a, b;
  grid_ends(LIN, net.fmin, fcut_eff(ev[EV_FCUT], net.fmax), &a, &b);
  const bool sum_mode = arm_sel < 0;
  const int narm = sum_mode ? net.n_sum_arms : 1;
  GridWalk<LIN> gw;
  gw.init(a, b, res, lane, nlanes);
  for (int base = 0; base < res; base += nlanes) {
#if defined(__HIP_DEVICE_COMPILE__) && !defined(GWF_NO_LOAD_BARRIER)
    // keep the per-event record in memory (scalar loads) instead of letting
    // loop-invariant loads occupy vector registers for the whole loop
    __asm__ volatile("" ::: "memory");
#endif
    const int i = base + lane;
    double f, wt, lnf;
    gw.step(base, res, &f, &wt, &lnf);
    if (i >= res) continue;
    wt *= 4.0;
#ifndef GWF_NO_PSD_PREFETCH
    // loads of the first two PSD tables' segments, consumed after the waveform
    PsdPre pre0, pre1;
    psd_prefetch(net.psd[0], f, lnf, pre0);
    if (net.n_psd > 1) psd_prefetch(net.psd[1], f, lnf, pre1);
#endif
    WfPoint<MODEL, NT> wp;
    wf_point<MODEL, NT, true>(w, rec, f, net.any_earth != 0, wp, lnf);
    const double tc = ev[EV_TC];
    double cdel, sdel;
    earth_delta(net.any_earth != 0, wp.tau.v, &cdel, &sdel);
    int cur_site = -1, cur_psd = -1;
    double Sf = 1.0;
    Geo g;
    if constexpr (MASKED) {
      (void)narm;
      for (int s = 0; s < net.n_sites; ++s) {
        const SiteK& S = net.site[s];
        const uint32_t all = (1u << S.narm) - 1u;
        const uint32_t bits = (mask >> S.arm0) & all;
        if (!bits) continue;
        const double* sc = rec + REC_SITE + s * SC_COUNT;
        if (S.psd != cur_psd) {
#ifndef GWF_NO_PSD_PREFETCH
          if (S.psd == 0)
            Sf = psd_finish(net.psd[0], f, lnf, pre0);
          else if (S.psd == 1)
            Sf = psd_finish(net.psd[1], f, lnf, pre1);
          else
#endif
            Sf = psd_at(net.psd[S.psd], f, lnf);
          cur_psd = S.psd;
        }
        site_geo<true>(S, sc, ev, site_tn(S.motion, tc, wp.tau.v), cdel, sdel, g);
        const bool full = bits == all;
        const int cnt = full ? S.nsum : S.narm;
        for (int j = 0; j < cnt; ++j) {
          if (!full && !((bits >> j) & 1u)) continue;
          const ArmK& arm = full ? net.sum_arm[S.sum0 + j] : net.arm[S.arm0 + j];
          double re[P], im[P], Ap, Ac;
          arm_derivs<MODEL, NT, P>(wp, ev, sc, g, arm, S.motion, f, re, im, &Ap, &Ac);
          gram_update<P>(G, re, im, wt * arm.wgt / Sf);
        }
      }
    } else {
      for (int k = 0; k < narm; ++k) {
        const ArmK& arm = sum_mode ? net.sum_arm[k] : net.arm[arm_sel];
        const int s = arm.site;
        const SiteK& S = net.site[s];
        const double* sc = rec + REC_SITE + s * SC_COUNT;
        if (s != cur_site) {
          if (S.psd != cur_psd) {
#ifndef GWF_NO_PSD_PREFETCH
            if (S.psd == 0)
              Sf = psd_finish(net.psd[0], f, lnf, pre0);
            else if (S.psd == 1)
              Sf = psd_finish(net.psd[1], f, lnf, pre1);
            else
#endif
              Sf = psd_at(net.psd[S.psd], f, lnf);
            cur_psd = S.psd;
          }
          site_geo<true>(S, sc, ev, site_tn(S.motion, tc, wp.tau.v), cdel, sdel, g);
          cur_site = s;
        }
        double re[P], im[P], Ap, Ac;
        arm_derivs<MODEL, NT, P>(wp, ev, sc, g, arm, S.motion, f, re, im, &Ap, &Ac);
        gram_update<P>(G, re, im, wt * arm.wgt / Sf);
      }
    }
  }
}

// ------------------------------------- Fisher body, producer / consumer split
// The per-frequency work of fisher_event in two halves that run in separate
// waves of one workgroup, exchanging 64-point chunks through LDS: the
// producer walks the grid and evaluates the PSDs, the waveform with its
// tangents, tau(f) and the Earth-rotation phase of tau; the consumer does the
// per-site geometry, the arm derivative rows and the Gram updates.  Neither
// wave holds the other's registers (the consumer's Gram accumulators are not
// live in the producer), so two waves fit on a SIMD.
// Per point: f, 4 x trapezoid weight, ln f, sqrt(4 wt / S) over the first
// two PSD tables (SW0, SW1: the consumer scales each arm's rows by it and by
// the arm's sqrt weight, instead of weighting the Gram), amplitude,
// tangents, tau and the Earth-rotation increment.
template <int NT>
struct PcLayout {
  static constexpr int F = 0, WT = 1, LNF = 2, SW0 = 3, SW1 = 4, AV = 5, AD = 6, PD = 6 + NT, TV = 6 + 2 * NT,
                       TD = 7 + 2 * NT, CD = 9 + 2 * NT, SD = 10 + 2 * NT, NV = 11 + 2 * NT;
};

// Amplitudes at or below this (a Planck taper's last points, a cut) count as
// zero in the pure-row factors c_p = A'/A - i Phi': 1/A overflows below
// 5.6e-309, while such a point's rows, all of the order of A and of its
// tangents (below ~1e-270 there, against ~1e-20 in band), add nothing to
// the Fisher matrix at double precision.
constexpr double AMP_TINY = 1e-280;

// Producer: point base + lane of the grid into buf[NV][64] (nothing past res).
// factored: rows AD hold A'/A (0 where A vanishes) and TD the factored
// rows' t_noloc tangents -tau'/day (0 without a rotating Earth), as
// pc_consume_semi reads them, instead of A' and tau'.
template <int MODEL, int NT, bool LIN>
GWF_HD void pc_produce(const gwf_waveform& w, const NetK& net, const double* rec, GridWalk<LIN>& gw, int base, int res,
                       int lane, double* buf, bool factored = false) {
  typedef PcLayout<NT> L;
#if defined(__HIP_DEVICE_COMPILE__) && !defined(GWF_NO_LOAD_BARRIER)
  // keep the record loads inside the chunk (scalar loads) rather than hoisted
  // into vector registers across the whole grid walk
  __asm__ volatile("" ::: "memory");
#endif
  double f, wt, lnf;
  gw.step(base, res, &f, &wt, &lnf);  // every lane: neighbours come by shuffles
  if (base + lane >= res) return;
  PsdPre pre0, pre1;
  psd_prefetch(net.psd[0], f, lnf, pre0);
  if (net.n_psd > 1) psd_prefetch(net.psd[1], f, lnf, pre1);
  WfPoint<MODEL, NT> wp;
  wf_point<MODEL, NT, true>(w, rec, f, net.any_earth != 0, wp, lnf);
  double cdel, sdel;
  earth_delta(net.any_earth != 0, wp.tau.v, &cdel, &sdel);
  const double s0 = psd_finish(net.psd[0], f, lnf, pre0);
  const double s1 = net.n_psd > 1 ? psd_finish(net.psd[1], f, lnf, pre1) : 1.0;
  buf[L::F * 64 + lane] = f;
  buf[L::WT * 64 + lane] = 4.0 * wt;
  buf[L::LNF * 64 + lane] = lnf;
  buf[L::SW0 * 64 + lane] = sqrt((4.0 * wt) / s0);
  buf[L::SW1 * 64 + lane] = sqrt((4.0 * wt) / s1);
  buf[L::AV * 64 + lane] = wp.A.v;
  const double ia = factored ? (fabs(wp.A.v) > AMP_TINY ? 1.0 / wp.A.v : 0.0) : 1.0;
  const double td = factored ? (net.any_earth ? -INV_DAY : 0.0) : 1.0;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    buf[(L::AD + k) * 64 + lane] = wp.A.d[k] * ia;
    buf[(L::PD + k) * 64 + lane] = wp.Phi.d[k];
  }
  buf[L::TV * 64 + lane] = wp.tau.v;
  buf[L::TD * 64 + lane] = wp.tau.d[0] * td;
  buf[(L::TD + 1) * 64 + lane] = wp.tau.d[1] * td;
  buf[L::CD * 64 + lane] = cdel;
  buf[L::SD * 64 + lane] = sdel;
}

// Consumer: the Gram updates of point base + lane (network sum arms; MASKED:
// the arms whose bit is set in mask, a site with all arms set through its
// sum arms, as fisher_event).
template <int MODEL, int NT, int P, bool MASKED = false>
GWF_HD void pc_consume(const NetK& net, const double* rec, int base, int res, int lane, const double* buf, double* G,
                       uint32_t mask = 0u) {
  typedef PcLayout<NT> L;
#if defined(__HIP_DEVICE_COMPILE__) && !defined(GWF_NO_LOAD_BARRIER)
  __asm__ volatile("" ::: "memory");
#endif
  if (base + lane >= res) return;
  const double* ev = rec + REC_EV;
  const double f = buf[L::F * 64 + lane];
  WfPoint<MODEL, NT> wp;
  wp.A.v = buf[L::AV * 64 + lane];
  wp.Phi.v = 0.0;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    wp.A.d[k] = buf[(L::AD + k) * 64 + lane];
    wp.Phi.d[k] = buf[(L::PD + k) * 64 + lane];
  }
  // values needed once per site or per arm are read from the chunk where
  // they are used, which leaves the compiler free to reload them instead of
  // spilling (the Gram accumulators and an arm's rows fill the registers)
  auto ld = [&](int slot) { return buf[slot * 64 + lane]; };
  const double tc = ev[EV_TC];
  int cur_site = -1, cur_psd = -1;
  double sw = 0.0;  // sqrt(4 wt / S) of the current PSD
  Geo g;
  // tables past the first two are interpolated here, once per site
  auto psd_sw = [&](int psd) {
    return psd == 0 ? buf[L::SW0 * 64 + lane]
                    : (psd == 1 ? buf[L::SW1 * 64 + lane]
                                : sqrt(ld(L::WT) / psd_at(net.psd[psd], f, ld(L::LNF))));
  };
  if constexpr (MASKED) {
    for (int s = 0; s < net.n_sites; ++s) {
      const SiteK& S = net.site[s];
      const uint32_t all = (1u << S.narm) - 1u;
      const uint32_t bits = (mask >> S.arm0) & all;
      if (!bits) continue;
      const double* sc = rec + REC_SITE + s * SC_COUNT;
      if (S.psd != cur_psd) {
        sw = psd_sw(S.psd);
        cur_psd = S.psd;
      }
      site_geo<true>(S, sc, ev, site_tn(S.motion, tc, ld(L::TV)), ld(L::CD), ld(L::SD), g);
      const bool full = bits == all;
      const int cnt = full ? S.nsum : S.narm;
      for (int j = 0; j < cnt; ++j) {
        if (!full && !((bits >> j) & 1u)) continue;
        const ArmK& arm = full ? net.sum_arm[S.sum0 + j] : net.arm[S.arm0 + j];
        double re[P], im[P], Ap, Ac;
        wp.tau.d[0] = ld(L::TD);
        wp.tau.d[1] = ld(L::TD + 1);
        arm_derivs<MODEL, NT, P>(wp, ev, sc, g, arm, S.motion, f, re, im, &Ap, &Ac, sw * arm.swgt);
        gram_update_unit<P>(G, re, im);
      }
    }
  } else {
    (void)mask;
    for (int k = 0; k < net.n_sum_arms; ++k) {
      const ArmK& arm = net.sum_arm[k];
      const int s = arm.site;
      const SiteK& S = net.site[s];
      const double* sc = rec + REC_SITE + s * SC_COUNT;
      if (s != cur_site) {
        if (S.psd != cur_psd) {
          sw = psd_sw(S.psd);
          cur_psd = S.psd;
        }
        site_geo<true>(S, sc, ev, site_tn(S.motion, tc, ld(L::TV)), ld(L::CD), ld(L::SD), g);
        cur_site = s;
      }
      double re[P], im[P], Ap, Ac;
      wp.tau.d[0] = ld(L::TD);
      wp.tau.d[1] = ld(L::TD + 1);
      arm_derivs<MODEL, NT, P>(wp, ev, sc, g, arm, S.motion, f, re, im, &Ap, &Ac, sw * arm.swgt);
      gram_update_unit<P>(G, re, im);
    }
  }
}

// Consumer for the 11-parameter models' network sum with the "pure" rows
// factored out.  Four of the 11 rows are the arm's response u = Ap + i Ac
// times a complex factor c_p that does not depend on the arm: dL (c = -1/dL),
// Phicoal (c = -i) and the two spins (c = A'/A - i Phi', the amplitude and
// phase tangents).  With T_j = sum_arms u conj(d_j) over the 7 other rows j
// and W = sum_arms |u|^2 at a point, their Fisher entries are Re(c_p T_j)
// and Re(c_p conj(c_q)) W: per arm that leaves the 28-entry Gram of the 7
// rows, T and W (86 FMAs instead of 132), and per point the 38 entries
// involving a pure row.  Rows scaled by sqrt(weight) make T and W weighted.

template <int P>
GWF_HD constexpr int pk(int a, int b) {  // packed upper-triangle index, a <= b
  return a * P - a * (a - 1) / 2 + (b - a);
}

template <int MODEL, int NT, bool MASKED = false>
GWF_HD void pc_consume_pure(const NetK& net, const double* rec, int base, int res, int lane, const double* buf,
                            double* G, uint32_t mask = 0u) {
  constexpr int P = 11;
  typedef PcLayout<NT> L;
#if defined(__HIP_DEVICE_COMPILE__) && !defined(GWF_NO_LOAD_BARRIER)
  __asm__ volatile("" ::: "memory");
#endif
  if (base + lane >= res) return;
  const double* ev = rec + REC_EV;
  const double f = buf[L::F * 64 + lane];
  WfPoint<MODEL, NT> wp;
  auto ld = [&](int slot) { return buf[slot * 64 + lane]; };
  // the arm rows need the (Mc, eta) tangents only; the spins' are read after
  // the arms, for their pure-row factors
  wp.A.v = ld(L::AV);
  wp.Phi.v = 0.0;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    wp.A.d[k] = k < 2 ? ld(L::AD + k) : 0.0;
    wp.Phi.d[k] = k < 2 ? ld(L::PD + k) : 0.0;
  }
  const double tc = ev[EV_TC];
  int cur_site = -1, cur_psd = -1;
  double sw = 0.0;
  Geo g;
  auto psd_sw = [&](int psd) {
    return psd == 0 ? ld(L::SW0)
                    : (psd == 1 ? ld(L::SW1) : sqrt(ld(L::WT) / psd_at(net.psd[psd], f, ld(L::LNF))));
  };
  constexpr int NP[7] = {0, 1, 3, 4, 5, 6, 7};  // the rows that are not pure
  double Tr[7], Ti[7], W = 0.0;
#pragma unroll
  for (int j = 0; j < 7; ++j) Tr[j] = Ti[j] = 0.0;
  // one arm of site s: the 7 rows' Gram, T and W
  auto do_arm = [&](const ArmK& arm, int s) {
    const SiteK& S = net.site[s];
    const double* sc = rec + REC_SITE + s * SC_COUNT;
    if (s != cur_site) {
      if (S.psd != cur_psd) {
        sw = psd_sw(S.psd);
        cur_psd = S.psd;
      }
      site_geo<true>(S, sc, ev, site_tn(S.motion, tc, ld(L::TV)), ld(L::CD), ld(L::SD), g);
      cur_site = s;
    }
    double re[P], im[P], Ap, Ac;
    wp.tau.d[0] = ld(L::TD);
    wp.tau.d[1] = ld(L::TD + 1);
    arm_derivs<MODEL, NT, P>(wp, ev, sc, g, arm, S.motion, f, re, im, &Ap, &Ac, sw * arm.swgt);
#pragma unroll
    for (int a = 0; a < 7; ++a) {
#pragma unroll
      for (int b = a; b < 7; ++b) {
        double& x = G[pk<P>(NP[a], NP[b])];
        x = fma(im[NP[a]], im[NP[b]], fma(re[NP[a]], re[NP[b]], x));
      }
      Tr[a] = fma(Ac, im[NP[a]], fma(Ap, re[NP[a]], Tr[a]));
      Ti[a] = fma(-Ap, im[NP[a]], fma(Ac, re[NP[a]], Ti[a]));
    }
    W = fma(Ac, Ac, fma(Ap, Ap, W));
  };
  if constexpr (MASKED) {  // the arms whose bit is set; a site with all set through its sum arms
    for (int s = 0; s < net.n_sites; ++s) {
      const SiteK& S = net.site[s];
      const uint32_t all = (1u << S.narm) - 1u;
      const uint32_t bits = (mask >> S.arm0) & all;
      if (!bits) continue;
      const bool full = bits == all;
      const int cnt = full ? S.nsum : S.narm;
      for (int j = 0; j < cnt; ++j) {
        if (!full && !((bits >> j) & 1u)) continue;
        do_arm(full ? net.sum_arm[S.sum0 + j] : net.arm[S.arm0 + j], s);
      }
    }
  } else {
    (void)mask;
    for (int k = 0; k < net.n_sum_arms; ++k) do_arm(net.sum_arm[k], net.sum_arm[k].site);
  }
  // c_p = cr + i ci: dL (-1/dL, 0), Phicoal (0, -1), chi1z, chi2z
  // (A'/A, -Phi'); a zero amplitude (past a cut) has u = 0 and c = 0 is used
  const double idL0 = ev[EV_IDL];
  const double ia = fabs(wp.A.v) > AMP_TINY ? 1.0 / wp.A.v : 0.0;
  const double a9 = ld(L::AD + 2) * ia, b9 = -ld(L::PD + 2);
  const double a10 = ld(L::AD + 3) * ia, b10 = -ld(L::PD + 3);
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    const int r = NP[j];
    double& g2 = G[r < 2 ? pk<P>(r, 2) : pk<P>(2, r)];
    g2 = fma(-idL0, Tr[j], g2);
    G[pk<P>(r, 8)] += Ti[j];
    double& g9 = G[pk<P>(r, 9)];
    g9 = fma(a9, Tr[j], fma(-b9, Ti[j], g9));
    double& g10 = G[pk<P>(r, 10)];
    g10 = fma(a10, Tr[j], fma(-b10, Ti[j], g10));
  }
  G[pk<P>(2, 2)] = fma(idL0 * idL0, W, G[pk<P>(2, 2)]);
  G[pk<P>(2, 9)] = fma(-idL0 * a9, W, G[pk<P>(2, 9)]);
  G[pk<P>(2, 10)] = fma(-idL0 * a10, W, G[pk<P>(2, 10)]);
  G[pk<P>(8, 8)] += W;
  G[pk<P>(8, 9)] = fma(-b9, W, G[pk<P>(8, 9)]);
  G[pk<P>(8, 10)] = fma(-b10, W, G[pk<P>(8, 10)]);
  G[pk<P>(9, 9)] = fma(fma(a9, a9, b9 * b9), W, G[pk<P>(9, 9)]);
  G[pk<P>(9, 10)] = fma(fma(a9, a10, b9 * b10), W, G[pk<P>(9, 10)]);
  G[pk<P>(10, 10)] = fma(fma(a10, a10, b10 * b10), W, G[pk<P>(10, 10)]);
}

// pc_consume_pure with the Mc, eta and tcoal rows factored as well.  Their
// Earth-rotation terms are one arm vector v times an arm-independent t_noloc
// tangent s_j, so every row outside theta, phi, iota, psi is
// d_j = u c_j + v s_j:  Mc, eta: c = A'/A - i Phi', s = -tau'/day (rotating
// Earth) or 0; tcoal: c = 2 pi i f, s = 1/day; dL, Phicoal, chi1z, chi2z:
// s = 0.  Per arm: the Gram of the 4 sky rows, Tu_k = sum d_k conj(u),
// Tv_k = Re sum d_k conj(v), sum |u|^2, sum |v|^2, sum u conj(v) (52 FMAs;
// 86 with only the pure rows factored); per point the other entries:
// F_kj = Re(Tu_k conj c_j) + s_j Tv_k,
// F_jl = Re(c_j conj c_l) W_uu + s_j s_l W_vv + s_l Re(c_j W_uv) + s_j Re(c_l W_uv).
// s_j must not vary between sites: not for a network that mixes rotating-
// and fixed-Earth sites (sites without motion have v = 0).
template <int MODEL, int NT, bool MASKED = false>
GWF_HD void pc_consume_semi(const NetK& net, const double* rec, int base, int res, int lane, const double* buf,
                            double* G, uint32_t mask = 0u) {
  constexpr int P = 11;
  typedef PcLayout<NT> L;
#if defined(__HIP_DEVICE_COMPILE__) && !defined(GWF_NO_LOAD_BARRIER)
  __asm__ volatile("" ::: "memory");
#endif
  if (base + lane >= res) return;
  const double* ev = rec + REC_EV;
  const double f = buf[L::F * 64 + lane];
  WfPoint<MODEL, NT> wp;
  auto ld = [&](int slot) { return buf[slot * 64 + lane]; };
  // the sky rows need no waveform or tau tangent
  wp.A.v = ld(L::AV);
  wp.Phi.v = wp.tau.v = wp.tau.d[0] = wp.tau.d[1] = 0.0;
#pragma unroll
  for (int k = 0; k < NT; ++k) wp.A.d[k] = wp.Phi.d[k] = 0.0;
  const double tc = ev[EV_TC];
  int cur_site = -1, cur_psd = -1;
  double sw = 0.0;
  Geo g;
  auto psd_sw = [&](int psd) {
    return psd == 0 ? ld(L::SW0)
                    : (psd == 1 ? ld(L::SW1) : sqrt(ld(L::WT) / psd_at(net.psd[psd], f, ld(L::LNF))));
  };
  constexpr int SK[4] = {3, 4, 5, 6};  // theta, phi, iota, psi
  double Tr[4], Ti[4], Tv[4], Wuu = 0.0, Wvv = 0.0, Wr = 0.0, Wi = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) Tr[j] = Ti[j] = Tv[j] = 0.0;
  auto do_arm = [&](const ArmK& arm, int s) {
    const SiteK& S = net.site[s];
    const double* sc = rec + REC_SITE + s * SC_COUNT;
    if (s != cur_site) {
      if (S.psd != cur_psd) {
        sw = psd_sw(S.psd);
        cur_psd = S.psd;
      }
      site_geo<true>(S, sc, ev, site_tn(S.motion, tc, ld(L::TV)), ld(L::CD), ld(L::SD), g);
      cur_site = s;
    }
    double re[P], im[P], Ap, Ac, v[2];
    arm_derivs<MODEL, NT, P>(wp, ev, sc, g, arm, S.motion, f, re, im, &Ap, &Ac, sw * arm.swgt, v);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int k = SK[a];
#pragma unroll
      for (int b = a; b < 4; ++b) {
        double& x = G[pk<P>(k, SK[b])];
        x = fma(im[k], im[SK[b]], fma(re[k], re[SK[b]], x));
      }
      Tr[a] = fma(Ac, im[k], fma(Ap, re[k], Tr[a]));
      Ti[a] = fma(-Ap, im[k], fma(Ac, re[k], Ti[a]));
      Tv[a] = fma(v[1], im[k], fma(v[0], re[k], Tv[a]));
    }
    Wuu = fma(Ac, Ac, fma(Ap, Ap, Wuu));
    Wvv = fma(v[1], v[1], fma(v[0], v[0], Wvv));
    Wr = fma(Ac, v[1], fma(Ap, v[0], Wr));   // Re(u conj v)
    Wi = fma(-Ap, v[1], fma(Ac, v[0], Wi));  // Im(u conj v)
  };
  if constexpr (MASKED) {
    for (int s = 0; s < net.n_sites; ++s) {
      const SiteK& S = net.site[s];
      const uint32_t all = (1u << S.narm) - 1u;
      const uint32_t bits = (mask >> S.arm0) & all;
      if (!bits) continue;
      const bool full = bits == all;
      const int cnt = full ? S.nsum : S.narm;
      for (int j = 0; j < cnt; ++j) {
        if (!full && !((bits >> j) & 1u)) continue;
        do_arm(full ? net.sum_arm[S.sum0 + j] : net.arm[S.arm0 + j], s);
      }
    }
  } else {
    (void)mask;
    for (int k = 0; k < net.n_sum_arms; ++k) do_arm(net.sum_arm[k], net.sum_arm[k].site);
  }
  // the factored rows: index, c_j = cr + i ci, s_j (HR / HI / HS: the parts
  // that are not identically zero); the producer wrote A'/A and s_j of Mc
  // and eta (pc_produce, factored)
  constexpr int FR[7] = {0, 1, 7, 2, 8, 9, 10};
  constexpr bool HR[7] = {true, true, false, true, false, true, true};
  constexpr bool HI[7] = {true, true, true, false, true, true, true};
  constexpr bool HS[7] = {true, true, true, false, false, false, false};
  double cr[7], ci[7], sj[7];
  cr[0] = ld(L::AD);
  ci[0] = -ld(L::PD);
  sj[0] = ld(L::TD);
  cr[1] = ld(L::AD + 1);
  ci[1] = -ld(L::PD + 1);
  sj[1] = ld(L::TD + 1);
  cr[2] = 0.0;
  ci[2] = 2. * PI * f;
  sj[2] = INV_DAY;
  cr[3] = -ev[EV_IDL];
  ci[3] = 0.0;
  cr[4] = 0.0;
  ci[4] = -1.0;
  cr[5] = ld(L::AD + 2);
  ci[5] = -ld(L::PD + 2);
  cr[6] = ld(L::AD + 3);
  ci[6] = -ld(L::PD + 3);
#pragma unroll
  for (int j = 3; j < 7; ++j) sj[j] = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int k = SK[a];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      const int r = FR[j];
      double& x = G[k < r ? pk<P>(k, r) : pk<P>(r, k)];
      double t = x;
      if (HR[j]) t = fma(cr[j], Tr[a], t);
      if (HI[j]) t = fma(-ci[j], Ti[a], t);
      if (HS[j]) t = fma(sj[j], Tv[a], t);
      x = t;
    }
  }
  // F_jl = y_j^T M y_l, y = (cr, ci, s), M = [[Wuu, 0, Wr], [0, Wuu, -Wi],
  // [Wr, -Wi, Wvv]]: m = M y_l once per row, then up to three FMAs per entry
  double m1[7], m2[7], m3[7];
#pragma unroll
  for (int l = 0; l < 7; ++l) {
    m1[l] = HR[l] ? (HS[l] ? fma(Wr, sj[l], Wuu * cr[l]) : Wuu * cr[l]) : (HS[l] ? Wr * sj[l] : 0.0);
    m2[l] = HI[l] ? (HS[l] ? fma(-Wi, sj[l], Wuu * ci[l]) : Wuu * ci[l]) : (HS[l] ? -Wi * sj[l] : 0.0);
    double t = 0.0;
    bool any = false;
    if (HR[l]) {
      t = Wr * cr[l];
      any = true;
    }
    if (HI[l]) {
      t = any ? fma(-Wi, ci[l], t) : -Wi * ci[l];
      any = true;
    }
    if (HS[l]) t = any ? fma(Wvv, sj[l], t) : Wvv * sj[l];
    m3[l] = t;
  }
#pragma unroll
  for (int j = 0; j < 7; ++j) {
#pragma unroll
    for (int l = j; l < 7; ++l) {
      const int r = FR[j], c = FR[l];
      double& x = G[r < c ? pk<P>(r, c) : pk<P>(c, r)];
      double t = x;
      if (HR[j] && (HR[l] || HS[l])) t = fma(cr[j], m1[l], t);
      if (HI[j] && (HI[l] || HS[l])) t = fma(ci[j], m2[l], t);
      if (HS[j]) t = fma(sj[j], m3[l], t);
      x = t;
    }
  }
}

// --------------------------------------------------------------- SNR body
// Per-arm trapz((Ap^2 + Ac^2)/S, f) of every arm of the network in one grid
// walk: the amplitude, tau(f) and the PSD segments are evaluated once per
// frequency and shared by the sites.  acc[a] accumulates physical arm a
// (a < MAXA, the network's arm count rounded up to a compile-time bound).
template <int MODEL, int MAXA = GWF_MAX_ARMS>
GWF_HD void snr_event(const gwf_waveform& w, const NetK& net, const double* rec, int res, int lane, int nlanes,
                      double acc[MAXA]) {
  const double* ev = rec + REC_EV;
  const double a = log10(net.fmin);
  const double b = log10(fcut_eff(ev[EV_FCUT], net.fmax));
  const double tc = ev[EV_TC], idL = ev[EV_IDL], cp = ev[EV_CP], cc = ev[EV_CC];
  const bool earth = net.any_earth != 0;
  GridWalk<false> gw;
  gw.init(a, b, res, lane, nlanes);
  for (int base = 0; base < res; base += nlanes) {
    const int i = base + lane;
    double f, wt, lnf;
    gw.step(base, res, &f, &wt, &lnf);
    if (i >= res) continue;
    PsdPre pre0, pre1;
    psd_prefetch(net.psd[0], f, lnf, pre0);
    if (net.n_psd > 1) psd_prefetch(net.psd[1], f, lnf, pre1);
    double tau;
    const double Av = wf_amp_tau_val<MODEL>(w, rec, f, lnf, earth, &tau) * idL;
    double cdel, sdel;
    earth_delta(earth, tau, &cdel, &sdel);
    int cur_site = -1, cur_psd = -1;
    double wS = 0.0, P1 = 0.0, P2 = 0.0;
    Geo g;
#pragma unroll
    for (int k = 0; k < MAXA; ++k) {
      if (k < net.n_arms) {
        const ArmK& A = net.arm[k];
        const int s = A.site;
        if (s != cur_site) {
          const SiteK& S = net.site[s];
          const double* sc = rec + REC_SITE + s * SC_COUNT;
          if (S.psd != cur_psd) {
            double Sf;
            if (S.psd == 0)
              Sf = psd_finish(net.psd[0], f, lnf, pre0);
            else if (S.psd == 1)
              Sf = psd_finish(net.psd[1], f, lnf, pre1);
            else
              Sf = psd_at(net.psd[S.psd], f, lnf);
            wS = wt / Sf;
            cur_psd = S.psd;
          }
          site_geo<false>(S, sc, ev, site_tn(S.motion, tc, tau), cdel, sdel, g);
          P1 = sc[SC_P1];
          P2 = sc[SC_P2];
          cur_site = s;
        }
        const double aa = A.s2x * g.Sa - A.c2x * g.Ca;
        const double bb = A.c2x * g.Cb + A.s2x * g.Sb;
        const double Ap = Av * (aa * P1 + bb * P2) * cp;
        const double Ac = Av * (bb * P1 - aa * P2) * cc;
        acc[k] += wS * (Ap * Ap + Ac * Ac);
      }
    }
  }
}

// ------------------------------------------------------------- strain body
// h = (Ap + i Ac) exp(i Psi) and the derivative rows dX h of one arm on a
// caller grid (signal.py:485-655, 1102-1301), Psi = 2 pi f tc 86400 - Phic -
// Phi + phiL (signal.py:483, 628-636).  f: this event's column, row stride
// fs.  h: complex pairs with point stride hs; J: complex pairs with parameter
// stride js (point stride 1 pair).
template <int MODEL, int NT, int P>
GWF_HD void strain_event(const gwf_waveform& w, const NetK& net, const double* rec, int armi, const double* f, int64_t fs,
                         int64_t nf, int lane, int nlanes, double* h, int64_t hs, double* J, int64_t js,
                         double* comp = nullptr, int64_t cis = 0, int64_t ccs = 0) {
  const double* ev = rec + REC_EV;
  const ArmK& arm = net.arm[armi];
  const int s = arm.site;
  const SiteK& S = net.site[s];
  const double* sc = rec + REC_SITE + s * SC_COUNT;
  const double tc = ev[EV_TC], phic = ev[EV_PHIC];
  for (int64_t i = lane; i < nf; i += nlanes) {
    const double fi = f[i * fs];
    WfPoint<MODEL, NT> wp;
    wf_point<MODEL, NT>(w, rec, fi, S.motion == MOTION_EARTH, wp);
    double cdel, sdel;
    earth_delta(S.motion == MOTION_EARTH, wp.tau.v, &cdel, &sdel);
    Geo g;
    site_geo<true>(S, sc, ev, site_tn(S.motion, tc, wp.tau.v), cdel, sdel, g);
    double re[P], im[P], Ap, Ac;
    arm_derivs<MODEL, NT, P>(wp, ev, sc, g, arm, S.motion, fi, re, im, &Ap, &Ac);
    const double psi = (2. * PI * fi * (tc * 3600. * 24.) - phic - wp.Phi.v) + 0. + (2. * PI * fi) * g.dt;
    if (comp) {  // GWstrain(return_single_comp=...): Ap, Ac and the phase Psi
      comp[i * cis] = Ap;
      comp[ccs + i * cis] = Ac;
      comp[2 * ccs + i * cis] = psi;
    }
    if (!h && !J) continue;
    double sp, cps;
    sincos_d(psi, &sp, &cps);
    if (h) {
      h[i * hs + 0] = Ap * cps - Ac * sp;
      h[i * hs + 1] = Ap * sp + Ac * cps;
    }
    if (J) {
#pragma unroll
      for (int X = 0; X < P; ++X) {
        J[X * js + 2 * i + 0] = re[X] * cps - im[X] * sp;
        J[X * js + 2 * i + 1] = re[X] * sp + im[X] * cps;
      }
    }
  }
}

}  // namespace gwf
