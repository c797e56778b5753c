// Host-side construction of the network argument block (sites, arms, PSD
// tables) from the ABI's detector descriptions (signal.py:113-147 for the
// geometry, 144-147 for the arm angle, 745-765 / 1057 for the triangle's
// null-stream third arm).
#pragma once
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <string>

#include "gwf_paths.h"

namespace gwf {

// Builds the network argument block (sites, arms, PSD tables).
inline int build_net(const gwf_detector* dets, int n_det, int res, NetK& net, std::string& err) {
  std::memset(&net, 0, sizeof(net));
  if (n_det < 1 || n_det > GWF_MAX_DETECTORS) { err = "number of detectors must be 1..8"; return -EINVAL; }
  net.n_sites = n_det;
  net.res = res;
  net.fmin = dets[0].fmin;
  net.fmax = dets[0].fmax;
  net.lin = dets[0].lin_spacing != 0;
  net.spin_limit = 1u << 22;  // the producer / consumer launcher sets its own
  int na = 0, np = 0;
  for (int d = 0; d < n_det; ++d) {
    const gwf_detector& g = dets[d];
    if (g.shape != GWF_SHAPE_L && g.shape != GWF_SHAPE_T) { err = "Enter valid detector configuration"; return -EINVAL; }
    if (!g.psd_f || !g.psd_S || g.n_psd < 2) { err = "Enter a valid PSD or ASD path"; return -EINVAL; }
    if (g.fmin != net.fmin || g.fmax != net.fmax)
      { err = "detectors of one call must share fmin and fmax"; return -EINVAL; }
    if ((g.lin_spacing != 0) != (net.lin != 0))
      { err = "detectors of one call must share the grid spacing"; return -EINVAL; }
    if (!(g.fmin > 0.0)) { err = "fmin must be positive"; return -EINVAL; }
    // PSD table index (tables shared between detectors are evaluated once per site anyway)
    int p = -1;
    for (int q = 0; q < np; ++q)
      if (net.psd[q].f == g.psd_f && net.psd[q].S == g.psd_S) p = q;
    if (p < 0) {
      p = np++;
      PsdTab& t = net.psd[p];
      t.f = g.psd_f;
      t.S = g.psd_S;
      t.n = g.n_psd;
      if (!(g.psd_f_first > 0.0) || !(g.psd_f_last > g.psd_f_first))
        { err = "PSD frequency bounds must be positive and increasing"; return -EINVAL; }
      t.lx0 = std::log(g.psd_f_first);
      t.idl = (double)(g.n_psd - 1) / (std::log(g.psd_f_last) - t.lx0);
      t.f_first = g.psd_f_first;
      t.f_last = g.psd_f_last;
    }
    SiteK& S = net.site[d];
    const double lat = g.lat_deg * M_PI / 180., lon = g.long_deg * M_PI / 180., xax = g.xax_deg * M_PI / 180.;
    S.clat = std::cos(lat);
    S.slat = std::sin(lat);
    S.s2lat = std::sin(2. * lat);
    S.c2lat = std::cos(2. * lat);
    S.lon = lon;
    S.sinzeta = std::sin(g.shape == GWF_SHAPE_L ? 0.5 * M_PI : M_PI / 3.);
    S.motion = g.no_motion ? MOTION_NONE : (g.earth_motion ? MOTION_EARTH : MOTION_FIXED);
    S.psd = p;
    if (S.motion == MOTION_EARTH) net.any_earth = 1;
    const int narms = g.shape == GWF_SHAPE_L ? 1 : 3;
    if (na + narms > GWF_MAX_ARMS) { err = "too many arms"; return -EINVAL; }
    S.arm0 = na;
    S.narm = narms;
    S.sum0 = net.n_sum_arms;
    double s2[3], c2[3];
    for (int k = 0; k < 3; ++k) {
      const double rot = (k * 60.) * M_PI / 180.;
      s2[k] = std::sin(2 * (xax + rot));
      c2[k] = std::cos(2 * (xax + rot));
    }
    for (int k = 0; k < narms; ++k) {
      ArmK& A = net.arm[na + k];
      A.site = d;
      if (k == 2 && g.compute2arms) {  // null stream: arm 3 = -(arm 1 + arm 2) (signal.py:751, 1057)
        A.s2x = -(s2[0] + s2[1]);
        A.c2x = -(c2[0] + c2[1]);
      } else {
        A.s2x = s2[k];
        A.c2x = c2[k];
      }
      A.wgt = 1.0;
      A.swgt = 1.0;
    }
    // network-sum arms of this site: eigenpairs of M = sum_k (s_k, c_k)^T (s_k, c_k)
    double m11 = 0., m12 = 0., m22 = 0.;
    for (int k = 0; k < narms; ++k) {
      const ArmK& A = net.arm[na + k];
      m11 += A.s2x * A.s2x;
      m12 += A.s2x * A.c2x;
      m22 += A.c2x * A.c2x;
    }
    if (narms == 1) {
      net.sum_arm[net.n_sum_arms++] = net.arm[na];
    } else {
      // Jacobi rotation diagonalising the 2x2 symmetric M
      double c = 1.0, s = 0.0;
      if (m12 != 0.0) {
        const double th = (m22 - m11) / (2.0 * m12);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        c = 1.0 / std::sqrt(t * t + 1.0);
        s = t * c;
      }
      const double l1 = m11 - (s / c) * m12 * (c != 0 ? 1.0 : 0.0);
      const double l2 = m22 + (s / c) * m12 * (c != 0 ? 1.0 : 0.0);
      ArmK A1, A2;
      A1.site = A2.site = d;
      A1.s2x = c; A1.c2x = -s; A1.wgt = l1; A1.swgt = std::sqrt(std::max(l1, 0.0));
      A2.s2x = s; A2.c2x = c; A2.wgt = l2; A2.swgt = std::sqrt(std::max(l2, 0.0));
      if (l1 != 0.0) net.sum_arm[net.n_sum_arms++] = A1;
      if (l2 != 0.0) net.sum_arm[net.n_sum_arms++] = A2;
    }
    S.nsum = net.n_sum_arms - S.sum0;
    na += narms;
  }
  net.n_arms = na;
  net.n_psd = np;
  bool any_fixed = false;
  for (int d = 0; d < n_det; ++d) any_fixed = any_fixed || net.site[d].motion == MOTION_FIXED;
  net.mixed_motion = net.any_earth && any_fixed;
  return 0;
}


}  // namespace gwf
