// libgwfast_hip.so: HIP kernels for gfx950 (MI355X) and the C ABI declared in
// include/gwfast_hip.h.
//
// Launch geometry: one 64-wide wavefront per event (4 events per 256-thread
// workgroup); the event index is wave-uniform so every per-event coefficient
// is read with scalar loads from the event record, and each lane walks the
// frequency grid with stride 64 (consecutive lanes = consecutive
// frequencies, so region branches of the piecewise models rarely diverge).
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "gwf_paths.h"
#include "gwf_net.h"

using namespace gwf;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int build_net_checked(const gwf_detector* dets, int n_det, int res, NetK& net) {
  std::string msg;
  const int rc = build_net(dets, n_det, res, net, msg);
  return rc ? fail(rc, msg) : 0;
}

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return fail(-EIO, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr int WAVES_PER_BLOCK = 4;
constexpr int BLOCK = 64 * WAVES_PER_BLOCK;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int64_t wave_event() {
  const int wave = threadIdx.x >> 6;
  const int64_t e = (int64_t)blockIdx.x * WAVES_PER_BLOCK + wave;
  return (int64_t)__builtin_amdgcn_readfirstlane((int)e);
}

// ------------------------------------------------------------------ kernels
template <int MODEL, int NT>
__global__ __launch_bounds__(256) void k_setup(gwf_waveform w, NetK net, const double* __restrict__ par, int64_t ld,
                                               int64_t n_ev, const double* __restrict__ fgrid, int64_t nf,
                                               double* __restrict__ rec) {
  // IMRPhenomD models: NT lanes per event, lane tk evaluating tangent tk
  // (setup_waveform_split); TaylorF2: one lane per event
  constexpr int LPE = is_pd(MODEL) ? NT : 1;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = tid / LPE;
  const int tk = (int)(tid - e * LPE);
  if (e >= n_ev) return;
  double f0;
  if (fgrid) {  // minimum of the caller's grid column (waveforms.py:1139)
    f0 = fgrid[e];
    for (int64_t i = 1; i < nf; ++i) f0 = fmin(f0, fgrid[i * n_ev + e]);
  } else {  // first grid point: 10**log10(fmin) for geomspace, fmin for linspace
    f0 = net.lin ? net.fmin : grid_f(log10(net.fmin), log10(net.fmin), 0, 1);
  }
  double* r = rec + e * (int64_t)REC_DOUBLES;
  if constexpr (is_pd(MODEL)) {
    setup_waveform_split<MODEL>(w, par, ld, e, f0, r, tk);
    if (tk == 0) setup_event_consts<MODEL>(net, par, ld, e, r);  // reads only lane 0's stores
  } else {
    setup_event<MODEL, NT>(w, net, par, ld, e, f0, r);
  }
}

#ifndef GWF_FISHER_MIN_BLOCKS
#define GWF_FISHER_MIN_BLOCKS 1
#endif

// Wave-reduces the packed Gram and writes the symmetric (P, P, n_ev) output.
template <int P>
__device__ __forceinline__ void store_packed(const double* G, double* __restrict__ F, int64_t n_ev, int64_t e, int lane) {
  int k = 0;
#pragma unroll
  for (int a = 0; a < P; ++a) {
#pragma unroll
    for (int b = a; b < P; ++b) {
      const double s = wave_sum(G[k]);
      if (lane == (k & 63)) {
        F[(int64_t)(a * P + b) * n_ev + e] = s;
        if (a != b) F[(int64_t)(b * P + a) * n_ev + e] = s;
      }
      ++k;
    }
  }
}

template <int MODEL, int NT, int P, bool LIN>
__global__ __launch_bounds__(256, GWF_FISHER_MIN_BLOCKS) void k_fisher(gwf_waveform w, NetK net, const double* __restrict__ rec, int64_t n_ev,
                                                int arm_sel, double* __restrict__ F) {
  const int64_t e = wave_event();
  if (e >= n_ev) return;  // wave-uniform exit
  const int lane = threadIdx.x & 63;
  constexpr int NG = P * (P + 1) / 2;
  double G[NG];
#pragma unroll
  for (int k = 0; k < NG; ++k) G[k] = 0.0;
  fisher_event<MODEL, NT, P, false, LIN>(w, net, rec + e * (int64_t)REC_DOUBLES, net.res, lane, 64, arm_sel, G);
  store_packed<P>(G, F, n_ev, e, lane);
}

// Events per block of the producer / consumer kernel k_fisher_pq (waves
// 0..EPB-1 produce, EPB..2 EPB-1 consume; two waves per SIMD).
#ifndef GWF_PC_EPB
#define GWF_PC_EPB 4  // measured: 4 events per 512-thread block 2.65 ms, 2 per 256-thread block 2.80 ms
#endif
constexpr int PC_EPB = GWF_PC_EPB;

// Producer / consumer pairs without block barriers: each pair (producer wave
// `slot`, consumer wave EPB + slot, one SIMD) hands 64-point chunks through a
// ring of PQ_NBUF LDS buffers, synchronised by two per-pair LDS counters
// (items produced, items consumed) with workgroup-scope release stores and
// acquire loads.  A pair runs its own events one after another, so pairs of
// one block never wait for each other: RAGGED (per-event grid lengths res_ev)
// takes events from the device counter counter[0]; otherwise pair p of the
// grid takes events p, p + n_pairs, ...  Every wait polls at most
// net.spin_limit times (with s_sleep; PQ_SPIN_LIMIT unless
// gwf_set_queue_spin_limit changed it); a wave whose wait runs out sets
// counter[1] and leaves, and k_pq_poison then fills the whole output with
// NaN, so the launch always ends and a failure cannot go unnoticed.
#ifndef GWF_PQ_NBUF
#define GWF_PQ_NBUF 3  // measured (IMRPhenomD, 10^4 x 2000): 3 buffers 2.38 ms, 2 buffers 2.44 ms, 4 as 3
#endif
constexpr int PQ_NBUF = GWF_PQ_NBUF;
constexpr int PQ_LDS_BYTES = 160 * 1024;  // LDS of one CU (one resident block)
constexpr unsigned PQ_SPIN_LIMIT = 1u << 22;  // >= 0.1 s of polling; a chunk takes ~10 us
unsigned g_pq_spin_limit = PQ_SPIN_LIMIT;      // gwf_set_queue_spin_limit

__device__ __forceinline__ void pq_release(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waits (wave-uniformly) until *p >= target; false after limit polls.
__device__ __forceinline__ bool pq_wait(int* p, int target, unsigned limit) {
  for (unsigned n = 0; n < limit; ++n) {
    const int v = __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (__builtin_amdgcn_readfirstlane(v) >= target) return true;
    __builtin_amdgcn_s_sleep(1);
  }
  return false;
}

__device__ __forceinline__ void pq_fail(unsigned int* counter, int lane) {
  if (lane == 0) atomicOr(counter + 1, 1u);
}

// Gram on the matrix cores: the arm rows of the chunk's 64 points, scaled by
// sqrt(w) (w = trapezoid weight x arm weight / S), go through a per-wave LDS
// tile X[point][param] (row stride PQ_XS doubles) and G += X^T X runs as 16
// v_mfma_f64_16x16x4f64 per (arm, re|im): lane l feeds X[4 s + l / 16][l % 16]
// as both operands (A[i][k] is lane i + 16 k, B[k][j] lane j + 16 k), and the
// 16 x 16 accumulator holds G[i][j] in lane j + 16 (i % 4), element i / 4.
// Points past res contribute zero rows; NACC accumulators rotate so that
// consecutive MFMAs do not wait for each other.
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int PQ_XS = 17;

// Deepest ring (<= PQ_NBUF buffers) whose chunks (nv doubles per point), the
// headers, the producers' record copies (rec_lds) and, with the matrix-core
// Gram, the transposition tiles fit the LDS.
constexpr int pq_ring_depth(int nv, bool mx, bool rec_lds) {
  int nb = PQ_NBUF;
  while (nb > 2 && PC_EPB * (nb * (nv * 64 * 8 + 16) + 8 + (mx ? 64 * PQ_XS * 8 : 0) + (rec_lds ? REC_SITE * 8 : 0)) >
                       PQ_LDS_BYTES)
    --nb;
  return nb;
}
#ifndef GWF_PQ_NACC
#define GWF_PQ_NACC 2  // measured: 4 accumulators no faster
#endif
constexpr int PQ_NACC = GWF_PQ_NACC;

template <int P, int NA = PQ_NACC>
__device__ __forceinline__ void mfma_rows(const double* re, const double* im, double* X, int lane, v4d* acc) {
  const int rd = (lane >> 4) * PQ_XS + (lane & 15);
#pragma unroll
  for (int part = 0; part < 2; ++part) {
    const double* v = part ? im : re;
#pragma unroll
    for (int i = 0; i < P; ++i) X[lane * PQ_XS + i] = v[i];
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const double x = X[4 * t * PQ_XS + rd];
      acc[t % NA] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, x, acc[t % NA], 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Consumer with every Gram update on the matrix cores (mfma_rows); dead lanes
// (points past res) take part with zero rows.  MODE PQ_SUM: the network-sum
// arms into acc[0..PQ_NACC); PQ_MASKED: the arms whose bit is set in mask, a
// site with all arms set through its sum arms; PQ_ARMS: physical arm a alone
// into acc[a] (a < PQ_ARMS_MAX), for the per-arm Fisher matrices.
constexpr int PQ_SUM = 0, PQ_MASKED = 1, PQ_ARMS = 2;
constexpr int PQ_ARMS_MAX = 8;
template <int MODEL, int NT, int P, int MODE>
__device__ __forceinline__ void pc_consume_mx(const NetK& net, const double* rec, int base, int res, int lane,
                                              const double* buf, double* X, v4d* acc, uint32_t mask) {
  typedef PcLayout<NT> L;
  __asm__ volatile("" ::: "memory");
  const bool live = base + lane < res;
  const double* ev = rec + REC_EV;
  WfPoint<MODEL, NT> wp;
  double f = 0.0, wt = 0.0, cdel = 1.0, sdel = 0.0;
  if (live) {
    f = buf[L::F * 64 + lane];
    wt = buf[L::WT * 64 + lane];
    wp.A.v = buf[L::AV * 64 + lane];
    wp.Phi.v = 0.0;
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      wp.A.d[k] = buf[(L::AD + k) * 64 + lane];
      wp.Phi.d[k] = buf[(L::PD + k) * 64 + lane];
    }
    wp.tau.v = buf[L::TV * 64 + lane];
    wp.tau.d[0] = buf[L::TD * 64 + lane];
    wp.tau.d[1] = buf[(L::TD + 1) * 64 + lane];
    cdel = buf[L::CD * 64 + lane];
    sdel = buf[L::SD * 64 + lane];
  }
  const double tc = ev[EV_TC];
  int cur_site = -1, cur_psd = -1;
  double sw = 0.0;  // sqrt(4 wt / S) of the current PSD
  Geo g;
  // the scaled rows of one arm of site s (zero rows on dead lanes)
  auto arm_rows = [&](const ArmK& arm, int s, double* re, double* im) {
    if (live) {
      const SiteK& S = net.site[s];
      const double* sc = rec + REC_SITE + s * SC_COUNT;
      if (s != cur_site) {
        if (S.psd != cur_psd) {
          sw = S.psd == 0 ? buf[L::SW0 * 64 + lane]
                          : (S.psd == 1 ? buf[L::SW1 * 64 + lane]
                                        : sqrt(wt / psd_at(net.psd[S.psd], f, buf[L::LNF * 64 + lane])));
          cur_psd = S.psd;
        }
        site_geo<true>(S, sc, ev, site_tn(S.motion, tc, wp.tau.v), cdel, sdel, g);
        cur_site = s;
      }
      double Ap, Ac;
      arm_derivs<MODEL, NT, P>(wp, ev, sc, g, arm, S.motion, f, re, im, &Ap, &Ac, sw * arm.swgt);
    } else {
#pragma unroll
      for (int i = 0; i < P; ++i) re[i] = im[i] = 0.0;
    }
  };
  if constexpr (MODE == PQ_ARMS) {
    (void)mask;
#pragma unroll
    for (int a = 0; a < PQ_ARMS_MAX; ++a) {
      if (a < net.n_arms) {
        double re[P], im[P];
        arm_rows(net.arm[a], net.arm[a].site, re, im);
        mfma_rows<P, 1>(re, im, X, lane, acc + a);
      }
    }
  } else if constexpr (MODE == PQ_MASKED) {
    for (int s = 0; s < net.n_sites; ++s) {
      const SiteK& S = net.site[s];
      const uint32_t all = (1u << S.narm) - 1u;
      const uint32_t bits = (mask >> S.arm0) & all;
      if (!bits) continue;
      const bool full = bits == all;
      const int cnt = full ? S.nsum : S.narm;
      for (int j = 0; j < cnt; ++j) {
        if (!full && !((bits >> j) & 1u)) continue;
        double re[P], im[P];
        arm_rows(full ? net.sum_arm[S.sum0 + j] : net.arm[S.arm0 + j], s, re, im);
        mfma_rows<P>(re, im, X, lane, acc);
      }
    }
  } else {
    (void)mask;
    for (int k = 0; k < net.n_sum_arms; ++k) {
      double re[P], im[P];
      arm_rows(net.sum_arm[k], net.sum_arm[k].site, re, im);
      mfma_rows<P>(re, im, X, lane, acc);
    }
  }
}

// Consumer for 12-14 parameters with the pure rows factored out (as
// pc_consume_pure: dL, Phicoal and the waveform-only rows 9..P-1 are the
// arm response u times c_p = A'/A - i Phi', T_j = sum_arms u conj(d_j) over
// the 7 other rows and W = sum_arms |u|^2).  On the VALU: per arm the
// 28-entry Gram of the 7 rows, T and W; per point the pure x pure entries
// Re(c_p conj(c_q)) W.  On the matrix cores: the pure x other entries
// sum_points (c_pr Tr_j - c_pi Ti_j) -- per chunk each lane writes its
// point's [c_r | Tr] and then [-c_i | Ti] into the wave's LDS tile (pure
// factors in columns 0..NPU-1, T in 8..14) and 16 v_mfma_f64_16x16x4f64 per
// part contract over the 64 points (A[p][k] = tile[4t + k][p],
// B[k][j] = tile[4t + k][8 + j]; D[p][j] in lane j + 16 (p % 4), item p / 4).
// 32 MFMAs per chunk instead of 32 per arm.  G holds the 28 entries of the
// 7 rows (packed, pk<7>) and then the pure x pure ones (pk<NPU>).
template <int MODEL, int NT, int P>
__device__ __forceinline__ void pc_consume_pure_mx(const NetK& net, const double* rec, int base, int res, int lane,
                                                    const double* buf, double* X, double* G, v4d& acc) {
  constexpr int NPU = P - 7;
  static_assert(NPU >= 2 && NPU <= 8, "pure rows: dL, Phicoal and up to six waveform-only rows");
  typedef PcLayout<NT> L;
  __asm__ volatile("" ::: "memory");
  const bool live = base + lane < res;
  const double* ev = rec + REC_EV;
  constexpr int NP[7] = {0, 1, 3, 4, 5, 6, 7};
  double Tr[7], Ti[7], cr[NPU], ci[NPU];
#pragma unroll
  for (int j = 0; j < 7; ++j) Tr[j] = Ti[j] = 0.0;
#pragma unroll
  for (int p = 0; p < NPU; ++p) cr[p] = ci[p] = 0.0;
  if (live) {
    auto ld = [&](int slot) { return buf[slot * 64 + lane]; };
    const double f = ld(L::F);
    WfPoint<MODEL, NT> wp;
    wp.A.v = ld(L::AV);
    wp.Phi.v = 0.0;
#pragma unroll
    for (int k = 0; k < NT; ++k) {  // the 7 rows need the (Mc, eta) tangents only
      wp.A.d[k] = k < 2 ? ld(L::AD + k) : 0.0;
      wp.Phi.d[k] = k < 2 ? ld(L::PD + k) : 0.0;
    }
    const double tc = ev[EV_TC];
    int cur_site = -1, cur_psd = -1;
    double sw = 0.0, W = 0.0;
    Geo g;
    for (int k = 0; k < net.n_sum_arms; ++k) {
      const ArmK& arm = net.sum_arm[k];
      const int s = arm.site;
      const SiteK& S = net.site[s];
      const double* sc = rec + REC_SITE + s * SC_COUNT;
      if (s != cur_site) {
        if (S.psd != cur_psd) {
          sw = S.psd == 0 ? ld(L::SW0)
                          : (S.psd == 1 ? ld(L::SW1) : sqrt(ld(L::WT) / psd_at(net.psd[S.psd], f, ld(L::LNF))));
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
          double& x = G[pk<7>(a, b)];
          x = fma(im[NP[a]], im[NP[b]], fma(re[NP[a]], re[NP[b]], x));
        }
        Tr[a] = fma(Ac, im[NP[a]], fma(Ap, re[NP[a]], Tr[a]));
        Ti[a] = fma(-Ap, im[NP[a]], fma(Ac, re[NP[a]], Ti[a]));
      }
      W = fma(Ac, Ac, fma(Ap, Ap, W));
    }
    // c_p: dL (-1/dL, 0), Phicoal (0, -1), rows 9.. (A'/A, -Phi') with the
    // tangent of row r in slot r - 7; a zero amplitude (past a cut) has u = 0
    const double ia = fabs(wp.A.v) > AMP_TINY ? 1.0 / wp.A.v : 0.0;
    cr[0] = -ev[EV_IDL];
    ci[1] = -1.0;
#pragma unroll
    for (int p = 2; p < NPU; ++p) {
      cr[p] = ld(L::AD + p) * ia;
      ci[p] = -ld(L::PD + p);
    }
#pragma unroll
    for (int p = 0; p < NPU; ++p)
#pragma unroll
      for (int q = p; q < NPU; ++q) {
        double& x = G[28 + pk<NPU>(p, q)];
        x = fma(fma(cr[p], cr[q], ci[p] * ci[q]), W, x);
      }
  }
  const int rd = (lane >> 4) * PQ_XS;
#pragma unroll
  for (int part = 0; part < 2; ++part) {
#pragma unroll
    for (int p = 0; p < NPU; ++p) X[lane * PQ_XS + p] = part ? -ci[p] : cr[p];
#pragma unroll
    for (int j = 0; j < 7; ++j) X[lane * PQ_XS + 8 + j] = part ? Ti[j] : Tr[j];
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const double a = X[4 * t * PQ_XS + rd + (lane & 15)];
      const double b = X[4 * t * PQ_XS + rd + 8 + (lane & 7)];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Stores the Fisher of pc_consume_pure_mx: the VALU entries wave-reduced,
// the pure x other entries from the matrix-core accumulator.
template <int P>
__device__ __forceinline__ void store_pure_mx(const double* G, const v4d& acc, double* __restrict__ F, int64_t n_ev,
                                              int64_t e, int lane) {
  constexpr int NPU = P - 7;
  constexpr int NP[7] = {0, 1, 3, 4, 5, 6, 7};
  auto PU = [](int p) { return p == 0 ? 2 : 7 + p; };  // 2, 8, 9, ..., P-1
  auto put = [&](int a, int b, double v) {
    F[(int64_t)(a * P + b) * n_ev + e] = v;
    if (a != b) F[(int64_t)(b * P + a) * n_ev + e] = v;
  };
  int k = 0;
#pragma unroll
  for (int a = 0; a < 7; ++a)
#pragma unroll
    for (int b = a; b < 7; ++b, ++k) {
      const double s = wave_sum(G[pk<7>(a, b)]);
      if (lane == (k & 63)) put(NP[a], NP[b], s);
    }
#pragma unroll
  for (int p = 0; p < NPU; ++p)
#pragma unroll
    for (int q = p; q < NPU; ++q, ++k) {
      const double s = wave_sum(G[28 + pk<NPU>(p, q)]);
      if (lane == (k & 63)) put(PU(p), PU(q), s);
    }
  const int j = lane & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int p = (lane >> 4) + 4 * r;
    if (p < NPU && j < 7) put(PU(p), NP[j], acc[r]);
  }
}

// Consumer for 12-14 parameters with the Mc, eta and tcoal rows factored as
// well (every row outside theta, phi, iota, psi is d_j = u c_j + v s_j, as
// pc_consume_semi for 11).  On the VALU: per arm the 4 sky rows' Gram and
// their sums Tu_k, Tv_k against u and v; per point the entries among the
// NF = P - 4 factored rows, y_j^T M y_l with y = (Re c, Im c, s) and M the
// point's 3 x 3 of the W sums, and s_j Tv_k of the three rows with s_j != 0.
// On the matrix cores: the rest of the sky x factored entries,
// sum_points (Re c_j Tr_k - Im c_j Ti_k), per chunk through the wave's LDS
// tile ([Re c | Tr], then [-Im c | Ti]: c in columns 0..NF-1, T in 10..13)
// with 32 v_mfma_f64_16x16x4f64, as pc_consume_pure_mx.
template <int P>
__device__ __forceinline__ constexpr int semi_row(int j) {  // factored row j: 0, 1, 7, 2, 8, 9, ..., P-1
  return j < 3 ? (j == 2 ? 7 : j) : (j == 3 ? 2 : 4 + j);
}
__device__ __forceinline__ constexpr bool semi_hr(int j) { return j != 2 && j != 4; }  // Re c_j not identically 0
__device__ __forceinline__ constexpr bool semi_hi(int j) { return j != 3; }            // Im c_j not identically 0
__device__ __forceinline__ constexpr bool semi_hs(int j) { return j < 3; }             // s_j not identically 0
template <int MODEL, int NT, int P, bool MASKED = false>
__device__ __forceinline__ void pc_consume_semi_vx(const NetK& net, const double* rec, int base, int res, int lane,
                                                    const double* buf, double* X, double* G, v4d& acc,
                                                    uint32_t mask = 0u) {
  constexpr int NF = P - 4;
  constexpr int NFF = NF * (NF + 1) / 2;  // G: [factored x factored | s x sky (12) | sky x sky (10)]
  static_assert(NF >= 8 && NF <= 10, "12-14 parameters");
  typedef PcLayout<NT> L;
  __asm__ volatile("" ::: "memory");
  const bool live = base + lane < res;
  const double* ev = rec + REC_EV;
  constexpr int SK[4] = {3, 4, 5, 6};
  double cr[NF], ci[NF], Tr[4], Ti[4];
#pragma unroll
  for (int j = 0; j < NF; ++j) cr[j] = ci[j] = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) Tr[a] = Ti[a] = 0.0;
  if (live) {
    auto ld = [&](int slot) { return buf[slot * 64 + lane]; };
    const double f = ld(L::F);
    WfPoint<MODEL, NT> wp;
    wp.A.v = ld(L::AV);
    wp.Phi.v = wp.tau.v = wp.tau.d[0] = wp.tau.d[1] = 0.0;
#pragma unroll
    for (int k = 0; k < NT; ++k) wp.A.d[k] = wp.Phi.d[k] = 0.0;
    const double tc = ev[EV_TC];
    int cur_site = -1, cur_psd = -1;
    double sw = 0.0;
    Geo g;
    double Tv[4], Wuu = 0.0, Wvv = 0.0, Wr = 0.0, Wi = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) Tv[a] = 0.0;
    auto do_arm = [&](const ArmK& arm, int s) {
      const SiteK& S = net.site[s];
      const double* sc = rec + REC_SITE + s * SC_COUNT;
      if (s != cur_site) {
        if (S.psd != cur_psd) {
          sw = S.psd == 0 ? ld(L::SW0)
                          : (S.psd == 1 ? ld(L::SW1) : sqrt(ld(L::WT) / psd_at(net.psd[S.psd], f, ld(L::LNF))));
          cur_psd = S.psd;
        }
        site_geo<true>(S, sc, ev, site_tn(S.motion, tc, ld(L::TV)), ld(L::CD), ld(L::SD), g);
        cur_site = s;
      }
      double re[P], im[P], Ap, Ac, v[2];
      arm_derivs<MODEL, NT, P>(wp, ev, sc, g, arm, S.motion, f, re, im, &Ap, &Ac, sw * arm.swgt, v);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int r = SK[a];
#pragma unroll
        for (int b = a; b < 4; ++b) {
          double& x = G[NFF + 12 + pk<4>(a, b)];
          x = fma(im[r], im[SK[b]], fma(re[r], re[SK[b]], x));
        }
        Tr[a] = fma(Ac, im[r], fma(Ap, re[r], Tr[a]));
        Ti[a] = fma(-Ap, im[r], fma(Ac, re[r], Ti[a]));
        Tv[a] = fma(v[1], im[r], fma(v[0], re[r], Tv[a]));
      }
      Wuu = fma(Ac, Ac, fma(Ap, Ap, Wuu));
      Wvv = fma(v[1], v[1], fma(v[0], v[0], Wvv));
      Wr = fma(Ac, v[1], fma(Ap, v[0], Wr));
      Wi = fma(-Ap, v[1], fma(Ac, v[0], Wi));
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
    // c_j, s_j; the producer wrote A'/A (tangent of row r >= 9 in slot
    // r - 7) and the Mc, eta t_noloc tangents (pc_produce, factored)
    double sj[3];
    cr[0] = ld(L::AD);
    ci[0] = -ld(L::PD);
    sj[0] = ld(L::TD);
    cr[1] = ld(L::AD + 1);
    ci[1] = -ld(L::PD + 1);
    sj[1] = ld(L::TD + 1);
    ci[2] = 2. * PI * f;
    sj[2] = INV_DAY;
    cr[3] = -ev[EV_IDL];
    ci[4] = -1.0;
#pragma unroll
    for (int j = 5; j < NF; ++j) {
      cr[j] = ld(L::AD + j - 3);
      ci[j] = -ld(L::PD + j - 3);
    }
    // s_j Tv_k of the rows with s_j != 0
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int a = 0; a < 4; ++a) G[NFF + 4 * j + a] = fma(sj[j], Tv[a], G[NFF + 4 * j + a]);
    // F_jl = y_j^T M y_l, M = [[Wuu, 0, Wr], [0, Wuu, -Wi], [Wr, -Wi, Wvv]]
    double m1[NF], m2[NF], m3[NF];
#pragma unroll
    for (int l = 0; l < NF; ++l) {
      const double s_l = semi_hs(l) ? sj[l < 3 ? l : 0] : 0.0;
      m1[l] = semi_hr(l) ? (semi_hs(l) ? fma(Wr, s_l, Wuu * cr[l]) : Wuu * cr[l]) : (semi_hs(l) ? Wr * s_l : 0.0);
      m2[l] = semi_hi(l) ? (semi_hs(l) ? fma(-Wi, s_l, Wuu * ci[l]) : Wuu * ci[l]) : (semi_hs(l) ? -Wi * s_l : 0.0);
      double t = 0.0;
      bool any = false;
      if (semi_hr(l)) {
        t = Wr * cr[l];
        any = true;
      }
      if (semi_hi(l)) {
        t = any ? fma(-Wi, ci[l], t) : -Wi * ci[l];
        any = true;
      }
      if (semi_hs(l)) t = any ? fma(Wvv, s_l, t) : Wvv * s_l;
      m3[l] = t;
    }
    int q = 0;
#pragma unroll
    for (int j = 0; j < NF; ++j) {
#pragma unroll
      for (int l = j; l < NF; ++l, ++q) {
        double t = G[q];
        if (semi_hr(j) && (semi_hr(l) || semi_hs(l))) t = fma(cr[j], m1[l], t);
        if (semi_hi(j) && (semi_hi(l) || semi_hs(l))) t = fma(ci[j], m2[l], t);
        if (semi_hs(j)) t = fma(sj[j < 3 ? j : 0], m3[l], t);
        G[q] = t;
      }
    }
  }
  // sum_points (Re c_j Tr_k - Im c_j Ti_k) on the matrix cores
  const int rd = (lane >> 4) * PQ_XS;
#pragma unroll
  for (int part = 0; part < 2; ++part) {
#pragma unroll
    for (int j = 0; j < NF; ++j) X[lane * PQ_XS + j] = part ? -ci[j] : cr[j];
#pragma unroll
    for (int a = 0; a < 4; ++a) X[lane * PQ_XS + 10 + a] = part ? Ti[a] : Tr[a];
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const double a = X[4 * t * PQ_XS + rd + (lane & 15)];
      const double b = X[4 * t * PQ_XS + rd + 10 + (lane & 3)];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Stores the Fisher of pc_consume_semi_vx.
template <int P>
__device__ __forceinline__ void store_semi_vx(const double* G, const v4d& acc, double* __restrict__ F, int64_t n_ev,
                                              int64_t e, int lane) {
  constexpr int NF = P - 4, NFF = NF * (NF + 1) / 2;
  constexpr int SK[4] = {3, 4, 5, 6};
  auto put = [&](int a, int b, double v) {
    F[(int64_t)(a * P + b) * n_ev + e] = v;
    if (a != b) F[(int64_t)(b * P + a) * n_ev + e] = v;
  };
  int q = 0;
#pragma unroll
  for (int j = 0; j < NF; ++j)
#pragma unroll
    for (int l = j; l < NF; ++l, ++q) {
      const double s = wave_sum(G[q]);
      if (lane == (q & 63)) put(semi_row<P>(j), semi_row<P>(l), s);
    }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = a; b < 4; ++b) {
      const double s = wave_sum(G[NFF + 12 + pk<4>(a, b)]);
      if (lane == ((NFF + pk<4>(a, b)) & 63)) put(SK[a], SK[b], s);
    }
  // sky x factored: the matrix-core part (element (j, k) in lane k + 16 (j % 4),
  // item j / 4) plus, for the rows with s_j != 0, the wave sum of s_j Tv_k
  double st[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) st[i] = wave_sum(G[NFF + i]);
  const int k = lane & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int j = (lane >> 4) + 4 * r;
    if (j < NF && k < 4) {
      double v = acc[r];
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
          if (j == jj && k == kk) v += st[4 * jj + kk];
      put(semi_row<P>(j), SK[k], v);
    }
  }
}

// Stores the event's Fisher from the NACC matrix-core accumulators (element
// (i, j) in lane j + 16 (i % 4), item i / 4): the upper-triangle element
// (i <= j) goes to both F[i][j] and F[j][i].
template <int P, int NA = PQ_NACC>
__device__ __forceinline__ void store_mx(const v4d* acc, double* __restrict__ F, int64_t n_ev, int64_t e, int lane) {
  const int j = lane & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = (lane >> 4) + 4 * r;
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < NA; ++q) v += acc[q][r];
    if (i < P && j < P && i <= j) {
      F[(int64_t)(i * P + j) * n_ev + e] = v;
      if (i != j) F[(int64_t)(j * P + i) * n_ev + e] = v;
    }
  }
}

// Models with at least GWF_PQ_MFMA_MIN_P parameters take the matrix-core Gram
// (their VALU Gram does not fit the consumer's registers: 79 VGPRs spilled at
// 13 parameters).  Measured per 10^4 events x 2000 points on the
// 5-interferometer network: IMRPhenomD_NRTidalv2 3.54 ms (k_fisher 4.22),
// TaylorF2 tidal 3.19 ms (3.36), eccentric TaylorF2 3.24 ms (VALU pairs 3.39);
// with 11 parameters the 16 x 16 tile wastes half the matrix-core work and the
// VALU pairs are faster (2.37 against 3.04 ms).
#ifndef GWF_PQ_MFMA_MIN_P
#define GWF_PQ_MFMA_MIN_P 12
#endif

// Consumer side of a pair: waits for each item in turn, runs item(e, t,
// res, chunk) on it, releases its buffer and, after the last chunk of event
// e, runs done(e); returns at the end marker or when a wait expires.
template <int NB, int CH, class Item, class Done>
__device__ __forceinline__ void pq_consumer(int* produced, int* consumed, int (*hd)[4], double (*bf)[CH],
                                            unsigned int* counter, unsigned limit, int lane, Item&& item,
                                            Done&& done) {
  for (int k = 0;; ++k) {
    const int s = k % NB;
    if (!pq_wait(produced, k + 1, limit)) {
      pq_fail(counter, lane);
      return;
    }
    const int e = __builtin_amdgcn_readfirstlane(hd[s][0]);
    if (e < 0) return;
    const int t = __builtin_amdgcn_readfirstlane(hd[s][1]);
    const int res = __builtin_amdgcn_readfirstlane(hd[s][2]);
    item(e, t, res, bf[s]);
    pq_release(consumed, k + 1);
    if ((t + 1) * 64 >= res) done(e);
  }
}

template <int MODEL, int NT, int P, bool LIN, bool RAGGED, int MODE>
__global__ __launch_bounds__(128 * PC_EPB, 4 / PC_EPB) void k_fisher_pq(gwf_waveform w, NetK net,
                                                                      const double* __restrict__ rec, int64_t n_ev,
                                                                      const int* __restrict__ res_ev,
                                                                      const uint32_t* __restrict__ arm_mask,
                                                                      unsigned int* counter, double* __restrict__ F) {
  typedef PcLayout<NT> L;
  constexpr bool MASKED = MODE == PQ_MASKED, ARMS = MODE == PQ_ARMS;
  constexpr bool MX = P >= GWF_PQ_MFMA_MIN_P || ARMS;  // Gram on the matrix cores (pc_consume_mx)
#ifdef GWF_NO_PURE_MX
  constexpr bool PURE_MX = false;
#else
  constexpr bool PURE_MX = true;  // 12-14 parameters: pure rows factored out (pc_consume_pure_mx)
#endif
  // The producer reads its event's waveform slots and constants from an LDS
  // copy, except for eccentric TaylorF2: its many phase coefficients, held in
  // VGPRs when read from LDS (SGPRs through the scalar cache), spill (95 and
  // 180 VGPRs at 12 and 14 parameters; 2.31 -> 2.36 and 2.44 -> 5.07 ms per
  // 10^4 events)
  constexpr bool REC_LDS = !(MODEL == GWF_MODEL_TAYLORF2 && (P == 12 || P == 14));
  // 12-13 parameters, network sum: pc_consume_semi_vx (pc_consume_pure_mx for
  // a network that mixes rotating- and fixed-Earth sites, and for 14)
#ifdef GWF_NO_SEMI_ROWS
  constexpr bool SEMI_VX = false;
#else
  constexpr bool SEMI_VX = MX && MODE == PQ_SUM && PURE_MX && P <= 13;
#endif
  // and with the per-event arm masks (duty cycle)
  constexpr bool SEMI_VX_M = MX && MASKED && P <= 13 && !ARMS
#ifdef GWF_NO_SEMI_ROWS
                             && false
#endif
      ;
  // ring depth: PQ_NBUF buffers, fewer where they would not fit the CU's LDS
  // beside the MFMA tiles
  constexpr int NB = pq_ring_depth(L::NV, MX, REC_LDS);
  __shared__ double buf[PC_EPB][NB][L::NV * 64];
  __shared__ int head[PC_EPB][NB][4];  // event, chunk, grid length of the item in each buffer
  __shared__ int cnt[PC_EPB][2];            // items produced, items consumed
  __shared__ double xt[PC_EPB][MX ? 64 * PQ_XS : 1];  // consumers' transposition tiles
  // each producer's copy of its event's waveform slots and event constants
  // (3.9 KB): read per chunk from LDS instead of through the scalar cache,
  // which the four pairs' records (5.6 KB each, both waves reading) thrash
  __shared__ double recl[PC_EPB][REC_LDS ? REC_SITE : 1];
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int slot = wave % PC_EPB;
  if (threadIdx.x < 2 * PC_EPB) cnt[threadIdx.x >> 1][threadIdx.x & 1] = 0;
  __syncthreads();
  // the consumer is pc_consume_semi (it reads the producer's rows factored)
#if defined(GWF_NO_PURE_ROWS) || defined(GWF_NO_SEMI_ROWS)
  const bool semi = false;
#else
  const bool semi = ((P == 11 && !MX) || SEMI_VX || SEMI_VX_M) && !net.mixed_motion;
#endif
  if (wave < PC_EPB) {
    int k = 0;  // items produced by this pair
    int64_t e = RAGGED ? 0 : (int64_t)blockIdx.x * PC_EPB + slot;
    const int64_t stride = (int64_t)gridDim.x * PC_EPB;
    for (;;) {
      if (RAGGED) {
        unsigned int t0 = 0;
        if (lane == 0) t0 = atomicAdd(counter, 1u);
        e = (int64_t)__builtin_amdgcn_readfirstlane((int)__shfl(t0, 0, 64));
      }
      const int s0 = k % NB;
      if (e >= n_ev) {  // end marker
        if (!pq_wait(&cnt[slot][1], k - NB + 1, net.spin_limit)) {
          pq_fail(counter, lane);
          return;
        }
        if (lane == 0) head[slot][s0][0] = -1;
        pq_release(&cnt[slot][0], k + 1);
        return;
      }
      int res = net.res;
      if (RAGGED) res = min(max(__builtin_amdgcn_readfirstlane(res_ev[e]), 2), 1 << 30);  // chunk counts stay in int
      const int nch = (res + 63) / 64;
      const double* rg = rec + e * (int64_t)REC_DOUBLES;
      const double* r = rg;
      if constexpr (REC_LDS) {
        double* rl = recl[slot];  // the previous event's chunks are all in the ring: free to overwrite
        for (int q = lane; q < REC_SITE; q += 64) rl[q] = rg[q];
        r = rl;
      }
      double a, b;
      grid_ends(LIN, net.fmin, fcut_eff(r[REC_EV + EV_FCUT], net.fmax), &a, &b);
      GridWalk<LIN> gw;
      gw.init(a, b, res, lane, 64);
      for (int t = 0; t < nch; ++t, ++k) {
        const int s = k % NB;
        if (!pq_wait(&cnt[slot][1], k - NB + 1, net.spin_limit)) {  // the buffer's previous item is consumed
          pq_fail(counter, lane);
          return;
        }
        pc_produce<MODEL, NT, LIN>(w, net, r, gw, t * 64, res, lane, buf[slot][s], semi);
        if (lane == 0) {
          head[slot][s][0] = (int)e;
          head[slot][s][1] = t;
          head[slot][s][2] = res;
        }
        pq_release(&cnt[slot][0], k + 1);
      }
      if (!RAGGED) e += stride;
    }
  } else if constexpr (MX && MODE == PQ_SUM && PURE_MX) {  // rows factored out, see pc_consume_semi_vx
    double* X = xt[slot];
    for (int q = lane; q < 64 * PQ_XS; q += 64) X[q] = 0.0;
#ifndef GWF_NO_SEMI_ROWS
    // the Mc, eta, tcoal rows factored as well, up to 13 parameters (at 14
    // the consumer spills 88 VGPRs: 2.44 -> 3.07 ms per 10^4 events)
    if (SEMI_VX && !net.mixed_motion) {
      constexpr int NF = P - 4, NG = NF * (NF + 1) / 2 + 12 + 10;
      double G[NG];
      v4d acc;
      pq_consumer<NB>(&cnt[slot][0], &cnt[slot][1], head[slot], buf[slot], counter, net.spin_limit, lane,
                      [&](int e, int t, int res, const double* chunk) {
                        if (t == 0) {
#pragma unroll
                          for (int q = 0; q < NG; ++q) G[q] = 0.0;
                          acc = v4d{0.0, 0.0, 0.0, 0.0};
                        }
                        pc_consume_semi_vx<MODEL, NT, P>(net, rec + (int64_t)e * REC_DOUBLES, t * 64, res, lane,
                                                         chunk, X, G, acc);
                      },
                      [&](int e) { store_semi_vx<P>(G, acc, F, n_ev, e, lane); });
      return;
    }
#endif
    constexpr int NPU = P - 7, NG = 28 + NPU * (NPU + 1) / 2;
    double G[NG];
    v4d acc;
    pq_consumer<NB>(&cnt[slot][0], &cnt[slot][1], head[slot], buf[slot], counter, net.spin_limit, lane,
                    [&](int e, int t, int res, const double* chunk) {
                      if (t == 0) {
#pragma unroll
                        for (int q = 0; q < NG; ++q) G[q] = 0.0;
                        acc = v4d{0.0, 0.0, 0.0, 0.0};
                      }
                      pc_consume_pure_mx<MODEL, NT, P>(net, rec + (int64_t)e * REC_DOUBLES, t * 64, res, lane, chunk,
                                                       X, G, acc);
                    },
                    [&](int e) { store_pure_mx<P>(G, acc, F, n_ev, e, lane); });
  } else if constexpr (MX) {  // Gram on the matrix cores
    double* X = xt[slot];
    for (int q = lane; q < 64 * PQ_XS; q += 64) X[q] = 0.0;  // padding columns stay zero
    if constexpr (SEMI_VX_M) {  // masked network sum, rows factored as in the unmasked one
      if (!net.mixed_motion) {
        constexpr int NF = P - 4, NG = NF * (NF + 1) / 2 + 12 + 10;
        double G[NG];
        v4d acc1;
        uint32_t mask = 0u;
        pq_consumer<NB>(&cnt[slot][0], &cnt[slot][1], head[slot], buf[slot], counter, net.spin_limit, lane,
                        [&](int e, int t, int res, const double* chunk) {
                          if (t == 0) {
#pragma unroll
                            for (int q = 0; q < NG; ++q) G[q] = 0.0;
                            acc1 = v4d{0.0, 0.0, 0.0, 0.0};
                            mask = (uint32_t)__builtin_amdgcn_readfirstlane((int)arm_mask[e]);
                          }
                          pc_consume_semi_vx<MODEL, NT, P, true>(net, rec + (int64_t)e * REC_DOUBLES, t * 64, res,
                                                                 lane, chunk, X, G, acc1, mask);
                        },
                        [&](int e) { store_semi_vx<P>(G, acc1, F, n_ev, e, lane); });
        return;
      }
    }
    constexpr int NA = ARMS ? PQ_ARMS_MAX : PQ_NACC;        // ARMS: one accumulator per physical arm
    v4d acc[NA];
    uint32_t mask = 0u;
    pq_consumer<NB>(&cnt[slot][0], &cnt[slot][1], head[slot], buf[slot], counter, net.spin_limit, lane,
                    [&](int e, int t, int res, const double* chunk) {
                      if (t == 0) {
#pragma unroll
                        for (int q = 0; q < NA; ++q) acc[q] = v4d{0.0, 0.0, 0.0, 0.0};
                        if (MASKED) mask = (uint32_t)__builtin_amdgcn_readfirstlane((int)arm_mask[e]);
                      }
                      pc_consume_mx<MODEL, NT, P, MODE>(net, rec + (int64_t)e * REC_DOUBLES, t * 64, res, lane, chunk,
                                                        X, acc, mask);
                    },
                    [&](int e) {
                      if constexpr (ARMS) {  // F: (n_arms, P, P, n_ev)
#pragma unroll
                        for (int a = 0; a < PQ_ARMS_MAX; ++a)
                          if (a < net.n_arms) store_mx<P, 1>(acc + a, F + (int64_t)a * P * P * n_ev, n_ev, e, lane);
                      } else {
                        store_mx<P>(acc, F, n_ev, e, lane);
                      }
                    });
  } else {  // Gram on the VALU, packed upper triangle per lane
    constexpr int NG = P * (P + 1) / 2;
    double G[NG];
    uint32_t mask = 0u;
    // KIND 0: no row factored, 1: the pure rows, 2: the pure rows and the
    // Mc, eta, tcoal rows (pc_consume_semi).  Chosen once per wave, so that
    // each consumer loop gets its own register allocation.
    auto consume = [&](auto kind) {
      constexpr int KIND = decltype(kind)::value;
      pq_consumer<NB>(&cnt[slot][0], &cnt[slot][1], head[slot], buf[slot], counter, net.spin_limit, lane,
                      [&](int e, int t, int res, const double* chunk) {
                        if (t == 0) {
#pragma unroll
                          for (int q = 0; q < NG; ++q) G[q] = 0.0;
                          if (MASKED) mask = (uint32_t)__builtin_amdgcn_readfirstlane((int)arm_mask[e]);
                        }
                        const double* r = rec + (int64_t)e * REC_DOUBLES;
                        if constexpr (KIND == 2)
                          pc_consume_semi<MODEL, NT, MASKED>(net, r, t * 64, res, lane, chunk, G, mask);
                        else if constexpr (KIND == 1)
                          pc_consume_pure<MODEL, NT, MASKED>(net, r, t * 64, res, lane, chunk, G, mask);
                        else
                          pc_consume<MODEL, NT, P, MASKED>(net, r, t * 64, res, lane, chunk, G, mask);
                      },
                      [&](int e) { store_packed<P>(G, F, n_ev, e, lane); });
    };
#ifdef GWF_NO_PURE_ROWS
    consume(std::integral_constant<int, 0>{});
#else
    if constexpr (P == 11) {
#ifndef GWF_NO_SEMI_ROWS
      if (!net.mixed_motion)
        consume(std::integral_constant<int, 2>{});
      else
#endif
        consume(std::integral_constant<int, 1>{});
    } else {
      consume(std::integral_constant<int, 0>{});
    }
#endif
  }
}

// After k_fisher_pq: NaN in every output element when a wave's wait ran out.
__global__ void k_pq_poison(const unsigned int* __restrict__ counter, double* __restrict__ F, int64_t n) {
  if (counter[1] == 0u) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    F[i] = __builtin_nan("");
}

// Network Fisher with per-event arm masks (duty cycle): bit a of arm_mask[e]
// switches physical arm a of the network on for event e.
template <int MODEL, int NT, int P, bool LIN>
__global__ __launch_bounds__(256, GWF_FISHER_MIN_BLOCKS) void k_fisher_masked(gwf_waveform w, NetK net,
                                                                             const double* __restrict__ rec,
                                                                             int64_t n_ev,
                                                                             const uint32_t* __restrict__ arm_mask,
                                                                             double* __restrict__ F) {
  const int64_t e = wave_event();
  if (e >= n_ev) return;  // wave-uniform exit
  const int lane = threadIdx.x & 63;
  const uint32_t mask = (uint32_t)__builtin_amdgcn_readfirstlane((int)arm_mask[e]);
  constexpr int NG = P * (P + 1) / 2;
  double G[NG];
#pragma unroll
  for (int k = 0; k < NG; ++k) G[k] = 0.0;
  fisher_event<MODEL, NT, P, true, LIN>(w, net, rec + e * (int64_t)REC_DOUBLES, net.res, lane, 64, -1, G, mask);
  store_packed<P>(G, F, n_ev, e, lane);
}

// Ragged grids: one persistent launch; each wavefront takes events from an
// atomic counter (lane 0, vector-memory atomic) until the batch is exhausted,
// so every wave reaches the exit; res_ev[e] is event e's grid length.
template <int MODEL, int NT, int P, bool LIN>
__global__ __launch_bounds__(256) void k_fisher_ragged(gwf_waveform w, NetK net, const double* __restrict__ rec,
                                                       int64_t n_ev, const int* __restrict__ res_ev, int arm_sel,
                                                       unsigned int* counter, double* __restrict__ F) {
  const int lane = threadIdx.x & 63;
  constexpr int NG = P * (P + 1) / 2;
  for (;;) {
    unsigned int t = 0;
    if (lane == 0) t = atomicAdd(counter, 1u);
    const int64_t e = (int64_t)__builtin_amdgcn_readfirstlane((int)__shfl(t, 0, 64));
    if (e >= n_ev) break;
    double G[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) G[k] = 0.0;
    const int res = min(max(res_ev[e], 2), 1 << 30);
    fisher_event<MODEL, NT, P, false, LIN>(w, net, rec + e * (int64_t)REC_DOUBLES, res, lane, 64, arm_sel, G);
    store_packed<P>(G, F, n_ev, e, lane);
  }
}

__global__ void k_zero_u32(unsigned int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// 3 waves per SIMD (<= 168 VGPRs): measured fastest for the 5-arm network
template <int MODEL, int MAXA>
__global__ __launch_bounds__(256, 3) void k_snr(gwf_waveform w, NetK net, const double* __restrict__ rec, int64_t n_ev,
                                             const int* __restrict__ res_ev, double* __restrict__ out) {
  const int64_t e = wave_event();
  if (e >= n_ev) return;
  const int lane = threadIdx.x & 63;
  // the event's record copied to LDS and read there per chunk: the scalar
  // cache cannot hold the records of the 12 events resident per CU (45.2 ->
  // 37.5 ms at 10^6 events x 2000 f)
  __shared__ double recs[WAVES_PER_BLOCK][REC_DOUBLES];
  double* rl = recs[threadIdx.x >> 6];
  const double* rg = rec + e * (int64_t)REC_DOUBLES;
  for (int q = lane; q < REC_DOUBLES; q += 64) rl[q] = rg[q];
  const double* r = rl;
  double acc[MAXA];
#pragma unroll
  for (int k = 0; k < MAXA; ++k) acc[k] = 0.0;
  snr_event<MODEL, MAXA>(w, net, r, res_ev ? min(max(res_ev[e], 2), 1 << 30) : net.res, lane, 64, acc);
#pragma unroll
  for (int k = 0; k < MAXA; ++k) {
    if (k < net.n_arms) {
      const double v = wave_sum(acc[k]);
      if (lane == 0) out[(int64_t)k * n_ev + e] = v;
    }
  }
}

template <int MODEL, int NT, int P>
__global__ __launch_bounds__(256) void k_strain(gwf_waveform w, NetK net, const double* __restrict__ rec, int64_t n_ev,
                                                int armi, const double* __restrict__ f, int64_t nf,
                                                double* __restrict__ h, double* __restrict__ J,
                                                double* __restrict__ comp) {
  const int64_t e = wave_event();
  if (e >= n_ev) return;
  const int lane = threadIdx.x & 63;
  strain_event<MODEL, NT, P>(w, net, rec + e * (int64_t)REC_DOUBLES, armi, f + e, n_ev, nf, lane, 64,
                             h ? h + 2 * e : nullptr, 2 * n_ev, J ? J + 2 * e * nf : nullptr, 2 * n_ev * nf,
                             comp ? comp + e : nullptr, n_ev, nf * n_ev);
}

__global__ void k_grid(NetK net, const double* __restrict__ rec, int64_t n_ev, int res, double* __restrict__ fo,
                       double* __restrict__ So) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)res * n_ev) return;
  const int64_t i = idx / n_ev, e = idx % n_ev;
  const double* ev = rec + e * (int64_t)REC_DOUBLES + REC_EV;
  double a, b;
  grid_ends(net.lin != 0, net.fmin, fcut_eff(ev[EV_FCUT], net.fmax), &a, &b);
  const double lin = grid_lin(a, b, (int)i, res - 1);
  const double f = net.lin ? lin : grid_pow10(lin, (int)i, res - 1);
  fo[idx] = f;
  if (So) So[idx] = psd_at(net.psd[0], f, net.lin ? log(f) : lin * 2.302585092994045684);
}

// Grid fmin..fcut[e] (geomspace, or linspace when net.lin) and the PSD on it.
__global__ void k_grid_fcut(NetK net, const double* __restrict__ fcut, int64_t n_ev, int res, double* __restrict__ fo,
                            double* __restrict__ So) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)res * n_ev) return;
  const int64_t i = idx / n_ev, e = idx % n_ev;
  double a, b;
  grid_ends(net.lin != 0, net.fmin, fcut[e], &a, &b);
  const double lin = grid_lin(a, b, (int)i, res - 1);
  const double f = net.lin ? lin : grid_pow10(lin, (int)i, res - 1);
  fo[idx] = f;
  if (So) So[idx] = psd_at(net.psd[0], f, net.lin ? log(f) : lin * 2.302585092994045684);
}

// Overlap integrals of two strains on one grid, one wave per event:
// out[0] += 4 trapz(Re(h1 conj h2)/S), out[1] += 4 trapz(|h1|^2/S),
// out[2] += 4 trapz(|h2|^2/S) (signal.py:1870-1920).  h: (nf, n_ev) complex.
__global__ __launch_bounds__(256) void k_overlap(const double* __restrict__ h1, const double* __restrict__ h2,
                                                 const double* __restrict__ f, const double* __restrict__ S,
                                                 int64_t n_ev, int64_t nf, int accumulate, double* __restrict__ out) {
  const int64_t e = wave_event();
  if (e >= n_ev) return;
  const int lane = threadIdx.x & 63;
  double o12 = 0.0, o11 = 0.0, o22 = 0.0;
  for (int64_t i = lane; i < nf; i += 64) {
    const double fi = f[i * n_ev + e];
    const double fl = i > 0 ? f[(i - 1) * n_ev + e] : fi;
    const double fr = i < nf - 1 ? f[(i + 1) * n_ev + e] : fi;
    const double w = 4.0 * 0.5 * (fr - fl) / S[i * n_ev + e];
    const double ar = h1[2 * (i * n_ev + e)], ai = h1[2 * (i * n_ev + e) + 1];
    const double br = h2[2 * (i * n_ev + e)], bi = h2[2 * (i * n_ev + e) + 1];
    o12 += w * (ar * br + ai * bi);
    o11 += w * (ar * ar + ai * ai);
    o22 += w * (br * br + bi * bi);
  }
  o12 = wave_sum(o12);
  o11 = wave_sum(o11);
  o22 = wave_sum(o22);
  if (lane == 0) {
    out[e] = accumulate ? out[e] + o12 : o12;
    out[n_ev + e] = accumulate ? out[n_ev + e] + o11 : o11;
    out[2 * n_ev + e] = accumulate ? out[2 * n_ev + e] + o22 : o22;
  }
}

template <int MODEL, int NT>
__global__ void k_wfeval(gwf_waveform w, const double* __restrict__ rec, int64_t n_ev, const double* __restrict__ f,
                         int64_t nf, double* __restrict__ Ao, double* __restrict__ Po, double* __restrict__ To) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nf * n_ev) return;
  const int64_t e = idx % n_ev;
  const double* r = rec + e * (int64_t)REC_DOUBLES;
  WfPoint<MODEL, NT> wp;
  wf_point<MODEL, NT>(w, r, f[idx], To != nullptr, wp);
  if (Ao) Ao[idx] = wp.A.v * r[REC_EV + EV_IDL];
  if (Po) Po[idx] = wp.Phi.v;
  if (To) To[idx] = wp.tau.v;
}

__global__ void k_fcut(gwf_waveform w, const double* __restrict__ par, int64_t ld, int64_t n_ev, double fmax,
                       double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_ev) return;
  const double Mc = par[e], eta = par[ld + e];
  double fc;
  if (w.model == GWF_MODEL_IMRPHENOMD)
    fc = w.fcut_par / (Mc * GMSUN_C3 / pow(eta, 3. / 5.));
  else if (w.model == GWF_MODEL_IMRPHENOMD_NRTIDALV2)
    fc = nrt_fcut(Mc, eta, par[11 * ld + e], par[12 * ld + e]);
  else
    fc = tf2_fcut(w, Mc, eta, par[9 * ld + e], par[10 * ld + e]);
  out[e] = fcut_eff(fc, fmax);
}

// Plain Gram over materialised derivatives (one wave per event).
template <int P>
__global__ __launch_bounds__(256) void k_gram(const double* __restrict__ J, const double* __restrict__ f,
                                              const double* __restrict__ S, int64_t n_ev, int64_t nf, int accumulate,
                                              double* __restrict__ F) {
  const int64_t e = wave_event();
  if (e >= n_ev) return;
  const int lane = threadIdx.x & 63;
  constexpr int NG = P * (P + 1) / 2;
  double G[NG];
#pragma unroll
  for (int k = 0; k < NG; ++k) G[k] = 0.0;
  const int64_t js = 2 * n_ev * nf;
  const double* Je = J + 2 * e * nf;
  for (int64_t i = lane; i < nf; i += 64) {
    const double fi = f[i * n_ev + e];
    const double fl = i > 0 ? f[(i - 1) * n_ev + e] : fi;
    const double fr = i < nf - 1 ? f[(i + 1) * n_ev + e] : fi;
    const double wS = 4.0 * 0.5 * (fr - fl) / S[i * n_ev + e];
    double re[P], im[P];
#pragma unroll
    for (int X = 0; X < P; ++X) {
      re[X] = Je[X * js + 2 * i];
      im[X] = Je[X * js + 2 * i + 1];
    }
    gram_update<P>(G, re, im, wS);
  }
  int k = 0;
#pragma unroll
  for (int a = 0; a < P; ++a) {
#pragma unroll
    for (int b = a; b < P; ++b) {
      const double s = wave_sum(G[k]);
      if (lane == (k & 63)) {
        const int64_t i1 = (int64_t)(a * P + b) * n_ev + e, i2 = (int64_t)(b * P + a) * n_ev + e;
        F[i1] = accumulate ? F[i1] + s : s;
        if (a != b) F[i2] = accumulate ? F[i2] + s : s;
      }
      ++k;
    }
  }
}

// The same Fisher integral on FP32 MFMA (v_mfma_f32_16x16x4_f32) with split
// operands and FP64 reduction.  Each weighted derivative x = J sqrt(w/S)
// (FP64) is split into hi = fp32(x) and lo = fp32(x - hi), so x = hi + lo to
// ~2^-48; F = hi^T hi + hi^T lo + (hi^T lo)^T (the lo^T lo term, ~2^-48
// relative, is dropped).  One wave per event; per step the wave takes 4
// frequencies, lane l holding parameter l%16 of frequency l/16, which is the
// A (row l%16, k l/16) and B (k l/16, column l%16) operand layout of the
// 16x16x4 MFMA, so A = B^T comes from the same register.  The real and
// imaginary parts are two k-steps.  The fp32 accumulators (8 samples each)
// are added into FP64 accumulators after every step, so the fp32 rounding
// is bounded by 8 * 2^-24 of sum_s |x_a x_b| <= sqrt(F_aa F_bb).
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

template <int P>
__global__ __launch_bounds__(256) void k_gram_split(const double* __restrict__ J, const double* __restrict__ f,
                                                    const double* __restrict__ S, int64_t n_ev, int64_t nf,
                                                    int accumulate, double* __restrict__ F) {
  __shared__ double red[WAVES_PER_BLOCK][2][16 * 16];
  const int64_t e = wave_event();
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const bool live = e < n_ev;  // whole wave uniform
  const int a = lane & 15, q = lane >> 4;
  double hh[4] = {0., 0., 0., 0.}, hl[4] = {0., 0., 0., 0.};
  if (live) {
    const double* Ja = J + 2 * ((int64_t)(a < P ? a : 0) * n_ev + e) * nf;
    for (int64_t i0 = 0; i0 < nf; i0 += 4) {
      const int64_t i = i0 + q;
      float hr = 0.f, lr = 0.f, hi_ = 0.f, li = 0.f;
      if (a < P && i < nf) {
        const double fi = f[i * n_ev + e];
        const double fl = i > 0 ? f[(i - 1) * n_ev + e] : fi;
        const double fr = i < nf - 1 ? f[(i + 1) * n_ev + e] : fi;
        const double sw = sqrt(4.0 * 0.5 * (fr - fl) / S[i * n_ev + e]);
        const double xr = Ja[2 * i] * sw, xi = Ja[2 * i + 1] * sw;
        hr = (float)xr;
        lr = (float)(xr - (double)hr);
        hi_ = (float)xi;
        li = (float)(xi - (double)hi_);
      }
      f32x4_t c_hh = {0.f, 0.f, 0.f, 0.f}, c_hl = {0.f, 0.f, 0.f, 0.f};
      c_hh = __builtin_amdgcn_mfma_f32_16x16x4f32(hr, hr, c_hh, 0, 0, 0);
      c_hl = __builtin_amdgcn_mfma_f32_16x16x4f32(hr, lr, c_hl, 0, 0, 0);
      c_hh = __builtin_amdgcn_mfma_f32_16x16x4f32(hi_, hi_, c_hh, 0, 0, 0);
      c_hl = __builtin_amdgcn_mfma_f32_16x16x4f32(hi_, li, c_hl, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        hh[r] += (double)c_hh[r];
        hl[r] += (double)c_hl[r];
      }
    }
    // lane l holds rows 4(l/16)..+3 of column l%16
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      red[wv][0][(4 * q + r) * 16 + a] = hh[r];
      red[wv][1][(4 * q + r) * 16 + a] = hl[r];
    }
  }
  __syncthreads();
  if (!live) return;
  constexpr int NG = P * (P + 1) / 2;
  for (int k = lane; k < NG; k += 64) {
    // k -> (r, c), r <= c, row-major over the upper triangle
    int r = 0, kk = k;
    while (kk >= P - r) {
      kk -= P - r;
      ++r;
    }
    const int c = r + kk;
    const double s = red[wv][0][r * 16 + c] + red[wv][1][r * 16 + c] + red[wv][1][c * 16 + r];
    const int64_t i1 = (int64_t)(r * P + c) * n_ev + e, i2 = (int64_t)(c * P + r) * n_ev + e;
    F[i1] = accumulate ? F[i1] + s : s;
    if (r != c) F[i2] = accumulate ? F[i2] + s : s;
  }
}

// VJP of the strain: one wave per (parameter, event) pair.
__global__ __launch_bounds__(256) void k_vjp(const double* __restrict__ J, const double* __restrict__ gh, int P,
                                             int64_t n_ev, int64_t nf, double tc_scale, double* __restrict__ g) {
  const int64_t pe = wave_event();
  if (pe >= (int64_t)P * n_ev) return;
  const int lane = threadIdx.x & 63;
  const int64_t p = pe / n_ev, e = pe % n_ev;
  const double* Jr = J + 2 * (p * n_ev + e) * nf;
  double acc = 0.0;
  for (int64_t i = lane; i < nf; i += 64)
    acc += gh[2 * (i * n_ev + e)] * Jr[2 * i] + gh[2 * (i * n_ev + e) + 1] * Jr[2 * i + 1];
  acc = wave_sum(acc);
  if (lane == 0) g[p * n_ev + e] = (p == 7) ? acc * tc_scale : acc;
}

// ------------------------------------------- covariance (one wave per event)
// fisherTools' per-event loops (CovMatr, CheckFisher) run one 64-lane wave
// per event.  The event's P x P matrices live in LDS (rows and columns
// beyond P kept zero); lanes work on matrix entries, rows or columns, and
// the workgroup is that one wave, so __syncthreads() orders the LDS steps
// without tying events together.  No private arrays: nothing goes to
// scratch memory.  PT > 0 fixes P at compile time (index arithmetic by
// constants, LDS sized for P); PT = 0 takes any P <= PMAX at run time.
constexpr int PMAX = 16;
constexpr double MP_TOL = 2.220446049250313e-16;  // 2^-52: mpmath's tolerance at 53 bits

template <int PT>
struct CovDim {
  static constexpr int N = PT > 0 ? PT + (PT & 1) : PMAX;  // kept rows / columns (an odd P adds a zero index)
  static constexpr int LD = N + 1;                         // odd row pitch (LDS banks)
  static constexpr int MAT = N * LD;
};

#define AT(M, i, j) (M)[(i) * LD + (j)]

// Blocks are dealt to the 8 XCDs round-robin (block b on XCD b % 8): this
// mapping gives each XCD one contiguous range of events, so the (P, P, N)
// loads and stores of neighbouring events meet in that XCD's L2.  The grid
// is a multiple of 8 blocks.
__device__ __forceinline__ int64_t xcd_event() {
  const int64_t b = blockIdx.x;
  return (b & 7) * (int64_t)(gridDim.x >> 3) + (b >> 3);
}

inline unsigned xcd_grid(int64_t n_ev) { return (unsigned)(((n_ev + 7) / 8) * 8); }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Round-robin pairing of the indices 0..m (m odd): in round r index m meets
// r and every other index i meets (2r - i) mod m; m rounds pair every index
// with every other one once.
__device__ __forceinline__ int rr_partner(int i, int r, int m) {
  if (i == m) return r;
  if (i == r) return m;
  const int v = 2 * r - i;  // in (-m, 2m)
  return v < 0 ? v + m : (v >= m ? v - m : v);
}

// Jacobi eigen-decomposition of the symmetric P x P matrix in E (destroyed):
// cyclic Jacobi in round-robin order, the P/2 disjoint rotations of a round
// applied together as E <- J^T E J (and V <- V J when WANT_V).  Each rotation
// is the classical one that zeroes E_pq (Numerical Recipes 11.1:
// t = sgn(theta) / (|theta| + sqrt(theta^2 + 1)), theta = (E_qq - E_pp) /
// (2 E_pq), here t = sgn(d) h / (|d| + sqrt(d^2 + h^2)) with d = E_qq - E_pp,
// h = 2 E_pq; E_pp -= t E_pq, E_qq += t E_pq, E_pq = 0), a pair rotated only
// while E_pq^2 > 2^-104 |E_pp E_qq| (the threshold of Demmel & Veselic's
// relatively accurate Jacobi).  A round: the lanes of the indices compute
// their pair's rotation and new diagonal, then one lane per upper
// off-diagonal entry applies the two-sided update (entries of a round's
// pairs never mix diagonal values in).  Sweeps until one rotates nothing
// or the off-diagonal mass is below 1e-36 of the diagonal's (at most 60).
// On return w[0..P) holds the eigenvalues in ascending order (mpmath.eigh's
// order; ties and NaNs keep index order, NaNs last) and, when WANT_V, the
// columns of V the eigenvectors.  E, E2, V, V2: LDS matrices with zero rows
// / columns beyond P (an odd P pairs with the zero index P); the pointers
// are swapped as the rounds alternate between them.  cc, ss: PMAX doubles of
// LDS.
template <int PT, bool WANT_V>
__device__ void wave_jacobi(double*& E, double*& E2, double*& V, double*& V2, int Prt, double* w, double* cc,
                            double* ss, int lane) {
  constexpr int LD = CovDim<PT>::LD;
  const int P = PT > 0 ? PT : Prt;
  if (WANT_V) {
    for (int t = lane; t < P * P; t += 64) {
      const int i = t / P, j = t - (t / P) * P;
      AT(V, i, j) = (i == j) ? 1.0 : 0.0;
    }
  }
  const int m = P + (P & 1) - 1;
  const int n_off = P * (P - 1) / 2;
  // this lane's upper off-diagonal entries k = lane, lane + 64, ... (row-major)
  constexpr int KMAX = (PMAX * (PMAX - 1) / 2 + 63) / 64;
  int oi[KMAX], oj[KMAX];
#pragma unroll
  for (int u = 0; u < KMAX; ++u) {
    int kk = lane + 64 * u, i = 0;
    while (i < P - 1 && kk >= P - 1 - i) {
      kk -= P - 1 - i;
      ++i;
    }
    oi[u] = i;
    oj[u] = i + 1 + kk;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int t = lane; t < P * P; t += 64) {
      const int i = t / P, j = t - (t / P) * P;
      const double a = AT(E, i, j);
      if (i == j) diag += a * a;
      else if (i < j) off += a * a;
    }
    off = wave_sum(off);
    diag = wave_sum(diag);
    if (!(off > 1e-36 * diag)) break;  // the same sums on every lane: uniform
    bool rotated = false;
    for (int r = 0; r < m; ++r) {
      bool rot = false;
      if (lane < P) {
        const int j = rr_partner(lane, r, m);
        double c = 1.0, s = 0.0, dnew = AT(E, lane, lane);
        if (j < P) {
          const int p = min(lane, j), q = max(lane, j);
          const double apq = AT(E, p, q);
          const double app = AT(E, p, p), aqq = AT(E, q, q);
          if (apq * apq > (MP_TOL * MP_TOL) * fabs(app * aqq)) {
            const double d = aqq - app, h = 2.0 * apq;
            const double num = d > 0.0 ? h : (d < 0.0 ? -h : fabs(h));
            const double tt = num / (fabs(d) + sqrt(d * d + h * h));
            c = rsqrt(tt * tt + 1.0);
            s = tt * c;
            dnew = lane == p ? app - tt * apq : aqq + tt * apq;
            rot = true;
          }
          if (lane == p) s = -s;  // J_pp = c, J_qp = -s, J_pq = s, J_qq = c
        }
        cc[lane] = c;
        ss[lane] = s;
        AT(E2, lane, lane) = dnew;
      }
      if (__ballot(rot) == 0) continue;  // nothing to rotate this round (uniform)
      rotated = true;
      __syncthreads();
#pragma unroll
      for (int u = 0; u < KMAX; ++u) {
        if (lane + 64 * u < n_off) {
          const int i = oi[u], j = oj[u];
          const int ip = rr_partner(i, r, m), jp = rr_partner(j, r, m);
          double v = 0.0;
          if (!(jp == i && ss[j] != 0.0)) {  // a rotated pair's own entry becomes 0
            const double ci = cc[i], si = ss[i], cj = cc[j], sj = ss[j];
            v = ci * (cj * AT(E, i, j) + sj * AT(E, i, jp)) + si * (cj * AT(E, ip, j) + sj * AT(E, ip, jp));
          }
          AT(E2, i, j) = v;
          AT(E2, j, i) = v;
        }
      }
      if (WANT_V) {
        for (int t = lane; t < P * P; t += 64) {
          const int i = t / P, j = t - (t / P) * P;
          AT(V2, i, j) = AT(V, i, j) * cc[j] + AT(V, i, rr_partner(j, r, m)) * ss[j];
        }
      }
      __syncthreads();
      double* t1 = E;
      E = E2;
      E2 = t1;
      if (WANT_V) {
        t1 = V;
        V = V2;
        V2 = t1;
      }
    }
    if (!rotated) break;
  }
  if (lane < P) {  // rank of each eigenvalue; the sorted eigenvectors go to V2
    const double wi = AT(E, lane, lane);
    const bool ni = isnan(wi);
    int rk = 0;
    for (int j = 0; j < P; ++j) {
      const double wj = AT(E, j, j);
      const bool nj = isnan(wj);
      const bool before = ni ? (nj ? j < lane : true) : (!nj && (wj < wi || (wj == wi && j < lane)));
      rk += before ? 1 : 0;
    }
    w[rk] = wi;
    if (WANT_V)
      for (int k = 0; k < P; ++k) AT(V2, k, rk) = AT(V, k, lane);
  }
  __syncthreads();
  if (WANT_V) {
    double* t1 = V;
    V = V2;
    V2 = t1;
  }
}

// ------------------------------------------------ covariance, one per lane
// CovMatr with invMethodIn = 'cho' (the reference's default) for P = 11-13,
// one event per lane with its matrices in registers: the upper triangle
// (P(P+1)/2 doubles) of the Jacobi-scaled matrix A, a cyclic Jacobi in
// serial order, unrolled (the same rotations and threshold as wave_jacobi),
// for the eigenvalues and the condition number; then mpmath's Cholesky,
// the inverse of the factor and C = Ci^T Ci, each in place, the scaled-back
// C written out, and eps = max|C F - I| with F read a column at a time.
// The events this path does not finish (a NaN or negative diagonal, a
// negative eigenvalue or a failed Cholesky, which take the alternative
// method) get the marker COV_PENDING as their last covariance element and
// are completed by k_cov<PT, true>, which skips every other event.
constexpr unsigned long long COV_PENDING = 0x7ff4c0de0000cafeull;  // a NaN payload arithmetic does not produce

template <int P>
__device__ __forceinline__ constexpr int upk(int i, int j) {  // packed upper triangle, any order of i, j
  return i <= j ? i * P - i * (i - 1) / 2 + (j - i) : j * P - j * (j - 1) / 2 + (i - j);
}
template <int P>
__device__ __forceinline__ constexpr int lok(int i, int j) {  // packed lower triangle, i >= j
  return i * (i + 1) / 2 + j;
}

// Cyclic Jacobi on one lane's packed upper triangle E (pairs in serial order
// p < q; the rotation formula and threshold of wave_jacobi; sweeps until one
// rotates nothing).  WANT_V: the rotations are also accumulated into V
// (row-major P x P, set to I first), so column k of V is the eigenvector of
// the eigenvalue left in E's diagonal slot k.  Fully unrolled: E and V stay
// in registers.
template <int P, bool WANT_V>
__device__ __forceinline__ void lane_jacobi(double* E, double* V) {
  if (WANT_V) {
#pragma unroll
    for (int i = 0; i < P * P; ++i) V[i] = (i / P == i % P) ? 1.0 : 0.0;
  }
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, dg = 0.0;
#pragma unroll
    for (int i = 0; i < P; ++i) {
      dg += E[upk<P>(i, i)] * E[upk<P>(i, i)];
#pragma unroll
      for (int j = i + 1; j < P; ++j) off += E[upk<P>(i, j)] * E[upk<P>(i, j)];
    }
    if (!(off > 1e-36 * dg)) break;
    bool rotated = false;
#pragma unroll
    for (int p = 0; p < P - 1; ++p) {
#pragma unroll
      for (int q = p + 1; q < P; ++q) {
        const double apq = E[upk<P>(p, q)], app = E[upk<P>(p, p)], aqq = E[upk<P>(q, q)];
        if (apq * apq > (MP_TOL * MP_TOL) * fabs(app * aqq)) {
          const double d = aqq - app, h = 2.0 * apq;
          const double num = d > 0.0 ? h : (d < 0.0 ? -h : fabs(h));
          const double tt = num / (fabs(d) + sqrt(d * d + h * h));
          const double c = rsqrt(tt * tt + 1.0), sn = tt * c;
#pragma unroll
          for (int k = 0; k < P; ++k) {
            if (k == p || k == q) continue;
            const double akp = E[upk<P>(k, p)], akq = E[upk<P>(k, q)];
            E[upk<P>(k, p)] = c * akp - sn * akq;
            E[upk<P>(k, q)] = sn * akp + c * akq;
          }
          E[upk<P>(p, p)] = app - tt * apq;
          E[upk<P>(q, q)] = aqq + tt * apq;
          E[upk<P>(p, q)] = 0.0;
          if (WANT_V) {
#pragma unroll
            for (int k = 0; k < P; ++k) {
              const double vkp = V[k * P + p], vkq = V[k * P + q];
              V[k * P + p] = c * vkp - sn * vkq;
              V[k * P + q] = sn * vkp + c * vkq;
            }
          }
          rotated = true;
        }
      }
    }
    if (!rotated) break;
  }
}

template <int P>
__global__ __launch_bounds__(64, P <= 12 ? 2 : 1) void k_cov_lane(const double* __restrict__ F, int64_t n_ev,
                                                 double* __restrict__ C, double* __restrict__ eps,
                                                 double* __restrict__ cond, int* __restrict__ method) {
  constexpr int NU = P * (P + 1) / 2;
  const int64_t e0 = (int64_t)blockIdx.x * 64;  // the wave's first event (uniform)
  const int lane = threadIdx.x;
  const int64_t e = e0 + lane;
  if (e >= n_ev) return;
  // element (i, j) of event e: a wave-uniform base plus the lane, so each
  // load or store needs a scalar address and one shared lane offset
  const double* Fb = F + e0;
  double* Cb = C + e0;
  // (the offset passes through an empty asm statement so that it is formed
  // in scalar registers at each use, not kept per element in vector ones)
  auto uoff = [&](int i, int j) {
    int64_t off = (int64_t)(i * P + j) * n_ev;
    __asm__ volatile("" : "+s"(off));
    return off;
  };
  auto f_at = [&](int i, int j) { return Fb[uoff(i, j) + lane]; };
  auto pending = [&]() { Cb[uoff(P - 1, P - 1) + lane] = __longlong_as_double((long long)COV_PENDING); };
  // NaN anywhere, or a negative diagonal: the wavefront kernel writes the NaN outputs
  bool bad = false, zero = false;
#pragma unroll
  for (int i = 0; i < P * P; ++i) bad |= isnan(f_at(i / P, i % P));
  double ws[P];
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const double d = f_at(i, i);
    bad |= d < 0.0;
    zero |= d == 0.0;
  }
  if (bad) {
    pending();
    return;
  }
#pragma unroll
  for (int i = 0; i < P; ++i) ws[i] = zero ? 1.0 : 1.0 / sqrt(f_at(i, i));
  __asm__ volatile("" ::: "memory");
  double E[NU];
#pragma unroll
  for (int i = 0; i < P; ++i)
#pragma unroll
    for (int j = i; j < P; ++j) E[upk<P>(i, j)] = ws[i] * f_at(i, j) * ws[j];
  lane_jacobi<P, false>(E, nullptr);
  double emax = 0.0, emin = INFINITY;
  bool neg = false;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const double w = E[upk<P>(i, i)];
    emax = fmax(emax, fabs(w));
    emin = fmin(emin, fabs(w));
    neg = neg || (w < 0.0);
  }
  if (cond) cond[e] = emax / emin;
  if (neg) {  // 'cho' switches to the alternative method
    pending();
    return;
  }
  // mpmath.cholesky of A (its lower triangle from F's lower triangle), in place.
  // (The barrier makes F be read again here rather than kept in registers
  // through the Jacobi sweeps.)
  __asm__ volatile("" ::: "memory");
  double L[NU];
#pragma unroll
  for (int i = 0; i < P; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) L[lok<P>(i, j)] = ws[i] * f_at(i, j) * ws[j];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    double d = L[lok<P>(j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= L[lok<P>(j, k)] * L[lok<P>(j, k)];
    ok = ok && (d >= MP_TOL);
    const double ljj = sqrt(d);
    L[lok<P>(j, j)] = ljj;
#pragma unroll
    for (int i = j + 1; i < P; ++i) {
      double s = L[lok<P>(i, j)];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= L[lok<P>(i, k)] * L[lok<P>(j, k)];
      L[lok<P>(i, j)] = s / ljj;
    }
  }
  if (!ok) {  // a pivot below 2^-52: not positive-definite for mpmath
    pending();
    return;
  }
  // Ci = L^{-1}, column by column upward, each column replacing L's
#pragma unroll
  for (int j = 0; j < P; ++j) {
    double col[P];
    col[j] = 1.0 / L[lok<P>(j, j)];
#pragma unroll
    for (int i = j + 1; i < P; ++i) {
      double s = 0.0;
#pragma unroll
      for (int k = j; k < i; ++k) s -= L[lok<P>(i, k)] * (k == j ? col[j] : col[k]);
      col[i] = s / L[lok<P>(i, i)];
    }
#pragma unroll
    for (int i = j; i < P; ++i) L[lok<P>(i, j)] = col[i];
  }
  // C = Ci^T Ci row by row, row i replacing Ci's column i; scaled back and written
#pragma unroll
  for (int i = 0; i < P; ++i) {
    double row[P];
#pragma unroll
    for (int j = i; j < P; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = j; k < P; ++k) s += L[lok<P>(k, i)] * L[lok<P>(k, j)];
      row[j] = s;
    }
#pragma unroll
    for (int j = i; j < P; ++j) {
      const double c = ws[i] * row[j] * ws[j];
      L[lok<P>(j, i)] = c;  // now C[i][j] (= C[j][i]) at the lower slot (j, i)
      Cb[uoff(i, j) + lane] = c;
      if (j != i) Cb[uoff(j, i) + lane] = c;
    }
  }
  if (eps) {  // max |C F - I|, F a column at a time
    __asm__ volatile("" ::: "memory");
    double mx = 0.0;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      double fc[P];
#pragma unroll
      for (int k = 0; k < P; ++k) fc[k] = f_at(k, j);
#pragma unroll
      for (int i = 0; i < P; ++i) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < P; ++k) s += L[i >= k ? lok<P>(i, k) : lok<P>(k, i)] * fc[k];
        mx = fmax(mx, fabs(s - (i == j ? 1.0 : 0.0)));
      }
    }
    eps[e] = mx;
  }
  if (method) method[e] = GWF_INV_CHO;
}

// fisherTools.CheckFisher (fisherTools.py:216-279) for P = 11, 12: one event
// per lane, the upper triangle of F and the eigenvectors in registers
// (lane_jacobi), eigenvalues ranked ascending as k_check ranks them.  Events
// with a NaN entry get the marker COV_PENDING as their condition number and
// are completed by k_check<PT, true>, which skips every other event.
template <int P, bool WANT_V>
__global__ __launch_bounds__(64, 1) void k_check_lane(const double* __restrict__ F, int64_t n_ev,
                                                       double* __restrict__ evals, double* __restrict__ evecs,
                                                       double* __restrict__ cond) {
  constexpr int NU = P * (P + 1) / 2;
  const int64_t e0 = (int64_t)blockIdx.x * 64;  // the wave's first event (uniform)
  const int lane = threadIdx.x;
  const int64_t e = e0 + lane;
  if (e >= n_ev) return;
  const double* Fb = F + e0;
  auto f_at = [&](int i, int j) {
    int64_t off = (int64_t)(i * P + j) * n_ev;  // scalar, as in k_cov_lane
    __asm__ volatile("" : "+s"(off));
    return Fb[off + lane];
  };
  bool bad = false;
#pragma unroll
  for (int i = 0; i < P * P; ++i) bad |= isnan(f_at(i / P, i % P));
  if (bad) {
    cond[e] = __longlong_as_double((long long)COV_PENDING);
    return;
  }
  double E[NU], V[WANT_V ? P * P : 1];
#pragma unroll
  for (int i = 0; i < P; ++i)
#pragma unroll
    for (int j = i; j < P; ++j) E[upk<P>(i, j)] = f_at(i, j);
  lane_jacobi<P, WANT_V>(E, V);
  double emax = 0.0, emin = INFINITY;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    emax = fmax(emax, fabs(E[upk<P>(i, i)]));
    emin = fmin(emin, fabs(E[upk<P>(i, i)]));
  }
  double* ev = evals + e * P;
  double* vv = WANT_V ? evecs + e * P * P : nullptr;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const double wi = E[upk<P>(i, i)];
    const bool ni = isnan(wi);
    int rk = 0;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const double wj = E[upk<P>(j, j)];
      const bool nj = isnan(wj);
      const bool before = ni ? (nj ? j < i : true) : (!nj && (wj < wi || (wj == wi && j < i)));
      rk += before ? 1 : 0;
    }
    ev[rk] = wi;
    if (WANT_V) {
#pragma unroll
      for (int k = 0; k < P; ++k) vv[k * P + rk] = V[k * P + i];
    }
  }
  cond[e] = emax / emin;
}

// fisherTools.CovMatr per event (fisherTools.py:53-196).  meth / alt:
// GWF_INV_* ids.  Steps: Jacobi scaling ws = 1/sqrt(F_ii) (a zero diagonal
// skips it, a negative one or any NaN gives NaN); eigenvalues of the scaled
// matrix A (condition number); a negative eigenvalue or a failed Cholesky
// switches 'cho' to alt; then the chosen inverse of A, symmetrised and
// unscaled; eps = max|C F - I| against F, or against the truncated scaled
// matrix when 'svd' truncates (as the reference stores it, 128-136).
template <int PT, bool PENDING_ONLY = false>
__global__ __launch_bounds__(64) void k_cov(const double* __restrict__ F, int Prt, int64_t n_ev, int meth_in, int alt,
                                            int truncate, double svals_thresh, double cond_max, double* __restrict__ C,
                                            double* __restrict__ eps, double* __restrict__ cond,
                                            int* __restrict__ method) {
  constexpr int LD = CovDim<PT>::LD, CMAT = CovDim<PT>::MAT;
  const int P = PT > 0 ? PT : Prt;
  __shared__ double sm[6 * CMAT];
  __shared__ double vw[PMAX], vws[PMAX], vcc[PMAX], vss[PMAX];
  __shared__ int vpiv[PMAX], vperm[PMAX];
  const int64_t e = xcd_event();
  if (e >= n_ev) return;  // the block is this one wave
  // PENDING_ONLY: only the events k_cov_lane left (marker in the last element)
  if (PENDING_ONLY && __double_as_longlong(C[(int64_t)(P * P - 1) * n_ev + e]) != (long long)COV_PENDING) return;
  const int lane = threadIdx.x;
  double *Fm = sm, *A = sm + CMAT, *E = sm + 2 * CMAT, *E2 = sm + 3 * CMAT, *V = sm + 4 * CMAT, *V2 = sm + 5 * CMAT;
  for (int t = lane; t < 6 * CMAT; t += 64) sm[t] = 0.0;
  __syncthreads();
  bool nan_l = false;
  for (int t = lane; t < P * P; t += 64) {
    const double v = F[(int64_t)t * n_ev + e];
    AT(Fm, t / P, t - (t / P) * P) = v;
    nan_l = nan_l || isnan(v);
  }
  bool neg_l = false, zero_l = false;
  __syncthreads();
  if (lane < P) {
    const double d = AT(Fm, lane, lane);
    zero_l = d == 0.0;
    neg_l = d < 0.0;
  }
  auto put_nan = [&]() {
    for (int t = lane; t < P * P; t += 64) C[(int64_t)t * n_ev + e] = NAN;
    if (lane == 0) {
      if (eps) eps[e] = NAN;
      if (cond) cond[e] = NAN;
      if (method) method[e] = -1;
    }
  };
  if (__ballot(nan_l || neg_l)) {
    put_nan();
    return;
  }
  // Jacobi scaling ws = diag(1/sqrt(F_ii)); a zero diagonal skips it
  // (fisherTools.py:81-94)
  const bool reweight = __ballot(zero_l) == 0;
  if (lane < P) vws[lane] = reweight ? 1.0 / sqrt(AT(Fm, lane, lane)) : 1.0;
  __syncthreads();
  for (int t = lane; t < P * P; t += 64) {
    const int i = t / P, j = t - (t / P) * P;
    const double a = vws[i] * AT(Fm, i, j) * vws[j];
    AT(A, i, j) = a;
    AT(E, i, j) = a;
  }
  __syncthreads();
  // PENDING_ONLY with 'cho': k_cov_lane has already found that Cholesky
  // does not apply, so the alternative method is taken; when that is an
  // SVD, one Jacobi run with eigenvectors serves both it and the condition
  // number
  const bool vecs_first = PENDING_ONLY && meth_in == GWF_INV_CHO && (alt == GWF_INV_SVD || alt == GWF_INV_SVD_REG);
  if (vecs_first)
    wave_jacobi<PT, true>(E, E2, V, V2, P, vw, vcc, vss, lane);
  else
    wave_jacobi<PT, false>(E, E2, V, V2, P, vw, vcc, vss, lane);
  double emax = 0.0, emin = INFINITY;
  bool neg = false;
  for (int i = 0; i < P; ++i) {
    emax = fmax(emax, fabs(vw[i]));
    emin = fmin(emin, fabs(vw[i]));
    neg = neg || (vw[i] < 0.0);
  }
  const double cnd = emax / emin;
  if (cond && lane == 0) cond[e] = cnd;
  // E, E2, V, V2 are free from here (the svd methods compute the eigenvectors,
  // or with vecs_first hold them in V)
  int m = meth_in;
  if (m == GWF_INV_CHO) {
    if (neg || PENDING_ONLY) {
      m = alt;
    } else {  // mpmath.cholesky: a pivot below 2^-52 is not positive-definite
      double* L = E;
      for (int t = lane; t < P * P; t += 64) AT(L, t / P, t - (t / P) * P) = 0.0;
      __syncthreads();
      for (int j = 0; j < P; ++j) {
        double d = AT(A, j, j);
        for (int k = 0; k < j; ++k) d -= AT(L, j, k) * AT(L, j, k);
        if (!(d >= MP_TOL)) {  // every lane computed the same d
          m = alt;
          break;
        }
        const double ljj = sqrt(d);
        if (lane == j) AT(L, j, j) = ljj;
        if (lane > j && lane < P) {
          double s = AT(A, lane, j);
          for (int k = 0; k < j; ++k) s -= AT(L, lane, k) * AT(L, j, k);
          AT(L, lane, j) = s / ljj;
        }
        __syncthreads();
      }
    }
  }
  double* Cc;
  if (m == GWF_INV_CHO) {
    // Ci = L^{-1} (lower triangular, one column per lane), Cc = Ci^T Ci
    const double* L = E;
    double* Ci = E2;
    if (lane < P) {
      const int j = lane;
      for (int i = 0; i < j; ++i) AT(Ci, i, j) = 0.0;
      AT(Ci, j, j) = 1.0 / AT(L, j, j);
      for (int i = j + 1; i < P; ++i) {
        double s = 0.0;
        for (int k = j; k < i; ++k) s -= AT(L, i, k) * AT(Ci, k, j);
        AT(Ci, i, j) = s / AT(L, i, i);
      }
    }
    __syncthreads();
    Cc = V2;
    for (int t = lane; t < P * P; t += 64) {
      const int i = t / P, j = t - (t / P) * P;
      double s = 0.0;
      for (int k = 0; k < P; ++k) s += AT(Ci, k, i) * AT(Ci, k, j);
      AT(Cc, i, j) = s;
    }
  } else if (m == GWF_INV_SVD || m == GWF_INV_SVD_REG) {
    // SVD of the symmetric A from its eigenpairs: singular values |w_k|,
    // A^+ = sum_k sign(w_k) v_k v_k^T / |w_k|.  The eigenvectors: the same
    // Jacobi run again on A, with V (the values come out the same)
    if (!vecs_first) {
      for (int t = lane; t < P * P; t += 64) AT(E, t / P, t - (t / P) * P) = AT(A, t / P, t - (t / P) * P);
      __syncthreads();
      wave_jacobi<PT, true>(E, E2, V, V2, P, vw, vcc, vss, lane);
    }
    double smax = 0.0;
    for (int k = 0; k < P; ++k) smax = fmax(smax, fabs(vw[k]));
    const bool trunc = (m == GWF_INV_SVD) && truncate && fabs(cnd) > cond_max;
    bool zero_sv = false;
    double* sinv = vcc;
    double* st = vss;
    if (lane < P) {
      const double s = fabs(vw[lane]), sg = vw[lane] < 0.0 ? -1.0 : 1.0;
      if (m == GWF_INV_SVD_REG) {
        sinv[lane] = s > svals_thresh ? sg / s : 0.0;  // singular values <= thresh left out
      } else if (trunc) {
        const bool keep = s / smax > svals_thresh;
        sinv[lane] = sg / (keep ? s : smax * svals_thresh);
        st[lane] = sg * (keep ? s : smax * svals_thresh);
      } else {
        zero_sv = s == 0.0;  // 1/0: no finite inverse
        sinv[lane] = zero_sv ? 0.0 : sg / s;
      }
    }
    if (__ballot(zero_sv)) {
      put_nan();
      return;
    }
    __syncthreads();
    Cc = E;
    for (int t = lane; t < P * P; t += 64) {
      const int i = t / P, j = t - (t / P) * P;
      double s = 0.0;
      for (int k = 0; k < P; ++k) s += AT(V, i, k) * AT(V, j, k) * sinv[k];
      AT(Cc, i, j) = s;
      if (trunc && i <= j) {  // the truncated (scaled) Fisher the inversion error is taken against
        double u = 0.0;
        for (int k = 0; k < P; ++k) u += AT(V, i, k) * AT(V, j, k) * st[k];
        AT(Fm, i, j) = u;
        AT(Fm, j, i) = u;
      }
    }
  } else {  // 'inv' (mpmath inverse) or 'lu' (U^{-1} (P L)^{-1}, fisherTools.py:156-161)
    // mpmath's LU_decomp: row pivot maximising |L[k][j]| / sum_{l>=j} |L[k][l]|;
    // numerically singular (NaN) when a row sum or a pivot is <= ||A||_1 2^-52
    double* L = E;
    for (int t = lane; t < P * P; t += 64) AT(L, t / P, t - (t / P) * P) = AT(A, t / P, t - (t / P) * P);
    __syncthreads();
    double n1 = 0.0;
    if (lane < P) {
      for (int i = 0; i < P; ++i) n1 += fabs(AT(L, i, lane));
    }
    const double tol = wave_max(n1) * MP_TOL;
    bool ok = true;
    for (int j = 0; j < P - 1 && ok; ++j) {
      bool bad = false;
      if (lane >= j && lane < P) {
        double s = 0.0;
        for (int l = j; l < P; ++l) s += fabs(AT(L, lane, l));
        bad = !(s > tol);
        vcc[lane] = (1.0 / s) * fabs(AT(L, lane, j));
      }
      __syncthreads();
      if (__ballot(bad)) {
        ok = false;
        break;
      }
      double biggest = 0.0;
      int pj = -1;
      for (int k = j; k < P; ++k) {
        if (vcc[k] > biggest) {
          biggest = vcc[k];
          pj = k;
        }
      }
      if (pj < 0) {
        ok = false;
        break;
      }
      if (lane == 0) vpiv[j] = pj;
      if (lane < P) {
        const double t = AT(L, j, lane);
        AT(L, j, lane) = AT(L, pj, lane);
        AT(L, pj, lane) = t;
      }
      __syncthreads();
      if (!(fabs(AT(L, j, j)) > tol)) {
        ok = false;
        break;
      }
      if (lane > j && lane < P) {
        const double lij = AT(L, lane, j) / AT(L, j, j);
        AT(L, lane, j) = lij;
        for (int k = j + 1; k < P; ++k) AT(L, lane, k) -= lij * AT(L, j, k);
      }
      __syncthreads();
    }
    if (!ok || !(fabs(AT(L, P - 1, P - 1)) > tol)) {
      put_nan();
      return;
    }
    // Ui = U^{-1} and Li = (unit L)^{-1}, one column per lane
    double *Ui = E2, *Li = V2, *X = V;
    if (lane < P) {
      const int j = lane;
      for (int i = j + 1; i < P; ++i) AT(Ui, i, j) = 0.0;
      AT(Ui, j, j) = 1.0 / AT(L, j, j);
      for (int i = j - 1; i >= 0; --i) {
        double s = 0.0;
        for (int k = i + 1; k <= j; ++k) s -= AT(L, i, k) * AT(Ui, k, j);
        AT(Ui, i, j) = s / AT(L, i, i);
      }
      for (int i = 0; i < j; ++i) AT(Li, i, j) = 0.0;
      AT(Li, j, j) = 1.0;
      for (int i = j + 1; i < P; ++i) {
        double s = 0.0;
        for (int k = j; k < i; ++k) s -= AT(L, i, k) * AT(Li, k, j);
        AT(Li, i, j) = s;
      }
    }
    if (lane == 0) {  // permutation of the row swaps: row r of Pm has its 1 in column perm[r]
      for (int r = 0; r < P; ++r) vperm[r] = r;
      for (int k = 0; k < P - 1; ++k) {
        const int t = vperm[k];
        vperm[k] = vperm[vpiv[k]];
        vperm[vpiv[k]] = t;
      }
    }
    __syncthreads();
    // X = L^{-1} M: 'inv' takes M = Pm (the inverse of A = Pm^T L U), 'lu'
    // takes M = Pm^T, as the reference's U^{-1} (Pm L)^{-1} does
    for (int t = lane; t < P * P; t += 64) {
      const int i = t / P, j = t - (t / P) * P;
      double s = 0.0;
      if (m == GWF_INV_INV) {
        for (int r = 0; r < P; ++r)
          if (vperm[r] == j) s = AT(Li, i, r);
      } else {
        s = AT(Li, i, vperm[j]);
      }
      AT(X, i, j) = s;
    }
    __syncthreads();
    Cc = V2;
    for (int t = lane; t < P * P; t += 64) {
      const int i = t / P, j = t - (t / P) * P;
      double s = 0.0;
      for (int k = 0; k < P; ++k) s += AT(Ui, i, k) * AT(X, k, j);
      AT(Cc, i, j) = s;
    }
  }
  __syncthreads();
  for (int t = lane; t < P * P; t += 64) {
    const int i = t / P, j = t - (t / P) * P;
    const double s = 0.5 * (AT(Cc, i, j) + AT(Cc, j, i));
    const double c = vws[i] * s * vws[j];
    C[(int64_t)t * n_ev + e] = c;
    AT(A, i, j) = c;
  }
  if (eps) {
    __syncthreads();
    double mx = 0.0;
    for (int t = lane; t < P * P; t += 64) {
      const int i = t / P, j = t - (t / P) * P;
      double s = 0.0;
      for (int k = 0; k < P; ++k) s += AT(A, i, k) * AT(Fm, k, j);
      mx = fmax(mx, fabs(s - (i == j ? 1.0 : 0.0)));
    }
    mx = wave_max(mx);
    if (lane == 0) eps[e] = mx;
  }
  if (method && lane == 0) method[e] = m;
}

// fisherTools.CheckFisher per event (fisherTools.py:216-279): eigenvalues
// (ascending), eigenvectors (evecs[e][i][k] = component i of vector k) and
// the condition number max|w| / min|w|; an all-NaN matrix gives NaN.
template <int PT, bool PENDING_ONLY = false>
__global__ __launch_bounds__(64) void k_check(const double* __restrict__ F, int Prt, int64_t n_ev,
                                              double* __restrict__ evals, double* __restrict__ evecs,
                                              double* __restrict__ cond) {
  constexpr int LD = CovDim<PT>::LD, CMAT = CovDim<PT>::MAT;
  const int P = PT > 0 ? PT : Prt;
  __shared__ double sm[4 * CMAT];
  __shared__ double vw[PMAX], vcc[PMAX], vss[PMAX];
  const int64_t e = xcd_event();
  if (e >= n_ev) return;
  // PENDING_ONLY: only the events k_check_lane left (marker as condition number)
  if (PENDING_ONLY && __double_as_longlong(cond[e]) != (long long)COV_PENDING) return;
  const int lane = threadIdx.x;
  double *E = sm, *E2 = sm + CMAT, *V = sm + 2 * CMAT, *V2 = sm + 3 * CMAT;
  for (int t = lane; t < 4 * CMAT; t += 64) sm[t] = 0.0;
  __syncthreads();
  bool num_l = false;
  for (int t = lane; t < P * P; t += 64) {
    const double v = F[(int64_t)t * n_ev + e];
    AT(E, t / P, t - (t / P) * P) = v;
    num_l = num_l || !isnan(v);
  }
  if (__ballot(num_l) == 0) {
    for (int t = lane; t < P; t += 64) evals[e * P + t] = NAN;
    if (evecs)
      for (int t = lane; t < P * P; t += 64) evecs[e * P * P + t] = NAN;
    if (lane == 0) cond[e] = NAN;
    return;
  }
  __syncthreads();
  wave_jacobi<PT, true>(E, E2, V, V2, P, vw, vcc, vss, lane);
  double emax = 0.0, emin = INFINITY;
  for (int i = 0; i < P; ++i) {
    emax = fmax(emax, fabs(vw[i]));
    emin = fmin(emin, fabs(vw[i]));
  }
  for (int t = lane; t < P; t += 64) evals[e * P + t] = vw[t];
  if (evecs)
    for (int t = lane; t < P * P; t += 64) evecs[e * P * P + t] = AT(V, t / P, t - (t / P) * P);
  if (lane == 0) cond[e] = emax / emin;
}

__global__ void k_inv_err(const double* __restrict__ F, const double* __restrict__ C, int P, int64_t n_ev,
                          double* __restrict__ eps) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_ev) return;
  double mx = 0.0;
  for (int i = 0; i < P; ++i)
    for (int j = 0; j < P; ++j) {
      double s = 0.0;
      for (int k = 0; k < P; ++k) s += C[(int64_t)(i * P + k) * n_ev + e] * F[(int64_t)(k * P + j) * n_ev + e];
      const double d = fabs(s - (i == j ? 1.0 : 0.0));
      mx = (d > mx || isnan(d)) ? d : mx;
    }
  eps[e] = mx;
}

// ------------------------------------------------------ post-Fisher batch
struct IdxMap {
  int idx[PMAX];
};

// out (Pn, Pn, n_ev) = rows/columns map.idx[0..Pn) of F (P, P, n_ev);
// diag (P,) optional: added to the kept diagonal first (addPrior)
__global__ void k_sub(const double* __restrict__ F, int P, int64_t n_ev, IdxMap map, int Pn,
                      const double* __restrict__ diag, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)Pn * Pn * n_ev) return;
  const int64_t e = t % n_ev;
  const int ab = (int)(t / n_ev);
  const int a = map.idx[ab / Pn], b = map.idx[ab % Pn];
  double v = F[(int64_t)(a * P + b) * n_ev + e];
  if (diag && a == b) v += diag[a];
  out[t] = v;
}

// sigma (P, n_ev) = sqrt(Cov_aa) and, if area != NULL, the sky area
// scale * 2 pi sqrt(C_thth C_phph - C_phth^2) |sin theta|
__global__ void k_sigma_area(const double* __restrict__ C, int P, int64_t n_ev, const double* __restrict__ theta,
                             int ith, int iph, double scale, double* __restrict__ sigma, double* __restrict__ area) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_ev) return;
  if (sigma)
    for (int a = 0; a < P; ++a) sigma[(int64_t)a * n_ev + e] = sqrt(C[(int64_t)(a * P + a) * n_ev + e]);
  if (area) {
    const double tt = C[(int64_t)(ith * P + ith) * n_ev + e], pp = C[(int64_t)(iph * P + iph) * n_ev + e];
    const double tp = C[(int64_t)(iph * P + ith) * n_ev + e];
    area[e] = scale * 2.0 * M_PI * sqrt(tt * pp - tp * tp) * fabs(sin(theta[e]));
  }
}

// Fisher matrices in (m1, m2) instead of (Mc, eta) and / or (chiS, chiA)
// instead of (chi1z, chi2z): F' = J^T F J with J the exact Jacobian of the
// old parameters with respect to the new ones at the event (the reference
// differentiates GWstrain in the new variables directly, signal.py:818-845;
// the chain rule makes the two equal up to rounding).
__global__ void k_reparam(const double* __restrict__ par, int64_t ld, double* __restrict__ F, int P, int64_t n_ev,
                          int use_m1m2, int use_chiSA) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_ev) return;
  // J restricted to the two 2x2 blocks: rows/cols (0, 1) and (9, 10)
  double j00 = 1., j01 = 0., j10 = 0., j11 = 1.;
  if (use_m1m2) {
    const double Mc = par[e], eta = par[ld + e];
    const double M = Mc / pow(eta, 3. / 5.);
    const double Seta = sqrt(eta < 0.25 ? 1.0 - 4.0 * eta : 0.0);
    const double m1 = 0.5 * M * (1.0 + Seta), m2 = 0.5 * M * (1.0 - Seta);
    // d(Mc, eta) / d(m1, m2)
    j00 = Mc * (3. / (5. * m1) - 1. / (5. * M));
    j01 = Mc * (3. / (5. * m2) - 1. / (5. * M));
    j10 = eta * (1. / m1 - 2. / M);
    j11 = eta * (1. / m2 - 2. / M);
  }
  double s00 = 1., s01 = 0., s10 = 0., s11 = 1.;
  if (use_chiSA) {  // chi1 = chiS + chiA, chi2 = chiS - chiA
    s00 = 1.;
    s01 = 1.;
    s10 = 1.;
    s11 = -1.;
  }
  auto at = [&](int a, int b) -> double& { return F[(int64_t)(a * P + b) * n_ev + e]; };
  // right multiplication F J (columns), then left J^T (rows), block by block
  for (int pass = 0; pass < 2; ++pass) {
    const int c0 = pass == 0 ? 0 : 9;
    const double a00 = pass == 0 ? j00 : s00, a01 = pass == 0 ? j01 : s01;
    const double a10 = pass == 0 ? j10 : s10, a11 = pass == 0 ? j11 : s11;
    if (a00 == 1. && a01 == 0. && a10 == 0. && a11 == 1.) continue;
    for (int r = 0; r < P; ++r) {  // columns: new_c0 = F[r][c0] a00 + F[r][c0+1] a10, new_c1 = F[r][c0] a01 + F[r][c0+1] a11
      const double x = at(r, c0), y = at(r, c0 + 1);
      at(r, c0) = x * a00 + y * a10;
      at(r, c0 + 1) = x * a01 + y * a11;
    }
    for (int c = 0; c < P; ++c) {  // rows
      const double x = at(c0, c), y = at(c0 + 1, c);
      at(c0, c) = a00 * x + a10 * y;
      at(c0 + 1, c) = a01 * x + a11 * y;
    }
  }
}

// out = A F B per event, A and B the identity except the block on rows /
// columns (i0, i1) (i1 < 0: the single entry i0), given per event as SoA
// (4, n_ev): L = (A[i0][i0], A[i0][i1], A[i1][i0], A[i1][i1]), R likewise for B.
// out may alias F: each thread owns one event.
__global__ void k_block_transform(const double* __restrict__ F, double* out, int P, int64_t n_ev, int i0, int i1,
                                  const double* __restrict__ L, const double* __restrict__ R) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_ev) return;
  auto at = [&](int a, int b) -> double& { return out[(int64_t)(a * P + b) * n_ev + e]; };
  if (F != out)
    for (int k = 0; k < P * P; ++k) out[(int64_t)k * n_ev + e] = F[(int64_t)k * n_ev + e];
  const double r00 = R[e], r01 = R[n_ev + e], r10 = R[2 * n_ev + e], r11 = R[3 * n_ev + e];
  const double l00 = L[e], l01 = L[n_ev + e], l10 = L[2 * n_ev + e], l11 = L[3 * n_ev + e];
  if (i1 < 0) {
    for (int r = 0; r < P; ++r) at(r, i0) *= r00;
    for (int c = 0; c < P; ++c) at(i0, c) *= l00;
    return;
  }
  for (int r = 0; r < P; ++r) {  // G = F B
    const double x = at(r, i0), y = at(r, i1);
    at(r, i0) = x * r00 + y * r10;
    at(r, i1) = x * r01 + y * r11;
  }
  for (int c = 0; c < P; ++c) {  // A G
    const double x = at(i0, c), y = at(i1, c);
    at(i0, c) = l00 * x + l01 * y;
    at(i1, c) = l10 * x + l11 * y;
  }
}

// ---------------------------------------------------------------- host side
int check_wf(const gwf_waveform* wf) {
  if (!wf) return fail(-EINVAL, "waveform description is NULL");
  if (wf->model != GWF_MODEL_TAYLORF2 && wf->model != GWF_MODEL_IMRPHENOMD &&
      wf->model != GWF_MODEL_IMRPHENOMD_NRTIDALV2)
    return fail(-EINVAL, "unknown waveform model");
  if (is_pd(wf->model) && (!wf->qnm_a || !wf->qnm_fring || !wf->qnm_fdamp || wf->n_qnm < 2))
    return fail(-EINVAL, "IMRPhenomD needs the QNM tables");
  if (wf->model == GWF_MODEL_IMRPHENOMD && wf->is_tidal) return fail(-EINVAL, "IMRPhenomD is not tidal");
  if (wf->is_eccentric && wf->model != GWF_MODEL_TAYLORF2)
    return fail(-EINVAL, "eccentricity is supported for TaylorF2 only");
  if (wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2 && !wf->is_tidal)
    return fail(-EINVAL, "IMRPhenomD_NRTidalv2 is tidal (13 parameters)");
  return 0;
}

int nt_of(const gwf_waveform* wf) {
  if (wf->model == GWF_MODEL_TAYLORF2 && wf->is_eccentric) return wf->is_tidal ? 7 : 5;  // 14 / 12 parameters
  return ((wf->model == GWF_MODEL_TAYLORF2 && wf->is_tidal) || wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2) ? 6 : 4;
}

// Model dispatch of the single-wave Fisher kernels for one grid spacing.
// Network sums run on the producer / consumer pairs of k_fisher_pq instead
// (GWF_FISHER_PQ / GWF_RAGGED_PQ = 0 switch that off); measured per 10^4
// events x 2000 points (IMRPhenomD, 5 interferometers): k_fisher_pq 2.37 ms,
// k_fisher 2.96 ms; ragged 2 x 10^4 events of 500..10^4 points: 10.6 ms
// against 14.2 ms.
#ifndef GWF_FISHER_PQ
#define GWF_FISHER_PQ 1
#endif
#ifndef GWF_RAGGED_PQ
#define GWF_RAGGED_PQ 1
#endif
#ifndef GWF_MASKED_PQ
#define GWF_MASKED_PQ 1  // duty-cycle (arm-mask) network sums on k_fisher_pq
#endif

template <bool LIN>
void launch_fisher(const gwf_waveform* wf, dim3 grid, dim3 block, hipStream_t st, const NetK& net, const double* rec,
                   int64_t n_ev, int arm, double* out) {
  if (wf->model == GWF_MODEL_IMRPHENOMD)
    hipLaunchKernelGGL((k_fisher<GWF_MODEL_IMRPHENOMD, 4, 11, LIN>), grid, block, 0, st, *wf, net, rec, n_ev, arm, out);
  else if (wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2)
    hipLaunchKernelGGL((k_fisher<GWF_MODEL_IMRPHENOMD_NRTIDALV2, 6, 13, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       arm, out);
  else if (nt_of(wf) == 5)
    hipLaunchKernelGGL((k_fisher<GWF_MODEL_TAYLORF2, 5, 12, LIN>), grid, block, 0, st, *wf, net, rec, n_ev, arm, out);
  else if (nt_of(wf) == 7)  // tidal and eccentric TaylorF2, 14 parameters
    hipLaunchKernelGGL((k_fisher<GWF_MODEL_TAYLORF2, 7, 14, LIN>), grid, block, 0, st, *wf, net, rec, n_ev, arm, out);
  else if (nt_of(wf) == 6)
    hipLaunchKernelGGL((k_fisher<GWF_MODEL_TAYLORF2, 6, 13, LIN>), grid, block, 0, st, *wf, net, rec, n_ev, arm, out);
  else
    hipLaunchKernelGGL((k_fisher<GWF_MODEL_TAYLORF2, 4, 11, LIN>), grid, block, 0, st, *wf, net, rec, n_ev, arm, out);
}

template <bool LIN>
void launch_fisher_masked(const gwf_waveform* wf, dim3 grid, dim3 block, hipStream_t st, const NetK& net,
                          const double* rec, int64_t n_ev, const uint32_t* mask, double* out) {
  if (wf->model == GWF_MODEL_IMRPHENOMD)
    hipLaunchKernelGGL((k_fisher_masked<GWF_MODEL_IMRPHENOMD, 4, 11, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       mask, out);
  else if (wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2)
    hipLaunchKernelGGL((k_fisher_masked<GWF_MODEL_IMRPHENOMD_NRTIDALV2, 6, 13, LIN>), grid, block, 0, st, *wf, net,
                       rec, n_ev, mask, out);
  else if (nt_of(wf) == 5)
    hipLaunchKernelGGL((k_fisher_masked<GWF_MODEL_TAYLORF2, 5, 12, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       mask, out);
  else if (nt_of(wf) == 7)  // tidal and eccentric TaylorF2, 14 parameters
    hipLaunchKernelGGL((k_fisher_masked<GWF_MODEL_TAYLORF2, 7, 14, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       mask, out);
  else if (nt_of(wf) == 6)
    hipLaunchKernelGGL((k_fisher_masked<GWF_MODEL_TAYLORF2, 6, 13, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       mask, out);
  else
    hipLaunchKernelGGL((k_fisher_masked<GWF_MODEL_TAYLORF2, 4, 11, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       mask, out);
}

template <bool LIN>
void launch_fisher_ragged(const gwf_waveform* wf, dim3 grid, dim3 block, hipStream_t st, const NetK& net,
                          const double* rec, int64_t n_ev, const int32_t* res_ev, int arm, unsigned int* counter,
                          double* out) {
  if (wf->model == GWF_MODEL_IMRPHENOMD)
    hipLaunchKernelGGL((k_fisher_ragged<GWF_MODEL_IMRPHENOMD, 4, 11, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       res_ev, arm, counter, out);
  else if (wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2)
    hipLaunchKernelGGL((k_fisher_ragged<GWF_MODEL_IMRPHENOMD_NRTIDALV2, 6, 13, LIN>), grid, block, 0, st, *wf, net,
                       rec, n_ev, res_ev, arm, counter, out);
  else if (nt_of(wf) == 5)
    hipLaunchKernelGGL((k_fisher_ragged<GWF_MODEL_TAYLORF2, 5, 12, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       res_ev, arm, counter, out);
  else if (nt_of(wf) == 7)  // tidal and eccentric TaylorF2, 14 parameters
    hipLaunchKernelGGL((k_fisher_ragged<GWF_MODEL_TAYLORF2, 7, 14, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       res_ev, arm, counter, out);
  else if (nt_of(wf) == 6)
    hipLaunchKernelGGL((k_fisher_ragged<GWF_MODEL_TAYLORF2, 6, 13, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       res_ev, arm, counter, out);
  else
    hipLaunchKernelGGL((k_fisher_ragged<GWF_MODEL_TAYLORF2, 4, 11, LIN>), grid, block, 0, st, *wf, net, rec, n_ev,
                       res_ev, arm, counter, out);
}

// The network Fisher on producer / consumer pairs (k_fisher_pq): one
// resident block per CU, then the NaN fill that reports an expired wait.
// counter: two words, zeroed here.
template <bool LIN, bool RAGGED, int MODE = PQ_SUM>
int launch_fisher_pq(const gwf_waveform* wf, hipStream_t st, const NetK& net_in, const double* rec, int64_t n_ev,
                     const int32_t* res_ev, unsigned int* counter, double* out, int P,
                     const uint32_t* arm_mask = nullptr) {
  int dev = 0, n_cu = 0;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  rc = hip_check(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, counter, 2);
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n_ev + PC_EPB - 1) / PC_EPB, n_cu));
  NetK net = net_in;
  net.spin_limit = g_pq_spin_limit;
  const dim3 g((unsigned)blocks), b(128 * PC_EPB);
  if (wf->model == GWF_MODEL_IMRPHENOMD)
    hipLaunchKernelGGL((k_fisher_pq<GWF_MODEL_IMRPHENOMD, 4, 11, LIN, RAGGED, MODE>), g, b, 0, st, *wf, net, rec,
                       n_ev, res_ev, arm_mask, counter, out);
  else if (wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2)
    hipLaunchKernelGGL((k_fisher_pq<GWF_MODEL_IMRPHENOMD_NRTIDALV2, 6, 13, LIN, RAGGED, MODE>), g, b, 0, st, *wf,
                       net, rec, n_ev, res_ev, arm_mask, counter, out);
  else if (nt_of(wf) == 5)
    hipLaunchKernelGGL((k_fisher_pq<GWF_MODEL_TAYLORF2, 5, 12, LIN, RAGGED, MODE>), g, b, 0, st, *wf, net, rec,
                       n_ev, res_ev, arm_mask, counter, out);
  else if (nt_of(wf) == 7)  // tidal and eccentric TaylorF2, 14 parameters
    hipLaunchKernelGGL((k_fisher_pq<GWF_MODEL_TAYLORF2, 7, 14, LIN, RAGGED, MODE>), g, b, 0, st, *wf, net, rec,
                       n_ev, res_ev, arm_mask, counter, out);
  else if (nt_of(wf) == 6)
    hipLaunchKernelGGL((k_fisher_pq<GWF_MODEL_TAYLORF2, 6, 13, LIN, RAGGED, MODE>), g, b, 0, st, *wf, net, rec,
                       n_ev, res_ev, arm_mask, counter, out);
  else
    hipLaunchKernelGGL((k_fisher_pq<GWF_MODEL_TAYLORF2, 4, 11, LIN, RAGGED, MODE>), g, b, 0, st, *wf, net, rec,
                       n_ev, res_ev, arm_mask, counter, out);
  const int64_t n_out = (int64_t)P * P * n_ev * (MODE == PQ_ARMS ? net.n_arms : 1);
  hipLaunchKernelGGL(k_pq_poison, dim3((unsigned)std::min<int64_t>((n_out + 255) / 256, 1024)), dim3(256), 0, st,
                     counter, out, n_out);
  return 0;
}

int launch_setup(const gwf_waveform* wf, const NetK& net, const double* par, int64_t ld, int64_t n_ev,
                 const double* fgrid, int64_t nf, double* rec, hipStream_t st) {
  // one event per thread; the setup is latency-bound (one dependent chain per
  // event), so small batches spread one wave per CU (measured at 10^4 events:
  // 64-thread blocks 0.095 ms, 256-thread blocks 0.132 ms; at 10^6 events
  // 4.46 against 4.37 ms)
  // (IMRPhenomD models: NT lanes per event, the tangent-split record)
  const int lpe = wf->model == GWF_MODEL_IMRPHENOMD ? 4 : (wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2 ? 6 : 1);
  const int64_t nt = n_ev * lpe;
  const int bs = nt <= 65536 * 4 ? 64 : 256;
  const dim3 grid((unsigned)((nt + bs - 1) / bs)), block(bs);
  if (wf->model == GWF_MODEL_IMRPHENOMD)
    hipLaunchKernelGGL((k_setup<GWF_MODEL_IMRPHENOMD, 4>), grid, block, 0, st, *wf, net, par, ld, n_ev, fgrid, nf, rec);
  else if (wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2)
    hipLaunchKernelGGL((k_setup<GWF_MODEL_IMRPHENOMD_NRTIDALV2, 6>), grid, block, 0, st, *wf, net, par, ld, n_ev, fgrid,
                       nf, rec);
  else if (nt_of(wf) == 5)
    hipLaunchKernelGGL((k_setup<GWF_MODEL_TAYLORF2, 5>), grid, block, 0, st, *wf, net, par, ld, n_ev, fgrid, nf, rec);
  else if (nt_of(wf) == 7)  // tidal and eccentric TaylorF2, 14 parameters
    hipLaunchKernelGGL((k_setup<GWF_MODEL_TAYLORF2, 7>), grid, block, 0, st, *wf, net, par, ld, n_ev, fgrid, nf, rec);
  else if (nt_of(wf) == 6)
    hipLaunchKernelGGL((k_setup<GWF_MODEL_TAYLORF2, 6>), grid, block, 0, st, *wf, net, par, ld, n_ev, fgrid, nf, rec);
  else
    hipLaunchKernelGGL((k_setup<GWF_MODEL_TAYLORF2, 4>), grid, block, 0, st, *wf, net, par, ld, n_ev, fgrid, nf, rec);
  return hip_check(hipGetLastError(), "setup launch");
}

int check_common(const gwf_waveform* wf, const double* par, int64_t ld, int64_t n_ev, void* ws, size_t ws_bytes) {
  int rc = check_wf(wf);
  if (rc) return rc;
  if (n_ev < 0) return fail(-EINVAL, "n_ev must be >= 0");
  if (n_ev > 0 && !par) return fail(-EINVAL, "params is NULL");
  if (ld < n_ev) return fail(-EINVAL, "ld must be >= n_ev");
  if (n_ev > (int64_t)0x7fffffff / WAVES_PER_BLOCK) return fail(-EINVAL, "too many events for one call");
  if (ws_bytes < (size_t)n_ev * REC_DOUBLES * sizeof(double) || (n_ev > 0 && !ws))
    return fail(-EINVAL, "workspace too small (see gwf_workspace_bytes)");
  return 0;
}

// Per-device scratch of gwf_fisher_noalloc: the event records and the queue
// counters of up to GWF_NOALLOC_MAX_EVENTS events in a file-scope device
// array (each GPU the library runs on holds its own copy until the process
// ends).  Host code never passes `scratch` itself to a kernel: its device
// address comes from hipGetSymbolAddress, once per device.
constexpr size_t NOALLOC_BYTES = (size_t)GWF_NOALLOC_MAX_EVENTS * REC_DOUBLES * sizeof(double) + 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_DEVICES] = {};

// The current device's address of `scratch`, looked up on the first call on
// that device with the device current.
int scratch_address(void** out) {
  int dev = 0;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= SCRATCH_DEVICES) return fail(-EINVAL, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    if (!p) return fail(-EIO, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

// the screen kernels (gwf_mlp.hip) report through the same gwf_last_error
namespace gwf_mlp {
void set_error(const std::string& s) { g_err = s; }
}  // namespace gwf_mlp

// ====================================================================== ABI
extern "C" {

int gwf_abi_version(void) { return GWF_ABI_VERSION; }

const char* gwf_last_error(void) { return g_err.c_str(); }

size_t gwf_workspace_bytes(int64_t n_ev) {
  // event records + 256 bytes of counters (ragged launches)
  return (size_t)(n_ev < 0 ? 0 : n_ev) * REC_DOUBLES * sizeof(double) + 256;
}

int gwf_num_arms(const gwf_detector* dets, int32_t n_det) {
  if (!dets || n_det < 0) return fail(-EINVAL, "invalid detector list");
  int n = 0;
  for (int d = 0; d < n_det; ++d) n += dets[d].shape == GWF_SHAPE_T ? 3 : 1;
  return n;
}

int gwf_fcut(const gwf_waveform* wf, const double* params, int64_t ld, int64_t n_ev, double fmax, double* fcut_out,
             void* stream) {
  int rc = check_wf(wf);
  if (rc) return rc;
  if (n_ev == 0) return 0;
  if (!params || !fcut_out || ld < n_ev) return fail(-EINVAL, "invalid arguments to gwf_fcut");
  hipLaunchKernelGGL(k_fcut, dim3((unsigned)((n_ev + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *wf, params, ld,
                     n_ev, fmax, fcut_out);
  return hip_check(hipGetLastError(), "gwf_fcut launch");
}

int gwf_grid(const gwf_waveform* wf, const gwf_detector* det, const double* params, int64_t ld, int64_t n_ev,
                int32_t res, double* f_out, double* S_out, void* workspace, size_t ws_bytes, void* stream) {
  int rc = check_common(wf, params, ld, n_ev, workspace, ws_bytes);
  if (rc) return rc;
  if (res < 2) return fail(-EINVAL, "res must be >= 2");
  if (n_ev == 0) return 0;
  if (!f_out) return fail(-EINVAL, "f_out is NULL");
  NetK net;
  rc = build_net_checked(det, 1, res, net);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  rc = launch_setup(wf, net, params, ld, n_ev, nullptr, 0, (double*)workspace, st);
  if (rc) return rc;
  const int64_t tot = (int64_t)res * n_ev;
  hipLaunchKernelGGL(k_grid, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, net, (const double*)workspace, n_ev,
                     (int)res, f_out, S_out);
  return hip_check(hipGetLastError(), "gwf_grid launch");
}

int gwf_grid_fcut(const gwf_detector* det, const double* fcut, int64_t n_ev, int32_t res, double* f_out,
                  double* S_out, void* stream) {
  if (res < 2) return fail(-EINVAL, "res must be >= 2");
  if (n_ev < 0) return fail(-EINVAL, "n_ev must be >= 0");
  if (n_ev == 0) return 0;
  if (!fcut || !f_out) return fail(-EINVAL, "NULL array");
  NetK net;
  int rc = build_net_checked(det, 1, res, net);
  if (rc) return rc;
  const int64_t tot = (int64_t)res * n_ev;
  hipLaunchKernelGGL(k_grid_fcut, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, (hipStream_t)stream, net, fcut,
                     n_ev, (int)res, f_out, S_out);
  return hip_check(hipGetLastError(), "gwf_grid_fcut launch");
}

int gwf_overlap(const double* h1, const double* h2, const double* f, const double* S, int64_t n_ev, int64_t n_f,
                int32_t accumulate, double* out, void* stream) {
  if (n_ev < 0 || n_f < 0) return fail(-EINVAL, "invalid sizes");
  if (n_ev == 0) return 0;
  if (!h1 || !h2 || !f || !S || !out) return fail(-EINVAL, "NULL array");
  if (n_ev > (int64_t)0x7fffffff / WAVES_PER_BLOCK) return fail(-EINVAL, "too many events for one call");
  hipLaunchKernelGGL(k_overlap, dim3((unsigned)((n_ev + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK)), dim3(BLOCK), 0,
                     (hipStream_t)stream, h1, h2, f, S, n_ev, n_f, (int)accumulate, out);
  return hip_check(hipGetLastError(), "gwf_overlap launch");
}

int gwf_waveform_eval(const gwf_waveform* wf, const double* params, int64_t ld, int64_t n_ev, const double* f,
                      int64_t n_f, double* ampl_out, double* phi_out, double* tau_out, void* workspace,
                      size_t ws_bytes, void* stream) {
  int rc = check_common(wf, params, ld, n_ev, workspace, ws_bytes);
  if (rc) return rc;
  if (n_ev == 0 || n_f == 0) return 0;
  if (!f || n_f < 1) return fail(-EINVAL, "frequency grid is NULL");
  NetK net;
  std::memset(&net, 0, sizeof(net));
  net.fmin = 1.0;
  hipStream_t st = (hipStream_t)stream;
  rc = launch_setup(wf, net, params, ld, n_ev, f, n_f, (double*)workspace, st);
  if (rc) return rc;
  const int64_t tot = n_f * n_ev;
  const dim3 grid((unsigned)((tot + 255) / 256)), block(256);
  const double* rec = (const double*)workspace;
  if (wf->model == GWF_MODEL_IMRPHENOMD)
    hipLaunchKernelGGL((k_wfeval<GWF_MODEL_IMRPHENOMD, 4>), grid, block, 0, st, *wf, rec, n_ev, f, n_f, ampl_out,
                       phi_out, tau_out);
  else if (wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2)
    hipLaunchKernelGGL((k_wfeval<GWF_MODEL_IMRPHENOMD_NRTIDALV2, 6>), grid, block, 0, st, *wf, rec, n_ev, f, n_f,
                       ampl_out, phi_out, tau_out);
  else if (nt_of(wf) == 5)
    hipLaunchKernelGGL((k_wfeval<GWF_MODEL_TAYLORF2, 5>), grid, block, 0, st, *wf, rec, n_ev, f, n_f, ampl_out, phi_out,
                       tau_out);
  else if (nt_of(wf) == 7)  // tidal and eccentric TaylorF2, 14 parameters
    hipLaunchKernelGGL((k_wfeval<GWF_MODEL_TAYLORF2, 7>), grid, block, 0, st, *wf, rec, n_ev, f, n_f, ampl_out, phi_out,
                       tau_out);
  else if (nt_of(wf) == 6)
    hipLaunchKernelGGL((k_wfeval<GWF_MODEL_TAYLORF2, 6>), grid, block, 0, st, *wf, rec, n_ev, f, n_f, ampl_out, phi_out,
                       tau_out);
  else
    hipLaunchKernelGGL((k_wfeval<GWF_MODEL_TAYLORF2, 4>), grid, block, 0, st, *wf, rec, n_ev, f, n_f, ampl_out, phi_out,
                       tau_out);
  return hip_check(hipGetLastError(), "gwf_waveform_eval launch");
}

int gwf_prepare(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params, int64_t ld,
                int64_t n_ev, int32_t res, void* workspace, size_t ws_bytes, void* stream) {
  int rc = check_common(wf, params, ld, n_ev, workspace, ws_bytes);
  if (rc) return rc;
  if (res < 2) return fail(-EINVAL, "res must be >= 2");
  NetK net;
  rc = build_net_checked(dets, n_det, res, net);
  if (rc) return rc;
  if (n_ev == 0) return 0;
  return launch_setup(wf, net, params, ld, n_ev, nullptr, 0, (double*)workspace, (hipStream_t)stream);
}

static int snr_prepared_impl(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, int64_t n_ev, int32_t res,
                      const int32_t* res_ev, double* snrsq_out, const void* workspace, size_t ws_bytes, void* stream) {
  int rc = check_wf(wf);
  if (rc) return rc;
  if (res < 2) return fail(-EINVAL, "res must be >= 2");
  if (n_ev < 0 || n_ev > (int64_t)0x7fffffff / WAVES_PER_BLOCK) return fail(-EINVAL, "invalid n_ev");
  if (ws_bytes < (size_t)n_ev * REC_DOUBLES * sizeof(double) || (n_ev > 0 && !workspace))
    return fail(-EINVAL, "workspace too small (see gwf_workspace_bytes)");
  NetK net;
  rc = build_net_checked(dets, n_det, res, net);
  if (rc) return rc;
  if (n_ev == 0) return 0;
  if (!snrsq_out) return fail(-EINVAL, "snrsq_out is NULL");
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)((n_ev + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK)), block(BLOCK);
  const double* rec = (const double*)workspace;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, block, 0, st, *wf, net, rec, n_ev, res_ev, snrsq_out);
  };
  const bool small = net.n_arms <= 8;  // per-arm accumulators held in registers
  if (wf->model == GWF_MODEL_IMRPHENOMD)
    small ? go(k_snr<GWF_MODEL_IMRPHENOMD, 8>) : go(k_snr<GWF_MODEL_IMRPHENOMD, GWF_MAX_ARMS>);
  else if (wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2)
    small ? go(k_snr<GWF_MODEL_IMRPHENOMD_NRTIDALV2, 8>) : go(k_snr<GWF_MODEL_IMRPHENOMD_NRTIDALV2, GWF_MAX_ARMS>);
  else
    small ? go(k_snr<GWF_MODEL_TAYLORF2, 8>) : go(k_snr<GWF_MODEL_TAYLORF2, GWF_MAX_ARMS>);
  return hip_check(hipGetLastError(), "gwf_snr launch");
}

int gwf_snr_prepared(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, int64_t n_ev, int32_t res,
                     double* snrsq_out, const void* workspace, size_t ws_bytes, void* stream) {
  return snr_prepared_impl(wf, dets, n_det, n_ev, res, nullptr, snrsq_out, workspace, ws_bytes, stream);
}

int gwf_fisher_prepared(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, int64_t n_ev, int32_t res,
                        int32_t arm, double* fisher_out, void* workspace, size_t ws_bytes, void* stream) {
  int rc = check_wf(wf);
  if (rc) return rc;
  if (res < 2) return fail(-EINVAL, "res must be >= 2");
  if (n_ev < 0 || n_ev > (int64_t)0x7fffffff / WAVES_PER_BLOCK) return fail(-EINVAL, "invalid n_ev");
  if (ws_bytes < (size_t)n_ev * REC_DOUBLES * sizeof(double) || (n_ev > 0 && !workspace))
    return fail(-EINVAL, "workspace too small (see gwf_workspace_bytes)");
  NetK net;
  rc = build_net_checked(dets, n_det, res, net);
  if (rc) return rc;
  if (arm >= net.n_arms) return fail(-EINVAL, "arm index out of range");
  if (n_ev == 0) return 0;
  if (!fisher_out) return fail(-EINVAL, "fisher_out is NULL");
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)((n_ev + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK)), block(BLOCK);
  const double* rec = (const double*)workspace;
  if (GWF_FISHER_PQ && arm < 0) {  // network sum on producer / consumer pairs
    if (ws_bytes < gwf_workspace_bytes(n_ev)) return fail(-EINVAL, "workspace too small (see gwf_workspace_bytes)");
    unsigned int* counter = (unsigned int*)((char*)workspace + (size_t)n_ev * REC_DOUBLES * sizeof(double));
    rc = net.lin ? launch_fisher_pq<true, false>(wf, st, net, rec, n_ev, nullptr, counter, fisher_out, 7 + nt_of(wf))
                 : launch_fisher_pq<false, false>(wf, st, net, rec, n_ev, nullptr, counter, fisher_out, 7 + nt_of(wf));
    if (rc) return rc;
    return hip_check(hipGetLastError(), "gwf_fisher launch");
  }
  if (net.lin)
    launch_fisher<true>(wf, grid, block, st, net, rec, n_ev, arm, fisher_out);
  else
    launch_fisher<false>(wf, grid, block, st, net, rec, n_ev, arm, fisher_out);
  return hip_check(hipGetLastError(), "gwf_fisher launch");
}

int gwf_fisher_masked(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params,
                      int64_t ld, int64_t n_ev, int32_t res, const uint32_t* arm_mask, double* fisher_out,
                      void* workspace, size_t ws_bytes, void* stream) {
  int rc = gwf_prepare(wf, dets, n_det, params, ld, n_ev, res, workspace, ws_bytes, stream);
  if (rc) return rc;
  NetK net;
  rc = build_net_checked(dets, n_det, res, net);
  if (rc) return rc;
  if (n_ev == 0) return 0;
  if (!fisher_out || !arm_mask) return fail(-EINVAL, "NULL array");
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)((n_ev + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK)), block(BLOCK);
  const double* rec = (const double*)workspace;
  if (GWF_MASKED_PQ) {  // producer / consumer pairs (needs the counter words)
    if (ws_bytes < gwf_workspace_bytes(n_ev)) return fail(-EINVAL, "workspace too small (see gwf_workspace_bytes)");
    unsigned int* counter = (unsigned int*)((char*)workspace + (size_t)n_ev * REC_DOUBLES * sizeof(double));
    rc = net.lin ? launch_fisher_pq<true, false, PQ_MASKED>(wf, st, net, rec, n_ev, nullptr, counter, fisher_out,
                                                        7 + nt_of(wf), arm_mask)
                 : launch_fisher_pq<false, false, PQ_MASKED>(wf, st, net, rec, n_ev, nullptr, counter, fisher_out,
                                                         7 + nt_of(wf), arm_mask);
    if (rc) return rc;
    return hip_check(hipGetLastError(), "gwf_fisher_masked launch");
  }
  if (net.lin)
    launch_fisher_masked<true>(wf, grid, block, st, net, rec, n_ev, arm_mask, fisher_out);
  else
    launch_fisher_masked<false>(wf, grid, block, st, net, rec, n_ev, arm_mask, fisher_out);
  return hip_check(hipGetLastError(), "gwf_fisher_masked launch");
}

int gwf_set_queue_spin_limit(uint32_t polls) {
  g_pq_spin_limit = polls ? polls : PQ_SPIN_LIMIT;
  return 0;
}

int gwf_fisher_arms(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params,
                    int64_t ld, int64_t n_ev, int32_t res, const int32_t* res_ev, double* fisher_out,
                    void* workspace, size_t ws_bytes, void* stream) {
  if (res_ev) res = 2;  // per-event grid lengths (the records do not depend on res)
  int rc = gwf_prepare(wf, dets, n_det, params, ld, n_ev, res, workspace, ws_bytes, stream);
  if (rc) return rc;
  NetK net;
  rc = build_net_checked(dets, n_det, res, net);
  if (rc) return rc;
  if (net.n_arms > PQ_ARMS_MAX) return fail(-EINVAL, "gwf_fisher_arms takes at most 8 arms (gwf_fisher per arm)");
  if (n_ev == 0) return 0;
  if (!fisher_out) return fail(-EINVAL, "fisher_out is NULL");
  if (ws_bytes < gwf_workspace_bytes(n_ev)) return fail(-EINVAL, "workspace too small (see gwf_workspace_bytes)");
  hipStream_t st = (hipStream_t)stream;
  const double* rec = (const double*)workspace;
  unsigned int* counter = (unsigned int*)((char*)workspace + (size_t)n_ev * REC_DOUBLES * sizeof(double));
  const int P = 7 + nt_of(wf);
  if (res_ev)
    rc = net.lin ? launch_fisher_pq<true, true, PQ_ARMS>(wf, st, net, rec, n_ev, res_ev, counter, fisher_out, P)
                 : launch_fisher_pq<false, true, PQ_ARMS>(wf, st, net, rec, n_ev, res_ev, counter, fisher_out, P);
  else
    rc = net.lin ? launch_fisher_pq<true, false, PQ_ARMS>(wf, st, net, rec, n_ev, nullptr, counter, fisher_out, P)
                 : launch_fisher_pq<false, false, PQ_ARMS>(wf, st, net, rec, n_ev, nullptr, counter, fisher_out, P);
  if (rc) return rc;
  return hip_check(hipGetLastError(), "gwf_fisher_arms launch");
}

int gwf_snr(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params, int64_t ld,
            int64_t n_ev, int32_t res, double* snrsq_out, void* workspace, size_t ws_bytes, void* stream) {
  int rc = gwf_prepare(wf, dets, n_det, params, ld, n_ev, res, workspace, ws_bytes, stream);
  if (rc) return rc;
  return gwf_snr_prepared(wf, dets, n_det, n_ev, res, snrsq_out, workspace, ws_bytes, stream);
}

int gwf_fisher(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params, int64_t ld,
               int64_t n_ev, int32_t res, int32_t arm, double* fisher_out, void* workspace, size_t ws_bytes,
               void* stream) {
  int rc = gwf_prepare(wf, dets, n_det, params, ld, n_ev, res, workspace, ws_bytes, stream);
  if (rc) return rc;
  return gwf_fisher_prepared(wf, dets, n_det, n_ev, res, arm, fisher_out, workspace, ws_bytes, stream);
}

int gwf_fisher_noalloc(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params,
                       int64_t ld, int64_t n_ev, int32_t res, int32_t arm, double* fisher_out, void* stream) {
  if (n_ev > GWF_NOALLOC_MAX_EVENTS)
    return fail(-EINVAL, "gwf_fisher_noalloc takes at most " + std::to_string(GWF_NOALLOC_MAX_EVENTS) +
                             " events per call (gwf_fisher with a workspace takes more)");
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc) return rc;
  return gwf_fisher(wf, dets, n_det, params, ld, n_ev, res, arm, fisher_out, ws, NOALLOC_BYTES, stream);
}

int gwf_snr_ragged(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params, int64_t ld,
                   int64_t n_ev, const int32_t* res_ev, double* snrsq_out, void* workspace, size_t ws_bytes,
                   void* stream) {
  if (n_ev > 0 && !res_ev) return fail(-EINVAL, "res_ev is NULL");
  int rc = gwf_prepare(wf, dets, n_det, params, ld, n_ev, 2, workspace, ws_bytes, stream);
  if (rc) return rc;
  return snr_prepared_impl(wf, dets, n_det, n_ev, 2, res_ev, snrsq_out, workspace, ws_bytes, stream);
}

int gwf_fisher_ragged(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params, int64_t ld,
                      int64_t n_ev, const int32_t* res_ev, int32_t arm, double* fisher_out, void* workspace,
                      size_t ws_bytes, void* stream) {
  if (n_ev > 0 && !res_ev) return fail(-EINVAL, "res_ev is NULL");
  int rc = gwf_prepare(wf, dets, n_det, params, ld, n_ev, 2, workspace, ws_bytes, stream);
  if (rc) return rc;
  if (ws_bytes < gwf_workspace_bytes(n_ev)) return fail(-EINVAL, "workspace too small (see gwf_workspace_bytes)");
  NetK net;
  rc = build_net_checked(dets, n_det, 2, net);
  if (rc) return rc;
  if (arm >= net.n_arms) return fail(-EINVAL, "arm index out of range");
  if (n_ev == 0) return 0;
  if (!fisher_out) return fail(-EINVAL, "fisher_out is NULL");
  hipStream_t st = (hipStream_t)stream;
  const double* rec = (const double*)workspace;
  unsigned int* counter = (unsigned int*)((char*)workspace + (size_t)n_ev * REC_DOUBLES * sizeof(double));
  if (GWF_RAGGED_PQ && arm < 0) {  // network sum: pairs on the event queue
    rc = net.lin ? launch_fisher_pq<true, true>(wf, st, net, rec, n_ev, res_ev, counter, fisher_out, 7 + nt_of(wf))
                 : launch_fisher_pq<false, true>(wf, st, net, rec, n_ev, res_ev, counter, fisher_out, 7 + nt_of(wf));
    if (rc) return rc;
    return hip_check(hipGetLastError(), "gwf_fisher_ragged launch");
  }
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, counter, 1);
  rc = hip_check(hipGetLastError(), "counter reset launch");
  if (rc) return rc;
  // persistent grid: enough waves to fill every SIMD several times, never more than the events
  int64_t blocks = std::min<int64_t>((n_ev + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK, 256 * 8);
  const dim3 grid((unsigned)blocks), block(BLOCK);
  if (net.lin)
    launch_fisher_ragged<true>(wf, grid, block, st, net, rec, n_ev, res_ev, arm, counter, fisher_out);
  else
    launch_fisher_ragged<false>(wf, grid, block, st, net, rec, n_ev, res_ev, arm, counter, fisher_out);
  return hip_check(hipGetLastError(), "gwf_fisher_ragged launch");
}

static int strain_impl(const gwf_waveform* wf, const gwf_detector* det, int32_t arm, const double* params, int64_t ld,
                       int64_t n_ev, const double* f, int64_t n_f, double* h_out, double* J_out, double* comp_out,
                       void* workspace, size_t ws_bytes, void* stream) {
  int rc = check_common(wf, params, ld, n_ev, workspace, ws_bytes);
  if (rc) return rc;
  NetK net;
  rc = build_net_checked(det, 1, 2, net);
  if (rc) return rc;
  if (arm < 0 || arm >= net.n_arms) return fail(-EINVAL, "arm index out of range");
  if (n_ev == 0 || n_f == 0) return 0;
  if (!f) return fail(-EINVAL, "frequency grid is NULL");
  hipStream_t st = (hipStream_t)stream;
  rc = launch_setup(wf, net, params, ld, n_ev, f, n_f, (double*)workspace, st);
  if (rc) return rc;
  const dim3 grid((unsigned)((n_ev + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK)), block(BLOCK);
  const double* rec = (const double*)workspace;
  if (wf->model == GWF_MODEL_IMRPHENOMD)
    hipLaunchKernelGGL((k_strain<GWF_MODEL_IMRPHENOMD, 4, 11>), grid, block, 0, st, *wf, net, rec, n_ev, (int)arm, f,
                       n_f, h_out, J_out, comp_out);
  else if (wf->model == GWF_MODEL_IMRPHENOMD_NRTIDALV2)
    hipLaunchKernelGGL((k_strain<GWF_MODEL_IMRPHENOMD_NRTIDALV2, 6, 13>), grid, block, 0, st, *wf, net, rec, n_ev,
                       (int)arm, f, n_f, h_out, J_out, comp_out);
  else if (nt_of(wf) == 5)
    hipLaunchKernelGGL((k_strain<GWF_MODEL_TAYLORF2, 5, 12>), grid, block, 0, st, *wf, net, rec, n_ev, (int)arm, f, n_f,
                       h_out, J_out, comp_out);
  else if (nt_of(wf) == 7)  // tidal and eccentric TaylorF2, 14 parameters
    hipLaunchKernelGGL((k_strain<GWF_MODEL_TAYLORF2, 7, 14>), grid, block, 0, st, *wf, net, rec, n_ev, (int)arm, f, n_f,
                       h_out, J_out, comp_out);
  else if (nt_of(wf) == 6)
    hipLaunchKernelGGL((k_strain<GWF_MODEL_TAYLORF2, 6, 13>), grid, block, 0, st, *wf, net, rec, n_ev, (int)arm, f, n_f,
                       h_out, J_out, comp_out);
  else
    hipLaunchKernelGGL((k_strain<GWF_MODEL_TAYLORF2, 4, 11>), grid, block, 0, st, *wf, net, rec, n_ev, (int)arm, f, n_f,
                       h_out, J_out, comp_out);
  return hip_check(hipGetLastError(), "gwf_strain launch");
}

int gwf_strain(const gwf_waveform* wf, const gwf_detector* det, int32_t arm, const double* params, int64_t ld,
               int64_t n_ev, const double* f, int64_t n_f, double* h_out, double* J_out, void* workspace,
               size_t ws_bytes, void* stream) {
  return strain_impl(wf, det, arm, params, ld, n_ev, f, n_f, h_out, J_out, nullptr, workspace, ws_bytes, stream);
}

int gwf_strain_components(const gwf_waveform* wf, const gwf_detector* det, int32_t arm, const double* params,
                          int64_t ld, int64_t n_ev, const double* f, int64_t n_f, double* comp_out, void* workspace,
                          size_t ws_bytes, void* stream) {
  if (!comp_out && n_ev > 0 && n_f > 0) return fail(-EINVAL, "comp_out is NULL");
  return strain_impl(wf, det, arm, params, ld, n_ev, f, n_f, nullptr, nullptr, comp_out, workspace, ws_bytes, stream);
}

int gwf_strain_vjp(const double* J, const double* gh, int32_t P, int64_t n_ev, int64_t n_f, double tcoal_scale,
                   double* g_out, void* stream) {
  if (P < 1 || P > 16 || n_ev < 0 || n_f < 0) return fail(-EINVAL, "invalid sizes");
  if ((int64_t)P * n_ev == 0) return 0;
  if (!J || !gh || !g_out) return fail(-EINVAL, "NULL array");
  if ((int64_t)P * n_ev > (int64_t)0x7fffffff / WAVES_PER_BLOCK) return fail(-EINVAL, "too many events for one call");
  const int64_t waves = (int64_t)P * n_ev;
  hipLaunchKernelGGL(k_vjp, dim3((unsigned)((waves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK)), dim3(BLOCK), 0,
                     (hipStream_t)stream, J, gh, (int)P, n_ev, n_f, tcoal_scale, g_out);
  return hip_check(hipGetLastError(), "gwf_strain_vjp launch");
}

int gwf_gram(const double* J, const double* f, const double* S, int32_t P, int64_t n_ev, int64_t n_f,
             int32_t accumulate, double* fisher_out, void* stream) {
  if (P < 11 || P > 14) return fail(-EINVAL, "gwf_gram supports P = 11 to 14");
  if (n_ev < 0 || n_f < 0) return fail(-EINVAL, "negative sizes");
  if (n_ev == 0) return 0;
  if (!J || !f || !S || !fisher_out) return fail(-EINVAL, "NULL array");
  if (n_ev > (int64_t)0x7fffffff / WAVES_PER_BLOCK) return fail(-EINVAL, "too many events for one call");
  const dim3 grid((unsigned)((n_ev + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK)), block(BLOCK);
  hipStream_t st = (hipStream_t)stream;
  if (P == 11)
    hipLaunchKernelGGL((k_gram<11>), grid, block, 0, st, J, f, S, n_ev, n_f, (int)accumulate, fisher_out);
  else if (P == 12)
    hipLaunchKernelGGL((k_gram<12>), grid, block, 0, st, J, f, S, n_ev, n_f, (int)accumulate, fisher_out);
  else if (P == 13)
    hipLaunchKernelGGL((k_gram<13>), grid, block, 0, st, J, f, S, n_ev, n_f, (int)accumulate, fisher_out);
  else
    hipLaunchKernelGGL((k_gram<14>), grid, block, 0, st, J, f, S, n_ev, n_f, (int)accumulate, fisher_out);
  return hip_check(hipGetLastError(), "gwf_gram launch");
}

int gwf_gram_split(const double* J, const double* f, const double* S, int32_t P, int64_t n_ev, int64_t n_f,
                   int32_t accumulate, double* fisher_out, void* stream) {
  if (P < 11 || P > 14) return fail(-EINVAL, "gwf_gram_split supports P = 11 to 14");
  if (n_ev < 0 || n_f < 0) return fail(-EINVAL, "negative sizes");
  if (n_ev == 0) return 0;
  if (!J || !f || !S || !fisher_out) return fail(-EINVAL, "NULL array");
  if (n_ev > (int64_t)0x7fffffff / WAVES_PER_BLOCK) return fail(-EINVAL, "too many events for one call");
  const dim3 grid((unsigned)((n_ev + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK)), block(BLOCK);
  hipStream_t st = (hipStream_t)stream;
  if (P == 11)
    hipLaunchKernelGGL((k_gram_split<11>), grid, block, 0, st, J, f, S, n_ev, n_f, (int)accumulate, fisher_out);
  else if (P == 12)
    hipLaunchKernelGGL((k_gram_split<12>), grid, block, 0, st, J, f, S, n_ev, n_f, (int)accumulate, fisher_out);
  else if (P == 13)
    hipLaunchKernelGGL((k_gram_split<13>), grid, block, 0, st, J, f, S, n_ev, n_f, (int)accumulate, fisher_out);
  else
    hipLaunchKernelGGL((k_gram_split<14>), grid, block, 0, st, J, f, S, n_ev, n_f, (int)accumulate, fisher_out);
  return hip_check(hipGetLastError(), "gwf_gram_split launch");
}

int gwf_cov_invert_ex(const double* F, int32_t P, int64_t n_ev, int32_t inv_method, int32_t alt_method,
                      int32_t truncate, double svals_thresh, double cond_numb_max, double* cov_out, double* eps_out,
                      double* cond_out, int32_t* method_out, void* stream) {
  if (P < 1 || P > PMAX) return fail(-EINVAL, "P must be 1..16");
  if (n_ev < 0) return fail(-EINVAL, "n_ev must be >= 0");
  if (inv_method < GWF_INV_CHO || inv_method > GWF_INV_LU || alt_method < GWF_INV_CHO || alt_method > GWF_INV_LU)
    return fail(-EINVAL, "unknown inversion method");
  if (inv_method == GWF_INV_CHO && alt_method == GWF_INV_CHO)
    return fail(-EINVAL, "alt_method has to be different from invMethodIn");
  if (n_ev == 0) return 0;
  if (!F || !cov_out) return fail(-EINVAL, "NULL array");
  hipStream_t st = (hipStream_t)stream;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(xcd_grid(n_ev)), dim3(64), 0, st, F, (int)P, n_ev, (int)inv_method,
                       (int)alt_method, (int)(truncate != 0), svals_thresh, cond_numb_max, cov_out, eps_out, cond_out,
                       (int*)method_out);
  };
  // 'cho' at the path's sizes: one event per lane (k_cov_lane), the events
  // it leaves (alternative method, NaN outputs) by the wavefront kernel.
  // The lane kernel reads F after writing C, so it needs them apart.
  const char *f0 = (const char*)F, *f1 = f0 + (size_t)P * P * n_ev * sizeof(double);
  const char *c0 = (const char*)cov_out, *c1 = c0 + (size_t)P * P * n_ev * sizeof(double);
  const bool lane_path = inv_method == GWF_INV_CHO && P >= 11 && P <= 13 && (f1 <= c0 || c1 <= f0);
  auto lane = [&](auto klane) {
    hipLaunchKernelGGL(klane, dim3((unsigned)((n_ev + 63) / 64)), dim3(64), 0, st, F, n_ev, cov_out, eps_out,
                       cond_out, (int*)method_out);
  };
  if (lane_path) {
    switch (P) {
      case 11: lane(k_cov_lane<11>); go(k_cov<11, true>); break;
      case 12: lane(k_cov_lane<12>); go(k_cov<12, true>); break;
      default: lane(k_cov_lane<13>); go(k_cov<13, true>);
    }
  } else {
    switch (P) {  // the path's sizes at compile time
      case 11: go(k_cov<11>); break;
      case 12: go(k_cov<12>); break;
      case 13: go(k_cov<13>); break;
      case 14: go(k_cov<14>); break;
      default: go(k_cov<0>);
    }
  }
  return hip_check(hipGetLastError(), "gwf_cov_invert launch");
}

int gwf_cov_invert(const double* F, int32_t P, int64_t n_ev, double* cov_out, double* eps_out, double* cond_out,
                   int32_t* method_out, void* stream) {
  return gwf_cov_invert_ex(F, P, n_ev, GWF_INV_CHO, GWF_INV_SVD, 0, 1e-15, 1e50, cov_out, eps_out, cond_out,
                           method_out, stream);
}

int gwf_inversion_error(const double* F, const double* C, int32_t P, int64_t n_ev, double* eps_out, void* stream) {
  if (P < 1 || P > PMAX) return fail(-EINVAL, "P must be 1..16");
  if (n_ev < 0) return fail(-EINVAL, "n_ev must be >= 0");
  if (n_ev == 0) return 0;
  if (!F || !C || !eps_out) return fail(-EINVAL, "NULL array");
  hipLaunchKernelGGL(k_inv_err, dim3((unsigned)((n_ev + 255) / 256)), dim3(256), 0, (hipStream_t)stream, F, C, (int)P,
                     n_ev, eps_out);
  return hip_check(hipGetLastError(), "gwf_inversion_error launch");
}

int gwf_check_fisher(const double* F, int32_t P, int64_t n_ev, double* evals_out, double* evecs_out, double* cond_out,
                     void* stream) {
  if (P < 1 || P > PMAX) return fail(-EINVAL, "P must be 1..16");
  if (n_ev < 0) return fail(-EINVAL, "n_ev must be >= 0");
  if (n_ev == 0) return 0;
  if (!F || !evals_out || !cond_out) return fail(-EINVAL, "NULL array");
  hipStream_t st = (hipStream_t)stream;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(xcd_grid(n_ev)), dim3(64), 0, st, F, (int)P, n_ev, evals_out, evecs_out,
                       cond_out);
  };
  // P = 11, 12: one event per lane (k_check_lane), the events with a NaN
  // entry by the wavefront kernel; the lane kernel needs F apart from the
  // outputs
  auto apart = [&](const void* p, size_t bytes) {
    const char *a0 = (const char*)F, *a1 = a0 + (size_t)P * P * n_ev * sizeof(double);
    const char *b0 = (const char*)p, *b1 = b0 + bytes;
    return !p || a1 <= b0 || b1 <= a0;
  };
  const size_t nb = (size_t)n_ev * sizeof(double);
  const bool lane_path = (P == 11 || P == 12) && apart(evals_out, nb * P) && apart(evecs_out, nb * P * P) &&
                         apart(cond_out, nb);
  auto lane = [&](auto klane) {
    hipLaunchKernelGGL(klane, dim3((unsigned)((n_ev + 63) / 64)), dim3(64), 0, st, F, n_ev, evals_out, evecs_out,
                       cond_out);
  };
  if (lane_path) {
    if (P == 11) {
      if (evecs_out) lane(k_check_lane<11, true>);
      else lane(k_check_lane<11, false>);
      go(k_check<11, true>);
    } else {
      if (evecs_out) lane(k_check_lane<12, true>);
      else lane(k_check_lane<12, false>);
      go(k_check<12, true>);
    }
  } else {
    switch (P) {
      case 11: go(k_check<11>); break;
      case 12: go(k_check<12>); break;
      case 13: go(k_check<13>); break;
      case 14: go(k_check<14>); break;
      default: go(k_check<0>);
    }
  }
  return hip_check(hipGetLastError(), "gwf_check_fisher launch");
}

int gwf_fix_params(const double* F, int32_t P, int64_t n_ev, uint32_t fix_mask, const double* prior_diag,
                   double* out, void* stream) {
  if (P < 1 || P > PMAX) return fail(-EINVAL, "P must be 1..16");
  if (n_ev < 0) return fail(-EINVAL, "n_ev must be >= 0");
  if (fix_mask >> P) return fail(-EINVAL, "fix_mask names a parameter >= P");
  IdxMap map;
  int Pn = 0;
  for (int a = 0; a < P; ++a)
    if (!((fix_mask >> a) & 1u)) map.idx[Pn++] = a;
  for (int a = Pn; a < PMAX; ++a) map.idx[a] = 0;
  if (Pn == 0) return fail(-EINVAL, "every parameter fixed");
  if (n_ev == 0) return 0;
  if (!F || !out) return fail(-EINVAL, "NULL array");
  const int64_t tot = (int64_t)Pn * Pn * n_ev;
  hipLaunchKernelGGL(k_sub, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, (hipStream_t)stream, F, (int)P, n_ev,
                     map, Pn, prior_diag, out);
  return hip_check(hipGetLastError(), "gwf_fix_params launch");
}

int gwf_sigma_sky_area(const double* C, int32_t P, int64_t n_ev, const double* theta, int32_t i_theta, int32_t i_phi,
                       double perc_level, int32_t sq_deg, double* sigma_out, double* area_out, void* stream) {
  if (P < 1 || P > PMAX) return fail(-EINVAL, "P must be 1..16");
  if (n_ev < 0) return fail(-EINVAL, "n_ev must be >= 0");
  if (!sigma_out && !area_out) return fail(-EINVAL, "no output requested");
  if (area_out && (!theta || i_theta < 0 || i_theta >= P || i_phi < 0 || i_phi >= P))
    return fail(-EINVAL, "sky area needs theta and the theta / phi indices");
  if (area_out && !(perc_level > 0.0 && perc_level < 100.0)) return fail(-EINVAL, "perc_level must be in (0, 100)");
  if (n_ev == 0) return 0;
  if (!C) return fail(-EINVAL, "NULL array");
  double scale = area_out ? -log(1.0 - perc_level / 100.0) : 0.0;
  if (sq_deg) scale *= (180.0 / M_PI) * (180.0 / M_PI);
  hipLaunchKernelGGL(k_sigma_area, dim3((unsigned)((n_ev + 255) / 256)), dim3(256), 0, (hipStream_t)stream, C, (int)P,
                     n_ev, theta, (int)i_theta, (int)i_phi, scale, sigma_out, area_out);
  return hip_check(hipGetLastError(), "gwf_sigma_sky_area launch");
}

int gwf_fisher_block_transform(const double* F, double* out, int32_t P, int64_t n_ev, int32_t i0, int32_t i1,
                               const double* L, const double* R, void* stream) {
  if (P < 1 || P > 64) return fail(-EINVAL, "P must be in [1, 64]");
  if (n_ev < 0) return fail(-EINVAL, "invalid sizes");
  if (i0 < 0 || i0 >= P || i1 >= P || i1 == i0) return fail(-EINVAL, "block indices out of range");
  if (n_ev == 0) return 0;
  if (!F || !out || !L || !R) return fail(-EINVAL, "NULL array");
  hipLaunchKernelGGL(k_block_transform, dim3((unsigned)((n_ev + 255) / 256)), dim3(256), 0, (hipStream_t)stream, F, out,
                     (int)P, n_ev, (int)i0, (int)(i1 < 0 ? -1 : i1), L, R);
  return hip_check(hipGetLastError(), "gwf_fisher_block_transform launch");
}

int gwf_fisher_reparam(const double* params, int64_t ld, double* F, int32_t P, int64_t n_ev, int32_t use_m1m2,
                       int32_t use_chiSA, void* stream) {
  if (P < 11 || P > 14) return fail(-EINVAL, "P must be 11 to 14");
  if (n_ev < 0 || ld < n_ev) return fail(-EINVAL, "invalid sizes");
  if (n_ev == 0 || (!use_m1m2 && !use_chiSA)) return 0;
  if (!params || !F) return fail(-EINVAL, "NULL array");
  hipLaunchKernelGGL(k_reparam, dim3((unsigned)((n_ev + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params, ld, F,
                     (int)P, n_ev, (int)use_m1m2, (int)use_chiSA);
  return hip_check(hipGetLastError(), "gwf_fisher_reparam launch");
}

}  // extern "C"
