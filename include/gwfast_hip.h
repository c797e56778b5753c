/*
 * gwfast_hip.h -- C ABI of libgwfast_hip.so, the MI355X (gfx950) kernels of
 * gwfast's event-batched waveform -> detector projection -> SNR / Fisher ->
 * covariance path.
 *
 * gwfast v1.1.2 is pure Python (JAX on CPU) and has no FFI of its own; its
 * plug-in point for this path is the WaveFormModel ABC consumed by GWSignal
 * (gwfast/waveforms.py:49-199, gwfast/signal.py:68-160).  Each entry point
 * below replaces one reference operation, cited at its declaration.  The
 * Python layer (gwfast_amd.waveforms / signal / network / fisherTools) keeps
 * the reference's class and method signatures and calls these through
 * ctypes.
 *
 * Conventions
 *   - All array arguments are DEVICE pointers (caller-owned, e.g. torch
 *     tensors); structs are passed by host pointer and copied by value.
 *   - Event parameters are FP64 SoA rows of length >= n_ev with row stride
 *     `ld` (elements), rows in gwfast's Fisher order (waveforms.py:78, 106-111,
 *     125-129):
 *        0 Mc [Msun, detector frame]  1 eta  2 dL [Gpc]  3 theta  4 phi
 *        5 iota  6 psi  7 tcoal [GMST days]  8 Phicoal  9 chi1z  10 chi2z
 *       11 LambdaTilde  12 deltaLambda            (rows 11-12 tidal only)
 *   - Fisher/covariance outputs use the reference layout (P, P, n_ev):
 *     element (a, b) of event e at [(a * P + b) * n_ev + e].
 *   - `stream` is a hipStream_t; every call is stream-ordered and returns
 *     before the work completes.  Calls never allocate device memory.
 *   - Return 0 on success, a negative errno-style code on failure; the
 *     message is available from gwf_last_error() (thread-local).
 */
#ifndef GWFAST_HIP_H
#define GWFAST_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GWF_ABI_VERSION 1

enum { GWF_MODEL_TAYLORF2 = 0, GWF_MODEL_IMRPHENOMD = 1, GWF_MODEL_IMRPHENOMD_NRTIDALV2 = 2 };
enum { GWF_SHAPE_L = 0, GWF_SHAPE_T = 1 };
#define GWF_MAX_DETECTORS 8
#define GWF_MAX_ARMS 24

/* Waveform model description: TaylorF2_RestrictedPN(...) (waveforms.py:726),
 * IMRPhenomD(fRef=None) (waveforms.py:973) or IMRPhenomD_NRTidalv2(fRef=None)
 * (waveforms.py:1339-1837; is_tidal = 1, QNM tables as IMRPhenomD). */
typedef struct gwf_waveform {
  int32_t model;            /* GWF_MODEL_* */
  int32_t is_tidal;         /* TaylorF2 is_tidal: 13 parameters */
  int32_t use_3p5PN_SpinHO; /* TaylorF2 */
  int32_t phiref_vlso;      /* TaylorF2 */
  int32_t kerr_isco;        /* TaylorF2 which_ISCO == 'Kerr' */
  int32_t use_quadmon;      /* TaylorF2 use_QuadMonTid */
  int32_t apply_fcut;       /* WaveFormModel apply_fcut (IMRPhenomD) */
  int32_t n_qnm;            /* IMRPhenomD QNM table length */
  double fcut_par;          /* TaylorF2 fHigh [Hz Msun]; IMRPhenomD 0.2 */
  double fref;              /* IMRPhenomD fixed fRef [Hz]; <= 0: grid minimum */
  const double* qnm_a;      /* device, IMRPhenomD QNMData_a */
  const double* qnm_fring;  /* device, QNMData_fring */
  const double* qnm_fdamp;  /* device, QNMData_fdamp */
  int32_t is_eccentric;     /* TaylorF2 is_eccentric (non-tidal): 12 parameters, row 11 = ecc */
  int32_t reserved;
  double fref_ecc;          /* TaylorF2 fRef_ecc [Hz]; <= 0: the grid minimum */
} gwf_waveform;

/* One detector site: GWSignal(...) (signal.py:68-82). */
typedef struct gwf_detector {
  int32_t shape;        /* GWF_SHAPE_L or GWF_SHAPE_T */
  int32_t earth_motion; /* useEarthMotion */
  int32_t no_motion;    /* noMotion */
  int32_t compute2arms; /* T only: third arm = -(arm1 + arm2) */
  double lat_deg, long_deg, xax_deg;
  double fmin;          /* Hz */
  double fmax;          /* Hz; <= 0: none */
  const double* psd_f;  /* device, strictly increasing frequencies [Hz] */
  const double* psd_S;  /* device, PSD (ASD^2 when the file is an ASD) */
  int32_t n_psd;
  int32_t lin_spacing;  /* Fisher grids (gwf_fisher*, gwf_grid): 0 geomspace, 1 linspace
                           (FisherMatr spacing, signal.py:893-896); SNR grids are always
                           geometric (signal.py:720-721) */
  double psd_f_first;   /* host copy of psd_f[0] */
  double psd_f_last;    /* host copy of psd_f[n_psd - 1] */
} gwf_detector;

int gwf_abi_version(void);
const char* gwf_last_error(void);

/* Bytes of device workspace the calls below need for n_ev events
 * (per-event coefficient records, then 256 bytes of queue counters). */
size_t gwf_workspace_bytes(int64_t n_ev);

/* Number of interferometer arms of a network: L -> 1, T -> 3. */
int gwf_num_arms(const gwf_detector* dets, int32_t n_det);

/* WaveFormModel.fcut(**ev) (waveforms.py:190-199, 903-953, 1324-1333),
 * clipped to fmax when fmax > 0 (signal.py:715-718). fcut_out: (n_ev,). */
int gwf_fcut(const gwf_waveform* wf, const double* params, int64_t ld, int64_t n_ev, double fmax,
             double* fcut_out, void* stream);

/* Frequency grid and PSD on it, as GWSignal.SNRInteg / FisherMatr build them
 * (signal.py:715-723, 884-901): jnp.geomspace(fmin, fcut, res) and
 * jnp.interp(f, psd_f, psd_S, left=1, right=1).  f_out, S_out: (res, n_ev);
 * S_out may be NULL. */
int gwf_grid(const gwf_waveform* wf, const gwf_detector* det, const double* params, int64_t ld, int64_t n_ev,
             int32_t res, double* f_out, double* S_out, void* workspace, size_t ws_bytes, void* stream);

/* WaveFormModel.Ampl / Phi / tau_star on a caller grid f (n_f, n_ev)
 * (waveforms.py:148-188).  Any output pointer may be NULL.  IMRPhenomD's
 * fRef is the minimum of each event's grid column (waveforms.py:1139). */
int gwf_waveform_eval(const gwf_waveform* wf, const double* params, int64_t ld, int64_t n_ev, const double* f,
                      int64_t n_f, double* ampl_out, double* phi_out, double* tau_out, void* workspace,
                      size_t ws_bytes, void* stream);

/* GWSignal.SNRInteg for every arm of a network (signal.py:658-777; network
 * sum in network.py:53-81 is left to the caller).  snrsq_out: (n_arms, n_ev)
 * holding trapz((Ap^2 + Ac^2)/S, f), i.e. SNR_arm = 2 sqrt(snrsq). */
int gwf_snr(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params, int64_t ld,
            int64_t n_ev, int32_t res, double* snrsq_out, void* workspace, size_t ws_bytes, void* stream);

/* GWSignal.FisherMatr summed over arms and detectors, i.e. DetNet.FisherMatr
 * (signal.py:782-1099, network.py:84-152), with derivatives in ParNums order
 * and the tcoal row in seconds (signal.py:919-920).  Fused: strain tangents
 * and the trapezoid Gram are computed in one pass and never written to
 * memory.  arm < 0 sums all arms; arm = k returns arm k only (network arm
 * order: detectors in order, T detectors contributing 3 arms).
 * fisher_out: (P, P, n_ev), P = 11 (13 when the model is tidal). */
int gwf_fisher(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params, int64_t ld,
               int64_t n_ev, int32_t res, int32_t arm, double* fisher_out, void* workspace, size_t ws_bytes,
               void* stream);

/* gwf_fisher without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * event records and queue counters live in a file-scope device array of the
 * library, sized for GWF_NOALLOC_MAX_EVENTS events (about 46.3 MB, held by
 * each GPU the library runs on until the process ends); larger n_ev returns
 * -EINVAL.  The call allocates nothing; the caller passes the inputs and
 * fisher_out as for gwf_fisher, and gets the same results.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (two
 * streams, two threads) race, so serialise them. */
#define GWF_NOALLOC_MAX_EVENTS 8192
int gwf_fisher_noalloc(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params,
                       int64_t ld, int64_t n_ev, int32_t res, int32_t arm, double* fisher_out, void* stream);

/* DetNet.FisherMatr with the duty-cycle masks of GWSignal(DutyFactor=...)
 * (signal.py:932-936, 968-969, 1004-1005, 1038-1039, 1071-1072; network.py:
 * 84-152) and of the catalogue driver (calculate_forecasts_from_catalog.py:
 * 423-447, 527-556): one fused pass summing, per event, the Fisher of the
 * physical arms whose bit is set in arm_mask[e] (bit a = network arm a, in
 * the arm order of gwf_fisher).  arm_mask: (n_ev,) device uint32. */
int gwf_fisher_masked(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params,
                      int64_t ld, int64_t n_ev, int32_t res, const uint32_t* arm_mask, double* fisher_out,
                      void* workspace, size_t ws_bytes, void* stream);

/* No reference counterpart; for tests of a failure path.  Each wait of the
 * producer / consumer Fisher kernels (gwf_fisher*, network sums and
 * gwf_fisher_arms) polls at most this many times (0 restores the default,
 * 2^22 polls, >= 0.1 s); a wait that runs out ends its wave and the launch
 * writes NaN over its whole output instead of hanging.  Process-wide. */
int gwf_set_queue_spin_limit(uint32_t polls);

/* GWSignal.FisherMatr / DetNet.FisherMatr with return_all=True
 * (signal.py:1087-1098, network.py:125-135; the catalogue driver always asks
 * for it, calculate_forecasts_from_catalog.py:507-513): the Fisher matrix of
 * every physical arm of the network from one launch (the waveform and its
 * tangents are evaluated once per frequency for all arms).  fisher_out:
 * (n_arms, P, P, n_ev), arm a at fisher_out + a * P * P * n_ev, arms in the
 * order of gwf_num_arms (detectors in order; a triangle's three arms, the
 * third -(first + second)).  At most 8 arms.  res_ev: NULL for one grid
 * length res, else per-event lengths as in gwf_fisher_ragged (res unused). */
int gwf_fisher_arms(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params,
                    int64_t ld, int64_t n_ev, int32_t res, const int32_t* res_ev, double* fisher_out,
                    void* workspace, size_t ws_bytes, void* stream);

/* Ragged per-event grids (no reference counterpart: gwfast uses one res for
 * all events, signal.py:890-898): event e uses jnp.geomspace(fmin, fcut_e,
 * res_ev[e]) with res_ev a DEVICE int32 array (values are clamped to [2, 2^30]).
 * gwf_fisher_ragged runs one persistent launch whose wavefronts pull events
 * from an atomic counter kept in the workspace, so uneven grids balance
 * across the GPU.  Outputs as gwf_snr / gwf_fisher. */
int gwf_snr_ragged(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params, int64_t ld,
                   int64_t n_ev, const int32_t* res_ev, double* snrsq_out, void* workspace, size_t ws_bytes,
                   void* stream);
int gwf_fisher_ragged(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params, int64_t ld,
                      int64_t n_ev, const int32_t* res_ev, int32_t arm, double* fisher_out, void* workspace,
                      size_t ws_bytes, void* stream);

/* The two halves of gwf_snr / gwf_fisher, for callers that reuse the
 * per-event records (e.g. SNR then Fisher of the same events) or time the
 * integral kernels alone.  gwf_prepare fills the workspace with the event
 * records of a (waveform, network, res) triple; the *_prepared calls must be
 * given the same triple and workspace. */
int gwf_prepare(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, const double* params, int64_t ld,
                int64_t n_ev, int32_t res, void* workspace, size_t ws_bytes, void* stream);
int gwf_snr_prepared(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, int64_t n_ev, int32_t res,
                     double* snrsq_out, const void* workspace, size_t ws_bytes, void* stream);
/* gwf_fisher_prepared also keeps two queue counters after the records (the
 * network sum's producer / consumer launch), so its workspace is writable and
 * at least gwf_workspace_bytes(n_ev). */
int gwf_fisher_prepared(const gwf_waveform* wf, const gwf_detector* dets, int32_t n_det, int64_t n_ev, int32_t res,
                        int32_t arm, double* fisher_out, void* workspace, size_t ws_bytes, void* stream);

/* GWSignal.GWstrain and the derivative rows of _SignalDerivatives for one arm
 * of one detector on the caller grid f (n_f, n_ev) (signal.py:485-655,
 * 1102-1301): h_out (n_f, n_ev) complex128 (interleaved re, im) and
 * J_out (P, n_ev, n_f) complex128 with the tcoal row divided by 86400, as
 * FisherMatr(return_derivatives=True) returns them.  arm: 0..2 (T) or 0 (L);
 * arm 2 of a compute2arms triangle is -(arm0 + arm1).  Either output may be
 * NULL. */
int gwf_strain(const gwf_waveform* wf, const gwf_detector* det, int32_t arm, const double* params, int64_t ld,
               int64_t n_ev, const double* f, int64_t n_f, double* h_out, double* J_out, void* workspace,
               size_t ws_bytes, void* stream);

/* GWstrain(..., return_single_comp=...) (signal.py:629-652): the plus and
 * cross amplitudes Ap, Ac and the phase Psi = 2 pi f tcoal - Phicoal -
 * Phi_GW + 2 pi f Delta t of one arm on the caller grid f (n_f, n_ev);
 * comp_out (3, n_f, n_ev) = (Ap, Ac, Psi). */
int gwf_strain_components(const gwf_waveform* wf, const gwf_detector* det, int32_t arm, const double* params,
                          int64_t ld, int64_t n_ev, const double* f, int64_t n_f, double* comp_out, void* workspace,
                          size_t ws_bytes, void* stream);

/* The common grid of GWSignal.WFOverlap (signal.py:1852-1863): geomspace
 * (linspace when det->lin_spacing) from det->fmin to fcut[e], and the PSD
 * interpolated on it.  f_out, S_out (res, n_ev). */
int gwf_grid_fcut(const gwf_detector* det, const double* fcut, int64_t n_ev, int32_t res, double* f_out,
                  double* S_out, void* stream);

/* The overlap integrals of GWSignal.WFOverlap for one arm (signal.py:
 * 1870-1920): out (3, n_ev) = 4 trapz(Re(h1 conj h2)/S), 4 trapz(|h1|^2/S),
 * 4 trapz(|h2|^2/S) on the grid f, S (n_f, n_ev), h1, h2 (n_f, n_ev) complex;
 * accumulate != 0 adds to out (sum over arms and detectors). */
int gwf_overlap(const double* h1, const double* h2, const double* f, const double* S, int64_t n_ev, int64_t n_f,
                int32_t accumulate, double* out, void* stream);

/* Vector-Jacobian product of the strain for reverse-mode callers (no
 * reference counterpart; JAX does this inside jacrev, signal.py:1178-1187):
 * g_out[p][e] = sum_f (gh[f][e].re J[p][e][f].re + gh[f][e].im J[p][e][f].im)
 * for J (P, n_ev, n_f) complex128 as written by gwf_strain and gh (n_f, n_ev)
 * complex128.  tcoal_scale multiplies row 7 (86400 gives d/dtcoal in days
 * from the per-second row). g_out: (P, n_ev). */
int gwf_strain_vjp(const double* J, const double* gh, int32_t P, int64_t n_ev, int64_t n_f, double tcoal_scale,
                   double* g_out, void* stream);

/* The reference's plain Fisher integral from materialised derivatives
 * (signal.py:922-931): F_ab = 4 trapz(Re(conj(J_a) J_b)/S, f) for J
 * (P, n_ev, n_f) complex128, f and S (n_f, n_ev).  fisher_out (P, P, n_ev);
 * accumulate != 0 adds into fisher_out. */
int gwf_gram(const double* J, const double* f, const double* S, int32_t P, int64_t n_ev, int64_t n_f,
             int32_t accumulate, double* fisher_out, void* stream);

/* The same integral as gwf_gram (signal.py:922-931) on FP32 MFMA with split
 * operands: x = J sqrt(w/S) is split into fp32 hi + lo, F = hi^T hi +
 * hi^T lo + lo^T hi, each 8-sample fp32 partial added into FP64.  Error
 * bound per element: ~8 * 2^-24 * sqrt(F_aa F_bb). */
int gwf_gram_split(const double* J, const double* f, const double* S, int32_t P, int64_t n_ev, int64_t n_f,
                   int32_t accumulate, double* fisher_out, void* stream);

/* fisherTools.CovMatr(F, invMethodIn='cho', alt_method='svd') and
 * compute_inversion_error (fisherTools.py:32-210): per event, Jacobi scaling
 * ws = diag(1/sqrt(F_ii)), eigenvalues of the scaled matrix, Cholesky
 * inverse when all eigenvalues are >= 0 and the factorisation succeeds,
 * otherwise the eigen-decomposition ('svd') inverse; symmetrised and
 * unscaled.  All-NaN input gives NaN output.
 *   cov_out (P, P, n_ev), eps_out (n_ev,) = max|Cov F - I|,
 *   cond_out (n_ev,) condition number of the scaled matrix,
 *   method_out (n_ev,) 0 cholesky, 1 eigen fallback, -1 NaN.
 * Any output except cov_out may be NULL. P <= 16. */
int gwf_cov_invert(const double* F, int32_t P, int64_t n_ev, double* cov_out, double* eps_out, double* cond_out,
                   int32_t* method_out, void* stream);

/* Inversion methods of fisherTools.CovMatr (invMethodIn / alt_method). */
#define GWF_INV_CHO 0
#define GWF_INV_SVD 1
#define GWF_INV_INV 2
#define GWF_INV_SVD_REG 3
#define GWF_INV_LU 4

/* fisherTools.CovMatr with every option (fisherTools.py:32-196): inv_method
 * and alt_method are GWF_INV_* ids ('cho' switches to alt_method on a
 * negative eigenvalue of the scaled matrix or a Cholesky pivot below 2^-52,
 * as mpmath's cholesky); 'svd' with truncate replaces singular values below
 * svals_thresh * max by that floor when the condition number exceeds
 * cond_numb_max, and eps is then taken against the truncated scaled matrix,
 * as the reference stores it; 'svd_reg' leaves out singular values
 * <= svals_thresh; 'inv' and 'lu' use mpmath's pivoted LU (numerically
 * singular -> NaN), 'lu' composing U^{-1} (P L)^{-1} as the reference does.
 * method_out holds the method used per event (-1: NaN). */
int gwf_cov_invert_ex(const double* F, int32_t P, int64_t n_ev, int32_t inv_method, int32_t alt_method,
                      int32_t truncate, double svals_thresh, double cond_numb_max, double* cov_out, double* eps_out,
                      double* cond_out, int32_t* method_out, void* stream);

/* fisherTools.compute_inversion_error (fisherTools.py:199-210) for given
 * Fisher and covariance (P, P, n_ev): eps_out (n_ev,) = max|C F - I|. */
int gwf_inversion_error(const double* F, const double* C, int32_t P, int64_t n_ev, double* eps_out, void* stream);

/* fisherTools.CheckFisher (fisherTools.py:216-279): eigenvalues ascending
 * evals_out (n_ev, P), eigenvectors evecs_out (n_ev, P, P) (column j is the
 * eigenvector of eigenvalue j), cond_out (n_ev,) = max|l| / min|l|.
 * evecs_out may be NULL. P <= 16. */
int gwf_check_fisher(const double* F, int32_t P, int64_t n_ev, double* evals_out, double* evecs_out, double* cond_out,
                     void* stream);

/* ------------------------------------------------------------------------
 * Learned SNR screen (no reference counterpart: gwfast evaluates every SNR
 * with GWSignal.SNRInteg, signal.py:658-777).  An 11 -> 256 -> 256 -> 256 -> 1
 * SiLU MLP with bf16 weights and activations and fp32 accumulation predicts
 * the network log SNR; events whose prediction falls inside a band around
 * the detection threshold are recomputed exactly.  bf16 arrays are passed
 * as void* (raw 16-bit patterns, round-to-nearest-even).
 * ---------------------------------------------------------------------- */

/* C (M x N, row stride ldc, fp32) = alpha * sum_k A(m, k) B(k, n) + beta * C,
 * A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn] (bf16), on MFMA.
 * split_k > 1 splits K over workgroups and adds the partial products
 * atomically (C pre-scaled by beta). */
int gwf_gemm_bf16(int32_t M, int32_t N, int32_t K, float alpha, const void* A, int64_t sam, int64_t sak,
                  const void* B, int64_t sbk, int64_t sbn, float beta, float* C, int64_t ldc, int32_t split_k,
                  void* stream);

/* Z = C + bias (per column), H = bf16(SiLU(Z)); C, Z (M, N) fp32, Z_out may be NULL. */
int gwf_bias_silu(const float* C, const float* bias, int64_t M, int32_t N, float* Z_out, void* H_out, void* stream);

/* dZ = dH * SiLU'(Z) over n elements; either output may be NULL (not both). */
int gwf_silu_backward(const float* dH, const float* Z, int64_t n, float* dZ_out, void* dZ_bf16_out, void* stream);

/* out[n] (+)= sum_m w[m] X[m, n] for X (M, N) given as fp32 X or bf16 X_bf16
 * (exactly one non-NULL); w NULL means weight 1; accumulate = 0 overwrites. */
int gwf_colsum(const float* X, const void* X_bf16, const float* w, int64_t M, int32_t N, float* out,
               int32_t accumulate, void* stream);

/* Output layer: y[m] = sum_k H[m, k] w[k] + b[0], H (M, 256) bf16, w fp32. */
int gwf_head_forward(const void* H, const float* w, const float* b, int64_t M, float* y, void* stream);

/* Output-layer backward: dH = g w^T (M, 256), dw = H^T g, db = sum g. */
int gwf_head_backward(const void* H, const float* w, const float* g, int64_t M, float* dH_out, float* dw_out,
                      float* db_out, void* stream);

/* Mean squared error: g = 2 (y - t) / n and *loss_accum += mean((y - t)^2). */
int gwf_mse_grad(const float* y, const float* t, int64_t n, float* g_out, float* loss_accum, void* stream);

/* Adam step `step` (>= 1, bias-corrected) on fp32 master weights w with
 * moments m, v; w_bf16_out (may be NULL) receives the bf16 copy. */
int gwf_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2, float eps,
             int32_t step, void* w_bf16_out, void* stream);

/* y = bf16(x) over n elements. */
int gwf_to_bf16(const float* x, int64_t n, void* y, void* stream);

/* Screen features of events given as FP64 parameter rows (>= 11 rows of
 * stride ld, Fisher order): log Mc, eta, log dL, cos theta,
 * cos/sin(phi - 2 pi tcoal), cos iota, cos/sin 2psi, chi1z, chi2z.
 * raw_out (n_ev, 11) fp32; X_out (n_ev, 16) bf16 standardised with
 * (f - mean) * istd (mean, istd NULL: unstandardised), columns 11..15 zero. */
int gwf_snr_features(const double* params, int64_t ld, int64_t n_ev, const float* mean, const float* istd,
                     float* raw_out, void* X_out, void* stream);

/* One-launch screen: features + MLP + decision.  W1 (256, 16), W2, W3
 * (256, 256) bf16 row-major (out, in), 16-byte aligned; b1..b3 (256,),
 * w4 (256,), b4 (1,) fp32.  log SNR = y * y_scale + y_shift;
 * class_out: 0 if SNR < snr_lo, 2 if SNR > snr_hi, else 1 (recompute).
 * Either output may be NULL (not both). */
int gwf_snr_screen(const double* params, int64_t ld, int64_t n_ev, const float* mean, const float* istd,
                   const void* W1, const float* b1, const void* W2, const float* b2, const void* W3, const float* b3,
                   const float* w4, const float* b4, float y_scale, float y_shift, float snr_lo, float snr_hi,
                   float* logsnr_out, int8_t* class_out, void* stream);

/* Bytes of workspace gwf_select needs for n events. */
size_t gwf_select_workspace_bytes(int64_t n);

/* Indices (ascending) of the events with cls[i] == want -> idx_out (room for
 * n), their number -> count_out[0] (device). */
int gwf_select(const int8_t* cls, int64_t n, int32_t want, int64_t* idx_out, int64_t* count_out, void* workspace,
               size_t workspace_bytes, void* stream);

/* out[r * ld_out + j] = params[r * ld + idx[j]] for r < n_rows, j < n_idx;
 * an index outside [0, ld) gives a NaN column. */
int gwf_gather_events(const double* params, int64_t ld, int32_t n_rows, const int64_t* idx, int64_t n_idx,
                      double* out, int64_t ld_out, void* stream);

/* GWSignal.FisherMatr(use_m1m2=True and / or use_chi1chi2=False)
 * (signal.py:818-845): rows / columns 0, 1 of F (P, P, n_ev) turned from
 * (Mc, eta) into the detector-frame (m1, m2), and 9, 10 from (chi1z, chi2z)
 * into (chiS, chiA), in place: F' = J^T F J with the exact Jacobian at each
 * event's params (Fisher-order rows of stride ld). */
int gwf_fisher_reparam(const double* params, int64_t ld, double* F, int32_t P, int64_t n_ev, int32_t use_m1m2,
                       int32_t use_chiSA, void* stream);

/* The matrix variable changes of fisherTools.py:410-459 (log dL <-> dL) and
 * 620-822 (Mc, eta <-> m1, m2; chi1z, chi2z -> chi_eff, Delta chi;
 * chiS, chiA -> chi1z, chi2z), batched over events: out = A F B for
 * F (P, P, n_ev), A and B the identity except on rows / columns (i0, i1),
 * where they hold the per-event blocks L and R, SoA (4, n_ev) in the order
 * [0][0], [0][1], [1][0], [1][1].  i1 = -1 scales row / column i0 alone by
 * L[0] / R[0].  out may be F. */
int gwf_fisher_block_transform(const double* F, double* out, int32_t P, int64_t n_ev, int32_t i0, int32_t i1,
                               const double* L, const double* R, void* stream);

/* ------------------------------------------------------------------------
 * Post-Fisher batch (the catalogue driver's consumers of CovMatr,
 * calculate_forecasts_from_catalog.py:573-630, 775).
 * ---------------------------------------------------------------------- */

/* fisherTools.fixParams (fisherTools.py:340-375) and addPrior (378-403):
 * out = F with the rows / columns whose bit is set in fix_mask removed,
 * (Pn, Pn, n_ev), Pn = P - popcount(fix_mask), kept order.  prior_diag
 * (device, (P,), may be NULL) is added to the diagonal first. */
int gwf_fix_params(const double* F, int32_t P, int64_t n_ev, uint32_t fix_mask, const double* prior_diag,
                   double* out, void* stream);

/* sigma_out (P, n_ev) = sqrt(diag Cov) (calculate_forecasts_from_catalog.py:775)
 * and area_out (n_ev,) = compute_localization_region (fisherTools.py:830-858):
 * -ln(1 - perc/100) 2 pi sqrt(C_thth C_phph - C_phth^2) |sin theta|,
 * in deg^2 when sq_deg != 0, else sr.  theta: (n_ev,) device.  Either
 * output may be NULL (not both). */
int gwf_sigma_sky_area(const double* C, int32_t P, int64_t n_ev, const double* theta, int32_t i_theta, int32_t i_phi,
                       double perc_level, int32_t sq_deg, double* sigma_out, double* area_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GWFAST_HIP_H */
