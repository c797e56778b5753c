// Kernels of the learned SNR screen: an 11 -> 256 -> 256 -> 256 -> 1 SiLU MLP
// with bf16 weights/activations and fp32 accumulation that predicts the
// network log SNR of an event from its parameters.  It has no reference
// counterpart: gwfast evaluates every SNR with GWSignal.SNRInteg
// (signal.py:658-777); the screen decides which events the exact SNR kernel
// must still evaluate.
//
// Training building blocks (C ABI in include/gwfast_hip.h):
//   gwf_gemm_bf16      C = alpha op(A) op(B) + beta C on MFMA
//                      (v_mfma_f32_16x16x16_bf16), fp32 accumulation, split-K
//   gwf_bias_silu      Z = C + b, H = bf16(SiLU(Z))
//   gwf_silu_backward  dZ = dH SiLU'(Z)
//   gwf_colsum         bias gradients
//   gwf_head_forward   y = H w + b (the 256 -> 1 output layer)
//   gwf_head_backward  dH = g w^T, dw = H^T g, db = sum g
//   gwf_mse_grad       loss and dL/dy of the mean squared error
//   gwf_adam           Adam on fp32 master weights + bf16 copies
//   gwf_snr_features   event parameters -> standardised features
// Inference: gwf_snr_screen runs features + the whole MLP + the threshold
// decision in one launch (128 events per workgroup, activations in LDS,
// weights from L2, v_mfma_f32_16x16x32_bf16 for the 256-wide layers).
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <string>

#include "gwfast_hip.h"

namespace gwf_mlp {

void set_error(const std::string& s);  // defined in gwfast_hip.hip (shared gwf_last_error)

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int NFEAT = 11;   // features of an event
constexpr int KFEAT = 16;   // padded input width of the first layer
constexpr int HID = 256;    // hidden width
constexpr int SCR_ROWS = 128;       // events per screen workgroup
constexpr int SCR_LD = HID + 8;     // LDS row stride (bf16 elements), 528 B: spreads rows over banks

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (finite inputs)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }

// The 11 features of an event from FP64 parameter rows (Fisher order
// Mc, eta, dL, theta, phi, iota, psi, tcoal, Phicoal, chi1z, chi2z):
//   log Mc, eta, log dL, cos theta, cos(phi - 2 pi tcoal), sin(phi - 2 pi tcoal),
//   cos iota, cos 2psi, sin 2psi, chi1z, chi2z.
// phi - 2 pi tcoal is the right ascension minus the Greenwich sidereal angle
// at coalescence (tcoal in GMST days), which is what the pattern functions see.
__device__ __forceinline__ void features(const double* __restrict__ par, int64_t ld, int64_t e, float* f) {
  const double Mc = par[e], eta = par[ld + e], dL = par[2 * ld + e], th = par[3 * ld + e], ph = par[4 * ld + e];
  const double io = par[5 * ld + e], ps = par[6 * ld + e], tc = par[7 * ld + e];
  const double c1 = par[9 * ld + e], c2 = par[10 * ld + e];
  const double ang = ph - 2.0 * M_PI * tc;
  f[0] = (float)log(Mc);
  f[1] = (float)eta;
  f[2] = (float)log(dL);
  f[3] = (float)cos(th);
  f[4] = (float)cos(ang);
  f[5] = (float)sin(ang);
  f[6] = (float)cos(io);
  f[7] = (float)cos(2.0 * ps);
  f[8] = (float)sin(2.0 * ps);
  f[9] = (float)c1;
  f[10] = (float)c2;
}

// ---------------------------------------------------------------- GEMM
// C[M x N] = alpha * sum_k A(m, k) B(k, n) + beta * C with
// A(m, k) = A[m sam + k sak], B(k, n) = B[k sbk + n sbn], C(m, n) = C[m ldc + n].
// 256 threads = 4 waves per 64 x 64 tile, each wave a 32 x 32 sub-tile of
// 2 x 2 MFMA 16x16x16 blocks.  Lane l of a block holds A(row l%16,
// k 4(l/16)..+3), B(k 4(l/16)..+3, col l%16) and C(rows 4(l/16)..+3, col l%16).
// blockIdx.z splits K into chunks of kc (a multiple of 16); with more than
// one chunk the partial products are added atomically into C, which the
// host has pre-scaled by beta.  AK / BK: A / B contiguous in k with rows
// 8-byte aligned, so a lane's 4 k values are one 8-byte load.
template <bool AK, bool BK>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, int kc, float alpha,
                                                   const uint16_t* __restrict__ A, int64_t sam, int64_t sak,
                                                   const uint16_t* __restrict__ B, int64_t sbk, int64_t sbn, float beta,
                                                   float* __restrict__ C, int64_t ldc) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
  const int r16 = lane & 15;
  const int kq = (lane >> 4) * 4;
  const int kb = blockIdx.z * kc;
  const int ke = min(K, kb + kc);
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 16) {
    const bool full = k0 + 16 <= ke;
    bf16x4 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int m = m0 + i * 16 + r16;
      if (AK && full) {
        uint2 u = make_uint2(0u, 0u);
        if (m < M) u = *(const uint2*)(A + (int64_t)m * sam + k0 + kq);
        a[i] = __builtin_bit_cast(bf16x4, u);
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = k0 + kq + q;
          a[i][q] = (m < M && k < ke) ? (short)A[(int64_t)m * sam + (int64_t)k * sak] : (short)0;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + j * 16 + r16;
      if (BK && full) {
        uint2 u = make_uint2(0u, 0u);
        if (n < N) u = *(const uint2*)(B + (int64_t)n * sbn + k0 + kq);
        b[j] = __builtin_bit_cast(bf16x4, u);
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = k0 + kq + q;
          b[j][q] = (n < N && k < ke) ? (short)B[(int64_t)k * sbk + (int64_t)n * sbn] : (short)0;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  const bool split = gridDim.z > 1;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + j * 16 + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + i * 16 + kq + r;
        if (m < M && n < N) {
          float* c = C + (int64_t)m * ldc + n;
          if (split)
            atomicAdd(c, alpha * acc[i][j][r]);
          else
            *c = alpha * acc[i][j][r] + (beta != 0.f ? beta * *c : 0.f);
        }
      }
    }
}

// C = beta C (beta == 0 writes zeros, so NaN garbage is cleared)
__global__ void k_scale_rows(float* __restrict__ C, int M, int N, int64_t ldc, float beta) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)M * N) return;
  float* c = C + (idx / N) * ldc + idx % N;
  *c = beta != 0.f ? beta * *c : 0.f;
}

// ---------------------------------------------------------- elementwise
__global__ void k_bias_silu(const float* __restrict__ Cin, const float* __restrict__ bias, int64_t M, int N,
                            float* __restrict__ Z, uint16_t* __restrict__ H) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const float z = Cin[idx] + bias[idx % N];
  if (Z) Z[idx] = z;
  H[idx] = f2bf(silu(z));
}

// dZ = dH SiLU'(Z), SiLU'(z) = s (1 + z (1 - s)), s = sigmoid(z)
__global__ void k_silu_bwd(const float* __restrict__ dH, const float* __restrict__ Z, int64_t n, float* __restrict__ dZ,
                           uint16_t* __restrict__ dZb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float z = Z[i];
  const float s = 1.f / (1.f + __expf(-z));
  const float g = dH[i] * (s * (1.f + z * (1.f - s)));
  if (dZ) dZ[i] = g;
  if (dZb) dZb[i] = f2bf(g);
}

__global__ void k_fill(float* __restrict__ x, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// out[n] += sum over a chunk of rows of w[m] X[m, n]  (w == NULL: weight 1);
// X is fp32 (Xf) or bf16 (Xb).  grid = (ceil(N / 256), row chunks).
constexpr int COLSUM_ROWS = 256;
__global__ void k_colsum(const float* __restrict__ Xf, const uint16_t* __restrict__ Xb, const float* __restrict__ w,
                         int64_t M, int N, float* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int64_t m0 = (int64_t)blockIdx.y * COLSUM_ROWS;
  const int64_t m1 = min(M, m0 + COLSUM_ROWS);
  float s = 0.f;
  for (int64_t m = m0; m < m1; ++m) {
    const float x = Xf ? Xf[m * N + n] : bf2f(Xb[m * N + n]);
    s += w ? w[m] * x : x;
  }
  atomicAdd(out + n, s);
}

// y[m] = sum_k H[m, k] w[k] + b  (H bf16, HID columns): one wave per row
__global__ __launch_bounds__(256) void k_head_fwd(const uint16_t* __restrict__ H, const float* __restrict__ w,
                                                  const float* __restrict__ b, int64_t M, float* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t m = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= M) return;
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < HID / 64; ++q) {
    const int k = q * 64 + lane;
    s += bf2f(H[m * HID + k]) * w[k];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) y[m] = s + b[0];
}

// dH[m, k] = g[m] w[k]
__global__ void k_head_bwd(const float* __restrict__ g, const float* __restrict__ w, int64_t M,
                           float* __restrict__ dH) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * HID) return;
  dH[i] = g[i / HID] * w[i % HID];
}

// g = 2 (y - t) / n, loss += sum (y - t)^2 / n
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, int64_t n, float* __restrict__ g,
                      float* __restrict__ loss) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float d = 0.f;
  if (i < n) {
    d = y[i] - t[i];
    g[i] = 2.f * d / (float)n;
  }
  red[threadIdx.x] = d * d;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)n);
}

// Adam with bias correction: w -= lr (m c1) / (sqrt(v c2) + eps)
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                       uint16_t* __restrict__ wb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float* __restrict__ x, int64_t n, uint16_t* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// raw features (n, 11) fp32 and standardised bf16 rows of 16 (5 zero pads)
__global__ void k_features(const double* __restrict__ par, int64_t ld, int64_t n, const float* __restrict__ mean,
                           const float* __restrict__ istd, float* __restrict__ raw, uint16_t* __restrict__ X) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float f[NFEAT];
  features(par, ld, e, f);
#pragma unroll
  for (int k = 0; k < NFEAT; ++k) {
    if (raw) raw[e * NFEAT + k] = f[k];
    if (X) X[e * KFEAT + k] = f2bf(mean ? (f[k] - mean[k]) * istd[k] : f[k]);
  }
  if (X)
#pragma unroll
    for (int k = NFEAT; k < KFEAT; ++k) X[e * KFEAT + k] = 0;
}

// ------------------------------------------------------- fused screen
struct ScreenArgs {
  const double* par;
  int64_t ld, n;
  const float* mean;  // (11,)
  const float* istd;  // (11,)
  const uint16_t* W1; // (256, 16) bf16
  const float* b1;
  const uint16_t* W2; // (256, 256)
  const float* b2;
  const uint16_t* W3; // (256, 256)
  const float* b3;
  const float* w4;    // (256,) fp32
  const float* b4;    // (1,)
  float y_scale, y_shift;   // log SNR = y * y_scale + y_shift
  float log_lo, log_hi;     // recompute band in log SNR
  float* logsnr_out;        // (n,) or NULL
  int8_t* class_out;        // (n,) 0 below band, 1 in band (or NaN), 2 above
};

// bias + SiLU of a wave's 128 x 64 accumulator block, written to LDS as bf16
__device__ __forceinline__ void screen_epilogue(const f32x4 (&acc)[8][4], const float* __restrict__ bias, int wave,
                                                int lane, uint16_t* H) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int col = 64 * wave + 16 * j + (lane & 15);
    const float bj = bias[col];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * i + 4 * (lane >> 4) + r;
        H[row * SCR_LD + col] = f2bf(silu(acc[i][j][r] + bj));
      }
  }
}

// one 256 -> 256 layer: acc = H W^T with H (128 x 256) in LDS, W (256 x 256)
// row-major (out, in) in global memory; wave w owns output columns 64w..64w+63.
// MFMA 16x16x32: lane l holds A(row l%16, k 8(l/16)..+7), B(k 8(l/16)..+7, col l%16).
__device__ __forceinline__ void screen_layer(const uint16_t* __restrict__ W, const uint16_t* H, int wave, int lane,
                                             f32x4 (&acc)[8][4]) {
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kq = 8 * (lane >> 4);
  for (int k0 = 0; k0 < HID; k0 += 32) {
    bf16x8 b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = 64 * wave + 16 * j + (lane & 15);
      b[j] = __builtin_bit_cast(bf16x8, *(const uint4*)(W + col * HID + k0 + kq));
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const bf16x8 a = __builtin_bit_cast(bf16x8, *(const uint4*)(H + (16 * i + (lane & 15)) * SCR_LD + k0 + kq));
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
    }
  }
}

__global__ __launch_bounds__(256, 2) void k_mlp_screen(ScreenArgs a) {
  __shared__ __align__(16) uint16_t H[SCR_ROWS * SCR_LD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t e0 = (int64_t)blockIdx.x * SCR_ROWS;
  if (tid < SCR_ROWS) {
    const int64_t e = e0 + tid;
    float f[NFEAT];
    if (e < a.n) {
      features(a.par, a.ld, e, f);
    } else {
#pragma unroll
      for (int k = 0; k < NFEAT; ++k) f[k] = 0.f;
    }
#pragma unroll
    for (int k = 0; k < NFEAT; ++k) H[tid * SCR_LD + k] = f2bf((f[k] - a.mean[k]) * a.istd[k]);
#pragma unroll
    for (int k = NFEAT; k < KFEAT; ++k) H[tid * SCR_LD + k] = 0;
  }
  __syncthreads();
  f32x4 acc[8][4];
  {  // first layer, K = 16: one MFMA 16x16x16 step
    const int kq = 4 * (lane >> 4);
    bf16x4 b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = 64 * wave + 16 * j + (lane & 15);
      b[j] = __builtin_bit_cast(bf16x4, *(const uint2*)(a.W1 + col * KFEAT + kq));
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const bf16x4 x = __builtin_bit_cast(bf16x4, *(const uint2*)(H + (16 * i + (lane & 15)) * SCR_LD + kq));
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x, b[j], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    }
  }
  __syncthreads();
  screen_epilogue(acc, a.b1, wave, lane, H);
  __syncthreads();
  screen_layer(a.W2, H, wave, lane, acc);
  __syncthreads();
  screen_epilogue(acc, a.b2, wave, lane, H);
  __syncthreads();
  screen_layer(a.W3, H, wave, lane, acc);
  __syncthreads();
  screen_epilogue(acc, a.b3, wave, lane, H);
  __syncthreads();
  // output layer: two threads per event, 128 columns each
  const int row = tid >> 1, half = tid & 1;
  float s = 0.f;
  const uint16_t* h = H + row * SCR_LD + half * (HID / 2);
  const float* w = a.w4 + half * (HID / 2);
#pragma unroll 8
  for (int k = 0; k < HID / 2; ++k) s += bf2f(h[k]) * w[k];
  s += __shfl_xor(s, 1);
  const int64_t e = e0 + row;
  if (half == 0 && e < a.n) {
    const float ls = (s + a.b4[0]) * a.y_scale + a.y_shift;
    if (a.logsnr_out) a.logsnr_out[e] = ls;
    if (a.class_out) {
      int8_t c = 1;
      if (ls < a.log_lo) c = 0;
      else if (ls > a.log_hi) c = 2;
      a.class_out[e] = c;
    }
  }
}

// Stream compaction of the events of one class, order preserved, in three
// launches: per-chunk counts, one-workgroup exclusive scan of the counts,
// scatter.  A chunk is one 256-thread workgroup's 256 events.
constexpr int SEL_CHUNK = 256;

__device__ __forceinline__ int chunk_rank(bool hit, int& total, int* wave_tot) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const unsigned long long mask = __ballot(hit);
  if (lane == 0) wave_tot[wave] = __popcll(mask);
  __syncthreads();
  int off = 0;
  total = 0;
  for (int w = 0; w < SEL_CHUNK / 64; ++w) {
    if (w < wave) off += wave_tot[w];
    total += wave_tot[w];
  }
  return off + __popcll(mask & ((1ull << lane) - 1ull));
}

__global__ __launch_bounds__(SEL_CHUNK) void k_sel_count(const int8_t* __restrict__ cls, int64_t n, int8_t want,
                                                         int64_t* __restrict__ counts) {
  __shared__ int wave_tot[SEL_CHUNK / 64];
  const int64_t i = (int64_t)blockIdx.x * SEL_CHUNK + threadIdx.x;
  int total;
  chunk_rank(i < n && cls[i] == want, total, wave_tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// exclusive scan of counts[0..nb) in place; total -> count_out
__global__ __launch_bounds__(1024) void k_sel_scan(int64_t* __restrict__ counts, int64_t nb,
                                                   int64_t* __restrict__ count_out) {
  __shared__ int64_t buf[1024];
  __shared__ int64_t carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < nb; c0 += 1024) {
    const int64_t i = c0 + tid;
    const int64_t v = i < nb ? counts[i] : 0;
    buf[tid] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
      const int64_t t = tid >= o ? buf[tid - o] : 0;
      __syncthreads();
      buf[tid] += t;
      __syncthreads();
    }
    if (i < nb) counts[i] = carry + buf[tid] - v;
    __syncthreads();
    if (tid == 0) carry += buf[1023];
    __syncthreads();
  }
  if (tid == 0) count_out[0] = carry;
}

__global__ __launch_bounds__(SEL_CHUNK) void k_sel_scatter(const int8_t* __restrict__ cls, int64_t n, int8_t want,
                                                           const int64_t* __restrict__ offsets,
                                                           int64_t* __restrict__ idx_out) {
  __shared__ int wave_tot[SEL_CHUNK / 64];
  const int64_t i = (int64_t)blockIdx.x * SEL_CHUNK + threadIdx.x;
  const bool hit = i < n && cls[i] == want;
  int total;
  const int r = chunk_rank(hit, total, wave_tot);
  if (hit) idx_out[offsets[blockIdx.x] + r] = i;
}

// out[r * ld_out + j] = par[r * ld + idx[j]] for r < n_rows, j < n_idx
__global__ void k_gather(const double* __restrict__ par, int64_t ld, int n_rows, const int64_t* __restrict__ idx,
                         int64_t n_idx, double* __restrict__ out, int64_t ld_out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_idx) return;
  const int64_t e = idx[j];
  if (e < 0 || e >= ld) {  // outside every row: a NaN column instead of a stray read
    for (int r = 0; r < n_rows; ++r) out[r * ld_out + j] = __builtin_nan("");
    return;
  }
  for (int r = 0; r < n_rows; ++r) out[r * ld_out + j] = par[r * ld + e];
}

int fail(int code, const char* msg) {
  set_error(msg);
  return code;
}

int launch_ok(const char* what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) return 0;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return -EIO;
}

unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace gwf_mlp

using namespace gwf_mlp;

extern "C" {

int gwf_gemm_bf16(int32_t M, int32_t N, int32_t K, float alpha, const void* A, int64_t sam, int64_t sak,
                  const void* B, int64_t sbk, int64_t sbn, float beta, float* C, int64_t ldc, int32_t split_k,
                  void* stream) {
  if (M < 0 || N < 0 || K < 0) return fail(-EINVAL, "negative GEMM size");
  if (split_k < 1 || split_k > 1024) return fail(-EINVAL, "split_k must be in [1, 1024]");
  if ((int64_t)M * N == 0) return 0;
  if (!A || !B || !C) return fail(-EINVAL, "NULL GEMM operand");
  if (ldc < N) return fail(-EINVAL, "ldc < N");
  if ((M + 63) / 64 > 65535) return fail(-EINVAL, "M too large for one launch");
  hipStream_t s = (hipStream_t)stream;
  int kc = (K + split_k - 1) / split_k;
  kc = (kc + 15) / 16 * 16;
  const int nz = K == 0 ? 1 : (K + kc - 1) / kc;
  if (nz > 1 || K == 0) {
    hipLaunchKernelGGL(k_scale_rows, dim3(blocks_for((int64_t)M * N, 256)), dim3(256), 0, s, C, (int)M, (int)N, ldc,
                       beta);
    if (K == 0) return launch_ok("gwf_gemm_bf16 launch");
  }
  const uint16_t* a = (const uint16_t*)A;
  const uint16_t* b = (const uint16_t*)B;
  const bool ak = sak == 1 && sam % 4 == 0 && ((uintptr_t)a & 7) == 0;
  const bool bk = sbk == 1 && sbn % 4 == 0 && ((uintptr_t)b & 7) == 0;
  const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64), (unsigned)nz), block(256);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, block, 0, s, (int)M, (int)N, (int)K, kc, alpha, a, sam, sak, b, sbk, sbn, beta, C,
                       ldc);
  };
  if (ak && bk)
    go(k_gemm_bf16<true, true>);
  else if (ak)
    go(k_gemm_bf16<true, false>);
  else if (bk)
    go(k_gemm_bf16<false, true>);
  else
    go(k_gemm_bf16<false, false>);
  return launch_ok("gwf_gemm_bf16 launch");
}

int gwf_bias_silu(const float* Cin, const float* bias, int64_t M, int32_t N, float* Z_out, void* H_out, void* stream) {
  if (M < 0 || N < 1) return fail(-EINVAL, "invalid sizes");
  if (M == 0) return 0;
  if (!Cin || !bias || !H_out) return fail(-EINVAL, "NULL array");
  hipLaunchKernelGGL(k_bias_silu, dim3(blocks_for(M * N, 256)), dim3(256), 0, (hipStream_t)stream, Cin, bias, M,
                     (int)N, Z_out, (uint16_t*)H_out);
  return launch_ok("gwf_bias_silu launch");
}

int gwf_silu_backward(const float* dH, const float* Z, int64_t n, float* dZ_out, void* dZ_bf16_out, void* stream) {
  if (n < 0) return fail(-EINVAL, "invalid size");
  if (n == 0) return 0;
  if (!dH || !Z || (!dZ_out && !dZ_bf16_out)) return fail(-EINVAL, "NULL array");
  hipLaunchKernelGGL(k_silu_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, dH, Z, n, dZ_out,
                     (uint16_t*)dZ_bf16_out);
  return launch_ok("gwf_silu_backward launch");
}

int gwf_colsum(const float* X, const void* X_bf16, const float* w, int64_t M, int32_t N, float* out,
               int32_t accumulate, void* stream) {
  if (M < 0 || N < 1) return fail(-EINVAL, "invalid sizes");
  if ((X == nullptr) == (X_bf16 == nullptr)) return fail(-EINVAL, "give exactly one of X and X_bf16");
  if (!out) return fail(-EINVAL, "NULL output");
  const int64_t chunks = (M + COLSUM_ROWS - 1) / COLSUM_ROWS;
  if (chunks > 65535) return fail(-EINVAL, "too many rows for one launch");
  hipStream_t s = (hipStream_t)stream;
  if (!accumulate) hipLaunchKernelGGL(k_fill, dim3(blocks_for(N, 256)), dim3(256), 0, s, out, (int64_t)N, 0.f);
  if (M > 0)
    hipLaunchKernelGGL(k_colsum, dim3(blocks_for(N, 256), (unsigned)chunks), dim3(256), 0, s, X,
                       (const uint16_t*)X_bf16, w, M, (int)N, out);
  return launch_ok("gwf_colsum launch");
}

int gwf_head_forward(const void* H, const float* w, const float* b, int64_t M, float* y, void* stream) {
  if (M < 0) return fail(-EINVAL, "invalid size");
  if (M == 0) return 0;
  if (!H || !w || !b || !y) return fail(-EINVAL, "NULL array");
  hipLaunchKernelGGL(k_head_fwd, dim3(blocks_for(M, 4)), dim3(256), 0, (hipStream_t)stream, (const uint16_t*)H, w, b,
                     M, y);
  return launch_ok("gwf_head_forward launch");
}

int gwf_head_backward(const void* H, const float* w, const float* g, int64_t M, float* dH_out, float* dw_out,
                      float* db_out, void* stream) {
  if (M < 1) return fail(-EINVAL, "invalid size");
  if (!H || !w || !g || !dH_out || !dw_out || !db_out) return fail(-EINVAL, "NULL array");
  hipLaunchKernelGGL(k_head_bwd, dim3(blocks_for(M * HID, 256)), dim3(256), 0, (hipStream_t)stream, g, w, M, dH_out);
  int rc = launch_ok("gwf_head_backward launch");
  if (rc) return rc;
  rc = gwf_colsum(nullptr, H, g, M, HID, dw_out, 0, stream);
  if (rc) return rc;
  return gwf_colsum(g, nullptr, nullptr, M, 1, db_out, 0, stream);
}

int gwf_mse_grad(const float* y, const float* t, int64_t n, float* g_out, float* loss_accum, void* stream) {
  if (n < 1) return fail(-EINVAL, "invalid size");
  if (!y || !t || !g_out || !loss_accum) return fail(-EINVAL, "NULL array");
  hipLaunchKernelGGL(k_mse, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, y, t, n, g_out, loss_accum);
  return launch_ok("gwf_mse_grad launch");
}

int gwf_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2, float eps,
             int32_t step, void* w_bf16_out, void* stream) {
  if (n < 0 || step < 1) return fail(-EINVAL, "invalid size or step");
  if (n == 0) return 0;
  if (!w || !g || !m || !v) return fail(-EINVAL, "NULL array");
  const float c1 = (float)(1.0 / (1.0 - pow((double)beta1, (double)step)));
  const float c2 = (float)(1.0 / (1.0 - pow((double)beta2, (double)step)));
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, n, lr, beta1,
                     beta2, eps, c1, c2, (uint16_t*)w_bf16_out);
  return launch_ok("gwf_adam launch");
}

int gwf_to_bf16(const float* x, int64_t n, void* y, void* stream) {
  if (n < 0) return fail(-EINVAL, "invalid size");
  if (n == 0) return 0;
  if (!x || !y) return fail(-EINVAL, "NULL array");
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, x, n, (uint16_t*)y);
  return launch_ok("gwf_to_bf16 launch");
}

int gwf_snr_features(const double* params, int64_t ld, int64_t n_ev, const float* mean, const float* istd,
                     float* raw_out, void* X_out, void* stream) {
  if (n_ev < 0 || ld < n_ev) return fail(-EINVAL, "invalid sizes");
  if (n_ev == 0) return 0;
  if (!params || (!raw_out && !X_out)) return fail(-EINVAL, "NULL array");
  if ((mean == nullptr) != (istd == nullptr)) return fail(-EINVAL, "mean and istd must both be given or both NULL");
  hipLaunchKernelGGL(k_features, dim3(blocks_for(n_ev, 256)), dim3(256), 0, (hipStream_t)stream, params, ld, n_ev,
                     mean, istd, raw_out, (uint16_t*)X_out);
  return launch_ok("gwf_snr_features launch");
}

int gwf_snr_screen(const double* params, int64_t ld, int64_t n_ev, const float* mean, const float* istd,
                   const void* W1, const float* b1, const void* W2, const float* b2, const void* W3, const float* b3,
                   const float* w4, const float* b4, float y_scale, float y_shift, float snr_lo, float snr_hi,
                   float* logsnr_out, int8_t* class_out, void* stream) {
  if (n_ev < 0 || ld < n_ev) return fail(-EINVAL, "invalid sizes");
  if (n_ev == 0) return 0;
  if (!params || !mean || !istd || !W1 || !b1 || !W2 || !b2 || !W3 || !b3 || !w4 || !b4)
    return fail(-EINVAL, "NULL array");
  if (!logsnr_out && !class_out) return fail(-EINVAL, "no output requested");
  if (!(snr_lo > 0.f) || !(snr_hi >= snr_lo)) return fail(-EINVAL, "need 0 < snr_lo <= snr_hi");
  if ((((uintptr_t)W1) | ((uintptr_t)W2) | ((uintptr_t)W3)) & 15) return fail(-EINVAL, "weights must be 16-byte aligned");
  if ((n_ev + SCR_ROWS - 1) / SCR_ROWS > 0x7fffffff) return fail(-EINVAL, "too many events for one call");
  ScreenArgs a;
  a.par = params;
  a.ld = ld;
  a.n = n_ev;
  a.mean = mean;
  a.istd = istd;
  a.W1 = (const uint16_t*)W1;
  a.b1 = b1;
  a.W2 = (const uint16_t*)W2;
  a.b2 = b2;
  a.W3 = (const uint16_t*)W3;
  a.b3 = b3;
  a.w4 = w4;
  a.b4 = b4;
  a.y_scale = y_scale;
  a.y_shift = y_shift;
  a.log_lo = logf(snr_lo);
  a.log_hi = logf(snr_hi);
  a.logsnr_out = logsnr_out;
  a.class_out = class_out;
  hipLaunchKernelGGL(k_mlp_screen, dim3(blocks_for(n_ev, SCR_ROWS)), dim3(256), 0, (hipStream_t)stream, a);
  return launch_ok("gwf_snr_screen launch");
}

size_t gwf_select_workspace_bytes(int64_t n) {
  return (size_t)((n < 0 ? 0 : n) + SEL_CHUNK - 1) / SEL_CHUNK * sizeof(int64_t);
}

int gwf_select(const int8_t* cls, int64_t n, int32_t want, int64_t* idx_out, int64_t* count_out, void* workspace,
               size_t workspace_bytes, void* stream) {
  if (n < 0) return fail(-EINVAL, "invalid size");
  if (!count_out) return fail(-EINVAL, "NULL count_out");
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(1024), 0, s, count_out, (int64_t)0, count_out);
    return launch_ok("gwf_select launch");
  }
  if (!cls || !idx_out || !workspace) return fail(-EINVAL, "NULL array");
  if (workspace_bytes < gwf_select_workspace_bytes(n)) return fail(-EINVAL, "workspace too small");
  const int64_t nb = (n + SEL_CHUNK - 1) / SEL_CHUNK;
  if (nb > 0x7fffffff) return fail(-EINVAL, "too many events for one call");
  int64_t* counts = (int64_t*)workspace;
  hipLaunchKernelGGL(k_sel_count, dim3((unsigned)nb), dim3(SEL_CHUNK), 0, s, cls, n, (int8_t)want, counts);
  hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(1024), 0, s, counts, nb, count_out);
  hipLaunchKernelGGL(k_sel_scatter, dim3((unsigned)nb), dim3(SEL_CHUNK), 0, s, cls, n, (int8_t)want,
                     (const int64_t*)counts, idx_out);
  return launch_ok("gwf_select launch");
}

int gwf_gather_events(const double* params, int64_t ld, int32_t n_rows, const int64_t* idx, int64_t n_idx,
                      double* out, int64_t ld_out, void* stream) {
  if (n_rows < 0 || n_idx < 0 || ld_out < n_idx) return fail(-EINVAL, "invalid sizes");
  if (n_idx == 0 || n_rows == 0) return 0;
  if (!params || !idx || !out) return fail(-EINVAL, "NULL array");
  hipLaunchKernelGGL(k_gather, dim3(blocks_for(n_idx, 256)), dim3(256), 0, (hipStream_t)stream, params, ld,
                     (int)n_rows, idx, n_idx, out, ld_out);
  return launch_ok("gwf_gather_events launch");
}

}  // extern "C"
