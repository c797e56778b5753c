// Forward-mode dual numbers for the per-event coefficient setup and the
// per-frequency tangent evaluation.  D<N> carries a value and N directional
// derivatives; tangent slots are fixed to
//   0 Mc, 1 eta, 2 chi1z, 3 chi2z, 4 LambdaTilde, 5 deltaLambda
// which are exactly the parameters gwfast differentiates with JAX
// (gwfast/signal.py:1153-1167, derivargs = (1,2,10,11,16,17)).
#pragma once
#include <type_traits>
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GWF_HD __host__ __device__ __forceinline__
#else
#define GWF_HD inline
#endif

namespace gwf {

template <int N>
struct D {
  double v;
  double d[N];
};

template <int N>
GWF_HD D<N> mkc(double v) {  // constant
  D<N> r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = 0.0;
  return r;
}

template <int N>
GWF_HD D<N> mkseed(double v, int slot) {
  D<N> r = mkc<N>(v);
  if (slot >= 0 && slot < N) r.d[slot] = 1.0;
  return r;
}

GWF_HD double val(double x) { return x; }
template <int N>
GWF_HD double val(const D<N>& x) { return x.v; }

// ---------------------------------------------------------------- arithmetic
template <int N>
GWF_HD D<N> operator+(const D<N>& a, const D<N>& b) {
  D<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N>
GWF_HD D<N> operator+(const D<N>& a, double b) {
  D<N> r = a;
  r.v = a.v + b;
  return r;
}
template <int N>
GWF_HD D<N> operator+(double a, const D<N>& b) {
  D<N> r = b;
  r.v = a + b.v;
  return r;
}
template <int N>
GWF_HD D<N> operator-(const D<N>& a, const D<N>& b) {
  D<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N>
GWF_HD D<N> operator-(const D<N>& a, double b) {
  D<N> r = a;
  r.v = a.v - b;
  return r;
}
template <int N>
GWF_HD D<N> operator-(double a, const D<N>& b) {
  D<N> r;
  r.v = a - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -b.d[k];
  return r;
}
template <int N>
GWF_HD D<N> operator-(const D<N>& a) {
  D<N> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N>
GWF_HD D<N> operator*(const D<N>& a, const D<N>& b) {
  D<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int N>
GWF_HD D<N> operator*(const D<N>& a, double b) {
  D<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int N>
GWF_HD D<N> operator*(double a, const D<N>& b) {
  return b * a;
}
template <int N>
GWF_HD D<N> operator/(const D<N>& a, const D<N>& b) {
  // value by a true division (the collocation and join formulas of the
  // setup are ill-conditioned: a reciprocal-multiply would perturb them)
  D<N> r;
  const double ib = 1.0 / b.v;
  r.v = a.v / b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
template <int N>
GWF_HD D<N> operator/(const D<N>& a, double b) {
  D<N> r;
  const double ib = 1.0 / b;
  r.v = a.v / b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * ib;
  return r;
}
template <int N>
GWF_HD D<N> operator/(double a, const D<N>& b) {
  D<N> r;
  const double ib = 1.0 / b.v;
  r.v = a / b.v;
  const double s = -r.v * ib;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = b.d[k] * s;
  return r;
}

// Mixed arithmetic: a D<M> with M < N stands for a D<N> whose last N - M
// tangents are zero (functions of M G f depend on Mc and eta only, the first
// two tangents); the result is a D<N> and the zero tangents are not
// multiplied.  The values equal those of the full D<N> arithmetic.
template <int N, int M>
using MixedIf = typename std::enable_if<(M < N), int>::type;

template <int N, int M, MixedIf<N, M> = 0>
GWF_HD D<N> operator+(const D<N>& a, const D<M>& b) {
  D<N> r = a;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < M; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N, int M, MixedIf<N, M> = 0>
GWF_HD D<N> operator+(const D<M>& a, const D<N>& b) {
  D<N> r = b;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < M; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N, int M, MixedIf<N, M> = 0>
GWF_HD D<N> operator-(const D<N>& a, const D<M>& b) {
  D<N> r = a;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < M; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N, int M, MixedIf<N, M> = 0>
GWF_HD D<N> operator-(const D<M>& a, const D<N>& b) {
  D<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < M; ++k) r.d[k] = a.d[k] - b.d[k];
#pragma unroll
  for (int k = M; k < N; ++k) r.d[k] = -b.d[k];
  return r;
}
template <int N, int M, MixedIf<N, M> = 0>
GWF_HD D<N> operator*(const D<N>& a, const D<M>& b) {
  D<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < M; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
#pragma unroll
  for (int k = M; k < N; ++k) r.d[k] = a.d[k] * b.v;
  return r;
}
template <int N, int M, MixedIf<N, M> = 0>
GWF_HD D<N> operator*(const D<M>& a, const D<N>& b) {
  D<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < M; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
#pragma unroll
  for (int k = M; k < N; ++k) r.d[k] = a.v * b.d[k];
  return r;
}
template <int N, int M, MixedIf<N, M> = 0>
GWF_HD D<N> operator/(const D<N>& a, const D<M>& b) {
  D<N> r;
  const double ib = 1.0 / b.v;
  r.v = a.v / b.v;
#pragma unroll
  for (int k = 0; k < M; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
#pragma unroll
  for (int k = M; k < N; ++k) r.d[k] = a.d[k] * ib;
  return r;
}
template <int N, int M, MixedIf<N, M> = 0>
GWF_HD D<N> operator/(const D<M>& a, const D<N>& b) {
  D<N> r;
  const double ib = 1.0 / b.v;
  r.v = a.v / b.v;
#pragma unroll
  for (int k = 0; k < M; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
#pragma unroll
  for (int k = M; k < N; ++k) r.d[k] = (-r.v * b.d[k]) * ib;
  return r;
}

// Horner step q <- c + u q with a D<N> accumulator and coefficient and a
// D<M> variable (M <= N; u's last N - M tangents are zero), in FMAs only:
// the sum-of-products form costs a multiply, an FMA and an add per full
// tangent and term.
template <int N, int M>
GWF_HD void dhorner(D<N>& q, const D<M>& u, const D<N>& c) {
  static_assert(M <= N, "variable with more tangents than the accumulator");
  const double qv = q.v;
  q.v = fma(u.v, qv, c.v);
#pragma unroll
  for (int k = 0; k < M; ++k) q.d[k] = fma(u.d[k], qv, fma(u.v, q.d[k], c.d[k]));
#pragma unroll
  for (int k = M; k < N; ++k) q.d[k] = fma(u.v, q.d[k], c.d[k]);
}

// ------------------------------------------------------------- elementaries
// Each takes the value-level result and its derivative factor.
template <int N>
GWF_HD D<N> chain(const D<N>& a, double v, double dfac) {
  D<N> r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * dfac;
  return r;
}

GWF_HD double dsqrt(double x) { return sqrt(x); }
GWF_HD double dlog(double x) { return log(x); }
GWF_HD double dexp(double x) { return exp(x); }
GWF_HD double datan(double x) { return atan(x); }
GWF_HD double dfabs(double x) { return fabs(x); }
GWF_HD double dpow(double x, double c) { return pow(x, c); }

template <int N>
GWF_HD D<N> dsqrt(const D<N>& a) {
  const double s = sqrt(a.v);
  return chain(a, s, 0.5 / s);  // 0 * (0.5/0) = NaN at the eta = 1/4 guard, as JAX
}
template <int N>
GWF_HD D<N> dlog(const D<N>& a) {
  return chain(a, log(a.v), 1.0 / a.v);
}
template <int N>
GWF_HD D<N> dexp(const D<N>& a) {
  const double e = exp(a.v);
  return chain(a, e, e);
}
template <int N>
GWF_HD D<N> datan(const D<N>& a) {
  return chain(a, atan(a.v), 1.0 / (1.0 + a.v * a.v));
}
template <int N>
GWF_HD D<N> dfabs(const D<N>& a) {
  const double s = (a.v > 0.0) ? 1.0 : ((a.v < 0.0) ? -1.0 : 0.0);
  return chain(a, fabs(a.v), s);
}
// x^(1/3) via cbrt; derivative (1/3) x^(1/3) / x
GWF_HD double dcbrt(double x) { return cbrt(x); }
template <int N>
GWF_HD D<N> dcbrt(const D<N>& a) {
  const double c = cbrt(a.v);
  return chain(a, c, c / (3.0 * a.v));
}

template <int N>
GWF_HD D<N> dpow(const D<N>& a, double c) {
  const double p = pow(a.v, c);
  return chain(a, p, c * pow(a.v, c - 1.0));
}

template <class T>
GWF_HD T sel(bool c, const T& a, const T& b) {
  return c ? a : b;
}

// Load / store a dual stored as (N+1) consecutive doubles.
template <int N>
GWF_HD D<N> ldd(const double* p) {
  D<N> r;
  r.v = p[0];
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = p[1 + k];
  return r;
}
template <int N>
GWF_HD void std_(double* p, const D<N>& x) {
  p[0] = x.v;
#pragma unroll
  for (int k = 0; k < N; ++k) p[1 + k] = x.d[k];
}
GWF_HD void std_(double* p, double x) { p[0] = x; }

// Store of one lane of a tangent-split evaluation: with tk >= 0 the D<1> x
// carries tangent tk only, and the lane writes that tangent (plus the value
// when tk == 0) into the full (value, tangents) slot at p; tk < 0 stores
// the whole dual.
template <int N>
GWF_HD void std_tk(double* p, const D<N>& x, int tk) {
  if (tk < 0) {
    std_(p, x);
  } else {
    if (tk == 0) p[0] = x.v;
    p[1 + tk] = x.d[0];
  }
}

}  // namespace gwf
