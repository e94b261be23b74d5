// dual.hpp -- forward-mode dual numbers for the generated model functions: a value
// and N partial derivatives (N = 0: a plain value), the elementary functions of
// the pdemodel language, and the small-matrix inverse and product it offers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

namespace exa {

template <int N>
struct Dual {
  double v;
  double d[N > 0 ? N : 1];
  __device__ __forceinline__ Dual() : v(0.0) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
  __device__ __forceinline__ explicit Dual(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
  __device__ __forceinline__ void seed(int k) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = (i == k) ? 1.0 : 0.0;
  }
  __device__ __forceinline__ double deriv(int k) const { return N > 0 ? d[k] : 0.0; }
};

// chain rule helper: r = f(a) with f'(a) = g
template <int N>
__device__ __forceinline__ Dual<N> dual_chain(const Dual<N>& a, double fv, double g) {
  Dual<N> r(fv);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = g * a.d[i];
  return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v + b.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v - b.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r(-a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v * b.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  const double q = a.v / b.v;
  Dual<N> r(q);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - q * b.d[i]) / b.v;
  return r;
}

template <int N> __device__ __forceinline__ Dual<N> dsin(const Dual<N>& a) { return dual_chain(a, sin(a.v), cos(a.v)); }
template <int N> __device__ __forceinline__ Dual<N> dcos(const Dual<N>& a) { return dual_chain(a, cos(a.v), -sin(a.v)); }
template <int N> __device__ __forceinline__ Dual<N> dtan(const Dual<N>& a) {
  const double t = tan(a.v);
  return dual_chain(a, t, 1.0 + t * t);
}
template <int N> __device__ __forceinline__ Dual<N> dexp(const Dual<N>& a) {
  const double e = exp(a.v);
  return dual_chain(a, e, e);
}
template <int N> __device__ __forceinline__ Dual<N> dlog(const Dual<N>& a) { return dual_chain(a, log(a.v), 1.0 / a.v); }
template <int N> __device__ __forceinline__ Dual<N> dsqrt(const Dual<N>& a) {
  const double r = sqrt(a.v);
  return dual_chain(a, r, 0.5 / r);
}
template <int N> __device__ __forceinline__ Dual<N> dtanh(const Dual<N>& a) {
  const double t = tanh(a.v);
  return dual_chain(a, t, 1.0 - t * t);
}
template <int N> __device__ __forceinline__ Dual<N> dsinh(const Dual<N>& a) { return dual_chain(a, sinh(a.v), cosh(a.v)); }
template <int N> __device__ __forceinline__ Dual<N> dcosh(const Dual<N>& a) { return dual_chain(a, cosh(a.v), sinh(a.v)); }
template <int N> __device__ __forceinline__ Dual<N> datan(const Dual<N>& a) {
  return dual_chain(a, atan(a.v), 1.0 / (1.0 + a.v * a.v));
}
template <int N> __device__ __forceinline__ Dual<N> dabs(const Dual<N>& a) {
  return dual_chain(a, fabs(a.v), a.v < 0.0 ? -1.0 : 1.0);
}
// a^b; for a constant exponent (all its derivatives zero) the derivative is b a^(b-1)
template <int N>
__device__ __forceinline__ Dual<N> dpow(const Dual<N>& a, const Dual<N>& b) {
  bool bconst = true;
#pragma unroll
  for (int i = 0; i < N; ++i) bconst = bconst && (b.d[i] == 0.0);
  const double pv = pow(a.v, b.v);
  Dual<N> r(pv);
  const double da = (b.v == 0.0) ? 0.0 : b.v * pow(a.v, b.v - 1.0);
  const double db = bconst ? 0.0 : pv * log(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = da * a.d[i] + db * b.d[i];
  return r;
}

// Ainv = A^{-1}: Gauss-Jordan on [A | I].  The pivot search compares row c with
// each later row in turn and swaps when the later entry is larger in magnitude,
// so row c ends with the largest |entry| of column c (partial pivoting) while
// every index stays a compile-time constant (the rows stay in registers).
template <typename S, int n>
__device__ __forceinline__ void mat_inv(const S (&A)[n][n], S (&Ainv)[n][n]) {
  S W[n][2 * n];
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = 0; j < 2 * n; ++j) W[i][j] = (j < n) ? A[i][j] : S(j - n == i ? 1.0 : 0.0);
#pragma unroll
  for (int c = 0; c < n; ++c) {
#pragma unroll
    for (int r = c + 1; r < n; ++r) {
      const bool sw = fabs(W[r][c].v) > fabs(W[c][c].v);
#pragma unroll
      for (int j = 0; j < 2 * n; ++j) {
        const S a = W[c][j], b = W[r][j];
        W[c][j] = sw ? b : a;
        W[r][j] = sw ? a : b;
      }
    }
    const S piv = W[c][c];
#pragma unroll
    for (int j = 0; j < 2 * n; ++j) W[c][j] = W[c][j] / piv;
#pragma unroll
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const S f = W[r][c];
#pragma unroll
      for (int j = 0; j < 2 * n; ++j) W[r][j] = W[r][j] - f * W[c][j];
    }
  }
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = 0; j < n; ++j) Ainv[i][j] = W[i][n + j];
}

// C = A B (C may be A or B)
template <typename S, int r, int k, int c>
__device__ __forceinline__ void mat_mul(const S (&A)[r][k], const S (&B)[k][c], S (&C)[r][c]) {
  S P[r][c];
#pragma unroll
  for (int i = 0; i < r; ++i)
#pragma unroll
    for (int j = 0; j < c; ++j) {
      S acc(0.0);
#pragma unroll
      for (int t = 0; t < k; ++t) acc = acc + A[i][t] * B[t][j];
      P[i][j] = acc;
    }
#pragma unroll
  for (int i = 0; i < r; ++i)
#pragma unroll
    for (int j = 0; j < c; ++j) C[i][j] = P[i][j];
}

}  // namespace exa
