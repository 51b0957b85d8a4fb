// Forward-mode dual numbers for the device physics.
//
// Dual<K> carries a value and K directional derivatives.  The value part of
// every operation is the same IEEE double operation the plain double code
// performs, so a function evaluated on Dual<K> takes exactly the branches the
// double evaluation takes and returns bit-identical values; the derivative
// parts are the tangents of that branch (the derivative torch autograd would
// propagate through the same where / clamp / maximum choices).
#pragma once

#include <hip/hip_runtime.h>

#define SRMHD_DEV __host__ __device__ __forceinline__

namespace srmhd {

template <int K>
struct Dual {
  double v;
  double d[K];
  SRMHD_DEV Dual() : v(0.0) {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  SRMHD_DEV Dual(double x) : v(x) {   // NOLINT: constants have zero tangent
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  // a value with unit tangent in direction `dir` (dir < 0: none)
  SRMHD_DEV static Dual seed(double x, int dir) {
    Dual r(x);
    if (dir >= 0 && dir < K) r.d[dir] = 1.0;
    return r;
  }
};

// value of a double or a Dual (branch decisions, comparisons)
SRMHD_DEV double val(double x) { return x; }
template <int K>
SRMHD_DEV double val(const Dual<K>& x) { return x.v; }

// ── arithmetic ──────────────────────────────────────────────────────────────
template <int K>
SRMHD_DEV Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator+(const Dual<K>& a, double b) {
  Dual<K> r = a;
  r.v = a.v + b;
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator+(double a, const Dual<K>& b) {
  Dual<K> r = b;
  r.v = a + b.v;
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator-(const Dual<K>& a, double b) {
  Dual<K> r = a;
  r.v = a.v - b;
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator-(double a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a - b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -b.d[k];
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator*(const Dual<K>& a, double b) {
  Dual<K> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator*(double a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a * b.d[k];
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator/(const Dual<K>& a, double b) {
  Dual<K> r;
  r.v = a.v / b;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / b;
  return r;
}
template <int K>
SRMHD_DEV Dual<K> operator/(double a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a / b.v;
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -r.v * b.d[k] * ib;
  return r;
}

// ── comparisons on values ───────────────────────────────────────────────────
#define SRMHD_DUAL_CMP(op)                                                              \
  template <int K>                                                                      \
  SRMHD_DEV bool operator op(const Dual<K>& a, const Dual<K>& b) { return a.v op b.v; } \
  template <int K>                                                                      \
  SRMHD_DEV bool operator op(const Dual<K>& a, double b) { return a.v op b; }           \
  template <int K>                                                                      \
  SRMHD_DEV bool operator op(double a, const Dual<K>& b) { return a op b.v; }
SRMHD_DUAL_CMP(<)
SRMHD_DUAL_CMP(>)
SRMHD_DUAL_CMP(<=)
SRMHD_DUAL_CMP(>=)
SRMHD_DUAL_CMP(==)
SRMHD_DUAL_CMP(!=)
#undef SRMHD_DUAL_CMP

// ── functions (the double forms live in srmhd_common.h) ─────────────────────
template <int K>
SRMHD_DEV Dual<K> fabs(const Dual<K>& a) {
  return a.v < 0.0 ? -a : (a.v > 0.0 ? a : Dual<K>(__builtin_fabs(a.v)));
}

}  // namespace srmhd
