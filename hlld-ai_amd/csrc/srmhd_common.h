// Point-wise SRMHD device functions for gfx950 (FP64 path).
//
// Arithmetic is written in the exact association order of the reference's
// float64 torch expressions and the library is compiled with
// -ffp-contract=off, so every value is bit-identical to the reference CPU
// path (AMDGPU f64 division and sqrt are correctly rounded).  NaN behaviour
// follows torch: clamp/maximum/minimum propagate NaN, sign(NaN) = 0.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

#include "../../include/srmhd.h"
#include "srmhd_dual.h"

namespace srmhd {

constexpr double TINY = 1e-300;  // hlld.py:29, c2p.py:29

// torch semantics for the element-wise helpers --------------------------------
// Every helper takes doubles or Dual numbers (srmhd_dual.h); on doubles each is
// the plain double expression.  Mixed arguments promote to the Dual type.
using ::fabs;
SRMHD_DEV double qnan() { return __builtin_nan(""); }
// torch.maximum / torch.minimum; at a tie of two Dual values the tangent is
// the mean of both (torch's convention for the gradient of a tie)
template <class R>
SRMHD_DEV R tie_mean(const R& a, const R& b) { return b; }
template <int K>
SRMHD_DEV Dual<K> tie_mean(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r = b;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = 0.5 * (a.d[k] + b.d[k]);
  return r;
}
template <class A, class B>
SRMHD_DEV auto tmax(const A& a, const B& b) -> decltype(a + b) {   // torch.maximum
  using R = decltype(a + b);
  if (val(a) == val(b)) return tie_mean(R(a), R(b));
  return (val(a) > val(b) || val(a) != val(a)) ? R(a) : R(b);
}
template <class A, class B>
SRMHD_DEV auto tmin(const A& a, const B& b) -> decltype(a + b) {   // torch.minimum
  using R = decltype(a + b);
  if (val(a) == val(b)) return tie_mean(R(a), R(b));
  return (val(a) < val(b) || val(a) != val(a)) ? R(a) : R(b);
}
template <class A, class B>
SRMHD_DEV auto clo(const A& x, const B& lo) -> decltype(x + lo) {  // clamp(min=)
  using R = decltype(x + lo);
  return val(x) < val(lo) ? R(lo) : R(x);
}
template <class A, class B>
SRMHD_DEV auto chi_(const A& x, const B& hi) -> decltype(x + hi) { // clamp(max=)
  using R = decltype(x + hi);
  return val(x) > val(hi) ? R(hi) : R(x);
}
template <class T>
SRMHD_DEV double tsign(const T& x) { return val(x) > 0.0 ? 1.0 : (val(x) < 0.0 ? -1.0 : 0.0); }
template <class T>
SRMHD_DEV T sq(const T& x) { return x * x; }                                       // x**2 == x*x
SRMHD_DEV bool finite_pos_or_zero(double x) { return x >= 0.0 && x <= DBL_MAX; }

// Correctly rounded square root.  gfx950 expands f64 sqrt as v_rsq_f64 plus
// two Newton steps, which can be one ulp off; the CPU's sqrt is correctly
// rounded.  Tuckerman's test picks the right neighbour: y = RN(sqrt(x)) iff
// y*pred(y) < x <= y*succ(y), and the sign of each exact product difference
// is the sign of one fma.  Inputs below 2^-900 (where the fma residual could
// underflow), zero, inf and NaN keep the hardware result.
SRMHD_DEV double csqrt(double x) {
  double y = __builtin_sqrt(x);
  if (!(x >= 0x1p-900 && x <= DBL_MAX)) return y;
  const long long b = __builtin_bit_cast(long long, y);
  const double up = __builtin_bit_cast(double, b + 1);
  const double dn = __builtin_bit_cast(double, b - 1);
  if (__builtin_fma(y, up, -x) < 0.0) {
    y = up;
  } else if (__builtin_fma(y, dn, -x) >= 0.0) {
    y = dn;
  }
  return y;
}
// n / d correctly rounded, where the compiler's IEEE f64 division scales
// nothing: its own instructions (reciprocal seed, two Newton steps, the
// quotient and one residual correction) without the scale and fix-up steps,
// which are identities for normal operands with 2^-900 < |n|, |d|, |n/d|
// < 2^900 (then 1/d is normal too).  Callers establish that range; the
// result is then the IEEE quotient bit for bit.  The host build divides.
SRMHD_DEV double div_nr(double n, double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q = n * r;
  e = __builtin_fma(-d, q, n);
  return __builtin_fma(e, r, q);
#else
  return n / d;
#endif
}

// n / d correctly rounded where 2^-100 <= |d| <= 2^100 and |n| <= 2^800:
// div_nr when |n| >= 2^-800 (the quotient is then inside div_nr's range) or
// n is a zero, the compiler's IEEE division for the rare tiny n.  A zero n
// gives a zero whose sign is IEEE's except for n = -0 with d > 0 (+0 here,
// -0 in IEEE): callers have no -0 dividend, or use the quotient only where
// the sign of a zero cannot matter.
SRMHD_DEV double div_lo(double n, double d) {
  if (fabs(n) >= 0x1p-800 || n == 0.0) return div_nr(n, d);
  return n / d;
}

// torch.maximum / torch.minimum (tmax, tmin) and clamp(min=) / clamp(max=)
// (clo, chi_) for operands that are not NaN and not zeros of opposite sign:
// the hardware max / min then returns the same double (for equal operands,
// the one value).  Callers establish both.
SRMHD_DEV double vmax(double a, double b) { return __builtin_fmax(a, b); }
SRMHD_DEV double vmin(double a, double b) { return __builtin_fmin(a, b); }

// csqrt(x) for 2^-760 <= x <= 2^900, bit for bit: the compiler's f64 sqrt
// expansion without its argument scaling (an identity in this range) --
// the hardware reciprocal square root and three Newton corrections -- then
// the same Tuckerman test.
SRMHD_DEV double csqrt_mid(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rsq(x);
  double g = x * r, h = 0.5 * r;
  const double e = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  double y = __builtin_fma(d, h, g);
  const long long b = __builtin_bit_cast(long long, y);
  const double up = __builtin_bit_cast(double, b + 1);
  const double dn = __builtin_bit_cast(double, b - 1);
  if (__builtin_fma(y, up, -x) < 0.0) {
    y = up;
  } else if (__builtin_fma(y, dn, -x) >= 0.0) {
    y = dn;
  }
  return y;
#else
  return csqrt(x);
#endif
}

// all active lanes of the wave agree that p holds (the host build: p)
SRMHD_DEV bool wave_all(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __all((int)p) != 0;
#else
  return p;
#endif
}

template <int K>
SRMHD_DEV Dual<K> csqrt(const Dual<K>& x) {
  Dual<K> r;
  r.v = csqrt(x.v);
  const double h = 0.5 / r.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = x.d[k] * h;
  return r;
}

SRMHD_DEV double tpow(double x, double g) { return pow(x, g); }
template <int K>
SRMHD_DEV Dual<K> tpow(const Dual<K>& x, double g) {
  Dual<K> r;
  r.v = pow(x.v, g);
  const double dp = g * pow(x.v, g - 1.0);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = x.d[k] * dp;
  return r;
}

// _sdiv: |den| < tiny -> +-tiny (zero counts as +)   hlld.py:45-54
template <class A, class B>
SRMHD_DEV auto sdiv(const A& num, const B& den) -> decltype(num / den) {
  using R = decltype(num / den);
  if (fabs(val(den)) < TINY) return R(num / (val(den) < 0.0 ? -TINY : TINY));
  return num / den;
}

// ── EOS (eos.py:15-76) ──────────────────────────────────────────────────────
// K*rho**gamma is 0 for K = 0 and finite rho >= 0; the pow is only taken
// where it can change the result (K != 0, or rho negative / non-finite).
// K rho^gamma off the inlined path: pow expands to a few hundred
// instructions, and inlined the compiler evaluates it on every lane of a
// hot loop even where the K = 0 shortcut below applies to all of them
__host__ __device__ __noinline__ inline double cold_press_pow(double K, double rho, double g) {
  return K * pow(rho, g);
}
SRMHD_DEV double cold_press(const srmhd_eos& e, double rho) {
  return cold_press_pow(e.K, rho, e.gamma);
}
template <int K>
SRMHD_DEV Dual<K> cold_press(const srmhd_eos& e, const Dual<K>& rho) {
  return e.K * tpow(rho, e.gamma);
}

// 0 * rho^gamma is +0 wherever rho^gamma is finite: rho in [0, 1e100] with
// 0 <= gamma <= 3 (the shortcut of the reference's K * rho**gamma at K = 0)
template <class T>
SRMHD_DEV void eos_cold(const srmhd_eos& e, const T& rho, T& pc, T& ec) {
  const bool zero_k = e.K == 0.0 && e.gamma >= 0.0 && e.gamma <= 3.0;
  if (zero_k && val(rho) >= 0.0 && val(rho) <= 1e100) {
    pc = T(0.0);
  } else {
    pc = cold_press(e, rho);
  }
  const T den = rho * (e.gamma - 1.0);
  ec = (val(pc) == 0.0 && val(den) > 0.0 && val(den) <= DBL_MAX) ? T(0.0) : T(pc / den);
}

template <class T>
SRMHD_DEV T eos_press(const srmhd_eos& e, const T& eps_in, const T& rho) {   // eos.py:32-39
  T pc, ec;
  eos_cold(e, rho, pc, ec);
  const T eps = tmax(eps_in, ec);
  if (e.is_cold) return pc;
  return pc + (eps - ec) * rho * (e.gamma_th - 1.0);
}

template <class T>
SRMHD_DEV T eos_eps(const srmhd_eos& e, const T& p, const T& rho) {         // eos.py:41-46
  T pc, ec;
  eos_cold(e, rho, pc, ec);
  if (e.is_cold) return ec;
  return ec + (p - pc) / (rho * (e.gamma_th - 1.0));
}

template <class T>
SRMHD_DEV T eos_eps_lo(const srmhd_eos& e, const T& rho) {                  // eos.py:48-51
  T pc, ec;
  eos_cold(e, rho, pc, ec);
  return ec;
}

template <class T>
SRMHD_DEV void eos_press_cs2(const srmhd_eos& e, const T& eps, const T& rho, T& p,
                             T& cs2) {                                      // eos.py:65-76
  p = eos_press(e, eps, rho);
  const T h = 1.0 + eps + p / rho;
  if (e.is_cold) {
    cs2 = e.gamma * p / (rho * h);
  } else {
    T pc, ec;
    eos_cold(e, rho, pc, ec);
    cs2 = (e.gamma * pc + e.gamma_th * (p - pc)) / (rho * h);
  }
}

// ── states ──────────────────────────────────────────────────────────────────
template <class T>
struct PrimT {
  T rho, v[3], p, eps, B[3];
};
using Prim = PrimT<double>;

// conserved / flux vector order: D, tau, Sx, Sy, Sz, Bx, By, Bz (hlld.py:298-332)
enum { CD = 0, CT = 1, CS = 2, CB = 5 };

// W from the clamped v^2   hlld.py:57-62
template <class T>
SRMHD_DEV T lorentz_W(const T v[3]) {
  const T v2 = sq(v[0]) + sq(v[1]) + sq(v[2]);
  return 1.0 / csqrt(1.0 - chi_(v2, 1.0 - 1e-10));
}

// hlld.py:84-90
template <class T>
SRMHD_DEV T b_squared(const T v[3], const T B[3], const T& W) {
  const T B2 = sq(B[0]) + sq(B[1]) + sq(B[2]);
  const T vdB = v[0] * B[0] + v[1] * B[1] + v[2] * B[2];
  return B2 / (W * W) + sq(vdB);
}

// Roots of the dispersion quadratic with one_over_alp2 = 1, betad = 0,
// gupdd = 1 kept in the arithmetic   hlld.py:100-113
template <class T>
SRMHD_DEV void char_speeds(const T& v02, const T& u0, const T& vd, T& cp, T& cm) {
  const T u0sq = u0 * u0;
  T a = u0sq * (1.0 - v02) + v02 * 1.0;
  const T b = 2.0 * ((0.0 * 1.0) * v02 - u0sq * vd * (1.0 - v02));
  const T c = u0sq * sq(vd) * (1.0 - v02) - v02 * 1.0;
  const T det = csqrt(clo(b * b - 4.0 * a * c, 0.0));
  a = clo(a, TINY);
  const T r1 = 0.5 * (det - b) / a;
  const T r2 = -0.5 * (det + b) / a;
  cp = tmax(r1, r2);
  cm = tmin(r1, r2);
}

// One side of compute_srmhd_fluxes   hlld.py:223-334
template <int N, class T>
SRMHD_DEV void side_state(const PrimT<T>& P, const srmhd_eos& e, T U[8], T F[8], T& cp, T& cm) {
  const T vn = P.v[N];
  const T W = lorentz_W(P.v);
  T p, cs2;
  eos_press_cs2(e, P.eps, P.rho, p, cs2);
  const T b2 = b_squared(P.v, P.B, W);
  const T vdB = P.v[0] * P.B[0] + P.v[1] * P.B[1] + P.v[2] * P.B[2];
  const T bt = W * vdB;
  T b[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i] = P.B[i] / W + vdB * P.v[i] * W;
  const T h = 1.0 + P.eps + p / P.rho;
  const T v02 = cs2 + (b2 / (P.rho * h + b2)) * (1.0 - cs2);
  char_speeds(v02, W, vn, cp, cm);
  const T d = P.rho * W;
  const T rhb2 = P.rho * h + b2;
  const T Pb2 = p + 0.5 * b2;
  const T bn = b[N];
  U[CD] = d;
  F[CD] = d * vn;
  const T Tt = rhb2 * (W * W) + Pb2 * (-1.0) - bt * bt;
  U[CT] = Tt - d;
  F[CT] = (rhb2 * (W * W) * vn - bt * bn) - d * vn;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const T ui = W * P.v[i];
    const double delta = (i == N) ? 1.0 : 0.0;
    U[CS + i] = rhb2 * W * ui - bt * b[i];
    F[CS + i] = rhb2 * (W * vn) * ui + Pb2 * delta - bn * b[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    U[CB + i] = P.B[i];
    F[CB + i] = vn * P.B[i] - P.v[i] * P.B[N];
  }
}

// hlld.py:161-168
template <class T>
SRMHD_DEV T hll_flux(const T& fl, const T& fr, const T& ul, const T& ur, const T& cm, const T& cx) {
  return (cx * fl + cm * fr - cx * cm * (ur - ul)) / (cx + cm);
}
template <class T>
SRMHD_DEV T hll_state(const T& fl, const T& fr, const T& ul, const T& ur, const T& cm, const T& cx) {
  return (cx * ur + cm * ul + fl - fr) / (cx + cm);
}

// Both sides plus the clamped HLL speeds   hlld.py:267-276
// (cps, cms: each side's speeds c+, c-, left first)
template <int N, class T>
SRMHD_DEV void srmhd_fluxes(const PrimT<T>& L, const PrimT<T>& R, const srmhd_eos& e, T uL[8],
                            T uR[8], T fL[8], T fR[8], T& cmax, T& cmin, T cps[2], T cms[2]) {
  side_state<N>(L, e, uL, fL, cps[0], cms[0]);
  side_state<N>(R, e, uR, fR, cps[1], cms[1]);
  cmax = clo(tmax(cps[0], cps[1]), 0.0);
  cmin = clo(tmax(-cms[0], -cms[1]), 0.0);
  if (cmax < 1e-12 && cmin < 1e-12) {
    cmax = T(1.0);
    cmin = T(1.0);
  }
}
template <int N, class T>
SRMHD_DEV void srmhd_fluxes(const PrimT<T>& L, const PrimT<T>& R, const srmhd_eos& e, T uL[8],
                            T uR[8], T fL[8], T fR[8], T& cmax, T& cmin) {
  T cps[2], cms[2];
  srmhd_fluxes<N>(L, R, e, uL, uR, fL, fR, cmax, cmin, cps, cms);
}

// primitive_to_conserved   hlld.py:185-217
template <class T>
SRMHD_DEV void prim_to_cons(const PrimT<T>& P, T c[8]) {
  const T v2 = sq(P.v[0]) + sq(P.v[1]) + sq(P.v[2]);
  const T W = 1.0 / csqrt(clo(1.0 - v2, 1e-10));
  const T vB = P.v[0] * P.B[0] + P.v[1] * P.B[1] + P.v[2] * P.B[2];
  const T B2 = sq(P.B[0]) + sq(P.B[1]) + sq(P.B[2]);
  const T h = 1.0 + P.eps + P.p / P.rho;
  const T rhW2 = P.rho * h * (W * W);
  const T D = P.rho * W;
  c[CD] = D;
#pragma unroll
  for (int i = 0; i < 3; ++i) c[CS + i] = (rhW2 + B2) * P.v[i] - vB * P.B[i];
  c[CT] = rhW2 - P.p - D + 0.5 * (B2 + B2 * v2 - vB * vB);
#pragma unroll
  for (int i = 0; i < 3; ++i) c[CB + i] = P.B[i];
}

}  // namespace srmhd
