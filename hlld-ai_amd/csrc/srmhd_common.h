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

#define SRMHD_DEV __host__ __device__ __forceinline__

namespace srmhd {

constexpr double TINY = 1e-300;  // hlld.py:29, c2p.py:29

// torch semantics for the element-wise helpers --------------------------------
SRMHD_DEV double qnan() { return __builtin_nan(""); }
SRMHD_DEV double tmax(double a, double b) { return (a > b || a != a) ? a : b; }   // torch.maximum
SRMHD_DEV double tmin(double a, double b) { return (a < b || a != a) ? a : b; }   // torch.minimum
SRMHD_DEV double clo(double x, double lo) { return x < lo ? lo : x; }             // clamp(min=)
SRMHD_DEV double chi_(double x, double hi) { return x > hi ? hi : x; }            // clamp(max=)
SRMHD_DEV double tsign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }
SRMHD_DEV double sq(double x) { return x * x; }                                    // x**2 == x*x
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

// _sdiv: |den| < tiny -> +-tiny (zero counts as +)   hlld.py:45-54
SRMHD_DEV double sdiv(double num, double den) {
  if (fabs(den) < TINY) return num / (den < 0.0 ? -TINY : TINY);
  return num / den;
}

// ── EOS (eos.py:15-76) ──────────────────────────────────────────────────────
// K*rho**gamma is 0 for K = 0 and finite rho >= 0; the pow is only taken
// where it can change the result (K != 0, or rho negative / non-finite).
SRMHD_DEV void eos_cold(const srmhd_eos& e, double rho, double& pc, double& ec) {
  if (e.K == 0.0 && finite_pos_or_zero(rho)) {
    pc = 0.0;
  } else {
    pc = e.K * pow(rho, e.gamma);
  }
  const double den = rho * (e.gamma - 1.0);
  ec = (pc == 0.0 && den > 0.0 && den <= DBL_MAX) ? 0.0 : pc / den;
}

SRMHD_DEV double eos_press(const srmhd_eos& e, double eps, double rho) {   // eos.py:32-39
  double pc, ec;
  eos_cold(e, rho, pc, ec);
  eps = tmax(eps, ec);
  if (e.is_cold) return pc;
  return pc + (eps - ec) * rho * (e.gamma_th - 1.0);
}

SRMHD_DEV double eos_eps(const srmhd_eos& e, double p, double rho) {      // eos.py:41-46
  double pc, ec;
  eos_cold(e, rho, pc, ec);
  if (e.is_cold) return ec;
  return ec + (p - pc) / (rho * (e.gamma_th - 1.0));
}

SRMHD_DEV double eos_eps_lo(const srmhd_eos& e, double rho) {             // eos.py:48-51
  double pc, ec;
  eos_cold(e, rho, pc, ec);
  return ec;
}

SRMHD_DEV void eos_press_cs2(const srmhd_eos& e, double eps, double rho, double& p,
                             double& cs2) {                                 // eos.py:65-76
  p = eos_press(e, eps, rho);
  const double h = 1.0 + eps + p / rho;
  if (e.is_cold) {
    cs2 = e.gamma * p / (rho * h);
  } else {
    double pc, ec;
    eos_cold(e, rho, pc, ec);
    cs2 = (e.gamma * pc + e.gamma_th * (p - pc)) / (rho * h);
  }
}

// ── states ──────────────────────────────────────────────────────────────────
struct Prim {
  double rho, v[3], p, eps, B[3];
};

// conserved / flux vector order: D, tau, Sx, Sy, Sz, Bx, By, Bz (hlld.py:298-332)
enum { CD = 0, CT = 1, CS = 2, CB = 5 };

// W from the clamped v^2   hlld.py:57-62
SRMHD_DEV double lorentz_W(const double v[3]) {
  const double v2 = sq(v[0]) + sq(v[1]) + sq(v[2]);
  return 1.0 / csqrt(1.0 - chi_(v2, 1.0 - 1e-10));
}

// hlld.py:84-90
SRMHD_DEV double b_squared(const double v[3], const double B[3], double W) {
  const double B2 = sq(B[0]) + sq(B[1]) + sq(B[2]);
  const double vdB = v[0] * B[0] + v[1] * B[1] + v[2] * B[2];
  return B2 / (W * W) + sq(vdB);
}

// Roots of the dispersion quadratic with one_over_alp2 = 1, betad = 0,
// gupdd = 1 kept in the arithmetic   hlld.py:100-113
SRMHD_DEV void char_speeds(double v02, double u0, double vd, double& cp, double& cm) {
  const double u0sq = u0 * u0;
  double a = u0sq * (1.0 - v02) + v02 * 1.0;
  const double b = 2.0 * ((0.0 * 1.0) * v02 - u0sq * vd * (1.0 - v02));
  const double c = u0sq * sq(vd) * (1.0 - v02) - v02 * 1.0;
  const double det = csqrt(clo(b * b - 4.0 * a * c, 0.0));
  a = clo(a, TINY);
  const double r1 = 0.5 * (det - b) / a;
  const double r2 = -0.5 * (det + b) / a;
  cp = tmax(r1, r2);
  cm = tmin(r1, r2);
}

// One side of compute_srmhd_fluxes   hlld.py:223-334
template <int N>
SRMHD_DEV void side_state(const Prim& P, const srmhd_eos& e, double U[8], double F[8],
                          double& cp, double& cm) {
  const double vn = P.v[N];
  const double W = lorentz_W(P.v);
  double p, cs2;
  eos_press_cs2(e, P.eps, P.rho, p, cs2);
  const double b2 = b_squared(P.v, P.B, W);
  const double vdB = P.v[0] * P.B[0] + P.v[1] * P.B[1] + P.v[2] * P.B[2];
  const double bt = W * vdB;
  double b[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i] = P.B[i] / W + vdB * P.v[i] * W;
  const double h = 1.0 + P.eps + p / P.rho;
  const double v02 = cs2 + (b2 / (P.rho * h + b2)) * (1.0 - cs2);
  char_speeds(v02, W, vn, cp, cm);
  const double d = P.rho * W;
  const double rhb2 = P.rho * h + b2;
  const double Pb2 = p + 0.5 * b2;
  const double bn = b[N];
  U[CD] = d;
  F[CD] = d * vn;
  const double Tt = rhb2 * (W * W) + Pb2 * (-1.0) - bt * bt;
  U[CT] = Tt - d;
  F[CT] = (rhb2 * (W * W) * vn - bt * bn) - d * vn;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double ui = W * P.v[i];
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
SRMHD_DEV double hll_flux(double fl, double fr, double ul, double ur, double cm, double cx) {
  return (cx * fl + cm * fr - cx * cm * (ur - ul)) / (cx + cm);
}
SRMHD_DEV double hll_state(double fl, double fr, double ul, double ur, double cm, double cx) {
  return (cx * ur + cm * ul + fl - fr) / (cx + cm);
}

// Both sides plus the clamped HLL speeds   hlld.py:267-276
template <int N>
SRMHD_DEV void srmhd_fluxes(const Prim& L, const Prim& R, const srmhd_eos& e, double uL[8],
                            double uR[8], double fL[8], double fR[8], double& cmax,
                            double& cmin) {
  double cpl, cml, cpr, cmr;
  side_state<N>(L, e, uL, fL, cpl, cml);
  side_state<N>(R, e, uR, fR, cpr, cmr);
  cmax = clo(tmax(cpl, cpr), 0.0);
  cmin = clo(tmax(-cml, -cmr), 0.0);
  if (cmax < 1e-12 && cmin < 1e-12) {
    cmax = 1.0;
    cmin = 1.0;
  }
}

// primitive_to_conserved   hlld.py:185-217
SRMHD_DEV void prim_to_cons(const Prim& P, double c[8]) {
  const double v2 = sq(P.v[0]) + sq(P.v[1]) + sq(P.v[2]);
  const double W = 1.0 / csqrt(clo(1.0 - v2, 1e-10));
  const double vB = P.v[0] * P.B[0] + P.v[1] * P.B[1] + P.v[2] * P.B[2];
  const double B2 = sq(P.B[0]) + sq(P.B[1]) + sq(P.B[2]);
  const double h = 1.0 + P.eps + P.p / P.rho;
  const double rhW2 = P.rho * h * (W * W);
  const double D = P.rho * W;
  c[CD] = D;
#pragma unroll
  for (int i = 0; i < 3; ++i) c[CS + i] = (rhW2 + B2) * P.v[i] - vB * P.B[i];
  c[CT] = rhW2 - P.p - D + 0.5 * (B2 + B2 * v2 - vB * vB);
#pragma unroll
  for (int i = 0; i < 3; ++i) c[CB + i] = P.B[i];
}

}  // namespace srmhd
