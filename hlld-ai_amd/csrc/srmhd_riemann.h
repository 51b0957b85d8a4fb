// HLLE / HLLC / HLLD point-wise fluxes (hlld.py:1005-1765), templated on
// the normal direction N so every component index is a compile-time
// constant (no register-array spills).
#pragma once

#include "srmhd_c2p.h"

namespace srmhd {

enum Solver { SOLVER_HLLE = 0, SOLVER_HLLC = 1, SOLVER_HLLD = 2 };
enum FaceStatus {
  FS_NOT_CONVERGED = 1, FS_WAVE_ORDER = 2, FS_BN_ZERO = 4, FS_FALLBACK = 8,
  FS_C2P_REDONE = 16   // the HLL-state con2prim needed the bisection continuation
};

constexpr double PSTAR_TOL = 1.0e-10;   // hlld.py:1080
constexpr int PSTAR_MAXIT = 30;         // hlld.py:1327

// HLL total-pressure estimate of hlle/hllc   hlld.py:1036-1042
template <class T>
SRMHD_DEV T p_hll_sides(const PrimT<T>& L, const PrimT<T>& R, const srmhd_eos& e) {
  T pl, pr, cs;
  eos_press_cs2(e, L.eps, L.rho, pl, cs);
  eos_press_cs2(e, R.eps, R.rho, pr, cs);
  const T b2l = b_squared(L.v, L.B, lorentz_W(L.v));
  const T b2r = b_squared(R.v, R.B, lorentz_W(R.v));
  return clo(0.5 * (pl + 0.5 * b2l + pr + 0.5 * b2r), 1e-30);
}

// hlld.py:1005-1044
template <int N, class T>
SRMHD_DEV void hlle(const PrimT<T>& L, const PrimT<T>& R, const srmhd_eos& e, T F[8], T U[8],
                    T& pstar) {
  T uL[8], uR[8], fL[8], fR[8], cmax, cmin;
  srmhd_fluxes<N>(L, R, e, uL, uR, fL, fR, cmax, cmin);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const T fh = hll_flux(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
    const T uh = hll_state(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
    F[k] = (cmax <= 0.0) ? fR[k] : ((cmin <= 0.0) ? fL[k] : fh);
    U[k] = (cmin <= 0.0) ? uL[k] : ((cmax <= 0.0) ? uR[k] : uh);
  }
  pstar = p_hll_sides(L, R, e);
}

// Mignone-style HLLC   hlld.py:1549-1765
template <int N, class T>
SRMHD_DEV void hllc(const PrimT<T>& L, const PrimT<T>& R, const srmhd_eos& e, T F[8], T U[8],
                    T& pstar) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  T uL[8], uR[8], fL[8], fR[8], cmax, cmin;
  srmhd_fluxes<N>(L, R, e, uL, uR, fL, fR, cmax, cmin);
  const T SL = -cmin, SR = cmax;
  T fH[8], uH[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    fH[k] = hll_flux(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
    uH[k] = hll_state(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
  }
  const T Bn = uL[CB + N];
  const bool hasB = fabs(Bn) > 1e-15;
  const T bt1 = uH[CB + T1], bt2 = uH[CB + T2];
  const T fb1 = fH[CB + T1], fb2 = fH[CB + T2];
  const T cross = hasB ? T(bt1 * fb1 + bt2 * fb2) : T(0.0);
  const T btan2 = hasB ? T(sq(bt1) + sq(bt2)) : T(0.0);
  const T fbtan2 = hasB ? T(sq(fb1) + sq(fb2)) : T(0.0);
  const T E = uH[CT] + uH[CD];
  const T fE = fH[CT] + fH[CD];
  const T qa = fE - cross;
  const T qb = -fH[CS + N] - E + btan2 + fbtan2;
  const T qc = uH[CS + N] - cross;
  const T disc = csqrt(clo(qb * qb - 4.0 * qa * qc, 0.0));
  const T vq = sdiv(-0.5 * (qb + disc), qa);
  const T vlin = sdiv(-qc, qb);
  const T vc = (fabs(qa) < 1e-15) ? ((fabs(qb) < 1e-15) ? T(0.0) : vlin) : vq;
  const T vt1 = hasB ? T(sdiv(uH[CB + T1] * vc - fH[CB + T1], Bn)) : T(0.0);
  const T vt2 = hasB ? T(sdiv(uH[CB + T2] * vc - fH[CB + T2], Bn)) : T(0.0);
  const T v2 = sq(vc) + sq(vt1) + sq(vt2);
  const T g = 1.0 / csqrt(1.0 - chi_(v2, 1.0 - 1e-10));
  const T vB = hasB ? T(vc * uH[CB + N] + vt1 * uH[CB + T1] + vt2 * uH[CB + T2]) : T(0.0);
  const T pB = fH[CS + N] + sq(Bn / g) - vc * (fH[CT] + fH[CD] - Bn * vB);
  const T p0 = fH[CS + N] - vc * (fH[CT] + fH[CD]);
  const T ps = hasB ? pB : p0;

  T ucL[8], ucR[8];
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const T* u = side ? uR : uL;
    const T* f = side ? fR : fL;
    T* s = side ? ucR : ucL;
    const T S = side ? SR : SL;
    const T vd = side ? R.v[N] : L.v[N];
    const T den = S - vc;
    s[CB + N] = Bn;
    s[CB + T1] = uH[CB + T1];
    s[CB + T2] = uH[CB + T2];
    s[CD] = sdiv(u[CD] * (S - vd), den);
    s[CT] = sdiv(S * u[CT] - f[CT] + ps * vc - vB * Bn, den);
    s[CS + N] = (s[CT] + s[CD] + ps) * vc - vB * Bn;
    const T w1 = sdiv(-Bn * (uH[CB + T1] / (g * g) + vB * vt1) - f[CS + T1] + S * u[CS + T1], den);
    const T n1 = sdiv(u[CS + T1] * (S - vd), den);
    s[CS + T1] = hasB ? w1 : n1;
    const T w2 = sdiv(-Bn * (uH[CB + T2] / (g * g) + vB * vt2) - f[CS + T2] + S * u[CS + T2], den);
    const T n2 = sdiv(u[CS + T2] * (S - vd), den);
    s[CS + T2] = hasB ? w2 : n2;
  }
  const bool degen = (vc <= SL) || (vc >= SR) || (v2 >= 1.0);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const T fsL = fL[k] + SL * (ucL[k] - uL[k]);
    const T fsR = fR[k] + SR * (ucR[k] - uR[k]);
    T f = fH[k], u = uH[k];
    if (0.0 <= SL) { f = fL[k]; u = uL[k]; }
    if ((SL < 0.0) && (0.0 < vc)) { f = fsL; u = ucL[k]; }
    if ((vc <= 0.0) && (0.0 < SR)) { f = fsR; u = ucR[k]; }
    if (0.0 >= SR) { f = fR[k]; u = uR[k]; }
    if (degen) { f = fH[k]; u = uH[k]; }
    F[k] = f;
    U[k] = u;
  }
  // non-finite safety, per key on (F, U) jointly   hlld.py:1750-1754
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bool bad = !(fabs(val(F[k])) <= DBL_MAX) || !(fabs(val(U[k])) <= DBL_MAX);
    if (bad) { F[k] = fH[k]; U[k] = uH[k]; }
  }
  pstar = p_hll_sides(L, R, e);
}

// ── HLLD (hlld.py:366-605, 669-746, 1202-1543) ──────────────────────────────

// R-vectors of HLLDComputation._build_R on the total-energy form
template <class T>
struct StarRT {
  T RSn[2], RS1[2], RS2[2], RBn[2], RB1[2], RB2[2], RT[2], RD[2];
  double Sg[2];
  T Bn[2], cm, cx;
};
using StarR = StarRT<double>;

// compute_all_variables at a trial total pressure p   hlld.py:431-583
template <class T>
struct StarOutT {
  T va[2], v1[2], v2[2], B1[2], B2[2], sq[2], Da[2], Ta[2], Sa[2], S1[2], S2[2];
  T la[2], K1L, K2L, Bc1, Bc2, KL2, KR2, KBL, KBR;
};
using StarOut = StarOutT<double>;

template <class T>
SRMHD_DEV void star_solve(const StarRT<T>& R, const T& p, StarOutT<T>& o) {
  const T cm = R.cm, cx = R.cx;
  const T BnL = R.Bn[0], BnR = R.Bn[1];
  const T RSnL = R.RSn[0], RSnR = R.RSn[1], RS1L = R.RS1[0], RS1R = R.RS1[1];
  const T RS2L = R.RS2[0], RS2R = R.RS2[1], RB1L = R.RB1[0], RB1R = R.RB1[1];
  const T RB2L = R.RB2[0], RB2R = R.RB2[1], RTL = R.RT[0], RTR = R.RT[1];

  const T AL = RSnL - (-cm) * RTL + p * (1.0 - cm * cm);
  const T AR = RSnR - cx * RTR + p * (1.0 - cx * cx);
  const T GL = sq(RB1L) + sq(RB2L);
  const T GR = sq(RB1R) + sq(RB2R);
  const T CL = RS1L * RB1L + RS2L * RB2L;
  const T CR = RS1R * RB1R + RS2R * RB2R;
  const T QL = -AL - GL + sq(BnL) * (1.0 - cm * cm);
  const T QR = -AR - GR + sq(BnR) * (1.0 - cx * cx);
  const T XL = BnL * (AL * (-cm) * BnL + CL) - (AL + GL) * ((-cm) * p + RTL);
  const T XR = BnR * (AR * cx * BnR + CR) - (AR + GR) * (cx * p + RTR);

  const T vaL = sdiv(BnL * (AL * BnL + (-cm) * CL) - (AL + GL) * (p + RSnL), XL);
  const T vaR = sdiv(BnR * (AR * BnR + cx * CR) - (AR + GR) * (p + RSnR), XR);
  const T v1L = sdiv(QL * RS1L + RB1L * (CL + BnL * ((-cm) * RSnL - RTL)), XL);
  const T v1R = sdiv(QR * RS1R + RB1R * (CR + BnR * (cx * RSnR - RTR)), XR);
  const T v2L = sdiv(QL * RS2L + RB2L * (CL + BnL * ((-cm) * RSnL - RTL)), XL);
  const T v2R = sdiv(QR * RS2R + RB2R * (CR + BnR * (cx * RSnR - RTR)), XR);

  const T dLa = -cm - vaL, dRa = cx - vaR;
  const T B1L = sdiv(RB1L - BnL * v1L, dLa);
  const T B1R = sdiv(RB1R - BnR * v1R, dRa);
  const T B2L = sdiv(RB2L - BnL * v2L, dLa);
  const T B2R = sdiv(RB2R - BnR * v2R, dRa);

  const T vRSL = vaL * RSnL + v1L * RS1L + v2L * RS2L;
  const T vRSR = vaR * RSnR + v1R * RS1R + v2R * RS2R;
  const T waL = p + sdiv(RTL - vRSL, dLa);
  const T waR = p + sdiv(RTR - vRSR, dRa);
  const T DaL = sdiv(R.RD[0], dLa);
  const T DaR = sdiv(R.RD[1], dRa);
  const T vBL = vaL * BnL + v1L * B1L + v2L * B2L;
  const T vBR = vaR * BnR + v1R * B1R + v2R * B2R;
  const T TaL = sdiv(RTL + p * vaL - vBL * BnL, dLa) - DaL;
  const T TaR = sdiv(RTR + p * vaR - vBR * BnR, dRa) - DaR;
  o.Sa[0] = (TaL + DaL + p) * vaL - vBL * BnL;
  o.Sa[1] = (TaR + DaR + p) * vaR - vBR * BnR;
  o.S1[0] = (TaL + DaL + p) * v1L - vBL * B1L;
  o.S1[1] = (TaR + DaR + p) * v1R - vBR * B1R;
  o.S2[0] = (TaL + DaL + p) * v2L - vBL * B2L;
  o.S2[1] = (TaR + DaR + p) * v2R - vBR * B2R;

  const T sqL = csqrt(clo(waL, 0.0));
  const T sqR = csqrt(clo(waR, 0.0));
  const double SgL = R.Sg[0], SgR = R.Sg[1];  // signs: no tangent
  const T dL = (-cm) * p + RTL - BnL * SgL * sqL;
  const T dR = cx * p + RTR + BnR * SgR * sqR;
  const T KnL = sdiv(RSnL + p - R.RBn[0] * SgL * sqL, dL);
  const T KnR = sdiv(RSnR + p + R.RBn[1] * SgR * sqR, dR);
  const T K1L = sdiv(RS1L - RB1L * SgL * sqL, dL);
  const T K1R = sdiv(RS1R + RB1R * SgR * sqR, dR);
  const T K2L = sdiv(RS2L - RB2L * SgL * sqL, dL);
  const T K2R = sdiv(RS2R + RB2R * SgR * sqR, dR);

  const T dll = KnR - KnL;
  const T Bc1 = sdiv((B1R * (KnR - vaR) + BnR * v1R) - (B1L * (KnL - vaL) + BnL * v1L), dll);
  const T Bc2 = sdiv((B2R * (KnR - vaR) + BnR * v2R) - (B2L * (KnL - vaL) + BnL * v2L), dll);
  const T Bcn = BnR;

  o.va[0] = vaL; o.va[1] = vaR;
  o.v1[0] = v1L; o.v1[1] = v1R;
  o.v2[0] = v2L; o.v2[1] = v2R;
  o.B1[0] = B1L; o.B1[1] = B1R;
  o.B2[0] = B2L; o.B2[1] = B2R;
  o.sq[0] = sqL; o.sq[1] = sqR;
  o.Da[0] = DaL; o.Da[1] = DaR;
  o.Ta[0] = TaL; o.Ta[1] = TaR;
  o.la[0] = KnL; o.la[1] = KnR;
  o.K1L = K1L; o.K2L = K2L;
  o.Bc1 = Bc1; o.Bc2 = Bc2;
  o.KL2 = sq(KnL) + sq(K1L) + sq(K2L);
  o.KR2 = sq(KnR) + sq(K1R) + sq(K2R);
  o.KBL = KnL * Bcn + K1L * Bc1 + K2L * Bc2;
  o.KBR = KnR * Bcn + K1R * Bc1 + K2R * Bc2;
}

// HLLDComputation.residual   hlld.py:586-602
template <class T>
SRMHD_DEV T star_residual(const StarRT<T>& R, const T& p) {
  StarOutT<T> o;
  star_solve(R, p, o);
  const T dR = R.Sg[1] * o.sq[1] - o.KBR;
  const T dL = R.Sg[0] * o.sq[0] + o.KBL;
  return (o.la[0] - o.la[1]) + R.Bn[0] * (sdiv(1.0 - o.KR2, dR) + sdiv(1.0 - o.KL2, dL));
}

// Everything of hlld_flux that precedes the con2prim of the HLL state.
template <int N, class T = double>
struct HlldPrep {
  T uL[8], uR[8], fL[8], fR[8], cmax, cmin, uH[8];
  SRMHD_DEV void run(const PrimT<T>& L, const PrimT<T>& R, const srmhd_eos& e) {
    T cps[2], cms[2];
    run(L, R, e, cps, cms);
  }
  // ... keeping each side's speeds c+ (cps) and c- (cms), left first
  SRMHD_DEV void run(const PrimT<T>& L, const PrimT<T>& R, const srmhd_eos& e, T cps[2],
                     T cms[2]) {
    srmhd_fluxes<N>(L, R, e, uL, uR, fL, fR, cmax, cmin, cps, cms);
#pragma unroll
    for (int k = 0; k < 8; ++k) uH[k] = hll_state(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
  }
};

// The side states and HLL speeds of a sweep's faces as HLLD's pass A forms
// them (HlldPrep::run), kept for its pass B, which would otherwise
// reconstruct the face states and form them again: HQ_N arrays of the
// sweep's nf faces -- uL, uR, fL, fR (8 each), cmax, cmin.  The HLL state
// uH is formed again on load (the same arithmetic as HlldPrep::run).
constexpr int HQ_N = 34;
template <int N>
SRMHD_DEV void side_q_store(double* hq, int64_t nf, int64_t t, const HlldPrep<N>& h) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    hq[(int64_t)k * nf + t] = h.uL[k];
    hq[(int64_t)(8 + k) * nf + t] = h.uR[k];
    hq[(int64_t)(16 + k) * nf + t] = h.fL[k];
    hq[(int64_t)(24 + k) * nf + t] = h.fR[k];
  }
  hq[(int64_t)32 * nf + t] = h.cmax;
  hq[(int64_t)33 * nf + t] = h.cmin;
}
template <int N>
SRMHD_DEV void side_q_load(const double* hq, int64_t nf, int64_t t, HlldPrep<N>& h) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    h.uL[k] = hq[(int64_t)k * nf + t];
    h.uR[k] = hq[(int64_t)(8 + k) * nf + t];
    h.fL[k] = hq[(int64_t)(16 + k) * nf + t];
    h.fR[k] = hq[(int64_t)(24 + k) * nf + t];
  }
  h.cmax = hq[(int64_t)32 * nf + t];
  h.cmin = hq[(int64_t)33 * nf + t];
#pragma unroll
  for (int k = 0; k < 8; ++k)
    h.uH[k] = hll_state(h.fL[k], h.fR[k], h.uL[k], h.uR[k], h.cmin, h.cmax);
}

// The face EMF flux of emf_mode "hll" (ct.py:63-88): the HLL average of the
// tangential-field flux the sweep keeps for the corner EMF (F^x(B_y) for
// N = 0, F^y(B_x) for N = 1) from the face's side states and HLL speeds
// (compute_srmhd_fluxes, sweep.py:106-118), whatever the solver.
template <int N, class T>
SRMHD_DEV T hll_emf_flux(const PrimT<T>& L, const PrimT<T>& R, const srmhd_eos& e) {
  static_assert(N == 0 || N == 1, "the 2D sweeps");
  T uL[8], uR[8], fL[8], fR[8], cmax, cmin;
  srmhd_fluxes<N>(L, R, e, uL, uR, fL, fR, cmax, cmin);
  constexpr int c = N == 0 ? CB + 1 : CB + 0;
  return hll_flux(fL[c], fR[c], uL[c], uR[c], cmin, cmax);
}

// Whether a face's HLL-state con2prim is inert in the 2D stage: it can then
// only reach the batch's break iteration, never the flux.  hlld_assemble
// falls back to the HLLE flux where |B_n| < 1e-14 on both sides (hlld.py:
// 1303, 1513) whatever p* is; p_hll (the con2prim's total pressure) enters
// only the p* bracket's use_p0 test, 0.5 (B_nL^2 + B_nR^2) / p_hll < 0.1
// (hlld.py:1295-1301), and the fallback sentinel -|p_hll| - 1e-10.  The test
// holds for every value the con2prim can return when the HLL state is
// moderate: with 1e-20 <= D <= 1e20 and every other component within 1e20
// the con2prim's arithmetic stays finite (mu in [0, 1], each term bounded),
// its velocity has v^2 <= 1 - 1e-10, so p_hll >= 0.5 b^2 >= 0.5 B^2 / W^2 >=
// 5e-11 B^2 (p >= 0), and B^2 >= 1e13 * 0.5 (B_nL^2 + B_nR^2), B^2 >= 1e-80
// put the ratio below 1e-2 with p_hll > 0.  The bracket then starts from p0,
// independent of p_hll, and the secant's iteration counts are the same.
template <int N>
SRMHD_DEV bool hll_c2p_inert(const HlldPrep<N>& h) {
  const double BnL = h.uL[CB + N], BnR = h.uR[CB + N];
  if (!((fabs(BnL) < 1e-14) && (fabs(BnR) < 1e-14))) return false;
  const double* u = h.uH;
  bool ok = u[CD] >= 1e-20 && u[CD] <= 1e20;
#pragma unroll
  for (int k = 1; k < 8; ++k) ok = ok && fabs(u[k]) <= 1e20;
  const double B2 = sq(u[CB]) + sq(u[CB + 1]) + sq(u[CB + 2]);
  const double Bn2 = 0.5 * (BnL * BnL + BnR * BnR);
  return ok && B2 >= 1e-80 && B2 >= 1e13 * Bn2;
}

// p_hll from the con2prim of the HLL state   hlld.py:1255-1263
template <class T>
SRMHD_DEV T p_hll_from_prim(const PrimT<T>& P, const srmhd_eos& e) {
  T p, cs;
  eos_press_cs2(e, P.eps, P.rho, p, cs);
  return p + 0.5 * b_squared(P.v, P.B, lorentz_W(P.v));
}

// R-vectors on the total-energy form (E = tau + D)   hlld.py:401-428
template <int N, class T>
SRMHD_DEV void hlld_rvec(const HlldPrep<N, T>& h, StarRT<T>& R) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  const T cmin = h.cmin, cmax = h.cmax;
  R.cm = cmin;
  R.cx = cmax;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const T* u = s ? h.uR : h.uL;
    const T* f = s ? h.fR : h.fL;
    const T uE = u[CT] + u[CD], fE = f[CT] + f[CD];
    if (s == 0) {
      R.RSn[0] = -cmin * u[CS + N] - f[CS + N];
      R.RS1[0] = -cmin * u[CS + T1] - f[CS + T1];
      R.RS2[0] = -cmin * u[CS + T2] - f[CS + T2];
      R.RBn[0] = -cmin * u[CB + N] - f[CB + N];
      R.RB1[0] = -cmin * u[CB + T1] - f[CB + T1];
      R.RB2[0] = -cmin * u[CB + T2] - f[CB + T2];
      R.RT[0] = -cmin * uE - fE;
      R.RD[0] = -cmin * u[CD] - f[CD];
    } else {
      R.RSn[1] = cmax * u[CS + N] - f[CS + N];
      R.RS1[1] = cmax * u[CS + T1] - f[CS + T1];
      R.RS2[1] = cmax * u[CS + T2] - f[CS + T2];
      R.RBn[1] = cmax * u[CB + N] - f[CB + N];
      R.RB1[1] = cmax * u[CB + T1] - f[CB + T1];
      R.RB2[1] = cmax * u[CB + T2] - f[CB + T2];
      R.RT[1] = cmax * uE - fE;
      R.RD[1] = cmax * u[CD] - f[CD];
    }
    R.Sg[s] = tsign(u[CB + N]);
    R.Bn[s] = u[CB + N];
  }
}

// The HLLD flux at total pressure ps (the Alfven and contact states, the
// region at xi = 0, the HLLE fallback)   hlld.py:1360-1534.  conv: the p*
// root finder's convergence flag.  Writes F (and U when non-null) and
// returns the face status bits; pstar gets ps, or the fallback sentinel
// -|p_hll| - 1e-10.
template <int N, class T>
SRMHD_DEV int hlld_assemble(const HlldPrep<N, T>& h, const StarRT<T>& R, const T fH[8],
                            const T& ps, bool conv, double p_hll, T F[8], T* U, T& pstar,
                            double* margins = nullptr) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  const T cmin = h.cmin, cmax = h.cmax;
  const T* uH = h.uH;
  const T BnL = h.uL[CB + N], BnR = h.uR[CB + N];
  const bool bn_zero = (fabs(BnL) < 1e-14) && (fabs(BnR) < 1e-14);

  StarOutT<T> o;
  star_solve(R, ps, o);

  // Alfven (a) states in the tau form, their fluxes   hlld.py:1360-1385
  T uaL[8], uaR[8], faL[8], faR[8];
  uaL[CB + N] = BnL; uaL[CB + T1] = o.B1[0]; uaL[CB + T2] = o.B2[0];
  uaL[CD] = o.Da[0]; uaL[CT] = o.Ta[0];
  uaL[CS + N] = o.Sa[0]; uaL[CS + T1] = o.S1[0]; uaL[CS + T2] = o.S2[0];
  uaR[CB + N] = BnR; uaR[CB + T1] = o.B1[1]; uaR[CB + T2] = o.B2[1];
  uaR[CD] = o.Da[1]; uaR[CT] = o.Ta[1];
  uaR[CS + N] = o.Sa[1]; uaR[CS + T1] = o.S1[1]; uaR[CS + T2] = o.S2[1];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    faL[k] = h.fL[k] + (-cmin) * (uaL[k] - h.uL[k]);
    faR[k] = h.fR[k] + cmax * (uaR[k] - h.uR[k]);
  }

  // contact velocity and states   hlld.py:1396-1445
  const T fac = sdiv(1.0 - o.KL2, -R.Sg[0] * o.sq[0] - o.KBL);
  const T laL = o.la[0], laR = o.la[1];
  const T vc = laL - BnR * fac;
  const T vt1 = o.K1L - o.Bc1 * fac;
  const T vt2 = o.K2L - o.Bc2 * fac;
  const T vBc = vc * BnR + vt1 * o.Bc1 + vt2 * o.Bc2;
  T ucL[8], ucR[8];
  ucL[CB + N] = BnR; ucL[CB + T1] = o.Bc1; ucL[CB + T2] = o.Bc2;
  ucR[CB + N] = BnR; ucR[CB + T1] = o.Bc1; ucR[CB + T2] = o.Bc2;
  ucL[CD] = uaL[CD] * sdiv(laL - o.va[0], laL - vc);
  ucR[CD] = uaR[CD] * sdiv(laR - o.va[1], laR - vc);
  ucL[CT] = sdiv(laL * uaL[CT] - faL[CT] + ps * vc - vBc * BnR, laL - vc);
  ucR[CT] = sdiv(laR * uaR[CT] - faR[CT] + ps * vc - vBc * BnR, laR - vc);
  {
    const T gl = ucL[CT] + ps + ucL[CD];
    const T gr = ucR[CT] + ps + ucR[CD];
    ucL[CS + N] = gl * vc - vBc * BnR;
    ucL[CS + T1] = gl * vt1 - vBc * o.Bc1;
    ucL[CS + T2] = gl * vt2 - vBc * o.Bc2;
    ucR[CS + N] = gr * vc - vBc * BnR;
    ucR[CS + T1] = gr * vt1 - vBc * o.Bc1;
    ucR[CS + T2] = gr * vt2 - vBc * o.Bc2;
  }

  const bool wave_bad = (o.va[1] > cmax) || (o.va[0] < -cmin) || (vc < laL) || (vc > laR);
  const bool failed = !conv || wave_bad || bn_zero;
  if (margins) {   // distances of the wave-order tests from flipping
    margins[0] = tmin(fabs(val(cmax) - val(o.va[1])), fabs(val(o.va[0]) + val(cmin)));
    margins[1] = tmin(fabs(val(vc) - val(laL)), fabs(val(laR) - val(vc)));
  }
  // region at xi = 0   hlld.py:1448-1480
  const bool mL = 0.0 <= -cmin, maL = (-cmin < 0.0) && (0.0 <= laL);
  const bool mcL = (laL < 0.0) && (0.0 < vc), mcR = (vc <= 0.0) && (0.0 < laR);
  const bool maR = (laR <= 0.0) && (0.0 < cmax), mR = 0.0 >= cmax;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    T f = fH[k], u = uH[k];
    if (mL) { f = h.fL[k]; u = h.uL[k]; }
    if (maL) { f = faL[k]; u = uaL[k]; }
    if (mcL) { f = faL[k] + laL * (ucL[k] - uaL[k]); u = ucL[k]; }
    if (mcR) { f = faR[k] + laR * (ucR[k] - uaR[k]); u = ucR[k]; }
    if (maR) { f = faR[k]; u = uaR[k]; }
    if (mR) { f = h.fR[k]; u = h.uR[k]; }
    if (failed) { f = fH[k]; u = uH[k]; }
    F[k] = f;
    if (U) U[k] = u;
  }
  pstar = failed ? T(-fabs(p_hll) - 1e-10) : ps;
  return (conv ? 0 : FS_NOT_CONVERGED) | (wave_bad ? FS_WAVE_ORDER : 0) |
         (bn_zero ? FS_BN_ZERO : 0) | (failed ? FS_FALLBACK : 0);
}

// p0, the p* bracket and its expansion   hlld.py:1265-1321.  Fills R and
// returns the bracket ends and their residuals.
template <int N>
SRMHD_DEV void hlld_bracket(const HlldPrep<N>& h, const double fH[8], double p_hll, StarR& R,
                            double& plo_out, double& phi_out, double& flo_out,
                            double& fhi_out) {
  const double* uH = h.uH;
  // p0 from the Bn = 0 quadratic   hlld.py:1265-1287
  const double a = fH[CT] + fH[CD];
  const double b = -fH[CS + N] - uH[CT] - uH[CD];
  const double c = uH[CS + N];
  const double disc = clo(b * b - 4.0 * a * c, 0.0);
  const bool big = fabs(a) > 1e-14;
  const double sa = big ? a : 1.0;
  const double sb = big ? b : 1.0;
  double vs = big ? (-b - csqrt(disc)) / (2.0 * sa) : -c / sb;
  vs = chi_(clo(vs, -(1.0 - 1e-10)), 1.0 - 1e-10);
  const double p0 = clo(fH[CS + N] - (fH[CT] + fH[CD]) * vs, 1e-30);

  const double BnL = h.uL[CB + N], BnR = h.uR[CB + N];
  const bool use_p0 = (0.5 * (BnL * BnL + BnR * BnR) / p_hll) < 0.1;
  double plo = use_p0 ? p0 * 0.99 : p_hll * 0.99;
  double phi = use_p0 ? p0 * 1.01 : p_hll * 1.01;

  hlld_rvec<N>(h, R);

  // bracket expansion   hlld.py:1311-1321
  double flo = star_residual(R, plo), fhi = star_residual(R, phi);
#pragma unroll 1
  for (int it = 0; it < 5; ++it) {
    if (!((flo * fhi) > 0.0)) break;
    plo = plo / 2.0;
    phi = phi * 2.0;
    flo = star_residual(R, plo);
    fhi = star_residual(R, phi);
  }
  plo_out = plo;
  phi_out = phi;
  flo_out = flo;
  fhi_out = fhi;
}

// The HLL flux of the face   hlld.py:1232-1235
template <int N, class T>
SRMHD_DEV void hll_fluxes(const HlldPrep<N, T>& h, T fH[8]) {
#pragma unroll
  for (int k = 0; k < 8; ++k)
    fH[k] = hll_flux(h.fL[k], h.fR[k], h.uL[k], h.uR[k], h.cmin, h.cmax);
}

// The rest of hlld_flux given p_hll   hlld.py:1265-1534.  Writes F (and U
// when non-null), p*, and returns the face status bits; *iters gets the
// secant iteration count (iters_used).  tol: the p* tolerance (hlld.py:1080).
template <int N>
SRMHD_DEV int hlld_finish(const HlldPrep<N>& h, const Prim& L, const Prim& Rp, double p_hll,
                          double F[8], double* U, double& pstar, int& iters,
                          double tol = PSTAR_TOL) {
  double fH[8];
  hll_fluxes<N>(h, fH);
  StarR R;
  double plo, phi, flo, fhi;
  hlld_bracket<N>(h, fH, p_hll, R, plo, phi, flo, fhi);

  // clamped secant with frozen converged lanes   hlld.py:669-746
  double x0 = plo, x1 = phi, f0 = flo, f1 = fhi;
  int used = PSTAR_MAXIT;
  bool conv = false;
  int it = 0;
#pragma unroll 1
  for (; it < PSTAR_MAXIT; ++it) {
    if (fabs(f1) < tol) {
      conv = true;
      used = it;
      break;
    }
    const double df = f1 - f0;
    const bool safe = fabs(df) > 1e-30 * fabs(f1);
    const double dx = safe ? f1 * (x1 - x0) / df : 0.0;
    const double x2 = clo(x1 - dx, 1e-30);
    const double f2 = star_residual(R, x2);
    x0 = x1;
    f0 = f1;
    x1 = x2;
    f1 = f2;
  }
  if (!conv && fabs(f1) < tol) {
    conv = true;
    used = PSTAR_MAXIT;
  }
  iters = used;
  return hlld_assemble<N>(h, R, fH, x1, conv, p_hll, F, U, pstar);
}

}  // namespace srmhd
