// The reference's non-default HLLD p* options (src/physics/hlld.py), read by
// hlld_flux at call time from module globals:
//   _PSTAR_METHOD  (hlld.py:1088)  secant | newton | newton_fd | newton_fd1 |
//                                   newton_safe   (newton_pstar, :751-904)
//   _PSTAR_BRACKET (hlld.py:1065)  legacy (HLL-state con2prim) | hllc
//                                   (_hllc_pstar, :1104-1199)
//   _PSTAR_TOL     (hlld.py:1080)  the residual tolerance
//   _PSTAR_FD_H    (hlld.py:1091)  relative step of the finite differences
#pragma once

#include "srmhd_riemann.h"

namespace srmhd {

struct HlldOpts {
  int method = SRMHD_PSTAR_SECANT;
  int bracket = SRMHD_BRACKET_LEGACY;
  double tol = PSTAR_TOL;
  double fd_h = 0.0;   // 0: the method's own default
};

// the options of a caller's struct (NULL or zeroed fields: the defaults)
inline HlldOpts hlld_opts(const srmhd_hlld_opts* o) {
  HlldOpts r;
  if (!o) return r;
  r.method = o->pstar_method;
  r.bracket = o->pstar_bracket;
  if (o->pstar_tol > 0.0) r.tol = o->pstar_tol;
  r.fd_h = o->pstar_fd_h;
  return r;
}
inline bool hlld_opts_ok(const HlldOpts& o) {
  return o.method >= SRMHD_PSTAR_SECANT && o.method <= SRMHD_PSTAR_NEWTON_SAFE &&
         (o.bracket == SRMHD_BRACKET_LEGACY || o.bracket == SRMHD_BRACKET_HLLC) &&
         o.tol > 0.0 && o.fd_h >= 0.0;
}
// the fused default kernels cover the secant on the legacy bracket (any tol)
inline bool hlld_opts_fused(const HlldOpts& o) {
  return o.method == SRMHD_PSTAR_SECANT && o.bracket == SRMHD_BRACKET_LEGACY;
}

// The HLLC star total pressure of the HLL flux and state   hlld.py:1104-1199
template <int N>
SRMHD_DEV double hllc_pstar(const double fH[8], const double uH[8]) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  const double Bn = uH[CB + N];
  const bool hasB = fabs(Bn) > 1e-15;
  const double Bt1 = uH[CB + T1], Bt2 = uH[CB + T2];
  const double fB1 = fH[CB + T1], fB2 = fH[CB + T2];
  const double cross = hasB ? Bt1 * fB1 + Bt2 * fB2 : 0.0;
  const double bt2 = hasB ? sq(Bt1) + sq(Bt2) : 0.0;
  const double fbt2 = hasB ? sq(fB1) + sq(fB2) : 0.0;
  const double E = uH[CT] + uH[CD];
  const double fE = fH[CT] + fH[CD];
  const double qa = fE - cross;
  const double qb = -fH[CS + N] - E + bt2 + fbt2;
  const double qc = uH[CS + N] - cross;
  const double disc = csqrt(clo(qb * qb - 4.0 * qa * qc, 0.0));
  const double vq = sdiv(-0.5 * (qb + disc), qa);
  const double vl = sdiv(-qc, qb);
  const double vc = (fabs(qa) < 1e-15) ? ((fabs(qb) < 1e-15) ? 0.0 : vl) : vq;
  const double vt1 = hasB ? sdiv(Bt1 * vc - fB1, Bn) : 0.0;
  const double vt2 = hasB ? sdiv(Bt2 * vc - fB2, Bn) : 0.0;
  const double v2 = sq(vc) + sq(vt1) + sq(vt2);
  const double g = 1.0 / csqrt(clo(1.0 - v2, 1e-10));
  const double vB = hasB ? vc * Bn + vt1 * Bt1 + vt2 * Bt2 : 0.0;
  const double pB = fH[CS + N] + sq(Bn / g) - vc * (fE - Bn * vB);
  const double p0 = fH[CS + N] - vc * fE;
  return clo(hasB ? pB : p0, 1e-30);
}

// _sdiv's denominator (hlld.py:45-54) and its reverse-mode step: the
// gradient torch's autograd passes to num and den of num / where(|den| <
// tiny, +-tiny, den) -- grad / safe, and -grad * ((num / safe) / safe)
// where the denominator is not replaced (0 where it is: the where sends the
// gradient to the constant branch).
SRMHD_DEV double sdiv_safe(double den) {
  return fabs(den) < TINY ? (den < 0.0 ? -TINY : TINY) : den;
}
SRMHD_DEV void sdiv_back(double num, double den, double g, double& gnum, double& gden) {
  const double safe = sdiv_safe(den);
  gnum += g / safe;
  if (!(fabs(den) < TINY)) gden += -g * ((num / safe) / safe);
}

// HLLDComputation.residual (hlld.py:586-602, with compute_all_variables
// :431-583) and its exact derivative d residual / d p as torch autograd
// forms it (torch.autograd.grad of the residual's sum, newton_pstar's
// derivative="ad", hlld.py:814-828): the forward is the residual's own
// float64 arithmetic; the backward visits the residual's operations in
// reverse order of creation and sums each value's incoming gradients in that
// order, which is the order torch's engine accumulates them in (nodes run
// from the latest-created), with torch's derivative formulas (mul: grad *
// other; div: grad / other and -grad * ((self / other) / other); pow(x, 2):
// grad * (2 * x); sqrt: grad / (2 * result); clamp(min): grad where
// x >= min).  The Alfven denominators -cm - va are separate operations at
// each use, as written in the reference.  Values unused by the residual
// (the Alfven D, tau, S) carry no gradient.
SRMHD_DEV void star_residual_grad(const StarR& R, double p, double& f, double& df) {
  const double cm = R.cm, cx = R.cx, ncm = -cm;
  const double BnL = R.Bn[0], BnR = R.Bn[1];
  const double RSiL = R.RSn[0], RSiR = R.RSn[1], RS1L = R.RS1[0], RS1R = R.RS1[1];
  const double RS2L = R.RS2[0], RS2R = R.RS2[1], RB1L = R.RB1[0], RB1R = R.RB1[1];
  const double RB2L = R.RB2[0], RB2R = R.RB2[1], RTL = R.RT[0], RTR = R.RT[1];
  const double SL = R.Sg[0], SR = R.Sg[1];
  // constants of the residual
  const double omL = 1.0 - cm * cm, omR = 1.0 - cx * cx;
  const double aL = RSiL - ncm * RTL, aR = RSiR - cx * RTR;
  const double GL = sq(RB1L) + sq(RB2L), GR = sq(RB1R) + sq(RB2R);
  const double CL = RS1L * RB1L + RS2L * RB2L, CR = RS1R * RB1R + RS2R * RB2R;
  const double bnL2om = sq(BnL) * (1.0 - cm * cm), bnR2om = sq(BnR) * (1.0 - cx * cx);
  const double ncmCL = ncm * CL, cxCR = cx * CR;
  const double K1L = RB1L * (CL + BnL * (ncm * RSiL - RTL));
  const double K1R = RB1R * (CR + BnR * (cx * RSiR - RTR));
  const double K2L = RB2L * (CL + BnL * (ncm * RSiL - RTL));
  const double K2R = RB2R * (CR + BnR * (cx * RSiR - RTR));
  const double BnSL = BnL * SL, BnSR = BnR * SR;
  const double RBSL = R.RBn[0] * SL, RBSR = R.RBn[1] * SR;
  const double R1SL = RB1L * SL, R1SR = RB1R * SR, R2SL = RB2L * SL, R2SR = RB2R * SR;

  // ── forward, in the reference's order of evaluation ──
  const double mAL = p * omL, AL = aL + mAL;
  const double mAR = p * omR, AR = aR + mAR;
  const double nAL = -AL, q1L = nAL - GL, QL = q1L + bnL2om;
  const double nAR = -AR, q1R = nAR - GR, QR = q1R + bnR2om;
  const double x1L = AL * ncm, x2L = x1L * BnL, x3L = x2L + CL, x4L = BnL * x3L;
  const double x5L = AL + GL, x6L = ncm * p, x7L = x6L + RTL, x8L = x5L * x7L, XL = x4L - x8L;
  const double x1R = AR * cx, x2R = x1R * BnR, x3R = x2R + CR, x4R = BnR * x3R;
  const double x5R = AR + GR, x6R = cx * p, x7R = x6R + RTR, x8R = x5R * x7R, XR = x4R - x8R;
  const double y1L = AL * BnL, y2L = y1L + ncmCL, y3L = BnL * y2L, y4L = AL + GL;
  const double y5L = p + RSiL, y6L = y4L * y5L, nvaL = y3L - y6L, vaL = sdiv(nvaL, XL);
  const double y1R = AR * BnR, y2R = y1R + cxCR, y3R = BnR * y2R, y4R = AR + GR;
  const double y5R = p + RSiR, y6R = y4R * y5R, nvaR = y3R - y6R, vaR = sdiv(nvaR, XR);
  const double z1L = QL * RS1L, n1L = z1L + K1L, v1L = sdiv(n1L, XL);
  const double z1R = QR * RS1R, n1R = z1R + K1R, v1R = sdiv(n1R, XR);
  const double z2L = QL * RS2L, n2L = z2L + K2L, v2L = sdiv(n2L, XL);
  const double z2R = QR * RS2R, n2R = z2R + K2R, v2R = sdiv(n2R, XR);
  const double w1L = BnL * v1L, w2L = RB1L - w1L, dA1L = ncm - vaL, B1L = sdiv(w2L, dA1L);
  const double w1R = BnR * v1R, w2R = RB1R - w1R, dA1R = cx - vaR, B1R = sdiv(w2R, dA1R);
  const double w3L = BnL * v2L, w4L = RB2L - w3L, dA2L = ncm - vaL, B2L = sdiv(w4L, dA2L);
  const double w3R = BnR * v2R, w4R = RB2R - w3R, dA2R = cx - vaR, B2R = sdiv(w4R, dA2R);
  const double r1L = vaL * RSiL, r2L = v1L * RS1L, r3L = r1L + r2L, r4L = v2L * RS2L;
  const double vRSL = r3L + r4L;
  const double r1R = vaR * RSiR, r2R = v1R * RS1R, r3R = r1R + r2R, r4R = v2R * RS2R;
  const double vRSR = r3R + r4R;
  const double u1L = RTL - vRSL, dA3L = ncm - vaL, u2L = sdiv(u1L, dA3L), waL = p + u2L;
  const double u1R = RTR - vRSR, dA3R = cx - vaR, u2R = sdiv(u1R, dA3R), waR = p + u2R;
  const double cL = clo(waL, 0.0), sqL = csqrt(cL);
  const double cR = clo(waR, 0.0), sqR = csqrt(cR);
  const double d1 = ncm * p, d2 = d1 + RTL, d4 = BnSL * sqL, dL = d2 - d4;
  const double e1 = cx * p, e2 = e1 + RTR, e4 = BnSR * sqR, dR = e2 + e4;
  const double k1 = RSiL + p, k3 = RBSL * sqL, k4 = k1 - k3, KaL = sdiv(k4, dL);
  const double l1 = RSiR + p, l3 = RBSR * sqR, l4 = l1 + l3, KaR = sdiv(l4, dR);
  const double m3 = R1SL * sqL, m4 = RS1L - m3, K1aL = sdiv(m4, dL);
  const double n3 = R1SR * sqR, n4 = RS1R + n3, K1aR = sdiv(n4, dR);
  const double m5 = R2SL * sqL, m6 = RS2L - m5, K2aL = sdiv(m6, dL);
  const double n5 = R2SR * sqR, n6 = RS2R + n5, K2aR = sdiv(n6, dR);
  const double dll = KaR - KaL;
  const double b1 = KaR - vaR, b2 = B1R * b1, b3 = BnR * v1R, b4 = b2 + b3;
  const double b5 = KaL - vaL, b6 = B1L * b5, b7 = BnL * v1L, b8 = b6 + b7, b9 = b4 - b8;
  const double Bc1 = sdiv(b9, dll);
  const double c1 = KaR - vaR, c2 = B2R * c1, c3 = BnR * v2R, c4 = c2 + c3;
  const double c5 = KaL - vaL, c6 = B2L * c5, c7 = BnL * v2L, c8 = c6 + c7, c9 = c4 - c8;
  const double Bc2 = sdiv(c9, dll);
  const double p1 = sq(KaL), p2 = sq(K1aL), p3 = p1 + p2, p4 = sq(K2aL), KL2 = p3 + p4;
  const double s1 = sq(KaR), s2 = sq(K1aR), s3 = s1 + s2, s4 = sq(K2aR), KR2 = s3 + s4;
  const double q1 = KaL * BnR, q2 = K1aL * Bc1, q3 = q1 + q2, q4 = K2aL * Bc2, KLB = q3 + q4;
  const double t1 = KaR * BnR, t2 = K1aR * Bc1, t3 = t1 + t2, t4 = K2aR * Bc2, KRB = t3 + t4;
  const double h1 = SR * sqR, denR = h1 - KRB;
  const double h2 = SL * sqL, denL = h2 + KLB;
  const double o1 = KaL - KaR;
  const double o2 = 1.0 - KR2, o3 = sdiv(o2, denR);
  const double o4 = 1.0 - KL2, o5 = sdiv(o4, denL);
  const double o6 = o3 + o5, o7 = BnL * o6;
  f = o1 + o7;

  // ── backward ──
  double gp = 0.0;
  double g_o1 = 0, g_o7 = 0, g_o6 = 0, g_o3 = 0, g_o5 = 0, g_o2 = 0, g_o4 = 0;
  double g_denR = 0, g_denL = 0, g_h1 = 0, g_h2 = 0, g_KRB = 0, g_KLB = 0;
  double g_KL2 = 0, g_KR2 = 0, g_KaL = 0, g_KaR = 0, g_K1aL = 0, g_K1aR = 0, g_K2aL = 0,
         g_K2aR = 0;
  double g_Bc1 = 0, g_Bc2 = 0, g_dll = 0, g_sqL = 0, g_sqR = 0;
  double g_vaL = 0, g_vaR = 0, g_v1L = 0, g_v1R = 0, g_v2L = 0, g_v2R = 0;
  double g_B1L = 0, g_B1R = 0, g_B2L = 0, g_B2R = 0;
  double g_XL = 0, g_XR = 0, g_QL = 0, g_QR = 0, g_AL = 0, g_AR = 0;
  double g_dL = 0, g_dR = 0;
  {  // residual
    const double g = 1.0;
    g_o1 += g;
    g_o7 += g;
    g_o6 += g_o7 * BnL;
    g_o3 += g_o6;
    g_o5 += g_o6;
    sdiv_back(o4, denL, g_o5, g_o4, g_denL);
    g_KL2 += -g_o4;
    sdiv_back(o2, denR, g_o3, g_o2, g_denR);
    g_KR2 += -g_o2;
    g_KaL += g_o1;
    g_KaR += -g_o1;
    g_h2 += g_denL;
    g_KLB += g_denL;
    g_sqL += g_h2 * SL;
    g_h1 += g_denR;
    g_KRB += -g_denR;
    g_sqR += g_h1 * SR;
  }
  {  // KaRBc, KaLBc
    const double g_t3 = g_KRB, g_t4 = g_KRB;
    g_K2aR += g_t4 * Bc2;
    g_Bc2 += g_t4 * K2aR;
    const double g_t1 = g_t3, g_t2 = g_t3;
    g_K1aR += g_t2 * Bc1;
    g_Bc1 += g_t2 * K1aR;
    g_KaR += g_t1 * BnR;
    const double g_q3 = g_KLB, g_q4 = g_KLB;
    g_K2aL += g_q4 * Bc2;
    g_Bc2 += g_q4 * K2aL;
    const double g_q1 = g_q3, g_q2 = g_q3;
    g_K1aL += g_q2 * Bc1;
    g_Bc1 += g_q2 * K1aL;
    g_KaL += g_q1 * BnR;
  }
  {  // KaR2, KaL2
    const double g_s3 = g_KR2, g_s4 = g_KR2;
    g_K2aR += g_s4 * (2.0 * K2aR);
    g_K1aR += g_s3 * (2.0 * K1aR);   // s2 (the sum passes the gradient through)
    g_KaR += g_s3 * (2.0 * KaR);     // s1
    const double g_p3 = g_KL2, g_p4 = g_KL2;
    g_K2aL += g_p4 * (2.0 * K2aL);
    g_K1aL += g_p3 * (2.0 * K1aL);
    g_KaL += g_p3 * (2.0 * KaL);
  }
  {  // Bt2c
    double g_c9 = 0;
    sdiv_back(c9, dll, g_Bc2, g_c9, g_dll);
    const double g_c4 = g_c9, g_c8 = -g_c9;
    const double g_c6 = g_c8, g_c7 = g_c8;
    g_v2L += g_c7 * BnL;
    g_B2L += g_c6 * c5;
    const double g_c5 = g_c6 * B2L;
    g_KaL += g_c5;
    g_vaL += -g_c5;
    const double g_c2 = g_c4, g_c3 = g_c4;
    g_v2R += g_c3 * BnR;
    g_B2R += g_c2 * c1;
    const double g_c1 = g_c2 * B2R;
    g_KaR += g_c1;
    g_vaR += -g_c1;
  }
  {  // Bt1c
    double g_b9 = 0;
    sdiv_back(b9, dll, g_Bc1, g_b9, g_dll);
    const double g_b4 = g_b9, g_b8 = -g_b9;
    const double g_b6 = g_b8, g_b7 = g_b8;
    g_v1L += g_b7 * BnL;
    g_B1L += g_b6 * b5;
    const double g_b5 = g_b6 * B1L;
    g_KaL += g_b5;
    g_vaL += -g_b5;
    const double g_b2 = g_b4, g_b3 = g_b4;
    g_v1R += g_b3 * BnR;
    g_B1R += g_b2 * b1;
    const double g_b1 = g_b2 * B1R;
    g_KaR += g_b1;
    g_vaR += -g_b1;
  }
  g_KaR += g_dll;
  g_KaL += -g_dll;
  {  // K vectors, last created first
    double g_n6 = 0;
    sdiv_back(n6, dR, g_K2aR, g_n6, g_dR);
    g_sqR += g_n6 * R2SR;   // n5 = R2SR * sqR
    double g_m6 = 0;
    sdiv_back(m6, dL, g_K2aL, g_m6, g_dL);
    g_sqL += -g_m6 * R2SL;  // m6 = RS2L - m5, m5 = R2SL * sqL
    double g_n4 = 0;
    sdiv_back(n4, dR, g_K1aR, g_n4, g_dR);
    g_sqR += g_n4 * R1SR;
    double g_m4 = 0;
    sdiv_back(m4, dL, g_K1aL, g_m4, g_dL);
    g_sqL += -g_m4 * R1SL;
    double g_l4 = 0;
    sdiv_back(l4, dR, g_KaR, g_l4, g_dR);
    g_sqR += g_l4 * RBSR;   // l3
    gp += g_l4;             // l1 = RSiR + p
    double g_k4 = 0;
    sdiv_back(k4, dL, g_KaL, g_k4, g_dL);
    g_sqL += -g_k4 * RBSL;  // k3
    gp += g_k4;             // k1 = RSiL + p
  }
  {  // dR = e2 + e4, dL = d2 - d4
    g_sqR += g_dR * BnSR;
    gp += g_dR * cx;        // e2 = e1 + RTR, e1 = cx * p
    g_sqL += -g_dL * BnSL;
    gp += g_dL * ncm;       // d1 = ncm * p
  }
  double g_waL = 0, g_waR = 0;
  {  // sqR, sqL
    const double g_cR = g_sqR / (2.0 * sqR);
    g_waR += (waR >= 0.0) ? g_cR : 0.0;
    const double g_cL = g_sqL / (2.0 * sqL);
    g_waL += (waL >= 0.0) ? g_cL : 0.0;
  }
  double g_vRSL = 0, g_vRSR = 0;
  {  // waR, waL
    gp += g_waR;
    double g_u1R = 0, g_dA3R = 0;
    sdiv_back(u1R, dA3R, g_waR, g_u1R, g_dA3R);
    g_vaR += -g_dA3R;
    g_vRSR += -g_u1R;
    gp += g_waL;
    double g_u1L = 0, g_dA3L = 0;
    sdiv_back(u1L, dA3L, g_waL, g_u1L, g_dA3L);
    g_vaL += -g_dA3L;
    g_vRSL += -g_u1L;
  }
  {  // vRS_AR, vRS_AL
    g_v2R += g_vRSR * RS2R;   // r4R
    g_v1R += g_vRSR * RS1R;   // r2R (r3R passes the gradient through)
    g_vaR += g_vRSR * RSiR;   // r1R
    g_v2L += g_vRSL * RS2L;
    g_v1L += g_vRSL * RS1L;
    g_vaL += g_vRSL * RSiL;
  }
  {  // Alfven B, last created first
    double g_w4R = 0, g_dA2R = 0;
    sdiv_back(w4R, dA2R, g_B2R, g_w4R, g_dA2R);
    g_vaR += -g_dA2R;
    g_v2R += -g_w4R * BnR;
    double g_w4L = 0, g_dA2L = 0;
    sdiv_back(w4L, dA2L, g_B2L, g_w4L, g_dA2L);
    g_vaL += -g_dA2L;
    g_v2L += -g_w4L * BnL;
    double g_w2R = 0, g_dA1R = 0;
    sdiv_back(w2R, dA1R, g_B1R, g_w2R, g_dA1R);
    g_vaR += -g_dA1R;
    g_v1R += -g_w2R * BnR;
    double g_w2L = 0, g_dA1L = 0;
    sdiv_back(w2L, dA1L, g_B1L, g_w2L, g_dA1L);
    g_vaL += -g_dA1L;
    g_v1L += -g_w2L * BnL;
  }
  {  // Alfven velocities, last created first
    double g_n2R = 0;
    sdiv_back(n2R, XR, g_v2R, g_n2R, g_XR);
    g_QR += g_n2R * RS2R;
    double g_n2L = 0;
    sdiv_back(n2L, XL, g_v2L, g_n2L, g_XL);
    g_QL += g_n2L * RS2L;
    double g_n1R = 0;
    sdiv_back(n1R, XR, g_v1R, g_n1R, g_XR);
    g_QR += g_n1R * RS1R;
    double g_n1L = 0;
    sdiv_back(n1L, XL, g_v1L, g_n1L, g_XL);
    g_QL += g_n1L * RS1L;
    double g_nvaR = 0;
    sdiv_back(nvaR, XR, g_vaR, g_nvaR, g_XR);
    {
      const double g_y3R = g_nvaR, g_y6R = -g_nvaR;
      const double g_y4R = g_y6R * y5R, g_y5R = g_y6R * y4R;
      gp += g_y5R;
      g_AR += g_y4R;
      g_AR += g_y3R * BnR * BnR;   // y3R = BnR * y2R, y2R = y1R + c, y1R = AR * BnR
    }
    double g_nvaL = 0;
    sdiv_back(nvaL, XL, g_vaL, g_nvaL, g_XL);
    {
      const double g_y3L = g_nvaL, g_y6L = -g_nvaL;
      const double g_y4L = g_y6L * y5L, g_y5L = g_y6L * y4L;
      gp += g_y5L;
      g_AL += g_y4L;
      g_AL += g_y3L * BnL * BnL;
    }
  }
  {  // XR, XL
    const double g_x4R = g_XR, g_x8R = -g_XR;
    const double g_x5R = g_x8R * x7R, g_x7R = g_x8R * x5R;
    gp += g_x7R * cx;              // x7R = x6R + RTR, x6R = cx * p
    g_AR += g_x5R;                 // x5R = AR + GR
    g_AR += g_x4R * BnR * BnR * cx;  // x4R = BnR * x3R, x3R = x2R + CR, x2R = x1R * BnR, x1R = AR * cx
    const double g_x4L = g_XL, g_x8L = -g_XL;
    const double g_x5L = g_x8L * x7L, g_x7L = g_x8L * x5L;
    gp += g_x7L * ncm;
    g_AL += g_x5L;
    g_AL += g_x4L * BnL * BnL * ncm;
  }
  {  // QR, QL, AR, AL
    g_AR += -g_QR;                 // QR = q1R + c, q1R = nAR - GR, nAR = -AR
    g_AL += -g_QL;
    gp += g_AR * omR;              // AR = aR + mAR, mAR = p * omR
    gp += g_AL * omL;
  }
  df = gp;
}

// The derivative forms of newton_pstar's f_and_df (hlld.py:814-842)
enum { DERIV_AD = 0, DERIV_FD = 1, DERIV_FD1 = 2 };
SRMHD_DEV void pstar_f_df(const StarR& R, double p, int deriv, double fd_h, double& f,
                          double& df) {
  if (deriv == DERIV_AD) {
    star_residual_grad(R, p, f, df);
    return;
  }
  const double h = fd_h * clo(fabs(p), 1e-30);
  if (deriv == DERIV_FD1) {
    const double f0 = star_residual(R, p);
    const double ph = p + h;
    df = (star_residual(R, ph) - f0) / (ph - p);
    f = f0;
    return;
  }
  const double fp = star_residual(R, p + h), fm = star_residual(R, clo(p - h, 1e-30));
  f = star_residual(R, p);
  df = (fp - fm) / (2.0 * h);
}

// newton_pstar for one face (hlld.py:751-904).  The batched loop's early
// exits (no active lane, no lane to halve) change nothing for a lane: a lane
// that is done (|f| < tol) or dead (derivative or residual non-finite, or
// |f'| <= 1e-300) is frozen for good, and one that needs no more halvings
// keeps its step.  conv: |f| < tol at the returned root; iters: iters_used.
SRMHD_DEV double newton_root(const StarR& R, double x, double tol, int deriv, double fd_h,
                             bool use_bracket, double lo, double hi, double flo, double fhi,
                             bool& conv, int& iters) {
  double f, df;
  pstar_f_df(R, x, deriv, fd_h, f, df);
  bool guarded = false;
  if (use_bracket) {
    const bool ord = lo <= hi;
    const double lo2 = ord ? lo : hi, hi2 = ord ? hi : lo;
    const double flo2 = ord ? flo : fhi, fhi2 = ord ? fhi : flo;
    lo = lo2;
    hi = hi2;
    flo = flo2;
    fhi = fhi2;
    guarded = (flo * fhi) < 0.0;
  }
  int used = PSTAR_MAXIT, it = 0;
  bool done = false, frozen = false;
#pragma unroll 1
  for (int k = 0; k < PSTAR_MAXIT; ++k) {
    ++it;
    if (fabs(f) < tol) {
      used = it - 1;
      done = true;
    }
    const bool dead = !(fabs(df) <= DBL_MAX) || (fabs(df) <= 1e-300) || !(fabs(f) <= DBL_MAX);
    if (done || dead) {
      frozen = true;
      break;
    }
    const double step = f / df;
    double lam = 1.0;
    double xn = clo(x - step, 1e-30);
    if (guarded && ((xn <= lo) || (xn >= hi))) xn = 0.5 * (lo + hi);
    double fn = star_residual(R, xn);
    bool worse = !guarded && !(fabs(fn) < fabs(f));
#pragma unroll 1
    for (int hh = 0; hh < 6 && worse; ++hh) {
      lam = lam * 0.5;
      xn = clo(x - lam * step, 1e-30);
      fn = star_residual(R, xn);
      worse = !(fabs(fn) < fabs(f));
    }
    x = xn;
    if (guarded) {
      if ((fn * flo) > 0.0) {
        lo = xn;
        flo = fn;
      } else {
        hi = xn;
        fhi = fn;
      }
    }
    pstar_f_df(R, x, deriv, fd_h, f, df);
  }
  conv = fabs(f) < tol;
  if (conv && !done && !frozen) used = it;
  iters = used;
  return x;
}

// hlld_flux with the options: p_hll from the caller (legacy bracket: the
// con2prim of the HLL state) or the HLLC star pressure (hllc bracket); the
// bracket and its expansion; the selected root finder; the assembly.
template <int N>
SRMHD_DEV int hlld_finish_opts(const HlldPrep<N>& h, double p_hll, const HlldOpts& o,
                               double F[8], double* U, double& pstar, int& iters) {
  double fH[8];
  hll_fluxes<N>(h, fH);
  if (o.bracket == SRMHD_BRACKET_HLLC) p_hll = clo(hllc_pstar<N>(fH, h.uH), 1e-30);
  StarR R;
  double plo, phi, flo, fhi;
  hlld_bracket<N>(h, fH, p_hll, R, plo, phi, flo, fhi);
  double ps;
  bool conv = false;
  int used = PSTAR_MAXIT;
  if (o.method == SRMHD_PSTAR_SECANT) {
    double x0 = plo, x1 = phi, f0 = flo, f1 = fhi;
    int it = 0;
#pragma unroll 1
    for (; it < PSTAR_MAXIT; ++it) {
      if (fabs(f1) < o.tol) {
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
    if (!conv && fabs(f1) < o.tol) {
      conv = true;
      used = PSTAR_MAXIT;
    }
    ps = x1;
  } else {
    // one start: the bracket end with the smaller residual   hlld.py:1330-1335
    const double start = (fabs(flo) <= fabs(fhi)) ? plo : phi;
    const int deriv = o.method == SRMHD_PSTAR_NEWTON_FD ? DERIV_FD
                      : (o.method == SRMHD_PSTAR_NEWTON_FD1 ? DERIV_FD1 : DERIV_AD);
    const double fd_h = o.fd_h > 0.0 ? o.fd_h : (deriv == DERIV_FD1 ? 1.5e-8 : 1e-6);
    ps = newton_root(R, start, o.tol, deriv, fd_h, o.method == SRMHD_PSTAR_NEWTON_SAFE, plo,
                     phi, flo, fhi, conv, used);
  }
  iters = used;
  return hlld_assemble<N>(h, R, fH, ps, conv, p_hll, F, U, pstar);
}

}  // namespace srmhd
