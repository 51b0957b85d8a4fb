// Vector-Jacobian products of the face flux of the 2D sweep (sweep.py:30-119,
// hlld.py:1202-1543) with respect to its 15 face inputs: rho, p, z_x, z_y,
// z_z, B_t1, B_t2 of the left face state (0..6), of the right one (7..13),
// and the normal field B_n (14).  Given the flux adjoint lf (8, flux-vector
// order) they return lq[m] = sum_k lf[k] dF_k / dq_m.
//
//  * flux_vjp_dual: one forward-mode dual pass per input through the whole
//    solver (every solver; HLLD's p* by implicit differentiation of the
//    residual at the taped root).
//  * flux_vjp_hlld: HLLD in reverse mode, run backwards once -- the HLL
//    flux, the R-vectors, the Alfven and contact states at p*, the region
//    choice, then each side's state and speeds (side_rev) -- with p*(inputs)
//    defined by residual(p*) = 0:
//        lambda^T dF/dz = lambda^T A_z + (o_bar + c Res_o)^T o_z + c Res_z,
//        o_bar = A_o^T lambda,   c = -(o_bar . o_p + lambda^T A_p) / Res_p',
//    where A is the flux assembly, o the star states (star_solve), Res the
//    residual and Res_p' its total derivative in p (one dual pass in p).
//    The branches are the forward's: the region at xi = 0, the sdiv guards,
//    the clamps, the HLL speed maxima (a tie splits the adjoint evenly, as
//    torch autograd does).
#pragma once

#include "srmhd_riemann.h"

namespace srmhd {

// x as a T; a dual gets a unit tangent in component `slot` (none when slot
// is out of range)
SRMHD_DEV double seeded(double x, int, double*) { return x; }
template <int K>
SRMHD_DEV Dual<K> seeded(double x, int slot, Dual<K>*) { return Dual<K>::seed(x, slot); }

// The face primitives of one side from its reconstructed variables (rho, p,
// z, B_t1, B_t2) and the face normal field, as face_prims2d forms them.  With
// duals, input m of (q[0..6], B_n) gets the tangent component m - base.
template <int N, class T>
SRMHD_DEV void side_prims_q(const double q[7], double Bn, const srmhd_eos& e, int base,
                            PrimT<T>& Q) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  T x[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) x[m] = seeded(q[m], m - base, (T*)nullptr);
  const T z2 = sq(x[2]) + sq(x[3]) + sq(x[4]);
  const T fac = 1.0 / csqrt(1.0 + z2);
  Q.v[0] = x[2] * fac;
  Q.v[1] = x[3] * fac;
  Q.v[2] = x[4] * fac;
  Q.rho = clo(x[0], 1e-12);
  Q.p = clo(x[1], 1e-14);
  Q.eps = eos_eps(e, Q.p, Q.rho);
  Q.B[N] = seeded(Bn, 7 - base, (T*)nullptr);
  Q.B[T1] = x[5];
  Q.B[T2] = x[6];
}

// ── forward mode, every solver ──────────────────────────────────────────────
template <int N, int SOL>
SRMHD_DEV void flux_vjp_dual(const double ql[7], const double qr[7], double Bn,
                             const srmhd_eos& e, double ps, const double lf[8], double lq[15]) {
  using D1 = Dual<1>;
  double Rp = 1.0;
#pragma unroll 1
  for (int d = (SOL == SOLVER_HLLD && ps >= 0.0) ? -1 : 0; d < 15; ++d) {
    PrimT<D1> L, R;
    side_prims_q<N, D1>(ql, Bn, e, d == 14 ? 7 : (d >= 0 && d < 7 ? d : -100), L);
    side_prims_q<N, D1>(qr, Bn, e, d == 14 ? 7 : (d >= 7 ? d - 7 : -100), R);
    D1 F[8];
    if (SOL == SOLVER_HLLE) {
      D1 U[8], pso;
      hlle<N>(L, R, e, F, U, pso);
    } else if (SOL == SOLVER_HLLC) {
      D1 U[8], pso;
      hllc<N>(L, R, e, F, U, pso);
    } else {
      HlldPrep<N, D1> h;
      h.run(L, R, e);
      D1 fH[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        fH[k] = hll_flux(h.fL[k], h.fR[k], h.uL[k], h.uR[k], h.cmin, h.cmax);
      if (ps < 0.0) {   // the forward fell back to the HLLE flux
#pragma unroll
        for (int k = 0; k < 8; ++k) F[k] = fH[k];
      } else {
        StarRT<D1> Rv;
        hlld_rvec<N>(h, Rv);
        if (d < 0) {   // d(residual)/dp* at the root
          Rp = star_residual(Rv, D1::seed(ps, 0)).d[0];
          continue;
        }
        const D1 r = star_residual(Rv, D1(ps));
        D1 psd(ps);
        psd.d[0] = -r.d[0] / Rp;
        D1 pso;
        hlld_assemble<N>(h, Rv, fH, psd, true, 0.0, F, (D1*)nullptr, pso);
      }
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += lf[k] * F[k].d[0];
    lq[d] = acc;
  }
}

// ── reverse-mode pieces ─────────────────────────────────────────────────────

// q = sdiv(a, b): adds the adjoints of a and b (b is a constant where the
// guard replaced it)
SRMHD_DEV void sdiv_rev(double b, double q, double lq, double& la, double& lb) {
  const bool guard = fabs(b) < TINY;
  const double bb = guard ? (b < 0.0 ? -TINY : TINY) : b;
  const double g = lq / bb;
  la += g;
  if (!guard) lb -= g * q;
}

// y = sqrt(clamp(x, min=0)): the adjoint of x (csqrt's tangent is 0.5 / y)
SRMHD_DEV void csqrt_clo_rev(double x, double y, double ly, double& lx) {
  if (!(x < 0.0)) lx += ly * (0.5 / y);
}

// star_solve (hlld.py:431-583) backwards: the adjoints lo of its outputs
// into the adjoints lR of the R-vectors, B_n and the HLL speeds (the
// adjoint of p is not formed: p is eliminated by the caller).
SRMHD_DEV void star_solve_rev(const StarR& R, double p, const StarOut& lo, StarR& lR) {
  const double cm = R.cm, cx = R.cx, BnL = R.Bn[0], BnR = R.Bn[1];
  const double RSnL = R.RSn[0], RSnR = R.RSn[1], RS1L = R.RS1[0], RS1R = R.RS1[1];
  const double RS2L = R.RS2[0], RS2R = R.RS2[1], RB1L = R.RB1[0], RB1R = R.RB1[1];
  const double RB2L = R.RB2[0], RB2R = R.RB2[1], RTL = R.RT[0], RTR = R.RT[1];
  const double RDL = R.RD[0], RDR = R.RD[1], RBnL = R.RBn[0], RBnR = R.RBn[1];
  const double SgL = R.Sg[0], SgR = R.Sg[1];
  // forward
  const double omcm2 = 1.0 - cm * cm, omcx2 = 1.0 - cx * cx;
  const double AL = RSnL - (-cm) * RTL + p * omcm2;
  const double AR = RSnR - cx * RTR + p * omcx2;
  const double GL = sq(RB1L) + sq(RB2L), GR = sq(RB1R) + sq(RB2R);
  const double CL = RS1L * RB1L + RS2L * RB2L, CR = RS1R * RB1R + RS2R * RB2R;
  const double QL = -AL - GL + sq(BnL) * omcm2, QR = -AR - GR + sq(BnR) * omcx2;
  const double AGL = AL + GL, AGR = AR + GR;
  const double mL1 = (-cm) * p + RTL, mR1 = cx * p + RTR;
  const double iXL = AL * (-cm) * BnL + CL, iXR = AR * cx * BnR + CR;
  const double XL = BnL * iXL - AGL * mL1, XR = BnR * iXR - AGR * mR1;
  const double iaL = AL * BnL + (-cm) * CL, iaR = AR * BnR + cx * CR;
  const double vaL = sdiv(BnL * iaL - AGL * (p + RSnL), XL);
  const double vaR = sdiv(BnR * iaR - AGR * (p + RSnR), XR);
  const double eL = (-cm) * RSnL - RTL, eR = cx * RSnR - RTR;
  const double EL = CL + BnL * eL, ER = CR + BnR * eR;
  const double v1L = sdiv(QL * RS1L + RB1L * EL, XL), v1R = sdiv(QR * RS1R + RB1R * ER, XR);
  const double v2L = sdiv(QL * RS2L + RB2L * EL, XL), v2R = sdiv(QR * RS2R + RB2R * ER, XR);
  const double dLa = -cm - vaL, dRa = cx - vaR;
  const double B1L = sdiv(RB1L - BnL * v1L, dLa), B1R = sdiv(RB1R - BnR * v1R, dRa);
  const double B2L = sdiv(RB2L - BnL * v2L, dLa), B2R = sdiv(RB2R - BnR * v2R, dRa);
  const double vRSL = vaL * RSnL + v1L * RS1L + v2L * RS2L;
  const double vRSR = vaR * RSnR + v1R * RS1R + v2R * RS2R;
  const double wqL = sdiv(RTL - vRSL, dLa), wqR = sdiv(RTR - vRSR, dRa);
  const double waL = p + wqL, waR = p + wqR;
  const double DaL = sdiv(RDL, dLa), DaR = sdiv(RDR, dRa);
  const double vBL = vaL * BnL + v1L * B1L + v2L * B2L;
  const double vBR = vaR * BnR + v1R * B1R + v2R * B2R;
  const double tqL = sdiv(RTL + p * vaL - vBL * BnL, dLa);
  const double tqR = sdiv(RTR + p * vaR - vBR * BnR, dRa);
  const double TaL = tqL - DaL, TaR = tqR - DaR;
  const double gL = TaL + DaL + p, gR = TaR + DaR + p;
  const double sqL = csqrt(clo(waL, 0.0)), sqR = csqrt(clo(waR, 0.0));
  const double dL = (-cm) * p + RTL - BnL * SgL * sqL;
  const double dR = cx * p + RTR + BnR * SgR * sqR;
  const double KnL = sdiv(RSnL + p - RBnL * SgL * sqL, dL);
  const double KnR = sdiv(RSnR + p + RBnR * SgR * sqR, dR);
  const double K1L = sdiv(RS1L - RB1L * SgL * sqL, dL), K1R = sdiv(RS1R + RB1R * SgR * sqR, dR);
  const double K2L = sdiv(RS2L - RB2L * SgL * sqL, dL), K2R = sdiv(RS2R + RB2R * SgR * sqR, dR);
  const double dll = KnR - KnL;
  const double Bc1 = sdiv((B1R * (KnR - vaR) + BnR * v1R) - (B1L * (KnL - vaL) + BnL * v1L), dll);
  const double Bc2 = sdiv((B2R * (KnR - vaR) + BnR * v2R) - (B2L * (KnL - vaL) + BnL * v2L), dll);
  const double Bcn = BnR;

  // backward: adjoints of the outputs
  double lvaL = lo.va[0], lvaR = lo.va[1], lv1L = lo.v1[0], lv1R = lo.v1[1];
  double lv2L = lo.v2[0], lv2R = lo.v2[1], lB1L = lo.B1[0], lB1R = lo.B1[1];
  double lB2L = lo.B2[0], lB2R = lo.B2[1], lsqL = lo.sq[0], lsqR = lo.sq[1];
  double lDaL = lo.Da[0], lDaR = lo.Da[1], lTaL = lo.Ta[0], lTaR = lo.Ta[1];
  const double lSaL = lo.Sa[0], lSaR = lo.Sa[1], lS1L = lo.S1[0], lS1R = lo.S1[1];
  const double lS2L = lo.S2[0], lS2R = lo.S2[1];
  double lKnL = lo.la[0], lKnR = lo.la[1], lK1L = lo.K1L, lK2L = lo.K2L, lK1R = 0.0, lK2R = 0.0;
  double lBc1 = lo.Bc1, lBc2 = lo.Bc2, lBcn = 0.0;
  const double lKL2 = lo.KL2, lKR2 = lo.KR2, lKBL = lo.KBL, lKBR = lo.KBR;
  double lcm = 0.0, lcx = 0.0, lBnL = 0.0, lBnR = 0.0;
  double lRSnL = 0.0, lRSnR = 0.0, lRS1L = 0.0, lRS1R = 0.0, lRS2L = 0.0, lRS2R = 0.0;
  double lRB1L = 0.0, lRB1R = 0.0, lRB2L = 0.0, lRB2R = 0.0, lRTL = 0.0, lRTR = 0.0;
  double lRDL = 0.0, lRDR = 0.0, lRBnL = 0.0, lRBnR = 0.0;

  // KBR, KBL, KR2, KL2
  lKnR += lKBR * Bcn; lBcn += lKBR * KnR; lK1R += lKBR * Bc1; lBc1 += lKBR * K1R;
  lK2R += lKBR * Bc2; lBc2 += lKBR * K2R;
  lKnL += lKBL * Bcn; lBcn += lKBL * KnL; lK1L += lKBL * Bc1; lBc1 += lKBL * K1L;
  lK2L += lKBL * Bc2; lBc2 += lKBL * K2L;
  lKnR += 2.0 * KnR * lKR2; lK1R += 2.0 * K1R * lKR2; lK2R += 2.0 * K2R * lKR2;
  lKnL += 2.0 * KnL * lKL2; lK1L += 2.0 * K1L * lKL2; lK2L += 2.0 * K2L * lKL2;
  lBnR += lBcn;
  // Bc2, Bc1
  double ldll = 0.0;
  {
    double ln = 0.0;
    sdiv_rev(dll, Bc2, lBc2, ln, ldll);
    lB2R += ln * (KnR - vaR); lKnR += ln * B2R; lvaR -= ln * B2R; lBnR += ln * v2R; lv2R += ln * BnR;
    lB2L -= ln * (KnL - vaL); lKnL -= ln * B2L; lvaL += ln * B2L; lBnL -= ln * v2L; lv2L -= ln * BnL;
  }
  {
    double ln = 0.0;
    sdiv_rev(dll, Bc1, lBc1, ln, ldll);
    lB1R += ln * (KnR - vaR); lKnR += ln * B1R; lvaR -= ln * B1R; lBnR += ln * v1R; lv1R += ln * BnR;
    lB1L -= ln * (KnL - vaL); lKnL -= ln * B1L; lvaL += ln * B1L; lBnL -= ln * v1L; lv1L -= ln * BnL;
  }
  lKnR += ldll;
  lKnL -= ldll;
  // K's (dL, dR shared)
  double ldL = 0.0, ldR = 0.0;
  {
    double ln = 0.0;
    sdiv_rev(dR, K2R, lK2R, ln, ldR);
    lRS2R += ln; lRB2R += ln * SgR * sqR; lsqR += ln * RB2R * SgR;
    ln = 0.0;
    sdiv_rev(dL, K2L, lK2L, ln, ldL);
    lRS2L += ln; lRB2L -= ln * SgL * sqL; lsqL -= ln * RB2L * SgL;
    ln = 0.0;
    sdiv_rev(dR, K1R, lK1R, ln, ldR);
    lRS1R += ln; lRB1R += ln * SgR * sqR; lsqR += ln * RB1R * SgR;
    ln = 0.0;
    sdiv_rev(dL, K1L, lK1L, ln, ldL);
    lRS1L += ln; lRB1L -= ln * SgL * sqL; lsqL -= ln * RB1L * SgL;
    ln = 0.0;
    sdiv_rev(dR, KnR, lKnR, ln, ldR);
    lRSnR += ln; lRBnR += ln * SgR * sqR; lsqR += ln * RBnR * SgR;   // (p: eliminated)
    ln = 0.0;
    sdiv_rev(dL, KnL, lKnL, ln, ldL);
    lRSnL += ln; lRBnL -= ln * SgL * sqL; lsqL -= ln * RBnL * SgL;
  }
  lcx += ldR * p; lRTR += ldR; lBnR += ldR * SgR * sqR; lsqR += ldR * BnR * SgR;
  lcm -= ldL * p; lRTL += ldL; lBnL -= ldL * SgL * sqL; lsqL -= ldL * BnL * SgL;
  double lwaL = 0.0, lwaR = 0.0;
  csqrt_clo_rev(waR, sqR, lsqR, lwaR);
  csqrt_clo_rev(waL, sqL, lsqL, lwaL);
  // S's and g's
  double lgL = 0.0, lgR = 0.0, lvBL = 0.0, lvBR = 0.0;
  lgR += lS2R * v2R; lv2R += lS2R * gR; lvBR -= lS2R * B2R; lB2R -= lS2R * vBR;
  lgR += lS1R * v1R; lv1R += lS1R * gR; lvBR -= lS1R * B1R; lB1R -= lS1R * vBR;
  lgR += lSaR * vaR; lvaR += lSaR * gR; lvBR -= lSaR * BnR; lBnR -= lSaR * vBR;
  lgL += lS2L * v2L; lv2L += lS2L * gL; lvBL -= lS2L * B2L; lB2L -= lS2L * vBL;
  lgL += lS1L * v1L; lv1L += lS1L * gL; lvBL -= lS1L * B1L; lB1L -= lS1L * vBL;
  lgL += lSaL * vaL; lvaL += lSaL * gL; lvBL -= lSaL * BnL; lBnL -= lSaL * vBL;
  lTaR += lgR; lDaR += lgR;
  lTaL += lgL; lDaL += lgL;
  // Ta = tq - Da
  double ldLa = 0.0, ldRa = 0.0;
  {
    const double ltqR = lTaR;
    lDaR -= lTaR;
    double ln = 0.0;
    sdiv_rev(dRa, tqR, ltqR, ln, ldRa);
    lRTR += ln; lvaR += ln * p; lvBR -= ln * BnR; lBnR -= ln * vBR;
    const double ltqL = lTaL;
    lDaL -= lTaL;
    ln = 0.0;
    sdiv_rev(dLa, tqL, ltqL, ln, ldLa);
    lRTL += ln; lvaL += ln * p; lvBL -= ln * BnL; lBnL -= ln * vBL;
  }
  // vB's
  lvaR += lvBR * BnR; lBnR += lvBR * vaR; lv1R += lvBR * B1R; lB1R += lvBR * v1R;
  lv2R += lvBR * B2R; lB2R += lvBR * v2R;
  lvaL += lvBL * BnL; lBnL += lvBL * vaL; lv1L += lvBL * B1L; lB1L += lvBL * v1L;
  lv2L += lvBL * B2L; lB2L += lvBL * v2L;
  // Da's
  sdiv_rev(dRa, DaR, lDaR, lRDR, ldRa);
  sdiv_rev(dLa, DaL, lDaL, lRDL, ldLa);
  // wa = p + sdiv(RT - vRS, da)
  {
    double ln = 0.0, lvRSR = 0.0, lvRSL = 0.0;
    sdiv_rev(dRa, wqR, lwaR, ln, ldRa);
    lRTR += ln; lvRSR -= ln;
    ln = 0.0;
    sdiv_rev(dLa, wqL, lwaL, ln, ldLa);
    lRTL += ln; lvRSL -= ln;
    lvaR += lvRSR * RSnR; lRSnR += lvRSR * vaR; lv1R += lvRSR * RS1R; lRS1R += lvRSR * v1R;
    lv2R += lvRSR * RS2R; lRS2R += lvRSR * v2R;
    lvaL += lvRSL * RSnL; lRSnL += lvRSL * vaL; lv1L += lvRSL * RS1L; lRS1L += lvRSL * v1L;
    lv2L += lvRSL * RS2L; lRS2L += lvRSL * v2L;
  }
  // B1, B2
  {
    double ln = 0.0;
    sdiv_rev(dRa, B2R, lB2R, ln, ldRa);
    lRB2R += ln; lBnR -= ln * v2R; lv2R -= ln * BnR;
    ln = 0.0;
    sdiv_rev(dLa, B2L, lB2L, ln, ldLa);
    lRB2L += ln; lBnL -= ln * v2L; lv2L -= ln * BnL;
    ln = 0.0;
    sdiv_rev(dRa, B1R, lB1R, ln, ldRa);
    lRB1R += ln; lBnR -= ln * v1R; lv1R -= ln * BnR;
    ln = 0.0;
    sdiv_rev(dLa, B1L, lB1L, ln, ldLa);
    lRB1L += ln; lBnL -= ln * v1L; lv1L -= ln * BnL;
  }
  lcx += ldRa; lvaR -= ldRa;
  lcm -= ldLa; lvaL -= ldLa;
  // v1, v2 (XL, XR shared; E's)
  double lXL = 0.0, lXR = 0.0, lQL = 0.0, lQR = 0.0, lEL = 0.0, lER = 0.0;
  {
    double ln = 0.0;
    sdiv_rev(XR, v2R, lv2R, ln, lXR);
    lQR += ln * RS2R; lRS2R += ln * QR; lRB2R += ln * ER; lER += ln * RB2R;
    ln = 0.0;
    sdiv_rev(XL, v2L, lv2L, ln, lXL);
    lQL += ln * RS2L; lRS2L += ln * QL; lRB2L += ln * EL; lEL += ln * RB2L;
    ln = 0.0;
    sdiv_rev(XR, v1R, lv1R, ln, lXR);
    lQR += ln * RS1R; lRS1R += ln * QR; lRB1R += ln * ER; lER += ln * RB1R;
    ln = 0.0;
    sdiv_rev(XL, v1L, lv1L, ln, lXL);
    lQL += ln * RS1L; lRS1L += ln * QL; lRB1L += ln * EL; lEL += ln * RB1L;
  }
  double lCL = 0.0, lCR = 0.0;
  lCR += lER; lBnR += lER * eR; lcx += lER * BnR * RSnR; lRSnR += lER * BnR * cx; lRTR -= lER * BnR;
  lCL += lEL; lBnL += lEL * eL; lcm -= lEL * BnL * RSnL; lRSnL -= lEL * BnL * cm; lRTL -= lEL * BnL;
  // va's
  double lAL = 0.0, lAR = 0.0, lAGL = 0.0, lAGR = 0.0;
  {
    double ln = 0.0;
    sdiv_rev(XR, vaR, lvaR, ln, lXR);
    lBnR += ln * iaR;
    const double li = ln * BnR;
    lAR += li * BnR; lBnR += li * AR; lcx += li * CR; lCR += li * cx;
    lAGR -= ln * (p + RSnR); lRSnR -= ln * AGR;
    ln = 0.0;
    sdiv_rev(XL, vaL, lvaL, ln, lXL);
    lBnL += ln * iaL;
    const double lj = ln * BnL;
    lAL += lj * BnL; lBnL += lj * AL; lcm -= lj * CL; lCL -= lj * cm;
    lAGL -= ln * (p + RSnL); lRSnL -= ln * AGL;
  }
  // X's
  {
    lBnR += lXR * iXR;
    const double li = lXR * BnR;
    lAR += li * cx * BnR; lcx += li * AR * BnR; lBnR += li * AR * cx; lCR += li;
    lAGR -= lXR * mR1;
    const double lmR1 = -lXR * AGR;
    lcx += lmR1 * p; lRTR += lmR1;
    lBnL += lXL * iXL;
    const double lj = lXL * BnL;
    lAL -= lj * cm * BnL; lcm -= lj * AL * BnL; lBnL -= lj * AL * cm; lCL += lj;
    lAGL -= lXL * mL1;
    const double lmL1 = -lXL * AGL;
    lcm -= lmL1 * p; lRTL += lmL1;
  }
  // AG = A + G; Q = -A - G + Bn^2 (1 - c^2)
  double lGL = 0.0, lGR = 0.0, lomcm2 = 0.0, lomcx2 = 0.0;
  lAR += lAGR; lGR += lAGR;
  lAL += lAGL; lGL += lAGL;
  lAR -= lQR; lGR -= lQR; lBnR += lQR * 2.0 * BnR * omcx2; lomcx2 += lQR * sq(BnR);
  lAL -= lQL; lGL -= lQL; lBnL += lQL * 2.0 * BnL * omcm2; lomcm2 += lQL * sq(BnL);
  // C, G
  lRS1R += lCR * RB1R; lRB1R += lCR * RS1R; lRS2R += lCR * RB2R; lRB2R += lCR * RS2R;
  lRS1L += lCL * RB1L; lRB1L += lCL * RS1L; lRS2L += lCL * RB2L; lRB2L += lCL * RS2L;
  lRB1R += 2.0 * RB1R * lGR; lRB2R += 2.0 * RB2R * lGR;
  lRB1L += 2.0 * RB1L * lGL; lRB2L += 2.0 * RB2L * lGL;
  // A
  lRSnR += lAR; lcx -= lAR * RTR; lRTR -= lAR * cx; lomcx2 += lAR * p;
  lRSnL += lAL; lcm += lAL * RTL; lRTL += lAL * cm; lomcm2 += lAL * p;
  lcx -= 2.0 * cx * lomcx2;
  lcm -= 2.0 * cm * lomcm2;

  lR.RSn[0] += lRSnL; lR.RSn[1] += lRSnR; lR.RS1[0] += lRS1L; lR.RS1[1] += lRS1R;
  lR.RS2[0] += lRS2L; lR.RS2[1] += lRS2R; lR.RB1[0] += lRB1L; lR.RB1[1] += lRB1R;
  lR.RB2[0] += lRB2L; lR.RB2[1] += lRB2R; lR.RT[0] += lRTL; lR.RT[1] += lRTR;
  lR.RD[0] += lRDL; lR.RD[1] += lRDR; lR.RBn[0] += lRBnL; lR.RBn[1] += lRBnR;
  lR.Bn[0] += lBnL; lR.Bn[1] += lBnR; lR.cm += lcm; lR.cx += lcx;
}

// the residual (hlld.py:586-602) backwards with output adjoint c: into the
// star-state adjoints lo and the direct B_n adjoint
SRMHD_DEV void residual_rev(const StarR& R, const StarOut& o, double c, StarOut& lo, StarR& lR) {
  const double dRr = R.Sg[1] * o.sq[1] - o.KBR, dLr = R.Sg[0] * o.sq[0] + o.KBL;
  const double t1 = sdiv(1.0 - o.KR2, dRr), t2 = sdiv(1.0 - o.KL2, dLr);
  lo.la[0] += c;
  lo.la[1] -= c;
  lR.Bn[0] += c * (t1 + t2);
  const double lt = c * R.Bn[0];
  double ln = 0.0, ld = 0.0;
  sdiv_rev(dRr, t1, lt, ln, ld);
  lo.KR2 -= ln; lo.sq[1] += ld * R.Sg[1]; lo.KBR -= ld;
  ln = 0.0;
  ld = 0.0;
  sdiv_rev(dLr, t2, lt, ln, ld);
  lo.KL2 -= ln; lo.sq[0] += ld * R.Sg[0]; lo.KBL += ld;
}

// hlld_rvec backwards: the R-vector adjoints into the side-state adjoints
template <int N>
SRMHD_DEV void rvec_rev(const HlldPrep<N>& h, const StarR& lR, double luL[8], double lfL[8],
                        double luR[8], double lfR[8], double& lcm, double& lcx) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  const double cmin = h.cmin, cmax = h.cmax;
  // left: R = -cmin u - f
  {
    const double* u = h.uL;
    const double uE = u[CT] + u[CD];
    const int idx[6] = {CS + N, CS + T1, CS + T2, CB + N, CB + T1, CB + T2};
    const double lv[6] = {lR.RSn[0], lR.RS1[0], lR.RS2[0], lR.RBn[0], lR.RB1[0], lR.RB2[0]};
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      luL[idx[a]] -= cmin * lv[a];
      lfL[idx[a]] -= lv[a];
      lcm -= u[idx[a]] * lv[a];
    }
    luL[CT] -= cmin * lR.RT[0]; luL[CD] -= cmin * lR.RT[0];
    lfL[CT] -= lR.RT[0]; lfL[CD] -= lR.RT[0];
    lcm -= uE * lR.RT[0];
    luL[CD] -= cmin * lR.RD[0];
    lfL[CD] -= lR.RD[0];
    lcm -= u[CD] * lR.RD[0];
    luL[CB + N] += lR.Bn[0];
  }
  // right: R = cmax u - f
  {
    const double* u = h.uR;
    const double uE = u[CT] + u[CD];
    const int idx[6] = {CS + N, CS + T1, CS + T2, CB + N, CB + T1, CB + T2};
    const double lv[6] = {lR.RSn[1], lR.RS1[1], lR.RS2[1], lR.RBn[1], lR.RB1[1], lR.RB2[1]};
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      luR[idx[a]] += cmax * lv[a];
      lfR[idx[a]] -= lv[a];
      lcx += u[idx[a]] * lv[a];
    }
    luR[CT] += cmax * lR.RT[1]; luR[CD] += cmax * lR.RT[1];
    lfR[CT] -= lR.RT[1]; lfR[CD] -= lR.RT[1];
    lcx += uE * lR.RT[1];
    luR[CD] += cmax * lR.RD[1];
    lfR[CD] -= lR.RD[1];
    lcx += u[CD] * lR.RD[1];
    luR[CB + N] += lR.Bn[1];
  }
  lcm += lR.cm;
  lcx += lR.cx;
}

// hlld_assemble's flux (not failed) backwards for flux adjoint lF: into the
// side-state, speed, star-state and R-vector adjoints and the direct adjoint
// of p*.  Returns the region at xi = 0 (hlld.py:1448-1480; the last true
// test wins): -1 none (the flux is the HLL flux: nothing is added), 0 / 5
// the left / right state, 1 / 4 the left / right Alfven state, 2 / 3 the
// left / right contact state.  Only the side s of the region (s = 1 for
// regions 3, 4) is formed: the other side's states reach no adjoint.
template <int N>
SRMHD_DEV int assemble_rev(const HlldPrep<N>& h, const StarR& R, const StarOut& o, double ps,
                           const double lF[8], double luL[8], double lfL[8], double luR[8],
                           double lfR[8], double& lcm, double& lcx, StarOut& lo, double& lps) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  const double cmin = h.cmin, cmax = h.cmax;
  const double BnR = h.uR[CB + N];
  const double dfac = -R.Sg[0] * o.sq[0] - o.KBL;
  const double fac = sdiv(1.0 - o.KL2, dfac);
  const double laL = o.la[0], laR = o.la[1];
  const double vc = laL - BnR * fac;
  const bool mL = 0.0 <= -cmin, maL = (-cmin < 0.0) && (0.0 <= laL);
  const bool mcL = (laL < 0.0) && (0.0 < vc), mcR = (vc <= 0.0) && (0.0 < laR);
  const bool maR = (laR <= 0.0) && (0.0 < cmax), mR = 0.0 >= cmax;
  int reg = -1;
  if (mL) reg = 0;
  if (maL) reg = 1;
  if (mcL) reg = 2;
  if (mcR) reg = 3;
  if (maR) reg = 4;
  if (mR) reg = 5;
  if (reg == 0 || reg == 5) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      lfL[k] += reg == 0 ? lF[k] : 0.0;
      lfR[k] += reg == 5 ? lF[k] : 0.0;
    }
  }
  if (!(reg >= 1 && reg <= 4)) return reg;
  // the side of the region: its Alfven state (ua), flux (fa), fast speed
  const bool s = reg >= 3, contact = reg == 2 || reg == 3;
  const double lam = s ? cmax : -cmin, la = s ? laR : laL, va = s ? o.va[1] : o.va[0];
  double hu[8], ua[8], fa[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) hu[k] = s ? h.uR[k] : h.uL[k];
  ua[CB + N] = hu[CB + N];
  ua[CB + T1] = s ? o.B1[1] : o.B1[0];
  ua[CB + T2] = s ? o.B2[1] : o.B2[0];
  ua[CD] = s ? o.Da[1] : o.Da[0];
  ua[CT] = s ? o.Ta[1] : o.Ta[0];
  ua[CS + N] = s ? o.Sa[1] : o.Sa[0];
  ua[CS + T1] = s ? o.S1[1] : o.S1[0];
  ua[CS + T2] = s ? o.S2[1] : o.S2[0];
#pragma unroll
  for (int k = 0; k < 8; ++k) fa[k] = (s ? h.fR[k] : h.fL[k]) + lam * (ua[k] - hu[k]);
  double lua[8], lfa[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    lfa[k] = lF[k];
    lua[k] = 0.0;
  }
  if (contact) {
    // the contact state of side s (forward)
    const double vt1 = o.K1L - o.Bc1 * fac;
    const double vt2 = o.K2L - o.Bc2 * fac;
    const double vBc = vc * BnR + vt1 * o.Bc1 + vt2 * o.Bc2;
    const double sD = sdiv(la - va, la - vc);
    double uc[8];
    uc[CB + N] = BnR; uc[CB + T1] = o.Bc1; uc[CB + T2] = o.Bc2;
    uc[CD] = ua[CD] * sD;
    const double nT = la * ua[CT] - fa[CT] + ps * vc - vBc * BnR;
    uc[CT] = sdiv(nT, la - vc);
    const double g = uc[CT] + ps + uc[CD];
    uc[CS + N] = g * vc - vBc * BnR;
    uc[CS + T1] = g * vt1 - vBc * o.Bc1;
    uc[CS + T2] = g * vt2 - vBc * o.Bc2;
    // F = fa + la (uc - ua)
    double lla = 0.0, luc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double l = lF[k];
      lla += l * (uc[k] - ua[k]);
      luc[k] = la * l;
      lua[k] -= la * l;
    }
    double lvc = 0.0, lvt1 = 0.0, lvt2 = 0.0, lvBc = 0.0, lBnR = 0.0, lfac = 0.0;
    double lBc1 = 0.0, lBc2 = 0.0, lg = 0.0;
    lg += luc[CS + N] * vc; lvc += luc[CS + N] * g; lvBc -= luc[CS + N] * BnR;
    lBnR -= luc[CS + N] * vBc;
    lg += luc[CS + T1] * vt1; lvt1 += luc[CS + T1] * g; lvBc -= luc[CS + T1] * o.Bc1;
    lBc1 -= luc[CS + T1] * vBc;
    lg += luc[CS + T2] * vt2; lvt2 += luc[CS + T2] * g; lvBc -= luc[CS + T2] * o.Bc2;
    lBc2 -= luc[CS + T2] * vBc;
    luc[CT] += lg; lps += lg; luc[CD] += lg;
    {
      double ln = 0.0, ld = 0.0;
      sdiv_rev(la - vc, uc[CT], luc[CT], ln, ld);
      lla += ln * ua[CT]; lua[CT] += ln * la; lfa[CT] -= ln; lps += ln * vc; lvc += ln * ps;
      lvBc -= ln * BnR; lBnR -= ln * vBc;
      lla += ld; lvc -= ld;
    }
    {
      lua[CD] += luc[CD] * sD;
      const double lsD = luc[CD] * ua[CD];
      double ln = 0.0, ld = 0.0;
      sdiv_rev(la - vc, sD, lsD, ln, ld);
      lla += ln;
      lo.va[0] -= s ? 0.0 : ln;
      lo.va[1] -= s ? ln : 0.0;
      lla += ld; lvc -= ld;
    }
    lBnR += luc[CB + N];
    lBc1 += luc[CB + T1];
    lBc2 += luc[CB + T2];
    lvc += lvBc * BnR; lBnR += lvBc * vc; lvt1 += lvBc * o.Bc1; lBc1 += lvBc * vt1;
    lvt2 += lvBc * o.Bc2; lBc2 += lvBc * vt2;
    lo.K2L += lvt2; lBc2 -= lvt2 * fac; lfac -= lvt2 * o.Bc2;
    lo.K1L += lvt1; lBc1 -= lvt1 * fac; lfac -= lvt1 * o.Bc1;
    // vc = la_L - B_nR fac
    double llaL = lvc;
    lBnR -= lvc * fac; lfac -= lvc * BnR;
    {
      double ln = 0.0, ld = 0.0;
      sdiv_rev(dfac, fac, lfac, ln, ld);
      lo.KL2 -= ln; lo.sq[0] -= ld * R.Sg[0]; lo.KBL -= ld;
    }
    lo.Bc1 += lBc1;
    lo.Bc2 += lBc2;
    llaL += s ? 0.0 : lla;
    lo.la[0] += llaL;
    lo.la[1] += s ? lla : 0.0;
    luR[CB + N] += lBnR;
  }
  // the Alfven flux fa = f + lam (ua - u) of side s
  double ll = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const double lu = -lam * lfa[k];
    lua[k] += lam * lfa[k];
    ll += (ua[k] - hu[k]) * lfa[k];
    lfL[k] += s ? 0.0 : lfa[k];
    lfR[k] += s ? lfa[k] : 0.0;
    luL[k] += s ? 0.0 : lu;
    luR[k] += s ? lu : 0.0;
  }
  lcm -= s ? 0.0 : ll;   // lam = -cmin
  lcx += s ? ll : 0.0;   // lam = cmax
  // the Alfven state of side s from the star states
  luL[CB + N] += s ? 0.0 : lua[CB + N];
  luR[CB + N] += s ? lua[CB + N] : 0.0;
  const double w0 = s ? 0.0 : 1.0, w1 = s ? 1.0 : 0.0;
  lo.B1[0] += w0 * lua[CB + T1]; lo.B1[1] += w1 * lua[CB + T1];
  lo.B2[0] += w0 * lua[CB + T2]; lo.B2[1] += w1 * lua[CB + T2];
  lo.Da[0] += w0 * lua[CD]; lo.Da[1] += w1 * lua[CD];
  lo.Ta[0] += w0 * lua[CT]; lo.Ta[1] += w1 * lua[CT];
  lo.Sa[0] += w0 * lua[CS + N]; lo.Sa[1] += w1 * lua[CS + N];
  lo.S1[0] += w0 * lua[CS + T1]; lo.S1[1] += w1 * lua[CS + T1];
  lo.S2[0] += w0 * lua[CS + T2]; lo.S2[1] += w1 * lua[CS + T2];
  return reg;
}

// d/dp of the star states and the residual's total derivative at p, by one
// dual pass in p
SRMHD_DEV void star_dp(const StarR& R, double p, StarOut& op, double& resp) {
  using D1 = Dual<1>;
  StarRT<D1> Rd;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    Rd.RSn[s] = D1(R.RSn[s]); Rd.RS1[s] = D1(R.RS1[s]); Rd.RS2[s] = D1(R.RS2[s]);
    Rd.RBn[s] = D1(R.RBn[s]); Rd.RB1[s] = D1(R.RB1[s]); Rd.RB2[s] = D1(R.RB2[s]);
    Rd.RT[s] = D1(R.RT[s]); Rd.RD[s] = D1(R.RD[s]); Rd.Sg[s] = R.Sg[s]; Rd.Bn[s] = D1(R.Bn[s]);
  }
  Rd.cm = D1(R.cm);
  Rd.cx = D1(R.cx);
  StarOutT<D1> od;
  const D1 ps = D1::seed(p, 0);
  star_solve(Rd, ps, od);
  const D1 dR = R.Sg[1] * od.sq[1] - od.KBR;
  const D1 dL = R.Sg[0] * od.sq[0] + od.KBL;
  resp = ((od.la[0] - od.la[1]) + Rd.Bn[0] * (sdiv(1.0 - od.KR2, dR) + sdiv(1.0 - od.KL2, dL))).d[0];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    op.va[s] = od.va[s].d[0]; op.v1[s] = od.v1[s].d[0]; op.v2[s] = od.v2[s].d[0];
    op.B1[s] = od.B1[s].d[0]; op.B2[s] = od.B2[s].d[0]; op.sq[s] = od.sq[s].d[0];
    op.Da[s] = od.Da[s].d[0]; op.Ta[s] = od.Ta[s].d[0]; op.Sa[s] = od.Sa[s].d[0];
    op.S1[s] = od.S1[s].d[0]; op.S2[s] = od.S2[s].d[0]; op.la[s] = od.la[s].d[0];
  }
  op.K1L = od.K1L.d[0]; op.K2L = od.K2L.d[0]; op.Bc1 = od.Bc1.d[0]; op.Bc2 = od.Bc2.d[0];
  op.KL2 = od.KL2.d[0]; op.KR2 = od.KR2.d[0]; op.KBL = od.KBL.d[0]; op.KBR = od.KBR.d[0];
}

SRMHD_DEV double star_dot(const StarOut& a, const StarOut& b) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 2; ++k)
    s += a.va[k] * b.va[k] + a.v1[k] * b.v1[k] + a.v2[k] * b.v2[k] + a.B1[k] * b.B1[k] +
         a.B2[k] * b.B2[k] + a.sq[k] * b.sq[k] + a.Da[k] * b.Da[k] + a.Ta[k] * b.Ta[k] +
         a.Sa[k] * b.Sa[k] + a.S1[k] * b.S1[k] + a.S2[k] * b.S2[k] + a.la[k] * b.la[k];
  return s + a.K1L * b.K1L + a.K2L * b.K2L + a.Bc1 * b.Bc1 + a.Bc2 * b.Bc2 + a.KL2 * b.KL2 +
         a.KR2 * b.KR2 + a.KBL * b.KBL + a.KBR * b.KBR;
}

SRMHD_DEV void star_zero(StarOut& o) {
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    o.va[k] = 0.0; o.v1[k] = 0.0; o.v2[k] = 0.0; o.B1[k] = 0.0; o.B2[k] = 0.0; o.sq[k] = 0.0;
    o.Da[k] = 0.0; o.Ta[k] = 0.0; o.Sa[k] = 0.0; o.S1[k] = 0.0; o.S2[k] = 0.0; o.la[k] = 0.0;
  }
  o.K1L = 0.0; o.K2L = 0.0; o.Bc1 = 0.0; o.Bc2 = 0.0;
  o.KL2 = 0.0; o.KR2 = 0.0; o.KBL = 0.0; o.KBR = 0.0;
}

SRMHD_DEV void rvec_zero(StarR& r) {
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    r.RSn[k] = 0.0; r.RS1[k] = 0.0; r.RS2[k] = 0.0; r.RBn[k] = 0.0; r.RB1[k] = 0.0;
    r.RB2[k] = 0.0; r.RT[k] = 0.0; r.RD[k] = 0.0; r.Sg[k] = 0.0; r.Bn[k] = 0.0;
  }
  r.cm = 0.0;
  r.cx = 0.0;
}

// ── reverse mode, one side ──────────────────────────────────────────────────
// The cold part of the EOS (eos_cold) with the derivatives of pc and ec in
// rho along the branch the value takes.
SRMHD_DEV void eos_cold_d(const srmhd_eos& e, double rho, double& pc, double& ec, double& dpc,
                          double& dec) {
  eos_cold(e, rho, pc, ec);
  const bool zero_k = e.K == 0.0 && e.gamma >= 0.0 && e.gamma <= 3.0;
  // d (K rho^gamma) / d rho = gamma pc / rho (rho > 0: the callers clamp it)
  dpc = (zero_k && rho >= 0.0 && rho <= 1e100) ? 0.0 : e.gamma * pc / rho;
  const double den = rho * (e.gamma - 1.0);
  dec = (pc == 0.0 && den > 0.0 && den <= DBL_MAX) ? 0.0 : (dpc - ec * (e.gamma - 1.0)) / den;
}

// The vector-Jacobian product of one side's state, flux and speeds
// (side_prims_q then side_state) with respect to its 8 inputs (q[0..6],
// B_n): lq8[m] = sum_k (lu[k] dU_k + lf[k] dF_k) / dx_m + lcp dc+/dx_m +
// lcn dc-/dx_m.  The forward's arithmetic once, then its reverse; the
// branches (the clamps, the EOS cases, the order of the two speeds, a tie
// splitting the adjoint evenly) as the forward takes them.
template <int N>
SRMHD_DEV void side_rev(const double q[7], double Bn, const srmhd_eos& e, const double lu[8],
                        const double lf[8], double lcp, double lcn, double lq8[8]) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  // forward: the face primitives
  const double z[3] = {q[2], q[3], q[4]};
  const double z2 = sq(z[0]) + sq(z[1]) + sq(z[2]);
  const double s1 = csqrt(1.0 + z2);
  const double fac = 1.0 / s1;
  double v[3], B[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] = z[i] * fac;
  B[N] = Bn;
  B[T1] = q[5];
  B[T2] = q[6];
  const double rho = clo(q[0], 1e-12), pin = clo(q[1], 1e-14);
  double pc, ec, dpc, dec;
  eos_cold_d(e, rho, pc, ec, dpc, dec);
  const double g1 = e.gamma_th - 1.0;
  const double eps = e.is_cold ? ec : ec + (pin - pc) / (rho * g1);
  // forward: side_state
  const double vn = v[N];
  const double v2 = sq(v[0]) + sq(v[1]) + sq(v[2]);
  const double sW = csqrt(1.0 - chi_(v2, 1.0 - 1e-10));
  const double W = 1.0 / sW, W2 = W * W;
  const double epsm = tmax(eps, ec);
  const double p = e.is_cold ? pc : pc + (epsm - ec) * rho * g1;
  const double hh = 1.0 + eps + p / rho;
  const double rh = rho * hh;
  const double cs2 = (e.is_cold ? e.gamma * p : e.gamma * pc + e.gamma_th * (p - pc)) / rh;
  const double B2 = sq(B[0]) + sq(B[1]) + sq(B[2]);
  const double vdB = v[0] * B[0] + v[1] * B[1] + v[2] * B[2];
  const double b2 = B2 / W2 + sq(vdB);
  const double bt = W * vdB;
  double b[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i] = B[i] / W + vdB * v[i] * W;
  const double rhb2 = rh + b2;
  const double fr = b2 / rhb2;
  const double v02 = cs2 + fr * (1.0 - cs2);
  const double om = 1.0 - v02;
  const double a = W2 * om + v02;
  const double bq = -2.0 * (W2 * vn * om);
  const double c = W2 * sq(vn) * om - v02;
  const double disc = bq * bq - 4.0 * a * c;
  const double det = csqrt(clo(disc, 0.0));
  const double ac = clo(a, TINY);
  const double r1 = 0.5 * (det - bq) / ac, r2 = -0.5 * (det + bq) / ac;
  const double d = rho * W;
  const double bn = b[N];
  // reverse: the conserved state and flux
  double gW = 0.0, gvn = 0.0, grhb2 = 0.0, gbt = 0.0, gbn = 0.0, gPb2 = 0.0, gd = 0.0;
  double gb[3], gv[3], gB[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    gb[i] = 0.0;
    gv[i] = 0.0;
    gB[i] = lu[CB + i] + lf[CB + i] * vn;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {   // F_B = vn B - v B_n
    gvn += lf[CB + i] * B[i];
    gv[i] -= lf[CB + i] * B[N];
    gB[N] -= lf[CB + i] * v[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {   // U_S = rhb2 W u_i - bt b_i, F_S likewise
    const double ui = W * v[i], ls = lu[CS + i], lfs = lf[CS + i];
    grhb2 += (ls * W + lfs * (W * vn)) * ui;
    gW += (ls + lfs * vn) * rhb2 * ui;
    gvn += lfs * rhb2 * W * ui;
    const double gui = (ls + lfs * vn) * rhb2 * W;
    gbt -= ls * b[i];
    gb[i] -= ls * bt + lfs * bn;
    gbn -= lfs * b[i];
    if (i == N) gPb2 += lfs;
    gW += gui * v[i];
    gv[i] += gui * W;
  }
  {   // tau and its flux
    const double l = lf[CT], m = lu[CT];
    grhb2 += (l * vn + m) * W2;
    gW += (l * vn + m) * rhb2 * 2.0 * W;
    gvn += l * (rhb2 * W2 - d);
    gbt -= l * bn + 2.0 * m * bt;
    gbn -= l * bt;
    gd -= l * vn + m;
    gPb2 -= m;
  }
  gd += lu[CD] + lf[CD] * vn;
  gvn += lf[CD] * d;
  gb[N] += gbn;
  double gp = gPb2, gb2 = 0.5 * gPb2;
  double grho = gd * W;
  gW += gd * rho;
  // reverse: the characteristic speeds (char_speeds)
  double gr1 = 0.0, gr2 = 0.0;
  if (r1 == r2) {
    gr1 = gr2 = 0.5 * (lcp + lcn);
  } else {
    if (r1 > r2 || r1 != r1) gr1 += lcp; else gr2 += lcp;
    if (r1 < r2 || r1 != r1) gr1 += lcn; else gr2 += lcn;
  }
  const double gdet = 0.5 * (gr1 - gr2) / ac;
  double gbq = -0.5 * (gr1 + gr2) / ac;
  double ga = a < TINY ? 0.0 : -(gr1 * r1 + gr2 * r2) / ac;
  const double gdisc = disc < 0.0 ? 0.0 : gdet * (0.5 / det);
  gbq += gdisc * 2.0 * bq;
  ga -= 4.0 * c * gdisc;
  const double gc = -4.0 * a * gdisc;
  const double gW2 = gc * sq(vn) * om - 2.0 * gbq * vn * om + ga * om;
  gvn += (gc * 2.0 * vn - 2.0 * gbq) * W2 * om;
  const double gom = gc * W2 * sq(vn) - 2.0 * gbq * W2 * vn + ga * W2;
  const double gv02 = -gc + ga - gom;
  gW += gW2 * 2.0 * W;
  // reverse: v02, rhb2, cs2
  const double gcs2 = gv02 * (1.0 - fr);
  const double gfr = gv02 * (1.0 - cs2);
  gb2 += gfr / rhb2;
  grhb2 -= gfr * fr / rhb2;
  gb2 += grhb2;
  const double grh = grhb2 - gcs2 * cs2 / rh, gnum = gcs2 / rh;
  double gpc = 0.0, gec = 0.0;
  if (e.is_cold) {
    gp += gnum * e.gamma;
  } else {
    gpc += gnum * (e.gamma - e.gamma_th);
    gp += gnum * e.gamma_th;
  }
  grho += grh * hh;
  const double ghh = grh * rho;
  // reverse: b, bt, b2
  double gvdB = gbt * W;
  gW += gbt * vdB;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    gB[i] += gb[i] / W;
    gW += gb[i] * (vdB * v[i] - B[i] / W2);
    gvdB += gb[i] * v[i] * W;
    gv[i] += gb[i] * vdB * W;
  }
  const double gB2 = gb2 / W2;
  gW -= 2.0 * gb2 * B2 / (W2 * W);
  gvdB += 2.0 * vdB * gb2;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    gB[i] += 2.0 * B[i] * gB2 + gvdB * v[i];
    gv[i] += gvdB * B[i];
  }
  // reverse: h = 1 + eps + p / rho, then p (eos_press)
  double geps = ghh;
  gp += ghh / rho;
  grho -= ghh * p / (rho * rho);
  gpc += gp;
  if (!e.is_cold) {
    const double ge = gp * rho * g1;
    gec -= ge;
    grho += gp * (epsm - ec) * g1;
    if (eps == ec) {
      geps += 0.5 * ge;
      gec += 0.5 * ge;
    } else if (eps > ec || eps != eps) {
      geps += ge;
    } else {
      gec += ge;
    }
  }
  {   // W = 1 / sqrt(1 - min(v2, 1 - 1e-10))
    const double gv2 = v2 > 1.0 - 1e-10 ? 0.0 : gW * W2 * (0.5 / sW);
#pragma unroll
    for (int i = 0; i < 3; ++i) gv[i] += 2.0 * v[i] * gv2;
  }
  gv[N] += gvn;
  // reverse: eps (eos_eps), the cold EOS, the face primitives
  double gpin = 0.0;
  gec += geps;
  if (!e.is_cold) {
    const double dd = rho * g1;
    gpin += geps / dd;
    gpc -= geps / dd;
    grho -= geps * (pin - pc) / (dd * rho);
  }
  grho += gpc * dpc + gec * dec;
  double gfac = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) gfac += gv[i] * z[i];
  const double gz2 = -gfac * fac * fac * (0.5 / s1);
#pragma unroll
  for (int i = 0; i < 3; ++i) lq8[2 + i] = gv[i] * fac + 2.0 * z[i] * gz2;
  lq8[0] = q[0] < 1e-12 ? 0.0 : grho;
  lq8[1] = q[1] < 1e-14 ? 0.0 : gpin;
  lq8[5] = gB[T1];
  lq8[6] = gB[T2];
  lq8[7] = gB[N];
}

// ── reverse mode, HLLD ──────────────────────────────────────────────────────
// Everything of flux_vjp_hlld but the side states' reverse: the adjoints of
// each side's state (luL, luR), flux (lfL, lfR) and speeds (lc: c+ left, c-
// left, c+ right, c- right).
template <int N>
SRMHD_DEV void hlld_core_rev(const double ql[7], const double qr[7], double Bn,
                             const srmhd_eos& e, double ps, const double lf[8], double luL[8],
                             double lfL[8], double luR[8], double lfR[8], double lc[4]) {
  Prim L, R;
  side_prims_q<N, double>(ql, Bn, e, 0, L);
  side_prims_q<N, double>(qr, Bn, e, 0, R);
  HlldPrep<N> h;
  double cp[2], cmn[2];
  h.run(L, R, e, cp, cmn);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    luL[k] = 0.0; lfL[k] = 0.0; luR[k] = 0.0; lfR[k] = 0.0;
  }
  double lcm = 0.0, lcx = 0.0;
  // the flux is the HLL flux where the forward fell back to it, or where no
  // region test holds
  bool hll = !(ps >= 0.0);
  StarR Rv;
  StarOut o, lo;
  double lps = 0.0;
  int reg = -1;
  if (!hll) {
    hlld_rvec<N>(h, Rv);
    star_solve(Rv, ps, o);
    star_zero(lo);
    reg = assemble_rev<N>(h, Rv, o, ps, lf, luL, lfL, luR, lfR, lcm, lcx, lo, lps);
    hll = reg < 0;
  }
  // the HLL fluxes (hlld.py:161-168)
  if (hll) {
    double fH[8];
    hll_fluxes<N>(h, fH);
    const double cm = h.cmin, cx = h.cmax, den = cx + cm;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double du = h.uR[k] - h.uL[k];
      const double ln = lf[k] / den;
      lfL[k] += cx * ln;
      lfR[k] += cm * ln;
      luR[k] -= cx * cm * ln;
      luL[k] += cx * cm * ln;
      lcx += (h.fL[k] - cm * du) * ln - ln * fH[k];
      lcm += (h.fR[k] - cx * du) * ln - ln * fH[k];
    }
  }
  // p* (regions 1 to 4; elsewhere lo and lps are 0: nothing reaches it)
  if (reg >= 1 && reg <= 4) {
    StarOut op;
    double resp;
    star_dp(Rv, ps, op, resp);
    const double c = -(star_dot(lo, op) + lps) / resp;
    StarR lR;
    rvec_zero(lR);
    residual_rev(Rv, o, c, lo, lR);
    star_solve_rev(Rv, ps, lo, lR);
    rvec_rev<N>(h, lR, luL, lfL, luR, lfR, lcm, lcx);
  }
  // the HLL speeds from the side speeds (srmhd_fluxes)
  double lcpl = 0.0, lcpr = 0.0, lcml = 0.0, lcmr = 0.0;
  {
    const double ax = tmax(cp[0], cp[1]), am = tmax(-cmn[0], -cmn[1]);
    const double cxr = clo(ax, 0.0), cmr = clo(am, 0.0);
    if (!(cxr < 1e-12 && cmr < 1e-12)) {
      const double la = ax < 0.0 ? 0.0 : lcx;
      if (cp[0] == cp[1]) {
        lcpl += 0.5 * la;
        lcpr += 0.5 * la;
      } else if (cp[0] > cp[1] || cp[0] != cp[0]) {
        lcpl += la;
      } else {
        lcpr += la;
      }
      const double lb = am < 0.0 ? 0.0 : lcm;
      if (-cmn[0] == -cmn[1]) {
        lcml -= 0.5 * lb;
        lcmr -= 0.5 * lb;
      } else if (-cmn[0] > -cmn[1] || cmn[0] != cmn[0]) {
        lcml -= lb;
      } else {
        lcmr -= lb;
      }
    }
  }
  lc[0] = lcpl;
  lc[1] = lcml;
  lc[2] = lcpr;
  lc[3] = lcmr;
}

// The VJP of the face flux with respect to the 15 face inputs: the core,
// then each side's state backwards (side_rev).
template <int N>
SRMHD_DEV void flux_vjp_hlld(const double ql[7], const double qr[7], double Bn,
                             const srmhd_eos& e, double ps, const double lf[8], double lq[15]) {
  double luL[8], lfL[8], luR[8], lfR[8], lc[4], l8[8];
  hlld_core_rev<N>(ql, qr, Bn, e, ps, lf, luL, lfL, luR, lfR, lc);
  side_rev<N>(ql, Bn, e, luL, lfL, lc[0], lc[1], l8);
#pragma unroll
  for (int m = 0; m < 7; ++m) lq[m] = l8[m];
  lq[14] = l8[7];
  side_rev<N>(qr, Bn, e, luR, lfR, lc[2], lc[3], l8);
#pragma unroll
  for (int m = 0; m < 7; ++m) lq[7 + m] = l8[m];
  lq[14] += l8[7];
}

}  // namespace srmhd
