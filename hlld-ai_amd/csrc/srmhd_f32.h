// FP32 point-wise SRMHD functions for the mixed-precision sweeps.
//
// The same algorithms as the FP64 headers (reconstruction.py:60-338,
// hlld.py:100-334, 1005-1765, c2p.py:41-310), evaluated in single precision
// with guards sized for it:
//   * TINY = 1e-30 (1e-300 underflows), v^2 capped at 1 - 1.2e-7 (1 - 1e-10
//     rounds to 1), WENO / MP epsilons 1e-30;
//   * the con2prim of the HLL state only seeds the p* bracket (p_lo, p_hi =
//     0.99, 1.01 p_hll), so it runs per element (no batch coupling) to a
//     tolerance of 1e-5, its secant bracketed by [0, 1) instead of mu_+;
//     mu_+ (14-step bisection: its upper end within 2^-14 above mu_+) serves the FP64 cell
//     con2prim of the mixed forms as its bracket end;
//   * the p* secant stops at |residual| < 1e-6 (its FP32 noise floor is
//     about 1e-7).  A face whose p* does not get there, or whose wave-order
//     test lies within 1e-4 of flipping, is flagged (FS_DOUBT) and recomputed
//     in FP64 by the caller: the HLLD -> HLLE fallback is a discontinuous
//     switch, and FP32 must not decide it differently from FP64.
// These are accuracy choices of the mixed form, not parity claims: results
// are compared with the FP64 path by L1 norms.
#pragma once

#include "srmhd_recon.h"
#include "srmhd_riemann.h"

namespace srmhd {
namespace f32 {

constexpr float TINY = 1e-30f;
constexpr float V2CAP = 1.0f - 1.2e-7f;
constexpr float C2P_TOL = 1e-5f;
constexpr int C2P_SECANT = 30;
// 14 steps: fewer leave the FP64 secant a wider bracket that costs more than
// the steps saved (measured on the 2048^2 rotor: 12-16 steps equal, 10 and
// fewer slower)
constexpr int MU_BISECT = 14;
constexpr int C2P_BISECT = 40;
constexpr float PSTAR_TOL = 1e-6f;
constexpr int PSTAR_MAXIT = 30;
constexpr float WAVE_MARGIN = 1e-4f;
constexpr float STRONG_BN = 0.05f;
// status bit of a face the FP64 fix-up pass recomputes
constexpr int FS_DOUBT = 32;

#define F32_DEV __host__ __device__ __forceinline__

F32_DEV float sq(float x) { return x * x; }
// sqrt(x) as the hardware square root (1 ulp), without the default
// lowering's rescaling of inputs below 2^-96
F32_DEV float fsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return __builtin_sqrtf(x);
#endif
}
F32_DEV float sgn(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
// n / d as the hardware reciprocal (1 ulp) times n.  The default FP32
// lowering wraps the reciprocal in exponent scaling (frexp, rcp, ldexp: 7
// instructions a division) that matters only near the denormal and overflow
// ranges; the physical magnitudes divided here stay far from both.
F32_DEV float fdiv(float n, float d) {
#if defined(__HIP_DEVICE_COMPILE__)
  return n * __builtin_amdgcn_rcpf(d);
#else
  return n / d;
#endif
}
// 1 / sqrt(x) in one instruction (1 ulp)
F32_DEV float frsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.f / __builtin_sqrtf(x);
#endif
}
// n / d with |d| raised to at least TINY, keeping its sign (d = -0 counts as
// negative)
F32_DEV float sdiv(float n, float d) { return fdiv(n, copysignf(fmaxf(fabsf(d), TINY), d)); }

struct Prim {
  float rho, v[3], p, eps, B[3];
};

// ── EOS (hybrid, eos.py:15-76) ──────────────────────────────────────────────
struct Eos {
  float K, g, gth;
  bool cold, k0;
  F32_DEV Eos(const srmhd_eos& e)
      : K((float)e.K), g((float)e.gamma), gth((float)e.gamma_th), cold(e.is_cold != 0),
        k0(e.K == 0.0) {}
  F32_DEV void cold_of(float rho, float& pc, float& ec) const {
    pc = k0 ? 0.f : K * powf(rho, g);
    ec = k0 ? 0.f : fdiv(pc, rho * (g - 1.f));
  }
  F32_DEV float press(float eps, float rho) const {
    float pc, ec;
    cold_of(rho, pc, ec);
    if (cold) return pc;
    return pc + (fmaxf(eps, ec) - ec) * rho * (gth - 1.f);
  }
  F32_DEV float eps_of(float p, float rho) const {
    float pc, ec;
    cold_of(rho, pc, ec);
    if (cold) return ec;
    return ec + fdiv(p - pc, rho * (gth - 1.f));
  }
  F32_DEV float eps_lo(float rho) const {
    float pc, ec;
    cold_of(rho, pc, ec);
    return ec;
  }
  F32_DEV void press_cs2(float eps, float rho, float& p, float& cs2) const {
    p = press(eps, rho);
    const float h = 1.f + eps + fdiv(p, rho);
    float pc = 0.f, ec;
    if (!k0) cold_of(rho, pc, ec);
    cs2 = fdiv(cold ? g * p : g * pc + gth * (p - pc), rho * h);
  }
};

// ── reconstruction ──────────────────────────────────────────────────────────
F32_DEV float minmod(float a, float b) {
  return (a * b > 0.f) ? (fabsf(a) < fabsf(b) ? a : b) : 0.f;
}
F32_DEV float mc_lim(float a, float b) {
  const float c = 0.5f * (a + b);
  return (a * b > 0.f) ? sgn(c) * fminf(2.f * fabsf(a), fminf(2.f * fabsf(b), fabsf(c))) : 0.f;
}
F32_DEV float weno5z(float vm2, float vm1, float v0, float vp1, float vp2) {
  const float p0 = (2.f * vm2 - 7.f * vm1 + 11.f * v0) * (1.f / 6.f);
  const float p1 = (-vm1 + 5.f * v0 + 2.f * vp1) * (1.f / 6.f);
  const float p2 = (2.f * v0 + 5.f * vp1 - vp2) * (1.f / 6.f);
  const float c1 = 13.f / 12.f, c2 = 0.25f;
  const float b0 = c1 * sq(vm2 - 2.f * vm1 + v0) + c2 * sq(vm2 - 4.f * vm1 + 3.f * v0);
  const float b1 = c1 * sq(vm1 - 2.f * v0 + vp1) + c2 * sq(vm1 - vp1);
  const float b2 = c1 * sq(v0 - 2.f * vp1 + vp2) + c2 * sq(3.f * v0 - 4.f * vp1 + vp2);
  const float t5 = fabsf(b0 - b2);
  // tau5 / (beta + eps) can exceed sqrt(FLT_MAX) where a beta vanishes; the
  // ratio is capped at 1e18 (the weights are then one-hot to FP32 precision,
  // as the FP64 weights are)
  const float a0 = 0.1f * (1.f + sq(fminf(fdiv(t5, b0 + 1e-30f), 1e18f)));
  const float a1 = 0.6f * (1.f + sq(fminf(fdiv(t5, b1 + 1e-30f), 1e18f)));
  const float a2 = 0.3f * (1.f + sq(fminf(fdiv(t5, b2 + 1e-30f), 1e18f)));
  return fdiv(a0 * p0 + a1 * p1 + a2 * p2, a0 + a1 + a2);
}
F32_DEV float mp_limit(float vl, float vm2, float vm1, float v0, float vp1, float vp2) {
  const float alpha = 4.f;
  const float vmp = v0 + minmod(vp1 - v0, alpha * (v0 - vm1));
  if (!((vl - v0) * (vl - vmp) > 1e-30f)) return vl;
  const float dm1 = vm2 - 2.f * vm1 + v0;
  const float d0 = vm1 - 2.f * v0 + vp1;
  const float dp1 = v0 - 2.f * vp1 + vp2;
  const float dm4p = minmod(minmod(4.f * d0 - dp1, 4.f * dp1 - d0), minmod(d0, dp1));
  const float dm4m = minmod(minmod(4.f * dm1 - d0, 4.f * d0 - dm1), minmod(dm1, d0));
  const float vul = v0 + alpha * (v0 - vm1);
  const float vav = 0.5f * (v0 + vp1);
  const float vmd = vav - 0.5f * dm4p;
  const float vlc = v0 + 0.5f * (v0 - vm1) + (4.f / 3.f) * dm4m;
  const float vmin = fmaxf(fminf(fminf(v0, vp1), vmd), fminf(fminf(v0, vul), vlc));
  const float vmax = fminf(fmaxf(fmaxf(v0, vp1), vmd), fmaxf(fmaxf(v0, vul), vlc));
  return vl + minmod(vmin - vl, vmax - vl);
}
F32_DEV float mp5(float vm2, float vm1, float v0, float vp1, float vp2) {
  const float vl = (2.f * vm2 - 13.f * vm1 + 47.f * v0 + 27.f * vp1 - 3.f * vp2) * (1.f / 60.f);
  return mp_limit(vl, vm2, vm1, v0, vp1, vp2);
}
F32_DEV float mp7(float vm3, float vm2, float vm1, float v0, float vp1, float vp2, float vp3) {
  const float vl = (-3.f * vm3 + 25.f * vm2 - 101.f * vm1 + 319.f * v0 + 214.f * vp1 -
                    38.f * vp2 + 4.f * vp3) * (1.f / 420.f);
  return mp_limit(vl, vm2, vm1, v0, vp1, vp2);
}

// right / left edge of cell c (q[k] = cell c + k - 3), as the FP64 cell_edge
template <int LIM>
F32_DEV float cell_edge(const float q[7], int c, int n, bool right) {
  constexpr int w = stencil_halfwidth(LIM);
  const float v0 = q[3];
  if (LIM == LIM_PCM) return v0;
  const bool high = (LIM == LIM_WENO5Z || LIM == LIM_MP5 || LIM == LIM_MP7);
  if (!high || n <= 2 * w || c < w || c >= n - w) {
    float s = 0.f;
    if (c >= 1 && c <= n - 2) {
      const float a = v0 - q[2], b = q[4] - v0;
      s = (LIM == LIM_MINMOD) ? minmod(a, b) : mc_lim(a, b);
    }
    return right ? v0 + 0.5f * s : v0 - 0.5f * s;
  }
  if (LIM == LIM_WENO5Z)
    return right ? weno5z(q[1], q[2], q[3], q[4], q[5]) : weno5z(q[5], q[4], q[3], q[2], q[1]);
  if (LIM == LIM_MP5)
    return right ? mp5(q[1], q[2], q[3], q[4], q[5]) : mp5(q[5], q[4], q[3], q[2], q[1]);
  return right ? mp7(q[0], q[1], q[2], q[3], q[4], q[5], q[6])
               : mp7(q[6], q[5], q[4], q[3], q[2], q[1], q[0]);
}

template <int LIM>
F32_DEV void face_states(const float q[8], int f, int n, float& ql, float& qr) {
  ql = (f == 0) ? cell_edge<LIM>(q + 1, 0, n, true) : cell_edge<LIM>(q, f - 1, n, true);
  qr = (f == n) ? cell_edge<LIM>(q, n - 1, n, false) : cell_edge<LIM>(q + 1, f, n, false);
}

// ── states and fluxes ───────────────────────────────────────────────────────
F32_DEV float lorentz_W(const float v[3]) {
  return frsqrt(1.f - fminf(sq(v[0]) + sq(v[1]) + sq(v[2]), V2CAP));
}
F32_DEV float b_squared(const float v[3], const float B[3], float W) {
  const float B2 = sq(B[0]) + sq(B[1]) + sq(B[2]);
  const float vdB = v[0] * B[0] + v[1] * B[1] + v[2] * B[2];
  return fdiv(B2, W * W) + sq(vdB);
}
F32_DEV void char_speeds(float v02, float u0, float vd, float& cp, float& cm) {
  const float u0sq = u0 * u0;
  const float a = fmaxf(u0sq * (1.f - v02) + v02, TINY);
  const float b = -2.f * u0sq * vd * (1.f - v02);
  const float c = u0sq * sq(vd) * (1.f - v02) - v02;
  const float det = fsqrt(fmaxf(b * b - 4.f * a * c, 0.f));
  const float r1 = fdiv(0.5f * (det - b), a);
  const float r2 = fdiv(-0.5f * (det + b), a);
  cp = fmaxf(r1, r2);
  cm = fminf(r1, r2);
}

template <int N>
F32_DEV void side_state(const Prim& P, const Eos& e, float U[8], float F[8], float& cp,
                        float& cm) {
  const float vn = P.v[N];
  const float W = lorentz_W(P.v);
  float p, cs2;
  e.press_cs2(P.eps, P.rho, p, cs2);
  const float b2 = b_squared(P.v, P.B, W);
  const float vdB = P.v[0] * P.B[0] + P.v[1] * P.B[1] + P.v[2] * P.B[2];
  const float bt = W * vdB;
  float b[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i] = fdiv(P.B[i], W) + vdB * P.v[i] * W;
  const float h = 1.f + P.eps + fdiv(p, P.rho);
  const float v02 = cs2 + fdiv(b2, P.rho * h + b2) * (1.f - cs2);
  char_speeds(v02, W, vn, cp, cm);
  const float d = P.rho * W;
  const float rhb2 = P.rho * h + b2;
  const float Pb2 = p + 0.5f * b2;
  const float bn = b[N];
  U[CD] = d;
  F[CD] = d * vn;
  U[CT] = rhb2 * (W * W) - Pb2 - bt * bt - d;
  F[CT] = (rhb2 * (W * W) * vn - bt * bn) - d * vn;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float ui = W * P.v[i];
    U[CS + i] = rhb2 * W * ui - bt * b[i];
    F[CS + i] = rhb2 * (W * vn) * ui + (i == N ? Pb2 : 0.f) - bn * b[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    U[CB + i] = P.B[i];
    F[CB + i] = vn * P.B[i] - P.v[i] * P.B[N];
  }
}

F32_DEV float hll_flux(float fl, float fr, float ul, float ur, float cm, float cx) {
  return fdiv(cx * fl + cm * fr - cx * cm * (ur - ul), cx + cm);
}
F32_DEV float hll_state(float fl, float fr, float ul, float ur, float cm, float cx) {
  return fdiv(cx * ur + cm * ul + fl - fr, cx + cm);
}

template <int N>
F32_DEV void srmhd_fluxes(const Prim& L, const Prim& R, const Eos& e, float uL[8], float uR[8],
                          float fL[8], float fR[8], float& cmax, float& cmin) {
  float cpl, cml, cpr, cmr;
  side_state<N>(L, e, uL, fL, cpl, cml);
  side_state<N>(R, e, uR, fR, cpr, cmr);
  cmax = fmaxf(fmaxf(cpl, cpr), 0.f);
  cmin = fmaxf(fmaxf(-cml, -cmr), 0.f);
  if (cmax < 1e-12f && cmin < 1e-12f) {
    cmax = 1.f;
    cmin = 1.f;
  }
}

template <int N>
F32_DEV void hlle(const Prim& L, const Prim& R, const Eos& e, float F[8]) {
  float uL[8], uR[8], fL[8], fR[8], cmax, cmin;
  srmhd_fluxes<N>(L, R, e, uL, uR, fL, fR, cmax, cmin);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float fh = hll_flux(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
    F[k] = (cmax <= 0.f) ? fR[k] : ((cmin <= 0.f) ? fL[k] : fh);
  }
}

template <int N>
F32_DEV void hllc(const Prim& L, const Prim& R, const Eos& e, float F[8]) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  float uL[8], uR[8], fL[8], fR[8], cmax, cmin;
  srmhd_fluxes<N>(L, R, e, uL, uR, fL, fR, cmax, cmin);
  const float SL = -cmin, SR = cmax;
  float fH[8], uH[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    fH[k] = hll_flux(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
    uH[k] = hll_state(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
  }
  const float Bn = uL[CB + N];
  const bool hasB = fabsf(Bn) > 1e-15f;
  const float bt1 = uH[CB + T1], bt2 = uH[CB + T2];
  const float fb1 = fH[CB + T1], fb2 = fH[CB + T2];
  const float cross = hasB ? bt1 * fb1 + bt2 * fb2 : 0.f;
  const float btan2 = hasB ? sq(bt1) + sq(bt2) : 0.f;
  const float fbtan2 = hasB ? sq(fb1) + sq(fb2) : 0.f;
  const float E = uH[CT] + uH[CD];
  const float fE = fH[CT] + fH[CD];
  const float qa = fE - cross;
  const float qb = -fH[CS + N] - E + btan2 + fbtan2;
  const float qc = uH[CS + N] - cross;
  const float disc = fsqrt(fmaxf(qb * qb - 4.f * qa * qc, 0.f));
  const float vq = sdiv(-0.5f * (qb + disc), qa);
  const float vlin = sdiv(-qc, qb);
  const float vc = (fabsf(qa) < 1e-15f) ? ((fabsf(qb) < 1e-15f) ? 0.f : vlin) : vq;
  const float vt1 = hasB ? sdiv(uH[CB + T1] * vc - fH[CB + T1], Bn) : 0.f;
  const float vt2 = hasB ? sdiv(uH[CB + T2] * vc - fH[CB + T2], Bn) : 0.f;
  const float v2 = sq(vc) + sq(vt1) + sq(vt2);
  const float g = frsqrt(1.f - fminf(v2, V2CAP));
  const float vB = hasB ? vc * uH[CB + N] + vt1 * uH[CB + T1] + vt2 * uH[CB + T2] : 0.f;
  const float pB = fH[CS + N] + sq(fdiv(Bn, g)) - vc * (fH[CT] + fH[CD] - Bn * vB);
  const float p0 = fH[CS + N] - vc * (fH[CT] + fH[CD]);
  const float ps = hasB ? pB : p0;
  float ucL[8], ucR[8];
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const float* u = side ? uR : uL;
    const float* f = side ? fR : fL;
    float* s = side ? ucR : ucL;
    const float S = side ? SR : SL;
    const float vd = side ? R.v[N] : L.v[N];
    const float den = S - vc;
    s[CB + N] = Bn;
    s[CB + T1] = uH[CB + T1];
    s[CB + T2] = uH[CB + T2];
    s[CD] = sdiv(u[CD] * (S - vd), den);
    s[CT] = sdiv(S * u[CT] - f[CT] + ps * vc - vB * Bn, den);
    s[CS + N] = (s[CT] + s[CD] + ps) * vc - vB * Bn;
    const float w1 = sdiv(-Bn * (fdiv(uH[CB + T1], g * g) + vB * vt1) - f[CS + T1] + S * u[CS + T1], den);
    const float w2 = sdiv(-Bn * (fdiv(uH[CB + T2], g * g) + vB * vt2) - f[CS + T2] + S * u[CS + T2], den);
    s[CS + T1] = hasB ? w1 : sdiv(u[CS + T1] * (S - vd), den);
    s[CS + T2] = hasB ? w2 : sdiv(u[CS + T2] * (S - vd), den);
  }
  const bool degen = (vc <= SL) || (vc >= SR) || (v2 >= 1.f);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    float f = fH[k];
    if (0.f <= SL) f = fL[k];
    if ((SL < 0.f) && (0.f < vc)) f = fL[k] + SL * (ucL[k] - uL[k]);
    if ((vc <= 0.f) && (0.f < SR)) f = fR[k] + SR * (ucR[k] - uR[k]);
    if (0.f >= SR) f = fR[k];
    if (degen || !(fabsf(f) <= 3.0e38f)) f = fH[k];
    F[k] = f;
  }
}

// ── con2prim of the HLL state (p_hll bracket only) ──────────────────────────
struct Kastaun {
  float D, q, r2, bt2, br, bperp2, v02, B[3], rv[3], btv[3];
  Eos e;
  F32_DEV Kastaun(const float c[8], const Eos& eos) : e(eos) {
    D = c[CD];
    float S[3] = {c[CS], c[CS + 1], c[CS + 2]};
    const float tau = fmaxf(c[CT], 0.f);
    B[0] = c[CB];
    B[1] = c[CB + 1];
    B[2] = c[CB + 2];
    float S2 = sq(S[0]) + sq(S[1]) + sq(S[2]);
    const float Sn = fsqrt(S2);
    const float Smax = D + tau;
    const float fac = (Sn > Smax) ? fdiv(0.9999f * Smax, Sn + TINY) : 1.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) S[i] *= fac;
    S2 = sq(S[0]) + sq(S[1]) + sq(S[2]);
    const float B2 = sq(B[0]) + sq(B[1]) + sq(B[2]);
    const float BS = B[0] * S[0] + B[1] * S[1] + B[2] * S[2];
    const float iD = fdiv(1.f, D + TINY);
    q = tau * iD;
#pragma unroll
    for (int i = 0; i < 3; ++i) rv[i] = S[i] * iD;
    r2 = fdiv(S2, D * D + TINY);
    const float sD = fsqrt(D + TINY);
    const float isD = fdiv(1.f, sD);
#pragma unroll
    for (int i = 0; i < 3; ++i) btv[i] = B[i] * isD;
    bt2 = B2 * iD;
    br = fdiv(BS, D * sD + TINY);
    bperp2 = r2 * bt2 - br * br;
    v02 = fdiv(r2, r2 + 1.f);
  }
  F32_DEV float chi(float mu) const { return fdiv(1.f, 1.f + mu * bt2); }
  F32_DEV float rbar2(float mu, float c) const {
    return r2 * (c * c) + mu * c * (1.f + c) * (br * br);
  }
  F32_DEV float fa(float mu) const { return mu * fsqrt(1.f + rbar2(mu, chi(mu))) - 1.f; }
  F32_DEV float master(float mu, float* rho_out = nullptr, float* eps_out = nullptr) const {
    const float c = chi(mu);
    const float mc = mu * c;
    const float qb = q - 0.5f * bt2 - 0.5f * (mc * mc) * bperp2;
    const float rb2 = rbar2(mu, c);
    const float vh2 = fminf(mu * mu * rb2, v02);
    const float iWh = fsqrt(fmaxf(1.f - vh2, 1.2e-7f));
    const float Wh = fdiv(1.f, iWh);
    const float rho = fminf(fmaxf(D * iWh, 1e-12f), 1e15f);
    const float eps0 = Wh * (qb - mu * rb2) + fdiv(vh2 * (Wh * Wh), 1.f + Wh);
    const float eps = fminf(fmaxf(eps0, e.eps_lo(rho)), 1e5f);
    const float p = e.press(eps, rho);
    const float a = fdiv(p, rho * (1.f + eps) + TINY);
    const float nu = fmaxf((1.f + a) * (1.f + eps) * iWh, (1.f + a) * (1.f + qb - mu * rb2));
    if (rho_out) *rho_out = rho;
    if (eps_out) *eps_out = eps;
    return mu - sdiv(1.f, nu + mu * rb2);
  }
  F32_DEV float mu_plus() const {
    float lo = 0.f, hi = 1.f;
    float flo = fa(lo);
#pragma unroll 1
    for (int it = 0; it < MU_BISECT; ++it) {
      const float mid = 0.5f * (lo + hi);
      const float fm = fa(mid);
      if ((flo * fm) <= 0.f) {
        hi = mid;
      } else {
        lo = mid;
        flo = fm;
      }
    }
    return hi;
  }
  // the total pressure of the state at its root (per element, early exit).
  // The bracket's upper end is 1/h_min = 1 rather than mu_+: the root lies
  // below mu_+ <= 1 and the master function is defined on all of (0, 1]
  // (v^2 clamped), so the secant converges from [0, 1) too, in ~3
  // evaluations instead of mu_+'s 14+ bisection steps (measured on rotor
  // states: every element converges, at most 14 evaluations).
  F32_DEV float total_pressure() const {
    const float mup = 1.f;
    float x0 = 0.f, x1 = mup * (1.f - 1e-6f);
    float f0 = master(x0), f1 = master(x1);
    const bool good = (f0 * f1) < 0.f;
    bool conv = false;
#pragma unroll 1
    for (int it = 0; it < C2P_SECANT; ++it) {
      if (fabsf(f1) < C2P_TOL) {
        conv = true;
        break;
      }
      const float df = f1 - f0;
      const float mid = 0.5f * (x0 + x1);
      float x2 = (fabsf(df) > 1e-30f) ? x1 - fdiv(f1 * (x1 - x0), df) : mid;
      x2 = ((x2 - x0) * (x2 - x1) < 0.f) ? x2 : mid;
      const float f2 = master(x2);
      if (!((f0 * f2) <= 0.f)) {
        x0 = x1;
        f0 = f1;
      }
      x1 = x2;
      f1 = f2;
    }
    float mu = x1;
    if (!conv && good) {
      float a = 0.f, b = mup * (1.f - 1e-6f), fa_ = master(a), fb = master(b);
#pragma unroll 1
      for (int it = 0; it < C2P_BISECT; ++it) {
        const float m = 0.5f * (a + b);
        const float fm = master(m);
        if ((fa_ * fm) <= 0.f) {
          b = m;
          fb = fm;
        } else {
          a = m;
          fa_ = fm;
        }
      }
      mu = fabsf(fa_) < fabsf(fb) ? a : b;
    }
    float rho, eps;
    master(mu, &rho, &eps);
    const float c = chi(mu);
    const float mc = mu * c;
    const float mbr = mu * br;
    float v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = mc * (rv[i] + mbr * btv[i]);
    const float v2 = sq(v[0]) + sq(v[1]) + sq(v[2]);
    const float lim = V2CAP;
    if (v2 > lim) {
      const float nrm = fsqrt(fdiv(lim, v2 + TINY));
#pragma unroll
      for (int i = 0; i < 3; ++i) v[i] *= nrm;
    }
    if (rho < 1e-10f) {
      rho = 1e-10f;
      eps = 1e-10f;
      v[0] = v[1] = v[2] = 0.f;
    }
    const float p = e.press(eps, rho);
    const float W = lorentz_W(v);
    return p + 0.5f * b_squared(v, B, W);
  }
};

// ── HLLD ────────────────────────────────────────────────────────────────────
struct StarR {
  float RSn[2], RS1[2], RS2[2], RBn[2], RB1[2], RB2[2], RT[2], RD[2], Sg[2], Bn[2], cm, cx;
};
struct StarOut {
  float va[2], v1[2], v2[2], B1[2], B2[2], sq_[2], Da[2], Ta[2], Sa[2], S1[2], S2[2];
  float la[2], K1L, K2L, Bc1, Bc2, KL2, KR2, KBL, KBR;
};

F32_DEV void star_solve(const StarR& R, float p, StarOut& o) {
  const float cm = R.cm, cx = R.cx;
  const float BnL = R.Bn[0], BnR = R.Bn[1];
  const float AL = R.RSn[0] + cm * R.RT[0] + p * (1.f - cm * cm);
  const float AR = R.RSn[1] - cx * R.RT[1] + p * (1.f - cx * cx);
  const float GL = sq(R.RB1[0]) + sq(R.RB2[0]);
  const float GR = sq(R.RB1[1]) + sq(R.RB2[1]);
  const float CL = R.RS1[0] * R.RB1[0] + R.RS2[0] * R.RB2[0];
  const float CR = R.RS1[1] * R.RB1[1] + R.RS2[1] * R.RB2[1];
  const float QL = -AL - GL + sq(BnL) * (1.f - cm * cm);
  const float QR = -AR - GR + sq(BnR) * (1.f - cx * cx);
  const float XL = BnL * (AL * (-cm) * BnL + CL) - (AL + GL) * ((-cm) * p + R.RT[0]);
  const float XR = BnR * (AR * cx * BnR + CR) - (AR + GR) * (cx * p + R.RT[1]);
  const float vaL = sdiv(BnL * (AL * BnL + (-cm) * CL) - (AL + GL) * (p + R.RSn[0]), XL);
  const float vaR = sdiv(BnR * (AR * BnR + cx * CR) - (AR + GR) * (p + R.RSn[1]), XR);
  const float v1L = sdiv(QL * R.RS1[0] + R.RB1[0] * (CL + BnL * ((-cm) * R.RSn[0] - R.RT[0])), XL);
  const float v1R = sdiv(QR * R.RS1[1] + R.RB1[1] * (CR + BnR * (cx * R.RSn[1] - R.RT[1])), XR);
  const float v2L = sdiv(QL * R.RS2[0] + R.RB2[0] * (CL + BnL * ((-cm) * R.RSn[0] - R.RT[0])), XL);
  const float v2R = sdiv(QR * R.RS2[1] + R.RB2[1] * (CR + BnR * (cx * R.RSn[1] - R.RT[1])), XR);
  const float dLa = -cm - vaL, dRa = cx - vaR;
  const float B1L = sdiv(R.RB1[0] - BnL * v1L, dLa);
  const float B1R = sdiv(R.RB1[1] - BnR * v1R, dRa);
  const float B2L = sdiv(R.RB2[0] - BnL * v2L, dLa);
  const float B2R = sdiv(R.RB2[1] - BnR * v2R, dRa);
  const float vRSL = vaL * R.RSn[0] + v1L * R.RS1[0] + v2L * R.RS2[0];
  const float vRSR = vaR * R.RSn[1] + v1R * R.RS1[1] + v2R * R.RS2[1];
  const float waL = p + sdiv(R.RT[0] - vRSL, dLa);
  const float waR = p + sdiv(R.RT[1] - vRSR, dRa);
  const float DaL = sdiv(R.RD[0], dLa);
  const float DaR = sdiv(R.RD[1], dRa);
  const float vBL = vaL * BnL + v1L * B1L + v2L * B2L;
  const float vBR = vaR * BnR + v1R * B1R + v2R * B2R;
  const float TaL = sdiv(R.RT[0] + p * vaL - vBL * BnL, dLa) - DaL;
  const float TaR = sdiv(R.RT[1] + p * vaR - vBR * BnR, dRa) - DaR;
  o.Sa[0] = (TaL + DaL + p) * vaL - vBL * BnL;
  o.Sa[1] = (TaR + DaR + p) * vaR - vBR * BnR;
  o.S1[0] = (TaL + DaL + p) * v1L - vBL * B1L;
  o.S1[1] = (TaR + DaR + p) * v1R - vBR * B1R;
  o.S2[0] = (TaL + DaL + p) * v2L - vBL * B2L;
  o.S2[1] = (TaR + DaR + p) * v2R - vBR * B2R;
  const float sqL = fsqrt(fmaxf(waL, 0.f));
  const float sqR = fsqrt(fmaxf(waR, 0.f));
  const float SgL = R.Sg[0], SgR = R.Sg[1];
  const float dL = (-cm) * p + R.RT[0] - BnL * SgL * sqL;
  const float dR = cx * p + R.RT[1] + BnR * SgR * sqR;
  const float KnL = sdiv(R.RSn[0] + p - R.RBn[0] * SgL * sqL, dL);
  const float KnR = sdiv(R.RSn[1] + p + R.RBn[1] * SgR * sqR, dR);
  const float K1L = sdiv(R.RS1[0] - R.RB1[0] * SgL * sqL, dL);
  const float K1R = sdiv(R.RS1[1] + R.RB1[1] * SgR * sqR, dR);
  const float K2L = sdiv(R.RS2[0] - R.RB2[0] * SgL * sqL, dL);
  const float K2R = sdiv(R.RS2[1] + R.RB2[1] * SgR * sqR, dR);
  const float dll = KnR - KnL;
  const float Bc1 = sdiv((B1R * (KnR - vaR) + BnR * v1R) - (B1L * (KnL - vaL) + BnL * v1L), dll);
  const float Bc2 = sdiv((B2R * (KnR - vaR) + BnR * v2R) - (B2L * (KnL - vaL) + BnL * v2L), dll);
  o.va[0] = vaL; o.va[1] = vaR;
  o.v1[0] = v1L; o.v1[1] = v1R;
  o.v2[0] = v2L; o.v2[1] = v2R;
  o.B1[0] = B1L; o.B1[1] = B1R;
  o.B2[0] = B2L; o.B2[1] = B2R;
  o.sq_[0] = sqL; o.sq_[1] = sqR;
  o.Da[0] = DaL; o.Da[1] = DaR;
  o.Ta[0] = TaL; o.Ta[1] = TaR;
  o.la[0] = KnL; o.la[1] = KnR;
  o.K1L = K1L; o.K2L = K2L;
  o.Bc1 = Bc1; o.Bc2 = Bc2;
  o.KL2 = sq(KnL) + sq(K1L) + sq(K2L);
  o.KR2 = sq(KnR) + sq(K1R) + sq(K2R);
  o.KBL = KnL * BnR + K1L * Bc1 + K2L * Bc2;
  o.KBR = KnR * BnR + K1R * Bc1 + K2R * Bc2;
}

F32_DEV float star_residual(const StarR& R, float p) {
  StarOut o;
  star_solve(R, p, o);
  const float dR = R.Sg[1] * o.sq_[1] - o.KBR;
  const float dL = R.Sg[0] * o.sq_[0] + o.KBL;
  return (o.la[0] - o.la[1]) + R.Bn[0] * (sdiv(1.f - o.KR2, dR) + sdiv(1.f - o.KL2, dL));
}

// The p* root of hlld_flux in FP32: HLL state, its con2prim (p_hll), p0,
// the bracket with its expansions and the clamped secant (hlld.py:1255-1356).
// Returns whether |residual| < PSTAR_TOL was reached; ps, p_hll out.
template <int N>
F32_DEV bool hlld_pstar_sides(const float uL[8], const float uR[8], const float fL[8],
                              const float fR[8], float cmax, float cmin, const Eos& e, float& ps,
                              float& p_hll) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  float fH[8], uH[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    fH[k] = hll_flux(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
    uH[k] = hll_state(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
  }
  p_hll = Kastaun(uH, e).total_pressure();

  const float a = fH[CT] + fH[CD];
  const float b = -fH[CS + N] - uH[CT] - uH[CD];
  const float c = uH[CS + N];
  const float disc = fmaxf(b * b - 4.f * a * c, 0.f);
  const bool big = fabsf(a) > 1e-14f;
  // with |a| small the reference divides -c by torch.where(big, b, 1) = 1
  float vs = big ? fdiv(-b - fsqrt(disc), 2.f * a) : -c;
  vs = fminf(fmaxf(vs, -V2CAP), V2CAP);
  const float p0 = fmaxf(fH[CS + N] - (fH[CT] + fH[CD]) * vs, 1e-30f);

  const float BnL = uL[CB + N], BnR = uR[CB + N];
  const bool use_p0 = fdiv(0.5f * (BnL * BnL + BnR * BnR), p_hll) < 0.1f;
  float plo = use_p0 ? p0 * 0.99f : p_hll * 0.99f;
  float phi = use_p0 ? p0 * 1.01f : p_hll * 1.01f;

  StarR R;
  R.cm = cmin;
  R.cx = cmax;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const float* u = s ? uR : uL;
    const float* f = s ? fR : fL;
    const float w = s ? cmax : -cmin;
    R.RSn[s] = w * u[CS + N] - f[CS + N];
    R.RS1[s] = w * u[CS + T1] - f[CS + T1];
    R.RS2[s] = w * u[CS + T2] - f[CS + T2];
    R.RBn[s] = w * u[CB + N] - f[CB + N];
    R.RB1[s] = w * u[CB + T1] - f[CB + T1];
    R.RB2[s] = w * u[CB + T2] - f[CB + T2];
    R.RT[s] = w * (u[CT] + u[CD]) - (f[CT] + f[CD]);
    R.RD[s] = w * u[CD] - f[CD];
    R.Sg[s] = sgn(u[CB + N]);
    R.Bn[s] = u[CB + N];
  }

  float flo = star_residual(R, plo), fhi = star_residual(R, phi);
#pragma unroll 1
  for (int it = 0; it < 5; ++it) {
    if (!((flo * fhi) > 0.f)) break;
    plo = plo * 0.5f;
    phi = phi * 2.f;
    flo = star_residual(R, plo);
    fhi = star_residual(R, phi);
  }
  float x0 = plo, x1 = phi, f0 = flo, f1 = fhi;
  bool conv = false;
#pragma unroll 1
  for (int it = 0; it < PSTAR_MAXIT; ++it) {
    if (fabsf(f1) < PSTAR_TOL) {
      conv = true;
      break;
    }
    const float df = f1 - f0;
    const float dx = (fabsf(df) > 1e-30f * fabsf(f1)) ? fdiv(f1 * (x1 - x0), df) : 0.f;
    const float x2 = fmaxf(x1 - dx, 1e-30f);
    const float f2 = star_residual(R, x2);
    x0 = x1;
    f0 = f1;
    x1 = x2;
    f1 = f2;
  }
  if (!conv && fabsf(f1) < PSTAR_TOL) conv = true;
  ps = x1;
  return conv;
}

// the same from the two face states
template <int N>
F32_DEV bool hlld_pstar(const Prim& L, const Prim& Rp, const Eos& e, float& ps, float& p_hll) {
  float uL[8], uR[8], fL[8], fR[8], cmax, cmin;
  srmhd_fluxes<N>(L, Rp, e, uL, uR, fL, fR, cmax, cmin);
  return hlld_pstar_sides<N>(uL, uR, fL, fR, cmax, cmin, e, ps, p_hll);
}

// HLLD flux (hlld.py:1202-1543) in FP32; returns the face status bits of the
// FP64 path (not converged, wave order, B_n = 0, fallback).
template <int N>
F32_DEV int hlld(const Prim& L, const Prim& Rp, const Eos& e, float F[8]) {
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  // B_n = 0 on both sides: the flux is the HLLE fallback whatever p* is
  // (hlld.py:1531), so the root is not sought
  if ((fabsf(L.B[N]) < 1e-14f) && (fabsf(Rp.B[N]) < 1e-14f)) {
    float uL[8], uR[8], fL[8], fR[8], cmax, cmin;
    srmhd_fluxes<N>(L, Rp, e, uL, uR, fL, fR, cmax, cmin);
#pragma unroll
    for (int k = 0; k < 8; ++k) F[k] = hll_flux(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
    return FS_BN_ZERO | FS_FALLBACK;
  }
  float uL[8], uR[8], fL[8], fR[8], cmax, cmin;
  srmhd_fluxes<N>(L, Rp, e, uL, uR, fL, fR, cmax, cmin);
  float ps, p_hll;
  const bool conv = hlld_pstar_sides<N>(uL, uR, fL, fR, cmax, cmin, e, ps, p_hll);
  float fH[8], uH[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    fH[k] = hll_flux(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
    uH[k] = hll_state(fL[k], fR[k], uL[k], uR[k], cmin, cmax);
  }
  const float BnL = uL[CB + N], BnR = uR[CB + N];
  const bool bn_zero = false;
  StarR R;
  R.cm = cmin;
  R.cx = cmax;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const float* u = s ? uR : uL;
    const float* f = s ? fR : fL;
    const float w = s ? cmax : -cmin;
    R.RSn[s] = w * u[CS + N] - f[CS + N];
    R.RS1[s] = w * u[CS + T1] - f[CS + T1];
    R.RS2[s] = w * u[CS + T2] - f[CS + T2];
    R.RBn[s] = w * u[CB + N] - f[CB + N];
    R.RB1[s] = w * u[CB + T1] - f[CB + T1];
    R.RB2[s] = w * u[CB + T2] - f[CB + T2];
    R.RT[s] = w * (u[CT] + u[CD]) - (f[CT] + f[CD]);
    R.RD[s] = w * u[CD] - f[CD];
    R.Sg[s] = sgn(u[CB + N]);
    R.Bn[s] = u[CB + N];
  }
  StarOut o;
  star_solve(R, ps, o);

  float uaL[8], uaR[8], faL[8], faR[8];
  uaL[CB + N] = BnL; uaL[CB + T1] = o.B1[0]; uaL[CB + T2] = o.B2[0];
  uaL[CD] = o.Da[0]; uaL[CT] = o.Ta[0];
  uaL[CS + N] = o.Sa[0]; uaL[CS + T1] = o.S1[0]; uaL[CS + T2] = o.S2[0];
  uaR[CB + N] = BnR; uaR[CB + T1] = o.B1[1]; uaR[CB + T2] = o.B2[1];
  uaR[CD] = o.Da[1]; uaR[CT] = o.Ta[1];
  uaR[CS + N] = o.Sa[1]; uaR[CS + T1] = o.S1[1]; uaR[CS + T2] = o.S2[1];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    faL[k] = fL[k] - cmin * (uaL[k] - uL[k]);
    faR[k] = fR[k] + cmax * (uaR[k] - uR[k]);
  }
  const float fac = sdiv(1.f - o.KL2, -R.Sg[0] * o.sq_[0] - o.KBL);
  const float laL = o.la[0], laR = o.la[1];
  const float vc = laL - BnR * fac;
  const float vt1 = o.K1L - o.Bc1 * fac;
  const float vt2 = o.K2L - o.Bc2 * fac;
  const float vBc = vc * BnR + vt1 * o.Bc1 + vt2 * o.Bc2;
  float ucL[8], ucR[8];
  ucL[CB + N] = BnR; ucL[CB + T1] = o.Bc1; ucL[CB + T2] = o.Bc2;
  ucR[CB + N] = BnR; ucR[CB + T1] = o.Bc1; ucR[CB + T2] = o.Bc2;
  ucL[CD] = uaL[CD] * sdiv(laL - o.va[0], laL - vc);
  ucR[CD] = uaR[CD] * sdiv(laR - o.va[1], laR - vc);
  ucL[CT] = sdiv(laL * uaL[CT] - faL[CT] + ps * vc - vBc * BnR, laL - vc);
  ucR[CT] = sdiv(laR * uaR[CT] - faR[CT] + ps * vc - vBc * BnR, laR - vc);
  {
    const float gl = ucL[CT] + ps + ucL[CD];
    const float gr = ucR[CT] + ps + ucR[CD];
    ucL[CS + N] = gl * vc - vBc * BnR;
    ucL[CS + T1] = gl * vt1 - vBc * o.Bc1;
    ucL[CS + T2] = gl * vt2 - vBc * o.Bc2;
    ucR[CS + N] = gr * vc - vBc * BnR;
    ucR[CS + T1] = gr * vt1 - vBc * o.Bc1;
    ucR[CS + T2] = gr * vt2 - vBc * o.Bc2;
  }
  const bool wave_bad = (o.va[1] > cmax) || (o.va[0] < -cmin) || (vc < laL) || (vc > laR);
  const bool failed = !conv || wave_bad || bn_zero;
  // faces whose fallback decision FP32 cannot settle: p* not converged, an
  // Alfven wave within FP32 noise of its fast wave, or of the contact where
  // |B_n| > 0.05 (below that the Alfven and contact waves merge and the
  // ordering test is decided by round-off in FP64 as well)
  const float fast_margin = fminf(fabsf(cmax - o.va[1]), fabsf(o.va[0] + cmin));
  const float contact_margin = fminf(fabsf(vc - laL), fabsf(laR - vc));
  const bool strong_bn = fminf(fabsf(BnL), fabsf(BnR)) > STRONG_BN;
  const bool doubt = !bn_zero && (!conv || !(fast_margin > WAVE_MARGIN) ||
                                  (strong_bn && !(contact_margin > WAVE_MARGIN)));
  const bool mL = 0.f <= -cmin, maL = (-cmin < 0.f) && (0.f <= laL);
  const bool mcL = (laL < 0.f) && (0.f < vc), mcR = (vc <= 0.f) && (0.f < laR);
  const bool maR = (laR <= 0.f) && (0.f < cmax), mR = 0.f >= cmax;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    float f = fH[k];
    if (mL) f = fL[k];
    if (maL) f = faL[k];
    if (mcL) f = faL[k] + laL * (ucL[k] - uaL[k]);
    if (mcR) f = faR[k] + laR * (ucR[k] - uaR[k]);
    if (maR) f = faR[k];
    if (mR) f = fR[k];
    if (failed) f = fH[k];
    F[k] = f;
  }
  return (conv ? 0 : FS_NOT_CONVERGED) | (wave_bad ? FS_WAVE_ORDER : 0) |
         (bn_zero ? FS_BN_ZERO : 0) | (failed ? FS_FALLBACK : 0) | (doubt ? FS_DOUBT : 0);
}

}  // namespace f32
}  // namespace srmhd
