// Kastaun con2prim on the device (c2p.py:41-363), batch-faithful.
//
// The reference's secant (c2p.py:214-233) keeps stepping EVERY element until
// all elements of the batch satisfy |f| < 1e-15, or 60 steps pass.  A
// converged element that keeps stepping is usually kicked out again (its
// next secant step rounds onto x1 and the bracket guard jumps to the
// midpoint), so the break iteration K is a property of the whole batch.
//
// It is reproduced in two passes with no grid-wide barrier:
//   pass A  every element runs the full 60 steps and records bit k of a
//           64-bit mask when it is NOT converged after k steps (k = 0..59);
//           the masks are OR-reduced into one word per batch; the state after
//           60 steps and mu_+ are stored (24 B per element);
//   pass B  K = index of the first zero bit of the OR (60 if none).  K = 60,
//           the common case for large batches, reuses the stored state;
//           K < 60 recomputes K steps.  Then _finish, the 200-step bisection
//           redo of unconverged elements with a good bracket, and the
//           atmosphere floor, exactly as conservative_to_primitive.
#pragma once

#include "srmhd_common.h"

namespace srmhd {

constexpr double C2P_TOL = 1e-15;
constexpr int C2P_SECANT = 60;
constexpr int C2P_BISECT = 200;
constexpr uint64_t C2P_ALL = (((uint64_t)1) << C2P_SECANT) - 1;   // every mask bit

// bitwise equality (NaNs of one pattern equal, +0 and -0 differ)
SRMHD_DEV bool same_bits(double a, double b) {
  return __builtin_bit_cast(unsigned long long, a) == __builtin_bit_cast(unsigned long long, b);
}

template <class T>
struct KastaunT {
  T D, q, r[3], r2, bt[3], bt2, br, bperp2, v02, B[3];
  srmhd_eos e;
  // The element's master function may be evaluated by master_fast: an
  // ideal-gas EOS and moderate coefficients, 1e-30 <= D <= 1e30, |q|, r^2,
  // b~^2, (B.r)^2-term <= 1e30 (all finite).  See master_fast for why this
  // puts five of its divisions and its square root in the range where the
  // shortened forms are exact.
  bool tame;

  // c2p.py:48-106; cons order D, tau, Sx, Sy, Sz, Bx, By, Bz
  SRMHD_DEV void init(const T c[8], const srmhd_eos& eos) {
    e = eos;
    D = c[CD];
    T S[3] = {c[CS], c[CS + 1], c[CS + 2]};
    const T tau = clo(c[CT], 0.0);
    B[0] = c[CB];
    B[1] = c[CB + 1];
    B[2] = c[CB + 2];
    T S2 = sq(S[0]) + sq(S[1]) + sq(S[2]);
    const T Sn = csqrt(S2);
    const T Smax = D + tau;
    const T fac = (Sn > Smax) ? T(0.9999 * Smax / (Sn + TINY)) : T(1.0);
    S[0] = S[0] * fac;
    S[1] = S[1] * fac;
    S[2] = S[2] * fac;
    S2 = sq(S[0]) + sq(S[1]) + sq(S[2]);
    const T B2 = sq(B[0]) + sq(B[1]) + sq(B[2]);
    const T BS = B[0] * S[0] + B[1] * S[1] + B[2] * S[2];
    q = tau / (D + TINY);
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = S[i] / (D + TINY);
    r2 = S2 / (D * D + TINY);
    const T sD = csqrt(D + TINY);
#pragma unroll
    for (int i = 0; i < 3; ++i) bt[i] = B[i] / sD;
    bt2 = B2 / (D + TINY);
    br = BS / (D * sD + TINY);
    bperp2 = r2 * bt2 - br * br;
    v02 = r2 / (r2 + 1.0);
    const double dD = val(D), dq = val(q), dr2 = val(r2), dbt2 = val(bt2), dbr = val(br);
    tame = e.K == 0.0 && e.gamma > 1.0 && e.gamma <= 3.0 &&
           (e.is_cold || (e.gamma_th > 1.0 && e.gamma_th <= 10.0)) && dD >= 1e-30 &&
           dD <= 1e30 && fabs(dq) <= 1e30 && dr2 >= 0.0 && dr2 <= 1e30 && dbt2 >= 0.0 &&
           dbt2 <= 1e30 && dbr * dbr <= 1e30;
  }

  template <class M>
  SRMHD_DEV auto chi(const M& mu) const -> decltype(mu + bt2) { return 1.0 / (1.0 + mu * bt2); }

  template <class M, class C>
  SRMHD_DEV auto rbar2(const M& mu, const C& c) const -> decltype(mu + c + r2) {
    return r2 * (c * c) + mu * c * (1.0 + c) * (br * br);
  }

  // mu sqrt(h_min^2 + rbar^2) - 1   c2p.py:124-127
  SRMHD_DEV T fa(const T& mu) const { return mu * csqrt(1.0 + rbar2(mu, chi(mu))) - 1.0; }

  // master function   c2p.py:129-158
  template <class M>
  SRMHD_DEV auto master(const M& mu, decltype(mu + bt2)* rho_out = nullptr,
                        decltype(mu + bt2)* eps_out = nullptr) const -> decltype(mu + bt2) {
    using R = decltype(mu + bt2);
    const R c = chi(mu);
    const R mc = mu * c;
    const R qb = q - 0.5 * bt2 - 0.5 * (mc * mc) * bperp2;
    const R rb2 = rbar2(mu, c);
    const R vh2 = chi_(mu * mu * rb2, v02);
    const R Wh = 1.0 / csqrt(clo(1.0 - vh2, 1e-10));
    const R rho = tmin(tmax(D / Wh, 1e-12), 1e15);
    const R eps0 = Wh * (qb - mu * rb2) + vh2 * (Wh * Wh) / (1.0 + Wh);
    const R eps = tmin(tmax(eps0, eos_eps_lo(e, rho)), 1e5);
    const R p = eos_press(e, eps, rho);
    const R a = p / (rho * (1.0 + eps) + TINY);
    const R nuA = (1.0 + a) * (1.0 + eps) / Wh;
    const R nuB = (1.0 + a) * (1.0 + qb - mu * rb2);
    const R nu = tmax(nuA, nuB);
    if (rho_out) *rho_out = rho;
    if (eps_out) *eps_out = eps;
    return mu - sdiv(1.0, nu + mu * rb2);
  }

  // master() for a tame element, bit for bit.  mu is in [0, 1] (every
  // secant and bisection iterate lies in the initial bracket [0, mu_+]).
  // Then, with the bounds of `tame`:
  //   chi = 1/(1 + mu b~^2):            divisor in [1, 1 + 1e30]
  //   vh^2 in [0, 1], so clamp(1 - vh^2, min=1e-10) is in [1e-10, 1]: its
  //   square root is in [1e-5, 1], and Wh = 1/sqrt in [1, 1e5]
  //   D / Wh:                           D in [1e-30, 1e30]
  //   rho in [1e-12, 1e15]; K = 0 gives eps_lo = 0 and p = (g_th - 1) rho
  //   eps >= 0 with eps in [0, 1e5], so a = p / (...) >= 0 and
  //   nuA = (1 + a)(1 + eps) / Wh:      dividend in [1, 1e7], divisor Wh
  //   nu = max(nuA, nuB) >= nuA >= 1e-5 and |qb|, rbar^2 <= 1e61, so
  //   1 / (nu + mu rbar^2):             divisor in [1e-5, 1e62]
  // -- all inside div_nr's range, and the square root inside csqrt_mid's.
  // The two divisions whose dividend can be arbitrarily small take div_lo:
  //   vh^2 Wh^2 / (1 + Wh):             dividend in [+0, 1e10], divisor in [2, 1e5 + 1]
  //   p / (rho (1 + eps) + tiny):       dividend in [+0, 1e21], divisor in [1e-12, 1e21]
  // (neither dividend is ever -0: see below).
  // Every value here is finite, so the clamps and maxima compare no NaN, and
  // none of them meets zeros of opposite sign: mu, vh^2 (and v0^2) and the
  // dividends are >= +0 and never -0 (mu = +0 or a sum of non-negative
  // terms; mu^2 rbar^2 a product of such), the bounds 1e-12, 1e15, 1e-10
  // and 1e5 are not zeros, eps0 is never -0 (its second term is >= +0, and
  // -0 + +0 = +0), and nuA >= 1e-5.  So vmax / vmin return the values of
  // tmax / tmin / clo / chi_.  With eps in [+0, 1e5] the EOS pressure
  // 0 + (max(eps, 0) - 0) rho (g_th - 1) is eps rho (g_th - 1), bit for bit.
  SRMHD_DEV double master_fast(double mu, double* rho_out = nullptr,
                               double* eps_out = nullptr) const {
    const double c = div_nr(1.0, 1.0 + mu * bt2);
    const double mc = mu * c;
    const double qb = q - 0.5 * bt2 - 0.5 * (mc * mc) * bperp2;
    const double rb2 = rbar2(mu, c);
    const double vh2 = vmin(mu * mu * rb2, v02);
    const double Wh = div_nr(1.0, csqrt_mid(vmax(1.0 - vh2, 1e-10)));
    const double rho = vmin(vmax(div_nr(D, Wh), 1e-12), 1e15);
    const double eps0 = Wh * (qb - mu * rb2) + div_lo(vh2 * (Wh * Wh), 1.0 + Wh);
    // the EOS at K = 0 with rho in [1e-12, 1e15]: cold pressure and energy
    // +0, so eps_lo = +0 and p = 0 + (eps - 0) rho (g_th - 1) (eos_press)
    const double eps = vmin(vmax(eps0, 0.0), 1e5);
    const double p = e.is_cold ? 0.0 : eps * rho * (e.gamma_th - 1.0);
    // rho (1 + eps) >= 1e-12, where adding TINY = 1e-300 rounds to itself
    const double a = div_lo(p, rho * (1.0 + eps));
    const double nuA = div_nr((1.0 + a) * (1.0 + eps), Wh);
    const double nuB = (1.0 + a) * (1.0 + qb - mu * rb2);
    const double nu = vmax(nuA, nuB);
    if (rho_out) *rho_out = rho;
    if (eps_out) *eps_out = eps;
    return mu - div_nr(1.0, nu + mu * rb2);
  }

  template <bool FAST>
  SRMHD_DEV double master_at(double mu) const { return FAST ? master_fast(mu) : master(mu); }

  // fa() for a tame element, bit for bit: 1 + mu b~^2 in [1, 1 + 1e30] and
  // 1 + rbar^2 in [1, 1 + 3e30]
  SRMHD_DEV double fa_fast(double mu) const {
    return mu * csqrt_mid(1.0 + rbar2(mu, div_nr(1.0, 1.0 + mu * bt2))) - 1.0;
  }

  // 50-step bisection of fa on [0, 1] (c2p.py:162-184).  The reference
  // re-evaluates fa(lo) every iteration; fa is deterministic, so carrying
  // the last value is bit-identical.
  SRMHD_DEV double mu_plus() const {
    double lo = 0.0, hi = 1.0 / (1.0 + TINY);
    double flo = fa(lo);
#pragma unroll 1
    for (int it = 0; it < 50; ++it) {
      const double mid = 0.5 * (lo + hi);
      const double fm = fa(mid);
      if ((flo * fm) <= 0.0) {
        hi = mid;
      } else {
        lo = mid;
        flo = fm;
      }
    }
    return hi;
  }

  // mu_plus(), bit for bit, evaluating fa only where the bisection's decision
  // is not already fixed.  The decision at mid is fa(mid) >= 0 (fa(lo) < 0
  // always; fa values are 0 or at least one ulp of 1 in magnitude, so the
  // product with fa(lo) cannot underflow).  fa is strictly increasing on
  // (0, 1]: mu^2 (1 + rbar^2) is a sum of products of positive increasing
  // factors (mu, mu chi, mu (1 + chi)), with slope at least 2 mu, so
  // fa' >= 1/sqrt(1 + rbar^2) >= 1/sqrt(1 + r^2) =: m0, and the root mu* lies
  // in [m0, 1] (rbar^2 <= r^2).  rbar^2 is a sum of non-negative terms and
  // the final "- 1" is exact near the root, so fa's rounding error is at most
  // ~10 ulp of max(1, mu sqrt(1 + rbar^2)) <= 2.3e-15 (1 + r^2)^(1/2).  Hence
  // wherever |mid - mu*| > 1e-13 (1 + r^2) the sign of fa(mid) -- the
  // decision -- is the sign of mid - mu*, with a margin of over 35.  mu* comes
  // from bracketed Newton steps on [m0, 1] (forward-mode derivative) to
  // |fa| <= 1e-14; anything non-finite, no convergence within 40 steps or
  // r^2 > 1e8 takes the full bisection.
  template <bool FAST = false>
  SRMHD_DEV double mu_plus_fast() const {
    if (!(r2 <= 1e8)) return mu_plus();
    double a = 1.0 / csqrt(1.0 + r2), b = 1.0 / (1.0 + TINY), x = a;
    bool ok = false;
#pragma unroll 1
    for (int it = 0; it < 40; ++it) {
      const Dual<1> m = Dual<1>::seed(x, 0);
      const Dual<1> f = m * csqrt(1.0 + rbar2(m, chi(m))) - 1.0;
      if (!(f.v == f.v) || !(f.d[0] > 0.0) || !(f.d[0] < 1e300)) break;
      // |fa(x)| <= 1e-14 puts x within (1e-14 + fa's rounding) / m0
      // <= 1.3e-14 (1 + r^2) of mu*, an eighth of delta
      if (fabs(f.v) <= 1e-14) {
        ok = true;
        break;
      }
      if (f.v < 0.0) a = x; else b = x;
      double xn = x - f.v / f.d[0];
      if (!(xn > a && xn < b)) xn = 0.5 * (a + b);
      x = xn;
    }
    if (!ok) return mu_plus();
    const double delta = 1e-13 * (1.0 + r2);
    double lo = 0.0, hi = 1.0 / (1.0 + TINY);
#pragma unroll 1
    for (int it = 0; it < 50; ++it) {
      const double mid = 0.5 * (lo + hi);
      const bool up = fabs(mid - x) > delta ? mid > x : (FAST ? fa_fast(mid) : fa(mid)) >= 0.0;
      if (up)
        hi = mid;
      else
        lo = mid;
    }
    return hi;
  }

  // one guarded secant step   c2p.py:219-232 (FAST: master_fast, tame
  // elements only).  For a tame element |f| <= 1 + 1e5 (1 / (nu + mu rbar^2)
  // <= 1e5), so the step's quotient has |dividend| <= 1 + 1e5 and divisor
  // 1e-30 < |df| <= 2 + 2e5: div_lo's range.  A zero dividend may be -0
  // (f1 = +0, x1 < x0); the quotient is then subtracted from x1, where a
  // zero of either sign leaves x1 unchanged.
  template <bool FAST = false>
  SRMHD_DEV void secant_step(double& x0, double& f0, double& x1, double& f1) const {
    const double df = f1 - f0;
    const double mid = 0.5 * (x0 + x1);
    double x2 = (fabs(df) > 1e-30)
                    ? x1 - (FAST ? div_lo(f1 * (x1 - x0), df) : f1 * (x1 - x0) / df)
                    : mid;
    x2 = ((x2 - x0) * (x2 - x1) < 0.0) ? x2 : mid;
    const double f2 = master_at<FAST>(x2);
    if (!((f0 * f2) <= 0.0)) {
      x0 = x1;
      f0 = f1;
    }
    x1 = x2;
    f1 = f2;
  }

  // secant_step, returning whether it left (x0, f0, x1, f1) unchanged bit for
  // bit: the walk depends on that state only, so it is then at a fixed point
  // and every later step leaves it unchanged too
  template <bool FAST = false>
  SRMHD_DEV bool secant_step_fixed(double& x0, double& f0, double& x1, double& f1) const {
    const double a0 = x0, b0 = f0, a1 = x1, b1 = f1;
    secant_step<FAST>(x0, f0, x1, f1);
    return same_bits(a0, x0) && same_bits(b0, f0) && same_bits(a1, x1) && same_bits(b1, f1);
  }

  // 200-step bisection continuation   c2p.py:238-266 -> (mu, |f| at mu).
  // Once the bracket has shrunk to adjacent doubles (about 55 halvings for a
  // root of order one) the midpoint is an end point and a step maps the state
  // to itself; the loop stops there, with the 200-step result.
  template <bool FAST = false>
  SRMHD_DEV double bisection(double mup, double& fmu) const {
    double x0 = 0.0, x1 = mup * (1.0 - 1e-10);
    double f0 = master_at<FAST>(x0), f1 = master_at<FAST>(x1);
#pragma unroll 1
    for (int it = 0; it < C2P_BISECT; ++it) {
      const double xm = 0.5 * (x0 + x1);
      const double fm = master_at<FAST>(xm);
      bool fixed;
      if ((f0 * fm) <= 0.0) {
        fixed = same_bits(x1, xm) && same_bits(f1, fm);
        x1 = xm;
        f1 = fm;
      } else {
        fixed = same_bits(x0, xm) && same_bits(f0, fm);
        x0 = xm;
        f0 = fm;
      }
      if (fixed) break;
    }
    const bool pick = fabs(f0) < fabs(f1);
    fmu = pick ? f0 : f1;
    return pick ? x0 : x1;
  }

  // primitives at mu   c2p.py:268-310; returns the v-clamp flag
  template <class M>
  SRMHD_DEV bool finish(const M& mu, PrimT<decltype(mu + bt2)>& P) const {
    using R = decltype(mu + bt2);
    R rho, eps;
    master(mu, &rho, &eps);
    const R mc = mu * chi(mu);
    const R mbr = mu * br;
    R v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = mc * (r[i] + mbr * bt[i]);
    const R v2 = sq(v[0]) + sq(v[1]) + sq(v[2]);
    const double lim = 1.0 - 1e-10;
    const bool clamped = v2 > lim;
    const R norm = clamped ? R(csqrt(chi_(v2, lim) / (v2 + TINY))) : R(1.0);
#pragma unroll
    for (int i = 0; i < 3; ++i) P.v[i] = v[i] * norm;
    P.rho = rho;
    P.eps = eps;
    P.p = eos_press(e, eps, rho);
    P.B[0] = B[0];
    P.B[1] = B[1];
    P.B[2] = B[2];
    return clamped;
  }
};
using Kastaun = KastaunT<double>;

// the pass-A walks test for a fixed point at steps 3, 7, 11, ...
SRMHD_DEV bool c2p_fix_test(int it) { return (it & 3) == 3; }

// per-element record of pass A: mu_+ (sign = good bracket), state after 60 steps
struct C2PRec {
  double mup_signed, x1, f1;
};

// Pass A.  Returns the element's not-converged mask (bit k: after k steps).
template <bool FAST>
SRMHD_DEV uint64_t c2p_pass_a_impl(const Kastaun& k, C2PRec& rec) {
  const double mup = k.mu_plus_fast<FAST>();   // bit-identical to k.mu_plus()
  double x0 = 0.0, x1 = mup * (1.0 - 1e-10);
  double f0 = k.master_at<FAST>(x0), f1 = k.master_at<FAST>(x1);
  const bool good = (f0 * f1) < 0.0;
  uint64_t mask = 0;
#pragma unroll 1
  for (int it = 0; it < C2P_SECANT; ++it) {
    const bool nc = !(fabs(f1) < C2P_TOL);
    if (nc) mask |= (uint64_t)1 << it;
    // a fixed point: every later step sees this state, so its bits repeat.
    // Tested every fourth step only (a state a step maps to itself stays so:
    // finding it later changes neither the mask nor the final state)
    if (c2p_fix_test(it) ? k.secant_step_fixed<FAST>(x0, f0, x1, f1)
                          : (k.secant_step<FAST>(x0, f0, x1, f1), false)) {
      if (nc) mask |= C2P_ALL & ~(((uint64_t)2 << it) - 1);
      break;
    }
  }
  rec.mup_signed = good ? mup : -mup;
  rec.x1 = x1;
  rec.f1 = f1;
  return mask;
}

// The wrappers pick the tame form (master_fast) when every active lane of
// the wave holds a tame element; it returns the same bits, in fewer
// instructions.
SRMHD_DEV uint64_t c2p_pass_a(const Kastaun& k, C2PRec& rec) {
  return wave_all(k.tame) ? c2p_pass_a_impl<true>(k, rec) : c2p_pass_a_impl<false>(k, rec);
}

// Pass-A record for batches that often break early (a tube's cells and
// faces).  Besides the final state it keeps where the element's walk became
// stationary and three intermediate secant states, so that pass B reaches
// the state after the batch's K steps with at most a few steps:
//  * jfix: the step from which the walk maps its state to itself (60: none).
//    The state after K >= jfix steps is the final one.
//  * ck[0]: the state after k steps, k = the element's first convergence
//    (|f| < tol; 60 when there is none).  The batch breaks at the first step
//    at which EVERY element has converged, so K >= k whenever K < 60.
//  * ck[1], ck[2]: the states after CK1 and CK2 steps (written only when the
//    walk is still moving there).  Batches that break, break late: after
//    ~49-59 steps on random Riemann tubes, with first convergence at ~5-13
//    steps.
// The walk depends only on (x0, f0, x1, f1), so every replay from a stored
// state is bit-identical to the uninterrupted walk.
constexpr int C2P_CK1 = 52, C2P_CK2 = 56;

struct C2PRecK {
  double mup_signed, x1, f1;       // as C2PRec: the state after 60 steps
  double ck[3][4];                 // (x0, f0, x1, f1) after k, CK1, CK2 steps
  int32_t k, jfix;
};

SRMHD_DEV void rec_put(double* c, double x0, double f0, double x1, double f1) {
  c[0] = x0;
  c[1] = f0;
  c[2] = x1;
  c[3] = f1;
}

template <bool FAST>
SRMHD_DEV uint64_t c2p_pass_a_impl(const Kastaun& k, C2PRecK& rec) {
  const double mup = k.mu_plus_fast<FAST>();
  double x0 = 0.0, x1 = mup * (1.0 - 1e-10);
  double f0 = k.master_at<FAST>(x0), f1 = k.master_at<FAST>(x1);
  const bool good = (f0 * f1) < 0.0;
  uint64_t mask = 0;
  int ke = C2P_SECANT, jf = C2P_SECANT;
#pragma unroll 1
  for (int it = 0; it < C2P_SECANT; ++it) {
    const bool nc = !(fabs(f1) < C2P_TOL);
    if (nc) {
      mask |= (uint64_t)1 << it;
    } else if (ke == C2P_SECANT) {
      ke = it;
      rec_put(rec.ck[0], x0, f0, x1, f1);
    }
    if (it == C2P_CK1) rec_put(rec.ck[1], x0, f0, x1, f1);
    if (it == C2P_CK2) rec_put(rec.ck[2], x0, f0, x1, f1);
    // (jf may then lie up to three steps past the first stationary step:
    // the state after any K >= jf steps is still the final one)
    if (c2p_fix_test(it) ? k.secant_step_fixed<FAST>(x0, f0, x1, f1)
                          : (k.secant_step<FAST>(x0, f0, x1, f1), false)) {
      if (nc) mask |= C2P_ALL & ~(((uint64_t)2 << it) - 1);
      jf = it;
      break;
    }
  }
  rec.mup_signed = good ? mup : -mup;
  rec.x1 = x1;
  rec.f1 = f1;
  rec.k = ke;
  rec.jfix = jf;
  return mask;
}

SRMHD_DEV uint64_t c2p_pass_a(const Kastaun& k, C2PRecK& rec) {
  return wave_all(k.tame) ? c2p_pass_a_impl<true>(k, rec) : c2p_pass_a_impl<false>(k, rec);
}

// break iteration of the batch from the OR of every element's mask
SRMHD_DEV int c2p_break_iter(uint64_t ormask) {
  const uint64_t zeros = ~ormask & ((((uint64_t)1) << C2P_SECANT) - 1);
  return zeros ? __builtin_ctzll(zeros) : C2P_SECANT;
}

// status bits written by the con2prim entry points
enum { C2P_BRACKET_OK = 1, C2P_CONVERGED = 2, C2P_V_CLAMPED = 4, C2P_ATMO = 8, C2P_REDONE = 16 };

// the secant state after the batch's K steps (pass A kept the 60-step one)
template <bool FAST>
SRMHD_DEV void c2p_secant_at_impl(const Kastaun& k, const C2PRec& rec, int K, double& x1,
                                  double& f1) {
  x1 = rec.x1;
  f1 = rec.f1;
  if (K < C2P_SECANT) {
    double x0 = 0.0;
    x1 = fabs(rec.mup_signed) * (1.0 - 1e-10);
    double f0 = k.master_at<FAST>(x0);
    f1 = k.master_at<FAST>(x1);
#pragma unroll 1
    for (int it = 0; it < K; ++it)
      if (k.secant_step_fixed<FAST>(x0, f0, x1, f1)) break;
  }
}

// the same from the checkpointed record: the final state when the walk was
// stationary by step K, else a replay from the latest stored state at or
// before K (K >= rec.k when K < 60)
template <bool FAST>
SRMHD_DEV void c2p_secant_at_impl(const Kastaun& k, const C2PRecK& rec, int K, double& x1,
                                  double& f1) {
  x1 = rec.x1;
  f1 = rec.f1;
  if (K < C2P_SECANT && K < rec.jfix) {
    const int slot = K >= C2P_CK2 ? 2 : (K >= C2P_CK1 ? 1 : 0);
    const int from = slot == 2 ? C2P_CK2 : (slot == 1 ? C2P_CK1 : rec.k);
    const double* c = rec.ck[slot];
    double x0 = c[0], f0 = c[1];
    x1 = c[2];
    f1 = c[3];
#pragma unroll 1
    for (int it = from; it < K; ++it)
      if (k.secant_step_fixed<FAST>(x0, f0, x1, f1)) break;
  }
}

template <class Rec>
SRMHD_DEV void c2p_secant_at(const Kastaun& k, const Rec& rec, int K, double& x1, double& f1) {
  if (wave_all(k.tame))
    c2p_secant_at_impl<true>(k, rec, K, x1, f1);
  else
    c2p_secant_at_impl<false>(k, rec, K, x1, f1);
}

// whether the element goes to the bisection redo: not converged after the
// batch's K secant steps but with a good bracket   c2p.py:338
template <class Rec>
SRMHD_DEV bool c2p_needs_redo(const Kastaun& k, const Rec& rec, int K) {
  double x1, f1;
  c2p_secant_at(k, rec, K, x1, f1);
  return !(fabs(f1) < C2P_TOL) && rec.mup_signed > 0.0;
}

// Pass B: full conservative_to_primitive of one element given its pass-A
// record and the secant state (x1, f1) after the batch's K steps
// (c2p_secant_at).   c2p.py:313-363
// force < 0: the element is redone by bisection when it needs it itself;
// force >= 0: the caller decides (force != 0 redoes), for batches whose redo
// set is wider than the elements that need it.
template <bool FAST, class Rec>
SRMHD_DEV int c2p_pass_at_impl(const Kastaun& k, const Rec& rec, double x1, double f1,
                               double atmo_rho, double atmo_eps, Prim& P, int force) {
  const bool good = rec.mup_signed > 0.0;
  const double mup = fabs(rec.mup_signed);
  bool conv = fabs(f1) < C2P_TOL;
  bool clamped;
  const bool redo = force < 0 ? (!conv && good) : (force != 0);
  if (redo) {
    double fmu;
    const double mu = k.bisection<FAST>(mup, fmu);
    conv = fabs(fmu) < C2P_TOL;
    clamped = k.finish(mu, P);
  } else {
    clamped = k.finish(x1, P);
  }
  const bool atmo = P.rho < atmo_rho;
  if (atmo) {
    P.rho = atmo_rho;
    P.eps = atmo_eps;
    P.v[0] = 0.0;
    P.v[1] = 0.0;
    P.v[2] = 0.0;
  }
  P.p = eos_press(k.e, P.eps, P.rho);
  return (good ? C2P_BRACKET_OK : 0) | (conv ? C2P_CONVERGED : 0) | (clamped ? C2P_V_CLAMPED : 0) |
         (atmo ? C2P_ATMO : 0) | (redo ? C2P_REDONE : 0);
}

template <class Rec>
SRMHD_DEV int c2p_pass_at(const Kastaun& k, const Rec& rec, double x1, double f1,
                          double atmo_rho, double atmo_eps, Prim& P, int force = -1) {
  return wave_all(k.tame) ? c2p_pass_at_impl<true>(k, rec, x1, f1, atmo_rho, atmo_eps, P, force)
                          : c2p_pass_at_impl<false>(k, rec, x1, f1, atmo_rho, atmo_eps, P, force);
}

// the same from the record and the batch break iteration K
template <class Rec>
SRMHD_DEV int c2p_pass_b(const Kastaun& k, const Rec& rec, int K, double atmo_rho,
                         double atmo_eps, Prim& P, int force = -1) {
  double x1, f1;
  c2p_secant_at(k, rec, K, x1, f1);
  return c2p_pass_at(k, rec, x1, f1, atmo_rho, atmo_eps, P, force);
}

// conservative_to_primitive of ONE element as its own batch (c2p.py:313-363
// with a batch of one): the secant stops as soon as this element converges.
// The mixed-precision form uses it for the cells (the batch coupling of the
// FP64 form only moves converged roots by a few ulps).  Returns the status
// bits of c2p_pass_b.
// mu_out (may be NULL) receives the root the primitives were finished at.
// mup_in >= 0 replaces the mu_+ bisection by a given upper bracket end (any
// value at or above mu_+ keeps the root bracketed; the secant converges to
// the same root from it).
template <bool FAST>
SRMHD_DEV int c2p_single_impl(const Kastaun& k, double atmo_rho, double atmo_eps, Prim& P,
                              double* mu_out, double mup_in) {
  const double mup = mup_in >= 0.0 ? mup_in : k.mu_plus_fast<FAST>();
  double x0 = 0.0, x1 = mup * (1.0 - 1e-10);
  double f0 = k.master_at<FAST>(x0), f1 = k.master_at<FAST>(x1);
  const bool good = (f0 * f1) < 0.0;
#pragma unroll 1
  for (int it = 0; it < C2P_SECANT; ++it) {
    if (fabs(f1) < C2P_TOL) break;
    if (k.secant_step_fixed<FAST>(x0, f0, x1, f1)) break;
  }
  bool conv = fabs(f1) < C2P_TOL;
  bool clamped;
  const bool redo = !conv && good;
  if (redo) {
    double fmu;
    const double mu = k.bisection<FAST>(mup, fmu);
    conv = fabs(fmu) < C2P_TOL;
    clamped = k.finish(mu, P);
    if (mu_out) *mu_out = mu;
  } else {
    clamped = k.finish(x1, P);
    if (mu_out) *mu_out = x1;
  }
  const bool atmo = P.rho < atmo_rho;
  if (atmo) {
    P.rho = atmo_rho;
    P.eps = atmo_eps;
    P.v[0] = 0.0;
    P.v[1] = 0.0;
    P.v[2] = 0.0;
  }
  P.p = eos_press(k.e, P.eps, P.rho);
  return (good ? C2P_BRACKET_OK : 0) | (conv ? C2P_CONVERGED : 0) | (clamped ? C2P_V_CLAMPED : 0) |
         (atmo ? C2P_ATMO : 0) | (redo ? C2P_REDONE : 0);
}

SRMHD_DEV int c2p_single(const Kastaun& k, double atmo_rho, double atmo_eps, Prim& P,
                         double* mu_out = nullptr, double mup_in = -1.0) {
  return wave_all(k.tame) ? c2p_single_impl<true>(k, atmo_rho, atmo_eps, P, mu_out, mup_in)
                          : c2p_single_impl<false>(k, atmo_rho, atmo_eps, P, mu_out, mup_in);
}

}  // namespace srmhd
