// Reconstruction device functions (reconstruction.py:60-338).
//
// Face f along an axis of n cells sits between cells f-1 and f.  QL[f] is the
// right edge of cell f-1 and QR[f] the left edge of cell f; faces 0 and n
// take the adjacent cell's own edge.  High-order cells within the stencil
// half-width of either end use PLM(mc) and the two outermost cells PCM, so
// the roll-around of the reference's shifted arrays is never read.
#pragma once

#include "srmhd_common.h"

namespace srmhd {

enum Limiter { LIM_PCM = 0, LIM_MC = 1, LIM_MINMOD = 2, LIM_WENO5Z = 3, LIM_MP5 = 4, LIM_MP7 = 5 };

SRMHD_DEV constexpr int stencil_halfwidth(int lim) {
  return lim == LIM_PCM ? 0 : (lim == LIM_MC || lim == LIM_MINMOD) ? 1 : (lim == LIM_MP7 ? 3 : 2);
}

template <class T>
SRMHD_DEV T minmod(const T& a, const T& b) {   // reconstruction.py:60-66
  return (a * b > 0.0) ? (fabs(a) < fabs(b) ? a : b) : T(0.0);
}

template <class T>
SRMHD_DEV T mc_lim(const T& a, const T& b) {   // reconstruction.py:69-77
  const T c = 0.5 * (a + b);
  return (a * b > 0.0) ? T(tsign(c) * tmin(2.0 * fabs(a), tmin(2.0 * fabs(b), fabs(c)))) : T(0.0);
}

// WENO5-Z value at the right face of v0   reconstruction.py:111-128
template <class T>
SRMHD_DEV T weno5z(const T& vm2, const T& vm1, const T& v0, const T& vp1, const T& vp2) {
  const T p0 = (2.0 * vm2 - 7.0 * vm1 + 11.0 * v0) / 6.0;
  const T p1 = (-vm1 + 5.0 * v0 + 2.0 * vp1) / 6.0;
  const T p2 = (2.0 * v0 + 5.0 * vp1 - vp2) / 6.0;
  const double c1 = 13.0 / 12.0, c2 = 0.25;
  const T b0 = c1 * sq(vm2 - 2.0 * vm1 + v0) + c2 * sq(vm2 - 4.0 * vm1 + 3.0 * v0);
  const T b1 = c1 * sq(vm1 - 2.0 * v0 + vp1) + c2 * sq(vm1 - vp1);
  const T b2 = c1 * sq(v0 - 2.0 * vp1 + vp2) + c2 * sq(3.0 * v0 - 4.0 * vp1 + vp2);
  const T t5 = fabs(b0 - b2);
  const double eps_w = 1e-40;
  const T a0 = 0.1 * (1.0 + sq(t5 / (b0 + eps_w)));
  const T a1 = 0.6 * (1.0 + sq(t5 / (b1 + eps_w)));
  const T a2 = 0.3 * (1.0 + sq(t5 / (b2 + eps_w)));
  return (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2);
}

// Suresh-Huynh monotonicity-preserving clip   reconstruction.py:136-162
template <class T>
SRMHD_DEV T mp_limit(const T& vl, const T& vm2, const T& vm1, const T& v0, const T& vp1,
                     const T& vp2) {
  const double alpha_mp = 4.0;
  const T vmp = v0 + minmod(vp1 - v0, alpha_mp * (v0 - vm1));
  if (!((vl - v0) * (vl - vmp) > 1e-40)) return vl;
  const T dm1 = vm2 - 2.0 * vm1 + v0;
  const T d0 = vm1 - 2.0 * v0 + vp1;
  const T dp1 = v0 - 2.0 * vp1 + vp2;
  const T dm4p = minmod(minmod(4.0 * d0 - dp1, 4.0 * dp1 - d0), minmod(d0, dp1));
  const T dm4m = minmod(minmod(4.0 * dm1 - d0, 4.0 * d0 - dm1), minmod(dm1, d0));
  const T vul = v0 + alpha_mp * (v0 - vm1);
  const T vav = 0.5 * (v0 + vp1);
  const T vmd = vav - 0.5 * dm4p;
  const T vlc = v0 + 0.5 * (v0 - vm1) + (4.0 / 3.0) * dm4m;
  const T vmin = tmax(tmin(tmin(v0, vp1), vmd), tmin(tmin(v0, vul), vlc));
  const T vmax = tmin(tmax(tmax(v0, vp1), vmd), tmax(tmax(v0, vul), vlc));
  return vl + minmod(vmin - vl, vmax - vl);
}

template <class T>
SRMHD_DEV T mp5(const T& vm2, const T& vm1, const T& v0, const T& vp1, const T& vp2) {
  const T vl = (2.0 * vm2 - 13.0 * vm1 + 47.0 * v0 + 27.0 * vp1 - 3.0 * vp2) / 60.0;
  return mp_limit(vl, vm2, vm1, v0, vp1, vp2);
}

template <class T>
SRMHD_DEV T mp7(const T& vm3, const T& vm2, const T& vm1, const T& v0, const T& vp1,
                const T& vp2, const T& vp3) {
  const T vl = (-3.0 * vm3 + 25.0 * vm2 - 101.0 * vm1 + 319.0 * v0 + 214.0 * vp1 -
                     38.0 * vp2 + 4.0 * vp3) / 420.0;
  return mp_limit(vl, vm2, vm1, v0, vp1, vp2);
}

// Right (right=true) or left edge of cell c along an axis of n cells.
// q[k] holds cell c + k - 3 for k = 0..6 (entries outside the axis unread).
template <int LIM, class T>
SRMHD_DEV T cell_edge(const T q[7], int c, int n, bool right) {
  constexpr int w = stencil_halfwidth(LIM);
  const T v0 = q[3];
  if (LIM == LIM_PCM) return right ? T(v0 + 0.5 * 0.0) : T(v0 - 0.5 * 0.0);
  const bool high = (LIM == LIM_WENO5Z || LIM == LIM_MP5 || LIM == LIM_MP7);
  if (!high || n <= 2 * w || c < w || c >= n - w) {
    // PLM; slope zero on the two outermost cells   reconstruction.py:225-236
    T s = T(0.0);
    if (c >= 1 && c <= n - 2) {
      const T a = v0 - q[2], b = q[4] - v0;
      s = (LIM == LIM_MINMOD) ? minmod(a, b) : mc_lim(a, b);
    }
    return right ? T(v0 + 0.5 * s) : T(v0 - 0.5 * s);
  }
  // mirrored stencil for the left edge   reconstruction.py:196-198
  if (LIM == LIM_WENO5Z)
    return right ? weno5z(q[1], q[2], q[3], q[4], q[5]) : weno5z(q[5], q[4], q[3], q[2], q[1]);
  if (LIM == LIM_MP5)
    return right ? mp5(q[1], q[2], q[3], q[4], q[5]) : mp5(q[5], q[4], q[3], q[2], q[1]);
  return right ? mp7(q[0], q[1], q[2], q[3], q[4], q[5], q[6])
               : mp7(q[6], q[5], q[4], q[3], q[2], q[1], q[0]);
}

// Left and right states of face f from the window q[k] = cell f + k - 4
// (k = 0..7), as reconstruct(): QL[f] = right edge of cell f-1, QR[f] = left
// edge of cell f; face 0 uses cell 0's right edge, face n cell n-1's left edge.
// (Both window offsets are compile-time constants, so q stays in registers.)
template <int LIM, class T>
SRMHD_DEV void face_states(const T q[8], int f, int n, T& ql, T& qr) {
  ql = (f == 0) ? cell_edge<LIM>(q + 1, 0, n, true) : cell_edge<LIM>(q, f - 1, n, true);
  qr = (f == n) ? cell_edge<LIM>(q, n - 1, n, false) : cell_edge<LIM>(q + 1, f, n, false);
}

}  // namespace srmhd
