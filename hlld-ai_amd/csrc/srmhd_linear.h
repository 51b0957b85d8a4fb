// The linearised (Roe-type) flux of one face (src/physics/linearised.py:
// 129-237), host- and device-compilable.
//
// In primitive variables V = (rho, p, v_n, v_t1, v_t2, B_t1, B_t2) with the
// normal field fixed at the mean of the sides: dU/dV and dF/dV of the
// single-state conserved variables and normal flux (side_state, the HLLD
// side-state arithmetic) by central differences at the mean state, A =
// (dU/dV)^-1 dF/dV, A R = R diag(lambda), V_R - V_L = R alpha, and the ray
// state V* = V_L + sum over lambda_k < 0 of alpha_k r_k, whose physical flux
// is the face's.  A face the decomposition cannot be trusted on keeps the
// fallback's answer (the caller's): a non-finite Jacobian or a non-positive
// density or pressure on either side, an exactly singular dU/dV or
// eigenvector matrix, a complex eigenvalue pair (the reference tolerates an
// imaginary part up to imag_tol (1 + |Re|) and then takes real parts of the
// complex eigenvectors; here a pair with any imaginary part falls back), an
// eigenvector matrix with 2-norm condition number >= cond_max, or a V* that
// is not physical.
//
// Where the mean state's spectrum is degenerate (B_n = 0, or B_t = 0, makes
// the Alfven and slow waves coincide with each other or with the contact)
// the eigenvectors of the repeated eigenvalue are any basis of its
// eigenspace.  V* does not depend on the basis, but cond(R) does, and the
// basis LAPACK's geev returns there is, face by face, either well
// conditioned or all but singular (measured on the rotor: cond 3-177 or
// 5e16-3e19, with no rule in the physics that tells which); this code's
// basis is well conditioned.  On such faces the two take the answered /
// fallback branches differently -- a documented difference from the
// reference, not a rounding one.
//
// The linear algebra is this file's own, on three 7x7 matrices per face held
// in LDS (Mat: element (r, c) of a thread's matrix at stride LIN_BLK, so the
// 64 lanes of a wave touch 64 consecutive words):
//   * LU with partial pivoting (the first largest pivot, as LAPACK's idamax);
//   * Householder reduction to Hessenberg form with the transformations
//     accumulated, then the Francis double-shift QR iteration to the real
//     Schur form and the eigenvectors by back-substitution in it, multiplied
//     by the accumulated transformations: the algorithm of the Handbook's
//     orthes / ortran / hqr2 (Wilkinson and Reinsch, Handbook for Automatic
//     Computation II, 1971), restricted to real spectra (no balancing);
//   * the condition number from the singular values of a one-sided
//     (Hestenes) Jacobi orthogonalisation.
// The results agree with LAPACK's (getrf/getrs, geev, gesdd) to rounding,
// not bit for bit.
#pragma once

#include "srmhd_riemann.h"

namespace srmhd {

constexpr int LIN_BLK = 64;   // faces per workgroup of the linear kernels
constexpr int LIN_NV = 7;

struct LinOpts {
  double fd_rel = 1e-5;     // linearised.py:130
  double imag_tol = 1e-6;   // linearised.py:131 (see the note on complex pairs)
  double cond_max = 1e12;   // linearised.py:131
};

// one thread's 7x7 matrix in LDS (or any memory with the same stride)
struct LinMat {
  double* b;
  SRMHD_DEV double& operator()(int r, int c) const { return b[(r * LIN_NV + c) * LIN_BLK]; }
};

// LU with partial pivoting in place; piv[k] = the row swapped into row k.
// false at an exactly zero pivot (LAPACK's info > 0, where torch.linalg.solve
// raises).
SRMHD_DEV bool lin_lu(const LinMat& M, int piv[LIN_NV]) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < LIN_NV; ++k) {
    int p = k;
    double big = fabs(M(k, k));
#pragma unroll
    for (int i = k + 1; i < LIN_NV; ++i) {
      const double a = fabs(M(i, k));
      if (a > big) {
        big = a;
        p = i;
      }
    }
    piv[k] = p;
    if (p != k) {
#pragma unroll 1
      for (int j = 0; j < LIN_NV; ++j) {
        const double t = M(k, j);
        M(k, j) = M(p, j);
        M(p, j) = t;
      }
    }
    const double d = M(k, k);
    if (d == 0.0) {
      ok = false;
      continue;
    }
#pragma unroll 1
    for (int i = k + 1; i < LIN_NV; ++i) {
      const double l = M(i, k) / d;
      M(i, k) = l;
#pragma unroll 1
      for (int j = k + 1; j < LIN_NV; ++j) M(i, j) = M(i, j) - l * M(k, j);
    }
  }
  return ok;
}

// x <- M^-1 x for one right-hand side, from lin_lu's factors
SRMHD_DEV void lin_lu_solve(const LinMat& M, const int piv[LIN_NV], double x[LIN_NV]) {
#pragma unroll
  for (int k = 0; k < LIN_NV; ++k) {
    const int p = piv[k];
    double t = x[k];
#pragma unroll
    for (int i = 0; i < LIN_NV; ++i)
      if (i == p && i != k) {
        x[k] = x[i];
        x[i] = t;
      }
    (void)t;
  }
#pragma unroll
  for (int i = 1; i < LIN_NV; ++i) {
    double s = x[i];
#pragma unroll
    for (int j = 0; j < i; ++j) s = s - M(i, j) * x[j];
    x[i] = s;
  }
#pragma unroll
  for (int i = LIN_NV - 1; i >= 0; --i) {
    double s = x[i];
#pragma unroll
    for (int j = i + 1; j < LIN_NV; ++j) s = s - M(i, j) * x[j];
    x[i] = s / M(i, i);
  }
}

// B <- M^-1 B column by column (B in LDS)
SRMHD_DEV void lin_lu_solve_mat(const LinMat& M, const int piv[LIN_NV], const LinMat& B) {
#pragma unroll 1
  for (int c = 0; c < LIN_NV; ++c) {
    double x[LIN_NV];
#pragma unroll
    for (int r = 0; r < LIN_NV; ++r) x[r] = B(r, c);
    lin_lu_solve(M, piv, x);
#pragma unroll
    for (int r = 0; r < LIN_NV; ++r) B(r, c) = x[r];
  }
}

// Householder reduction of H to upper Hessenberg form (orthes) and the
// accumulated transformation in V (ortran); ort: scratch vector (row 0 of a
// free matrix).  Below the subdiagonal H is left zero.
SRMHD_DEV void lin_hessenberg(const LinMat& H, const LinMat& V, const LinMat& W) {
  constexpr int n = LIN_NV;
  auto ort = [&](int i) -> double& { return W(0, i); };
#pragma unroll 1
  for (int m = 1; m < n - 1; ++m) {
    double scale = 0.0;
#pragma unroll 1
    for (int i = m; i < n; ++i) scale = scale + fabs(H(i, m - 1));
    if (scale != 0.0) {
      double h = 0.0;
#pragma unroll 1
      for (int i = n - 1; i >= m; --i) {
        ort(i) = H(i, m - 1) / scale;
        h = h + ort(i) * ort(i);
      }
      double g = csqrt(h);
      if (ort(m) > 0.0) g = -g;
      h = h - ort(m) * g;
      ort(m) = ort(m) - g;
#pragma unroll 1
      for (int j = m; j < n; ++j) {
        double f = 0.0;
#pragma unroll 1
        for (int i = n - 1; i >= m; --i) f = f + ort(i) * H(i, j);
        f = f / h;
#pragma unroll 1
        for (int i = m; i < n; ++i) H(i, j) = H(i, j) - f * ort(i);
      }
#pragma unroll 1
      for (int i = 0; i < n; ++i) {
        double f = 0.0;
#pragma unroll 1
        for (int j = n - 1; j >= m; --j) f = f + ort(j) * H(i, j);
        f = f / h;
#pragma unroll 1
        for (int j = m; j < n; ++j) H(i, j) = H(i, j) - f * ort(j);
      }
      ort(m) = scale * ort(m);
      H(m, m - 1) = scale * g;
    }
  }
#pragma unroll 1
  for (int i = 0; i < n; ++i)
#pragma unroll 1
    for (int j = 0; j < n; ++j) V(i, j) = (i == j) ? 1.0 : 0.0;
#pragma unroll 1
  for (int m = n - 2; m >= 1; --m) {
    if (H(m, m - 1) != 0.0) {
#pragma unroll 1
      for (int i = m + 1; i < n; ++i) ort(i) = H(i, m - 1);
#pragma unroll 1
      for (int j = m; j < n; ++j) {
        double g = 0.0;
#pragma unroll 1
        for (int i = m; i < n; ++i) g = g + ort(i) * V(i, j);
        g = (g / ort(m)) / H(m, m - 1);
#pragma unroll 1
        for (int i = m; i < n; ++i) V(i, j) = V(i, j) + g * ort(i);
      }
    }
  }
#pragma unroll 1
  for (int i = 2; i < n; ++i)
#pragma unroll 1
    for (int j = 0; j < i - 1; ++j) H(i, j) = 0.0;
}

// Francis double-shift QR of the Hessenberg H to real Schur form, with the
// transformations accumulated into Z (hqr2), then the eigenvectors: on
// return Z holds them as columns (not normalised) and wr(k) = W(0, k) the
// eigenvalues.  false at a complex pair or without convergence (60
// iterations on one eigenvalue).
SRMHD_DEV bool lin_hqr2_real(const LinMat& H, const LinMat& Z, const LinMat& W) {
  constexpr int n = LIN_NV;
  constexpr double eps = 0x1p-52;
  auto wr = [&](int i) -> double& { return W(0, i); };
  double norm = 0.0;
#pragma unroll 1
  for (int i = 0; i < n; ++i)
#pragma unroll 1
    for (int j = (i > 0 ? i - 1 : 0); j < n; ++j) norm = norm + fabs(H(i, j));
  if (!(norm < 1e300)) return false;
  double exshift = 0.0, p = 0.0, q = 0.0, r = 0.0, s = 0.0, z = 0.0, w, x, y;
  int nn = n - 1, iter = 0;
#pragma unroll 1
  while (nn >= 0) {
    int l = nn;
#pragma unroll 1
    while (l > 0) {
      s = fabs(H(l - 1, l - 1)) + fabs(H(l, l));
      if (s == 0.0) s = norm;
      if (fabs(H(l, l - 1)) < eps * s) break;
      --l;
    }
    if (l == nn) {   // one root
      H(nn, nn) = H(nn, nn) + exshift;
      wr(nn) = H(nn, nn);
      --nn;
      iter = 0;
    } else if (l == nn - 1) {   // two roots
      w = H(nn, nn - 1) * H(nn - 1, nn);
      p = (H(nn - 1, nn - 1) - H(nn, nn)) / 2.0;
      q = p * p + w;
      z = csqrt(fabs(q));
      H(nn, nn) = H(nn, nn) + exshift;
      H(nn - 1, nn - 1) = H(nn - 1, nn - 1) + exshift;
      x = H(nn, nn);
      if (!(q >= 0.0)) return false;   // a complex pair (or NaN)
      z = (p >= 0.0) ? p + z : p - z;
      wr(nn - 1) = x + z;
      wr(nn) = wr(nn - 1);
      if (z != 0.0) wr(nn) = x - w / z;
      x = H(nn, nn - 1);
      s = fabs(x) + fabs(z);
      p = x / s;
      q = z / s;
      r = csqrt(p * p + q * q);
      p = p / r;
      q = q / r;
#pragma unroll 1
      for (int j = nn - 1; j < n; ++j) {
        z = H(nn - 1, j);
        H(nn - 1, j) = q * z + p * H(nn, j);
        H(nn, j) = q * H(nn, j) - p * z;
      }
#pragma unroll 1
      for (int i = 0; i <= nn; ++i) {
        z = H(i, nn - 1);
        H(i, nn - 1) = q * z + p * H(i, nn);
        H(i, nn) = q * H(i, nn) - p * z;
      }
#pragma unroll 1
      for (int i = 0; i < n; ++i) {
        z = Z(i, nn - 1);
        Z(i, nn - 1) = q * z + p * Z(i, nn);
        Z(i, nn) = q * Z(i, nn) - p * z;
      }
      nn -= 2;
      iter = 0;
    } else {   // no root yet: a double-shift step on rows l..nn
      x = H(nn, nn);
      y = H(nn - 1, nn - 1);
      w = H(nn, nn - 1) * H(nn - 1, nn);
      if (iter == 10) {   // Wilkinson's exceptional shift
        exshift = exshift + x;
#pragma unroll 1
        for (int i = 0; i <= nn; ++i) H(i, i) = H(i, i) - x;
        s = fabs(H(nn, nn - 1)) + fabs(H(nn - 1, nn - 2));
        x = y = 0.75 * s;
        w = -0.4375 * s * s;
      }
      if (iter == 30) {   // a second exceptional shift
        s = (y - x) / 2.0;
        s = s * s + w;
        if (s > 0.0) {
          s = csqrt(s);
          if (y < x) s = -s;
          s = x - w / ((y - x) / 2.0 + s);
#pragma unroll 1
          for (int i = 0; i <= nn; ++i) H(i, i) = H(i, i) - s;
          exshift = exshift + s;
          x = y = w = 0.964;
        }
      }
      if (++iter > 60) return false;
      int m = nn - 2;
#pragma unroll 1
      while (m >= l) {
        z = H(m, m);
        r = x - z;
        s = y - z;
        p = (r * s - w) / H(m + 1, m) + H(m, m + 1);
        q = H(m + 1, m + 1) - z - r - s;
        r = H(m + 2, m + 1);
        s = fabs(p) + fabs(q) + fabs(r);
        p = p / s;
        q = q / s;
        r = r / s;
        if (m == l) break;
        if (fabs(H(m, m - 1)) * (fabs(q) + fabs(r)) <
            eps * (fabs(p) * (fabs(H(m - 1, m - 1)) + fabs(z) + fabs(H(m + 1, m + 1)))))
          break;
        --m;
      }
#pragma unroll 1
      for (int i = m + 2; i <= nn; ++i) {
        H(i, i - 2) = 0.0;
        if (i > m + 2) H(i, i - 3) = 0.0;
      }
#pragma unroll 1
      for (int k = m; k <= nn - 1; ++k) {
        const bool notlast = k != nn - 1;
        if (k != m) {
          p = H(k, k - 1);
          q = H(k + 1, k - 1);
          r = notlast ? H(k + 2, k - 1) : 0.0;
          x = fabs(p) + fabs(q) + fabs(r);
          if (x == 0.0) continue;
          p = p / x;
          q = q / x;
          r = r / x;
        }
        s = csqrt(p * p + q * q + r * r);
        if (p < 0.0) s = -s;
        if (s != 0.0) {
          if (k != m)
            H(k, k - 1) = -s * x;
          else if (l != m)
            H(k, k - 1) = -H(k, k - 1);
          p = p + s;
          x = p / s;
          y = q / s;
          z = r / s;
          q = q / p;
          r = r / p;
#pragma unroll 1
          for (int j = k; j < n; ++j) {
            p = H(k, j) + q * H(k + 1, j);
            if (notlast) {
              p = p + r * H(k + 2, j);
              H(k + 2, j) = H(k + 2, j) - p * z;
            }
            H(k, j) = H(k, j) - p * x;
            H(k + 1, j) = H(k + 1, j) - p * y;
          }
          const int ie = nn < k + 3 ? nn : k + 3;
#pragma unroll 1
          for (int i = 0; i <= ie; ++i) {
            p = x * H(i, k) + y * H(i, k + 1);
            if (notlast) {
              p = p + z * H(i, k + 2);
              H(i, k + 2) = H(i, k + 2) - p * r;
            }
            H(i, k) = H(i, k) - p;
            H(i, k + 1) = H(i, k + 1) - p * q;
          }
#pragma unroll 1
          for (int i = 0; i < n; ++i) {
            p = x * Z(i, k) + y * Z(i, k + 1);
            if (notlast) {
              p = p + z * Z(i, k + 2);
              Z(i, k + 2) = Z(i, k + 2) - p * r;
            }
            Z(i, k) = Z(i, k) - p;
            Z(i, k + 1) = Z(i, k + 1) - p * q;
          }
        }
      }
    }
  }
  // eigenvectors of the (upper triangular) Schur form, in place in its upper
  // triangle, then times Z
#pragma unroll 1
  for (nn = n - 1; nn >= 0; --nn) {
    p = wr(nn);
    H(nn, nn) = 1.0;
#pragma unroll 1
    for (int i = nn - 1; i >= 0; --i) {
      w = H(i, i) - p;
      r = 0.0;
#pragma unroll 1
      for (int j = i + 1; j <= nn; ++j) r = r + H(i, j) * H(j, nn);
      H(i, nn) = (w != 0.0) ? -r / w : -r / (eps * norm);
      const double t = fabs(H(i, nn));
      if ((eps * t) * t > 1.0) {
#pragma unroll 1
        for (int j = i; j <= nn; ++j) H(j, nn) = H(j, nn) / t;
      }
    }
  }
#pragma unroll 1
  for (int j = n - 1; j >= 0; --j) {
#pragma unroll 1
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
#pragma unroll 1
      for (int k = 0; k <= j; ++k) acc = acc + Z(i, k) * H(k, j);
      Z(i, j) = acc;
    }
  }
  return true;
}

// 2-norm condition number of M (destroyed): one-sided Jacobi rotations until
// every column pair is orthogonal to 1e-15, then the ratio of the largest to
// the smallest column norm (the singular values)
SRMHD_DEV double lin_cond(const LinMat& M) {
  constexpr int n = LIN_NV;
#pragma unroll 1
  for (int sweep = 0; sweep < 40; ++sweep) {
    bool rotated = false;
#pragma unroll 1
    for (int i = 0; i < n - 1; ++i) {
#pragma unroll 1
      for (int j = i + 1; j < n; ++j) {
        double a = 0.0, b = 0.0, g = 0.0;
#pragma unroll 1
        for (int k = 0; k < n; ++k) {
          const double ci = M(k, i), cj = M(k, j);
          a = a + ci * ci;
          b = b + cj * cj;
          g = g + ci * cj;
        }
        if (!(fabs(g) > 1e-15 * csqrt(a * b))) continue;
        rotated = true;
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + csqrt(1.0 + zeta * zeta));
        const double c = 1.0 / csqrt(1.0 + t * t), sn = c * t;
#pragma unroll 1
        for (int k = 0; k < n; ++k) {
          const double ci = M(k, i), cj = M(k, j);
          M(k, i) = c * ci - sn * cj;
          M(k, j) = sn * ci + c * cj;
        }
      }
    }
    if (!rotated) break;
  }
  double smax = 0.0, smin = 1e308;
#pragma unroll 1
  for (int j = 0; j < n; ++j) {
    double a = 0.0;
#pragma unroll 1
    for (int k = 0; k < n; ++k) a = a + M(k, j) * M(k, j);
    const double sv = csqrt(a);
    smax = sv > smax ? sv : smax;
    smin = sv < smin ? sv : smin;
  }
  return smin > 0.0 ? smax / smin : INFINITY;
}

// the tangential axes of the linearised solver (linearised.py:62-67): NOT
// the cyclic order of the HLLD code -- for N = 1 the first tangent is x
template <int N>
struct LinAxes {
  static constexpr int t1 = N == 0 ? 1 : 0;
  static constexpr int t2 = N == 2 ? 1 : 2;
};

// V -> (U, F) of the seven variables (linearised.py:70-87)
template <int N>
SRMHD_DEV void lin_uf(const double V[LIN_NV], double Bn, const srmhd_eos& e, double Uo[LIN_NV],
                      double Fo[LIN_NV]) {
  constexpr int t1 = LinAxes<N>::t1, t2 = LinAxes<N>::t2;
  Prim P;
  P.rho = V[0];
  P.p = V[1];
  P.v[N] = V[2];
  P.v[t1] = V[3];
  P.v[t2] = V[4];
  P.B[N] = Bn;
  P.B[t1] = V[5];
  P.B[t2] = V[6];
  P.eps = eos_eps(e, V[1], V[0]);
  double u[8], f[8], cp, cm;
  side_state<N>(P, e, u, f, cp, cm);
  const int idx[LIN_NV] = {CD, CT, CS + N, CS + t1, CS + t2, CB + t1, CB + t2};
#pragma unroll
  for (int k = 0; k < LIN_NV; ++k) {
    Uo[k] = u[idx[k]];
    Fo[k] = f[idx[k]];
  }
}

// The face.  M1, M2, M3: three 7x7 work matrices.  true: F, U (8 each, the
// flux vector's order) and ps hold the linearised answer; false: the caller
// keeps its fallback.
template <int N>
SRMHD_DEV bool linear_face(const Prim& L, const Prim& R, const srmhd_eos& e, const LinOpts& o,
                           const LinMat& M1, const LinMat& M2, const LinMat& M3, double F[8],
                           double U[8], double& ps) {
  constexpr int t1 = LinAxes<N>::t1, t2 = LinAxes<N>::t2;
  const double VL[LIN_NV] = {L.rho, L.p, L.v[N], L.v[t1], L.v[t2], L.B[t1], L.B[t2]};
  const double VR[LIN_NV] = {R.rho, R.p, R.v[N], R.v[t1], R.v[t2], R.B[t1], R.B[t2]};
  const double Bn = 0.5 * (L.B[N] + R.B[N]);
  double Vm[LIN_NV], h[LIN_NV];
#pragma unroll
  for (int k = 0; k < LIN_NV; ++k) Vm[k] = 0.5 * (VL[k] + VR[k]);
  {
    const double rp = csqrt(fabs(Vm[1]));
    const double sc[LIN_NV] = {fabs(Vm[0]), fabs(Vm[1]), 1.0, 1.0, 1.0, rp + fabs(Vm[5]),
                               rp + fabs(Vm[6])};
#pragma unroll
    for (int k = 0; k < LIN_NV; ++k) h[k] = o.fd_rel * clo(sc[k], 1e-300);
  }
  // the two Jacobians, column j from V_m +- h_j e_j
  bool finite = true;
#pragma unroll 1
  for (int j = 0; j < LIN_NV; ++j) {
    double Vp[LIN_NV], Vq[LIN_NV], hj = 0.0;
#pragma unroll
    for (int k = 0; k < LIN_NV; ++k) {
      Vp[k] = (k == j) ? Vm[k] + h[k] : Vm[k];
      Vq[k] = (k == j) ? Vm[k] - h[k] : Vm[k];
      if (k == j) hj = h[k];
    }
    double Up[LIN_NV], Fp[LIN_NV], Uq[LIN_NV], Fq[LIN_NV];
    lin_uf<N>(Vp, Bn, e, Up, Fp);
    lin_uf<N>(Vq, Bn, e, Uq, Fq);
    const double d = 2.0 * hj;
#pragma unroll
    for (int r = 0; r < LIN_NV; ++r) {
      const double du = (Up[r] - Uq[r]) / d, df = (Fp[r] - Fq[r]) / d;
      finite = finite && (fabs(du) <= DBL_MAX) && (fabs(df) <= DBL_MAX);
      M1(r, j) = du;
      M2(r, j) = df;
    }
  }
  if (!finite || !(VL[0] > 0.0) || !(VR[0] > 0.0) || !(VL[1] > 0.0) || !(VR[1] > 0.0))
    return false;
  // A = (dU/dV)^-1 dF/dV in M2
  int piv[LIN_NV];
  if (!lin_lu(M1, piv)) return false;
  lin_lu_solve_mat(M1, piv, M2);
  // eigen-decomposition: vectors in M1, eigenvalues in row 0 of M3
  lin_hessenberg(M2, M1, M3);
  if (!lin_hqr2_real(M2, M1, M3)) return false;
  double lam[LIN_NV];
#pragma unroll
  for (int k = 0; k < LIN_NV; ++k) lam[k] = M3(0, k);
  // unit columns, the largest component positive (linearised.py:192-200)
#pragma unroll 1
  for (int c = 0; c < LIN_NV; ++c) {
    double big = -1.0, ph = 0.0;
#pragma unroll 1
    for (int r = 0; r < LIN_NV; ++r) {
      const double a = fabs(M1(r, c));
      if (a > big) {
        big = a;
        ph = M1(r, c);
      }
    }
    const double sgn = ph / clo(fabs(ph), 1e-300);
    double nrm = 0.0;
#pragma unroll 1
    for (int r = 0; r < LIN_NV; ++r) {
      M1(r, c) = M1(r, c) * sgn;
      nrm = nrm + M1(r, c) * M1(r, c);
    }
    nrm = clo(csqrt(nrm), 1e-300);
#pragma unroll 1
    for (int r = 0; r < LIN_NV; ++r) M1(r, c) = M1(r, c) / nrm;
  }
  // conditioning of R (a copy in M2)
#pragma unroll 1
  for (int r = 0; r < LIN_NV; ++r)
#pragma unroll 1
    for (int c = 0; c < LIN_NV; ++c) M2(r, c) = M1(r, c);
  const double cond = lin_cond(M2);
  if (!(cond <= DBL_MAX) || !(cond < o.cond_max)) return false;
  // the jump in the eigenvector basis (an LU of a copy in M2)
#pragma unroll 1
  for (int r = 0; r < LIN_NV; ++r)
#pragma unroll 1
    for (int c = 0; c < LIN_NV; ++c) M2(r, c) = M1(r, c);
  if (!lin_lu(M2, piv)) return false;
  double al[LIN_NV];
#pragma unroll
  for (int k = 0; k < LIN_NV; ++k) al[k] = VR[k] - VL[k];
  lin_lu_solve(M2, piv, al);
  // the state on the ray xi = 0: V_L plus the waves running left
  double Vs[LIN_NV];
#pragma unroll
  for (int k = 0; k < LIN_NV; ++k) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < LIN_NV; ++j) acc = acc + M1(k, j) * (al[j] * (lam[j] < 0.0 ? 1.0 : 0.0));
    Vs[k] = VL[k] + acc;
  }
  const double v2 = Vs[2] * Vs[2] + Vs[3] * Vs[3] + Vs[4] * Vs[4];
  bool phys = (Vs[0] > 0.0) && (Vs[1] > 0.0) && (v2 < 1.0);
#pragma unroll
  for (int k = 0; k < LIN_NV; ++k) phys = phys && (fabs(Vs[k]) <= DBL_MAX);
  if (!phys) return false;
  // its flux (the physical flux of V*) and total pressure
  Prim P;
  P.rho = Vs[0];
  P.p = Vs[1];
  P.v[N] = Vs[2];
  P.v[t1] = Vs[3];
  P.v[t2] = Vs[4];
  P.B[N] = Bn;
  P.B[t1] = Vs[5];
  P.B[t2] = Vs[6];
  P.eps = eos_eps(e, Vs[1], Vs[0]);
  double cp, cm;
  side_state<N>(P, e, U, F, cp, cm);
  const double W2 = 1.0 / (1.0 - v2);
  const double eta = Bn * Vs[2] + Vs[5] * Vs[3] + Vs[6] * Vs[4];
  const double b2 = (Bn * Bn + Vs[5] * Vs[5] + Vs[6] * Vs[6]) / W2 + eta * eta;
  ps = Vs[1] + 0.5 * b2;
  return true;
}

}  // namespace srmhd
