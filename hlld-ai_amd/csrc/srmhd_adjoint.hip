// Reverse-mode derivative of the FP64 2D stage (include/srmhd.h,
// srmhd2d_stage_adjoint): the differentiable rollout of rk_step_ct
// (src/physics/driver2d.py:221-249) with dt held fixed.
//
// Structure.  The stage is a composition of local maps; its adjoint runs them
// backwards:
//   con2prim of every cell      -> per cell, implicit differentiation of the
//                                  Kastaun root mu (c2p.py:129-158):
//                                  dmu = -(df/dU) / (df/dmu);
//   BC fills                    -> ghost adjoints folded onto their sources;
//   RK combination              -> scaled copies;
//   flux divergence, ct_rhs     -> differences of the cell / face adjoints;
//   Gardiner-Stone corner EMF   -> linear with the taped upwind signs;
//   reconstruction + Riemann    -> per face, the Jacobian of the flux with
//                                  respect to the 15 face-state inputs and of
//                                  the face states with respect to the stencil
//                                  (limiter branches as the forward took them);
//                                  HLLD's p* by implicit differentiation of the
//                                  residual at the taped root (hlld.py:586-602),
//                                  the HLLE flux where the forward fell back;
//                                  HLLD's part after the side states in reverse
//                                  mode (srmhd_flux_adj.h).
// The other local Jacobians come from forward-mode dual numbers evaluated on
// the same templated device code as the forward (srmhd_common.h ..
// srmhd_riemann.h), compiled with -ffp-contract=off like the forward, so every
// value, and therefore every branch, is the forward's own; the reverse-mode
// code recomputes the forward's values with the same functions and takes the
// same branches.
//
// Every adjoint accumulation is a gather in a fixed order (no atomics), so the
// gradient is deterministic.
#include "srmhd_grid2d.h"
#include "srmhd_recon.h"
#include "srmhd_riemann.h"
#include "srmhd_flux_adj.h"

using namespace srmhd;

namespace {

using D1 = Dual<1>;

#ifndef SRMHD_C2P_ADJ_W
#define SRMHD_C2P_ADJ_W 8
#endif
constexpr int C2P_ADJ_W = SRMHD_C2P_ADJ_W;   // 1, 2, 4 or 8
static_assert(8 % C2P_ADJ_W == 0, "the 8 conserved inputs in whole passes");

struct CFlx {  // read-only flux adjoints: D, Sx, Sy, Sz, tau, Bz, E-component
  const double* p[7];
};

// ── BC preimages ────────────────────────────────────────────────────────────
// Indices g of an axis whose BC source is i (i itself first); 0 when i is a
// ghost (its value is a copy, its adjoint is folded elsewhere).
__device__ inline int pre_cell(int i, int n, int ng, int bc, int out[8]) {
  if (src_cell(i, n, ng, bc) != i) return 0;
  int c = 0;
  out[c++] = i;
  for (int q = 0; q < ng; ++q)
    if (src_cell(q, n, ng, bc) == i && c < 8) out[c++] = q;
  for (int q = ng + n; q < n + 2 * ng; ++q)
    if (src_cell(q, n, ng, bc) == i && c < 8) out[c++] = q;
  return c;
}
__device__ inline int pre_face(int f, int n, int ng, int bc, int out[8]) {
  if (src_face(f, n, ng, bc) != f) return 0;
  int c = 0;
  out[c++] = f;
  for (int q = 0; q < ng; ++q)
    if (q != f && src_face(q, n, ng, bc) == f && c < 8) out[c++] = q;
  for (int q = ng + n; q <= n + 2 * ng; ++q)
    if (q != f && src_face(q, n, ng, bc) == f && c < 8) out[c++] = q;
  return c;
}

__global__ void k_zero(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK)
    p[i] = 0.0;
}

// ── con2prim (c2p.py:313-363) ───────────────────────────────────────────────
// lamP: the adjoint of the stage-output primitives (all cells; ghost entries
// are copies of sources and are gathered here); adds the adjoint of the cell
// cons to lamU and writes the adjoint of the cell-centred Bx, By (the face
// averages the con2prim reads) to lBcx, lBcy, interior cells only.
__global__ __launch_bounds__(BLK) void k_adj_c2p(G2 g, St2 U, srmhd_eos e, double atmo_rho,
                                                 double atmo_eps, Pr2 lamP, St2Out lamU,
                                                 double* lBcx, double* lBcy) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (t >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(t / g.ny) + g.ng, j = (int)(t % g.ny) + g.ng;
  const int64_t c = (int64_t)i * g.nyt + j;
  int px[8], py[8];
  const int nxp = pre_cell(i, g.nx, g.ng, g.bcx, px);
  const int nyp = pre_cell(j, g.ny, g.ng, g.bcy, py);
  double lp[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int a = 0; a < nxp; ++a)
    for (int b = 0; b < nyp; ++b) {
      const int64_t q = (int64_t)px[a] * g.nyt + py[b];
#pragma unroll
      for (int k = 0; k < 6; ++k) lp[k] += lamP.p[k][q];
    }
  double u[8];
  cell_cons(g, U, i, j, u);
  Kastaun k;
  k.init(u, e);
  Prim q;
  double mu;
  c2p_single(k, atmo_rho, atmo_eps, q, &mu);
  double lu[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (!(q.rho < atmo_rho)) {   // the atmosphere floor is constant in U
    const double fmu = k.master(D1::seed(mu, 0)).d[0];
    constexpr int CW = C2P_ADJ_W;   // tangent directions of U per pass
    using DW = Dual<CW>;
#pragma unroll 1
    for (int base = 0; base < 8; base += CW) {
      DW ud[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) ud[m] = DW::seed(u[m], m - base);
      KastaunT<DW> kd;
      kd.init(ud, e);
      const DW fth = kd.master(mu);
      DW mud(mu);
#pragma unroll
      for (int c = 0; c < CW; ++c) mud.d[c] = -fth.d[c] / fmu;
      PrimT<DW> Pd;
      kd.finish(mud, Pd);
#pragma unroll
      for (int c = 0; c < CW; ++c)
        lu[base + c] = lp[0] * Pd.rho.d[c] + lp[1] * Pd.v[0].d[c] + lp[2] * Pd.v[1].d[c] +
                       lp[3] * Pd.v[2].d[c] + lp[4] * Pd.p.d[c] + lp[5] * Pd.eps.d[c];
    }
  }
  lamU.p[0][c] += lu[CD];
  lamU.p[1][c] += lu[CS];
  lamU.p[2][c] += lu[CS + 1];
  lamU.p[3][c] += lu[CS + 2];
  lamU.p[4][c] += lu[CT];
  lamU.p[5][c] += lu[CB + 2];
  lBcx[c] = lu[CB];
  lBcy[c] = lu[CB + 1];
}

// Face adjoints from the cell-centred-B adjoints (ct.py:44-54 transposed),
// plus the normal-B adjoints of the sweeps when given.
__global__ __launch_bounds__(BLK) void k_adj_bc2face(G2 g, const double* lBcx, const double* lBcy,
                                                     const double* lBnx, const double* lBny,
                                                     double* lamBxf, double* lamByf) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const int nc = g.nyt + 1;
  if (t >= (int64_t)(g.nxt + 1) * nc) return;
  const int i = (int)(t / nc), j = (int)(t - (int64_t)i * nc);
  if (j < g.nyt) {   // x-face (i, j)
    const int64_t f = (int64_t)i * g.nyt + j;
    double s = 0.0;
    if (i >= 1) s += lBcx[(int64_t)(i - 1) * g.nyt + j];
    if (i < g.nxt) s += lBcx[(int64_t)i * g.nyt + j];
    lamBxf[f] += 0.5 * s + (lBnx ? lBnx[f] : 0.0);
  }
  if (i < g.nxt) {   // y-face (i, j)
    const int64_t f = (int64_t)i * (g.nyt + 1) + j;
    double s = 0.0;
    if (j >= 1) s += lBcy[(int64_t)i * g.nyt + j - 1];
    if (j < g.nyt) s += lBcy[(int64_t)i * g.nyt + j];
    lamByf[f] += 0.5 * s + (lBny ? lBny[f] : 0.0);
  }
}

// Fold the ghost adjoints of a state onto their BC sources (the face and cell
// BCs of sync_state, grid.py:93-143, as k_update applies them): out = the
// adjoint with respect to the source elements, 0 at ghosts.
__global__ __launch_bounds__(BLK) void k_adj_fold(G2 g, St2 lam, St2Out out) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const int nc = g.nyt + 1;
  if (t >= (int64_t)(g.nxt + 1) * nc) return;
  const int i = (int)(t / nc), j = (int)(t - (int64_t)i * nc);
  int pa[8], pb[8];
  if (i < g.nxt && j < g.nyt) {
    const int na = pre_cell(i, g.nx, g.ng, g.bcx, pa), nb = pre_cell(j, g.ny, g.ng, g.bcy, pb);
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int a = 0; a < na; ++a)
      for (int b = 0; b < nb; ++b) {
        const int64_t q = (int64_t)pa[a] * g.nyt + pb[b];
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] += lam.p[k][q];
      }
    const int64_t c = (int64_t)i * g.nyt + j;
#pragma unroll
    for (int k = 0; k < 6; ++k) out.p[k][c] = s[k];
  }
  if (j < g.nyt) {
    const int na = pre_face(i, g.nx, g.ng, g.bcx, pa), nb = pre_cell(j, g.ny, g.ng, g.bcy, pb);
    double s = 0.0;
    for (int a = 0; a < na; ++a)
      for (int b = 0; b < nb; ++b) s += lam.p[6][(int64_t)pa[a] * g.nyt + pb[b]];
    out.p[6][(int64_t)i * g.nyt + j] = s;
  }
  if (i < g.nxt) {
    const int na = pre_cell(i, g.nx, g.ng, g.bcx, pa), nb = pre_face(j, g.ny, g.ng, g.bcy, pb);
    double s = 0.0;
    for (int a = 0; a < na; ++a)
      for (int b = 0; b < nb; ++b) s += lam.p[7][(int64_t)pa[a] * (g.nyt + 1) + pb[b]];
    out.p[7][(int64_t)i * (g.nyt + 1) + j] = s;
  }
}

// RK combination (state.py:77-103) transposed: lamU0 += a LE, lamU += b LE.
__global__ __launch_bounds__(BLK) void k_adj_combine(G2 g, St2 LE, double a, double b,
                                                     St2Out lamU0, St2Out lamU) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const int nc = g.nyt + 1;
  if (t >= (int64_t)(g.nxt + 1) * nc) return;
  const int i = (int)(t / nc), j = (int)(t - (int64_t)i * nc);
  if (i < g.nxt && j < g.nyt) {
    const int64_t c = (int64_t)i * g.nyt + j;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double l = LE.p[k][c];
      lamU0.p[k][c] += a * l;
      lamU.p[k][c] += b * l;
    }
  }
  if (j < g.nyt) {
    const int64_t f = (int64_t)i * g.nyt + j;
    const double l = LE.p[6][f];
    lamU0.p[6][f] += a * l;
    lamU.p[6][f] += b * l;
  }
  if (i < g.nxt) {
    const int64_t f = (int64_t)i * (g.nyt + 1) + j;
    const double l = LE.p[7][f];
    lamU0.p[7][f] += a * l;
    lamU.p[7][f] += b * l;
  }
}

// ct_rhs (ct.py:180-185) inside the combination, transposed: the corner-EMF
// adjoint from the face adjoints.  w = c dt.
__global__ __launch_bounds__(BLK) void k_adj_ez(G2 g, St2 LE, const double* dt_dev, double c,
                                                double* lEz) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const int nc = g.nyt + 1;
  if (t >= (int64_t)(g.nxt + 1) * nc) return;
  const int i = (int)(t / nc), j = (int)(t - (int64_t)i * nc);
  if (i == 0 || j == 0 || i == g.nxt || j == g.nyt) {   // the ring of zeros
    lEz[t] = 0.0;
    return;
  }
  const double w = c * (*dt_dev);
  const double* Lx = LE.p[6];
  const double* Ly = LE.p[7];
  const double sx = Lx[(int64_t)i * g.nyt + j] - Lx[(int64_t)i * g.nyt + j - 1];
  const double sy = Ly[(int64_t)(i - 1) * (g.nyt + 1) + j] - Ly[(int64_t)i * (g.nyt + 1) + j];
  lEz[t] = w * (sx / g.dy) + w * (sy / g.dx);
}

__device__ inline double usign(double F, double th) { return (fabs(F) > th) ? tsign(F) : 0.0; }

// corner (ci, cj) of the interior ring?
__device__ inline bool ring_in(const G2& g, int ci, int cj) {
  return ci >= 1 && cj >= 1 && ci <= g.nxt - 1 && cj <= g.nyt - 1;
}

// Flux adjoints of the x sweep: the divergence part for the six cell keys,
// the corner-EMF part for F^x(By) (Efx = -F^x(By), ct.py:119-177).
__global__ __launch_bounds__(BLK) void k_adj_fx(G2 g, St2 LE, const double* lEz,
                                                const double* FDx, const double* FDy,
                                                const double* th, int upwind,
                                                const double* dt_dev, double c, Flx lF) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (t >= (int64_t)(g.nxt + 1) * g.nyt) return;
  const int i = (int)(t / g.nyt), j = (int)(t - (int64_t)i * g.nyt);
  const double w = c * (*dt_dev);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double r = i < g.nxt ? LE.p[k][(int64_t)i * g.nyt + j] : 0.0;
    const double l = i >= 1 ? LE.p[k][(int64_t)(i - 1) * g.nyt + j] : 0.0;
    lF.p[k][t] = w * (r / g.dx) - w * (l / g.dx);
  }
  const int nc = g.nyt + 1;
  double s = 0.0;
  // exf of corner (i, j)
  if (ring_in(g, i, j)) {
    double d = 0.25;
    if (upwind) {
      const double Sy = usign(FDy[(int64_t)i * (g.nyt + 1) + j], th[1]);
      const double Sym = usign(FDy[(int64_t)(i - 1) * (g.nyt + 1) + j], th[1]);
      d = 0.25 + 0.125 * ((1.0 - Sym) + (1.0 - Sy));
    }
    s += lEz[(int64_t)i * nc + j] * d;
  }
  // exfm of corner (i, j + 1)
  if (ring_in(g, i, j + 1)) {
    double d = 0.25;
    if (upwind) {
      const double Sy = usign(FDy[(int64_t)i * (g.nyt + 1) + j + 1], th[1]);
      const double Sym = usign(FDy[(int64_t)(i - 1) * (g.nyt + 1) + j + 1], th[1]);
      d = 0.25 + 0.125 * ((1.0 + Sym) + (1.0 + Sy));
    }
    s += lEz[(int64_t)i * nc + j + 1] * d;
  }
  lF.p[6][t] = -s;
  (void)FDx;
}

// Flux adjoints of the y sweep (Efy = +F^y(Bx)).
__global__ __launch_bounds__(BLK) void k_adj_fy(G2 g, St2 LE, const double* lEz,
                                                const double* FDx, const double* FDy,
                                                const double* th, int upwind,
                                                const double* dt_dev, double c, Flx lF) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (t >= (int64_t)g.nxt * (g.nyt + 1)) return;
  const int i = (int)(t / (g.nyt + 1)), j = (int)(t - (int64_t)i * (g.nyt + 1));
  const double w = c * (*dt_dev);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double r = j < g.nyt ? LE.p[k][(int64_t)i * g.nyt + j] : 0.0;
    const double l = j >= 1 ? LE.p[k][(int64_t)i * g.nyt + j - 1] : 0.0;
    lF.p[k][t] = w * (r / g.dy) - w * (l / g.dy);
  }
  const int nc = g.nyt + 1;
  double s = 0.0;
  // eyf of corner (i, j)
  if (ring_in(g, i, j)) {
    double d = 0.25;
    if (upwind) {
      const double Sx = usign(FDx[(int64_t)i * g.nyt + j], th[0]);
      const double Sxm = usign(FDx[(int64_t)i * g.nyt + j - 1], th[0]);
      d = 0.25 + 0.125 * ((1.0 - Sxm) + (1.0 - Sx));
    }
    s += lEz[(int64_t)i * nc + j] * d;
  }
  // eyfm of corner (i + 1, j)
  if (ring_in(g, i + 1, j)) {
    double d = 0.25;
    if (upwind) {
      const double Sx = usign(FDx[(int64_t)(i + 1) * g.nyt + j], th[0]);
      const double Sxm = usign(FDx[(int64_t)(i + 1) * g.nyt + j - 1], th[0]);
      d = 0.25 + 0.125 * ((1.0 + Sxm) + (1.0 + Sx));
    }
    s += lEz[(int64_t)(i + 1) * nc + j] * d;
  }
  lF.p[6][t] = s;
  (void)FDy;
}

// Cell-EMF adjoint (E_c = Bcx vy - Bcy vx, ct.py:57-61) from the four
// corners the cell touches; adds to lamP (vx, vy) and writes lBcx, lBcy.
__global__ __launch_bounds__(BLK) void k_adj_ec(G2 g, St2 S, Pr2 P, const double* lEz,
                                                const double* FDx, const double* FDy,
                                                const double* th, int upwind, Pr2Out lamP,
                                                double* lBcx, double* lBcy) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (t >= (int64_t)g.nxt * g.nyt) return;
  const int ci = (int)(t / g.nyt), cj = (int)(t - (int64_t)ci * g.nyt);
  double lEc = 0.0;
  if (upwind) {
    const int nc = g.nyt + 1;
    auto sx = [&](int a, int b) { return usign(FDx[(int64_t)a * g.nyt + b], th[0]); };
    auto sy = [&](int a, int b) { return usign(FDy[(int64_t)a * (g.nyt + 1) + b], th[1]); };
    if (ring_in(g, ci, cj))   // NE of corner (ci, cj)
      lEc += lEz[(int64_t)ci * nc + cj] * (-0.125 * ((1.0 - sx(ci, cj)) + (1.0 - sy(ci, cj))));
    if (ring_in(g, ci + 1, cj))   // NW of corner (ci + 1, cj)
      lEc += lEz[(int64_t)(ci + 1) * nc + cj] *
             (-0.125 * ((1.0 + sx(ci + 1, cj)) + (1.0 - sy(ci, cj))));
    if (ring_in(g, ci, cj + 1))   // SE of corner (ci, cj + 1)
      lEc += lEz[(int64_t)ci * nc + cj + 1] *
             (-0.125 * ((1.0 - sx(ci, cj)) + (1.0 + sy(ci, cj + 1))));
    if (ring_in(g, ci + 1, cj + 1))   // SW of corner (ci + 1, cj + 1)
      lEc += lEz[(int64_t)(ci + 1) * nc + cj + 1] *
             (-0.125 * ((1.0 + sx(ci + 1, cj)) + (1.0 + sy(ci, cj + 1))));
  }
  const double vx = P.p[1][t], vy = P.p[2][t];
  const double Bx = bcx(g, S.p[6], ci, cj), By = bcy(g, S.p[7], ci, cj);
  lamP.p[1][t] += -(lEc * By);
  lamP.p[2][t] += lEc * Bx;
  lBcx[t] = lEc * vy;
  lBcy[t] = -(lEc * vx);
}

// The stencil window of face t of the sweep along N, with face_prims2d's
// arithmetic: a[v][k] for the reconstructed variables (rho, p, z_x, z_y,
// z_z, B_t1, B_t2) at slots k0..k1; the face index f along the sweep, the
// line length n and whether the face took the PCM fallback.
template <int N, int LIM>
__device__ inline void adj_window(const G2& g, const St2& S, const Pr2& P, int64_t t,
                                  double a[7][8], int& f, int& n, bool& bad, double ql[7],
                                  double qr[7], double& Bn) {
  constexpr int hw = stencil_halfwidth(LIM);
  constexpr int k0 = 3 - hw, k1 = 4 + hw;
  int i, j;
  face_ij<N>(g, t, i, j);
  n = N == 0 ? g.nxt : g.nyt;
  f = N == 0 ? i : j;
#pragma unroll
  for (int v = 0; v < 7; ++v)
#pragma unroll
    for (int k = 0; k < 8; ++k) a[v][k] = 0.0;
#pragma unroll
  for (int k = k0; k <= k1; ++k) {
    const int cc = clampi(f + k - 4, 0, n - 1);
    const int64_t c = N == 0 ? (int64_t)cc * g.nyt + j : (int64_t)i * g.nyt + cc;
    const int ci = N == 0 ? cc : i, cj = N == 0 ? j : cc;
    a[0][k] = P.p[0][c];
    a[1][k] = P.p[4][c];
    const double vx = P.p[1][c], vy = P.p[2][c], vz = P.p[3][c];
    const double Wk = 1.0 / csqrt(clo(1.0 - (sq(vx) + sq(vy) + sq(vz)), 1e-14));
    a[2][k] = Wk * vx;
    a[3][k] = Wk * vy;
    a[4][k] = Wk * vz;
    // N=0: t1 = y (faces), t2 = z (cons);  N=1: t1 = z (cons), t2 = x (faces)
    a[5][k] = N == 0 ? bcy(g, S.p[7], ci, cj) : S.p[5][c];
    a[6][k] = N == 0 ? S.p[5][c] : bcx(g, S.p[6], ci, cj);
  }
#pragma unroll
  for (int v = 0; v < 2; ++v) face_states<LIM>(a[v], f, n, ql[v], qr[v]);
  bad = (ql[0] <= 1e-12) || (qr[0] <= 1e-12) || (ql[1] <= 1e-14) || (qr[1] <= 1e-14);
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    if (bad) {
      ql[v] = a[v][f == 0 ? 4 : 3] + 0.5 * 0.0;
      qr[v] = a[v][f == n ? 3 : 4] - 0.5 * 0.0;
    } else if (v >= 2) {
      face_states<LIM>(a[v], f, n, ql[v], qr[v]);
    }
  }
  Bn = N == 0 ? S.p[6][(int64_t)i * g.nyt + j] : S.p[7][(int64_t)i * (g.nyt + 1) + j];
}

// Per face of the sweep along N, the adjoint of the flux (sweep.py:30-119)
// with respect to its 15 face inputs: LQ[m * nf + t] for the 14 face-state
// inputs (left 0..6, right 7..13) and lBn[t] for the normal field.  lamF: the
// seven flux adjoints of the face.  HLLD in reverse mode; HLLE, HLLC by one
// dual pass per input.
template <int N, int LIM, int SOL>
__global__ __launch_bounds__(BLK) void k_adj_flux(G2 g, St2 S, Pr2 P, srmhd_eos e,
                                                  const double* pstar, CFlx lamF, double* LQ,
                                                  double* lBn) {
  const int64_t nf = n_faces<N>(g);
  const int64_t t = xcd_block() * BLK + threadIdx.x;
  if (t >= nf) return;
  double ql[7], qr[7], Bn;
  {
    double a[7][8];
    int f, n;
    bool bad;
    adj_window<N, LIM>(g, S, P, t, a, f, n, bad, ql, qr, Bn);
  }
  // flux adjoints in flux-vector order (the normal-B flux is zeroed: none)
  double lf[8];
  lf[CD] = lamF.p[0][t];
  lf[CS] = lamF.p[1][t];
  lf[CS + 1] = lamF.p[2][t];
  lf[CS + 2] = lamF.p[3][t];
  lf[CT] = lamF.p[4][t];
  lf[CB + N] = 0.0;
  lf[CB + 2] = lamF.p[5][t];
  lf[N == 0 ? CB + 1 : CB + 0] = lamF.p[6][t];
  const double ps = SOL == SOLVER_HLLD ? pstar[t] : -1.0;
  double lq[15];
  if constexpr (SOL == SOLVER_HLLD) {
    // the face states wait in LDS (each lane its own slots) while the core
    // runs, so that they hold no registers there
    __shared__ double qs[15][BLK];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      qs[m][threadIdx.x] = ql[m];
      qs[7 + m][threadIdx.x] = qr[m];
    }
    qs[14][threadIdx.x] = Bn;
    double luL[8], lfL[8], luR[8], lfR[8], lc[4], l8[8];
    hlld_core_rev<N>(ql, qr, Bn, e, ps, lf, luL, lfL, luR, lfR, lc);
    __asm__ volatile("" ::: "memory");   // reload below, no forwarding
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      ql[m] = qs[m][threadIdx.x];
      qr[m] = qs[7 + m][threadIdx.x];
    }
    Bn = qs[14][threadIdx.x];
    side_rev<N>(ql, Bn, e, luL, lfL, lc[0], lc[1], l8);
#pragma unroll
    for (int m = 0; m < 7; ++m) lq[m] = l8[m];
    lq[14] = l8[7];
    side_rev<N>(qr, Bn, e, luR, lfR, lc[2], lc[3], l8);
#pragma unroll
    for (int m = 0; m < 7; ++m) lq[7 + m] = l8[m];
    lq[14] += l8[7];
  } else {
    flux_vjp_dual<N, SOL>(ql, qr, Bn, e, ps, lf, lq);
  }
#pragma unroll
  for (int m = 0; m < 14; ++m) LQ[(int64_t)m * nf + t] = lq[m];
  lBn[t] = lq[14];
}

// Per face of the sweep along N: the adjoint of reconstruct_prims
// (reconstruction.py:272-338) from the face-state adjoints LQ -- the face
// states with respect to the stencil window, every limiter branch as the
// forward took it (all slots at once); writes W[(v*KW + s)*nf + t], the
// adjoint of variable v of (rho, p, z_x, z_y, z_z, B_t1, B_t2) at window slot
// k0 + s.
template <int N, int LIM>
__global__ __launch_bounds__(BLK) void k_adj_recon(G2 g, St2 S, Pr2 P, const double* LQ,
                                                   double* W) {
  constexpr int hw = stencil_halfwidth(LIM);
  constexpr int k0 = 3 - hw, k1 = 4 + hw, KW = k1 - k0 + 1;
  const int64_t nf = n_faces<N>(g);
  const int64_t t = xcd_block() * BLK + threadIdx.x;
  if (t >= nf) return;
  double a[7][8], ql[7], qr[7], Bn;
  int f, n;
  bool bad;
  adj_window<N, LIM>(g, S, P, t, a, f, n, bad, ql, qr, Bn);
  // unrolled: a[v][.] with a run-time v would put the window in scratch
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    Dual<KW> aa[8];
#pragma unroll
    for (int k = k0; k <= k1; ++k) aa[k] = Dual<KW>::seed(a[v][k], k - k0);
    Dual<KW> dl, dr;
    if (bad) {
      dl = aa[f == 0 ? 4 : 3] + 0.5 * 0.0;
      dr = aa[f == n ? 3 : 4] - 0.5 * 0.0;
    } else {
      face_states<LIM>(aa, f, n, dl, dr);
    }
    const double wl = LQ[(int64_t)v * nf + t], wr = LQ[(int64_t)(7 + v) * nf + t];
#pragma unroll
    for (int s = 0; s < KW; ++s) W[(int64_t)(v * KW + s) * nf + t] = wl * dl.d[s] + wr * dr.d[s];
  }
}

// Per cell: gather the window adjoints of both sweeps, turn the z = W v
// adjoint into the v adjoint, add to lamP (rho, vx, vy, vz, p), to the cons Bz
// adjoint and to the cell-centred B adjoints.
template <int LIM>
__global__ __launch_bounds__(BLK) void k_adj_gather(G2 g, Pr2 P, const double* Wx,
                                                    const double* Wy, Pr2Out lamP, double* lamBz,
                                                    double* lBcx, double* lBcy) {
  constexpr int hw = stencil_halfwidth(LIM);
  constexpr int k0 = 3 - hw, k1 = 4 + hw, KW = k1 - k0 + 1;
  const int64_t t = xcd_block() * BLK + threadIdx.x;
  if (t >= (int64_t)g.nxt * g.nyt) return;
  const int ci = (int)(t / g.nyt), cj = (int)(t - (int64_t)ci * g.nyt);
  double lq[2][7];
#pragma unroll
  for (int N = 0; N < 2; ++N) {
    const int n = N == 0 ? g.nxt : g.nyt;
    const int c = N == 0 ? ci : cj;
    const int64_t nf = N == 0 ? (int64_t)(g.nxt + 1) * g.nyt : (int64_t)g.nxt * (g.nyt + 1);
    const double* Wd = N == 0 ? Wx : Wy;
#pragma unroll
    for (int v = 0; v < 7; ++v) lq[N][v] = 0.0;
    for (int s = 0; s < KW; ++s) {
      const int k = k0 + s;
      // faces f in [0, n] whose slot k reads cell c: clampi(f + k - 4, 0, n-1) == c
      int flo = c + 4 - k, fhi = c + 4 - k;
      if (c == 0) flo = 0;
      if (c == n - 1) fhi = n;
      flo = flo < 0 ? 0 : flo;
      fhi = fhi > n ? n : fhi;
      for (int fa = flo; fa <= fhi; ++fa) {
        if (clampi(fa + k - 4, 0, n - 1) != c) continue;
        const int64_t ft = N == 0 ? (int64_t)fa * g.nyt + cj : (int64_t)ci * (g.nyt + 1) + fa;
#pragma unroll
        for (int v = 0; v < 7; ++v) lq[N][v] += Wd[(int64_t)(v * KW + s) * nf + ft];
      }
    }
  }
  const double vx = P.p[1][t], vy = P.p[2][t], vz = P.p[3][t];
  const double om = 1.0 - (sq(vx) + sq(vy) + sq(vz));
  const double Wl = 1.0 / csqrt(clo(om, 1e-14));
  const double lz[3] = {lq[0][2] + lq[1][2], lq[0][3] + lq[1][3], lq[0][4] + lq[1][4]};
  const double zv = om < 1e-14 ? 0.0 : (lz[0] * vx + lz[1] * vy + lz[2] * vz) * (Wl * Wl * Wl);
  lamP.p[0][t] += lq[0][0] + lq[1][0];
  lamP.p[4][t] += lq[0][1] + lq[1][1];
  lamP.p[1][t] += Wl * lz[0] + zv * vx;
  lamP.p[2][t] += Wl * lz[1] + zv * vy;
  lamP.p[3][t] += Wl * lz[2] + zv * vz;
  lamBz[t] += lq[0][6] + lq[1][5];
  lBcy[t] += lq[0][5];
  lBcx[t] += lq[1][6];
}

// primitive_to_conserved (hlld.py:185-217) transposed, point-wise.
struct P9 {
  const double* p[9];
};
struct P9Out {
  double* p[9];
};
struct C8 {
  const double* p[8];
};
__global__ __launch_bounds__(BLK) void k_adj_p2c(int64_t n, P9 P, C8 lc, P9Out lp) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  double l[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) l[k] = lc.p[k][i];
#pragma unroll 1
  for (int d = 0; d < 9; ++d) {
    PrimT<D1> q;
    q.rho = D1::seed(P.p[0][i], d == 0 ? 0 : -1);
#pragma unroll
    for (int m = 0; m < 3; ++m) q.v[m] = D1::seed(P.p[1 + m][i], d == 1 + m ? 0 : -1);
    q.p = D1::seed(P.p[4][i], d == 4 ? 0 : -1);
    q.eps = D1::seed(P.p[5][i], d == 5 ? 0 : -1);
#pragma unroll
    for (int m = 0; m < 3; ++m) q.B[m] = D1::seed(P.p[6 + m][i], d == 6 + m ? 0 : -1);
    D1 c[8];
    prim_to_cons(q, c);
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += l[k] * c[k].d[0];
    lp.p[d][i] += acc;
  }
}

// ── host side ───────────────────────────────────────────────────────────────

G2 make_g2a(const srmhd_grid2d* g) {
  G2 r;
  r.nx = g->nx;
  r.ny = g->ny;
  r.ng = g->ng;
  r.nxt = g->nx + 2 * g->ng;
  r.nyt = g->ny + 2 * g->ng;
  r.bcx = g->bc_x;
  r.bcy = g->bc_y;
  r.dx = g->dx;
  r.dy = g->dy;
  return r;
}

bool grid_ok_a(const srmhd_grid2d* g, const srmhd_scheme2d* s) {
  if (!g || !s || g->nx < 1 || g->ny < 1 || g->ng < 1) return false;
  if (s->precision != SRMHD_PREC_FP64 || s->emf_mode != SRMHD_EMF_SOLVER || s->solver < 0 ||
      s->solver > 2 || s->limiter < 0 || s->limiter > 5 ||
      (s->flags & ~SRMHD_STAGE_SERIAL_SWEEPS) != 0)
    return false;
  // the reference's default p* root (zeroed options): the one the forward taped
  if (s->hlld.pstar_method != SRMHD_PSTAR_SECANT ||
      s->hlld.pstar_bracket != SRMHD_BRACKET_LEGACY ||
      !(s->hlld.pstar_tol <= 0.0 || s->hlld.pstar_tol == 1e-10))
    return false;
  const int need = s->limiter == SRMHD_PCM ? 0
                   : (s->limiter <= SRMHD_MINMOD ? 1 : (s->limiter == SRMHD_MP7 ? 3 : 2));
  if (g->ng < (need + 1 > 2 ? need + 1 : 2) && s->limiter != SRMHD_PCM) return false;
  for (int bc : {g->bc_x, g->bc_y})
    if (bc != SRMHD_BC_OUTFLOW && bc != SRMHD_BC_PERIODIC) return false;
  if (g->bc_x == SRMHD_BC_PERIODIC && g->nx <= g->ng) return false;
  if (g->bc_y == SRMHD_BC_PERIODIC && g->ny <= g->ng) return false;
  return g->dx > 0.0 && g->dy > 0.0;
}

int window_width(int lim) {
  const int hw = lim == SRMHD_PCM ? 0 : (lim <= SRMHD_MINMOD ? 1 : (lim == SRMHD_MP7 ? 3 : 2));
  return 2 * hw + 2;
}

struct WsA {
  double *lBcx, *lBcy, *lEz, *lBnx, *lBny, *Wx, *Wy, *LQx, *LQy;
  double* LE[8];
  double* lFx[7];
  double* lFy[7];
  int64_t bytes;
};

WsA carve_adj(const G2& g, int lim, void* work) {
  Carver c{(char*)work};
  WsA w;
  const int64_t nfx = (int64_t)(g.nxt + 1) * g.nyt, nfy = (int64_t)g.nxt * (g.nyt + 1);
  const int64_t ncell = (int64_t)g.nxt * g.nyt, ncor = (int64_t)(g.nxt + 1) * (g.nyt + 1);
  const int KW = window_width(lim);
  w.lBcx = c.take<double>(ncell);
  w.lBcy = c.take<double>(ncell);
  w.lEz = c.take<double>(ncor);
  w.lBnx = c.take<double>(nfx);
  w.lBny = c.take<double>(nfy);
  for (int k = 0; k < 6; ++k) w.LE[k] = c.take<double>(ncell);
  w.LE[6] = c.take<double>(nfx);
  w.LE[7] = c.take<double>(nfy);
  for (int k = 0; k < 7; ++k) w.lFx[k] = c.take<double>(nfx);
  for (int k = 0; k < 7; ++k) w.lFy[k] = c.take<double>(nfy);
  w.Wx = c.take<double>(7 * KW * nfx);
  w.Wy = c.take<double>(7 * KW * nfy);
  // the face-state adjoints of each sweep (the two sweeps may run at once)
  w.LQx = c.take<double>(14 * nfx);
  w.LQy = c.take<double>(14 * nfy);
  w.bytes = c.off;
  return w;
}

template <int N, int LIM>
void launch_sweep(int sol, const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e,
                  const double* ps, const CFlx& lF, double* LQ, double* W, double* lBn,
                  hipStream_t s) {
  const int nb = blocks_for(N == 0 ? (int64_t)(g.nxt + 1) * g.nyt : (int64_t)g.nxt * (g.nyt + 1));
  switch (sol) {
    case SOLVER_HLLE: k_adj_flux<N, LIM, SOLVER_HLLE><<<nb, BLK, 0, s>>>(g, S, P, e, ps, lF, LQ, lBn); break;
    case SOLVER_HLLC: k_adj_flux<N, LIM, SOLVER_HLLC><<<nb, BLK, 0, s>>>(g, S, P, e, ps, lF, LQ, lBn); break;
    default: k_adj_flux<N, LIM, SOLVER_HLLD><<<nb, BLK, 0, s>>>(g, S, P, e, ps, lF, LQ, lBn);
  }
  k_adj_recon<N, LIM><<<nb, BLK, 0, s>>>(g, S, P, LQ, W);
}

template <int LIM>
void sweeps_adj(int sol, const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e,
                const srmhd2d_tape* tape, const WsA& w, Pr2Out lamP, double* lamBz,
                hipStream_t s, bool par) {
  CFlx lx, ly;
  for (int k = 0; k < 7; ++k) {
    lx.p[k] = w.lFx[k];
    ly.p[k] = w.lFy[k];
  }
  // par: the y sweep beside the x sweep on the device's second stream (they
  // share no buffer), joined before the gather that reads both
  std::lock_guard<std::mutex> lock(g_aux_mu);
  AuxStream* ax = par ? aux_stream() : nullptr;
  const bool forked = ax && hipEventRecord(ax->fork, s) == hipSuccess &&
                      hipStreamWaitEvent(ax->s, ax->fork, 0) == hipSuccess;
  timing_begin(SRMHD_K_ADJ_SWEEP_X, s);
  launch_sweep<0, LIM>(sol, g, S, P, e, tape->pstar_x, lx, w.LQx, w.Wx, w.lBnx, s);
  timing_end(SRMHD_K_ADJ_SWEEP_X, s);
  launch_sweep<1, LIM>(sol, g, S, P, e, tape->pstar_y, ly, w.LQy, w.Wy, w.lBny,
                       forked ? ax->s : s);
  if (forked) {
    if (hipEventRecord(ax->join, ax->s) == hipSuccess)
      (void)hipStreamWaitEvent(s, ax->join, 0);
    else
      (void)hipStreamSynchronize(ax->s);   // no join event: wait on the host
  }
  k_adj_gather<LIM><<<blocks_for((int64_t)g.nxt * g.nyt), BLK, 0, s>>>(g, P, w.Wx, w.Wy, lamP,
                                                                       lamBz, w.lBcx, w.lBcy);
}

void sweeps_adj_dispatch(int lim, int sol, const G2& g, const St2& S, const Pr2& P,
                         const srmhd_eos& e, const srmhd2d_tape* tape, const WsA& w, Pr2Out lamP,
                         double* lamBz, hipStream_t s, bool par) {
  switch (lim) {
    case LIM_PCM: sweeps_adj<LIM_PCM>(sol, g, S, P, e, tape, w, lamP, lamBz, s, par); break;
    case LIM_MC: sweeps_adj<LIM_MC>(sol, g, S, P, e, tape, w, lamP, lamBz, s, par); break;
    case LIM_MINMOD: sweeps_adj<LIM_MINMOD>(sol, g, S, P, e, tape, w, lamP, lamBz, s, par); break;
    case LIM_WENO5Z: sweeps_adj<LIM_WENO5Z>(sol, g, S, P, e, tape, w, lamP, lamBz, s, par); break;
    case LIM_MP5: sweeps_adj<LIM_MP5>(sol, g, S, P, e, tape, w, lamP, lamBz, s, par); break;
    default: sweeps_adj<LIM_MP7>(sol, g, S, P, e, tape, w, lamP, lamBz, s, par);
  }
}

// con2prim + BC part shared by the stage and sync adjoints: lamU (8, the
// adjoint of the synced state) += the con2prim adjoint, then folded into out.
void c2p_fold_adj(const G2& g, const srmhd_scheme2d* sp, const srmhd_eos& e, const St2& U,
                  const Pr2& lamP, const St2Out& lamU, const WsA& w, const St2Out& out,
                  hipStream_t s) {
  const int64_t ncell = (int64_t)g.nxt * g.nyt, ncor = (int64_t)(g.nxt + 1) * (g.nyt + 1);
  k_zero<<<blocks_for(ncell) < 1024 ? blocks_for(ncell) : 1024, BLK, 0, s>>>(w.lBcx, ncell);
  k_zero<<<blocks_for(ncell) < 1024 ? blocks_for(ncell) : 1024, BLK, 0, s>>>(w.lBcy, ncell);
  k_adj_c2p<<<blocks_for((int64_t)g.nx * g.ny), BLK, 0, s>>>(g, U, e, sp->atmo_rho, sp->atmo_eps,
                                                             lamP, lamU, w.lBcx, w.lBcy);
  k_adj_bc2face<<<blocks_for(ncor), BLK, 0, s>>>(g, w.lBcx, w.lBcy, nullptr, nullptr, lamU.p[6],
                                                 lamU.p[7]);
  St2 lu;
  for (int k = 0; k < 8; ++k) lu.p[k] = lamU.p[k];
  k_adj_fold<<<blocks_for(ncor), BLK, 0, s>>>(g, lu, out);
}

}  // namespace

extern "C" {

int64_t srmhd2d_adjoint_workspace_bytes(const srmhd_grid2d* g, const srmhd_scheme2d* s) {
  if (!grid_ok_a(g, s)) return -1;
  return carve_adj(make_g2a(g), s->limiter, nullptr).bytes;
}

int srmhd2d_stage_adjoint(const srmhd_grid2d* gp, const srmhd_scheme2d* sp, const srmhd_eos* eos,
                          double a, double b, double c, const double* dt_dev,
                          const double* const* U, const double* const* P,
                          const double* const* Uout, const srmhd2d_tape* tape,
                          double* const* lamUout, double* const* lamPout,
                          double* const* lamU0, double* const* lamU, double* const* lamP,
                          void* work, int64_t work_bytes, void* stream) {
  if (!grid_ok_a(gp, sp) || !eos_ok(eos) || !dt_dev || !U || !P || !Uout || !tape || !lamUout ||
      !lamPout || !lamU0 || !lamU || !lamP || !work)
    return SRMHD_EARG;
  if (!tape->pstar_x || !tape->pstar_y || !tape->FD_x || !tape->FD_y || !tape->thresh)
    return SRMHD_EARG;
  const G2 g = make_g2a(gp);
  const WsA w = carve_adj(g, sp->limiter, work);
  if (work_bytes < w.bytes) return SRMHD_EARG;
  St2 su, suo, le;
  St2Out lo, l0, l1, leo;
  Pr2 sp6, lpo;
  Pr2Out lp;
  for (int k = 0; k < 8; ++k) {
    su.p[k] = U[k];
    suo.p[k] = Uout[k];
    lo.p[k] = lamUout[k];
    l0.p[k] = lamU0[k];
    l1.p[k] = lamU[k];
    le.p[k] = w.LE[k];
    leo.p[k] = w.LE[k];
    if (!su.p[k] || !suo.p[k] || !lo.p[k] || !l0.p[k] || !l1.p[k]) return SRMHD_EARG;
  }
  for (int k = 0; k < 6; ++k) {
    sp6.p[k] = P[k];
    lpo.p[k] = lamPout[k];
    lp.p[k] = lamP[k];
    if (!sp6.p[k] || !lpo.p[k] || !lp.p[k]) return SRMHD_EARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t ncor = (int64_t)(g.nxt + 1) * (g.nyt + 1);
  const int64_t nfx = (int64_t)(g.nxt + 1) * g.nyt, nfy = (int64_t)g.nxt * (g.nyt + 1);
  // con2prim of the output, its BCs: the adjoint of the combined state
  c2p_fold_adj(g, sp, *eos, suo, lpo, lo, w, leo, s);
  // RK combination
  k_adj_combine<<<blocks_for(ncor), BLK, 0, s>>>(g, le, a, b, l0, l1);
  // ct_rhs, corner EMF, flux divergence
  k_adj_ez<<<blocks_for(ncor), BLK, 0, s>>>(g, le, dt_dev, c, w.lEz);
  Flx fx, fy;
  for (int k = 0; k < 7; ++k) {
    fx.p[k] = w.lFx[k];
    fy.p[k] = w.lFy[k];
  }
  k_adj_fx<<<blocks_for(nfx), BLK, 0, s>>>(g, le, w.lEz, tape->FD_x, tape->FD_y, tape->thresh,
                                           sp->upwind, dt_dev, c, fx);
  k_adj_fy<<<blocks_for(nfy), BLK, 0, s>>>(g, le, w.lEz, tape->FD_x, tape->FD_y, tape->thresh,
                                           sp->upwind, dt_dev, c, fy);
  k_adj_ec<<<blocks_for((int64_t)g.nxt * g.nyt), BLK, 0, s>>>(
      g, su, sp6, w.lEz, tape->FD_x, tape->FD_y, tape->thresh, sp->upwind, lp, w.lBcx, w.lBcy);
  // the sweeps; a launch error so far is reported first (creating the second
  // stream they use may clear the error state when it fails)
  if (const int rc = launch_status()) return rc;
  sweeps_adj_dispatch(sp->limiter, sp->solver, g, su, sp6, *eos, tape, w, lp, lamU[5], s,
                      !(sp->flags & SRMHD_STAGE_SERIAL_SWEEPS));
  k_adj_bc2face<<<blocks_for(ncor), BLK, 0, s>>>(g, w.lBcx, w.lBcy, w.lBnx, w.lBny, lamU[6],
                                                 lamU[7]);
  return launch_status();
}

int srmhd2d_sync_adjoint(const srmhd_grid2d* gp, const srmhd_scheme2d* sp, const srmhd_eos* eos,
                         const double* const* U, double* const* lamU, double* const* lamP,
                         double* const* lamUin, void* work, int64_t work_bytes, void* stream) {
  if (!grid_ok_a(gp, sp) || !eos_ok(eos) || !U || !lamU || !lamP || !lamUin || !work)
    return SRMHD_EARG;
  const G2 g = make_g2a(gp);
  const WsA w = carve_adj(g, sp->limiter, work);
  if (work_bytes < w.bytes) return SRMHD_EARG;
  St2 su;
  St2Out lu, li;
  Pr2 lp;
  for (int k = 0; k < 8; ++k) {
    su.p[k] = U[k];
    lu.p[k] = lamU[k];
    li.p[k] = lamUin[k];
    if (!su.p[k] || !lu.p[k] || !li.p[k]) return SRMHD_EARG;
  }
  for (int k = 0; k < 6; ++k)
    if (!(lp.p[k] = lamP[k])) return SRMHD_EARG;
  c2p_fold_adj(g, sp, *eos, su, lp, lu, w, li, (hipStream_t)stream);
  return launch_status();
}

int srmhd_p2c_adjoint(int64_t n, const double* const* prim, const double* const* lam_cons,
                      double* const* lam_prim, void* stream) {
  if (n < 0 || !prim || !lam_cons || !lam_prim) return SRMHD_EARG;
  if (n == 0) return SRMHD_OK;
  P9 P;
  C8 lc;
  P9Out lp;
  for (int k = 0; k < 9; ++k)
    if (!(P.p[k] = prim[k]) || !(lp.p[k] = lam_prim[k])) return SRMHD_EARG;
  for (int k = 0; k < 8; ++k)
    if (!(lc.p[k] = lam_cons[k])) return SRMHD_EARG;
  k_adj_p2c<<<blocks_for(n), BLK, 0, (hipStream_t)stream>>>(n, P, lc, lp);
  return launch_status();
}

}  // extern "C"
