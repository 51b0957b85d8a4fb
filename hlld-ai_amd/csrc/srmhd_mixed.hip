// Mixed-precision sweeps of the 2D stage.  Form MIXED: reconstruction and
// the Riemann solver in FP32 (HLLD with its HLL-state con2prim per face).
// Form MIXED_R64: reconstruction in FP64, HLLD's p* root found in FP32 and
// polished, and its flux assembled, in FP64.  Everything else of the stage
// (corner EMF, CT face-B update, RK accumulation, con2prim) is FP64.
//
// One kernel per sweep (no pass A: the FP32 HLL-state con2prim is per
// element), writing the same seven FP64 flux arrays as the FP64 sweeps, so
// the downstream kernels are shared.  MIXED reconstructs from FP32 copies of
// each cell's reconstructed variables (rho, p, W v, cell-centred B), formed
// once per stage by k_cells_f32; MIXED_R64 reads the FP64 primitives.
// Built with FP32 contraction and approximate FP32 division / sqrt (this
// translation unit only).
#include "srmhd_mixed.h"

#include "srmhd_c2p.h"
#include "srmhd_f32.h"
#include "srmhd_grid2d.h"
#include "srmhd_recon.h"

namespace srmhd {
namespace {

// the reconstructed variables of every cell in FP32 (the W v and cell-centred
// B of reconstruction.py:272-300 and sweep.py:82-89, formed in FP64 and
// rounded once)
__global__ __launch_bounds__(BLK) void k_cells_f32(G2 g, St2 S, Pr2 P, Cells32 C) {
  const int64_t c = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (c >= (int64_t)g.nxt * g.nyt) return;
  const int i = (int)(c / g.nyt), j = (int)(c - (int64_t)i * g.nyt);
  const double vx = P.p[1][c], vy = P.p[2][c], vz = P.p[3][c];
  const double W = 1.0 / __builtin_sqrt(fmax(1.0 - (vx * vx + vy * vy + vz * vz), 1e-14));
  C.p[0][c] = (float)P.p[0][c];
  C.p[1][c] = (float)P.p[4][c];
  C.p[2][c] = (float)(W * vx);
  C.p[3][c] = (float)(W * vy);
  C.p[4][c] = (float)(W * vz);
  C.p[5][c] = (float)bcx(g, S.p[6], i, j);
  C.p[6][c] = (float)bcy(g, S.p[7], i, j);
  C.p[7][c] = (float)S.p[5][c];
}

// face_prims2d (reconstruction.py:272-338 with vel_var "Wv", then B_n := face
// value, sweep.py:82-89) in FP32 from the cell arrays C: rho, p, W v and the
// two tangential B components reconstructed by the limiter, the PCM fallback
// where a reconstructed rho or p is not positive.
template <int N, int LIM>
__device__ void face_prims_c32(const G2& g, const St2& S, const Cells32& C, const f32::Eos& e,
                               int i, int j, f32::Prim& L, f32::Prim& R) {
  constexpr int w = stencil_halfwidth(LIM);
  constexpr int k0 = 3 - w, k1 = 4 + w;
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  // C arrays of the recon variables rho, p, zx, zy, zz, B_t1, B_t2
  constexpr int src[7] = {0, 1, 2, 3, 4, N == 0 ? 6 : 7, N == 0 ? 7 : 5};
  const int n = N == 0 ? g.nxt : g.nyt;
  const int f = N == 0 ? i : j;
  const int64_t stride = N == 0 ? g.nyt : 1;
  const int64_t base = N == 0 ? (int64_t)j : (int64_t)i * g.nyt;
  int64_t idx[8];
#pragma unroll
  for (int k = k0; k <= k1; ++k) idx[k] = base + (int64_t)clampi(f + k - 4, 0, n - 1) * stride;
  float ql[7], qr[7], q[8];
#pragma unroll
  for (int v = 0; v < 2; ++v) {
#pragma unroll
    for (int k = k0; k <= k1; ++k) q[k] = C.p[src[v]][idx[k]];
    f32::face_states<LIM>(q, f, n, ql[v], qr[v]);
  }
  const bool bad = (ql[0] <= 1e-12f) || (qr[0] <= 1e-12f) || (ql[1] <= 1e-14f) || (qr[1] <= 1e-14f);
  if (bad) {
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      ql[v] = C.p[src[v]][idx[f == 0 ? 4 : 3]];
      qr[v] = C.p[src[v]][idx[f == n ? 3 : 4]];
    }
  }
#pragma unroll
  for (int v = 2; v < 7; ++v) {
#pragma unroll
    for (int k = k0; k <= k1; ++k) q[k] = C.p[src[v]][idx[k]];
    if (bad) {
      ql[v] = f == 0 ? q[4] : q[3];
      qr[v] = f == n ? q[3] : q[4];
    } else {
      f32::face_states<LIM>(q, f, n, ql[v], qr[v]);
    }
  }
  const float Bn = (float)(N == 0 ? S.p[6][(int64_t)i * g.nyt + j]
                                  : S.p[7][(int64_t)i * (g.nyt + 1) + j]);
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const float* x = s ? qr : ql;
    f32::Prim& Q = s ? R : L;
    const float fac = f32::frsqrt(1.f + (x[2] * x[2] + x[3] * x[3] + x[4] * x[4]));
    Q.v[0] = x[2] * fac;
    Q.v[1] = x[3] * fac;
    Q.v[2] = x[4] * fac;
    Q.rho = fmaxf(x[0], 1e-12f);
    Q.p = fmaxf(x[1], 1e-14f);
    Q.eps = e.eps_of(Q.p, Q.rho);
    Q.B[N] = Bn;
    Q.B[T1] = x[5];
    Q.B[T2] = x[6];
  }
}

// face_prims2d in FP64 (the FP64 path's reconstruction) for MIXED_R64: the
// face states in FP64 and their FP32 copies.
template <int N, int LIM>
__device__ void face_prims_r64(const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e64,
                               const f32::Eos& e, int i, int j, f32::Prim& L, f32::Prim& R,
                               Prim& L64, Prim& R64) {
  face_prims2d<N, LIM>(g, S, P, e64, i, j, L64, R64);
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const Prim& Q64 = s ? R64 : L64;
    f32::Prim& Q = s ? R : L;
    Q.v[0] = (float)Q64.v[0];
    Q.v[1] = (float)Q64.v[1];
    Q.v[2] = (float)Q64.v[2];
    Q.rho = (float)Q64.rho;
    Q.p = (float)Q64.p;
    Q.eps = (float)Q64.eps;
    Q.B[0] = (float)Q64.B[0];
    Q.B[1] = (float)Q64.B[1];
    Q.B[2] = (float)Q64.B[2];
  }
}

// HLLD of the mixed form: the p* root is found in FP32 (the HLL-state
// con2prim that seeds the bracket, the bracket and the secant: the iterative,
// division-heavy part), then polished by FP64 secant steps on the FP64
// residual from the FP64 face states, and the star states and the flux are
// assembled in FP64 (hlld_assemble, the FP64 path's own code), so the flux
// carries FP64 rounding.  Faces whose FP32 root does not converge, or whose
// polish does not reach the FP64 tolerance within 8 steps, are flagged for
// the full FP64 recomputation.
// distance below which the FP64 wave-order test of a polished face is
// treated as undecided (p* is polished to the FP64 tolerance, 1e-10 in the
// residual; the flux assembly here is contracted, the FP64 path's is not)
constexpr double WAVE_MARGIN64 = 1e-7;

template <int N>
__device__ int hlld_mixed(const f32::Prim& l, const f32::Prim& r, const f32::Eos& e,
                          const Prim& L64, const Prim& R64, const srmhd_eos& e64, double F[8]) {
  float ps32, p_hll32;
  const bool conv32 = f32::hlld_pstar<N>(l, r, e, ps32, p_hll32);
  HlldPrep<N> h;
  h.run(L64, R64, e64);
  StarR R;
  hlld_rvec<N>(h, R);
  double fH[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) fH[k] = hll_flux(h.fL[k], h.fR[k], h.uL[k], h.uR[k], h.cmin, h.cmax);
  double x0 = (double)ps32 * (1.0 - 1e-6), x1 = (double)ps32;
  double f0 = star_residual(R, x0), f1 = star_residual(R, x1);
  bool conv = fabs(f1) < PSTAR_TOL;
#pragma unroll 1
  for (int it = 0; it < 8 && !conv; ++it) {
    const double df = f1 - f0;
    const double dx = (fabs(df) > 1e-30 * fabs(f1)) ? f1 * (x1 - x0) / df : 0.0;
    const double x2 = fmax(x1 - dx, 1e-30);
    const double f2 = star_residual(R, x2);
    x0 = x1;
    f0 = f1;
    x1 = x2;
    f1 = f2;
    conv = fabs(f1) < PSTAR_TOL;
  }
  double ps, margin[2];
  const int st = hlld_assemble<N>(h, R, fH, x1, conv32 && conv, (double)p_hll32, F,
                                  (double*)nullptr, ps, margin);
  // The wave-order test is decided by round-off where an Alfven wave meets
  // its fast wave or the contact (e.g. a vanishing tangential field); such
  // faces take the FP64 path's own arithmetic, so the HLLE-fallback switch
  // is decided as the FP64 form decides it.
  const bool bn_zero = st & FS_BN_ZERO;
  const bool strong_bn = fmin(fabs(h.uL[CB + N]), fabs(h.uR[CB + N])) > f32::STRONG_BN;
  const bool doubt = !bn_zero && (!(conv32 && conv) || !(margin[0] > WAVE_MARGIN64) ||
                                  (strong_bn && !(margin[1] > WAVE_MARGIN64)));
  return doubt ? (st | f32::FS_DOUBT) : st;
}

// fix: count + indices of the faces to recompute in FP64 (fix[0] = count,
// indices from fix[2]); capacity >= the number of faces.
// The face EMF of emf_mode "hll" in FP32: the HLL average of the
// tangential-field flux from the FP32 side states (hll_emf_flux's FP32 form).
template <int N>
__device__ float hll_emf_flux32(const f32::Prim& L, const f32::Prim& R, const f32::Eos& e) {
  float uL[8], uR[8], fL[8], fR[8], cmax, cmin;
  f32::srmhd_fluxes<N>(L, R, e, uL, uR, fL, fR, cmax, cmin);
  constexpr int c = N == 0 ? CB + 1 : CB + 0;
  return hll_flux(fL[c], fR[c], uL[c], uR[c], cmin, cmax);
}

// at least two waves per SIMD: the MIXED_R64 instance otherwise spills two
// registers into AGPRs and drops to one wave
template <int N, int LIM, int SOL, bool R64>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(2))) void k_sweep_f32(
    G2 g, St2 S, Pr2 P, Cells32 C, srmhd_eos e64, srmhd_flux_diag* d, Flx F, unsigned* fix,
    int emf_hll) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const bool active = t < n_faces<N>(g);
  int st = 0;
  double fD = 0.0;
  if (active) {
    int i, j;
    face_ij<N>(g, t, i, j);
    const f32::Eos e(e64);
    f32::Prim l, r;
    Prim l64, r64;
    if (R64)
      face_prims_r64<N, LIM>(g, S, P, e64, e, i, j, l, r, l64, r64);
    else
      face_prims_c32<N, LIM>(g, S, C, e, i, j, l, r);
    // the HLL face EMF first, while only the face states are live
    double emf_h = 0.0;
    if (emf_hll) emf_h = R64 ? hll_emf_flux<N>(l64, r64, e64) : (double)hll_emf_flux32<N>(l, r, e);
    double f[8];
    if (SOL == SOLVER_HLLD && R64) {
      st = hlld_mixed<N>(l, r, e, l64, r64, e64, f);
      if (st & f32::FS_DOUBT) fix[2 + atomicAdd(fix, 1u)] = (unsigned)t;
    } else if (SOL == SOLVER_HLLD) {
      float f3[8];
      st = f32::hlld<N>(l, r, e, f3);
      if (st & f32::FS_DOUBT) fix[2 + atomicAdd(fix, 1u)] = (unsigned)t;
#pragma unroll
      for (int k = 0; k < 8; ++k) f[k] = (double)f3[k];
    } else {
      float f3[8];
      if (SOL == SOLVER_HLLE) {
        f32::hlle<N>(l, r, e, f3);
      } else {
        f32::hllc<N>(l, r, e, f3);
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) f[k] = (double)f3[k];
    }
    F.p[0][t] = f[CD];
    F.p[1][t] = f[CS];
    F.p[2][t] = f[CS + 1];
    F.p[3][t] = f[CS + 2];
    F.p[4][t] = f[CT];
    F.p[5][t] = f[CB + 2];
    F.p[6][t] = emf_hll ? emf_h : (N == 0 ? f[CB + 1] : f[CB + 0]);
    fD = f[CD];
  }
  diag_accumulate(diag_shard(d), active, st, 0, dbits(fabs(fD)));
}

template <int N, int LIM, bool R64>
void launch_lim(int sol, const G2& g, const St2& S, const Pr2& P, const Cells32& C,
                const srmhd_eos& e, srmhd_flux_diag* d, const Flx& F, unsigned* fix,
                int emf_hll, hipStream_t s) {
  const int nb = blocks_for(n_faces_host(g, N));
  switch (sol) {
    case SOLVER_HLLE:
      k_sweep_f32<N, LIM, SOLVER_HLLE, R64><<<nb, BLK, 0, s>>>(g, S, P, C, e, d, F, fix, emf_hll);
      break;
    case SOLVER_HLLC:
      k_sweep_f32<N, LIM, SOLVER_HLLC, R64><<<nb, BLK, 0, s>>>(g, S, P, C, e, d, F, fix, emf_hll);
      break;
    default: k_sweep_f32<N, LIM, SOLVER_HLLD, R64><<<nb, BLK, 0, s>>>(g, S, P, C, e, d, F, fix, emf_hll);
  }
}

template <int N, bool R64>
void launch_dir(int lim, int sol, const G2& g, const St2& S, const Pr2& P, const Cells32& C,
                const srmhd_eos& e, srmhd_flux_diag* d, const Flx& F, unsigned* fix,
                int emf_hll, hipStream_t s) {
  switch (lim) {
    case LIM_PCM: launch_lim<N, LIM_PCM, R64>(sol, g, S, P, C, e, d, F, fix, emf_hll, s); break;
    case LIM_MC: launch_lim<N, LIM_MC, R64>(sol, g, S, P, C, e, d, F, fix, emf_hll, s); break;
    case LIM_MINMOD: launch_lim<N, LIM_MINMOD, R64>(sol, g, S, P, C, e, d, F, fix, emf_hll, s); break;
    case LIM_WENO5Z: launch_lim<N, LIM_WENO5Z, R64>(sol, g, S, P, C, e, d, F, fix, emf_hll, s); break;
    case LIM_MP5: launch_lim<N, LIM_MP5, R64>(sol, g, S, P, C, e, d, F, fix, emf_hll, s); break;
    default: launch_lim<N, LIM_MP7, R64>(sol, g, S, P, C, e, d, F, fix, emf_hll, s);
  }
}

// con2prim of the physical cells, each cell its own batch, with mu_+ from
// the FP32 bisection: mu_+ only closes the secant's bracket, which any value
// at or above it does.  The bisection's upper end lies above mu_+ (within
// 2^-14) except where FP32 rounding flips a decision next to the root (an
// error of ~2e-7); inflated by 4e-6 relative plus 4e-7 it stands in for the
// 50 FP64 bisection steps, and the FP64 secant converges to the same root,
// |f| < 1e-15.
__global__ __launch_bounds__(BLK) void k_c2p2d_mixed(G2 g, St2 S, srmhd_eos e, double atmo_rho,
                                                     double atmo_eps, Pr2Out P,
                                                     unsigned long long* redo_count) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  bool redo = false;
  if (t < (int64_t)g.nx * g.ny) {
    const int pi = (int)(t / g.ny), pj = (int)(t - (int64_t)pi * g.ny);
    const int i = pi + g.ng, j = pj + g.ng;
    double u[8];
    cell_cons(g, S, i, j, u);
    float u32[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) u32[k] = (float)u[k];
    const f32::Kastaun k32(u32, f32::Eos(e));
    const double mup = fmin((double)k32.mu_plus() * (1.0 + 4e-6) + 4e-7, 1.0 / (1.0 + TINY));
    Kastaun k;
    k.init(u, e);
    Prim q;
    const int st = c2p_single(k, atmo_rho, atmo_eps, q, nullptr, mup);
    redo = st & C2P_REDONE;
    const int64_t c = (int64_t)i * g.nyt + j;
    P.p[0][c] = q.rho;
    P.p[1][c] = q.v[0];
    P.p[2][c] = q.v[1];
    P.p[3][c] = q.v[2];
    P.p[4][c] = q.p;
    P.p[5][c] = q.eps;
  }
  const unsigned long long nr = __popcll(__ballot(redo));
  if (lane0() && nr) atomicAdd(redo_count, nr);
}

}  // namespace

void c2p_cells_mixed(const G2& g, const St2& S, const srmhd_eos& e, double atmo_rho,
                     double atmo_eps, const Pr2Out& P, unsigned long long* redo_count,
                     hipStream_t s) {
  k_c2p2d_mixed<<<blocks_for((int64_t)g.nx * g.ny), BLK, 0, s>>>(g, S, e, atmo_rho, atmo_eps, P,
                                                                  redo_count);
}

void cells_f32(const G2& g, const St2& S, const Pr2& P, const Cells32& C, hipStream_t s) {
  k_cells_f32<<<blocks_for((int64_t)g.nxt * g.nyt), BLK, 0, s>>>(g, S, P, C);
}

void sweep_mixed(int dir, int lim, int sol, bool r64, const G2& g, const St2& S, const Pr2& P,
                 const Cells32& C, const srmhd_eos& e, srmhd_flux_diag* d, const Flx& F,
                 unsigned* fix, int emf_hll, hipStream_t s) {
  if (dir == 0) {
    if (r64) launch_dir<0, true>(lim, sol, g, S, P, C, e, d, F, fix, emf_hll, s);
    else launch_dir<0, false>(lim, sol, g, S, P, C, e, d, F, fix, emf_hll, s);
  } else {
    if (r64) launch_dir<1, true>(lim, sol, g, S, P, C, e, d, F, fix, emf_hll, s);
    else launch_dir<1, false>(lim, sol, g, S, P, C, e, d, F, fix, emf_hll, s);
  }
}

}  // namespace srmhd
