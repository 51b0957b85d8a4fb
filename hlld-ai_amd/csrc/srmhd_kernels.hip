// Kernels and C-ABI entry points of the gfx950 SRMHD step (see include/srmhd.h).
//
// Launch shape: one thread per face / cell, 256-thread blocks (4 wave64s).
// Reductions are wave-level (shuffles, ballots), then block-level, then one
// atomic per block on 64-bit words (the 2D sweeps' counters into one of 64
// line-sized shards, folded after the sweep); every reduced quantity is an
// integer, a bit mask or a max, so results are deterministic.
#include <mutex>

#include "srmhd_grid2d.h"
#include "srmhd_mixed.h"
#include "srmhd_linear_launch.h"
#include "srmhd_pstar_launch.h"
#include "srmhd_recon.h"
#include "srmhd_riemann.h"

using namespace srmhd;

namespace {

struct PrimPtrs {
  const double* p[9];
};
struct ConsPtrs {
  double* p[8];
};

__device__ inline void load_prim(const PrimPtrs& s, int64_t i, Prim& P) {
  P.rho = s.p[0][i];
  P.v[0] = s.p[1][i];
  P.v[1] = s.p[2][i];
  P.v[2] = s.p[3][i];
  P.p = s.p[4][i];
  P.eps = s.p[5][i];
  P.B[0] = s.p[6][i];
  P.B[1] = s.p[7][i];
  P.B[2] = s.p[8][i];
}

__global__ void k_zero_u64(unsigned long long* w, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = 0ull;
}

// up to five regions cleared by one launch (a grid covering the longest);
// a null region is skipped
struct ZeroSet {
  unsigned long long* p[5];
  int n[5];
};
__global__ void k_zero_set(ZeroSet z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 5; ++k)
    if (z.p[k] && i < z.n[k]) z.p[k][i] = 0ull;
}

// ── flat flux (flux_fn) ─────────────────────────────────────────────────────

template <int N>
__global__ __launch_bounds__(BLK) void k_flux_a(int64_t n, PrimPtrs L, PrimPtrs R, srmhd_eos e,
                                                C2PRec* rec, srmhd_flux_diag* d) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  uint64_t m = 0;
  if (i < n) {
    Prim l, r;
    load_prim(L, i, l);
    load_prim(R, i, r);
    HlldPrep<N> h;
    h.run(l, r, e);
    Kastaun k;
    k.init(h.uH, e);
    m = c2p_pass_a(k, rec[i]);
  }
  block_or_atomic(&d->c2p_mask, m);
}

template <int N, int S>
__global__ __launch_bounds__(BLK) void k_flux_b(int64_t n, PrimPtrs L, PrimPtrs R, srmhd_eos e,
                                                const C2PRec* rec, srmhd_flux_diag* d,
                                                ConsPtrs F, ConsPtrs U, int want_u,
                                                double* pstar, uint8_t* status, double tol) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const bool active = i < n;
  int st = 0, iters = 0;
  if (active) {
    Prim l, r;
    load_prim(L, i, l);
    load_prim(R, i, r);
    double f[8], u[8], ps;
    if (S == SOLVER_HLLE) {
      hlle<N>(l, r, e, f, u, ps);
    } else if (S == SOLVER_HLLC) {
      hllc<N>(l, r, e, f, u, ps);
    } else {
      HlldPrep<N> h;
      h.run(l, r, e);
      Kastaun k;
      k.init(h.uH, e);
      Prim ph;
      const int cst = c2p_pass_b(k, rec[i], c2p_break_iter(d->c2p_mask), 1e-10, 1e-10, ph);
      st = hlld_finish<N>(h, l, r, p_hll_from_prim(ph, e), f, u, ps, iters, tol);
      if (cst & C2P_REDONE) st |= FS_C2P_REDONE;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) F.p[k][i] = f[k];
    if (want_u) {
#pragma unroll
      for (int k = 0; k < 8; ++k) U.p[k][i] = u[k];
    }
    if (pstar) pstar[i] = ps;
    if (status) status[i] = (uint8_t)st;
  }
  if (S == SOLVER_HLLD) diag_accumulate(d, active, st, iters, 0ull);
}

// ── flat con2prim / prim2con ─────────────────────────────────────────────────

struct CConsPtrs {
  const double* p[8];
};

__device__ inline void load_cons(const CConsPtrs& c, int64_t i, double u[8]) {
#pragma unroll
  for (int k = 0; k < 8; ++k) u[k] = c.p[k][i];
}

__global__ __launch_bounds__(BLK) void k_c2p_a(int64_t n, CConsPtrs C, srmhd_eos e, C2PRec* rec,
                                               unsigned long long* mask) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  uint64_t m = 0;
  if (i < n) {
    double u[8];
    load_cons(C, i, u);
    Kastaun k;
    k.init(u, e);
    m = c2p_pass_a(k, rec[i]);
  }
  block_or_atomic(mask, m);
}

struct OutPrimPtrs {
  double* p[9];
};

__global__ __launch_bounds__(BLK) void k_c2p_b(int64_t n, CConsPtrs C, srmhd_eos e,
                                               const C2PRec* rec, const unsigned long long* mask,
                                               double atmo_rho, double atmo_eps, OutPrimPtrs P,
                                               uint8_t* status) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  double u[8];
  load_cons(C, i, u);
  Kastaun k;
  k.init(u, e);
  Prim q;
  const int st = c2p_pass_b(k, rec[i], c2p_break_iter(*mask), atmo_rho, atmo_eps, q);
  P.p[0][i] = q.rho;
  P.p[1][i] = q.v[0];
  P.p[2][i] = q.v[1];
  P.p[3][i] = q.v[2];
  P.p[4][i] = q.p;
  P.p[5][i] = q.eps;
  P.p[6][i] = q.B[0];
  P.p[7][i] = q.B[1];
  P.p[8][i] = q.B[2];
  if (status) status[i] = (uint8_t)st;
}

__global__ __launch_bounds__(BLK) void k_p2c(int64_t n, PrimPtrs P, ConsPtrs C) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  Prim q;
  load_prim(P, i, q);
  double c[8];
  prim_to_cons(q, c);
#pragma unroll
  for (int k = 0; k < 8; ++k) C.p[k][i] = c[k];
}


// Pass A of the sweep's HLL-state con2prims.  DEFER: faces whose con2prim
// is inert (hll_c2p_inert: B_n = 0 on both sides, so the flux is the HLLE
// one) are left out; their result matters only when the other faces'
// masks leave the batch's break iteration below 60, and k_sweep_a_rescue
// then runs them.  On the rotor ~95 % of the y faces are such faces.
// hq (non-NULL): the side states are kept there for pass B (side_q_store).
template <int N, int LIM, bool DEFER = false>
__global__ __launch_bounds__(BLK) void k_sweep_a(G2 g, St2 S, Pr2 P, srmhd_eos e, C2PRec* rec,
                                                 srmhd_flux_diag* dsh, double* hq = nullptr) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  uint64_t m = 0;
  if (t < n_faces<N>(g)) {
    int i, j;
    face_ij<N>(g, t, i, j);
    Prim l, r;
    face_prims2d<N, LIM>(g, S, P, e, i, j, l, r);
    HlldPrep<N> h;
    h.run(l, r, e);
    if (hq) side_q_store<N>(hq, n_faces<N>(g), t, h);
    if (!(DEFER && hll_c2p_inert<N>(h))) {
      Kastaun k;
      k.init(h.uH, e);
      m = c2p_pass_a(k, rec[t]);
    }
  }
  block_or_atomic(&diag_shard(dsh)->c2p_mask, m);
}

SRMHD_DEV bool c2p_mask_full(uint64_t m) { return (m & C2P_ALL) == C2P_ALL; }

// Pass A of the faces k_sweep_a<N, LIM, true> left out, when the other
// faces' mask d->c2p_mask leaves the break iteration below 60 (otherwise
// every workgroup returns at once); their mask bits go to *mres.
template <int N, int LIM>
// hq (non-NULL): the side states pass A kept, instead of forming them again.
__global__ __launch_bounds__(BLK) void k_sweep_a_rescue(G2 g, St2 S, Pr2 P, srmhd_eos e,
                                                        C2PRec* rec, const srmhd_flux_diag* d,
                                                        unsigned long long* mres,
                                                        const double* hq = nullptr) {
  if (c2p_mask_full(d->c2p_mask)) return;
  const int64_t nf = n_faces<N>(g);
  uint64_t m = 0;
  for (int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x; t < nf;
       t += (int64_t)gridDim.x * BLK) {
    HlldPrep<N> h;
    if (hq) {
      side_q_load<N>(hq, nf, t, h);
    } else {
      int i, j;
      face_ij<N>(g, t, i, j);
      Prim l, r;
      face_prims2d<N, LIM>(g, S, P, e, i, j, l, r);
      h.run(l, r, e);
    }
    if (hll_c2p_inert<N>(h)) {
      Kastaun k;
      k.init(h.uH, e);
      m |= c2p_pass_a(k, rec[t]);
    }
  }
  block_or_atomic(mres, m);
}

// LOCAL: the con2prim of each face's HLL state is solved as its own batch
// (c2p_single) instead of finishing pass A's batch-coupled iteration.
// mres (non-NULL after k_sweep_a<N, LIM, true>): the mask of the deferred
// faces.  When the other faces' mask alone is full, the deferred faces have
// no pass-A record and skip their con2prim: p_hll then reaches only the
// fallback sentinel of p* (written as -1 - 1e-10 there: negative, as every
// fallback's), and the rest of the solver is unchanged.
// HQ: the side states come from pass A's copy hq (side_q_load) instead of
// the reconstruction and HlldPrep::run; the HLL face EMF from them too.
template <int N, int LIM, int SOL, bool LOCAL = false, bool HQ = false>
__global__ __launch_bounds__(BLK) void k_sweep_b(G2 g, St2 S, Pr2 P, srmhd_eos e,
                                                 const C2PRec* rec, const srmhd_flux_diag* d,
                                                 srmhd_flux_diag* dsh, Flx F, double* pstar_out,
                                                 int emf_hll, double tol,
                                                 const unsigned long long* mres = nullptr,
                                                 const double* hq = nullptr) {
  static_assert(!LOCAL || SOL == SOLVER_HLLD, "the local con2prim concerns HLLD only");
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const bool active = t < n_faces<N>(g);
  int st = 0, iters = 0;
  double fD = 0.0;
  if (active) {
    int i, j;
    face_ij<N>(g, t, i, j);
    Prim l{}, r{};   // (HQ: not formed; the solver reads them not)
    HlldPrep<N> h;
    double emf_h = 0.0;
    if (HQ) {
      static_assert(!HQ || SOL == SOLVER_HLLD, "the side states come from HLLD's pass A");
      side_q_load<N>(hq, n_faces<N>(g), t, h);
      // hll_emf_flux's arithmetic on the same side states
      constexpr int c = N == 0 ? CB + 1 : CB + 0;
      if (emf_hll) emf_h = hll_flux(h.fL[c], h.fR[c], h.uL[c], h.uR[c], h.cmin, h.cmax);
    } else {
      face_prims2d<N, LIM>(g, S, P, e, i, j, l, r);
      // the HLL face EMF first, while only the face states are live
      emf_h = emf_hll ? hll_emf_flux<N>(l, r, e) : 0.0;
    }
    double f[8], u[8], ps;
    if (SOL == SOLVER_HLLE) {
      hlle<N>(l, r, e, f, u, ps);
    } else if (SOL == SOLVER_HLLC) {
      hllc<N>(l, r, e, f, u, ps);
    } else {
      if (!HQ) h.run(l, r, e);
      double p_hll = 1.0;
      int cst = 0;
      if (!(!LOCAL && mres && c2p_mask_full(d->c2p_mask) && hll_c2p_inert<N>(h))) {
        Kastaun k;
        k.init(h.uH, e);
        Prim ph;
        const uint64_t mask = d->c2p_mask | (mres ? *mres : 0ull);
        cst = LOCAL ? c2p_single(k, 1e-10, 1e-10, ph)
                    : c2p_pass_b(k, rec[t], c2p_break_iter(mask), 1e-10, 1e-10, ph);
        p_hll = p_hll_from_prim(ph, e);
      }
      st = hlld_finish<N>(h, l, r, p_hll, f, nullptr, ps, iters, tol);
      if (cst & C2P_REDONE) st |= FS_C2P_REDONE;
    }
    F.p[0][t] = f[CD];
    F.p[1][t] = f[CS];
    F.p[2][t] = f[CS + 1];
    F.p[3][t] = f[CS + 2];
    F.p[4][t] = f[CT];
    F.p[5][t] = f[CB + 2];
    // F^x(By) / F^y(Bx): the face EMF
    F.p[6][t] = emf_hll ? emf_h : (N == 0 ? f[CB + 1] : f[CB + 0]);
    fD = f[CD];
    if (pstar_out) pstar_out[t] = ps;
  }
  // the deadband scale of the upwind EMF is needed for every solver
  diag_accumulate(diag_shard(dsh), active, st, iters, dbits(fabs(fD)));
}

// d[k] = the fold of the DIAG_SHARDS copies of sweep k (k < blockDim.x / 64):
// sums of the counts, maxima, OR of the con2prim mask.  One wave per sweep.
__global__ void k_fold_diag(const srmhd_flux_diag* dsh, srmhd_flux_diag* d) {
  static_assert(DIAG_SHARDS == 64, "one lane per shard");
  const int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const srmhd_flux_diag& c = *(const srmhd_flux_diag*)((const char*)dsh +
                                                       (int64_t)DIAG_SHARD_BYTES *
                                                           (k * DIAG_SHARDS + lane));
  srmhd_flux_diag r;
  r.n_interfaces = wave_sum(c.n_interfaces);
  r.n_not_converged = wave_sum(c.n_not_converged);
  r.n_wave_order_bad = wave_sum(c.n_wave_order_bad);
  r.n_bn_zero = wave_sum(c.n_bn_zero);
  r.n_hlle_fallback = wave_sum(c.n_hlle_fallback);
  r.sum_iters_converged = wave_sum(c.sum_iters_converged);
  r.n_converged = wave_sum(c.n_converged);
  r.max_iters = wave_max(c.max_iters);
  r.c2p_mask = wave_or(c.c2p_mask);
  r.absF_D_max = wave_max(c.absF_D_max);
  r.n_c2p_redo = wave_sum(c.n_c2p_redo);
  if (lane == 0) d[k] = r;
}

// FP64 recomputation of the faces the mixed-precision sweep flagged (its
// list: fix[0] = count, indices from fix[2]): the FP64 face states and HLLD,
// with the HLL-state con2prim of the face as its own batch.
// The list is short (tens of faces on a 2048^2 rotor stage) and each face is
// a long serial FP64 computation whose iteration counts differ from face to
// face, and the x fix-up runs beside the y sweep, which fills every CU.  So:
// one-wave workgroups, one face per wave while the list fits the grid (lanes
// on different faces would serialise each other's branches), at most 256
// registers (the compiler's own choice is 282, which fits beside only one of
// the sweep's 118-register waves on a SIMD; 256 fits beside two), and raised
// issue priority once resident.
constexpr int FIXUP_BLK = 64;
template <int N, int LIM>
__global__ __launch_bounds__(FIXUP_BLK) __attribute__((amdgpu_waves_per_eu(2)))
void k_fixup64(G2 g, St2 S, Pr2 P, srmhd_eos e, const unsigned* fix, Flx F, int emf_hll) {
  const unsigned cnt = fix[0];
  const unsigned waves = gridDim.x;
  const unsigned per = (cnt + waves - 1) / waves;   // faces per wave
  const unsigned lanes = per < FIXUP_BLK ? per : FIXUP_BLK;
  if (threadIdx.x >= lanes) return;
  __builtin_amdgcn_s_setprio(3);
  for (unsigned q = blockIdx.x * lanes + threadIdx.x; q < cnt; q += waves * lanes) {
    const int64_t t = fix[2 + q];
    int i, j;
    face_ij<N>(g, t, i, j);
    Prim l, r;
    face_prims2d<N, LIM>(g, S, P, e, i, j, l, r);
    const double emf_h = emf_hll ? hll_emf_flux<N>(l, r, e) : 0.0;
    HlldPrep<N> h;
    h.run(l, r, e);
    Kastaun k;
    k.init(h.uH, e);
    Prim ph;
    c2p_single(k, 1e-10, 1e-10, ph);   // the face's HLL state as its own batch
    double f[8], ps;
    int it;
    hlld_finish<N>(h, l, r, p_hll_from_prim(ph, e), f, nullptr, ps, it);
    F.p[0][t] = f[CD];
    F.p[1][t] = f[CS];
    F.p[2][t] = f[CS + 1];
    F.p[3][t] = f[CS + 2];
    F.p[4][t] = f[CT];
    F.p[5][t] = f[CB + 2];
    F.p[6][t] = emf_hll ? emf_h : (N == 0 ? f[CB + 1] : f[CB + 0]);
  }
}

// Gardiner-Stone corner EMF (ct.py:119-177) with the ring of zeros.  A block
// takes a tile of CT_TI x CT_TJ corners (j fastest); the cell EMFs
// E_c = Bx vy - By vx of the (CT_TI + 1) x (CT_TJ + 1) cells around it are
// formed once into LDS and each corner reads its four from there.
constexpr int CT_TI = 8, CT_TJ = 32;
static_assert(CT_TI * CT_TJ == BLK, "one corner per thread");
constexpr int CT_LJ = CT_TJ + 1, CT_LN = (CT_TI + 1) * CT_LJ;
__global__ __launch_bounds__(BLK) void k_corner(G2 g, St2 S, Pr2 P, const double* FxD,
                                                const double* FxB, const double* FyD,
                                                const double* FyB, const srmhd_flux_diag* dg,
                                                int upwind, double* Ez) {
  __shared__ double L[CT_LN];
  const int i0 = blockIdx.y * CT_TI, j0 = blockIdx.x * CT_TJ;
  const double* Bxf = S.p[6];
  const double* Byf = S.p[7];
  // cells (i0 - 1 + e / CT_LJ, j0 - 1 + e % CT_LJ); those outside [0, nxt) x
  // [0, nyt) only neighbour ring corners, which do not read them
  for (int e = threadIdx.x; e < CT_LN; e += BLK) {
    const int ci = i0 - 1 + e / CT_LJ, cj = j0 - 1 + e % CT_LJ;
    double v = 0.0;
    if (ci >= 0 && ci < g.nxt && cj >= 0 && cj < g.nyt) {
      const int64_t c = (int64_t)ci * g.nyt + cj;
      v = bcx(g, Bxf, ci, cj) * P.p[2][c] - bcy(g, Byf, ci, cj) * P.p[1][c];
    }
    L[e] = v;
  }
  __syncthreads();
  const int ti = threadIdx.x / CT_TJ, tj = threadIdx.x % CT_TJ;
  const int i = i0 + ti, j = j0 + tj;
  if (i > g.nxt || j > g.nyt) return;
  const int64_t t = (int64_t)i * (g.nyt + 1) + j;
  if (i == 0 || j == 0 || i == g.nxt || j == g.nyt) {
    Ez[t] = 0.0;
    return;
  }
  const int64_t yf = (int64_t)i * (g.nyt + 1) + j, yfm = (int64_t)(i - 1) * (g.nyt + 1) + j;
  const int64_t xf = (int64_t)i * g.nyt + j, xfm = (int64_t)i * g.nyt + j - 1;
  const double eyf = FyB[yf], eyfm = FyB[yfm];
  const double exf = -FxB[xf], exfm = -FxB[xfm];
  const double NE = L[(ti + 1) * CT_LJ + tj + 1], NW = L[ti * CT_LJ + tj + 1];
  const double SE = L[(ti + 1) * CT_LJ + tj], SW = L[ti * CT_LJ + tj];
  const double avg = 0.25 * ((eyf + eyfm) + (exf + exfm));
  if (!upwind) {
    Ez[t] = avg;
    return;
  }
  const double thx = 1e-9 * __longlong_as_double(dg[0].absF_D_max);
  const double thy = 1e-9 * __longlong_as_double(dg[1].absF_D_max);
  auto sgn = [](double F, double th) { return (fabs(F) > th) ? tsign(F) : 0.0; };
  const double Sx = sgn(FxD[xf], thx), Sxm = sgn(FxD[xfm], thx);
  const double Sy = sgn(FyD[yf], thy), Sym = sgn(FyD[yfm], thy);
  const double dN = (1.0 - Sx) * (NE - eyf) + (1.0 + Sx) * (NW - eyfm);
  const double dS = (1.0 - Sxm) * (eyf - SE) + (1.0 + Sxm) * (eyfm - SW);
  const double dEdy = 0.125 * (dS - dN);
  const double dE = (1.0 - Sy) * (NE - exf) + (1.0 + Sy) * (SE - exfm);
  const double dW = (1.0 - Sym) * (exf - NW) + (1.0 + Sym) * (exfm - SW);
  const double dEdx = 0.125 * (dW - dE);
  Ez[t] = avg + (dEdy + dEdx);
}

// out = a*U0 + b*U + (c*dt)*L(U) with the flux divergence (sweep.py:122-135),
// ct_rhs (ct.py:180-185), state.combine (state.py:77-103) and the face then
// cell BCs of sync_state (grid.py:229-255), each ghost taking its source.
__global__ __launch_bounds__(BLK) void k_update(G2 g, St2 U0, St2 U, const Flx Fx, const Flx Fy,
                                                const double* Ez, double a, double b, double c,
                                                const double* dt_dev, St2Out out) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const int nc = g.nyt + 1;
  if (t >= (int64_t)(g.nxt + 1) * nc) return;
  const int i = (int)(t / nc), j = (int)(t - (int64_t)i * nc);
  const double w = c * (*dt_dev);
  if (i < g.nxt && j < g.nyt) {
    const int si = src_cell(i, g.nx, g.ng, g.bcx), sj = src_cell(j, g.ny, g.ng, g.bcy);
    const int64_t cs = (int64_t)si * g.nyt + sj;
    const int64_t xlo = (int64_t)si * g.nyt + sj, xhi = (int64_t)(si + 1) * g.nyt + sj;
    const int64_t ylo = (int64_t)si * (g.nyt + 1) + sj, yhi = ylo + 1;
    const int64_t cd = (int64_t)i * g.nyt + j;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double rx = -(Fx.p[k][xhi] - Fx.p[k][xlo]) / g.dx;
      const double ry = -(Fy.p[k][yhi] - Fy.p[k][ylo]) / g.dy;
      out.p[k][cd] = a * U0.p[k][cs] + b * U.p[k][cs] + w * (rx + ry);
    }
  }
  if (j < g.nyt) {   // x-face (i, j), i in [0, nxt]
    const int fi = src_face(i, g.nx, g.ng, g.bcx), sj = src_cell(j, g.ny, g.ng, g.bcy);
    const int64_t s = (int64_t)fi * g.nyt + sj;
    const double dB = (Ez[(int64_t)fi * nc + sj] - Ez[(int64_t)fi * nc + sj + 1]) / g.dy;
    out.p[6][(int64_t)i * g.nyt + j] = a * U0.p[6][s] + b * U.p[6][s] + w * dB;
  }
  if (i < g.nxt) {   // y-face (i, j), j in [0, nyt]
    const int si = src_cell(i, g.nx, g.ng, g.bcx), fj = src_face(j, g.ny, g.ng, g.bcy);
    const int64_t s = (int64_t)si * (g.nyt + 1) + fj;
    const double dB = (Ez[(int64_t)(si + 1) * nc + fj] - Ez[(int64_t)si * nc + fj]) / g.dx;
    out.p[7][(int64_t)i * (g.nyt + 1) + j] = a * U0.p[7][s] + b * U.p[7][s] + w * dB;
  }
}

// ghost faces and cells of a state in place (the sync path)
__global__ __launch_bounds__(BLK) void k_bc_state(G2 g, St2Out U) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const int nc = g.nyt + 1;
  if (t >= (int64_t)(g.nxt + 1) * nc) return;
  const int i = (int)(t / nc), j = (int)(t - (int64_t)i * nc);
  if (i < g.nxt && j < g.nyt) {
    const int si = src_cell(i, g.nx, g.ng, g.bcx), sj = src_cell(j, g.ny, g.ng, g.bcy);
    if (si != i || sj != j) {
#pragma unroll
      for (int k = 0; k < 6; ++k)
        U.p[k][(int64_t)i * g.nyt + j] = U.p[k][(int64_t)si * g.nyt + sj];
    }
  }
  if (j < g.nyt) {
    const int fi = src_face(i, g.nx, g.ng, g.bcx), sj = src_cell(j, g.ny, g.ng, g.bcy);
    if (fi != i || sj != j) U.p[6][(int64_t)i * g.nyt + j] = U.p[6][(int64_t)fi * g.nyt + sj];
  }
  if (i < g.nxt) {
    const int si = src_cell(i, g.nx, g.ng, g.bcx), fj = src_face(j, g.ny, g.ng, g.bcy);
    if (si != i || fj != j)
      U.p[7][(int64_t)i * (g.nyt + 1) + j] = U.p[7][(int64_t)si * (g.nyt + 1) + fj];
  }
}

__global__ __launch_bounds__(BLK) void k_c2p2d_a(G2 g, St2 S, srmhd_eos e, C2PRec* rec,
                                                 unsigned long long* mask) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  uint64_t m = 0;
  if (t < (int64_t)g.nxt * g.nyt) {
    const int i = (int)(t / g.nyt), j = (int)(t - (int64_t)i * g.nyt);
    double u[8];
    cell_cons(g, S, i, j, u);
    Kastaun k;
    k.init(u, e);
    m = c2p_pass_a(k, rec[t]);
  }
  block_or_atomic(mask, m);
}

// pass B on the physical cells only: ghost primitives are overwritten by the
// primitive BCs (driver2d.py:216)
__global__ __launch_bounds__(BLK) void k_c2p2d_b(G2 g, St2 S, srmhd_eos e, const C2PRec* rec,
                                                 const unsigned long long* mask, double atmo_rho,
                                                 double atmo_eps, Pr2Out P,
                                                 unsigned long long* redo_count) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (t >= (int64_t)g.nx * g.ny) return;
  const int pi = (int)(t / g.ny), pj = (int)(t - (int64_t)pi * g.ny);
  const int i = pi + g.ng, j = pj + g.ng;
  double u[8];
  cell_cons(g, S, i, j, u);
  Kastaun k;
  k.init(u, e);
  Prim q;
  const int64_t c = (int64_t)i * g.nyt + j;
  const int st = c2p_pass_b(k, rec[c], c2p_break_iter(*mask), atmo_rho, atmo_eps, q);
  if (st & C2P_REDONE) atomicAdd(redo_count, 1ull);
  P.p[0][c] = q.rho;
  P.p[1][c] = q.v[0];
  P.p[2][c] = q.v[1];
  P.p[3][c] = q.v[2];
  P.p[4][c] = q.p;
  P.p[5][c] = q.eps;
}

// con2prim of the physical cells, each cell its own batch (the FP64_LOCAL form)
__global__ __launch_bounds__(BLK) void k_c2p2d_single(G2 g, St2 S, srmhd_eos e, double atmo_rho,
                                                      double atmo_eps, Pr2Out P,
                                                      unsigned long long* redo_count) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  bool redo = false;
  if (t < (int64_t)g.nx * g.ny) {
    const int pi = (int)(t / g.ny), pj = (int)(t - (int64_t)pi * g.ny);
    const int i = pi + g.ng, j = pj + g.ng;
    double u[8];
    cell_cons(g, S, i, j, u);
    Kastaun k;
    k.init(u, e);
    Prim q;
    const int st = c2p_single(k, atmo_rho, atmo_eps, q);
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

__global__ __launch_bounds__(BLK) void k_prim_bc(G2 g, Pr2Out P) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (t >= (int64_t)g.nxt * g.nyt) return;
  const int i = (int)(t / g.nyt), j = (int)(t - (int64_t)i * g.nyt);
  const int si = src_cell(i, g.nx, g.ng, g.bcx), sj = src_cell(j, g.ny, g.ng, g.bcy);
  if (si == i && sj == j) return;
  const int64_t s = (int64_t)si * g.nyt + sj;
#pragma unroll
  for (int k = 0; k < 6; ++k) P.p[k][t] = P.p[k][s];
}

// compute_dt_2d (driver2d.py:52-82): wave maxima of |v_x| + c_f, |v_y| + c_f
// over the physical cells, as non-negative double bit patterns
struct Pr9 {  // rho, vx, vy, vz, p, eps, Bx, By, Bz (Bx, By, Bz may be null)
  const double* p[9];
};

__global__ __launch_bounds__(BLK) void k_dt_reduce(G2 g, St2 S, Pr9 P, srmhd_eos e,
                                                   unsigned long long* mx) {
  unsigned long long bx = 0, by = 0;
  const int64_t n = (int64_t)g.nx * g.ny;
  for (int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x; t < n; t += (int64_t)gridDim.x * BLK) {
    const int i = (int)(t / g.ny) + g.ng, j = (int)(t % g.ny) + g.ng;
    const int64_t c = (int64_t)i * g.nyt + j;
    const double rho = P.p[0][c], vx = P.p[1][c], vy = P.p[2][c], vz = P.p[3][c];
    const double eps = P.p[5][c];
    const double Bx = P.p[6] ? P.p[6][c] : bcx(g, S.p[6], i, j);
    const double By = P.p[7] ? P.p[7][c] : bcy(g, S.p[7], i, j);
    const double Bz = P.p[8] ? P.p[8][c] : S.p[5][c];
    double p, cs2;
    eos_press_cs2(e, eps, rho, p, cs2);
    const double v2 = sq(vx) + sq(vy) + sq(vz);
    const double W = 1.0 / csqrt(clo(1.0 - v2, 1e-10));
    const double B2 = sq(Bx) + sq(By) + sq(Bz);
    const double vB = vx * Bx + vy * By + vz * Bz;
    const double b2 = (B2 + vB * vB) / (W * W);
    const double h = 1.0 + eps + p / rho;
    const double v02 = cs2 + (b2 / (rho * h + b2)) * (1.0 - cs2);
    const double cf = csqrt(chi_(clo(v02, 0.0), 1.0));
    // |.| keeps a NaN above +inf in the unsigned order (torch.max propagates NaN)
    const unsigned long long ax = dbits(fabs(vx) + cf) & 0x7fffffffffffffffull;
    const unsigned long long ay = dbits(fabs(vy) + cf) & 0x7fffffffffffffffull;
    bx = ax > bx ? ax : bx;
    by = ay > by ? ay : by;
  }
  block_max_atomic2(&mx[0], bx, &mx[1], by);
}

__global__ void k_dt_final(const unsigned long long* mx, double cfl, double dx, double dy,
                           double* dt) {
  if (threadIdx.x || blockIdx.x) return;
  const double sx = clo(__longlong_as_double(mx[0]), 1e-14);
  const double sy = clo(__longlong_as_double(mx[1]), 1e-14);
  // torch evaluates `cfl / tensor` as reciprocal(tensor) * cfl: two roundings
  *dt = (1.0 / (sx / dx + sy / dy)) * cfl;
}

__global__ void k_copy_diag(const srmhd_flux_diag* src, srmhd_flux_diag* dst, int n) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ── host side ───────────────────────────────────────────────────────────────

// Optional live timing: HIP event pairs around every launch of one kernel
// class, on the launch stream (srmhd_timing_enable / srmhd_timing_read).
struct Timing {
  int cls = -1;
  int cap = 0, used = 0;
  hipEvent_t* ev = nullptr;   // 2*cap events
} g_timing;

inline void tbegin(int cls, hipStream_t s) {
  if (cls == g_timing.cls && g_timing.used < g_timing.cap)
    (void)hipEventRecord(g_timing.ev[2 * g_timing.used], s);
}
inline void tend(int cls, hipStream_t s) {
  if (cls == g_timing.cls && g_timing.used < g_timing.cap) {
    (void)hipEventRecord(g_timing.ev[2 * g_timing.used + 1], s);
    ++g_timing.used;
  }
}

}  // namespace

// the live-timing hooks for the other translation units (srmhd_launch.h)
void srmhd::timing_begin(int cls, hipStream_t s) { tbegin(cls, s); }
void srmhd::timing_end(int cls, hipStream_t s) { tend(cls, s); }

namespace {


template <int N>
int flux_dispatch(int solver, int64_t n, const PrimPtrs& L, const PrimPtrs& R, const srmhd_eos& e,
                  C2PRec* rec, srmhd_flux_diag* d, const ConsPtrs& F, const ConsPtrs& U,
                  int want_u, double* pstar, uint8_t* status, hipStream_t s,
                  const HlldOpts& o = HlldOpts()) {
  const int nb = blocks_for(n);
  switch (solver) {
    case SOLVER_HLLE:
      k_flux_b<N, SOLVER_HLLE><<<nb, BLK, 0, s>>>(n, L, R, e, rec, d, F, U, want_u, pstar, status,
                                                  o.tol);
      break;
    case SOLVER_HLLC:
      k_flux_b<N, SOLVER_HLLC><<<nb, BLK, 0, s>>>(n, L, R, e, rec, d, F, U, want_u, pstar, status,
                                                  o.tol);
      break;
    default:
      // the HLL-state con2prim batch: the legacy bracket's p_hll
      if (o.bracket == SRMHD_BRACKET_LEGACY) k_flux_a<N><<<nb, BLK, 0, s>>>(n, L, R, e, rec, d);
      if (hlld_opts_fused(o)) {
        k_flux_b<N, SOLVER_HLLD><<<nb, BLK, 0, s>>>(n, L, R, e, rec, d, F, U, want_u, pstar,
                                                    status, o.tol);
      } else {
        if (const int rc = launch_status()) return rc;
        return launch_flux_opts_b(N, n, L.p, R.p, e, rec, d, F.p, want_u ? U.p : nullptr, pstar,
                                  status, o, s);
      }
  }
  return launch_status();
}

G2 make_g2(const srmhd_grid2d* g) {
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

bool grid_ok(const srmhd_grid2d* g, int limiter) {
  if (!g || g->nx < 1 || g->ny < 1 || g->ng < 1) return false;
  const int need = limiter == SRMHD_PCM ? 0 : (limiter <= SRMHD_MINMOD ? 1 : (limiter == SRMHD_MP7 ? 3 : 2));
  if (g->ng < (need + 1 > 2 ? need + 1 : 2) && limiter != SRMHD_PCM) return false;
  for (int bc : {g->bc_x, g->bc_y})
    if (bc != SRMHD_BC_OUTFLOW && bc != SRMHD_BC_PERIODIC) return false;
  if (g->bc_x == SRMHD_BC_PERIODIC && g->nx <= g->ng) return false;
  if (g->bc_y == SRMHD_BC_PERIODIC && g->ny <= g->ng) return false;
  return g->dx > 0.0 && g->dy > 0.0;
}

constexpr int N_STAGE_WORDS = 10;

struct Ws2 {
  double* psx = nullptr;        // optional p* tape of the x and y sweeps (not carved)
  double* psy = nullptr;
  Flx Fx, Fy;
  double* Ez;
  C2PRec *recx, *recy, *recc;
  srmhd_flux_diag* diag;        // [0] x sweep, [1] y sweep
  unsigned long long* words;    // [0] cell c2p mask, [1..2] dt maxima, [3] c2p status OR,
                                // [4..7] the linear solver's |F_D| maxima and counts,
                                // [8..9] the deferred faces' con2prim masks (x, y)
  Cells32 cf{};                 // FP32 cell variables of the MIXED form (carved for it only)
  double* hq[2] = {nullptr, nullptr};   // pass A's side states for pass B (FP64 form only)
  srmhd_flux_diag* dsh;         // DIAG_SHARDS copies of diag[0], then of diag[1]
  int emf_hll = 0;              // the scheme's emf_mode is SRMHD_EMF_HLL
  HlldOpts o;                   // the scheme's HLLD p* options
  int64_t bytes;
};

Ws2 carve2d(const G2& g, void* work, bool cells32 = false, bool sideq = false) {
  Carver c{(char*)work};
  Ws2 w;
  const int64_t nfx = (int64_t)(g.nxt + 1) * g.nyt, nfy = (int64_t)g.nxt * (g.nyt + 1);
  const int64_t ncell = (int64_t)g.nxt * g.nyt, ncor = (int64_t)(g.nxt + 1) * (g.nyt + 1);
  w.diag = c.take<srmhd_flux_diag>(2);
  w.words = c.take<unsigned long long>(N_STAGE_WORDS);
  for (int k = 0; k < 7; ++k) w.Fx.p[k] = c.take<double>(nfx);
  for (int k = 0; k < 7; ++k) w.Fy.p[k] = c.take<double>(nfy);
  w.Ez = c.take<double>(ncor);
  w.recx = c.take<C2PRec>(nfx);
  w.recy = c.take<C2PRec>(nfy);
  w.recc = c.take<C2PRec>(ncell);
  w.dsh = (srmhd_flux_diag*)c.take<char>(2 * DIAG_SHARDS * DIAG_SHARD_BYTES);
  if (cells32)
    for (int k = 0; k < 8; ++k) w.cf.p[k] = c.take<float>(ncell);
  if (sideq) {
    w.hq[0] = c.take<double>(HQ_N * nfx);
    w.hq[1] = c.take<double>(HQ_N * nfy);
  }
  w.bytes = c.off;
  return w;
}

constexpr int N_WORDS = (int)(2 * sizeof(srmhd_flux_diag) / 8);
constexpr int DIAG_SHARD_WORDS = 2 * DIAG_SHARDS * DIAG_SHARD_BYTES / 8;
constexpr int DIAG_ZERO_BLOCKS = (DIAG_SHARD_WORDS + 255) / 256;
static_assert(N_WORDS <= DIAG_ZERO_BLOCKS * 256, "the stage's zeroing grid covers the diagnostics");

// workgroups of k_sweep_a_rescue (grid-stride over the faces)
constexpr int RESCUE_BLOCKS = 1024;

// the sweep along N of ct_stage for one (limiter, solver), on stream s.
// HLLD: pass A with the inert faces deferred, the batch mask folded, the
// deferred faces' pass A only if that mask is not full, then pass B.  Each
// direction has its own shards, records, masks and fluxes.
template <int N, int LIM, int SOL>
void sweep_dir(const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e, const Ws2& w,
               hipStream_t s) {
  const int nb = blocks_for(N == 0 ? (int64_t)(g.nxt + 1) * g.nyt : (int64_t)g.nxt * (g.nyt + 1));
  srmhd_flux_diag* sh = (srmhd_flux_diag*)((char*)w.dsh + N * DIAG_SHARDS * DIAG_SHARD_BYTES);
  C2PRec* rec = N == 0 ? w.recx : w.recy;
  const bool hd = SOL == SOLVER_HLLD;
  unsigned long long* mres = hd ? w.words + 8 + N : nullptr;
  // HLLD: pass B takes the side states pass A formed (when the workspace
  // holds them) instead of reconstructing the faces and forming them again
  double* hq = hd ? w.hq[N] : nullptr;
  if (hd) {
    tbegin(N == 0 ? SRMHD_K_SWEEP_X_A : SRMHD_K_SWEEP_Y_A, s);
    k_sweep_a<N, LIM, true><<<nb, BLK, 0, s>>>(g, S, P, e, rec, sh, hq);
    tend(N == 0 ? SRMHD_K_SWEEP_X_A : SRMHD_K_SWEEP_Y_A, s);
    k_fold_diag<<<1, 64, 0, s>>>(sh, w.diag + N);   // the batch mask for pass B
    k_sweep_a_rescue<N, LIM><<<nb < RESCUE_BLOCKS ? nb : RESCUE_BLOCKS, BLK, 0, s>>>(
        g, S, P, e, rec, w.diag + N, mres, hq);
  }
  tbegin(N == 0 ? SRMHD_K_SWEEP_X_B : SRMHD_K_SWEEP_Y_B, s);
  if (hq)
    k_sweep_b<N, LIM, SOLVER_HLLD, false, true><<<nb, BLK, 0, s>>>(
        g, S, P, e, rec, w.diag + N, sh, N == 0 ? w.Fx : w.Fy, N == 0 ? w.psx : w.psy,
        w.emf_hll, w.o.tol, mres, hq);
  else
    k_sweep_b<N, LIM, SOL><<<nb, BLK, 0, s>>>(g, S, P, e, rec, w.diag + N, sh,
                                               N == 0 ? w.Fx : w.Fy, N == 0 ? w.psx : w.psy,
                                               w.emf_hll, w.o.tol, mres);
  tend(N == 0 ? SRMHD_K_SWEEP_X_B : SRMHD_K_SWEEP_Y_B, s);
}

// the two sweeps of ct_stage for one (limiter, solver).  ax (non-NULL: the
// caller holds g_aux_mu): the y sweep runs on the auxiliary stream beside
// the x sweep, joined before the fold of both; the results are the same
// bit for bit either way.
template <int LIM, int SOL>
void sweeps(const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e, const Ws2& w,
            hipStream_t s, AuxStream* ax = nullptr) {
  const bool forked = ax && hipEventRecord(ax->fork, s) == hipSuccess &&
                      hipStreamWaitEvent(ax->s, ax->fork, 0) == hipSuccess;
  sweep_dir<0, LIM, SOL>(g, S, P, e, w, s);
  sweep_dir<1, LIM, SOL>(g, S, P, e, w, forked ? ax->s : s);
  if (forked) {
    if (hipEventRecord(ax->join, ax->s) == hipSuccess)
      (void)hipStreamWaitEvent(s, ax->join, 0);
    else
      (void)hipStreamSynchronize(ax->s);   // no join event: wait on the host
  }
  k_fold_diag<<<1, 128, 0, s>>>(w.dsh, w.diag);
}

// HLLD sweeps under non-default p* options: the default first pass where the
// legacy bracket needs the HLL-state con2prim, then the option kernels
template <int LIM>
void sweeps_opts(const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e, const Ws2& w,
                 hipStream_t s) {
  const int bx = blocks_for((int64_t)(g.nxt + 1) * g.nyt);
  const int by = blocks_for((int64_t)g.nxt * (g.nyt + 1));
  srmhd_flux_diag* shx = w.dsh;
  srmhd_flux_diag* shy = (srmhd_flux_diag*)((char*)w.dsh + DIAG_SHARDS * DIAG_SHARD_BYTES);
  const bool legacy = w.o.bracket == SRMHD_BRACKET_LEGACY;
  if (legacy) {
    k_sweep_a<0, LIM><<<bx, BLK, 0, s>>>(g, S, P, e, w.recx, shx);
    k_fold_diag<<<1, 64, 0, s>>>(shx, w.diag + 0);
  }
  tbegin(SRMHD_K_SWEEP_X_B, s);
  launch_sweep_opts_b<0, LIM>(g, S, P, e, w.recx, w.diag + 0, shx, w.Fx, w.psx, w.emf_hll, w.o,
                              s);
  tend(SRMHD_K_SWEEP_X_B, s);
  if (legacy) {
    k_sweep_a<1, LIM><<<by, BLK, 0, s>>>(g, S, P, e, w.recy, shy);
    k_fold_diag<<<1, 64, 0, s>>>(shy, w.diag + 1);
  }
  tbegin(SRMHD_K_SWEEP_Y_B, s);
  launch_sweep_opts_b<1, LIM>(g, S, P, e, w.recy, w.diag + 1, shy, w.Fy, w.psy, w.emf_hll, w.o,
                              s);
  tend(SRMHD_K_SWEEP_Y_B, s);
  k_fold_diag<<<1, 128, 0, s>>>(w.dsh, w.diag);
}

// HLLD sweeps with each face's HLL-state con2prim solved locally (no pass A)
template <int LIM>
void sweeps_local(const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e, const Ws2& w,
                  hipStream_t s) {
  const int bx = blocks_for((int64_t)(g.nxt + 1) * g.nyt);
  const int by = blocks_for((int64_t)g.nxt * (g.nyt + 1));
  srmhd_flux_diag* shx = w.dsh;
  srmhd_flux_diag* shy = (srmhd_flux_diag*)((char*)w.dsh + DIAG_SHARDS * DIAG_SHARD_BYTES);
  tbegin(SRMHD_K_SWEEP_X_B, s);
  k_sweep_b<0, LIM, SOLVER_HLLD, true><<<bx, BLK, 0, s>>>(g, S, P, e, w.recx, w.diag + 0, shx,
                                                          w.Fx, w.psx, w.emf_hll, w.o.tol);
  tend(SRMHD_K_SWEEP_X_B, s);
  tbegin(SRMHD_K_SWEEP_Y_B, s);
  k_sweep_b<1, LIM, SOLVER_HLLD, true><<<by, BLK, 0, s>>>(g, S, P, e, w.recy, w.diag + 1, shy,
                                                          w.Fy, w.psy, w.emf_hll, w.o.tol);
  tend(SRMHD_K_SWEEP_Y_B, s);
  k_fold_diag<<<1, 128, 0, s>>>(w.dsh, w.diag);
}

template <int LIM>
void sweeps_solver(int sol, bool local, const G2& g, const St2& S, const Pr2& P,
                   const srmhd_eos& e, const Ws2& w, hipStream_t s, AuxStream* ax) {
  switch (sol) {
    case SOLVER_HLLE: sweeps<LIM, SOLVER_HLLE>(g, S, P, e, w, s, ax); break;
    case SOLVER_HLLC: sweeps<LIM, SOLVER_HLLC>(g, S, P, e, w, s, ax); break;
    default:
      if (local)
        sweeps_local<LIM>(g, S, P, e, w, s);
      else if (!hlld_opts_fused(w.o))
        sweeps_opts<LIM>(g, S, P, e, w, s);
      else
        sweeps<LIM, SOLVER_HLLD>(g, S, P, e, w, s, ax);
  }
}

// the mixed-precision sweep along dir and the FP64 fix-up of its flagged faces
// (the count is on the device: a grid of 1024 one-wave blocks strides over it)
constexpr int FIXUP_BLOCKS = 1024;
template <int N, int LIM>
void fixup_lim(const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e, const unsigned* fix,
               const Flx& F, int emf_hll, hipStream_t s) {
  k_fixup64<N, LIM><<<FIXUP_BLOCKS, FIXUP_BLK, 0, s>>>(g, S, P, e, fix, F, emf_hll);
}

template <int N>
void fixup_dispatch(int lim, const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e,
                    const unsigned* fix, const Flx& F, int emf_hll, hipStream_t s) {
  switch (lim) {
    case LIM_PCM: fixup_lim<N, LIM_PCM>(g, S, P, e, fix, F, emf_hll, s); break;
    case LIM_MC: fixup_lim<N, LIM_MC>(g, S, P, e, fix, F, emf_hll, s); break;
    case LIM_MINMOD: fixup_lim<N, LIM_MINMOD>(g, S, P, e, fix, F, emf_hll, s); break;
    case LIM_WENO5Z: fixup_lim<N, LIM_WENO5Z>(g, S, P, e, fix, F, emf_hll, s); break;
    case LIM_MP5: fixup_lim<N, LIM_MP5>(g, S, P, e, fix, F, emf_hll, s); break;
    default: fixup_lim<N, LIM_MP7>(g, S, P, e, fix, F, emf_hll, s);
  }
}

int sweeps_mixed(int lim, int sol, bool r64, const G2& g, const St2& S, const Pr2& P,
                 const srmhd_eos& e, const Ws2& w, hipStream_t s) {
  // the C2PRec areas are unused in this form: they hold the fix-up lists
  // (their counts, the first word of each, are cleared by the stage's zeroing launch)
  unsigned* fx = (unsigned*)w.recx;
  unsigned* fy = (unsigned*)w.recy;
  if (!r64) cells_f32(g, S, P, w.cf, s);
  srmhd_flux_diag* shx = w.dsh;
  srmhd_flux_diag* shy = (srmhd_flux_diag*)((char*)w.dsh + DIAG_SHARDS * DIAG_SHARD_BYTES);
  tbegin(SRMHD_K_SWEEP_X_B, s);
  sweep_mixed(0, lim, sol, r64, g, S, P, w.cf, e, shx, w.Fx, fx, w.emf_hll, s);
  tend(SRMHD_K_SWEEP_X_B, s);
  // a launch error so far is reported now: creating the auxiliary stream
  // may clear the error state when it fails
  if (const int rc = launch_status()) return rc;
  // the x fix-up reads the primitives and the state and writes only Fx: it
  // runs beside the y sweep, joined before anything reads Fx
  std::lock_guard<std::mutex> lock(g_aux_mu);
  AuxStream* ax = sol == SOLVER_HLLD ? aux_stream() : nullptr;
  const bool forked = ax && hipEventRecord(ax->fork, s) == hipSuccess &&
                      hipStreamWaitEvent(ax->s, ax->fork, 0) == hipSuccess;
  if (sol == SOLVER_HLLD)
    fixup_dispatch<0>(lim, g, S, P, e, fx, w.Fx, w.emf_hll, forked ? ax->s : s);
  const bool joinable = forked && hipEventRecord(ax->join, ax->s) == hipSuccess;
  tbegin(SRMHD_K_SWEEP_Y_B, s);
  sweep_mixed(1, lim, sol, r64, g, S, P, w.cf, e, shy, w.Fy, fy, w.emf_hll, s);
  tend(SRMHD_K_SWEEP_Y_B, s);
  if (sol == SOLVER_HLLD) fixup_dispatch<1>(lim, g, S, P, e, fy, w.Fy, w.emf_hll, s);
  if (joinable) {
    (void)hipStreamWaitEvent(s, ax->join, 0);
  } else if (forked) {
    (void)hipStreamSynchronize(ax->s);   // no join event: wait on the host
  }
  k_fold_diag<<<1, 128, 0, s>>>(w.dsh, w.diag);
  return 0;
}

void sweeps_dispatch(int lim, int sol, bool local, const G2& g, const St2& S, const Pr2& P,
                     const srmhd_eos& e, const Ws2& w, hipStream_t s, AuxStream* ax = nullptr) {
  switch (lim) {
    case LIM_PCM: sweeps_solver<LIM_PCM>(sol, local, g, S, P, e, w, s, ax); break;
    case LIM_MC: sweeps_solver<LIM_MC>(sol, local, g, S, P, e, w, s, ax); break;
    case LIM_MINMOD: sweeps_solver<LIM_MINMOD>(sol, local, g, S, P, e, w, s, ax); break;
    case LIM_WENO5Z: sweeps_solver<LIM_WENO5Z>(sol, local, g, S, P, e, w, s, ax); break;
    case LIM_MP5: sweeps_solver<LIM_MP5>(sol, local, g, S, P, e, w, s, ax); break;
    default: sweeps_solver<LIM_MP7>(sol, local, g, S, P, e, w, s, ax);
  }
}

// con2prim of every cell of state S as one batch, then primitive BCs;
// single: each cell its own batch (the mixed form)
void c2p_cells(const G2& g, const St2& S, const srmhd_eos& e, double atmo_rho, double atmo_eps,
               const Ws2& w, const Pr2Out& P, hipStream_t s, bool single = false,
               bool mup32 = false) {
  const int64_t ncell = (int64_t)g.nxt * g.nyt;
  if (single) {
    tbegin(SRMHD_K_C2P_A, s);
    if (mup32)
      c2p_cells_mixed(g, S, e, atmo_rho, atmo_eps, P, w.words + 3, s);
    else
      k_c2p2d_single<<<blocks_for((int64_t)g.nx * g.ny), BLK, 0, s>>>(g, S, e, atmo_rho,
                                                                      atmo_eps, P, w.words + 3);
    tend(SRMHD_K_C2P_A, s);
    k_prim_bc<<<blocks_for(ncell), BLK, 0, s>>>(g, P);
    return;
  }
  tbegin(SRMHD_K_C2P_A, s);
  k_c2p2d_a<<<blocks_for(ncell), BLK, 0, s>>>(g, S, e, w.recc, w.words + 0);
  tend(SRMHD_K_C2P_A, s);
  tbegin(SRMHD_K_C2P_B, s);
  k_c2p2d_b<<<blocks_for((int64_t)g.nx * g.ny), BLK, 0, s>>>(g, S, e, w.recc, w.words + 0,
                                                             atmo_rho, atmo_eps, P, w.words + 3);
  tend(SRMHD_K_C2P_B, s);
  k_prim_bc<<<blocks_for(ncell), BLK, 0, s>>>(g, P);
}

// ── 1D shock tube (driver.py:135-236) ────────────────────────────────────────

struct G1 {
  int n, ng, nt, bc;   // bc: 0 outflow, 1 periodic, 2 constant
  double dx;
};

// source of a 1D ghost cell for outflow / periodic (grid.py:49-71)
__device__ inline int src1(int i, const G1& g) {
  if (i < g.ng) return g.bc == SRMHD_BC_PERIODIC ? i + g.n : g.ng;
  if (i >= g.ng + g.n) return g.bc == SRMHD_BC_PERIODIC ? i - g.n : g.ng + g.n - 1;
  return i;
}

// get_interface_states (reconstruction.py:354-373): every primitive key is
// reconstructed directly (no W v, no floors, eps too) on the physical faces
// ng .. ng+n of the whole array.
template <int LIM>
__global__ __launch_bounds__(BLK) void k_recon1d(G1 g, PrimPtrs P, OutPrimPtrs L, OutPrimPtrs R) {
  const int j = blockIdx.x * BLK + threadIdx.x;
  if (j > g.n) return;
  const int f = g.ng + j;
  constexpr int w = stencil_halfwidth(LIM);
  double q[8];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
#pragma unroll
    for (int m = 3 - w; m <= 4 + w; ++m) q[m] = P.p[k][clampi(f + m - 4, 0, g.nt - 1)];
    double ql, qr;
    face_states<LIM>(q, f, g.nt, ql, qr);
    L.p[k][j] = ql;
    R.p[k][j] = qr;
  }
}

struct CC8 {
  const double* p[8];
};

// Heun stage (driver.py:219-232): stage 1 U1 = U + dt k1, stage 2
// U' = 0.5 (U + U1 + dt k2), with k = (F[j] - F[j+1])/dx on physical cells
// (0 on ghosts), F[Bx] := 0, then the BC of every ghost and Bx := U's Bx.
// Cons order: D, tau, Sx, Sy, Sz, Bx, By, Bz.
__global__ __launch_bounds__(BLK) void k_heun1d(G1 g, int stage, const double* dt_dev, CC8 U,
                                                CC8 U1, CC8 F, ConsPtrs out) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= g.nt) return;
  const double dt = *dt_dev;
  const bool ghost = i < g.ng || i >= g.ng + g.n;
  const int s = (g.bc == 2) ? i : src1(i, g);
  const bool sphys = s >= g.ng && s < g.ng + g.n;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    double v;
    if (k == CB) {
      v = U.p[k][i];
    } else if (ghost && g.bc == 2) {
      v = U.p[k][i];
    } else {
      double rhs = 0.0;
      if (sphys) {
        const int j = s - g.ng;
        rhs = (F.p[k][j] - F.p[k][j + 1]) / g.dx;
      }
      v = stage == 1 ? U.p[k][s] + dt * rhs : 0.5 * (U.p[k][s] + U1.p[k][s] + dt * rhs);
    }
    out.p[k][i] = v;
  }
}

// primitive BCs of the 1D driver; constant takes the old primitives
__global__ __launch_bounds__(BLK) void k_primbc1d(G1 g, OutPrimPtrs P, PrimPtrs Pold) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= g.nt) return;
  if (i >= g.ng && i < g.ng + g.n) return;
  if (g.bc == 2) {
#pragma unroll
    for (int k = 0; k < 9; ++k) P.p[k][i] = Pold.p[k][i];
    return;
  }
  const int s = src1(i, g);
#pragma unroll
  for (int k = 0; k < 9; ++k) P.p[k][i] = P.p[k][s];
}

// compute_dt (driver.py:77-127): (cfl dx) / max(|vx| + v_fast), torch's
// float / tensor being reciprocal(tensor) * float
__global__ __launch_bounds__(BLK) void k_dt1d(G1 g, PrimPtrs P, srmhd_eos e, double cfl,
                                              int speed_of_light, double* dt) {
  unsigned long long bx = 0;
  for (int t = threadIdx.x; t < g.n; t += BLK) {
    const int c = t + g.ng;
    const double rho = P.p[0][c], vx = P.p[1][c], vy = P.p[2][c], vz = P.p[3][c];
    const double eps = P.p[5][c], Bx = P.p[6][c], By = P.p[7][c], Bz = P.p[8][c];
    double p, cs2;
    eos_press_cs2(e, eps, rho, p, cs2);
    const double v2 = sq(vx) + sq(vy) + sq(vz);
    const double W = 1.0 / csqrt(clo(1.0 - v2, 1e-10));
    const double B2 = sq(Bx) + sq(By) + sq(Bz);
    const double vB = vx * Bx + vy * By + vz * Bz;
    const double b2 = (B2 + vB * vB) / (W * W);
    const double h = 1.0 + eps + p / rho;
    const double v02 = cs2 + (b2 / (rho * h + b2)) * (1.0 - cs2);
    const double vf = csqrt(chi_(clo(v02, 0.0), 1.0));
    const unsigned long long a = dbits(fabs(vx) + vf) & 0x7fffffffffffffffull;
    bx = a > bx ? a : bx;
  }
  __shared__ unsigned long long sh[BLK / 64];
  bx = wave_max(bx);
  if (lane0()) sh[threadIdx.x >> 6] = bx;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = 0;
    for (int w = 0; w < BLK / 64; ++w) m = sh[w] > m ? sh[w] : m;
    double smax = clo(__longlong_as_double(m), 1e-14);
    if (speed_of_light) smax = 1.0;
    *dt = (1.0 / smax) * (cfl * g.dx);
  }
}

struct Ws1 {
  OutPrimPtrs L, R, P1;
  ConsPtrs F, U1;
  C2PRec *recf, *recc;
  srmhd_flux_diag* diag;    // [0] stage-1 flux, [1] stage-2 flux
  unsigned long long* mask;  // cell con2prim batch mask, one per stage
  int64_t bytes;
};

Ws1 carve1d(const G1& g, void* work) {
  Carver c{(char*)work};
  Ws1 w;
  const int64_t nf = g.n + 1, nt = g.nt;
  w.diag = c.take<srmhd_flux_diag>(2);
  w.mask = c.take<unsigned long long>(2);
  for (int k = 0; k < 9; ++k) w.L.p[k] = c.take<double>(nf);
  for (int k = 0; k < 9; ++k) w.R.p[k] = c.take<double>(nf);
  for (int k = 0; k < 9; ++k) w.P1.p[k] = c.take<double>(nt);
  for (int k = 0; k < 8; ++k) w.F.p[k] = c.take<double>(nf);
  for (int k = 0; k < 8; ++k) w.U1.p[k] = c.take<double>(nt);
  w.recf = c.take<C2PRec>(nf);
  w.recc = c.take<C2PRec>(nt);
  w.bytes = c.off;
  return w;
}

template <int LIM>
void recon1d(const G1& g, const PrimPtrs& P, const Ws1& w, hipStream_t s) {
  k_recon1d<LIM><<<blocks_for(g.n + 1), BLK, 0, s>>>(g, P, w.L, w.R);
}

void recon1d_dispatch(int lim, const G1& g, const PrimPtrs& P, const Ws1& w, hipStream_t s) {
  switch (lim) {
    case LIM_PCM: recon1d<LIM_PCM>(g, P, w, s); break;
    case LIM_MC: recon1d<LIM_MC>(g, P, w, s); break;
    case LIM_MINMOD: recon1d<LIM_MINMOD>(g, P, w, s); break;
    case LIM_WENO5Z: recon1d<LIM_WENO5Z>(g, P, w, s); break;
    case LIM_MP5: recon1d<LIM_MP5>(g, P, w, s); break;
    default: recon1d<LIM_MP7>(g, P, w, s);
  }
}

PrimPtrs as_in(const OutPrimPtrs& o) {
  PrimPtrs p;
  for (int k = 0; k < 9; ++k) p.p[k] = o.p[k];
  return p;
}

// one stage: reconstruction, the flux on the n+1 physical faces, the Heun
// combination with BCs, con2prim of every cell, primitive BCs
int stage1d(const G1& g, int solver, int lim, const srmhd_eos& e, double atmo_rho, int stage,
            const double* dt_dev, const CC8& U, const CC8& U1, const PrimPtrs& P,
            const PrimPtrs& Pold, const ConsPtrs& out, const OutPrimPtrs& Pout, const Ws1& w,
            hipStream_t s, const HlldOpts& o) {
  recon1d_dispatch(lim, g, P, w, s);
  ConsPtrs none;
  for (int k = 0; k < 8; ++k) none.p[k] = nullptr;
  srmhd_flux_diag* d = w.diag + (stage - 1);
  const bool linear = solver == SRMHD_LINEAR;
  const int rc = flux_dispatch<0>(linear ? SOLVER_HLLD : solver, g.n + 1, as_in(w.L),
                                  as_in(w.R), e, w.recf, d, w.F, none, 0, nullptr, nullptr, s, o);
  if (rc) return rc;
  if (linear) {
    // the linearisation where it answers, over the HLLD fallback
    const int rl = launch_linear_flat(0, g.n + 1, as_in(w.L).p, as_in(w.R).p, e, LinOpts(),
                                      w.F.p, nullptr, nullptr, nullptr, nullptr, s);
    if (rl) return rl;
  }
  CC8 F;
  for (int k = 0; k < 8; ++k) F.p[k] = w.F.p[k];
  k_heun1d<<<blocks_for(g.nt), BLK, 0, s>>>(g, stage, dt_dev, U, U1, F, out);
  CConsPtrs C;
  for (int k = 0; k < 8; ++k) C.p[k] = out.p[k];
  unsigned long long* mask = w.mask + (stage - 1);
  k_c2p_a<<<blocks_for(g.nt), BLK, 0, s>>>(g.nt, C, e, w.recc, mask);
  k_c2p_b<<<blocks_for(g.nt), BLK, 0, s>>>(g.nt, C, e, w.recc, mask, atmo_rho, 1e-10, Pout,
                                           nullptr);
  k_primbc1d<<<blocks_for(g.nt), BLK, 0, s>>>(g, Pout, Pold);
  return launch_status();
}

bool g1_ok(const srmhd_grid1d* g) {
  return g && g->ncells >= 1 && g->ng >= 1 && g->bc >= 0 && g->bc <= 2 && g->dx > 0.0 &&
         !(g->bc == SRMHD_BC_PERIODIC && g->ncells < g->ng);
}

G1 make_g1(const srmhd_grid1d* g) {
  G1 r;
  r.n = g->ncells;
  r.ng = g->ng;
  r.nt = g->ncells + 2 * g->ng;
  r.bc = g->bc;
  r.dx = g->dx;
  return r;
}

// ── the allocation-free flux's scratch ──────────────────────────────────────
// srmhd_flux_noalloc carves srmhd_flux's workspace (the diag words, then one
// C2PRec per face) from this array instead of a caller's buffer.  Every call
// on a device shares that device's copy, so calls that overlap on one device
// race; the diag words are cleared by a kernel at every call and each record
// is written by pass A before pass B reads it.  In host code `scratch` names
// a host-side stand-in: the device address comes from hipGetSymbolAddress,
// once per device, and only that address reaches a kernel.
constexpr int64_t NOALLOC_BYTES = 256 + SRMHD_NOALLOC_MAX_FACES * (int64_t)sizeof(C2PRec);
static_assert(sizeof(srmhd_flux_diag) <= 256, "the diag words fit the first 256 bytes");

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
void* g_scratch[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// the current device's copy of `scratch` (nullptr when the device cannot be
// named or the symbol lookup fails)
void* scratch_ptr() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();   // reported as the entry point's own error
      return nullptr;
    }
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

}  // namespace

// ── C ABI ───────────────────────────────────────────────────────────────────

extern "C" {

int64_t srmhd_flux_workspace_bytes(int solver, int64_t n) {
  if (n < 0) return -1;
  return align256(sizeof(srmhd_flux_diag)) +
         (solver == SRMHD_HLLD ? align256(n * (int64_t)sizeof(C2PRec)) : 0);
}

int srmhd_flux(int solver, int idir, int64_t n, const double* const* primL,
               const double* const* primR, const srmhd_eos* eos, double* const* F,
               double* const* U, double* pstar, uint8_t* status, srmhd_flux_diag* diag,
               void* work, int64_t work_bytes, void* stream) {
  return srmhd_flux_ex(solver, idir, n, primL, primR, eos, nullptr, F, U, pstar, status, diag,
                       work, work_bytes, stream);
}

int srmhd_flux_ex(int solver, int idir, int64_t n, const double* const* primL,
                  const double* const* primR, const srmhd_eos* eos,
                  const srmhd_hlld_opts* opts, double* const* F, double* const* U,
                  double* pstar, uint8_t* status, srmhd_flux_diag* diag, void* work,
                  int64_t work_bytes, void* stream) {
  if (solver < 0 || solver > 2 || idir < 0 || idir > 2 || n < 0 || !primL || !primR || !F ||
      !eos_ok(eos) || !work)
    return SRMHD_EARG;
  const HlldOpts o = hlld_opts(opts);
  if (!hlld_opts_ok(o)) return SRMHD_EARG;
  if (work_bytes < srmhd_flux_workspace_bytes(solver, n)) return SRMHD_EARG;
  hipStream_t s = (hipStream_t)stream;
  Carver c{(char*)work};
  srmhd_flux_diag* d = c.take<srmhd_flux_diag>(1);
  C2PRec* rec = solver == SRMHD_HLLD ? c.take<C2PRec>(n) : nullptr;
  k_zero_u64<<<1, 64, 0, s>>>((unsigned long long*)d, (int)(sizeof(srmhd_flux_diag) / 8));
  if (n > 0) {
    PrimPtrs L, R;
    for (int k = 0; k < 9; ++k) {
      L.p[k] = primL[k];
      R.p[k] = primR[k];
      if (!L.p[k] || !R.p[k]) return SRMHD_EARG;
    }
    ConsPtrs Fp, Up;
    for (int k = 0; k < 8; ++k) {
      Fp.p[k] = F[k];
      Up.p[k] = U ? U[k] : nullptr;
      if (!Fp.p[k] || (U && !Up.p[k])) return SRMHD_EARG;
    }
    const int want_u = U ? 1 : 0;
    int rc;
    if (idir == 0)
      rc = flux_dispatch<0>(solver, n, L, R, *eos, rec, d, Fp, Up, want_u, pstar, status, s, o);
    else if (idir == 1)
      rc = flux_dispatch<1>(solver, n, L, R, *eos, rec, d, Fp, Up, want_u, pstar, status, s, o);
    else
      rc = flux_dispatch<2>(solver, n, L, R, *eos, rec, d, Fp, Up, want_u, pstar, status, s, o);
    if (rc) return rc;
  }
  if (diag) k_copy_diag<<<1, 64, 0, s>>>(d, diag, 1);
  return launch_status();
}

int srmhd_flux_noalloc(int solver, int idir, int64_t n, const double* const* primL,
                       const double* const* primR, const srmhd_eos* eos,
                       const srmhd_hlld_opts* opts, double* const* F, double* const* U,
                       double* pstar, uint8_t* status, srmhd_flux_diag* diag, void* stream) {
  if (n > SRMHD_NOALLOC_MAX_FACES) return SRMHD_ESIZE;
  if (n < 0) return SRMHD_EARG;
  if (stream) {
    // the scratch is the current device's: the stream must be on it
    int cur = 0, sd = 0;
    if (hipGetDevice(&cur) != hipSuccess ||
        hipStreamGetDevice((hipStream_t)stream, &sd) != hipSuccess || sd != cur) {
      (void)hipGetLastError();
      return SRMHD_EARG;
    }
  }
  void* work = scratch_ptr();
  if (!work) return SRMHD_EHIP + (int)hipErrorInvalidSymbol;
  return srmhd_flux_ex(solver, idir, n, primL, primR, eos, opts, F, U, pstar, status, diag, work,
                       NOALLOC_BYTES, stream);
}

int64_t srmhd_c2p_workspace_bytes(int64_t n) {
  if (n < 0) return -1;
  return align256(8 * 8) + align256(n * (int64_t)sizeof(C2PRec));
}

int srmhd_c2p(int64_t n, const double* const* cons, const srmhd_eos* eos, double atmo_rho,
              double atmo_eps, double* const* prim, uint8_t* status, void* work,
              int64_t work_bytes, void* stream) {
  if (n < 0 || !cons || !prim || !eos_ok(eos) || !work) return SRMHD_EARG;
  if (work_bytes < srmhd_c2p_workspace_bytes(n)) return SRMHD_EARG;
  hipStream_t s = (hipStream_t)stream;
  Carver c{(char*)work};
  unsigned long long* mask = c.take<unsigned long long>(8);
  C2PRec* rec = c.take<C2PRec>(n);
  if (n == 0) return SRMHD_OK;
  CConsPtrs C;
  OutPrimPtrs P;
  for (int k = 0; k < 8; ++k)
    if (!(C.p[k] = cons[k])) return SRMHD_EARG;
  for (int k = 0; k < 9; ++k)
    if (!(P.p[k] = prim[k])) return SRMHD_EARG;
  k_zero_u64<<<1, 64, 0, s>>>(mask, 1);
  k_c2p_a<<<blocks_for(n), BLK, 0, s>>>(n, C, *eos, rec, mask);
  k_c2p_b<<<blocks_for(n), BLK, 0, s>>>(n, C, *eos, rec, mask, atmo_rho, atmo_eps, P, status);
  return launch_status();
}

int srmhd_p2c(int64_t n, const double* const* prim, double* const* cons, void* stream) {
  if (n < 0 || !prim || !cons) return SRMHD_EARG;
  if (n == 0) return SRMHD_OK;
  PrimPtrs P;
  ConsPtrs C;
  for (int k = 0; k < 9; ++k)
    if (!(P.p[k] = prim[k])) return SRMHD_EARG;
  for (int k = 0; k < 8; ++k)
    if (!(C.p[k] = cons[k])) return SRMHD_EARG;
  k_p2c<<<blocks_for(n), BLK, 0, (hipStream_t)stream>>>(n, P, C);
  return launch_status();
}

int64_t srmhd2d_workspace_bytes(const srmhd_grid2d* g, const srmhd_scheme2d* s) {
  if (!g || !s || !grid_ok(g, s->limiter)) return -1;
  return carve2d(make_g2(g), nullptr, s->precision == SRMHD_PREC_MIXED,
                 s->precision == SRMHD_PREC_FP64).bytes;
}

// the HLLD p* options are the reference's default
static bool hlld_default(const HlldOpts& o) {
  return o.method == SRMHD_PSTAR_SECANT && o.bracket == SRMHD_BRACKET_LEGACY &&
         o.tol == PSTAR_TOL;
}

static bool scheme_ok(const srmhd_scheme2d* s) {
  if (!(s && s->solver >= 0 && s->solver <= 3 && s->limiter >= 0 && s->limiter <= 5 &&
        (s->flags & ~SRMHD_STAGE_SERIAL_SWEEPS) == 0 &&
        (s->precision == SRMHD_PREC_FP64 || s->precision == SRMHD_PREC_MIXED ||
         s->precision == SRMHD_PREC_MIXED_R64 || s->precision == SRMHD_PREC_FP64_LOCAL) &&
        (s->emf_mode == SRMHD_EMF_SOLVER || s->emf_mode == SRMHD_EMF_HLL)))
    return false;
  // the linearised flux is an FP64 solver
  if (s->solver == SRMHD_LINEAR && s->precision != SRMHD_PREC_FP64) return false;
  const HlldOpts o = hlld_opts(&s->hlld);
  // the options are the FP64 form's (the other forms have their own roots)
  return hlld_opts_ok(o) && (s->precision == SRMHD_PREC_FP64 || hlld_default(o));
}

// the deadband thresholds of the upwind EMF, as k_corner forms them
__global__ void k_tape_thresh(const srmhd_flux_diag* dg, double* th) {
  if (threadIdx.x || blockIdx.x) return;
  th[0] = 1e-9 * __longlong_as_double(dg[0].absF_D_max);
  th[1] = 1e-9 * __longlong_as_double(dg[1].absF_D_max);
}

static int stage_impl(const srmhd_grid2d* gp, const srmhd_scheme2d* sp, const srmhd_eos* eos,
                      double a, double b, double c, const double* dt_dev,
                      const double* const* U0, const double* const* U, const double* const* P,
                      double* const* out, double* const* Pout, srmhd_flux_diag* diag_xy,
                      const srmhd2d_tape* tape, void* work, int64_t work_bytes, void* stream) {
  if (!scheme_ok(sp) || !grid_ok(gp, sp->limiter) || !eos_ok(eos) || !dt_dev || !U0 || !U ||
      !P || !out || !Pout || !work)
    return SRMHD_EARG;
  if (tape && (sp->precision != SRMHD_PREC_FP64 || !tape->pstar_x || !tape->pstar_y ||
               !tape->FD_x || !tape->FD_y || !tape->thresh))
    return SRMHD_EARG;
  // the adjoint differentiates the reference's default root (of HLLD) only
  if (tape && (!hlld_default(hlld_opts(&sp->hlld)) || sp->solver == SRMHD_LINEAR))
    return SRMHD_EARG;
  const G2 g = make_g2(gp);
  Ws2 w = carve2d(g, work, sp->precision == SRMHD_PREC_MIXED,
                  sp->precision == SRMHD_PREC_FP64);
  if (work_bytes < w.bytes) return SRMHD_EARG;
  w.emf_hll = sp->emf_mode == SRMHD_EMF_HLL;
  w.o = hlld_opts(&sp->hlld);
  if (tape) {
    w.psx = tape->pstar_x;
    w.psy = tape->pstar_y;
  }
  St2 s0, s1, so;
  St2Out o;
  Pr2 pr, pro;
  Pr2Out po;
  for (int k = 0; k < 8; ++k) {
    s0.p[k] = U0[k];
    s1.p[k] = U[k];
    o.p[k] = out[k];
    so.p[k] = out[k];
    if (!s0.p[k] || !s1.p[k] || !o.p[k]) return SRMHD_EARG;
    if (o.p[k] == s0.p[k] || o.p[k] == s1.p[k]) return SRMHD_EARG;
  }
  for (int k = 0; k < 6; ++k) {
    pr.p[k] = P[k];
    po.p[k] = Pout[k];
    pro.p[k] = Pout[k];
    if (!pr.p[k] || !po.p[k] || po.p[k] == pr.p[k]) return SRMHD_EARG;
  }
  (void)pro;
  hipStream_t st = (hipStream_t)stream;
  const bool mixed = sp->precision == SRMHD_PREC_MIXED || sp->precision == SRMHD_PREC_MIXED_R64;
  {
    // the folded diagnostics, the sweeps' diagnostic shards, the words and
    // (mixed forms) the two fix-up list counts, in one launch
    ZeroSet z;
    z.p[0] = (unsigned long long*)w.diag;
    z.n[0] = N_WORDS;
    z.p[1] = (unsigned long long*)w.dsh;
    z.n[1] = DIAG_SHARD_WORDS;
    z.p[2] = w.words;
    z.n[2] = N_STAGE_WORDS;
    z.p[3] = mixed ? (unsigned long long*)w.recx : nullptr;
    z.n[3] = 1;
    z.p[4] = mixed ? (unsigned long long*)w.recy : nullptr;
    z.n[4] = 1;
    k_zero_set<<<DIAG_ZERO_BLOCKS, 256, 0, st>>>(z);
  }
  const bool local = sp->precision != SRMHD_PREC_FP64;   // per-element con2prim
  if (mixed) {
    const int rc = sweeps_mixed(sp->limiter, sp->solver, sp->precision == SRMHD_PREC_MIXED_R64,
                                g, s1, pr, *eos, w, st);
    if (rc) return rc;
  } else if (sp->solver == SRMHD_LINEAR) {
    // HLLD on every face, then the linearisation where it answers
    sweeps_dispatch(sp->limiter, SOLVER_HLLD, local, g, s1, pr, *eos, w, st);
    const int rc = launch_linear_sweeps(sp->limiter, g, s1, pr, *eos, LinOpts(), w.Fx, w.Fy,
                                        w.emf_hll, w.diag, w.words + 4, st);
    if (rc) return rc;
  } else if (local || (sp->flags & SRMHD_STAGE_SERIAL_SWEEPS)) {
    sweeps_dispatch(sp->limiter, sp->solver, local, g, s1, pr, *eos, w, st);
  } else {
    // the y sweep beside the x sweep on the device's auxiliary stream; a
    // launch error so far is reported first (creating that stream may clear
    // the error state when it fails)
    if (const int rc = launch_status()) return rc;
    std::lock_guard<std::mutex> lock(g_aux_mu);
    sweeps_dispatch(sp->limiter, sp->solver, local, g, s1, pr, *eos, w, st, aux_stream());
  }
  if (tape) {
    const int64_t nfx = (int64_t)(g.nxt + 1) * g.nyt, nfy = (int64_t)g.nxt * (g.nyt + 1);
    if (hipMemcpyAsync(tape->FD_x, w.Fx.p[0], nfx * sizeof(double), hipMemcpyDeviceToDevice,
                       st) != hipSuccess ||
        hipMemcpyAsync(tape->FD_y, w.Fy.p[0], nfy * sizeof(double), hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
      return launch_status();
    k_tape_thresh<<<1, 1, 0, st>>>(w.diag, tape->thresh);
  }
  const int64_t ncor = (int64_t)(g.nxt + 1) * (g.nyt + 1);
  tbegin(SRMHD_K_CORNER, st);
  const dim3 cgrid((g.nyt + 1 + CT_TJ - 1) / CT_TJ, (g.nxt + 1 + CT_TI - 1) / CT_TI);
  k_corner<<<cgrid, BLK, 0, st>>>(g, s1, pr, w.Fx.p[0], w.Fx.p[6], w.Fy.p[0], w.Fy.p[6], w.diag,
                                  sp->upwind, w.Ez);
  tend(SRMHD_K_CORNER, st);
  tbegin(SRMHD_K_UPDATE, st);
  k_update<<<blocks_for(ncor), BLK, 0, st>>>(g, s0, s1, w.Fx, w.Fy, w.Ez, a, b, c, dt_dev, o);
  tend(SRMHD_K_UPDATE, st);
  c2p_cells(g, so, *eos, sp->atmo_rho, sp->atmo_eps, w, po, st, local, mixed);
  if (diag_xy) k_copy_diag<<<1, 64, 0, st>>>(w.diag, diag_xy, 2);
  return launch_status();
}

int srmhd2d_stage(const srmhd_grid2d* gp, const srmhd_scheme2d* sp, const srmhd_eos* eos,
                  double a, double b, double c, const double* dt_dev,
                  const double* const* U0, const double* const* U, const double* const* P,
                  double* const* out, double* const* Pout, srmhd_flux_diag* diag_xy,
                  void* work, int64_t work_bytes, void* stream) {
  return stage_impl(gp, sp, eos, a, b, c, dt_dev, U0, U, P, out, Pout, diag_xy, nullptr, work,
                    work_bytes, stream);
}

int srmhd2d_stage_taped(const srmhd_grid2d* gp, const srmhd_scheme2d* sp, const srmhd_eos* eos,
                        double a, double b, double c, const double* dt_dev,
                        const double* const* U0, const double* const* U, const double* const* P,
                        double* const* out, double* const* Pout, const srmhd2d_tape* tape,
                        void* work, int64_t work_bytes, void* stream) {
  if (!tape) return SRMHD_EARG;
  return stage_impl(gp, sp, eos, a, b, c, dt_dev, U0, U, P, out, Pout, nullptr, tape, work,
                    work_bytes, stream);
}

int srmhd2d_sync(const srmhd_grid2d* gp, const srmhd_scheme2d* sp, const srmhd_eos* eos,
                 double* const* U, double* const* P, void* work, int64_t work_bytes,
                 void* stream) {
  if (!scheme_ok(sp) || !grid_ok(gp, sp->limiter) || !eos_ok(eos) || !U || !P || !work)
    return SRMHD_EARG;
  const G2 g = make_g2(gp);
  const Ws2 w = carve2d(g, work);
  if (work_bytes < w.bytes) return SRMHD_EARG;
  St2 s;
  St2Out so;
  Pr2Out po;
  for (int k = 0; k < 8; ++k) {
    s.p[k] = U[k];
    so.p[k] = U[k];
    if (!U[k]) return SRMHD_EARG;
  }
  for (int k = 0; k < 6; ++k)
    if (!(po.p[k] = P[k])) return SRMHD_EARG;
  hipStream_t st = (hipStream_t)stream;
  const int64_t ncor = (int64_t)(g.nxt + 1) * (g.nyt + 1);
  k_zero_u64<<<1, 64, 0, st>>>(w.words, 4);
  k_bc_state<<<blocks_for(ncor), BLK, 0, st>>>(g, so);
  c2p_cells(g, s, *eos, sp->atmo_rho, sp->atmo_eps, w, po, st);
  return launch_status();
}

int srmhd2d_dt(const srmhd_grid2d* gp, const srmhd_eos* eos, double cfl, const double* const* U,
               const double* const* P, double* dt_dev, void* work, int64_t work_bytes,
               void* stream) {
  if (!gp || gp->nx < 1 || gp->ny < 1 || !eos_ok(eos) || !P || !dt_dev || !work)
    return SRMHD_EARG;
  const G2 g = make_g2(gp);
  const Ws2 w = carve2d(g, work);
  if (work_bytes < w.bytes) return SRMHD_EARG;
  St2 s;
  Pr9 pr;
  for (int k = 0; k < 9; ++k) pr.p[k] = P[k];
  for (int k = 0; k < 6; ++k)
    if (!pr.p[k]) return SRMHD_EARG;
  const bool need_u = !pr.p[6] || !pr.p[7] || !pr.p[8];
  for (int k = 0; k < 8; ++k) {
    s.p[k] = U ? U[k] : nullptr;
    if (need_u && !s.p[k]) return SRMHD_EARG;
  }
  hipStream_t st = (hipStream_t)stream;
  k_zero_u64<<<1, 64, 0, st>>>(w.words + 1, 2);
  const int nb = blocks_for((int64_t)g.nx * g.ny);
  k_dt_reduce<<<nb < 2048 ? nb : 2048, BLK, 0, st>>>(g, s, pr, *eos, w.words + 1);
  k_dt_final<<<1, 64, 0, st>>>(w.words + 1, cfl, g.dx, g.dy, dt_dev);
  return launch_status();
}

int64_t srmhd1d_workspace_bytes(const srmhd_grid1d* g) {
  if (!g1_ok(g)) return -1;
  return carve1d(make_g1(g), nullptr).bytes;
}

int srmhd1d_rk2_step(const srmhd_grid1d* gp, int solver, int limiter, const srmhd_eos* eos,
                     double atmo_rho, const double* dt_dev, const double* const* cons,
                     const double* const* prim, double* const* cons_out, double* const* prim_out,
                     srmhd_flux_diag* diag2, void* work, int64_t work_bytes, void* stream) {
  return srmhd1d_rk2_step_ex(gp, solver, limiter, eos, nullptr, atmo_rho, dt_dev, cons, prim,
                             cons_out, prim_out, diag2, work, work_bytes, stream);
}

int srmhd1d_rk2_step_ex(const srmhd_grid1d* gp, int solver, int limiter, const srmhd_eos* eos,
                        const srmhd_hlld_opts* opts, double atmo_rho, const double* dt_dev,
                        const double* const* cons, const double* const* prim,
                        double* const* cons_out, double* const* prim_out,
                        srmhd_flux_diag* diag2, void* work, int64_t work_bytes, void* stream) {
  if (!g1_ok(gp) || solver < 0 || solver > 3 || limiter < 0 || limiter > 5 || !eos_ok(eos) ||
      !dt_dev || !cons || !prim || !cons_out || !prim_out || !work)
    return SRMHD_EARG;
  const HlldOpts o = hlld_opts(opts);
  if (!hlld_opts_ok(o)) return SRMHD_EARG;
  const G1 g = make_g1(gp);
  const Ws1 w = carve1d(g, work);
  if (work_bytes < w.bytes) return SRMHD_EARG;
  CC8 U, U1;
  ConsPtrs out;
  PrimPtrs P;
  OutPrimPtrs Pout;
  for (int k = 0; k < 8; ++k) {
    U.p[k] = cons[k];
    U1.p[k] = w.U1.p[k];
    out.p[k] = cons_out[k];
    if (!U.p[k] || !out.p[k] || out.p[k] == U.p[k]) return SRMHD_EARG;
  }
  for (int k = 0; k < 9; ++k) {
    P.p[k] = prim[k];
    Pout.p[k] = prim_out[k];
    if (!P.p[k] || !Pout.p[k] || Pout.p[k] == P.p[k]) return SRMHD_EARG;
  }
  hipStream_t s = (hipStream_t)stream;
  k_zero_u64<<<1, 64, 0, s>>>((unsigned long long*)w.diag, N_WORDS);
  k_zero_u64<<<1, 64, 0, s>>>(w.mask, 2);
  int rc =
      stage1d(g, solver, limiter, *eos, atmo_rho, 1, dt_dev, U, U1, P, P, w.U1, w.P1, w, s, o);
  if (rc) return rc;
  rc = stage1d(g, solver, limiter, *eos, atmo_rho, 2, dt_dev, U, U1, as_in(w.P1), P, out, Pout,
               w, s, o);
  if (rc) return rc;
  if (diag2) k_copy_diag<<<1, 64, 0, s>>>(w.diag, diag2, 2);
  return launch_status();
}

int srmhd1d_dt(const srmhd_grid1d* gp, const srmhd_eos* eos, double cfl, int speed_of_light,
               const double* const* prim, double* dt_dev, void* stream) {
  if (!g1_ok(gp) || !eos_ok(eos) || !prim || !dt_dev) return SRMHD_EARG;
  const G1 g = make_g1(gp);
  PrimPtrs P;
  for (int k = 0; k < 9; ++k)
    if (!(P.p[k] = prim[k])) return SRMHD_EARG;
  k_dt1d<<<1, BLK, 0, (hipStream_t)stream>>>(g, P, *eos, cfl, speed_of_light, dt_dev);
  return launch_status();
}

int srmhd_timing_enable(int kernel_class, int max_launches) {
  if (max_launches < 0 || max_launches > (1 << 20)) return SRMHD_EARG;
  if (g_timing.ev) {
    for (int i = 0; i < 2 * g_timing.cap; ++i) (void)hipEventDestroy(g_timing.ev[i]);
    delete[] g_timing.ev;
    g_timing.ev = nullptr;
  }
  g_timing.cls = -1;
  g_timing.cap = g_timing.used = 0;
  if (kernel_class < 0 || max_launches == 0) return SRMHD_OK;
  g_timing.ev = new hipEvent_t[2 * max_launches];
  for (int i = 0; i < 2 * max_launches; ++i) {
    const hipError_t e = hipEventCreate(&g_timing.ev[i]);
    if (e != hipSuccess) return hip_status(e);
  }
  g_timing.cap = max_launches;
  g_timing.cls = kernel_class;
  return SRMHD_OK;
}

int srmhd_timing_read(double* total_ms, int64_t* launches) {
  if (!total_ms || !launches) return SRMHD_EARG;
  double sum = 0.0;
  for (int i = 0; i < g_timing.used; ++i) {
    hipError_t e = hipEventSynchronize(g_timing.ev[2 * i + 1]);
    if (e != hipSuccess) return hip_status(e);
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, g_timing.ev[2 * i], g_timing.ev[2 * i + 1]);
    if (e != hipSuccess) return hip_status(e);
    sum += ms;
  }
  *total_ms = sum;
  *launches = g_timing.used;
  g_timing.used = 0;
  return SRMHD_OK;
}

}  // extern "C"
