// HLLD under the reference's non-default p* options (srmhd_pstar.h): the
// second pass of the flat flux (srmhd_flux_ex, the 1D step) and of the 2D
// sweeps.  The first pass -- the batch-coupled con2prim of every face's HLL
// state, needed by the legacy bracket only -- is the default path's own
// kernel (srmhd_kernels.hip), launched before these.  One kernel per normal
// direction (and limiter in 2D) with the method and bracket chosen at run
// time: these are not the production path, so register pressure is not
// tuned for.
#include "srmhd_grid2d.h"
#include "srmhd_pstar.h"
#include "srmhd_pstar_launch.h"

using namespace srmhd;

namespace {

__device__ inline void load9(const double* const (&s)[9], int64_t i, Prim& P) {
  P.rho = s[0][i];
  P.v[0] = s[1][i];
  P.v[1] = s[2][i];
  P.v[2] = s[3][i];
  P.p = s[4][i];
  P.eps = s[5][i];
  P.B[0] = s[6][i];
  P.B[1] = s[7][i];
  P.B[2] = s[8][i];
}

struct FluxArgs {
  const double* L[9];
  const double* R[9];
  double* F[8];
  double* U[8];
};

// p_hll of the legacy bracket: the HLL state's con2prim finished from its
// pass-A record and the batch's break iteration   hlld.py:1253-1263
template <int N>
__device__ inline double legacy_p_hll(const HlldPrep<N>& h, const srmhd_eos& e,
                                      const C2PRec& rec, const srmhd_flux_diag* d, int& cst) {
  Kastaun k;
  k.init(h.uH, e);
  Prim ph;
  cst = c2p_pass_b(k, rec, c2p_break_iter(d->c2p_mask), 1e-10, 1e-10, ph);
  return p_hll_from_prim(ph, e);
}

template <int N>
__global__ __launch_bounds__(BLK) void k_flux_opts_b(int64_t n, FluxArgs a, srmhd_eos e,
                                                     const C2PRec* rec, srmhd_flux_diag* d,
                                                     int want_u, double* pstar,
                                                     uint8_t* status, HlldOpts o) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const bool active = i < n;
  int st = 0, iters = 0;
  if (active) {
    Prim l, r;
    load9(a.L, i, l);
    load9(a.R, i, r);
    HlldPrep<N> h;
    h.run(l, r, e);
    int cst = 0;
    const double p_hll =
        o.bracket == SRMHD_BRACKET_LEGACY ? legacy_p_hll<N>(h, e, rec[i], d, cst) : 0.0;
    double f[8], u[8], ps;
    st = hlld_finish_opts<N>(h, p_hll, o, f, u, ps, iters);
    if (cst & C2P_REDONE) st |= FS_C2P_REDONE;
#pragma unroll
    for (int k = 0; k < 8; ++k) a.F[k][i] = f[k];
    if (want_u) {
#pragma unroll
      for (int k = 0; k < 8; ++k) a.U[k][i] = u[k];
    }
    if (pstar) pstar[i] = ps;
    if (status) status[i] = (uint8_t)st;
  }
  diag_accumulate(d, active, st, iters, 0ull);
}

template <int N, int LIM>
__global__ __launch_bounds__(BLK) void k_sweep_opts_b(G2 g, St2 S, Pr2 P, srmhd_eos e,
                                                      const C2PRec* rec,
                                                      const srmhd_flux_diag* d,
                                                      srmhd_flux_diag* dsh, Flx F,
                                                      double* pstar_out, int emf_hll,
                                                      HlldOpts o) {
  const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const bool active = t < n_faces<N>(g);
  int st = 0, iters = 0;
  double fD = 0.0;
  if (active) {
    int i, j;
    face_ij<N>(g, t, i, j);
    Prim l, r;
    face_prims2d<N, LIM>(g, S, P, e, i, j, l, r);
    const double emf_h = emf_hll ? hll_emf_flux<N>(l, r, e) : 0.0;
    HlldPrep<N> h;
    h.run(l, r, e);
    int cst = 0;
    const double p_hll =
        o.bracket == SRMHD_BRACKET_LEGACY ? legacy_p_hll<N>(h, e, rec[t], d, cst) : 0.0;
    double f[8], ps;
    st = hlld_finish_opts<N>(h, p_hll, o, f, nullptr, ps, iters);
    if (cst & C2P_REDONE) st |= FS_C2P_REDONE;
    F.p[0][t] = f[CD];
    F.p[1][t] = f[CS];
    F.p[2][t] = f[CS + 1];
    F.p[3][t] = f[CS + 2];
    F.p[4][t] = f[CT];
    F.p[5][t] = f[CB + 2];
    F.p[6][t] = emf_hll ? emf_h : (N == 0 ? f[CB + 1] : f[CB + 0]);
    fD = f[CD];
    if (pstar_out) pstar_out[t] = ps;
  }
  diag_accumulate(diag_shard(dsh), active, st, iters, dbits(fabs(fD)));
}

}  // namespace

namespace srmhd {

int launch_flux_opts_b(int idir, int64_t n, const double* const* L, const double* const* R,
                       const srmhd_eos& e, const C2PRec* rec, srmhd_flux_diag* d,
                       double* const* F, double* const* U, double* pstar, uint8_t* status,
                       const HlldOpts& o, hipStream_t s) {
  FluxArgs a;
  for (int k = 0; k < 9; ++k) {
    a.L[k] = L[k];
    a.R[k] = R[k];
  }
  for (int k = 0; k < 8; ++k) {
    a.F[k] = F[k];
    a.U[k] = U ? U[k] : nullptr;
  }
  const int want_u = U ? 1 : 0;
  const int nb = blocks_for(n);
  if (idir == 0)
    k_flux_opts_b<0><<<nb, BLK, 0, s>>>(n, a, e, rec, d, want_u, pstar, status, o);
  else if (idir == 1)
    k_flux_opts_b<1><<<nb, BLK, 0, s>>>(n, a, e, rec, d, want_u, pstar, status, o);
  else
    k_flux_opts_b<2><<<nb, BLK, 0, s>>>(n, a, e, rec, d, want_u, pstar, status, o);
  return launch_status();
}

template <int N, int LIM>
void launch_sweep_opts_b(const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e,
                         const C2PRec* rec, const srmhd_flux_diag* d, srmhd_flux_diag* dsh,
                         const Flx& F, double* pstar_out, int emf_hll, const HlldOpts& o,
                         hipStream_t s) {
  k_sweep_opts_b<N, LIM><<<blocks_for(n_faces<N>(g)), BLK, 0, s>>>(g, S, P, e, rec, d, dsh, F,
                                                                   pstar_out, emf_hll, o);
}

#define SRMHD_SWEEP_OPTS(N, LIM)                                                           \
  template void launch_sweep_opts_b<N, LIM>(const G2&, const St2&, const Pr2&,             \
                                            const srmhd_eos&, const C2PRec*,               \
                                            const srmhd_flux_diag*, srmhd_flux_diag*,      \
                                            const Flx&, double*, int, const HlldOpts&,     \
                                            hipStream_t);
SRMHD_SWEEP_OPTS(0, LIM_PCM)
SRMHD_SWEEP_OPTS(1, LIM_PCM)
SRMHD_SWEEP_OPTS(0, LIM_MC)
SRMHD_SWEEP_OPTS(1, LIM_MC)
SRMHD_SWEEP_OPTS(0, LIM_MINMOD)
SRMHD_SWEEP_OPTS(1, LIM_MINMOD)
SRMHD_SWEEP_OPTS(0, LIM_WENO5Z)
SRMHD_SWEEP_OPTS(1, LIM_WENO5Z)
SRMHD_SWEEP_OPTS(0, LIM_MP5)
SRMHD_SWEEP_OPTS(1, LIM_MP5)
SRMHD_SWEEP_OPTS(0, LIM_MP7)
SRMHD_SWEEP_OPTS(1, LIM_MP7)
#undef SRMHD_SWEEP_OPTS

}  // namespace srmhd
