// The linearised flux on the HIP path (srmhd_linear_flux, include/srmhd.h):
// the fallback solver's flux of every face (srmhd_flux_ex), then one thread
// per face runs linear_face (srmhd_linear.h) with its three 7x7 matrices in
// LDS and overwrites the faces it answers.  64-face workgroups (one wave):
// 3 x 49 doubles x 64 lanes = 73.5 KiB of LDS, two workgroups per CU.
#include "srmhd_launch.h"
#include "srmhd_linear_launch.h"

using namespace srmhd;

namespace {

struct LPrim {
  const double* p[9];
};
struct LCons {
  double* p[8];
};

__global__ void k_zero_count(unsigned long long* c) {
  if (threadIdx.x == 0) *c = 0ull;
}

template <int N>
__global__ __launch_bounds__(LIN_BLK) void k_linear(int64_t n, LPrim Lp, LPrim Rp, srmhd_eos e,
                                                    LinOpts o, LCons F, LCons U, double* pstar,
                                                    uint8_t* mask, unsigned long long* count) {
  __shared__ double sm[3 * LIN_NV * LIN_NV * LIN_BLK];
  const int64_t i = (int64_t)blockIdx.x * LIN_BLK + threadIdx.x;
  bool ok = false;
  if (i < n) {
    Prim L, R;
    L.rho = Lp.p[0][i];
    L.v[0] = Lp.p[1][i];
    L.v[1] = Lp.p[2][i];
    L.v[2] = Lp.p[3][i];
    L.p = Lp.p[4][i];
    L.eps = Lp.p[5][i];
    L.B[0] = Lp.p[6][i];
    L.B[1] = Lp.p[7][i];
    L.B[2] = Lp.p[8][i];
    R.rho = Rp.p[0][i];
    R.v[0] = Rp.p[1][i];
    R.v[1] = Rp.p[2][i];
    R.v[2] = Rp.p[3][i];
    R.p = Rp.p[4][i];
    R.eps = Rp.p[5][i];
    R.B[0] = Rp.p[6][i];
    R.B[1] = Rp.p[7][i];
    R.B[2] = Rp.p[8][i];
    const LinMat M1{sm + threadIdx.x}, M2{sm + LIN_NV * LIN_NV * LIN_BLK + threadIdx.x},
        M3{sm + 2 * LIN_NV * LIN_NV * LIN_BLK + threadIdx.x};
    double f[8], u[8], ps;
    ok = linear_face<N>(L, R, e, o, M1, M2, M3, f, u, ps);
    if (ok) {
#pragma unroll
      for (int k = 0; k < 8; ++k) F.p[k][i] = f[k];
      if (U.p[0]) {
#pragma unroll
        for (int k = 0; k < 8; ++k) U.p[k][i] = u[k];
      }
      if (pstar) pstar[i] = ps;
    }
    if (mask) mask[i] = ok ? 1 : 0;
  }
  const unsigned long long b = __ballot(ok);
  if ((threadIdx.x & 63) == 0 && b && count) atomicAdd(count, (unsigned long long)__popcll(b));
}

// the 2D sweep along N: the face states of face_prims2d (the sweep's own),
// the linearisation, the seven sweep fluxes (D, Sx, Sy, Sz, tau, Bz, E) where
// it answers, and the wave maxima of the final |F_D| and the answered count
template <int N, int LIM>
__global__ __launch_bounds__(LIN_BLK) void k_sweep_linear(G2 g, St2 S, Pr2 P, srmhd_eos e,
                                                          LinOpts o, Flx F, int emf_hll,
                                                          unsigned long long* words) {
  __shared__ double sm[3 * LIN_NV * LIN_NV * LIN_BLK];
  const int64_t t = (int64_t)blockIdx.x * LIN_BLK + threadIdx.x;
  const bool active = t < n_faces<N>(g);
  bool ok = false;
  double fD = 0.0;
  if (active) {
    int i, j;
    face_ij<N>(g, t, i, j);
    Prim l, r;
    face_prims2d<N, LIM>(g, S, P, e, i, j, l, r);
    const LinMat M1{sm + threadIdx.x}, M2{sm + LIN_NV * LIN_NV * LIN_BLK + threadIdx.x},
        M3{sm + 2 * LIN_NV * LIN_NV * LIN_BLK + threadIdx.x};
    double f[8], u[8], ps;
    ok = linear_face<N>(l, r, e, o, M1, M2, M3, f, u, ps);
    if (ok) {
      F.p[0][t] = f[CD];
      F.p[1][t] = f[CS];
      F.p[2][t] = f[CS + 1];
      F.p[3][t] = f[CS + 2];
      F.p[4][t] = f[CT];
      F.p[5][t] = f[CB + 2];
      if (!emf_hll) F.p[6][t] = N == 0 ? f[CB + 1] : f[CB + 0];
      fD = f[CD];
    } else {
      fD = F.p[0][t];
    }
  }
  const unsigned long long m = wave_max(active ? dbits(fabs(fD)) : 0ull);
  const unsigned long long b = __ballot(ok);
  if ((threadIdx.x & 63) == 0) {
    if (m) atomicMax(&words[N], m);
    if (b) atomicAdd(&words[2 + N], (unsigned long long)__popcll(b));
  }
}

__global__ void k_zero4(unsigned long long* w) {
  if (threadIdx.x < 4) w[threadIdx.x] = 0ull;
}

__global__ void k_linear_fd(srmhd_flux_diag* d, const unsigned long long* w) {
  if (threadIdx.x < 2) d[threadIdx.x].absF_D_max = w[threadIdx.x];
}

template <int LIM>
void sweeps_linear(const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e, const LinOpts& o,
                   const Flx& Fx, const Flx& Fy, int emf_hll, unsigned long long* words,
                   hipStream_t s) {
  const int bx = (int)((n_faces<0>(g) + LIN_BLK - 1) / LIN_BLK);
  const int by = (int)((n_faces<1>(g) + LIN_BLK - 1) / LIN_BLK);
  k_sweep_linear<0, LIM><<<bx, LIN_BLK, 0, s>>>(g, S, P, e, o, Fx, emf_hll, words);
  k_sweep_linear<1, LIM><<<by, LIN_BLK, 0, s>>>(g, S, P, e, o, Fy, emf_hll, words);
}

}  // namespace

namespace srmhd {

int launch_linear_flat(int idir, int64_t n, const double* const* L, const double* const* R,
                       const srmhd_eos& e, const LinOpts& o, double* const* F, double* const* U,
                       double* pstar, uint8_t* mask, unsigned long long* count, hipStream_t s) {
  if (n <= 0) return launch_status();
  LPrim Lp, Rp;
  for (int k = 0; k < 9; ++k) {
    Lp.p[k] = L[k];
    Rp.p[k] = R[k];
  }
  LCons Fp, Up;
  for (int k = 0; k < 8; ++k) {
    Fp.p[k] = F[k];
    Up.p[k] = U ? U[k] : nullptr;
  }
  const int nb = (int)((n + LIN_BLK - 1) / LIN_BLK);
  if (idir == 0)
    k_linear<0><<<nb, LIN_BLK, 0, s>>>(n, Lp, Rp, e, o, Fp, Up, pstar, mask, count);
  else if (idir == 1)
    k_linear<1><<<nb, LIN_BLK, 0, s>>>(n, Lp, Rp, e, o, Fp, Up, pstar, mask, count);
  else
    k_linear<2><<<nb, LIN_BLK, 0, s>>>(n, Lp, Rp, e, o, Fp, Up, pstar, mask, count);
  return launch_status();
}

int launch_linear_sweeps(int lim, const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e,
                         const LinOpts& o, const Flx& Fx, const Flx& Fy, int emf_hll,
                         srmhd_flux_diag* diag, unsigned long long* words, hipStream_t s) {
  k_zero4<<<1, 64, 0, s>>>(words);
  switch (lim) {
    case LIM_PCM: sweeps_linear<LIM_PCM>(g, S, P, e, o, Fx, Fy, emf_hll, words, s); break;
    case LIM_MC: sweeps_linear<LIM_MC>(g, S, P, e, o, Fx, Fy, emf_hll, words, s); break;
    case LIM_MINMOD: sweeps_linear<LIM_MINMOD>(g, S, P, e, o, Fx, Fy, emf_hll, words, s); break;
    case LIM_WENO5Z: sweeps_linear<LIM_WENO5Z>(g, S, P, e, o, Fx, Fy, emf_hll, words, s); break;
    case LIM_MP5: sweeps_linear<LIM_MP5>(g, S, P, e, o, Fx, Fy, emf_hll, words, s); break;
    default: sweeps_linear<LIM_MP7>(g, S, P, e, o, Fx, Fy, emf_hll, words, s);
  }
  k_linear_fd<<<1, 64, 0, s>>>(diag, words);
  return launch_status();
}

}  // namespace srmhd

extern "C" {

int64_t srmhd_linear_workspace_bytes(int64_t n) {
  return srmhd_flux_workspace_bytes(SRMHD_HLLD, n);
}

int srmhd_linear_flux(int idir, int64_t n, const double* const* primL,
                      const double* const* primR, const srmhd_eos* eos,
                      const srmhd_linear_opts* opts, double* const* F, double* const* U,
                      double* pstar, uint8_t* linear_mask, srmhd_flux_diag* diag,
                      unsigned long long* n_linear, void* work, int64_t work_bytes,
                      void* stream) {
  LinOpts o;
  int fb = SRMHD_HLLD;
  if (opts) {
    if (opts->fallback == SRMHD_LINEAR_FALLBACK_HLLE)
      fb = SRMHD_HLLE;
    else if (opts->fallback == SRMHD_LINEAR_FALLBACK_HLLC)
      fb = SRMHD_HLLC;
    else if (opts->fallback != SRMHD_LINEAR_FALLBACK_HLLD)
      return SRMHD_EARG;
    if (opts->fd_rel > 0.0) o.fd_rel = opts->fd_rel;
    if (opts->imag_tol > 0.0) o.imag_tol = opts->imag_tol;
    if (opts->cond_max > 0.0) o.cond_max = opts->cond_max;
    if (!(o.fd_rel < 1.0) || !(o.cond_max <= DBL_MAX)) return SRMHD_EARG;
  }
  if (work_bytes < srmhd_linear_workspace_bytes(n)) return SRMHD_EARG;
  // the fallback's answer on every face (it validates the common arguments)
  const int rc = srmhd_flux_ex(fb, idir, n, primL, primR, eos,
                               (opts && fb == SRMHD_HLLD) ? &opts->hlld : nullptr, F, U, pstar,
                               nullptr, diag, work, work_bytes, stream);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (n_linear) k_zero_count<<<1, 64, 0, s>>>(n_linear);
  return launch_linear_flat(idir, n, primL, primR, *eos, o, F, U, pstar, linear_mask, n_linear,
                            s);
}

}  // extern "C"
