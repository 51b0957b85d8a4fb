// 2D grid structures, BC index maps and the block-level reductions shared by
// the FP64 and the mixed-precision kernel translation units.
#pragma once

#include "srmhd_launch.h"
#include "srmhd_riemann.h"

namespace srmhd {

// Block-level combination of per-wave partials: lane 0 of every wave parks
// its values in LDS, thread 0 folds them and issues one atomic per value.
// Every thread of the block must call these (no early return before).
__device__ inline void block_or_atomic(unsigned long long* dst, unsigned long long v) {
  __shared__ unsigned long long sh[BLK / 64];
  v = wave_or(v);
  if (lane0()) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
#pragma unroll
    for (int w = 0; w < BLK / 64; ++w) t |= sh[w];
    if (t) atomicOr(dst, t);
  }
}

__device__ inline void block_max_atomic2(unsigned long long* dst0, unsigned long long v0,
                                         unsigned long long* dst1, unsigned long long v1) {
  __shared__ unsigned long long sh[2][BLK / 64];
  v0 = wave_max(v0);
  v1 = wave_max(v1);
  if (lane0()) {
    sh[0][threadIdx.x >> 6] = v0;
    sh[1][threadIdx.x >> 6] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0;
#pragma unroll
    for (int w = 0; w < BLK / 64; ++w) {
      a = sh[0][w] > a ? sh[0][w] : a;
      b = sh[1][w] > b ? sh[1][w] : b;
    }
    if (a) atomicMax(dst0, a);
    if (b) atomicMax(dst1, b);
  }
}

// Per-face counters of one flux call (hlld.py:1516-1529) plus the max |F_D|
// of the sweep, block-aggregated.
__device__ void diag_accumulate(srmhd_flux_diag* d, bool active, int st, int iters,
                                unsigned long long absFD_bits) {
  constexpr int NS = 8;
  unsigned long long v[NS + 2];
  v[0] = __popcll(__ballot(active));
  v[1] = __popcll(__ballot(active && (st & FS_NOT_CONVERGED)));
  v[2] = __popcll(__ballot(active && (st & FS_WAVE_ORDER)));
  v[3] = __popcll(__ballot(active && (st & FS_BN_ZERO)));
  v[4] = __popcll(__ballot(active && (st & FS_FALLBACK)));
  const bool conv = active && !(st & FS_NOT_CONVERGED);
  v[5] = __popcll(__ballot(conv));
  v[6] = wave_sum(conv ? (unsigned long long)iters : 0ull);
  v[7] = __popcll(__ballot(active && (st & FS_C2P_REDONE)));
  v[NS] = wave_max(active ? (unsigned long long)iters : 0ull);
  v[NS + 1] = wave_max(active ? absFD_bits : 0ull);
  __shared__ unsigned long long sh[NS + 2][BLK / 64];
  if (lane0()) {
#pragma unroll
    for (int k = 0; k < NS + 2; ++k) sh[k][threadIdx.x >> 6] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < NS + 2) {
    const int k = threadIdx.x;
    unsigned long long t = 0;
    for (int w = 0; w < BLK / 64; ++w) t = (k < NS) ? t + sh[k][w] : (sh[k][w] > t ? sh[k][w] : t);
    if (t) {
      switch (k) {
        case 0: atomicAdd(&d->n_interfaces, t); break;
        case 1: atomicAdd(&d->n_not_converged, t); break;
        case 2: atomicAdd(&d->n_wave_order_bad, t); break;
        case 3: atomicAdd(&d->n_bn_zero, t); break;
        case 4: atomicAdd(&d->n_hlle_fallback, t); break;
        case 5: atomicAdd(&d->n_converged, t); break;
        case 6: atomicAdd(&d->sum_iters_converged, t); break;
        case 7: atomicAdd(&d->n_c2p_redo, t); break;
        case 8: atomicMax(&d->max_iters, t); break;
        default: atomicMax(&d->absF_D_max, t); break;
      }
    }
  }
}

// ── 2D ──────────────────────────────────────────────────────────────────────

struct G2 {
  int nx, ny, ng, nxt, nyt, bcx, bcy;
  double dx, dy;
};

struct St2 {  // D, Sx, Sy, Sz, tau, Bz, Bxf, Byf
  const double* p[8];
};
struct St2Out {
  double* p[8];
};
struct Pr2 {  // rho, vx, vy, vz, p, eps
  const double* p[6];
};
struct Pr2Out {
  double* p[6];
};
struct Flx {  // D, Sx, Sy, Sz, tau, Bz, E-component
  double* p[7];
};

// BC source index of a cell-centred axis (grid.py:93-113)
__device__ inline int src_cell(int i, int n, int ng, int bc) {
  if (i < ng) return bc == SRMHD_BC_PERIODIC ? i + n : ng;
  if (i >= ng + n) return bc == SRMHD_BC_PERIODIC ? i - n : ng + n - 1;
  return i;
}
// BC source index of a face-staggered axis along its normal (grid.py:116-143)
__device__ inline int src_face(int f, int n, int ng, int bc) {
  if (bc == SRMHD_BC_PERIODIC) {
    if (f < ng) return f + n;
    if (f == ng + n) return ng;
    if (f > ng + n) return f - n;
    return f;
  }
  if (f < ng) return ng;
  if (f > ng + n) return ng + n;
  return f;
}


// cell-centred B from the staggered field (ct.py:44-54)
__device__ inline double bcx(const G2& g, const double* Bxf, int i, int j) {
  return 0.5 * (Bxf[(int64_t)i * g.nyt + j] + Bxf[(int64_t)(i + 1) * g.nyt + j]);
}
__device__ inline double bcy(const G2& g, const double* Byf, int i, int j) {
  return 0.5 * (Byf[(int64_t)i * (g.nyt + 1) + j] + Byf[(int64_t)i * (g.nyt + 1) + j + 1]);
}

// con2prim input of cell (i, j): cons with B from the faces (driver2d.py:208-212)
__device__ inline void cell_cons(const G2& g, const St2& S, int i, int j, double u[8]) {
  const int64_t c = (int64_t)i * g.nyt + j;
  u[CD] = S.p[0][c];
  u[CS] = S.p[1][c];
  u[CS + 1] = S.p[2][c];
  u[CS + 2] = S.p[3][c];
  u[CT] = S.p[4][c];
  u[CB] = bcx(g, S.p[6], i, j);
  u[CB + 1] = bcy(g, S.p[7], i, j);
  u[CB + 2] = S.p[5][c];
}

template <int N>
__device__ inline void face_ij(const G2& g, int64_t t, int& i, int& j) {
  const int ny = N == 0 ? g.nyt : g.nyt + 1;
  i = (int)(t / ny);
  j = (int)(t - (int64_t)i * ny);
}

template <int N>
__device__ inline int64_t n_faces(const G2& g) {
  return N == 0 ? (int64_t)(g.nxt + 1) * g.nyt : (int64_t)g.nxt * (g.nyt + 1);
}

}  // namespace srmhd
