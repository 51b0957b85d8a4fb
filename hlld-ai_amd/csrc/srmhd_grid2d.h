// 2D grid structures, BC index maps and the block-level reductions shared by
// the FP64 and the mixed-precision kernel translation units.
#pragma once

#include "srmhd_launch.h"
#include "srmhd_recon.h"
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

// The counters of a 2D sweep are accumulated in DIAG_SHARDS copies, one
// 128-byte line each, picked by block index: every block ends with up to ten
// atomics, and ~16 k blocks on one line serialise in the L2 (the sweep whose
// faces mostly have B_n = 0 ran 2.4x longer on them).  k_fold_diag folds
// the copies into the srmhd_flux_diag the downstream kernels read.
constexpr int DIAG_SHARDS = 64;
constexpr int DIAG_SHARD_BYTES = 128;
static_assert(sizeof(srmhd_flux_diag) <= DIAG_SHARD_BYTES, "a shard holds one diag");

// this block's copy among the shards starting at base
__device__ inline srmhd_flux_diag* diag_shard(srmhd_flux_diag* base) {
  return (srmhd_flux_diag*)((char*)base +
                            (int64_t)DIAG_SHARD_BYTES * (blockIdx.x & (DIAG_SHARDS - 1)));
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
__host__ __device__ inline int64_t n_faces(const G2& g) {
  return N == 0 ? (int64_t)(g.nxt + 1) * g.nyt : (int64_t)g.nxt * (g.nyt + 1);
}

// Reconstructed face states of face (i, j) of the sweep along N, as
// reconstruct_prims (reconstruction.py:272-338, vel_var "Wv") followed by
// B_n := face value (sweep.py:82-89).
template <int N, int LIM>
__device__ void face_prims2d(const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e, int i,
                             int j, Prim& L, Prim& R) {
  constexpr int w = stencil_halfwidth(LIM);
  constexpr int k0 = 3 - w, k1 = 4 + w;
  const int n = N == 0 ? g.nxt : g.nyt;
  const int f = N == 0 ? i : j;
  auto cidx = [&](int k) -> int64_t {
    const int c = clampi(f + k - 4, 0, n - 1);
    return N == 0 ? (int64_t)c * g.nyt + j : (int64_t)i * g.nyt + c;
  };
  auto ccoord = [&](int k, int& ci, int& cj) {
    const int c = clampi(f + k - 4, 0, n - 1);
    ci = N == 0 ? c : i;
    cj = N == 0 ? j : c;
  };
  double q[8];
  double ql[7], qr[7];   // rho, p, zx, zy, zz, Bt1, Bt2

  // rho and p first: they decide the PCM fallback of every variable
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    const double* a = P.p[v == 0 ? 0 : 4];
#pragma unroll
    for (int k = k0; k <= k1; ++k) q[k] = a[cidx(k)];
    face_states<LIM>(q, f, n, ql[v], qr[v]);
  }
  const bool bad = (ql[0] <= 1e-12) || (qr[0] <= 1e-12) || (ql[1] <= 1e-14) || (qr[1] <= 1e-14);
  if (bad) {
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const double* a = P.p[v == 0 ? 0 : 4];
      ql[v] = a[cidx(f == 0 ? 4 : 3)] + 0.5 * 0.0;
      qr[v] = a[cidx(f == n ? 3 : 4)] - 0.5 * 0.0;
    }
  }
  // W v with W from clamp(1 - v^2, 1e-14)
  double W[8];
#pragma unroll
  for (int k = k0; k <= k1; ++k) {
    const int64_t c = cidx(k);
    const double vx = P.p[1][c], vy = P.p[2][c], vz = P.p[3][c];
    W[k] = 1.0 / csqrt(clo(1.0 - (sq(vx) + sq(vy) + sq(vz)), 1e-14));
  }
  // 5 more variables: zx, zy, zz, B_t1, B_t2
#pragma unroll
  for (int v = 2; v < 7; ++v) {
#pragma unroll
    for (int k = k0; k <= k1; ++k) {
      double val;
      if (v <= 4) {
        val = W[k] * P.p[v - 1][cidx(k)];
      } else {
        // N=0: t1 = y (faces), t2 = z (cons);  N=1: t1 = z (cons), t2 = x (faces)
        const bool fromBz = (N == 0) ? (v == 6) : (v == 5);
        if (fromBz) {
          val = S.p[5][cidx(k)];
        } else {
          int ci, cj;
          ccoord(k, ci, cj);
          val = (N == 0) ? bcy(g, S.p[7], ci, cj) : bcx(g, S.p[6], ci, cj);
        }
      }
      q[k] = val;
    }
    if (bad) {
      // PCM: the window slots of cells max(f-1, 0) and min(f, n-1)
      ql[v] = (f == 0 ? q[4] : q[3]) + 0.5 * 0.0;
      qr[v] = (f == n ? q[3] : q[4]) - 0.5 * 0.0;
    } else {
      face_states<LIM>(q, f, n, ql[v], qr[v]);
    }
  }
  constexpr int T1 = (N + 1) % 3, T2 = (N + 2) % 3;
  const double Bn = N == 0 ? S.p[6][(int64_t)i * g.nyt + j] : S.p[7][(int64_t)i * (g.nyt + 1) + j];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const double* x = s ? qr : ql;
    Prim& Q = s ? R : L;
    const double z2 = sq(x[2]) + sq(x[3]) + sq(x[4]);
    const double fac = 1.0 / csqrt(1.0 + z2);
    Q.v[0] = x[2] * fac;
    Q.v[1] = x[3] * fac;
    Q.v[2] = x[4] * fac;
    Q.rho = clo(x[0], 1e-12);
    Q.p = clo(x[1], 1e-14);
    Q.eps = eos_eps(e, Q.p, Q.rho);
    Q.B[N] = Bn;
    Q.B[T1] = x[5];
    Q.B[T2] = x[6];
  }
}

}  // namespace srmhd
