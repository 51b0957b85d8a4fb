// Batched 1D Riemann tubes on gfx950 (see include/srmhd.h, srmhd_tubes_*): the
// kernels of the limiter SRMHD_TUBE_LIM (compiled once per limiter).
//
// Launch shape: one workgroup per tube (64..512 threads, a tunable), grid =
// number of tubes, dispatched through an order table (longest tube first).
// Ragged layout: tube t's cells are [off_c[t], off_c[t+1]) (L_t physical + 2
// ng ghosts), its faces [off_c[t] + t, off_c[t+1] + t + 1).  Evolved keys D,
// Sx, Sy, Sz, tau, By, Bz; Bx = B_n is constant per tube; primitives rho, vx,
// vy, vz, p, eps.
//
// Two forms share the per-phase device code:
//  * per-tube dt (srmhd_tubes_run): a tube is one con2prim batch and has its
//    own dt, so every reduction is block-level and any number of steps is ONE
//    launch (k_tubes);
//  * shared dt (srmhd_tubes_shared_step, tube_seed.py:93-144): one dt and one
//    con2prim batch over all tubes, so the batch masks and the dt max are
//    grid-wide and a step is a short chain of launches (k_ts_*).
#include "srmhd_recon.h"
#include "srmhd_riemann.h"
#include "srmhd_tubes.h"

using namespace srmhd;

#ifndef SRMHD_TUBE_LIM
#error "compile with -DSRMHD_TUBE_LIM=<limiter id>"
#endif

namespace {

// The tube a workgroup owns, with its pointers offset to the tube's start.
struct Tube {
  int id, L, nt, g;
  int64_t oc;
  double dx, Bn;
  double* P[6];
  double* F[7];
  C2PRecK* recf;
  C2PRecK* recc;
  __device__ Tube(const TubeArgs& A) {
    id = A.order ? A.order[blockIdx.x] : (int)blockIdx.x;
    oc = A.off_c[id];
    L = A.ncells[id];
    nt = (int)(A.off_c[id + 1] - oc);
    g = (nt - L) / 2;
    dx = A.dx[id];
    Bn = A.Bn[id];
#pragma unroll
    for (int k = 0; k < 6; ++k) P[k] = A.P[k] + oc;
#pragma unroll
    for (int k = 0; k < 7; ++k) F[k] = A.F[k] + oc + id;
    recf = A.recf + oc + id;
    recc = A.recc + oc;
  }
  __device__ double* U(const TubeArgs& A, int b, int k) const { return A.U[b][k] + oc; }
};

template <int NT>
__device__ inline uint64_t tblock_or(uint64_t v, uint64_t* sh) {
  v = wave_or(v);
  __syncthreads();
  if (lane0()) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  uint64_t t = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t |= sh[w];
  return t;
}

template <int NT>
__device__ inline uint64_t tblock_max(uint64_t v, uint64_t* sh) {
  v = wave_max(v);
  __syncthreads();
  if (lane0()) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  uint64_t t = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t = sh[w] > t ? sh[w] : t;
  return t;
}

// reconstruct_prims along the tube (reconstruction.py:272-338, vel_var "Wv")
// with L[Bx] = R[Bx] = B_n (tube_seed.py:81-84); the same per-face algebra as
// the 2D sweeps' face states.
template <int LIM>
__device__ void tube_face(const double* const* P, const double* By, const double* Bz, double Bn,
                          int f, int n, const srmhd_eos& e, Prim& L, Prim& R) {
  constexpr int w = stencil_halfwidth(LIM);
  constexpr int k0 = 3 - w, k1 = 4 + w;
  auto ci = [&](int k) { return clampi(f + k - 4, 0, n - 1); };
  double q[8], ql[7], qr[7];   // rho, p, zx, zy, zz, By, Bz
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    const double* a = P[v == 0 ? 0 : 4];
#pragma unroll
    for (int k = k0; k <= k1; ++k) q[k] = a[ci(k)];
    face_states<LIM>(q, f, n, ql[v], qr[v]);
  }
  const bool bad = (ql[0] <= 1e-12) || (qr[0] <= 1e-12) || (ql[1] <= 1e-14) || (qr[1] <= 1e-14);
  if (bad) {
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const double* a = P[v == 0 ? 0 : 4];
      ql[v] = a[ci(f == 0 ? 4 : 3)] + 0.5 * 0.0;
      qr[v] = a[ci(f == n ? 3 : 4)] - 0.5 * 0.0;
    }
  }
  double W[8];
#pragma unroll
  for (int k = k0; k <= k1; ++k) {
    const int c = ci(k);
    W[k] = 1.0 / csqrt(clo(1.0 - (sq(P[1][c]) + sq(P[2][c]) + sq(P[3][c])), 1e-14));
  }
#pragma unroll
  for (int v = 2; v < 7; ++v) {
#pragma unroll
    for (int k = k0; k <= k1; ++k) {
      const int c = ci(k);
      q[k] = v <= 4 ? W[k] * P[v - 1][c] : (v == 5 ? By[c] : Bz[c]);
    }
    if (bad) {
      ql[v] = (f == 0 ? q[4] : q[3]) + 0.5 * 0.0;
      qr[v] = (f == n ? q[3] : q[4]) - 0.5 * 0.0;
    } else {
      face_states<LIM>(q, f, n, ql[v], qr[v]);
    }
  }
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
    Q.B[0] = Bn;
    Q.B[1] = x[5];
    Q.B[2] = x[6];
  }
}

// ── phases (every thread of the workgroup calls each) ───────────────────────

// HLLD pass A over the tube's faces; returns this thread's partial mask.
template <int LIM, int NT>
__device__ uint64_t ph_flux_a(const Tube& T, const double* By, const double* Bz,
                              const srmhd_eos& e) {
  uint64_t m = 0;
  const double* const* Pc = T.P;
  for (int f = threadIdx.x; f <= T.nt; f += NT) {
    Prim l, r;
    tube_face<LIM>(Pc, By, Bz, T.Bn, f, T.nt, e, l, r);
    HlldPrep<0> h;
    h.run(l, r, e);
    Kastaun k;
    k.init(h.uH, e);
    m |= c2p_pass_a(k, T.recf[f]);
  }
  return m;
}

// The flux on every face of the tube into T.F (D, Sx, Sy, Sz, tau, By, Bz).
// frow (shared form only, else NULL): the face rows whose HLL-state
// con2prim is redone by bisection.
template <int LIM, int SOL, int NT>
__device__ void ph_flux_b(const Tube& T, const double* By, const double* Bz, const srmhd_eos& e,
                          uint64_t mf, const unsigned* frow) {
  const double* const* Pc = T.P;
  const int K = c2p_break_iter(mf);
  for (int f = threadIdx.x; f <= T.nt; f += NT) {
    Prim l, r;
    tube_face<LIM>(Pc, By, Bz, T.Bn, f, T.nt, e, l, r);
    double fl[8], u[8], ps;
    if (SOL == SOLVER_HLLE) {
      hlle<0>(l, r, e, fl, u, ps);
    } else if (SOL == SOLVER_HLLC) {
      hllc<0>(l, r, e, fl, u, ps);
    } else {
      HlldPrep<0> h;
      h.run(l, r, e);
      Kastaun k;
      k.init(h.uH, e);
      Prim ph;
      c2p_pass_b(k, T.recf[f], K, 1e-10, 1e-10, ph, frow ? (int)frow[f] : -1);
      int it;
      hlld_finish<0>(h, l, r, p_hll_from_prim(ph, e), fl, nullptr, ps, it);
    }
    T.F[0][f] = fl[CD];
    T.F[1][f] = fl[CS];
    T.F[2][f] = fl[CS + 1];
    T.F[3][f] = fl[CS + 2];
    T.F[4][f] = fl[CT];
    T.F[5][f] = fl[CB + 1];
    T.F[6][f] = fl[CB + 2];
  }
}

// Stage combination on every cell with L = -(F[j+1] - F[j]) / dx
// (tube_seed.py:90).  form 0: a U0 + b U + wdt L (state.combine); form 1
// (Heun, tube_seed.py:134-139): stage 0 U0 + dt L, stage 1 0.5 (U0 + U + dt L).
// Element-wise, so Uo may alias U0.
template <int NT>
__device__ void ph_update(const Tube& T, const TubeArgs& A, int b0, int bu, int bo, int form,
                          int s, double a, double b, double wdt, double dt) {
  for (int c = threadIdx.x; c < T.nt; c += NT) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const double rhs = -(T.F[k][c + 1] - T.F[k][c]) / T.dx;
      const double u0 = T.U(A, b0, k)[c];
      const double uu = T.U(A, bu, k)[c];
      double v;
      if (form == 0) {
        v = a * u0 + b * uu + wdt * rhs;
      } else if (s == 0) {
        v = u0 + dt * rhs;
      } else {
        v = 0.5 * (u0 + uu + dt * rhs);
      }
      T.U(A, bo, k)[c] = v;
    }
  }
}

__device__ inline void tube_cons(const Tube& T, const TubeArgs& A, int bo, int c, double u[8]) {
  u[CD] = T.U(A, bo, 0)[c];
  u[CT] = T.U(A, bo, 4)[c];
  u[CS] = T.U(A, bo, 1)[c];
  u[CS + 1] = T.U(A, bo, 2)[c];
  u[CS + 2] = T.U(A, bo, 3)[c];
  u[CB] = T.Bn;
  u[CB + 1] = T.U(A, bo, 5)[c];
  u[CB + 2] = T.U(A, bo, 6)[c];
}

// con2prim pass A of every cell of the tube (tube_seed.py:102-109).
template <int NT>
__device__ uint64_t ph_c2p_a(const Tube& T, const TubeArgs& A, int bo, const srmhd_eos& e) {
  uint64_t m = 0;
  for (int c = threadIdx.x; c < T.nt; c += NT) {
    double u[8];
    tube_cons(T, A, bo, c, u);
    Kastaun k;
    k.init(u, e);
    m |= c2p_pass_a(k, T.recc[c]);
  }
  return m;
}

// Pass B on the physical cells (the ghosts are overwritten next); crow as
// frow above, for the cell rows.
template <int NT>
__device__ void ph_c2p_b(const Tube& T, const TubeArgs& A, int bo, const srmhd_eos& e,
                         uint64_t mc, double atmo_rho, const unsigned* crow) {
  const int K = c2p_break_iter(mc);
  for (int c = T.g + threadIdx.x; c < T.g + T.L; c += NT) {
    double u[8];
    tube_cons(T, A, bo, c, u);
    Kastaun k;
    k.init(u, e);
    Prim q;
    c2p_pass_b(k, T.recc[c], K, atmo_rho, 1e-10, q, crow ? (int)crow[c] : -1);
    T.P[0][c] = q.rho;
    T.P[1][c] = q.v[0];
    T.P[2][c] = q.v[1];
    T.P[3][c] = q.v[2];
    T.P[4][c] = q.p;
    T.P[5][c] = q.eps;
  }
}

// Two-cell-copy ghosts (tube_seed.py:72-75).  They also reach the conserved
// By, Bz: the reference's primitive dict holds the same tensors
// (tube_seed.py:105-108).
template <int NT>
__device__ void ph_ghosts(const Tube& T, const TubeArgs& A, int bo) {
  for (int c = threadIdx.x; c < 2 * T.g; c += NT) {
    const int dst = c < T.g ? c : T.L + c;
    const int src = c < T.g ? T.g : T.g + T.L - 1;
#pragma unroll
    for (int k = 0; k < 6; ++k) T.P[k][dst] = T.P[k][src];
    T.U(A, bo, 5)[dst] = T.U(A, bo, 5)[src];
    T.U(A, bo, 6)[dst] = T.U(A, bo, 6)[src];
  }
}

// Per-tube dt: compute_dt (driver.py:77-127) over the physical cells, torch's
// float / tensor being reciprocal(tensor) * float.
template <int NT>
__device__ double ph_dt(const Tube& T, const TubeArgs& A, double cfl, const srmhd_eos& e,
                        uint64_t* sh, double* sdt) {
  const double* By = T.U(A, 0, 5);
  const double* Bz = T.U(A, 0, 6);
  const double Bn = T.Bn;
  uint64_t m = 0;
  for (int c = T.g + threadIdx.x; c < T.g + T.L; c += NT) {
    const double rho = T.P[0][c], vx = T.P[1][c], vy = T.P[2][c], vz = T.P[3][c];
    const double eps = T.P[5][c];
    double p, cs2;
    eos_press_cs2(e, eps, rho, p, cs2);
    const double v2 = sq(vx) + sq(vy) + sq(vz);
    const double W = 1.0 / csqrt(clo(1.0 - v2, 1e-10));
    const double B2 = sq(Bn) + sq(By[c]) + sq(Bz[c]);
    const double vB = vx * Bn + vy * By[c] + vz * Bz[c];
    const double b2 = (B2 + vB * vB) / (W * W);
    const double h = 1.0 + eps + p / rho;
    const double v02 = cs2 + (b2 / (rho * h + b2)) * (1.0 - cs2);
    const double vf = csqrt(chi_(clo(v02, 0.0), 1.0));
    const uint64_t a = dbits(fabs(vx) + vf) & 0x7fffffffffffffffull;
    m = a > m ? a : m;
  }
  m = tblock_max<NT>(m, sh);
  if (threadIdx.x == 0) *sdt = (1.0 / clo(__longlong_as_double(m), 1e-14)) * (cfl * T.dx);
  __syncthreads();
  return *sdt;
}

// SSP weights (driver2d.py:104-105)
__host__ __device__ inline void rk_weights(int rk3, int s, double& a, double& b, double& c) {
  if (rk3) {
    a = s == 0 ? 1.0 : (s == 1 ? 0.75 : 1.0 / 3.0);
    b = s == 0 ? 0.0 : (s == 1 ? 0.25 : 2.0 / 3.0);
    c = s == 0 ? 1.0 : (s == 1 ? 0.25 : 2.0 / 3.0);
  } else {
    a = s == 0 ? 1.0 : 0.5;
    b = s == 0 ? 0.0 : 0.5;
    c = s == 0 ? 1.0 : 0.5;
  }
}

// ── per-tube dt: every step of every tube in one launch ─────────────────────
// Buffer roles: 0 is the step-start state and also receives the last stage
// (the combination is element-wise); stage s < last writes buffer 1 + s % 2.
template <int LIM, int SOL, int NT>
__global__ __launch_bounds__(NT) void k_tubes(TubeArgs A, srmhd_eos e, int steps, double cfl,
                                              int rk3, double atmo_rho) {
  __shared__ uint64_t sh[NT / 64];
  __shared__ double sdt;
  const Tube T(A);
  const int nst = rk3 ? 3 : 2;
  double t = 0.0;
#pragma unroll 1
  for (int step = 0; step < steps; ++step) {
    const double dt = ph_dt<NT>(T, A, cfl, e, sh, &sdt);
    int bu = 0;
#pragma unroll 1
    for (int s = 0; s < nst; ++s) {
      const int bo = (s == nst - 1) ? 0 : 1 + (s % 2);
      const double* By = T.U(A, bu, 5);
      const double* Bz = T.U(A, bu, 6);
      uint64_t mf = 0;
      if (SOL == SOLVER_HLLD) mf = tblock_or<NT>(ph_flux_a<LIM, NT>(T, By, Bz, e), sh);
      ph_flux_b<LIM, SOL, NT>(T, By, Bz, e, mf, nullptr);
      __syncthreads();
      double a, b, c;
      rk_weights(rk3, s, a, b, c);
      ph_update<NT>(T, A, 0, bu, bo, 0, s, a, b, c * dt, dt);
      __syncthreads();
      const uint64_t mc = tblock_or<NT>(ph_c2p_a<NT>(T, A, bo, e), sh);
      ph_c2p_b<NT>(T, A, bo, e, mc, atmo_rho, nullptr);
      __syncthreads();
      ph_ghosts<NT>(T, A, bo);
      __syncthreads();
      bu = bo;
    }
    t += dt;
  }
  if (threadIdx.x == 0) A.t_out[T.id] += t;
}

// ── per-tube dt, phased: one grid-wide kernel per phase ─────────────────────
// The same per-element arithmetic and the same batches (each tube's faces,
// each tube's cells) as k_tubes, but every phase is its own launch over all
// tubes' elements at once: each kernel keeps only its own registers live
// (k_tubes holds every phase's at once and runs one workgroup per CU), and
// no workgroup waits at a barrier for its slowest wave.  A tube's con2prim
// masks and its dt live in per-tube words between the launches.  The
// elements whose con2prim takes the bisection redo are finished by their own
// launches from work lists.  A step is 1 + 6 x stages launches with HLLD
// (1 + 4 x stages otherwise); a run is one stream of them (one graph when
// captured).

// the tube of global cell i (off_c[t] <= i < off_c[t+1]) or global face i
// (off_c[t] + t <= i < off_c[t+1] + t + 1): binary search of the offsets
template <bool FACE>
__device__ inline int tube_of(const int64_t* off_c, int n, int64_t i) {
  int lo = 0, hi = n - 1;
#pragma unroll 1
  for (int it = 0; it < 32 && lo < hi; ++it) {
    const int mid = (lo + hi + 1) >> 1;
    if (off_c[mid] + (FACE ? mid : 0) <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// OR of m into dst[tube] for the active lanes of the wave: one atomic per
// distinct tube among them (a wave spans one tube or a few).  Every lane of
// the wave calls it.
__device__ inline void tube_or(unsigned long long* dst, int tube, unsigned long long m,
                               bool active) {
  unsigned long long todo = __ballot(active);
#pragma unroll 1
  for (int it = 0; it < 64 && todo; ++it) {
    const int leader = __builtin_ctzll(todo);
    const int lt = __shfl(tube, leader);
    const bool mine = active && tube == lt;
    const unsigned long long mm = wave_or(mine ? m : 0ull);
    const unsigned long long who = __ballot(mine);
    if ((int)(threadIdx.x & 63) == leader && mm) atomicOr(&dst[lt], mm);
    todo &= ~who;
  }
}

__global__ __launch_bounds__(BLK) void k_tp_init(TubeState S, int n) {
  const int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= n) return;
  S.mf[t] = 0ull;
  S.mc[t] = 0ull;
  S.tacc[t] = 0.0;
}

// per-tube dt (one workgroup per tube) and the tube's clock
template <int NT>
__global__ __launch_bounds__(NT) void k_tp_dt(TubeArgs A, TubeState S, srmhd_eos e, double cfl) {
  __shared__ uint64_t sh[NT / 64];
  __shared__ double sdt;
  const Tube T(A);
  const double dt = ph_dt<NT>(T, A, cfl, e, sh, &sdt);
  if (threadIdx.x == 0) {
    S.dt[T.id] = dt;
    S.tacc[T.id] = S.tacc[T.id] + dt;
  }
}

__device__ inline void tube_prims(const TubeArgs& A, int64_t oc, const double* (&P)[6]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) P[k] = A.P[k] + oc;
}

// HLLD pass A on every face of every tube
template <int LIM>
__global__ __launch_bounds__(BLK) void k_tp_flux_a(TubeArgs A, TubeState S, int n, int64_t nf,
                                                   int bu, srmhd_eos e) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i == 0) S.qn[0] = 0ull;   // the face work list of this stage starts empty
  const bool active = i < nf;
  int tb = 0;
  uint64_t m = 0;
  if (active) {
    tb = tube_of<true>(A.off_c, n, i);
    const int64_t oc = A.off_c[tb];
    const int nt = (int)(A.off_c[tb + 1] - oc);
    const int f = (int)(i - oc - tb);
    const double* P[6];
    tube_prims(A, oc, P);
    Prim l, r;
    tube_face<LIM>(P, A.U[bu][5] + oc, A.U[bu][6] + oc, A.Bn[tb], f, nt, e, l, r);
    HlldPrep<0> h;
    h.run(l, r, e);
    side_q_store<0>(A.hq, nf, i, h);   // for pass B and the redo
    Kastaun k;
    k.init(h.uH, e);
    m = c2p_pass_a(k, A.recf[i]);
  }
  tube_or(S.mf, tb, m, active);
}

__device__ inline void put_flux(const TubeArgs& A, int64_t i, const double fl[8]) {
  A.F[0][i] = fl[CD];
  A.F[1][i] = fl[CS];
  A.F[2][i] = fl[CS + 1];
  A.F[3][i] = fl[CS + 2];
  A.F[4][i] = fl[CT];
  A.F[5][i] = fl[CB + 1];
  A.F[6][i] = fl[CB + 2];
}

// the tube and the face states of global face i
template <int LIM>
__device__ inline int tp_face(const TubeArgs& A, int n, int64_t i, int bu, const srmhd_eos& e,
                              Prim& l, Prim& r, int& f) {
  const int tb = tube_of<true>(A.off_c, n, i);
  const int64_t oc = A.off_c[tb];
  const int nt = (int)(A.off_c[tb + 1] - oc);
  f = (int)(i - oc - tb);
  const double* P[6];
  tube_prims(A, oc, P);
  tube_face<LIM>(P, A.U[bu][5] + oc, A.U[bu][6] + oc, A.Bn[tb], f, nt, e, l, r);
  return tb;
}

// the flux on every face of every tube; the first face of a tube clears the
// tube's cell mask for this stage.  HLLD faces whose HLL-state con2prim takes
// the bisection redo (unconverged after the batch's K secant steps, with a
// good bracket) go to the face work list instead: up to half of a tube's
// elements take it, and a wave would otherwise run the ~55-step bisection
// for its few lanes that need it.
template <int LIM, int SOL>
__global__ __launch_bounds__(BLK) void k_tp_flux_b(TubeArgs A, TubeState S, int n, int64_t nf,
                                                   int bu, srmhd_eos e) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= nf) return;
  Prim l{}, r{};   // (HLLD: not formed; its solver reads them not)
  int f, tb;
  if (SOL == SOLVER_HLLD) {   // the side states pass A kept
    tb = tube_of<true>(A.off_c, n, i);
    f = (int)(i - A.off_c[tb] - tb);
  } else {
    tb = tp_face<LIM>(A, n, i, bu, e, l, r, f);
  }
  if (f == 0) S.mc[tb] = 0ull;
  double fl[8], u[8], ps;
  if (SOL == SOLVER_HLLE) {
    hlle<0>(l, r, e, fl, u, ps);
  } else if (SOL == SOLVER_HLLC) {
    hllc<0>(l, r, e, fl, u, ps);
  } else {
    HlldPrep<0> h;
    side_q_load<0>(A.hq, nf, i, h);
    Kastaun k;
    k.init(h.uH, e);
    const C2PRecK& rec = A.recf[i];
    double x1, f1;
    c2p_secant_at(k, rec, c2p_break_iter(S.mf[tb]), x1, f1);
    const bool redo = !(fabs(f1) < C2P_TOL) && rec.mup_signed > 0.0;
    queue_push(S.qf, &S.qn[0], i, redo);
    if (redo) return;
    Prim ph;
    c2p_pass_at(k, rec, x1, f1, 1e-10, 1e-10, ph, 0);
    int it;
    hlld_finish<0>(h, l, r, p_hll_from_prim(ph, e), fl, nullptr, ps, it);
  }
  put_flux(A, i, fl);
}

// the HLLD flux of the faces on the work list (the bisection redo of their
// HLL-state con2prim, then the rest of the solver)
// (two waves per SIMD: the compiler's own choice is 258 registers, one wave)
template <int LIM>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(2)))
void k_tp_flux_redo(TubeArgs A, TubeState S, int n, int bu, srmhd_eos e) {
  const int64_t cnt = (int64_t)S.qn[0];
  for (int64_t q = (int64_t)blockIdx.x * BLK + threadIdx.x; q < cnt;
       q += (int64_t)gridDim.x * BLK) {
    const int64_t i = S.qf[q];
    Prim l{}, r{};   // not read by the solver
    HlldPrep<0> h;
    side_q_load<0>(A.hq, A.cells + n, i, h);   // kept by pass A
    Kastaun k;
    k.init(h.uH, e);
    const C2PRecK& rec = A.recf[i];
    Prim ph;
    c2p_pass_at(k, rec, rec.x1, rec.f1, 1e-10, 1e-10, ph, 1);
    double fl[8], ps;
    int it;
    hlld_finish<0>(h, l, r, p_hll_from_prim(ph, e), fl, nullptr, ps, it);
    put_flux(A, i, fl);
  }
}

// the stage combination a U0 + b U + (c dt) L of every cell, then its
// con2prim pass A
__global__ __launch_bounds__(BLK) void k_tp_update_c2pa(TubeArgs A, TubeState S, int n,
                                                        int64_t nc, int bu, int bo, double a,
                                                        double b, double cw, srmhd_eos e) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i == 0) S.qn[1] = 0ull;   // the cell work list of this stage starts empty
  const bool active = i < nc;
  int tb = 0;
  uint64_t m = 0;
  if (active) {
    tb = tube_of<false>(A.off_c, n, i);
    const double dx = A.dx[tb];
    const double wdt = cw * S.dt[tb];
    double v[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const double rhs = -(A.F[k][i + tb + 1] - A.F[k][i + tb]) / dx;
      v[k] = a * A.U[0][k][i] + b * A.U[bu][k][i] + wdt * rhs;
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) A.U[bo][k][i] = v[k];
    double u[8];
    u[CD] = v[0];
    u[CT] = v[4];
    u[CS] = v[1];
    u[CS + 1] = v[2];
    u[CS + 2] = v[3];
    u[CB] = A.Bn[tb];
    u[CB + 1] = v[5];
    u[CB + 2] = v[6];
    Kastaun k;
    k.init(u, e);
    m = c2p_pass_a(k, A.recc[i]);
  }
  tube_or(S.mc, tb, m, active);
}

// con2prim pass B of every physical cell; a ghost cell takes its source
// cell's (two-cell-copy ghosts, tube_seed.py:72-75: the same values, formed
// again) and copies its conserved By, Bz; the first cell of a tube clears
// the tube's face mask for the next stage
// the conserved state of global cell i's con2prim source cell (itself, or
// for a ghost the physical cell it copies) -> the tube; *si = that cell
__device__ inline int tp_cell(const TubeArgs& A, int n, int64_t i, int bo, double u[8],
                              int64_t* si, bool* ghost, bool* first) {
  const int tb = tube_of<false>(A.off_c, n, i);
  const int64_t oc = A.off_c[tb];
  const int nt = (int)(A.off_c[tb + 1] - oc);
  const int L = A.ncells[tb];
  const int g = (nt - L) / 2;
  const int c = (int)(i - oc);
  const int src = c < g ? g : (c >= g + L ? g + L - 1 : c);
  *si = oc + src;
  *ghost = src != c;
  *first = c == 0;
  u[CD] = A.U[bo][0][*si];
  u[CT] = A.U[bo][4][*si];
  u[CS] = A.U[bo][1][*si];
  u[CS + 1] = A.U[bo][2][*si];
  u[CS + 2] = A.U[bo][3][*si];
  u[CB] = A.Bn[tb];
  u[CB + 1] = A.U[bo][5][*si];
  u[CB + 2] = A.U[bo][6][*si];
  return tb;
}

__device__ inline void put_prims(const TubeArgs& A, int64_t i, const Prim& q) {
  A.P[0][i] = q.rho;
  A.P[1][i] = q.v[0];
  A.P[2][i] = q.v[1];
  A.P[3][i] = q.v[2];
  A.P[4][i] = q.p;
  A.P[5][i] = q.eps;
}

// Cells whose con2prim takes the bisection redo go to the cell work list
// (as the faces in k_tp_flux_b).
__global__ __launch_bounds__(BLK) void k_tp_c2p_b(TubeArgs A, TubeState S, int n, int64_t nc,
                                                  int bo, double atmo_rho, srmhd_eos e) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= nc) return;
  double u[8];
  int64_t si;
  bool ghost, first;
  const int tb = tp_cell(A, n, i, bo, u, &si, &ghost, &first);
  if (ghost) {
    A.U[bo][5][i] = u[CB + 1];
    A.U[bo][6][i] = u[CB + 2];
  }
  Kastaun k;
  k.init(u, e);
  const C2PRecK& rec = A.recc[si];
  double x1, f1;
  c2p_secant_at(k, rec, c2p_break_iter(S.mc[tb]), x1, f1);
  if (first) S.mf[tb] = 0ull;
  const bool redo = !(fabs(f1) < C2P_TOL) && rec.mup_signed > 0.0;
  queue_push(S.qc, &S.qn[1], i, redo);
  if (redo) return;
  Prim q;
  c2p_pass_at(k, rec, x1, f1, atmo_rho, 1e-10, q, 0);
  put_prims(A, i, q);
}

// con2prim of the cells on the work list (the bisection redo)
__global__ __launch_bounds__(BLK) void k_tp_c2p_redo(TubeArgs A, TubeState S, int n, int bo,
                                                     double atmo_rho, srmhd_eos e) {
  const int64_t cnt = (int64_t)S.qn[1];
  for (int64_t q = (int64_t)blockIdx.x * BLK + threadIdx.x; q < cnt;
       q += (int64_t)gridDim.x * BLK) {
    const int64_t i = S.qc[q];
    double u[8];
    int64_t si;
    bool ghost, first;
    tp_cell(A, n, i, bo, u, &si, &ghost, &first);
    Kastaun k;
    k.init(u, e);
    const C2PRecK& rec = A.recc[si];
    Prim p;
    c2p_pass_at(k, rec, rec.x1, rec.f1, atmo_rho, 1e-10, p, 1);
    put_prims(A, i, p);
  }
}

__global__ __launch_bounds__(BLK) void k_tp_final(TubeArgs A, TubeState S, int n) {
  const int t = blockIdx.x * BLK + threadIdx.x;
  if (t < n) A.t_out[t] += S.tacc[t];
}

// ── shared dt: one step = a chain of grid-wide phases ────────────────────────
//
// The reference evolves the tubes as the columns of (rows, N) arrays and hands
// those 2-D arrays to conservative_to_primitive (the cells) and to the HLLD
// flux (whose HLL-state con2prim sees the (rows + 1, N) faces).  Its bisection
// redo gathers `torch.nonzero(redo).reshape(-1)` (c2p.py:339-346): the row AND
// column index of every element that needs the redo, both used as ROW
// indices, and re-solves those whole rows by bisection.  The marks below
// reproduce that set: a flagged element (row r, tube c) marks rows r and c.
// A column index past the last row is an IndexError in the reference; here it
// is counted in Clock::index_errors and the host raises.



__device__ inline unsigned* face_marks(Clock* c, int rows, int s) {
  return (unsigned*)(c + 1) + s * (rows + 1);
}
__device__ inline unsigned* cell_marks(Clock* c, int rows, int s) {
  return (unsigned*)(c + 1) + 2 * (rows + 1) + s * rows;
}

// The ideal-gas bound |vx| + min(c_s + 0.5, 1) over every cell of the tube
// (ghosts included, tube_seed.py:125-130), maximised as nonnegative double
// bit patterns (NaN orders last, as torch's max propagates it).
template <int NT>
__global__ __launch_bounds__(NT) void k_ts_smax(TubeArgs A, double gamma, Clock* clk) {
  __shared__ uint64_t sh[NT / 64];
  const Tube T(A);
  uint64_t m = 0;
  for (int c = threadIdx.x; c < T.nt; c += NT) {
    const double rho = T.P[0][c], vx = T.P[1][c], p = T.P[4][c];
    const double cs = csqrt(clo((gamma * p) / clo(rho + (gamma * p) / (gamma - 1.0), 1e-30), 0.0));
    const uint64_t a = dbits(fabs(vx) + chi_(cs + 0.5, 1.0)) & 0x7fffffffffffffffull;
    m = a > m ? a : m;
  }
  m = tblock_max<NT>(m, sh);
  if (threadIdx.x == 0) atomicMax(&clk->smax, (unsigned long long)m);
}

// dt = min(cfl dx / max(smax, 1e-8), tend - t) with Python's max/min (the
// first argument wins unless the second compares strictly greater/less), the
// loop test t < tend - 1e-14, and the clock; clears the masks, the max and
// the row marks for this step.
__global__ __launch_bounds__(256) void k_ts_clock(Clock* clk, double cfl, double dx, double tend,
                                                 int rows) {
  unsigned* marks = (unsigned*)(clk + 1);
  for (int i = threadIdx.x; i < 4 * rows + 2; i += 256) marks[i] = 0u;
  if (threadIdx.x != 0) return;
  const double smax = __longlong_as_double(clk->smax);
  clk->smax = 0ull;
  clk->mf[0] = clk->mf[1] = clk->mc[0] = clk->mc[1] = 0ull;
  const double t = clk->t;
  if (!(t < tend - 1e-14)) {
    clk->active = 0.0;
    return;
  }
  const double s = (1e-8 > smax) ? 1e-8 : smax;
  const double x = cfl * dx / s;
  const double rem = tend - t;
  const double dt = (rem < x) ? rem : x;
  clk->dt = dt;
  clk->active = 1.0;
  clk->t = t + dt;
  clk->steps += 1.0;
}

__device__ inline void mark_rows(unsigned* marks, int r, int col, int nrows, Clock* clk) {
  marks[r] = 1u;
  if (col < nrows)
    marks[col] = 1u;
  else
    atomicAdd(&clk->index_errors, 1ull);
}

template <int LIM, int NT>
__global__ __launch_bounds__(NT) void k_ts_flux_a(TubeArgs A, srmhd_eos e, int s, Clock* clk) {
  __shared__ uint64_t sh[NT / 64];
  if (clk->active == 0.0) return;
  const Tube T(A);
  const int bu = s == 0 ? 0 : 1;
  const uint64_t m = tblock_or<NT>(ph_flux_a<LIM, NT>(T, T.U(A, bu, 5), T.U(A, bu, 6), e), sh);
  if (threadIdx.x == 0 && m) atomicOr(&clk->mf[s], (unsigned long long)m);
}

// the faces whose HLL-state con2prim needs the redo -> face-row marks
template <int LIM, int NT>
__global__ __launch_bounds__(NT) void k_ts_face_rows(TubeArgs A, srmhd_eos e, int s, int rows,
                                                     Clock* clk) {
  if (clk->active == 0.0) return;
  const Tube T(A);
  const int bu = s == 0 ? 0 : 1;
  const int K = c2p_break_iter(clk->mf[s]);
  unsigned* marks = face_marks(clk, rows, s);
  const double* const* Pc = T.P;
  for (int f = threadIdx.x; f <= T.nt; f += NT) {
    Prim l, r;
    tube_face<LIM>(Pc, T.U(A, bu, 5), T.U(A, bu, 6), T.Bn, f, T.nt, e, l, r);
    HlldPrep<0> h;
    h.run(l, r, e);
    Kastaun k;
    k.init(h.uH, e);
    if (c2p_needs_redo(k, T.recf[f], K)) mark_rows(marks, f, T.id, rows + 1, clk);
  }
}

// flux pass B, the Heun combination and con2prim pass A
template <int LIM, int SOL, int NT>
__global__ __launch_bounds__(NT) void k_ts_mid(TubeArgs A, srmhd_eos e, int s, int rows,
                                               Clock* clk) {
  __shared__ uint64_t sh[NT / 64];
  if (clk->active == 0.0) return;
  const Tube T(A);
  const int bu = s == 0 ? 0 : 1, bo = s == 0 ? 1 : 0;
  ph_flux_b<LIM, SOL, NT>(T, T.U(A, bu, 5), T.U(A, bu, 6), e, clk->mf[s],
                          SOL == SOLVER_HLLD ? face_marks(clk, rows, s) : nullptr);
  __syncthreads();
  ph_update<NT>(T, A, 0, bu, bo, 1, s, 0.0, 0.0, 0.0, clk->dt);
  __syncthreads();
  const uint64_t m = tblock_or<NT>(ph_c2p_a<NT>(T, A, bo, e), sh);
  if (threadIdx.x == 0 && m) atomicOr(&clk->mc[s], (unsigned long long)m);
}

// the cells whose con2prim needs the redo -> cell-row marks
template <int NT>
__global__ __launch_bounds__(NT) void k_ts_cell_rows(TubeArgs A, srmhd_eos e, int s, int rows,
                                                     Clock* clk) {
  if (clk->active == 0.0) return;
  const Tube T(A);
  const int bo = s == 0 ? 1 : 0;
  const int K = c2p_break_iter(clk->mc[s]);
  unsigned* marks = cell_marks(clk, rows, s);
  for (int c = threadIdx.x; c < T.nt; c += NT) {
    double u[8];
    tube_cons(T, A, bo, c, u);
    Kastaun k;
    k.init(u, e);
    if (c2p_needs_redo(k, T.recc[c], K)) mark_rows(marks, c, T.id, rows, clk);
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_ts_end(TubeArgs A, srmhd_eos e, int s, int rows,
                                               double atmo_rho, Clock* clk) {
  if (clk->active == 0.0) return;
  const Tube T(A);
  const int bo = s == 0 ? 1 : 0;
  ph_c2p_b<NT>(T, A, bo, e, clk->mc[s], atmo_rho, cell_marks(clk, rows, s));
  __syncthreads();
  ph_ghosts<NT>(T, A, bo);
}

// ── launch plumbing ─────────────────────────────────────────────────────────

template <int LIM, int SOL, int NT>
int run_nt(int n, const TubeArgs& A, const srmhd_eos& e, int steps, double cfl, int rk3,
           double atmo, hipStream_t st) {
  k_tubes<LIM, SOL, NT><<<n, NT, 0, st>>>(A, e, steps, cfl, rk3, atmo);
  return launch_status();
}

// workgroups of the work-list kernels (grid-stride over the list, whose
// length is known only on the device)
constexpr int REDO_BLOCKS = 2048;

template <int LIM, int SOL>
int run_phased(int n, const TubeArgs& A, const srmhd_eos& e, int steps, double cfl, int rk3,
               double atmo, hipStream_t st) {
  const TubeState& S = A.ts;
  const int64_t cells = A.cells;
  const int64_t faces = cells + n;
  const int bf = blocks_for(faces), bc = blocks_for(cells);
  const int nst = rk3 ? 3 : 2;
  k_tp_init<<<blocks_for(n), BLK, 0, st>>>(S, n);
  for (int step = 0; step < steps; ++step) {
    k_tp_dt<256><<<n, 256, 0, st>>>(A, S, e, cfl);
    int bu = 0;
    for (int s = 0; s < nst; ++s) {
      const int bo = (s == nst - 1) ? 0 : 1 + (s % 2);
      if (SOL == SOLVER_HLLD) k_tp_flux_a<LIM><<<bf, BLK, 0, st>>>(A, S, n, faces, bu, e);
      k_tp_flux_b<LIM, SOL><<<bf, BLK, 0, st>>>(A, S, n, faces, bu, e);
      if (SOL == SOLVER_HLLD)
        k_tp_flux_redo<LIM><<<(bf < REDO_BLOCKS ? bf : REDO_BLOCKS), BLK, 0, st>>>(A, S, n, bu, e);
      double a, b, c;
      rk_weights(rk3, s, a, b, c);
      k_tp_update_c2pa<<<bc, BLK, 0, st>>>(A, S, n, cells, bu, bo, a, b, c, e);
      k_tp_c2p_b<<<bc, BLK, 0, st>>>(A, S, n, cells, bo, atmo, e);
      k_tp_c2p_redo<<<(bc < REDO_BLOCKS ? bc : REDO_BLOCKS), BLK, 0, st>>>(A, S, n, bo, atmo, e);
      bu = bo;
    }
    if (const int rc = launch_status()) return rc;
  }
  k_tp_final<<<blocks_for(n), BLK, 0, st>>>(A, S, n);
  return launch_status();
}

// rows: cells per tube (ghosts included), the same for every tube
template <int LIM, int SOL, int NT>
int shared_nt(int n, int rows, const TubeArgs& A, const srmhd_eos& e, double cfl, double dx,
              double tend, double atmo, Clock* clk, hipStream_t st) {
  k_ts_smax<NT><<<n, NT, 0, st>>>(A, e.gamma, clk);
  k_ts_clock<<<1, 256, 0, st>>>(clk, cfl, dx, tend, rows);
  for (int s = 0; s < 2; ++s) {
    if (SOL == SOLVER_HLLD) {
      k_ts_flux_a<LIM, NT><<<n, NT, 0, st>>>(A, e, s, clk);
      k_ts_face_rows<LIM, NT><<<n, NT, 0, st>>>(A, e, s, rows, clk);
    }
    k_ts_mid<LIM, SOL, NT><<<n, NT, 0, st>>>(A, e, s, rows, clk);
    k_ts_cell_rows<NT><<<n, NT, 0, st>>>(A, e, s, rows, clk);
    k_ts_end<NT><<<n, NT, 0, st>>>(A, e, s, rows, atmo, clk);
  }
  return launch_status();
}

// mode 0: per-tube run; mode 1: one shared-dt step.  block 0 (per-tube run
// only): the phased form
template <int LIM, int SOL>
int by_block(int mode, int block, int n, const TubeArgs& A, const srmhd_eos& e, int steps,
             int rows, double cfl, int rk3, double dx, double tend, double atmo, Clock* clk,
             hipStream_t st) {
  switch (block) {
    case 0:
      return mode ? SRMHD_EARG : run_phased<LIM, SOL>(n, A, e, steps, cfl, rk3, atmo, st);
    case 64:
      return mode ? shared_nt<LIM, SOL, 64>(n, rows, A, e, cfl, dx, tend, atmo, clk, st)
                  : run_nt<LIM, SOL, 64>(n, A, e, steps, cfl, rk3, atmo, st);
    case 128:
      return mode ? shared_nt<LIM, SOL, 128>(n, rows, A, e, cfl, dx, tend, atmo, clk, st)
                  : run_nt<LIM, SOL, 128>(n, A, e, steps, cfl, rk3, atmo, st);
    case 256:
      return mode ? shared_nt<LIM, SOL, 256>(n, rows, A, e, cfl, dx, tend, atmo, clk, st)
                  : run_nt<LIM, SOL, 256>(n, A, e, steps, cfl, rk3, atmo, st);
    case 512:
      return mode ? shared_nt<LIM, SOL, 512>(n, rows, A, e, cfl, dx, tend, atmo, clk, st)
                  : run_nt<LIM, SOL, 512>(n, A, e, steps, cfl, rk3, atmo, st);
    default:
      return SRMHD_EARG;
  }
}

template <int LIM>
int by_solver(int sol, int mode, int block, int n, const TubeArgs& A, const srmhd_eos& e,
              int steps, int rows, double cfl, int rk3, double dx, double tend, double atmo, Clock* clk,
              hipStream_t st) {
  switch (sol) {
    case SOLVER_HLLE:
      return by_block<LIM, SOLVER_HLLE>(mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend, atmo, clk, st);
    case SOLVER_HLLC:
      return by_block<LIM, SOLVER_HLLC>(mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend, atmo, clk, st);
    default:
      return by_block<LIM, SOLVER_HLLD>(mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend, atmo, clk, st);
  }
}

}  // namespace

namespace srmhd {

template <>
int tubes_dispatch<SRMHD_TUBE_LIM>(int sol, int mode, int block, int n, const TubeArgs& A,
                                   const srmhd_eos& e, int steps, int rows, double cfl, int rk3,
                                   double dx, double tend, double atmo, Clock* clk,
                                   hipStream_t st) {
  return by_solver<SRMHD_TUBE_LIM>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend,
                                   atmo, clk, st);
}

}  // namespace srmhd
