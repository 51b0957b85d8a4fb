// Batched 1D Riemann tubes on gfx950 (see include/srmhd.h, srmhd_tubes_*).
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
#include "srmhd_launch.h"
#include "srmhd_recon.h"
#include "srmhd_riemann.h"

using namespace srmhd;

namespace {

struct TubeArgs {
  const int64_t* off_c;
  const int* ncells;      // physical cells L_t
  const double* dx;
  const double* Bn;
  const int* order;       // workgroup -> tube (NULL: identity)
  double* t_out;          // per-tube elapsed time
  double* U[3][7];        // U[0]: caller's state (in/out); U[1..2]: stage buffers
  double* P[6];
  double* F[7];
  C2PRec* recf;
  C2PRec* recc;
};

// The tube a workgroup owns, with its pointers offset to the tube's start.
struct Tube {
  int id, L, nt, g;
  int64_t oc;
  double dx, Bn;
  double* P[6];
  double* F[7];
  C2PRec* recf;
  C2PRec* recc;
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
__device__ inline void rk_weights(int rk3, int s, double& a, double& b, double& c) {
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

struct Clock {
  double t, steps, dt, active;
  unsigned long long mf[2], mc[2], smax, index_errors;
  // then unsigned marks: face rows [2][rows + 1], cell rows [2][rows]
};

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

// mode 0: per-tube run; mode 1: one shared-dt step
template <int LIM, int SOL>
int by_block(int mode, int block, int n, const TubeArgs& A, const srmhd_eos& e, int steps,
             int rows, double cfl, int rk3, double dx, double tend, double atmo, Clock* clk,
             hipStream_t st) {
  switch (block) {
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

int dispatch(int lim, int sol, int mode, int block, int n, const TubeArgs& A, const srmhd_eos& e,
             int steps, int rows, double cfl, int rk3, double dx, double tend, double atmo, Clock* clk,
             hipStream_t st) {
  switch (lim) {
    case SRMHD_PCM: return by_solver<LIM_PCM>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend, atmo, clk, st);
    case SRMHD_MC: return by_solver<LIM_MC>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend, atmo, clk, st);
    case SRMHD_MINMOD: return by_solver<LIM_MINMOD>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend, atmo, clk, st);
    case SRMHD_WENO5Z: return by_solver<LIM_WENO5Z>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend, atmo, clk, st);
    case SRMHD_MP5: return by_solver<LIM_MP5>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend, atmo, clk, st);
    default: return by_solver<LIM_MP7>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend, atmo, clk, st);
  }
}

struct WsT {
  double* U[2][7];
  double* F[7];
  C2PRec *recf, *recc;
  int64_t bytes;
};

WsT carve_tubes(int64_t cells, int64_t faces, void* work) {
  Carver c{(char*)work};
  WsT w;
  for (int b = 0; b < 2; ++b)
    for (int k = 0; k < 7; ++k) w.U[b][k] = c.take<double>(cells);
  for (int k = 0; k < 7; ++k) w.F[k] = c.take<double>(faces);
  w.recf = c.take<C2PRec>(faces);
  w.recc = c.take<C2PRec>(cells);
  w.bytes = c.off;
  return w;
}

bool tube_lim_ok(int lim) {
  return lim == SRMHD_PCM || lim == SRMHD_MC || lim == SRMHD_MINMOD || lim == SRMHD_WENO5Z ||
         lim == SRMHD_MP5 || lim == SRMHD_MP7;
}

int tubes_args(int n_tubes, int64_t total_cells, const srmhd_tube_layout* lay,
               double* const* U, double* const* P, void* work, int64_t work_bytes, TubeArgs& A) {
  if (n_tubes < 0 || total_cells < 0 || !lay || !lay->off_c || !lay->ncells || !lay->dx ||
      !lay->Bn || !U || !P || !work)
    return SRMHD_EARG;
  const WsT w = carve_tubes(total_cells, total_cells + n_tubes, work);
  if (work_bytes < w.bytes) return SRMHD_EARG;
  A.off_c = lay->off_c;
  A.ncells = lay->ncells;
  A.dx = lay->dx;
  A.Bn = lay->Bn;
  A.order = lay->order;
  A.t_out = nullptr;
  for (int k = 0; k < 7; ++k) {
    if (!U[k]) return SRMHD_EARG;
    A.U[0][k] = U[k];
    A.U[1][k] = w.U[0][k];
    A.U[2][k] = w.U[1][k];
    A.F[k] = w.F[k];
  }
  for (int k = 0; k < 6; ++k) {
    if (!P[k]) return SRMHD_EARG;
    A.P[k] = P[k];
  }
  A.recf = w.recf;
  A.recc = w.recc;
  return SRMHD_OK;
}

bool block_ok(int b) { return b == 64 || b == 128 || b == 256 || b == 512; }

}  // namespace

extern "C" {

int64_t srmhd_tubes_workspace_bytes(int n_tubes, int64_t total_cells) {
  if (n_tubes < 0 || total_cells < 0) return -1;
  return carve_tubes(total_cells, total_cells + n_tubes, nullptr).bytes;
}

int srmhd_tubes_run(int n_tubes, int64_t total_cells, const srmhd_tube_layout* lay,
                    const srmhd_eos* eos, int solver, int limiter, int rk3, int steps, double cfl,
                    double atmo_rho, int block, double* const* U, double* const* P, double* t_out,
                    void* work, int64_t work_bytes, void* stream) {
  if (!eos_ok(eos) || solver < 0 || solver > 2 || !tube_lim_ok(limiter) || steps < 0 ||
      !t_out || !block_ok(block))
    return SRMHD_EARG;
  TubeArgs A;
  const int rc = tubes_args(n_tubes, total_cells, lay, U, P, work, work_bytes, A);
  if (rc) return rc;
  if (n_tubes == 0 || steps == 0) return SRMHD_OK;
  A.t_out = t_out;
  return dispatch(limiter, solver, 0, block, n_tubes, A, *eos, steps, 0, cfl, rk3 ? 1 : 0, 0.0,
                  0.0, atmo_rho, nullptr, (hipStream_t)stream);
}

int64_t srmhd_tubes_clock_bytes(int64_t rows) {
  if (rows < 1) return -1;
  return (int64_t)sizeof(Clock) + (4 * rows + 2) * (int64_t)sizeof(unsigned);
}

int srmhd_tubes_shared_step(int n_tubes, int64_t total_cells, const srmhd_tube_layout* lay,
                            const srmhd_eos* eos, int solver, int limiter, double cfl, double tend,
                            double atmo_rho, int block, double* const* U, double* const* P,
                            void* clock, void* work, int64_t work_bytes, void* stream) {
  if (!eos_ok(eos) || solver < 0 || solver > 2 || !tube_lim_ok(limiter) || !clock ||
      !block_ok(block) || !lay || !(lay->dx_shared > 0.0) || n_tubes < 0 ||
      (n_tubes > 0 && total_cells % n_tubes != 0) || total_cells / (n_tubes ? n_tubes : 1) > INT32_MAX)
    return SRMHD_EARG;
  TubeArgs A;
  const int rc = tubes_args(n_tubes, total_cells, lay, U, P, work, work_bytes, A);
  if (rc) return rc;
  if (n_tubes == 0) return SRMHD_OK;
  const int rows = (int)(total_cells / n_tubes);
  return dispatch(limiter, solver, 1, block, n_tubes, A, *eos, 1, rows, cfl, 0, lay->dx_shared,
                  tend, atmo_rho, (Clock*)clock, (hipStream_t)stream);
}

}  // extern "C"
