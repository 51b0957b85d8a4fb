// Types shared by the batched-tube translation units (srmhd_tubes.hip, the
// C ABI; srmhd_tubes_lim.hip, the kernels of one limiter each).
#pragma once

#include "srmhd_launch.h"
#include "srmhd_c2p.h"
#include "srmhd_recon.h"
#include "srmhd_riemann.h"

namespace srmhd {

struct TubeState {
  double* dt;                // this step's dt per tube
  unsigned long long* mf;    // HLL-state con2prim mask of the stage, per tube
  unsigned long long* mc;    // cell con2prim mask of the stage, per tube
  double* tacc;              // elapsed time of this run, per tube
  // work lists of the elements whose con2prim takes the bisection redo
  // (faces, cells) and their counts qn[0], qn[1]
  int64_t* qf;
  int64_t* qc;
  unsigned long long* qn;
};

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
  C2PRecK* recf;
  C2PRecK* recc;
  double* hq;             // the phased form's side states, pass A to pass B (HQ_N x faces)
  TubeState ts;           // the phased form's per-tube words
  int64_t cells;          // all tubes' cells
};

// Clock of the shared-dt form (srmhd_tubes_shared_step)
struct Clock {
  double t, steps, dt, active;
  unsigned long long mf[2], mc[2], smax, index_errors;
  // then unsigned marks: face rows [2][rows + 1], cell rows [2][rows]
};

// The launches of one limiter (srmhd_tubes_lim.hip, one object per limiter):
// mode 0, the per-tube run (block 64..512: k_tubes; 0: the phased form);
// mode 1, one shared-dt step.
template <int LIM>
int tubes_dispatch(int sol, int mode, int block, int n, const TubeArgs& A, const srmhd_eos& e,
                   int steps, int rows, double cfl, int rk3, double dx, double tend, double atmo,
                   Clock* clk, hipStream_t st);
#define SRMHD_TUBES_DISPATCH_DECL(L)                                                        \
  template <>                                                                              \
  int tubes_dispatch<L>(int sol, int mode, int block, int n, const TubeArgs& A,            \
                        const srmhd_eos& e, int steps, int rows, double cfl, int rk3,      \
                        double dx, double tend, double atmo, Clock* clk, hipStream_t st);
SRMHD_TUBES_DISPATCH_DECL(LIM_PCM)
SRMHD_TUBES_DISPATCH_DECL(LIM_MC)
SRMHD_TUBES_DISPATCH_DECL(LIM_MINMOD)
SRMHD_TUBES_DISPATCH_DECL(LIM_WENO5Z)
SRMHD_TUBES_DISPATCH_DECL(LIM_MP5)
SRMHD_TUBES_DISPATCH_DECL(LIM_MP7)
#undef SRMHD_TUBES_DISPATCH_DECL

}  // namespace srmhd
