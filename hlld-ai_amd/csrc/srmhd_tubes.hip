// Batched 1D Riemann tubes on gfx950: the C ABI (include/srmhd.h,
// srmhd_tubes_*).  The kernels of each limiter are in their own object
// (srmhd_tubes_lim.hip, compiled once per limiter).
#include "srmhd_tubes.h"
#include "srmhd_recon.h"

using namespace srmhd;

namespace {

int dispatch(int lim, int sol, int mode, int block, int n, const TubeArgs& A, const srmhd_eos& e,
             int steps, int rows, double cfl, int rk3, double dx, double tend, double atmo,
             Clock* clk, hipStream_t st) {
  switch (lim) {
    case SRMHD_PCM:
      return tubes_dispatch<LIM_PCM>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend,
                                     atmo, clk, st);
    case SRMHD_MC:
      return tubes_dispatch<LIM_MC>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend,
                                    atmo, clk, st);
    case SRMHD_MINMOD:
      return tubes_dispatch<LIM_MINMOD>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx,
                                        tend, atmo, clk, st);
    case SRMHD_WENO5Z:
      return tubes_dispatch<LIM_WENO5Z>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx,
                                        tend, atmo, clk, st);
    case SRMHD_MP5:
      return tubes_dispatch<LIM_MP5>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend,
                                     atmo, clk, st);
    default:
      return tubes_dispatch<LIM_MP7>(sol, mode, block, n, A, e, steps, rows, cfl, rk3, dx, tend,
                                     atmo, clk, st);
  }
}

struct WsT {
  double* U[2][7];
  double* F[7];
  C2PRecK *recf, *recc;
  double* hq;
  TubeState ts;
  int64_t bytes;
};

WsT carve_tubes(int64_t cells, int64_t faces, void* work) {
  Carver c{(char*)work};
  WsT w;
  const int64_t n = faces - cells;
  for (int b = 0; b < 2; ++b)
    for (int k = 0; k < 7; ++k) w.U[b][k] = c.take<double>(cells);
  for (int k = 0; k < 7; ++k) w.F[k] = c.take<double>(faces);
  w.recf = c.take<C2PRecK>(faces);
  w.recc = c.take<C2PRecK>(cells);
  w.ts.dt = c.take<double>(n);
  w.ts.mf = c.take<unsigned long long>(n);
  w.ts.mc = c.take<unsigned long long>(n);
  w.ts.tacc = c.take<double>(n);
  w.ts.qf = c.take<int64_t>(faces);
  w.ts.qc = c.take<int64_t>(cells);
  w.ts.qn = c.take<unsigned long long>(2);
  w.hq = c.take<double>(HQ_N * faces);
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
  A.hq = w.hq;
  A.ts = w.ts;
  A.cells = total_cells;
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
      !t_out || !(block == 0 || block_ok(block)))
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
