// Mixed-precision sweep launcher shared between translation units.
#pragma once

#include "srmhd_grid2d.h"

namespace srmhd {

inline int64_t n_faces_host(const G2& g, int dir) {
  return dir == 0 ? (int64_t)(g.nxt + 1) * g.nyt : (int64_t)g.nxt * (g.nyt + 1);
}

// The reconstructed variables of every cell (ghosts included) in FP32, one
// array each: rho, p, W vx, W vy, W vz, cell-centred Bx, By, and Bz.
struct Cells32 {
  float* p[8];
};

// Fill C from the primitives P and the state S (face B) of one stage.
void cells_f32(const G2& g, const St2& S, const Pr2& P, const Cells32& C, hipStream_t s);

// One mixed-precision sweep along dir (0: x, 1: y) into the FP64 flux arrays
// F, with the flux counters (and the |F_D| max of the upwind EMF) into the
// DIAG_SHARDS copies starting at d (folded by the caller),
// and the faces to recompute in FP64 appended to fix (fix[0] = count, zeroed
// by the caller; indices from fix[2]; room for every face).  r64: the
// reconstruction in FP64 from P (else in FP32 from C, filled by cells_f32).
// emf_hll: the EMF slot of F holds the HLL average of the tangential-field
// flux (emf_mode "hll") instead of the solver's.
void sweep_mixed(int dir, int lim, int sol, bool r64, const G2& g, const St2& S, const Pr2& P,
                 const Cells32& C, const srmhd_eos& e, srmhd_flux_diag* d, const Flx& F,
                 unsigned* fix, int emf_hll, hipStream_t s);

// con2prim of the physical cells of S into P for the mixed forms: each cell
// its own batch, mu_+ bracketed from FP32 (no primitive BCs; the caller
// applies them)
void c2p_cells_mixed(const G2& g, const St2& S, const srmhd_eos& e, double atmo_rho,
                     double atmo_eps, const Pr2Out& P, unsigned long long* redo_count,
                     hipStream_t s);

}  // namespace srmhd
