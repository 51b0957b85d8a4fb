// Launchers of the linearised flux for the 2D stage and the 1D step
// (srmhd_linear.hip), called after the HLLD fallback has filled the fluxes.
#pragma once

#include "srmhd_grid2d.h"
#include "srmhd_linear.h"

namespace srmhd {

// flat faces (the 1D step's interface arrays): overwrite F (and U, pstar
// when non-null) where the linearisation answers; mask / count may be null
int launch_linear_flat(int idir, int64_t n, const double* const* L, const double* const* R,
                       const srmhd_eos& e, const LinOpts& o, double* const* F, double* const* U,
                       double* pstar, uint8_t* mask, unsigned long long* count, hipStream_t s);

// both sweeps of a 2D stage: re-form each face's states as the sweep does,
// overwrite the seven sweep fluxes where the linearisation answers, and
// replace diag[0..1].absF_D_max (the upwind-EMF deadband scale) by the
// largest |F_D| of the final fluxes.  words: 4 device words, zeroed here
// (0..1 the |F_D| maxima, 2..3 the answered counts of x and y).
int launch_linear_sweeps(int lim, const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e,
                         const LinOpts& o, const Flx& Fx, const Flx& Fy, int emf_hll,
                         srmhd_flux_diag* diag, unsigned long long* words, hipStream_t s);

}  // namespace srmhd
