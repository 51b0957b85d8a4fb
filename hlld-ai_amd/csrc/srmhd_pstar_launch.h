// Host launchers of the HLLD p* option kernels (srmhd_pstar.hip), called by
// the entry points of srmhd_kernels.hip after their first pass.
#pragma once

#include "srmhd_grid2d.h"
#include "srmhd_pstar.h"

namespace srmhd {

// second pass of the flat flux with options (HLLD): L, R 9 pointers, F 8,
// U 8 or NULL; rec / d: the first pass's records and folded mask (legacy
// bracket; unused with the hllc bracket)
int launch_flux_opts_b(int idir, int64_t n, const double* const* L, const double* const* R,
                       const srmhd_eos& e, const C2PRec* rec, srmhd_flux_diag* d,
                       double* const* F, double* const* U, double* pstar, uint8_t* status,
                       const HlldOpts& o, hipStream_t s);

// second pass of a 2D HLLD sweep along N with options
template <int N, int LIM>
void launch_sweep_opts_b(const G2& g, const St2& S, const Pr2& P, const srmhd_eos& e,
                         const C2PRec* rec, const srmhd_flux_diag* d, srmhd_flux_diag* dsh,
                         const Flx& F, double* pstar_out, int emf_hll, const HlldOpts& o,
                         hipStream_t s);

}  // namespace srmhd
