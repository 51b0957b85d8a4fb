/*
 * srmhd.h — C ABI of the MI355X (gfx950) SRMHD finite-volume step.
 *
 * Every entry point takes plain device pointers (float64 structure-of-arrays,
 * contiguous), sizes, and a HIP stream (hipStream_t passed as void*; NULL =
 * the null stream).  The caller owns every buffer, including the workspace
 * whose size the matching *_workspace_bytes() returns (srmhd_flux_noalloc
 * takes none: it uses a fixed device array of the library's); no device
 * memory is allocated inside.  One exception holds host-side HIP objects: the first
 * call on a device that uses a second stream -- srmhd2d_stage in the FP64
 * form and srmhd2d_stage_adjoint, whose y sweep runs beside the x sweep
 * unless the scheme's flags hold SRMHD_STAGE_SERIAL_SWEEPS; srmhd2d_stage in
 * the mixed-precision forms with HLLD, whose FP64 fix-up of the x faces runs
 * beside the y sweep -- creates one non-blocking stream and two events for
 * that device (default priority), shared by these calls and kept until the
 * process ends; if their creation fails the call runs on the caller's
 * stream alone.  All calls are asynchronous on the stream (the second stream
 * is forked from and joined back into it with events).  Return value: 0 on
 * success, SRMHD_EARG for a bad argument, otherwise SRMHD_EHIP + the
 * hipError_t of the failing launch.  Calls that share a workspace must not
 * overlap.
 *
 * Variable orders (the reference's dict keys):
 *   primitives  rho, vx, vy, vz, p, eps, Bx, By, Bz            (9)
 *   conserved   D, tau, Sx, Sy, Sz, Bx, By, Bz                 (8, tau form)
 *   2D state    D, Sx, Sy, Sz, tau, Bz, Bxf, Byf               (8: the reference's
 *               EVOLVED_KEYS, src/physics/state.py:41, then the staggered field)
 *   2D prims    rho, vx, vy, vz, p, eps                         (6; B is derived
 *               from the faces and from Bz)
 *
 * References are to kemi2805/HLLD-AI.
 */
#ifndef SRMHD_H
#define SRMHD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SRMHD_OK 0
#define SRMHD_EARG 1
#define SRMHD_ESIZE 2   /* n above SRMHD_NOALLOC_MAX_FACES (srmhd_flux_noalloc) */
#define SRMHD_EHIP 1000

/* hybrid_eos (src/physics/eos.py:15-20): p = K rho^g + (eps - eps_cold) rho (g_th - 1);
 * is_cold = 1 when gamma_th is None or 1. */
typedef struct srmhd_eos {
  double K, gamma, gamma_th;
  int32_t is_cold;
  int32_t reserved;
} srmhd_eos;

/* Device-resident counters of one flux call (the reference's hlld.LAST_DIAG,
 * src/physics/hlld.py:1516-1529).  Filled by srmhd_flux and per sweep by
 * srmhd2d_stage; read with a device-to-host copy when wanted. */
typedef struct srmhd_flux_diag {
  unsigned long long n_interfaces;
  unsigned long long n_not_converged;
  unsigned long long n_wave_order_bad;
  unsigned long long n_bn_zero;
  unsigned long long n_hlle_fallback;
  unsigned long long sum_iters_converged;
  unsigned long long n_converged;
  unsigned long long max_iters;
  unsigned long long c2p_mask;      /* OR of the HLL-state con2prim masks */
  unsigned long long absF_D_max;    /* bits of max |F_D| (upwind-EMF deadband scale) */
  unsigned long long n_c2p_redo;    /* HLL-state con2prims finished by the bisection redo */
} srmhd_flux_diag;

/* SRMHD_LINEAR (the linearised flux with the HLLD fallback, the reference's
 * --solver linear) is a solver of srmhd2d_stage and srmhd1d_rk2_step in the
 * FP64 form; the flat flux entry point for it is srmhd_linear_flux. */
enum { SRMHD_HLLE = 0, SRMHD_HLLC = 1, SRMHD_HLLD = 2, SRMHD_LINEAR = 3 };
enum { SRMHD_PCM = 0, SRMHD_MC = 1, SRMHD_MINMOD = 2, SRMHD_WENO5Z = 3, SRMHD_MP5 = 4,
       SRMHD_MP7 = 5 };
enum { SRMHD_BC_OUTFLOW = 0, SRMHD_BC_PERIODIC = 1 };

/* ── HLLD p* options ─────────────────────────────────────────────────────────
 * The reference's hlld_flux reads four module globals at call time
 * (src/physics/hlld.py:1065-1092): the root finder of the star pressure
 * (_PSTAR_METHOD, from RMHD_PSTAR_METHOD: "secant", the production clamped
 * secant, or one of the four newton_pstar forms, hlld.py:751-904), the
 * bracket source (_PSTAR_BRACKET: "legacy", the con2prim of the HLL state,
 * or "hllc", the HLLC star pressure _hllc_pstar, hlld.py:1104-1199), the
 * residual tolerance (_PSTAR_TOL, 1e-10) and the relative step of the finite
 * differences (_PSTAR_FD_H, from RMHD_PSTAR_FD_H; None = 1.5e-8 for
 * newton_fd1, 1e-6 for newton_fd).  A zeroed struct (or NULL) is the
 * reference's default: secant, legacy, 1e-10. */
enum { SRMHD_PSTAR_SECANT = 0, SRMHD_PSTAR_NEWTON = 1, SRMHD_PSTAR_NEWTON_FD = 2,
       SRMHD_PSTAR_NEWTON_FD1 = 3, SRMHD_PSTAR_NEWTON_SAFE = 4 };
enum { SRMHD_BRACKET_LEGACY = 0, SRMHD_BRACKET_HLLC = 1 };
typedef struct srmhd_hlld_opts {
  int32_t pstar_method;    /* SRMHD_PSTAR_* */
  int32_t pstar_bracket;   /* SRMHD_BRACKET_* */
  double pstar_tol;        /* <= 0: 1e-10 */
  double pstar_fd_h;       /* <= 0: the method's default */
} srmhd_hlld_opts;

/* ── point-wise flux: the flux_fn plugin ──────────────────────────────────────
 * Replaces hlle_flux / hllc_flux / hlld_flux (src/physics/hlld.py:1005, 1549,
 * 1202) behind the contract flux_fn(sL, sR, eos, idir) -> (F, U, p_star)
 * (src/physics/sweep.py:96, driver.py:168).  n interfaces; primL/primR: 9
 * pointers; F, U: 8 pointers (U may be NULL); pstar, status (uint8 per face:
 * bit0 not converged, bit1 wave order, bit2 B_n = 0, bit3 HLLE fallback, bit4
 * HLL-state con2prim redone by bisection) may
 * be NULL; diag (device) may be NULL.  HLLD runs the reference's
 * batch-coupled con2prim of the HLL state over all n faces as one batch. */
int64_t srmhd_flux_workspace_bytes(int solver, int64_t n);
int srmhd_flux(int solver, int idir, int64_t n, const double* const* primL,
               const double* const* primR, const srmhd_eos* eos, double* const* F,
               double* const* U, double* pstar, uint8_t* status, srmhd_flux_diag* diag,
               void* work, int64_t work_bytes, void* stream);

/* srmhd_flux with the HLLD p* options (hlld_flux under the reference's
 * _PSTAR_METHOD / _PSTAR_BRACKET / _PSTAR_TOL / _PSTAR_FD_H, hlld.py:1065-1092,
 * 1246-1343); opts NULL = srmhd_flux.  The options concern HLLD only.  The
 * workspace is srmhd_flux_workspace_bytes(solver, n) whatever the options. */
int srmhd_flux_ex(int solver, int idir, int64_t n, const double* const* primL,
                  const double* const* primR, const srmhd_eos* eos,
                  const srmhd_hlld_opts* opts, double* const* F, double* const* U,
                  double* pstar, uint8_t* status, srmhd_flux_diag* diag, void* work,
                  int64_t work_bytes, void* stream);

/* srmhd_flux_ex for callers that cannot allocate device memory or hand over
 * a workspace: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  No counterpart in the reference.  The workspace is
 * a fixed device array inside the library (one copy per device, 256 bytes +
 * 24 bytes per face: 24 MiB + 256 bytes for SRMHD_NOALLOC_MAX_FACES faces,
 * held until the process ends); n above SRMHD_NOALLOC_MAX_FACES returns
 * SRMHD_ESIZE.  The scratch used is the CURRENT device's, so a non-NULL
 * stream must belong to the current device (else SRMHD_EARG).  Every call on
 * a device shares that device's array: calls that overlap on one device
 * (concurrent streams, threads) race and must be ordered by the caller.  No
 * allocation of any kind is made; the results equal srmhd_flux_ex's on the
 * same inputs, bit for bit. */
#define SRMHD_NOALLOC_MAX_FACES (1 << 20)
int srmhd_flux_noalloc(int solver, int idir, int64_t n, const double* const* primL,
                       const double* const* primR, const srmhd_eos* eos,
                       const srmhd_hlld_opts* opts, double* const* F, double* const* U,
                       double* pstar, uint8_t* status, srmhd_flux_diag* diag, void* stream);

/* ── the linearised flux: --solver linear ────────────────────────────────────
 * Replaces linearised_flux (src/physics/linearised.py:129-237), the same
 * flux_fn contract: the fallback solver's (F, U, p*) on every face, then on
 * each face the linearisation answers, the physical flux of the state on the
 * ray xi = 0 from the 7x7 eigen-decomposition of the mean state's
 * quasi-linear matrix (Jacobians by central differences with relative step
 * fd_rel), its conserved state and total pressure.  The faces kept by the
 * fallback: non-finite Jacobians or a non-positive density or pressure on a
 * side, a singular dU/dV or eigenvector matrix, complex eigenvalues (any
 * imaginary part: imag_tol bounds only the rounding the reference allows),
 * an eigenvector matrix of 2-norm condition >= cond_max, or an unphysical
 * ray state.  linear_mask (uint8 per face, may be NULL): 1 where the
 * linearisation answered; n_linear (one device word, may be NULL): their
 * count; diag, status words of the fallback as srmhd_flux's.  The eigen
 * decomposition is the library's own (Hessenberg reduction and Francis QR),
 * so the results agree with the reference's LAPACK ones to rounding. */
enum { SRMHD_LINEAR_FALLBACK_HLLD = 0, SRMHD_LINEAR_FALLBACK_HLLE = 1,
       SRMHD_LINEAR_FALLBACK_HLLC = 2 };
typedef struct srmhd_linear_opts {
  int32_t fallback;        /* SRMHD_LINEAR_FALLBACK_* (linearised.py:130 fallback=hlld_flux) */
  int32_t reserved;
  double fd_rel;           /* <= 0: 1e-5 */
  double imag_tol;         /* <= 0: 1e-6 */
  double cond_max;         /* <= 0: 1e12 */
  srmhd_hlld_opts hlld;    /* the HLLD fallback's p* options (zeroed: the default) */
} srmhd_linear_opts;
int64_t srmhd_linear_workspace_bytes(int64_t n);
int srmhd_linear_flux(int idir, int64_t n, const double* const* primL,
                      const double* const* primR, const srmhd_eos* eos,
                      const srmhd_linear_opts* opts, double* const* F, double* const* U,
                      double* pstar, uint8_t* linear_mask, srmhd_flux_diag* diag,
                      unsigned long long* n_linear, void* work, int64_t work_bytes,
                      void* stream);

/* ── con2prim ────────────────────────────────────────────────────────────────
 * Replaces conservative_to_primitive (src/physics/c2p.py:313-363) over one
 * batch of n cells: cons 8 pointers in, prim 9 pointers out, status (uint8:
 * bit0 bracket_ok, bit1 converged, bit2 v_clamped, bit3 atmo_floored, bit4 finished
 * by the bisection redo; may be NULL). */
int64_t srmhd_c2p_workspace_bytes(int64_t n);
int srmhd_c2p(int64_t n, const double* const* cons, const srmhd_eos* eos, double atmo_rho,
              double atmo_eps, double* const* prim, uint8_t* status, void* work,
              int64_t work_bytes, void* stream);

/* primitive_to_conserved (src/physics/hlld.py:185-217): prim 9 -> cons 8. */
int srmhd_p2c(int64_t n, const double* const* prim, double* const* cons, void* stream);

/* ── 2D constrained-transport step ───────────────────────────────────────────
 * Grid2D (src/physics/grid.py:146-206): cells (nxt, nyt) with nxt = nx+2ng,
 * x-faces (nxt+1, nyt), y-faces (nxt, nyt+1), row-major with y contiguous. */
typedef struct srmhd_grid2d {
  int32_t nx, ny, ng, bc_x, bc_y, reserved;
  double dx, dy;
} srmhd_grid2d;

/* precision: SRMHD_PREC_FP64 (the reference's float64 arithmetic, bit-faithful)
 * or one of three forms with no counterpart in the reference:
 *   SRMHD_PREC_FP64_LOCAL FP64 arithmetic throughout, but every con2prim (of a
 *                         cell, and of the HLL state that seeds HLLD's p*) is
 *                         solved as its own batch and stops at its own
 *                         convergence (|f| < 1e-15), instead of iterating with
 *                         its batch until the whole batch has converged;
 *   SRMHD_PREC_MIXED      reconstruction and the Riemann solver in FP32 (HLLD
 *                         with the con2prim of its HLL state solved per face);
 *                         faces whose HLLE-fallback decision FP32 cannot settle
 *                         are recomputed in FP64;
 *   SRMHD_PREC_MIXED_R64  reconstruction in FP64; HLLD's p* root found in FP32,
 *                         polished and its flux assembled in FP64.
 * In both mixed forms, corner EMF, CT face-B update and RK accumulation are
 * FP64 (div B is untouched) and the cell con2prim is FP64 with each cell its
 * own batch. */
enum {
  SRMHD_PREC_FP64 = 0,
  SRMHD_PREC_MIXED = 1,
  SRMHD_PREC_MIXED_R64 = 2,
  SRMHD_PREC_FP64_LOCAL = 3
};

/* emf_mode: the face EMF fed to the corner EMF (ct.py:63-88, selected by
 * run_2d.py --emf-mode): SRMHD_EMF_SOLVER, the solver's own flux of the
 * tangential field (-F^x(B_y), F^y(B_x)); SRMHD_EMF_HLL, the HLL average of
 * that flux from the side states and HLL speeds of the face
 * (compute_srmhd_fluxes, sweep.py:106-118), whatever the solver.  The
 * adjoint entry points take SRMHD_EMF_SOLVER only. */
enum { SRMHD_EMF_SOLVER = 0, SRMHD_EMF_HLL = 1 };

/* hlld: the HLLD p* options (zeroed: the reference's default); non-default
 * methods and the hllc bracket are taken by the FP64 precision only (the
 * other forms and the adjoint entry points refuse them, SRMHD_EARG). */
/* flags: SRMHD_STAGE_SERIAL_SWEEPS runs the FP64 form's y sweep (in
 * srmhd2d_stage and srmhd2d_stage_adjoint) after its x sweep on the caller's
 * stream; without it the y sweep runs beside the x sweep on a second stream
 * of the device (created at first use, default priority; forked from and
 * joined back into the caller's stream with events).  The results are the
 * same bit for bit. */
enum { SRMHD_STAGE_SERIAL_SWEEPS = 1 };
typedef struct srmhd_scheme2d {
  int32_t solver, limiter, upwind, precision;
  int32_t emf_mode, flags;   /* flags: SRMHD_STAGE_* bits, 0 by default */
  double atmo_rho, atmo_eps;
  srmhd_hlld_opts hlld;
} srmhd_scheme2d;

int64_t srmhd2d_workspace_bytes(const srmhd_grid2d* g, const srmhd_scheme2d* s);

/* One RK stage of rk_step_ct (src/physics/driver2d.py:241-247 = ct_stage +
 * state.combine + sync_state): out = a*U0 + b*U + (c*dt)*L(U, P), then face
 * and cell BCs, con2prim of every cell (one batch), primitive BCs.  dt is read
 * from device memory (*dt_dev).  U0, U, out: 8 pointers (2D state order);
 * P, Pout: 6 pointers.  out/Pout must not alias U0/U/P.  diag_xy (device, may
 * be NULL) receives the x and y sweep counters (2 structs). */
int srmhd2d_stage(const srmhd_grid2d* g, const srmhd_scheme2d* s, const srmhd_eos* eos,
                  double a, double b, double c, const double* dt_dev,
                  const double* const* U0, const double* const* U, const double* const* P,
                  double* const* out, double* const* Pout, srmhd_flux_diag* diag_xy,
                  void* work, int64_t work_bytes, void* stream);

/* The stage's record for the adjoint (srmhd2d_stage_adjoint): HLLD's p* of
 * every x-face ((nxt+1) x nyt) and y-face (nxt x (nyt+1)), negative where the
 * face fell back to HLLE (the reference's sentinel, hlld.py:1531); the mass
 * fluxes of both sweeps and the two deadband thresholds 1e-9 max|F_D| of the
 * upwind EMF (ct.py:96-116).  Device memory, caller-owned. */
typedef struct srmhd2d_tape {
  double* pstar_x;
  double* pstar_y;
  double* FD_x;
  double* FD_y;
  double* thresh;  /* 2 doubles */
} srmhd2d_tape;

/* srmhd2d_stage (FP64 precision only) that also fills *tape.  No counterpart
 * in the reference: the forward half of the adjoint. */
int srmhd2d_stage_taped(const srmhd_grid2d* g, const srmhd_scheme2d* s, const srmhd_eos* eos,
                        double a, double b, double c, const double* dt_dev,
                        const double* const* U0, const double* const* U, const double* const* P,
                        double* const* out, double* const* Pout, const srmhd2d_tape* tape,
                        void* work, int64_t work_bytes, void* stream);

/* ── adjoint of the 2D step (no counterpart in the reference) ───────────────
 * Reverse-mode derivative of one FP64 stage out = a U0 + b U + (c dt) L(U, P)
 * followed by its BCs and con2prim, for the differentiable rollout of
 * driver2d.rk_step_ct (src/physics/driver2d.py:221-249) with dt held fixed.
 * Transposed by hand: the con2prim (implicit differentiation of the Kastaun
 * root, c2p.py:129-158), the BC fills, the RK combination, the flux
 * divergence, ct_rhs and the Gardiner-Stone corner EMF with the taped upwind
 * signs (ct.py:119-185), and per face the reconstruction (limiter branches as
 * taken) and the Riemann solver (HLLD: implicit differentiation of p* at the
 * taped root, hlld.py:586-602; the HLLE fallback where taped).
 *   U, P      the stage inputs U^(k) (8) and P^(k) (6);
 *   Uout      the stage output U^(k+1) (8), BCs applied;
 *   lamUout   adjoint of U^(k+1) (8), lamPout adjoint of P^(k+1) (6): inputs,
 *             overwritten (used as scratch);
 *   lamU0, lamU, lamP: accumulated (+=) adjoints of U0, U^(k), P^(k); lamU0
 *             and lamU may be the same buffers (stage 0, where U = U0).
 * Every accumulation is a gather in a fixed order: results are deterministic. */
int64_t srmhd2d_adjoint_workspace_bytes(const srmhd_grid2d* g, const srmhd_scheme2d* s);
int srmhd2d_stage_adjoint(const srmhd_grid2d* g, const srmhd_scheme2d* s, const srmhd_eos* eos,
                          double a, double b, double c, const double* dt_dev,
                          const double* const* U, const double* const* P,
                          const double* const* Uout, const srmhd2d_tape* tape,
                          double* const* lamUout, double* const* lamPout,
                          double* const* lamU0, double* const* lamU, double* const* lamP,
                          void* work, int64_t work_bytes, void* stream);

/* Adjoint of srmhd2d_sync: U is the synced state (8), lamU (8) and lamP (6)
 * the adjoints of the synced state and primitives (overwritten); lamUin (8)
 * receives (=) the adjoint of the state before the sync (ghost entries 0). */
int srmhd2d_sync_adjoint(const srmhd_grid2d* g, const srmhd_scheme2d* s, const srmhd_eos* eos,
                         const double* const* U, double* const* lamU, double* const* lamP,
                         double* const* lamUin, void* work, int64_t work_bytes, void* stream);

/* Adjoint of srmhd_p2c: lam_prim (9) += J^T lam_cons (8) at each of n points. */
int srmhd_p2c_adjoint(int64_t n, const double* const* prim, const double* const* lam_cons,
                      double* const* lam_prim, void* stream);

/* sync_state (src/physics/driver2d.py:196-218) in place on state U (face BCs,
 * cell BCs) and into P: the initial-data and restart path. */
int srmhd2d_sync(const srmhd_grid2d* g, const srmhd_scheme2d* s, const srmhd_eos* eos,
                 double* const* U, double* const* P, void* work, int64_t work_bytes,
                 void* stream);

/* compute_dt_2d (src/physics/driver2d.py:52-82) into *dt_dev.  P: 9 pointers
 * rho, vx, vy, vz, p, eps, Bx, By, Bz (cell-centred).  When P[6] and P[7] are
 * NULL, Bx and By are the face averages of U's staggered field and Bz is U's
 * Bz; U may be NULL otherwise. */
int srmhd2d_dt(const srmhd_grid2d* g, const srmhd_eos* eos, double cfl,
               const double* const* U, const double* const* P, double* dt_dev, void* work,
               int64_t work_bytes, void* stream);

/* ── 1D shock tube (driver.run) ─────────────────────────────────────────────
 * Grid1D (src/physics/grid.py:23-43): ncells physical cells, ng ghosts per
 * side; bc 0 outflow, 1 periodic, 2 constant (ghosts keep the step-start
 * values, grid.py:49-87). */
typedef struct srmhd_grid1d {
  int32_t ncells, ng, bc, reserved;
  double dx;
} srmhd_grid1d;

int64_t srmhd1d_workspace_bytes(const srmhd_grid1d* g);

/* One Heun SSP-RK2 step of driver.rk2_step (src/physics/driver.py:192-236):
 * get_interface_states on the ncells+1 physical faces (every primitive key
 * reconstructed directly, reconstruction.py:354-373), the flux (HLLD's
 * HLL-state con2prim batch = those faces), F[Bx] = 0, the BCs, Bx restored,
 * con2prim of every cell as one batch, primitive BCs.  cons/cons_out: 8
 * conserved pointers (D, tau, Sx, Sy, Sz, Bx, By, Bz) of ncells+2ng; prim /
 * prim_out: 9 primitive pointers.  dt from device memory.  diag2 (device,
 * may be NULL) gets the two stages' flux counters. */
int srmhd1d_rk2_step(const srmhd_grid1d* g, int solver, int limiter, const srmhd_eos* eos,
                     double atmo_rho, const double* dt_dev, const double* const* cons,
                     const double* const* prim, double* const* cons_out, double* const* prim_out,
                     srmhd_flux_diag* diag2, void* work, int64_t work_bytes, void* stream);

/* srmhd1d_rk2_step with the HLLD p* options (NULL = srmhd1d_rk2_step). */
int srmhd1d_rk2_step_ex(const srmhd_grid1d* g, int solver, int limiter, const srmhd_eos* eos,
                        const srmhd_hlld_opts* opts, double atmo_rho, const double* dt_dev,
                        const double* const* cons, const double* const* prim,
                        double* const* cons_out, double* const* prim_out,
                        srmhd_flux_diag* diag2, void* work, int64_t work_bytes, void* stream);

/* compute_dt (src/physics/driver.py:77-127) into *dt_dev. */
int srmhd1d_dt(const srmhd_grid1d* g, const srmhd_eos* eos, double cfl, int speed_of_light,
               const double* const* prim, double* dt_dev, void* stream);

/* ── batched 1D Riemann tubes (src/physics/tube_seed.py:72-144) ─────────────
 * n_tubes tubes of possibly different lengths in one ragged layout.  Tube t
 * owns cells [off_c[t], off_c[t+1]) = ncells[t] physical cells plus ng ghosts
 * per side (ng = off_c[t+1] - off_c[t] - ncells[t] halves), and faces
 * [off_c[t] + t, off_c[t+1] + t + 1).  All arrays below are device memory.
 * U: 7 conserved pointers D, Sx, Sy, Sz, tau, By, Bz (B_n is constant per
 * tube); P: 6 primitive pointers rho, vx, vy, vz, p, eps; both over
 * total_cells = off_c[n_tubes] and updated in place.  order (may be NULL):
 * workgroup -> tube, e.g. longest first.  dx_shared: the one dx of the
 * shared-dt form (host value). */
typedef struct srmhd_tube_layout {
  const int64_t* off_c;   /* n_tubes + 1 */
  const int32_t* ncells;  /* n_tubes */
  const double* dx;       /* n_tubes */
  const double* Bn;       /* n_tubes */
  const int32_t* order;   /* n_tubes or NULL */
  double dx_shared;
} srmhd_tube_layout;

int64_t srmhd_tubes_workspace_bytes(int n_tubes, int64_t total_cells);

/* `steps` steps of every tube in ONE launch, each tube with its own dt:
 * compute_dt of the 1D driver (src/physics/driver.py:77-127) over its
 * physical cells; SSP-RK2 (rk3 = 0) or SSP-RK3 (rk3 = 1) with the weights of
 * src/physics/driver2d.py:104-105 in combine form a U0 + b U + (c dt) L;
 * each stage as tube_seed.py:81-109: reconstruct_prims along the tube with
 * L[Bx] = R[Bx] = B_n, the flux on every face of the tube (HLLD's HLL-state
 * con2prim batch = the tube's faces), L = -(F[j+1] - F[j]) / dx, con2prim of
 * every cell of the tube as one batch, two-cell-copy ghosts (which also reach
 * U's By, Bz, tube_seed.py:72-75,105-108).  t_out[t] += the tube's elapsed
 * time.  block: 64, 128, 256 or 512 -- one persistent launch, a workgroup of
 * that many threads per tube walking its tube phase by phase; or 0 -- the
 * phased form, one launch per phase over every tube's faces or cells
 * (2 + 4 x stages launches a step, one graph when captured), with the same
 * arithmetic and batches, bit for bit.  HLLD runs with the reference's
 * default p* options (secant, legacy bracket, 1e-10). */
int srmhd_tubes_run(int n_tubes, int64_t total_cells, const srmhd_tube_layout* layout,
                    const srmhd_eos* eos, int solver, int limiter, int rk3, int steps, double cfl,
                    double atmo_rho, int block, double* const* U, double* const* P, double* t_out,
                    void* work, int64_t work_bytes, void* stream);

/* The shared-dt form of run_tubes (tube_seed.py:93-144): n_tubes tubes of
 * equal length (rows = total_cells / n_tubes cells each, ghosts included),
 * one dt and one con2prim batch over every tube.  clock: device memory of
 * srmhd_tubes_clock_bytes(rows) bytes, zeroed by the caller before the first
 * step; it starts with doubles {t, steps, dt, active} and 64-bit words
 * {mask x4, smax, index_errors}.  Each call is one iteration of the
 * reference's loop: while t < tend - 1e-14, dt = min(cfl dx / max(smax,
 * 1e-8), tend - t), smax the max over every cell of every tube of |vx| +
 * min(c_s + 0.5, 1) (ideal-gas c_s from eos->gamma), then one Heun step (U1 =
 * U + dt L, U' = 0.5 (U + U1 + dt L)) with the stage algebra of
 * srmhd_tubes_run; after tend a call does nothing, so max_steps calls can be
 * captured in one graph.  The reference hands its (rows, N) arrays to
 * conservative_to_primitive unflattened, whose bisection redo then re-solves
 * every row named by the row OR the column index of an element that needs it
 * (c2p.py:339-346); that set is reproduced.  A column index past the last row
 * (an IndexError in the reference) increments index_errors.
 * layout->dx_shared is the dx of every tube. */
int64_t srmhd_tubes_clock_bytes(int64_t rows);
int srmhd_tubes_shared_step(int n_tubes, int64_t total_cells, const srmhd_tube_layout* layout,
                            const srmhd_eos* eos, int solver, int limiter, double cfl, double tend,
                            double atmo_rho, int block, double* const* U, double* const* P,
                            void* clock, void* work, int64_t work_bytes, void* stream);

/* ── live kernel timing (measurement only; no counterpart in the reference) ──
 * srmhd_timing_enable(cls, n): from now on srmhd2d_stage records a HIP event
 * pair around each launch of kernel class cls, on its stream, for up to n
 * launches (cls < 0 or n = 0 disables).  srmhd_timing_read() waits for the
 * recorded events and returns their summed duration and count, then resets.
 * Not thread-safe: one timing session per process. */
enum {
  SRMHD_K_SWEEP_X_A = 0, SRMHD_K_SWEEP_X_B = 1, SRMHD_K_SWEEP_Y_A = 2, SRMHD_K_SWEEP_Y_B = 3,
  SRMHD_K_CORNER = 4, SRMHD_K_UPDATE = 5, SRMHD_K_C2P_A = 6, SRMHD_K_C2P_B = 7,
  SRMHD_K_ADJ_SWEEP_X = 8  /* srmhd2d_stage_adjoint: the x-sweep adjoint */
};
int srmhd_timing_enable(int kernel_class, int max_launches);
int srmhd_timing_read(double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* SRMHD_H */
