"""2D constrained-transport driver on the HIP path: drop-in for
src/physics/driver2d.py (``rk_step_ct``, ``sync_state``, ``compute_dt_2d``,
``prims_to_cons_2d``, ``cons_to_prims_2d``).

One RK stage is one C-ABI call (``srmhd2d_stage``): both sweeps with the
HLL-state con2prim, the Gardiner-Stone corner EMF, the flux divergence and
CT update with the RK combination and the BCs, and the con2prim of every
cell, all on the device.  ``Stepper2D`` keeps every buffer resident (three
states and two primitive sets, ping-ponged) for the step loop; the
functional ``rk_step_ct`` allocates new output tensors per stage like the
reference's ``combine``.
"""

import ctypes

import torch

from . import _lib
from .c2p import conservative_to_primitive
from .eos import as_c
from .hlld import LAST_DIAG, _LazyDiag, diag_solver, primitive_to_conserved, solver_name
from .state import EVOLVED_KEYS, PRIM6, State2D

RK = {"rk2": ((1.0, 0.0, 1.0), (0.5, 0.5, 0.5)),          # driver2d.py:104-105
      "rk3": ((1.0, 0.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0))}


def _c_grid(grid, bc_x, bc_y):
    return _lib.Grid2D(grid.nx, grid.ny, grid.ng, _lib.BC_IDS[bc_x], _lib.BC_IDS[bc_y], 0,
                       float(grid.dx), float(grid.dy))


class Stepper2D:
    """Device buffers, workspace and launch parameters for one grid + scheme."""

    def __init__(self, grid, eos, *, solver="hlld", limiter="mc", bc_x="outflow",
                 bc_y="outflow", upwind=True, atmo_rho=1e-10, atmo_eps=1e-10, device=None,
                 precision="fp64", emf_mode="solver", hlld_opts_from_globals=True,
                 serial_sweeps=False):
        self.grid = grid
        self.device = torch.device(device) if device is not None else grid.device
        if self.device.type != "cuda":
            raise RuntimeError("Stepper2D needs a HIP device (no CPU fallback)")
        self.lib = _lib.lib()
        self.eos = eos
        self.ceos = as_c(eos)
        self.solver = solver_name(solver)
        self.limiter = limiter
        if limiter not in _lib.LIMITER_IDS:
            raise ValueError(f"unknown limiter {limiter!r}")
        self.bc = (bc_x, bc_y)
        self.cg = _c_grid(grid, bc_x, bc_y)
        if precision not in _lib.PRECISION_IDS:
            raise ValueError(f"precision must be one of {tuple(_lib.PRECISION_IDS)}")
        self.precision = precision
        if emf_mode not in _lib.EMF_MODE_IDS:
            raise ValueError(f"unknown emf mode {emf_mode!r}")
        self.emf_mode = emf_mode
        self.cs = _lib.Scheme2D(_lib.SOLVER_IDS[self.solver], _lib.LIMITER_IDS[limiter],
                                1 if upwind else 0, _lib.PRECISION_IDS[precision],
                                _lib.EMF_MODE_IDS[emf_mode],
                                _lib.STAGE_SERIAL_SWEEPS if serial_sweeps else 0,
                                float(atmo_rho), float(atmo_eps))
        # HLLD's p* switches: srmhd.hlld's module globals, read at every stage
        # like the reference's hlld_flux does (False: the default root, fixed)
        self.hlld_opts_from_globals = bool(hlld_opts_from_globals)
        nb = self.lib.srmhd2d_workspace_bytes(ctypes.byref(self.cg), ctypes.byref(self.cs))
        if nb < 0:
            raise ValueError("grid / scheme rejected by srmhd2d_workspace_bytes")
        self.ws = _lib.workspace(nb, self.device)
        self.dt = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.diag = torch.zeros(2 * _lib.DIAG_WORDS, dtype=torch.int64, device=self.device)

    # ── buffers ─────────────────────────────────────────────────────────────
    def new_state(self):
        g, d = self.grid, self.device
        cells = [torch.empty(g.nxt, g.nyt, dtype=torch.float64, device=d) for _ in EVOLVED_KEYS]
        return cells + [torch.empty(g.nxt + 1, g.nyt, dtype=torch.float64, device=d),
                        torch.empty(g.nxt, g.nyt + 1, dtype=torch.float64, device=d)]

    def new_prims(self):
        g, d = self.grid, self.device
        return [torch.empty(g.nxt, g.nyt, dtype=torch.float64, device=d) for _ in PRIM6]

    # ── C-ABI calls ─────────────────────────────────────────────────────────
    def _refresh_hlld_opts(self):
        if not (self.hlld_opts_from_globals and self.solver in ("hlld", "linear")):
            return
        from . import hlld as H
        o = H.hlld_opts()
        if self.precision != "fp64" and not H.hlld_opts_default(o):
            raise ValueError(f"precision {self.precision!r} has its own p* root: the p* switches "
                             "(_PSTAR_METHOD, _PSTAR_BRACKET, _PSTAR_TOL) apply to fp64 only")
        self.cs.hlld = o

    def stage(self, a, b, c, U0, U, P, out, Pout):
        self._refresh_hlld_opts()
        a0, k0 = _lib.ptrs(U0)
        a1, k1 = _lib.ptrs(U)
        ap, kp = _lib.ptrs(P)
        ao, ko = _lib.ptrs(out)
        apo, kpo = _lib.ptrs(Pout)
        rc = self.lib.srmhd2d_stage(ctypes.byref(self.cg), ctypes.byref(self.cs),
                                    ctypes.byref(self.ceos), float(a), float(b), float(c),
                                    self.dt.data_ptr(), a0, a1, ap, ao, apo,
                                    self.diag.data_ptr(), self.ws.data_ptr(), self.ws.numel(),
                                    _lib.stream_ptr(self.device))
        _lib.check(rc, "srmhd2d_stage")

    def sync(self, U, P):
        au, ku = _lib.ptrs(U)
        ap, kp = _lib.ptrs(P)
        rc = self.lib.srmhd2d_sync(ctypes.byref(self.cg), ctypes.byref(self.cs),
                                   ctypes.byref(self.ceos), au, ap, self.ws.data_ptr(),
                                   self.ws.numel(), _lib.stream_ptr(self.device))
        _lib.check(rc, "srmhd2d_sync")

    def compute_dt(self, U, P, cfl, P9=None):
        """compute_dt_2d into ``self.dt`` (device); P9 optionally supplies
        cell-centred Bx, By, Bz."""
        plist = list(P) + ([None, None, None] if P9 is None else list(P9))
        au, ku = _lib.ptrs(U if U is not None else [None] * 8)
        ap, kp = _lib.ptrs(plist)
        rc = self.lib.srmhd2d_dt(ctypes.byref(self.cg), ctypes.byref(self.ceos), float(cfl),
                                 au if U is not None else None, ap, self.dt.data_ptr(),
                                 self.ws.data_ptr(), self.ws.numel(),
                                 _lib.stream_ptr(self.device))
        _lib.check(rc, "srmhd2d_dt")
        return self.dt

    def set_dt(self, dt):
        if torch.is_tensor(dt):
            self.dt.copy_(dt.reshape(1))
        else:
            self.dt.fill_(float(dt))

    def diag_dicts(self):
        dx, dy = _LazyDiag(), _LazyDiag()
        w = self.diag.clone()
        # the FP32-flux forms count their own solver's decisions, not the
        # reference's secant iterations
        counted = self.precision in ("fp64", "fp64_local")
        dx._set_pending(w[:_lib.DIAG_WORDS], diag_solver(self.solver), 0, iters_counted=counted)
        dy._set_pending(w[_lib.DIAG_WORDS:], diag_solver(self.solver), 1, iters_counted=counted)
        return {"x": {"diag": dx}, "y": {"diag": dy}}

    # ── resident step loop ─────────────────────────────────────────────────
    def load(self, U, P):
        """Adopt state U (8 tensors) and prims P (6) as the resident buffers."""
        self.U = [t.contiguous() for t in U]
        self.P = [t.contiguous() for t in P]
        self._spare = [self.new_state(), self.new_state()]
        self._pspare = [self.new_prims(), self.new_prims()]

    def step(self, scheme="rk3"):
        """One RK step with dt already in ``self.dt``.  Stage outputs
        alternate between two spare states (and two spare primitive sets);
        the step-start state U0 is only read."""
        U0, P0 = self.U, self.P
        s, ps = self._spare, self._pspare
        U, P = U0, P0
        stages = RK[scheme]
        for k, (a, b, c) in enumerate(stages):
            out, Pout = s[k % 2], ps[k % 2]
            self.stage(a, b, c, U0, U, P, out, Pout)
            U, P = out, Pout
        last = (len(stages) - 1) % 2
        self._spare = [U0, s[1 - last]]
        self._pspare = [P0, ps[1 - last]]
        self.U, self.P = U, P


_STEPPERS = {}


def _stepper(grid, eos, solver, limiter, bc_x, bc_y, upwind, atmo_rho, device,
             precision="fp64", emf_mode="solver"):
    key = (id(grid), grid.nx, grid.ny, grid.ng, float(grid.dx), float(grid.dy),
           float(eos.K), float(eos.gamma), getattr(eos, "gamma_th", None), solver, limiter,
           bc_x, bc_y, bool(upwind), float(atmo_rho), str(device), precision, emf_mode)
    st = _STEPPERS.get(key)
    if st is None:
        if len(_STEPPERS) > 8:
            _STEPPERS.clear()
        st = Stepper2D(grid, eos, solver=solver, limiter=limiter, bc_x=bc_x, bc_y=bc_y,
                       upwind=upwind, atmo_rho=atmo_rho, device=device, precision=precision,
                       emf_mode=emf_mode)
        _STEPPERS[key] = st
    return st


def _cell_B(state, grid, bc_x, bc_y):
    """prims Bx, By, Bz of sync_state (driver2d.py:208-216): face averages,
    then the cell BCs; Bz is the conserved Bz itself."""
    Bcx = 0.5 * (state.Bxf[:-1, :] + state.Bxf[1:, :])
    Bcy = 0.5 * (state.Byf[:, :-1] + state.Byf[:, 1:])
    grid.apply_bc_cc({"Bx": Bcx, "By": Bcy}, bc_x, bc_y)
    return Bcx, Bcy, state.cons["Bz"]


def _finish_prims(state, P6, grid, bc_x, bc_y):
    prims = {k: P6[i] for i, k in enumerate(PRIM6)}
    prims["Bx"], prims["By"], prims["Bz"] = _cell_B(state, grid, bc_x, bc_y)
    state.prims = prims
    return state


def _check_options(vel_var, emf_mode, recorder):
    if vel_var != "Wv":
        raise ValueError("the HIP path reconstructs W v (vel_var='Wv')")
    if emf_mode not in ("solver", "hll"):
        raise ValueError(f"unknown emf mode {emf_mode!r}")
    if recorder is not None:
        raise ValueError("interface recorders are not supported on the HIP path")


def sync_state(state, grid, eos, bc_x="outflow", bc_y="outflow", atmo_rho=1e-10,
               return_status=False, limiter="mc"):
    """driver2d.py:196-218: face BCs, cell BCs, con2prim of every cell as one
    batch, primitive BCs (in place on the state's arrays)."""
    if return_status:
        grid.apply_bc_face(state.Bxf, state.Byf, bc_x, bc_y)
        grid.apply_bc_cc(state.cons, bc_x, bc_y)
        Bcx = 0.5 * (state.Bxf[:-1, :] + state.Bxf[1:, :])
        Bcy = 0.5 * (state.Byf[:, :-1] + state.Byf[:, 1:])
        cons = {**state.cons, "Bx": Bcx, "By": Bcy}
        prims, status = cons_to_prims_2d(cons, grid, eos, atmo_rho, return_status=True)
        prims["Bx"], prims["By"], prims["Bz"] = Bcx, Bcy, cons["Bz"]
        grid.apply_bc_cc(prims, bc_x, bc_y)
        state.prims = prims
        return state, status
    st = _stepper(grid, eos, "hlld", limiter, bc_x, bc_y, True, atmo_rho, grid.device)
    U = [t.contiguous() for t in state.state_list()]
    for k, t in zip(EVOLVED_KEYS, U[:6]):
        state.cons[k] = t
    state.Bxf, state.Byf = U[6], U[7]
    P = st.new_prims()
    st.sync(U, P)
    return _finish_prims(state, P, grid, bc_x, bc_y)


def compute_dt_2d(prims, grid, eos, cfl, speed_of_light=False):
    """driver2d.py:52-82 (a host float, like the reference)."""
    if speed_of_light:
        return float(cfl / (1.0 / grid.dx + 1.0 / grid.dy))
    st = _stepper(grid, eos, "hlld", "mc", "outflow", "outflow", True, 1e-10, grid.device)
    P = [prims[k].contiguous() for k in PRIM6]
    B = [prims[k].contiguous() for k in ("Bx", "By", "Bz")]
    return float(st.compute_dt(None, P, cfl, P9=B).item())


def rk_step_ct(state, grid, eos, dt, *, scheme="rk3", bc_x="outflow", bc_y="outflow",
               flux_fn="hlld", limiter="mc", vel_var="Wv", emf_mode="solver", upwind=True,
               atmo_rho=1e-10, recorder=None, precision="fp64"):
    """One SSP-RK step with constrained transport (driver2d.py:221-249).
    ``flux_fn`` is one of this package's flux plugins, the reference's, or a
    name; ``dt`` a float or a device scalar tensor.  ``precision="mixed"``
    (no counterpart in the reference) runs reconstruction and the Riemann
    solver in FP32 and the rest of the stage in FP64."""
    _check_options(vel_var, emf_mode, recorder)
    st = _stepper(grid, eos, solver_name(flux_fn), limiter, bc_x, bc_y, upwind, atmo_rho,
                  grid.device, precision, emf_mode)
    st.set_dt(dt)
    U0 = [t.contiguous() for t in state.state_list()]
    U, P = U0, [state.prims[k].contiguous() for k in PRIM6]
    for a, b, c in RK[scheme]:
        out, Pout = st.new_state(), st.new_prims()
        st.stage(a, b, c, U0, U, P, out, Pout)
        U, P = out, Pout
    S = State2D({k: U[i] for i, k in enumerate(EVOLVED_KEYS)}, U[6], U[7])
    _finish_prims(S, P, grid, bc_x, bc_y)
    diag = st.diag_dicts()
    LAST_DIAG._set_pending(st.diag.clone()[_lib.DIAG_WORDS:], diag_solver(st.solver), 1,
                           iters_counted=st.precision in ("fp64", "fp64_local"))
    return S, diag


def prims_to_cons_2d(prims, grid):
    return primitive_to_conserved(prims)


def cons_to_prims_2d(cons, grid, eos, atmo_rho=1e-10, return_status=False):
    """driver2d.py:85-94: one batch over every cell."""
    return conservative_to_primitive(cons, eos, atmo_rho=atmo_rho, return_status=return_status)
