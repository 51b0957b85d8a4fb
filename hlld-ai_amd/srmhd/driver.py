"""1D shock-tube driver on the HIP path: drop-in for src/physics/driver.py
(``run(cfg)``, ``rk2_step``, ``compute_dt``) with the same YAML-derived config
dict (grid / eos / run / output blocks).

Each Heun step is one C-ABI call (``srmhd1d_rk2_step``); the time step is a
device reduction read back once per step, as the reference's ``.item()``.
"""

import ctypes
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .eos import as_c, hybrid_eos
from .hlld import LAST_DIAG, diag_solver, primitive_to_conserved, solver_name

PRIM_KEYS = ("rho", "vx", "vy", "vz", "p", "eps", "Bx", "By", "Bz")
CONS_C = ("D", "tau", "Sx", "Sy", "Sz", "Bx", "By", "Bz")      # the C ABI's order
_HYDRO_PRIM_KEYS = ["rho", "vx", "vy", "vz", "p", "eps"]
_HYDRO_CONS_KEYS = ["D", "Sx", "Sy", "Sz", "tau"]
_MAG_KEYS = ["Bx", "By", "Bz"]


class Grid1D:
    """grid.py:23-47 (coordinates on the CPU, state on ``device``)."""

    def __init__(self, xmin, xmax, ncells, ng=2, device="cuda"):
        self.xmin, self.xmax, self.ncells, self.ng = xmin, xmax, ncells, ng
        self.device = torch.device(device)
        self.dx = (xmax - xmin) / ncells
        self.ntotal = ncells + 2 * ng
        i = torch.arange(self.ntotal, dtype=torch.float64)
        self.x = xmin + (i - ng + 0.5) * self.dx
        self.phys = slice(ng, ng + ncells)

    @property
    def x_phys(self):
        return self.x[self.phys]

    def apply_outflow_bc(self, U):
        ng, n = self.ng, self.ncells
        for v in U.values():
            v[:ng] = v[ng]
            v[ng + n:] = v[ng + n - 1]
        return U


class Stepper1D:
    def __init__(self, grid, eos, solver="hlld", limiter="pcm", bc="outflow", atmo_rho=1e-10):
        self.grid = grid
        self.lib = _lib.lib()
        self.ceos = as_c(eos)
        self.solver = solver_name(solver)
        if limiter not in _lib.LIMITER_IDS:
            raise ValueError(f"unknown limiter {limiter!r}")
        self.limiter = limiter
        self.cg = _lib.Grid1D(grid.ncells, grid.ng, _lib.BC_IDS[bc], 0, float(grid.dx))
        nb = self.lib.srmhd1d_workspace_bytes(ctypes.byref(self.cg))
        if nb < 0:
            raise ValueError("grid rejected by srmhd1d_workspace_bytes")
        self.ws = _lib.workspace(nb, grid.device)
        self.dt = torch.zeros(1, dtype=torch.float64, device=grid.device)
        self.diag = torch.zeros(2 * _lib.DIAG_WORDS, dtype=torch.int64, device=grid.device)
        self.atmo_rho = float(atmo_rho)

    def compute_dt(self, P, cfl, speed_of_light):
        a, keep = _lib.ptrs(P)
        rc = self.lib.srmhd1d_dt(ctypes.byref(self.cg), ctypes.byref(self.ceos), float(cfl),
                                 1 if speed_of_light else 0, a, self.dt.data_ptr(),
                                 _lib.stream_ptr(self.grid.device))
        _lib.check(rc, "srmhd1d_dt")
        return self.dt

    def step(self, C, P):
        """One Heun step from (C, P) into new tensors; dt from ``self.dt``."""
        d = self.grid.device
        n = self.grid.ntotal
        Co = [torch.empty(n, dtype=torch.float64, device=d) for _ in CONS_C]
        Po = [torch.empty(n, dtype=torch.float64, device=d) for _ in PRIM_KEYS]
        ac, kc = _lib.ptrs(C)
        ap, kp = _lib.ptrs(P)
        aco, kco = _lib.ptrs(Co)
        apo, kpo = _lib.ptrs(Po)
        from . import hlld as H
        o = H.hlld_opts()   # the p* switches at call time, as hlld_flux reads them
        rc = self.lib.srmhd1d_rk2_step_ex(ctypes.byref(self.cg), _lib.SOLVER_IDS[self.solver],
                                          _lib.LIMITER_IDS[self.limiter], ctypes.byref(self.ceos),
                                          ctypes.byref(o), self.atmo_rho, self.dt.data_ptr(), ac,
                                          ap, aco, apo, self.diag.data_ptr(), self.ws.data_ptr(),
                                          self.ws.numel(), _lib.stream_ptr(d))
        _lib.check(rc, "srmhd1d_rk2_step_ex")
        return Co, Po


def init_shocktube(grid, primL, primR, x_interface):
    """initial_data.init_shocktube (initial_data.py:100-127) on the CPU."""
    mask = grid.x < x_interface
    return {k: torch.where(mask, primL[k].expand_as(mask), primR[k].expand_as(mask))
            .to(torch.float64) for k in primL}


def generic(primL, primR, x_interface=0.5, eos=None):
    """initial_data.generic (initial_data.py:78-95)."""
    pL = {k: torch.tensor(v, dtype=torch.float64) for k, v in primL.items()}
    pR = {k: torch.tensor(v, dtype=torch.float64) for k, v in primR.items()}
    if eos is not None:
        for P in (pL, pR):
            if "eps" not in P:
                P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    return pL, pR, x_interface


def save_snapshot(path, t, grid, prims, cons):
    """driver.py:245-268 (physical region, npz)."""
    ph = grid.phys
    data = {"t": t, "x": grid.x_phys.numpy()}
    for k in _HYDRO_PRIM_KEYS:
        data[f"prim_{k}"] = prims[k][ph].cpu().numpy()
    for k in _HYDRO_CONS_KEYS:
        data[f"cons_{k}"] = cons[k][ph].cpu().numpy()
    for k in _MAG_KEYS:
        data[k] = cons[k][ph].cpu().numpy()
    np.savez_compressed(path, **data)


def run(cfg, write_snapshots=True):
    """driver.run (driver.py:274-436) with the device taken from run.device
    (default cuda).  Returns (t, grid, prims, cons)."""
    gc, ec, rc, oc = cfg["grid"], cfg["eos"], cfg["run"], cfg.get("output", {})
    device = rc.get("device", "cuda")
    if torch.device(device).type != "cuda":
        raise RuntimeError("srmhd.driver.run runs on a HIP device (run.device: cuda)")
    eos = hybrid_eos(K=ec["K"], gamma=ec["gamma"], gamma_th=ec.get("gamma_th", None))
    grid = Grid1D(gc["xmin"], gc["xmax"], gc["ncells"], gc.get("ng", 2), device=device)
    problem = rc.get("problem", "generic")
    if problem != "generic":
        raise ValueError("the HIP 1D driver takes run.problem: generic (primL / primR blocks)")
    pL, pR, x_int = generic(rc["primL"], rc["primR"], rc.get("x_interface", 0.5), eos=eos)
    prims = {k: v.to(grid.device) for k, v in init_shocktube(grid, pL, pR, x_int).items()}
    cons = {k: v.clone() for k, v in primitive_to_conserved(prims).items()}
    grid.apply_outflow_bc(prims)
    grid.apply_outflow_bc(cons)
    t_end, cfl = rc["t_end"], rc.get("cfl", 0.4)
    atmo_rho = rc.get("atmo_rho", 1e-10)
    solver = rc.get("solver", "hlld")
    if solver in ("exact", "exact_batched"):
        raise ValueError(f"solver {solver!r} is not on the HIP path (hlle, hllc, hlld, linear)")
    solver = solver if solver in ("hlle", "hllc", "hlld", "linear") else "hlld"   # driver.py:379-380
    limiter = rc.get("limiter", "pcm")
    dt_c = bool(rc.get("dt_speed_of_light", True))
    out_dir = Path(oc.get("dir", "output"))
    out_every = oc.get("every_n_steps", 100)
    out_dt_max = oc.get("dt_snapshot", None)
    if write_snapshots:
        out_dir.mkdir(parents=True, exist_ok=True)
    stp = Stepper1D(grid, eos, solver, limiter, "outflow", atmo_rho)
    C = [cons[k].contiguous() for k in CONS_C]
    P = [prims[k].contiguous() for k in PRIM_KEYS]
    t, step, t_next = 0.0, 0, 0.0
    while t < t_end:
        dt = float(stp.compute_dt(P, cfl, dt_c).item())
        dt = min(dt, t_end - t)
        if dt <= 0.0:
            break
        if step == 0 and write_snapshots:
            save_snapshot(out_dir / f"snap_{step:06d}.npz", t, grid,
                          dict(zip(PRIM_KEYS, P)), dict(zip(CONS_C, C)))
        stp.dt.fill_(dt)
        C, P = stp.step(C, P)
        t += dt
        step += 1
        do_snap = (step % out_every == 0) or (t >= t_end)
        if out_dt_max is not None:
            do_snap = do_snap or (t >= t_next)
        if do_snap and write_snapshots:
            save_snapshot(out_dir / f"snap_{step:06d}.npz", t, grid,
                          dict(zip(PRIM_KEYS, P)), dict(zip(CONS_C, C)))
            t_next = t + (out_dt_max or 0.0)
    LAST_DIAG._set_pending(stp.diag.clone()[_lib.DIAG_WORDS:], diag_solver(stp.solver), 0)
    return t, grid, dict(zip(PRIM_KEYS, P)), dict(zip(CONS_C, C))
