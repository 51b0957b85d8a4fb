"""1D shock-tube driver (oracle restatement of src/physics/driver.py and
src/physics/initial_data.py:78-127)."""

import torch

from .c2p import cons_to_prim
from .eos import HybridEOS
from .grid import Grid1D
from .physics import prim_to_cons, sqrt
from .recon import interface_states_1d
from .riemann import SOLVERS, hlld_flux

CONS_KEYS = ["D", "Sx", "Sy", "Sz", "tau", "Bx", "By", "Bz"]     # driver.py:43-48


def compute_dt(prims, eos, dx, cfl, ng, ncells, speed_of_light=False):
    """driver.py:77-127 (0-d tensor)."""
    ph = slice(ng, ng + ncells)
    rho, eps = prims["rho"][ph], prims["eps"][ph]
    vx, vy, vz = prims["vx"][ph], prims["vy"][ph], prims["vz"][ph]
    Bx, By, Bz = prims["Bx"][ph], prims["By"][ph], prims["Bz"][ph]
    p, cs2 = eos.press_and_cs2(eps, rho)
    v2 = vx ** 2 + vy ** 2 + vz ** 2
    W = 1.0 / sqrt(torch.clamp(1.0 - v2, min=1e-10))
    B2 = Bx ** 2 + By ** 2 + Bz ** 2
    vB = vx * Bx + vy * By + vz * Bz
    b2 = (B2 + vB ** 2) / W ** 2
    h = 1.0 + eps + p / rho
    v02 = cs2 + (b2 / (rho * h + b2)) * (1.0 - cs2)
    vf = sqrt(torch.clamp(v02, min=0.0, max=1.0))
    smax = torch.clamp((vx.abs() + vf).max(), min=1e-14)
    if speed_of_light:
        smax = torch.ones_like(smax)
    return (cfl * dx) / smax


def compute_rhs(prims, grid, eos, flux_fn=hlld_flux, limiter="pcm"):
    """driver.py:135-186: physical faces only, F[Bx] = 0, ghosts get 0."""
    ng, n = grid.ng, grid.ncells
    PL, PR = interface_states_1d(prims, ng, n, limiter)
    F, _, _ = flux_fn(PL, PR, eos, idir=0)
    F["Bx"] = torch.zeros_like(F["Bx"])
    rhs = {}
    for k in CONS_KEYS:
        f = F[k]
        arr = torch.zeros(grid.ntotal, dtype=f.dtype)
        arr[ng:ng + n] = +((f[:-1] - f[1:]) / grid.dx)
        rhs[k] = arr
    return rhs


def rk2_step(cons, prims, grid, eos, dt, atmo_rho, bc="outflow", flux_fn=hlld_flux,
             limiter="pcm"):
    """Heun form   driver.py:192-236."""
    def _bc(U, U_old):
        if bc == "outflow":
            return grid.outflow(U)
        if bc == "constant":
            return grid.constant(U, U_old)
        return grid.periodic(U)

    k1 = compute_rhs(prims, grid, eos, flux_fn, limiter)
    c1 = _bc({k: cons[k] + dt * k1[k] for k in CONS_KEYS}, cons)
    c1["Bx"] = cons["Bx"].clone()
    p1 = _bc(cons_to_prim(c1, eos, atmo_rho=atmo_rho), prims)
    k2 = compute_rhs(p1, grid, eos, flux_fn, limiter)
    cn = _bc({k: 0.5 * (cons[k] + c1[k] + dt * k2[k]) for k in CONS_KEYS}, cons)
    cn["Bx"] = cons["Bx"].clone()
    pn = _bc(cons_to_prim(cn, eos, atmo_rho=atmo_rho), prims)
    return cn, pn


def init_generic(grid, eos, primL, primR, x_int=0.5):
    """initial_data.generic + init_shocktube (initial_data.py:78-127)."""
    pL = {k: torch.tensor(v, dtype=torch.float64) for k, v in primL.items()}
    pR = {k: torch.tensor(v, dtype=torch.float64) for k, v in primR.items()}
    for P in (pL, pR):
        if "eps" not in P:
            P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    mask = grid.x < x_int
    return {k: torch.where(mask, pL[k].expand_as(mask), pR[k].expand_as(mask)).to(torch.float64)
            for k in pL}


def run(cfg, max_steps=None):
    """driver.run(cfg) without snapshots (driver.py:274-436): the YAML keys
    grid/eos/run are read the same way; an unknown solver name selects HLLD.
    Returns (t, grid, prims, cons, steps)."""
    gc, ec, rc = cfg["grid"], cfg["eos"], cfg["run"]
    eos = HybridEOS(ec["K"], ec["gamma"], ec.get("gamma_th", None))
    grid = Grid1D(gc["xmin"], gc["xmax"], gc["ncells"], gc.get("ng", 2))
    prims = init_generic(grid, eos, rc["primL"], rc["primR"], rc.get("x_interface", 0.5))
    cons = prim_to_cons(prims)
    grid.outflow(prims)
    grid.outflow(cons)
    t_end, cfl = rc["t_end"], rc.get("cfl", 0.4)
    atmo = rc.get("atmo_rho", 1e-10)
    if rc.get("solver") == "linear":              # driver.py:378-380
        from .linear import linearised_flux as flux_fn
    else:
        flux_fn = SOLVERS.get(rc.get("solver", "hlld"), hlld_flux)
    limiter = rc.get("limiter", "pcm")
    dt_c = bool(rc.get("dt_speed_of_light", True))
    t, step = 0.0, 0
    while t < t_end:
        if max_steps is not None and step >= max_steps:
            break
        dt = compute_dt(prims, eos, grid.dx, cfl, grid.ng, grid.ncells, dt_c).item()
        dt = min(dt, t_end - t)
        if dt <= 0.0:
            break
        cons, prims = rk2_step(cons, prims, grid, eos, torch.tensor(dt, dtype=torch.float64),
                               atmo, flux_fn=flux_fn, limiter=limiter)
        t += dt
        step += 1
    return t, grid, prims, cons, step


def balsara1_cfg(solver="hlld", limiter="mc"):
    """configs/Giacomazzo/Balsara1.yaml with the solver and limiter set."""
    return {"grid": {"xmin": 0.0, "xmax": 1.0, "ncells": 400, "ng": 2},
            "eos": {"K": 0.0, "gamma": 2.0, "gamma_th": 2.0},
            "run": {"problem": "generic", "t_end": 0.4, "cfl": 0.4, "atmo_rho": 1.0e-10,
                    "solver": solver, "limiter": limiter,
                    "primL": {"rho": 1.0, "p": 1.0, "vx": 0.0, "vy": 0.0, "vz": 0.0,
                              "Bx": 0.5, "By": 1.0, "Bz": 0.0},
                    "primR": {"rho": 0.125, "p": 0.1, "vx": 0.0, "vy": 0.0, "vz": 0.0,
                              "Bx": 0.5, "By": -1.0, "Bz": 0.0},
                    "x_interface": 0.5}}
