"""2D initial data.  Evaluated in float64 on the CPU from the grid's
coordinates (the same arithmetic on every host), then moved to the device by
``init_state``.  Every problem returns a corner-valued Az; the staggered
field is its discrete curl, so div B = 0 at t = 0 (initial_data2d.py:24-29).
"""

import math

import numpy as np
import torch

from .driver2d import prims_to_cons_2d, sync_state
from .state import EVOLVED_KEYS, State2D


def b_from_potential(Az, dx, dy):
    return (Az[:, 1:] - Az[:, :-1]) / dy, -(Az[1:, :] - Az[:-1, :]) / dx


def _corners(grid):
    XF = grid.xf.unsqueeze(1).expand(grid.nxt + 1, grid.nyt + 1)
    YF = grid.yf.unsqueeze(0).expand(grid.nxt + 1, grid.nyt + 1)
    return XF, YF


def magnetic_rotor(grid, eos, r0=0.1, rho_in=10.0, rho_out=1.0, omega=9.95, press=1.0, B0=1.0):
    """The reference's sharp relativistic rotor (initial_data2d.py:88-125)."""
    X, Y = grid.X, grid.Y
    inside = X ** 2 + Y ** 2 <= r0 * r0 * (1.0 + 16 * torch.finfo(torch.float64).eps)
    z = torch.zeros(grid.shape, dtype=torch.float64)
    P = {"rho": torch.where(inside, torch.full_like(z, rho_in), torch.full_like(z, rho_out)),
         "p": torch.full_like(z, press),
         "vx": torch.where(inside, -omega * Y, z), "vy": torch.where(inside, omega * X, z),
         "vz": z.clone(), "Bx": torch.full_like(z, B0), "By": z.clone(), "Bz": z.clone()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    return P, B0 * _corners(grid)[1]


def rotor_tapered(grid, eos, r0=0.1, r1=0.115, rho_in=10.0, rho_out=1.0, omega=9.95, press=1.0,
                  B0=1.0):
    """Tapered rotor: rho = 10 and omega = 9.95 for r <= r0, a linear taper
    f = (r1 - r)/(r1 - r0) of (rho - 1) and omega over r0 < r < r1, rho = 1
    at rest outside; p = 1 everywhere, B = (B0, 0, 0) from Az = B0 y."""
    X, Y = grid.X, grid.Y
    r = torch.from_numpy(np.sqrt((X * X + Y * Y).numpy()))   # correctly rounded
    f = torch.where(r <= r0, torch.ones_like(r),
                    torch.where(r < r1, (r1 - r) / (r1 - r0), torch.zeros_like(r)))
    om = omega * f
    z = torch.zeros(grid.shape, dtype=torch.float64)
    P = {"rho": rho_out + (rho_in - rho_out) * f, "p": torch.full_like(z, press),
         "vx": -om * Y, "vy": om * X, "vz": z.clone(),
         "Bx": torch.full_like(z, B0), "By": z.clone(), "Bz": z.clone()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    return P, B0 * _corners(grid)[1]


def orszag_tang_2pi(grid, eos, A=1.0, vmax=0.99, rho=1.0, press=1.0):
    """Relativistic Orszag-Tang on [0, 2 pi]^2: rho = p = 1,
    v = vmax/sqrt2 (-sin y, sin x, 0), Az = A (cos y + cos(2x)/2), so that
    B = A (-sin y, sin 2x, 0)."""
    X, Y = grid.X, grid.Y
    amp = vmax / math.sqrt(2.0)
    z = torch.zeros(grid.shape, dtype=torch.float64)
    P = {"rho": torch.full_like(z, rho), "p": torch.full_like(z, press),
         "vx": -amp * torch.sin(Y), "vy": amp * torch.sin(X), "vz": z.clone(),
         "Bx": -A * torch.sin(Y), "By": A * torch.sin(2.0 * X), "Bz": z.clone()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    XF, YF = _corners(grid)
    return P, A * (torch.cos(YF) + 0.5 * torch.cos(2.0 * XF))


PROBLEMS = {"rotor": magnetic_rotor, "rotor_tapered": rotor_tapered,
            "orszag_tang_2pi": orszag_tang_2pi}


def init_state(grid, eos, prims, Az, bc_x="outflow", bc_y="outflow", limiter="mc"):
    """run_2d.py:227-231: faces from Az, prim -> cons, sync, on the device."""
    dev = grid.device
    Bxf, Byf = b_from_potential(Az, grid.dx, grid.dy)
    Pd = {k: v.contiguous().to(dev) for k, v in prims.items()}
    c = prims_to_cons_2d(Pd, grid)
    st = State2D({k: c[k].contiguous() for k in EVOLVED_KEYS}, Bxf.contiguous().to(dev),
                 Byf.contiguous().to(dev), Pd)
    return sync_state(st, grid, eos, bc_x, bc_y, limiter=limiter)
