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


def libm(fn, v):
    """fn (math.sin or math.cos, the C library's) of every element of the 1-D
    float64 tensor v.  torch's vectorised CPU sin/cos round differently on
    different CPUs; the C library gives the same values on every host."""
    return torch.tensor([fn(a) for a in v.tolist()], dtype=torch.float64)


def orszag_tang_2pi(grid, eos, A=1.0, vmax=0.99, rho=1.0, press=1.0):
    """Relativistic Orszag-Tang on [0, 2 pi]^2: rho = p = 1,
    v = vmax/sqrt2 (-sin y, sin x, 0), Az = A (cos y + cos(2x)/2), so that
    B = A (-sin y, sin 2x, 0).  The sines and cosines are taken on the 1-D
    coordinate vectors with the C library (the same on every host) and
    broadcast."""
    nx, ny = grid.nxt, grid.nyt
    sx = libm(math.sin, grid.x)[:, None].expand(nx, ny)
    sy = libm(math.sin, grid.y)[None, :].expand(nx, ny)
    s2x = libm(math.sin, 2.0 * grid.x)[:, None].expand(nx, ny)
    amp = vmax / math.sqrt(2.0)
    z = torch.zeros(grid.shape, dtype=torch.float64)
    P = {"rho": torch.full_like(z, rho), "p": torch.full_like(z, press),
         "vx": -amp * sy, "vy": amp * sx, "vz": z.clone(),
         "Bx": -A * sy, "By": A * s2x, "Bz": z.clone()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    cyf = libm(math.cos, grid.yf)[None, :].expand(nx + 1, ny + 1)
    c2xf = libm(math.cos, 2.0 * grid.xf)[:, None].expand(nx + 1, ny + 1)
    return P, A * (cyf + 0.5 * c2xf)


def orszag_tang(grid, eos, v_max=0.99, B0=None, press=None, rho=None):
    """The reference's relativistic Orszag-Tang vortex on the periodic unit
    square (initial_data2d.py:194-253): rho = 25/(36 pi), p = 5/(12 pi),
    B0 = 1/sqrt(4 pi), v = v_max/sqrt2 (-sin 2 pi y, sin 2 pi x, 0),
    B = B0 (-sin 2 pi y, sin 4 pi x, 0) from
    Az = B0 (cos(4 pi x)/(4 pi) + cos(2 pi y)/(2 pi))."""
    if B0 is None:
        B0 = 1.0 / math.sqrt(4.0 * math.pi)
    if rho is None:
        rho = 25.0 / (36.0 * math.pi)
    if press is None:
        press = 5.0 / (12.0 * math.pi)
    amp = v_max / math.sqrt(2.0)
    X, Y = grid.X, grid.Y
    z = torch.zeros(grid.shape, dtype=torch.float64)
    P = {"rho": torch.full_like(z, rho), "p": torch.full_like(z, press),
         "vx": -amp * torch.sin(2 * math.pi * Y), "vy": amp * torch.sin(2 * math.pi * X),
         "vz": z.clone(), "Bx": -B0 * torch.sin(2 * math.pi * Y),
         "By": B0 * torch.sin(4 * math.pi * X), "Bz": z.clone()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    XF, YF = _corners(grid)
    Az = B0 * (torch.cos(4 * math.pi * XF) / (4 * math.pi)
               + torch.cos(2 * math.pi * YF) / (2 * math.pi))
    return P, Az


def orszag_tang_bs(grid, eos):
    """Beckwith & Stone's flow speed, 0.5 per component
    (initial_data2d.py:256-270)."""
    return orszag_tang(grid, eos, v_max=0.5 * math.sqrt(2.0))


def riemann2d(grid, eos, B0=0.5, angle_deg=45.0, angle_z_deg=0.0):
    """The four-quadrant Riemann problem threaded by a uniform field
    B = B0 (cos a cos b, sin a cos b, sin b) (initial_data2d.py:128-191):
    (rho, p, vx, vy) top left (0.1, 1, 0.99, 0), top right (0.1, 0.01, 0,
    0), bottom left (0.5, 1, 0, 0), bottom right (0.1, 1, 0, 0.99); at
    45 degrees sin a is taken equal to cos a (the problem's x <-> y
    symmetry)."""
    X, Y = grid.X, grid.Y
    z = torch.zeros(grid.shape, dtype=torch.float64)
    right, top = X > 0.0, Y > 0.0

    def full(v):
        return torch.full_like(z, float(v))

    def pick(tl, tr, bl, br):
        return torch.where(top, torch.where(right, full(tr), full(tl)),
                           torch.where(right, full(br), full(bl)))

    a, b = math.radians(angle_deg), math.radians(angle_z_deg)
    ca, sa = math.cos(a), math.sin(a)
    if angle_deg == 45.0:
        sa = ca
    bx, by, bz = B0 * ca * math.cos(b), B0 * sa * math.cos(b), B0 * math.sin(b)
    P = {"rho": pick(0.1, 0.1, 0.5, 0.1), "p": pick(1.0, 0.01, 1.0, 1.0),
         "vx": pick(0.99, 0.0, 0.0, 0.0), "vy": pick(0.0, 0.0, 0.0, 0.99), "vz": z.clone(),
         "Bx": full(bx), "By": full(by), "Bz": full(bz)}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    XF, YF = _corners(grid)
    return P, bx * YF - by * XF


PROBLEMS = {"rotor": magnetic_rotor, "rotor_tapered": rotor_tapered,
            "orszag_tang_2pi": orszag_tang_2pi, "orszag_tang": orszag_tang,
            "orszag_tang_bs": orszag_tang_bs, "riemann2d": riemann2d}


def init_state(grid, eos, prims, Az, bc_x="outflow", bc_y="outflow", limiter="mc"):
    """run_2d.py:227-231: faces from Az, prim -> cons, sync, on the device."""
    dev = grid.device
    Bxf, Byf = b_from_potential(Az, grid.dx, grid.dy)
    Pd = {k: v.contiguous().to(dev) for k, v in prims.items()}
    c = prims_to_cons_2d(Pd, grid)
    st = State2D({k: c[k].contiguous() for k in EVOLVED_KEYS}, Bxf.contiguous().to(dev),
                 Byf.contiguous().to(dev), Pd)
    return sync_state(st, grid, eos, bc_x, bc_y, limiter=limiter)
