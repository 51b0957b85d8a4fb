"""2D SRMHD step with constrained transport (oracle restatement of
src/physics/{sweep,ct,state,driver2d,initial_data2d}.py)."""

import math
from dataclasses import dataclass, field

import torch

from . import physics as PH
from .c2p import cons_to_prim
from .grid import Grid2D
from .physics import hll_flux, prim_to_cons, sqrt, srmhd_fluxes
from .recon import reconstruct_prims
from .riemann import LAST_DIAG, hlld_flux

EVOLVED_KEYS = ["D", "Sx", "Sy", "Sz", "tau", "Bz"]      # state.py:41
_BKEY = ("Bx", "By", "Bz")
RK = {"rk2": ((1.0, 0.0, 1.0), (0.5, 0.5, 0.5)),          # driver2d.py:104-105
      "rk3": ((1.0, 0.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0))}


def flat(d):
    return {k: v.reshape(-1) for k, v in d.items()}


def unflat(d, shape):
    return {k: v.reshape(*shape) for k, v in d.items()}


@dataclass
class State2D:
    """state.py:55-74."""
    cons: dict
    Bxf: torch.Tensor
    Byf: torch.Tensor
    prims: dict = field(default_factory=dict)

    def clone(self):
        return State2D({k: v.clone() for k, v in self.cons.items()}, self.Bxf.clone(),
                       self.Byf.clone(), {k: v.clone() for k, v in self.prims.items()})


def combine(states, weights):
    """sum_k w_k state_k, left to right, cells and faces alike (state.py:77-103)."""
    w0, s0 = weights[0], states[0]
    cons = {k: w0 * s0.cons[k] for k in s0.cons}
    Bxf, Byf = w0 * s0.Bxf, w0 * s0.Byf
    for w, s in zip(weights[1:], states[1:]):
        for k in cons:
            cons[k] = cons[k] + w * s.cons[k]
        Bxf = Bxf + w * s.Bxf
        Byf = Byf + w * s.Byf
    return State2D(cons, Bxf, Byf)


def sweep(prims, eos, axis, flux_fn=hlld_flux, limiter="mc", Bnf=None, want_hll_aux=False):
    """reconstruct -> B_n := face value -> flux_fn on flattened faces ->
    zero the normal-B flux   sweep.py:30-119.  Returns (F, p_star, diag), and
    with want_hll_aux the side states and HLL speeds of compute_srmhd_fluxes
    (sweep.py:106-118) as a fourth item."""
    L, R = reconstruct_prims(prims, axis, eos, limiter)
    shape = L["rho"].shape
    if Bnf is not None:
        if Bnf.shape != shape:
            raise ValueError(f"Bnf shape {tuple(Bnf.shape)} != face shape {tuple(shape)}")
        L[_BKEY[axis]] = Bnf
        R[_BKEY[axis]] = Bnf
    F, _, ps = flux_fn(flat(L), flat(R), eos, idir=axis)
    F = unflat(F, shape)
    F[_BKEY[axis]] = torch.zeros_like(F[_BKEY[axis]])
    if not want_hll_aux:
        return F, ps.reshape(shape), dict(LAST_DIAG)
    diag = dict(LAST_DIAG)
    uL, uR, fL, fR, cmax, cmin = srmhd_fluxes(flat(L), flat(R), eos, axis)
    hll = {"uL": unflat(uL, shape), "uR": unflat(uR, shape), "fL": unflat(fL, shape),
           "fR": unflat(fR, shape), "cmax": cmax.reshape(shape), "cmin": cmin.reshape(shape)}
    return F, ps.reshape(shape), diag, hll


def face_emf_x(Fx, hll=None, mode="solver"):
    """E^z on x-faces   ct.py:63-74: -F^x(B_y), or with mode "hll" minus the
    HLL average of B_y's flux from the side states."""
    if mode == "solver":
        return -Fx["By"]
    if mode == "hll":
        return -hll_flux(hll["fL"]["By"], hll["fR"]["By"], hll["uL"]["By"], hll["uR"]["By"],
                         hll["cmin"], hll["cmax"])
    raise ValueError(f"unknown emf mode {mode!r}")


def face_emf_y(Fy, hll=None, mode="solver"):
    """E^z on y-faces   ct.py:77-88: F^y(B_x), or its HLL average."""
    if mode == "solver":
        return Fy["Bx"]
    if mode == "hll":
        return hll_flux(hll["fL"]["Bx"], hll["fR"]["Bx"], hll["uL"]["Bx"], hll["uR"]["Bx"],
                        hll["cmin"], hll["cmax"])
    raise ValueError(f"unknown emf mode {mode!r}")


def flux_divergence(F, axis, delta, keys):
    """-(F[i+1] - F[i]) / delta over the full block   sweep.py:122-135."""
    lo = [slice(None)] * 2
    hi = [slice(None)] * 2
    lo[axis], hi[axis] = slice(None, -1), slice(1, None)
    return {k: -(F[k][tuple(hi)] - F[k][tuple(lo)]) / delta for k in keys}


def cell_centred_B(Bxf, Byf):
    """ct.py:44-54."""
    return 0.5 * (Bxf[:-1, :] + Bxf[1:, :]), 0.5 * (Byf[:, :-1] + Byf[:, 1:])


def upwind_sign(F, rel_tol):
    """sign(F) with a deadband scaled by the sweep-wide max |F| (ct.py:96-116)."""
    scale = F.abs().max()
    return torch.where(F.abs() > rel_tol * scale, torch.sign(F), torch.zeros_like(F))


def corner_emf(Ec, Efx, Efy, Fx_rho, Fy_rho, upwind=True, tol=1e-9):
    """Gardiner-Stone corner EMF, ring of zeros   ct.py:119-177."""
    nxt, nyt = Ec.shape
    Ez = torch.zeros(nxt + 1, nyt + 1, dtype=Ec.dtype)
    I, Im = slice(1, nxt), slice(0, nxt - 1)
    J, Jm = slice(1, nyt), slice(0, nyt - 1)
    eyf, eyfm = Efy[I, J], Efy[Im, J]
    exf, exfm = Efx[I, J], Efx[I, Jm]
    NE, NW, SE, SW = Ec[I, J], Ec[Im, J], Ec[I, Jm], Ec[Im, Jm]
    avg = 0.25 * ((eyf + eyfm) + (exf + exfm))
    if not upwind:
        Ez[I, J] = avg
        return Ez
    sx = upwind_sign(Fx_rho, tol)
    Sx, Sxm = sx[I, J], sx[I, Jm]
    dN = (1.0 - Sx) * (NE - eyf) + (1.0 + Sx) * (NW - eyfm)
    dS = (1.0 - Sxm) * (eyf - SE) + (1.0 + Sxm) * (eyfm - SW)
    dEdy = 0.125 * (dS - dN)
    sy = upwind_sign(Fy_rho, tol)
    Sy, Sym = sy[I, J], sy[Im, J]
    dE = (1.0 - Sy) * (NE - exf) + (1.0 + Sy) * (SE - exfm)
    dW = (1.0 - Sym) * (exf - NW) + (1.0 + Sym) * (exfm - SW)
    dEdx = 0.125 * (dW - dE)
    Ez[I, J] = avg + (dEdy + dEdx)
    return Ez


def ct_rhs(Ez, dx, dy):
    """ct.py:180-185."""
    return (Ez[:, :-1] - Ez[:, 1:]) / dy, (Ez[1:, :] - Ez[:-1, :]) / dx


def div_b(Bxf, Byf, dx, dy):
    """ct.py:196-199."""
    return (Bxf[1:, :] - Bxf[:-1, :]) / dx + (Byf[:, 1:] - Byf[:, :-1]) / dy


def compute_dt_2d(prims, grid, eos, cfl, speed_of_light=False):
    """Unsplit CFL with the reference's b2 = (B^2 + (v.B)^2)/W^2
    (driver2d.py:52-82).  Returned as a Python float, like the reference."""
    ph = grid.phys
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
    cf = sqrt(torch.clamp(v02, min=0.0, max=1.0))
    if speed_of_light:
        sx = sy = torch.ones((), dtype=rho.dtype)
    else:
        sx = torch.clamp((vx.abs() + cf).max(), min=1e-14)
        sy = torch.clamp((vy.abs() + cf).max(), min=1e-14)
    return float(cfl / (sx / grid.dx + sy / grid.dy))


def prims_to_cons_2d(prims, grid):
    return unflat(prim_to_cons(flat(prims)), grid.shape)


def cons_to_prims_2d(cons, grid, eos, atmo_rho=1e-10, return_status=False):
    """driver2d.py:85-94."""
    out = cons_to_prim(flat(cons), eos, atmo_rho=atmo_rho, return_status=return_status)
    if return_status:
        P, S = out
        return unflat(P, grid.shape), {k: v.reshape(*grid.shape) for k, v in S.items()}
    return unflat(out, grid.shape)


def ct_stage(state, grid, eos, flux_fn=hlld_flux, limiter="mc", upwind=True, emf_mode="solver"):
    """driver2d.py:162-193.  Returns (rhs, dBxf, dByf, aux) where aux carries
    the face fluxes and diag."""
    if emf_mode not in ("solver", "hll"):
        raise ValueError(f"unknown emf mode {emf_mode!r}")
    Bcx, Bcy = cell_centred_B(state.Bxf, state.Byf)
    prims = dict(state.prims)
    prims["Bx"], prims["By"] = Bcx, Bcy
    Ec = Bcx * prims["vy"] - Bcy * prims["vx"]
    want = emf_mode == "hll"
    ox = sweep(prims, eos, 0, flux_fn, limiter, Bnf=state.Bxf, want_hll_aux=want)
    oy = sweep(prims, eos, 1, flux_fn, limiter, Bnf=state.Byf, want_hll_aux=want)
    (Fx, psx, dgx), (Fy, psy, dgy) = ox[:3], oy[:3]
    Efx = face_emf_x(Fx, ox[3] if want else None, emf_mode)
    Efy = face_emf_y(Fy, oy[3] if want else None, emf_mode)
    Ez = corner_emf(Ec, Efx, Efy, Fx["D"], Fy["D"], upwind=upwind)
    rx = flux_divergence(Fx, 0, grid.dx, EVOLVED_KEYS)
    ry = flux_divergence(Fy, 1, grid.dy, EVOLVED_KEYS)
    rhs = {k: rx[k] + ry[k] for k in EVOLVED_KEYS}
    dBx, dBy = ct_rhs(Ez, grid.dx, grid.dy)
    return rhs, dBx, dBy, {"Fx": Fx, "Fy": Fy, "Ez": Ez, "pstar_x": psx, "pstar_y": psy,
                           "x": {"diag": dgx}, "y": {"diag": dgy}}


def sync_state(state, grid, eos, bc_x="outflow", bc_y="outflow", atmo_rho=1e-10):
    """Face BCs -> cell BCs -> B_c -> con2prim on ALL cells -> prim BCs
    (driver2d.py:196-218)."""
    grid.bc_faces(state.Bxf, state.Byf, bc_x, bc_y)
    grid.bc_cells(state.cons, bc_x, bc_y)
    Bcx, Bcy = cell_centred_B(state.Bxf, state.Byf)
    cons = dict(state.cons)
    cons["Bx"], cons["By"] = Bcx, Bcy
    # copies: the BC fill below writes in place, and the con2prim keeps its
    # inputs and outputs for torch autograd (same values as the reference's
    # aliases)
    prims = {k: v.clone() for k, v in cons_to_prims_2d(cons, grid, eos, atmo_rho).items()}
    prims["Bx"], prims["By"] = Bcx.clone(), Bcy.clone()
    prims["Bz"] = cons["Bz"].clone()
    grid.bc_cells(prims, bc_x, bc_y)
    state.prims = prims
    return state


def rk_step_ct(state, grid, eos, dt, scheme="rk3", bc_x="outflow", bc_y="outflow",
               flux_fn=hlld_flux, limiter="mc", upwind=True, atmo_rho=1e-10, emf_mode="solver"):
    """driver2d.py:221-249."""
    S0 = State2D({k: state.cons[k].clone() for k in EVOLVED_KEYS}, state.Bxf.clone(),
                 state.Byf.clone(), state.prims)
    S = state
    aux = {}
    for a, b, c in RK[scheme]:
        rhs, dBx, dBy, aux = ct_stage(S, grid, eos, flux_fn, limiter, upwind, emf_mode)
        S = combine([S0, S, State2D(rhs, dBx, dBy)], [a, b, c * dt])
        S = sync_state(S, grid, eos, bc_x, bc_y, atmo_rho)
    return S, aux


# ── initial data ─────────────────────────────────────────────────────────────

def b_from_potential(Az, dx, dy):
    """Discrete curl of a corner potential   initial_data2d.py:24-29."""
    return (Az[:, 1:] - Az[:, :-1]) / dy, -(Az[1:, :] - Az[:-1, :]) / dx


def _corners(grid):
    XF = grid.xf.unsqueeze(1).expand(grid.nxt + 1, grid.nyt + 1)
    YF = grid.yf.unsqueeze(0).expand(grid.nxt + 1, grid.nyt + 1)
    return XF, YF


def rotor_sharp(grid, eos, r0=0.1, rho_in=10.0, rho_out=1.0, omega=9.95, press=1.0, B0=1.0):
    """The reference's sharp rotor   initial_data2d.py:88-125."""
    X, Y = grid.X, grid.Y
    inside = X ** 2 + Y ** 2 <= r0 * r0 * (1.0 + 16 * torch.finfo(torch.float64).eps)
    z = torch.zeros(grid.shape, dtype=torch.float64)
    P = {"rho": torch.where(inside, torch.full_like(z, rho_in), torch.full_like(z, rho_out)),
         "p": torch.full_like(z, press),
         "vx": torch.where(inside, -omega * Y, z), "vy": torch.where(inside, omega * X, z),
         "vz": z.clone(), "Bx": torch.full_like(z, B0), "By": z.clone(), "Bz": z.clone()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    return P, B0 * _corners(grid)[1]


def rotor_tapered(grid, eos, r0=0.1, r1=0.115, rho_in=10.0, rho_out=1.0, omega=9.95,
                  press=1.0, B0=1.0):
    """Tapered relativistic rotor: rho = 10, omega = 9.95 for r <= r0, a
    linear taper f = (r1 - r)/(r1 - r0) of (rho - 1) and omega on
    r0 < r < r1, rho = 1 at rest outside; p = 1, B = (B0, 0, 0) from Az = B0 y."""
    X, Y = grid.X, grid.Y
    r = sqrt(X * X + Y * Y)
    f = torch.where(r <= r0, torch.ones_like(r),
                    torch.where(r < r1, (r1 - r) / (r1 - r0), torch.zeros_like(r)))
    rho = rho_out + (rho_in - rho_out) * f
    om = omega * f
    z = torch.zeros(grid.shape, dtype=torch.float64)
    P = {"rho": rho, "p": torch.full_like(z, press), "vx": -om * Y, "vy": om * X,
         "vz": z.clone(), "Bx": torch.full_like(z, B0), "By": z.clone(), "Bz": z.clone()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    return P, B0 * _corners(grid)[1]


def orszag_tang_2pi(grid, eos, A=1.0, vmax=0.99, rho=1.0, press=1.0):
    """Relativistic Orszag-Tang on [0, 2 pi]^2: rho = p = 1,
    v = vmax/sqrt2 (-sin y, sin x, 0), Az = A (cos y + cos(2x)/2) so that
    B = A (-sin y, sin 2x, 0).  In the reference's rounding mode
    (physics.SQRT_MODE "torch") sin and cos are torch's CPU ones, as the
    reference evaluates them; otherwise the C library's on the 1-D
    coordinates (torch's vectorised CPU sin/cos round differently on
    different CPUs), as the device path's initialiser does."""
    X, Y = grid.X, grid.Y
    amp = vmax / math.sqrt(2.0)
    z = torch.zeros(grid.shape, dtype=torch.float64)
    XF, YF = _corners(grid)
    if PH.SQRT_MODE == "torch":
        # the reference host's own torch CPU sin/cos (the pinning mode)
        sy, sx, s2x = torch.sin(Y), torch.sin(X), torch.sin(2.0 * X)
        cyf, c2xf = torch.cos(YF), torch.cos(2.0 * XF)
    else:
        # the C library's, on the 1-D coordinates: the same on every host
        def libm(fn, v):
            return torch.tensor([fn(a) for a in v.tolist()], dtype=torch.float64)
        nx, ny = grid.nxt, grid.nyt
        sx = libm(math.sin, grid.x)[:, None].expand(nx, ny)
        sy = libm(math.sin, grid.y)[None, :].expand(nx, ny)
        s2x = libm(math.sin, 2.0 * grid.x)[:, None].expand(nx, ny)
        cyf = libm(math.cos, grid.yf)[None, :].expand(nx + 1, ny + 1)
        c2xf = libm(math.cos, 2.0 * grid.xf)[:, None].expand(nx + 1, ny + 1)
    P = {"rho": torch.full_like(z, rho), "p": torch.full_like(z, press),
         "vx": -amp * sy, "vy": amp * sx, "vz": z.clone(),
         "Bx": -A * sy, "By": A * s2x, "Bz": z.clone()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    return P, A * (cyf + 0.5 * c2xf)


def init_state(grid, eos, prims, Az, bc=("outflow", "outflow")):
    """run_2d.py:227-231: faces from Az, cons from prims, sync."""
    Bxf, Byf = b_from_potential(Az, grid.dx, grid.dy)
    # copies: the sync below fills BCs in place (torch autograd keeps the
    # primitive_to_conserved outputs)
    c = {k: v.clone() for k, v in prims_to_cons_2d(prims, grid).items()}
    st = State2D({k: c[k] for k in EVOLVED_KEYS}, Bxf.contiguous(), Byf.contiguous(), prims)
    return sync_state(st, grid, eos, bc[0], bc[1])


def make_grid(n, box, limiter):
    from .recon import ghosts_needed
    return Grid2D(box[0], box[1], n, box[0], box[1], n, ng=ghosts_needed(limiter))
