"""Batched 1D Riemann tubes.

``run_tubes_shared`` restates src/physics/tube_seed.py:44-144 (equal-length
tubes as columns of one array, ONE dt for all tubes, SSP-RK2 in Heun form,
con2prim batched over every cell of every tube).

``run_tube`` is the per-tube form the batched GPU runner reproduces: one
tube is one batch, with its own dt from the 1D fast-speed bound
(driver.py:77-127), SSP-RK2/3 weights of driver2d.py:104-105 applied as in
state.combine, and the tube_seed building blocks (reconstruct_prims along
the tube, constant B_n, -(F[i+1]-F[i])/dx on D, Sx, Sy, Sz, tau, By, Bz, a
con2prim of every cell, two-cell-copy ghosts).  A loop of ``run_tube`` over
tubes is the per-tube baseline.
"""

import math

import numpy as np
import torch

from .c2p import cons_to_prim
from .physics import prim_to_cons, sqrt
from .recon import ghosts_needed, reconstruct_prims
from .riemann import SOLVERS
from .step1d import compute_dt
from .step2d import RK

KEYS = ("D", "Sx", "Sy", "Sz", "tau")


def _bc(d, ng):
    """tube_seed.py:72-75."""
    for v in d.values():
        v[:ng] = v[ng:ng + 1]
        v[-ng:] = v[-ng - 1:-ng]


def _rhs(prims, eos, dx, limiter, flux):
    """tube_seed.py:81-90."""
    L, R = reconstruct_prims(prims, axis=0, eos=eos, limiter=limiter)
    L["Bx"] = R["Bx"] = prims["Bx"][:1].expand(L["rho"].shape).clone()
    F, _, _ = SOLVERS[flux](L, R, eos, idir=0)
    return {k: -(F[k][1:] - F[k][:-1]) / dx for k in KEYS + ("By", "Bz")}


def _sync(c, Bx, eos, ng):
    full = {k: c[k] for k in KEYS}
    full["Bx"], full["By"], full["Bz"] = Bx, c["By"], c["Bz"]
    p = cons_to_prim(full, eos)
    p["Bx"], p["By"], p["Bz"] = Bx, c["By"], c["Bz"]
    _bc(p, ng)
    return p


def prims_from_total(UL, UR, Bn, eos, ncells, ng):
    """tube_seed.py:44-69: solver-frame states (rho, P_tot, vx, vy, vz, By, Bz)
    split at the middle of an (ncells + 2 ng, N) grid."""
    n = Bn.shape[0]
    nt = ncells + 2 * ng
    half = nt // 2
    t = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float64)  # noqa: E731
    P = {}
    for k, j in (("rho", 0), ("vx", 2), ("vy", 3), ("vz", 4), ("By", 5), ("Bz", 6)):
        col = torch.empty((nt, n), dtype=torch.float64)
        col[:half] = t(UL[:, j])
        col[half:] = t(UR[:, j])
        P[k] = col
    P["Bx"] = torch.empty((nt, n), dtype=torch.float64)
    P["Bx"][:] = t(Bn)
    for U, sl in ((UL, slice(None, half)), (UR, slice(half, None))):
        v2 = U[:, 2] ** 2 + U[:, 3] ** 2 + U[:, 4] ** 2
        W2 = 1.0 / np.maximum(1.0 - v2, 1e-300)
        eta = Bn * U[:, 2] + U[:, 5] * U[:, 3] + U[:, 6] * U[:, 4]
        b2 = (Bn ** 2 + U[:, 5] ** 2 + U[:, 6] ** 2) / W2 + eta ** 2
        P.setdefault("p", torch.empty((nt, n), dtype=torch.float64))
        P["p"][sl] = t(U[:, 1] - 0.5 * b2)
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    return P


def run_tubes_shared(UL, UR, Bn, eos, ncells=192, tend=0.18, cfl=0.3, limiter="mc",
                     ng=None, max_steps=400, flux="hlld"):
    """tube_seed.py:93-144."""
    if ng is None:
        ng = ghosts_needed(limiter)
    dx = 1.0 / ncells
    P = prims_from_total(np.asarray(UL, float), np.asarray(UR, float),
                         np.asarray(Bn, float), eos, ncells, ng)
    _bc(P, ng)
    c = prim_to_cons(P)
    cons = {k: c[k].clone() for k in KEYS}
    cons["By"], cons["Bz"] = P["By"].clone(), P["Bz"].clone()
    Bx = P["Bx"]
    t, step = 0.0, 0
    while t < tend - 1e-14 and step < max_steps:
        cs = sqrt(torch.clamp(
            eos.gamma * P["p"] / torch.clamp(
                P["rho"] + eos.gamma * P["p"] / (eos.gamma - 1.0), min=1e-30), min=0.0))
        smax = float((P["vx"].abs() + torch.clamp(cs + 0.5, max=1.0)).max())
        dt = min(cfl * dx / max(smax, 1e-8), tend - t)
        k1 = _rhs(P, eos, dx, limiter, flux)
        c1 = {k: cons[k] + dt * k1[k] for k in cons}
        p1 = _sync(c1, Bx, eos, ng)
        k2 = _rhs(p1, eos, dx, limiter, flux)
        cons = {k: 0.5 * (cons[k] + c1[k] + dt * k2[k]) for k in cons}
        P = _sync(cons, Bx, eos, ng)
        t += dt
        step += 1
    ph = slice(ng, ng + ncells)
    return {k: v[ph] for k, v in P.items()}, t


# ── per-tube form (the batched GPU runner's contract) ───────────────────────

def tube_initial(ncells, ng, eos, left, right, Bn):
    """Cell-centred primitives of one tube on x in [0, 1], interface at 0.5;
    ``left``/``right`` give rho, p, vx, vy, vz, By, Bz."""
    dx = 1.0 / ncells
    x = (torch.arange(ncells + 2 * ng, dtype=torch.float64) - ng + 0.5) * dx
    m = x < 0.5
    P = {k: torch.where(m, torch.full_like(x, float(left[k])), torch.full_like(x, float(right[k])))
         for k in ("rho", "p", "vx", "vy", "vz", "By", "Bz")}
    P["Bx"] = torch.full_like(x, float(Bn))
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    return P


def run_tube(P0, eos, ncells, ng, steps, cfl=0.4, limiter="mp5", scheme="rk3", flux="hlld"):
    """Advance one tube ``steps`` steps; returns (prims incl. ghosts, t)."""
    dx = 1.0 / ncells
    P = {k: v.clone() for k, v in P0.items()}
    c = prim_to_cons(P)
    cons = {k: c[k].clone() for k in KEYS}
    cons["By"], cons["Bz"] = P["By"].clone(), P["Bz"].clone()
    Bx = P["Bx"]
    t = 0.0
    for _ in range(steps):
        dt = float(compute_dt(P, eos, dx, cfl, ng, ncells, speed_of_light=False))
        U0 = cons
        U = cons
        for a, b, cc in RK[scheme]:
            L = _rhs(P, eos, dx, limiter, flux)
            w = cc * dt
            U = {k: a * U0[k] + b * U[k] + w * L[k] for k in U0}
            P = _sync(U, Bx, eos, ng)
        cons = U
        t += dt
    return P, t


def draw_tubes(n_tubes, seed=0, lmin=128, lmax=2048):
    """The batched-tube workload: lengths U{lmin..lmax}; per tube rho_L,
    rho_R log-uniform [0.1, 10], p_L, p_R log-uniform [0.01, 10], velocity
    direction uniform on the sphere with |v| ~ U[0, 0.9] per side, By, Bz ~
    U[-2, 2] per side, one Bx ~ U[-1, 1] shared.  Drawn with a CPU
    torch.Generator seeded with ``seed``."""
    g = torch.Generator().manual_seed(seed)

    def u(lo, hi, shape=(n_tubes,)):
        return lo + (hi - lo) * torch.rand(shape, generator=g, dtype=torch.float64)

    lengths = torch.randint(lmin, lmax + 1, (n_tubes,), generator=g)
    sides = []
    for _ in range(2):
        rho = 10.0 ** u(-1.0, 1.0)
        p = 10.0 ** u(-2.0, 1.0)
        ct = u(-1.0, 1.0)
        ph = u(0.0, 2.0 * math.pi)
        sp = u(0.0, 0.9)
        st = sqrt(1.0 - ct * ct)
        sides.append({"rho": rho, "p": p, "vx": sp * st * torch.cos(ph),
                      "vy": sp * st * torch.sin(ph), "vz": sp * ct,
                      "By": u(-2.0, 2.0), "Bz": u(-2.0, 2.0)})
    Bn = u(-1.0, 1.0)
    return lengths, sides[0], sides[1], Bn
