"""Point-wise SRMHD physics (oracle restatement of src/physics/hlld.py:45-360).

Conserved keys: D, tau, Sx, Sy, Sz, Bx, By, Bz (tau = T^00 - D).
Primitive keys: rho, vx, vy, vz, p, eps, Bx, By, Bz.
"""

import numpy as np
import torch

TINY = 1e-300          # hlld.py:29, c2p.py:29

# Square root used by every oracle function.  torch's CPU float64 sqrt is not
# correctly rounded on every host (it goes through a vectorised math library:
# on the Xeon that wrote tests/golden about 0.8 % of values are one ulp off,
# deterministically per value), which is why the reference's runs differ
# between hosts.  "ieee" (default) is the correctly rounded sqrt, the
# arithmetic the HIP path implements; "torch" is the reference's own
# torch.sqrt, used to pin the oracle bitwise to the golden vectors on a host
# whose torch.sqrt matches the one that wrote them.
SQRT_MODE = "ieee"


def set_sqrt_mode(mode):
    global SQRT_MODE
    if mode not in ("ieee", "torch"):
        raise ValueError(mode)
    SQRT_MODE = mode


def _ieee_sqrt(x):
    with np.errstate(invalid="ignore"):
        return torch.from_numpy(np.asarray(np.sqrt(x.detach().numpy()), dtype=np.float64))


class _IeeeSqrt(torch.autograd.Function):
    """The correctly rounded square root with the derivative 1 / (2 sqrt x),
    so that torch autograd runs through the IEEE-mode oracle."""

    @staticmethod
    def forward(ctx, x):
        y = _ieee_sqrt(x)
        ctx.save_for_backward(y)
        return y

    @staticmethod
    def backward(ctx, g):
        (y,) = ctx.saved_tensors
        return g / (2.0 * y)


def sqrt(x):
    if SQRT_MODE == "torch":
        return torch.sqrt(x)
    if x.requires_grad:
        return _IeeeSqrt.apply(x)
    return _ieee_sqrt(x)


CONS_KEYS = ("D", "tau", "Sx", "Sy", "Sz", "Bx", "By", "Bz")   # hlld.py:298-332 order
VEL = ("vx", "vy", "vz")
BK = ("Bx", "By", "Bz")
SK = ("Sx", "Sy", "Sz")


def sdiv(num, den, tiny=TINY):
    """num / den with |den| < tiny replaced by +-tiny (0 counts as +)
    hlld.py:45-54, c2p.py:32-36."""
    sgn = torch.sign(den)
    sgn = torch.where(sgn == 0.0, torch.ones_like(sgn), sgn)
    return num / torch.where(den.abs() < tiny, sgn * tiny, den)


def lorentz(P):
    """hlld.py:57-62."""
    v2 = P["vx"] ** 2 + P["vy"] ** 2 + P["vz"] ** 2
    return 1.0 / sqrt(1.0 - torch.clamp(v2, max=1.0 - 1e-10)), v2


def smallb(v, B, W):
    """b^mu (Minkowski)   hlld.py:65-81."""
    vdB = v[0] * B[0] + v[1] * B[1] + v[2] * B[2]
    bt = W * vdB
    return (bt,) + tuple(B[i] / W + vdB * v[i] * W for i in range(3))


def b_squared(v, B, W):
    """hlld.py:84-90."""
    B2 = B[0] ** 2 + B[1] ** 2 + B[2] ** 2
    vdB = v[0] * B[0] + v[1] * B[1] + v[2] * B[2]
    return B2 / (W ** 2) + vdB ** 2


def fast_v2_and_h(cs2, b2, rho, eps, p):
    """hlld.py:93-97."""
    h = 1.0 + eps + p / rho
    return cs2 + (b2 / (rho * h + b2)) * (1.0 - cs2), h


def char_speeds(v02, u0, vd):
    """Roots of the flat-space dispersion quadratic, hlld.py:100-113 with its
    defaults one_over_alp2 = 1, betad = 0, gupdd = 1 kept in the arithmetic."""
    ooa2, betad, gupdd = 1.0, 0.0, 1.0
    u0sq = u0 ** 2
    a = u0sq * (1.0 - v02) + v02 * ooa2
    b = 2.0 * (betad * ooa2 * v02 - u0sq * vd * (1.0 - v02))
    c = u0sq * vd ** 2 * (1.0 - v02) - v02 * (gupdd - betad ** 2 * ooa2)
    det = sqrt(torch.clamp(b ** 2 - 4.0 * a * c, min=0.0))
    a = torch.clamp(a, min=TINY)
    r1 = 0.5 * (det - b) / a
    r2 = -0.5 * (det + b) / a
    return torch.max(r1, r2), torch.min(r1, r2)


def prim_to_cons(P):
    """hlld.py:185-217."""
    vx, vy, vz = P["vx"], P["vy"], P["vz"]
    Bx, By, Bz = P["Bx"], P["By"], P["Bz"]
    rho, p = P["rho"], P["p"]
    v2 = vx ** 2 + vy ** 2 + vz ** 2
    W = 1.0 / sqrt(torch.clamp(1.0 - v2, min=1e-10))
    vB = vx * Bx + vy * By + vz * Bz
    B2 = Bx ** 2 + By ** 2 + Bz ** 2
    h = 1.0 + P["eps"] + p / rho
    rhW2 = rho * h * W ** 2
    D = rho * W
    return {
        "D": D,
        "Sx": (rhW2 + B2) * vx - vB * Bx,
        "Sy": (rhW2 + B2) * vy - vB * By,
        "Sz": (rhW2 + B2) * vz - vB * Bz,
        "tau": rhW2 - p - D + 0.5 * (B2 + B2 * v2 - vB ** 2),
        "Bx": Bx, "By": By, "Bz": Bz,
    }


def side_state(P, eos, idir):
    """One side of ``compute_srmhd_fluxes`` (hlld.py:223-334): returns
    (U, F, c_plus, c_minus) with U, F keyed by CONS_KEYS."""
    vn = P[VEL[idir]]
    W, _ = lorentz(P)
    p, cs2 = eos.press_and_cs2(P["eps"], P["rho"])
    v = (P["vx"], P["vy"], P["vz"])
    B = (P["Bx"], P["By"], P["Bz"])
    b2 = b_squared(v, B, W)
    sb = smallb(v, B, W)
    v02, h = fast_v2_and_h(cs2, b2, P["rho"], P["eps"], p)
    cp, cm = char_speeds(v02, W, vn)
    d = P["rho"] * W
    rhb2 = P["rho"] * h + b2
    Pb2 = p + 0.5 * b2
    bn = sb[1 + idir]
    U, F = {}, {}
    U["D"] = d
    F["D"] = d * vn
    Tt = rhb2 * W ** 2 + Pb2 * (-1.0) - sb[0] ** 2
    U["tau"] = Tt - d
    F["tau"] = (rhb2 * W ** 2 * vn - sb[0] * bn) - d * vn
    for ii in range(3):
        ui = W * v[ii]
        U[SK[ii]] = rhb2 * W * ui - sb[0] * sb[1 + ii]
        F[SK[ii]] = rhb2 * (W * vn) * ui + Pb2 * (1.0 if ii == idir else 0.0) - bn * sb[1 + ii]
    Bn = B[idir]
    for ii in range(3):
        U[BK[ii]] = B[ii]
        F[BK[ii]] = vn * B[ii] - v[ii] * Bn
    return U, F, cp, cm


def srmhd_fluxes(PL, PR, eos, idir=0):
    """hlld.py:223-334: (uL, uR, fL, fR, cmax, cmin)."""
    uL, fL, cpl, cml = side_state(PL, eos, idir)
    uR, fR, cpr, cmr = side_state(PR, eos, idir)
    cmax = torch.clamp(torch.max(cpl, cpr), min=0.0)
    cmin = torch.clamp(torch.max(-cml, -cmr), min=0.0)
    small = (cmax < 1e-12) & (cmin < 1e-12)
    cmax = torch.where(small, torch.ones_like(cmax), cmax)
    cmin = torch.where(small, torch.ones_like(cmin), cmin)
    return uL, uR, fL, fR, cmax, cmin


def hll_flux(fl, fr, ul, ur, cm, cx):
    """hlld.py:161-163."""
    return (cx * fl + cm * fr - cx * cm * (ur - ul)) / (cx + cm)


def hll_state(fl, fr, ul, ur, cm, cx):
    """hlld.py:166-168."""
    return (cx * ur + cm * ul + fl - fr) / (cx + cm)


def energy_form(u, f):
    """tau -> E = tau + D on a (state, flux) pair   hlld.py:337-360."""
    return {**u, "tau": u["tau"] + u["D"]}, {**f, "tau": f["tau"] + f["D"]}
