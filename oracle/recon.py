"""Reconstruction (oracle restatement of src/physics/reconstruction.py).

Face convention (reconstruction.py:35-47): face i sits between cell i-1 and
cell i; an axis of n cells has n+1 faces.  QL[i] is the right edge of cell
i-1, QR[i] the left edge of cell i; faces 0 and n take the adjacent cell's
own edge value.
"""

import torch

from .physics import sqrt

STENCIL = {"pcm": 0, "mc": 1, "minmod": 1, "weno5z": 2, "mp5": 2, "mp7": 3}
HIGH_ORDER = ("weno5z", "mp5", "mp7")
RECON_KEYS = ("rho", "p", "vx", "vy", "vz", "Bx", "By", "Bz")   # reconstruction.py:269


def ghosts_needed(limiter):
    """reconstruction.py:89-98."""
    if limiter not in STENCIL:
        raise ValueError(f"unknown limiter {limiter!r}")
    return max(2, STENCIL[limiter] + 1)


def minmod(a, b):
    """reconstruction.py:60-66."""
    return torch.where(a * b > 0.0, torch.where(a.abs() < b.abs(), a, b), torch.zeros_like(a))


def mc(a, b):
    """reconstruction.py:69-77."""
    c = 0.5 * (a + b)
    return torch.where(a * b > 0.0,
                       torch.sign(c) * torch.minimum(2.0 * a.abs(),
                                                     torch.minimum(2.0 * b.abs(), c.abs())),
                       torch.zeros_like(a))


LIMITERS = {"mc": mc, "minmod": minmod}


def _sh(Q, k, axis):
    """entry i holds Q[i+k] (wrapping; wrapped entries are overwritten)."""
    return torch.roll(Q, shifts=-k, dims=axis)


def weno5z_right(v, axis):
    """reconstruction.py:111-128."""
    vm2, vm1, v0, vp1, vp2 = (_sh(v, k, axis) for k in (-2, -1, 0, 1, 2))
    p0 = (2.0 * vm2 - 7.0 * vm1 + 11.0 * v0) / 6.0
    p1 = (-vm1 + 5.0 * v0 + 2.0 * vp1) / 6.0
    p2 = (2.0 * v0 + 5.0 * vp1 - vp2) / 6.0
    c1, c2 = 13.0 / 12.0, 0.25
    b0 = c1 * (vm2 - 2.0 * vm1 + v0) ** 2 + c2 * (vm2 - 4.0 * vm1 + 3.0 * v0) ** 2
    b1 = c1 * (vm1 - 2.0 * v0 + vp1) ** 2 + c2 * (vm1 - vp1) ** 2
    b2 = c1 * (v0 - 2.0 * vp1 + vp2) ** 2 + c2 * (3.0 * v0 - 4.0 * vp1 + vp2) ** 2
    t5 = (b0 - b2).abs()
    eps = 1e-40
    a0 = 0.1 * (1.0 + (t5 / (b0 + eps)) ** 2)
    a1 = 0.6 * (1.0 + (t5 / (b1 + eps)) ** 2)
    a2 = 0.3 * (1.0 + (t5 / (b2 + eps)) ** 2)
    return (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2)


def _median(a, b, c):
    return a + minmod(b - a, c - a)


def mp_limit(vl, v, axis, alpha=4.0):
    """Suresh-Huynh clip   reconstruction.py:136-162."""
    vm2, vm1, v0, vp1, vp2 = (_sh(v, k, axis) for k in (-2, -1, 0, 1, 2))
    vmp = v0 + minmod(vp1 - v0, alpha * (v0 - vm1))
    need = (vl - v0) * (vl - vmp) > 1e-40
    dm1 = vm2 - 2.0 * vm1 + v0
    d0 = vm1 - 2.0 * v0 + vp1
    dp1 = v0 - 2.0 * vp1 + vp2
    dm4p = minmod(minmod(4.0 * d0 - dp1, 4.0 * dp1 - d0), minmod(d0, dp1))
    dm4m = minmod(minmod(4.0 * dm1 - d0, 4.0 * d0 - dm1), minmod(dm1, d0))
    vul = v0 + alpha * (v0 - vm1)
    vav = 0.5 * (v0 + vp1)
    vmd = vav - 0.5 * dm4p
    vlc = v0 + 0.5 * (v0 - vm1) + (4.0 / 3.0) * dm4m
    vmin = torch.maximum(torch.minimum(torch.minimum(v0, vp1), vmd),
                         torch.minimum(torch.minimum(v0, vul), vlc))
    vmax = torch.minimum(torch.maximum(torch.maximum(v0, vp1), vmd),
                         torch.maximum(torch.maximum(v0, vul), vlc))
    return torch.where(need, _median(vl, vmin, vmax), vl)


def mp5_right(v, axis):
    """reconstruction.py:165-168."""
    vm2, vm1, v0, vp1, vp2 = (_sh(v, k, axis) for k in (-2, -1, 0, 1, 2))
    vl = (2.0 * vm2 - 13.0 * vm1 + 47.0 * v0 + 27.0 * vp1 - 3.0 * vp2) / 60.0
    return mp_limit(vl, v, axis)


def mp7_right(v, axis):
    """reconstruction.py:171-176."""
    vm3, vm2, vm1, v0, vp1, vp2, vp3 = (_sh(v, k, axis) for k in (-3, -2, -1, 0, 1, 2, 3))
    vl = (-3.0 * vm3 + 25.0 * vm2 - 101.0 * vm1 + 319.0 * v0
          + 214.0 * vp1 - 38.0 * vp2 + 4.0 * vp3) / 420.0
    return mp_limit(vl, v, axis)


FACE_FN = {"weno5z": weno5z_right, "mp5": mp5_right, "mp7": mp7_right}


def _along(axis, nd, s):
    idx = [slice(None)] * nd
    idx[axis] = s
    return tuple(idx)


def cell_slopes(Q, axis=0, limiter="mc"):
    """reconstruction.py:225-236: zero on the two outermost cells."""
    lim = LIMITERS.get(limiter)
    if lim is None:
        raise ValueError(f"unknown limiter {limiter!r}")
    nd = Q.ndim
    fwd = Q[_along(axis, nd, slice(1, None))] - Q[_along(axis, nd, slice(None, -1))]
    s = torch.zeros_like(Q)
    s[_along(axis, nd, slice(1, -1))] = lim(fwd[_along(axis, nd, slice(None, -1))],
                                            fwd[_along(axis, nd, slice(1, None))])
    return s


def edges(Q, axis, limiter):
    """(right edge, left edge) of every cell   reconstruction.py:182-213,
    239-258.  The left edge of a high-order scheme is the right edge of the
    mirrored array."""
    if limiter in FACE_FN:
        fn = FACE_FN[limiter]
        right = fn(Q, axis)
        left = torch.flip(fn(torch.flip(Q, dims=(axis,)), axis), dims=(axis,))
        w = STENCIL[limiter]
        s = cell_slopes(Q, axis, "mc")
        pr, pl = Q + 0.5 * s, Q - 0.5 * s
        n = Q.shape[axis]
        if n <= 2 * w:
            return pr, pl
        e = torch.zeros(n, dtype=torch.bool)
        e[:w] = True
        e[n - w:] = True
        shape = [1] * Q.ndim
        shape[axis] = n
        e = e.reshape(shape)
        return torch.where(e, pr, right), torch.where(e, pl, left)
    s = torch.zeros_like(Q) if limiter == "pcm" else cell_slopes(Q, axis, limiter)
    return Q + 0.5 * s, Q - 0.5 * s


def reconstruct(Q, axis=0, limiter="pcm"):
    """(QL, QR) on the faces   reconstruction.py:239-264."""
    nd = Q.ndim
    r, l = edges(Q, axis, limiter)
    QL = torch.cat([r[_along(axis, nd, slice(0, 1))], r], dim=axis)
    QR = torch.cat([l, l[_along(axis, nd, slice(-1, None))]], dim=axis)
    return QL, QR


def reconstruct_prims(prims, axis, eos, limiter="mc", floor_rho=1e-12, floor_p=1e-14):
    """reconstruction.py:272-338 with vel_var = "Wv": reconstruct W v, PCM
    fallback on faces whose rho or p would cross the floor, floors, and an
    EOS-consistent eps."""
    work = {k: prims[k] for k in RECON_KEYS if k in prims}
    v2 = prims["vx"] ** 2 + prims["vy"] ** 2 + prims["vz"] ** 2
    W = 1.0 / sqrt(torch.clamp(1.0 - v2, min=1e-14))
    for c in ("vx", "vy", "vz"):
        work[c] = W * prims[c]
    L, R = {}, {}
    for k, v in work.items():
        L[k], R[k] = reconstruct(v, axis=axis, limiter=limiter)
    if limiter != "pcm":
        bad = (L["rho"] <= floor_rho) | (R["rho"] <= floor_rho)
        bad = bad | (L["p"] <= floor_p) | (R["p"] <= floor_p)
        if bool(bad.any()):
            for k in work:
                cl, cr = reconstruct(work[k], axis=axis, limiter="pcm")
                L[k] = torch.where(bad, cl, L[k])
                R[k] = torch.where(bad, cr, R[k])
    for side in (L, R):
        z2 = side["vx"] ** 2 + side["vy"] ** 2 + side["vz"] ** 2
        fac = 1.0 / sqrt(1.0 + z2)
        for c in ("vx", "vy", "vz"):
            side[c] = side[c] * fac
        side["rho"] = torch.clamp(side["rho"], min=floor_rho)
        side["p"] = torch.clamp(side["p"], min=floor_p)
        side["eps"] = eos.eps__press_rho(side["p"], side["rho"])
    return L, R


def interface_states_1d(prims, ng, ncells, limiter="pcm"):
    """Legacy 1D reconstruction of ALL primitive keys (no W v, no floors, eps
    reconstructed directly), physical faces only   reconstruction.py:354-373."""
    keys = list(prims.keys())
    Q = torch.stack([prims[k] for k in keys], dim=1)
    QL, QR = reconstruct(Q, axis=0, limiter=limiter)
    sl = slice(ng, ng + ncells + 1)
    return ({k: QL[sl, j] for j, k in enumerate(keys)},
            {k: QR[sl, j] for j, k in enumerate(keys)})
