"""HLLE / HLLC / HLLD fluxes (oracle restatement of src/physics/hlld.py).

All three keep the reference's plugin contract
``flux_fn(sL, sR, eos, idir) -> (F, U, p_star)`` with dicts keyed
D, tau, Sx, Sy, Sz, Bx, By, Bz, and update the module-global ``LAST_DIAG``
like hlld.py:1516-1529.
"""

import torch

from .c2p import cons_to_prim
from .physics import (BK, CONS_KEYS, SK, VEL, TINY, b_squared, energy_form,
                      hll_flux, hll_state, lorentz, sdiv, sqrt, srmhd_fluxes)

LAST_DIAG: dict = {}
PSTAR_TOL = 1.0e-10       # hlld.py:1080
PSTAR_MAXIT = 30          # hlld.py:1327
# The reference's call-time switches of hlld_flux (module globals, read on
# every call): the p* root finder (hlld.py:1088, RMHD_PSTAR_METHOD), the
# bracket source (hlld.py:1065) and the relative finite-difference step of
# the newton_fd / newton_fd1 derivatives (hlld.py:1091-1092, None = the
# method's own default).
PSTAR_METHOD = "secant"
PSTAR_BRACKET = "legacy"
PSTAR_FD_H = None
PSTAR_METHODS = ("secant", "newton", "newton_fd", "newton_fd1", "newton_safe")


def _p_hll_sides(sL, sR, eos):
    """The HLL total-pressure estimate returned by hlle/hllc
    (hlld.py:1036-1042, 1757-1763)."""
    out = []
    for s in (sL, sR):
        p, _ = eos.press_and_cs2(s["eps"], s["rho"])
        W, _ = lorentz(s)
        out.append((p, b_squared((s["vx"], s["vy"], s["vz"]), (s["Bx"], s["By"], s["Bz"]), W)))
    (pl, b2l), (pr, b2r) = out
    return torch.clamp(0.5 * (pl + 0.5 * b2l + pr + 0.5 * b2r), min=1e-30)


def hlle_flux(sL, sR, eos, idir=0):
    """hlld.py:1005-1044."""
    uL, uR, fL, fR, cmax, cmin = srmhd_fluxes(sL, sR, eos, idir)
    F, U = {}, {}
    for k in CONS_KEYS:
        fh = hll_flux(fL[k], fR[k], uL[k], uR[k], cmin, cmax)
        uh = hll_state(fL[k], fR[k], uL[k], uR[k], cmin, cmax)
        F[k] = torch.where(cmax <= 0, fR[k], torch.where(cmin <= 0, fL[k], fh))
        U[k] = torch.where(cmin <= 0, uL[k], torch.where(cmax <= 0, uR[k], uh))
    return F, U, _p_hll_sides(sL, sR, eos)


def hllc_flux(sL, sR, eos, idir=0):
    """Mignone-style HLLC   hlld.py:1549-1765."""
    uL, uR, fL, fR, cmax, cmin = srmhd_fluxes(sL, sR, eos, idir)
    n, t1, t2 = idir, (idir + 1) % 3, (idir + 2) % 3
    SL, SR = -cmin, cmax
    fH = {k: hll_flux(fL[k], fR[k], uL[k], uR[k], cmin, cmax) for k in CONS_KEYS}
    uH = {k: hll_state(fL[k], fR[k], uL[k], uR[k], cmin, cmax) for k in CONS_KEYS}
    Bn = uL[BK[n]]
    hasB = Bn.abs() > 1e-15
    z = torch.zeros_like(Bn)
    bt1, bt2 = uH[BK[t1]], uH[BK[t2]]
    fb1, fb2 = fH[BK[t1]], fH[BK[t2]]
    cross = torch.where(hasB, bt1 * fb1 + bt2 * fb2, z)
    btan2 = torch.where(hasB, bt1 ** 2 + bt2 ** 2, z)
    fbtan2 = torch.where(hasB, fb1 ** 2 + fb2 ** 2, z)
    E = uH["tau"] + uH["D"]
    fE = fH["tau"] + fH["D"]
    qa = fE - cross
    qb = -fH[SK[n]] - E + btan2 + fbtan2
    qc = uH[SK[n]] - cross
    disc = sqrt(torch.clamp(qb ** 2 - 4.0 * qa * qc, min=0.0))
    vq = sdiv(-0.5 * (qb + disc), qa)
    vlin = sdiv(-qc, qb)
    vc = torch.where(qa.abs() < 1e-15,
                     torch.where(qb.abs() < 1e-15, torch.zeros_like(qa), vlin), vq)
    vt1 = torch.where(hasB, sdiv(uH[BK[t1]] * vc - fH[BK[t1]], Bn), torch.zeros_like(vc))
    vt2 = torch.where(hasB, sdiv(uH[BK[t2]] * vc - fH[BK[t2]], Bn), torch.zeros_like(vc))
    v2 = vc ** 2 + vt1 ** 2 + vt2 ** 2
    g = 1.0 / sqrt(1.0 - torch.clamp(v2, max=1.0 - 1e-10))
    vB = torch.where(hasB, vc * uH[BK[n]] + vt1 * uH[BK[t1]] + vt2 * uH[BK[t2]],
                     torch.zeros_like(vc))
    pB = fH[SK[n]] + (Bn / g) ** 2 - vc * (fH["tau"] + fH["D"] - Bn * vB)
    p0 = fH[SK[n]] - vc * (fH["tau"] + fH["D"])
    ps = torch.where(hasB, pB, p0)

    def star(u, f, S, vd):
        den = S - vc
        s = {BK[n]: Bn, BK[t1]: uH[BK[t1]], BK[t2]: uH[BK[t2]]}
        s["D"] = sdiv(u["D"] * (S - vd), den)
        s["tau"] = sdiv(S * u["tau"] - f["tau"] + ps * vc - vB * Bn, den)
        s[SK[n]] = (s["tau"] + s["D"] + ps) * vc - vB * Bn
        for t, vt in ((t1, vt1), (t2, vt2)):
            withB = sdiv(-Bn * (uH[BK[t]] / g ** 2 + vB * vt) - f[SK[t]] + S * u[SK[t]], den)
            noB = sdiv(u[SK[t]] * (S - vd), den)
            s[SK[t]] = torch.where(hasB, withB, noB)
        return s

    ucL = star(uL, fL, SL, sL[VEL[n]])
    ucR = star(uR, fR, SR, sR[VEL[n]])
    vi = torch.zeros_like(cmin)
    F, U = {}, {}
    for k in CONS_KEYS:
        fsL = fL[k] + SL * (ucL[k] - uL[k])
        fsR = fR[k] + SR * (ucR[k] - uR[k])
        f, u = fH[k].clone(), uH[k].clone()
        m = (vi <= SL, (SL < vi) & (vi < vc), (vc <= vi) & (vi < SR), vi >= SR)
        for mask, fv, uv in zip(m, (fL[k], fsL, fsR, fR[k]), (uL[k], ucL[k], ucR[k], uR[k])):
            f = torch.where(mask, fv, f)
            u = torch.where(mask, uv, u)
        F[k], U[k] = f, u
    degen = (vc <= SL) | (vc >= SR) | (v2 >= 1.0)
    for k in CONS_KEYS:
        F[k] = torch.where(degen, fH[k], F[k])
        U[k] = torch.where(degen, uH[k], U[k])
    for k in CONS_KEYS:
        bad = ~torch.isfinite(F[k]) | ~torch.isfinite(U[k])
        F[k] = torch.where(bad, fH[k], F[k])
        U[k] = torch.where(bad, uH[k], U[k])
    return F, U, _p_hll_sides(sL, sR, eos)


class StarStates:
    """Alfven/contact state algebra of HLLDComputation (hlld.py:366-605),
    built on the total-energy (E = tau + D) form of both sides."""

    def __init__(self, fL, fR, uL, uR, sL, sR, cmin, cmax, idir):
        self.n, self.t1, self.t2 = idir, (idir + 1) % 3, (idir + 2) % 3
        self.uL, self.uR, self.sL, self.sR = uL, uR, sL, sR
        self.cm, self.cx = cmin, cmax
        cm, cx = cmin, cmax
        n, t1, t2 = self.n, self.t1, self.t2

        def rl(k):
            return -cm * uL[k] - fL[k]

        def rr(k):
            return cx * uR[k] - fR[k]
        # _build_R   hlld.py:401-428
        self.RSnL, self.RSnR = rl(SK[n]), rr(SK[n])
        self.RS1L, self.RS1R = rl(SK[t1]), rr(SK[t1])
        self.RS2L, self.RS2R = rl(SK[t2]), rr(SK[t2])
        self.RBnL, self.RBnR = rl(BK[n]), rr(BK[n])
        self.RB1L, self.RB1R = rl(BK[t1]), rr(BK[t1])
        self.RB2L, self.RB2R = rl(BK[t2]), rr(BK[t2])
        self.RTL, self.RTR = rl("tau"), rr("tau")
        self.RDL, self.RDR = rl("D"), rr("D")
        self.SgL = torch.sign(uL[BK[n]])
        self.SgR = torch.sign(uR[BK[n]])

    def solve(self, p):
        """compute_all_variables   hlld.py:431-583."""
        cm, cx = self.cm, self.cx
        BnL, BnR = self.uL[BK[self.n]], self.uR[BK[self.n]]
        RSnL, RSnR, RS1L, RS1R = self.RSnL, self.RSnR, self.RS1L, self.RS1R
        RS2L, RS2R, RB1L, RB1R = self.RS2L, self.RS2R, self.RB1L, self.RB1R
        RB2L, RB2R, RTL, RTR = self.RB2L, self.RB2R, self.RTL, self.RTR

        AL = RSnL - (-cm) * RTL + p * (1 - cm * cm)
        AR = RSnR - cx * RTR + p * (1 - cx * cx)
        GL = RB1L ** 2 + RB2L ** 2
        GR = RB1R ** 2 + RB2R ** 2
        CL = RS1L * RB1L + RS2L * RB2L
        CR = RS1R * RB1R + RS2R * RB2R
        QL = -AL - GL + BnL ** 2 * (1 - cm * cm)
        QR = -AR - GR + BnR ** 2 * (1 - cx * cx)
        XL = BnL * (AL * (-cm) * BnL + CL) - (AL + GL) * ((-cm) * p + RTL)
        XR = BnR * (AR * cx * BnR + CR) - (AR + GR) * (cx * p + RTR)

        vaL = sdiv(BnL * (AL * BnL + (-cm) * CL) - (AL + GL) * (p + RSnL), XL)
        vaR = sdiv(BnR * (AR * BnR + cx * CR) - (AR + GR) * (p + RSnR), XR)
        v1L = sdiv(QL * RS1L + RB1L * (CL + BnL * ((-cm) * RSnL - RTL)), XL)
        v1R = sdiv(QR * RS1R + RB1R * (CR + BnR * (cx * RSnR - RTR)), XR)
        v2L = sdiv(QL * RS2L + RB2L * (CL + BnL * ((-cm) * RSnL - RTL)), XL)
        v2R = sdiv(QR * RS2R + RB2R * (CR + BnR * (cx * RSnR - RTR)), XR)

        # the Alfven denominators -cm - va, cx - va are formed afresh at every
        # use, as the reference writes them: the values are the same either
        # way, but torch autograd (the exact-derivative Newton p* root)
        # accumulates their gradients use by use
        B1L = sdiv(RB1L - BnL * v1L, -cm - vaL)
        B1R = sdiv(RB1R - BnR * v1R, cx - vaR)
        B2L = sdiv(RB2L - BnL * v2L, -cm - vaL)
        B2R = sdiv(RB2R - BnR * v2R, cx - vaR)

        vRSL = vaL * RSnL + v1L * RS1L + v2L * RS2L
        vRSR = vaR * RSnR + v1R * RS1R + v2R * RS2R
        waL = p + sdiv(RTL - vRSL, -cm - vaL)
        waR = p + sdiv(RTR - vRSR, cx - vaR)
        DaL = sdiv(self.RDL, -cm - vaL)
        DaR = sdiv(self.RDR, cx - vaR)
        vBL = vaL * BnL + v1L * B1L + v2L * B2L
        vBR = vaR * BnR + v1R * B1R + v2R * B2R
        TaL = sdiv(RTL + p * vaL - vBL * BnL, -cm - vaL) - DaL
        TaR = sdiv(RTR + p * vaR - vBR * BnR, cx - vaR) - DaR
        SaL = (TaL + DaL + p) * vaL - vBL * BnL
        SaR = (TaR + DaR + p) * vaR - vBR * BnR
        S1L = (TaL + DaL + p) * v1L - vBL * B1L
        S1R = (TaR + DaR + p) * v1R - vBR * B1R
        S2L = (TaL + DaL + p) * v2L - vBL * B2L
        S2R = (TaR + DaR + p) * v2R - vBR * B2R

        sqL = sqrt(torch.clamp(waL, min=0.0))
        sqR = sqrt(torch.clamp(waR, min=0.0))
        SgL, SgR = self.SgL, self.SgR
        dL = (-cm) * p + RTL - BnL * SgL * sqL
        dR = cx * p + RTR + BnR * SgR * sqR
        KnL = sdiv(RSnL + p - self.RBnL * SgL * sqL, dL)
        KnR = sdiv(RSnR + p + self.RBnR * SgR * sqR, dR)
        K1L = sdiv(RS1L - RB1L * SgL * sqL, dL)
        K1R = sdiv(RS1R + RB1R * SgR * sqR, dR)
        K2L = sdiv(RS2L - RB2L * SgL * sqL, dL)
        K2R = sdiv(RS2R + RB2R * SgR * sqR, dR)

        dll = KnR - KnL
        Bc1 = sdiv((B1R * (KnR - vaR) + BnR * v1R) - (B1L * (KnL - vaL) + BnL * v1L), dll)
        Bc2 = sdiv((B2R * (KnR - vaR) + BnR * v2R) - (B2L * (KnL - vaL) + BnL * v2L), dll)
        Bcn = BnR
        K2Lsq = KnL ** 2 + K1L ** 2 + K2L ** 2
        K2Rsq = KnR ** 2 + K1R ** 2 + K2R ** 2
        KBL = KnL * Bcn + K1L * Bc1 + K2L * Bc2
        KBR = KnR * Bcn + K1R * Bc1 + K2R * Bc2
        self.__dict__.update(
            vaL=vaL, vaR=vaR, v1L=v1L, v1R=v1R, v2L=v2L, v2R=v2R,
            B1L=B1L, B1R=B1R, B2L=B2L, B2R=B2R, sqL=sqL, sqR=sqR,
            DaL=DaL, DaR=DaR, TaL=TaL, TaR=TaR, SaL=SaL, SaR=SaR,
            S1L=S1L, S1R=S1R, S2L=S2L, S2R=S2R, laL=KnL, laR=KnR,
            K1L=K1L, K2L=K2L, Bcn=Bcn, Bc1=Bc1, Bc2=Bc2,
            KL2=K2Lsq, KR2=K2Rsq, KBL=KBL, KBR=KBR)

    def residual(self, p):
        """hlld.py:586-602."""
        BnL = self.uL[BK[self.n]]
        self.solve(p)
        dR = self.SgR * self.sqR - self.KBR
        dL = self.SgL * self.sqL + self.KBL
        return (self.laL - self.laR) + BnL * (sdiv(1.0 - self.KR2, dR) + sdiv(1.0 - self.KL2, dL))


def clamped_secant(func, x0, x1, tol, max_iter):
    """``safe_secant_bisection`` (hlld.py:669-746): a clamped secant with
    converged lanes frozen.  Returns (root, err, iters_used)."""
    x0, x1 = x0.clone(), x1.clone()
    f0, f1 = func(x0), func(x1)
    err = torch.full(x0.shape, max_iter, dtype=torch.int32)
    used = torch.full(x0.shape, max_iter, dtype=torch.int32)
    done = torch.zeros(x0.shape, dtype=torch.bool)
    it = 0
    for _ in range(max_iter):
        it += 1
        conv = f1.abs() < tol
        err = torch.where(conv, torch.zeros_like(err), err)
        used = torch.where(conv & ~done, torch.full_like(used, it - 1), used)
        done = done | conv
        if torch.all(conv):
            break
        df = f1 - f0
        safe = df.abs() > 1e-30 * f1.abs()
        dx = torch.where(safe, f1 * (x1 - x0) / torch.where(safe, df, torch.ones_like(df)),
                         torch.zeros_like(f1))
        x2 = torch.clamp(x1 - dx, min=1e-30)
        x2 = torch.where(conv, x1, x2)
        f2 = torch.where(conv, f1, func(x2))
        x0, f0, x1, f1 = x1, f1, x2, f2
    conv = f1.abs() < tol
    err = torch.where(conv, torch.zeros_like(err), err)
    used = torch.where(conv & ~done, torch.full_like(used, it), used)
    return x1, err, used


def newton_pstar(func, x0, tol=1e-12, max_iter=30, derivative="ad", fd_h=1e-6,
                 max_halvings=6, bracket=None):
    """Batched safeguarded Newton for p* (hlld.py:751-904): (root, err,
    iters_used, cost).  derivative "ad": exact f' by reverse-mode autograd
    through the residual; "fd": central difference, two extra residuals;
    "fd1": forward difference from f(p).  Lanes with a non-finite or
    vanishing derivative freeze; a step that does not reduce |f| is halved up
    to max_halvings times (unguarded lanes only); with a sign-changing
    bracket the step that would leave it bisects instead and the bracket
    shrinks around the sign change.  Converged lanes freeze."""
    x = x0.detach().clone()
    cost = dict(fevals=0, backward=0, halvings=0)

    def f_only(p):
        cost["fevals"] += 1
        with torch.no_grad():
            return func(p)

    def f_and_df(p):
        if derivative == "ad":
            q = p.detach().clone().requires_grad_(True)
            with torch.enable_grad():
                f = func(q)
                g = (torch.autograd.grad(f.sum(), q, allow_unused=True)[0]
                     if f.requires_grad else None)
            cost["fevals"] += 1
            cost["backward"] += 1
            if g is None:
                g = torch.zeros_like(q)
            return f.detach(), g.detach()
        h = fd_h * p.abs().clamp(min=1e-30)
        if derivative == "fd1":
            f0 = f_only(p)
            ph = p + h
            return f0, (f_only(ph) - f0) / (ph - p)
        fp, fm = f_only(p + h), f_only((p - h).clamp(min=1e-30))
        return f_only(p), (fp - fm) / (2.0 * h)

    f, df = f_and_df(x)
    n = x.shape
    if bracket is not None:
        lo, hi, flo, fhi = (b.detach().clone().reshape(n) for b in bracket)
        lo, hi, flo, fhi = (torch.where(lo <= hi, lo, hi), torch.where(lo <= hi, hi, lo),
                            torch.where(lo <= hi, flo, fhi), torch.where(lo <= hi, fhi, flo))
        guarded = (flo * fhi) < 0
    else:
        guarded = torch.zeros(n, dtype=torch.bool)
    err = torch.full(n, max_iter, dtype=torch.int32)
    iters_used = torch.full(n, max_iter, dtype=torch.int32)
    done = torch.zeros(n, dtype=torch.bool)
    dead = torch.zeros(n, dtype=torch.bool)
    it = 0
    for _ in range(max_iter):
        it += 1
        conv = f.abs() < tol
        newly = conv & ~done
        iters_used = torch.where(newly, torch.full_like(iters_used, it - 1), iters_used)
        done = done | conv
        dead = dead | ~torch.isfinite(df) | (df.abs() <= 1e-300) | ~torch.isfinite(f)
        active = ~done & ~dead
        if not bool(active.any()):
            break
        step = torch.where(active, f / torch.where(active, df, torch.ones_like(df)),
                           torch.zeros_like(f))
        lam = torch.ones_like(x)
        x_new = torch.clamp(x - step, min=1e-30)
        if bracket is not None:
            out = guarded & ((x_new <= lo) | (x_new >= hi))
            x_new = torch.where(out, 0.5 * (lo + hi), x_new)
        f_new = f_only(x_new)
        worse = active & ~guarded & ~(f_new.abs() < f.abs())
        for _h in range(max_halvings):
            if not bool(worse.any()):
                break
            cost["halvings"] += 1
            lam = torch.where(worse, lam * 0.5, lam)
            x_try = torch.clamp(x - lam * step, min=1e-30)
            f_try = f_only(x_try)
            x_new = torch.where(worse, x_try, x_new)
            f_new = torch.where(worse, f_try, f_new)
            worse = worse & ~(f_new.abs() < f.abs())
        x = torch.where(active, x_new, x)
        if bracket is not None:
            left = guarded & active & ((f_new * flo) > 0)
            right = guarded & active & ~left
            lo = torch.where(left, x_new, lo)
            flo = torch.where(left, f_new, flo)
            hi = torch.where(right, x_new, hi)
            fhi = torch.where(right, f_new, fhi)
        f_next, df_next = f_and_df(x)
        f = torch.where(active, f_next, f)
        df = torch.where(active, df_next, df)
    conv = f.abs() < tol
    err = torch.where(conv, torch.zeros_like(err), err)
    iters_used = torch.where(conv & ~done, torch.full_like(iters_used, it), iters_used)
    return x, err, iters_used, cost


def hllc_pstar(fH, uH, idir=0):
    """The HLLC star total pressure from the HLL flux and state, the "hllc"
    bracket source of hlld_flux (hlld.py:1104-1199); returns (p*, Bt1, Bt2)."""
    n, t1, t2 = idir, (idir + 1) % 3, (idir + 2) % 3
    Bn = uH[BK[n]]
    hasB = Bn.abs() > 1e-15
    Bt1, Bt2 = uH[BK[t1]], uH[BK[t2]]
    fB1, fB2 = fH[BK[t1]], fH[BK[t2]]
    z = torch.zeros_like(Bn)
    cross = torch.where(hasB, Bt1 * fB1 + Bt2 * fB2, z)
    bt2 = torch.where(hasB, Bt1 ** 2 + Bt2 ** 2, z)
    fbt2 = torch.where(hasB, fB1 ** 2 + fB2 ** 2, z)
    E = uH["tau"] + uH["D"]
    fE = fH["tau"] + fH["D"]
    qa = fE - cross
    qb = -fH[SK[n]] - E + bt2 + fbt2
    qc = uH[SK[n]] - cross
    disc = sqrt(torch.clamp(qb ** 2 - 4.0 * qa * qc, min=0.0))
    vq = sdiv(-0.5 * (qb + disc), qa)
    vl = sdiv(-qc, qb)
    vc = torch.where(qa.abs() < 1e-15, torch.where(qb.abs() < 1e-15, torch.zeros_like(qa), vl),
                     vq)
    vt1 = torch.where(hasB, sdiv(Bt1 * vc - fB1, Bn), torch.zeros_like(vc))
    vt2 = torch.where(hasB, sdiv(Bt2 * vc - fB2, Bn), torch.zeros_like(vc))
    v2 = vc ** 2 + vt1 ** 2 + vt2 ** 2
    g = 1.0 / sqrt(torch.clamp(1.0 - v2, min=1e-10))
    vB = torch.where(hasB, vc * Bn + vt1 * Bt1 + vt2 * Bt2, torch.zeros_like(vc))
    pB = fH[SK[n]] + (Bn / g) ** 2 - vc * (fE - Bn * vB)
    p0 = fH[SK[n]] - vc * fE
    ps = torch.clamp(torch.where(hasB, pB, p0), min=1e-30)
    return ps, Bt1, Bt2


def hlld_flux(sL, sR, eos, idir=0):
    """Mignone-Ugliano-Bodo relativistic HLLD as the reference implements it
    (hlld.py:1202-1543): legacy p_hll bracket from a full con2prim of the HLL
    state, Bn = 0 p0 estimate, clamped-secant p*, HLLE fallback where the
    root-find failed, the wave order is bad or B_n vanishes."""
    uL, uR, fL, fR, cmax, cmin = srmhd_fluxes(sL, sR, eos, idir)
    n = idir
    fH = {k: hll_flux(fL[k], fR[k], uL[k], uR[k], cmin, cmax) for k in CONS_KEYS}
    uH = {k: hll_state(fL[k], fR[k], uL[k], uR[k], cmin, cmax) for k in CONS_KEYS}
    EL, GL = energy_form(uL, fL)
    ER, GR = energy_form(uR, fR)

    if PSTAR_BRACKET == "hllc":
        # the HLLC star pressure of the HLL state   hlld.py:1246-1251
        p_hll, _, _ = hllc_pstar(fH, uH, idir)
        p_hll = torch.clamp(p_hll, min=1e-30)
    else:
        # bracket source: total pressure of the HLL state   hlld.py:1253-1263
        ph = cons_to_prim(uH, eos)
        p_hll, _ = eos.press_and_cs2(ph["eps"], ph["rho"])
        wh, _ = lorentz(ph)
        p_hll = p_hll + 0.5 * b_squared((ph["vx"], ph["vy"], ph["vz"]),
                                         (ph["Bx"], ph["By"], ph["Bz"]), wh)

    # p0 from the Bn = 0 quadratic   hlld.py:1265-1287
    a = fH["tau"] + fH["D"]
    b = -fH[SK[n]] - uH["tau"] - uH["D"]
    c = uH[SK[n]]
    disc = torch.clamp(b ** 2 - 4.0 * a * c, min=0.0)
    big = a.abs() > 1e-14
    sa = torch.where(big, a, torch.ones_like(a))
    sb = torch.where(big, b, torch.ones_like(b))
    vs = torch.where(big, (-b - sqrt(disc)) / (2.0 * sa), -c / sb)
    vs = torch.clamp(vs, min=-(1.0 - 1e-10), max=(1.0 - 1e-10))
    p0 = torch.clamp(fH[SK[n]] - (fH["tau"] + fH["D"]) * vs, min=1e-30)

    BnL, BnR = uL[BK[n]], uR[BK[n]]
    use_p0 = (0.5 * (BnL ** 2 + BnR ** 2) / p_hll) < 0.1
    p_lo = torch.where(use_p0, p0 * 0.99, p_hll * 0.99)
    p_hi = torch.where(use_p0, p0 * 1.01, p_hll * 1.01)
    bn_zero = (BnL.abs() < 1e-14) & (BnR.abs() < 1e-14)

    calc = StarStates(GL, GR, EL, ER, sL, sR, cmin, cmax, idir)
    f_lo, f_hi = calc.residual(p_lo), calc.residual(p_hi)
    for _ in range(5):                      # bracket expansion  hlld.py:1314-1321
        bad = (f_lo * f_hi) > 0.0
        if not torch.any(bad):
            break
        p_lo = torch.where(bad, p_lo / 2.0, p_lo)
        p_hi = torch.where(bad, p_hi * 2.0, p_hi)
        f_lo = torch.where(bad, calc.residual(p_lo), f_lo)
        f_hi = torch.where(bad, calc.residual(p_hi), f_hi)
    method, cost = PSTAR_METHOD, None
    if method == "secant":
        ps, err, used = clamped_secant(calc.residual, p_lo, p_hi, PSTAR_TOL, PSTAR_MAXIT)
    elif method in PSTAR_METHODS:
        # one start: the bracket end with the smaller residual   hlld.py:1330-1343
        start = torch.where(f_lo.abs() <= f_hi.abs(), p_lo, p_hi)
        deriv = {"newton_fd": "fd", "newton_fd1": "fd1"}.get(method, "ad")
        fd_h = PSTAR_FD_H or (1.5e-8 if deriv == "fd1" else 1e-6)
        ps, err, used, cost = newton_pstar(
            calc.residual, start, tol=PSTAR_TOL, max_iter=PSTAR_MAXIT, derivative=deriv,
            fd_h=fd_h, bracket=((p_lo, p_hi, f_lo, f_hi) if method == "newton_safe" else None))
    else:
        raise ValueError("PSTAR_METHOD=%r: expected one of %s" % (method, PSTAR_METHODS))
    calc.solve(ps)

    t1, t2 = calc.t1, calc.t2
    uaL = {BK[n]: BnL, BK[t1]: calc.B1L, BK[t2]: calc.B2L, "D": calc.DaL,
           "tau": calc.TaL, SK[n]: calc.SaL, SK[t1]: calc.S1L, SK[t2]: calc.S2L}
    uaR = {BK[n]: BnR, BK[t1]: calc.B1R, BK[t2]: calc.B2R, "D": calc.DaR,
           "tau": calc.TaR, SK[n]: calc.SaR, SK[t1]: calc.S1R, SK[t2]: calc.S2R}
    faL = {k: fL[k] + (-cmin) * (uaL[k] - uL[k]) for k in CONS_KEYS}
    faR = {k: fR[k] + cmax * (uaR[k] - uR[k]) for k in CONS_KEYS}

    # contact velocity   hlld.py:1396-1402
    fac = sdiv(1.0 - calc.KL2, -calc.SgL * calc.sqL - calc.KBL)
    laL, laR = calc.laL, calc.laR
    vc = laL - calc.Bcn * fac
    vt1 = calc.K1L - calc.Bc1 * fac
    vt2 = calc.K2L - calc.Bc2 * fac
    Bc = (calc.Bcn, calc.Bc1, calc.Bc2)
    vBc = vc * Bc[0] + vt1 * Bc[1] + vt2 * Bc[2]

    ucL = {BK[n]: Bc[0], BK[t1]: Bc[1], BK[t2]: Bc[2]}
    ucR = dict(ucL)
    ucL["D"] = uaL["D"] * sdiv(laL - calc.vaL, laL - vc)
    ucR["D"] = uaR["D"] * sdiv(laR - calc.vaR, laR - vc)
    ucL["tau"] = sdiv(laL * uaL["tau"] - faL["tau"] + ps * vc - vBc * Bc[0], laL - vc)
    ucR["tau"] = sdiv(laR * uaR["tau"] - faR["tau"] + ps * vc - vBc * Bc[0], laR - vc)
    for uc in (ucL, ucR):
        for key, vv, bb in zip((SK[n], SK[t1], SK[t2]), (vc, vt1, vt2), Bc):
            uc[key] = (uc["tau"] + ps + uc["D"]) * vv - vBc * bb

    # region selection at xi = 0   hlld.py:1448-1480
    vi = torch.zeros_like(cmin)
    masks = (vi <= -cmin, (-cmin < vi) & (vi <= laL), (laL < vi) & (vi < vc),
             (vc <= vi) & (vi < laR), (laR <= vi) & (vi < cmax), vi >= cmax)
    F, U = {}, {}
    for k in CONS_KEYS:
        fcL = faL[k] + laL * (ucL[k] - uaL[k])
        fcR = faR[k] + laR * (ucR[k] - uaR[k])
        f, u = fH[k].clone(), uH[k].clone()
        for m, fv, uv in zip(masks, (fL[k], faL[k], fcL, fcR, faR[k], fR[k]),
                             (uL[k], uaL[k], ucL[k], ucR[k], uaR[k], uR[k])):
            f = torch.where(m, fv, f)
            u = torch.where(m, uv, u)
        F[k], U[k] = f, u

    wave_bad = (calc.vaR > cmax) | (calc.vaL < -cmin) | (vc < laL) | (vc > laR)
    failed = (err > 0) | wave_bad | bn_zero
    nf = err.numel()
    conv = err == 0
    LAST_DIAG.update(
        solver="hlld", idir=idir, n_interfaces=nf,
        n_not_converged=int((err > 0).sum()), n_wave_order_bad=int(wave_bad.sum()),
        n_bn_zero=int(bn_zero.sum()), n_hlle_fallback=int(failed.sum()),
        frac_hlle_fallback=float(failed.sum()) / max(nf, 1),
        mean_iters=(float(used[conv].double().mean()) if bool(conv.any()) else float("nan")),
        pstar_method=method, pstar_cost=cost, max_iters=int(used.max()) if nf else 0)
    ps = torch.where(failed, -torch.abs(p_hll) - 1e-10, ps)
    for k in CONS_KEYS:
        F[k] = torch.where(failed, fH[k], F[k])
        U[k] = torch.where(failed, uH[k], U[k])
    return F, U, ps


SOLVERS = {"hlle": hlle_flux, "hllc": hllc_flux, "hlld": hlld_flux}
