"""Kastaun con2prim (oracle restatement of src/physics/c2p.py).

The secant of ``invert`` (c2p.py:188-236) is BATCH-COUPLED: the loop keeps
stepping every cell until all cells of the batch satisfy |f| < tol, or 60
iterations pass, so a cell's result depends on its batch mates.  That is
reproduced here exactly (``coupled=True``, the default).  ``coupled=False``
runs each cell as a batch of one -- used only to bound the coupling effect.
"""

import torch

from .physics import TINY, sdiv, sqrt


class Kastaun:
    """Per-batch precomputation of c2p.py:48-106."""

    def __init__(self, cons, eos):
        D = cons["D"]
        Sx, Sy, Sz = cons["Sx"], cons["Sy"], cons["Sz"]
        tau = torch.clamp(cons["tau"], min=0.0)
        Bx, By, Bz = cons["Bx"], cons["By"], cons["Bz"]
        S2 = Sx ** 2 + Sy ** 2 + Sz ** 2
        Sn = sqrt(S2)
        Smax = D + tau
        fac = torch.where(Sn > Smax, 0.9999 * Smax / (Sn + TINY), torch.ones_like(Sn))
        Sx, Sy, Sz = Sx * fac, Sy * fac, Sz * fac
        S2 = Sx ** 2 + Sy ** 2 + Sz ** 2
        B2 = Bx ** 2 + By ** 2 + Bz ** 2
        BS = Bx * Sx + By * Sy + Bz * Sz
        self.D, self.eos = D, eos
        self.B = (Bx, By, Bz)
        self.q = tau / (D + TINY)
        self.r = tuple(s / (D + TINY) for s in (Sx, Sy, Sz))
        self.r2 = S2 / (D ** 2 + TINY)
        sD = sqrt(D + TINY)
        self.bt = tuple(b / sD for b in (Bx, By, Bz))
        self.bt2 = B2 / (D + TINY)
        self.br = BS / (D * sD + TINY)
        self.bperp2 = self.r2 * self.bt2 - self.br ** 2
        self.v02 = self.r2 / (self.r2 + 1.0)      # h_min = 1  (c2p.py:101-104)

    def chi(self, mu):
        return 1.0 / (1.0 + mu * self.bt2)

    def rbar2(self, mu):
        chi = self.chi(mu)
        return self.r2 * chi ** 2 + mu * chi * (1.0 + chi) * self.br ** 2

    def qbar(self, mu):
        chi = self.chi(mu)
        return self.q - 0.5 * self.bt2 - 0.5 * (mu * chi) ** 2 * self.bperp2

    def fa(self, mu):
        """mu sqrt(h_min^2 + rbar^2) - 1   c2p.py:124-127."""
        return mu * sqrt(1.0 + self.rbar2(mu)) - 1.0

    def master(self, mu):
        """c2p.py:129-158 -> (f, rho, eps, W, v^2)."""
        qb = self.qbar(mu)
        rb2 = self.rbar2(mu)
        vh2 = torch.clamp(mu ** 2 * rb2, max=self.v02)
        Wh = 1.0 / sqrt(torch.clamp(1.0 - vh2, min=1e-10))
        rho = torch.clamp(self.D / Wh, min=torch.full_like(Wh, 1e-12),
                          max=torch.full_like(Wh, 1e15))
        eps0 = Wh * (qb - mu * rb2) + vh2 * Wh ** 2 / (1.0 + Wh)
        elo, ehi = self.eos.eps_range__rho(rho)
        eps = torch.clamp(eps0, min=elo, max=ehi)
        p = self.eos.press__eps_rho(eps, rho)
        a = p / (rho * (1.0 + eps) + TINY)
        nuA = (1 + a) * (1.0 + eps) / Wh
        nuB = (1 + a) * (1.0 + qb - mu * rb2)
        nu = torch.maximum(nuA, nuB)
        return mu - sdiv(1.0, nu + mu * rb2), rho, eps, Wh, vh2

    def mu_plus(self):
        """50-step bisection of fa on [0, 1]  (c2p.py:162-184).  Every bracket
        starts at width 1 and halves exactly, so the early stop always fires
        at the 50th iteration: the loop is batch-independent."""
        lo = torch.zeros_like(self.D)
        hi = 1.0 / (torch.ones_like(self.D) + TINY)
        for _ in range(50):
            mid = 0.5 * (lo + hi)
            left = (self.fa(lo) * self.fa(mid)) <= 0.0
            hi = torch.where(left, mid, hi)
            lo = torch.where(left, lo, mid)
        return hi

    def secant(self, n_iter=60, tol=1e-15):
        """c2p.py:188-236 -> (mu, converged, good_bracket)."""
        x1 = self.mu_plus() * (1.0 - 1e-10)
        x0 = torch.zeros_like(self.D)
        f0 = self.master(x0)[0]
        f1 = self.master(x1)[0]
        good = (f0 * f1) < 0.0
        for _ in range(n_iter):
            if torch.all(f1.abs() < tol):
                break
            df = f1 - f0
            safe = df.abs() > 1e-30
            mid = 0.5 * (x0 + x1)
            x2 = torch.where(safe, x1 - f1 * (x1 - x0) / torch.where(safe, df, torch.ones_like(df)), mid)
            x2 = torch.where((x2 - x0) * (x2 - x1) < 0.0, x2, mid)
            f2 = self.master(x2)[0]
            left = (f0 * f2) <= 0.0
            x0 = torch.where(left, x0, x1)
            f0 = torch.where(left, f0, f1)
            x1, f1 = x2, f2
        return x1, f1.abs() < tol, good

    def bisection(self, n_iter=200, tol=1e-15):
        """c2p.py:238-266 (batch-independent) -> (mu, converged, good)."""
        x1 = self.mu_plus() * (1.0 - 1e-10)
        x0 = torch.zeros_like(self.D)
        f0 = self.master(x0)[0]
        f1 = self.master(x1)[0]
        good = (f0 * f1) < 0.0
        for _ in range(n_iter):
            xm = 0.5 * (x0 + x1)
            fm = self.master(xm)[0]
            left = (f0 * fm) <= 0.0
            x1 = torch.where(left, xm, x1)
            f1 = torch.where(left, fm, f1)
            x0 = torch.where(left, x0, xm)
            f0 = torch.where(left, f0, fm)
        pick = f0.abs() < f1.abs()
        return torch.where(pick, x0, x1), torch.where(pick, f0, f1).abs() < tol, good

    def finish(self, mu):
        """c2p.py:268-310 -> primitive dict and the v-clamp mask."""
        _, rho, eps, _, _ = self.master(mu)
        mc = mu * self.chi(mu)
        v = [mc * (self.r[i] + (mu * self.br) * self.bt[i]) for i in range(3)]
        v2 = torch.stack(v, dim=1).pow(2).sum(dim=1)
        lim = 1.0 - 1e-10
        norm = torch.where(v2 > lim, sqrt(torch.clamp(v2, max=lim) / (v2 + TINY)),
                           torch.ones_like(v2))
        v = [c * norm for c in v]
        p = self.eos.press__eps_rho(eps, rho)
        P = {"rho": rho, "vx": v[0], "vy": v[1], "vz": v[2], "p": p, "eps": eps,
             "Bx": self.B[0], "By": self.B[1], "Bz": self.B[2]}
        return P, v2 > lim


_C2P_IN = ("D", "Sx", "Sy", "Sz", "tau", "Bx", "By", "Bz")


def cons_to_prim(cons, eos, atmo_rho=1e-10, atmo_eps=1e-10, return_status=False,
                 coupled=True):
    """c2p.py:313-363: secant, bisection redo of unconverged cells with a good
    bracket, atmosphere floor."""
    if not coupled:
        return _per_cell(cons, eos, atmo_rho, atmo_eps, return_status)
    k = Kastaun(cons, eos)
    mu, conv, good = k.secant()
    P, vcl = k.finish(mu)
    redo = ~conv & good
    if bool(redo.any()):
        idx = torch.nonzero(redo).reshape(-1)
        ks = Kastaun({c: cons[c][idx] for c in _C2P_IN}, eos)
        mu2, conv2, _ = ks.bisection()
        P2, vcl2 = ks.finish(mu2)
        P = {c: v.clone() for c, v in P.items()}   # out of place for torch autograd
        for c in P:
            P[c][idx] = P2[c]
        conv = conv.clone()
        conv[idx] = conv2
        vcl[idx] = vcl2
    atmo = P["rho"] < atmo_rho
    P["rho"] = torch.where(atmo, torch.full_like(P["rho"], atmo_rho), P["rho"])
    P["eps"] = torch.where(atmo, torch.full_like(P["eps"], atmo_eps), P["eps"])
    for c in ("vx", "vy", "vz"):
        P[c] = torch.where(atmo, torch.zeros_like(P[c]), P[c])
    P["p"] = eos.press__eps_rho(P["eps"], P["rho"])
    if return_status:
        return P, {"bracket_ok": good, "converged": conv, "v_clamped": vcl,
                   "atmo_floored": atmo}
    return P


def _per_cell(cons, eos, atmo_rho, atmo_eps, return_status):
    n = cons["D"].numel()
    outs = [cons_to_prim({c: cons[c].reshape(-1)[i:i + 1] for c in _C2P_IN}, eos,
                         atmo_rho, atmo_eps, True) for i in range(n)]
    P = {c: torch.cat([o[0][c] for o in outs]) for c in outs[0][0]}
    S = {c: torch.cat([o[1][c] for o in outs]) for c in outs[0][1]}
    return (P, S) if return_status else P
