"""The linearised (Roe-type) flux: oracle restatement of
src/physics/linearised.py:129-237.  Test infrastructure only -- the product
path never imports it.

Per face, in primitive variables V = (rho, p, v_n, v_t1, v_t2, B_t1, B_t2)
with the normal field fixed at the mean of the two sides: the Jacobians
dU/dV and dF/dV of the single-state conserved variables and normal flux (the
same side-state arithmetic as HLLD's, hlld.py:223-334) by central
differences at the mean state, A = (dU/dV)^-1 dF/dV, its eigen-decomposition
A R = R diag(lambda), the jump V_R - V_L = R alpha, and the state on the ray
xi = 0, V* = V_L + sum over lambda_k < 0 of alpha_k r_k; the flux is the
physical flux of V*.  Faces where the decomposition cannot be trusted keep
the fallback's (HLLD's) answer: non-finite Jacobians or a non-positive
density or pressure on either side, a singular dU/dV, an eigenvalue or an
eigenvector with a non-negligible imaginary part, an eigenvector matrix with
2-norm condition number >= cond_max, or a V* that is not physical.

The linear algebra is the reference's own choice of library routines
(torch.linalg.solve, eig and cond, i.e. LAPACK getrf/getrs, geev and gesdd
on the CPU), called on the same operands in the same order, so the oracle
reproduces the reference bit for bit on a host with the same torch build.
"""

import torch

from .physics import CONS_KEYS, side_state, sqrt
from .riemann import hlld_flux

LAST_LINEAR: dict = {}

# the tangential axes of each normal direction (linearised.py:62-67): NOT the
# cyclic order of the HLLD code -- for idir = 1 the first tangent is x
_TANGENTS = ((1, 2), (0, 2), (0, 1))
_V = ("vx", "vy", "vz")
_B = ("Bx", "By", "Bz")
_S = ("Sx", "Sy", "Sz")


def axes(idir):
    """(velocity keys, field keys, momentum keys), each (normal, t1, t2)."""
    t1, t2 = _TANGENTS[idir]
    return tuple(_V[a] for a in (idir, t1, t2)), tuple(_B[a] for a in (idir, t1, t2)), \
        tuple(_S[a] for a in (idir, t1, t2))


def _prims(V, Bn, eos, vk, bk):
    """The primitive dict of the columns V (linearised.py:70-75)."""
    d = {"rho": V[0], "p": V[1], vk[0]: V[2], vk[1]: V[3], vk[2]: V[4],
         bk[0]: Bn, bk[1]: V[5], bk[2]: V[6]}
    d["eps"] = eos.eps__press_rho(V[1], V[0])
    return d


def conserved_and_flux(V, Bn, eos, idir):
    """The seven conserved variables and normal fluxes (7, N) of the
    primitive columns (linearised.py:78-87): D, tau, S_n, S_t1, S_t2, B_t1,
    B_t2."""
    vk, bk, sk = axes(idir)
    u, f, _, _ = side_state(_prims(V, Bn, eos, vk, bk), eos, idir)
    order = ("D", "tau", sk[0], sk[1], sk[2], bk[1], bk[2])
    return torch.stack([u[k] for k in order]), torch.stack([f[k] for k in order])


def solve_each(M, B):
    """M^-1 B per element with a per-element flag (linearised.py:90-104):
    LAPACK refuses a whole batch when one matrix is exactly singular, so the
    batch is retried element by element."""
    try:
        return torch.linalg.solve(M, B), torch.ones(M.shape[0], dtype=torch.bool)
    except RuntimeError:
        pass
    X = torch.zeros_like(B)
    ok = torch.zeros(M.shape[0], dtype=torch.bool)
    for e in range(M.shape[0]):
        try:
            X[e] = torch.linalg.solve(M[e], B[e])
            ok[e] = True
        except RuntimeError:
            pass
    return X, ok


def eig_each(A):
    """torch.linalg.eig per element with a flag (linearised.py:107-126)."""
    n = A.shape[0]
    try:
        w, R = torch.linalg.eig(A)
        return w, R, torch.ones(n, dtype=torch.bool)
    except RuntimeError:
        pass
    w = torch.zeros(n, A.shape[1], dtype=torch.complex128)
    R = torch.zeros(n, A.shape[1], A.shape[2], dtype=torch.complex128)
    ok = torch.zeros(n, dtype=torch.bool)
    for e in range(n):
        try:
            w[e], R[e] = torch.linalg.eig(A[e])
            ok[e] = True
        except RuntimeError:
            pass
    return w, R, ok


def side_vectors(sL, sR, idir):
    """V_L, V_R (7, N) and the face's normal field, the mean of the sides'."""
    vk, bk, _ = axes(idir)
    keys = ("rho", "p", vk[0], vk[1], vk[2], bk[1], bk[2])
    VL = torch.stack([sL[k].reshape(-1).to(torch.float64) for k in keys])
    VR = torch.stack([sR[k].reshape(-1).to(torch.float64) for k in keys])
    Bn = 0.5 * (sL[bk[0]].reshape(-1).to(torch.float64) + sR[bk[0]].reshape(-1).to(torch.float64))
    return VL, VR, Bn


def jacobians(Vm, Bn, eos, idir, fd_rel=1e-5):
    """dU/dV and dF/dV (N, 7, 7) at the mean state by central differences
    with the per-variable step fd_rel * scale (linearised.py:156-174)."""
    one = torch.ones_like(Vm[2])
    rp = sqrt(Vm[1].abs())
    scale = torch.stack([Vm[0].abs(), Vm[1].abs(), one, one, one, rp + Vm[5].abs(),
                         rp + Vm[6].abs()])
    h = fd_rel * scale.clamp(min=1e-300)
    N = Vm.shape[1]
    dU = torch.zeros(N, 7, 7, dtype=torch.float64)
    dF = torch.zeros(N, 7, 7, dtype=torch.float64)
    for j in range(7):
        Vp, Vq = Vm.clone(), Vm.clone()
        Vp[j] = Vm[j] + h[j]
        Vq[j] = Vm[j] - h[j]
        Up, Fp = conserved_and_flux(Vp, Bn, eos, idir)
        Uq, Fq = conserved_and_flux(Vq, Bn, eos, idir)
        dU[:, :, j] = ((Up - Uq) / (2.0 * h[j])).T
        dF[:, :, j] = ((Fp - Fq) / (2.0 * h[j])).T
    return dU, dF


def decompose(A, imag_tol=1e-6, cond_max=1e12):
    """Eigenvalues (n, 7), unit-norm real eigenvectors as columns (n, 7, 7)
    and the per-element trust flag (linearised.py:185-204): the phase of a
    computed eigenvector is taken off by its largest component before its
    real part is kept."""
    w, Rc, ok = eig_each(A)
    ok = ok & (w.imag.abs() <= imag_tol * (1.0 + w.real.abs())).all(dim=1)
    big = torch.argmax(Rc.abs(), dim=1, keepdim=True)
    ph = torch.gather(Rc, 1, big)
    Rc = Rc * (ph.conj() / ph.abs().clamp(min=1e-300))
    ok = ok & (Rc.imag.abs() <= imag_tol * (1.0 + Rc.real.abs())).all(dim=(1, 2))
    lam = w.real
    Rm = Rc.real
    Rm = Rm / torch.linalg.norm(Rm, dim=1, keepdim=True).clamp(min=1e-300)
    cond = torch.linalg.cond(Rm)
    return lam, Rm, ok & torch.isfinite(cond) & (cond < cond_max)


def ray_state(VL, VR, lam, Rm, idx):
    """V* of the faces idx, V_L where the eigenvector solve fails, and the
    solve flags (linearised.py:207-217)."""
    dV = (VR - VL)[:, idx].T.unsqueeze(-1)
    alpha, solved = solve_each(Rm, dV)
    alpha = alpha.squeeze(-1)
    left = (lam < 0.0).to(torch.float64)
    step = torch.einsum("nkj,nj->nk", Rm, alpha * left)
    return torch.where(solved[None, :], VL[:, idx] + step.T, VL[:, idx]), solved


def star_pressure(Vi, Bn):
    """Total pressure of the ray state (linearised.py:231-234)."""
    W2 = 1.0 / (1.0 - (Vi[2] ** 2 + Vi[3] ** 2 + Vi[4] ** 2))
    eta = Bn * Vi[2] + Vi[5] * Vi[3] + Vi[6] * Vi[4]
    b2 = (Bn ** 2 + Vi[5] ** 2 + Vi[6] ** 2) / W2 + eta ** 2
    return Vi[1] + 0.5 * b2


def linearised_flux(sL, sR, eos, idir=0, fallback=hlld_flux, fd_rel=1e-5, imag_tol=1e-6,
                    cond_max=1e12, return_mask=False):
    """linearised.py:129-237: (F, U, p_star) for every face, plus the mask of
    the faces this solver answered when return_mask."""
    F0, U0, p0 = fallback(sL, sR, eos, idir=idir)
    shape = sL["rho"].shape
    VL, VR, Bn = side_vectors(sL, sR, idir)
    N = VL.shape[1]
    Vm = 0.5 * (VL + VR)
    dU, dF = jacobians(Vm, Bn, eos, idir, fd_rel)

    ok = torch.isfinite(dU).all(dim=(1, 2)) & torch.isfinite(dF).all(dim=(1, 2))
    ok &= (VL[0] > 0) & (VR[0] > 0) & (VL[1] > 0) & (VR[1] > 0)
    A = torch.zeros(N, 7, 7, dtype=torch.float64)
    if ok.any():
        i = torch.nonzero(ok).reshape(-1)
        A[i], solved = solve_each(dU[i], dF[i])
        ok[i] = solved

    lam = torch.zeros(N, 7, dtype=torch.float64)
    Rm = torch.zeros(N, 7, 7, dtype=torch.float64)
    if ok.any():
        i = torch.nonzero(ok).reshape(-1)
        lam[i], Rm[i], ok[i] = decompose(A[i], imag_tol, cond_max)

    Vs = VL.clone()
    if ok.any():
        i = torch.nonzero(ok).reshape(-1)
        Vs[:, i], solved = ray_state(VL, VR, lam[i], Rm[i], i)
        ok[i] &= solved
    v2 = Vs[2] ** 2 + Vs[3] ** 2 + Vs[4] ** 2
    ok &= (Vs[0] > 0) & (Vs[1] > 0) & (v2 < 1.0) & torch.isfinite(Vs).all(dim=0)

    F = {k: v.clone() for k, v in F0.items()}
    U = {k: v.clone() for k, v in U0.items()}
    p_star = p0.clone()
    if ok.any():
        i = torch.nonzero(ok).reshape(-1)
        Vi = Vs[:, i]
        vk, bk, _ = axes(idir)
        u, f, _, _ = side_state(_prims(Vi, Bn[i], eos, vk, bk), eos, idir)
        for k in CONS_KEYS:
            F[k].reshape(-1)[i] = f[k]
            U[k].reshape(-1)[i] = u[k]
        p_star.reshape(-1)[i] = star_pressure(Vi, Bn[i])
    n_lin = int(ok.sum())
    LAST_LINEAR.clear()
    LAST_LINEAR.update(n=N, n_linear=n_lin, frac_fallback=1.0 - n_lin / max(N, 1))
    if return_mask:
        return F, U, p_star, ok.reshape(shape)
    return F, U, p_star
