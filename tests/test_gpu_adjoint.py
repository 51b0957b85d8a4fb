"""The hand-written adjoint of the 2D step (srmhd2d_stage_adjoint) on the
differentiable relativistic Orszag-Tang problem: PLM (MC) + HLLD + upwind CT
+ SSP-RK2, periodic, dt frozen from the A = 1 forward run, loss =
sum (rho_final - rho_target)^2 with the target from A = 1.05.

Checks: dL/dA and a directional derivative in the initial density field
against central finite differences of the same FP64 rollout (h = 1e-6),
bitwise independence of the checkpoint interval, and determinism.

The density direction is taken at a generic base density (no two cells of
a stencil equal): at the problem's uniform rho = 1, and at any base field
symmetric about the cell pairs around its extrema, PLM slopes sit at the
limiter's kink (MC of two zero differences), where the loss has only a
directional derivative that is nonlinear in the direction (finite
differences see mc(d rho); the adjoint, like torch autograd of the
reference, the branch the forward took).  dL/dA is taken at the problem's
own base state: scaling B is a radial direction, along which the limiters
(positively homogeneous) are linear.
"""

import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _problem(n, steps, K, limiter="mc", solver="hlld", scheme="rk2"):
    from srmhd import adjoint as AD, eos as E, grid as G
    dev = torch.device("cuda:0")
    eos = E.hybrid_eos(0.0, 4.0 / 3.0, 4.0 / 3.0)
    ng = 3 if limiter in ("weno5z", "mp5") else 2
    g = G.Grid2D(0.0, 2 * math.pi, n, 0.0, 2 * math.pi, n, ng=ng, device=dev)
    roll = AD.Rollout2D(g, eos, solver=solver, limiter=limiter, bc_x="periodic",
                        bc_y="periodic", scheme=scheme, checkpoint_every=K)
    prob = AD.OTProblem(g, eos)
    rho0 = torch.ones(n, n, dtype=torch.float64, device=dev)
    one = torch.tensor(1.0, dtype=torch.float64, device=dev)
    U, P, _ = prob.initial_state(roll, rho0, one)
    dts = AD.frozen_dts(roll, U, P, steps, 0.4)
    target = AD.ot_final_rho(prob, roll, rho0, one * 1.05, dts)
    return AD, roll, prob, rho0, one, dts, target


def _grad(AD, roll, prob, rho0, A0, dts, target):
    rho = rho0.clone().requires_grad_(True)
    A = A0.clone().requires_grad_(True)
    rf = AD.OTRollout.apply(rho, A, prob, roll, dts)
    L = ((rf - target) ** 2).sum()
    L.backward()
    torch.cuda.synchronize()
    return float(L.detach()), rho.grad.clone(), A.grad.clone()


def _loss(AD, roll, prob, rho, A, dts, target):
    r = AD.ot_final_rho(prob, roll, rho, A, dts)
    return float(((r - target) ** 2).sum())


def _generic(n, device):
    x = (torch.arange(n, dtype=torch.float64, device=device) + 0.5) * (2 * math.pi / n)
    base = 1.0 + 0.1 * torch.sin(x + 0.3)[:, None] * torch.sin(x + 0.7)[None, :]
    dirn = (0.3 * torch.sin(x + 0.1)[:, None] * torch.cos(2 * x + 0.2)[None, :]
            + 0.2 * torch.cos(x + 0.4)[None, :])
    return base, dirn


@pytest.mark.parametrize("limiter,solver", [("mc", "hlld"), ("mc", "hlle"), ("weno5z", "hlld")])
def test_adjoint_dLdA_matches_central_differences(limiter, solver):
    """dL/dA at the problem's base state (rho = 1, A = 1): measured 9e-10
    (MC/HLLD), 1.6e-7 (MC/HLLE), 4e-8 (WENO5-Z/HLLD)."""
    _dev()
    print(f"dL/dA adjoint vs FD: OT 32^2, 8 RK2 steps, {limiter} {solver}", flush=True)
    AD, roll, prob, rho0, one, dts, target = _problem(32, 8, 3, limiter, solver)
    L, g_rho, g_A = _grad(AD, roll, prob, rho0, one, dts, target)
    h = 1e-6
    fd_A = (_loss(AD, roll, prob, rho0, one + h, dts, target)
            - _loss(AD, roll, prob, rho0, one - h, dts, target)) / (2 * h)
    eA = abs(float(g_A) - fd_A) / abs(fd_A)
    print(f"L = {L:.6e}; dL/dA adjoint {float(g_A):.12e} FD {fd_A:.12e} rel {eA:.2e}", flush=True)
    assert eA < 1e-6


@pytest.mark.parametrize("limiter,solver,tol", [("weno5z", "hlld", 1e-6), ("mc", "hlld", 1e-4)])
def test_adjoint_density_direction_matches_central_differences(limiter, solver, tol):
    """A smooth direction in the initial density field at a generic base.
    WENO5-Z is smooth there: measured 1.9e-7.  MC is piecewise linear, and
    the finite difference straddles some of its switches: the error varies
    erratically with h (MC/HLLE at 8 steps: 1e-3, 4e-4, 7e-5, 4e-4, 7e-7,
    4e-5 for h = 1e-3 .. 1e-8) instead of converging, the signature of kinks
    and of the con2prim's own 1e-15 convergence noise, not of a wrong
    derivative; MC is pinned exactly by the torch-autograd test below."""
    _dev()
    print(f"density-direction adjoint vs FD: OT 32^2, 8 RK2 steps, {limiter} {solver}",
          flush=True)
    AD, roll, prob, rho0, one, dts, target = _problem(32, 8, 3, limiter, solver)
    base, dirn = _generic(32, rho0.device)
    _, g_rb, _ = _grad(AD, roll, prob, base, one, dts, target)
    h = 1e-6
    fd_r = (_loss(AD, roll, prob, base + h * dirn, one, dts, target)
            - _loss(AD, roll, prob, base - h * dirn, one, dts, target)) / (2 * h)
    ad_r = float((g_rb * dirn).sum())
    er = abs(ad_r - fd_r) / abs(fd_r)
    print(f"rho-direction adjoint {ad_r:.12e} FD {fd_r:.12e} rel {er:.2e}", flush=True)
    assert er < tol


def test_adjoint_checkpoint_interval_and_determinism():
    _dev()
    print("adjoint checkpoint interval 1 / 4 / 10", flush=True)
    out = []
    for K in (1, 4, 10, 4):
        AD, roll, prob, rho0, one, dts, target = _problem(24, 10, K)
        out.append(_grad(AD, roll, prob, rho0, one, dts, target))
    for L, gr, gA in out[1:]:
        assert L == out[0][0]
        assert torch.equal(gr, out[0][1]) and torch.equal(gA, out[0][2])


def _oracle_grad(n, ng, limiter, solver, dts, target, rho_phys, A):
    """Torch autograd through the oracle's step (CPU, IEEE arithmetic: the
    trajectory, and so every branch, is bit-identical to the device's)."""
    import oracle.eos as OE
    import oracle.riemann as OR
    import oracle.step2d as OS
    from oracle.grid import Grid2D
    eos = OE.ideal(4.0 / 3.0)
    g = Grid2D(0.0, 2 * math.pi, n, 0.0, 2 * math.pi, n, ng=ng)
    r = rho_phys.detach().cpu().clone().requires_grad_(True)
    a = A.detach().cpu().clone().requires_grad_(True)
    rho = torch.nn.functional.pad(r[None, None], (ng, ng, ng, ng), mode="circular")[0, 0]
    amp = 0.99 / math.sqrt(2.0)
    z = torch.zeros(g.shape, dtype=torch.float64)
    P = {"rho": rho, "p": torch.ones_like(z), "vx": -amp * torch.sin(g.Y),
         "vy": amp * torch.sin(g.X), "vz": z, "Bx": a * (-torch.sin(g.Y)),
         "By": a * torch.sin(2.0 * g.X), "Bz": z}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    XF = g.xf.unsqueeze(1).expand(g.nxt + 1, g.nyt + 1)
    YF = g.yf.unsqueeze(0).expand(g.nxt + 1, g.nyt + 1)
    Az0 = torch.cos(YF) + 0.5 * torch.cos(2.0 * XF)
    Bxf = a * ((Az0[:, 1:] - Az0[:, :-1]) / g.dy)
    Byf = a * (-(Az0[1:, :] - Az0[:-1, :]) / g.dx)
    c = {k: v.clone() for k, v in OS.prims_to_cons_2d(P, g).items()}
    st = OS.State2D({k: c[k] for k in OS.EVOLVED_KEYS}, Bxf, Byf, P)
    st = OS.sync_state(st, g, eos, "periodic", "periodic")
    flux = {"hlld": OR.hlld_flux, "hlle": OR.hlle_flux, "hllc": OR.hllc_flux}[solver]
    for dt in dts:
        st, _ = OS.rk_step_ct(st, g, eos, dt, scheme="rk2", bc_x="periodic", bc_y="periodic",
                              flux_fn=flux, limiter=limiter)
    rf = st.prims["rho"][ng:ng + n, ng:ng + n]
    L = ((rf - target.cpu()) ** 2).sum()
    L.backward()
    return float(L.detach()), rf.detach(), r.grad, a.grad


@pytest.mark.parametrize("limiter,solver", [("mc", "hlld"), ("weno5z", "hlld")])
def test_adjoint_matches_torch_autograd_of_the_oracle(limiter, solver):
    """At the problem's own base state (rho = 1, A = 1), where the MC slopes
    of rho sit at the limiter's kink: torch autograd of the oracle takes the
    same branches as the device's forward (bit-identical trajectory), and the
    adjoint follows torch's conventions at non-differentiable points (the
    branch taken; the mean of both tangents at a tie of maximum / minimum),
    so it must reproduce the oracle's gradient.  The oracle differentiates
    through the iterations (HLLD's secant to 1e-10, the batch-coupled 60-step
    con2prim) where the adjoint uses the implicit derivative at the root:
    measured agreement 5.6e-12 (dL/dA) and 3.6e-11 (max over the density
    field) with MC, 1.6e-9 and 1.3e-10 with WENO5-Z."""
    _dev()
    print(f"adjoint vs torch autograd of the oracle: OT 12^2, 3 RK2 steps, {limiter} {solver}",
          flush=True)
    n, steps = 12, 3
    AD, roll, prob, rho0, one, dts, target = _problem(n, steps, 2, limiter, solver)
    L, g_rho, g_A = _grad(AD, roll, prob, rho0, one, dts, target)
    ng = roll.grid.ng
    Lo, rfo, go_rho, go_A = _oracle_grad(n, ng, limiter, solver, dts, target, rho0, one)
    rf = AD.ot_final_rho(prob, roll, rho0, one, dts)
    assert torch.equal(rf.cpu(), rfo), "forward trajectories differ"
    eA = abs(float(g_A) - float(go_A)) / abs(float(go_A))
    er = float((g_rho.cpu() - go_rho).abs().max() / go_rho.abs().max())
    print(f"L {L:.12e} oracle {Lo:.12e}; dL/dA {float(g_A):.12e} oracle {float(go_A):.12e} "
          f"rel {eA:.2e}; max |d rho| rel {er:.2e}", flush=True)
    assert abs(L - Lo) <= 1e-13 * Lo   # the sums reduce in different orders
    assert eA < 1e-8 and er < 1e-8
