"""Write the golden vectors that pin the oracle to the reference solver.

Run in a container that holds the reference checkout (default
/root/reference, or the path in $HLLD_AI_REFERENCE):

    python tests/golden/make_golden.py

It imports the reference's own functions for exactly the parts that no
closed form or invariant pins (hlld_flux/hllc_flux branch decisions and
values, the batch-coupled Kastaun con2prim,
reconstruct_prims with its PCM fallback, rk_step_ct, driver.run on
Balsara 1, run_tubes), feeds them inputs generated here from fixed seeds,
and stores inputs and outputs as .npz data.  No reference source is copied;
the fixtures are data only.  The GPU tests never need the reference: they
read these files.
"""

import os
import sys
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
REF = os.environ.get("HLLD_AI_REFERENCE", "/root/reference")
sys.path.insert(0, REF)

from src.physics import hlld as R_hlld                               # noqa: E402
from src.physics.c2p import conservative_to_primitive                # noqa: E402
from src.physics.driver import run as R_run                          # noqa: E402
from src.physics.driver2d import compute_dt_2d, prims_to_cons_2d, rk_step_ct, sync_state  # noqa: E402
from src.physics.eos import hybrid_eos                               # noqa: E402
from src.physics.grid import Grid2D                                  # noqa: E402
from src.physics.initial_data2d import b_from_potential, magnetic_rotor  # noqa: E402
from src.physics.reconstruction import ghosts_needed, reconstruct_prims  # noqa: E402
from src.physics.state import EVOLVED_KEYS, State2D                  # noqa: E402
from src.physics.tube_seed import run_tubes                          # noqa: E402

PRIM9 = ("rho", "vx", "vy", "vz", "p", "eps", "Bx", "By", "Bz")
CONS8 = ("D", "tau", "Sx", "Sy", "Sz", "Bx", "By", "Bz")


def _np(t):
    return t.detach().cpu().numpy()


def random_prims(n, eos, rng, max_W=5.0, bscale=3.0):
    """Admissible random primitive states, sampled through the Lorentz factor."""
    rho = 10.0 ** rng.uniform(-1.0, 1.0, n)
    p = 10.0 ** rng.uniform(-1.5, 1.0, n)
    W = 1.0 + 10.0 ** rng.uniform(-3.0, np.log10(max_W - 1.0), n)
    sp = np.sqrt(1.0 - 1.0 / W ** 2)
    th = rng.uniform(0.0, np.pi, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    B = rng.uniform(-bscale, bscale, (3, n))
    s = {"rho": rho, "p": p, "vx": sp * np.sin(th) * np.cos(ph),
         "vy": sp * np.sin(th) * np.sin(ph), "vz": sp * np.cos(th),
         "Bx": B[0], "By": B[1], "Bz": B[2]}
    s = {k: torch.from_numpy(v.copy()) for k, v in s.items()}
    s["eps"] = eos.eps__press_rho(s["p"], s["rho"])
    return s


def gen_flux():
    eos = hybrid_eos(K=0.0, gamma=5.0 / 3.0, gamma_th=5.0 / 3.0)
    rng = np.random.default_rng(20260824)
    n = 384
    sL = random_prims(n, eos, rng)
    sR = random_prims(n, eos, rng)
    # mild jumps: a second batch near-continuous, where HLLD rarely falls back
    sR2 = {k: v.clone() for k, v in sL.items()}
    for k in ("rho", "p"):
        sR2[k] = sL[k] * torch.from_numpy(1.0 + 0.2 * rng.uniform(-1, 1, n))
    for k in ("vx", "vy", "vz"):
        sR2[k] = sL[k] * torch.from_numpy(1.0 + 0.1 * rng.uniform(-1, 1, n))
    sR2["eps"] = eos.eps__press_rho(sR2["p"], sR2["rho"])
    for k in ("Bx", "By", "Bz"):
        sR[k] = sL[k]
        sR2[k] = sL[k]
    out = {}
    for tag, right in (("a", sR), ("b", sR2)):
        for k in PRIM9:
            out[f"{tag}_L_{k}"] = _np(sL[k])
            out[f"{tag}_R_{k}"] = _np(right[k])
        for name, fn in (("hlle", R_hlld.hlle_flux), ("hllc", R_hlld.hllc_flux),
                         ("hlld", R_hlld.hlld_flux)):
            for idir in (0, 1, 2):
                F, U, ps = fn(sL, right, eos, idir=idir)
                for k in CONS8:
                    out[f"{tag}_{name}{idir}_F_{k}"] = _np(F[k])
                    out[f"{tag}_{name}{idir}_U_{k}"] = _np(U[k])
                out[f"{tag}_{name}{idir}_pstar"] = _np(ps)
                if name == "hlld":
                    d = R_hlld.LAST_DIAG
                    out[f"{tag}_{name}{idir}_diag"] = np.array(
                        [d["n_not_converged"], d["n_wave_order_bad"], d["n_bn_zero"],
                         d["n_hlle_fallback"], d["max_iters"]], dtype=np.int64)
                    out[f"{tag}_{name}{idir}_mean_iters"] = np.array(d["mean_iters"])
    np.savez_compressed(HERE / "flux_random.npz", **out)


def gen_pstar():
    """hlld_flux with the non-default p* root finders and bracket
    (hlld.py:751-905, 1046-1098, 1104-1199, 1323-1343), selected through the
    module globals the reference reads at call time, on the flux_random
    interface pairs."""
    g = np.load(HERE / "flux_random.npz")
    eos = hybrid_eos(K=0.0, gamma=5.0 / 3.0, gamma_th=5.0 / 3.0)
    variants = [("newton", "legacy", 1e-10), ("newton_fd", "legacy", 1e-10),
                ("newton_fd1", "legacy", 1e-10), ("newton_safe", "legacy", 1e-10),
                ("secant", "hllc", 1e-10), ("newton_safe", "hllc", 1e-10),
                ("secant", "legacy", 1e-13)]
    out = {}
    saved = (R_hlld._PSTAR_METHOD, R_hlld._PSTAR_BRACKET, R_hlld._PSTAR_TOL)
    try:
        for method, bracket, tol in variants:
            R_hlld._PSTAR_METHOD, R_hlld._PSTAR_BRACKET, R_hlld._PSTAR_TOL = method, bracket, tol
            for tag in ("a", "b"):
                sL = {k: torch.from_numpy(g[f"{tag}_L_{k}"].copy()) for k in PRIM9}
                sR = {k: torch.from_numpy(g[f"{tag}_R_{k}"].copy()) for k in PRIM9}
                for idir in (0, 1, 2):
                    F, U, ps = R_hlld.hlld_flux(sL, sR, eos, idir=idir)
                    pre = f"{method}_{bracket}_{tol:g}_{tag}{idir}"
                    for k in CONS8:
                        out[f"{pre}_F_{k}"] = _np(F[k])
                        if idir == 0:        # the state comes from the same assembly
                            out[f"{pre}_U_{k}"] = _np(U[k])
                    out[f"{pre}_pstar"] = _np(ps)
                    d = R_hlld.LAST_DIAG
                    out[f"{pre}_diag"] = np.array(
                        [d["n_not_converged"], d["n_wave_order_bad"], d["n_bn_zero"],
                         d["n_hlle_fallback"], d["max_iters"]], dtype=np.int64)
                    out[f"{pre}_mean_iters"] = np.array(d["mean_iters"])
    finally:
        R_hlld._PSTAR_METHOD, R_hlld._PSTAR_BRACKET, R_hlld._PSTAR_TOL = saved
    np.savez_compressed(HERE / "pstar_methods.npz", **out)


LINEAR_SHRINK = (1.0, 0.1, 0.01)


def shrink_pair(sL, sR, s, eos_eps):
    """The jump of every primitive (rho, p, v, B) scaled by ``s`` about the
    pair's mean; eps from the EOS."""
    L2, R2 = {}, {}
    for k in ("rho", "p", "vx", "vy", "vz", "Bx", "By", "Bz"):
        m = 0.5 * (sL[k] + sR[k])
        L2[k] = m + s * (sL[k] - m)
        R2[k] = m + s * (sR[k] - m)
    L2["eps"] = eos_eps(L2["p"], L2["rho"])
    R2["eps"] = eos_eps(R2["p"], R2["rho"])
    return L2, R2


def gen_linear():
    """linearised_flux (src/physics/linearised.py:129-237, HLLD fallback,
    default fd_rel / imag_tol / cond_max) on the flux_random interface pairs
    with their jumps scaled by 1, 0.1 and 0.01 about the mean (the weak
    problems it is meant for), every direction: F, U, p* and the mask of the
    faces it answered."""
    from src.physics.linearised import LAST_LINEAR, linearised_flux
    g = np.load(HERE / "flux_random.npz")
    eos = hybrid_eos(K=0.0, gamma=5.0 / 3.0, gamma_th=5.0 / 3.0)
    out = {}
    for tag in ("a", "b"):
        sL = {k: torch.from_numpy(g[f"{tag}_L_{k}"].copy()) for k in PRIM9}
        sR = {k: torch.from_numpy(g[f"{tag}_R_{k}"].copy()) for k in PRIM9}
        for si, sc in enumerate(LINEAR_SHRINK):
            L2, R2 = shrink_pair(sL, sR, sc, eos.eps__press_rho)
            for k in PRIM9:
                out[f"{tag}{si}_L_{k}"] = _np(L2[k])
                out[f"{tag}{si}_R_{k}"] = _np(R2[k])
            for idir in (0, 1, 2):
                F, U, ps, ok = linearised_flux(L2, R2, eos, idir=idir, return_mask=True)
                pre = f"{tag}{si}_{idir}"
                for k in CONS8:
                    out[f"{pre}_F_{k}"] = _np(F[k])
                    out[f"{pre}_U_{k}"] = _np(U[k])
                out[f"{pre}_pstar"] = _np(ps)
                out[f"{pre}_ok"] = _np(ok)
                out[f"{pre}_n_linear"] = np.array(LAST_LINEAR["n_linear"], dtype=np.int64)
    np.savez_compressed(HERE / "linear_flux.npz", **out)


def gen_c2p():
    eos = hybrid_eos(K=0.0, gamma=5.0 / 3.0, gamma_th=5.0 / 3.0)
    rng = np.random.default_rng(7)
    parts = [random_prims(96, eos, rng, max_W=2.0), random_prims(96, eos, rng, max_W=10.0),
             random_prims(32, eos, rng, max_W=2.0, bscale=0.0)]
    hot = {"rho": torch.full((16,), 0.1), "p": torch.full((16,), 1.0),
           "vx": torch.full((16,), 0.99), "vy": torch.zeros(16), "vz": torch.zeros(16),
           "Bx": torch.from_numpy(rng.uniform(-0.5, 0.5, 16)), "By": torch.zeros(16),
           "Bz": torch.zeros(16)}
    hot = {k: v.double() for k, v in hot.items()}
    hot["eps"] = eos.eps__press_rho(hot["p"], hot["rho"])
    parts.append(hot)
    P = {k: torch.cat([q[k] for q in parts]) for k in PRIM9}
    cons = R_hlld.primitive_to_conserved(P)
    cons = {k: v.clone() for k, v in cons.items()}
    # a few cells pushed off the admissible set (|S| > D + tau, tau < 0)
    cons["Sx"][:4] = cons["Sx"][:4] * 3.0
    cons["tau"][4:8] = -1e-3
    out = {f"in_{k}": _np(v) for k, v in cons.items()}
    prims, status = conservative_to_primitive({k: v.clone() for k, v in cons.items()}, eos,
                                              return_status=True)
    for k in PRIM9:
        out[f"out_{k}"] = _np(prims[k])
    for k, v in status.items():
        out[f"st_{k}"] = _np(v)
    # the hot batch alone: a batch where the secant leaves cells unconverged
    sub = {k: v[-16:].clone() for k, v in cons.items()}
    p2, s2 = conservative_to_primitive(sub, eos, return_status=True)
    for k in PRIM9:
        out[f"hot_{k}"] = _np(p2[k])
    out["hot_converged"] = _np(s2["converged"])
    np.savez_compressed(HERE / "c2p_batch.npz", **out)


def gen_recon():
    eos = hybrid_eos(K=0.0, gamma=5.0 / 3.0, gamma_th=5.0 / 3.0)
    rng = np.random.default_rng(11)
    shp = (13, 11)
    P = {"rho": 10.0 ** rng.uniform(-1, 1, shp), "p": 10.0 ** rng.uniform(-1, 1, shp),
         "vx": rng.uniform(-0.6, 0.6, shp), "vy": rng.uniform(-0.6, 0.6, shp),
         "vz": rng.uniform(-0.3, 0.3, shp), "Bx": rng.uniform(-1, 1, shp),
         "By": rng.uniform(-1, 1, shp), "Bz": rng.uniform(-1, 1, shp)}
    P["rho"][6, 4] = 1e-13          # forces the PCM fallback around it
    P = {k: torch.from_numpy(v.copy()) for k, v in P.items()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    out = {f"in_{k}": _np(v) for k, v in P.items()}
    for lim in ("pcm", "mc", "minmod", "weno5z", "mp5", "mp7"):
        for axis in (0, 1):
            L, Rr = reconstruct_prims(P, axis=axis, eos=eos, limiter=lim)
            for k in PRIM9:
                out[f"{lim}{axis}_L_{k}"] = _np(L[k])
                out[f"{lim}{axis}_R_{k}"] = _np(Rr[k])
    np.savez_compressed(HERE / "recon_block.npz", **out)


def _tapered(g, eos):
    X, Y = g.X, g.Y
    r = torch.sqrt(X * X + Y * Y)
    f = torch.where(r <= 0.1, torch.ones_like(r),
                    torch.where(r < 0.115, (0.115 - r) / 0.015, torch.zeros_like(r)))
    z = torch.zeros(g.shape, dtype=torch.float64)
    om = 9.95 * f
    P = {"rho": 1.0 + 9.0 * f, "p": torch.full_like(z, 1.0), "vx": -om * Y, "vy": om * X,
         "vz": z.clone(), "Bx": torch.full_like(z, 1.0), "By": z.clone(), "Bz": z.clone()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    YF = g.yf.unsqueeze(0).expand(g.nxt + 1, g.nyt + 1)
    return P, 1.0 * YF


def _ot2pi(g, eos, A=1.0):
    X, Y = g.X, g.Y
    amp = 0.99 / np.sqrt(2.0)
    z = torch.zeros(g.shape, dtype=torch.float64)
    P = {"rho": torch.full_like(z, 1.0), "p": torch.full_like(z, 1.0),
         "vx": -amp * torch.sin(Y), "vy": amp * torch.sin(X), "vz": z.clone(),
         "Bx": -A * torch.sin(Y), "By": A * torch.sin(2.0 * X), "Bz": z.clone()}
    P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
    XF = g.xf.unsqueeze(1).expand(g.nxt + 1, g.nyt + 1)
    YF = g.yf.unsqueeze(0).expand(g.nxt + 1, g.nyt + 1)
    return P, A * (torch.cos(YF) + 0.5 * torch.cos(2.0 * XF))


def _state_arrays(prefix, st, out):
    for k in EVOLVED_KEYS:
        out[f"{prefix}cons_{k}"] = _np(st.cons[k])
    out[f"{prefix}Bxf"] = _np(st.Bxf)
    out[f"{prefix}Byf"] = _np(st.Byf)
    for k in PRIM9:
        out[f"{prefix}prim_{k}"] = _np(st.prims[k])


def _run2d(name, problem, n, limiter, solver, scheme, steps, cfl, gamma, box, bc, first=False):
    eos = hybrid_eos(K=0.0, gamma=gamma, gamma_th=gamma)
    g = Grid2D(box[0], box[1], n, box[0], box[1], n, ng=ghosts_needed(limiter))
    if problem == "rotor":
        prims, Az = magnetic_rotor(g, eos)
    elif problem == "tapered":
        prims, Az = _tapered(g, eos)
    else:
        prims, Az = _ot2pi(g, eos)
    Bxf, Byf = b_from_potential(Az, g.dx, g.dy)
    c = prims_to_cons_2d(prims, g)
    st = State2D(cons={k: c[k] for k in EVOLVED_KEYS}, Bxf=Bxf, Byf=Byf, prims=prims)
    st = sync_state(st, g, eos, *bc)
    out = {}
    _state_arrays("init_", st, out)
    from src.physics.linearised import linearised_flux
    flux = {"hlld": R_hlld.hlld_flux, "hllc": R_hlld.hllc_flux, "hlle": R_hlld.hlle_flux,
            "linear": linearised_flux}[solver]
    dts = []
    for i in range(steps):
        dt = compute_dt_2d(st.prims, g, eos, cfl)
        dts.append(dt)
        st, _ = rk_step_ct(st, g, eos, dt, scheme=scheme, bc_x=bc[0], bc_y=bc[1],
                           flux_fn=flux, limiter=limiter)
        if first and i == 0:
            _state_arrays("s1_", st, out)
    out["dts"] = np.array(dts)
    _state_arrays("fin_", st, out)
    np.savez_compressed(HERE / f"{name}.npz", **out)


def gen_2d():
    r = (-0.5, 0.5)
    oo = ("outflow", "outflow")
    _run2d("rotor_sharp_mc_hlld", "rotor", 24, "mc", "hlld", "rk3", 2, 0.4, 5.0 / 3.0, r, oo)
    _run2d("rotor_tapered_weno5z_hlld", "tapered", 20, "weno5z", "hlld", "rk3", 1, 0.4,
           5.0 / 3.0, r, oo)
    _run2d("rotor_tapered_mp5_hllc", "tapered", 16, "mp5", "hllc", "rk3", 1, 0.4, 5.0 / 3.0, r, oo)
    _run2d("rotor_sharp_mc_hlle", "rotor", 16, "mc", "hlle", "rk2", 1, 0.4, 5.0 / 3.0, r, oo)
    _run2d("ot2pi_mc_hlld", "ot", 16, "mc", "hlld", "rk2", 2, 0.4, 4.0 / 3.0,
           (0.0, 2.0 * np.pi), ("periodic", "periodic"))


def gen_2d_linear():
    """rk_step_ct with flux_fn=linearised_flux (run_2d.py --solver linear):
    the tapered rotor, MC and WENO5-Z, two RK3 steps."""
    r = (-0.5, 0.5)
    oo = ("outflow", "outflow")
    _run2d("rotor_tapered_mc_linear", "tapered", 24, "mc", "linear", "rk3", 2, 0.4, 5.0 / 3.0,
           r, oo, first=True)
    _run2d("rotor_tapered_weno5z_linear", "tapered", 20, "weno5z", "linear", "rk3", 1, 0.4,
           5.0 / 3.0, r, oo)
    # no face of the 2 pi Orszag-Tang problem has a degenerate spectrum
    _run2d("ot2pi_mc_linear", "ot", 24, "mc", "linear", "rk2", 3, 0.4, 4.0 / 3.0,
           (0.0, 2.0 * np.pi), ("periodic", "periodic"))


def gen_2d_long():
    """100 steps at 32^2: the north star's multi-step criterion (L1 in rho
    after 100 steps) needs the reference's own trajectory, not one step."""
    r = (-0.5, 0.5)
    _run2d("rotor_tapered_weno5z_hlld_100", "tapered", 32, "weno5z", "hlld", "rk3", 100, 0.4,
           5.0 / 3.0, r, ("outflow", "outflow"), first=True)
    _run2d("ot2pi_mc_hlld_100", "ot", 32, "mc", "hlld", "rk2", 100, 0.4, 4.0 / 3.0,
           (0.0, 2.0 * np.pi), ("periodic", "periodic"), first=True)


def gen_dt1d():
    """driver.compute_dt on random 1D primitives, with the fast-speed bound
    (speed_of_light=False) and the legacy c = 1 form; the 1D driver's own
    runs only ever take the latter."""
    from src.physics.driver import compute_dt
    out = {}
    for tag, gamma, seed, vmax, bsc in (("a", 5.0 / 3.0, 3, 0.6, 1.0), ("b", 4.0 / 3.0, 4, 0.95, 4.0),
                                        ("c", 2.0, 5, 0.3, 0.0)):
        eos = hybrid_eos(K=0.0, gamma=gamma, gamma_th=gamma)
        rng = np.random.default_rng(seed)
        n, ng = 61, 3
        nt = n + 2 * ng
        v = rng.uniform(-1.0, 1.0, (3, nt))
        v = v / np.linalg.norm(v, axis=0) * rng.uniform(0.0, vmax, nt)
        P = {"rho": 10.0 ** rng.uniform(-2, 1, nt), "p": 10.0 ** rng.uniform(-2, 1, nt),
             "vx": v[0], "vy": v[1], "vz": v[2], "Bx": bsc * rng.uniform(-1, 1, nt),
             "By": bsc * rng.uniform(-1, 1, nt), "Bz": bsc * rng.uniform(-1, 1, nt)}
        P = {k: torch.from_numpy(x.copy()) for k, x in P.items()}
        P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
        for k in PRIM9:
            out[f"{tag}_{k}"] = _np(P[k])
        out[f"{tag}_gamma"] = np.array(gamma)
        dx = 1.0 / n
        for sol in (False, True):
            out[f"{tag}_dt_{int(sol)}"] = _np(compute_dt(P, eos, dx, 0.4, ng, n,
                                                         speed_of_light=sol))
    np.savez_compressed(HERE / "dt1d.npz", **out)


def gen_tubes_pertube():
    """The per-tube-dt SSP-RK form of the batched tubes, assembled from the
    reference's own pieces: tube_seed's _rhs/_bc and con2prim sync
    (tube_seed.py:72-90,111-117), driver.compute_dt with the fast-speed
    bound (driver.py:77-127), the RK weights of driver2d._RK2/_RK3 applied by
    state.combine (state.py:77-103).  Tubes drawn as the batched workload
    draws them (seeded CPU generator), shortened so the CPU run is quick."""
    from src.physics.driver import compute_dt
    from src.physics.driver2d import _RK2, _RK3
    from src.physics.state import combine
    from src.physics.tube_seed import _bc, _rhs
    eos = hybrid_eos(K=0.0, gamma=5.0 / 3.0, gamma_th=5.0 / 3.0)
    keys = ("D", "Sx", "Sy", "Sz", "tau")
    g = torch.Generator().manual_seed(0)
    nt_ = 3

    def u(lo, hi):
        return lo + (hi - lo) * torch.rand((nt_,), generator=g, dtype=torch.float64)

    lengths = torch.randint(24, 65, (nt_,), generator=g)
    sides = []
    for _ in range(2):
        rho = 10.0 ** u(-1.0, 1.0)
        p = 10.0 ** u(-2.0, 1.0)
        ct = u(-1.0, 1.0)
        ph = u(0.0, 2.0 * np.pi)
        sp = u(0.0, 0.9)
        st = torch.sqrt(1.0 - ct * ct)
        sides.append({"rho": rho, "p": p, "vx": sp * st * torch.cos(ph),
                      "vy": sp * st * torch.sin(ph), "vz": sp * ct,
                      "By": u(-2.0, 2.0), "Bz": u(-2.0, 2.0)})
    Bn = u(-1.0, 1.0)
    out = {"lengths": _np(lengths), "Bn": _np(Bn)}
    for s, side in zip(("L", "R"), sides):
        for k, v in side.items():
            out[f"{s}_{k}"] = _np(v)
    for lim, flux, scheme, steps, ntubes in (("mp5", "hlld", "rk3", 20, nt_),
                                             ("mc", "hllc", "rk2", 15, 2)):
        ng = ghosts_needed(lim)
        rk = {"rk2": _RK2, "rk3": _RK3}[scheme]
        for i in range(ntubes):
            n = int(lengths[i])
            dx = 1.0 / n
            x = (torch.arange(n + 2 * ng, dtype=torch.float64) - ng + 0.5) * dx
            m = x < 0.5
            P = {k: torch.where(m, torch.full_like(x, float(sides[0][k][i])),
                                torch.full_like(x, float(sides[1][k][i])))
                 for k in ("rho", "p", "vx", "vy", "vz", "By", "Bz")}
            P["Bx"] = torch.full_like(x, float(Bn[i]))
            P["eps"] = eos.eps__press_rho(P["p"], P["rho"])
            c = R_hlld.primitive_to_conserved(P)
            cons = {k: c[k].clone() for k in keys}
            cons["By"], cons["Bz"] = P["By"].clone(), P["Bz"].clone()
            Bx = P["Bx"]

            def sync(cc):
                full = {k: cc[k] for k in keys}
                full["Bx"], full["By"], full["Bz"] = Bx, cc["By"], cc["Bz"]
                p = conservative_to_primitive(full, eos)
                p["Bx"], p["By"], p["Bz"] = Bx, cc["By"], cc["Bz"]
                _bc(p, ng)
                return p

            z = torch.zeros(1, dtype=torch.float64)
            t, dts = 0.0, []
            for _ in range(steps):
                dt = float(compute_dt(P, eos, dx, 0.4, ng, n, speed_of_light=False))
                dts.append(dt)
                S0 = State2D(cons={k: v for k, v in cons.items()}, Bxf=z, Byf=z)
                S = S0
                for a, b, cw in rk:
                    L = _rhs(P, eos, dx, ng, lim, flux)
                    S = combine([S0, S, State2D(cons=L, Bxf=z, Byf=z)], [a, b, cw * dt])
                    P = sync(S.cons)
                cons = S.cons
                t += dt
            pre = f"{lim}_{flux}_{scheme}_{i}"
            out[f"{pre}_dts"] = np.array(dts)
            for k in PRIM9:
                out[f"{pre}_{k}"] = _np(P[k])
        out[f"{lim}_{flux}_{scheme}_steps"] = np.array(steps)
        out[f"{lim}_{flux}_{scheme}_ntubes"] = np.array(ntubes)
    np.savez_compressed(HERE / "tubes_pertube.npz", **out)


def gen_run2d():
    """scripts/run_2d.py itself on a 32^2 sharp rotor (diag.csv every 2 steps,
    the final snapshot), and its non-finite guard at CFL 40
    (tests/test_high_order_2d.py:196-215)."""
    import subprocess
    import tempfile
    out = {}
    base = ["--problem", "rotor", "--n", "32", "--tend", "0.02", "--cfl", "0.4",
            "--log-every", "2", "--nsnap", "2", "--solver", "hlld"]
    env = dict(os.environ, TORCH_THREADS="4")
    for lim in ("mc", "weno5z"):
        with tempfile.TemporaryDirectory() as d:
            r = subprocess.run([sys.executable, "scripts/run_2d.py", *base, "--limiter", lim,
                                "--out", d], cwd=REF, capture_output=True, text=True,
                               timeout=1800, env=env)
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            out[f"{lim}_diag"] = np.array(open(os.path.join(d, "diag.csv")).read().splitlines())
            f = np.load(os.path.join(d, "snap_fin.npz"))
            for k in ("t", "rho", "p", "vx", "vy", "Bx", "By", "divB"):
                out[f"{lim}_fin_{k}"] = f[k]
    with tempfile.TemporaryDirectory() as d:
        boom = os.path.join(d, "boom")
        r = subprocess.run([sys.executable, "scripts/run_2d.py", "--problem", "rotor", "--n", "32",
                            "--solver", "hlld", "--limiter", "mc", "--cfl", "40.0", "--nsnap",
                            "2", "--out", boom], cwd=REF, capture_output=True, text=True,
                           timeout=1800, env=env)
        msg = (r.stdout + r.stderr)
        line = [ln for ln in msg.splitlines() if "run failed" in ln]
        out["cfl40_returncode"] = np.array(r.returncode)
        out["cfl40_message"] = np.array(line[-1] if line else "")
        out["cfl40_snap_fin"] = np.array(os.path.exists(os.path.join(boom, "snap_fin.npz")))
    np.savez_compressed(HERE / "run2d_rotor32.npz", **out)


def gen_1d():
    cfg = {"grid": {"xmin": 0.0, "xmax": 1.0, "ncells": 400, "ng": 2},
           "eos": {"K": 0.0, "gamma": 2.0, "gamma_th": 2.0},
           "run": {"problem": "generic", "t_end": 0.4, "cfl": 0.4, "atmo_rho": 1.0e-10,
                   "solver": "hlld", "limiter": "mc", "device": "cpu",
                   "primL": {"rho": 1.0, "p": 1.0, "vx": 0.0, "vy": 0.0, "vz": 0.0,
                             "Bx": 0.5, "By": 1.0, "Bz": 0.0},
                   "primR": {"rho": 0.125, "p": 0.1, "vx": 0.0, "vy": 0.0, "vz": 0.0,
                             "Bx": 0.5, "By": -1.0, "Bz": 0.0},
                   "x_interface": 0.5},
           "output": {"dir": str(HERE / "_tmp_out"), "every_n_steps": 100000}}
    out = {}
    for solver, t_end in (("hlld", 0.4), ("hllc", 0.05), ("hlle", 0.05)):
        cfg["run"]["solver"] = solver
        cfg["run"]["t_end"] = t_end
        t, grid, prims, cons = R_run(cfg)
        out[f"{solver}_t"] = np.array(t)
        for k in PRIM9:
            out[f"{solver}_prim_{k}"] = _np(prims[k])
        for k in ("D", "Sx", "Sy", "Sz", "tau"):
            out[f"{solver}_cons_{k}"] = _np(cons[k])
    np.savez_compressed(HERE / "balsara1.npz", **out)
    import shutil
    shutil.rmtree(HERE / "_tmp_out", ignore_errors=True)


def gen_1d_linear():
    """driver.run on Balsara 1 with solver "linear" (driver.py:378-380), MC,
    to t = 0.02."""
    cfg = {"grid": {"xmin": 0.0, "xmax": 1.0, "ncells": 400, "ng": 2},
           "eos": {"K": 0.0, "gamma": 2.0, "gamma_th": 2.0},
           "run": {"problem": "generic", "t_end": 0.02, "cfl": 0.4, "atmo_rho": 1.0e-10,
                   "solver": "linear", "limiter": "mc", "device": "cpu",
                   "primL": {"rho": 1.0, "p": 1.0, "vx": 0.0, "vy": 0.0, "vz": 0.0,
                             "Bx": 0.5, "By": 1.0, "Bz": 0.0},
                   "primR": {"rho": 0.125, "p": 0.1, "vx": 0.0, "vy": 0.0, "vz": 0.0,
                             "Bx": 0.5, "By": -1.0, "Bz": 0.0},
                   "x_interface": 0.5},
           "output": {"dir": str(HERE / "_tmp_out"), "every_n_steps": 100000}}
    out = {}
    t, grid, prims, cons = R_run(cfg)
    out["t"] = np.array(t)
    for k in PRIM9:
        out[f"prim_{k}"] = _np(prims[k])
    # a non-coplanar drifting tube: no face's mean state has B_t = 0 or a
    # contact at rest, so no spectrum is degenerate
    cfg["run"]["primL"] = {"rho": 1.0, "p": 1.0, "vx": 0.1, "vy": 0.05, "vz": 0.0,
                           "Bx": 0.5, "By": 1.0, "Bz": 0.4}
    cfg["run"]["primR"] = {"rho": 0.125, "p": 0.1, "vx": 0.1, "vy": 0.0, "vz": 0.05,
                           "Bx": 0.5, "By": -1.0, "Bz": 0.2}
    t, grid, prims, cons = R_run(cfg)
    out["nc_t"] = np.array(t)
    for k in PRIM9:
        out[f"nc_prim_{k}"] = _np(prims[k])
    np.savez_compressed(HERE / "balsara1_linear.npz", **out)
    import shutil
    shutil.rmtree(HERE / "_tmp_out", ignore_errors=True)


def gen_tubes():
    eos = hybrid_eos(K=0.0, gamma=5.0 / 3.0, gamma_th=5.0 / 3.0)
    BW_L, BW_R = [1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0], [0.125, 0.6, 0.0, 0.0, 0.0, -1.0, 0.0]
    QUIET = [1.0, 0.02, 0.0, 0.0, 0.0, 0.0, 0.0]
    HOT_L, HOT_R = [1.0, 30.0, 0.5, 0.0, 0.0, 2.0, 0.0], [0.1, 0.1, -0.5, 0.0, 0.0, -2.0, 0.0]
    UL = np.array([BW_L, QUIET, HOT_L])
    UR = np.array([BW_R, QUIET, HOT_R])
    Bn = np.array([0.5, 0.3, 0.5])
    out = {"UL": UL, "UR": UR, "Bn": Bn}
    for lim in ("mc", "mp5"):
        prof, t = run_tubes(UL, UR, Bn, eos, ncells=48, tend=0.04, limiter=lim)
        out[f"{lim}_t"] = np.array(t)
        for k, v in prof.items():
            out[f"{lim}_{k}"] = _np(v)
    np.savez_compressed(HERE / "tubes_shared.npz", **out)


def gen_init2d():
    """The reference's remaining 2D initial data (initial_data2d.py:128-270):
    unit-square Orszag-Tang, Beckwith & Stone's variant and the
    four-quadrant Riemann problem (coplanar 45 degrees and tilted), on a 24^2
    grid with two ghosts: cell primitives and the corner potential."""
    from src.physics.initial_data2d import orszag_tang, orszag_tang_bs, riemann2d
    out = {}
    cases = {"ot": (orszag_tang, (0.0, 1.0), 4.0 / 3.0, {}),
             "otbs": (orszag_tang_bs, (0.0, 1.0), 5.0 / 3.0, {}),
             "r2d": (riemann2d, (-0.5, 0.5), 5.0 / 3.0, {}),
             "r2dtilt": (riemann2d, (-0.5, 0.5), 5.0 / 3.0, {"B0": 0.7, "angle_z_deg": 30.0})}
    for name, (fn, box, gamma, kw) in cases.items():
        eos = hybrid_eos(K=0.0, gamma=gamma, gamma_th=gamma)
        g = Grid2D(box[0], box[1], 24, box[0], box[1], 24, ng=2)
        P, Az = fn(g, eos, **kw)
        for k, v in P.items():
            out[f"{name}_{k}"] = _np(v)
        out[f"{name}_Az"] = _np(Az)
    np.savez_compressed(HERE / "init2d.npz", **out)


def gen_host():
    """torch.sqrt of fixed inputs on the host that wrote the fixtures: torch's
    CPU float64 sqrt is not correctly rounded on every host, so bitwise
    comparisons against these fixtures only hold where it matches."""
    x = np.random.default_rng(5).uniform(0.0, 4.0, 8192)
    np.savez_compressed(HERE / "host_sqrt.npz", x=x, torch_sqrt=_np(torch.sqrt(torch.from_numpy(x))))


if __name__ == "__main__":
    torch.set_num_threads(4)
    which = sys.argv[1:] or ["flux", "c2p", "recon", "2d", "1d", "tubes", "host", "init2d",
                             "2d_long", "dt1d", "tubes_pertube", "run2d", "pstar", "linear", "2d_linear", "1d_linear"]
    for w in which:
        print("generating", w, flush=True)
        globals()[f"gen_{w}"]()
