"""Write the oracle's own 100-step trajectories of the two 32^2 problems of
the reference's 100-step fixtures, in the oracle's IEEE mode (correctly
rounded sqrt, the arithmetic of the HIP path):

    python tests/golden/make_oracle_long.py

The reference wrote rotor_tapered_weno5z_hlld_100.npz and
ot2pi_mc_hlld_100.npz with torch's CPU float64 sqrt, which is one ulp off
the correctly rounded value on ~0.8 % of inputs; the oracle in the
reference's sqrt mode reproduces both bit for bit
(tests/test_oracle_golden.py).  These files hold the same runs with IEEE
sqrt, so the device can be compared bit for bit over all 100 steps; they are
the oracle's output, not the reference's (a 100-step oracle run takes
minutes on the CPU).  Per step: dt, and the final state.  Likewise the
per-tube-dt tube runs of tubes_pertube.npz (oracle_ieee_tubes_pertube.npz),
and three steps of the 2 pi Orszag-Tang problem with the linearised flux
(oracle_ieee_ot2pi_mc_linear.npz).
"""

import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import eos as E, physics as PH, riemann as R, step2d as S2  # noqa: E402

PRIM9 = ("rho", "vx", "vy", "vz", "p", "eps", "Bx", "By", "Bz")
CASES = [
    ("rotor_tapered_weno5z_hlld_100", "tapered", 32, "weno5z", "hlld", "rk3", 5 / 3,
     (-0.5, 0.5), "outflow"),
    ("ot2pi_mc_hlld_100", "ot", 32, "mc", "hlld", "rk2", 4 / 3, (0.0, 2 * np.pi), "periodic"),
    # --solver linear: its eigen-decompositions go through the host's LAPACK,
    # whose results differ between hosts, so the device compares with this
    # file rather than with the oracle run on its own host (3 steps)
    ("ot2pi_mc_linear", "ot", 24, "mc", "linear", "rk2", 4 / 3, (0.0, 2 * np.pi), "periodic"),
]
STEPS = {"ot2pi_mc_linear": 3}


def _flux(solver):
    if solver == "linear":
        from oracle import linear as OL
        return OL.linearised_flux
    return R.SOLVERS[solver]


def run(case, steps=100):
    name, problem, n, lim, solver, scheme, gamma, box, bc = case
    eos = E.ideal(gamma)
    grid = S2.make_grid(n, box, lim)
    init = {"tapered": S2.rotor_tapered, "ot": S2.orszag_tang_2pi}[problem]
    P, Az = init(grid, eos)
    st = S2.init_state(grid, eos, P, Az, (bc, bc))
    dts = []
    for _ in range(steps):
        dt = S2.compute_dt_2d(st.prims, grid, eos, 0.4)
        dts.append(dt)
        st, _ = S2.rk_step_ct(st, grid, eos, dt, scheme=scheme, bc_x=bc, bc_y=bc,
                              flux_fn=_flux(solver), limiter=lim)
    out = {"dts": np.array(dts)}
    for k in S2.EVOLVED_KEYS:
        out[f"fin_cons_{k}"] = st.cons[k].numpy()
    out["fin_Bxf"] = st.Bxf.numpy()
    out["fin_Byf"] = st.Byf.numpy()
    for k in PRIM9:
        out[f"fin_prim_{k}"] = st.prims[k].numpy()
    np.savez_compressed(os.path.join(HERE, f"oracle_ieee_{name}.npz"), **out)


def run_tubes():
    """The per-tube-dt tube runs of tubes_pertube.npz (the reference's
    composition) by the oracle in IEEE mode: oracle_ieee_tubes_pertube.npz."""
    from oracle import recon as RC, tubes as T
    g = np.load(os.path.join(HERE, "tubes_pertube.npz"))
    eos = E.ideal(5.0 / 3.0)
    out = {}
    for lim, flux, scheme in (("mp5", "hlld", "rk3"), ("mc", "hllc", "rk2")):
        pre = f"{lim}_{flux}_{scheme}"
        steps, ng = int(g[f"{pre}_steps"]), RC.ghosts_needed(lim)
        for i in range(int(g[f"{pre}_ntubes"])):
            n = int(g["lengths"][i])
            keys = ("rho", "p", "vx", "vy", "vz", "By", "Bz")
            left = {k: float(g[f"L_{k}"][i]) for k in keys}
            right = {k: float(g[f"R_{k}"][i]) for k in keys}
            P0 = T.tube_initial(n, ng, eos, left, right, float(g["Bn"][i]))
            P, t = T.run_tube(P0, eos, n, ng, steps, cfl=0.4, limiter=lim, scheme=scheme,
                              flux=flux)
            out[f"{pre}_{i}_t"] = np.array(t)
            for k in PRIM9:
                out[f"{pre}_{i}_{k}"] = P[k].numpy()
    np.savez_compressed(os.path.join(HERE, "oracle_ieee_tubes_pertube.npz"), **out)


if __name__ == "__main__":
    PH.set_sqrt_mode("ieee")
    torch.set_num_threads(4)
    for c in CASES:
        if not sys.argv[1:] or c[0] in sys.argv[1:]:
            print("oracle (ieee)", c[0], flush=True)
            run(c, STEPS.get(c[0], 100))
    if not sys.argv[1:] or "tubes" in sys.argv[1:]:
        print("oracle (ieee) tubes_pertube", flush=True)
        run_tubes()
