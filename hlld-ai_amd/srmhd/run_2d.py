"""scripts/run_2d.py on the HIP path: a 2D SRMHD run with the reference's
command line, output files and failure behaviour.

    python -m srmhd.run_2d --problem rotor --n 512 --limiter weno5z --cfl 0.4

Files written into ``--out`` (the reference's schema, run_2d.py:170-334):
  run_meta.json   what made the run (rewritten at the end with the cost);
  diag.csv        every --log-every steps: step, t, dt, divB_max, divB_l2,
                  sym_err (rotor), rho_max, rho_min, p_min, W_max,
                  hlle_frac_x, hlle_frac_y, mean_iters, c2p_bad, n_exact_x,
                  n_verified_x (-1: no exact solver here);
  snap_NNN.npz    physical-region rho, p, vx, vy, Bx, By, divB at t = k
                  tend / nsnap, and snap_fin.npz at the end;
  restart.npz     the complete solver state (cell cons and staggered
                  field, ghosts included) every --restart-every steps; a run
                  resumed from it continues with the same arithmetic (the
                  primitives are rebuilt by sync_state, as after every step),
                  so it is bit-identical to the uninterrupted run.

The step itself is ``driver2d.Stepper2D`` (one C-ABI call per RK stage, the
state resident on the device).  ``--solver linear`` is the linearised flux
(srmhd.linearised) over the HLLD fallback, in the FP64 form.  The exact
solver of the reference stays on its CPU path, and its interface recorders
are not offered here.
"""

import argparse
import json
import math
import os
import socket
import sys
import tempfile
import time

import numpy as np
import torch

from . import _lib
from . import driver2d as D2
from . import eos as E
from . import grid as G
from . import initial_data2d as I
from .reconstruction import ghosts_needed
from .state import EVOLVED_KEYS, State2D

# problem -> domain, gamma, default t_end, BCs, pi-rotation symmetry
# (run_2d.py:30-44; rotor_tapered and orszag_tang_2pi are this package's)
PROBLEMS = {
    "rotor": dict(box=(-0.5, 0.5), gamma=5.0 / 3.0, tend=0.4, bc=("outflow", "outflow"), sym=True),
    "orszag_tang": dict(box=(0.0, 1.0), gamma=4.0 / 3.0, tend=1.0, bc=("periodic", "periodic"),
                        sym=False),
    "orszag_tang_bs": dict(box=(0.0, 1.0), gamma=5.0 / 3.0, tend=1.0,
                           bc=("periodic", "periodic"), sym=False),
    "riemann2d": dict(box=(-0.5, 0.5), gamma=5.0 / 3.0, tend=0.4, bc=("outflow", "outflow"),
                      sym=False),
    "rotor_tapered": dict(box=(-0.5, 0.5), gamma=5.0 / 3.0, tend=0.4,
                          bc=("outflow", "outflow"), sym=True),
    "orszag_tang_2pi": dict(box=(0.0, 2.0 * math.pi), gamma=4.0 / 3.0, tend=1.0,
                            bc=("periodic", "periodic"), sym=False),
}
SOLVERS = ("hlld", "hlle", "hllc", "linear")
DIAG_HEADER = ("step,t,dt,divB_max,divB_l2,sym_err,rho_max,rho_min,p_min,W_max,hlle_frac_x,"
               "hlle_frac_y,mean_iters,c2p_bad,n_exact_x,n_verified_x\n")


def div_b(Bxf, Byf, dx, dy):
    """ct.py:196-199."""
    return (Bxf[1:, :] - Bxf[:-1, :]) / dx + (Byf[:, 1:] - Byf[:, :-1]) / dy


def symmetry_error(prims, g):
    """Rotor invariance under a pi-rotation with B -> -B (run_2d.py:65-87)."""
    ph = g.phys
    worst = 0.0
    for k, parity in (("rho", +1), ("p", +1), ("vx", -1), ("vy", -1), ("Bx", +1), ("By", +1)):
        a = prims[k][ph]
        b = torch.flip(a, dims=(0, 1))
        worst = max(worst, float((a - parity * b).abs().max()) / max(float(a.abs().max()), 1e-30))
    return worst


class RunDir:
    """The bookkeeping of one run directory: run_meta.json (each update
    rewrites the whole file through a temporary one, so a reader never sees
    a partial file), the check that the directory does not already hold a
    different run, and when snapshots and restart files fall due."""

    # the keys that make two runs different runs (run_2d.py:240-261)
    IDENTITY = ("limiter", "n", "solver", "problem")

    def __init__(self, out, meta, tend, nsnap, restart_every):
        self.out, self.meta = out, meta
        self.path = os.path.join(out, "run_meta.json")
        # snapshot k is due once t reaches k tend / nsnap (to 1e-14)
        self.snap_times = [k * tend / nsnap for k in range(1, nsnap + 1)]
        self.restart_every = restart_every

    def refuse_other_run(self):
        if not os.path.exists(self.path):
            return
        with open(self.path) as fh:
            held = json.load(fh)
        diff = [f"{k} {held.get(k)!r} != {self.meta[k]!r}" for k in self.IDENTITY
                if held.get(k) != self.meta[k]]
        if diff:
            raise SystemExit(f"refusing to write into {self.out}: it holds a different run "
                             f"({', '.join(diff)}).  Pass --out, or --restart to continue "
                             f"that one.")

    def record(self, **fields):
        self.meta.update(fields)
        fd, tmp = tempfile.mkstemp(prefix=".run_meta.", suffix=".json", dir=self.out)
        with os.fdopen(fd, "w") as fh:
            json.dump(self.meta, fh, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def record_cost(self, step, steps_here, wall, **fields):
        self.record(steps=step, steps_here=steps_here, wall_s=round(wall, 1),
                    s_per_step=round(wall / max(steps_here, 1), 3),
                    finished=time.strftime("%Y-%m-%dT%H:%M:%S"), **fields)

    def snap_due(self, t, next_snap):
        """Snapshot number next_snap (1-based) is due at time t."""
        return next_snap <= len(self.snap_times) and t >= self.snap_times[next_snap - 1] - 1e-14

    def restart_due(self, step, done):
        return bool(self.restart_every) and not done and step % self.restart_every == 0


def _parser():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--problem", default="rotor", choices=list(PROBLEMS))
    ap.add_argument("--b0", type=float, default=0.5)
    ap.add_argument("--bz-deg", type=float, default=0.0)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--tend", type=float, default=None)
    ap.add_argument("--cfl", type=float, default=0.25)
    ap.add_argument("--solver", default="hlld", choices=SOLVERS)
    ap.add_argument("--emf-mode", default="solver", choices=["solver", "hll"])
    ap.add_argument("--limiter", default="mc")
    ap.add_argument("--no-upwind-emf", action="store_true")
    ap.add_argument("--precision", default="fp64", choices=tuple(_lib.PRECISION_IDS),
                    help="fp64 (the reference's arithmetic) or a mixed form")
    ap.add_argument("--out", default=None)
    ap.add_argument("--restart", default=None)
    ap.add_argument("--restart-every", type=int, default=10)
    ap.add_argument("--log-every", type=int, default=10)
    ap.add_argument("--max-steps", type=int, default=None)
    ap.add_argument("--nsnap", type=int, default=5)
    return ap


def main(argv=None):
    a = _parser().parse_args(argv)
    if not torch.cuda.is_available():
        raise SystemExit("srmhd.run_2d runs on a HIP device")
    dev = torch.device("cuda:0")
    P = PROBLEMS[a.problem]
    tend = a.tend if a.tend is not None else P["tend"]
    out = a.out or (f"results/{a.problem}_{a.n}_{a.solver}"
                    + ("" if a.limiter == "mc" else f"_{a.limiter}"))
    bc_x, bc_y = P["bc"]
    os.makedirs(out, exist_ok=True)
    meta = dict(problem=a.problem, b0=(a.b0 if a.problem == "riemann2d" else None),
                bz_deg=(a.bz_deg if a.problem == "riemann2d" else None), n=a.n,
                solver=a.solver, limiter=a.limiter, ng=ghosts_needed(a.limiter), cfl=a.cfl,
                tend=tend, nsnap=a.nsnap, emf_mode=a.emf_mode, upwind_emf=not a.no_upwind_emf,
                bc=[bc_x, bc_y], gamma=P["gamma"], host=socket.gethostname(),
                device=torch.cuda.get_device_name(dev), precision=a.precision,
                started=time.strftime("%Y-%m-%dT%H:%M:%S"))
    rd = RunDir(out, meta, tend, a.nsnap, a.restart_every)
    if not a.restart:
        rd.refuse_other_run()
    eos = E.hybrid_eos(K=0.0, gamma=P["gamma"], gamma_th=P["gamma"])
    lo, hi = P["box"]
    g = G.Grid2D(lo, hi, a.n, lo, hi, a.n, ng=ghosts_needed(a.limiter), device=dev)
    if a.problem == "riemann2d":
        prims, Az = I.riemann2d(g, eos, B0=a.b0, angle_z_deg=a.bz_deg)
    else:
        prims, Az = I.PROBLEMS[a.problem](g, eos)
    st = I.init_state(g, eos, prims, Az, bc_x=bc_x, bc_y=bc_y, limiter=a.limiter)

    gen = 0
    R = None
    if a.restart:
        R = np.load(a.restart)
        want = (g.nxt, g.nyt)
        got = tuple(R["cons_D"].shape)
        if "limiter" in R.files and str(R["limiter"]) != a.limiter:
            raise SystemExit("restart %s was written with --limiter %s, this run says %s"
                             % (a.restart, str(R["limiter"]), a.limiter))
        if got != want:
            raise SystemExit("restart %s holds %s arrays, this grid is %s (ng = %d for "
                             "--limiter %s): the limiter does not match"
                             % (a.restart, got, want, g.ng, a.limiter))
        st = State2D({k: torch.from_numpy(R["cons_" + k].copy()).to(dev) for k in EVOLVED_KEYS},
                     torch.from_numpy(R["Bxf"].copy()).to(dev),
                     torch.from_numpy(R["Byf"].copy()).to(dev))
        st = D2.sync_state(st, g, eos, bc_x, bc_y, limiter=a.limiter)
        gen = int(R["gen"]) + 1
    rd.record()

    S = D2.Stepper2D(g, eos, solver=a.solver, limiter=a.limiter, bc_x=bc_x, bc_y=bc_y,
                     upwind=not a.no_upwind_emf, precision=a.precision,
                     emf_mode=a.emf_mode)
    S.load(st.state_list(), st.prim_list())

    def current():
        """The state and the reference's primitive dict (B from the faces
        with the cell BCs, Bz the conserved Bz)."""
        cur = State2D({k: S.U[i] for i, k in enumerate(EVOLVED_KEYS)}, S.U[6], S.U[7])
        return D2._finish_prims(cur, list(S.P), g, bc_x, bc_y)

    def save(tag, t):
        cur = current()
        ph = g.phys
        np.savez_compressed(
            os.path.join(out, f"snap_{tag}.npz"), t=t, x=g.x[ph[0]].numpy(),
            y=g.y[ph[1]].numpy(),
            **{k: cur.prims[k][ph].cpu().numpy() for k in ("rho", "p", "vx", "vy", "Bx", "By")},
            divB=div_b(cur.Bxf, cur.Byf, g.dx, g.dy)[ph].cpu().numpy(),
            limiter=a.limiter, ng=g.ng, n=a.n, solver=a.solver)

    print(f"{a.problem} {a.n}^2  solver={a.solver}  emf={a.emf_mode}"
          f"{' (flux-CT)' if a.no_upwind_emf else ''}  limiter={a.limiter}  cfl={a.cfl}  "
          f"gamma={P['gamma']:.4f}  bc=({bc_x},{bc_y})  t_end={tend}  precision={a.precision}",
          flush=True)
    if a.restart:
        print(f"  restart from {a.restart}: t={float(R['t']):.6f} step={int(R['step'])} "
              f"generation={gen}", flush=True)
        log = open(os.path.join(out, "diag.csv"), "a")
    else:
        save("000", 0.0)
        log = open(os.path.join(out, "diag.csv"), "w")
        log.write(DIAG_HEADER)

    t, step, t0 = 0.0, 0, time.time()
    next_snap = 1
    if a.restart:
        t, step, next_snap = float(R["t"]), int(R["step"]), int(R["next_snap"])

    def save_restart():
        path = os.path.join(out, "restart.npz")
        tmp = path + ".tmp.npz"
        np.savez(tmp, t=t, step=step, next_snap=next_snap, gen=gen, limiter=a.limiter, ng=g.ng,
                 n=a.n, solver=a.solver, harvest_sweeps=0,
                 Bxf=S.U[6].cpu().numpy(), Byf=S.U[7].cpu().numpy(),
                 **{"cons_" + k: S.U[i].cpu().numpy() for i, k in enumerate(EVOLVED_KEYS)})
        os.replace(tmp, path)

    steps_here = 0
    while t < tend - 1e-14:
        dt = min(float(S.compute_dt(S.U, S.P, a.cfl).item()), tend - t)
        if not np.isfinite(dt) or dt <= 0.0:
            raise SystemExit("run failed at step %d, t = %.6f: dt = %r.  The state went "
                             "non-finite (rho_max %.3e, p_min %.3e); no snap_fin is written and "
                             "the exit status is non-zero."
                             % (step, t, dt, float(S.P[0].max()), float(S.P[4].min())))
        S.set_dt(dt)
        S.step("rk3")
        t += dt
        step += 1
        if not np.isfinite(float(S.U[0].sum())):
            raise SystemExit("run failed at step %d, t = %.6f: the conserved state is not "
                             "finite.  No snap_fin is written and the exit status is non-zero."
                             % (step, t))
        steps_here += 1
        if a.max_steps is not None:
            print(f"  step {step:5d} t={t:.4f} dt={dt:.2e} ({time.time() - t0:.0f}s)", flush=True)

        if step % max(1, a.log_every) == 0 or t >= tend - 1e-14:
            cur = current()
            pr = cur.prims
            ph = g.phys
            d = div_b(cur.Bxf, cur.Byf, g.dx, g.dy)[ph]
            v2 = pr["vx"][ph] ** 2 + pr["vy"][ph] ** 2 + pr["vz"][ph] ** 2
            W = float((1.0 / torch.sqrt(torch.clamp(1 - v2, min=1e-16))).max())
            dg = S.diag_dicts()
            dx_, dy_ = dict(dg["x"]["diag"]), dict(dg["y"]["diag"])
            _, status = D2.cons_to_prims_2d({**cur.cons, "Bx": pr["Bx"], "By": pr["By"]}, g, eos,
                                            1e-10, return_status=True)
            c2p_bad = int((~status["converged"][ph]).sum())
            sym = symmetry_error(pr, g) if P["sym"] else float("nan")
            log.write(f"{step},{t:.6e},{dt:.3e},{float(d.abs().max()):.3e},"
                      f"{float(d.pow(2).mean().sqrt()):.3e},{sym:.3e},"
                      f"{float(pr['rho'][ph].max()):.4f},{float(pr['rho'][ph].min()):.4e},"
                      f"{float(pr['p'][ph].min()):.4e},{W:.3f},"
                      f"{dx_.get('frac_hlle_fallback', float('nan')):.4f},"
                      f"{dy_.get('frac_hlle_fallback', float('nan')):.4f},"
                      f"{dx_.get('mean_iters', float('nan')):.2f},{c2p_bad},-1,-1\n")
            log.flush()
            extra = f"sym={sym:.2e} " if P["sym"] else ""
            print(f"  step {step:5d} t={t:.4f} dt={dt:.2e} divB={float(d.abs().max()):.2e} "
                  f"{extra}rho=[{float(pr['rho'][ph].min()):.3f},{float(pr['rho'][ph].max()):.3f}]"
                  f" c2p_bad={c2p_bad} ({time.time() - t0:.0f}s)", flush=True)

        if rd.snap_due(t, next_snap):
            save(f"{next_snap:03d}", t)
            next_snap += 1
        done = t >= tend - 1e-14
        if rd.restart_due(step, done):
            save_restart()
        if a.max_steps is not None and steps_here >= a.max_steps and not done:
            save_restart()
            log.close()
            wall = time.time() - t0
            rd.record_cost(step, steps_here, wall, stopped_early=True)
            print(f"stopped after {steps_here} steps at t={t:.6f} (step {step}); restart file "
                  f"written ({wall:.0f}s)", flush=True)
            return out

    save("fin", t)
    log.close()
    wall = time.time() - t0
    rd.record_cost(step, steps_here, wall)
    print(f"done in {wall:.0f}s, {step} steps -> {out}", flush=True)
    return out


if __name__ == "__main__":
    main(sys.argv[1:])
