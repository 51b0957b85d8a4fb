"""Differentiable 2D rollout: the SRMHD step with a hand-written adjoint.

The reference differentiates nothing; its step is src/physics/driver2d.py:
rk_step_ct (221-249).  ``Rollout2D`` runs that step (the FP64, bit-faithful
form) for a fixed sequence of dt values and returns the reverse-mode
derivative of a loss on the final primitives with respect to the initial
state, through the C ABI's ``srmhd2d_stage_adjoint`` (per RK stage:
con2prim by implicit differentiation of the Kastaun root, BCs, the RK
combination, flux divergence and constrained transport, and per face the
reconstruction and the Riemann solver, HLLD's p* by implicit differentiation
at the recorded root).

Memory: the forward keeps a checkpoint (state + primitives) every
``checkpoint_every`` steps and runs the last segment with the stage records
(``srmhd2d_stage_taped``); the backward walks that segment's records, then
re-runs every earlier segment with records and walks it backwards.  The
re-run of the next segment (a second stepper, on a side stream) goes on
while the current one is walked, so the two share the device; the resident
set is three segments of stage records plus the checkpoints (at 512^2,
about 77 MB of records per RK2 step).  ``overlap=False`` runs the re-runs
in line on the caller's stream instead (two segments resident).

``OTRollout`` is the differentiable relativistic Orszag-Tang problem: the
initial density field (physical cells) and the magnetic amplitude A are the
parameters; a ``torch.autograd.Function`` returns the final density.

There is no CPU fallback: every call runs HIP kernels.
"""

import ctypes
import math

import torch

from . import _lib
from .driver2d import RK, Stepper2D
from .hlld import primitive_to_conserved
from .state import EVOLVED_KEYS, PRIM6

CONS8 = ("D", "tau", "Sx", "Sy", "Sz", "Bx", "By", "Bz")   # srmhd_p2c order


class Rollout2D:
    """Forward and reverse sweeps of ``steps`` RK steps with frozen dt values
    on one grid.  ``dts``: a sequence of floats (one per step)."""

    def __init__(self, grid, eos, *, solver="hlld", limiter="mc", bc_x="outflow",
                 bc_y="outflow", upwind=True, scheme="rk2", checkpoint_every=20,
                 atmo_rho=1e-10, atmo_eps=1e-10, device=None, overlap=True):
        # the adjoint differentiates the reference's default p* root, so the
        # rollout keeps the default options whatever srmhd.hlld's switches say
        kw = dict(solver=solver, limiter=limiter, bc_x=bc_x, bc_y=bc_y, upwind=upwind,
                  atmo_rho=atmo_rho, atmo_eps=atmo_eps, device=device, precision="fp64",
                  hlld_opts_from_globals=False)
        self.st = Stepper2D(grid, eos, **kw)
        # the backward re-runs the next segment while it walks the current
        # one: that re-run has its own stepper (workspace, dt) and stream,
        # and keeps its two sweeps in sequence on it (the device's second
        # stream then serves the walk's sweeps alone)
        self.st_rerun = Stepper2D(grid, eos, serial_sweeps=bool(overlap), **kw)
        self.overlap = bool(overlap)
        self.side = torch.cuda.Stream(device=self.st.device) if self.overlap else None
        self.grid = grid
        self.device = self.st.device
        self.lib = self.st.lib
        self.scheme = scheme
        if scheme not in RK:
            raise ValueError(f"scheme must be one of {tuple(RK)}")
        self.K = int(checkpoint_every)
        if self.K < 1:
            raise ValueError("checkpoint_every must be >= 1")
        nb = self.lib.srmhd2d_adjoint_workspace_bytes(ctypes.byref(self.st.cg),
                                                      ctypes.byref(self.st.cs))
        if nb < 0:
            raise ValueError("grid / scheme rejected by srmhd2d_adjoint_workspace_bytes")
        self.aws = _lib.workspace(nb, self.device)
        self.ckpt = {}
        self.tail, self.tail_start = [], 0
        self.dts = None

    # ── buffers ─────────────────────────────────────────────────────────────
    def _zeros_state(self):
        return [torch.zeros_like(t) for t in self.st.new_state()]

    def _zeros_prims(self):
        return [torch.zeros_like(t) for t in self.st.new_prims()]

    def _zeros_group(self, n_states, n_prims):
        """n_states state lists and n_prims primitive lists of zeros, cut
        from one buffer so that one fill clears them all."""
        g = self.grid
        cell = (g.nxt, g.nyt)
        state = [cell] * len(EVOLVED_KEYS) + [(g.nxt + 1, g.nyt), (g.nxt, g.nyt + 1)]
        prims = [cell] * len(PRIM6)
        shapes = state * n_states + prims * n_prims
        flat = torch.zeros(sum(a * b for a, b in shapes), dtype=torch.float64,
                           device=self.device)
        views, o = [], 0
        for a, b in shapes:
            views.append(flat[o:o + a * b].view(a, b))
            o += a * b
        ns = len(state)
        states = [views[i * ns:(i + 1) * ns] for i in range(n_states)]
        base = n_states * ns
        np_ = len(prims)
        return states, [views[base + i * np_:base + (i + 1) * np_] for i in range(n_prims)]

    def _new_tape(self):
        g, d = self.grid, self.device
        nfx, nfy = (g.nxt + 1) * g.nyt, g.nxt * (g.nyt + 1)
        bufs = [torch.empty(nfx, dtype=torch.float64, device=d),
                torch.empty(nfy, dtype=torch.float64, device=d),
                torch.empty(nfx, dtype=torch.float64, device=d),
                torch.empty(nfy, dtype=torch.float64, device=d),
                torch.empty(2, dtype=torch.float64, device=d)]
        return bufs, _lib.Tape2D(*[b.data_ptr() for b in bufs])

    # ── forward ─────────────────────────────────────────────────────────────
    def _step(self, U, P, dt, record=None, st=None):
        """One RK step from (U, P) -> new (U, P); with ``record`` (a list) the
        stage inputs and tapes are appended to it.  ``st``: the stepper (the
        rollout's own by default)."""
        st = self.st if st is None else st
        st.set_dt(dt)
        U0 = U
        for a, b, c in RK[self.scheme]:
            out, Pout = st.new_state(), st.new_prims()
            if record is None:
                st.stage(a, b, c, U0, U, P, out, Pout)
            else:
                bufs, tape = self._new_tape()
                a0, k0 = _lib.ptrs(U0)
                a1, k1 = _lib.ptrs(U)
                ap, kp = _lib.ptrs(P)
                ao, ko = _lib.ptrs(out)
                apo, kpo = _lib.ptrs(Pout)
                rc = self.lib.srmhd2d_stage_taped(
                    ctypes.byref(st.cg), ctypes.byref(st.cs), ctypes.byref(st.ceos), float(a),
                    float(b), float(c), st.dt.data_ptr(), a0, a1, ap, ao, apo, ctypes.byref(tape),
                    st.ws.data_ptr(), st.ws.numel(), _lib.stream_ptr(self.device))
                _lib.check(rc, "srmhd2d_stage_taped")
                record.append((U, P, out, bufs, tape))
            U, P = out, Pout
        return U, P

    def forward(self, U, P, dts):
        """Run len(dts) steps from the synced state (U: 8 tensors, P: 6),
        keeping checkpoints and the last segment's stage records; returns
        the final (U, P)."""
        self.dts = [float(x) for x in dts]
        N = len(self.dts)
        # the last segment starts at the last multiple of K below N
        self.tail_start = ((N - 1) // self.K) * self.K if N else 0
        self.tail = []
        U = [t.clone() for t in U]
        P = [t.clone() for t in P]
        self.ckpt = {0: ([t.clone() for t in U], [t.clone() for t in P])}
        for n, dt in enumerate(self.dts):
            if n >= self.tail_start:
                rec = []
                U, P = self._step(U, P, dt, record=rec)
                self.tail.append(rec)
            else:
                U, P = self._step(U, P, dt)
            if (n + 1) % self.K == 0 and n + 1 < self.tail_start:
                self.ckpt[n + 1] = ([t.clone() for t in U], [t.clone() for t in P])
        return U, P

    # ── reverse ─────────────────────────────────────────────────────────────
    def _stage_adjoint(self, a, b, c, U, P, Uout, tape, lamUout, lamPout, lamU0, lamU, lamP):
        st = self.st
        arrs = [_lib.ptrs(x) for x in (U, P, Uout, lamUout, lamPout, lamU0, lamU, lamP)]
        rc = self.lib.srmhd2d_stage_adjoint(
            ctypes.byref(st.cg), ctypes.byref(st.cs), ctypes.byref(st.ceos), float(a), float(b),
            float(c), st.dt.data_ptr(), arrs[0][0], arrs[1][0], arrs[2][0], ctypes.byref(tape),
            arrs[3][0], arrs[4][0], arrs[5][0], arrs[6][0], arrs[7][0], self.aws.data_ptr(),
            self.aws.numel(), _lib.stream_ptr(self.device))
        _lib.check(rc, "srmhd2d_stage_adjoint")

    def backward(self, lamU, lamP):
        """Adjoints of the final state (8) and primitives (6) -> adjoints of
        the initial synced state and primitives.  The inputs are consumed."""
        if self.dts is None:
            raise RuntimeError("backward() needs a forward() first")
        N = len(self.dts)
        stages = RK[self.scheme]
        lamU = [t.contiguous() for t in lamU]
        lamP = [t.contiguous() for t in lamP]
        S = len(stages)
        main = torch.cuda.current_stream(self.device)
        side = self.side
        if side is not None:
            side.wait_stream(main)   # the checkpoints and the tail are written
        starts = list(reversed(range(0, N, self.K)))
        pending = None
        for idx, s0 in enumerate(starts):
            s1 = min(s0 + self.K, N)
            if s0 == self.tail_start:
                recs = self.tail     # recorded by forward()
            elif pending is not None:
                recs, done = pending   # re-run on the side stream
                main.wait_event(done)
            else:
                recs = self._rerun(s0)
            # the next segment's re-run goes on beside this segment's walk
            pending = None
            if side is not None and idx + 1 < len(starts):
                pending = self._rerun_side(starts[idx + 1], main)
            for n in reversed(range(s0, s1)):
                self.st.set_dt(self.dts[n])
                rec = recs[n - s0]
                # acc (the step input's adjoint), the adjoint of stage k > 0's
                # input state and every stage's primitive adjoint
                states, prims = self._zeros_group(S, S)
                acc = states[0]
                for k in reversed(range(S)):
                    a, b, c = stages[k]
                    Uk, Pk, Uout, _bufs, tape = rec[k]
                    newU = acc if k == 0 else states[k]
                    newP = prims[k]
                    self._stage_adjoint(a, b, c, Uk, Pk, Uout, tape, lamU, lamP, acc, newU, newP)
                    lamU, lamP = newU, newP
                lamU = acc
            del recs
        if side is not None:
            main.wait_stream(side)
        return lamU, lamP

    def _rerun(self, s0, st=None):
        """The stage records of the segment starting at step s0, re-run from
        its checkpoint."""
        U, P = self.ckpt[s0]
        recs = []
        for n in range(s0, min(s0 + self.K, len(self.dts))):
            rec = []
            U, P = self._step(U, P, self.dts[n], record=rec, st=st)
            recs.append(rec)
        return recs

    def _rerun_side(self, s0, main):
        """_rerun on the side stream with the second stepper; returns the
        records and the event that marks them written.  Every record tensor
        is marked as used on the main stream, so that its memory is not
        handed out again before the walk that reads it has run."""
        with torch.cuda.stream(self.side):
            recs = self._rerun(s0, st=self.st_rerun)
            done = torch.cuda.Event()
            done.record(self.side)
        for rec in recs:
            for Uk, Pk, Uout, bufs, _tape in rec:
                for t in (*Uk, *Pk, *Uout, *bufs):
                    t.record_stream(main)
        return recs, done

    def sync_adjoint(self, U, lamU, lamP):
        """Adjoint of sync_state: -> the adjoint of the state before the
        sync (ghost entries 0).  lamU, lamP are consumed."""
        st = self.st
        out = self._zeros_state()
        au, ku = _lib.ptrs(U)
        al, kl = _lib.ptrs(lamU)
        ap, kp = _lib.ptrs(lamP)
        ao, ko = _lib.ptrs(out)
        rc = self.lib.srmhd2d_sync_adjoint(ctypes.byref(st.cg), ctypes.byref(st.cs),
                                           ctypes.byref(st.ceos), au, al, ap, ao,
                                           self.aws.data_ptr(), self.aws.numel(),
                                           _lib.stream_ptr(self.device))
        _lib.check(rc, "srmhd2d_sync_adjoint")
        return out


def p2c_adjoint(prims9, lam_cons8, device):
    """lam_prim (9, rho vx vy vz p eps Bx By Bz) = J^T lam_cons for
    primitive_to_conserved, point-wise (srmhd_p2c_adjoint)."""
    lib = _lib.lib()
    n = prims9[0].numel()
    P = [t.contiguous().reshape(-1) for t in prims9]
    L = [t.contiguous().reshape(-1) for t in lam_cons8]
    out = [torch.zeros(n, dtype=torch.float64, device=device) for _ in range(9)]
    ap, kp = _lib.ptrs(P)
    al, kl = _lib.ptrs(L)
    ao, ko = _lib.ptrs(out)
    _lib.check(lib.srmhd_p2c_adjoint(n, ap, al, ao, _lib.stream_ptr(device)), "srmhd_p2c_adjoint")
    return [o.reshape(prims9[0].shape) for o in out]


# ── the differentiable Orszag-Tang problem ──────────────────────────────────

class OTProblem:
    """Relativistic Orszag-Tang on [0, 2 pi]^2, periodic: rho = rho_field
    (the physical cells; ghosts by periodic wrap), p = 1,
    v = vmax/sqrt2 (-sin y, sin x, 0), B = A (-sin y, sin 2x, 0) with the
    staggered field the curl of Az = A (cos y + cos(2x)/2) (div B = 0)."""

    def __init__(self, grid, eos, vmax=0.99, press=1.0):
        self.grid, self.eos = grid, eos
        g = grid
        X, Y = g.X, g.Y
        amp = vmax / math.sqrt(2.0)
        d = g.device
        self.vx = (-amp * torch.sin(Y)).to(d)
        self.vy = (amp * torch.sin(X)).to(d)
        self.bx0 = (-torch.sin(Y)).to(d)
        self.by0 = torch.sin(2.0 * X).to(d)
        self.press = press
        XF = g.xf.unsqueeze(1).expand(g.nxt + 1, g.nyt + 1)
        YF = g.yf.unsqueeze(0).expand(g.nxt + 1, g.nyt + 1)
        Az = torch.cos(YF) + 0.5 * torch.cos(2.0 * XF)
        self.Bxf0 = ((Az[:, 1:] - Az[:, :-1]) / g.dy).contiguous().to(d)
        self.Byf0 = (-(Az[1:, :] - Az[:-1, :]) / g.dx).contiguous().to(d)

    def wrap(self, rho_phys):
        """Physical-cell field -> all cells (periodic ghosts)."""
        ng = self.grid.ng
        return torch.nn.functional.pad(rho_phys[None, None], (ng, ng, ng, ng),
                                       mode="circular")[0, 0]

    def prims9(self, rho_phys, A):
        """The nine initial primitives (differentiable torch ops)."""
        rho = self.wrap(rho_phys)
        p = torch.full_like(rho, self.press)
        z = torch.zeros_like(rho)
        eps = self.eos.eps__press_rho(p, rho)
        return [rho, self.vx, self.vy, z, p, eps, A * self.bx0, A * self.by0, z]

    def initial_state(self, roll, rho_phys, A):
        """init_state (run_2d.py:227-231): cons from the primitives, faces
        from Az, sync -> (U, P, prims9)."""
        with torch.no_grad():
            P9 = [t.contiguous() for t in self.prims9(rho_phys, A)]
            cons = primitive_to_conserved(dict(zip(("rho", "vx", "vy", "vz", "p", "eps",
                                                    "Bx", "By", "Bz"), P9)))
            U = [cons[k].contiguous() for k in EVOLVED_KEYS]
            U += [(A * self.Bxf0).contiguous(), (A * self.Byf0).contiguous()]
            P = roll.st.new_prims()
            roll.st.sync(U, P)
        return U, P, P9

    def init_backward(self, roll, U, lamU, lamP, rho_phys, A):
        """Adjoints of the synced initial state -> (d/d rho_phys, d/dA)."""
        lam_pre = roll.sync_adjoint(U, lamU, lamP)
        idx = {k: i for i, k in enumerate(EVOLVED_KEYS)}
        zero = torch.zeros_like(lam_pre[0])
        lam_cons = [lam_pre[idx[k]] if k in idx else zero for k in CONS8]
        with torch.no_grad():
            P9 = [t.contiguous() for t in self.prims9(rho_phys, A)]
        lam9 = p2c_adjoint(P9, lam_cons, roll.device)
        with torch.enable_grad():
            r = rho_phys.detach().requires_grad_(True)
            a = A.detach().requires_grad_(True)
            P9g = self.prims9(r, a)
            faces = [a * self.Bxf0, a * self.Byf0]
            outs = [t for t in P9g if t.requires_grad] + faces
            grads = [l for t, l in zip(P9g, lam9) if t.requires_grad] + [lam_pre[6], lam_pre[7]]
            g_rho, g_A = torch.autograd.grad(outs, (r, a), grads, allow_unused=True)
        if g_rho is None:
            g_rho = torch.zeros_like(rho_phys)
        if g_A is None:
            g_A = torch.zeros_like(A)
        return g_rho, g_A


def frozen_dts(roll, U, P, steps, cfl):
    """The dt sequence of a plain forward run (compute_dt_2d each step)."""
    st = roll.st
    st.load([t.clone() for t in U], [t.clone() for t in P])
    dts = []
    for _ in range(steps):
        st.compute_dt(st.U, st.P, cfl)
        dts.append(float(st.dt.item()))
        st.step(roll.scheme)
    return dts


class OTRollout(torch.autograd.Function):
    """rho_final (physical cells) = rollout(rho_phys, A) with the hand-written
    adjoint as its backward.  Non-tensor arguments: the OTProblem, the
    Rollout2D and the frozen dt sequence."""

    @staticmethod
    def forward(ctx, rho_phys, A, prob, roll, dts):
        U, P, _ = prob.initial_state(roll, rho_phys, A)
        Uf, Pf = roll.forward(U, P, dts)
        ctx.prob, ctx.roll, ctx.U0 = prob, roll, U
        ctx.save_for_backward(rho_phys, A)
        g = roll.grid
        return Pf[0][g.ng:g.ng + g.nx, g.ng:g.ng + g.ny].clone()

    @staticmethod
    def backward(ctx, grad_rho):
        rho_phys, A = ctx.saved_tensors
        roll, prob = ctx.roll, ctx.prob
        g = roll.grid
        lamU = roll._zeros_state()
        lamP = roll._zeros_prims()
        lamP[0][g.ng:g.ng + g.nx, g.ng:g.ng + g.ny] = grad_rho.to(torch.float64)
        lamU0, lamP0 = roll.backward(lamU, lamP)
        g_rho, g_A = prob.init_backward(roll, ctx.U0, lamU0, lamP0, rho_phys, A)
        return g_rho, g_A, None, None, None


def ot_final_rho(prob, roll, rho_phys, A, dts):
    """The final density (physical cells) of the rollout, no gradient."""
    with torch.no_grad():
        U, P, _ = prob.initial_state(roll, rho_phys, A)
        for dt in dts:
            U, P = roll._step(U, P, dt)
        g = roll.grid
        return P[0][g.ng:g.ng + g.nx, g.ng:g.ng + g.ny].clone()
