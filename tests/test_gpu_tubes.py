"""Batched Riemann tubes on the HIP path.

* the per-tube-dt runner (one persistent launch, ragged lengths) bitwise
  against the oracle's per-tube loop, for several limiter / solver / scheme
  combinations, independent of the batch it runs in, of the workgroup size
  and of graph capture;
* the shared-dt form of the reference's run_tubes bitwise against the
  oracle's restatement and within 1e-12 of the reference's own output.
"""

import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

PRIM9 = ("rho", "vx", "vy", "vz", "p", "eps", "Bx", "By", "Bz")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _same(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return (a == b) | (np.isnan(a) & np.isnan(b))


def _workload(n, seed, lmin, lmax):
    from srmhd import tubes as T
    return T.draw_tubes(n, seed=seed, lmin=lmin, lmax=lmax)


def _side(s, i):
    return {k: float(v[i]) for k, v in s.items()}


def _batch(lengths, left, right, Bn, limiter, solver, **kw):
    from srmhd import eos as E, tubes as T
    eos = E.hybrid_eos(0.0, 5.0 / 3.0, 5.0 / 3.0)
    return T.TubeBatch(lengths, left, right, Bn, eos, limiter=limiter, solver=solver,
                       device="cuda", **kw)


def _oracle_tube(L, ng, left, right, Bn, steps, limiter, solver, scheme):
    import oracle.eos as OE
    import oracle.tubes as OT
    eos = OE.ideal(5.0 / 3.0)
    P0 = OT.tube_initial(L, ng, eos, left, right, Bn)
    return OT.run_tube(P0, eos, L, ng, steps, cfl=0.4, limiter=limiter, scheme=scheme,
                       flux=solver)


@pytest.mark.parametrize("limiter,solver,scheme", [("mp5", "hlld", "rk3"), ("mc", "hlle", "rk2"),
                                                   ("weno5z", "hllc", "rk3"),
                                                   ("pcm", "hlld", "rk2")])
def test_per_tube_bitwise_vs_oracle(limiter, solver, scheme):
    _dev()
    print(f"tubes per-tube {limiter} {solver} {scheme}", flush=True)
    lengths, left, right, Bn = _workload(5, seed=11, lmin=16, lmax=48)
    steps = 4
    runs = []
    for block in (None, 0):        # the persistent kernel and the phased form
        b = _batch(lengths, left, right, Bn, limiter, solver)
        b.run(steps, cfl=0.4, scheme=scheme, block=block)
        torch.cuda.synchronize()
        runs.append(b)
    for i in range(runs[0].n_tubes):
        P, t = _oracle_tube(int(lengths[i]), runs[0].ng, _side(left, i), _side(right, i),
                            float(Bn[i]), steps, limiter, solver, scheme)
        for b in runs:
            assert b.t.cpu().numpy()[i] == t, (i, b.block, t)
            got = b.tube(i)
            for k in PRIM9:
                s = _same(got[k].cpu().numpy(), P[k].numpy())
                assert s.all(), f"tube {i} {k}: {int((~s).sum())} differ"


def test_batch_independence_block_size_and_graph():
    """A tube's result does not depend on the other tubes, on the workgroup
    size or on replaying a captured graph (per-tube batches, deterministic
    block reductions)."""
    _dev()
    print("tubes independence / block / graph", flush=True)
    lengths, left, right, Bn = _workload(9, seed=5, lmin=20, lmax=300)
    ref = _batch(lengths, left, right, Bn, "mp5", "hlld")
    ref.run(3, block=256)
    torch.cuda.synchronize()
    for block in (64, 128, 512, 0):   # 0: the phased form
        b = _batch(lengths, left, right, Bn, "mp5", "hlld")
        b.run(3, block=block)
        torch.cuda.synchronize()
        for u, v in zip(b.P + b.U, ref.P + ref.U):
            assert _same(u.cpu().numpy(), v.cpu().numpy()).all(), f"block {block} differs"
    # tube 4 alone
    i = 4
    one = _batch(lengths[i:i + 1], {k: v[i:i + 1] for k, v in left.items()},
                 {k: v[i:i + 1] for k, v in right.items()}, Bn[i:i + 1], "mp5", "hlld")
    one.run(3)
    torch.cuda.synchronize()
    a, c = ref.tube(i), one.tube(0)
    for k in PRIM9:
        assert _same(a[k].cpu().numpy(), c[k].cpu().numpy()).all(), k
    # graph replays: each replay from the same start state equals the eager run
    b = _batch(lengths, left, right, Bn, "mp5", "hlld")
    snap = b.snapshot()
    g = b.capture(3)
    for rep in range(3):
        b.restore(snap)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for u, v in zip(b.P + b.U, ref.P + ref.U):
            assert _same(u.cpu().numpy(), v.cpu().numpy()).all(), f"replay {rep} differs"
        assert torch.equal(b.t, ref.t)


def test_autotune_picks_a_valid_block():
    _dev()
    from srmhd import tubes as T
    lengths, left, right, Bn = _workload(64, seed=2, lmin=64, lmax=256)
    b = _batch(lengths, left, right, Bn, "mp5", "hlld")
    snap = b.snapshot()
    best, times = T.autotune_block(b, steps=1)
    assert best in (0, 64, 128, 256, 512) and set(times) == {0, 64, 128, 256, 512}
    for u, v in zip(b.U + b.P, snap[0] + snap[1]):   # state restored
        assert torch.equal(u, v)
    assert T.autotune_block(b, steps=1)[0] == best   # cached


# Bound against the reference's own output.  mc: 1e-12 relative (max).  mp5:
# at step 10 one face's wave-order test (the HLLE fallback of hlld.py) falls
# on opposite sides between IEEE sqrt and the torch CPU sqrt the fixture was
# written with (not correctly rounded on every host); the IEEE oracle shows the
# same gap (L1(rho) 9.2e-6) and the oracle in torch-sqrt mode reproduces the
# fixture bitwise, so the bound there is L1(rho) <= 2e-5.
REF_BOUND = {"mc": ("max", 1e-12), "mp5": ("l1", 2e-5)}


@pytest.mark.parametrize("lim", ["mc", "mp5"])
def test_shared_dt_bitwise_vs_oracle_and_close_to_reference(golden_dir, lim):
    _dev()
    import oracle.eos as OE
    import oracle.tubes as OT
    from srmhd import eos as E, tubes as T
    print(f"tubes shared {lim}", flush=True)
    g = np.load(os.path.join(golden_dir, "tubes_shared.npz"))
    eos = E.hybrid_eos(0.0, 5.0 / 3.0, 5.0 / 3.0)
    prof, t = T.run_tubes(g["UL"], g["UR"], g["Bn"], eos, ncells=48, tend=0.04, limiter=lim,
                          max_steps=400)
    po, to = OT.run_tubes_shared(g["UL"], g["UR"], g["Bn"], OE.ideal(5.0 / 3.0), ncells=48,
                                 tend=0.04, limiter=lim)
    assert t == to
    for k in PRIM9:
        s = _same(prof[k].cpu().numpy(), po[k].numpy())
        assert s.all(), f"{k}: {int((~s).sum())} differ from the oracle"
    assert abs(t - float(g[f"{lim}_t"])) <= 1e-14
    kind, bound = REF_BOUND[lim]
    if kind == "max":
        for k in PRIM9:
            a, ref = prof[k].cpu().numpy(), g[f"{lim}_{k}"]
            err = np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300)
            assert err <= bound, (k, err)
    else:
        a, ref = prof["rho"].cpu().numpy(), g[f"{lim}_rho"]
        l1 = np.abs(a - ref).mean() / np.abs(ref).mean()
        print(f"L1(rho) vs reference = {l1:.3e}", flush=True)
        assert l1 <= bound
    # eager loop = graph replay
    prof2, t2 = T.run_tubes(g["UL"], g["UR"], g["Bn"], eos, ncells=48, tend=0.04, limiter=lim,
                            max_steps=400, graph=False)
    assert t2 == t
    for k in PRIM9:
        assert _same(prof2[k].cpu().numpy(), prof[k].cpu().numpy()).all(), k


@pytest.mark.parametrize("lim,flux", [("minmod", "hlle"), ("weno5z", "hllc")])
def test_shared_dt_other_solvers_bitwise_vs_oracle(golden_dir, lim, flux):
    _dev()
    import oracle.eos as OE
    import oracle.tubes as OT
    from srmhd import eos as E, tubes as T
    print(f"tubes shared {lim} {flux}", flush=True)
    g = np.load(os.path.join(golden_dir, "tubes_shared.npz"))
    eos = E.hybrid_eos(0.0, 5.0 / 3.0, 5.0 / 3.0)
    prof, t = T.run_tubes(g["UL"], g["UR"], g["Bn"], eos, ncells=40, tend=0.03, limiter=lim,
                          max_steps=6, flux=flux)
    po, to = OT.run_tubes_shared(g["UL"], g["UR"], g["Bn"], OE.ideal(5.0 / 3.0), ncells=40,
                                 tend=0.03, limiter=lim, max_steps=6, flux=flux)
    assert t == to
    for k in PRIM9:
        s = _same(prof[k].cpu().numpy(), po[k].numpy())
        assert s.all(), f"{k}: {int((~s).sum())} differ from the oracle"


def _oracle_continue(U, P, Bn, ncells, ng, eos, limiter, solver, scheme):
    """One step of the oracle's per-tube loop (oracle/tubes.py:run_tube) from
    a given conserved state U (D, Sx, Sy, Sz, tau, By, Bz) and its
    primitives P (rho, vx, vy, vz, p, eps, Bx, By, Bz), ghosts included."""
    import oracle.tubes as OT
    dx = 1.0 / ncells
    cons = dict(U)
    Bx = P["Bx"]
    dt = float(OT.compute_dt(P, eos, dx, 0.4, ng, ncells, speed_of_light=False))
    U0, Uk = cons, cons
    for a, b, c in OT.RK[scheme]:
        L = OT._rhs(P, eos, dx, limiter, solver)
        w = c * dt
        Uk = {k: a * U0[k] + b * Uk[k] + w * L[k] for k in U0}
        P = OT._sync(Uk, Bx, eos, ng)
    return Uk, P


def test_nonfinite_tubes_blow_up_on_the_oracle_too():
    """The tubes of the benchmark workload (4096 tubes of seed 0, MP5 + HLLD +
    SSP-RK3, per-tube dt, 500 steps) that end non-finite: each is run alone
    (the per-tube dt makes it independent of the batch), one step at a time,
    to its first non-finite step S; from its device state three steps
    before S, the oracle's per-tube step continues in lockstep, equal bit for
    bit after steps S - 2 and S - 1 and non-finite after step S too.  The
    blow-ups are the algorithm's, not the device path's."""
    _dev()
    import oracle.eos as OE
    from srmhd import tubes as T
    lengths, left, right, Bn = _workload(4096, seed=0, lmin=128, lmax=2048)
    b = _batch(lengths, left, right, Bn, "mp5", "hlld")
    b.run(500, cfl=0.4, block=0)   # the phased form (bit-identical, faster here)
    torch.cuda.synchronize()
    fin = torch.ones(b.total_cells, dtype=torch.bool, device=b.U[0].device)
    for p in b.P:
        fin &= torch.isfinite(p)
    off = b.off_host
    bad = [i for i in range(b.n_tubes) if not bool(fin[int(off[i]):int(off[i + 1])].all())]
    print(f"non-finite tubes after 500 steps: {bad}", flush=True)
    assert 0 < len(bad) <= 64
    eos = OE.ideal(5.0 / 3.0)
    for i in bad:
        one = _batch(lengths[i:i + 1], {k: v[i:i + 1] for k, v in left.items()},
                     {k: v[i:i + 1] for k, v in right.items()}, Bn[i:i + 1], "mp5", "hlld")
        ng, L = one.ng, int(lengths[i])
        hist = []   # (U, P) on the host after the last few steps
        S = None
        for s in range(1, 501):
            one.run(1, cfl=0.4)
            U = {k: u.cpu() for k, u in zip(T.U_KEYS, one.U)}
            tb = one.tube(0)
            P = {k: v.cpu().clone() for k, v in tb.items()}
            if not all(bool(torch.isfinite(P[k]).all()) for k in T.P_KEYS):
                S = s
                break
            hist = (hist + [(U, P)])[-3:]
        assert S is not None and S >= 4, (i, S)
        Uo, Po = hist[0]                      # after step S - 3
        for j, s in enumerate((S - 2, S - 1, S)):
            Uo, Po = _oracle_continue(Uo, Po, float(Bn[i]), L, ng, eos, "mp5", "hlld", "rk3")
            if s < S:
                Ug, Pg = hist[j + 1]
                for k in T.P_KEYS:
                    same = _same(Pg[k].numpy(), Po[k].numpy())
                    assert same.all(), f"tube {i} step {s} {k}: {int((~same).sum())} differ"
            else:
                assert not all(bool(torch.isfinite(Po[k]).all()) for k in T.P_KEYS), \
                    f"tube {i}: the oracle stays finite at step {S}"
        print(f"tube {i} ({L} cells): non-finite from step {S} on both", flush=True)
