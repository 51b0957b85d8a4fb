"""Throughput of the SRMHD step on one MI355X: zone-cycles/s.

Workload (the headline configuration): the tapered relativistic magnetic
rotor on [-0.5, 0.5]^2 at 2048^2, Gamma = 5/3, WENO5-Z + HLLD + upwind
constrained transport + Kastaun con2prim + SSP-RK3, CFL 0.4, outflow BCs.
A step is one full RK3 step including the CFL reduction; one zone-cycle is
one physical cell advanced one step.  Each precision form (float64, and the
FP32-flux forms) is timed over K steps from the same initial state and then
continued, untimed, to step 200, where its L1(rho, p, |B|) distance from the
float64 form and div B are taken.  ``value`` is the fastest form whose
L1(rho) there is <= 1e-6 (the float64 form always qualifies).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints ONE JSON line.  ``roofline`` is for the headline form's dominant
kernel (measured live with HIP events around its launches inside the timed
region); ``cpu_baseline`` times the CPU oracle (torch, float64) on a bounded
sample of the same problem.  The same line carries the other two
configurations as keys: ``tubes`` (4096 ragged Riemann tubes, MP5/HLLD/RK3,
500 steps in one persistent launch) and ``adjoint`` (the gradient of the
differentiable 512^2 Orszag-Tang rollout, 200 steps, checked against central
differences), each with its own roofline and CPU baseline.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "hlld-ai_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
# steps (from the initial state, sweeps in sequence) over which the rotor
# forms' dominant kernel class is timed with HIP events
KERNEL_TIMING_STEPS = 20
FP64_PEAK_TFLOPS = 78.6

# FP64 VALU work per element of the dominant kernels, from rocprofv3 --pmc
# SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 on the 2048^2 rotor
# (profiles/rotor2048_counters.md): flops = 64 lanes x (add + mul + 2 fma +
# trans) per wave, divided by the elements of the launch.
FP64_FLOPS_PER_ELEMENT = {("fp64", "sweep_x_a"): 13802.0, ("fp64", "sweep_y_a"): 4131.0,
                          ("fp64", "c2p_a"): 10388.0}

# the kernel class each form spends most time in (rocprofv3 --stats of the
# 2048^2 rotor: profiles/rotor2048_weno5z_hlld_kernel_stats.md)
DOMINANT_KERNEL = {"fp64": "sweep_x_a", "mixed": "sweep_x_b", "mixed_r64": "sweep_x_b",
                   "fp64_local": "sweep_x_b"}

# wave-level VALU instructions per launch (rocprofv3 --pmc SQ_INSTS_VALU on the
# 2048^2 rotor, profiles/rotor2048_counters.md)
VALU_INSTS_PER_LAUNCH = {("mixed", "sweep_x_b", 2048): 1.836e8, ("fp64", "sweep_x_a", 2048): 9.097e8}

# measured VALU busy fraction of the launch (same profile): SQ_ACTIVE_INST_VALU
# (quad-cycles in which a SIMD issued a VALU instruction, summed over the
# chip) / (256 CUs x GRBM_GUI_ACTIVE cycles of one XCD) -- 1 means every SIMD
# issued a vector instruction every quad-cycle of the launch
VALU_BUSY = {("fp64", "sweep_x_a", 2048): 0.961}

# the adjoint's x sweep at 512^2 (k_adj_flux<0, 1, 2> + k_adj_recon<0, 1>),
# rocprofv3 --pmc on the Orszag-Tang gradient (profiles/adjoint_ot512_counters.md):
# FP64 flops per face (64 lanes x (add + mul + 2 fma + trans) per wave over
# the launch's faces) and HBM bytes per launch pair (FETCH_SIZE scaled by the
# copy kernel's 2.0, plus WRITE_SIZE)
ADJ_SWEEP_FLOPS_PER_FACE = {512: 6783.0}

# measured HBM bytes per cell and RK3 step of the batched tubes' phased form
# (4096 tubes, MP5 + HLLD; rocprofv3 --pmc FETCH_SIZE, scaled by the copy
# kernel's 2.0, and WRITE_SIZE, summed over one step's launches:
# profiles/tubes4096_counters.md), keyed by (limiter, solver)
TUBES_TRAFFIC_PER_CELL_STEP = {("mp5", "hlld"): 4968.2}
ADJ_SWEEP_TRAFFIC = {512: 269.3e6}

# measured HBM bytes per launch (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE,
# scaled by a plain copy kernel; profiles/rotor2048_counters.md), keyed by
# (form, kernel class, n)
MEASURED_TRAFFIC = {("mixed", "sweep_x_b", 2048): 420.0e6, ("fp64", "sweep_x_a", 2048): 1588.6e6,
                    ("fp64", "c2p_a", 2048): 376.3e6, ("mixed", "c2p_a", 2048): 490.1e6}


def algorithmic_bytes(kernel, n, ng, precision="fp64"):
    """Compulsory HBM bytes of one launch of ``kernel`` on an n^2 grid with
    ng ghosts: every input array element read once, every output written
    once (float64)."""
    nt = n + 2 * ng
    cells = nt * nt
    faces = (nt + 1) * nt
    if precision == "mixed" and kernel in ("sweep_x_b", "sweep_y_b"):
        # the FP32 sweep: 7 FP32 cell arrays (rho, p, W v, two tangential B)
        # and the FP64 normal face field in, 7 FP64 flux arrays out
        return 4 * 7 * cells + 8 * faces + 56 * faces
    if precision in ("fp64_local", "mixed_r64") and kernel in ("sweep_x_b", "sweep_y_b"):
        # the FP64-reconstructing sweep without a pass-A record: prims rho,
        # p, vx, vy, vz + cons Bz + both face-B arrays in, 7 flux arrays out
        return 8 * (6 * cells + 2 * faces) + 56 * faces
    if kernel in ("sweep_x_b", "sweep_y_b"):
        # HLLD's pass B: pass A's side states and HLL speeds (34 doubles/face)
        # and record (3 doubles/face) in; 7 flux arrays out
        return (34 * 8 + 24) * faces + 56 * faces
    if kernel in ("sweep_x_a", "sweep_y_a"):
        # prims rho, p, vx, vy, vz + cons Bz + both face-B arrays in; the
        # record (3 doubles/face) and the side states (34 doubles/face) out
        return 8 * (6 * cells + 2 * faces) + 24 * faces + 34 * 8 * faces
    if precision != "fp64" and kernel == "c2p_a":
        # the per-cell con2prim of the mixed forms: physical cells only, the
        # 6 cell + 2 face state arrays in, 6 primitives out
        return n * n * (8 * 8 + 6 * 8)
    if kernel == "c2p_a":
        # cons D, Sx, Sy, Sz, tau, Bz + both face arrays in; record out
        return 8 * (6 * cells + 2 * faces) + 24 * cells
    if kernel == "c2p_b":
        phys = n * n
        return phys * (8 * 8 + 24 + 6 * 8)
    raise ValueError(kernel)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n_sample=256, limiter="weno5z", steps=2, threads=None):
    """The oracle (torch CPU float64, the reference's algorithm and op order)
    on the same problem at n_sample^2: one untimed RK3 step, then ``steps``
    timed RK3 steps, at each thread count of ``threads`` (default: this
    process's share, torch.get_num_threads(), and the reference's production
    default of 8).  The fastest thread count is the reported value."""
    import torch
    import oracle.eos as OE
    import oracle.step2d as OS
    share = torch.get_num_threads()
    if threads is None:
        threads = sorted({share, 8})
    eos = OE.ideal(5.0 / 3.0)
    g = OS.make_grid(n_sample, (-0.5, 0.5), limiter)
    P, Az = OS.rotor_tapered(g, eos)
    st0 = OS.init_state(g, eos, P, Az)
    dt = OS.compute_dt_2d(st0.prims, g, eos, 0.4)
    runs = {}
    for T in threads:
        torch.set_num_threads(T)
        st, _ = OS.rk_step_ct(st0, g, eos, dt, limiter=limiter)   # warm-up
        t0 = time.perf_counter()
        for _ in range(steps):
            st, _ = OS.rk_step_ct(st, g, eos, dt, limiter=limiter)
        wall = time.perf_counter() - t0
        runs[T] = {"zone_cycles_per_s": n_sample * n_sample * steps / wall, "wall_s": wall}
    torch.set_num_threads(share)
    best = max(runs, key=lambda T: runs[T]["zone_cycles_per_s"])
    return {"value": runs[best]["zone_cycles_per_s"], "unit": "zone-cycles/s", "cores": best,
            "kind": "port", "by_threads": {str(T): r for T, r in runs.items()},
            "host": {"cpu_model": _cpu_model(), "os_cpu_count": os.cpu_count(),
                     "torch_threads_share": share, "torch": torch.__version__},
            "sample": (f"CPU oracle (torch {torch.__version__} float64, OMP/MKL threads as set, "
                       f"torch threads {threads}): the same tapered rotor, WENO5-Z/HLLD/CT, at "
                       f"{n_sample}^2, one warm-up then {steps} timed RK3 steps per thread count")}


def cpu_baseline_tubes(lengths, left, right, Bn, steps=2, limiter="mp5"):
    """The oracle's per-tube loop (torch CPU float64) on the first tube of
    the same workload for a few steps, timed."""
    import torch
    import oracle.eos as OE
    import oracle.tubes as OT
    threads = torch.get_num_threads()
    eos = OE.ideal(5.0 / 3.0)
    L = int(lengths[0])
    side = lambda s: {k: float(v[0]) for k, v in s.items()}  # noqa: E731
    P0 = OT.tube_initial(L, 3, eos, side(left), side(right), float(Bn[0]))
    t0 = time.perf_counter()
    OT.run_tube(P0, eos, L, 3, steps, cfl=0.4, limiter=limiter, scheme="rk3", flux="hlld")
    wall = time.perf_counter() - t0
    return {"value": L * steps / wall, "unit": "zone-cycles/s", "cores": threads, "kind": "port",
            "sample": (f"CPU oracle (torch {torch.__version__} float64, {threads} threads): "
                       f"{steps} MP5/HLLD/RK3 steps of tube 0 ({L} cells), the per-tube loop "
                       f"form ({wall:.1f} s)")}


def bench_tubes(a):
    """4096 ragged Riemann tubes, MP5 + HLLD + SSP-RK3, per-tube dt, the
    timed steps captured as one HIP graph of the autotuned form (the phased
    form's per-phase launches over every tube, or one persistent launch with
    a workgroup per tube).  Returns the result dict."""
    import torch
    from srmhd import eos as E, tubes as T
    dev = torch.device("cuda:0")
    eos = E.hybrid_eos(0.0, 5.0 / 3.0, 5.0 / 3.0)
    lengths, left, right, Bn = T.draw_tubes(a.tubes, seed=0, lmin=128, lmax=2048)
    batch = T.TubeBatch(lengths, left, right, Bn, eos, limiter=a.limiter, solver=a.solver,
                        device=dev)
    zones = int(lengths.sum())
    best, tune = T.autotune_block(batch, steps=2, cfl=a.cfl)
    print(f"autotune: block {best} ({tune})", file=sys.stderr, flush=True)
    snap = batch.snapshot()
    if a.warmup:
        batch.run(a.warmup, cfl=a.cfl)
        torch.cuda.synchronize()
        batch.restore(snap)
    graph = batch.capture(a.steps, cfl=a.cfl)
    torch.cuda.synchronize()
    print(f"warmup done ({a.warmup} steps), graph captured", file=sys.stderr, flush=True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    graph.replay()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1)
    zc = zones * a.steps / wall
    fin = torch.ones(batch.total_cells, dtype=torch.bool, device=dev)
    for p in batch.P:
        fin &= torch.isfinite(p)
    seg = torch.repeat_interleave(torch.arange(a.tubes, device=dev),
                                  (batch.off_host[1:] - batch.off_host[:-1]).to(dev))
    bad = torch.zeros(a.tubes, dtype=torch.int64, device=dev).index_add_(
        0, seg, (~fin).to(torch.int64))
    nonfinite = [int(t) for t in torch.nonzero(bad).reshape(-1).tolist()]
    # the per-tube loop on the GPU: the first tubes one launch each, same steps
    n_loop = min(a.loop_tubes, a.tubes)
    sub = [T.TubeBatch(lengths[i:i + 1], {k: v[i:i + 1] for k, v in left.items()},
                       {k: v[i:i + 1] for k, v in right.items()}, Bn[i:i + 1], eos,
                       limiter=a.limiter, solver=a.solver, device=dev) for i in range(n_loop)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b in sub:
        b.run(a.steps, cfl=a.cfl, block=best)
        torch.cuda.synchronize()
    loop_wall = time.perf_counter() - t0
    loop_zc = int(lengths[:n_loop].sum()) * a.steps / loop_wall
    # compulsory bytes per cell and RK stage, each phase reading and writing
    # its arrays once: HLLD pass A 64 in (P, By, Bz) + 24 (record) + 272 (the
    # side states and HLL speeds kept for pass B) out; pass B 296 in (side
    # states, record) + 56 out (7 fluxes); update 168 in (F, U0, U) + 56 out;
    # cell con2prim pass A 56 in + 24 out; pass B 80 in + 48 out (P) --
    # 1144 B; plus 64 B per cell and step for the CFL reduction
    tube_bytes = float(zones) * a.steps * (3 * 1144 + 64)
    out = {
        "metric": "zone-cycles/s, batched 1D Riemann tubes (4096 ragged tubes, MP5/HLLD/RK3)",
        "value": zc, "unit": "zone-cycles/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{a.tubes} Riemann tubes, lengths U[128, 2048] (seed 0, "
                               f"{zones} cells), {a.limiter.upper()} + {a.solver.upper()} + "
                               f"SSP-RK3, per-tube dt, CFL {a.cfl}, float64, the steps "
                               "captured as one HIP graph",
                   "tubes": a.tubes, "cells": zones, "block": best},
        "kernel_ms": kernel_ms,
        "roofline": {"bound": "hbm", "achieved": tube_bytes / (kernel_ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": tube_bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": (TUBES_TRAFFIC_PER_CELL_STEP[(a.limiter, a.solver)] * zones *
                                 a.steps if best == 0 and (a.limiter, a.solver) in
                                 TUBES_TRAFFIC_PER_CELL_STEP else None),
                     "kernel": ("the phased form's launches, every step" if best == 0 else
                                f"k_tubes (the persistent launch, block {best}, every step)"),
                     "avg_launch_ms": kernel_ms, "launches": 1,
                     "algorithmic_bytes_per_launch": tube_bytes},
        "per_tube_loop": {"value": loop_zc, "unit": "zone-cycles/s", "tubes_timed": n_loop,
                          "speedup_of_batch": zc / loop_zc},
        "autotune_ms": tune,
        # MP5 + HLLD blows up on 11 of the random Riemann problems of seed 0
        # (500 steps); each blows up at the same step on the oracle's
        # per-tube step, continued bit for bit from the device state
        # (tests/test_gpu_tubes.py)
        "nonfinite_tubes": {"count": len(nonfinite), "first": nonfinite[:16]},
    }
    if not a.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline_tubes(lengths, left, right, Bn, limiter=a.limiter)
        except Exception as exc:  # the baseline never hides the GPU result
            out["cpu_baseline"] = {"error": repr(exc)}
    return out


def _rotor_form(a, precision, kernel, U_init, P_init):
    """Warm up, then time a.steps RK3 steps of one precision form from the
    initial state; live HIP-event timing of one kernel class."""
    import ctypes
    import torch
    from srmhd import _lib
    from srmhd import driver2d as D2, eos as E

    eos = E.hybrid_eos(0.0, 5.0 / 3.0, 5.0 / 3.0)
    step = D2.Stepper2D(a.grid, eos, solver=a.solver, limiter=a.limiter, precision=precision)

    def one_step():
        step.compute_dt(step.U, step.P, a.cfl)
        step.step("rk3")

    step.load([t.clone() for t in U_init], [t.clone() for t in P_init])
    for _ in range(a.warmup):
        one_step()
    # the kernel class's own launch time: KERNEL_TIMING_STEPS steps from the
    # initial state with the sweeps one after the other (in the timed run
    # the FP64 form's y sweep shares the device with its x sweep)
    lib = step.lib
    step.load([t.clone() for t in U_init], [t.clone() for t in P_init])
    step.cs.flags = _lib.STAGE_SERIAL_SWEEPS
    ktime_steps = min(KERNEL_TIMING_STEPS, a.steps)
    _lib.check(lib.srmhd_timing_enable(_lib.KERNEL_CLASSES[kernel], 3 * ktime_steps),
               "srmhd_timing_enable")
    for _ in range(ktime_steps):
        one_step()
    torch.cuda.synchronize()
    ms = ctypes.c_double()
    cnt = ctypes.c_int64()
    _lib.check(lib.srmhd_timing_read(ctypes.byref(ms), ctypes.byref(cnt)), "srmhd_timing_read")
    lib.srmhd_timing_enable(-1, 0)
    step.cs.flags = 0
    step.load([t.clone() for t in U_init], [t.clone() for t in P_init])
    torch.cuda.synchronize()
    print(f"{precision}: warmup done ({a.warmup} steps, {ktime_steps} kernel-timing steps)",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        one_step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"{precision}: {a.steps} steps in {wall:.3f} s", file=sys.stderr, flush=True)
    avg_ms = ms.value / max(cnt.value, 1)
    n, ng = a.n, a.ng
    kb = algorithmic_bytes(kernel, n, ng, precision)
    nt = n + 2 * ng
    kelems = nt * nt if kernel.startswith("c2p") else (nt + 1) * nt
    achieved = kb / (avg_ms * 1e-3) / 1e9
    zc = n * n * a.steps / wall
    traffic = MEASURED_TRAFFIC.get((precision, kernel, n))
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
            "avg_launch_ms": avg_ms, "launches": cnt.value, "algorithmic_bytes_per_launch": kb}
    flops = FP64_FLOPS_PER_ELEMENT.get((precision, kernel))
    valu = None
    if flops is not None:
        tf = flops * kelems / (avg_ms * 1e-3) / 1e12
        valu = {"kernel": kernel, "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tf / FP64_PEAK_TFLOPS}
    insts = VALU_INSTS_PER_LAUNCH.get((precision, kernel, n))
    if insts is not None:
        roof["valu"] = {"insts_per_launch": insts,
                        "issue_rate": insts / (avg_ms * 1e-3), "unit": "wave-instructions/s",
                        "busy_measured": VALU_BUSY.get((precision, kernel, n)),
                        "busy_source": "rocprofv3 --pmc SQ_ACTIVE_INST_VALU GRBM_GUI_ACTIVE "
                                       "(profiles/rotor2048_counters.md)"}
    return {"step": step, "zc": zc, "wall": wall, "roofline": roof, "fp64_valu": valu,
            "ms_per_step": 1e3 * wall / a.steps}


def _rotor_l1_floor(a, U_init, P_init):
    """The FP64 form again, from the initial state with the conserved density
    D raised by one ulp (nextafter) in every cell: how far 200 steps of this
    problem carry a round-off-sized difference.  Untimed."""
    import torch
    from srmhd import driver2d as D2, eos as E

    eos = E.hybrid_eos(0.0, 5.0 / 3.0, 5.0 / 3.0)
    step = D2.Stepper2D(a.grid, eos, solver=a.solver, limiter=a.limiter, precision="fp64")
    U = [t.clone() for t in U_init]
    U[0] = torch.nextafter(U[0], torch.full_like(U[0], float("inf")))
    step.load(U, [t.clone() for t in P_init])
    for _ in range(a.gate_at):
        step.compute_dt(step.U, step.P, a.cfl)
        step.step("rk3")
    torch.cuda.synchronize()
    print(f"fp64 (D + 1 ulp): {a.gate_at} steps done", file=sys.stderr, flush=True)
    return step


def _rotor_fields(step, n, ng):
    """rho, p, |B| (cell-centred from the faces) over the physical cells and
    div B: max |div B| and max |div B| / (max |B_face| / dx)."""
    import torch
    sl = (slice(ng, ng + n), slice(ng, ng + n))
    Bxf, Byf, Bz = step.U[6], step.U[7], step.U[5]
    Bx = 0.5 * (Bxf[ng:ng + n, ng:ng + n] + Bxf[ng + 1:ng + n + 1, ng:ng + n])
    By = 0.5 * (Byf[ng:ng + n, ng:ng + n] + Byf[ng:ng + n, ng + 1:ng + n + 1])
    Bm = torch.sqrt(Bx * Bx + By * By + Bz[sl] * Bz[sl])
    dx, dy = step.grid.dx, step.grid.dy
    div = ((Bxf[ng + 1:ng + n + 1, ng:ng + n] - Bxf[ng:ng + n, ng:ng + n]) / dx
           + (Byf[ng:ng + n, ng + 1:ng + n + 1] - Byf[ng:ng + n, ng:ng + n]) / dy)
    dmax = float(div.abs().max())
    bmax = max(float(Bxf.abs().max()), float(Byf.abs().max()))
    return step.P[0][sl], step.P[4][sl], Bm, {"max_abs": dmax,
                                               "max_normalised": dmax / (bmax / dx)}


def _l1(a, b):
    """The reference's l1_rel (scripts/rotor_compare.py:57-58):
    mean |a - b| / mean |b|, b the FP64 form."""
    return float((a - b).abs().mean() / b.abs().mean())


def cpu_baseline_adjoint(n_sample=16, steps=2):
    """torch autograd through the CPU oracle (float64, IEEE arithmetic) on the
    same differentiable Orszag-Tang problem at n_sample^2, ``steps`` RK2
    steps with the first dt of the device run repeated: forward + backward,
    timed."""
    import torch
    import oracle.eos as OE
    import oracle.step2d as OS
    from oracle.grid import Grid2D
    threads = torch.get_num_threads()
    eos = OE.ideal(4.0 / 3.0)
    ng = 2
    g = Grid2D(0.0, 2 * math.pi, n_sample, 0.0, 2 * math.pi, n_sample, ng=ng)
    t0 = time.perf_counter()
    r = torch.ones(n_sample, n_sample, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
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
    c = {k: v.clone() for k, v in OS.prims_to_cons_2d(P, g).items()}
    st = OS.State2D({k: c[k] for k in OS.EVOLVED_KEYS}, a * ((Az0[:, 1:] - Az0[:, :-1]) / g.dy),
                    a * (-(Az0[1:, :] - Az0[:-1, :]) / g.dx), P)
    st = OS.sync_state(st, g, eos, "periodic", "periodic")
    dt = OS.compute_dt_2d(st.prims, g, eos, 0.4)
    for _ in range(steps):
        st, _ = OS.rk_step_ct(st, g, eos, dt, scheme="rk2", bc_x="periodic", bc_y="periodic",
                              limiter="mc")
    L = (st.prims["rho"][ng:ng + n_sample, ng:ng + n_sample] ** 2).sum()
    L.backward()
    wall = time.perf_counter() - t0
    return {"value": n_sample * n_sample * steps / wall, "unit": "zone-cycles/s (forward + backward)",
            "cores": threads, "kind": "port",
            "sample": (f"torch {torch.__version__} autograd through the CPU oracle (float64, "
                       f"{threads} threads): the same differentiable Orszag-Tang problem at "
                       f"{n_sample}^2, {steps} MC/HLLD/CT RK2 steps, forward + backward "
                       f"({wall:.1f} s)")}


def bench_adjoint(a):
    """Config 5: the gradient of L = sum (rho_final - rho_target)^2 of the
    relativistic Orszag-Tang problem (512^2 periodic, MC + HLLD + upwind CT
    + SSP-RK2, 200 steps, dt frozen from the A = 1 run, target from A = 1.05)
    with respect to the initial density field and the field amplitude A,
    through the hand-written adjoint with a checkpoint every 20 steps; dL/dA
    checked against central finite differences (h = 1e-6)."""
    import ctypes
    import torch
    from srmhd import _lib, adjoint as AD, eos as E, grid as G
    dev = torch.device("cuda:0")
    eos = E.hybrid_eos(0.0, 4.0 / 3.0, 4.0 / 3.0)
    n, steps, K = a.n, a.steps, a.checkpoint
    g = G.Grid2D(0.0, 2 * math.pi, n, 0.0, 2 * math.pi, n, ng=2, device=dev)
    roll = AD.Rollout2D(g, eos, solver="hlld", limiter="mc", bc_x="periodic", bc_y="periodic",
                        scheme="rk2", checkpoint_every=K)
    prob = AD.OTProblem(g, eos)
    rho0 = torch.ones(n, n, dtype=torch.float64, device=dev)
    one = torch.tensor(1.0, dtype=torch.float64, device=dev)
    U, P, _ = prob.initial_state(roll, rho0, one)
    dts = AD.frozen_dts(roll, U, P, steps, a.cfl)
    target = AD.ot_final_rho(prob, roll, rho0, one * 1.05, dts)
    torch.cuda.synchronize()
    print(f"dt sequence and target ready ({steps} steps)", file=sys.stderr, flush=True)

    def gradient(dt_seq):
        rho = rho0.clone().requires_grad_(True)
        A = one.clone().requires_grad_(True)
        rf = AD.OTRollout.apply(rho, A, prob, roll, dt_seq)
        L = ((rf - target[:rf.shape[0], :rf.shape[1]]) ** 2).sum()
        L.backward()
        return L.detach(), rho.grad, A.grad

    for _ in range(a.warmup):
        gradient(dts[:2])
    torch.cuda.synchronize()
    lib = roll.lib
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    L, g_rho, g_A = gradient(dts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    peak = torch.cuda.max_memory_allocated()
    print(f"gradient in {wall:.2f} s", file=sys.stderr, flush=True)
    # the x-sweep adjoint's own launch time, from the gradient of one
    # checkpoint segment with the sweeps in sequence (its records come from
    # the forward, so no re-run shares the device with the walk, and no
    # y sweep runs beside it, as they do in the run above)
    seg = min(K, steps)
    roll.st.cs.flags = _lib.STAGE_SERIAL_SWEEPS
    _lib.check(lib.srmhd_timing_enable(_lib.KERNEL_CLASSES["adj_sweep_x"], 2 * seg),
               "srmhd_timing_enable")
    gradient(dts[:seg])
    torch.cuda.synchronize()
    ms = ctypes.c_double()
    cnt = ctypes.c_int64()
    _lib.check(lib.srmhd_timing_read(ctypes.byref(ms), ctypes.byref(cnt)), "srmhd_timing_read")
    lib.srmhd_timing_enable(-1, 0)
    roll.st.cs.flags = 0
    t1 = time.perf_counter()
    h = 1e-6

    def loss(A):
        r = AD.ot_final_rho(prob, roll, rho0, A, dts)
        return float(((r - target) ** 2).sum())

    fd = (loss(one + h) - loss(one - h)) / (2 * h)
    fwd_wall = (time.perf_counter() - t1) / 2
    avg_ms = ms.value / max(cnt.value, 1)
    nfx = (n + 5) * (n + 4)
    # the x-sweep adjoint per face (its two launches: the flux adjoint and
    # the reconstruction adjoint): 8 state/primitive arrays of the stencil and
    # p* in, 7 flux adjoints in, 7 x 4 window adjoints + the B_n adjoint out
    # (the 14 face-state adjoints passed between the launches are not counted)
    kb = nfx * (8 * 8 + 8 + 7 * 8 + 29 * 8)
    achieved = kb / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": "zone-cycles/s of the differentiable rollout (forward + hand-written adjoint), "
                  "relativistic Orszag-Tang 512^2",
        "value": n * n * steps / wall, "unit": "zone-cycles/s", "n_gpus": 1, "steps": steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * wall / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"relativistic Orszag-Tang {n}^2 periodic, MC + HLLD + upwind CT "
                               f"+ SSP-RK2, {steps} steps, dt frozen from the A = 1 run, loss "
                               f"sum (rho - rho_target)^2 with the target from A = 1.05, "
                               f"gradient w.r.t. the initial rho field and A, checkpoint every "
                               f"{K} steps", "n": n, "checkpoint_every": K},
        "forward_only_ms_per_step": 1e3 * fwd_wall / steps,
        "adjoint_over_forward": wall / fwd_wall,
        "loss": float(L), "dLdA": {"adjoint": float(g_A), "central_fd_h1e-6": fd,
                                   "rel": abs(float(g_A) - fd) / abs(fd)},
        "grad_rho_l2": float(g_rho.norm()),
        "peak_device_memory_GB": peak / 1e9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": ADJ_SWEEP_TRAFFIC.get(n),
                     "kernel": "k_adj_flux<0, 1, 2> + k_adj_recon<0, 1> (x-sweep adjoint)",
                     "avg_launch_ms": avg_ms, "launches": cnt.value,
                     "algorithmic_bytes_per_launch": kb},
    }
    flops = ADJ_SWEEP_FLOPS_PER_FACE.get(n)
    if flops is not None:
        tf = flops * nfx / (avg_ms * 1e-3) / 1e12
        out["roofline"]["fp64_valu"] = {
            "flops_per_face": flops, "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tf / FP64_PEAK_TFLOPS,
            "source": "rocprofv3 --pmc SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 "
                      "(profiles/adjoint_ot512_counters.md)"}
    if not a.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline_adjoint()
        except Exception as exc:  # the baseline never hides the GPU result
            out["cpu_baseline"] = {"error": repr(exc)}
    return out


GATE_STEPS = 200   # the configuration's L1 / div B comparison point
L1_GATE = 1e-6


def _continue_to(a, step, done, total):
    """Advance a form's stepper from ``done`` to ``total`` RK3 steps (untimed)."""
    import torch
    for _ in range(max(total - done, 0)):
        step.compute_dt(step.U, step.P, a.cfl)
        step.step("rk3")
    torch.cuda.synchronize()


def bench_rotor(a):
    """The headline: the tapered rotor at a.n^2 in each precision form, timed
    over a.steps RK3 steps; L1 distances and div B after GATE_STEPS steps."""
    import torch
    from srmhd import eos as E, grid as G, initial_data2d as I
    from srmhd.reconstruction import ghosts_needed

    dev = torch.device("cuda:0")
    eos = E.hybrid_eos(0.0, 5.0 / 3.0, 5.0 / 3.0)
    a.ng = ng = ghosts_needed(a.limiter)
    a.grid = G.Grid2D(-0.5, 0.5, a.n, -0.5, 0.5, a.n, ng=ng, device=dev)
    P, Az = I.rotor_tapered(a.grid, eos)
    st = I.init_state(a.grid, eos, P, Az, limiter=a.limiter)
    U_init, P_init = st.state_list(), st.prim_list()
    del P, Az
    forms = tuple(f.strip() for f in a.precision.split(",") if f.strip())
    for f in forms:
        if f not in DOMINANT_KERNEL:
            raise SystemExit(f"unknown precision form {f!r}")
    if "fp64" not in forms:
        raise SystemExit("the fp64 form is the reference every other form is measured against")
    res = {}
    a.gate_at = max(GATE_STEPS, a.steps)
    for prec in forms:
        kernel = a.kernel or DOMINANT_KERNEL[prec]
        res[prec] = _rotor_form(a, prec, kernel, U_init, P_init)
        # the comparison point of the configuration, whatever was timed
        _continue_to(a, res[prec]["step"], a.steps, a.gate_at)
        print(f"{prec}: continued to {a.gate_at} steps", file=sys.stderr, flush=True)

    if a.copy_calibration:
        # an elementwise kernel of known traffic (torch routes copy_ through
        # the DMA engines, which the counters do not see): 512 MiB read,
        # 512 MiB written per launch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    fields = {p: _rotor_fields(r["step"], a.n, ng) for p, r in res.items()}
    finite = all(bool(torch.isfinite(fields[p][0]).all().item()) for p in fields)
    ra, pa, ba, _ = fields["fp64"]

    def l1_of_fields(fb):
        rb, pb, bb, _ = fb
        return {"rho": _l1(rb, ra), "p": _l1(pb, pa), "absB": _l1(bb, ba)}

    l1 = {f: l1_of_fields(fields[f]) for f in forms if f != "fp64"}
    meets = {f: l1[f]["rho"] <= L1_GATE for f in l1}
    # the headline is the fastest form that meets the gate; FP64 always does
    head = max([f for f in forms if f == "fp64" or meets[f]], key=lambda f: res[f]["zc"])
    h = res[head]
    bytes_per_zc = 752.0  # RK3 2D algorithmic bytes per zone-cycle
    form_desc = {"fp64": "float64 throughout (plain form, bit-faithful to the reference)",
                 "mixed": "reconstruction + Riemann-solver root finding in FP32; HLLD flux "
                          "assembly, con2prim, CT and RK in FP64",
                 "mixed_r64": "Riemann-solver root finding in FP32; reconstruction, HLLD flux "
                              "assembly, con2prim, CT and RK in FP64",
                 "fp64_local": "float64 throughout; every con2prim solved per element to its "
                               "own convergence instead of batch-coupled"}
    f64 = res["fp64"]
    out = {
        "metric": "zone-cycles/s, 2D rotor 2048^2 (FP64 vs FP32-flux @ L1<=1e-6); HBM GB/s",
        "value": h["zc"], "unit": "zone-cycles/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": h["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head in ("fp64", "fp64_local") else "f32+f64 (mixed)",
        "data": "synthetic",
        "config": {"workload": f"tapered relativistic rotor {a.n}^2, {a.limiter.upper()} + "
                               f"{a.solver.upper()} + CT + Kastaun C2P + SSP-RK3, CFL {a.cfl}, "
                               f"{a.steps} timed steps, {form_desc[head]}",
                   "n": a.n, "ng": ng, "limiter": a.limiter, "solver": a.solver, "scheme": "rk3",
                   "precision": head},
        "headline_form": head,
        "headline_rule": (f"the fastest form whose L1(rho) vs the FP64 form after {a.gate_at} "
                          f"steps is <= {L1_GATE:g} (the FP64 form itself qualifies)"),
        "roofline": h["roofline"],
        "fp64_valu": h["fp64_valu"],
        "step_hbm": {"bytes_per_zone_cycle": bytes_per_zc,
                     "achieved_GBs": h["zc"] * bytes_per_zc / 1e9,
                     "frac": h["zc"] * bytes_per_zc / 1e9 / HBM_PEAK_GBS},
        "gate_steps": a.gate_at,
        "div_b": fields[head][3],
        "state_finite": finite,
        "fp64_form": {"value": f64["zc"], "ms_per_step": f64["ms_per_step"],
                      "roofline": f64["roofline"], "fp64_valu": f64["fp64_valu"],
                      "div_b": fields["fp64"][3]},
        "other_forms": {
            form: {"value": res[form]["zc"], "ms_per_step": res[form]["ms_per_step"],
                   "speedup_vs_fp64": res[form]["zc"] / f64["zc"],
                   f"l1_vs_fp64_after_{a.gate_at}_steps": l1[form],
                   "l1_rho_gate_1e-6_met": meets[form], "div_b": fields[form][3],
                   "roofline": res[form]["roofline"], "description": form_desc[form]}
            for form in forms if form != "fp64"},
    }
    if a.tol_study:
        # the FP64 form with HLLD's p* tolerance (the reference's _PSTAR_TOL,
        # hlld.py:1080) tightened: how far the reference's own stopping rule
        # leaves the answer free
        from srmhd import driver2d as D2, hlld as H
        study = {}
        for tol in (1e-11, 1e-12, 1e-13):
            H._PSTAR_TOL = tol
            try:
                stp = D2.Stepper2D(a.grid, E.hybrid_eos(0.0, 5.0 / 3.0, 5.0 / 3.0),
                                   solver=a.solver, limiter=a.limiter, precision="fp64")
                stp.load([t.clone() for t in U_init], [t.clone() for t in P_init])
                _continue_to(a, stp, 0, a.gate_at)
                study[f"{tol:g}"] = l1_of_fields(_rotor_fields(stp, a.n, ng))
                del stp
            finally:
                H._PSTAR_TOL = 1e-10
            print(f"tol study {tol:g} done", file=sys.stderr, flush=True)
        out["pstar_tol_study"] = {
            f"l1_vs_fp64_after_{a.gate_at}_steps": study,
            "description": "the FP64 form with HLLD's p* residual tolerance tightened from the "
                           "reference's 1e-10: its distance from the FP64 form is the freedom "
                           "the reference's own stopping rule leaves"}
    if not a.no_floor:
        fl = _rotor_l1_floor(a, U_init, P_init)
        out["l1_floor"] = {
            f"l1_vs_fp64_after_{a.gate_at}_steps": l1_of_fields(_rotor_fields(fl, a.n, ng)),
            "description": "the FP64 form from the initial state with D raised by one ulp in "
                           "every cell: the distance round-off alone reaches on this problem"}
        del fl
    for r in res.values():
        del r["step"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="all", choices=("all", "rotor", "tubes", "adjoint"),
                    help="all (default): the 2048^2 rotor headline, with the batched tubes and "
                         "the differentiable Orszag-Tang rollout as keys of the same line; "
                         "rotor / tubes / adjoint: one workload alone")
    ap.add_argument("--checkpoint", type=int, default=20,
                    help="adjoint workload: steps between checkpoints")
    ap.add_argument("--precision", default="fp64,mixed,mixed_r64",
                    help="comma-separated rotor forms to time (fp64 required; mixed, "
                         "mixed_r64, fp64_local); each is compared with fp64 after 200 steps")
    ap.add_argument("--tubes", type=int, default=4096)
    ap.add_argument("--loop-tubes", type=int, default=None,
                    help="tubes timed one launch each for the per-tube-loop comparison "
                         "(default 16 alone, 8 with --workload all)")
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps of the workload (rotor default 200; with --workload "
                         "all the tubes and adjoint keys keep their configurations' 500 and 200)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--limiter", default=None)
    ap.add_argument("--solver", default="hlld")
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--kernel", default=None,
                    help="kernel class timed live for the roofline (default: the dominant one "
                         "of each form)")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tol-study", action="store_true",
                    help="rotor: also run the FP64 form with HLLD's p* tolerance at 1e-11, "
                         "1e-12, 1e-13 and report each one's L1 from the FP64 form")
    ap.add_argument("--no-floor", action="store_true",
                    help="skip the untimed FP64 run from the one-ulp-perturbed initial state")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed region, copy a 512 MiB float64 buffer three times "
                         "with an elementwise kernel (the plain copy kernel that scales the "
                         "FETCH/WRITE counters)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (the step would "
              "shard as x-slabs with halo exchange; only the one-GPU kernel is built)",
              file=sys.stderr)
        sys.exit(2)

    def sub(workload, steps, n, limiter, loop_tubes):
        b = argparse.Namespace(**vars(a))
        b.steps = steps
        b.n = n
        b.limiter = limiter
        b.loop_tubes = loop_tubes
        return b

    lim = a.limiter
    if a.workload == "tubes":
        out = bench_tubes(sub("tubes", a.steps or 500, a.n, lim or "mp5", a.loop_tubes or 16))
    elif a.workload == "adjoint":
        out = bench_adjoint(sub("adjoint", a.steps or 200, a.n or 512, lim, 0))
    else:
        out = bench_rotor(sub("rotor", a.steps or GATE_STEPS, a.n or 2048, lim or "weno5z", 0))
        if not a.no_cpu:
            try:
                out["cpu_baseline"] = cpu_baseline(a.cpu_sample, lim or "weno5z")
            except Exception as exc:  # the baseline never hides the GPU result
                out["cpu_baseline"] = {"error": repr(exc)}
        if a.workload == "all":
            # the other two configurations, each with its own roofline and CPU
            # baseline, at their own step counts
            out["tubes"] = bench_tubes(sub("tubes", 500, None, "mp5", a.loop_tubes or 8))
            print("tubes done", file=sys.stderr, flush=True)
            out["adjoint"] = bench_adjoint(sub("adjoint", 200, 512, None, 0))
            print("adjoint done", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
