"""The 1D driver (driver.run) on the HIP path: Balsara 1 (config 1 of the
workload list) bitwise against the IEEE oracle, and against the reference's
400-step run within the north-star bound (L1 of rho <= 1e-10)."""

import os
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

PRIM9 = ("rho", "vx", "vy", "vz", "p", "eps", "Bx", "By", "Bz")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _cfg(solver, limiter, t_end, tmp_path):
    import oracle.step1d as S1
    cfg = S1.balsara1_cfg(solver=solver, limiter=limiter)
    cfg["run"]["t_end"] = t_end
    cfg["run"]["device"] = "cuda"
    cfg["output"] = {"dir": str(tmp_path / "out"), "every_n_steps": 1000}
    return cfg


@pytest.mark.parametrize("solver,limiter", [("hlld", "mc"), ("hllc", "mc"), ("hlle", "pcm"),
                                            ("hlld", "minmod")])
def test_balsara1_bitwise_vs_oracle(solver, limiter, tmp_path):
    _dev()
    import oracle.step1d as S1
    from srmhd import driver as D
    print(f"1d {solver} {limiter}", flush=True)
    cfg = _cfg(solver, limiter, 0.03, tmp_path)
    t, grid, P, C = D.run(cfg)
    to, go, Po, Co, steps = S1.run(cfg)
    assert t == to and steps == 30
    for k in PRIM9:
        a, b = P[k].cpu().numpy(), Po[k].numpy()
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        assert same.all(), f"{k}: {int((~same).sum())} differ"


def test_balsara1_400_steps_close_to_reference(golden_dir, tmp_path):
    _dev()
    from srmhd import driver as D
    print("1d balsara1 hlld mc 400 steps vs reference", flush=True)
    g = np.load(os.path.join(golden_dir, "balsara1.npz"))
    t, grid, P, C = D.run(_cfg("hlld", "mc", 0.4, tmp_path))
    assert t == float(g["hlld_t"])
    rho, ref = P["rho"].cpu().numpy(), g["hlld_prim_rho"]
    l1 = np.abs(rho - ref).mean() / np.abs(ref).mean()
    print(f"L1(rho) vs reference = {l1:.3e}", flush=True)
    assert l1 <= 1e-10
    assert os.path.exists(tmp_path / "out" / "snap_000000.npz")
    assert os.path.exists(tmp_path / "out" / "snap_000400.npz")
