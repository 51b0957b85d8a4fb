"""The host-side 2D initial data (srmhd.initial_data2d, evaluated on the CPU
like the reference's) against the reference's own outputs
(tests/golden/init2d.npz, written by make_golden.py gen_init2d): bitwise."""

import os

import numpy as np
import pytest
import torch

GOLD = os.path.join(os.path.dirname(__file__), "golden", "init2d.npz")


@pytest.mark.parametrize("name,problem,box,gamma,kw", [
    ("ot", "orszag_tang", (0.0, 1.0), 4.0 / 3.0, {}),
    ("otbs", "orszag_tang_bs", (0.0, 1.0), 5.0 / 3.0, {}),
    ("r2d", "riemann2d", (-0.5, 0.5), 5.0 / 3.0, {}),
    ("r2dtilt", "riemann2d", (-0.5, 0.5), 5.0 / 3.0, {"B0": 0.7, "angle_z_deg": 30.0}),
])
def test_initial_data_bitwise_vs_reference(name, problem, box, gamma, kw):
    from srmhd import eos as E, grid as G, initial_data2d as I
    gold = np.load(GOLD)
    eos = E.hybrid_eos(0.0, gamma, gamma)
    g = G.Grid2D(box[0], box[1], 24, box[0], box[1], 24, ng=2, device="cpu")
    P, Az = I.PROBLEMS[problem](g, eos, **kw)
    for k in ("rho", "p", "vx", "vy", "vz", "Bx", "By", "Bz", "eps"):
        assert np.array_equal(P[k].numpy(), gold[f"{name}_{k}"]), k
    assert np.array_equal(Az.numpy(), gold[f"{name}_Az"])
