"""srmhd.run_2d, the HIP-path scripts/run_2d.py: output files with the
reference's schema, and a restarted run bit-identical to the uninterrupted
one (the reference's own restart property, tests/test_high_order_2d.py:101-126)."""

import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


ARGS = ["--problem", "rotor", "--n", "32", "--tend", "0.02", "--cfl", "0.4", "--log-every", "2",
        "--nsnap", "2"]


@pytest.mark.parametrize("limiter", ["mc", "weno5z"])
def test_restart_is_bitwise_and_files_have_the_reference_schema(tmp_path, limiter):
    _dev()
    from srmhd import run_2d
    print(f"run_2d rotor 32^2 {limiter}: uninterrupted vs restarted", flush=True)
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    run_2d.main(ARGS + ["--limiter", limiter, "--out", a, "--restart-every", "0"])
    run_2d.main(ARGS + ["--limiter", limiter, "--out", b, "--max-steps", "3"])
    assert os.path.exists(os.path.join(b, "restart.npz"))
    assert not os.path.exists(os.path.join(b, "snap_fin.npz"))
    run_2d.main(ARGS + ["--limiter", limiter, "--out", b, "--restart",
                        os.path.join(b, "restart.npz")])
    fa = np.load(os.path.join(a, "snap_fin.npz"))
    fb = np.load(os.path.join(b, "snap_fin.npz"))
    assert sorted(fa.files) == sorted(["t", "x", "y", "rho", "p", "vx", "vy", "Bx", "By", "divB",
                                       "limiter", "ng", "n", "solver"])
    for k in fa.files:
        assert np.array_equal(fa[k], fb[k]), k
    la = open(os.path.join(a, "diag.csv")).read().splitlines()
    lb = open(os.path.join(b, "diag.csv")).read().splitlines()
    assert la[0] == run_2d.DIAG_HEADER.strip()
    assert la == lb
    assert len(la) > 2
    cols = la[-1].split(",")
    assert len(cols) == 16
    assert float(cols[3]) < 1e-10          # div B max
    assert int(cols[13]) == 0              # c2p_bad
    r = np.load(os.path.join(b, "restart.npz"))
    for k in ("t", "step", "next_snap", "gen", "limiter", "ng", "n", "solver", "Bxf", "Byf",
              "cons_D", "cons_Sx", "cons_Sy", "cons_Sz", "cons_tau", "cons_Bz"):
        assert k in r.files, k
    assert os.path.exists(os.path.join(a, "snap_000.npz"))
    assert os.path.exists(os.path.join(a, "snap_001.npz"))


def test_refuses_to_overwrite_a_different_run(tmp_path):
    _dev()
    from srmhd import run_2d
    a = str(tmp_path / "a")
    run_2d.main(ARGS + ["--out", a, "--tend", "0.002"])
    with pytest.raises(SystemExit):
        run_2d.main(ARGS + ["--out", a, "--tend", "0.002", "--limiter", "weno5z"])
