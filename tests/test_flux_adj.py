"""The reverse-mode vector-Jacobian product of the HLLD face flux
(srmhd_flux_adj.h: flux_vjp_hlld, used by the adjoint's sweeps) against the
forward-mode one (flux_vjp_dual: one dual pass per face input) on 20 000
random faces of both normal directions at the root p* of the forward solver
(HLLE-fallback faces included): the device header compiled for the host.
CPU only (no GPU code runs)."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_reverse_mode_hlld_vjp_equals_forward_mode(tmp_path):
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "flux_adj_check")
    src = os.path.join(HERE, "native", "flux_adj_check.cpp")
    cmd = [HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "hlld-ai_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           "-o", exe, src]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([exe, "20000", "20261204"], check=False, capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "bad 0" in out.stdout
