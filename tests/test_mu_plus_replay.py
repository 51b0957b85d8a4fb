"""mu_plus_fast() (the mu_+ bisection with the decisions far from the root
taken without evaluating fa) against the reference's 50-step bisection,
c2p.py:162-184, bit for bit: the device header compiled for the host with
hipcc, 300 000 random conserved states over the whole admissible range and
beyond.  CPU only (no GPU code runs)."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_mu_plus_fast_is_bit_identical(tmp_path):
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "mu_plus_check")
    src = os.path.join(HERE, "native", "mu_plus_check.cpp")
    cmd = [HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "hlld-ai_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           "-o", exe, src]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([exe, "300000", "20261203"], check=False, capture_output=True,
                         text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0" in out.stdout
