"""CPU oracle for the SRMHD finite-volume step (TEST INFRASTRUCTURE ONLY).

This package restates, in float64 PyTorch on the CPU, the algorithm of the
reference solver kemi2805/HLLD-AI for exactly the hot path that the HIP
library accelerates: EOS, primitive<->conserved maps, HLLE/HLLC/HLLD fluxes,
Kastaun con2prim (with its batch-coupled secant), PCM/PLM/WENO5-Z/MP5/MP7
reconstruction, the directional sweep, Gardiner-Stone constrained transport,
the SSP-RK2/3 stage, the 1D driver and the batched tube runner.

Every function cites the reference file:line it follows.  Arithmetic is
written in the same association order as the reference's torch expressions
(no fused multiply-add exists in element-wise torch CPU ops), so the oracle
is bit-faithful to the reference and can be pinned bitwise against the
golden vectors under ``tests/golden``.

Who may use it: ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` -- as the checker and as the timed CPU
baseline only.  The product path (``hlld-ai_amd/srmhd``) never imports it.
"""

from . import eos, physics, c2p, riemann, recon, grid, step2d, step1d, tubes  # noqa: F401
