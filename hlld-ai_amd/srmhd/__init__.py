"""srmhd: the SRMHD finite-volume step of kemi2805/HLLD-AI on MI355X (gfx950).

Python host code mirroring the reference's plugin and driver interfaces
(``hlld.hlld_flux`` / ``hllc_flux`` / ``hlle_flux`` with ``LAST_DIAG``,
``c2p.conservative_to_primitive``, ``driver2d.rk_step_ct`` and friends,
``driver.run``), calling hand-written HIP kernels through the C ABI in
include/srmhd.h.  There is no CPU fallback.
"""

from . import _lib, eos, hlld, c2p, grid, state, driver, driver2d, initial_data2d  # noqa: F401

__all__ = ["eos", "hlld", "c2p", "grid", "state", "driver", "driver2d", "initial_data2d"]
