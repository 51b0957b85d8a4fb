"""Reconstruction metadata of the HIP path (reconstruction.py:85-98 of the
reference): stencil half-widths and ghost counts.  The reconstruction itself
runs inside the fused sweep kernels (csrc/srmhd_recon.h)."""

STENCIL = {"pcm": 0, "mc": 1, "minmod": 1, "weno5z": 2, "mp5": 2, "mp7": 3}
HIGH_ORDER = ("weno5z", "mp5", "mp7")


def ghosts_needed(limiter):
    if limiter not in STENCIL:
        raise ValueError(f"unknown limiter {limiter!r}")
    return max(2, STENCIL[limiter] + 1)
