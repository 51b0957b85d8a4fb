"""State container (src/physics/state.py:41-74)."""

from dataclasses import dataclass, field

import torch

EVOLVED_KEYS = ["D", "Sx", "Sy", "Sz", "tau", "Bz"]
PRIM6 = ("rho", "vx", "vy", "vz", "p", "eps")


@dataclass
class State2D:
    """Cell-centred conserved variables + staggered B (+ primitives)."""

    cons: dict
    Bxf: torch.Tensor
    Byf: torch.Tensor
    prims: dict = field(default_factory=dict)

    def clone(self):
        return State2D({k: v.clone() for k, v in self.cons.items()}, self.Bxf.clone(),
                       self.Byf.clone(), {k: v.clone() for k, v in self.prims.items()})

    def state_list(self):
        """The C-ABI order: D, Sx, Sy, Sz, tau, Bz, Bxf, Byf."""
        return [self.cons[k] for k in EVOLVED_KEYS] + [self.Bxf, self.Byf]

    def prim_list(self):
        return [self.prims[k] for k in PRIM6]
