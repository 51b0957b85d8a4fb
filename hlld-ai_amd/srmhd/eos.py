"""The reference's EOS object (src/physics/eos.py:15-20), as the parameters the
device code reads.  Host-side helpers are element-wise torch on whatever
device the tensors live on; the hot path evaluates the EOS inside the HIP
kernels."""

import torch

from ._lib import Eos


class hybrid_eos:  # noqa: N801  (the reference's class name: a drop-in argument)
    def __init__(self, K, gamma, gamma_th=None):
        self.K = K
        self.gamma = gamma
        self.gamma_th = gamma_th
        self.is_cold = gamma_th is None or gamma_th == 1.0

    def to_c(self):
        return Eos(float(self.K), float(self.gamma),
                   float(self.gamma if self.gamma_th is None else self.gamma_th),
                   1 if self.is_cold else 0, 0)

    def press_cold_eps_cold__rho(self, rho):
        pc = self.K * rho ** self.gamma
        return pc, pc / (rho * (self.gamma - 1))

    def press__eps_rho(self, eps, rho):
        pc, ec = self.press_cold_eps_cold__rho(rho)
        eps = torch.maximum(eps, ec)
        if self.is_cold:
            return pc
        return pc + (eps - ec) * rho * (self.gamma_th - 1)

    def eps__press_rho(self, press, rho):
        pc, ec = self.press_cold_eps_cold__rho(rho)
        if self.is_cold:
            return ec
        return ec + (press - pc) / (rho * (self.gamma_th - 1))

    def press_and_cs2(self, eps, rho):
        p = self.press__eps_rho(eps, rho)
        h = 1.0 + eps + p / rho
        if self.is_cold:
            return p, self.gamma * p / (rho * h)
        pc, _ = self.press_cold_eps_cold__rho(rho)
        return p, (self.gamma * pc + self.gamma_th * (p - pc)) / (rho * h)


def as_c(eos):
    """Accept this class, the reference's hybrid_eos or any object with
    K / gamma / gamma_th attributes."""
    if hasattr(eos, "to_c"):
        return eos.to_c()
    gth = getattr(eos, "gamma_th", None)
    cold = gth is None or gth == 1.0
    return Eos(float(eos.K), float(eos.gamma), float(eos.gamma if gth is None else gth),
               1 if cold else 0, 0)
