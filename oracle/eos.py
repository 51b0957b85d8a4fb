"""Hybrid cold-polytrope + thermal ideal-gas EOS (oracle restatement).

Follows src/physics/eos.py:15-76 of the reference.  With K = 0 and
gamma_th = gamma (every configuration of the hot path) it is an ideal gas,
but the cold part ``K * rho**gamma`` is still evaluated, exactly as the
reference does (eos.py:23,49).
"""

import torch


class HybridEOS:
    """p = K rho^g + (eps - eps_cold) rho (g_th - 1)   (eos.py:15-20)."""

    def __init__(self, K, gamma, gamma_th=None):
        self.K = K
        self.gamma = gamma
        self.gamma_th = gamma_th
        self.is_cold = gamma_th is None or gamma_th == 1.0

    def cold(self, rho):
        """(p_cold, eps_cold)   eos.py:22-25."""
        pc = self.K * rho ** self.gamma
        return pc, pc / (rho * (self.gamma - 1))

    # the reference's method names are kept: the drop-in API passes this
    # object where the reference passes ``hybrid_eos``
    def press_cold_eps_cold__rho(self, rho):
        return self.cold(rho)

    def press__eps_rho(self, eps, rho):
        """eos.py:32-39."""
        pc, ec = self.cold(rho)
        eps = torch.maximum(eps, ec)
        if self.is_cold:
            return pc
        return pc + (eps - ec) * rho * (self.gamma_th - 1)

    def eps__press_rho(self, press, rho):
        """eos.py:41-46."""
        pc, ec = self.cold(rho)
        if self.is_cold:
            return ec
        return ec + (press - pc) / (rho * (self.gamma_th - 1))

    def eps_range__rho(self, rho):
        """eos.py:48-51: eps in [eps_cold, 1e5]."""
        pc = self.K * rho ** self.gamma
        return pc / (rho * (self.gamma - 1)), torch.full_like(rho, 1e5)

    def press_and_cs2(self, eps, rho):
        """eos.py:65-76."""
        p = self.press__eps_rho(eps, rho)
        h = 1.0 + eps + p / rho
        if self.is_cold:
            cs2 = self.gamma * p / (rho * h)
        else:
            pc, _ = self.cold(rho)
            cs2 = (self.gamma * pc + self.gamma_th * (p - pc)) / (rho * h)
        return p, cs2


def ideal(gamma):
    """The EOS every configuration uses: K = 0, gamma_th = gamma
    (scripts/run_2d.py:213)."""
    return HybridEOS(0.0, gamma, gamma)
