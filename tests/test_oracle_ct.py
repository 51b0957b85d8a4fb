"""The oracle's Gardiner-Stone corner EMF (oracle.step2d.corner_emf,
restating src/physics/ct.py:119-177) against known answers that do not
restate it:

* upwind=False is the plain four-face average (the reference's own test,
  tests/test_ct.py:173-182) -- and on a field linear in x and y that average
  is the exact corner value, whatever the upwind signs: a wrong face index
  moves it by O(dx);
* E depending on one coordinate only reduces the upwinded corner EMF to the
  face value of that coordinate (tests/test_ct.py:94-137), for either sign
  of the mass flux and for zero flux;
* with zero mass flux the upwind weights are all one, and the correction is
  the centred transverse-derivative form (tests/test_ct.py:140-166).
"""

import pytest
import torch

from oracle import step2d as S2

NXT, NYT = 11, 9


def _gen(seed):
    return torch.Generator().manual_seed(seed)


def test_upwind_false_is_the_four_face_average():
    g = _gen(1)
    Ec = torch.randn(NXT, NYT, generator=g, dtype=torch.float64)
    Efx = torch.randn(NXT + 1, NYT, generator=g, dtype=torch.float64)
    Efy = torch.randn(NXT, NYT + 1, generator=g, dtype=torch.float64)
    Fx = torch.randn(NXT + 1, NYT, generator=g, dtype=torch.float64)
    Fy = torch.randn(NXT, NYT + 1, generator=g, dtype=torch.float64)
    Ez = S2.corner_emf(Ec, Efx, Efy, Fx, Fy, upwind=False)
    want = 0.25 * ((Efy[1:NXT, 1:NYT] + Efy[0:NXT - 1, 1:NYT])
                   + (Efx[1:NXT, 1:NYT] + Efx[1:NXT, 0:NYT - 1]))
    assert torch.equal(Ez[1:NXT, 1:NYT], want)
    # the boundary ring stays zero
    assert float(Ez[0].abs().max()) == 0.0 and float(Ez[:, 0].abs().max()) == 0.0
    assert float(Ez[NXT].abs().max()) == 0.0 and float(Ez[:, NYT].abs().max()) == 0.0


@pytest.mark.parametrize("upwind", [False, True])
def test_linear_field_gives_the_exact_corner_value(upwind):
    """E = a x + b y sampled where each array lives (cells at (x_i, y_j),
    x faces at (x_{i-1/2}, y_j), y faces at (x_i, y_{j-1/2})): the corner
    (x_{i-1/2}, y_{j-1/2}) value is exact for any mass-flux signs."""
    a, b, dx, dy = 0.7, -1.3, 0.1, 0.25
    xc = (torch.arange(NXT, dtype=torch.float64) + 0.5) * dx
    yc = (torch.arange(NYT, dtype=torch.float64) + 0.5) * dy
    xf = torch.arange(NXT + 1, dtype=torch.float64) * dx
    yf = torch.arange(NYT + 1, dtype=torch.float64) * dy
    Ec = a * xc[:, None] + b * yc[None, :]
    Efx = a * xf[:, None] + b * yc[None, :]
    Efy = a * xc[:, None] + b * yf[None, :]
    g = _gen(2)
    Fx = torch.randn(NXT + 1, NYT, generator=g, dtype=torch.float64)
    Fy = torch.randn(NXT, NYT + 1, generator=g, dtype=torch.float64)
    Ez = S2.corner_emf(Ec, Efx, Efy, Fx, Fy, upwind=upwind)
    want = a * xf[1:NXT, None] + b * yf[None, 1:NYT]
    assert float((Ez[1:NXT, 1:NYT] - want).abs().max()) < 1e-14


@pytest.mark.parametrize("sgn", [-1.0, 0.0, 1.0])
@pytest.mark.parametrize("axis", ["y", "x"])
def test_one_coordinate_dependence_reduces_to_the_face_value(sgn, axis):
    g = _gen(3)
    if axis == "y":
        e = torch.randn(NYT, generator=g, dtype=torch.float64)
        Ec = e.unsqueeze(0).expand(NXT, NYT).contiguous()
        Efx = e.unsqueeze(0).expand(NXT + 1, NYT).contiguous()
        Efy = torch.randn(NYT + 1, generator=g, dtype=torch.float64).unsqueeze(0).expand(
            NXT, NYT + 1).contiguous()
    else:
        e = torch.randn(NXT, generator=g, dtype=torch.float64)
        Ec = e.unsqueeze(1).expand(NXT, NYT).contiguous()
        Efy = e.unsqueeze(1).expand(NXT, NYT + 1).contiguous()
        Efx = torch.randn(NXT + 1, generator=g, dtype=torch.float64).unsqueeze(1).expand(
            NXT + 1, NYT).contiguous()
    Fx = torch.full((NXT + 1, NYT), sgn, dtype=torch.float64)
    Fy = torch.full((NXT, NYT + 1), sgn, dtype=torch.float64)
    Ez = S2.corner_emf(Ec, Efx, Efy, Fx, Fy, upwind=True)
    want = (Efy if axis == "y" else Efx)[1:NXT, 1:NYT]
    err = float((Ez[1:NXT, 1:NYT] - want).abs().max()) / float(want.abs().max())
    assert err < 1e-14


def test_zero_mass_flux_is_the_centred_correction():
    g = _gen(4)
    Ec = torch.randn(NXT, NYT, generator=g, dtype=torch.float64)
    Efx = torch.randn(NXT + 1, NYT, generator=g, dtype=torch.float64)
    Efy = torch.randn(NXT, NYT + 1, generator=g, dtype=torch.float64)
    up = S2.corner_emf(Ec, Efx, Efy, torch.zeros(NXT + 1, NYT, dtype=torch.float64),
                       torch.zeros(NXT, NYT + 1, dtype=torch.float64), upwind=True)
    I, Im, J, Jm = slice(1, NXT), slice(0, NXT - 1), slice(1, NYT), slice(0, NYT - 1)
    yf, yfm, xf, xfm = Efy[I, J], Efy[Im, J], Efx[I, J], Efx[I, Jm]
    NE, NW, SE, SW = Ec[I, J], Ec[Im, J], Ec[I, Jm], Ec[Im, Jm]
    avg = 0.25 * ((yf + yfm) + (xf + xfm))
    dEdy = 0.125 * ((yf - SE) + (yfm - SW) - (NE - yf) - (NW - yfm))
    dEdx = 0.125 * ((xf - NW) + (xfm - SW) - (NE - xf) - (SE - xfm))
    assert float((up[I, J] - (avg + (dEdy + dEdx))).abs().max()) <= 1e-15
