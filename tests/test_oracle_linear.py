"""The oracle's linearised flux (oracle/linear.py, restating
src/physics/linearised.py:129-237) pinned two ways:

* bitwise against the reference's own outputs (tests/golden/linear_flux.npz,
  tests/golden/make_golden.py gen_linear): the flux_random pairs with their
  jumps scaled by 1, 0.1 and 0.01, every direction -- F, U, p*, the mask of
  answered faces and the answered count;
* against what the mathematics fixes, independently of the formula's text:
  a zero jump is the physical flux (the reference's own test,
  tests/test_linearised.py:81-87); the same face swept along x or y gives the
  same flux (tests/test_linearised.py:90-103); a face whose waves all run
  right takes the left state's flux; a jump along one eigenvector r_k of the
  mean state's A crosses to the right state exactly when lambda_k < 0; and
  the central-difference Jacobians agree with autograd of the side-state
  arithmetic to the truncation error.
"""

import os

import numpy as np
import pytest
import torch

from oracle import eos as E
from oracle import linear as L
from oracle import physics as PH

PRIM9 = ("rho", "vx", "vy", "vz", "p", "eps", "Bx", "By", "Bz")
CONS8 = ("D", "tau", "Sx", "Sy", "Sz", "Bx", "By", "Bz")
EOS = E.ideal(5.0 / 3.0)


@pytest.fixture
def reference_sqrt(golden_dir):
    """torch's CPU sqrt, as the reference ran, where this host rounds like
    the fixtures' host (the bitwise comparison needs it)."""
    path = os.path.join(golden_dir, "host_sqrt.npz")
    if not os.path.exists(path):
        pytest.skip("host_sqrt.npz not generated")
    h = np.load(path)
    if not (torch.sqrt(torch.from_numpy(h["x"])).numpy() == h["torch_sqrt"]).all():
        pytest.skip("this host's torch.sqrt rounds differently from the fixtures' host")
    PH.set_sqrt_mode("torch")
    yield
    PH.set_sqrt_mode("ieee")


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    assert same.all(), f"{what}: {int((~same).sum())}/{a.size} differ"


@pytest.mark.parametrize("si", [0, 1, 2])
@pytest.mark.parametrize("tag", ["a", "b"])
def test_linear_flux_bitwise_vs_reference(golden_dir, reference_sqrt, tag, si):
    g = np.load(os.path.join(golden_dir, "linear_flux.npz"))
    sL = {k: torch.from_numpy(g[f"{tag}{si}_L_{k}"]) for k in PRIM9}
    sR = {k: torch.from_numpy(g[f"{tag}{si}_R_{k}"]) for k in PRIM9}
    for idir in (0, 1, 2):
        F, U, ps, ok = L.linearised_flux(sL, sR, EOS, idir=idir, return_mask=True)
        pre = f"{tag}{si}_{idir}"
        for k in CONS8:
            _same(F[k].numpy(), g[f"{pre}_F_{k}"], f"{pre} F[{k}]")
            _same(U[k].numpy(), g[f"{pre}_U_{k}"], f"{pre} U[{k}]")
        _same(ps.numpy(), g[f"{pre}_pstar"], f"{pre} p*")
        assert (ok.numpy() == g[f"{pre}_ok"]).all()
        assert L.LAST_LINEAR["n_linear"] == int(g[f"{pre}_n_linear"])


def _states(n, seed, vn=None):
    g = torch.Generator().manual_seed(seed)
    u = lambda lo, hi: lo + (hi - lo) * torch.rand(n, generator=g, dtype=torch.float64)
    s = {"rho": u(0.5, 2.0), "p": u(0.2, 2.0), "vx": u(-0.3, 0.3), "vy": u(-0.3, 0.3),
         "vz": u(-0.3, 0.3), "Bx": u(-1.0, 1.0), "By": u(-1.0, 1.0), "Bz": u(-1.0, 1.0)}
    if vn is not None:
        s["vx"] = torch.full((n,), vn, dtype=torch.float64)
        s["vy"] = u(-0.05, 0.05)
        s["vz"] = u(-0.05, 0.05)
    s["eps"] = EOS.eps__press_rho(s["p"], s["rho"])
    return s


def test_zero_jump_is_the_physical_flux():
    s = _states(64, 1)
    for idir in (0, 1, 2):
        F, U, ps, ok = L.linearised_flux(s, s, EOS, idir=idir, return_mask=True)
        assert ok.all()
        u, f, _, _ = PH.side_state(s, EOS, idir)
        for k in CONS8:
            assert torch.equal(F[k], f[k]), (idir, k)
            assert torch.equal(U[k], u[k]), (idir, k)


def test_the_same_face_along_x_or_y():
    a, b = _states(48, 2), _states(48, 3)
    sw = {"rho": "rho", "p": "p", "eps": "eps", "vx": "vy", "vy": "vx", "vz": "vz",
          "Bx": "By", "By": "Bx", "Bz": "Bz"}
    Fx = L.linearised_flux(a, b, EOS, 0)[0]
    Fy = L.linearised_flux({sw[k]: v for k, v in a.items()}, {sw[k]: v for k, v in b.items()},
                           EOS, 1)[0]
    fmap = {"D": "D", "tau": "tau", "Sx": "Sy", "Sy": "Sx", "Sz": "Sz", "Bx": "By",
            "By": "Bx", "Bz": "Bz"}
    for k in CONS8:
        x, y = Fx[k], Fy[fmap[k]]
        assert torch.allclose(x, y, rtol=1e-12, atol=1e-14), k


def test_all_waves_to_the_right_take_the_left_flux():
    """v_n = 0.97 with a weak field: every characteristic speed of the mean
    state is positive, so V* = V_L and the flux is V_L's own."""
    a = _states(32, 4, vn=0.97)
    b = {k: v.clone() for k, v in a.items()}
    for k in ("rho", "p"):
        b[k] = a[k] * 1.01
    b["eps"] = EOS.eps__press_rho(b["p"], b["rho"])
    for s in (a, b):
        for k in ("Bx", "By", "Bz"):
            s[k] = 0.05 * s[k]
    b["Bx"] = a["Bx"].clone()
    F, U, ps, ok = L.linearised_flux(a, b, EOS, 0, return_mask=True)
    assert ok.all()
    VL, VR, Bn = L.side_vectors(a, b, 0)
    A = torch.linalg.solve(*L.jacobians(0.5 * (VL + VR), Bn, EOS, 0))
    assert (torch.linalg.eigvals(A).real > 0).all()
    u, f, _, _ = PH.side_state(a, EOS, 0)
    for k in CONS8:
        assert torch.equal(F[k], f[k]), k


@pytest.mark.parametrize("idir", [0, 1, 2])
def test_a_jump_along_one_eigenvector_is_one_wave(idir):
    """V_R - V_L = eps r_k about a mean state: the ray state is V_R when the
    wave runs left (lambda_k < 0) and V_L when it runs right."""
    s = _states(16, 5 + idir)
    VL0, _, Bn = L.side_vectors(s, s, idir)
    lam, R, ok = L.decompose(torch.linalg.solve(*L.jacobians(VL0, Bn, EOS, idir)))
    assert ok.all()
    vk, bk, _ = L.axes(idir)
    keys = ("rho", "p", vk[0], vk[1], vk[2], bk[1], bk[2])
    hits = 0
    for k in range(7):
        eps = 1e-4 * VL0[0]
        VL = VL0 - 0.5 * eps * R[:, :, k].T
        VR = VL0 + 0.5 * eps * R[:, :, k].T
        sl = {**{key: VL[m] for m, key in enumerate(keys)}, bk[0]: Bn}
        sr = {**{key: VR[m] for m, key in enumerate(keys)}, bk[0]: Bn}
        for d in (sl, sr):
            d["eps"] = EOS.eps__press_rho(d["p"], d["rho"])
        VLs, VRs, Bns = L.side_vectors(sl, sr, idir)
        lam2, R2, ok2 = L.decompose(torch.linalg.solve(*L.jacobians(0.5 * (VLs + VRs), Bns,
                                                                    EOS, idir)))
        Vs, solved = L.ray_state(VLs, VRs, lam2, R2, torch.arange(VLs.shape[1]))
        want = torch.where((lam[:, k] < 0.0)[None, :], VRs, VLs)
        # a wave speed within the linearisation's change of the mean is ambiguous
        clear = (lam[:, k].abs() > 1e-3) & ok2 & solved
        err = ((Vs - want).abs().max(dim=0).values / VLs.abs().max(dim=0).values)[clear]
        assert (err < 1e-8).all(), (k, float(err.max()))
        hits += int(clear.sum())
    assert hits > 80


@pytest.mark.parametrize("idir", [0, 1, 2])
def test_central_difference_jacobians_match_autograd(idir):
    s = _states(8, 9)
    VL, _, Bn = L.side_vectors(s, s, idir)
    dU, dF = L.jacobians(VL, Bn, EOS, idir)
    for e in range(VL.shape[1]):
        v = VL[:, e].clone()
        JU = torch.autograd.functional.jacobian(
            lambda x: L.conserved_and_flux(x[:, None], Bn[e:e + 1], EOS, idir)[0][:, 0], v)
        JF = torch.autograd.functional.jacobian(
            lambda x: L.conserved_and_flux(x[:, None], Bn[e:e + 1], EOS, idir)[1][:, 0], v)
        for a, b in ((dU[e], JU), (dF[e], JF)):
            assert float((a - b).abs().max()) <= 1e-7 * max(1.0, float(b.abs().max()))
