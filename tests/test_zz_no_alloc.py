"""srmhd_flux_noalloc, the flux entry point that allocates nothing (its
workspace is a fixed device array inside the library), against
srmhd_flux_ex with a caller's workspace on the same inputs: bit for bit
(fluxes, HLL states, p*, per-face status, the diag words), for every solver,
on two calls in a row with the second smaller -- the scratch keeps the first
call's records, which the second must not see -- and under a non-default
HLLD p* option.  Sizes above the fixed maximum are refused with SRMHD_ESIZE
before anything is launched.
"""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _states(n, seed, idir, dev):
    """Random admissible primitive pairs (rho, v, p, eps, B), |v| < 0.9, one
    normal field per face."""
    g = torch.Generator().manual_seed(seed)

    def side():
        rho = 0.1 + 10.0 * torch.rand(n, generator=g, dtype=torch.float64)
        p = 0.01 + 10.0 * torch.rand(n, generator=g, dtype=torch.float64)
        v = torch.rand(3, n, generator=g, dtype=torch.float64) * 2.0 - 1.0
        v = v * (0.9 * torch.rand(n, generator=g, dtype=torch.float64) /
                 v.norm(dim=0).clamp_min(1e-12))
        B = 3.0 * (torch.rand(3, n, generator=g, dtype=torch.float64) * 2.0 - 1.0)
        eps = p / ((GAMMA - 1.0) * rho)
        return [t.contiguous().to(dev) for t in (rho, v[0], v[1], v[2], p, eps, B[0], B[1], B[2])]

    L, R = side(), side()
    # a few faces with B_n = 0 and identical states (the degenerate branches)
    for k in range(9):
        R[k][:7] = L[k][:7]
    L[6 + idir][7:13] = 0.0
    R[6 + idir][:] = L[6 + idir]
    return L, R


def _run(lib, entry, solver, idir, L, R, opts):
    from srmhd import _lib
    dev = L[0].device
    n = L[0].numel()
    F = [torch.full((n,), float("nan"), dtype=torch.float64, device=dev) for _ in range(8)]
    U = [torch.full((n,), float("nan"), dtype=torch.float64, device=dev) for _ in range(8)]
    ps = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    st = torch.full((n,), 255, dtype=torch.uint8, device=dev)
    diag = torch.full((_lib.DIAG_WORDS,), -1, dtype=torch.int64, device=dev)
    a_l, k1 = _lib.ptrs(L)
    a_r, k2 = _lib.ptrs(R)
    a_f, k3 = _lib.ptrs(F)
    a_u, k4 = _lib.ptrs(U)
    e = _lib.Eos(0.0, GAMMA, GAMMA, 0, 0)
    s = _lib.stream_ptr(dev)
    if entry == "noalloc":
        rc = lib.srmhd_flux_noalloc(solver, idir, n, a_l, a_r, ctypes.byref(e),
                                    ctypes.byref(opts), a_f, a_u, ps.data_ptr(), st.data_ptr(),
                                    diag.data_ptr(), s)
    else:
        ws = _lib.workspace(lib.srmhd_flux_workspace_bytes(solver, n), dev)
        rc = lib.srmhd_flux_ex(solver, idir, n, a_l, a_r, ctypes.byref(e), ctypes.byref(opts),
                               a_f, a_u, ps.data_ptr(), st.data_ptr(), diag.data_ptr(),
                               ws.data_ptr(), ws.numel(), s)
    assert rc == 0, (entry, rc)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in F + U + [ps]], st.cpu().numpy(), diag.cpu().numpy()


def _same(a, b, what):
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    assert same.all(), f"{what}: {int((~same).sum())}/{a.size} differ"


@pytest.mark.parametrize("solver", [0, 1, 2], ids=["hlle", "hllc", "hlld"])
def test_noalloc_flux_equals_the_workspace_flux_on_two_calls(solver):
    dev = _dev()
    from srmhd import _lib
    lib = _lib.lib()
    opts = _lib.HlldOpts(0, 0, 1.0e-10, 0.0)
    for call, (n, seed, idir) in enumerate([(60000, 11, 1), (3001, 12, 0)]):
        print(f"noalloc solver {solver} call {call}: n {n} idir {idir}", flush=True)
        L, R = _states(n, seed, idir, dev)
        got = _run(lib, "noalloc", solver, idir, L, R, opts)
        want = _run(lib, "ws", solver, idir, L, R, opts)
        for i, (a, b) in enumerate(zip(got[0], want[0])):
            _same(a, b, f"call {call} array {i}")
        assert (got[1] == want[1]).all(), f"call {call} status"
        assert (got[2] == want[2]).all(), f"call {call} diag"
        assert np.isfinite(got[0][0]).all()


def test_noalloc_flux_passes_the_pstar_options_through():
    dev = _dev()
    from srmhd import _lib
    lib = _lib.lib()
    opts = _lib.HlldOpts(_lib.PSTAR_METHOD_IDS["newton_safe"], _lib.PSTAR_BRACKET_IDS["hllc"],
                         1.0e-12, 0.0)
    for call, n in enumerate([20000, 999]):
        print(f"noalloc hlld newton_safe/hllc call {call}: n {n}", flush=True)
        L, R = _states(n, 20 + call, 2, dev)
        got = _run(lib, "noalloc", 2, 2, L, R, opts)
        want = _run(lib, "ws", 2, 2, L, R, opts)
        for i, (a, b) in enumerate(zip(got[0], want[0])):
            _same(a, b, f"call {call} array {i}")
        assert (got[1] == want[1]).all() and (got[2] == want[2]).all()


def test_noalloc_flux_refuses_sizes_above_its_scratch():
    dev = _dev()
    from srmhd import _lib
    lib = _lib.lib()
    n = _lib.NOALLOC_MAX_FACES + 1
    null = (ctypes.c_void_p * 9)()   # never read: the size is checked first
    e = _lib.Eos(0.0, GAMMA, GAMMA, 0, 0)
    opts = _lib.HlldOpts(0, 0, 1.0e-10, 0.0)
    rc = lib.srmhd_flux_noalloc(2, 0, n, null, null, ctypes.byref(e), ctypes.byref(opts), null,
                                None, None, None, None, _lib.stream_ptr(dev))
    assert rc == _lib.ESIZE
    with pytest.raises(ValueError, match="faces"):
        _lib.check(rc, "srmhd_flux_noalloc")
    # the empty call is valid and launches only the zeroing of the diag words
    diag = torch.full((_lib.DIAG_WORDS,), -1, dtype=torch.int64, device=dev)
    rc = lib.srmhd_flux_noalloc(2, 0, 0, null, null, ctypes.byref(e), ctypes.byref(opts), null,
                                None, None, None, diag.data_ptr(), _lib.stream_ptr(dev))
    assert rc == 0
    torch.cuda.synchronize()
    assert int(diag.abs().sum()) == 0
