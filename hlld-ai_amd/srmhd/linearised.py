"""The linearised (Roe-type) flux on the HIP path: drop-in replacement of
the reference's ``linearised_flux`` (src/physics/linearised.py:129-237),
the solver ``--solver linear`` selects (scripts/run_2d.py:26-27,
src/physics/driver.py:378-380):

    linearised_flux(sL, sR, eos, idir=0, fallback=hlld_flux, fd_rel=1e-5,
                    imag_tol=1e-6, cond_max=1e12, return_mask=False)
        -> (F, U, p_star) [, mask]

Every face first takes the fallback's answer (HLLD by default, with the p*
switches of ``srmhd.hlld``; ``hlld.LAST_DIAG`` holds its counters, as the
reference's does after the fallback call); the faces the linearisation can
be trusted on then take the physical flux of the state on the ray xi = 0.
``LAST_LINEAR`` holds n, n_linear and frac_fallback (read from the device
on first access).  The fallback must be one of this package's hlld_flux,
hlle_flux or hllc_flux (or their names): it runs inside the same C call.
"""

import ctypes

import torch

from . import _lib
from . import hlld as H
from .eos import as_c

PRIM_KEYS = H.PRIM_KEYS
CONS_KEYS = H.CONS_KEYS


class _LazyLinear(dict):
    """LAST_LINEAR whose n_linear comes from the device when first read."""

    def __init__(self):
        super().__init__()
        self._pending = None

    def _set_pending(self, n, count):
        self._pending = (n, count)

    def _fill(self):
        if self._pending is not None:
            n, count = self._pending
            self._pending = None
            k = int(count.cpu()[0])
            dict.clear(self)
            dict.update(self, n=n, n_linear=k, frac_fallback=1.0 - k / max(n, 1))

    def __getitem__(self, k):
        self._fill()
        return dict.__getitem__(self, k)

    def get(self, k, default=None):
        self._fill()
        return dict.get(self, k, default)

    def __iter__(self):
        self._fill()
        return dict.__iter__(self)

    def __len__(self):
        self._fill()
        return dict.__len__(self)

    def __contains__(self, k):
        self._fill()
        return dict.__contains__(self, k)

    def keys(self):
        self._fill()
        return dict.keys(self)

    def items(self):
        self._fill()
        return dict.items(self)


LAST_LINEAR = _LazyLinear()

_FALLBACKS = {H.hlld_flux: "hlld", H.hlle_flux: "hlle", H.hllc_flux: "hllc",
              "hlld": "hlld", "hlle": "hlle", "hllc": "hllc"}


def linear_opts(fallback="hlld", fd_rel=1e-5, imag_tol=1e-6, cond_max=1e12):
    """srmhd_linear_opts of the arguments (ValueError for a fallback that
    is not one of the HIP solvers)."""
    name = _FALLBACKS.get(fallback)
    if name is None:
        raise ValueError("linearised_flux on the HIP path takes hlld_flux, hlle_flux or hllc_flux "
                         f"as its fallback, got {fallback!r}")
    if not (0.0 < float(fd_rel) < 1.0) or not float(cond_max) > 0.0 or not float(imag_tol) > 0.0:
        raise ValueError("fd_rel must be in (0, 1); imag_tol and cond_max positive")
    hopts = H.hlld_opts() if name == "hlld" else _lib.HlldOpts(0, 0, 1.0e-10, 0.0)
    return _lib.LinearOpts(_lib.LINEAR_FALLBACK_IDS[name], 0, float(fd_rel), float(imag_tol),
                           float(cond_max), hopts)


def linearised_flux(sL, sR, eos, idir=0, fallback=H.hlld_flux, fd_rel=1e-5, imag_tol=1e-6,
                    cond_max=1e12, return_mask=False):
    """linearised.py:129-237: (F, U, p_star) for every face; with
    ``return_mask`` also the bool mask of the faces this solver answered."""
    if idir not in (0, 1, 2):
        raise ValueError(f"idir must be 0, 1 or 2, got {idir!r}")
    o = linear_opts(fallback, fd_rel, imag_tol, cond_max)
    L0 = sL["rho"]
    dev = L0.device
    if dev.type != "cuda":
        raise RuntimeError("srmhd linearised_flux needs device tensors (no CPU fallback)")
    shape = L0.shape
    n = L0.numel()
    lib = _lib.lib()
    pl = [_lib.f64(sL[k], dev).reshape(-1) for k in PRIM_KEYS]
    pr = [_lib.f64(sR[k], dev).reshape(-1) for k in PRIM_KEYS]
    for t in pl + pr:
        if t.numel() != n:
            raise ValueError("all state tensors must have the same number of elements")
    F = [torch.empty(n, dtype=torch.float64, device=dev) for _ in CONS_KEYS]
    U = [torch.empty(n, dtype=torch.float64, device=dev) for _ in CONS_KEYS]
    ps = torch.empty(n, dtype=torch.float64, device=dev)
    mask = torch.empty(n, dtype=torch.uint8, device=dev)
    diag = torch.zeros(_lib.DIAG_WORDS, dtype=torch.int64, device=dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    ws = _lib.workspace(lib.srmhd_linear_workspace_bytes(n), dev)
    a_l, k1 = _lib.ptrs(pl)
    a_r, k2 = _lib.ptrs(pr)
    a_f, k3 = _lib.ptrs(F)
    a_u, k4 = _lib.ptrs(U)
    e = as_c(eos)
    with torch.cuda.device(dev):
        rc = lib.srmhd_linear_flux(idir, n, a_l, a_r, ctypes.byref(e), ctypes.byref(o), a_f, a_u,
                                   ps.data_ptr(), mask.data_ptr(), diag.data_ptr(),
                                   count.data_ptr(), ws.data_ptr(), ws.numel(),
                                   _lib.stream_ptr(dev))
    _lib.check(rc, "srmhd_linear_flux")
    fb = _FALLBACKS[fallback]
    if fb == "hlld":
        H.LAST_DIAG._set_pending(diag, "hlld", idir, pstar_method=H._PSTAR_METHOD)
    LAST_LINEAR._set_pending(n, count)
    Fd = {k: F[i].reshape(shape) for i, k in enumerate(CONS_KEYS)}
    Ud = {k: U[i].reshape(shape) for i, k in enumerate(CONS_KEYS)}
    out = (Fd, Ud, ps.reshape(shape))
    return out + (mask.reshape(shape).bool(),) if return_mask else out
