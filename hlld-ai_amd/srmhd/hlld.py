"""The flux_fn plugins on the HIP path: drop-in replacements of the
reference's ``hlle_flux`` / ``hllc_flux`` / ``hlld_flux``
(src/physics/hlld.py:1005, 1549, 1202) with the same contract

    flux_fn(sL, sR, eos, idir=0) -> (F, U, p_star)

sL, sR: dicts rho, vx, vy, vz, p, eps, Bx, By, Bz of equal-shape float64
device tensors; F, U: dicts D, tau, Sx, Sy, Sz, Bx, By, Bz; p_star < 0 marks
an HLLE fallback.  ``LAST_DIAG`` is refreshed by every hlld_flux call like
the reference's module global (hlld.py:1516-1529); its values are read from
the device on first access, so a call does not synchronise the stream.

HLLD's star-pressure switches are the reference's module globals, read at
every call (hlld.py:1065-1092) by hlld_flux and by the 1D and 2D drivers:
``_PSTAR_METHOD`` (from RMHD_PSTAR_METHOD: secant, newton, newton_fd,
newton_fd1, newton_safe), ``_PSTAR_BRACKET`` (legacy, hllc), ``_PSTAR_TOL``
and ``_PSTAR_FD_H`` (from RMHD_PSTAR_FD_H; None: the method's default).
An unknown method raises ValueError, as the reference does.
"""

import ctypes
import os

import torch

from . import _lib
from .eos import as_c

PRIM_KEYS = ("rho", "vx", "vy", "vz", "p", "eps", "Bx", "By", "Bz")
CONS_KEYS = ("D", "tau", "Sx", "Sy", "Sz", "Bx", "By", "Bz")

_PSTAR_BRACKET = "legacy"
_PSTAR_TOL = 1.0e-10
_PSTAR_METHOD = os.environ.get("RMHD_PSTAR_METHOD", "secant")
_PSTAR_FD_H = (float(os.environ["RMHD_PSTAR_FD_H"]) if os.environ.get("RMHD_PSTAR_FD_H")
               else None)


def hlld_opts():
    """The current p* switches as the C ABI's srmhd_hlld_opts (ValueError on
    an unknown method or bracket, hlld.py:1342-1344)."""
    if _PSTAR_METHOD not in _lib.PSTAR_METHOD_IDS:
        raise ValueError("RMHD_PSTAR_METHOD=%r: expected secant, newton, newton_safe, newton_fd "
                         "or newton_fd1" % (_PSTAR_METHOD,))
    if _PSTAR_BRACKET not in _lib.PSTAR_BRACKET_IDS:
        raise ValueError(f"_PSTAR_BRACKET={_PSTAR_BRACKET!r}: expected legacy or hllc")
    if not float(_PSTAR_TOL) > 0.0:
        raise ValueError(f"_PSTAR_TOL must be positive, got {_PSTAR_TOL!r}")
    return _lib.HlldOpts(_lib.PSTAR_METHOD_IDS[_PSTAR_METHOD],
                         _lib.PSTAR_BRACKET_IDS[_PSTAR_BRACKET], float(_PSTAR_TOL),
                         float(_PSTAR_FD_H) if _PSTAR_FD_H else 0.0)


def hlld_opts_default(o):
    return o.pstar_method == 0 and o.pstar_bracket == 0 and o.pstar_tol == 1.0e-10


class _LazyDiag(dict):
    """LAST_DIAG whose values come from device counters when first read."""

    def __init__(self):
        super().__init__()
        self._pending = None

    def _set_pending(self, words, solver, idir, **kw):
        self._pending = (words, solver, idir, kw)

    def _fill(self):
        if self._pending is not None:
            words, solver, idir, kw = self._pending
            self._pending = None
            dict.clear(self)
            dict.update(self, _lib.diag_to_dict(words.cpu(), solver, idir, **kw))

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

    def values(self):
        self._fill()
        return dict.values(self)

    def copy(self):
        self._fill()
        return dict(dict.items(self))

    def __repr__(self):
        self._fill()
        return dict.__repr__(self)

    def update(self, *a, **kw):
        self._fill()
        dict.update(self, *a, **kw)


LAST_DIAG = _LazyDiag()


def _flux(solver, sL, sR, eos, idir, want_status=False):
    if idir not in (0, 1, 2):
        raise ValueError(f"idir must be 0, 1 or 2, got {idir!r}")
    L = sL["rho"]
    dev = L.device
    if dev.type != "cuda":
        raise RuntimeError("srmhd flux_fn needs device tensors (the HIP path has no CPU fallback)")
    shape = L.shape
    n = L.numel()
    lib = _lib.lib()
    pl = [_lib.f64(sL[k], dev).reshape(-1) for k in PRIM_KEYS]
    pr = [_lib.f64(sR[k], dev).reshape(-1) for k in PRIM_KEYS]
    for t in pl + pr:
        if t.numel() != n:
            raise ValueError("all state tensors must have the same number of elements")
    F = [torch.empty(n, dtype=torch.float64, device=dev) for _ in CONS_KEYS]
    U = [torch.empty(n, dtype=torch.float64, device=dev) for _ in CONS_KEYS]
    ps = torch.empty(n, dtype=torch.float64, device=dev)
    st = torch.empty(n, dtype=torch.uint8, device=dev)
    diag = torch.zeros(_lib.DIAG_WORDS, dtype=torch.int64, device=dev)
    sid = _lib.SOLVER_IDS[solver]
    ws = _lib.workspace(lib.srmhd_flux_workspace_bytes(sid, n), dev)
    a_l, keep_l = _lib.ptrs(pl)
    a_r, keep_r = _lib.ptrs(pr)
    a_f, keep_f = _lib.ptrs(F)
    a_u, keep_u = _lib.ptrs(U)
    e = as_c(eos)
    o = hlld_opts() if solver == "hlld" else _lib.HlldOpts(0, 0, 1.0e-10, 0.0)
    with torch.cuda.device(dev):
        rc = lib.srmhd_flux_ex(sid, idir, n, a_l, a_r, ctypes.byref(e), ctypes.byref(o), a_f,
                               a_u, ps.data_ptr(), st.data_ptr(), diag.data_ptr(), ws.data_ptr(),
                               ws.numel(), _lib.stream_ptr(dev))
    _lib.check(rc, f"srmhd_flux_ex({solver})")
    if solver == "hlld":
        LAST_DIAG._set_pending(diag, "hlld", idir, pstar_method=_PSTAR_METHOD)
    Fd = {k: F[i].reshape(shape) for i, k in enumerate(CONS_KEYS)}
    Ud = {k: U[i].reshape(shape) for i, k in enumerate(CONS_KEYS)}
    out = (Fd, Ud, ps.reshape(shape))
    return out + (st.reshape(shape),) if want_status else out


def hlle_flux(sL, sR, eos, idir=0):
    """HLLE (hlld.py:1005-1044); p_star is the HLL total-pressure estimate."""
    return _flux("hlle", sL, sR, eos, idir)


def hllc_flux(sL, sR, eos, idir=0):
    """HLLC (hlld.py:1549-1765)."""
    return _flux("hllc", sL, sR, eos, idir)


def hlld_flux(sL, sR, eos, idir=0, lo_fac=0.1, hi_fac=10):
    """HLLD (hlld.py:1202-1543).  ``lo_fac``/``hi_fac`` are accepted and,
    as in the reference, unused."""
    return _flux("hlld", sL, sR, eos, idir)


def flux_with_status(solver, sL, sR, eos, idir=0):
    """(F, U, p_star, status) with per-face status bits: 1 p* not converged,
    2 wave order, 4 B_n = 0, 8 HLLE fallback (HLLD only)."""
    return _flux(solver, sL, sR, eos, idir, want_status=True)


def primitive_to_conserved(prim):
    """hlld.py:185-217 on the device."""
    L = prim["rho"]
    dev = L.device
    if dev.type != "cuda":
        raise RuntimeError("srmhd primitive_to_conserved needs device tensors")
    shape, n = L.shape, L.numel()
    p = [_lib.f64(prim[k], dev).reshape(-1) for k in PRIM_KEYS]
    c = [torch.empty(n, dtype=torch.float64, device=dev) for _ in CONS_KEYS]
    a_p, kp = _lib.ptrs(p)
    a_c, kc = _lib.ptrs(c)
    with torch.cuda.device(dev):
        rc = _lib.lib().srmhd_p2c(n, a_p, a_c, _lib.stream_ptr(dev))
    _lib.check(rc, "srmhd_p2c")
    return {k: c[i].reshape(shape) for i, k in enumerate(CONS_KEYS)}


SOLVERS = {"hlle": hlle_flux, "hllc": hllc_flux, "hlld": hlld_flux}


def solver_name(flux_fn):
    """Map a flux_fn (this package's, the reference's, or a name) to its
    name: hlle, hllc, hlld or linear (linearised_flux with its default HLLD
    fallback, linearised.py:129)."""
    if isinstance(flux_fn, str):
        name = flux_fn
    else:
        name = getattr(flux_fn, "__name__", "")
        name = {"hlle_flux": "hlle", "hllc_flux": "hllc", "hlld_flux": "hlld",
                "linearised_flux": "linear"}.get(name, name)
    if name not in SOLVERS and name != "linear":
        raise ValueError(f"unsupported flux_fn {flux_fn!r}: expected hlle, hllc, hlld or linear")
    return name


def diag_solver(name):
    """The solver whose counters LAST_DIAG reports: the linearised flux's
    are its HLLD fallback's, as in the reference (linearised.py:141)."""
    return "hlld" if name == "linear" else name
