"""conservative_to_primitive on the HIP path (src/physics/c2p.py:313-363).

The whole input is ONE batch, exactly as in the reference: the secant's
break iteration is shared by every cell of the call (see the device code's
two-pass scheme).
"""

import ctypes

import torch

from . import _lib
from .eos import as_c
from .hlld import PRIM_KEYS

_C2P_IN = ("D", "tau", "Sx", "Sy", "Sz", "Bx", "By", "Bz")
STATUS_BITS = {"bracket_ok": 1, "converged": 2, "v_clamped": 4, "atmo_floored": 8}


def conservative_to_primitive(cons, eos, atmo_rho=1e-10, atmo_eps=1e-10, n_iter=60, tol=1e-15,
                              return_status=False, n_iter_bisect=200):
    if (n_iter, tol, n_iter_bisect) != (60, 1e-15, 200):
        raise ValueError("the HIP con2prim implements n_iter=60, tol=1e-15, n_iter_bisect=200")
    D = cons["D"]
    dev = D.device
    if dev.type != "cuda":
        raise RuntimeError("srmhd conservative_to_primitive needs device tensors")
    shape, n = D.shape, D.numel()
    lib = _lib.lib()
    c = [_lib.f64(cons[k], dev).reshape(-1) for k in _C2P_IN]
    p = [torch.empty(n, dtype=torch.float64, device=dev) for _ in PRIM_KEYS]
    st = torch.empty(n, dtype=torch.uint8, device=dev)
    ws = _lib.workspace(lib.srmhd_c2p_workspace_bytes(n), dev)
    a_c, kc = _lib.ptrs(c)
    a_p, kp = _lib.ptrs(p)
    with torch.cuda.device(dev):
        rc = lib.srmhd_c2p(n, a_c, ctypes.byref(as_c(eos)), float(atmo_rho), float(atmo_eps), a_p,
                           st.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_ptr(dev))
    _lib.check(rc, "srmhd_c2p")
    prims = {k: p[i].reshape(shape) for i, k in enumerate(PRIM_KEYS)}
    if return_status:
        status = {k: ((st & b) != 0).reshape(shape) for k, b in STATUS_BITS.items()}
        return prims, status
    return prims

