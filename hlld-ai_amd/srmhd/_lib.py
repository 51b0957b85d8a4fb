"""ctypes binding of the C ABI in include/srmhd.h.

The library is loaded from ``_lib/libsrmhd_hip.so`` next to this file.  There
is no fallback: if the library or a HIP device is missing, every entry point
raises ``RuntimeError``.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libsrmhd_hip.so")

SOLVER_IDS = {"hlle": 0, "hllc": 1, "hlld": 2, "linear": 3}
LIMITER_IDS = {"pcm": 0, "mc": 1, "minmod": 2, "weno5z": 3, "mp5": 4, "mp7": 5}
BC_IDS = {"outflow": 0, "periodic": 1, "constant": 2}
PRECISION_IDS = {"fp64": 0, "mixed": 1, "mixed_r64": 2, "fp64_local": 3}
EMF_MODE_IDS = {"solver": 0, "hll": 1}

EARG = 1
ESIZE = 2
EHIP = 1000
NOALLOC_MAX_FACES = 1 << 20


class Eos(ctypes.Structure):
    _fields_ = [("K", ctypes.c_double), ("gamma", ctypes.c_double), ("gamma_th", ctypes.c_double),
                ("is_cold", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class FluxDiag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_ulonglong) for n in (
        "n_interfaces", "n_not_converged", "n_wave_order_bad", "n_bn_zero", "n_hlle_fallback",
        "sum_iters_converged", "n_converged", "max_iters", "c2p_mask", "absF_D_max",
        "n_c2p_redo")]


class Grid2D(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("ng", ctypes.c_int32),
                ("bc_x", ctypes.c_int32), ("bc_y", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double)]


class Grid1D(ctypes.Structure):
    _fields_ = [("ncells", ctypes.c_int32), ("ng", ctypes.c_int32), ("bc", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("dx", ctypes.c_double)]


class HlldOpts(ctypes.Structure):
    _fields_ = [("pstar_method", ctypes.c_int32), ("pstar_bracket", ctypes.c_int32),
                ("pstar_tol", ctypes.c_double), ("pstar_fd_h", ctypes.c_double)]


class LinearOpts(ctypes.Structure):
    _fields_ = [("fallback", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("fd_rel", ctypes.c_double), ("imag_tol", ctypes.c_double),
                ("cond_max", ctypes.c_double), ("hlld", HlldOpts)]


LINEAR_FALLBACK_IDS = {"hlld": 0, "hlle": 1, "hllc": 2}
PSTAR_METHOD_IDS = {"secant": 0, "newton": 1, "newton_fd": 2, "newton_fd1": 3, "newton_safe": 4}
PSTAR_BRACKET_IDS = {"legacy": 0, "hllc": 1}


# srmhd_scheme2d.flags: the FP64 form's y sweep after its x sweep, on the
# caller's stream (default: beside it, on the device's second stream)
STAGE_SERIAL_SWEEPS = 1


class Scheme2D(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int32), ("limiter", ctypes.c_int32), ("upwind", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("emf_mode", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("atmo_rho", ctypes.c_double),
                ("atmo_eps", ctypes.c_double), ("hlld", HlldOpts)]


class Tape2D(ctypes.Structure):
    _fields_ = [("pstar_x", ctypes.c_void_p), ("pstar_y", ctypes.c_void_p),
                ("FD_x", ctypes.c_void_p), ("FD_y", ctypes.c_void_p),
                ("thresh", ctypes.c_void_p)]


class TubeLayout(ctypes.Structure):
    _fields_ = [("off_c", ctypes.c_void_p), ("ncells", ctypes.c_void_p), ("dx", ctypes.c_void_p),
                ("Bn", ctypes.c_void_p), ("order", ctypes.c_void_p),
                ("dx_shared", ctypes.c_double)]


DIAG_WORDS = ctypes.sizeof(FluxDiag) // 8

_lib = None
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_PP = ctypes.POINTER(ctypes.c_void_p)

_SIGS = {
    "srmhd_flux_workspace_bytes": (_I64, [ctypes.c_int, _I64]),
    "srmhd_flux": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _I64, _PP, _PP, ctypes.POINTER(Eos),
                                  _PP, _PP, _P, _P, _P, _P, _I64, _P]),
    "srmhd_flux_ex": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _I64, _PP, _PP,
                                     ctypes.POINTER(Eos), ctypes.POINTER(HlldOpts), _PP, _PP, _P,
                                     _P, _P, _P, _I64, _P]),
    "srmhd_flux_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _I64, _PP, _PP,
                                          ctypes.POINTER(Eos), ctypes.POINTER(HlldOpts), _PP,
                                          _PP, _P, _P, _P, _P]),
    "srmhd_linear_workspace_bytes": (_I64, [_I64]),
    "srmhd_linear_flux": (ctypes.c_int, [ctypes.c_int, _I64, _PP, _PP, ctypes.POINTER(Eos),
                                         ctypes.POINTER(LinearOpts), _PP, _PP, _P, _P, _P, _P,
                                         _P, _I64, _P]),
    "srmhd_c2p_workspace_bytes": (_I64, [_I64]),
    "srmhd_c2p": (ctypes.c_int, [_I64, _PP, ctypes.POINTER(Eos), _D, _D, _PP, _P, _P, _I64, _P]),
    "srmhd_p2c": (ctypes.c_int, [_I64, _PP, _PP, _P]),
    "srmhd2d_workspace_bytes": (_I64, [ctypes.POINTER(Grid2D), ctypes.POINTER(Scheme2D)]),
    "srmhd2d_stage": (ctypes.c_int, [ctypes.POINTER(Grid2D), ctypes.POINTER(Scheme2D),
                                     ctypes.POINTER(Eos), _D, _D, _D, _P, _PP, _PP, _PP, _PP, _PP,
                                     _P, _P, _I64, _P]),
    "srmhd2d_sync": (ctypes.c_int, [ctypes.POINTER(Grid2D), ctypes.POINTER(Scheme2D),
                                    ctypes.POINTER(Eos), _PP, _PP, _P, _I64, _P]),
    "srmhd2d_stage_taped": (ctypes.c_int, [ctypes.POINTER(Grid2D), ctypes.POINTER(Scheme2D),
                                           ctypes.POINTER(Eos), _D, _D, _D, _P, _PP, _PP, _PP,
                                           _PP, _PP, ctypes.POINTER(Tape2D), _P, _I64, _P]),
    "srmhd2d_adjoint_workspace_bytes": (_I64, [ctypes.POINTER(Grid2D),
                                               ctypes.POINTER(Scheme2D)]),
    "srmhd2d_stage_adjoint": (ctypes.c_int, [ctypes.POINTER(Grid2D), ctypes.POINTER(Scheme2D),
                                             ctypes.POINTER(Eos), _D, _D, _D, _P, _PP, _PP, _PP,
                                             ctypes.POINTER(Tape2D), _PP, _PP, _PP, _PP, _PP, _P,
                                             _I64, _P]),
    "srmhd2d_sync_adjoint": (ctypes.c_int, [ctypes.POINTER(Grid2D), ctypes.POINTER(Scheme2D),
                                            ctypes.POINTER(Eos), _PP, _PP, _PP, _PP, _P, _I64,
                                            _P]),
    "srmhd_p2c_adjoint": (ctypes.c_int, [_I64, _PP, _PP, _PP, _P]),
    "srmhd2d_dt": (ctypes.c_int, [ctypes.POINTER(Grid2D), ctypes.POINTER(Eos), _D, _PP, _PP, _P,
                                  _P, _I64, _P]),
    "srmhd1d_workspace_bytes": (_I64, [ctypes.POINTER(Grid1D)]),
    "srmhd1d_rk2_step": (ctypes.c_int, [ctypes.POINTER(Grid1D), ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(Eos), _D, _P, _PP, _PP, _PP, _PP, _P, _P,
                                        _I64, _P]),
    "srmhd1d_rk2_step_ex": (ctypes.c_int, [ctypes.POINTER(Grid1D), ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(Eos), ctypes.POINTER(HlldOpts), _D, _P,
                                           _PP, _PP, _PP, _PP, _P, _P, _I64, _P]),
    "srmhd1d_dt": (ctypes.c_int, [ctypes.POINTER(Grid1D), ctypes.POINTER(Eos), _D, ctypes.c_int,
                                  _PP, _P, _P]),
    "srmhd_tubes_workspace_bytes": (_I64, [ctypes.c_int, _I64]),
    "srmhd_tubes_run": (ctypes.c_int, [ctypes.c_int, _I64, ctypes.POINTER(TubeLayout),
                                       ctypes.POINTER(Eos), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_int, _PP, _PP,
                                       _P, _P, _I64, _P]),
    "srmhd_tubes_clock_bytes": (_I64, [_I64]),
    "srmhd_tubes_shared_step": (ctypes.c_int, [ctypes.c_int, _I64, ctypes.POINTER(TubeLayout),
                                               ctypes.POINTER(Eos), ctypes.c_int, ctypes.c_int,
                                               _D, _D, _D, ctypes.c_int, _PP, _PP, _P, _P, _I64,
                                               _P]),
    "srmhd_timing_enable": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "srmhd_timing_read": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_int64)]),
}
KERNEL_CLASSES = {"sweep_x_a": 0, "sweep_x_b": 1, "sweep_y_a": 2, "sweep_y_b": 3, "corner": 4,
                  "update": 5, "c2p_a": 6, "c2p_b": 7, "adj_sweep_x": 8}
EXPORTED = tuple(_SIGS)


def load_library(path=LIB_PATH):
    """Load the shared library and declare the prototypes (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def lib():
    """The library, for a process that is about to launch kernels."""
    if not torch.cuda.is_available():
        raise RuntimeError("srmhd HIP path needs a HIP device (torch.cuda.is_available() is False)")
    return load_library()


def check(rc, what):
    if rc == 0:
        return
    if rc == EARG:
        raise ValueError(f"{what}: invalid argument")
    if rc == ESIZE:
        raise ValueError(f"{what}: more than {NOALLOC_MAX_FACES} faces for the fixed scratch")
    raise RuntimeError(f"{what}: HIP error {rc - EHIP}")


def ptrs(tensors):
    """A C array of device pointers (kept alive by the returned object)."""
    arr = (ctypes.c_void_p * len(tensors))(*[None if t is None else t.data_ptr() for t in tensors])
    return ctypes.cast(arr, _PP), arr


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def f64(t, device):
    """A contiguous float64 tensor on ``device`` (no copy when already so)."""
    if not torch.is_tensor(t):
        t = torch.as_tensor(t, dtype=torch.float64)
    return t.to(device=device, dtype=torch.float64).contiguous()


def workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


def diag_to_dict(words, solver, idir, iters_counted=True, pstar_method="secant"):
    """FluxDiag words (a host int64 tensor) -> the reference's LAST_DIAG keys
    (hlld.py:1516-1529).  iters_counted False (the FP32-flux forms, whose
    root finder is not the reference's): mean_iters is NaN and max_iters -1,
    and the fallback counts are those of the FP32 solver's decisions."""
    w = [int(x) & 0xFFFFFFFFFFFFFFFF for x in words.tolist()]
    n, nnc, nwo, nbz, nfb, si, ncv, mx = w[:8]
    mean = (si / ncv) if (ncv and iters_counted) else float("nan")
    return dict(solver=solver, idir=idir, n_interfaces=n, n_not_converged=nnc,
                n_wave_order_bad=nwo, n_bn_zero=nbz, n_hlle_fallback=nfb,
                frac_hlle_fallback=nfb / max(n, 1),
                mean_iters=mean, pstar_method=pstar_method, pstar_cost=None,
                max_iters=mx if iters_counted else -1,
                n_c2p_redo=w[10] if len(w) > 10 else 0)
