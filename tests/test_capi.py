"""The C-ABI library loads and exports every entry point include/srmhd.h
declares (no GPU compute)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "srmhd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t)\s+(srmhd\w+)\s*\(", txt, re.M)))


def test_header_declares_the_expected_entry_points():
    names = _declared()
    for n in ("srmhd_flux", "srmhd_c2p", "srmhd_p2c", "srmhd2d_stage", "srmhd2d_sync",
              "srmhd2d_dt", "srmhd_flux_workspace_bytes", "srmhd_c2p_workspace_bytes",
              "srmhd2d_workspace_bytes", "srmhd2d_stage_taped", "srmhd2d_stage_adjoint",
              "srmhd2d_sync_adjoint", "srmhd_p2c_adjoint", "srmhd2d_adjoint_workspace_bytes"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from srmhd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load_library()
    for n in _declared():
        assert hasattr(lib, n), n
    assert set(_declared()) == set(_lib.EXPORTED)


def test_workspace_queries_without_gpu():
    from srmhd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load_library()
    assert lib.srmhd_flux_workspace_bytes(2, 1000) >= 1000 * 24
    assert lib.srmhd_flux_workspace_bytes(0, 1000) > 0
    assert lib.srmhd_c2p_workspace_bytes(-1) == -1
    g = _lib.Grid2D(64, 64, 3, 0, 0, 0, 1 / 64, 1 / 64)
    s = _lib.Scheme2D(2, 3, 1, 0, 0, 0, 1e-10, 1e-10)
    nb = lib.srmhd2d_workspace_bytes(ctypes.byref(g), ctypes.byref(s))
    assert nb > 70 * 70 * 8 * 14
    bad = _lib.Grid2D(64, 64, 3, 7, 0, 0, 1 / 64, 1 / 64)
    assert lib.srmhd2d_workspace_bytes(ctypes.byref(bad), ctypes.byref(s)) == -1


def test_product_path_refuses_cpu_tensors():
    import torch
    from srmhd import eos as E, hlld as H
    s = {k: torch.ones(4, dtype=torch.float64) for k in
         ("rho", "vx", "vy", "vz", "p", "eps", "Bx", "By", "Bz")}
    with pytest.raises(RuntimeError):
        H.hlld_flux(s, s, E.hybrid_eos(0.0, 5 / 3, 5 / 3))


def test_adjoint_workspace_query_without_gpu():
    """FP64 only, every limiter; the window adjoints grow with the stencil."""
    from srmhd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load_library()
    g = _lib.Grid2D(64, 64, 4, 1, 1, 0, 0.1, 0.1)   # ng 4: room for MP7
    sizes = []
    for lim in range(6):
        s = _lib.Scheme2D(2, lim, 1, 0, 0, 0, 1e-10, 1e-10)
        sizes.append(lib.srmhd2d_adjoint_workspace_bytes(ctypes.byref(g), ctypes.byref(s)))
    assert all(b > 0 for b in sizes)
    assert sizes[0] < sizes[1] < sizes[3] < sizes[5]
    mixed = _lib.Scheme2D(2, 1, 1, 1, 0, 0, 1e-10, 1e-10)
    assert lib.srmhd2d_adjoint_workspace_bytes(ctypes.byref(g), ctypes.byref(mixed)) == -1
    emf_hll = _lib.Scheme2D(2, 1, 1, 0, 1, 0, 1e-10, 1e-10)   # the adjoint is of the solver EMF
    assert lib.srmhd2d_adjoint_workspace_bytes(ctypes.byref(g), ctypes.byref(emf_hll)) == -1
