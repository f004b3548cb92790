"""The C-ABI library loads and exports every entry point include/ppf_hip.h
declares (no compute calls: this runs without a GPU)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"^(?:int|int64_t|const char\*)\s+(ppf_\w+)\s*\(", txt, flags=re.M)
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "ppf_warren_step" in names and "ppf_ncomp_step" in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    from pyphasefield_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libppf_hip.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    lib = _lib.lib()
    assert lib.ppf_abi_version() == _lib.ABI_VERSION


def test_struct_layouts_match_header():
    """ctypes mirrors of the C structs have the sizes the C compiler gives."""
    from pyphasefield_amd import _lib
    assert ctypes.sizeof(_lib.Grid) == 4 * 6 + 8 * 3 + 8 * 4
    assert ctypes.sizeof(_lib.Temperature) == 8 + 6 * 8 + 8 + 16
    assert ctypes.sizeof(_lib.WarrenParams) == 17 * 8
    assert ctypes.sizeof(_lib.NcompParams) == 8 + 15 * 8 + 6 * 4 * 8
    assert ctypes.sizeof(_lib.RKPhase) == 8 + 8 + 4 * 5 * 8 + 6 * 3 * 2 * 8
    assert ctypes.sizeof(_lib.Thermo) == 2 * ctypes.sizeof(_lib.RKPhase) + 16
    assert ctypes.sizeof(_lib.TPoly) == 440
    assert ctypes.sizeof(_lib.TdbTables) == 2 * 22 * 440
    from pyphasefield_amd.ensemble import EnsInstance
    assert ctypes.sizeof(EnsInstance) == 88
    from pyphasefield_amd._lib_ext import MlpNet
    assert ctypes.sizeof(MlpNet) == 2 * 8 + 2 * 8 + 6 * 4 + 12 * 4 + 4 * 4 + 8


def test_invalid_arguments_are_rejected_without_gpu_work():
    """Argument validation happens on the host before any launch."""
    from pyphasefield_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libppf_hip.so not built")
    L = _lib.lib()
    g = _lib.Grid()
    g.ndim = 4
    g.dtype = _lib.PPF_F64
    g.nfields = 1
    rc = L.ppf_diffusion_step(g, None, None, None, 1.0, 0.1, None)
    assert rc == 1000
    assert b"ndim" in L.ppf_last_error()
    g.ndim = 2
    g.n[0] = g.n[1] = 4
    g.sy = 3  # smaller than nx + 2
    g.dx = 1.0
    assert L.ppf_diffusion_step(g, None, None, None, 1.0, 0.1, None) == 1000
