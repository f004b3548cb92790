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
    assert ctypes.sizeof(_lib.Temperature) == 8 + 6 * 8 + 8 + 16 + 8 + 16
    assert ctypes.sizeof(_lib.WarrenParams) == 17 * 8
    assert ctypes.sizeof(_lib.NcompParams) == 8 + 15 * 8 + 6 * 4 * 8
    assert ctypes.sizeof(_lib.RKPhase) == 8 + 8 + 4 * 5 * 8 + 6 * 3 * 2 * 8
    assert ctypes.sizeof(_lib.Thermo) == 2 * ctypes.sizeof(_lib.RKPhase) + 16
    assert ctypes.sizeof(_lib.TPoly) == 440
    assert ctypes.sizeof(_lib.TdbTables) == 2 * (4 + 6 * 5 + 4 * 3) * 440
    from pyphasefield_amd.ensemble import EnsInstance
    assert ctypes.sizeof(EnsInstance) == 88
    from pyphasefield_amd._lib_ext import MlpNet
    assert ctypes.sizeof(MlpNet) == 2 * 8 + 2 * 8 + 6 * 4 + 12 * 4 + 4 * 4 + 8 + 8


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


def _grid2(L, n=(8, 6), bc=(0, 0), nf=1):
    from pyphasefield_amd import _lib
    g = _lib.Grid()
    g.ndim = 2
    g.dtype = _lib.PPF_F32
    g.nfields = nf
    g.n[0], g.n[1], g.n[2] = n[0], n[1], 1
    g.bc[0], g.bc[1], g.bc[2] = bc[0], bc[1], 0
    g.sy = n[0] + 2
    g.sz = 0
    g.sf = (n[0] + 2) * (n[1] + 2)
    g.dx = 1.0
    return g


def test_edge_case_arguments_are_rejected():
    """Empty axes, bad boundary kinds, aliased or missing buffers, too few
    fields, bad reduction kinds, negative counts and malformed ensembles are
    refused with PPF_EINVAL and a message, before any launch (placeholder
    device pointers are never dereferenced)."""
    from pyphasefield_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libppf_hip.so not built")
    L = _lib.lib()
    EINVAL = 1000
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)

    def rejected(rc, word):
        assert rc == EINVAL
        assert word.encode() in L.ppf_last_error(), L.ppf_last_error()

    g = _grid2(L, n=(0, 6))
    rejected(L.ppf_cahnhilliard_step(g, a, b, None, 1.0, 1.0, 0.01, None), "empty axis")
    g = _grid2(L, bc=(0, 7))
    rejected(L.ppf_allencahn_step(g, a, b, None, 1.0, 1.0, 1.0, 0.01, None), "boundary kind")
    g = _grid2(L)
    g.dx = 0.0
    rejected(L.ppf_diffusion_step(g, a, b, None, 1.0, 0.1, None), "dx")
    g = _grid2(L)
    rejected(L.ppf_diffusion_step(g, a, a, None, 1.0, 0.1, None), "distinct")
    rejected(L.ppf_cahnhilliard_step(g, None, b, None, 1.0, 1.0, 0.01, None), "distinct")
    g3 = _grid2(L)
    g3.ndim = 3
    g3.n[2] = 4
    g3.sz = 10  # smaller than a plane
    rejected(L.ppf_cahnhilliard_step(g3, a, b, None, 1.0, 1.0, 0.01, None), "plane stride")
    p = _lib.WarrenParams()
    rejected(L.ppf_warren_step(_grid2(L, nf=1), a, b, None, p, 1, 1, None), "needs fields")
    rejected(L.ppf_warren_step(_grid2(L, nf=2), a, b, None, None, 1, 1, None), "params")
    out = (ctypes.c_double * 2)()
    rejected(L.ppf_reduce(_grid2(L), a, None, 5, 1.0, out, b, None), "kind")
    th = _lib.Thermo()
    rejected(L.ppf_thermo_eval(th, -1, a, a, a, None), "negative")
    shape = (ctypes.c_int32 * 4)()
    rejected(L.ppf_ensemble_config(L.ppf_ensemble_num_configs(), shape), "config")
    seg = (ctypes.c_int32 * 9)(*([0] * 8 + [4]))
    rejected(L.ppf_ensemble_step(a, 1, a, 4, seg, 0, 8, a, a, None), "aliased")
    rejected(L.ppf_ensemble_step(a, 0, a, 4, seg, 0, 8, a, b, None), "empty")
    rejected(L.ppf_ensemble_step(a, 1, a, 4, seg, -1, 8, a, b, None), "tile configuration")
    bad = (ctypes.c_int32 * 9)(0, 3, 2, 4, 4, 4, 4, 4, 4)
    rejected(L.ppf_ensemble_step(a, 1, a, 4, bad, 0, 8, a, b, None), "non-decreasing")
    short = (ctypes.c_int32 * 9)(*([0] * 8 + [3]))
    rejected(L.ppf_ensemble_step(a, 1, a, 4, short, 0, 8, a, b, None), "cover")
    rejected(L.ppf_mlp_gemm(0, 4, 4, a, 1, 1, a, 1, 1, b, 1, 1.0, 0.0, 1, 0, None), "bad shape")
    rejected(L.ppf_mlp_gemm(4, 4, 4, a, 1, 1, a, 1, 1, b, 1, 1.0, 0.0, 1, 3, None), "operand format")


def test_ncomp_flags_and_cahnhilliard_line_arguments_are_rejected():
    """Unknown N-component flags, the legacy flags on the surrogate pass, and
    Cahn-Hilliard line solves off a periodic 1D grid of at least 5 cells are
    refused before any launch."""
    from pyphasefield_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libppf_hip.so not built")
    L = _lib.lib()
    a, b, w = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), ctypes.c_void_p(0x3000)

    def rejected(rc, word):
        assert rc == 1000
        assert word.encode() in L.ppf_last_error(), L.ppf_last_error()

    p = _lib.NcompParams()
    p.ncomp = 3
    p.flags = 16
    g = _grid2(L, nf=5)
    t = _lib.Temperature()
    th = _lib.Thermo()
    rejected(L.ppf_ncomp_fused_step(g, a, b, None, t, p, th, 1, 1, None), "flags")
    tg = _grid2(L, nf=5)
    rejected(L.ppf_ncomp_step(g, a, b, None, t, tg, w, p, 1, 1, None), "flags")
    p.flags = 1
    from pyphasefield_amd._lib_ext import MlpNet
    rejected(L.ppf_ncomp_thermo_mlp(g, a, None, t, tg, w, p, th, MlpNet(), 1, 1, None), "legacy")
    g1 = _grid2(L, n=(4, 1))
    g1.ndim = 1
    rejected(L.ppf_cahnhilliard_line_solve(g1, a, b, 0.1, 1.0, w, None), "5 cells")
    g1 = _grid2(L, n=(16, 1))
    g1.ndim = 1
    rejected(L.ppf_cahnhilliard_line_solve(g1, a, a, 0.1, 1.0, w, None), "differ")
    g2 = _grid2(L)
    rejected(L.ppf_cahnhilliard_line_solve(g2, a, b, 0.1, 1.0, w, None), "1D")
    gn = _grid2(L, n=(16, 1), bc=(1, 0))
    gn.ndim = 1
    rejected(L.ppf_cahnhilliard_line_solve(gn, a, b, 0.1, 1.0, w, None), "PERIODIC")


def test_multistep_limits_and_rejections():
    """ppf_multistep_max: 0 for grids the several-steps kernel does not take
    (3D, a Dirichlet side, bc values), else min(8, each side); the multistep
    entry points refuse nsteps outside 1..max before any launch."""
    from pyphasefield_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libppf_hip.so not built")
    L = _lib.lib()
    EINVAL = 1000
    P, N, D = _lib.PPF_PERIODIC, _lib.PPF_NEUMANN, _lib.PPF_DIRICHLET
    assert L.ppf_multistep_max(_grid2(L, (64, 40), (P, N)), None) == 8
    assert L.ppf_multistep_max(_grid2(L, (5, 40), (P, P)), None) == 5
    assert L.ppf_multistep_max(_grid2(L, (64, 40), (P, D)), None) == 0
    assert L.ppf_multistep_max(_grid2(L, (64, 40), (N, N)), 16) == 0  # bc values
    g3 = _grid2(L, (16, 16), (P, P))
    g3.ndim, g3.n[2], g3.sz = 3, 16, 18 * 18
    g3.sf = g3.sz * 18
    assert L.ppf_multistep_max(g3, None) == 0
    g = _grid2(L, (64, 40), (P, N))
    for k in (0, 9):
        assert L.ppf_diffusion_multistep(g, 16, 32, None, 1.0, 0.1, k, None) == EINVAL
        assert b"nsteps" in L.ppf_last_error()
        assert L.ppf_allencahn_multistep(g, 16, 32, None, 1.0, 1.0, 1.0, 0.01, k, None) == EINVAL
    assert L.ppf_diffusion_multistep(g, 16, 16, None, 1.0, 0.1, 2, None) == EINVAL  # aliased buffers


def _grid3(L, n, bc, dtype):
    g = _grid2(L, (n[0], n[1]), (bc[0], bc[1]))
    g.ndim, g.n[2], g.bc[2], g.dtype = 3, n[2], bc[2], dtype
    g.sz = (n[0] + 2) * (n[1] + 2)
    g.sf = g.sz * (n[2] + 2)
    return g


def test_cahnhilliard_multistep_limits_and_rejections():
    """ppf_cahnhilliard_multistep_max: 2 only for FP32 3D grids with every
    side periodic and at least 8 cells a side, else 1; nsteps outside 1..max
    is refused before any launch."""
    from pyphasefield_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libppf_hip.so not built")
    L = _lib.lib()
    EINVAL = 1000
    P, N = _lib.PPF_PERIODIC, _lib.PPF_NEUMANN
    F32, F64 = _lib.PPF_F32, _lib.PPF_F64
    assert L.ppf_cahnhilliard_multistep_max(_grid3(L, (16, 9, 8), (P, P, P), F32)) == 2
    assert L.ppf_cahnhilliard_multistep_max(_grid3(L, (16, 7, 8), (P, P, P), F32)) == 1  # a side below 8
    assert L.ppf_cahnhilliard_multistep_max(_grid3(L, (16, 16, 16), (P, N, P), F32)) == 1
    assert L.ppf_cahnhilliard_multistep_max(_grid3(L, (16, 16, 16), (P, P, P), F64)) == 1
    assert L.ppf_cahnhilliard_multistep_max(_grid2(L, (64, 64), (P, P))) == 1
    g = _grid3(L, (16, 16, 16), (P, P, P), F32)
    for k in (0, 3):
        assert L.ppf_cahnhilliard_multistep(g, 16, 32, None, 1.0, 1.0, 0.01, k, None) == EINVAL
        assert b"nsteps" in L.ppf_last_error()
    gn = _grid3(L, (16, 16, 16), (P, N, P), F32)
    assert L.ppf_cahnhilliard_multistep(gn, 16, 32, None, 1.0, 1.0, 0.01, 2, None) == EINVAL
    assert L.ppf_cahnhilliard_multistep(g, 16, 16, None, 1.0, 1.0, 0.01, 2, None) == EINVAL  # aliased buffers
