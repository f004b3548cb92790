"""ctypes binding of libppf_hip.so (declarations in include/ppf_hip.h).

The library is loaded from the package directory; there is no fallback: if it
is missing or fails to load, every engine raises.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PPF_LIB") or os.path.join(HERE, "libppf_hip.so")

PPF_F32, PPF_F64 = 0, 1
PPF_PERIODIC, PPF_NEUMANN, PPF_DIRICHLET = 0, 1, 2
PPF_T_CONST, PPF_T_LINEAR, PPF_T_ARRAY, PPF_T_FRAMES = 0, 1, 2, 3
PPF_DG_FORWARD_DIFF, PPF_DG_ANALYTIC = 0, 1
PPF_MAX_COMP = 4
ABI_VERSION = 1

c_i32, c_i64, c_u64, c_f64, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p


class Grid(ctypes.Structure):
    _fields_ = [("ndim", c_i32), ("dtype", c_i32), ("nfields", c_i32), ("bc", c_i32 * 3),
                ("n", c_i64 * 3), ("sy", c_i64), ("sz", c_i64), ("sf", c_i64), ("dx", c_f64)]


class Temperature(ctypes.Structure):
    _fields_ = [("kind", c_i32), ("pad_", c_i32), ("T0", c_f64), ("dTdx", c_f64), ("dTdy", c_f64),
                ("dTdz", c_f64), ("dTdt", c_f64), ("dt", c_f64), ("data", c_vp), ("sy", c_i64), ("sz", c_i64),
                ("data1", c_vp), ("t_start", c_f64), ("t_end", c_f64)]


class WarrenParams(ctypes.Structure):
    _fields_ = [(n, c_f64) for n in ("T", "T_mA", "T_mB", "L_A", "L_B", "D_L", "D_S", "v_m", "R", "y_e",
                                     "alpha", "ebar", "W_A", "W_B", "M_A", "M_B", "dt")]


_NC_SCALARS = ("dx", "d", "v_m", "M_qmax", "H", "y_e", "beta", "D_L", "D_S", "dt", "ebar", "eqbar",
               "noise_phi", "noise_c", "noise_q")
_NC_VECTORS = ("L", "T_M", "S", "B", "W", "M")


class NcompParams(ctypes.Structure):
    _fields_ = ([("ncomp", c_i32), ("flags", c_i32)] + [(n, c_f64) for n in _NC_SCALARS]
                + [(n, c_f64 * PPF_MAX_COMP) for n in _NC_VECTORS])


class RKPhase(ctypes.Structure):
    _fields_ = [("ncomp", c_i32), ("pad_", c_i32), ("R", c_f64), ("g0", (c_f64 * 5) * PPF_MAX_COMP),
                ("L", ((c_f64 * 2) * 3) * 6)]


PPF_TPOLY_TERMS = 8
PPF_TPOLY_RANGES = 6


class TPoly(ctypes.Structure):
    _fields_ = [("nranges", c_i32), ("pad_", c_i32), ("T_hi", c_f64 * PPF_TPOLY_RANGES),
                ("c", (c_f64 * PPF_TPOLY_TERMS) * PPF_TPOLY_RANGES)]


PPF_RK_ORDERS = 5


class TdbPhase(ctypes.Structure):
    _fields_ = [("g0", TPoly * PPF_MAX_COMP), ("L", (TPoly * PPF_RK_ORDERS) * 6), ("T3", (TPoly * 3) * 4)]


class TdbTables(ctypes.Structure):
    _fields_ = [("liquid", TdbPhase), ("solid", TdbPhase)]


class Thermo(ctypes.Structure):
    _fields_ = [("liquid", RKPhase), ("solid", RKPhase), ("dg_mode", c_i32), ("pad_", c_i32),
                ("tables", ctypes.c_void_p)]


class AnisoParams(ctypes.Structure):
    _fields_ = [(n, c_f64) for n in ("T", "M_qmax", "H", "T_mA", "T_mB", "L_A", "L_B", "D_L", "D_S", "v_m", "R",
                                     "y_e", "beta", "alpha", "ebar", "eqbar", "W_A", "W_B", "M_A", "M_B", "dt")]


_lib = None


def lib():
    """Load (once) and return the library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libppf_hip.so is not built; run `python -m pyphasefield_amd.build` "
                           "(expected at %s)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    sigs = {
        "ppf_abi_version": ([], c_i32),
        "ppf_last_error": ([], ctypes.c_char_p),
        "ppf_apply_bc": ([P(Grid), c_vp, c_vp, c_vp], c_i32),
        "ppf_apply_bc_temperature": ([P(Grid), c_vp, c_vp], c_i32),
        "ppf_fill_temperature": ([P(Grid), c_vp, P(Temperature), c_i64, c_vp], c_i32),
        "ppf_diffusion_step": ([P(Grid), c_vp, c_vp, c_vp, c_f64, c_f64, c_vp], c_i32),
        "ppf_allencahn_step": ([P(Grid), c_vp, c_vp, c_vp, c_f64, c_f64, c_f64, c_f64, c_vp], c_i32),
        "ppf_multistep_max": ([P(Grid), c_vp], c_i32),
        "ppf_diffusion_multistep": ([P(Grid), c_vp, c_vp, c_vp, c_f64, c_f64, c_i32, c_vp], c_i32),
        "ppf_allencahn_multistep": ([P(Grid), c_vp, c_vp, c_vp, c_f64, c_f64, c_f64, c_f64, c_i32, c_vp], c_i32),
        "ppf_cahnhilliard_multistep_max": ([P(Grid)], c_i32),
        "ppf_cahnhilliard_multistep": ([P(Grid), c_vp, c_vp, c_vp, c_f64, c_f64, c_f64, c_i32, c_vp], c_i32),
        "ppf_cahnhilliard_step": ([P(Grid), c_vp, c_vp, c_vp, c_f64, c_f64, c_f64, c_vp], c_i32),
        "ppf_warren_step": ([P(Grid), c_vp, c_vp, c_vp, P(WarrenParams), c_u64, c_i64, c_vp], c_i32),
        "ppf_ncomp_thermo": ([P(Grid), c_vp, c_vp, P(Temperature), P(Grid), c_vp, P(NcompParams), P(Thermo),
                              c_u64, c_i64, c_vp], c_i32),
        "ppf_ncomp_step": ([P(Grid), c_vp, c_vp, c_vp, P(Temperature), P(Grid), c_vp, P(NcompParams), c_u64,
                            c_i64, c_vp], c_i32),
        "ppf_warren_step_dev": ([P(Grid), c_vp, c_vp, c_vp, P(WarrenParams), c_u64, c_vp, c_i64, c_vp], c_i32),
        "ppf_ncomp_thermo_dev": ([P(Grid), c_vp, c_vp, P(Temperature), P(Grid), c_vp, P(NcompParams), P(Thermo),
                                  c_u64, c_vp, c_i64, c_vp], c_i32),
        "ppf_ncomp_step_dev": ([P(Grid), c_vp, c_vp, c_vp, P(Temperature), P(Grid), c_vp, P(NcompParams), c_u64,
                                c_vp, c_i64, c_vp], c_i32),
        "ppf_ncomp_fused_step": ([P(Grid), c_vp, c_vp, c_vp, P(Temperature), P(NcompParams), P(Thermo), c_u64,
                                  c_i64, c_vp], c_i32),
        "ppf_ncomp_fused_step_dev": ([P(Grid), c_vp, c_vp, c_vp, P(Temperature), P(NcompParams), P(Thermo), c_u64,
                                      c_vp, c_i64, c_vp], c_i32),
        "ppf_aniso_fused_step_dev": ([P(Grid), c_vp, c_vp, c_vp, P(AnisoParams), c_u64, c_vp, c_i64, c_vp], c_i32),
        "ppf_step_counter_add": ([c_vp, c_i64, c_vp], c_i32),
        "ppf_reduce_work_bytes": ([P(Grid)], c_i64),
        "ppf_reduce": ([P(Grid), c_vp, c_vp, ctypes.c_int, c_f64, c_vp, c_vp, c_vp], c_i32),
        "ppf_thermo_eval": ([P(Thermo), c_i64, c_vp, c_vp, c_vp, c_vp], c_i32),
        "ppf_aniso_helper": ([P(Grid), c_vp, c_vp, P(Grid), c_vp, P(AnisoParams), c_vp], c_i32),
        "ppf_aniso_step": ([P(Grid), c_vp, c_vp, c_vp, P(Grid), c_vp, P(AnisoParams), c_u64, c_i64, c_vp], c_i32),
        "ppf_aniso_fused_step": ([P(Grid), c_vp, c_vp, c_vp, P(AnisoParams), c_u64, c_i64, c_vp], c_i32),
        "ppf_scan_table_len": ([c_i64], c_i64),
        "ppf_scan_table": ([c_i64, c_f64, c_i64, c_vp], c_i32),
        "ppf_line_scan_work_bytes": ([P(Grid), ctypes.c_int, c_i64], c_i64),
        "ppf_diffusion_line_solve_scan": ([P(Grid), c_vp, c_vp, ctypes.c_int, c_vp, c_i64, ctypes.c_int, c_f64, c_vp,
                                           c_vp], c_i32),
        "ppf_cyclic_table_len": ([c_i64], c_i64),
        "ppf_cyclic_table": ([c_i64, c_f64, c_vp], c_i32),
        "ppf_diffusion_line_solve": ([P(Grid), c_vp, c_vp, ctypes.c_int, c_vp, ctypes.c_int, c_f64, c_vp], c_i32),
        "ppf_diffusion_cg_work_bytes": ([P(Grid)], c_i64),
        "ppf_allencahn_line_work_bytes": ([P(Grid)], c_i64),
        "ppf_allencahn_line_solve": ([P(Grid), c_vp, c_vp, ctypes.c_int, c_f64, ctypes.c_int, c_f64, c_f64, c_f64,
                                      c_vp, c_vp], c_i32),
        "ppf_allencahn_cg_begin": ([P(Grid), c_vp, c_vp, c_f64, c_f64, c_vp, c_vp], c_i32),
        "ppf_allencahn_cg_iterate": ([P(Grid), c_vp, c_vp, c_f64, c_f64, c_f64, c_i64, ctypes.c_int, c_vp, c_vp],
                                     c_i32),
        "ppf_cahnhilliard_line_work_bytes": ([P(Grid)], c_i64),
        "ppf_cahnhilliard_line_solve": ([P(Grid), c_vp, c_vp, c_f64, c_f64, c_vp, c_vp], c_i32),
        "ppf_diffusion_cg_begin": ([P(Grid), c_vp, c_vp, c_f64, c_f64, c_vp, c_vp], c_i32),
        "ppf_diffusion_cg_iterate": ([P(Grid), c_vp, c_f64, c_f64, c_i64, ctypes.c_int, c_vp, c_vp], c_i32),
    }
    optional = {}
    try:
        from . import _lib_ext  # noqa: F401  (additional signatures: surrogate, ensemble)
        optional = dict(_lib_ext.SIGNATURES)
    except ImportError:
        pass
    if hasattr(L, "ppf_reduce_noalloc"):
        optional["ppf_reduce_noalloc"] = ([P(Grid), c_vp, c_vp, ctypes.c_int, c_f64, c_vp, c_vp], c_i32)
    for name, (args, res) in list(sigs.items()) + list(optional.items()):
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    if L.ppf_abi_version() != ABI_VERSION:
        raise RuntimeError("libppf_hip.so ABI version mismatch")
    _lib = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().ppf_last_error().decode(errors="replace")
        raise RuntimeError("%s failed (status %d): %s" % (what, rc, msg))


# ---------------------------------------------------------------- helpers --

BC_CODES = {"PERIODIC": PPF_PERIODIC, "NEUMANN": PPF_NEUMANN, "DIRCHLET": PPF_DIRICHLET,
            "DIRICHLET": PPF_DIRICHLET}


def bc_codes(bc, ndim):
    """Per-axis codes in the order [x, y, z] (ppf_gpu_utils.py:431-436)."""
    if bc is None:
        bc = "PERIODIC"
    if isinstance(bc, str):
        bc = [bc] * ndim
    bc = list(bc)
    if len(bc) < ndim:
        raise ValueError("boundary_conditions needs one entry per axis")
    out = []
    for b in bc[:ndim]:
        key = str(b).upper()
        if key not in BC_CODES:
            raise ValueError("unknown boundary condition %r" % (b,))
        out.append(BC_CODES[key])
    return out


def make_grid(stack, ndim, bc, dx):
    """Grid descriptor of a padded stack tensor [F, (nz+2,) (ny+2,) nx+2]."""
    g = Grid()
    g.ndim = ndim
    import torch
    if stack.dtype == torch.float32:
        g.dtype = PPF_F32
    elif stack.dtype == torch.float64:
        g.dtype = PPF_F64
    else:
        raise TypeError("fields must be float32 or float64")
    if stack.stride(-1) != 1:
        raise ValueError("the x axis must be contiguous")
    g.nfields = stack.shape[0]
    shape = stack.shape[1:]
    if len(shape) != ndim:
        raise ValueError("stack rank does not match ndim")
    n = [shape[-1] - 2, shape[-2] - 2 if ndim >= 2 else 1, shape[-3] - 2 if ndim >= 3 else 1]
    for a in range(3):
        g.n[a] = n[a]
    codes = bc_codes(bc, ndim) + [PPF_PERIODIC] * (3 - ndim)
    for a in range(3):
        g.bc[a] = codes[a]
    g.sy = stack.stride(-2) if ndim >= 2 else 0
    g.sz = stack.stride(-3) if ndim >= 3 else 0
    g.sf = stack.stride(0)
    g.dx = float(dx)
    return g


def alloc_stack(nfields, interior_shape, dtype, device, align_bytes=256):
    """Pitched padded stack [F, *(s+2)] whose interior rows start 256-byte aligned.

    Row pitch is a multiple of the alignment and the storage offset puts padded
    column 1 (the first interior cell) on an aligned address, so every
    wavefront-wide row load of the interior is aligned.
    """
    import torch
    esize = torch.empty((), dtype=dtype).element_size()
    A = max(1, align_bytes // esize)
    padded = [s + 2 for s in interior_shape]
    nx = padded[-1]
    pitch = ((nx + A - 1) // A) * A
    strides = [1]
    if len(padded) >= 2:
        strides.insert(0, pitch)
    if len(padded) >= 3:
        strides.insert(0, pitch * padded[-2])
    per_field = strides[0] * padded[0] if len(padded) >= 2 else pitch
    per_field = ((per_field + A - 1) // A) * A
    total = nfields * per_field + A
    buf = torch.empty(total, dtype=dtype, device=device)
    return buf.as_strided([nfields] + padded, [per_field] + strides, A - 1)


def stream_handle():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def np_dtype(precision):
    return np.float32 if precision == "FP32" else np.float64
