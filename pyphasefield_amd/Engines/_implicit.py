"""Implicit and Crank-Nicolson diffusion steps on the GPU
(Engines/Diffusion.py:210-633 and the dispatch of :671-733).

The reference's matrices are periodic (diffusion_matrix_1d/2d/3d), so these
solvers require PERIODIC boundaries on every axis.  With a = D*dt/dx^2
(Crank-Nicolson: a/2 for both halves):

* 1D and the alternating-direction variants (user_data["adi"]) are sequences
  of cyclic tridiagonal line solves, each optionally preceded by an explicit
  half step, in the reference's axis order;
* the full 2D/3D systems run matrix-free conjugate gradients to a relative
  residual of user_data["cg_tol"] (default 1e-12 in FP64, 1e-6 in FP32).

user_data["gmres"] selects scipy's approximate GMRES in the reference; both
settings solve the same system here, to the stated tolerance.
"""

import numpy as np

from .. import _lib

# (solve axis, explicit axis) in ABI axis numbers (0 = x, 1 = y, 2 = z)
_LINE_PLANS = {
    ("implicit", 1): [(0, -1)],
    ("crank-nicolson", 1): [(0, 0)],
    # Diffusion.py:413-433 / :456-483: x lines, then y, then z
    ("implicit", 2): [(0, -1), (1, -1)],
    ("implicit", 3): [(0, -1), (1, -1), (2, -1)],
    # :512-537: explicit y + y lines, then explicit x + x lines
    ("crank-nicolson", 2): [(1, 1), (0, 0)],
    # :565-599: explicit z + y lines, explicit y + z lines, explicit x + x lines
    ("crank-nicolson", 3): [(1, 2), (2, 1), (0, 0)],
}

_STATE = np.dtype([("bb", "<f8"), ("rr", "<f8"), ("iterations", "<i4"), ("converged", "<i4")])


def _check(sim):
    bcs = sim._bc_list()
    if any(b != _lib.PPF_PERIODIC for b in bcs):
        raise ValueError("implicit and Crank-Nicolson diffusion use the reference's periodic matrices "
                         "(Diffusion.py:18-196): set PERIODIC boundary conditions")
    if any(n < 3 for n in sim.dimensions):
        raise ValueError("implicit diffusion needs at least 3 cells per axis")


def _table(sim, n, a):
    """Device copy of the cyclic-system coefficients for (n, a), cached."""
    import torch
    cache = sim.__dict__.setdefault("_implicit_tables", {})
    key = (int(n), float(a))
    t = cache.get(key)
    if t is None:
        L = _lib.lib()
        host = np.zeros(int(L.ppf_cyclic_table_len(int(n))), dtype=np.float64)
        _lib.check(L.ppf_cyclic_table(int(n), float(a), host.ctypes.data), "cyclic_table")
        t = torch.from_numpy(host).to(sim._fields_gpu_device.device)
        cache[key] = t
    return t


def _passive(sim):
    F = sim._fields_gpu_device.shape[0]
    if F > 1:
        sim._fields_out_gpu_device[1:].copy_(sim._fields_gpu_device[1:])


def line_steps(sim, solver, a):
    nd = len(sim.dimensions)
    am = a if solver == "implicit" else 0.5 * a
    L = _lib.lib()
    for axis, e in _LINE_PLANS[(solver, nd)]:
        n = sim._grid.n[axis]
        tab = _table(sim, n, am)
        rc = L.ppf_diffusion_line_solve(sim._grid, _lib.ptr(sim._fields_gpu_device),
                                        _lib.ptr(sim._fields_out_gpu_device), axis, _lib.ptr(tab), e,
                                        am if e >= 0 else 0.0, _lib.stream_handle())
        _lib.check(rc, "diffusion_line_solve")
        _passive(sim)
        sim._swap_buffers()


def cg_step(sim, solver, a):
    import torch
    L = _lib.lib()
    am = a if solver == "implicit" else 0.5 * a
    b = 0.0 if solver == "implicit" else am
    nbytes = int(L.ppf_diffusion_cg_work_bytes(sim._grid))
    work = getattr(sim, "_cg_work", None)
    if work is None or work.numel() != nbytes or work.device != sim._fields_gpu_device.device:
        work = torch.empty(nbytes, dtype=torch.uint8, device=sim._fields_gpu_device.device)
        sim._cg_work = work
    u = sim.user_data
    tol = float(u.get("cg_tol", 1e-12 if sim._precision == "FP64" else 1e-6))
    max_it = int(u.get("cg_max_iterations", 5000))
    chunk = int(u.get("cg_check_every", 32))
    s = _lib.stream_handle()
    _lib.check(L.ppf_diffusion_cg_begin(sim._grid, _lib.ptr(sim._fields_gpu_device),
                                        _lib.ptr(sim._fields_out_gpu_device), am, b, _lib.ptr(work), s),
               "diffusion_cg_begin")
    k = 0
    while True:
        _lib.check(L.ppf_diffusion_cg_iterate(sim._grid, _lib.ptr(sim._fields_out_gpu_device), am, tol, k, chunk,
                                              _lib.ptr(work), s), "diffusion_cg_iterate")
        k += chunk
        st = np.frombuffer(work[:_STATE.itemsize].cpu().numpy().tobytes(), dtype=_STATE)[0]
        if st["converged"]:
            sim._cg_iterations = int(st["iterations"])
            break
        if k >= max_it:
            raise RuntimeError("conjugate gradients did not reach a relative residual of %g in %d iterations "
                               "(residual %g)" % (tol, k, np.sqrt(st["rr"] / st["bb"]) if st["bb"] else np.inf))
    _passive(sim)
    sim._swap_buffers()


def step(sim):
    """One implicit / Crank-Nicolson step (Diffusion.py:671-733 dispatch)."""
    u = sim.user_data
    solver = u.get("solver", "explicit")
    if solver not in ("implicit", "crank-nicolson"):
        raise ValueError("unknown diffusion solver %r" % (solver,))
    _check(sim)
    a = float(u["D"]) * float(sim.dt) / float(sim.dx) ** 2
    if len(sim.dimensions) == 1 or u.get("adi", False):
        line_steps(sim, solver, a)
    else:
        cg_step(sim, solver, a)
