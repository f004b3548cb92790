"""Implicit and Crank-Nicolson diffusion steps on the GPU
(Engines/Diffusion.py:210-633 and the dispatch of :671-733).

The reference's matrices are periodic (diffusion_matrix_1d/2d/3d), so these
solvers require PERIODIC boundaries on every axis.  With a = D*dt/dx^2
(Crank-Nicolson: a/2 for both halves):

* 1D and the alternating-direction variants (user_data["adi"]) are sequences
  of cyclic tridiagonal line solves, each optionally preceded by an explicit
  half step, in the reference's axis order (by default the scan-based solver,
  parallel along each line; user_data["line_solver"] = "thomas" selects one
  thread per line);
* the full 2D/3D systems run matrix-free conjugate gradients to a relative
  residual of user_data["cg_tol"] (default 1e-12 in FP64, 1e-6 in FP32),
  reading the device-side convergence state after 8, 16, 32, ... iterations.

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


# cells per chunk of the scan-based line solve
_SCAN_CHUNK = 32


def _scan_table(sim, n, a, m):
    """Device copy of the scan solver's coefficients for (n, a, m), cached."""
    import torch
    cache = sim.__dict__.setdefault("_scan_tables", {})
    key = (int(n), float(a), int(m))
    t = cache.get(key)
    if t is None:
        L = _lib.lib()
        host = np.zeros(int(L.ppf_scan_table_len(int(m))), dtype=np.float64)
        _lib.check(L.ppf_scan_table(int(n), float(a), int(m), host.ctypes.data), "scan_table")
        t = torch.from_numpy(host).to(sim._fields_gpu_device.device)
        cache[key] = t
    return t


def line_steps(sim, solver, a):
    """user_data["line_solver"]: "scan" (default; parallel along the lines) or
    "thomas" (one thread per line)."""
    nd = len(sim.dimensions)
    am = a if solver == "implicit" else 0.5 * a
    L = _lib.lib()
    scan = sim.user_data.get("line_solver", "scan") != "thomas"
    for axis, e in _LINE_PLANS[(solver, nd)]:
        n = sim._grid.n[axis]
        b = am if e >= 0 else 0.0
        if scan:
            m = int(sim.user_data.get("scan_chunk", _SCAN_CHUNK))
            tab = _scan_table(sim, n, am, m)
            work = _work(sim, "_scan_work_%d" % axis, int(L.ppf_line_scan_work_bytes(sim._grid, axis, m)))
            rc = L.ppf_diffusion_line_solve_scan(sim._grid, _lib.ptr(sim._fields_gpu_device),
                                                 _lib.ptr(sim._fields_out_gpu_device), axis, _lib.ptr(tab), m, e, b,
                                                 _lib.ptr(work), _lib.stream_handle())
            _lib.check(rc, "diffusion_line_solve_scan")
        else:
            tab = _table(sim, n, am)
            rc = L.ppf_diffusion_line_solve(sim._grid, _lib.ptr(sim._fields_gpu_device),
                                            _lib.ptr(sim._fields_out_gpu_device), axis, _lib.ptr(tab), e, b,
                                            _lib.stream_handle())
            _lib.check(rc, "diffusion_line_solve")
        _passive(sim)
        sim._swap_buffers()


def cg_step(sim, solver, a):
    L = _lib.lib()
    am = a if solver == "implicit" else 0.5 * a
    b = 0.0 if solver == "implicit" else am
    work = _work(sim, "_cg_work", int(L.ppf_diffusion_cg_work_bytes(sim._grid)))
    u = sim.user_data
    tol = float(u.get("cg_tol", 1e-12 if sim._precision == "FP64" else 1e-6))
    s = _lib.stream_handle()
    _lib.check(L.ppf_diffusion_cg_begin(sim._grid, _lib.ptr(sim._fields_gpu_device),
                                        _lib.ptr(sim._fields_out_gpu_device), am, b, _lib.ptr(work), s),
               "diffusion_cg_begin")
    out = _lib.ptr(sim._fields_out_gpu_device)
    _cg_loop(sim, lambda k, n: _lib.check(L.ppf_diffusion_cg_iterate(sim._grid, out, am, tol, k, n, _lib.ptr(work),
                                                                        s), "diffusion_cg_iterate"), tol)
    _passive(sim)
    sim._swap_buffers()


def _work(sim, attr, nbytes):
    import torch
    work = getattr(sim, attr, None)
    if work is None or work.numel() != nbytes or work.device != sim._fields_gpu_device.device:
        work = torch.empty(nbytes, dtype=torch.uint8, device=sim._fields_gpu_device.device)
        setattr(sim, attr, work)
    return work


def _cg_loop(sim, iterate, tol):
    """Run CG iterations in chunks until the device state reports convergence.
    The chunks grow (8, 16, 32, ... up to user_data["cg_check_every"], default
    256): few host reads of the state, few iterations launched past
    convergence (those return at once)."""
    u = sim.user_data
    max_it = int(u.get("cg_max_iterations", 5000))
    cap = int(u.get("cg_check_every", 256))
    chunk = min(8, cap)
    work = sim._cg_work
    k = 0
    while True:
        iterate(k, chunk)
        k += chunk
        chunk = min(2 * chunk, cap)
        st = np.frombuffer(work[:_STATE.itemsize].cpu().numpy().tobytes(), dtype=_STATE)[0]
        if st["converged"]:
            sim._cg_iterations = int(st["iterations"])
            return
        if k >= max_it:
            raise RuntimeError("conjugate gradients did not reach a relative residual of %g in %d iterations "
                               "(residual %g)" % (tol, k, np.sqrt(st["rr"] / st["bb"]) if st["bb"] else np.inf))


# Allen-Cahn: (solve axis, explicit axis) and the fraction of K = 16 W M dt per sweep
_AC_PLANS = {
    ("implicit", 2): ([(0, -1), (1, -1)], 0.5),          # CahnAllen.py:179-196
    ("implicit", 3): ([(0, -1), (1, -1), (2, -1)], 1 / 3.),  # :240-265
    ("crank-nicolson", 2): ([(0, 0), (1, 1)], 0.5),       # :198-218
}


def allencahn_step(sim, where="CahnAllen.py:320-362"):
    """One semi-implicit Allen-Cahn step with the reference's dispatch
    (CahnAllen.py:320-362): implicit 1D, 2D (gmres and not adi: full system;
    otherwise alternating directions), 3D (alternating directions); imex 1D;
    crank-nicolson 2D.  The reference's remaining branches do nothing
    (crank-nicolson 1D returns, imex 2D/3D and crank-nicolson 3D print a
    message); they raise here."""
    u = sim.user_data
    solver = u.get("solver", "explicit")
    nd = len(sim.dimensions)
    _check(sim)
    M, W, eps = float(u["M"]), float(u["W"]), float(u["epsilon"])
    dt, dx = float(sim.dt), float(sim.dx)
    a = M * dt * eps * eps / (dx * dx)
    K = 16. * W * M * dt
    L = _lib.lib()
    s = _lib.stream_handle()
    lw = _work(sim, "_ac_line_work", int(L.ppf_allencahn_line_work_bytes(sim._grid)))

    def line(axis, e, aa, b, src, imex):
        rc = L.ppf_allencahn_line_solve(sim._grid, _lib.ptr(sim._fields_gpu_device),
                                        _lib.ptr(sim._fields_out_gpu_device), axis, aa, e, b, src, imex,
                                        _lib.ptr(lw), s)
        _lib.check(rc, "allencahn_line_solve")
        _passive(sim)
        sim._swap_buffers()

    if solver == "implicit" and nd == 1:
        line(0, -1, a, 0.0, K, 0.0)
    elif solver == "imex" and nd == 1:
        line(0, -1, a, 0.0, 0.0, K)
    elif solver == "implicit" and nd == 2 and u.get("gmres", False) and not u.get("adi", False):
        tol = float(u.get("cg_tol", 1e-12 if sim._precision == "FP64" else 1e-6))
        sim._cg_work = _work(sim, "_cg_work", int(L.ppf_diffusion_cg_work_bytes(sim._grid)))
        fin, fout = _lib.ptr(sim._fields_gpu_device), _lib.ptr(sim._fields_out_gpu_device)
        _lib.check(L.ppf_allencahn_cg_begin(sim._grid, fin, fout, a, K, _lib.ptr(sim._cg_work), s),
                   "allencahn_cg_begin")
        _cg_loop(sim, lambda k, n: _lib.check(L.ppf_allencahn_cg_iterate(sim._grid, fin, fout, a, K, tol, k, n,
                                                                            _lib.ptr(sim._cg_work), s),
                                              "allencahn_cg_iterate"), tol)
        _passive(sim)
        sim._swap_buffers()
    elif (solver, nd) in _AC_PLANS:
        plan, frac = _AC_PLANS[(solver, nd)]
        aa = a if solver == "implicit" else 0.5 * a
        for axis, e in plan:
            line(axis, e, aa, aa if e >= 0 else 0.0, K * frac, 0.0)
    else:
        raise NotImplementedError("the reference's engine has no working %r solver for %dD (%s)"
                                  % (solver, nd, where))


def cahnhilliard_step(sim):
    """One semi-implicit Cahn-Hilliard step with the reference's dispatch
    (CahnHilliard.py:382-421).

    * implicit 1D (with or without gmres, :148-174): the periodic
      pentadiagonal system of implicit_matrix_1d (:20-70), alpha = M dt/dx^2,
      e = eps^2/dx^2, s from the current field;
    * crank-nicolson 1D (:176-222): the explicit step with dt/2 (the
      reference omits dt there), then the same solve with dt/2 and s from the
      updated field;
    * implicit 2D/3D, imex 1D, crank-nicolson 2D (:224-341): the bodies of
      these functions are those of the Allen-Cahn engine's (CahnAllen.py:
      165-280), built with a three-argument implicit_matrix_1d /
      implicit_matrix_2d that this module's four- and five-argument builders
      reject; they run here as written, i.e. as the Allen-Cahn schemes with
      this engine's M, W, epsilon (``allencahn_step``);
    * imex 2D/3D and crank-nicolson 3D print a message and do nothing in the
      reference; they raise here."""
    u = sim.user_data
    solver = u.get("solver", "explicit")
    if len(sim.dimensions) != 1 or solver not in ("implicit", "crank-nicolson"):
        allencahn_step(sim, where="CahnHilliard.py:382-421")
        return
    _check(sim)
    if sim.dimensions[0] < 5:
        raise ValueError("the periodic pentadiagonal Cahn-Hilliard system needs at least 5 cells")
    from . import _gpu
    M, eps = float(u["M"]), float(u["epsilon"])
    dt, dx = float(sim.dt), float(sim.dx)
    h = dt
    if solver == "crank-nicolson":
        h = 0.5 * dt
        _gpu.stencil_step(sim, "ppf_cahnhilliard_step", M, eps, h, nactive=1)
    L = _lib.lib()
    work = _work(sim, "_ch_line_work", int(L.ppf_cahnhilliard_line_work_bytes(sim._grid)))
    rc = L.ppf_cahnhilliard_line_solve(sim._grid, _lib.ptr(sim._fields_gpu_device),
                                       _lib.ptr(sim._fields_out_gpu_device), M * h / (dx * dx),
                                       eps * eps / (dx * dx), _lib.ptr(work), _lib.stream_handle())
    _lib.check(rc, "cahnhilliard_line_solve")
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
