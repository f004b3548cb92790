"""Implicit / Crank-Nicolson diffusion without a GPU: the oracle pinned by
closed forms and by its independent spectral formulation, the host-built
coefficient tables of the cyclic line solves, and argument checks of the C
ABI (Engines/Diffusion.py:18-633)."""

import ctypes

import numpy as np
import pytest

from oracle import implicit as oi
from pyphasefield_amd import _lib

EINVAL = 1000  # PPF_EINVAL of the C ABI


def test_matrix_1d_closed_form():
    """diffusion_matrix_1d(5, c, n): c on the diagonal, n on both cyclic
    neighbours, zero elsewhere."""
    want = np.array([[3., -1, 0, 0, -1],
                     [-1, 3, -1, 0, 0],
                     [0, -1, 3, -1, 0],
                     [0, 0, -1, 3, -1],
                     [-1, 0, 0, -1, 3]])
    assert np.array_equal(oi.matrix_1d(5, 3.0, -1.0), want)


def test_matrix_nd_couples_the_periodic_neighbours():
    """Row of cell (z, y, x) = (1, 0, 3) of a (3, 4, 5) grid: neighbours
    (1,0,2), (1,0,4), (1,3,3) (wrap in y), (1,1,3), (0,0,3), (2,0,3)."""
    shape = (3, 4, 5)
    A = oi.matrix_nd(shape, 7.0, -1.0)
    idx = np.ravel_multi_index((1, 0, 3), shape)
    row = A[idx].reshape(shape)
    want = np.zeros(shape)
    want[1, 0, 3] = 7.0
    for nb in [(1, 0, 2), (1, 0, 4), (1, 3, 3), (1, 1, 3), (0, 0, 3), (2, 0, 3)]:
        want[nb] = -1.0
    assert np.array_equal(row, want)
    assert np.array_equal(A, A.T)


@pytest.mark.parametrize("solver", ["implicit", "crank-nicolson"])
@pytest.mark.parametrize("shape,adi", [((11,), False), ((6, 9), False), ((6, 9), True),
                                       ((4, 5, 6), False), ((4, 5, 6), True)])
@pytest.mark.parametrize("a", [0.3, 2.5])
def test_dense_steps_equal_their_fourier_multipliers(solver, shape, adi, a):
    """The reference's solve/inverse sequences equal the Fourier multipliers
    of the same periodic operators (a wrong neighbour, centre, half step or
    sweep axis changes the multiplier)."""
    c = np.random.default_rng(len(shape) + int(10 * a)).random(shape)
    dense = oi.step(c, a, solver, adi)
    spec = oi.spectral_step(c, a, solver, adi)
    assert np.allclose(dense, spec, rtol=0, atol=1e-13)
    assert abs(dense.sum() - c.sum()) <= 1e-12 * c.size  # rows sum to one: mass is conserved
    const = np.full(shape, 0.7)
    assert np.allclose(oi.step(const, a, solver, adi), const, rtol=0, atol=1e-14)


def test_spectral_steps_decay_fourier_modes_by_the_textbook_factors():
    """A single Fourier mode sin(2 pi k x / n) decays by 1/(1 + 4a sin^2(pi k/n))
    per implicit step and by (1 - 4b s)/(1 + 4b s) per Crank-Nicolson step."""
    n, k, a = 32, 3, 1.7
    x = np.arange(n)
    mode = np.sin(2 * np.pi * k * x / n)
    s = np.sin(np.pi * k / n) ** 2
    assert np.allclose(oi.spectral_step(mode, a, "implicit"), mode / (1 + 4 * a * s), atol=1e-14)
    b = a / 2
    assert np.allclose(oi.spectral_step(mode, a, "crank-nicolson"), mode * (1 - 4 * b * s) / (1 + 4 * b * s),
                       atol=1e-14)


# ------------------------------------------------------- coefficient table --
def _table(n, a):
    L = _lib.lib()
    t = np.zeros(int(L.ppf_cyclic_table_len(n)), dtype=np.float64)
    assert L.ppf_cyclic_table(n, a, t.ctypes.data) == 0
    return t


@pytest.mark.parametrize("n,a", [(3, 0.5), (7, 2.0), (64, 0.01), (200, 40.0), (5, 0.0)])
def test_cyclic_table_factorises_the_corner_free_matrix(n, a):
    """The table holds, for M = diffusion_matrix_1d(n, 1+2a, -a) written as
    M' + u v^T (u = (g, 0.., -a), v = (1, 0.., -a/g), g = -(1+2a)): the LU
    factors of M' (L: diagonal 1/den, sub-diagonal -a; U: unit diagonal,
    super-diagonal cp), q = M'^-1 u, r = -a/g and fac = 1/(1 + v.q)."""
    t = _table(n, a)
    assert t.size == 4 + 3 * n
    off, fac, r = t[0], t[1], t[2]
    cp, den, q = t[4:4 + n], t[4 + n:4 + 2 * n], t[4 + 2 * n:]
    M = oi.matrix_1d(n, 1 + 2 * a, -a)
    g = -(1 + 2 * a)
    u = np.zeros(n)
    u[0], u[-1] = g, -a
    v = np.zeros(n)
    v[0], v[-1] = 1.0, -a / g
    Mp = M - np.outer(u, v)
    Lf = np.diag(1.0 / den) + np.diag(np.full(n - 1, -a), -1)
    Uf = np.eye(n) + np.diag(cp[:-1], 1)
    assert off == -a and r == pytest.approx(-a / g, rel=1e-15)
    assert np.allclose(Lf @ Uf, Mp, rtol=0, atol=1e-12 * (1 + a))
    assert np.allclose(Mp @ q, u, rtol=0, atol=1e-12 * (1 + a))
    assert fac == pytest.approx(1.0 / (1.0 + v @ q), rel=1e-12)
    # and the Sherman-Morrison solve with these pieces is the cyclic solve
    d = np.random.default_rng(n).random(n)
    y = np.linalg.solve(Mp, d)
    x = y - (y[0] + r * y[-1]) * fac * q
    assert np.allclose(x, np.linalg.solve(M, d), rtol=0, atol=1e-12)


def test_cyclic_table_rejects_bad_arguments():
    L = _lib.lib()
    t = np.zeros(16)
    assert L.ppf_cyclic_table_len(2) == -1
    assert L.ppf_cyclic_table(2, 1.0, t.ctypes.data) == EINVAL
    assert L.ppf_cyclic_table(4, -1.0, t.ctypes.data) == EINVAL
    assert L.ppf_cyclic_table(4, float("nan"), t.ctypes.data) == EINVAL
    assert L.ppf_cyclic_table(4, 1.0, None) == EINVAL


def _grid(bc, n=(8, 8, 1), ndim=2):
    g = _lib.Grid()
    g.ndim, g.dtype, g.nfields = ndim, _lib.PPF_F64, 1
    for i in range(3):
        g.bc[i] = bc
        g.n[i] = n[i]
    g.sy = n[0] + 2
    g.sz = g.sy * (n[1] + 2)
    g.sf = g.sz * (n[2] + 2)
    g.dx = 1.0
    return g


def test_solver_entry_points_reject_before_touching_the_device():
    L = _lib.lib()
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    tab = ctypes.c_void_p(0x3000)
    g = _grid(_lib.PPF_NEUMANN)
    assert L.ppf_diffusion_line_solve(g, a, b, 0, tab, -1, 0.0, None) == EINVAL
    assert b"periodic" in L.ppf_last_error()
    assert L.ppf_diffusion_cg_begin(g, a, b, 1.0, 0.0, tab, None) == EINVAL
    gp = _grid(_lib.PPF_PERIODIC)
    assert L.ppf_diffusion_line_solve(gp, a, a, 0, tab, -1, 0.0, None) == EINVAL
    assert L.ppf_diffusion_line_solve(gp, a, b, 2, tab, -1, 0.0, None) == EINVAL
    assert L.ppf_diffusion_line_solve(gp, a, b, 0, tab, 2, 0.5, None) == EINVAL
    assert L.ppf_diffusion_cg_begin(gp, a, a, 1.0, 0.0, tab, None) == EINVAL
    assert L.ppf_diffusion_cg_iterate(gp, b, 1.0, 1e-12, -1, 4, tab, None) == EINVAL
    small = _grid(_lib.PPF_PERIODIC, n=(2, 8, 1))
    assert L.ppf_diffusion_line_solve(small, a, b, 0, tab, -1, 0.0, None) == EINVAL
    assert L.ppf_diffusion_cg_work_bytes(gp) == 256 + 4 * 1024 * 8 + 4 * 512


def test_engine_rejects_non_periodic_implicit_runs():
    import pyphasefield_amd.Engines as E
    from pyphasefield_amd.Engines import _implicit
    sim = E.Diffusion(dimensions=[8, 8])
    sim.set_boundary_conditions("NEUMANN")
    sim.set_user_data({"D": 1.0, "solver": "implicit"})
    with pytest.raises(ValueError, match="PERIODIC"):
        _implicit._check(sim)
    sim.set_user_data({"D": 1.0, "solver": "crank-nicolson"})
    assert sim._graph_chunk() == 0


# ------------------------------------------------- Allen-Cahn semi-implicit --
def _uniform_ac(phi0, K, sweeps):
    """Closed form for a spatially uniform field: each sweep solves
    (1 + K s(phi)) phi_new = phi (the neighbour couplings cancel)."""
    p = phi0
    for _ in range(sweeps):
        p = p / (1 + K / sweeps * (4 * p * p - 6 * p + 2))
    return p


@pytest.mark.parametrize("shape,solver,adi,sweeps", [((9,), "implicit", False, 1), ((6, 7), "implicit", False, 1),
                                                     ((6, 7), "implicit", True, 2), ((4, 5, 6), "implicit", True, 3),
                                                     ((6, 7), "crank-nicolson", True, 2)])
def test_allencahn_uniform_state_closed_form(shape, solver, adi, sweeps):
    K = 1.3
    for phi0 in (0.2, 0.75, 0.93):
        got = oi.ac_step(np.full(shape, phi0), 0.7, K, solver, adi)
        assert np.allclose(got, _uniform_ac(phi0, K, sweeps), rtol=1e-14, atol=0)


def test_allencahn_imex_uniform_state_closed_form():
    K, phi0 = 0.4, 0.3
    got = oi.ac_step(np.full((8,), phi0), 0.9, K, "imex")
    assert np.allclose(got, phi0 - K * (4 * phi0 ** 3 - 6 * phi0 ** 2 + 2 * phi0), rtol=1e-14, atol=0)


@pytest.mark.parametrize("shape,solver,adi", [((11,), "implicit", False), ((6, 9), "implicit", False),
                                              ((6, 9), "implicit", True), ((4, 5, 6), "implicit", True),
                                              ((6, 9), "crank-nicolson", True)])
def test_allencahn_without_double_well_is_the_diffusion_step(shape, solver, adi):
    """K = 0 leaves the linear operator: the same Fourier multiplier as the
    diffusion solver with a = M dt eps^2 / dx^2 (Crank-Nicolson: a/2 per half)."""
    c = np.random.default_rng(9).random(shape)
    a = 0.8
    got = oi.ac_step(c, a / 2 if solver == "crank-nicolson" else a, 0.0, solver, adi)
    assert np.allclose(got, oi.spectral_step(c, a, solver, adi), rtol=0, atol=1e-13)


@pytest.mark.parametrize("n,a,m", [(3, 0.5, 1), (37, 2.0, 8), (2048, 4.0, 32), (100, 0.0, 7)])
def test_scan_table_holds_the_factorisation(n, a, m):
    """r solves a r^2 - (1+2a) r + a = 0 with r < 1, so that
    (1+2a) I - a (S + S^T) = (a/r)(I - r S)(I - r S^T); then r/a, r^n and
    the powers r^0..r^m (a = 0: r = 0, r/a -> 1, the identity)."""
    L = _lib.lib()
    t = np.zeros(int(L.ppf_scan_table_len(m)))
    assert L.ppf_scan_table(n, a, m, t.ctypes.data) == 0
    r, scale, rn, mm = t[:4]
    assert mm == m
    if a == 0:
        assert r == 0 and scale == 1
    else:
        assert 0 < r < 1 and abs(a * r * r - (1 + 2 * a) * r + a) <= 1e-15 * (1 + 2 * a)
        assert scale == pytest.approx(r / a, rel=1e-15)
        S = np.roll(np.eye(min(n, 40)), 1, axis=0)
        k = S.shape[0]
        lhs = (a / r) * (np.eye(k) - r * S) @ (np.eye(k) - r * S.T)
        assert np.allclose(lhs, oi.matrix_1d(k, 1 + 2 * a, -a), rtol=0, atol=1e-12 * (1 + a))
    assert rn == pytest.approx(r ** n, rel=1e-12, abs=1e-300)
    assert np.allclose(t[4:], r ** np.arange(m + 1), rtol=1e-14, atol=0)
    assert L.ppf_scan_table(n, a, 0, t.ctypes.data) == EINVAL


def test_new_entry_points_reject_bad_arguments_without_a_device():
    """Allen-Cahn line/CG solves, the scan line solve, the AnisoDorr fused step
    and the work-size queries validate before any launch."""
    L = _lib.lib()
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    w = ctypes.c_void_p(0x3000)
    gp = _grid(_lib.PPF_PERIODIC)
    gn = _grid(_lib.PPF_NEUMANN)
    # Allen-Cahn
    assert L.ppf_allencahn_line_solve(gn, a, b, 0, 1.0, -1, 0.0, 1.0, 0.0, w, None) == EINVAL
    assert L.ppf_allencahn_line_solve(gp, a, a, 0, 1.0, -1, 0.0, 1.0, 0.0, w, None) == EINVAL
    assert L.ppf_allencahn_line_solve(gp, a, b, 0, float("inf"), -1, 0.0, 1.0, 0.0, w, None) == EINVAL
    assert L.ppf_allencahn_line_solve(gp, a, b, 0, 1.0, -1, 0.0, float("nan"), 0.0, w, None) == EINVAL
    assert L.ppf_allencahn_line_solve(gp, a, b, 0, 1.0, -1, 0.0, 1.0, 0.0, None, None) == EINVAL
    assert L.ppf_allencahn_cg_begin(gp, a, a, 1.0, 1.0, w, None) == EINVAL
    assert L.ppf_allencahn_cg_iterate(gp, a, b, 1.0, 1.0, 1e-12, 0, -1, w, None) == EINVAL
    assert L.ppf_allencahn_line_work_bytes(gp) == 2 * gp.sf * 8
    # scan line solve
    assert L.ppf_diffusion_line_solve_scan(gp, a, b, 0, w, 0, -1, 0.0, w, None) == EINVAL  # chunk length
    assert L.ppf_diffusion_line_solve_scan(gp, a, b, 0, w, 32, -1, 0.0, None, None) == EINVAL
    assert L.ppf_diffusion_line_solve_scan(gn, a, b, 0, w, 32, -1, 0.0, w, None) == EINVAL
    assert L.ppf_line_scan_work_bytes(gp, 0, 32) == 2 * 8 * 1 * 8  # 8 lines x 1 chunk x (E, K) doubles
    assert L.ppf_line_scan_work_bytes(gp, 2, 32) == -1
    # AnisoDorr fused step: 4-field stack, 2D only, positive T/R/beta, distinct buffers
    p = _lib.AnisoParams()
    p.T, p.R, p.beta = 1574.0, 8.314, 1.5
    g4 = _grid(_lib.PPF_PERIODIC)
    g4.nfields = 4
    assert L.ppf_aniso_fused_step(gp, a, b, None, p, 0, 1, None) == EINVAL   # 1 field
    assert L.ppf_aniso_fused_step(g4, a, a, None, p, 0, 1, None) == EINVAL   # aliased
    p.beta = 0.0
    assert L.ppf_aniso_fused_step(g4, a, b, None, p, 0, 1, None) == EINVAL
    g3 = _grid(_lib.PPF_PERIODIC, n=(8, 8, 8), ndim=3)
    g3.nfields = 4
    p.beta = 1.5
    assert L.ppf_aniso_fused_step(g3, a, b, None, p, 0, 1, None) == EINVAL   # 3D


def test_engine_registry_exports_the_reference_engine_classes():
    """Engines/__init__.py discovery: every engine class of the reference is
    importable by name; the legacy NCGPU shares NCGPU_new's machinery."""
    import pyphasefield_amd.Engines as E
    for name in ("AnisoDorrGPU", "CahnAllen", "CahnHilliard", "Diffusion", "DiffusionGPU", "NCGPU", "NCGPU_new",
                 "Template", "Warren1995"):
        assert isinstance(getattr(E, name), type), name
    assert issubclass(E.NCGPU, E.NCGPU_new)
    s = E.NCGPU(dimensions=[8, 8])
    assert s.user_data["d_ratio"] == 4.
    s = E.DiffusionGPU(dimensions=[8, 12])
    assert s.user_data["D"] == 0.1 and s.get_framework() == "GPU_SERIAL"
    s.initialize_fields_and_imported_data()
    c = s.fields[0].get_cells()
    assert c.sum() == 4 * 6 and c[2:6, 3:9].min() == 1.0
