"""Implicit and Crank-Nicolson diffusion on the GPU (Engines/Diffusion.py:210-733)
against the oracle: the reference's dense solves at small sizes, the
spectral form of the same operators at larger ones."""

import numpy as np
import pytest

from oracle import implicit as oi
from tests.oracle_runner import rel_l2

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _sim(shape, solver, adi, precision="FP64", a=2.5, gmres=False, seed=0):
    import pyphasefield_amd.Engines as E
    sim = E.Diffusion(dimensions=list(shape))
    sim.set_user_data({"D": 1.0, "solver": solver, "adi": adi, "gmres": gmres})
    sim.set_framework("GPU_SERIAL")
    sim.set_dx(1.0)
    sim.set_dt(a)  # a = D dt / dx^2
    sim.set_boundary_conditions("PERIODIC")
    sim.set_precision(precision)
    sim.initialize_fields_and_imported_data()
    c0 = np.random.default_rng(seed).random(shape)
    sim.fields[0].data[tuple(slice(1, -1) for _ in shape)] = c0
    return sim, c0


CASES = [((40,), False), ((18, 22), False), ((18, 22), True), ((7, 9, 11), False), ((7, 9, 11), True)]


@pytest.mark.parametrize("solver", ["implicit", "crank-nicolson"])
@pytest.mark.parametrize("shape,adi", CASES)
def test_implicit_steps_match_reference_dense_solves(gpu, solver, shape, adi):
    sim, c0 = _sim(shape, solver, adi)
    steps = 3
    sim.simulate(steps)
    want = c0
    for _ in range(steps):
        want = oi.step(want, 2.5, solver, adi)
    assert rel_l2(sim.fields[0].get_cells(), want) <= FP64_TOL


def test_gmres_flag_solves_the_same_system(gpu):
    sim, c0 = _sim((18, 22), "crank-nicolson", False, gmres=True)
    sim.simulate(2)
    want = oi.step(oi.step(c0, 2.5, "crank-nicolson"), 2.5, "crank-nicolson")
    assert rel_l2(sim.fields[0].get_cells(), want) <= FP64_TOL


@pytest.mark.parametrize("solver,adi,shape,a", [("crank-nicolson", False, (384, 640), 3.0),
                                                ("implicit", False, (48, 40, 56), 1.5),
                                                ("implicit", True, (48, 40, 56), 4.0),
                                                ("crank-nicolson", True, (256, 320), 0.7)])
def test_implicit_steps_at_larger_sizes_match_fourier_multipliers(gpu, solver, adi, shape, a):
    """Beyond dense-matrix sizes: the spectral form of the same operators;
    mass is conserved to rounding."""
    sim, c0 = _sim(shape, solver, adi, a=a, seed=3)
    sim.simulate(2)
    got = sim.fields[0].get_cells()
    want = oi.spectral_step(oi.spectral_step(c0, a, solver, adi), a, solver, adi)
    assert rel_l2(got, want) <= FP64_TOL
    assert abs(got.sum() - c0.sum()) <= 1e-11 * abs(c0.sum())


def test_implicit_fp32_conjugate_gradients(gpu):
    sim, c0 = _sim((256, 256), "implicit", False, precision="FP32", a=2.0, seed=5)
    sim.simulate(2)
    want = oi.spectral_step(oi.spectral_step(c0, 2.0, "implicit"), 2.0, "implicit")
    assert rel_l2(sim.fields[0].get_cells(), want) <= 1e-5


def test_fp32_line_solves(gpu):
    sim, c0 = _sim((64, 48, 40), "crank-nicolson", True, precision="FP32", a=1.2, seed=6)
    sim.simulate(2)
    want = oi.spectral_step(oi.spectral_step(c0, 1.2, "crank-nicolson", True), 1.2, "crank-nicolson", True)
    assert rel_l2(sim.fields[0].get_cells(), want) <= 1e-5


def test_conjugate_gradients_stop_early_and_report(gpu):
    """The device-side convergence test stops the iterations (later launches
    do nothing) and records the iteration count; more iterations than needed
    leave the solution unchanged."""
    sim, c0 = _sim((64, 64), "implicit", False, a=0.5, seed=7)
    sim.user_data["cg_check_every"] = 200
    sim.simulate(1)
    it = sim._cg_iterations
    assert 0 < it < 200
    want = oi.spectral_step(c0, 0.5, "implicit")
    assert rel_l2(sim.fields[0].get_cells(), want) <= FP64_TOL


# ------------------------------------------------- Allen-Cahn semi-implicit --
def _ac_sim(shape, solver, adi=False, gmres=False, precision="FP64", M=1.0, W=0.2, eps=1.0, dt=0.5, seed=4):
    import pyphasefield_amd.Engines as E
    sim = E.CahnAllen(dimensions=list(shape))
    sim.set_user_data({"M": M, "W": W, "epsilon": eps, "solver": solver, "adi": adi, "gmres": gmres})
    sim.set_framework("GPU_SERIAL")
    sim.set_dx(1.0)
    sim.set_dt(dt)
    sim.set_boundary_conditions("PERIODIC")
    sim.set_precision(precision)
    sim.initialize_fields_and_imported_data()
    phi0 = 0.5 + 0.35 * (np.random.default_rng(seed).random(shape) - 0.5)
    sim.fields[0].data[tuple(slice(1, -1) for _ in shape)] = phi0
    return sim, phi0


AC_CASES = [((40,), "implicit", False, False), ((40,), "imex", False, False),
            ((18, 22), "implicit", False, True), ((18, 22), "implicit", False, False),
            ((18, 22), "implicit", True, True), ((7, 9, 11), "implicit", False, False),
            ((18, 22), "crank-nicolson", True, False)]


@pytest.mark.parametrize("shape,solver,adi,gmres", AC_CASES)
def test_allencahn_semi_implicit_matches_dense_solves(gpu, shape, solver, adi, gmres):
    sim, phi = _ac_sim(shape, solver, adi, gmres)
    a, K = 1.0 * 0.5 * 1.0, 16 * 0.2 * 1.0 * 0.5
    full = solver == "implicit" and len(shape) == 2 and gmres and not adi
    oadi = len(shape) > 1 and not full
    sim.simulate(3)
    for _ in range(3):
        phi = oi.ac_step(phi, a / 2 if solver == "crank-nicolson" else a, K, solver, oadi)
    assert rel_l2(sim.fields[0].get_cells(), phi) <= FP64_TOL


def test_allencahn_semi_implicit_fp32(gpu):
    sim, phi = _ac_sim((96, 80), "implicit", precision="FP32")
    sim.simulate(3)
    for _ in range(3):
        phi = oi.ac_step(phi, 0.5, 1.6, "implicit", True)
    assert rel_l2(sim.fields[0].get_cells(), phi) <= 1e-5


def test_allencahn_missing_reference_branches_raise(gpu):
    sim, _ = _ac_sim((20,), "crank-nicolson")
    with pytest.raises(NotImplementedError):
        sim.simulate(1)


@pytest.mark.parametrize("shape,solver", [((333, 250), "crank-nicolson"), ((45, 37, 70), "implicit"),
                                          ((1000,), "implicit")])
def test_scan_line_solver_equals_thomas(gpu, shape, solver):
    """The chunked-scan line solver (default) and the one-thread-per-line
    Thomas solver agree to rounding (lengths that are not multiples of the
    32-cell chunk included)."""
    a_, c0 = _sim(shape, solver, True, a=3.0, seed=8)
    b_, _ = _sim(shape, solver, True, a=3.0, seed=8)
    b_.user_data["line_solver"] = "thomas"
    a_.simulate(3)
    b_.simulate(3)
    assert rel_l2(a_.fields[0].get_cells(), b_.fields[0].get_cells()) <= 1e-13
    want = c0
    for _ in range(3):
        want = oi.spectral_step(want, 3.0, solver, True)
    assert rel_l2(a_.fields[0].get_cells(), want) <= FP64_TOL


# ---------------------------------------------- Cahn-Hilliard semi-implicit --
def _ch_sim(shape, solver, adi=False, gmres=False, precision="FP64", M=0.7, W=0.2, eps=1.1, dt=0.3, seed=6):
    import pyphasefield_amd.Engines as E
    sim = E.CahnHilliard(dimensions=list(shape))
    sim.set_user_data({"M": M, "W": W, "epsilon": eps, "solver": solver, "adi": adi, "gmres": gmres})
    sim.set_framework("GPU_SERIAL")
    sim.set_dx(1.0)
    sim.set_dt(dt)
    sim.set_boundary_conditions("PERIODIC")
    sim.set_precision(precision)
    sim.initialize_fields_and_imported_data()
    c0 = 0.5 + 0.35 * (np.random.default_rng(seed).random(shape) - 0.5)
    sim.fields[0].data[tuple(slice(1, -1) for _ in shape)] = c0
    return sim, c0


@pytest.mark.parametrize("solver,gmres,n", [("implicit", False, 64), ("implicit", True, 5),
                                            ("crank-nicolson", False, 64), ("crank-nicolson", True, 301),
                                            ("implicit", False, 2000)])
def test_cahnhilliard_1d_matches_dense_solves(gpu, solver, gmres, n):
    """The pentadiagonal solve (banded elimination + rank-4 corner
    correction) against np.linalg.solve of the reference's matrix, over
    several steps, s refrozen each step; n = 5 is the smallest system whose
    corners do not overlap the band."""
    sim, c = _ch_sim((n,), solver, gmres=gmres)
    steps = 4
    sim.simulate(steps)
    for _ in range(steps):
        c = oi.ch_step_1d(c, 0.7, 1.1, 0.3, 1.0, solver)
    assert rel_l2(sim.fields[0].get_cells(), c) <= FP64_TOL


def test_cahnhilliard_1d_fp32_storage(gpu):
    """FP32 fields: the solve runs in FP64 on FP32 inputs and rounds once."""
    sim, c = _ch_sim((256,), "crank-nicolson", precision="FP32")
    sim.simulate(3)
    for _ in range(3):
        c = oi.ch_step_1d(c, 0.7, 1.1, 0.3, 1.0, "crank-nicolson")
    assert rel_l2(sim.fields[0].get_cells(), c) <= 1e-6


CH_AC_CASES = [((40,), "imex", False, False), ((18, 22), "implicit", False, True),
               ((18, 22), "implicit", False, False), ((18, 22), "implicit", True, True),
               ((7, 9, 11), "implicit", False, False), ((18, 22), "crank-nicolson", True, False)]


@pytest.mark.parametrize("shape,solver,adi,gmres", CH_AC_CASES)
def test_cahnhilliard_other_schemes_run_the_allencahn_bodies(gpu, shape, solver, adi, gmres):
    """imex 1D, implicit 2D/3D and crank-nicolson 2D: the reference's bodies
    (CahnHilliard.py:224-341) are the Allen-Cahn schemes with this engine's
    M, W, epsilon: a = M dt eps^2 / dx^2, K = 16 W M dt."""
    M, W, eps, dt = 0.7, 0.2, 1.1, 0.3
    sim, phi = _ch_sim(shape, solver, adi, gmres, M=M, W=W, eps=eps, dt=dt)
    a, K = M * dt * eps * eps, 16 * W * M * dt
    full = solver == "implicit" and len(shape) == 2 and gmres and not adi
    oadi = len(shape) > 1 and not full
    sim.simulate(3)
    for _ in range(3):
        phi = oi.ac_step(phi, a / 2 if solver == "crank-nicolson" else a, K, solver, oadi)
    assert rel_l2(sim.fields[0].get_cells(), phi) <= FP64_TOL


@pytest.mark.parametrize("shape,solver", [((12, 10), "imex"), ((6, 7, 8), "crank-nicolson"), ((6, 7, 8), "imex")])
def test_cahnhilliard_missing_reference_branches_raise(gpu, shape, solver):
    sim, _ = _ch_sim(shape, solver)
    with pytest.raises(NotImplementedError):
        sim.simulate(1)
