"""Pins of the Cahn-Hilliard semi-implicit 1D restatement
(oracle.implicit.ch_matrix_1d / ch_step_1d, CahnHilliard.py:20-70, 148-222):

* the matrix applied to any vector equals the operator
  x - dt M [s_i lap x - eps^2 lap^2 x] built independently from shifted
  copies (the row-own frozen factor s_i = 4c_i^2 - 6c_i + 2 of
  f'(c) = c s(c));
* on a uniform state the system is circulant, and a single Fourier mode k
  decays by exactly 1 / (1 + dt M (lam s0 + eps^2 lam^2)),
  lam = 4 sin^2(pi k / n) / dx^2 -- the linearised Cahn-Hilliard symbol;
* the uniform state is a fixed point (mass conserved); Crank-Nicolson takes
  its explicit half step with dt / 2 before the solve.
"""

import numpy as np
import pytest

from oracle import implicit as oi


def _operator(c, x, alpha, e):
    """x - alpha s_i (x_{i+1} + x_{i-1} - 2 x_i) + alpha e lap^2 x, from rolls."""
    s = oi.ch_s(c)
    lap = np.roll(x, 1) + np.roll(x, -1) - 2 * x
    lap2 = np.roll(lap, 1) + np.roll(lap, -1) - 2 * lap
    return x - alpha * s * lap + alpha * e * lap2


@pytest.mark.parametrize("n", [5, 6, 17])
def test_matrix_rows_are_the_frozen_operator(n):
    rng = np.random.default_rng(n)
    c = rng.uniform(0.1, 0.9, n)
    x = rng.standard_normal(n)
    alpha, e = 0.37, 1.9
    A = oi.ch_matrix_1d(c, alpha, e)
    np.testing.assert_allclose(A @ x, _operator(c, x, alpha, e), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("c0", [0.2, 0.5, 0.75])
def test_single_mode_decays_by_the_linearised_symbol(c0):
    n, k, amp = 64, 5, 1e-3
    M, eps, dt, dx = 0.7, 1.3, 0.05, 0.9
    x = np.arange(n)
    c = c0 + amp * np.cos(2 * np.pi * k * x / n)
    A = oi.ch_matrix_1d(np.full(n, c0), M * dt / dx ** 2, eps ** 2 / dx ** 2)
    out = np.linalg.solve(A, c)
    lam = 4 * np.sin(np.pi * k / n) ** 2 / dx ** 2
    s0 = 4 * c0 ** 2 - 6 * c0 + 2
    g = 1.0 / (1.0 + dt * M * (lam * s0 + eps ** 2 * lam ** 2))
    np.testing.assert_allclose(out - c0, g * (c - c0), rtol=1e-10, atol=1e-15)


def test_uniform_state_is_fixed_and_mass_conserved():
    c = np.full(20, 0.43)
    for solver in ("implicit", "crank-nicolson"):
        out = oi.ch_step_1d(c, 0.5, 1.1, 0.1, 1.0, solver)
        np.testing.assert_allclose(out, c, rtol=1e-14)


def test_crank_nicolson_takes_a_half_explicit_step_first():
    """With eps = 0 and a mode in the linear regime the scheme's factor is the
    product of the explicit half step (1 - (dt/2) M lam s0) and the implicit
    half (1 / (1 + (dt/2) M lam s0))."""
    n, k, amp = 32, 3, 1e-6
    M, dt, dx, c0 = 0.6, 0.02, 1.0, 0.2
    x = np.arange(n)
    c = c0 + amp * np.cos(2 * np.pi * k * x / n)
    out = oi.ch_step_1d(c, M, 0.0, dt, dx, "crank-nicolson")
    lam = 4 * np.sin(np.pi * k / n) ** 2 / dx ** 2
    fp2 = 12 * c0 ** 2 - 12 * c0 + 2        # explicit half: f''(c0) of the full f'(c)
    s0 = 4 * c0 ** 2 - 6 * c0 + 2           # implicit half: the frozen factor
    g = (1 - dt / 2 * M * lam * fp2) / (1 + dt / 2 * M * lam * s0)
    np.testing.assert_allclose(out - c0, g * (c - c0), rtol=1e-5, atol=1e-13)
