"""Pins of the AnisoDorr vertex terms (oracle.aniso.helper, restating
AnisoDorr_helper_kernel, Engines/AnisoDorrGPU.py:262-375) by the energy they
come from, derived here symbolically and independently of the oracle's
expressions.

With the rotated gradient psi = R(q) grad(phi), R(q) the rotation of
a = q1^2 - q4^2, b = 2 q1 q4, the four-fold anisotropy energy density is

    e = 2 y_e (psi_x^4 + psi_y^4) / |psi|^2        (times ebar^2).

* In the orientation q = (1, 0) the phi fluxes are its grad(phi)-gradient:
  (pf_comp_x, pf_comp_y) = de / d grad(phi).
* The orientation driving terms (t1_temp, t4_temp) are -ebar^2 de/dq along
  the unit circle: the orientation update uses only their component
  orthogonal to q (dq/dt = M_q (t - q (q . t)), AnisoDorrGPU.py:222-232), so
  that component is pinned, t . tau = -ebar^2 de/dq . tau, tau = (-q4, q1).

The helper evaluates them at the cell vertices (i - 1/2, j - 1/2) from
staggered differences and four-cell averages; on smooth fields sampled at two
resolutions the error against the continuum expressions at the vertex
positions must fall by 4 (second order).  A wrong prefactor (4 y_e,
16 ebar^2 y_e), a swapped sign or an x/y index slip of the staggered stencil
leaves an O(1) error and fails.
"""

import numpy as np
import sympy as sp

from oracle import aniso as oa

LX = 1.0
Y_E, EBAR = 0.12, 0.7


def _energy_derivatives():
    px, py, q1, q4, D = sp.symbols("px py q1 q4 D", real=True)
    a, b = q1 ** 2 - q4 ** 2, 2 * q1 * q4
    sx, sy = a * px - b * py, a * py + b * px
    e = 2 * Y_E * (sx ** 4 + sy ** 4) / (sx ** 2 + sy ** 2)
    # driving force along the unit circle: -ebar^2 de/dq . tau, tau = (-q4, q1)
    tang = -EBAR ** 2 * (-q4 * sp.diff(e, q1) + q1 * sp.diff(e, q4))
    fx = sp.diff(e, px).subs({q1: 1, q4: 0})
    fy = sp.diff(e, py).subs({q1: 1, q4: 0})
    mk = lambda ex, args: sp.lambdify(args, ex, "numpy")  # noqa: E731
    return mk(tang, (px, py, q1, q4)), mk(fx, (px, py)), mk(fy, (px, py))


def _phi(x, y):
    return 0.5 + 0.3 * np.sin(2 * np.pi * x / LX) * np.cos(2 * np.pi * y / LX) + 0.2 * x * y


def _phi_grad(x, y):
    k = 2 * np.pi / LX
    gx = 0.3 * k * np.cos(k * x) * np.cos(k * y) + 0.2 * y
    gy = -0.3 * k * np.sin(k * x) * np.sin(k * y) + 0.2 * x
    return gx, gy


def _theta(x, y):
    return 0.6 + 0.4 * np.sin(2 * np.pi * (x + 2 * y) / LX)


def _fields(n, rotate):
    """Padded (n+2)^2 fields; padded cell (i, j) sits at x = (j - 1/2) dx,
    y = (i - 1/2) dx, so vertex (i - 1/2, j - 1/2) sits at ((j-1) dx, (i-1) dx)."""
    dx = LX / n
    jj, ii = np.meshgrid(np.arange(n + 2), np.arange(n + 2))
    x, y = (jj - 0.5) * dx, (ii - 0.5) * dx
    th = _theta(x, y) if rotate else np.zeros_like(x)
    f = np.stack([_phi(x, y), np.cos(th / 2), np.sin(th / 2), np.full_like(x, 0.4)])
    return f, dx


def _vertex_errors(n, rotate, funcs, mutate=None):
    tang, fx, fy = funcs
    f, dx = _fields(n, rotate)
    p = oa.derived_params(dx)
    p.update(y_e=Y_E, ebar=EBAR)
    tr = oa.helper(f, p, dx)
    if mutate is not None:
        tr = mutate(tr)
    # interior vertices only (away from the padded edge)
    sl = (slice(2, n), slice(2, n))
    jj, ii = np.meshgrid(np.arange(n + 2), np.arange(n + 2))
    xv, yv = ((jj - 1) * dx)[sl], ((ii - 1) * dx)[sl]
    gx, gy = _phi_grad(xv, yv)
    th = _theta(xv, yv) if rotate else np.zeros_like(xv)
    q1, q4 = np.cos(th / 2), np.sin(th / 2)
    if rotate:
        # the helper's tangent: that of its four-cell average q
        qa1 = 0.25 * (f[1][1:, 1:] + f[1][:-1, 1:] + f[1][1:, :-1] + f[1][:-1, :-1])
        qa4 = 0.25 * (f[2][1:, 1:] + f[2][:-1, 1:] + f[2][1:, :-1] + f[2][:-1, :-1])
        qn = np.sqrt(qa1 * qa1 + qa4 * qa4)
        # qa[k][l] is the vertex (k + 1/2, l + 1/2), stored by the helper at [k+1][l+1]
        qv1, qv4 = np.zeros_like(tr[2]), np.zeros_like(tr[2])
        qv1[1:, 1:], qv4[1:, 1:] = qa1 / qn, qa4 / qn
        got = -qv4[sl] * tr[2][sl] + qv1[sl] * tr[3][sl]
        return (np.abs(got - tang(gx, gy, q1, q4)).max(),)
    return np.abs(tr[0][sl] - fx(gx, gy)).max(), np.abs(tr[1][sl] - fy(gx, gy)).max()


def test_phi_fluxes_are_energy_gradient_second_order():
    funcs = _energy_derivatives()
    e1 = _vertex_errors(64, False, funcs)
    e2 = _vertex_errors(128, False, funcs)
    for a, b in zip(e1, e2):
        assert 3.2 < a / b < 4.8, (a, b)
        assert b < 2e-3


def test_orientation_terms_are_energy_q_derivatives_second_order():
    funcs = _energy_derivatives()
    e1 = _vertex_errors(64, True, funcs)
    e2 = _vertex_errors(128, True, funcs)
    for a, b in zip(e1, e2):
        assert 3.2 < a / b < 4.8, (a, b)
        assert b < 5e-3  # the terms themselves are O(1) here


def test_gradient_floor_at_flat_vertices():
    """|psi|^2 is floored at 1e-9 (AnisoDorrGPU.py:308-310): a flat field gives
    zero terms, not a division by zero."""
    f, dx = _fields(8, True)
    f[0] = 0.3
    tr = oa.helper(f, oa.derived_params(dx), dx)
    assert np.all(np.isfinite(tr)) and np.all(tr[:4] == 0.0)


def _converges(rotate, funcs, mutate):
    e1 = _vertex_errors(64, rotate, funcs, mutate)
    e2 = _vertex_errors(128, rotate, funcs, mutate)
    return all(3.2 < a / b < 4.8 and b < 5e-3 for a, b in zip(e1, e2))


def _scaled(idx, k):
    def m(tr):
        tr = tr.copy()
        tr[idx] *= k
        return tr
    return m


def _swapped(tr):
    return tr[[1, 0, 2, 3, 4, 5]]


def _negated4(tr):
    tr = tr.copy()
    tr[3] = -tr[3]
    return tr


def test_pins_catch_planted_slips():
    """Slips planted in the helper's output: a halved 4 y_e prefactor, a
    doubled 16 ebar^2 y_e prefactor, x and y fluxes exchanged, a sign flip
    of t4_temp.  Each breaks second-order convergence to the energy."""
    funcs = _energy_derivatives()
    assert not _converges(False, funcs, _scaled(0, 0.5))
    assert not _converges(False, funcs, _swapped)
    assert not _converges(True, funcs, _scaled(2, 2.0))
    assert not _converges(True, funcs, _negated4)
    assert _converges(False, funcs, None) and _converges(True, funcs, None)
