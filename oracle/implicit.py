"""Implicit and Crank-Nicolson diffusion, FP64 numpy restatement (oracle).

Test infrastructure only: the product path never imports this module.

Follows Engines/Diffusion.py:18-633.  The reference works on interior arrays
(axes (z, y, x), x last) with periodic matrices; its defects D1 (`self` for
`sim`) and the padded/interior size mismatch of `c.data` are resolved by
operating on the interior cells.  a = D*dt/dx^2 throughout; Crank-Nicolson
uses b = a/2 for both halves.

Two independent formulations:
  * ``step``: dense matrices built like diffusion_matrix_1d/2d/3d and the
    reference's solve / inverse-per-line sequence (small sizes);
  * ``spectral_step``: the same linear maps as Fourier multipliers (any size),
    which pins the dense construction (neighbour offsets, wrap, centre value,
    axis order of the alternating-direction sweeps) and serves as the oracle
    at sizes where dense matrices do not fit.
"""

import numpy as np


def _shift_sum(n):
    """S + S^T for the cyclic shift S of size n (n >= 3)."""
    e = np.eye(n)
    return np.roll(e, 1, axis=1) + np.roll(e, -1, axis=1)


def matrix_1d(n, center, neighbor):
    """diffusion_matrix_1d (Diffusion.py:18-64): center on the diagonal,
    neighbor at (i, i+-1 mod n)."""
    if n < 3:
        raise ValueError("periodic diffusion matrices need n >= 3")
    return center * np.eye(n) + neighbor * _shift_sum(n)


def matrix_nd(shape, center, neighbor):
    """diffusion_matrix_2d / _3d (Diffusion.py:66-196): row-major flattening
    (x fastest), neighbor coupling along every axis with periodic wrap."""
    if any(n < 3 for n in shape):
        raise ValueError("periodic diffusion matrices need n >= 3 on every axis")
    size = int(np.prod(shape))
    A = center * np.eye(size)
    for ax in range(len(shape)):
        term = np.ones((1, 1))
        for k, n in enumerate(shape):
            term = np.kron(term, _shift_sum(n) if k == ax else np.eye(n))
        A = A + neighbor * term
    return A


def explicit_axis(c, b, axis):
    """(1-2b) c + b (roll(c, 1, axis) + roll(c, -1, axis))  (Diffusion.py:321, 526, 532, 583...)."""
    return (1 - 2 * b) * c + b * (np.roll(c, 1, axis) + np.roll(c, -1, axis))


def explicit_all(c, b):
    """(1-2*nd*b) c + b * sum of the 2*nd rolled copies (Diffusion.py:355, 389)."""
    nd = c.ndim
    s = np.roll(c, 1, 0) + np.roll(c, -1, 0)
    for ax in range(1, nd):
        s = s + np.roll(c, 1, ax) + np.roll(c, -1, ax)
    return (1 - 2 * nd * b) * c + b * s


def _apply_lines(c, inv, axis):
    """np.dot(inv, line) for every line along `axis` (the reference loops)."""
    return np.moveaxis(np.tensordot(inv, np.moveaxis(c, axis, 0), axes=(1, 0)), 0, axis)


def step(c, a, solver, adi=False):
    """One step of the reference's implicit / Crank-Nicolson engines on an
    interior array c (1D, 2D or 3D); np.linalg.solve and gmres variants solve
    the same system, so `gmres` has no counterpart here."""
    nd = c.ndim
    if solver == "implicit":
        if nd == 1:  # engine_ImplicitDiffusion1D, :210-224
            return np.linalg.solve(matrix_1d(c.shape[0], 1 + 2 * a, -a), c)
        if not adi:  # :243-257, :276-290
            A = matrix_nd(c.shape, 1 + 2 * nd * a, -a)
            return np.linalg.solve(A, c.ravel()).reshape(c.shape)
        # :413-433 (x lines, then y lines), :456-483 (x, y, z)
        out = c.copy()
        for axis in range(nd - 1, -1, -1):
            inv = np.linalg.inv(matrix_1d(c.shape[axis], 1 + 2 * a, -a))
            out = _apply_lines(out, inv, axis)
        return out
    if solver == "crank-nicolson":
        b = 0.5 * a
        if nd == 1:  # :309-324
            rhs = explicit_axis(c, b, 0)
            return np.linalg.solve(matrix_1d(c.shape[0], 1 + 2 * b, -b), rhs)
        if not adi:  # :344-358, :378-393
            A = matrix_nd(c.shape, 1 + 2 * nd * b, -b)
            return np.linalg.solve(A, explicit_all(c, b).ravel()).reshape(c.shape)
        inv = [np.linalg.inv(matrix_1d(n, 1 + 2 * b, -b)) for n in c.shape]
        if nd == 2:  # :512-537: explicit axis 0 + lines along axis 0, then axis 1 + lines along 1
            out = _apply_lines(explicit_axis(c, b, 0), inv[0], 0)
            return _apply_lines(explicit_axis(out, b, 1), inv[1], 1)
        # :565-599: explicit axis 0 + lines along 1; explicit 1 + lines along 0;
        # explicit 2 + lines along 2
        out = _apply_lines(explicit_axis(c, b, 0), inv[1], 1)
        out = _apply_lines(explicit_axis(out, b, 1), inv[0], 0)
        return _apply_lines(explicit_axis(out, b, 2), inv[2], 2)
    raise ValueError("solver must be 'implicit' or 'crank-nicolson'")


def _symbols(shape):
    """2 (1 - cos(2 pi k / n)) per axis, broadcast to the FFT grid."""
    out = []
    for ax, n in enumerate(shape):
        s = 2.0 * (1.0 - np.cos(2.0 * np.pi * np.arange(n) / n))
        sh = [1] * len(shape)
        sh[ax] = n
        out.append(s.reshape(sh))
    return out


def _product(factors):
    out = factors[0]
    for f in factors[1:]:
        out = out * f
    return out


def spectral_step(c, a, solver, adi=False):
    """The same step as a Fourier multiplier: with s_ax = 2(1 - cos theta_ax),
    implicit: 1 / (1 + a sum s); implicit ADI: prod 1 / (1 + a s_ax);
    Crank-Nicolson (b = a/2): (1 - b sum s) / (1 + b sum s);
    CN ADI: prod (1 - b s_ax) / (1 + b s_ax)  (all factors commute on a
    periodic grid, so the sweep order only changes rounding)."""
    s = _symbols(c.shape)
    split = adi or c.ndim == 1
    if solver == "implicit":
        g = _product([1.0 / (1.0 + a * x) for x in s]) if split else 1.0 / (1.0 + a * sum(s))
    elif solver == "crank-nicolson":
        b = 0.5 * a
        if split:
            g = _product([(1.0 - b * x) / (1.0 + b * x) for x in s])
        else:
            g = (1.0 - b * sum(s)) / (1.0 + b * sum(s))
    else:
        raise ValueError("solver must be 'implicit' or 'crank-nicolson'")
    return np.real(np.fft.ifftn(np.fft.fftn(c) * g))


# ---------------------------------------------------------------------------
# Allen-Cahn semi-implicit schemes (Engines/CahnAllen.py:18-280)
# ---------------------------------------------------------------------------
# The double-well force f(phi) = 16 W (4 phi^3 - 6 phi^2 + 2 phi) is taken
# implicitly through its linearisation f ~ 16 W s(phi_old) phi_new with
# s(phi) = 4 phi^2 - 6 phi + 2, which puts K s(phi) on the matrix diagonal
# (CahnAllen.py:146-148).  The reference writes that diagonal with
# np.fill_diagonal on matrices built block-wise, so in 2D and 3D it keeps
# only the first row's values; the restatement uses every cell's own value.
# K = 16 W M dt for a full solve, 8 W M dt per sweep of the 2D alternating
# directions (:188, :193), 16 W M dt / 3 per sweep in 3D (:249).

def ac_s(phi):
    return 4. * phi ** 2 - 6. * phi + 2.


def _cyclic_var(n, diag, off):
    A = np.diag(diag) + off * _shift_sum(n)
    return A


def _solve_lines_var(c, a, K, axis):
    """Every line along `axis`: (diag 1 + 2a + K s(c), off -a) x = c."""
    out = np.empty_like(c)
    cm = np.moveaxis(c, axis, -1)
    om = np.moveaxis(out, axis, -1)
    n = cm.shape[-1]
    for idx in np.ndindex(cm.shape[:-1]):
        line = cm[idx]
        om[idx] = np.linalg.solve(_cyclic_var(n, 1 + 2 * a + K * ac_s(line), -a), line)
    return out


def ac_step(phi, a, K, solver, adi=False, imex_K=None):
    """One semi-implicit Allen-Cahn step on an interior array.

    a = M dt eps^2 / dx^2 (Crank-Nicolson: half of it), K = 16 W M dt.
    implicit 1D (:137-149), 2D full (:165-177: the gmres engine), 2D ADI
    (:179-196: rows then columns, K/2 from the current field), 3D ADI
    (:240-265: x, y, z lines, K/3); crank-nicolson 2D ADI (:198-218:
    explicit x then rows, explicit y then columns, K/2 from the field after
    the explicit half); imex 1D (:267-280: rhs = phi - K(4 phi^3 - 6 phi^2 +
    2 phi), the explicit engine's sign of the 2 phi term)."""
    nd = phi.ndim
    if solver == "imex":
        if nd != 1:
            raise ValueError("imex is 1D")
        rhs = phi - K * (4 * phi ** 3 - 6 * phi ** 2 + 2 * phi)
        return np.linalg.solve(matrix_1d(phi.shape[0], 1 + 2 * a, -a), rhs)
    if solver == "implicit":
        if nd == 1:
            return np.linalg.solve(_cyclic_var(phi.shape[0], 1 + 2 * a + K * ac_s(phi), -a), phi)
        if nd == 2 and not adi:
            A = matrix_nd(phi.shape, 0.0, -a) + np.diag((1 + 4 * a + K * ac_s(phi)).ravel())
            return np.linalg.solve(A, phi.ravel()).reshape(phi.shape)
        k = K / 2 if nd == 2 else K / 3
        out = phi
        for axis in range(nd - 1, -1, -1):  # x lines, then y, then z
            out = _solve_lines_var(out, a, k, axis)
        return out
    if solver == "crank-nicolson":
        if nd != 2:
            raise ValueError("crank-nicolson is 2D")
        out = explicit_axis(phi, a, 1)
        out = _solve_lines_var(out, a, K / 2, 1)
        out = explicit_axis(out, a, 0)
        return _solve_lines_var(out, a, K / 2, 0)
    raise ValueError("unknown solver")


# ---------------------------------------------------------------------------
# Cahn-Hilliard semi-implicit 1D (Engines/CahnHilliard.py:20-70, 148-222)
# ---------------------------------------------------------------------------

def ch_s(c):
    """Frozen factor of the linearised chemical potential, f'(c) = c s(c)."""
    return 4 * c ** 2 - 6 * c + 2


def ch_matrix_1d(c, alpha, e2idx2):
    """The periodic pentadiagonal system implicit_matrix_1d builds for
    engine_CahnHilliardImplicit1D (CahnHilliard.py:20-70, 148-160): after its
    sequence of fill_diagonal / roll operations, row i holds its own
    coefficients

      (i, i)       1 + alpha (2 s_i + 6 e)
      (i, i -+ 1)  -alpha (s_i + 4 e)
      (i, i -+ 2)  alpha e            (indices mod n; e = eps^2 / dx^2)

    with s_i = s(c_i) and alpha = M dt / dx^2: the row of
    c - dt M [s_i lap c - eps^2 lap^2 c] at cell i."""
    n = len(c)
    if n < 5:
        raise ValueError("the periodic pentadiagonal system needs n >= 5")
    s = ch_s(c)
    A = np.zeros((n, n))
    i = np.arange(n)
    A[i, i] = 1 + alpha * (2 * s + 6 * e2idx2)
    for d in (1, -1):
        A[i, (i + d) % n] = -alpha * (s + 4 * e2idx2)
    for d in (2, -2):
        A[i, (i + d) % n] = alpha * e2idx2
    return A


def ch_step_1d(c, M, epsilon, dt, dx, solver):
    """One step of engine_CahnHilliardImplicit1D ("implicit") or
    engine_CahnHilliardCrankNicolson1D ("crank-nicolson"), CahnHilliard.py:
    148-200, on the interior cells of a periodic line; the GMRES variants
    (:162-174, :202-222) solve the same systems.  Crank-Nicolson first takes
    the explicit half step c += (dt/2) M lap(mu) (the reference omits dt
    there, defect D3) and then solves with dt/2 and s from the updated c."""
    from . import engines as _oe
    from . import bc as _obc
    c = np.asarray(c, dtype=np.float64)
    h = dt
    if solver == "crank-nicolson":
        h = dt / 2
        p = _obc.pad(c)
        _obc.apply_bc(p[None], "PERIODIC", None, dx)
        c = _oe.cahnhilliard_step(p, M, epsilon, h, dx, "PERIODIC")[1:-1]
    elif solver != "implicit":
        raise ValueError("solver must be 'implicit' or 'crank-nicolson'")
    A = ch_matrix_1d(c, M * h / dx ** 2, epsilon ** 2 / dx ** 2)
    return np.linalg.solve(A, c)
